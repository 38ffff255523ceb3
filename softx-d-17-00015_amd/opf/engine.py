"""Python mirror of the reference's time-step interface over libopf.so.

``Context`` is one resolution class (N, K, M, Mp) holding ``capacity``
device-resident states.  Method names follow the reference procedures they
stand for (velocity.f90, timestep.f90, transform.fftw3.f90, io.f90); spectral
fields are complex numpy arrays ``a[n, h]`` of shape (N, H) (radial point,
harmonic in parallel.h:47-56 order), physical fields real arrays
``p[n, th, z]`` matching the reference's Re(0:3K-1, 0:3M-1, N) layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

LDP, LDM, LDZ, LNP = 0, 1, 2, 3
DIAG_FIELDS = ("step", "t", "E", "Ek0", "Em0", "Ub", "Uc", "Ufr", "ncorr", "dterr")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _split(a, N, H):
    a = np.asarray(a, dtype=np.complex128)
    if a.shape != (N, H):
        raise ValueError(f"field shape {a.shape}, expected {(N, H)}")
    # coll layout Re(1:N, 0:H-1), radial index fastest
    return np.ascontiguousarray(a.real.T), np.ascontiguousarray(a.imag.T)


def _join(re, im):
    return (re + 1j * im).T.copy()


def _path(p):
    return str(p).encode()


def state_file_info(path):
    """Grid and scalars of a state file: N, K, M, Mp, t, Re, alpha, dt."""
    lib = _lib.load()
    iv = [C.c_int() for _ in range(4)]
    dv = [C.c_double() for _ in range(4)]
    check(lib.opf_state_file_info(_path(path), *[C.byref(x) for x in iv], *[C.byref(x) for x in dv]))
    return dict(zip(("N", "K", "M", "Mp"), [x.value for x in iv]),
                **dict(zip(("t", "Re", "alpha", "dt"), [x.value for x in dv])))


def write_state_file(path, r, K, M, Mp, ur, ut, uz, t=0.0, Re=0.0, alpha=0.0, dt=0.0, dtcor=0.0,
                     dtcfl=0.0):
    """io_save_state (io.f90:295-392) of host fields a[n, h] on the radial points r."""
    lib = _lib.load()
    r = np.ascontiguousarray(r, dtype=np.float64)
    N = r.size
    H = (2 * K - 1) * M - (K - 1)
    parts = []
    for a in (ur, ut, uz):
        parts.extend(_split(a, N, H))
    check(lib.opf_state_file_write(_path(path), N, K, M, Mp, _dptr(r), float(t), float(Re), float(alpha),
                                   float(dt), float(dtcor), float(dtcfl), *[_dptr(p) for p in parts]))


def read_state_file(path, N, K, M, Mp=1, r=None):
    """io_load_state (io.f90:98-232) onto the grid (N, K, M, Mp, r) on the host.

    Returns a dict with ur, ut, uz (complex (N, H)) and t, Re, alpha, dt, dtcor,
    dtcfl.  r = None is the default mesh."""
    lib = _lib.load()
    H = (2 * K - 1) * M - (K - 1)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    parts = [np.zeros((H, N)) for _ in range(6)]
    sc = [C.c_double() for _ in range(6)]
    check(lib.opf_state_file_read(_path(path), N, K, M, Mp, _dptr(rr) if rr is not None else None,
                                  *[C.byref(x) for x in sc], *[_dptr(p) for p in parts]))
    out = dict(zip(("t", "Re", "alpha", "dt", "dtcor", "dtcfl"), [x.value for x in sc]))
    out.update(ur=_join(parts[0], parts[1]), ut=_join(parts[2], parts[3]), uz=_join(parts[4], parts[5]))
    return out


class Context:
    def __init__(self, N, K, M, Mp=1, capacity=1, r=None, pairs=0, tm=0, tz=0):
        """One resolution class with `capacity` slots; pairs / tm / tz > 0 fix
        the launch shapes (opf_ctx_create_tiles), 0 takes the tuning table's
        or the built-in choice."""
        self.lib = _lib.load()
        h = C.c_void_p()
        rr = None
        if r is not None:
            rr = np.ascontiguousarray(r, dtype=np.float64)
        check(self.lib.opf_ctx_create_tiles(C.byref(h), N, K, M, Mp, capacity,
                                            _dptr(rr) if rr is not None else None, int(pairs), int(tm), int(tz)))
        self.h = h
        dims = [C.c_int() for _ in range(6)]
        nbytes = C.c_size_t()
        check(self.lib.opf_ctx_dims(h, *[C.byref(d) for d in dims], C.byref(nbytes)))
        self.N, self.K, self.M, self.Mp, self.H, self.capacity = [d.value for d in dims]
        self.bytes = nbytes.value

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.opf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- mesh (mes_D) -------------------------------------------------------
    def mesh(self):
        N = self.N
        r = np.zeros(N)
        intrdr = np.zeros(N)
        dr0 = np.zeros(25)
        dr1 = np.zeros(25)
        d1 = np.zeros(9 * N)
        lap = np.zeros(9 * N)
        check(self.lib.opf_mesh_get(self.h, _dptr(r), _dptr(intrdr), _dptr(dr0), _dptr(dr1),
                                    _dptr(d1), _dptr(lap)))
        return dict(r=r, intrdr=intrdr, dr0=dr0.reshape(5, 5).T, dr1=dr1.reshape(5, 5).T,
                    dr1band=d1.reshape(N, 9).T, radlap=lap.reshape(N, 9).T)

    # -- set-up and state ---------------------------------------------------------
    def setup(self, slot, Re, alpha, dt, implicit=0.51, const_flux=True):
        """vel_matrices (velocity.f90:120-146) for one slot."""
        check(self.lib.opf_slot_setup(self.h, slot, float(Re), float(alpha), float(dt),
                                      float(implicit), int(bool(const_flux))))

    def put_state(self, slot, ur, ut, uz, t=0.0):
        parts = []
        for a in (ur, ut, uz):
            parts.extend(_split(a, self.N, self.H))
        check(self.lib.opf_put_state(self.h, slot, *[_dptr(p) for p in parts], float(t)))

    def get_state(self, slot):
        parts = [np.zeros((self.H, self.N)) for _ in range(6)]
        t = C.c_double()
        check(self.lib.opf_get_state(self.h, slot, *[_dptr(p) for p in parts], C.byref(t)))
        return (_join(parts[0], parts[1]), _join(parts[2], parts[3]), _join(parts[4], parts[5]),
                t.value)

    # -- state files (io.f90) ------------------------------------------------------
    def save_state(self, slot, path):
        """io_save_state (io.f90:295-392) of one slot to a netCDF classic file."""
        check(self.lib.opf_state_save(self.h, slot, _path(path)))

    def load_state(self, slot, path, t=-1.0):
        """io_load_state (io.f90:98-232) into a set-up slot; t < 0 keeps the file's time."""
        check(self.lib.opf_state_load(self.h, slot, _path(path), float(t)))

    # -- Newton-Krylov clients (newton.f90) -----------------------------------------
    def shift(self, slot, dth, dz):
        """var_coll_shift (variables.f90:487-528) of the slot's state."""
        check(self.lib.opf_vel_shift(self.h, slot, float(dth), float(dz)))

    def steporbit(self, slots, ndts, T, sth, sz):
        """steporbit (newton.f90:721-775) for several slots in one launch
        sequence; scalars broadcast over the slots."""
        slots = [int(s) for s in np.atleast_1d(slots)]
        n = len(slots)

        def vec(v, dtype):
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dtype), (n,)))
            return a

        nd = vec(ndts, np.int32)
        Tv, sv, zv = vec(T, np.float64), vec(sth, np.float64), vec(sz, np.float64)
        sl = np.ascontiguousarray(slots, dtype=np.int32)
        ip = C.POINTER(C.c_int)
        check(self.lib.opf_steporbit(self.h, n, sl.ctypes.data_as(ip), nd.ctypes.data_as(ip), _dptr(Tv),
                                     _dptr(sv), _dptr(zv)))

    # -- time loop ------------------------------------------------------------
    def run(self, nsteps, corr_fixed=0, diag_every=0, slots=None):
        """main.f90:18-46 for `nsteps` steps on the device."""
        if slots is None:
            check(self.lib.opf_run(self.h, int(nsteps), int(corr_fixed), int(diag_every)))
        else:
            arr = (C.c_int * len(slots))(*slots)
            check(self.lib.opf_run_slots(self.h, arr, len(slots), int(nsteps), int(corr_fixed),
                                         int(diag_every)))

    def status(self, slot):
        step = C.c_int()
        t = C.c_double()
        st = C.c_int()
        nc = C.c_int()
        de = C.c_double()
        pr = C.c_double()
        check(self.lib.opf_slot_status(self.h, slot, C.byref(step), C.byref(t), C.byref(st),
                                       C.byref(nc), C.byref(de), C.byref(pr)))
        return dict(step=step.value, t=t.value, status=st.value, last_ncorr=nc.value,
                    last_dterr=de.value, Pr0=pr.value)

    def diag(self, slot, maxrec=256):
        buf = np.zeros((maxrec, 10))
        n = C.c_int()
        check(self.lib.opf_diag_get(self.h, slot, _dptr(buf), maxrec, C.byref(n)))
        return buf[:n.value]

    # -- reference procedures, one call each -------------------------------------
    def imposesym(self, slot=0):
        check(self.lib.opf_vel_imposesym(self.h, slot))

    def transform(self, slot=0):
        check(self.lib.opf_vel_transform(self.h, slot))

    def nonlinear(self, slot=0):
        check(self.lib.opf_vel_nonlinear(self.h, slot))

    def predictor(self, slot=0):
        check(self.lib.opf_vel_predictor(self.h, slot))

    def corrector(self, slot=0):
        d = C.c_double()
        check(self.lib.opf_vel_corrector(self.h, slot, C.byref(d)))
        return d.value

    def get_nonlinear(self, slot=0):
        parts = [np.zeros((self.H, self.N)) for _ in range(6)]
        check(self.lib.opf_get_nonlinear(self.h, slot, *[_dptr(p) for p in parts]))
        return (_join(parts[0], parts[1]), _join(parts[2], parts[3]), _join(parts[4], parts[5]))

    def var_norm(self, slot=0, comp=0):
        """var_coll_norm (variables.f90:718-751) of one component of the
        current state: (E, Ek[0:K], Em[0:M])."""
        E = C.c_double()
        Ek = np.zeros(self.K)
        Em = np.zeros(self.M)
        check(self.lib.opf_var_norm(self.h, slot, comp, C.byref(E), _dptr(Ek), _dptr(Em)))
        return E.value, Ek, Em

    def coll_norm(self, field, slot=0):
        """var_coll_norm (variables.f90:718-751) of any field (N x H complex),
        with the slot's alpha, on the device: (E, Ek[0:K], Em[0:M])."""
        re, im = _split(field, self.N, self.H)
        E = C.c_double()
        Ek = np.zeros(self.K)
        Em = np.zeros(self.M)
        check(self.lib.opf_var_coll_norm(self.h, slot, _dptr(re), _dptr(im), C.byref(E), _dptr(Ek), _dptr(Em)))
        return E.value, Ek, Em

    def friction(self, slot=0):
        """io_write_friction (io.f90:627-639): (Ub, Uc, u_tau)."""
        v = [C.c_double() for _ in range(3)]
        check(self.lib.opf_friction(self.h, slot, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def get_phys(self, slot=0, which=0):
        """vel_r/t/z (0..2), vel_curlr/t/z (3..5) of the current state ->
        phys[n, th, z]."""
        Z, T = 3 * self.K, 3 * self.M
        p = np.zeros((self.N, T, Z))
        check(self.lib.opf_get_phys(self.h, slot, which, _dptr(p)))
        return p

    def add_nonlinear(self, slot, nr, nt, nz):
        """vel_N* += (nr, nt, nz): forcing of var_null(2) hooks."""
        parts = []
        for a in (nr, nt, nz):
            parts.extend(_split(a, self.N, self.H))
        check(self.lib.opf_add_nonlinear(self.h, slot, *[_dptr(p) for p in parts]))

    def coll2phys(self, a):
        """tra_coll2phys of one field -> phys[n, th, z]."""
        re, im = _split(a, self.N, self.H)
        Z, T = 3 * self.K, 3 * self.M
        p = np.zeros((self.N, T, Z))
        check(self.lib.opf_tra_coll2phys(self.h, _dptr(re), _dptr(im), _dptr(p)))
        return p

    def phys2coll(self, p):
        Z, T = 3 * self.K, 3 * self.M
        p = np.ascontiguousarray(p, dtype=np.float64)
        if p.shape != (self.N, T, Z):
            raise ValueError("phys shape")
        re = np.zeros((self.H, self.N))
        im = np.zeros((self.H, self.N))
        check(self.lib.opf_tra_phys2coll(self.h, _dptr(p), _dptr(re), _dptr(im)))
        return _join(re, im)

    def lumesh_get(self, slot, which):
        """tim_lumesh_init result: (ab[h, 13, n], ipiv[h, n]) like the reference."""
        ab = np.zeros((self.H, self.N, 13))
        ipiv = np.zeros((self.H, self.N), dtype=np.int32)
        check(self.lib.opf_tim_lumesh_get(self.h, slot, which, _dptr(ab),
                                          ipiv.ctypes.data_as(C.POINTER(C.c_int))))
        return ab.transpose(0, 2, 1).copy(), ipiv

    def lumesh_invert(self, slot, which, b, BC=-1):
        """tim_lumesh_invert(BC, A, b) with factor set `which` of the slot as A
        (BC < 0: the BC the set was built with)."""
        re, im = _split(b, self.N, self.H)
        check(self.lib.opf_tim_lumesh_invert(self.h, slot, which, int(BC), _dptr(re), _dptr(im)))
        return _join(re, im)

    def lumesh_init(self, slot, which, PM, BC, c1, c2):
        """tim_lumesh_init(PM, BC, c1, c2, A) into factor set `which` of the slot."""
        check(self.lib.opf_tim_lumesh_init(self.h, slot, which, int(PM), int(BC), float(c1), float(c2)))

    # -- timing ------------------------------------------------------------------
    def kernel_timing(self, enable=True):
        check(self.lib.opf_kernel_timing(self.h, int(bool(enable))))

    # -- adaptive time step (main.f90:24-29) -----------------------------------------
    def adaptive(self, slot, enable=True, maxdt=1e99, courant=0.5):
        """d_timestep < 0 for the slot: vel_maxtstep + tim_new_tstep (and
        vel_matrices on a change) before every predictor of run()."""
        check(self.lib.opf_slot_adaptive(self.h, slot, int(bool(enable)), float(maxdt), float(courant)))

    def symmetry(self, slot, mirror=False, shiftrefl=False, shiftrott=False):
        """b_mirrorsym, b_shiftrefl, b_shiftrott: the symmetries vel_imposesym
        imposes at the start of every step (var_imposesym modes 1, 2, 3)."""
        check(self.lib.opf_slot_symmetry(self.h, slot, int(bool(mirror)), int(bool(shiftrefl)),
                                         int(bool(shiftrott))))

    def coriolis(self, slot, Ek0=None, phi=0.0):
        """The Coriolis plugin (Coriolis.f90:54-71) for the slot: Ekman number
        Ek0 and angle phi between pipe and rotation axes; Ek0=None removes it."""
        on = Ek0 is not None
        check(self.lib.opf_slot_coriolis(self.h, slot, int(on), float(Ek0 or 0.0), float(phi)))

    def nonnewtonian(self, slot, rho=None, radius=1.0, mu_0=1.0, mu_inf=0.0, lam_CY=0.0, n_CY=0.0, a_CY=2.0):
        """The non-Newtonian plugin (nonnewtonian.f90, Carreau-Yasuda model of
        params_model.in) for the slot; rho=None removes it."""
        on = rho is not None
        check(self.lib.opf_slot_nonnewtonian(self.h, slot, int(on), float(rho or 0.0), float(radius), float(mu_0),
                                             float(mu_inf), float(lam_CY), float(n_CY), float(a_CY)))

    def les(self, slot, model=None, C_s=0.025, filter=1, A_plus=25.0):
        """The LES plugin (LES.f90): model 1 Smagorinsky or 2 Vreman with the
        params_model.in values C_s, filter, A_plus; model=None removes it."""
        on = model is not None
        check(self.lib.opf_slot_les(self.h, slot, int(on), int(model or 1), float(C_s), int(filter),
                                    float(A_plus)))

    def timestep(self, slot):
        """tim_dt, tim_corr_dt, tim_cfl_dt, tim_cfl_dir (io_write_timestep)."""
        dt, cd, cf, di = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        check(self.lib.opf_slot_timestep(self.h, slot, C.byref(dt), C.byref(cd), C.byref(cf), C.byref(di)))
        return dict(dt=dt.value, corr_dt=cd.value, cfl_dt=cf.value, cfl_dir=di.value)

    def vel_maxtstep(self, slot):
        cf, di = C.c_double(), C.c_int()
        check(self.lib.opf_vel_maxtstep(self.h, slot, C.byref(cf), C.byref(di)))
        return cf.value, di.value

    def vel_matrices(self, slot, dt):
        check(self.lib.opf_vel_matrices(self.h, slot, float(dt)))

    def step_config(self):
        """(pairs per workgroup, threads per workgroup, LDS bytes) of the step kernel."""
        gp, thr, lds = C.c_int(), C.c_int(), C.c_longlong()
        check(self.lib.opf_step_config(self.h, C.byref(gp), C.byref(thr), C.byref(lds)))
        return gp.value, thr.value, lds.value

    def step_pairs(self):
        return self.step_config()[0]

    def tile_config(self):
        """(step pairs per workgroup, z-transform m columns, theta z-lines) in use."""
        v = [C.c_int() for _ in range(3)]
        check(self.lib.opf_tile_config(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def launch_mode(self, graph=-1, compact=-1, streams=0):
        """Set the time loop's launch modes (iterations per captured graph,
        compacted tail launches, one or two streams; -1 / 0 leaves one as it
        is) and return the modes now and the counters (graph_launches,
        graph_captures, compact_iterations)."""
        g, cp, ns = C.c_int(), C.c_int(), C.c_int()
        v = [C.c_longlong() for _ in range(3)]
        check(self.lib.opf_launch_mode(self.h, int(graph), int(compact), int(streams), C.byref(g), C.byref(cp),
                                       C.byref(ns), *[C.byref(x) for x in v]))
        return dict(graph=g.value, compact=cp.value, streams=ns.value, graph_launches=v[0].value,
                    graph_captures=v[1].value, compact_iterations=v[2].value)

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int()
        check(self.lib.opf_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def bytes_needed(N, K, M, capacity=1):
    lib = _lib.load()
    return int(lib.opf_ctx_bytes_needed(N, K, M, capacity))


def host_mesh(N, r=None):
    """mes_precompute on the host (no device needed): r, intrdr, dr0[q, i],
    dr1[q, i], dr1band[row, col], radlap[row, col]."""
    lib = _lib.load()
    out = dict(r=np.zeros(N), intrdr=np.zeros(N), dr0=np.zeros(25), dr1=np.zeros(25),
               dr1band=np.zeros(9 * N), radlap=np.zeros(9 * N))
    rin = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    check(lib.opf_mesh_build(int(N), _dptr(rin) if rin is not None else None,
                             *[_dptr(out[k]) for k in ("r", "intrdr", "dr0", "dr1", "dr1band", "radlap")]))
    out["dr0"] = out["dr0"].reshape(5, 5).T.copy()
    out["dr1"] = out["dr1"].reshape(5, 5).T.copy()
    out["dr1band"] = out["dr1band"].reshape(N, 9).T.copy()
    out["radlap"] = out["radlap"].reshape(N, 9).T.copy()
    return out
