"""Python host of the request server (opf_server_* in include/opf.h).

``Server`` holds one device context per resolution class and a native
scheduler thread that continuously batches integration requests: a request
is admitted into a free slot of its class, advanced together with every
other occupied slot of every class, and retired when done.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .engine import _dptr, _join, _split, bytes_needed


class ServerError(RuntimeError):
    pass


def _check(lib, rc):
    if rc != 0:
        raise ServerError(f"{_lib.ERRORS.get(rc, rc)}: {lib.opf_server_last_error().decode(errors='replace')}")


class Server:
    def __init__(self, classes, capacity, chunk=50):
        """classes: list of (N, K, M); capacity: slots per class."""
        self.lib = _lib.load()
        self.classes = [tuple(int(v) for v in c) for c in classes]
        nkm = (C.c_int * (3 * len(classes)))(*[v for c in self.classes for v in c])
        cap = (C.c_int * len(classes))(*[int(v) for v in capacity])
        h = C.c_void_p()
        _check(self.lib, self.lib.opf_server_create(C.byref(h), len(classes), nkm, cap, int(chunk)))
        self.h = h
        self.H = [(2 * K - 1) * M - (K - 1) for (_, K, M) in self.classes]

    @staticmethod
    def bytes_needed(classes, capacity):
        return sum(bytes_needed(N, K, M, c) for (N, K, M), c in zip(classes, capacity))

    def close(self):
        if getattr(self, "h", None):
            self.lib.opf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cls, Re, alpha, dt, nsteps, state, corr_fixed=1):
        """Queue an integration of `state` (ur, ut, uz) for nsteps steps."""
        if not 0 <= int(cls) < len(self.classes):
            raise ServerError(f"OPF_ERR_ARG: no resolution class {cls}")
        N, _, _ = self.classes[cls]
        parts = []
        for a in state[:3]:
            parts.extend(_split(a, N, self.H[cls]))
        rid = C.c_longlong()
        _check(self.lib, self.lib.opf_server_submit(self.h, int(cls), float(Re), float(alpha), float(dt),
                                                    int(nsteps), int(corr_fixed),
                                                    *[_dptr(p) for p in parts], C.byref(rid)))
        return rid.value

    def poll(self, rid):
        done, st = C.c_int(), C.c_int()
        ts, ta, td = C.c_double(), C.c_double(), C.c_double()
        _check(self.lib, self.lib.opf_server_poll(self.h, C.c_longlong(rid), C.byref(done), C.byref(st),
                                                  C.byref(ts), C.byref(ta), C.byref(td)))
        return dict(done=bool(done.value), status=st.value, t_submit=ts.value, t_admit=ta.value,
                    t_done=td.value)

    def result(self, rid, cls):
        N, _, _ = self.classes[cls]
        H = self.H[cls]
        parts = [np.zeros((H, N)) for _ in range(6)]
        t = C.c_double()
        _check(self.lib, self.lib.opf_server_result(self.h, C.c_longlong(rid), *[_dptr(p) for p in parts],
                                                    C.byref(t)))
        return (_join(parts[0], parts[1]), _join(parts[2], parts[3]), _join(parts[4], parts[5]), t.value)

    def stats(self):
        q, r = C.c_int(), C.c_int()
        ss = C.c_longlong()
        _check(self.lib, self.lib.opf_server_stats(self.h, C.byref(q), C.byref(r), C.byref(ss)))
        return dict(queued=q.value, running=r.value, state_steps=ss.value)
