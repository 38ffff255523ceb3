"""Python host of the request server (opf_server_* in include/opf.h).

``Server`` holds a device memory pool shared by the resolution classes and a
native scheduler thread that continuously batches integration requests: a
request is admitted into a free slot of its class (a new block of slots is
taken from the pool when none is free) between two iterations of the other
slots, advanced with every running slot of its block, and retired when done.
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
    def __init__(self, classes, block_slots, pool_bytes, depth=3):
        """classes: list of (N, K, M); block_slots: slots per block of each
        class; pool_bytes: the device memory the server may hold; depth:
        iterations kept queued per block."""
        self.lib = _lib.load()
        self.classes = [tuple(int(v) for v in c) for c in classes]
        nkm = (C.c_int * (3 * len(classes)))(*[v for c in self.classes for v in c])
        bs = (C.c_int * len(classes))(*[int(v) for v in block_slots])
        h = C.c_void_p()
        _check(self.lib, self.lib.opf_server_create(C.byref(h), len(classes), nkm, bs, int(pool_bytes),
                                                    int(depth)))
        self.h = h
        self.H = [(2 * K - 1) * M - (K - 1) for (_, K, M) in self.classes]
        self.block_bytes = [bytes_needed(N, K, M, b) for (N, K, M), b in zip(self.classes, block_slots)]

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

    def pool(self):
        used, limit = C.c_size_t(), C.c_size_t()
        nb = C.c_int()
        _check(self.lib, self.lib.opf_server_pool(self.h, C.byref(used), C.byref(limit), C.byref(nb)))
        return dict(used=used.value, limit=limit.value, blocks=nb.value)
