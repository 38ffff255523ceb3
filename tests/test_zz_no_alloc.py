"""opf_var_coll_norm_noalloc (the allocation-free var_coll_norm on device
arrays) against opf_var_coll_norm (host arrays, library-allocated staging)
through the C ABI: two calls in a row, the second on a smaller context, each
bit for bit equal to the host-array entry point; and the size limit of its
static buffer.  Device arrays come from the HIP runtime the library itself
links (libamdhip64), so no second runtime is loaded."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libopf.so links."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0

    def get(self, src, n):
        out = np.zeros(n)
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), src, out.nbytes, self.D2H) == 0
        return out

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def _norm_noalloc(hip, ctx, field, slot=0):
    """The allocation-free entry point on device copies of field's coll planes."""
    re = np.ascontiguousarray(field.real.T)
    im = np.ascontiguousarray(field.imag.T)
    nout = 1 + ctx.K + ctx.M
    d_re, d_im, d_out = hip.malloc(re.nbytes), hip.malloc(im.nbytes), hip.malloc(8 * nout)
    try:
        hip.put(d_re, re)
        hip.put(d_im, im)
        hip.put(d_out, np.full(nout, np.nan))
        rc = ctx.lib.opf_var_coll_norm_noalloc(ctx.h, slot, d_re, d_im, d_out, None)
        assert rc == 0, ctx.lib.opf_last_error()
        out = hip.get(d_out, nout)
    finally:
        for p in (d_re, d_im, d_out):
            hip.free(p)
    return out[0], out[1:1 + ctx.K], out[1 + ctx.K:]


def test_norm_noalloc_equals_host_entry_point():
    import opf
    hip = _Hip()
    # first call: (32,24,12); second: the smaller (16,6,4), on the same
    # device buffer that still holds the first call's partials
    for name in ("c0", "mini"):
        g = load_golden(name)
        ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]))
        ctx.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
        f = g["ic"][0] + 2.0 * g["ic"][2]
        print(f"noalloc norm at {name}: H = {ctx.H}", flush=True)
        e, ek, em = _norm_noalloc(hip, ctx, f)
        e0, ek0, em0 = ctx.coll_norm(f, 0)
        assert e == e0, name
        assert np.array_equal(ek, ek0) and np.array_equal(em, em0), name
        ctx.close()


def test_norm_noalloc_rejects_oversize_and_unset_slots():
    """H above the buffer's 65536 harmonics is refused before any launch;
    so is a slot without parameters."""
    import opf
    hip = _Hip()
    big = opf.Context(9, 192, 192)        # H = 383 * 192 - 191 = 73345
    assert big.H > 65536
    d = hip.malloc(8)
    try:
        rc = big.lib.opf_var_coll_norm_noalloc(big.h, 0, d, d, d, None)
        assert rc == 1 and b"65536" in big.lib.opf_last_error()
        small = opf.Context(16, 6, 4)
        rc = small.lib.opf_var_coll_norm_noalloc(small.h, 0, d, d, d, None)
        assert rc == 5
        small.close()
    finally:
        hip.free(d)
    big.close()
