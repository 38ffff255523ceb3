"""Index bookkeeping of the whole-plane evaluation kernel (k_plane_t in
csrc/kernels_fft.hip), checked on the host for every grid it is built for.

The kernel keeps the six z-transformed fields of a radial plane in LDS
(field f, mode m at (f M + m) Zp, z at pidx(z)) and writes the nonlinear
terms' z columns, in the linear input layout (z at z), over fields 0..2 one
chunk of PL z-lines at a time.  That is safe only if no entry a chunk writes
is read later by the pack of a following chunk, and every access stays
inside its region.  The bars: no such hazard, all reads and writes in
bounds, and the LDS request within 160 KiB.
"""
import numpy as np
import pytest


def padded_len(L):
    return (L + (L - 1) // 16 + 1) | 1


def pidx(i):
    return i + (i >> 4)


# (K, M, PL): the instantiated grids of the kernel's table
GRIDS = [(24, 12, 24)]


@pytest.mark.parametrize("K,M,PL", GRIDS)
def test_chunk_writes_never_reach_unread_entries(K, M, PL):
    Z, T = 3 * K, 3 * M
    Zp, Tp = padded_len(Z), padded_len(T)
    assert Z % PL == 0 and PL % 2 == 0
    size_a = 6 * M * Zp
    written = np.zeros(size_a, dtype=bool)
    for z0 in range(0, Z, PL):
        # pack of this chunk: fields 2pr and 2pr + 1 of every mode at its z-lines
        reads = [(2 * pr + h) * M * Zp + mm * Zp + pidx(z0 + line)
                 for line in range(PL) for pr in range(3) for h in (0, 1) for mm in range(M)]
        reads = np.array(reads)
        assert reads.min() >= 0 and reads.max() < size_a
        # padded z positions stay inside a column
        assert pidx(z0 + PL - 1) < Zp
        assert not written[reads].any(), z0
        # unpack of this chunk (after the pack, across a barrier)
        writes = np.array([(f * M + mm) * Zp + z0 + line
                           for line in range(PL) for f in range(3) for mm in range(M)])
        assert writes.min() >= 0 and writes.max() < 3 * M * Zp     # fields 0..2 only
        written[writes] = True
    # at the end every linear z of fields 0..2 holds a nonlinear term
    cols = np.arange(3 * M)[:, None] * Zp + np.arange(Z)[None, :]
    assert written[cols].all()


@pytest.mark.parametrize("K,M,PL", GRIDS)
def test_theta_chunk_and_lds_request(K, M, PL):
    Z, T = 3 * K, 3 * M
    Zp, Tp = padded_len(Z), padded_len(T)
    # the chunk's transforms: 3 PL lines of T values (linear input, padded
    # output), the N_z transforms PL .. PL + PL/2 - 1
    assert pidx(T - 1) < Tp
    assert PL + PL // 2 <= 3 * PL
    lds = (6 * M * Zp + 3 * PL * Tp + Z + T) * 16
    static = 3 * (1024 // 64) * 8            # the block's maxima
    assert lds + static <= 160 * 1024, lds
