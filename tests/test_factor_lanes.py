"""Index bookkeeping of the band factorisation kernel (k_lu_factor in
csrc/kernels_step.hip), checked on the host.

Four lanes share a system.  In an elimination step lane q interchanges and
updates the window columns j + c with c = q + 4t (t < 3, and c <= 8); a new
column's rows are built by lane q at rows q + 4t (t < 3) and row 12 by all
four; a finished column leaves as LAPACK rows 4q .. 4q + 3 from lane q, the
forward-copy pair q from lanes 0..2 and the backward-copy pairs q (and 4
from lane 0).  The backward pairs and pivot bytes are written eight columns
at a time: flush(j - j % 8, j % 8 + 1) after column j when j % 8 == 7 or
j == N - 1.  The bars: every window column, band row, stored row and pair is
covered by exactly one lane (row 12 and the all-lane reads aside), and the
flushes write every column once, in bounds, for any N.
"""
import collections

import pytest

KL, NLU, LUS, FW_Q, BW_Q, LPS, SG = 4, 13, 16, 3, 5, 4, 8
KV = 2 * KL


def test_window_columns_split_once():
    cols = collections.Counter()
    for q in range(LPS):
        for t in range(3):
            c = q + LPS * t
            if t == 2 and q != 0:
                continue            # c = 8 is lane 0's alone
            assert c <= KV
            cols[c] += 1
    assert sorted(cols) == list(range(KV + 1)) and set(cols.values()) == {1}


@pytest.mark.parametrize("jp", range(KL + 1))
def test_update_rows_stay_in_the_window(jp):
    # rows touched for column j + c: KV + jp - c and KV - c (interchange),
    # KV + i - c for i = 1..KL (multipliers for c = 0, update for c > 0)
    for c in range(KV + 1):
        rows = [KV + jp - c, KV - c] + [KV + i - c for i in range(1, KL + 1)]
        assert min(rows) >= 0 and max(rows) <= NLU - 1, (c, rows)


def test_built_rows_cover_the_column():
    rows = collections.Counter()
    for q in range(LPS):
        for t in range(4):
            rows[(q, LPS * t + q if t < 3 else NLU - 1)] += 1
    by_row = collections.Counter(r for (_, r) in rows)
    assert sorted(by_row) == list(range(NLU))
    assert all(by_row[r] == 1 for r in range(NLU - 1)) and by_row[NLU - 1] == LPS
    # the mesh rows each lane loads: band row r uses lap[r - KL], clamped at 0
    for q in range(LPS):
        for t in range(4):
            r = LPS * t + q if t < 3 else NLU - 1
            assert 0 <= max(r - KL, 0) <= 2 * KL


def test_stored_rows_and_pairs_once():
    lapack = collections.Counter(LPS * q + r for q in range(LPS) for r in range(LPS))
    assert sorted(lapack) == list(range(LUS)) and set(lapack.values()) == {1}
    fw = [q for q in range(LPS) if q < FW_Q]
    assert fw == list(range(FW_Q))
    bw = list(range(LPS)) + [LPS]          # pair q from lane q, pair 4 from lane 0
    assert sorted(bw) == list(range(BW_Q))
    # backward pair q >= 1 = (row 9 - 2q, row 8 - 2q): rows 0..7 each once
    rows = [r for q in range(1, BW_Q) for r in (KV + 1 - 2 * q, KV - 2 * q)]
    assert sorted(rows) == list(range(KV))


@pytest.mark.parametrize("N", [16, 32, 33, 48, 63, 64, 96, 255])
def test_flushes_write_every_column_once(N):
    written = collections.Counter()
    pvs = collections.Counter()
    for j in range(N):
        if j % SG == SG - 1 or j == N - 1:
            c0, n = j - j % SG, j % SG + 1
            for cc in range(n):
                col = c0 + cc
                assert 0 <= col < N
                written[col] += 1
                pvs[col] += 1
    assert sorted(written) == list(range(N)) and set(written.values()) == {1}
    assert pvs == written


@pytest.mark.parametrize("N,j", [(32, 1), (32, 9), (64, 17), (33, 32)])
def test_singular_flush_stays_in_bounds(N, j):
    # on a zero pivot at column j: the pairs of columns j - j % 8 .. j - 1
    # and the pivot bytes of columns j - j % 8 .. j
    c0 = j - j % SG
    assert all(0 <= c0 + cc < N for cc in range(j % SG + 1))
