"""Staged-gather sets of a SELL pattern (tc_sell_stage, host code; no GPU).

The builder's output is checked against its definition recomputed here:
block v owns slices [v*per, (v+1)*per), per = ceil(n_slices / blocks); its
set is the sorted distinct columns of those slices' entries; own[v] is the
position of the block's first row in the set; lidx[e] is the position of
entry e's column in its block's set.  A block referencing more than
TC_STAGE_MAX_COLS columns is refused with TC_ERR_SIZE.
"""
import ctypes as C

import numpy as np
import pytest


def _sell(n, rng, spread):
    """A SELL-64 layout of a random symmetric-ish pattern with a diagonal:
    slice_ptr, and per-row columns column-major inside each slice, padding
    repeating the row's own index."""
    rows = []
    for i in range(n):
        k = int(rng.integers(3, 12))
        c = np.clip(i + rng.integers(-spread, spread + 1, k), 0, n - 1)
        rows.append(np.unique(np.concatenate([[i], c])))
    ns = (n + 63) // 64
    sp = [0]
    cols = []
    for s in range(ns):
        rr = list(range(s * 64, min(n, (s + 1) * 64)))
        w = max(len(rows[i]) for i in rr)
        blk = np.empty((w, 64), np.int32)
        for lane in range(64):
            i = s * 64 + lane
            if i < n:
                r = rows[i]
                blk[:, lane] = np.concatenate([r, np.full(w - len(r), i)])
            else:
                blk[:, lane] = n - 1
        cols.append(blk.reshape(-1))
        sp.append(sp[-1] + w * 64)
    return np.array(sp, np.int32), np.concatenate(cols).astype(np.int32)


def _stage(n, sp, cols, G, chunk=0):
    import torchcor_amd._lib as L
    ns = sp.shape[0] - 1
    per = -(-ns // G)
    units = G * (-(-per // chunk) if chunk else 1)
    col_ptr = np.empty(units + 1, np.int32)
    own = np.empty(units, np.int32)
    out = np.empty(cols.shape[0], np.int32)
    lidx = np.empty(cols.shape[0], np.uint16)
    mx = C.c_int32(0)
    rc = L.lib.tc_sell_stage(n, ns, sp.ctypes.data, cols.ctypes.data, G, chunk,
                             col_ptr.ctypes.data, own.ctypes.data, out.ctypes.data,
                             lidx.ctypes.data, C.byref(mx))
    return rc, col_ptr, own, out, lidx, mx.value


@pytest.mark.parametrize("n,G,spread", [(1000, 3, 40), (5000, 16, 300), (64 * 7 + 5, 4, 10),
                                        (130, 8, 5)])
def test_stage_matches_definition(n, G, spread):
    import torchcor_amd._lib as L
    rng = np.random.default_rng(n + G)
    sp, cols = _sell(n, rng, spread)
    rc, col_ptr, own, out, lidx, mx = _stage(n, sp, cols, G)
    assert rc == 0, L.lib.tc_last_error()
    ns = sp.shape[0] - 1
    per = -(-ns // G)
    biggest = 0
    for v in range(G):
        s0, s1 = v * per, min(ns, (v + 1) * per)
        got = out[col_ptr[v]:col_ptr[v + 1]]
        if s0 >= s1:
            assert got.size == 0
            continue
        ent = cols[sp[s0]:sp[s1]]
        want = np.unique(ent)
        assert np.array_equal(got, want), v
        assert want[own[v]] == s0 * 64
        assert np.array_equal(want[lidx[sp[s0]:sp[s1]]], ent)
        biggest = max(biggest, want.size)
    assert mx == biggest


def test_stage_refuses_large_sets():
    import torchcor_amd._lib as L
    rng = np.random.default_rng(3)
    n = 20000
    sp, cols = _sell(n, rng, 9000)     # scattered columns: sets far above the limit
    rc, *_ , mx = _stage(n, sp, cols, 2)
    assert rc == L.TC_ERR_SIZE and mx > 4096
    assert b"columns" in L.lib.tc_last_error()


def test_stage_rejects_bad_arguments():
    import torchcor_amd._lib as L
    sp = np.array([0, 64], np.int32)
    cols = np.full(64, 99, np.int32)   # column out of range for n = 10
    rc, *_ = _stage(10, sp, cols, 1)
    assert rc == L.TC_ERR_ARG


@pytest.mark.parametrize("n,G,chunk,spread", [(5000, 3, 8, 300), (64 * 41 + 9, 2, 16, 60),
                                              (3000, 5, 16, 100)])
def test_stage_units_match_definition(n, G, chunk, spread):
    """Blocks staged in units of `chunk` slices: unit c of block v covers
    slices v*per + c*chunk .. (clipped to the block), its set the distinct
    columns of those slices' entries."""
    import torchcor_amd._lib as L
    rng = np.random.default_rng(n + G + chunk)
    sp, cols = _sell(n, rng, spread)
    rc, col_ptr, own, out, lidx, mx = _stage(n, sp, cols, G, chunk)
    assert rc == 0, L.lib.tc_last_error()
    ns = sp.shape[0] - 1
    per = -(-ns // G)
    upb = -(-per // chunk)
    biggest, covered = 0, 0
    for v in range(G):
        b0, b1 = v * per, min(ns, (v + 1) * per)
        for c in range(upb):
            u = v * upb + c
            s0, s1 = b0 + c * chunk, min(b0 + (c + 1) * chunk, b1)
            got = out[col_ptr[u]:col_ptr[u + 1]]
            if s0 >= s1:
                assert got.size == 0
                continue
            covered += s1 - s0
            ent = cols[sp[s0]:sp[s1]]
            want = np.unique(ent)
            assert np.array_equal(got, want), (v, c)
            assert want[own[u]] == s0 * 64
            assert np.array_equal(want[lidx[sp[s0]:sp[s1]]], ent)
            biggest = max(biggest, want.size)
    assert covered == ns and mx == biggest


def test_stage_rejects_bad_chunk():
    import torchcor_amd._lib as L
    rng = np.random.default_rng(3)
    sp, cols = _sell(500, rng, 20)
    assert _stage(500, sp, cols, 2, 12)[0] == L.TC_ERR_ARG
    assert _stage(500, sp, cols, 2, 24)[0] == L.TC_ERR_ARG
