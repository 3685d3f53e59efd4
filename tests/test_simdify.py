"""SELL-C-sigma helpers: oracle pinned by a hand-traced known answer and
invariants, the product's host implementation (through the C ABI) bit-exact
against the oracle.  Host only."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M
from oracle import race_oracle as O


def _kat_input():
    """src/simdify.cpp:41-77: C = simdWidth = 4, four rows of length 4, every
    row holds columns 0, 1, 2, 3 (col[i*4 + j] = i, i.e. position i of every
    lane is column i)."""
    col = np.zeros(16, np.int32)
    val = np.ones(16)
    for i in range(4):
        for j in range(4):
            col[i * 4 + j] = i
    return col, val, np.array([0, 17], np.int32), np.array([4], np.int32)


# traced by hand through include/simdify.h:70-357: lane rows of the result
KAT_LANES = [[0, 1, 2, 3], [1, 0, 3, 2], [2, 3, 0, 1], [3, 2, 1, 0]]


def _lanes(col, C, w, lanes=4):
    return [[int(col[j * C + i]) for j in range(w)] for i in range(lanes)]


def test_oracle_kat():
    col, val, cs, clp = _kat_input()
    assert O.simdify_resolve(4, 4, 4, col, cs, clp, val)
    assert _lanes(col, 4, 4) == KAT_LANES


def test_product_kat():
    col, val, cs, clp = _kat_input()
    assert R.simdify_resolve(4, 4, 4, col, cs, clp, val)
    assert _lanes(col, 4, 4) == KAT_LANES


def _rows(cs, rl, clp, col, val, C, n):
    out = []
    for r in range(n):
        base = cs[r // C] + r % C
        out.append(sorted((int(col[base + j * C]), float(val[base + j * C])) for j in range(clp[r // C])))
    return out


def _random_sell(seed, n, C, minlen, maxlen, dtype=np.float64):
    rng = np.random.default_rng(seed)
    rp = [0]
    cols = []
    for r in range(n):
        k = int(rng.integers(minlen, maxlen + 1))
        c = np.sort(rng.choice(n, size=min(k, n), replace=False))
        cols.extend(c.tolist())
        rp.append(len(cols))
    rp = np.array(rp, np.int32)
    cols = np.array(cols, np.int32)
    vals = rng.uniform(-1, 1, len(cols)).astype(dtype)
    return M.to_sell(rp, cols, vals, C)


CASES = [(0, 37, 8, 4, 6, 9), (1, 64, 8, 8, 8, 12), (2, 100, 16, 4, 5, 8), (3, 30, 4, 2, 3, 5),
         (4, 96, 64, 8, 10, 14), (5, 21, 4, 4, 4, 4)]


@pytest.mark.parametrize("seed,n,C,w,lo,hi", CASES)
@pytest.mark.parametrize("diag_first", [False, True])
def test_simdify_invariants_and_product(seed, n, C, w, lo, hi, diag_first):
    cs, rl, clp, col, val = _random_sell(seed, n, C, lo, hi)
    # distinct padding like sell_c_sigmize with one leaf per SIMD group
    for r in range(n):
        base = cs[r // C] + r % C
        for j in range(rl[r], clp[r // C]):
            col[base + j * C] = (r // w) * w + (r % w)
    before = _rows(cs, rl, clp, col, val, C, n)
    oc, ov = col.copy(), val.copy()
    ok_o = O.simdify_resolve(w, C, n, oc, cs, clp, ov, diag_first)
    pc, pv = col.copy(), val.copy()
    ok_p = R.simdify_resolve(w, C, n, pc, cs, clp, pv, diag_first)
    assert ok_o == ok_p
    assert np.array_equal(oc, pc) and np.array_equal(ov, pv)
    # rows keep their entries
    assert _rows(cs, rl, clp, oc, ov, C, n) == before
    if ok_o:
        s = 1 if diag_first else 0
        for r0 in range(0, n, w):
            c = r0 // C
            lanes = [r for r in range(r0, min(r0 + w, n))]
            for j in range(s, clp[c]):
                got = [int(oc[cs[c] + r % C + j * C]) for r in lanes]
                assert len(set(got)) == len(got), (r0, j, got)
        if diag_first:
            for r in range(n):
                assert oc[cs[r // C] + r % C] == col[cs[r // C] + r % C]


def test_simdify_reports_impossible_groups():
    # four rows of length 2 sharing both columns cannot be made distinct
    # over four lanes: both sides return false
    rp = np.array([0, 2, 4, 6, 8], np.int32)
    cols = np.array([0, 1] * 4, np.int32)
    cs, rl, clp, col, val = M.to_sell(rp, cols, np.ones(8), 4)
    a, b = col.copy(), val.copy()
    assert not O.simdify_resolve(4, 4, 4, a, cs, clp, b)
    a, b = col.copy(), val.copy()
    assert not R.simdify_resolve(4, 4, 4, a, cs, clp, b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d2c", [False, True])
@pytest.mark.parametrize("n,C,w", [(64, 8, 4), (61, 8, 4), (50, 16, 8), (33, 4, 4), (65, 64, 8)])
def test_simdify_d1(n, C, w, d2c, dtype):
    # 1D Laplacian plus a few far columns: every lane depends on its
    # predecessors through the sub-diagonal
    rng = np.random.default_rng(n * 7 + C)
    rp = [0]
    cols = []
    for r in range(n):
        c = {r}
        for k in range(1, w):
            if r - k >= 0:
                c.add(r - k)
        c.update(rng.choice(n, size=2, replace=False).tolist())
        cols.extend(sorted(c))
        rp.append(len(cols))
    rp = np.array(rp, np.int32)
    cols = np.array(cols, np.int32)
    vals = rng.uniform(-1, 1, len(cols)).astype(dtype)
    cs, rl, clp, col, val = M.to_sell(rp, cols, vals, C, pad_col=-1)
    before = _rows(cs, rl, clp, col, val, C, n)
    oc, ov = col.copy(), val.copy()
    O.simdify_d1(w, C, n, oc, cs, clp, ov, d2c)
    pc, pv = col.copy(), val.copy()
    assert R.simdifyD1(w, C, n, pc, cs, rl, clp, pv, d2c)
    assert np.array_equal(oc, pc) and np.array_equal(ov, pv)
    assert _rows(cs, rl, clp, oc, ov, C, n) == before



def test_sell_c_sigmize_oracle():
    rp, cols = M.stencil7(4)
    cs, rl, clp, col, val = M.to_sell(rp, cols, np.ones(len(cols)), 8, pad_col=-1)
    calls = [(0, 5), (5, 23), (23, 64)]
    O.sell_c_sigmize(calls, 4, 8, col, cs, rl, clp)
    for s, e in calls:
        for r in range(s, e):
            base = cs[r // 8] + r % 8
            for j in range(clp[r // 8]):
                v = col[base + j * 8]
                if j >= rl[r]:
                    assert v == s + (r - s) % 4
                else:
                    assert v == cols[rp[r] + j]


# include/simdify.h:389-447 traced by hand on a 4x4 block whose rows all
# hold columns 0..3 (C = simdWidth = 4).  Without dist2Compatible the result
# is the picture of include/simdify.h:367-373 (lane m's row at position
# 3 - m).  With it, lane 3 ends as [1, 0, 3, 2], not the pictured
# [., 0, 1, 2]: after the swap at position 0 the value moved in (1) is only
# compared with the remaining dependencies, as in the reference loop.
D1_KAT = {False: [[0, 1, 2, 3], [3, 1, 2, 0], [3, 2, 1, 0], [3, 2, 1, 0]],
          True: [[0, 1, 2, 3], [3, 1, 2, 0], [2, 3, 0, 1], [1, 0, 3, 2]]}


@pytest.mark.parametrize("d2c", [False, True])
def test_simdify_d1_kat(d2c):
    for impl in ("oracle", "product"):
        col = np.array([j for j in range(4) for _ in range(4)], np.int32)  # col[4j + i] = j
        val = col.astype(np.float64) + 0.5
        cs, clp, rl = np.array([0, 16], np.int32), np.array([4], np.int32), np.full(4, 4, np.int32)
        if impl == "oracle":
            O.simdify_d1(4, 4, 4, col, cs, clp, val, d2c)
        else:
            R.simdifyD1(4, 4, 4, col, cs, rl, clp, val, d2c)
        assert _lanes(col, 4, 4) == D1_KAT[d2c], impl
        assert np.array_equal(val, col + 0.5), impl  # values travel with their columns

