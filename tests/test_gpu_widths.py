"""SELL row widths around the group sizes of the row loop.  Rows of at most 8
entries take the narrow kernels (one group of 8 per row); wider rows take
whole groups of 8 and then groups of 4, the last of which re-reads the
row's last entry with value 0.  SpMV over uniform widths (chunks without
records) and mixed widths (chunk records), and the MPK over a 27-point
stencil (the wide MPK step), against the oracle and scipy."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12

# distinct column offsets: row r holds columns (r + s) mod n for the first
# w offsets
OFFSETS = [0, 1, 3, 7, 12, 20, 33, 54, 88, 143, 232, 376, 609, 985, 1597, 2584, 4181, 6765, 10946, 17711,
           28657]


def _width(widths, r):
    """one width per 64-row chunk, cycling through widths: a list of several
    widths gives chunks of different widths"""
    return widths[(r // 64) % len(widths)]


def _rows_of_widths(n, widths, seed):
    rows = [sorted({(r + s) % n for s in OFFSETS[:_width(widths, r)]}) for r in range(n)]
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    col = np.array([c for cs in rows for c in cs], np.int32)
    val = np.random.default_rng(seed).uniform(-1, 1, len(col))
    return rp, col, val


# [8]: narrow, uniform; [3, 8, 5]: narrow chunks of three widths (chunk
# records); [9], [13], [16], [21]: wide, uniform (8 + 1, 8 + 4 + 1, 8 + 8,
# 8 + 8 + 4 + 1); [5, 9, 13, 6, 11]: chunks of five widths, narrow and wide
# (chunk records)
@pytest.mark.parametrize("widths", [[8], [3, 8, 5], [9], [13], [16], [21], [5, 9, 13, 6, 11]])
def test_spmv_row_widths(oracle, widths):
    print("case spmv widths", widths, flush=True)
    n = 30000
    rp, col, val = _rows_of_widths(n, widths, seed=len(widths) + widths[0])
    assert np.all(np.diff(rp) == [_width(widths, r) for r in range(n)])
    # several widths: padding every chunk to the widest would cost more than
    # 1/8 extra entries, so the layout keeps one record per chunk
    padded = max(widths) * n
    assert (padded * 8 > int(rp[-1]) * 9) == (len(widths) > 1)
    x = M.uniform_vec(n, 21)
    ce = R.Interface(n, 1, R.ONE, rp, col, symm_hint=False)
    dx, dy = R.DeviceArray.from_host(x), R.DeviceArray.from_host(np.zeros(n))
    fid = ce.registerDevice(R.K_SPMV, rp, col, val, dx, dy)
    ce.executeFunction(fid)
    got = dy.download()
    ref = np.zeros(n)
    oracle.run_calls(oracle.K_SPMV, rp, col, val, x, ref, [(0, n)])
    assert relerr(got, ref) <= TOL
    assert relerr(got, M.to_scipy(rp, col, val) @ x) <= TOL


def test_mpk_wide_rows(oracle):
    """27-point stencil (27 entries per interior row): the MPK step with
    whole groups of 8 and a group of 4, over 7 level groups."""
    rp, col = M.stencil27(10)
    val = M.jittered_values(rp, col, 0.1, seed=12) / 26.0
    n, k = len(rp) - 1, 3
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, 256 * 1024.0)
    assert len(ce.mpkGroups()) > 3
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    x0 = M.uniform_vec(n, 2)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    print("case mpk 27-pt", flush=True)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL, p
