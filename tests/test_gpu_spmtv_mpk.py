"""GPU parity of SpMTV and the matrix-power variants against the CPU oracle.

SpMTV (SPMTV_KERNEL examples/src/helpers/kernels.cpp:106-124) runs as a
gather over the transpose: FP64 within 1e-12 normwise of the oracle's
row-by-row scatter.  The only-highest MPK (MAT_SPMV_KERNEL_only_highest
kernels.cpp:706-734) runs the same per-row sums as the full MPK, so its two
slots are bit-equal to the full MPK's last two powers on the device.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _unsymmetric(n, density, seed):
    rng = np.random.default_rng(seed)
    A = (sp.random(n, n, density=density, random_state=rng, format="csr") + sp.eye(n)).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy(), A


@pytest.mark.parametrize("nrhs", [1, 3])
@pytest.mark.parametrize("colored", [False, True])
def test_spmtv_matches_oracle(oracle, nrhs, colored):
    print("case spmtv", nrhs, colored, flush=True)
    n = 3000
    rp, col, val, A = _unsymmetric(n, 0.002, 11 + nrhs)
    ce = R.Interface(n, 8, R.TWO, rp, col, symm_hint=False)
    if colored:
        ce.RACEColor()
        perm, inv = ce.getPerm(), ce.getInvPerm()
        rp, col, val = M.permute(rp, col, val, perm, inv)
        A = M.to_scipy(rp, col, val)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, nrhs * n)
    b0 = rng.uniform(-1, 1, nrhs * n)
    dx, db = R.DeviceArray.from_host(x), R.DeviceArray.from_host(b0)
    fid = ce.registerDevice(R.K_SPMTV, rp, col, val, dx, db, nrhs=nrhs)
    ce.executeFunction(fid)
    got = db.download()
    ref = b0.copy()
    oracle.run_calls(oracle.K_SPMTV, rp, col, val, x, ref, [(0, n)], nrhs=nrhs)
    for k in range(nrhs):
        s = slice(k * n, (k + 1) * n)
        assert relerr(got[s], ref[s]) <= TOL, k
        assert relerr(got[s], b0[s] + A.T @ x[s]) <= TOL, k
    # a second call accumulates again
    ce.executeFunction(fid)
    oracle.run_calls(oracle.K_SPMTV, rp, col, val, x, ref, [(0, n)], nrhs=nrhs)
    assert relerr(db.download(), ref) <= TOL


def test_spmtv_empty_rows_and_columns(oracle):
    """Rows with no entries and columns no row references (empty rows of A^T)."""
    n = 200
    rows = [[] if r % 7 == 3 else sorted({r, (3 * r + 1) % n, (5 * r) % n} - {n - 1}) for r in range(n)]
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    col = np.array([c for cs in rows for c in cs], np.int32)
    val = np.random.default_rng(2).uniform(-1, 1, len(col))
    assert not (col == n - 1).any()
    ce = R.Interface(n, 1, R.ONE, rp, col, symm_hint=False)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    dx, db = R.DeviceArray.from_host(x), R.DeviceArray.from_host(np.zeros(n))
    fid = ce.registerDevice(R.K_SPMTV, rp, col, val, dx, db)
    ce.executeFunction(fid)
    ref = np.zeros(n)
    oracle.run_calls(oracle.K_SPMTV, rp, col, val, x, ref, [(0, n)])
    got = db.download()
    assert relerr(got, ref) <= TOL
    assert got[n - 1] == 0.0


def _power_setup(g, k, cache, seed):
    rp, col = M.stencil7(g)
    val = M.jittered_values(rp, col, 0.1, seed=seed)
    n = len(rp) - 1
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, cache)
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    return ce, n, (prp, pcol, pval)


@pytest.mark.parametrize("k", [1, 4, 5])
def test_mpk_highest_matches_full_mpk(oracle, k):
    print("case mpk_highest", k, flush=True)
    ce, n, (prp, pcol, pval) = _power_setup(16, k, 64 * 1024.0, 3)
    assert len(ce.mpkGroups()) > 3
    x0 = M.uniform_vec(n, 4)
    full = np.zeros((k + 1) * n)
    full[:n] = x0
    dfull = R.DeviceArray.from_host(full)
    two = np.zeros(2 * n)
    two[:n] = x0
    dtwo = R.DeviceArray.from_host(two)
    f_full = ce.registerDevice(R.K_MPK, prp, pcol, pval, dfull, power=k)
    f_two = ce.registerDevice(R.K_MPK_HIGHEST, prp, pcol, pval, dtwo, power=k)
    ce.executeFunction(f_full)
    ce.executeFunction(f_two)
    gf = dfull.download().reshape(k + 1, n)
    gt = dtwo.download().reshape(2, n)
    assert np.array_equal(gt[k % 2], gf[k])
    assert np.array_equal(gt[(k - 1) % 2], gf[k - 1])
    ref = oracle.mpk_highest_serial(prp, pcol, pval, x0, k)
    assert relerr(gt[k % 2], ref[k % 2]) <= TOL
    assert relerr(gt[(k - 1) % 2], ref[(k - 1) % 2]) <= TOL


def test_mpk_many_powers(oracle):
    """k = 20 over 44 level groups: wavefront steps with up to 20 segments,
    more than the 16 that travel in the kernel arguments, so those steps read
    the device segment table; full and only-highest against the oracle."""
    k = 20
    ce, n, (prp, pcol, pval) = _power_setup(16, k, 64 * 1024.0, 5)
    assert len(ce.mpkGroups()) - 1 >= 33
    pval = pval / 6.0  # spectral radius ~1: twenty powers stay O(1)
    x0 = M.uniform_vec(n, 9)
    full = np.zeros((k + 1) * n)
    full[:n] = x0
    dfull = R.DeviceArray.from_host(full)
    two = np.zeros(2 * n)
    two[:n] = x0
    dtwo = R.DeviceArray.from_host(two)
    f_full = ce.registerDevice(R.K_MPK, prp, pcol, pval, dfull, power=k)
    f_two = ce.registerDevice(R.K_MPK_HIGHEST, prp, pcol, pval, dtwo, power=k)
    print("case mpk k=20", flush=True)
    ce.executeFunction(f_full)
    ce.executeFunction(f_two)
    got = dfull.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL, p
    gt = dtwo.download().reshape(2, n)
    assert np.array_equal(gt[k % 2], got[k]) and np.array_equal(gt[(k - 1) % 2], got[k - 1])


def test_mpk_set_power(oracle):
    """setPower(p) on a device MPK computes powers 1..p only; above the
    registered power it is rejected."""
    k = 4
    ce, n, (prp, pcol, pval) = _power_setup(12, k, 32 * 1024.0, 6)
    x0 = M.uniform_vec(n, 8)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    ce.setPower(fid, 2)
    assert ce.getPower(fid) == 2
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    assert relerr(got[1], ref[1]) <= TOL and relerr(got[2], ref[2]) <= TOL
    assert not got[3].any() and not got[4].any()
    with pytest.raises(R.RaceError):
        ce.setPower(fid, k + 1)
    assert ce.getPower(fid) == 2
    best = ce.tuneFunction(fid)
    assert 1 <= best <= k and ce.getPower(fid) == best
    ce.setPower(fid, k)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL, p
