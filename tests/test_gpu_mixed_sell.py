"""FP32 / BF16 storage with FP32 accumulation for the SELL kernels (SpMV,
SymmSpMV, SpMTV, MPK).

Two checks per result: against the FP64 product of the same values rounded
on the host to the storage precision (what the device stores), which leaves
only the FP32 accumulation rounding (<= 2e-6 normwise); and against the FP64
oracle with the unrounded values, the error a user of the mixed path sees
(<= 1e-6 for FP32 storage, <= 1e-2 for BF16)."""
import numpy as np
import pytest
import scipy.sparse as sp

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

PREC = {"fp32": (R.FP32, 1e-6), "bf16": (R.BF16, 1e-2)}


def rounded(v, prec):
    """values as the device stores them (float32, or bfloat16 by round to
    nearest even), back in float64"""
    f = np.asarray(v, np.float64).astype(np.float32)
    if prec == "fp32":
        return f.astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _dev(a):
    return R.DeviceArray.from_host(np.asarray(a, np.float32))


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("nrhs", [1, 2])
def test_spmv_and_spmtv_low_precision(prec, nrhs):
    print("case spmv/spmtv", prec, nrhs, flush=True)
    rp, col = M.stencil27(16)
    val = M.gs_values(rp, col, seed=4) * np.repeat(np.linspace(0.5, 1.5, len(rp) - 1), np.diff(rp))
    n = len(rp) - 1
    vp, tol = PREC[prec]
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    Ar = sp.csr_matrix((rounded(val, prec), col, rp), shape=(n, n))
    x = M.uniform_vec(n, 3, nrhs=nrhs).astype(np.float32).astype(np.float64)
    b0 = M.uniform_vec(n, 4, nrhs=nrhs).astype(np.float32).astype(np.float64)
    ce = R.Interface(n, 1, R.ONE, rp, col, symm_hint=True)
    dx, dy = _dev(x), _dev(np.zeros(nrhs * n))
    fid = ce.registerDevice(R.K_SPMV, rp, col, val, dx, dy, nrhs=nrhs, val_precision=vp, vec_precision=R.FP32)
    ce.executeFunction(fid)
    dz = _dev(b0)
    tid = ce.registerDevice(R.K_SPMTV, rp, col, val, dx, dz, nrhs=nrhs, val_precision=vp, vec_precision=R.FP32)
    ce.executeFunction(tid)
    y, z = dy.download(), dz.download()
    for k in range(nrhs):
        s = slice(k * n, (k + 1) * n)
        assert relerr(y[s], Ar @ x[s]) <= 2e-6
        assert relerr(y[s], A @ x[s]) <= tol
        assert relerr(z[s], b0[s] + Ar.T @ x[s]) <= 2e-6
        assert relerr(z[s] - b0[s], A.T @ x[s]) <= tol


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_symmspmv_low_precision(oracle, prec):
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=2)
    n = len(rp) - 1
    vp, tol = PREC[prec]
    ce = R.Interface(n, 16, R.TWO, rp, col, symm_hint=True)
    ce.RACEColor()
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
    urp, ucol, uval = M.upper_triangle(prp, pcol, pval)
    x = M.uniform_vec(n, 5).astype(np.float32).astype(np.float64)
    dx, db = _dev(x), _dev(np.zeros(n))
    fid = ce.registerDevice(R.K_SYMMSPMV, urp, ucol, uval, dx, db, val_precision=vp, vec_precision=R.FP32)
    ce.executeFunction(fid)
    got = db.download()
    ref_r = np.zeros(n)
    oracle.run_calls(oracle.K_SYMMSPMV, urp, ucol, rounded(uval, prec), x.copy(), ref_r, [(0, n)])
    ref = np.zeros(n)
    oracle.run_calls(oracle.K_SYMMSPMV, urp, ucol, uval, x.copy(), ref, [(0, n)])
    assert relerr(got, ref_r) <= 2e-6
    assert relerr(got, ref) <= tol
    # the atomic one-launch variant is FP64 only
    ce.setOption("symm_mode", 2)
    with pytest.raises(R.RaceError, match="NOT_IMPLEMENTED"):
        ce.registerDevice(R.K_SYMMSPMV, urp, ucol, uval, dx, db, val_precision=vp, vec_precision=R.FP32)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_mpk_low_precision(oracle, prec):
    print("case mpk", prec, flush=True)
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=6) / 6.0   # spectral radius ~1: powers stay O(1)
    n, k = len(rp) - 1, 4
    vp, tol = PREC[prec]
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, 64 * 1024.0)
    assert len(ce.mpkGroups()) > 3
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    x0 = M.uniform_vec(n, 7).astype(np.float32).astype(np.float64)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = _dev(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k, val_precision=vp, vec_precision=R.FP32)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref_r = oracle.mpk_serial(prp, pcol, rounded(pval, prec), x0, k)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    for p in range(1, k + 1):
        assert relerr(got[p], ref_r[p]) <= 2e-6 * p, p
        assert relerr(got[p], ref[p]) <= tol * p, p
    # only-highest: the same sums, two slots
    two = np.zeros(2 * n)
    two[:n] = x0
    d2 = _dev(two)
    hid = ce.registerDevice(R.K_MPK_HIGHEST, prp, pcol, pval, d2, power=k, val_precision=vp,
                            vec_precision=R.FP32)
    ce.executeFunction(hid)
    g2 = d2.download().reshape(2, n)
    assert np.array_equal(g2[k % 2], got[k]) and np.array_equal(g2[(k - 1) % 2], got[k - 1])
