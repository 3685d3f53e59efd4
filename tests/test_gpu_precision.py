"""Mixed precision on the GPU: stated error of FP32/BF16 storage against the
FP64 oracle, and FP64 iterative refinement reaching the FP64 residual."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

# stated errors of one forward+backward Gauss-Seidel sweep against the FP64
# oracle in the same order: FP32 storage and arithmetic ~ 1e-7 relative,
# BF16 storage (8-bit mantissa) with FP32 arithmetic ~ 1e-3
TOL_FP32 = 1e-6
TOL_BF16 = 5e-3


def _setup(oracle, g=14, T=256):
    rp, col = M.stencil27(g)
    val = M.gs_values(rp, col, seed=11)
    n = len(rp) - 1
    ce = R.Interface(n, T, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    G = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(G, T, 1)
    return ce, (prp, pcol, pval), zt, n


@pytest.mark.parametrize("prec,tol", [(R.FP32, TOL_FP32), (R.BF16, TOL_BF16)])
def test_low_precision_gs_error_is_within_stated_bound(oracle, prec, tol):
    ce, (prp, pcol, pval), zt, n = _setup(oracle)
    b = M.uniform_vec(n, 3)
    dx = R.DeviceArray(n, np.float32)
    dx.zero()
    db = R.DeviceArray.from_host(b.astype(np.float32))
    fid = ce.registerDevice(R.K_GS, prp, pcol, pval, dx, db, val_precision=prec, vec_precision=R.FP32)
    ce.executeFunction(fid)
    ce.executeFunction(fid, True)
    xo = np.zeros(n)
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt))
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, True))
    err = relerr(dx.download(), xo)
    print("relative error", err)
    assert err <= tol


@pytest.mark.parametrize("inner_prec", [R.FP64, R.FP32, R.BF16])
def test_refinement_reaches_fp64_residual(oracle, inner_prec):
    ce, (prp, pcol, pval), zt, n = _setup(oracle, g=12, T=64)
    xs = M.uniform_vec(n, 5)
    A = M.to_scipy(prp, pcol, pval)
    b = A @ xs
    vecp = R.FP64 if inner_prec == R.FP64 else R.FP32
    dt = np.float64 if vecp == R.FP64 else np.float32
    wx, wb = R.DeviceArray(n, dt), R.DeviceArray(n, dt)
    inner = ce.registerDevice(R.K_GS, prp, pcol, pval, wx, wb, val_precision=inner_prec,
                              vec_precision=vecp)
    x = R.DeviceArray.from_host(np.zeros(n))
    db = R.DeviceArray.from_host(b)
    rid = ce.registerRefinement(R.K_GS, prp, pcol, pval, x, db, inner, sweeps=4)
    steps, sweeps, rr = ce.refine(rid, 1e-10, max_outer=500)
    print("steps", steps, "sweeps", sweeps, "relres", rr)
    assert rr <= 1e-10
    assert relerr(x.download(), xs) <= 1e-8
    assert ce.residual(rid) == pytest.approx(rr, rel=1e-6)


def test_kaczmarz_refinement_converges(oracle):
    rp, col = M.stencil7(8)
    val = M.jittered_values(rp, col, 0.1, seed=4)
    n = len(rp) - 1
    ce = R.Interface(n, 16, R.TWO, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    xs = M.uniform_vec(n, 6)
    b = M.to_scipy(prp, pcol, pval) @ xs
    wx, wb = R.DeviceArray(n, np.float32), R.DeviceArray(n, np.float32)
    inner = ce.registerDevice(R.K_KACZ, prp, pcol, pval, wx, wb, val_precision=R.FP32,
                              vec_precision=R.FP32)
    x = R.DeviceArray.from_host(np.zeros(n))
    db = R.DeviceArray.from_host(b)
    rid = ce.registerRefinement(R.K_KACZ, prp, pcol, pval, x, db, inner, sweeps=20)
    steps, sweeps, rr = ce.refine(rid, 1e-9, max_outer=400)
    print("steps", steps, "sweeps", sweeps, "relres", rr)
    assert rr <= 1e-9
