"""race_refine_noalloc: FP64 iterative refinement with its workspace in the
library's static device array (no allocation in the call), against
race_refine on the same problems: two calls in a row, the second on a
smaller matrix, each bit for bit in x and within 1e-12 in the relative
residual; and a problem larger than the array is refused."""
import ctypes

import numpy as np
import pytest

import race_amd as R
from race_amd import _lib as L
from race_amd import matgen as M

pytestmark = pytest.mark.gpu


def _problem(g, T, inner_prec, seed):
    rp, col = M.stencil27(g)
    val = M.gs_values(rp, col, seed=seed)
    n = len(rp) - 1
    ce = R.Interface(n, T, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
    xs = M.uniform_vec(n, seed + 1)
    b = M.to_scipy(prp, pcol, pval) @ xs
    vp, xp, dt = inner_prec
    wx, wb = R.DeviceArray(n, dt), R.DeviceArray(n, dt)
    inner = ce.registerDevice(R.K_GS, prp, pcol, pval, wx, wb, val_precision=vp, vec_precision=xp)
    x = R.DeviceArray.from_host(np.zeros(n))
    db = R.DeviceArray.from_host(b)
    rid = ce.registerRefinement(R.K_GS, prp, pcol, pval, x, db, inner, sweeps=2)
    return ce, rid, x, (wx, wb, db)


def _noalloc(ce, rid, tol, max_outer):
    o, sw, rr = ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
    L.check(L.load().race_refine_noalloc(ce.h, rid, float(tol), int(max_outer), ctypes.byref(o),
                                         ctypes.byref(sw), ctypes.byref(rr)), "race_refine_noalloc")
    return o.value, sw.value, rr.value


@pytest.mark.parametrize("prec", [(R.FP32, R.FP32, np.float32), (R.FP64, R.FP64, np.float64)],
                         ids=["fp32", "fp64"])
def test_refine_noalloc_equals_refine(prec):
    for g, T in ((24, 256), (12, 64)):  # the second call on a smaller matrix
        print("case refine_noalloc", prec[0], g, T, flush=True)
        ce, rid, x, keep = _problem(g, T, prec, seed=g)
        # a fixed number of steps (tol 0): both run the same kernels in order
        ref = ce.refine(rid, 0.0, max_outer=6)
        x_ref = x.download()
        x.upload(np.zeros(len(x_ref)))
        got = _noalloc(ce, rid, 0.0, 6)
        assert got[0] == ref[0] == 6 and got[1] == ref[1]
        assert np.array_equal(x.download(), x_ref)
        assert got[2] == pytest.approx(ref[2], rel=1e-12)
        # to a tolerance: the same steps, the same solution
        x.upload(np.zeros(len(x_ref)))
        a = ce.refine(rid, 1e-10, max_outer=500)
        xa = x.download()
        x.upload(np.zeros(len(x_ref)))
        b = _noalloc(ce, rid, 1e-10, 500)
        assert a[0] == b[0] and b[2] <= 1e-10
        assert np.array_equal(x.download(), xa)


def test_refine_noalloc_refuses_larger_problems():
    cap = ctypes.c_size_t()
    L.check(L.load().race_refine_noalloc_max_bytes(ctypes.byref(cap)), "max_bytes")
    assert cap.value == 64 << 20
    # FP64 inner sweeps on 2 x 5^3 ... : pick nrhs so that 2 n nrhs 8 bytes
    # exceed the array while the matrix stays small
    g = 10
    rp, col = M.stencil7(g)
    val = M.gs_values(rp, col, seed=1)
    n = len(rp) - 1
    nrhs = cap.value // (2 * 8 * n) + 1
    ce = R.Interface(n, 8, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
    wx, wb = R.DeviceArray(n * nrhs), R.DeviceArray(n * nrhs)
    inner = ce.registerDevice(R.K_GS, prp, pcol, pval, wx, wb, nrhs=nrhs)
    x = R.DeviceArray.from_host(np.zeros(n * nrhs))
    db = R.DeviceArray.from_host(np.ones(n * nrhs))
    rid = ce.registerRefinement(R.K_GS, prp, pcol, pval, x, db, inner, sweeps=1, nrhs=nrhs)
    with pytest.raises(R.RaceError):
        _noalloc(ce, rid, 1e-8, 3)
    # the allocating entry point still handles it
    steps, _, rr = ce.refine(rid, 1e-8, max_outer=3)
    assert steps >= 1
