"""Batched execution (one function on many matrices, fused launches) and the
layout tuner, against per-matrix execution and the CPU oracle."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _mats(count=6, seed=3, n_lo=300, n_hi=3000):
    return list(M.batch_matrices(count, seed, n_lo, n_hi))


def _colour(rp, col, T, dist):
    n = len(rp) - 1
    ce = R.Interface(n, T, dist, rp, col, symm_hint=True)
    ce.RACEColor()
    return ce


def _prep(oracle, rp, col, val, T, dist):
    ce = _colour(rp, col, T, dist)
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    g = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(g, T, dist)
    assert np.array_equal(g.perm, perm)
    return ce, prp, pcol, pval, zt


@pytest.mark.parametrize("kind,dist", [(R.K_KACZ, R.TWO), (R.K_GS, R.ONE)])
def test_batch_fp64_matches_oracle(oracle, kind, dist):
    """FP64 batch of banded and Chung-Lu matrices (hub rows, leaves larger
    than a bundle), nrhs 2, forward then backward sweep, each matrix against
    the oracle in RACE order."""
    okind = oracle.K_KACZ if kind == R.K_KACZ else oracle.K_GS
    batch = R.Batch(kind, R.FP64, R.FP64, nrhs=2)
    keep = []
    for k, (name, n, mean, rp, col, val) in enumerate(_mats()):
        T = 16 if k % 3 else 64
        print("matrix", k, name, n, T, flush=True)
        ce, prp, pcol, pval, zt = _prep(oracle, rp, col, val, T, dist)
        x0 = M.uniform_vec(n, 20 + k, nrhs=2)
        b = M.uniform_vec(n, 40 + k, nrhs=2)
        dx, db = R.DeviceArray.from_host(x0), R.DeviceArray.from_host(b)
        batch.add(ce, prp, pcol, pval, dx, db)
        keep.append((ce, prp, pcol, pval, zt, x0.copy(), b, dx))
    batch.build()
    for rev in (False, True):
        batch.execute(rev)
        for ce, prp, pcol, pval, zt, xo, b, dx in keep:
            oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev), nrhs=2)
            assert relerr(dx.download(), xo) <= TOL


def test_batch_fp32_equals_single_matrix_path(oracle):
    """The fused launches run the same bundles as per-matrix registration:
    FP32 KACZ with FP64 row norms, 8 right-hand sides, bit for bit."""
    batch = R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs=8)
    singles = []
    for k, (name, n, mean, rp, col, val) in enumerate(_mats(5, 9)):
        ce, prp, pcol, pval, zt = _prep(oracle, rp, col, val, 32, R.TWO)
        b = M.uniform_vec(n, 60 + k, nrhs=8).astype(np.float32)
        xa, ba = R.DeviceArray.from_host(np.zeros(8 * n, np.float32)), R.DeviceArray.from_host(b)
        xs, bs = R.DeviceArray.from_host(np.zeros(8 * n, np.float32)), R.DeviceArray.from_host(b)
        batch.add(ce, prp, pcol, pval, xa, ba)
        ce.setOption("layout", 2)
        fid = ce.registerDevice(R.K_KACZ, prp, pcol, pval, xs, bs, nrhs=8, val_precision=R.FP32,
                                vec_precision=R.FP32)
        singles.append((ce, fid, xa, xs))
    batch.build()
    for rev in (False, True, False):
        batch.execute(rev)
        for ce, fid, xa, xs in singles:
            ce.executeFunction(fid, rev)
            assert np.array_equal(xa.download(), xs.download())


@pytest.mark.parametrize("prec", [R.FP64, R.FP32])
def test_batch_symmspmv_all_layouts(oracle, prec):
    """SymmSpMV b += A x (upper triangle in, expanded on the device) for the
    12 SELL layouts of the tuner, 3 right-hand sides, against scipy."""
    dt = np.float64 if prec == R.FP64 else np.float32
    tol = TOL if prec == R.FP64 else 2e-6
    mats = []
    for k, (name, n, mean, rp, col, val) in enumerate(_mats(4, 5)):
        ce = _colour(rp, col, 8, R.TWO)
        perm, inv = ce.getPerm(), ce.getInvPerm()
        prp, pcol, pval = M.permute(rp, col, val, perm, inv)
        prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
        urp, ucol, uval = M.upper_triangle(prp, pcol, pval)
        if prec == R.FP32:
            pval = pval.astype(np.float32).astype(np.float64)
            uval = uval.astype(np.float32).astype(np.float64)
        x = M.uniform_vec(n, 70 + k, nrhs=3).astype(dt)
        A = M.to_scipy(prp, pcol, pval)
        ref = np.concatenate([A @ x[q * n:(q + 1) * n].astype(np.float64) for q in range(3)])
        mats.append((ce, urp, ucol, uval, x, ref))
    for cand in range(12):
        batch = R.Batch(R.K_SYMMSPMV, prec, prec, nrhs=3)
        outs = []
        for ce, urp, ucol, uval, x, ref in mats:
            dx, db = R.DeviceArray.from_host(x), R.DeviceArray.from_host(np.zeros(len(x), dt))
            batch.add(ce, urp, ucol, uval, dx, db)
            outs.append((dx, db, ref))
        batch.useCandidate(cand)
        batch.execute()
        for dx, db, ref in outs:
            assert relerr(db.download().astype(np.float64), ref) <= tol, (cand, R.Batch.candidate(R.K_SYMMSPMV, cand))


def test_batch_autotune_keeps_results(oracle):
    """The tuner times the 12 bundle layouts and keeps the fastest; results
    stay those of the oracle."""
    batch = R.Batch(R.K_KACZ, R.FP64, R.FP64, nrhs=1)
    keep = []
    for k, (name, n, mean, rp, col, val) in enumerate(_mats(4, 11)):
        ce, prp, pcol, pval, zt = _prep(oracle, rp, col, val, 16, R.TWO)
        b = M.uniform_vec(n, 80 + k)
        dx, db = R.DeviceArray.from_host(np.zeros(n)), R.DeviceArray.from_host(b)
        batch.add(ce, prp, pcol, pval, dx, db)
        keep.append((prp, pcol, pval, zt, b, dx))
    best, times = batch.autotune(reps=2)
    assert 0 <= best < 12 and times[best] > 0
    assert all(t < 0 or t >= times[best] for t in times)
    for _, _, _, _, _, dx in keep:
        dx.upload(np.zeros(dx.n))
    batch.execute(False)
    for prp, pcol, pval, zt, b, dx in keep:
        xo = np.zeros(len(b))
        oracle.run_calls(oracle.K_KACZ, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, False))
        assert relerr(dx.download(), xo) <= TOL
