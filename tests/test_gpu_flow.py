"""Gauss-Seidel in one launch per sweep (the "flow" option): the chains of all
colour phases are handed out in phase order and each waits for the earlier
chains it conflicts with through per-chain completion flags.  The result must
be RACE's execution order exactly: against the oracle run in RACE order, and
bit for bit against the phase-by-phase launches of the same bundles."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr
from .test_gpu_parity import _prepared

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _register(ce, prp, pcol, pval, x0, b, flow, prec=R.FP64):
    ce.setOption("flow", flow)
    xdt = np.float64 if prec == R.FP64 else np.float32
    dx = R.DeviceArray.from_host(x0.astype(xdt))
    db = R.DeviceArray.from_host(b.astype(xdt))
    fid = ce.registerDevice(R.K_GS, prp, pcol, pval, dx, db, val_precision=prec,
                            vec_precision=R.FP64 if prec == R.FP64 else R.FP32)
    return fid, dx


# (T, bundle_rows, bundle_local): the last ones cut leaves into chains of
# several bundles
@pytest.mark.parametrize("T,rows,local", [(1, 0, 0), (4, 0, 0), (64, 0, 0), (1024, 128, 0),
                                          (4, 32, 256), (16, 16, 128)])
def test_flow_gs_matches_race_order(oracle, T, rows, local):
    rp, col = M.stencil27(12)
    val = M.gs_values(rp, col, seed=7)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, R.ONE, oracle)
    if rows:
        ce.setOption("bundle_rows", rows)
    if local:
        ce.setOption("bundle_local", local)
    b = M.uniform_vec(n, 11)
    x0 = M.uniform_vec(n, 12)
    fid, dx = _register(ce, prp, pcol, pval, x0, b, 1)
    assert ce.launchCount(fid) == 1 and ce.launchCount(fid, True) == 1
    xo = x0.copy()
    for rev in (False, True, False, True):
        print("flow sweep rev", rev, "T", T, rows, local, flush=True)
        ce.executeFunction(fid, rev)
        oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL
    assert ce.flowAborts(fid) == 0


@pytest.mark.parametrize("prec", [R.FP64, R.FP32, R.BF16])
@pytest.mark.parametrize("g,T", [(24, 256), (40, 4096)])
def test_flow_bit_identical_to_phase_launches(prec, g, T):
    rp, col = M.stencil27(g)
    val = M.gs_values(rp, col, seed=3)
    n = len(rp) - 1
    ce = R.Interface(n, T, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    b = M.uniform_vec(n, 4)
    x0 = M.uniform_vec(n, 5)
    fa, xa = _register(ce, prp, pcol, pval, x0, b, 0, prec)
    fb, xb = _register(ce, prp, pcol, pval, x0, b, 1, prec)
    assert ce.launchCount(fa) > 1 and ce.launchCount(fb) == 1
    for s in range(6):
        rev = bool(s % 2)
        print("flow vs phases", prec, g, T, "sweep", s, flush=True)
        ce.executeFunction(fa, rev)
        ce.executeFunction(fb, rev)
        assert np.array_equal(xa.download(), xb.download()), s
    assert ce.flowAborts(fb) == 0


def test_flow_graph_replays():
    """The one-launch sweep keeps its state on the device (ticket, epoch,
    flags), so a captured graph replays it: every replay, with eager work and
    synchronisation between, equals the eager phase launches bit for bit."""
    rp, col = M.stencil27(20)
    val = M.gs_values(rp, col, seed=9)
    n = len(rp) - 1
    ce = R.Interface(n, 512, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    b = M.uniform_vec(n, 6)
    x0 = M.uniform_vec(n, 7)
    fa, xa = _register(ce, prp, pcol, pval, x0, b, 0)
    ce.setOption("graph", 1)
    fb, xb = _register(ce, prp, pcol, pval, x0, b, 1)
    for s in range(5):
        rev = bool(s % 2)
        print("flow graph replay", s, flush=True)
        ce.executeFunction(fb, rev)
        ce.setOption("graph", 0)
        ce.executeFunction(fa, rev)
        ce.setOption("graph", 1)
        assert np.array_equal(xa.download(), xb.download()), s
    assert ce.flowAborts(fb) == 0
