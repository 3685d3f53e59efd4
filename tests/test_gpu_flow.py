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


# --- Kaczmarz (distance-2 colouring): every column a chain touches is read
# and written, so a chain waits for the previous chain touching each column

@pytest.mark.parametrize("T,rows,local,nrhs,bs", [(8, 0, 0, 1, 256), (64, 0, 0, 3, 256), (32, 32, 256, 2, 512),
                                                 (256, 0, 0, 1, 1024)])
def test_flow_kacz_matches_race_order(oracle, T, rows, local, nrhs, bs):
    rp, col = M.stencil7(14)
    val = M.jittered_values(rp, col, 0.1, seed=3)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, R.TWO, oracle)
    ce.setOption("bundle_threads", bs)
    if rows:
        ce.setOption("bundle_rows", rows)
    if local:
        ce.setOption("bundle_local", local)
    ce.setOption("flow", 1)
    b = M.uniform_vec(n, 5, nrhs=nrhs)
    dx = R.DeviceArray.from_host(np.zeros(nrhs * n))
    db = R.DeviceArray.from_host(b)
    fid = ce.registerDevice(R.K_KACZ, prp, pcol, pval, dx, db, nrhs=nrhs)
    assert ce.launchCount(fid) == 1
    xo = np.zeros(nrhs * n)
    for rev in (False, True, False):
        print("flow kacz rev", rev, T, rows, local, nrhs, bs, flush=True)
        ce.executeFunction(fid, rev)
        oracle.run_calls(oracle.K_KACZ, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev), nrhs=nrhs)
        assert relerr(dx.download(), xo) <= TOL
    assert ce.flowAborts(fid) == 0


@pytest.mark.parametrize("prec", [R.FP64, R.FP32, R.BF16])
def test_flow_kacz_bit_identical_with_long_rows(prec):
    """A Chung-Lu Laplacian (hub rows: variable lanes per row, leaves larger
    than a bundle): one-launch sweeps equal the phase launches bit for bit."""
    rp, col, val = M.chung_lu_laplacian(6000, 12, seed=4)
    n = len(rp) - 1
    ce = R.Interface(n, 32, R.TWO, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    xdt = np.float64 if prec == R.FP64 else np.float32
    vp = R.FP64 if prec == R.FP64 else R.FP32
    b = M.uniform_vec(n, 8, nrhs=2).astype(xdt)
    xs = []
    for flow in (0, 1):
        ce.setOption("flow", flow)
        dx = R.DeviceArray.from_host(np.zeros(2 * n, xdt))
        db = R.DeviceArray.from_host(b)
        fid = ce.registerDevice(R.K_KACZ, prp, pcol, pval, dx, db, nrhs=2, val_precision=prec, vec_precision=vp)
        xs.append((fid, dx, db))
    (fa, xa, _), (fb, xb, _) = xs
    assert ce.launchCount(fb) == 1
    for s in range(4):
        rev = bool(s % 2)
        print("flow kacz long rows", prec, "sweep", s, flush=True)
        ce.executeFunction(fa, rev)
        ce.executeFunction(fb, rev)
        assert np.array_equal(xa.download(), xb.download()), s
    assert ce.flowAborts(fb) == 0


@pytest.mark.parametrize("bs", [512, 1024])
def test_flow_batch_kacz_equals_fused_phases(bs):
    """A batch of banded and Chung-Lu matrices (several kernel shapes,
    variable lanes), FP32 values and vectors, FP64 row norms, 8 right-hand
    sides: every matrix advancing through its own phases in one launch per
    shape group equals the fused phase launches bit for bit."""
    mats = list(M.batch_matrices(8, 11, 500, 6000))
    batches = [R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs=8) for _ in range(2)]
    keep = []
    for k, (name, n, mean, rp, col, val) in enumerate(mats):
        ce = R.Interface(n, max(4, n // 256), R.TWO, rp, col, symm_hint=True)
        ce.RACEColor()
        perm, inv = ce.getPerm(), ce.getInvPerm()
        prp, pcol, pval = M.permute(rp, col, val, perm, inv)
        prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
        b = M.uniform_vec(n, 70 + k, nrhs=8).astype(np.float32)
        pair = []
        for bt in batches:
            dx = R.DeviceArray.from_host(np.zeros(8 * n, np.float32))
            db = R.DeviceArray.from_host(b)
            bt.add(ce, prp, pcol, pval, dx, db)
            pair.append(dx)
        keep.append((ce, pair))
    for bt, flow in zip(batches, (0, 1)):
        bt.setOption("bundle_threads", bs)
        bt.setOption("bundle_rows", bs // 2)
        bt.setOption("flow", flow)
        bt.build()
    # one launch per kernel-shape group against one per group and phase
    assert batches[1].launchCount() < batches[0].launchCount()
    for s in range(4):
        rev = bool(s % 2)
        print("flow batch kacz", bs, "sweep", s, flush=True)
        for bt in batches:
            bt.execute(rev)
        for ce, (xa, xb) in keep:
            assert np.array_equal(xa.download(), xb.download()), s
    assert batches[1].flowAborts() == 0


def test_flow_batch_async_sweeps_in_order():
    """Several shape groups, so each sweep forks the groups' launches onto
    streams of the batch and joins them back into the batch's stream: four
    sweeps queued back to back without synchronisation (each sweep's launches
    must start after the previous sweep's joins) equal four synchronous
    sweeps bit for bit, and the last synchronous call waits for every
    group."""
    mats = list(M.batch_matrices(8, 11, 500, 6000))
    batches = [R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs=8) for _ in range(2)]
    keep = []
    for k, (name, n, mean, rp, col, val) in enumerate(mats):
        ce = R.Interface(n, max(4, n // 256), R.TWO, rp, col, symm_hint=True)
        ce.RACEColor()
        perm, inv = ce.getPerm(), ce.getInvPerm()
        prp, pcol, pval = M.permute(rp, col, val, perm, inv)
        prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
        b = M.uniform_vec(n, 90 + k, nrhs=8).astype(np.float32)
        pair = []
        for bt in batches:
            dx = R.DeviceArray.from_host(np.zeros(8 * n, np.float32))
            db = R.DeviceArray.from_host(b)
            bt.add(ce, prp, pcol, pval, dx, db)
            pair.append(dx)
        keep.append((ce, pair))
    for bt in batches:
        bt.setOption("bundle_threads", 512)
        bt.setOption("bundle_rows", 256)
        bt.setOption("flow", 1)
        bt.build()
    assert batches[0].launchCount() > 1, "the case needs several shape groups"
    print("flow batch async: 4 queued sweeps", flush=True)
    for s in range(3):
        batches[0].executeAsync(bool(s % 2))
    batches[0].execute(True)
    for s in range(4):
        batches[1].execute(bool(s % 2))
    for ce, (xa, xb) in keep:
        assert np.array_equal(xa.download(), xb.download())
    assert batches[0].flowAborts() == 0 and batches[1].flowAborts() == 0
