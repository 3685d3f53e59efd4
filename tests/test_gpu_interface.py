"""The rest of the RACE::Interface surface on the device: matrix types "L" /
"U" of the POWER preprocessing, device matrix powers with sub-powers, host
callbacks on OpenMP threads (colouring phases and power wavefront), the
serial mode, NUMA first-touch helpers, and the refinement's argument
checks; plus the error of FP32 Kaczmarz with FP64 row norms (the batched
workload's arithmetic) against the FP64 oracle."""
import threading

import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def banded(n, offsets, seed):
    """symmetric banded pattern with the given positive offsets"""
    rows, cols = [], []
    for o in [0] + list(offsets):
        r = np.arange(n - o)
        rows += [r, r + o] if o else [r]
        cols += [r + o, r] if o else [r]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(r, minlength=n), out=rp[1:])
    val = np.where(r == c, 2.0, -0.1) * (1 + 0.05 * np.random.default_rng(seed).uniform(-1, 1, len(c)))
    return rp, c.astype(np.int32), val


@pytest.mark.parametrize("mtx", ["L", "U"])
def test_mtx_type_row_order_levels(oracle, mtx):
    """mtxType L / U (src/traverse_GAP_graphSoA.cpp:316-335): every row is a
    level in row order (L) or reversed row order (U), no BFS.  The
    permutation is the identity / the reversal, the groups equal the
    oracle's, and the powers equal repeated SpMV (a banded matrix keeps its
    entries within neighbouring groups)."""
    n, k = 6000, 4
    rp, col, val = banded(n, [1, 2, 5], seed=3)
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, 16 * 1024.0, mtxType=mtx)
    g = oracle.Graph(rp, col, symm_hint=True)
    part = oracle.mpk_partition(g, k, 16 * 1024.0, mtx_type=mtx)
    ident = np.arange(n)
    assert np.array_equal(ce.getPerm(), ident if mtx == "L" else ident[::-1])
    assert np.array_equal(ce.getPerm(), g.perm)
    assert np.array_equal(ce.mpkGroups(), part["groupPtr"])
    assert len(part["groupPtr"]) > 10
    assert ce.checkPowerCoverage()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    x0 = M.uniform_vec(n, 5)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    A = M.to_scipy(prp, pcol, pval)
    y = x0.copy()
    for p in range(1, k + 1):
        y = A @ y
        assert relerr(got[p], y) <= TOL, p


def test_mtx_type_row_order_refuses_far_entries():
    """With row order as levels, a 7-point stencil's +-g^2 entries cross many
    small groups: the slope-2 wavefront would read powers not yet computed,
    so registration refuses the matrix (RACE_ERR_INCOMPATIBILITY)."""
    rp, col = M.stencil7(12)
    val = M.laplace7_values(rp, col)
    n = len(rp) - 1
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(3, 2 * 1024.0, mtxType="L")
    dx = R.DeviceArray(4 * n)
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_MPK, rp, col, val, dx, power=3)
    with pytest.raises(R.RaceError):
        ce2 = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
        ce2.RACEColor(3, 2 * 1024.0, mtxType="X")


@pytest.mark.parametrize("power,sub", [(2, 3), (4, 1), (1, 5)])
def test_device_mpk_sub_powers(oracle, power, sub):
    """registerFunction(..., power, subPower) (include/interface.h:153-157):
    the executor visits every (pow, subPow) pair in order; the device MPK
    does one product with A per pair, product t = (pow-1) subPower + subPow
    in slot t.  Equal to power * subPower repeated SpMVs; setSubPower /
    setPower accept at most the registered products."""
    rp, col = M.stencil7(14)
    val = M.jittered_values(rp, col, 0.1, seed=2) / 6.0
    n = len(rp) - 1
    k = power * sub
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(power, 32 * 1024.0, highestSubPower=sub)
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    x0 = M.uniform_vec(n, 7)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=power, subPower=sub)
    assert ce.getPower(fid) == power and ce.getSubPower(fid) == sub
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    for t in range(1, k + 1):
        assert relerr(got[t], ref[t]) <= TOL, t
    with pytest.raises(R.RaceError):
        ce.setSubPower(fid, sub + 1)
    with pytest.raises(R.RaceError):
        ce.setPower(fid, power + 1)
    # fewer products: only the first slots are rewritten
    if k > 1:
        dx.upload(np.concatenate([x0, np.zeros(k * n)]))
        ce.setSubPower(fid, 1)
        ce.executeFunction(fid)
        got = dx.download().reshape(k + 1, n)
        for t in range(1, power + 1):
            assert relerr(got[t], ref[t]) <= TOL, t
        assert not got[power + 1:].any()


def test_host_colour_callbacks_run_on_threads(oracle):
    """Host Gauss-Seidel callbacks: the leaves of a colour phase run on
    several OpenMP threads (seen from the callback), and the result equals
    the oracle in RACE order; setSerial runs the same on one thread with the
    same result (src/functionManager.cpp:199-265, 1561-1580)."""
    rp, col = M.stencil27(12)
    val = M.gs_values(rp, col, seed=5)
    n = len(rp) - 1
    ce = R.Interface(n, 16, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    g = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(g, 16, oracle.ONE)
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
    b = M.uniform_vec(n, 9)
    threads = set()

    def gs(start, end, x):
        threads.add(threading.get_ident())
        step = -1 if start > end else 1
        for r in range(start, end, step):
            s, e = prp[r], prp[r + 1]
            x[r] = (b[r] - np.dot(pval[s + 1:e], x[pcol[s + 1:e]])) / pval[s]

    x = np.zeros(n)
    fid = ce.registerFunction(gs, x)
    ce.executeFunction(fid)
    ce.executeFunction(fid, rev=True)
    xo = np.zeros(n)
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt))
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, True))
    assert relerr(x, xo) <= TOL
    assert len(threads) > 1
    threads.clear()
    x1 = np.zeros(n)
    sid = ce.registerFunction(gs, x1)
    ce.setSerial(sid)
    ce.executeFunction(sid)
    ce.executeFunction(sid, rev=True)
    assert len(threads) == 1
    assert relerr(x1, xo) <= TOL


def test_host_power_callbacks_with_sub_powers(oracle):
    """Power callbacks on threads over the slope-2 wavefront: every row sees
    (pow, subPow) in order exactly once (the check of
    MAT_NUM_VEC_ACCESSES_w_subPower, examples/src/helpers/sparsemat.cpp:
    1399-1412), and a host matrix-power callback equals repeated SpMV."""
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=8) / 6.0
    n = len(rp) - 1
    power, sub = 3, 3
    ce = R.Interface(n, 8, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(power, 32 * 1024.0, highestSubPower=sub)
    seen = np.zeros(n, np.int64)
    bad = []

    def count(start, end, pw, sp, numa, _):
        want = (pw - 1) * sub + sp
        if np.any(seen[start:end] != want - 1):
            bad.append((start, end, pw, sp))
        seen[start:end] = want

    fid = ce.registerFunction(count, None, power=power, subPower=sub)
    ce.executeFunction(fid)
    assert not bad and np.all(seen == power * sub)
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    A = M.to_scipy(prp, pcol, pval).tocsr()
    k = power * sub
    xs = np.zeros((k + 1, n))
    xs[0] = M.uniform_vec(n, 3)

    def mpk(start, end, pw, sp, numa, _):
        t = (pw - 1) * sub + sp
        xs[t, start:end] = A[start:end] @ xs[t - 1]

    mid = ce.registerFunction(mpk, None, power=power, subPower=sub)
    ce.executeFunction(mid)
    y = xs[0].copy()
    for t in range(1, k + 1):
        y = A @ y
        assert relerr(xs[t], y) <= TOL, t


def test_numa_helpers_and_no_op_methods():
    """numaInitRowPtr / numaInitMtxVec zero their arrays through a power
    callback (src/interface.cpp:852-1040); getNumaSplitting is one domain
    [0, n]; pinThread, sleep and resetTime are accepted; the NUMA helpers
    refuse colouring handles, as the reference does."""
    rp, col = M.stencil7(10)
    n = len(rp) - 1
    ce = R.Interface(n, 4, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(2, 16 * 1024.0)
    assert ce.getNumaSplitting().tolist() == [0, n]
    r2 = np.full(n + 1, 7, np.int32)
    ce.numaInitRowPtr(r2)
    assert not r2.any()
    c2 = np.full(len(col), 5, np.int32)
    v2 = np.full(len(col), 3.0)
    x2 = np.full(3 * n, 2.0)
    ce.numaInitMtxVec(rp, c2, v2, x2, power=2)
    assert not c2.any() and not v2.any() and not x2[:2 * n].any() and np.all(x2[2 * n:] == 2.0)
    ce.pinThread(0)
    ce.sleep()
    ce.resetTime()
    cc = R.Interface(n, 4, R.ONE, rp, col, symm_hint=True)
    cc.RACEColor()
    with pytest.raises(R.RaceError):
        cc.numaInitRowPtr(r2)
    with pytest.raises(R.RaceError):
        cc.getNumaSplitting()


def test_refinement_checks_the_inner_function(oracle):
    """The refinement's workspace has the refinement's rows and right-hand
    sides: an inner function with another nrhs is refused (it would index
    past the workspace).  With nrhs 3 the refinement converges every
    right-hand side, and afterwards the inner function still runs on its own
    vectors."""
    rp, col = M.stencil27(10)
    val = M.gs_values(rp, col, seed=2)
    n = len(rp) - 1
    ce = R.Interface(n, 32, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
    nr = 3
    wx = R.DeviceArray(n * nr, np.float32)
    wb = R.DeviceArray.from_host(np.ones(n * nr, np.float32))
    inner = ce.registerDevice(R.K_GS, prp, pcol, pval, wx, wb, nrhs=nr, val_precision=R.FP32,
                              vec_precision=R.FP32)
    xs = M.uniform_vec(n, 4, nrhs=nr)
    A = M.to_scipy(prp, pcol, pval)
    b = np.concatenate([A @ xs[k * n:(k + 1) * n] for k in range(nr)])
    x = R.DeviceArray.from_host(np.zeros(n * nr))
    db = R.DeviceArray.from_host(b)
    with pytest.raises(R.RaceError):
        ce.registerRefinement(R.K_GS, prp, pcol, pval, x, db, inner, sweeps=2, nrhs=1)
    rid = ce.registerRefinement(R.K_GS, prp, pcol, pval, x, db, inner, sweeps=2, nrhs=nr)
    steps, sweeps, rr = ce.refine(rid, 1e-10, max_outer=300)
    assert rr <= 1e-10
    assert relerr(x.download(), xs) <= 1e-8
    # the inner function is back on (wx, wb): one sweep from 0 with b = 1
    wx.upload(np.zeros(n * nr, np.float32))
    ce.executeFunction(inner)
    got = wx.download().astype(np.float64)
    g = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(g, 32, oracle.ONE)
    xo = np.zeros(n * nr)
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, np.ones(n * nr), oracle.leaf_calls(zt), nrhs=nr)
    assert relerr(got, xo) <= 1e-6


def test_fp32_kaczmarz_fp64_norms_error_vs_fp64_oracle(oracle):
    """The batched workload's arithmetic: Kaczmarz sweeps with FP32 values
    and vectors, FP32 accumulation and FP64 row norms, 8 right-hand sides,
    10 forward + backward sweep pairs on a banded and a Chung-Lu matrix.
    Stated error against the FP64 oracle on the original FP64 data in RACE
    order (storage rounding included): normwise relative error <= 2e-5 for
    every right-hand side block."""
    nr = 8
    mats = list(M.batch_matrices(2, 11, 3000, 6000))
    batch = R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs=nr)
    keep = []
    for k, (name, n, mean, rp, col, val) in enumerate(mats):
        ce = R.Interface(n, 64, R.TWO, rp, col, symm_hint=True)
        ce.RACEColor()
        prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
        g = oracle.Graph(rp, col, symm_hint=True)
        zt = oracle.level_recursion(g, 64, oracle.TWO)
        x0 = M.uniform_vec(n, 60 + k, nrhs=nr)
        b = M.uniform_vec(n, 70 + k, nrhs=nr)
        dx = R.DeviceArray.from_host(x0.astype(np.float32))
        db = R.DeviceArray.from_host(b.astype(np.float32))
        batch.add(ce, prp, pcol, pval, dx, db)
        keep.append((name, n, ce, prp, pcol, pval, zt, x0.copy(), b, dx))
    batch.build()
    for _ in range(10):
        batch.execute(False)
        batch.execute(True)
    worst = 0.0
    for name, n, ce, prp, pcol, pval, zt, xo, b, dx in keep:
        for _ in range(10):
            oracle.run_calls(oracle.K_KACZ, prp, pcol, pval, xo, b, oracle.leaf_calls(zt), nrhs=nr)
            oracle.run_calls(oracle.K_KACZ, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, True), nrhs=nr)
        got = dx.download().astype(np.float64)
        for k in range(nr):
            e = relerr(got[k * n:(k + 1) * n], xo[k * n:(k + 1) * n])
            worst = max(worst, e)
        print(name, n, "worst rhs error vs FP64 oracle", worst, flush=True)
    assert worst <= 2e-5


def _grid_hubs(g, hubs):
    n0 = g * g
    rows = [[] for _ in range(n0 + hubs)]
    for i in range(g):
        for j in range(g):
            r = i * g + j
            for di, dj in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
                a, b = i + di, j + dj
                if 0 <= a < g and 0 <= b < g:
                    rows[r].append(a * g + b)
    for h in range(hubs):
        rows[n0 + h].append(n0 + h)
        for r in range(0, n0, 3):  # every third grid row
            rows[r].append(n0 + h)
            rows[n0 + h].append(r)
    rp = np.concatenate([[0], np.cumsum([len(x) for x in rows])]).astype(np.int32)
    col = np.array([c for x in rows for c in sorted(x)], np.int32)
    return rp, col


@pytest.mark.parametrize("dist,kind,T", [(R.ONE, R.K_GS, 32), (R.ONE, R.K_GS, 4)])
def test_dense_rows_run_as_a_serial_tail(oracle, monkeypatch, dist, kind, T):
    """RACE_MAX_DENSE: the hub rows leave the colouring and run last, in
    order, in both sweep directions.  Permutation bit for bit with the
    oracle; device sweeps and a host callback against the oracle's leaf
    calls plus the serial tail."""
    monkeypatch.setenv("RACE_MAX_DENSE", "4")
    rp, col = _grid_hubs(40, 3)
    n = len(rp) - 1
    val = M.gs_values(rp, col, seed=6)
    ce = R.Interface(n, T, dist, rp, col, symm_hint=True)
    ce.RACEColor()
    G = oracle.Graph(rp, col, symm_hint=True, max_dense=4)
    assert G.dense_rows == [1600, 1601, 1602]
    zt = oracle.level_recursion(G, T, dist)
    perm = ce.getPerm()
    assert np.array_equal(perm, G.full_perm())
    ns = G.n
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, perm, ce.getInvPerm()))
    b = M.uniform_vec(n, 4)
    x0 = M.uniform_vec(n, 5)
    dx, db = R.DeviceArray.from_host(x0), R.DeviceArray.from_host(b)
    fid = ce.registerDevice(kind, prp, pcol, pval, dx, db)
    okind = oracle.K_GS if kind == R.K_GS else oracle.K_KACZ
    xo = x0.copy()
    xh = x0.copy()

    def host(start, end, x):
        oracle.run_calls(okind, prp, pcol, pval, x, b, [(start, end)])

    hid = ce.registerFunction(host, xh)
    ce.setSerial(hid)
    for rev in (False, True):
        ce.executeFunction(fid, rev)
        ce.executeFunction(hid, rev)
        oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev, serial_part=(ns, n)))
        assert relerr(dx.download(), xo) <= TOL
        assert relerr(xh, xo) <= TOL
    with pytest.raises(R.RaceError):
        ce.saveSchedule()


def test_dense_rows_refuse_column_writing_kernels(oracle, monkeypatch):
    """Kaczmarz writes every column of its row: two concurrent rows of a
    distance-2 phase may share a dense column once dense rows are no graph
    vertices, so the device function is refused (the colouring itself and
    the permutation still follow the reference)."""
    monkeypatch.setenv("RACE_MAX_DENSE", "4")
    rp, col = _grid_hubs(20, 2)
    n = len(rp) - 1
    val = M.gs_values(rp, col, seed=2)
    ce = R.Interface(n, 8, R.TWO, rp, col, symm_hint=True)
    ce.RACEColor()
    G = oracle.Graph(rp, col, symm_hint=True, max_dense=4)
    oracle.level_recursion(G, 8, oracle.TWO)
    assert np.array_equal(ce.getPerm(), G.full_perm())
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm()))
    dx, db = R.DeviceArray(n), R.DeviceArray(n)
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_KACZ, prp, pcol, pval, dx, db)
