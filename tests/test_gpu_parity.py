"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer schedule data (levels, permutations, zone trees, phases, MPK groups)
must be bit-exact; FP64 kernels must agree to <= 1e-12 normwise relative
error (summation order differs: atomics, lane-group reductions, FMA).
"""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


def graphs():
    """(name, rowPtr, col) of small test graphs."""
    out = []
    rp, c = M.stencil7(12)
    out.append(("7pt-12", rp, c))
    rp, c = M.stencil27(10)
    out.append(("27pt-10", rp, c))
    # two disconnected 7-point blocks (island restart) with a dangling row
    rp1, c1 = M.stencil7(6)
    n1 = len(rp1) - 1
    rp2 = np.concatenate([rp1, rp1[1:] + rp1[-1], [2 * rp1[-1] + 1]]).astype(np.int32)
    c2 = np.concatenate([c1, c1 + n1, [2 * n1]]).astype(np.int32)
    out.append(("islands", rp2, c2))
    # 40 rows, each joined only to a private row: with distance 2 the range
    # [0, 40) needs three levels per row
    from .test_oracle import private_bridge_graph
    rp3, c3 = private_bridge_graph(40)
    out.append(("private-bridges", rp3, c3))
    return out


def _iface(rp, col, T, dist, symm=True, bfs_block_max=None):
    ce = R.Interface(len(rp) - 1, T, dist, rp, col, symm_hint=symm)
    if bfs_block_max is not None:
        ce.setOption("bfs_block_max", bfs_block_max)
    return ce


# traversal paths: default (ranges up to 2^17 rows in one workgroup),
# 0 = every range level by level over the device, 200 = both, by range size
BFS_PATHS = [None, 0, 200]


@pytest.mark.parametrize("bmax", BFS_PATHS)
@pytest.mark.parametrize("dist", [R.ONE, R.TWO])
def test_stage0_perm_matches_oracle(oracle, dist, bmax):
    for name, rp, col in graphs():
        print("case", name, dist, bmax, flush=True)
        ce = _iface(rp, col, 1, dist, bfs_block_max=bmax)
        ce.RACEColor()
        g = oracle.Graph(rp, col, symm_hint=True)
        g.traverse(dist, 0, g.n)
        assert np.array_equal(ce.getPerm(), g.perm), name
        assert np.array_equal(ce.getInvPerm(), g.inv), name


def test_spin_symmetrized_perm(oracle, spin):
    rp, col, _ = spin
    for dist in (R.ONE, R.TWO):
        ce = R.Interface(len(rp) - 1, 1, dist, rp, col, symm_hint=False)
        ce.RACEColor()
        g = oracle.Graph(rp, col, symm_hint=False)
        g.traverse(dist, 0, g.n)
        assert np.array_equal(ce.getPerm(), g.perm)


@pytest.mark.parametrize("T", [4, 32])
def test_spin_fixture_kernels(oracle, spin, T):
    """The reference's own matrix (examples/matrices/spinSZ12.mm.gz, 924 rows,
    400 without a diagonal) through every device kernel: GS at distance 1,
    KACZ and SymmSpMV at distance 2, MPK under POWER, each against the
    oracle in the RACE order of the oracle's tree.  A missing diagonal is
    inserted (as the harness's makeDiagFirst does) as 1 + the row's
    absolute sum, so that GS and KACZ are defined."""
    rp, col, val = spin
    n = len(rp) - 1
    for dist in (R.ONE, R.TWO):
        print("case spin", T, dist, flush=True)
        ce = R.Interface(n, T, dist, rp, col, symm_hint=False)
        ce.RACEColor()
        g = oracle.Graph(rp, col, symm_hint=False)
        zt = oracle.level_recursion(g, T, dist)
        perm, inv = ce.getPerm(), ce.getInvPerm()
        assert np.array_equal(perm, g.perm)
        prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, perm, inv))
        pval = pval.copy()
        for r in range(n):
            pval[prp[r]] = 1.0 + np.abs(pval[prp[r] + 1:prp[r + 1]]).sum()
        b = M.uniform_vec(n, 17)
        if dist == R.ONE:
            kinds = [(R.K_GS, oracle.K_GS)]
        else:
            kinds = [(R.K_KACZ, oracle.K_KACZ)]
        for kind, okind in kinds:
            dx, db = R.DeviceArray.from_host(np.zeros(n)), R.DeviceArray.from_host(b)
            fid = ce.registerDevice(kind, prp, pcol, pval, dx, db)
            xo = np.zeros(n)
            for rev in (False, True):
                ce.executeFunction(fid, rev)
                oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
                assert relerr(dx.download(), xo) <= TOL, (kind, rev)
        if dist == R.TWO:
            urp, ucol, uval = M.upper_triangle(prp, pcol, pval)
            x = M.uniform_vec(n, 18)
            dx, db = R.DeviceArray.from_host(x), R.DeviceArray.from_host(np.zeros(n))
            fid = ce.registerDevice(R.K_SYMMSPMV, urp, ucol, uval, dx, db)
            ce.executeFunction(fid)
            bo = np.zeros(n)
            oracle.run_calls(oracle.K_SYMMSPMV, urp, ucol, uval, x.copy(), bo, oracle.leaf_calls(zt))
            assert relerr(db.download(), bo) <= TOL
    k = 4
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=False)
    ce.RACEColor(k, 16 * 1024.0)
    g = oracle.Graph(rp, col, symm_hint=False)
    part = oracle.mpk_partition(g, k, 16 * 1024.0)
    assert np.array_equal(ce.getPerm(), g.perm)
    assert np.array_equal(ce.mpkGroups(), part["groupPtr"])
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    xs = np.zeros((k + 1) * n)
    xs[:n] = M.uniform_vec(n, 19)
    dx = R.DeviceArray.from_host(xs)
    ce.executeFunction(ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k))
    ref = oracle.mpk_serial(prp, pcol, pval, xs[:n], k)
    got = dx.download().reshape(k + 1, n)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL, p


@pytest.mark.parametrize("bmax", BFS_PATHS)
@pytest.mark.parametrize("T", [2, 3, 8, 20, 64])
@pytest.mark.parametrize("dist", [R.ONE, R.TWO])
def test_recursion_tree_and_schedule(oracle, T, dist, bmax):
    for name, rp, col in graphs()[:2]:
        print("case", name, T, dist, bmax, flush=True)
        ce = _iface(rp, col, T, dist, bfs_block_max=bmax)
        ce.RACEColor()
        g = oracle.Graph(rp, col, symm_hint=True)
        zt = oracle.level_recursion(g, T, dist)
        assert np.array_equal(ce.getPerm(), g.perm), name
        assert ce.getNumThreads() == zt.available_threads
        tr = ce.tree()
        ref = np.array([[l.valueZ[0], l.valueZ[-1], l.parentZ, l.nthreadsZ, l.stage, l.effRowZ]
                        for l in zt.tree], np.int32)
        assert np.array_equal(tr, ref), name
        for rev in (False, True):
            pp, lo, hi = ce.schedule(rev)
            ph = [[(a, b) for (a, b) in p if (b > a if not rev else a > b)] for p in oracle.phases(zt, rev)]
            ph = [p for p in ph if p]
            assert len(pp) - 1 == len(ph)
            for k, p in enumerate(ph):
                mine = list(zip(lo[pp[k]:pp[k + 1]], hi[pp[k]:pp[k + 1]]))
                want = [(b + 1, a + 1) if rev else (a, b) for (a, b) in p]
                assert mine == want, (name, rev, k)


def _prepared(rp, col, val, T, dist, oracle):
    """colour, permute (diagonal first), return the interface and oracle tree."""
    n = len(rp) - 1
    ce = R.Interface(n, T, dist, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    g = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(g, T, dist)
    assert np.array_equal(g.perm, perm)
    return ce, (prp, pcol, pval), zt


# (T, layout, bundle_rows, bundle_local): the last three force leaves larger
# than a bundle, run as consecutive segments in sub-phases
@pytest.mark.parametrize("T,layout,rows,local", [(1, 1, 0, 0), (4, 1, 0, 0), (16, 1, 0, 0), (64, 1, 0, 0),
                                                 (64, 0, 0, 0), (256, 2, 128, 0), (1024, 2, 128, 0),
                                                 (1, 2, 64, 512), (4, 2, 32, 256), (16, 2, 16, 128)])
def test_gs_matches_race_order(oracle, T, layout, rows, local):
    rp, col = M.stencil27(12)
    val = M.gs_values(rp, col, seed=7)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, R.ONE, oracle)
    b = M.uniform_vec(n, 11)
    x0 = M.uniform_vec(n, 12)
    dx = R.DeviceArray.from_host(x0)
    db = R.DeviceArray.from_host(b)
    ce.setOption("layout", layout)
    if rows:
        ce.setOption("bundle_rows", rows)
    if local:
        ce.setOption("bundle_local", local)
    fid = ce.registerDevice(R.K_GS, prp, pcol, pval, dx, db)
    xo = x0.copy()
    for rev in (False, True, False):
        print("sweep rev", rev, "T", T, flush=True)
        ce.executeFunction(fid, rev)
        oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL


@pytest.mark.parametrize("T,layout,rows,local", [(1, 1, 0, 0), (8, 1, 0, 0), (32, 1, 0, 0), (32, 0, 0, 0),
                                                 (256, 2, 256, 2048), (1, 2, 64, 512), (8, 2, 32, 256)])
def test_kacz_matches_race_order(oracle, T, layout, rows, local):
    rp, col = M.stencil7(14)
    val = M.jittered_values(rp, col, 0.1, seed=3)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, R.TWO, oracle)
    b = M.uniform_vec(n, 5)
    dx = R.DeviceArray.from_host(np.zeros(n))
    db = R.DeviceArray.from_host(b)
    ce.setOption("layout", layout)
    if rows:
        ce.setOption("bundle_rows", rows)
    if local:
        ce.setOption("bundle_local", local)
    fid = ce.registerDevice(R.K_KACZ, prp, pcol, pval, dx, db)
    xo = np.zeros(n)
    for rev in (False, True):
        ce.executeFunction(fid, rev)
        oracle.run_calls(oracle.K_KACZ, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL


@pytest.mark.parametrize("kind", ["gs", "kacz"])
@pytest.mark.parametrize("threads", [64, 128, 256, 512, 1024])
def test_bundle_workgroup_sizes(oracle, kind, threads):
    """Every bundle workgroup size (64 = one wavefront, rows per bundle
    capped by the lanes) against the oracle in RACE order."""
    rp, col = M.stencil27(12) if kind == "gs" else M.stencil7(14)
    val = M.gs_values(rp, col, seed=9)
    n = len(rp) - 1
    dist = R.ONE if kind == "gs" else R.TWO
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, 256, dist, oracle)
    b = M.uniform_vec(n, 13)
    dx = R.DeviceArray.from_host(np.zeros(n))
    db = R.DeviceArray.from_host(b)
    ce.setOption("layout", 2)
    ce.setOption("bundle_threads", threads)
    ce.setOption("bundle_rows", 1024)
    k = R.K_GS if kind == "gs" else R.K_KACZ
    okind = oracle.K_GS if kind == "gs" else oracle.K_KACZ
    fid = ce.registerDevice(k, prp, pcol, pval, dx, db)
    xo = np.zeros(n)
    for rev in (False, True):
        ce.executeFunction(fid, rev)
        oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL


@pytest.mark.parametrize("kind", ["gs", "kacz"])
@pytest.mark.parametrize("T,rows", [(4, 0), (16, 32)])
def test_long_rows_variable_lanes(oracle, kind, T, rows):
    """A Chung-Lu graph with hub rows of up to ~300 entries: rows longer
    than 64 entries select variable lanes per row (LDS row accumulators);
    large leaves run as chains.  FP64 against the oracle in RACE order."""
    rp, col, val = M.chung_lu_laplacian(2500, 12.0, seed=4, max_deg=300)
    assert np.diff(rp).max() > 65
    n = len(rp) - 1
    dist = R.ONE if kind == "gs" else R.TWO
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, dist, oracle)
    b = M.uniform_vec(n, 31)
    x0 = M.uniform_vec(n, 32)
    dx = R.DeviceArray.from_host(x0)
    db = R.DeviceArray.from_host(b)
    ce.setOption("layout", 2)
    if rows:
        ce.setOption("bundle_rows", rows)
    k = R.K_GS if kind == "gs" else R.K_KACZ
    okind = oracle.K_GS if kind == "gs" else oracle.K_KACZ
    fid = ce.registerDevice(k, prp, pcol, pval, dx, db)
    xo = x0.copy()
    for rev in (False, True):
        ce.executeFunction(fid, rev)
        oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL


def _block_diagonal(nblk, blk, seed):
    """nblk dense blk x blk blocks on the diagonal, diagonally dominant"""
    rng = np.random.default_rng(seed)
    rp, col, val = [0], [], []
    for bk in range(nblk):
        for r in range(blk):
            for c in range(blk):
                col.append(bk * blk + c)
                val.append(blk + 1.0 + rng.random() if c == r else -rng.random())
            rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val)


@pytest.mark.parametrize("kind", ["gs", "kacz"])
@pytest.mark.parametrize("blk,T", [(1, 16), (3, 16), (70, 4)])
def test_bundles_without_outside_columns(oracle, kind, blk, T):
    """Block-diagonal matrices: bundles whose rows reach no column outside
    the bundle (no outside-column gathers) next to bundles that do; blk = 1
    (a diagonal matrix) leaves Gauss-Seidel with no entries at all, only the
    zero entry that lanes past a bundle's width read; blk = 70 takes the
    variable-lanes kernels."""
    rp, col, val = _block_diagonal(600 // blk if blk < 70 else 12, blk, seed=blk)
    n = len(rp) - 1
    dist = R.ONE if kind == "gs" else R.TWO
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, dist, oracle)
    b = M.uniform_vec(n, 17)
    x0 = M.uniform_vec(n, 18)
    dx = R.DeviceArray.from_host(x0)
    db = R.DeviceArray.from_host(b)
    ce.setOption("layout", 2)
    k = R.K_GS if kind == "gs" else R.K_KACZ
    okind = oracle.K_GS if kind == "gs" else oracle.K_KACZ
    fid = ce.registerDevice(k, prp, pcol, pval, dx, db)
    xo = x0.copy()
    for rev in (False, True):
        print("case", kind, "blk", blk, "rev", rev, flush=True)
        ce.executeFunction(fid, rev)
        oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL


# symm_mode: 0 default, 1 colour phases, 2 SELL atomics, 3 expanded full matrix
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("T", [1, 8, 64])
def test_symmspmv(oracle, T, mode):
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=5)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, R.TWO, oracle)
    urp, ucol, uval = M.upper_triangle(prp, pcol, pval)
    x = M.uniform_vec(n, 1234)
    dx = R.DeviceArray.from_host(x)
    db = R.DeviceArray.from_host(np.zeros(n))
    if mode:
        ce.setOption("symm_mode", mode)
    fid = ce.registerDevice(R.K_SYMMSPMV, urp, ucol, uval, dx, db)
    ce.executeFunction(fid)
    bo = np.zeros(n)
    oracle.run_calls(oracle.K_SYMMSPMV, urp, ucol, uval, x.copy(), bo, oracle.leaf_calls(zt))
    full = np.zeros(n)
    oracle.run_calls(oracle.K_SPMV, prp, pcol, pval, x.copy(), full, [(0, n)])
    got = db.download()
    assert relerr(got, bo) <= TOL
    assert relerr(got, full) <= TOL


@pytest.mark.parametrize("bmax", BFS_PATHS)
def test_mpk_partition_all_graphs(oracle, bmax):
    """POWER traversals (adjacent island levels) and level groups, every graph."""
    for name, rp, col in graphs():
        print("case", name, bmax, flush=True)
        ce = _iface(rp, col, 1, R.POWER, bfs_block_max=bmax)
        ce.RACEColor(3, 16 * 1024.0)
        g = oracle.Graph(rp, col, symm_hint=True)
        part = oracle.mpk_partition(g, 3, 16 * 1024.0)
        assert np.array_equal(ce.getPerm(), g.perm), name
        assert np.array_equal(ce.mpkGroups(), part["groupPtr"]), name


def test_mpk_groups_and_powers(oracle):
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=9)
    n = len(rp) - 1
    k = 4
    cache = 64 * 1024.0  # small cache so there are many level groups
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, cache)
    g = oracle.Graph(rp, col, symm_hint=True)
    part = oracle.mpk_partition(g, k, cache)
    assert np.array_equal(ce.getPerm(), g.perm)
    assert np.array_equal(ce.mpkGroups(), part["groupPtr"])
    assert len(part["groupPtr"]) > 3
    assert ce.checkPowerCoverage()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    x0 = M.uniform_vec(n, 2)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL, p


@pytest.mark.parametrize("dist,T", [(R.ONE, 64), (R.TWO, 16)])
def test_rcm_initial_permutation(oracle, dist, T):
    """The harness path: a relabelled matrix, its reverse Cuthill-McKee order
    as initPerm (examples/src/helpers/sparsemat.cpp:715-771), then
    RACEColor: the permutation equals the oracle's bit for bit and GS / KACZ
    follow the RACE order of the oracle's tree."""
    import scipy.sparse as sp
    rp, col = M.stencil27(10)
    val = M.gs_values(rp, col, seed=3)
    n = len(rp) - 1
    q = np.random.default_rng(9).permutation(n)
    A = sp.csr_matrix((val, col, rp), shape=(n, n))[q][:, q].tocsr()
    A.sort_indices()
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    perm0, inv0 = R.rcm(rp, col)
    ce = R.Interface(n, T, dist, rp, col, symm_hint=True, initPerm=perm0, initInvPerm=inv0)
    ce.RACEColor()
    g = oracle.Graph(rp, col, symm_hint=True, init_perm=perm0, init_inv_perm=inv0)
    zt = oracle.level_recursion(g, T, dist)
    perm, inv = ce.getPerm(), ce.getInvPerm()
    assert np.array_equal(perm, g.perm)
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, perm, inv))
    b = M.uniform_vec(n, 2)
    dx, db = R.DeviceArray.from_host(np.zeros(n)), R.DeviceArray.from_host(b)
    kind, okind = (R.K_GS, oracle.K_GS) if dist == R.ONE else (R.K_KACZ, oracle.K_KACZ)
    fid = ce.registerDevice(kind, prp, pcol, pval, dx, db)
    xo = np.zeros(n)
    for rev in (False, True):
        ce.executeFunction(fid, rev)
        oracle.run_calls(okind, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL


def test_attach_communication(oracle):
    """attachCommunicationToFunction: accepted for POWER functions, stored and
    never called (one device has no remote rows, as the reference's non-MPI
    build); NOT_IMPLEMENTED for colouring handles; the two-argument form
    needs highestSubPower 1."""
    rp, col = M.stencil7(10)
    val = M.laplace7_values(rp, col)
    n, k = len(rp) - 1, 3
    calls = []
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, 64 * 1024.0)
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    xs = np.zeros((k + 1) * n)
    xs[:n] = M.uniform_vec(n, 1)
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    ce.attachCommunicationToFunction(fid, lambda p, sp, a: calls.append((p, sp)))
    ce.attachCommunicationToFunction(fid, lambda p, a: calls.append(p), simple=True)
    hid = ce.registerFunction(lambda s, e, p, sp, nu, a: None, power=k)
    ce.attachCommunicationToFunction(hid, lambda p, sp, a: calls.append((p, sp)))
    ce.executeFunction(fid)
    ce.executeFunction(hid)
    assert calls == []
    ref = oracle.mpk_serial(prp, pcol, pval, xs[:n], k)
    assert relerr(dx.download().reshape(k + 1, n)[k], ref[k]) <= TOL
    with pytest.raises(R.RaceError):
        ce.attachCommunicationToFunction(99, lambda p, sp, a: None)
    sub = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    sub.RACEColor(2, 64 * 1024.0, highestSubPower=2)
    sid = sub.registerFunction(lambda s, e, p, sp, nu, a: None, power=2, subPower=2)
    sub.attachCommunicationToFunction(sid, lambda p, sp, a: None)
    with pytest.raises(R.RaceError, match="INCOMPATIBILITY"):
        sub.attachCommunicationToFunction(sid, lambda p, a: None, simple=True)
    col1 = _iface(rp, col, 4, R.ONE)
    col1.RACEColor()
    cid = col1.registerFunction(lambda s, e, a: None)
    with pytest.raises(R.RaceError, match="NOT_IMPLEMENTED"):
        col1.attachCommunicationToFunction(cid, lambda p, sp, a: None)


def test_registered_vectors_stay_alive(oracle):
    """A device array passed to registerDevice without another reference
    must live as long as the registration (the kernel writes through it)."""
    import gc
    rp, col = M.stencil7(8)
    val = M.laplace7_values(rp, col)
    n = len(rp) - 1
    ce = _iface(rp, col, 1, R.ONE)
    ce.RACEColor()
    x = M.uniform_vec(n, 3)
    dx = R.DeviceArray.from_host(x)
    fid = ce.registerDevice(R.K_SPMV, rp, col, val, dx, R.DeviceArray(n))
    gc.collect()
    kept = ce._vecs.get(fid)
    if kept is None or kept[1] is None or not kept[1].ptr:
        pytest.fail("registered output array was not retained")
    ce.executeFunction(fid)
    ref = M.to_scipy(rp, col, val) @ x
    assert relerr(kept[1].download(), ref) <= TOL


def test_host_callback_runs_race_order(oracle):
    """registerFunction with a CPU callback: GS through the host executor."""
    rp, col = M.stencil27(8)
    val = M.gs_values(rp, col, seed=1)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, 8, R.ONE, oracle)
    b = M.uniform_vec(n, 3)
    x = np.zeros(n)
    calls = []

    def gs_cb(s, e, args):
        calls.append((s, e))
        oracle.run_calls(oracle.K_GS, prp, pcol, pval, x, b, [(s, e)])

    fid = ce.registerFunction(gs_cb)
    ce.executeFunction(fid)
    xo = np.zeros(n)
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt))
    assert relerr(x, xo) == 0.0


def test_cpp_drop_in_program():
    """A C++ program written against RACE::Interface (include/race_amd_interface.hpp)
    runs a CPU Gauss-Seidel callback and the GPU kernel and compares them."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "drop_in")
    assert os.path.exists(exe), "build with python -m race_amd.build"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    assert "relerr=" in p.stdout


@pytest.mark.parametrize("T,C,w", [(8, 8, 4), (64, 16, 8), (20, 64, 4)])
def test_sell_c_sigmize_and_simdify(oracle, T, C, w):
    """Padding by the colouring leaves (the reference runs sell_c_sigmize as
    a RACE callback), then the distance-2 rearrangement, against the oracle
    on the same leaf calls."""
    rp, col = M.stencil27(6)
    val = M.laplace7_values(rp, col)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, T, R.TWO, oracle)
    calls = oracle.leaf_calls(zt)
    cs, rl, clp, scol, sval = M.to_sell(prp, pcol, pval, C, pad_col=-1)
    a = scol.copy()
    ce.sell_c_sigmize(w, C, a, cs, rl, clp)
    b = scol.copy()
    oracle.sell_c_sigmize(calls, w, C, b, cs, rl, clp)
    assert np.array_equal(a, b)
    pc, pv = scol.copy(), sval.copy()
    ok = ce.simdify(w, C, n, pc, cs, rl, clp, pv, diagFirst=True)
    oc, ov = scol.copy(), sval.copy()
    oracle.sell_c_sigmize(calls, w, C, oc, cs, rl, clp)
    assert ok == oracle.simdify_resolve(w, C, n, oc, cs, clp, ov, True)
    assert np.array_equal(pc, oc) and np.array_equal(pv, ov)


@pytest.mark.parametrize("bs", [256, 512])
def test_gs_bundles_with_wide_halos(oracle, bs):
    """Bundles of many small leaves (27-point 24^3 at 2048 RACE threads):
    more than 2 x workgroup-size outside columns per bundle, gathered in one
    batch (6 x 256 or 4 x 512 id slots, the rest round by round), against
    the oracle in RACE order."""
    rp, col = M.stencil27(24)
    val = M.gs_values(rp, col, seed=17)
    n = len(rp) - 1
    ce, (prp, pcol, pval), zt = _prepared(rp, col, val, 2048, R.ONE, oracle)
    ce.setOption("bundle_threads", bs)
    b = M.uniform_vec(n, 3)
    x0 = M.uniform_vec(n, 4)
    dx, db = R.DeviceArray.from_host(x0), R.DeviceArray.from_host(b)
    fid = ce.registerDevice(R.K_GS, prp, pcol, pval, dx, db)
    xo = x0.copy()
    for rev in (False, True, False):
        print("wide halo GS", bs, rev, flush=True)
        ce.executeFunction(fid, rev)
        oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt, rev))
        assert relerr(dx.download(), xo) <= TOL
