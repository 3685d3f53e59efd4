"""Pinning the CPU oracle (oracle/race_oracle.c, oracle/race_oracle.py).

The reference cannot be built in this image, and its tree holds no golden
output vectors (only examples/matrices/spinSZ12.mm.gz).  Every oracle
function is therefore pinned by checks a plausible mistake would fail:
closed forms, textbook reductions computed independently with scipy,
hand-traced cases, brute-force invariants.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
from scipy.sparse.linalg import spsolve_triangular

from race_amd import matgen as M

from .conftest import relerr


def grid_coords(g):
    idx = np.arange(g ** 3)
    return idx // (g * g), (idx // g) % g, idx % g


# ---------------------------------------------------------------------------
# traversal
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("g", [3, 5, 8])
def test_bfs_closed_form_7pt(oracle, g):
    rp, col = M.stencil7(g)
    G = oracle.Graph(rp, col, symm_hint=True)
    lr, ln, dist = G.traverse(oracle.TWO, 0, G.n, want_dist=True)
    i, j, k = grid_coords(g)
    lvl = i + j + k
    assert len(lr) == 3 * g - 2
    assert np.array_equal(dist, lvl)
    assert np.array_equal(lr, np.bincount(lvl))
    rowlen = np.diff(rp)
    assert np.array_equal(ln, np.bincount(lvl, weights=rowlen).astype(int))
    # stage-0 permutation = rows stably sorted by level
    assert np.array_equal(G.perm, np.argsort(lvl, kind="stable"))
    assert np.array_equal(G.inv[G.perm], np.arange(G.n))


@pytest.mark.parametrize("dist", [1, 2])
def test_init_perm_composes(oracle, dist):
    """initPerm (src/graph_SoA.cpp:64-79 with RACE_PERMUTE_ON_FLY): the
    traversal runs on the initially permuted graph and composes into it.  A
    grid relabelled by q and handed back its natural order as initPerm
    must give perm = qinv[perm of the natural grid]."""
    rp, col = M.stencil7(6)
    n = len(rp) - 1
    q = np.random.default_rng(4).permutation(n)
    qinv = np.empty(n, np.int64)
    qinv[q] = np.arange(n)
    B = M.to_scipy(rp, col, np.ones(len(col)))[q][:, q].tocsr()
    B.sort_indices()
    ga = oracle.Graph(rp, col, symm_hint=True)
    ga.traverse(dist, 0, n)
    gb = oracle.Graph(B.indptr, B.indices, symm_hint=True, init_perm=qinv, init_inv_perm=q)
    gb.traverse(dist, 0, n)
    assert np.array_equal(gb.perm, qinv[ga.perm])
    assert np.array_equal(gb.inv[gb.perm], np.arange(n))


@pytest.mark.parametrize("g", [3, 6])
def test_bfs_closed_form_27pt(oracle, g):
    rp, col = M.stencil27(g)
    G = oracle.Graph(rp, col, symm_hint=True)
    lr, ln, dist = G.traverse(oracle.ONE, 0, G.n, want_dist=True)
    i, j, k = grid_coords(g)
    lvl = np.maximum(np.maximum(i, j), k)
    assert np.array_equal(dist, lvl)
    L = np.arange(g)
    assert np.array_equal(lr, 3 * L * L + 3 * L + 1)
    assert np.array_equal(G.perm, np.argsort(lvl, kind="stable"))


def _random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=p, random_state=rng, format="csr")
    A = ((A + A.T) != 0).astype(np.float64)
    A.setdiag(1.0)
    A = A.tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_bfs_levels_are_shortest_paths_with_island_rule(oracle, seed):
    """Connected part: levels = unweighted shortest path from row 0.
    Islands: restart at the smallest unvisited row, one empty level apart
    for colouring, adjacent for POWER (computed here with scipy)."""
    n = 300
    rp, col = _random_graph(n, 0.006 + 0.002 * seed, seed)
    A = sp.csr_matrix((np.ones(len(col)), col, rp), shape=(n, n))
    for dist_t in (oracle.ONE, oracle.POWER):
        G = oracle.Graph(rp, col, symm_hint=True)
        lr, ln, dist = G.traverse(dist_t, 0, n, want_dist=True)
        want = np.full(n, -1)
        base = 0
        prev_max = None
        while (want < 0).any():
            root = int(np.nonzero(want < 0)[0][0])
            d = csg.shortest_path(A, unweighted=True, indices=root)
            comp = np.isfinite(d)
            if prev_max is None:
                base = 0
            else:
                base = prev_max + (2 if dist_t != oracle.POWER else 1)
            want[comp] = base + d[comp].astype(int)
            prev_max = want[comp].max()
        assert np.array_equal(dist, want), dist_t
        assert len(lr) == want.max() + 1


def test_bfs_distance2_bridges_through_rows_outside_the_range(oracle):
    """Hand-traced: range [0,4) of a 6-row graph in which rows 2 and 3 are
    reached from row 1 only through row 5 (outside the range).  Distance-2
    bridging puts 2 and 3 both at level 3 (level 2 holds only the bridge);
    distance-1 cannot cross row 5: rows 2 and 3 are islands (no edge between
    them), each restarted one empty level after the previous one: 3 and 5."""
    edges = [(0, 1), (1, 5), (5, 2), (5, 3), (4, 4)]
    n = 6
    A = sp.lil_matrix((n, n))
    for a, b in edges:
        A[a, b] = 1
        A[b, a] = 1
    for r in range(n):
        A[r, r] = 1
    A = A.tocsr()
    A.sort_indices()
    rp, col = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    G = oracle.Graph(rp, col, symm_hint=True)
    _, _, d2 = G.traverse(oracle.TWO, 0, 4, want_dist=True)
    assert d2.tolist() == [0, 1, 3, 3]
    G = oracle.Graph(rp, col, symm_hint=True)
    _, _, d1 = G.traverse(oracle.ONE, 0, 4, want_dist=True)
    assert d1.tolist() == [0, 1, 3, 5]


def private_bridge_graph(k):
    """Rows 0..k-1 (the range) each joined only to its own row k+i outside
    the range."""
    n = 2 * k
    A = sp.lil_matrix((n, n))
    for i in range(k):
        A[i, k + i] = A[k + i, i] = 1
    for r in range(n):
        A[r, r] = 1
    A = A.tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


def test_bfs_distance2_three_levels_per_isolated_row(oracle):
    """Hand-traced: every row of the range [0,k) is an island whose only
    neighbour is a private row outside the range.  Distance 2: the root's
    level, one level expanding to its bridge, one expanding the bridge
    (nothing new), then the next island restarts one empty level later --
    row i at level 3i, 3k-1 levels in all (more than 2k+2)."""
    k = 40
    rp, col = private_bridge_graph(k)
    G = oracle.Graph(rp, col, symm_hint=True)
    lr, ln, d = G.traverse(oracle.TWO, 0, k, want_dist=True)
    assert d.tolist() == [3 * i for i in range(k)]
    assert len(lr) == 3 * k - 1
    assert lr.tolist() == [1 if L % 3 == 0 else 0 for L in range(3 * k - 1)]
    # distance 1 ignores the outside rows: islands two levels apart
    G = oracle.Graph(rp, col, symm_hint=True)
    lr, ln, d = G.traverse(oracle.ONE, 0, k, want_dist=True)
    assert d.tolist() == [2 * i for i in range(k)]


def test_subrange_traversal_composes_permutation(oracle):
    rp, col = M.stencil7(6)
    G = oracle.Graph(rp, col, symm_hint=True)
    G.traverse(oracle.ONE, 0, G.n)
    before = G.perm.copy()
    lo, hi = 40, 120
    lr, ln = G.traverse(oracle.ONE, lo, hi)
    assert lr.sum() == hi - lo
    assert np.array_equal(G.perm[:lo], before[:lo]) and np.array_equal(G.perm[hi:], before[hi:])
    assert sorted(G.perm[lo:hi]) == sorted(before[lo:hi])
    assert np.array_equal(G.inv[G.perm], np.arange(G.n))


def test_spin_fixture_and_symmetrisation(oracle, spin):
    """examples/matrices/spinSZ12.mm.gz: 924 x 924, 6572 entries (its header).
    The symmetrised pattern must be the pattern of A + A^T exactly."""
    rp, col, val = spin
    n = len(rp) - 1
    assert n == 924 and rp[-1] == 6572
    A = sp.csr_matrix((np.ones(len(col)), col, rp), shape=(n, n))
    G = oracle.Graph(rp, col, symm_hint=False)
    S = sp.csr_matrix((np.ones(len(G.col)), G.col, G.rowPtr), shape=(n, n))
    want = ((A + A.T) != 0)
    assert (S != want).nnz == 0
    assert len(G.col) == want.nnz  # no duplicates added


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
def _gs_matrix(g=6, seed=3):
    rp, col = M.stencil27(g)
    val = M.gs_values(rp, col, seed=seed)
    rp, col, val = M.make_diag_first(rp, col, val)
    return rp, col, val


def test_gs_one_leaf_is_textbook_gauss_seidel(oracle):
    rp, col, val = _gs_matrix()
    n = len(rp) - 1
    A = M.to_scipy(rp, col, val).tocsr()
    b = M.uniform_vec(n, 1)
    x0 = M.uniform_vec(n, 2)
    Lw = sp.tril(A, k=0, format="csr")
    Up = sp.triu(A, k=1, format="csr")
    fwd = spsolve_triangular(Lw, b - Up @ x0, lower=True)
    x = x0.copy()
    oracle.run_calls(oracle.K_GS, rp, col, val, x, b, [(0, n)])
    assert relerr(x, fwd) < 1e-13
    Uw = sp.triu(A, k=0, format="csr")
    Lo = sp.tril(A, k=-1, format="csr")
    bwd = spsolve_triangular(Uw, b - Lo @ fwd, lower=False)
    oracle.run_calls(oracle.K_GS, rp, col, val, x, b, [(n - 1, -1)])
    assert relerr(x, bwd) < 1e-13


def test_kacz_step_projects_onto_row(oracle):
    """After the Kaczmarz step of row r, a_r . x == b_r, and x moved along a_r."""
    rp, col, val = _gs_matrix()
    n = len(rp) - 1
    A = M.to_scipy(rp, col, val).tocsr()
    b = M.uniform_vec(n, 4)
    x = M.uniform_vec(n, 5)
    rng = np.random.default_rng(0)
    for r in rng.integers(0, n, 20):
        x_old = x.copy()
        oracle.run_calls(oracle.K_KACZ, rp, col, val, x, b, [(int(r), int(r) + 1)])
        ar = A.getrow(int(r)).toarray().ravel()
        assert abs(ar @ x - b[r]) < 1e-12 * (abs(b[r]) + np.abs(ar) @ np.abs(x))
        dx = x - x_old
        # dx is parallel to a_r
        assert np.linalg.norm(dx - (dx @ ar) / (ar @ ar) * ar) < 1e-12 * max(1.0, np.linalg.norm(dx))


def test_kacz_sweep_converges_to_solution(oracle):
    rp, col, val = _gs_matrix(5, 8)
    n = len(rp) - 1
    A = M.to_scipy(rp, col, val).tocsr()
    xs = M.uniform_vec(n, 9)
    b = A @ xs
    x = np.zeros(n)
    # each step projects onto a hyperplane containing xs: the error never grows
    prev = np.linalg.norm(xs)
    e0 = prev
    for _ in range(10):
        oracle.run_calls(oracle.K_KACZ, rp, col, val, x, b, [(0, n)])
        oracle.run_calls(oracle.K_KACZ, rp, col, val, x, b, [(n - 1, -1)])
        e = np.linalg.norm(x - xs)
        assert e <= prev * (1 + 1e-12)
        prev = e
    assert prev < 0.5 * e0


def test_symmspmv_equals_full_spmv(oracle):
    rp, col = M.stencil7(7)
    val = M.jittered_values(rp, col, 0.1, seed=2)
    rp, col, val = M.make_diag_first(rp, col, val)
    n = len(rp) - 1
    urp, ucol, uval = M.upper_triangle(rp, col, val)
    x = M.uniform_vec(n, 3)
    b = np.zeros(n)
    # any split of the rows into sequential leaves gives the same sum
    calls = [(0, 100), (100, 101), (101, n)]
    oracle.run_calls(oracle.K_SYMMSPMV, urp, ucol, uval, x, b, calls)
    A = M.to_scipy(rp, col, val)
    assert relerr(b, A @ x) < 1e-14


def test_mpk_equals_repeated_spmv(oracle):
    rp, col = M.stencil7(6)
    val = M.jittered_values(rp, col, 0.1, seed=5)
    n = len(rp) - 1
    A = M.to_scipy(rp, col, val)
    x0 = M.uniform_vec(n, 6)
    X = oracle.mpk_serial(rp, col, val, x0, 5)
    y = x0.copy()
    for p in range(1, 6):
        y = A @ y
        assert relerr(X[p], y) < 1e-14


def test_spmtv_equals_transpose_product(oracle):
    # unsymmetric values and pattern, so A x and A^T x differ
    rng = np.random.default_rng(8)
    n = 300
    A = (sp.random(n, n, density=0.03, random_state=rng, format="csr") + sp.eye(n)).tocsr()
    A.sort_indices()
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    x = rng.uniform(-1, 1, 2 * n)
    b0 = rng.uniform(-1, 1, 2 * n)
    assert relerr(A @ x[:n], A.T @ x[:n]) > 1e-2
    b = b0.copy()
    # leaves in any order: every b entry sums the same terms
    oracle.run_calls(oracle.K_SPMTV, rp, col, val, x, b, [(150, n), (0, 150)], nrhs=2)
    for k in range(2):
        ref = b0[k * n:(k + 1) * n] + A.T @ x[k * n:(k + 1) * n]
        assert relerr(b[k * n:(k + 1) * n], ref) < 1e-14


@pytest.mark.parametrize("k", [1, 2, 5])
def test_mpk_highest_is_last_two_powers(oracle, k):
    rp, col = M.stencil7(5)
    val = M.jittered_values(rp, col, 0.1, seed=4)
    x0 = M.uniform_vec(len(rp) - 1, 9)
    X = oracle.mpk_serial(rp, col, val, x0, k)
    H = oracle.mpk_highest_serial(rp, col, val, x0, k)
    assert np.array_equal(H[k % 2], X[k])
    assert np.array_equal(H[(k - 1) % 2], X[k - 1])


def test_multi_rhs_columns_are_independent(oracle):
    rp, col, val = _gs_matrix(4, 1)
    n = len(rp) - 1
    b = M.uniform_vec(n, 1, nrhs=3)
    x = np.zeros(3 * n)
    oracle.run_calls(oracle.K_GS, rp, col, val, x, b, [(0, n)], nrhs=3)
    for k in range(3):
        xk = np.zeros(n)
        oracle.run_calls(oracle.K_GS, rp, col, val, xk, b[k * n:(k + 1) * n].copy(), [(0, n)])
        assert np.array_equal(xk, x[k * n:(k + 1) * n])


# ---------------------------------------------------------------------------
# load balancer, zone tree, schedule
# ---------------------------------------------------------------------------
def _phase_labels(n, phases):
    """per phase: leaf label of every row (-1 when the row is not in the phase)"""
    out = []
    for p in phases:
        lab = np.full(n, -1)
        for k, (a, b) in enumerate(p):
            lo, hi = (a, b) if b >= a else (b + 1, a + 1)
            lab[lo:hi] = k
        out.append(lab)
    return out


@pytest.mark.parametrize("T", [2, 5, 16, 64, 200])
@pytest.mark.parametrize("dist", [1, 2])
def test_schedule_is_a_valid_distance_k_colouring(oracle, T, dist):
    """Leaves run concurrently in one phase share no edge (distance 1) and,
    for distance 2, no common neighbour; every row runs exactly once."""
    rp, col = (M.stencil27(7) if dist == 1 else M.stencil7(9))
    G = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(G, T, dist)
    n = G.n
    # permuted adjacency
    prp, pcol, _ = M.permute(rp, col, None, G.perm, G.inv)
    A = sp.csr_matrix((np.ones(len(pcol)), pcol, prp), shape=(n, n))
    if dist == 2:
        A = ((A @ A) != 0).astype(np.float64)
    A = A.tocoo()
    covered = np.zeros(n, int)
    for rev in (False, True):
        ph = oracle.phases(zt, rev)
        for lab in _phase_labels(n, ph):
            if not rev:
                covered += lab >= 0
            same_phase = (lab[A.row] >= 0) & (lab[A.col] >= 0)
            assert np.all(lab[A.row][same_phase] == lab[A.col][same_phase])
    assert np.all(covered == 1)
    # the serial leaf order visits every row once too
    calls = oracle.leaf_calls(zt)
    assert sum(b - a for a, b in calls) == n


@pytest.mark.parametrize("seed", range(40))
def test_split_zones_invariants(oracle, seed):
    rng = np.random.default_rng(seed)
    L = int(rng.integers(8, 200))
    lr = rng.integers(1, 3000, L)
    ln = lr * rng.integers(2, 28, L)
    T = int(rng.integers(1, max(2, L // 4)))
    dist = int(rng.choice([1, 2]))
    r = oracle.split_zones(lr, ln, T, 40.0, dist)
    lp = r["levelPtr"]
    assert lp[0] == 0 and lp[-1] == L
    assert all(b >= a for a, b in zip(lp, lp[1:]))
    assert sum(r["zone_rows"]) == lr.sum()
    assert len(lp) - 1 == r["bpt"] * r["threads"]
    assert sum(r["scale"]) in (T - 1, T, T + 1) or r["threads"] < T


@pytest.mark.parametrize("T,dist,m", [(2, 1, 4), (3, 1, 3), (4, 2, 2), (5, 2, 3), (8, 1, 2), (6, 1, 5),
                                      (7, 2, 4), (16, 1, 3), (10, 2, 2)])
def test_split_zones_uniform_levels_known_answer(oracle, T, dist, m):
    """A known answer fixed by symmetry: every level carries the same nnz and
    the level count is T x blocks-per-thread x m.  The zone split
    (LB::splitZones, src/lb.cpp:296-391) then closes a zone every T-th of
    the levels, the blocks of a zone split it evenly, every block carries
    the same work (zero variance), and no move of the variance heuristic
    (src/lb.cpp:408-566) can lower that, so the result is the uniform split
    (tests/test_host_logic.py checks the product's load balancer on the same
    cases)."""
    bpt = oracle.block_per_thread(dist)
    L = T * bpt * m
    lr, ln = np.full(L, 7), np.full(L, 21)
    expect = [k * m for k in range(T * bpt + 1)]
    a = oracle.split_zones(lr, ln, T, 40.0, dist)
    assert a["threads"] == T and list(a["levelPtr"]) == expect
    assert a["scale"] == [1] * T


def test_level_recursion_thread_accounting(oracle):
    """nthreads of every inner node = sum over threads of the max child
    nthreads per colour; the root reports the available threads."""
    rp, col = M.stencil27(8)
    for T in (3, 12, 50):
        G = oracle.Graph(rp, col, symm_hint=True)
        zt = oracle.level_recursion(G, T, 1)
        for idx, leaf in enumerate(zt.tree):
            ch = leaf.children
            if not ch:
                continue
            best = 0
            for sb in range(len(ch) // 2):
                nt = (ch[2 * sb + 1] - ch[2 * sb]) // zt.bpt
                best = max(best, sum(max(zt.tree[ch[2 * sb] + zt.bpt * t + b].nthreadsZ
                                         for b in range(zt.bpt)) for t in range(nt)))
            assert leaf.nthreadsZ == best
        assert zt.available_threads == zt.tree[0].nthreadsZ
        assert np.array_equal(np.sort(G.perm), np.arange(G.n))


def test_one_thread_is_the_serial_order(oracle):
    """One RACE thread: the root's two blocks (src/lb.cpp:384-391) run one
    after the other, so the non-empty leaf calls tile [0, n) in row order
    and concatenate to the serial sweep; reversed, they tile it backwards."""
    rp, col = M.stencil27(6)
    G = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(G, 1, 1)
    calls = [(a, b) for a, b in oracle.leaf_calls(zt) if b > a]
    assert calls[0][0] == 0 and calls[-1][1] == G.n
    assert all(b == c for (_, b), (c, _) in zip(calls, calls[1:]))
    rcalls = [(a, b) for a, b in oracle.leaf_calls(zt, rev=True) if a > b]
    assert rcalls[0][0] == G.n - 1 and rcalls[-1][1] == -1
    assert all(b == c for (_, b), (c, _) in zip(rcalls, rcalls[1:]))


def test_mpk_partition_respects_cache_bound(oracle):
    rp, col = M.stencil7(12)
    G = oracle.Graph(rp, col, symm_hint=True)
    k, cache = 4, 96 * 1024.0
    part = oracle.mpk_partition(G, k, cache)
    gl = part["groupLevel"]
    ln = part["levelNnz"]
    elem = [2.0 * (k + 1) * int(v) for v in ln]
    assert gl[0] == 0 and gl[-1] == len(ln)
    for a, b in zip(gl, gl[1:]):
        assert b > a
        # a group is cut before the level that reaches the bound
        if b - a > 1:
            assert sum(elem[a:b - 1]) < part["cache_elem"]
    # consecutive groups: adding the next group's first level reaches the bound
    for a, b, c in zip(gl, gl[1:], gl[2:]):
        assert sum(elem[a:b + 1]) >= part["cache_elem"]
    assert part["groupPtr"][-1] == G.n


@pytest.mark.parametrize("kind,dist,T", [("gs", 1, 24), ("kacz", 2, 24), ("symmspmv", 2, 40)])
def test_parallel_phases_equal_the_serial_leaf_order(oracle, kind, dist, T):
    """or_run_phases (the leaves of a colour phase on OpenMP threads, the
    CPU baseline) gives bit for bit what the serial leaf order gives: leaves
    of one phase are independent, so their interleaving cannot matter."""
    rp, col = M.stencil27(12) if dist == 1 else M.stencil7(16)
    val = M.gs_values(rp, col, seed=3)
    n = len(rp) - 1
    G = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(G, T, dist)
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, G.perm, G.inv))
    okind = {"gs": oracle.K_GS, "kacz": oracle.K_KACZ, "symmspmv": oracle.K_SYMMSPMV}[kind]
    if kind == "symmspmv":
        prp, pcol, pval = M.upper_triangle(prp, pcol, pval)
    b = M.uniform_vec(n, 4, nrhs=2)
    x0 = M.uniform_vec(n, 5, nrhs=2)
    for rev in (False, True):
        xa, ba = x0.copy(), b.copy()
        oracle.run_calls(okind, prp, pcol, pval, xa, ba, oracle.leaf_calls(zt, rev), nrhs=2)
        xb, bb = x0.copy(), b.copy()
        pp, calls = oracle.phase_calls(zt, rev)
        assert len(pp) > 3
        nt = oracle.run_phases(okind, prp, pcol, pval, xb, bb, pp, calls, nrhs=2, nthreads=8)
        assert nt == 8
        assert np.array_equal(xa, xb) and np.array_equal(ba, bb)


def _grid2d(g):
    rp, col = [0], []
    for i in range(g):
        for j in range(g):
            for di, dj in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
                a, b = i + di, j + dj
                if 0 <= a < g and 0 <= b < g:
                    col.append(a * g + b)
            rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32)


@pytest.mark.parametrize("g,lv,radius", [(12, (6, 11), 1), (10, (4, 9), 2), (9, (3, 3), 1)])
def test_traverse_boundary_is_a_bfs_of_the_region_and_its_boundary(oracle, g, lv, radius):
    """The recursion stage's traversal (Traverse with boundary regions,
    src/traverse_GAP_graphSoA.cpp:76-88, 184-197, 449-575), checked against
    an independent computation: after a stage-0 traversal of a 2D grid, the
    rows of levels lv[0]..lv[1] are traversed again with `radius` levels on
    each side as boundary regions.  Levels = unweighted shortest paths from
    the range's first row inside the subgraph of range + boundary rows
    (scipy); every region is stably sorted by level; the histograms count
    the range and each region."""
    rp, col = _grid2d(g)
    G = oracle.Graph(rp, col, symm_hint=True)
    lr, ln = G.traverse(oracle.POWER, 0, G.n)
    lptr = np.concatenate([[0], np.cumsum(lr)])
    lo, hi = int(lptr[lv[0]]), int(lptr[lv[1] + 1])
    bounds = []
    for r in range(-radius, radius + 1):
        L = (lv[0] + r) if r < 0 else (lv[1] + r)
        if r != 0 and 0 <= L < len(lr):
            bounds.append((int(lptr[L]), int(lptr[L + 1])))
    before = G.perm.copy()
    # the permuted pattern before the call, restricted to range + boundaries
    prp, pcol, _ = M.permute(rp, col, None, before, G.inv)
    nodes = np.concatenate([np.arange(lo, hi)] + [np.arange(a, b) for a, b in bounds])
    idx = {int(v): i for i, v in enumerate(nodes)}
    rows, cols = [], []
    for v in nodes:
        for c in pcol[prp[v]:prp[v + 1]]:
            if int(c) in idx:
                rows.append(idx[int(v)])
                cols.append(idx[int(c)])
    A = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(len(nodes), len(nodes)))
    d = csg.shortest_path(A, unweighted=True, indices=idx[lo])
    assert np.all(np.isfinite(d[:hi - lo]))  # the band is connected
    level = {int(v): int(d[i]) if np.isfinite(d[i]) else -1 for i, v in enumerate(nodes)}
    top = max(level.values()) + 1  # totalLevel: one past the deepest level
    unreached = [v for v in level if level[v] < 0]
    for v in unreached:
        level[v] = top
    total = top + (1 if unreached else 0)
    clr, cln, regs = G.traverse_boundary(lo, hi, bounds)
    assert len(clr) == total
    lens = np.diff(rp)
    exp_lr = np.bincount([level[v] for v in range(lo, hi)], minlength=total)
    exp_ln = np.bincount([level[v] for v in range(lo, hi)], weights=[lens[before[v]] for v in range(lo, hi)],
                         minlength=total)
    assert np.array_equal(clr, exp_lr) and np.array_equal(cln, exp_ln.astype(int))
    for (a, b), (blr, bln) in zip(bounds, regs):
        assert np.array_equal(blr, np.bincount([level[v] for v in range(a, b)], minlength=total))
    # every region stably sorted by level, composed into the permutation
    expect = before.copy()
    for a, b in [(lo, hi)] + bounds:
        order = sorted(range(a, b), key=lambda v: level[v])
        expect[a:b] = before[order]
    assert np.array_equal(G.perm, expect)
    assert np.array_equal(G.inv[G.perm], np.arange(G.n))


def test_traverse_boundary_path_by_hand(oracle):
    """A path of 20 rows, range [2, 18) with boundary rows {1} and {18}:
    from row 2, row 1 (boundary) and row 3 are level 1, row r is level r-2,
    row 18 (boundary) is level 16; row 0 is outside both and never reached.
    17 levels (the last holds only row 18), the permutation is unchanged."""
    n = 20
    rp, col = [0], []
    for r in range(n):
        col += [c for c in (r - 1, r, r + 1) if 0 <= c < n]
        rp.append(len(col))
    G = oracle.Graph(np.array(rp, np.int32), np.array(col, np.int32), symm_hint=True)
    clr, cln, regs = G.traverse_boundary(2, 18, [(1, 2), (18, 19)])
    assert list(clr) == [1] * 16 + [0]
    assert list(cln) == [3] * 16 + [0]
    assert list(regs[0][0]) == [0, 1] + [0] * 15
    assert list(regs[1][0]) == [0] * 16 + [1]
    assert np.array_equal(G.perm, np.arange(n))


def _grid_with_hubs(g, hubs):
    """a g x g 5-point grid plus `hubs` rows joined to every grid row"""
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
        for r in range(n0):
            rows[r].append(n0 + h)
            rows[n0 + h].append(r)
    rp = np.concatenate([[0], np.cumsum([len(x) for x in rows])]).astype(np.int32)
    col = np.array([c for x in rows for c in sorted(x)], np.int32)
    return rp, col


@pytest.mark.parametrize("reverse_init", [False, True])
def test_dense_rows_move_to_a_serial_tail(oracle, reverse_init):
    """RACE_MAX_DENSE (Graph::getStatistics / permuteAndRemoveSerialPart,
    src/graph_SoA.cpp:321-435): a 10 x 10 grid with two hub rows joined to
    all 100 grid rows has 862 entries over 102 rows, mean 862 // 102 = 8;
    with RACE_MAX_DENSE = 4 the hubs (101 entries > 32) leave the graph, the
    grid rows keep at most 7 entries.  The remaining graph is the plain grid:
    its one-thread BFS order is (i + j, index); the hubs follow in the order
    they have under the initial permutation."""
    rp, col = _grid_with_hubs(10, 2)
    n = len(rp) - 1
    assert int(rp[-1]) == 862
    ip = np.arange(n, dtype=np.int32)[::-1].copy() if reverse_init else None
    G = oracle.Graph(rp, col, symm_hint=True, init_perm=ip, init_inv_perm=ip, max_dense=4)
    assert G.n == 100 and G.nfull == 102
    assert G.dense_rows == ([101, 100] if reverse_init else [100, 101])
    zt = oracle.level_recursion(G, 1, oracle.ONE)
    grid = np.arange(100)
    if reverse_init:
        grid = grid[::-1]  # graph row q is original grid row 99 - q
    q = np.arange(100)
    i, j = (grid[q] // 10), (grid[q] % 10)
    if reverse_init:
        # BFS from graph row 0 = original (9, 9): level (9 - i) + (9 - j)
        lvl = (9 - i) + (9 - j)
    else:
        lvl = i + j
    expect = grid[np.lexsort((q, lvl))]
    assert np.array_equal(G.full_perm(), np.concatenate([expect, G.dense_rows]))
    calls = oracle.leaf_calls(zt, serial_part=(100, 102))
    assert calls[-1] == (100, 102)
    assert oracle.leaf_calls(zt, rev=True, serial_part=(100, 102))[-1] == (101, 99)
