"""Layout tuner: the oracle MLP pinned by finite differences, tuner features
pinned by closed forms on stencils (host only)."""
import numpy as np

from race_amd import matgen as M
from race_amd import tuner as TU
from oracle import race_oracle as O

N_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 12 + 12


def _problem(seed, B=24):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, N_PARAMS)
    X = rng.normal(0, 1, (B, 16))
    T = rng.normal(0, 1, (B, 12))
    mask = (rng.uniform(size=(B, 12)) > 0.2).astype(np.float64)
    return p, X, T, mask


def test_oracle_mlp_gradient_matches_finite_differences():
    p, X, T, mask = _problem(1)
    loss, g = O.mlp_loss_grad(p, X, T, mask)
    rng = np.random.default_rng(2)
    for k in rng.choice(N_PARAMS, 60, replace=False):
        e = np.zeros(N_PARAMS)
        e[k] = 1e-6
        fd = (O.mlp_loss_grad(p + e, X, T, mask)[0] - O.mlp_loss_grad(p - e, X, T, mask)[0]) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k


def test_oracle_mlp_forward_closed_form():
    # identity-like weights: W1 = [I 0], W2 = I, W3 = first 12 columns -> y = relu(x)[:12]
    p = np.zeros(N_PARAMS)
    W1 = np.zeros((16, 64)); W1[:, :16] = np.eye(16)
    W2 = np.eye(64)
    W3 = np.zeros((64, 12)); W3[:12, :12] = np.eye(12)
    p[:1024] = W1.reshape(-1)
    p[1024 + 64:1024 + 64 + 4096] = W2.reshape(-1)
    o = 1024 + 64 + 4096 + 64
    p[o:o + 768] = W3.reshape(-1)
    X = np.random.default_rng(3).normal(0, 1, (5, 16))
    assert np.allclose(O.mlp_forward(p, X), np.maximum(X, 0)[:, :12])


def _bfs_levels_reference(rowPtr, col):
    """Level sizes of a breadth-first search from row 0 (islands restarted at
    their smallest row), as numpy frontier expansion."""
    n = len(rowPtr) - 1
    rp = np.asarray(rowPtr, np.int64)
    c = np.asarray(col, np.int64)
    level = np.full(n, -1, np.int64)
    sizes = []
    nxt = 0
    while True:
        rest = np.nonzero(level < 0)[0]
        if rest.size == 0:
            break
        front = rest[:1]
        level[front] = nxt
        while front.size:
            sizes.append(front.size)
            starts, lens = rp[front], rp[front + 1] - rp[front]
            excl = np.cumsum(lens) - lens
            idx = np.arange(lens.sum()) + np.repeat(starts - excl, lens)
            nb = np.unique(c[idx])
            nb = nb[level[nb] < 0]
            nxt += 1
            level[nb] = nxt
            front = nb
    return np.array(sizes, np.int64)


def _features_reference(rowPtr, col):
    """The 16 features written out in numpy (the library computes them in C++)."""
    rp = np.asarray(rowPtr, np.int64)
    c = np.asarray(col, np.int64)
    n = len(rp) - 1
    nnz = int(rp[-1])
    rl = np.diff(rp).astype(np.float64)
    mean = rl.mean()
    r = np.repeat(np.arange(n), np.diff(rp))
    dist = np.abs(c - r).astype(np.float64)
    lv = _bfs_levels_reference(rp, c)
    srt = np.sort(rl)
    gini = 1.0 - 2.0 * np.sum(np.cumsum(srt)) / (n * srt.sum()) + 1.0 / n
    return np.array([
        np.log10(n), np.log10(max(nnz, 1)), mean, rl.std() / mean, np.log10(rl.max()), rl.min(),
        np.mean(rl > 2 * mean), np.mean(rl > 32), dist.mean() / n, dist.max() / n, np.mean(dist <= 64),
        np.log10(len(lv)), lv.max() / n, lv.mean() / lv.max(), gini, np.log10(nnz / len(lv)),
    ], np.float64)


def test_library_features_equal_numpy_reference():
    """race_matrix_features against the numpy restatement on the tuner's own
    matrix families (banded, Chung-Lu with hubs, stencils, islands)."""
    mats = [(rp, col) for _, rp, col, _ in TU.tuner_matrices(24, 5, n_lo=300, n_hi=4000)]
    rp1, c1 = M.stencil27(4)
    n1 = len(rp1) - 1
    mats.append((np.concatenate([rp1, rp1[1:] + rp1[-1]]).astype(np.int32),
                 np.concatenate([c1, c1 + n1]).astype(np.int32)))
    for rp, col in mats:
        assert np.allclose(TU.features(rp, col), _features_reference(rp, col), rtol=1e-10, atol=1e-12)


def test_features_closed_form_7pt():
    g = 10
    rp, col = M.stencil7(g)
    f = TU.features(rp, col)
    n = g ** 3
    assert np.isclose(f[0], np.log10(n))
    assert np.isclose(f[2], 7 - 6 / g)          # mean row length
    assert f[4] == np.log10(7) and f[5] == 4    # longest / shortest row
    assert np.isclose(f[9], g * g / n)          # widest column offset g^2
    assert np.isclose(f[11], np.log10(3 * g - 2))  # BFS levels from row 0: i+j+k
    # the widest level of i+j+k = L over [0,g)^3
    cnt = np.zeros(3 * g - 2)
    for i in range(g):
        for j in range(g):
            for k in range(g):
                cnt[i + j + k] += 1
    assert np.isclose(f[12], cnt.max() / n)


def test_features_levels_with_islands():
    # two disconnected 27-point blocks: levels restart on the second block
    rp1, c1 = M.stencil27(4)
    n1 = len(rp1) - 1
    rp = np.concatenate([rp1, rp1[1:] + rp1[-1]]).astype(np.int32)
    col = np.concatenate([c1, c1 + n1]).astype(np.int32)
    lv = _bfs_levels_reference(rp, col)
    assert lv.tolist() == [1, 7, 19, 37] * 2  # 3L^2 + 3L + 1
    f = TU.features(rp, col)
    assert np.isclose(f[11], np.log10(8)) and np.isclose(f[12], 37 / (2 * n1))


def test_targets():
    y, m = TU.targets([[2.0, 1.0, -1.0, 4.0]])
    assert m.tolist() == [[1, 1, 0, 1]]
    assert np.allclose(y, [[np.log(2), 0, 0, np.log(4)]])
    assert TU.score([[2.0, 1.0, -1.0]], [0]) == 0.5
