"""Synthetic sparse matrices and CRS utilities (numpy, vectorised).

The matrices are the synthetic workloads the RACE engine is measured on:
3D 7-point and 27-point stencils on g^3 grids (row index (i*g + j)*g + k,
columns sorted within a row), random banded matrices and Chung-Lu power-law
graph Laplacians.  Matrix values that must be symmetric are drawn from a
counter-based hash of the unordered pair (min(r,c), max(r,c)), so A[r,c] ==
A[c,r] exactly without any communication between rows.

CRS helpers mirror what the reference's example harness does to a matrix
before and after colouring (examples/src/helpers/sparsemat.cpp):
make_diag_first (:432-545), upper_triangle / computeSymmData (:581-631) and
permute (symmetric permutation, :990+).
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def pair_uniform(r, c, n, seed):
    """U[0,1) drawn from the unordered pair {r,c}: symmetric by construction."""
    r = np.asarray(r, np.int64)
    c = np.asarray(c, np.int64)
    lo = np.minimum(r, c).astype(np.uint64)
    hi = np.maximum(r, c).astype(np.uint64)
    with np.errstate(over="ignore"):
        key = lo * np.uint64(n) + hi
        h = _splitmix64(key ^ _splitmix64(np.uint64(seed)))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _stencil(g, offsets):
    n = g * g * g
    idx = np.arange(n, dtype=np.int64)
    i, j, k = idx // (g * g), (idx // g) % g, idx % g
    cols = []
    valid = []
    for (di, dj, dk) in offsets:
        ii, jj, kk = i + di, j + dj, k + dk
        ok = (ii >= 0) & (ii < g) & (jj >= 0) & (jj < g) & (kk >= 0) & (kk < g)
        cols.append((ii * g + jj) * g + kk)
        valid.append(ok)
    cols = np.stack(cols, axis=1)
    valid = np.stack(valid, axis=1)
    rowlen = valid.sum(axis=1)
    rowPtr = np.zeros(n + 1, np.int64)
    np.cumsum(rowlen, out=rowPtr[1:])
    col = cols[valid]  # row-major, offsets sorted => columns ascending
    return rowPtr.astype(np.int32), col.astype(np.int32)


def stencil7(g):
    offs = sorted([(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0)])
    return _stencil(g, offs)


def stencil27(g):
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    return _stencil(g, offs)


def row_of(rowPtr):
    n = len(rowPtr) - 1
    return np.repeat(np.arange(n, dtype=np.int64), np.diff(rowPtr))


def laplace7_values(rowPtr, col, diag=6.0, off=-1.0):
    r = row_of(rowPtr)
    return np.where(col == r, diag, off).astype(np.float64)


def jittered_values(rowPtr, col, jitter, seed, diag=6.0, off=-1.0):
    """diag/off values times (1 + U[-jitter, jitter]) drawn symmetrically."""
    n = len(rowPtr) - 1
    r = row_of(rowPtr)
    u = pair_uniform(r, col, n, seed)
    base = np.where(col == r, diag, off)
    return (base * (1.0 + jitter * (2.0 * u - 1.0))).astype(np.float64)


def gs_values(rowPtr, col, seed, dominance=1.05):
    """Off-diagonals U[-1,0) (symmetric), diagonal = dominance * sum|off|."""
    n = len(rowPtr) - 1
    r = row_of(rowPtr)
    u = pair_uniform(r, col, n, seed)
    v = -1.0 + u  # in [-1, 0)
    isd = col == r
    v = np.where(isd, 0.0, v)
    rs = np.zeros(n, np.float64)
    np.add.at(rs, r, np.abs(v))
    v = np.where(isd, dominance * rs[r], v)
    return v.astype(np.float64)


def uniform_vec(n, seed, lo=-1.0, hi=1.0, nrhs=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=n * nrhs).astype(np.float64)


# --------------------------------------------------------------------------
# CRS transformations (mirroring the reference's example harness)
# --------------------------------------------------------------------------
def make_diag_first(rowPtr, col, val, missing_diag=0.0):
    """Diagonal entry first in every row, the others keep their order; a
    missing diagonal is inserted with value `missing_diag`
    (sparsemat::makeDiagFirst examples/src/helpers/sparsemat.cpp:432-545)."""
    n = len(rowPtr) - 1
    r = row_of(rowPtr)
    has = np.zeros(n, bool)
    has[r[col == r]] = True
    if not has.all():
        miss = np.nonzero(~has)[0]
        rl = np.diff(rowPtr).astype(np.int64)
        rl[miss] += 1
        rp2 = np.zeros(n + 1, np.int64)
        np.cumsum(rl, out=rp2[1:])
        c2 = np.empty(rp2[-1], np.int32)
        v2 = np.empty(rp2[-1], np.float64)
        pos = rp2[:-1][r] + (np.arange(len(col)) - rowPtr[:-1][r])
        c2[pos] = col
        v2[pos] = val
        c2[rp2[1:][miss] - 1] = miss
        v2[rp2[1:][miss] - 1] = missing_diag
        rowPtr, col, val, r = rp2.astype(np.int32), c2, v2, row_of(rp2)
    isd = col == r
    # sort key within a row: diagonal first (0), others keep order (1 + pos)
    pos = np.arange(len(col), dtype=np.int64) - rowPtr[:-1][r]
    key = np.where(isd, 0, pos + 1)
    order = np.lexsort((key, r))
    return rowPtr.astype(np.int32), col[order].astype(np.int32), val[order].astype(np.float64)


def upper_triangle(rowPtr, col, val):
    """Entries with col >= row, order kept (computeSymmData :581-631)."""
    r = row_of(rowPtr)
    keep = col >= r
    n = len(rowPtr) - 1
    cnt = np.bincount(r[keep], minlength=n)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=rp[1:])
    return rp.astype(np.int32), col[keep].astype(np.int32), val[keep].astype(np.float64)


def permute(rowPtr, col, val, perm, inv_perm):
    """A' = P A P^T: new row r is old row perm[r]; columns via inv_perm; the
    order inside a row is kept (sparsemat::permute :990+)."""
    perm = np.asarray(perm, np.int64)
    inv_perm = np.asarray(inv_perm, np.int64)
    n = len(rowPtr) - 1
    rl = np.diff(rowPtr).astype(np.int64)[perm]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(rl, out=rp[1:])
    nnz = int(rp[-1])
    starts = rowPtr[:-1].astype(np.int64)[perm]
    src = np.repeat(starts - rp[:-1], rl) + np.arange(nnz, dtype=np.int64)
    newcol = inv_perm[col[src]].astype(np.int32)
    newval = None if val is None else val[src]
    return rp.astype(np.int32), newcol, newval


def permute_vec(v, perm, nrhs=1):
    n = len(perm)
    v = np.asarray(v).reshape(nrhs, n)
    return np.ascontiguousarray(v[:, perm].reshape(-1))


def to_scipy(rowPtr, col, val):
    import scipy.sparse as sp
    n = len(rowPtr) - 1
    return sp.csr_matrix((val, col, rowPtr), shape=(n, n))


def to_sell(rowPtr, col, val, C, pad_col=0):
    """SELL-C arrays (sigma = 1) of a CRS matrix, the layout the SELL-C-sigma
    helpers take: entry j of row r at chunkStart[r // C] + r % C + j * C.
    Returns chunkStart, rowLen, chunkLen (widest row of each chunk), col and
    val of length chunkStart[-1] + chunkLen[-1] * C; padding slots hold
    column pad_col and value 0."""
    rowPtr = np.asarray(rowPtr, np.int64)
    n = len(rowPtr) - 1
    rl = np.diff(rowPtr).astype(np.int32)
    nch = (n + C - 1) // C
    clp = np.zeros(nch, np.int32)
    for c in range(nch):
        clp[c] = rl[c * C:(c + 1) * C].max() if n else 0
    cs = np.zeros(nch + 1, np.int32)
    cs[1:] = np.cumsum(clp.astype(np.int64) * C)
    total = int(cs[-1])
    scol = np.full(total, pad_col, np.int32)
    sval = np.zeros(total, np.asarray(val).dtype)
    for r in range(n):
        base = cs[r // C] + r % C
        for j in range(rl[r]):
            scol[base + j * C] = col[rowPtr[r] + j]
            sval[base + j * C] = val[rowPtr[r] + j]
    return cs, rl, clp, scol, sval


def _symmetric_csr(n, rows, cols):
    """Sorted, duplicate-free symmetric pattern with the diagonal, from an
    edge list (both directions are added)."""
    r = np.concatenate([rows, cols, np.arange(n)]).astype(np.int64)
    c = np.concatenate([cols, rows, np.arange(n)]).astype(np.int64)
    key = np.unique(r * n + c)
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32)


def banded_random(n, mean_nnz, seed, band_frac=0.01):
    """Random banded symmetric matrix: about mean_nnz entries per row (half of
    them drawn per row, the rest from the transpose) at column offsets
    uniform in [-bw, bw], bw = ceil(band_frac * n); off-diagonals U[-1,0)
    (symmetric), diagonal 1 + sum |off| (strictly diagonally dominant)."""
    rng = np.random.default_rng(seed)
    bw = max(1, int(np.ceil(band_frac * n)))
    k = max(1, int(round((mean_nnz - 1) / 2)))
    rows = np.repeat(np.arange(n), k)
    off = rng.integers(1, bw + 1, size=rows.size) * rng.choice([-1, 1], size=rows.size)
    cols = rows + off
    keep = (cols >= 0) & (cols < n)
    rp, col = _symmetric_csr(n, rows[keep], cols[keep])
    return rp, col, _dominant_values(rp, col, seed)


def chung_lu_laplacian(n, mean_deg, seed, gamma=2.3, max_deg=None):
    """Chung-Lu random graph with power-law expected degrees (exponent
    gamma, mean mean_deg; expected degree capped at max_deg, default
    min(sqrt(sum w), 400)), returned as the pattern of L + I with L the graph
    Laplacian: off-diagonals -1, diagonal degree + 1."""
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.float64)
    w = (i + 1.0) ** (-1.0 / (gamma - 1.0))
    w *= mean_deg * n / w.sum()
    cap = max_deg if max_deg is not None else min(np.sqrt(w.sum()), 400.0)
    w = np.minimum(w, cap)
    w *= mean_deg * n / w.sum()
    w = np.minimum(w, cap)
    m = int(w.sum() / 2)
    p = w / w.sum()
    a = rng.choice(n, size=m, p=p)
    b = rng.choice(n, size=m, p=p)
    keep = a != b
    perm = rng.permutation(n)  # hubs spread over the numbering
    rp, col = _symmetric_csr(n, perm[a[keep]], perm[b[keep]])
    r = row_of(rp)
    deg = np.diff(rp) - 1
    val = np.where(col == r, deg[r] + 1.0, -1.0)
    return rp, col, val.astype(np.float64)


def _dominant_values(rp, col, seed):
    n = len(rp) - 1
    r = row_of(rp)
    u = pair_uniform(r, col, n, seed)
    off = np.where(col == r, 0.0, -u)
    s = np.zeros(n)
    np.add.at(s, r, np.abs(off))
    return np.where(col == r, 1.0 + s[r], off)


def batch_sizes(count, seed, n_lo=2000, n_hi=200000):
    """(kind, n, mean) of batch_matrices(count, seed) without building them
    (the same random draws)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        mean = float(rng.uniform(5, 40))
        rng.integers(1 << 30)
        out.append(("banded" if k % 2 == 0 else "chung-lu", n, mean))
    return out


def batch_matrices(count, seed, n_lo=2000, n_hi=200000):
    """The batched workload: count matrices, n log-uniform in [n_lo, n_hi],
    even indices random banded (band 1 % of n), odd indices Chung-Lu
    Laplacians (exponent 2.3); mean entries per row uniform in [5, 40].
    Yields (kind, n, mean, rowPtr, col, val)."""
    rng = np.random.default_rng(seed)
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        mean = float(rng.uniform(5, 40))
        s = int(rng.integers(1 << 30))
        if k % 2 == 0:
            rp, col, val = banded_random(n, mean, s)
            yield "banded", n, mean, rp, col, val
        else:
            rp, col, val = chung_lu_laplacian(n, mean, s)
            yield "chung-lu", n, mean, rp, col, val
