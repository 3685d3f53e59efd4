"""Reverse Cuthill-McKee initial permutation (race_rcm, host only).

The reference's harness takes it from SpMP (absent from the reference
tree), so the exact tie-breaking is parity unpinned.  What the algorithm
fixes is checked instead: a bijection with its inverse; Cuthill-McKee is a
BFS, so in the un-reversed order every row but a component's first has an
earlier neighbour, components are contiguous and BFS levels from the
component's first row never decrease; the first row of a component is at
least as eccentric as the component's smallest row (George-Liu); closed
form on a relabelled path (bandwidth 1); on grids the bandwidth is within
a small factor of scipy's RCM; unsymmetric patterns are refused."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import race_amd as R
from race_amd import matgen as M

from .conftest import GOLDEN
import os


def _csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


def _relabel(rp, col, seed):
    n = len(rp) - 1
    p = np.random.default_rng(seed).permutation(n)
    A = sp.csr_matrix((np.ones(len(col)), col, rp), shape=(n, n))
    return _csr(A[p][:, p])


def _bandwidth(rp, col, perm, inv):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    return int(np.abs(inv[rows] - inv[col]).max()) if len(col) else 0


def _grid2d(nx, ny):
    I = sp.identity
    T = lambda k: sp.diags([1, 1, 1], [-1, 0, 1], shape=(k, k))  # noqa: E731
    return _csr(sp.kron(I(ny), T(nx)) + sp.kron(T(ny), I(nx)))


def _check_cm(rp, col, perm, inv):
    n = len(rp) - 1
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert np.array_equal(inv[perm], np.arange(n))
    A = sp.csr_matrix((np.ones(len(col)), col, rp), shape=(n, n))
    cm = perm[::-1]                         # Cuthill-McKee order
    pos = np.empty(n, int)
    pos[cm] = np.arange(n)
    ncomp, label = csg.connected_components(A, directed=False)
    first = {}
    for i, v in enumerate(cm):
        first.setdefault(label[v], i)
    # components contiguous
    starts = sorted(first.values())
    for c, s in first.items():
        assert np.all(label[cm[s:s + np.sum(label == c)]] == c)
    for v in range(n):
        if pos[v] == first[label[v]]:
            continue
        nb = A.indices[A.indptr[v]:A.indptr[v + 1]]
        assert (pos[nb] < pos[v]).any(), v
    # BFS levels from each component's first row never decrease along the order
    for c, s in first.items():
        root = cm[s]
        d = csg.shortest_path(A, unweighted=True, indices=root, directed=False)
        members = cm[s:s + np.sum(label == c)]
        assert np.all(np.diff(d[members]) >= 0)
        lo = np.min(np.where(label == c)[0])
        dlo = csg.shortest_path(A, unweighted=True, indices=lo, directed=False)
        assert d[members].max() >= dlo[members].max()
    return starts


def test_rcm_relabelled_path_has_bandwidth_one():
    n = 200
    rp, col = _csr(sp.diags([1, 1, 1], [-1, 0, 1], shape=(n, n)))
    rp, col = _relabel(rp, col, 3)
    perm, inv = R.rcm(rp, col)
    _check_cm(rp, col, perm, inv)
    assert _bandwidth(rp, col, perm, inv) == 1


@pytest.mark.parametrize("nx,ny,seed", [(30, 40, 1), (17, 5, 2), (64, 64, 3)])
def test_rcm_grid_bandwidth(nx, ny, seed):
    rp, col = _relabel(*_grid2d(nx, ny), seed)
    perm, inv = R.rcm(rp, col)
    _check_cm(rp, col, perm, inv)
    bw = _bandwidth(rp, col, perm, inv)
    A = sp.csr_matrix((np.ones(len(col)), col, rp), shape=(len(rp) - 1,) * 2)
    sperm = csg.reverse_cuthill_mckee(A, symmetric_mode=True)
    sinv = np.empty_like(sperm)
    sinv[sperm] = np.arange(len(sperm))
    assert bw <= min(nx, ny) + 2
    assert bw <= 1.25 * _bandwidth(rp, col, sperm, sinv) + 2


def test_rcm_components_and_isolated_rows():
    rp1, c1 = _grid2d(6, 7)
    rp2, c2 = _grid2d(5, 3)
    n1, n2 = len(rp1) - 1, len(rp2) - 1
    A = sp.block_diag([sp.csr_matrix((np.ones(len(c1)), c1, rp1)), sp.identity(3),
                       sp.csr_matrix((np.ones(len(c2)), c2, rp2))])
    rp, col = _relabel(*_csr(A), 5)
    perm, inv = R.rcm(rp, col)
    starts = _check_cm(rp, col, perm, inv)
    assert len(starts) == 5


def test_rcm_stencil_and_reference_fixture(spin):
    rp, col = M.stencil7(9)
    rp, col = _relabel(rp, col, 7)
    perm, inv = R.rcm(rp, col)
    _check_cm(rp, col, perm, inv)
    assert _bandwidth(rp, col, perm, inv) <= 2 * 81
    srp, scol, _ = spin
    perm, inv = R.rcm(srp, scol)
    _check_cm(srp, scol, perm, inv)
    ident = np.arange(len(srp) - 1)
    assert _bandwidth(srp, scol, perm, inv) <= _bandwidth(srp, scol, ident, ident)


def test_rcm_refuses_unsymmetric_pattern():
    rp = np.array([0, 2, 3, 4], np.int32)
    col = np.array([0, 1, 1, 2], np.int32)   # (0,1) without (1,0)
    with pytest.raises(R.RaceError, match="MATRIX_SYMM"):
        R.rcm(rp, col)
    with pytest.raises(R.RaceError):
        R.rcm(np.array([0, 1], np.int32), np.array([5], np.int32))
