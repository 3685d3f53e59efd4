"""Edge cases on the device: malformed input rejected on the host before any
kernel runs, a 1 x 1 matrix, more RACE threads than rows, batch misuse."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu


def _small():
    rp, col = M.stencil7(4)
    return rp, col, M.gs_values(rp, col, seed=3)


def test_malformed_matrix_is_rejected_before_any_kernel():
    rp, col, val = _small()
    n = len(rp) - 1
    ce = R.Interface(n, 4, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    dx, db = R.DeviceArray(n), R.DeviceArray(n)
    bad = col.copy()
    bad[5] = n  # column outside the matrix
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_GS, rp, bad, val, dx, db)
    bad = col.copy()
    bad[7] = -1
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_KACZ, rp, bad, val, dx, db)
    brp = rp.copy()
    brp[3], brp[4] = brp[4], brp[3]  # rowPtr not non-decreasing
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_GS, brp, col, val, dx, db)
    # Gauss-Seidel needs the diagonal first in every row (it divides by it)
    swp = col.copy()
    swp[rp[4]], swp[rp[4] + 1] = swp[rp[4] + 1], swp[rp[4]]
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_GS, rp, swp, val, dx, db)
    # an empty row has no diagonal (a GS row would read past its entries)
    erp = rp.copy()
    erp[-1] = erp[-2]
    with pytest.raises(R.RaceError):
        ce.registerDevice(R.K_GS, erp, col, val, dx, db)
    # the handle still works afterwards
    fid = ce.registerDevice(R.K_SPMV, rp, col, val, dx, db)
    ce.executeFunction(fid)


@pytest.mark.parametrize("dist,kind", [(R.ONE, R.K_GS), (R.TWO, R.K_KACZ), (R.TWO, R.K_SYMMSPMV)])
def test_one_by_one_matrix(dist, kind):
    rp, col, val = np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0])
    ce = R.Interface(1, 8, dist, rp, col, symm_hint=True)
    ce.RACEColor()
    assert ce.getPerm().tolist() == [0]
    dx, db = R.DeviceArray.from_host(np.array([3.0])), R.DeviceArray.from_host(np.array([2.0]))
    fid = ce.registerDevice(kind, rp, col, val, dx, db)
    ce.executeFunction(fid)
    if kind == R.K_SYMMSPMV:
        assert db.download().tolist() == [2.0 + 4.0 * 3.0]
    else:  # GS: x = b / a; KACZ: x -= (a x - b) a / a^2 -> b / a
        assert np.allclose(dx.download(), [0.5])


def test_more_threads_than_rows(oracle):
    rp, col, val = _small()
    n = len(rp) - 1
    T = 4 * n
    ce = R.Interface(n, T, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    g = oracle.Graph(rp, col, symm_hint=True)
    zt = oracle.level_recursion(g, T, oracle.ONE)
    assert np.array_equal(ce.getPerm(), g.perm)
    assert ce.getNumThreads() == zt.available_threads <= n
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    b = M.uniform_vec(n, 4)
    dx, db = R.DeviceArray.from_host(np.zeros(n)), R.DeviceArray.from_host(b)
    fid = ce.registerDevice(R.K_GS, prp, pcol, pval, dx, db)
    ce.executeFunction(fid)
    xo = np.zeros(n)
    oracle.run_calls(oracle.K_GS, prp, pcol, pval, xo, b, oracle.leaf_calls(zt))
    assert relerr(dx.download(), xo) <= 1e-12


def test_batch_misuse_is_rejected():
    rp, col, val = _small()
    n = len(rp) - 1
    b = R.Batch(R.K_KACZ, R.FP64, R.FP64, 1)
    with pytest.raises(R.RaceError):
        b.build()  # nothing added
    ce = R.Interface(n, 4, R.TWO, rp, col, symm_hint=True)
    with pytest.raises(R.RaceError):  # not coloured yet
        b.add(ce, rp, col, val, R.DeviceArray(n), R.DeviceArray(n))
    ce.RACEColor()
    with pytest.raises(R.RaceError):
        b.setOption("bundle_tpr", 3)
    with pytest.raises(R.RaceError):
        b.useCandidate(12)
