"""Schedule save / load through the C ABI: a handle that loads a saved
schedule instead of running RACEColor has the same permutation, zone tree,
level groups and thread count, and its device kernels give bit-identical
results (same schedule, same layout, same summation order)."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dist,T", [(R.ONE, 64), (R.TWO, 16)])
def test_colour_schedule_round_trip(dist, T):
    print("case colour", dist, T, flush=True)
    rp, col = M.stencil27(12) if dist == R.ONE else M.stencil7(14)
    val = M.gs_values(rp, col, seed=3)
    n = len(rp) - 1
    a = R.Interface(n, T, dist, rp, col, symm_hint=True)
    a.RACEColor()
    blob = a.saveSchedule()
    assert R.schedule_info(blob) == {"nrow": n, "dist": dist, "nthreads": a.getNumThreads()}
    b = R.Interface(n, T, dist, rp, col, symm_hint=True)
    b.loadSchedule(blob)
    assert np.array_equal(a.getPerm(), b.getPerm())
    assert np.array_equal(a.getInvPerm(), b.getInvPerm())
    assert np.array_equal(a.tree(), b.tree())
    for rev in (False, True):
        for u, v in zip(a.schedule(rev), b.schedule(rev)):
            assert np.array_equal(u, v)
    assert a.getNumThreads() == b.getNumThreads()
    assert a.getMaxStageDepth() == b.getMaxStageDepth()
    assert a.getEfficiency() == b.getEfficiency()
    assert b.saveSchedule() == blob
    perm, inv = a.getPerm(), a.getInvPerm()
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, perm, inv))
    rhs = M.uniform_vec(n, 4)
    out = []
    for h in (a, b):
        x, bb = R.DeviceArray.from_host(np.zeros(n)), R.DeviceArray.from_host(rhs)
        kind = R.K_GS if dist == R.ONE else R.K_KACZ
        fid = h.registerDevice(kind, prp, pcol, pval, x, bb)
        h.executeFunction(fid)
        h.executeFunction(fid, rev=True)
        out.append(x.download())
    assert np.array_equal(out[0], out[1])
    assert np.abs(out[0]).max() > 0


def test_power_schedule_round_trip():
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=2)
    n, k = len(rp) - 1, 3
    a = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    a.RACEColor(k, 64 * 1024.0)
    blob = a.saveSchedule()
    b = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    b.loadSchedule(blob)
    assert np.array_equal(a.mpkGroups(), b.mpkGroups()) and len(a.mpkGroups()) > 3
    assert np.array_equal(a.getPerm(), b.getPerm())
    assert b.getHighestPower() == k
    assert b.checkPowerCoverage()
    prp, pcol, pval = M.permute(rp, col, val, a.getPerm(), a.getInvPerm())
    xs = np.zeros((k + 1) * n)
    xs[:n] = M.uniform_vec(n, 1)
    out = []
    for h in (a, b):
        dx = R.DeviceArray.from_host(xs)
        h.executeFunction(h.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k))
        out.append(dx.download())
    assert np.array_equal(out[0], out[1])


def test_schedule_load_rejects_mismatches():
    rp, col = M.stencil7(8)
    n = len(rp) - 1
    a = R.Interface(n, 8, R.ONE, rp, col, symm_hint=True)
    with pytest.raises(R.RaceError):
        a.saveSchedule()  # not coloured yet
    a.RACEColor()
    blob = a.saveSchedule()
    for T, dist in [(4, R.ONE), (8, R.TWO)]:
        with pytest.raises(R.RaceError):
            R.Interface(n, T, dist, rp, col, symm_hint=True).loadSchedule(blob)
    # same size, other pattern
    rp2, col2 = M.stencil27(8)
    with pytest.raises(R.RaceError):
        R.Interface(n, 8, R.ONE, rp2, col2, symm_hint=True).loadSchedule(blob)
    # damaged bytes
    bad = bytearray(blob)
    bad[len(bad) // 2] ^= 1
    with pytest.raises(R.RaceError):
        R.Interface(n, 8, R.ONE, rp, col, symm_hint=True).loadSchedule(bytes(bad))
    # functions already registered against the handle's schedule
    c = R.Interface(n, 8, R.ONE, rp, col, symm_hint=True)
    c.loadSchedule(blob)
    val = M.laplace7_values(rp, col)
    c.registerDevice(R.K_SPMV, rp, col, val, R.DeviceArray(n), R.DeviceArray(n))
    with pytest.raises(R.RaceError):
        c.loadSchedule(blob)
