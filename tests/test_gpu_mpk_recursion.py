"""MPK with hopeless regions: levels that exceed the cache by more than the
cut-off (src/matrixPower.cpp:505-604), merged into hopeless regions
(identifyHopelessRegions :147-502), each region one level group with k-1
single-level groups around it (consolidatePartition :685-880), and every
hopeless group traversed again with its boundary groups
(mtxPowerRecursive::recursivePartition src/matrixPowerRecursive.cpp:156-338).
Product (GPU traversals) against the oracle: permutation, groups and the
recursion-stage leaves bit for bit; the powers against repeated SpMV."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

from .conftest import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12
INT_MAX = 2 ** 31 - 1


@pytest.mark.parametrize("g,k,cache,env,cut0,cut1", [
    (16, 4, 4096.0, None, 50, INT_MAX),
    (16, 4, 2048.0, None, 50, INT_MAX),
    # RACE_CACHE_VIOLATION_CUTOFF=1: stage 0 cut-off 1, the recursion stage 50
    (20, 3, 8192.0, "1", 1, 50),
    (24, 6, 30000.0, "2", 2, 50),
])
def test_recursive_mpk_matches_oracle(oracle, monkeypatch, g, k, cache, env, cut0, cut1):
    if env is not None:
        monkeypatch.setenv("RACE_CACHE_VIOLATION_CUTOFF", env)
    print("case recursion", g, k, cache, env, flush=True)
    rp, col = M.stencil7(g)
    val = M.jittered_values(rp, col, 0.1, seed=g) / 6.0
    n = len(rp) - 1
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, cache)
    G = oracle.Graph(rp, col, symm_hint=True)
    part = oracle.mpk_partition(G, k, cache, cutoff=cut0, cutoff1=cut1)
    assert part["hopelessGroups"], "the case must have a hopeless region"
    assert np.array_equal(ce.getPerm(), G.perm)
    assert np.array_equal(ce.mpkGroups(), part["groupPtr"])
    kids = [(c["lo"], c["hi"], c["group"], len(c["groupPtr"]) - 1) for c in part["children"]]
    assert ce.mpkChildren() == kids
    assert ce.checkPowerCoverage()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    x0 = M.uniform_vec(n, 3)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL, p


def test_schedule_round_trip_keeps_the_recursive_permutation(oracle):
    """A saved MPK schedule with a hopeless group loads into a fresh handle:
    same permutation (the recursion's), same groups, same powers."""
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=5) / 6.0
    n, k = len(rp) - 1, 4
    a = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    a.RACEColor(k, 4096.0)
    assert a.mpkChildren()
    data = a.saveSchedule()
    b = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    b.loadSchedule(data)
    assert np.array_equal(a.getPerm(), b.getPerm())
    assert np.array_equal(a.mpkGroups(), b.mpkGroups())
    prp, pcol, pval = M.permute(rp, col, val, b.getPerm(), b.getInvPerm())
    xs = np.zeros((k + 1) * n)
    xs[:n] = M.uniform_vec(n, 8)
    dx = R.DeviceArray.from_host(xs)
    fid = b.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    b.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    ref = oracle.mpk_serial(prp, pcol, pval, xs[:n], k)
    for p in range(1, k + 1):
        assert relerr(got[p], ref[p]) <= TOL
