"""The "graph" option: a function's launch sequence captured once as a HIP
graph and replayed.  Every replay must equal the directly launched sequence
bit for bit (same kernels, same order), across replays separated by other
work, after the vectors change (a new capture), and for another power."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M

pytestmark = pytest.mark.gpu


def _pair(rp, col, T, dist):
    n = len(rp) - 1
    a = R.Interface(n, T, dist, rp, col, symm_hint=True)
    a.RACEColor()
    b = R.Interface(n, T, dist, rp, col, symm_hint=True)
    b.loadSchedule(a.saveSchedule())
    b.setOption("graph", 1)
    return a, b


@pytest.mark.parametrize("kind,dist,nrhs", [(R.K_GS, R.ONE, 1), (R.K_KACZ, R.TWO, 3)])
def test_graph_replays_equal_direct_launches(kind, dist, nrhs):
    print("case graph", kind, nrhs, flush=True)
    rp, col = M.stencil27(14)
    val = M.gs_values(rp, col, seed=5)
    n = len(rp) - 1
    a, g = _pair(rp, col, 64, dist)
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, a.getPerm(), a.getInvPerm()))
    rhs = M.uniform_vec(n, 6, nrhs=nrhs)
    xa, ba = R.DeviceArray.from_host(np.zeros(nrhs * n)), R.DeviceArray.from_host(rhs)
    xg, bg = R.DeviceArray.from_host(np.zeros(nrhs * n)), R.DeviceArray.from_host(rhs)
    fa = a.registerDevice(kind, prp, pcol, pval, xa, ba, nrhs=nrhs)
    fg = g.registerDevice(kind, prp, pcol, pval, xg, bg, nrhs=nrhs)
    for it in range(3):
        for rev in (False, True):
            a.executeFunction(fa, rev=rev)
            g.executeFunction(fg, rev=rev)
        ref = xa.download()
        assert np.array_equal(xg.download(), ref), it
        assert np.abs(ref).max() > 0
    # new vectors: the next call captures a new graph
    xg2, bg2 = R.DeviceArray.from_host(xg.download()), R.DeviceArray.from_host(rhs)
    g.setVectors(fg, xg2, bg2)
    g.executeFunction(fg)
    a.executeFunction(fa)
    assert np.array_equal(xg2.download(), xa.download())


def test_graph_mpk_follows_set_power(oracle):
    rp, col = M.stencil7(16)
    val = M.jittered_values(rp, col, 0.1, seed=4)
    n, k = len(rp) - 1, 4
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, 64 * 1024.0)
    ce.setOption("graph", 1)
    prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
    x0 = M.uniform_vec(n, 3)
    xs = np.zeros((k + 1) * n)
    xs[:n] = x0
    dx = R.DeviceArray.from_host(xs)
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    ref = oracle.mpk_serial(prp, pcol, pval, x0, k)
    ce.setPower(fid, 2)
    ce.executeFunction(fid)
    got = dx.download().reshape(k + 1, n)
    assert not got[3].any() and not got[4].any()
    ce.setPower(fid, k)
    for _ in range(2):
        ce.executeFunction(fid)
        got = dx.download().reshape(k + 1, n)
        for p in range(1, k + 1):
            assert np.linalg.norm(got[p] - ref[p]) <= 1e-12 * np.linalg.norm(ref[p]), p


@pytest.mark.parametrize("what", ["symm_phases", "gs_leaf_workgroups"])
def test_graph_other_phase_launchers(what):
    """The colour-phase SymmSpMV (symm_mode 1) and the workgroup-per-leaf GS
    layout (layout 1) replayed as graphs equal their direct launches."""
    rp, col = M.stencil7(14)
    val = M.gs_values(rp, col, seed=2)
    n = len(rp) - 1
    dist = R.TWO if what == "symm_phases" else R.ONE
    a, g = _pair(rp, col, 32, dist)
    for h in (a, g):
        h.setOption("symm_mode" if what == "symm_phases" else "layout", 1)
    prp, pcol, pval = M.make_diag_first(*M.permute(rp, col, val, a.getPerm(), a.getInvPerm()))
    if what == "symm_phases":
        prp, pcol, pval = M.upper_triangle(prp, pcol, pval)
        kind, x0, b0 = R.K_SYMMSPMV, M.uniform_vec(n, 1), np.zeros(n)
    else:
        kind, x0, b0 = R.K_GS, np.zeros(n), M.uniform_vec(n, 1)
    xa, ba = R.DeviceArray.from_host(x0), R.DeviceArray.from_host(b0)
    xg, bg = R.DeviceArray.from_host(x0), R.DeviceArray.from_host(b0)
    fa = a.registerDevice(kind, prp, pcol, pval, xa, ba)
    fg = g.registerDevice(kind, prp, pcol, pval, xg, bg)
    assert g.launchCount(fg) > 1
    for it in range(3):
        a.executeFunction(fa)
        g.executeFunction(fg)
        out_a = (ba if what == "symm_phases" else xa).download()
        out_g = (bg if what == "symm_phases" else xg).download()
        assert np.array_equal(out_g, out_a), it
