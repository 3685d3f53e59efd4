#!/usr/bin/env python3
"""RACE Gauss-Seidel benchmark on one MI355X.

Workload: symmetric Gauss-Seidel (forward + backward sweep) under a RACE
distance-1 colouring on the 3D 27-point stencil, 128^3 grid (2,097,152 rows,
55,742,968 stored entries), off-diagonals U[-1,0), diagonal 1.05 x row |sum|,
b ~ U[-1,1].  One step = one forward + one backward sweep over the whole
matrix, with the matrix and vectors already resident in HBM.

Prints one JSON line: GB/s of algorithmic traffic (per sweep: nnz*(8+4) +
n*(4 + 3*8) bytes for FP64), the roofline of the Gauss-Seidel kernel against
the 8 TB/s HBM peak, the CPU oracle (the same sweeps in RACE order, single
thread) timed on this host beside it, the time to a relative residual of
1e-10 of FP64 iterative refinement around FP64, FP32 and BF16 inner sweeps,
and ("other_kernels") SymmSpMV on the 7-pt 64^3 grid (hot and cold caches),
the level-blocked matrix powers, k = 8, on the 7-pt 256^3 grid (all powers
kept, and the only-highest variant with two vector slots), and SpMTV on the
27-pt 128^3 grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SymmSpMV/GS/MPK GB/s vs HBM peak; mixed vs FP64 GPU at stated err; CPU RACE"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """CPU model and core count of this host (lscpu, else /proc/cpuinfo)."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return info


def gs_bytes(n, nnz, sv=8, sx=8, si=4):
    # values + column indices, row pointer, x read + x write + b read
    return nnz * (sv + si) + n * (si + 3 * sx)


def bench_symmspmv(R, M, g=64, threads=64, reps=200, O=None):
    """SymmSpMV b += A x, 7-pt g^3 (diag 6, off -1), upper triangle with the
    diagonal first, RACE distance-2 colouring.  Hot: back-to-back calls
    (working set in the caches); cold: each call after a 1 GiB copy.  With
    the oracle O: the FP64 result against the oracle's SymmSpMV in its own
    RACE order, and the FP32 / BF16 errors against that."""
    rp, col = M.stencil7(g)
    val = M.laplace7_values(rp, col)
    n = len(rp) - 1
    ce = R.Interface(n, threads, R.TWO, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    urp, ucol, uval = M.upper_triangle(prp, pcol, pval)
    nnzu = int(urp[-1])
    byts = nnzu * 12 + n * (4 + 3 * 8) + 4
    dx = R.DeviceArray.from_host(M.uniform_vec(n, 1234))
    db = R.DeviceArray.from_host(np.zeros(n))
    fid = ce.registerDevice(R.K_SYMMSPMV, urp, ucol, uval, dx, db)
    for _ in range(5):
        ce.executeAsync(fid)
    ce.timerBegin()
    for _ in range(reps):
        ce.executeAsync(fid)
    hot = ce.timerEnd() / reps
    fa, fb = R.DeviceArray(1 << 27), R.DeviceArray(1 << 27)
    cold = []
    st = ce.stream()
    for _ in range(20):
        # the 1 GiB copy on the handle's stream right before the call: the
        # timer starts when the copy ends, with the call already queued
        R.interface.L.check(R.interface.L.load().race_device_copy_async(fb.ptr, fa.ptr, fa.nbytes, st), "copy")
        ce.timerBegin()
        ce.executeAsync(fid)
        cold.append(ce.timerEnd())
    cold = float(np.median(cold))
    db.upload(np.zeros(n))
    ce.executeFunction(fid)
    ref = db.download()
    parity = None
    if O is not None:
        G = O.Graph(rp, col, symm_hint=True)
        zt = O.level_recursion(G, threads, O.TWO)
        bo = np.zeros(n)
        O.run_calls(O.K_SYMMSPMV, urp, ucol, uval, M.uniform_vec(n, 1234), bo, O.leaf_calls(zt))
        perr = float(np.linalg.norm(ref - bo) / np.linalg.norm(bo))
        peq = bool(np.array_equal(G.perm, perm))
        parity = {"rel_err_vs_oracle": perr, "perm_equal_oracle": peq, "tol": 1e-12,
                  "ok": bool(peq and perr <= 1e-12)}
        log(f"parity: SymmSpMV {g}^3 FP64 rel err vs oracle {perr:.2e}, perm equal {peq}")
        ref = bo
    # FP32 / BF16 values with FP32 vectors and accumulation: same calls,
    # bytes with the smaller values and vectors, error against the FP64 result
    low = {}
    for name, vp, sv in (("fp32", R.FP32, 4), ("bf16", R.BF16, 2)):
        lx = R.DeviceArray.from_host(dx.download().astype(np.float32))
        lb = R.DeviceArray.from_host(np.zeros(n, np.float32))
        lid = ce.registerDevice(R.K_SYMMSPMV, urp, ucol, uval, lx, lb, val_precision=vp, vec_precision=R.FP32)
        for _ in range(5):
            ce.executeAsync(lid)
        lb.upload(np.zeros(n, np.float32))
        ce.executeFunction(lid)
        err = float(np.linalg.norm(lb.download() - ref) / np.linalg.norm(ref))
        ce.timerBegin()
        for _ in range(reps):
            ce.executeAsync(lid)
        lh = ce.timerEnd() / reps
        lc = []
        for _ in range(20):
            R.interface.L.check(R.interface.L.load().race_device_copy_async(fb.ptr, fa.ptr, fa.nbytes, st), "copy")
            ce.timerBegin()
            ce.executeAsync(lid)
            lc.append(ce.timerEnd())
        lc = float(np.median(lc))
        lbyts = nnzu * (sv + 4) + n * (4 + 3 * 4) + 4
        low[name] = {"bytes_per_call": lbyts, "hot_us": round(lh * 1e3, 2), "hot_gbs": round(lbyts / lh / 1e6, 1),
                     "cold_us": round(lc * 1e3, 2), "cold_gbs": round(lbyts / lc / 1e6, 1),
                     "speedup_hot_vs_fp64": round(hot / lh, 3), "speedup_cold_vs_fp64": round(cold / lc, 3),
                     ("rel_err_vs_oracle" if parity else "rel_err_vs_fp64"): err}
        log(f"SymmSpMV {g}^3 {name}: hot {low[name]['hot_us']} us, cold {low[name]['cold_us']} us, "
            f"err {err:.1e}")
        del lx, lb
    out = {"workload": f"SymmSpMV 7-pt {g}^3 (n={n}, nnz_upper={nnzu}), distance-2 colouring, "
                       f"{threads} RACE threads, FP64",
           "bytes_per_call": byts, "hot_us": round(hot * 1e3, 2), "hot_gbs": round(byts / hot / 1e6, 1),
           "cold_us": round(cold * 1e3, 2), "cold_gbs": round(byts / cold / 1e6, 1),
           "cold_method": "1 GiB device copy enqueued right before each timed call",
           "cold_frac_hbm": round(byts / cold / 1e6 / HBM_PEAK_GBS, 4),
           "low_precision": low, "parity": parity}
    log(f"SymmSpMV {g}^3: hot {out['hot_us']} us ({out['hot_gbs']} GB/s), cold {out['cold_us']} us "
        f"({out['cold_gbs']} GB/s)")
    return out


def bench_mpk(R, M, g=256, k=8, cache_mb=256.0, reps=10, O=None):
    """Matrix powers x_q = A x_{q-1}, q = 1..k, 7-pt g^3 (diag 6, off -1,
    jitter 10 %), RACE level groups for a cache of cache_mb MB (the 256 MB
    Infinity Cache), safety 2.  With the oracle O: the permutation and level
    groups against the oracle's, every FP64 power against the oracle's
    repeated SpMV, and the FP32 / BF16 errors against that."""
    rp, col = M.stencil7(g)
    val = M.jittered_values(rp, col, 0.1, seed=9)
    n = len(rp) - 1
    nnz = int(rp[-1])
    ce = R.Interface(n, 1, R.POWER, rp, col, symm_hint=True)
    ce.RACEColor(k, cache_mb * 1e6)
    groups = len(ce.mpkGroups()) - 1
    perm, inv = ce.getPerm(), ce.getInvPerm()
    part_equal = None
    if O is not None:
        G = O.Graph(rp, col, symm_hint=True)
        part = O.mpk_partition(G, k, cache_mb * 1e6)
        part_equal = bool(np.array_equal(G.perm, perm) and np.array_equal(ce.mpkGroups(), part["groupPtr"]))
        del G, part
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    del rp, col, val
    xs = np.zeros((k + 1) * n)
    xs[:n] = M.uniform_vec(n, 2)
    dx = R.DeviceArray.from_host(xs)
    del xs
    fid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dx, power=k)
    for _ in range(2):
        ce.executeAsync(fid)
    ce.timerBegin()
    for _ in range(reps):
        ce.executeAsync(fid)
    ms = ce.timerEnd() / reps
    eff = k * (nnz * 12 + n * (4 + 2 * 8))
    floor = nnz * 12 + n * 4 + (k + 1) * n * 8
    # only-highest variant: two vector slots instead of k + 1
    x2 = np.zeros(2 * n)
    x2[:n] = dx.download()[:n]
    dx2 = R.DeviceArray.from_host(x2)
    del x2
    hid = ce.registerDevice(R.K_MPK_HIGHEST, prp, pcol, pval, dx2, power=k)
    ce.executeAsync(hid)
    ce.timerBegin()
    for _ in range(reps):
        ce.executeAsync(hid)
    msh = ce.timerEnd() / reps
    del dx2
    # FP32 / BF16 values, FP32 powers and accumulation: time and the error
    # of every power against the FP64 path (itself within 1e-12 of the oracle)
    ref = dx.download().reshape(k + 1, n)
    parity = None
    if O is not None:
        oref = O.mpk_serial(prp, pcol, pval, ref[0], k)
        perrs = [float(np.linalg.norm(ref[p] - oref[p]) / np.linalg.norm(oref[p])) for p in range(1, k + 1)]
        parity = {"rel_err_per_power_vs_oracle": perrs, "groups_and_perm_equal_oracle": part_equal,
                  "tol": 1e-12, "ok": bool(part_equal and max(perrs) <= 1e-12)}
        log(f"parity: MPK {g}^3 k={k} FP64 worst power rel err vs oracle {max(perrs):.2e}, "
            f"groups/perm equal {part_equal}")
        ref = oref
        del oref
    low = {}
    for name, vp, sv in (("fp32", R.FP32, 4), ("bf16", R.BF16, 2)):
        xl = np.zeros((k + 1) * n, np.float32)
        xl[:n] = ref[0]
        dl = R.DeviceArray.from_host(xl)
        del xl
        lid = ce.registerDevice(R.K_MPK, prp, pcol, pval, dl, power=k, val_precision=vp, vec_precision=R.FP32)
        ce.executeFunction(lid)
        got = dl.download().reshape(k + 1, n)
        errs = [float(np.linalg.norm(got[p] - ref[p]) / np.linalg.norm(ref[p])) for p in range(1, k + 1)]
        del got
        ce.timerBegin()
        for _ in range(reps):
            ce.executeAsync(lid)
        ml = ce.timerEnd() / reps
        leff = k * (nnz * (sv + 4) + n * (4 + 2 * 4))
        low[name] = {"ms_per_call": round(ml, 4), "effective_gbs": round(leff / ml / 1e6, 1),
                     "speedup_vs_fp64": round(ms / ml, 3),
                     ("rel_err_per_power_vs_oracle" if parity else "rel_err_per_power_vs_fp64"): errs}
        log(f"MPK {g}^3 k={k} {name}: {ml:.3f} ms/call ({low[name]['effective_gbs']} GB/s), "
            f"error power 1 / {k}: {errs[0]:.1e} / {errs[-1]:.1e}")
        del dl
    # the same k products as k separate SpMVs, for comparison
    dy = R.DeviceArray(n)
    sid = ce.registerDevice(R.K_SPMV, prp, pcol, pval, dx, dy)
    ce.executeAsync(sid)
    ce.timerBegin()
    for _ in range(reps):
        ce.executeAsync(sid)
    ms1 = ce.timerEnd() / reps
    out = {"workload": f"MPK 7-pt {g}^3 (n={n}, nnz={nnz}), k={k}, cache {cache_mb:.0f} MB, "
                       f"{groups} level groups, FP64",
           "ms_per_call": round(ms, 4), "effective_gbs": round(eff / ms / 1e6, 1),
           "effective_frac_hbm": round(eff / ms / 1e6 / HBM_PEAK_GBS, 4),
           "compulsory_gbs": round(floor / ms / 1e6, 1),
           "only_highest_ms": round(msh, 4), "only_highest_effective_gbs": round(eff / msh / 1e6, 1),
           "spmv_ms": round(ms1, 4),
           "spmv_gbs": round((nnz * 12 + n * 20) / ms1 / 1e6, 1), "speedup_vs_k_spmv": round(k * ms1 / ms, 3),
           "low_precision": low, "parity": parity}
    log(f"MPK {g}^3 k={k}: {ms:.3f} ms/call, effective {out['effective_gbs']} GB/s; "
        f"only-highest {msh:.3f} ms; SpMV {ms1:.3f} ms ({out['spmv_gbs']} GB/s)")
    return out


def bench_spmtv(R, M, g=128, reps=20):
    """SpMTV b += A^T x, 27-pt g^3 (GS values, unsymmetric after a row
    scaling), no colouring: one gather launch over the transpose."""
    rp, col = M.stencil27(g)
    val = M.gs_values(rp, col, seed=7)
    n = len(rp) - 1
    nnz = int(rp[-1])
    val = val * np.repeat(np.linspace(0.5, 1.5, n), np.diff(rp))
    ce = R.Interface(n, 1, R.ONE, rp, col, symm_hint=True)
    dx = R.DeviceArray.from_host(M.uniform_vec(n, 5))
    db = R.DeviceArray.from_host(np.zeros(n))
    fid = ce.registerDevice(R.K_SPMTV, rp, col, val, dx, db)
    del rp, col, val
    for _ in range(3):
        ce.executeAsync(fid)
    ce.timerBegin()
    for _ in range(reps):
        ce.executeAsync(fid)
    ms = ce.timerEnd() / reps
    byts = nnz * 12 + n * (4 + 3 * 8)
    out = {"workload": f"SpMTV 27-pt {g}^3 (n={n}, nnz={nnz}), FP64", "bytes_per_call": byts,
           "ms_per_call": round(ms, 4), "gbs": round(byts / ms / 1e6, 1),
           "frac_hbm": round(byts / ms / 1e6 / HBM_PEAK_GBS, 4)}
    log(f"SpMTV 27-pt {g}^3: {ms:.3f} ms/call, {out['gbs']} GB/s")
    return out


def bench_batched(args, R, M):
    """Batched workload: args.batch_count matrices (n log-uniform in
    [2000, 200000], half random banded, half Chung-Lu Laplacians), each
    coloured by RACE (distance 2); one step = SymmSpMV + args.batch_sweeps
    Kaczmarz sweeps (alternating forward / backward) on all of them, FP32
    values and vectors, FP64 row norms, 8 right-hand sides.  Layouts are
    tuned on the 64 matrices with the most entries (a fused launch lasts as
    long as its largest matrices' chains) and applied to the whole batch;
    fused against per-matrix launches is compared on the first 64."""
    nrhs, sweeps, count = 8, args.batch_sweeps, args.batch_count
    tune_n = min(64, count)
    sizes = [(k, n * mean) for k, (kd, n, mean) in enumerate(M.batch_sizes(count, args.batch_seed))
             if args.batch_kinds in ("both", kd)]
    tune_set = {k for k, _ in sorted(sizes, key=lambda t: -t[1])[:tune_n]}
    kacz = R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs)
    symm = R.Batch(R.K_SYMMSPMV, R.FP32, R.FP32, nrhs)
    tk = R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs)
    ts = R.Batch(R.K_SYMMSPMV, R.FP32, R.FP32, nrhs)
    fk = R.Batch(R.K_KACZ, R.FP32, R.FP32, nrhs)   # the first tune_n matrices
    bytes_symm = bytes_kacz = 0
    tot_n = tot_nnz = 0
    kinds = {}
    singles = []
    cpu_sample = []
    kacz_x = []
    mlp_feats = []
    if args.batch_layouts == "mlp":
        from race_amd import tuner as TU
    t0 = time.time()
    gen = ((kd, n, mean, rp, col, val) for kd, n, mean, rp, col, val in M.batch_matrices(count, args.batch_seed)
           if args.batch_kinds in ("both", kd))
    kept = [k for k, _ in sizes]
    for k, (kind, n, mean, rp, col, val) in enumerate(gen):
        T = max(4, min(1024, n // args.batch_rows_per_thread))
        ce = R.Interface(n, T, R.TWO, rp, col, symm_hint=True)
        ce.RACEColor()
        perm, inv = ce.getPerm(), ce.getInvPerm()
        prp, pcol, pval = M.permute(rp, col, val, perm, inv)
        prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
        urp, ucol, uval = M.upper_triangle(prp, pcol, pval)
        nnz, nnzu = int(prp[-1]), int(urp[-1])
        bvec = M.uniform_vec(n, 1000 + k, nrhs=nrhs).astype(np.float32)
        xvec = M.uniform_vec(n, 5000 + k, nrhs=nrhs).astype(np.float32)
        dx, db = R.DeviceArray.from_host(np.zeros(nrhs * n, np.float32)), R.DeviceArray.from_host(bvec)
        sx, sb = R.DeviceArray.from_host(xvec), R.DeviceArray.from_host(np.zeros(nrhs * n, np.float32))
        kacz.add(ce, prp, pcol, pval, dx, db)
        kacz_x.append(dx)
        if args.batch_layouts == "mlp":
            mlp_feats.append(TU.features(rp, col))
        symm.add(ce, urp, ucol, uval, sx, sb)
        if kept[k] in tune_set:
            tk.add(ce, prp, pcol, pval, dx, db)
            ts.add(ce, urp, ucol, uval, sx, sb)
        if k < tune_n:
            fk.add(ce, prp, pcol, pval, dx, db)
            singles.append((ce, prp, pcol, pval, n))
        if k < 8:
            cpu_sample.append((ce, prp, pcol, pval, n, bvec.astype(np.float64), dx))
        bytes_symm += nnzu * 8 + n * (4 + nrhs * 3 * 4)
        bytes_kacz += nnz * 8 + n * (4 + 8 + nrhs * 3 * 4)
        tot_n += n
        tot_nnz += nnz
        kinds[kind] = kinds.get(kind, 0) + 1
        if (k + 1) % 32 == 0:
            log(f"batched: {k + 1} matrices prepared ({time.time() - t0:.0f}s, {tot_nnz / 1e6:.0f}M nnz)")
    count = k + 1
    tune_n = min(tune_n, count)
    t_prep = time.time() - t0
    t0 = time.time()
    for bt in (kacz, tk, fk):
        bt.setOption("flow", args.batch_flow)
    if args.batch_layouts == "mlp":
        # One layout serves the whole batch.  The tuner ranks the candidates
        # by their mean predicted log-time over the matrices; its models were
        # trained on single matrices, and in a batch the matrices also run
        # concurrently, so the three best ranked Kaczmarz layouts are timed
        # on the tuning sample (instead of all twelve)
        F = np.array(mlp_feats)
        rank_k = np.argsort(TU.Tuner.pretrained("kacz").scores(F).mean(axis=0))
        bs = int(np.argmin(TU.Tuner.pretrained("symmspmv").scores(F).mean(axis=0)))
        times_k = []
        for c in rank_k[:3]:
            tk.useCandidate(int(c))
            tk.build()
            times_k.append(tk.time(False, reps=2))
            log(f"tuner-ranked KACZ candidate {int(c)}: {times_k[-1]:.2f} ms")
        bk = int(rank_k[int(np.argmin(times_k))])
        times_s = []
    elif args.batch_layouts:
        bk, bs = (int(v) for v in args.batch_layouts.split(","))
        times_k, times_s = [], []
    else:
        def prog(name):
            return lambda i, ms: log(f"tuning {name}: candidate {i} {ms:.2f} ms")
        bk, times_k = tk.autotune(False, reps=2, progress=prog("KACZ"))
        bs, times_s = ts.autotune(False, reps=3, progress=prog("SymmSpMV"))
    log(f"tuned on {tune_n} matrices in {time.time() - t0:.1f}s: KACZ layout {bk} "
        f"{R.Batch.candidate(R.K_KACZ, bk)}, SymmSpMV layout {bs} {R.Batch.candidate(R.K_SYMMSPMV, bs)}")
    # batched against per-matrix launches on the first tune_n matrices
    # (forward sweep, tuned layout)
    fk.useCandidate(bk)
    t_batch_sample = fk.time(False, reps=3)
    t_single = 0.0
    for ce, prp, pcol, pval, n in singles:
        wx = R.DeviceArray(nrhs * n, np.float32)
        wb = R.DeviceArray.from_host(np.ones(nrhs * n, np.float32))
        fid = ce.registerDevice(R.K_KACZ, prp, pcol, pval, wx, wb, nrhs=nrhs, val_precision=R.FP32,
                                vec_precision=R.FP32)
        ce.executeFunction(fid)
        ce.timerBegin()
        for _ in range(3):
            ce.executeAsync(fid)
        t_single += ce.timerEnd() / 3
    del tk, ts, fk
    t0 = time.time()
    kacz.useCandidate(bk)
    symm.useCandidate(bs)
    kacz.build()
    symm.build()
    t_build = time.time() - t0
    t_s = symm.time(False, reps=5)
    t_f = kacz.time(False, reps=3)
    t_b = kacz.time(True, reps=3)
    launches = (kacz.launchCount(False), kacz.launchCount(True))
    aborts = kacz.flowAborts() if args.batch_flow else 0
    ms = t_s + sweeps / 2 * (t_f + t_b)
    total = bytes_symm + sweeps * bytes_kacz
    gbs = total / (ms * 1e-3) / 1e9
    log(f"batched step: SymmSpMV {t_s:.3f} ms + {sweeps} x KACZ ({t_f:.3f} / {t_b:.3f} ms) = {ms:.2f} ms, "
        f"{gbs:.1f} GB/s; per-matrix launches on the sample: {t_single:.2f} ms vs batched {t_batch_sample:.2f} ms")
    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import race_oracle as O
        O.build()
        cb = ct = 0.0
        # parity: one forward sweep of the batch from x = 0 (FP32 values and
        # vectors, FP64 row norms) against the oracle's FP64 sweep in RACE
        # order on the sample matrices
        for xd in kacz_x:
            xd.zero()
        kacz.execute(False)
        errs = []
        for ce, prp, pcol, pval, n, bv, dxs in cpu_sample:
            pp, lo, hi = ce.schedule(False)
            calls = np.stack([lo, hi], axis=1).reshape(-1)
            xo = np.zeros(nrhs * n)
            t1 = time.time()
            O.run_calls(O.K_KACZ, prp, pcol, pval, xo, bv, calls, nrhs=nrhs)
            ct += time.time() - t1
            xg = dxs.download().astype(np.float64)
            errs.append(float(np.linalg.norm(xg - xo) / np.linalg.norm(xo)))
            cb += int(prp[-1]) * 8 + n * (4 + 8 + nrhs * 3 * 4)
        parity = {"kacz_fp32_rel_err_vs_oracle_fp64": max(errs), "matrices": len(errs), "tol": 1e-5,
                  "ok": bool(max(errs) <= 1e-5)}
        log(f"parity: one forward KACZ sweep, FP32 batch vs FP64 oracle, worst rel err {max(errs):.2e} "
            f"on {len(errs)} matrices")
        cpu = {"value": round(cb / ct / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"one forward KACZ sweep, 8 right-hand sides, on the first {len(cpu_sample)} matrices "
                         f"of the batch (oracle C restatement in FP64, 1 thread, {ct * 1e3:.0f} ms)"}
    return {
        "metric": "batched SymmSpMV + 20 KACZ sweeps GB/s (FP32, 8 RHS)",
        "value": round(gbs, 3), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{count} matrices (seed {args.batch_seed}; {kinds}; RACE threads n/"
                               f"{args.batch_rows_per_thread}), n log-uniform in "
                               f"[2000, 200000], total n={tot_n}, nnz={tot_nnz}; SymmSpMV + {sweeps} KACZ sweeps, "
                               f"distance-2 colouring, FP32 values/vectors, FP64 row norms, {nrhs} RHS",
                   "matrices": count, "rows": tot_n, "nnz": tot_nnz},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None},
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": {"symmspmv_ms": round(t_s, 4), "kacz_fwd_ms": round(t_f, 4), "kacz_bwd_ms": round(t_b, 4),
                    "tuned_kacz_layout": list(R.Batch.candidate(R.K_KACZ, bk)),
                    "tuned_symmspmv_layout": list(R.Batch.candidate(R.K_SYMMSPMV, bs)),
                    "tuner_kacz_ms": [round(t, 4) for t in times_k],
                    "tuner_symmspmv_ms": [round(t, 4) for t in times_s],
                    "flow": args.batch_flow, "kacz_launches_fwd_bwd": list(launches), "flow_aborts": aborts,
                    "sample_matrices": tune_n, "sample_per_matrix_launches_ms": round(t_single, 4),
                    "sample_batched_ms": round(t_batch_sample, 4),
                    "batching_speedup": round(t_single / t_batch_sample, 3),
                    "prepare_s": round(t_prep, 1), "build_s": round(t_build, 1)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--threads", type=int, default=int(os.environ.get("RACE_BENCH_THREADS", "131072")))
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--inner-sweeps", type=int, default=4)
    ap.add_argument("--workload", choices=["gs", "batched", "symmspmv", "mpk"], default="gs",
                    help="gs: the Gauss-Seidel line (default); batched: many matrices at once; "
                         "symmspmv: only the SymmSpMV 7-pt 64^3 measurement (every precision); "
                         "mpk: only the MPK 7-pt 256^3 k=8 measurement")
    ap.add_argument("--mpk-cache-mb", type=float, default=256.0,
                    help="mpk workload: the cache size RACEColor blocks the level groups for")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--batch-sweeps", type=int, default=20)
    ap.add_argument("--batch-seed", type=int, default=2026)
    ap.add_argument("--batch-rows-per-thread", type=int, default=256,
                    help="RACE threads requested per matrix: n / this (4 .. 1024)")
    ap.add_argument("--batch-kinds", choices=["both", "banded", "chung-lu"], default="both",
                    help="keep only one matrix family of the batch")
    ap.add_argument("--batch-flow", type=int, choices=[0, 1], default=1,
                    help="Kaczmarz batch: one launch per sweep and shape group, every matrix through its own "
                         "phases (1), or one fused launch per phase index (0)")
    ap.add_argument("--batch-layouts", default="",
                    help="KACZ,SymmSpMV candidate layout indices, or mlp (the tuner's SymmSpMV layout and the "
                         "fastest of its three best ranked KACZ layouts); either skips the full autotune")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched workload (512 matrices, tuner layouts) in other_kernels")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the SymmSpMV 64^3 and MPK 256^3 measurements")
    ap.add_argument("--graph", type=int, choices=[0, 1], default=0,
                    help="replay each sweep's phase launches as a HIP graph")
    ap.add_argument("--xcd-order", type=int, choices=[0, 1], default=0,
                    help="GS bundles: consecutive bundles of a phase on the same XCD")
    ap.add_argument("--bundle-rows", type=int, default=0, help="GS bundles: rows per bundle (0 = default)")
    ap.add_argument("--bundle-threads", type=int, default=0, help="GS bundles: workgroup size (0 = default)")
    ap.add_argument("--flow", type=int, choices=[0, 1], default=0,
                    help="one launch per sweep with per-chain dependency flags instead of one per colour phase")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run a plain copy kernel of this many MiB first (byte-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (single-GPU build)",
              file=sys.stderr)
        sys.exit(2)

    import race_amd as R
    from race_amd import matgen as M

    if args.workload == "batched":
        out = bench_batched(args, R, M)
        print(json.dumps(out), flush=True)
        if out.get("parity") and not out["parity"]["ok"]:
            log("PARITY FAILED for the batched Kaczmarz sweep")
            sys.exit(3)
        return
    if args.workload == "mpk":
        r = bench_mpk(R, M, cache_mb=args.mpk_cache_mb, reps=max(1, args.steps))
        print(json.dumps({"metric": "MPK 7-pt 256^3 k=8 effective GB/s (FP64)", "value": r["effective_gbs"],
                          "unit": "GB/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": r["workload"]}, "mpk": r}), flush=True)
        return
    if args.workload == "symmspmv":
        r = bench_symmspmv(R, M)
        print(json.dumps({"metric": "SymmSpMV 7-pt 64^3 GB/s (cold, FP64)", "value": r["cold_gbs"], "unit": "GB/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": r["workload"]}, "symmspmv": r}), flush=True)
        return

    g = args.grid
    t0 = time.time()
    rp, col = M.stencil27(g)
    val = M.gs_values(rp, col, seed=2024)
    n = len(rp) - 1
    nnz = int(rp[-1])
    b = M.uniform_vec(n, 7)
    log(f"matrix 27-pt {g}^3: n={n} nnz={nnz} ({time.time() - t0:.1f}s)")
    # what the learned tuner proposes for RACE's thread count on this matrix
    # (its training matrices have at most ~100k rows: an extrapolation here)
    race_tuner = None
    try:
        from race_amd import tuner as TU
        t0 = time.time()
        feats = TU.features(rp, col)
        rpt = TU.RACE_ROWS_PER_THREAD[int(TU.Tuner.pretrained("race").predict(feats)[0])]
        eps = TU.EPS_VALUES[int(TU.Tuner.pretrained("eps").predict(feats)[0])]
        race_tuner = {"proposed_rows_per_thread": rpt, "proposed_threads": max(1, n // rpt),
                      "proposed_efficiency": eps, "seconds": round(time.time() - t0, 2),
                      "threads_used": args.threads, "efficiency_used": 40}
        log(f"tuner proposes {rpt} rows per RACE thread ({n // rpt} threads), efficiency {eps}; "
            f"measured best 16 (131072) at efficiency 40")
    except OSError:
        pass

    t0 = time.time()
    ce = R.Interface(n, args.threads, R.ONE, rp, col, symm_hint=True)
    ce.RACEColor()
    t_color = time.time() - t0
    ce.setOption("graph", args.graph)
    ce.setOption("flow", args.flow)
    ce.setOption("xcd_order", args.xcd_order)
    if args.bundle_rows:
        ce.setOption("bundle_rows", args.bundle_rows)
    if args.bundle_threads:
        ce.setOption("bundle_threads", args.bundle_threads)
    bfs_ms, ntrav = ce.preprocessStats()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    log(f"RACEColor T={args.threads}: {t_color:.2f}s (GPU BFS {bfs_ms:.0f} ms in {ntrav} traversals), "
        f"threads={ce.getNumThreads()} eff={ce.getEfficiency():.3f} stages={ce.getMaxStageDepth()}")
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    pb = M.permute_vec(b, perm)
    del val  # the pattern stays for the oracle's own colouring

    dx = R.DeviceArray.from_host(np.zeros(n))
    db = R.DeviceArray.from_host(pb)
    t0 = time.time()
    fid = ce.registerDevice(R.K_GS, prp, pcol, pval, dx, db)
    log(f"register GS: {time.time() - t0:.2f}s, waves fwd={ce.waveCount(fid)} bwd={ce.waveCount(fid, True)}")
    pp, _, _ = ce.schedule()

    if args.copy_mib:
        nb = args.copy_mib << 20
        src = R.DeviceArray(nb // 8)
        dst = R.DeviceArray(nb // 8)
        R.interface.L.check(R.interface.L.load().race_device_copy(dst.ptr, src.ptr, nb), "copy")
        log(f"reference copy kernel: {args.copy_mib} MiB read + {args.copy_mib} MiB written")
        del src, dst

    for _ in range(args.warmup):
        ce.executeAsync(fid, False)
        ce.executeAsync(fid, True)
    R.interface.L.load().race_device_synchronize()
    ce.timerBegin()
    for _ in range(args.steps):
        ce.executeAsync(fid, False)
        ce.executeAsync(fid, True)
    ms = ce.timerEnd()
    R.interface.L.load().race_device_synchronize()
    x_gpu = dx.download()  # after warmup + steps fwd+bwd steps from x = 0
    ms_step = ms / args.steps
    bytes_step = 2 * gs_bytes(n, nnz)
    gbs = bytes_step / (ms_step * 1e-3) / 1e9
    launches_per_step = ce.launchCount(fid) + ce.launchCount(fid, True)
    log(f"GPU: {ms_step:.3f} ms/step, {gbs:.1f} GB/s, {launches_per_step} launches/step")

    # mixed precision vs FP64 to the same residual: FP64 iterative refinement
    # around inner symmetric GS sweeps in FP64 / FP32 / BF16 storage
    mixed = None
    if not args.no_mixed:
        A = M.to_scipy(prp, pcol, pval)
        xs = M.uniform_vec(n, 99)
        bs = A @ xs
        mixed = {"tol": args.tol, "inner_sweeps": args.inner_sweeps}
        for name, vp, xp, dt in (("fp64", R.FP64, R.FP64, np.float64), ("fp32", R.FP32, R.FP32, np.float32),
                                 ("bf16", R.BF16, R.FP32, np.float32)):
            wx, wb = R.DeviceArray(n, dt), R.DeviceArray(n, dt)
            inner = ce.registerDevice(R.K_GS, prp, pcol, pval, wx, wb, val_precision=vp, vec_precision=xp)
            xd = R.DeviceArray.from_host(np.zeros(n))
            bd = R.DeviceArray.from_host(bs)
            rid = ce.registerRefinement(R.K_GS, prp, pcol, pval, xd, bd, inner, sweeps=args.inner_sweeps)
            ce.refine(rid, 1e-2, max_outer=1)  # warm-up step
            xd.upload(np.zeros(n))
            t0 = time.time()
            steps, sweeps, rr = ce.refine(rid, args.tol, max_outer=2000)
            tt = time.time() - t0
            err = float(np.linalg.norm(xd.download() - xs) / np.linalg.norm(xs))
            mixed[name] = {"seconds": round(tt, 5), "refinements": steps, "sweeps": sweeps,
                           "relres": rr, "rel_err_vs_solution": err}
            log(f"refinement {name}: {tt * 1e3:.1f} ms, {steps} steps, {sweeps} sweeps, relres {rr:.2e}, "
                f"err {err:.2e}")
            del wx, wb, xd, bd
        mixed["speedup_fp32_vs_fp64"] = round(mixed["fp64"]["seconds"] / mixed["fp32"]["seconds"], 3)
        mixed["speedup_bf16_vs_fp64"] = round(mixed["fp64"]["seconds"] / mixed["bf16"]["seconds"], 3)

    extra = None
    if not args.no_extra:
        Ox = None
        if not args.no_cpu:
            from oracle import race_oracle as Ox
            Ox.build()
        extra = {"symmspmv": bench_symmspmv(R, M, O=Ox), "mpk": bench_mpk(R, M, O=Ox),
                 "spmtv": bench_spmtv(R, M)}
        if not args.no_batched:
            # the batched workload with the tuner's layouts (no autotune)
            bargs = argparse.Namespace(**vars(args))
            bargs.batch_layouts, bargs.batch_flow = "mlp", 1
            extra["batched"] = bench_batched(bargs, R, M)

    # CPU oracle beside it, and the parity of the timed GPU state: the oracle
    # builds its own RACE colouring of the same matrix (LevelRecursion in the
    # oracle restatement), runs the same warmup + steps fwd+bwd sweeps from
    # x = 0 in its RACE order on one thread and compares; then the same sweeps
    # phase by phase with the leaves of a phase on all host cores (RACE's own
    # OpenMP execution) for the CPU baseline.
    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import race_oracle as O
        O.build()
        t0 = time.time()
        G = O.Graph(rp, col, symm_hint=True)
        zt = O.level_recursion(G, args.threads, O.ONE)
        t_orc = time.time() - t0
        perm_equal = bool(np.array_equal(G.perm, perm))
        fcalls = np.asarray(O.leaf_calls(zt), np.int32).reshape(-1)
        rcalls = np.asarray(O.leaf_calls(zt, True), np.int32).reshape(-1)
        xo = np.zeros(n)
        nsteps = args.warmup + args.steps
        t0 = time.time()
        for _ in range(nsteps):
            O.run_calls(O.K_GS, prp, pcol, pval, xo, pb, fcalls)
            O.run_calls(O.K_GS, prp, pcol, pval, xo, pb, rcalls)
        t1 = (time.time() - t0) / nsteps
        err = float(np.linalg.norm(x_gpu - xo) / np.linalg.norm(xo))
        parity = {"gs_rel_err_vs_oracle": err, "steps": nsteps, "perm_equal_oracle": perm_equal,
                  "oracle_threads": zt.available_threads, "gpu_threads": ce.getNumThreads(),
                  "tol": 1e-12, "ok": bool(perm_equal and err <= 1e-12)}
        log(f"parity: GS {g}^3 after {nsteps} steps, rel err vs oracle {err:.2e}, perm equal {perm_equal} "
            f"(oracle colouring {t_orc:.1f}s)")
        ppf, cf = O.phase_calls(zt)
        ppr, cr = O.phase_calls(zt, True)  # reversed leaves (hi-1, lo-1), colours in reverse
        xo2 = np.zeros(n)
        nt = O.run_phases(O.K_GS, prp, pcol, pval, xo2, pb, ppf, cf)
        O.run_phases(O.K_GS, prp, pcol, pval, xo2, pb, ppr, cr)
        times = []
        t0 = time.time()
        while time.time() - t0 < args.cpu_seconds or len(times) < 3:
            ts = time.time()
            O.run_phases(O.K_GS, prp, pcol, pval, xo2, pb, ppf, cf)
            O.run_phases(O.K_GS, prp, pcol, pval, xo2, pb, ppr, cr)
            times.append(time.time() - ts)
        tc = float(np.median(times))
        cpu = {"value": round(bytes_step / tc / 1e9, 3), "unit": "GB/s", "cores": nt, "kind": "port",
               "sample": f"{len(times)} fwd+bwd GS steps on the full {g}^3 27-pt matrix, the oracle's RACE "
                         f"colouring ({zt.available_threads} threads), leaves of a phase on {nt} OpenMP "
                         f"threads, median {tc * 1e3:.1f} ms/step (oracle C restatement)",
               "host": host_info(),
               "single_thread": {"value": round(bytes_step / t1 / 1e9, 3), "ms_per_step": round(t1 * 1e3, 1),
                                 "steps": nsteps}}
        log(f"CPU oracle: {tc * 1e3:.1f} ms/step on {nt} threads, {t1 * 1e3:.1f} ms/step on 1 thread")
        del xo, xo2
    del rp, col

    # HBM bytes per launch measured by the FETCH_SIZE/WRITE_SIZE passes recorded
    # under profiles/ for this exact configuration (null when it differs)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "gs_bundle_27pt128_traffic.json")
    if os.path.exists(tfile):
        rec = json.load(open(tfile))
        if rec.get("grid") == g and rec.get("race_threads") == args.threads:
            traffic = rec["hbm_bytes_per_launch"]

    out = {
        "metric": METRIC,
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"RACE Gauss-Seidel fwd+bwd sweep, 3D 27-pt {g}^3 (n={n}, nnz={nnz}), "
                               f"distance-1 colouring, {args.threads} RACE threads, FP64",
                   "rows": n, "nnz": nnz, "race_threads": args.threads,
                   "launches_per_step": launches_per_step},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": traffic},
        "cpu_baseline": cpu,
        "parity": parity,
        "race_tuner": race_tuner,
        "mixed_precision": mixed,
        "other_kernels": extra,
    }
    print(json.dumps(out), flush=True)
    bad = [k for k, v in (("gs", parity), ("symmspmv", (extra or {}).get("symmspmv", {}).get("parity")),
                          ("mpk", (extra or {}).get("mpk", {}).get("parity")),
                          ("batched", (extra or {}).get("batched", {}).get("parity"))) if v and not v.get("ok")]
    if bad:
        log(f"PARITY FAILED for {bad}: the GPU result differs from the oracle")
        sys.exit(3)


if __name__ == "__main__":
    main()
