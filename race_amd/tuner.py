"""Layout tuner: a 16-64-64-12 MLP that picks one of the 12 candidate layouts
of a kernel family (race_batch_candidate) from 16 structural features of
the matrix, instead of timing all 12 (race_batch_autotune).

The model runs on the device through the C ABI (race_mlp_*): FP32 training
(full-batch Adam), BF16 inference on the matrix cores.  Its target for a
matrix is the relative log-time log(t_i / t_best) of every candidate i
(masked where a layout does not fit); the prediction is the candidate with
the smallest predicted value.

For the matrix powers (kind "mpk", k = 4) the candidates are the cache
budgets MPK_CACHE_MB handed to RACEColor(4, budget): the MPK block depth
(level groups per wavefront step), which sets the number of launches.
For RACE's own preprocessing (kind "race") the candidates are the rows per
RACE thread RACE_ROWS_PER_THREAD of a distance-1 colouring for a symmetric
Gauss-Seidel step: the thread count sets the recursion depth (stages,
colour phases) and the leaf sizes (dependency waves).  For RACE's efficiency
threshold (kind "eps") the candidates are EPS_VALUES at n // 64 threads.

    python -m race_amd.tuner sweep  --kind symmspmv --count 1200 --out data.npz
    python -m race_amd.tuner train  --data data.npz --out model.npz
    python -m race_amd.tuner score  --data data.npz --model model.npz
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib as L
from . import matgen as M
from .interface import (BF16, FP32, FP64, K_GS, K_KACZ, K_MPK, K_SYMMSPMV, ONE, POWER, TWO, Batch, DeviceArray,
                        Interface)

N_FEATURES, N_CANDIDATES = 16, 12
KINDS = {"symmspmv": K_SYMMSPMV, "gs": K_GS, "kacz": K_KACZ, "mpk": K_MPK, "race": -1, "eps": -2}
# MPK candidates: cache budgets in MB; the last is the Infinity Cache size,
# the fixed choice the tuner is compared with
MPK_CACHE_MB = [0.125, 0.25, 0.5, 1, 2, 4, 8, 16, 32, 64, 128, 256]
MPK_POWER = 4
# RACE threads = max(1, n // rows per thread); the fixed choice compared
# with is 256 rows per thread (the batched workload's rule)
RACE_ROWS_PER_THREAD = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
# RACE's efficiency threshold (RACE_EFFICIENCY, read by LevelRecursion at
# each stage, src/level_recursion.cpp:48,62-70) for a distance-1 colouring
# with n // EPS_ROWS_PER_THREAD threads; the fixed choice is RACE's default 40
EPS_VALUES = [5, 10, 20, 30, 40, 50, 60, 70, 80, 85, 90, 95]
EPS_ROWS_PER_THREAD = 64
DEFAULT_CANDIDATE = {"symmspmv": 0, "gs": 0, "kacz": 0, "mpk": len(MPK_CACHE_MB) - 1,
                     "race": RACE_ROWS_PER_THREAD.index(256), "eps": EPS_VALUES.index(40)}


def _fp(a):
    return a.ctypes.data_as(L.f32p)


class MLP:
    """The tuner's model (race_mlp_* in the C ABI)."""

    def __init__(self, seed=0):
        self._L = L.load()
        h = ctypes.c_void_p()
        L.check(self._L.race_mlp_create(ctypes.byref(h), ctypes.c_uint(seed)), "race_mlp_create")
        self.h = h
        d = [ctypes.c_int() for _ in range(4)]
        L.check(self._L.race_mlp_dims(*[ctypes.byref(v) for v in d]), "race_mlp_dims")
        self.n_in, self.n_hidden, self.n_out, self.n_params = (v.value for v in d)

    def params(self):
        p = np.empty(self.n_params, np.float32)
        L.check(self._L.race_mlp_get_params(self.h, _fp(p), self.n_params), "race_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        L.check(self._L.race_mlp_set_params(self.h, _fp(p), len(p)), "race_mlp_set_params")

    def forward(self, x, precision=FP32):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((len(x), self.n_out), np.float32)
        L.check(self._L.race_mlp_forward(self.h, _fp(x), len(x), _fp(y), int(precision)), "race_mlp_forward")
        return y

    def gradients(self, x, t, mask):
        x, t, mask = (np.ascontiguousarray(a, np.float32) for a in (x, t, mask))
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        L.check(self._L.race_mlp_gradients(self.h, _fp(x), _fp(t), _fp(mask), len(x), _fp(g), ctypes.byref(loss)),
                "race_mlp_gradients")
        return g, float(loss.value)

    def train(self, x, t, mask, epochs, lr):
        x, t, mask = (np.ascontiguousarray(a, np.float32) for a in (x, t, mask))
        loss = ctypes.c_float()
        L.check(self._L.race_mlp_train(self.h, _fp(x), _fp(t), _fp(mask), len(x), int(epochs), ctypes.c_float(lr),
                                       ctypes.byref(loss)), "race_mlp_train")
        return float(loss.value)

    def close(self):
        if getattr(self, "h", None):
            self._L.race_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# features
def features(rowPtr, col):
    """16 structural features of a square sparse pattern (CRS), computed on
    the host by the library (race_matrix_features): size, row-length
    moments, column offsets, breadth-first level profile."""
    rp = np.ascontiguousarray(rowPtr, np.int32)
    c = np.ascontiguousarray(col, np.int32)
    out = np.empty(N_FEATURES, np.float64)
    L.check(L.load().race_matrix_features(len(rp) - 1, rp.ctypes.data_as(L.i32p), c.ctypes.data_as(L.i32p),
                                          out.ctypes.data_as(L.f64p)), "race_matrix_features")
    return out


# ---------------------------------------------------------------------------
# sweeps (training data): every candidate timed on one matrix
def tuner_matrices(count, seed, n_lo=2000, n_hi=100000, big=0):
    """The batched workload's generators plus 3D 7- and 27-point stencils;
    with big > 0, every (count // big)-th matrix instead a large stencil of
    1-4.1M rows (7-point 100^3 .. 27-point 160^3)."""
    rng = np.random.default_rng(seed)
    gen = M.batch_matrices(count, seed + 1, n_lo, n_hi)
    every = count // big if big else 0
    for k in range(count):
        if every and k % every == every - 1:
            q = k // every
            g = int(100 + 60 * q / max(1, big - 1)) if big > 1 else 160
            seven = q % 2 == 0
            rp, col = (M.stencil7 if seven else M.stencil27)(g)
            val = M.gs_values(rp, col, seed=k)
            next(gen)
            yield ("stencil7" if seven else "stencil27"), rp, col, val
        elif k % 8 == 7:
            g = int(rng.integers(14, 48))
            rp, col = (M.stencil7 if k % 16 == 7 else M.stencil27)(g)
            val = M.gs_values(rp, col, seed=k)
            next(gen)
            yield ("stencil7" if k % 16 == 7 else "stencil27"), rp, col, val
        else:
            name, n, mean, rp, col, val = next(gen)
            yield name, rp, col, val


def sweep_mpk(rp, col, val, reps=3, stats=None):
    """(features, times[12]) of the matrix powers, k = MPK_POWER, FP64, for
    every cache budget of MPK_CACHE_MB; -1 where RACEColor refuses one."""
    n = len(rp) - 1
    x0 = M.uniform_vec(n, 7)
    times = np.full(N_CANDIDATES, -1.0)
    t0 = time.time()
    for i, mb in enumerate(MPK_CACHE_MB):
        ce = Interface(n, 1, POWER, rp, col, symm_hint=True)
        try:
            ce.RACEColor(MPK_POWER, mb * 1e6)
        except Exception:
            continue
        prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
        xs = np.zeros((MPK_POWER + 1) * n)
        xs[:n] = x0
        dx = DeviceArray.from_host(xs)
        fid = ce.registerDevice(K_MPK, prp, pcol, pval, dx, power=MPK_POWER)
        ce.executeAsync(fid)
        ce.timerBegin()
        for _ in range(reps):
            ce.executeAsync(fid)
        times[i] = ce.timerEnd() / reps
        ce.close()
    t1 = time.time()
    f = features(rp, col)
    if stats is not None:
        stats["autotune"] = t1 - t0
        stats["features"] = time.time() - t1
    return f, times


def sweep_race(rp, col, val, reps=3, stats=None):
    """(features, times[12]) of a symmetric Gauss-Seidel step (forward +
    backward sweep, FP64, default layout) for every RACE thread count
    n // RACE_ROWS_PER_THREAD[i], colouring at distance 1."""
    n = len(rp) - 1
    b = M.uniform_vec(n, 8)
    times = np.full(N_CANDIDATES, -1.0)
    t0 = time.time()
    for i, rpt in enumerate(RACE_ROWS_PER_THREAD):
        if n > 200000:
            print(f"  race candidate {i} ({rpt} rows per thread) of n = {n}", flush=True)
        ce = Interface(n, max(1, n // rpt), ONE, rp, col, symm_hint=True)
        ce.RACEColor()
        prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
        prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
        x, bd = DeviceArray.from_host(np.zeros(n)), DeviceArray.from_host(b)
        fid = ce.registerDevice(K_GS, prp, pcol, pval, x, bd)
        ce.executeAsync(fid)
        ce.executeAsync(fid, True)
        ce.timerBegin()
        for _ in range(reps):
            ce.executeAsync(fid)
            ce.executeAsync(fid, True)
        times[i] = ce.timerEnd() / reps
        ce.close()
    t1 = time.time()
    f = features(rp, col)
    if stats is not None:
        stats["autotune"] = t1 - t0
        stats["features"] = time.time() - t1
    return f, times


def sweep_eps(rp, col, val, reps=3, stats=None):
    """(features, times[12]) of a symmetric Gauss-Seidel step (forward +
    backward sweep, FP64, default layout) for every efficiency threshold
    EPS_VALUES (environment variable RACE_EFFICIENCY, as RACE reads it) at
    n // EPS_ROWS_PER_THREAD threads, colouring at distance 1."""
    n = len(rp) - 1
    b = M.uniform_vec(n, 8)
    times = np.full(N_CANDIDATES, -1.0)
    saved = os.environ.get("RACE_EFFICIENCY")
    t0 = time.time()
    try:
        for i, eps in enumerate(EPS_VALUES):
            os.environ["RACE_EFFICIENCY"] = str(eps)
            ce = Interface(n, max(1, n // EPS_ROWS_PER_THREAD), ONE, rp, col, symm_hint=True)
            ce.RACEColor()
            prp, pcol, pval = M.permute(rp, col, val, ce.getPerm(), ce.getInvPerm())
            prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
            x, bd = DeviceArray.from_host(np.zeros(n)), DeviceArray.from_host(b)
            fid = ce.registerDevice(K_GS, prp, pcol, pval, x, bd)
            ce.executeAsync(fid)
            ce.executeAsync(fid, True)
            ce.timerBegin()
            for _ in range(reps):
                ce.executeAsync(fid)
                ce.executeAsync(fid, True)
            times[i] = ce.timerEnd() / reps
            ce.close()
    finally:
        if saved is None:
            os.environ.pop("RACE_EFFICIENCY", None)
        else:
            os.environ["RACE_EFFICIENCY"] = saved
    t1 = time.time()
    f = features(rp, col)
    if stats is not None:
        stats["autotune"] = t1 - t0
        stats["features"] = time.time() - t1
    return f, times


def sweep_one(kind, rp, col, val, reps=3, stats=None):
    """(features, times[12]) of one matrix; times -1 where a layout does not
    fit.  Matrix coloured by RACE with max(4, n / 256) threads.  stats, if
    given, receives the seconds spent on features and on timing layouts.
    kind K_MPK sweeps cache budgets, -1 RACE thread counts, -2 efficiency
    thresholds; K_KACZ times the batched workload's Kaczmarz sweeps (FP32,
    8 right-hand sides, one-launch sweeps), the others FP64 with one."""
    if kind == K_MPK:
        return sweep_mpk(rp, col, val, reps, stats)
    if kind == -1:
        return sweep_race(rp, col, val, reps, stats)
    if kind == -2:
        return sweep_eps(rp, col, val, reps, stats)
    n = len(rp) - 1
    T = max(4, min(1024, n // 256))
    dist = ONE if kind == K_GS else TWO
    ce = Interface(n, T, dist, rp, col, symm_hint=True)
    ce.RACEColor()
    perm, inv = ce.getPerm(), ce.getInvPerm()
    prp, pcol, pval = M.permute(rp, col, val, perm, inv)
    prp, pcol, pval = M.make_diag_first(prp, pcol, pval)
    if kind == K_SYMMSPMV:
        prp, pcol, pval = M.upper_triangle(prp, pcol, pval)
    if kind == K_KACZ:
        # the batched workload's arithmetic: FP32 values and vectors, FP64
        # row norms, 8 right-hand sides, readiness-driven one-launch sweeps
        x = DeviceArray.from_host(np.zeros(8 * n, np.float32))
        b = DeviceArray.from_host(M.uniform_vec(n, 8, nrhs=8).astype(np.float32))
        bt = Batch(kind, FP32, FP32, 8)
        bt.setOption("flow", 1)
    else:
        x = DeviceArray.from_host(M.uniform_vec(n, 7))
        b = DeviceArray.from_host(M.uniform_vec(n, 8) if kind != K_SYMMSPMV else np.zeros(n))
        bt = Batch(kind, FP64, FP64, 1)
    bt.add(ce, prp, pcol, pval, x, b)
    times = np.full(N_CANDIDATES, -1.0)
    t0 = time.time()
    for i in range(N_CANDIDATES):
        bt.useCandidate(i)
        try:
            times[i] = bt.time(False, reps)
        except Exception:
            times[i] = -1.0
    t1 = time.time()
    f = features(rp, col)
    if stats is not None:
        stats["autotune"] = t1 - t0
        stats["features"] = time.time() - t1
    return f, times


def targets(times, penalty=None):
    """log(t_i / t_best) per matrix and the mask of the layouts that fit.
    With a penalty, a candidate that does not fit gets that target instead
    of being masked, so the model learns to avoid it."""
    t = np.atleast_2d(np.asarray(times, np.float64))
    ok = t > 0
    best = np.where(ok, t, np.inf).min(axis=1, keepdims=True)
    best = np.where(np.isfinite(best), best, 1.0)
    y = np.where(ok, np.log(np.where(ok, t, 1.0) / best), 0.0)
    if penalty is None:
        return y, ok.astype(np.float64)
    return np.where(ok, y, penalty), np.ones_like(y)


class Tuner:
    """Feature normalisation + MLP; predict() picks a candidate index."""

    def __init__(self, seed=0):
        self.mlp = MLP(seed)
        self.mu = np.zeros(N_FEATURES)
        self.sd = np.ones(N_FEATURES)

    def _norm(self, F):
        return ((np.asarray(F, np.float64) - self.mu) / self.sd).astype(np.float32)

    def fit(self, F, times, epochs=3000, lr=3e-3, penalty=None):
        F = np.asarray(F, np.float64)
        self.mu = F.mean(axis=0)
        self.sd = F.std(axis=0) + 1e-9
        y, m = targets(times, penalty)
        return self.mlp.train(self._norm(F), y, m, epochs, lr)

    def predict(self, F, precision=BF16):
        return np.argmin(self.scores(F, precision), axis=1)

    def scores(self, F, precision=BF16):
        """predicted log(t_i / t_best) of every candidate, per matrix"""
        return self.mlp.forward(self._norm(np.atleast_2d(F)), precision)

    @classmethod
    def pretrained(cls, kind):
        """The model shipped in race_amd/tuner_data for "symmspmv", "gs",
        "kacz", "mpk", "race" or "eps" (trained on the sweeps stored beside
        it)."""
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuner_data", f"{kind}_model.npz")
        return cls(0).load(path)

    def save(self, path):
        np.savez(path, params=self.mlp.params(), mu=self.mu, sd=self.sd)

    def load(self, path):
        z = np.load(path)
        self.mlp.set_params(z["params"])
        self.mu, self.sd = z["mu"], z["sd"]
        return self


def score(times, choice):
    """Geometric mean over matrices of t_best / t_chosen (1 = always the
    fastest layout)."""
    t = np.asarray(times, np.float64)
    best = np.where(t > 0, t, np.inf).min(axis=1)
    chosen = t[np.arange(len(t)), choice]
    chosen = np.where(chosen > 0, chosen, np.inf)
    return float(np.exp(np.mean(np.log(best / chosen))))


def _main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m race_amd.tuner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--kind", choices=sorted(KINDS), default="symmspmv")
    s.add_argument("--count", type=int, default=1200)
    s.add_argument("--seed", type=int, default=11)
    s.add_argument("--n-hi", type=int, default=100000)
    s.add_argument("--big", type=int, default=0, help="large stencils (1-4.1M rows) among the count matrices")
    s.add_argument("--stencils", default="",
                   help="sweep only these 3D stencils instead, e.g. 7:60,27:128 (points:grid size)")
    s.add_argument("--start", type=int, default=0, help="sweep only matrices [start, stop) of the sequence")
    s.add_argument("--stop", type=int, default=None)
    s.add_argument("--out", required=True)
    mg = sub.add_parser("merge")
    mg.add_argument("--parts", nargs="+", required=True)
    mg.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--holdout", type=int, default=200)
    t.add_argument("--epochs", type=int, default=3000)
    t.add_argument("--epochs-grid", default="",
                   help="comma-separated epoch counts: pick the one scoring best on the last --val training "
                        "matrices (trained on the rest), then train on all training matrices with it")
    t.add_argument("--val", type=int, default=0)
    t.add_argument("--penalty", type=float, default=None,
                   help="target for candidates that do not fit (default: masked)")
    t.add_argument("--out", required=True)
    c = sub.add_parser("score")
    c.add_argument("--data", required=True)
    c.add_argument("--model", required=True)
    c.add_argument("--holdout", type=int, default=200)
    a = ap.parse_args(argv)
    if a.cmd == "sweep":
        kind = KINDS[a.kind]
        F, T, names, secs, fsecs = [], [], [], [], []
        t0 = time.time()
        stop = a.count if a.stop is None else a.stop
        seq = tuner_matrices(a.count, a.seed, n_hi=a.n_hi, big=a.big)
        if a.stencils:
            def stencil_seq(spec):
                for q, item in enumerate(spec.split(",")):
                    pts, g = (int(v) for v in item.split(":"))
                    rp, col = (M.stencil7 if pts == 7 else M.stencil27)(g)
                    yield f"stencil{pts}", rp, col, M.gs_values(rp, col, seed=1000 + q)
            seq = stencil_seq(a.stencils)
            stop = len(a.stencils.split(","))
            a.count = stop
        for k, (name, rp, col, val) in enumerate(seq):
            if k < a.start or k >= stop:
                continue
            st = {}
            f, tm = sweep_one(kind, rp, col, val, stats=st)
            secs.append(st["autotune"])
            fsecs.append(st["features"])
            F.append(f)
            T.append(tm)
            names.append(name)
            if (k + 1) % 25 == 0 or len(rp) > 1000000:
                print(f"sweep {a.kind}: {k + 1}/{a.count} ({time.time() - t0:.0f}s)", flush=True)
        np.savez(a.out, features=np.array(F), times=np.array(T), names=np.array(names), kind=a.kind,
                 sweep_seconds=np.array(secs), feature_seconds=np.array(fsecs))
    elif a.cmd == "merge":
        zs = [np.load(p) for p in a.parts]
        np.savez(a.out, **{k: np.concatenate([z[k] for z in zs]) for k in
                           ("features", "times", "names", "sweep_seconds", "feature_seconds")},
                 kind=zs[0]["kind"])
    elif a.cmd == "train":
        z = np.load(a.data)
        F, T = z["features"], z["times"]
        tr = slice(0, len(F) - a.holdout)
        epochs = a.epochs
        if a.epochs_grid and a.val > 0:
            ntr = len(F) - a.holdout
            fit_part, val_part = slice(0, ntr - a.val), slice(ntr - a.val, ntr)
            best = None
            for ep in (int(v) for v in a.epochs_grid.split(",")):
                tv = Tuner(0)
                tv.fit(F[fit_part], T[fit_part], epochs=ep, penalty=a.penalty)
                sc = score(T[val_part], tv.predict(F[val_part]))
                print(f"epochs {ep}: validation score {sc:.4f} on {a.val} matrices", flush=True)
                if best is None or sc > best[0]:
                    best = (sc, ep)
            epochs = best[1]
            print(f"chosen epochs {epochs}", flush=True)
        tu = Tuner(0)
        t0 = time.time()
        loss = tu.fit(F[tr], T[tr], epochs=epochs, penalty=a.penalty)
        print(f"trained on {len(F) - a.holdout} matrices, loss {loss:.4f}, {time.time() - t0:.1f}s", flush=True)
        tu.save(a.out)
    else:
        z = np.load(a.data)
        F, T = z["features"], z["times"]
        ho = slice(len(F) - a.holdout, len(F))
        tu = Tuner(0).load(a.model)
        t0 = time.time()
        pick = tu.predict(F[ho])
        t_pred = time.time() - t0
        res = {"kind": str(z["kind"]), "heldout": int(len(F[ho])),
               "score_mlp_bf16": round(score(T[ho], pick), 4),
               "score_mlp_fp32": round(score(T[ho], tu.predict(F[ho], FP32)), 4),
               "score_default_layout": round(score(T[ho], np.full(len(F[ho]), DEFAULT_CANDIDATE[str(z["kind"])])), 4),
               "predict_seconds": round(t_pred, 4),
               "autotune_seconds": round(float(z["sweep_seconds"][ho].sum()), 2)}
        if "feature_seconds" in z:
            fs = float(z["feature_seconds"][ho].sum())
            res["feature_seconds"] = round(fs, 2)
            res["cost_tuner_vs_autotune"] = round((fs + t_pred) / float(z["sweep_seconds"][ho].sum()), 4)
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    _main()
