"""The tuner's MLP on the device: FP32 gradients and forward against the
oracle (float64 back-propagation, itself pinned by finite differences),
the BF16 matrix-core forward within BF16 rounding, training, and one sweep."""
import numpy as np
import pytest

import race_amd as R
from race_amd import matgen as M
from race_amd import tuner as TU

from .conftest import relerr

pytestmark = pytest.mark.gpu


def _problem(seed, B):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (B, 16)).astype(np.float32)
    T = rng.normal(0, 1, (B, 12)).astype(np.float32)
    mask = (rng.uniform(size=(B, 12)) > 0.2).astype(np.float32)
    return X, T, mask


@pytest.mark.parametrize("B", [1, 17, 200])
def test_mlp_gradients_match_oracle(oracle, B):
    m = TU.MLP(seed=5)
    p = m.params().astype(np.float64)
    X, T, mask = _problem(B, B)
    g, loss = m.gradients(X, T, mask)
    lo, go = oracle.mlp_loss_grad(p, X, T, mask)
    assert abs(loss - lo) <= 1e-5 * max(1.0, lo)
    assert relerr(g.astype(np.float64), go) <= 1e-5


@pytest.mark.parametrize("B", [1, 16, 33, 500])
def test_mlp_forward_fp32_and_bf16(oracle, B):
    m = TU.MLP(seed=7)
    p = m.params().astype(np.float64)
    X, _, _ = _problem(100 + B, B)
    ref = oracle.mlp_forward(p, X)
    assert relerr(m.forward(X, R.FP32).astype(np.float64), ref) <= 1e-5
    # BF16 weights and activations (8-bit mantissa), FP32 accumulation
    assert relerr(m.forward(X, R.BF16).astype(np.float64), ref) <= 3e-2


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (256, 16)).astype(np.float32)
    A = rng.normal(0, 0.5, (16, 12))
    T = np.tanh(X @ A).astype(np.float32)
    mask = np.ones_like(T)
    m = TU.MLP(seed=1)
    _, l0 = m.gradients(X, T, mask)
    l1 = m.train(X, T, mask, epochs=800, lr=3e-3)
    assert l1 < 0.1 * l0


def test_sweep_and_predict():
    rp, col = M.stencil7(12)
    val = M.gs_values(rp, col, seed=1)
    f, t = TU.sweep_one(R.K_SYMMSPMV, rp, col, val)
    assert f.shape == (16,) and t.shape == (12,) and (t > 0).all()
    tu = TU.Tuner(0)
    F = np.stack([f + np.random.default_rng(k).normal(0, 0.01, 16) for k in range(8)])
    T = np.stack([t] * 8)
    tu.fit(F, T, epochs=300)
    pick = tu.predict(F)
    assert pick.shape == (8,) and ((0 <= pick) & (pick < 12)).all()


def test_pretrained_mpk_tuner_scores_on_held_out():
    """The shipped MPK model (cache budget of RACEColor(4, budget)) on the
    held-out tail of its sweep: within 10 % of the autotuned best, and never
    a budget RACEColor refused."""
    import os
    z = np.load(os.path.join(os.path.dirname(TU.__file__), "tuner_data", "mpk_sweep.npz"))
    F, T = z["features"], z["times"]
    ho = slice(len(F) - 80, len(F))
    tu = TU.Tuner.pretrained("mpk")
    pick = tu.predict(F[ho])
    assert (T[ho][np.arange(80), pick] > 0).all()
    assert TU.score(T[ho], pick) >= 0.9


def test_mpk_sweep_one_matrix():
    rp, col = M.stencil7(10)
    val = M.laplace7_values(rp, col)
    f, t = TU.sweep_one(R.K_MPK, rp, col, val)
    assert f.shape == (16,) and t.shape == (12,)
    assert (t[-4:] > 0).all()


def test_pretrained_race_threads_tuner_scores_on_held_out():
    """The shipped RACE-threads model (rows per RACE thread for a symmetric
    GS step) on the held-out tail of its sweep: well above the fixed
    256-rows-per-thread rule."""
    import os
    z = np.load(os.path.join(os.path.dirname(TU.__file__), "tuner_data", "race_sweep.npz"))
    F, T = z["features"], z["times"]
    ho = slice(len(F) - 48, len(F))
    tu = TU.Tuner.pretrained("race")
    pick = tu.predict(F[ho])
    s_mlp = TU.score(T[ho], pick)
    s_fixed = TU.score(T[ho], np.full(48, TU.DEFAULT_CANDIDATE["race"]))
    assert s_mlp >= 0.85 and s_mlp > s_fixed + 0.1


@pytest.mark.parametrize("kind,holdout,floor", [("symmspmv", 200, 0.95), ("gs", 200, 0.95), ("kacz", 100, 0.95),
                                                ("eps", 60, 0.9)])
def test_pretrained_heads_beat_the_fixed_choice_on_held_out(kind, holdout, floor):
    """The shipped SymmSpMV / GS layout models (2000 training matrices, the
    held-out 200 include stencils of 3.8M and 4.1M rows), the Kaczmarz layout
    model of the batched workload (FP32, 8 right-hand sides, one-launch
    sweeps) and the efficiency model on the held-out tails of their sweeps:
    at least the floor of the autotuned best, and better than the fixed
    choice (layout 0, efficiency 40)."""
    import os
    z = np.load(os.path.join(os.path.dirname(TU.__file__), "tuner_data", f"{kind}_sweep.npz"))
    F, T = z["features"], z["times"]
    ho = slice(len(F) - holdout, len(F))
    pick = TU.Tuner.pretrained(kind).predict(F[ho])
    s_mlp = TU.score(T[ho], pick)
    s_fixed = TU.score(T[ho], np.full(holdout, TU.DEFAULT_CANDIDATE[kind]))
    print(kind, "held-out score", s_mlp, "fixed", s_fixed, flush=True)
    assert s_mlp >= floor and s_mlp > s_fixed
