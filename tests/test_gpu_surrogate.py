"""Return-map surrogate on the GPU: BF16-MFMA inference, FP32 gradients and
Adam against the FP64 oracle network, deterministic training, and what the
trained network saves in Newton iterations without changing the solution.

Tolerances: inference |y - y_ref| <= 0.03 (1 + |y_ref|) (BF16 operands,
FP32 accumulation); gradient relative L2 <= 1e-4 (FP32 vs FP64); Adam
weights relative L2 <= 1e-5 after 3 steps; solution fields with and
without the surrogate relative L2 <= 1e-8 (both Newton solves stop at
|residual| <= 1e-11 s_trial).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from weldformgpu_amd import _abi, scenarios

from harness import randomize, rel_l2, small_c1, to_gpu

pytestmark = pytest.mark.gpu


def _dom(wflib):
    return to_gpu(small_c1(0.3), wflib)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.3, 1249).astype(np.float32)
    X = rng.uniform(0, 1.5, (n, 4)).astype(np.float32)
    Y = rng.uniform(0, 1, n).astype(np.float32)
    return w, X, Y


def test_mfma_inference_matches_oracle_network(wflib):
    w, X, _ = _data(0, 1000)
    with _dom(wflib) as d:
        d.set_mlp_params(w)
        y = d.mlp_infer(X)
    ref = np.array([po.mlp_forward(w.astype(np.float64), x.astype(np.float64)) for x in X])
    err = np.abs(y - ref)
    print(f"[case] mfma inference: max err {err.max():.3e}, ref scale {np.abs(ref).max():.3e}")
    assert np.all(err <= 0.03 * (1.0 + np.abs(ref)))
    assert np.corrcoef(y, ref)[0, 1] > 0.999


def test_mlp_gradient_matches_oracle(wflib):
    w, X, Y = _data(1, 3000)
    with _dom(wflib) as d:
        d.set_mlp_params(w)
        g, loss = d.mlp_grad(X, Y)
    g_ref, loss_ref = po.mlp_loss_grad(w.astype(np.float64), X.astype(np.float64),
                                       Y.astype(np.float64))
    print(f"[case] grad: rel L2 {rel_l2(g, g_ref):.2e}, loss {loss:.6f} vs {loss_ref:.6f}")
    assert rel_l2(g, g_ref) < 1e-4
    assert abs(loss - loss_ref) < 1e-5 * max(1.0, loss_ref)


def test_adam_matches_textbook(wflib):
    w, X, Y = _data(2, 500)
    lr = 1e-3
    with _dom(wflib) as d:
        d.set_mlp_params(w)
        d.mlp_reset_adam()
        d.mlp_fit(X, Y, 3, lr)
        w_gpu = d.mlp_params()
    # Kingma & Ba (2015), Algorithm 1, FP64, gradients from the oracle network
    wr = w.astype(np.float64)
    m = np.zeros(1249)
    v = np.zeros(1249)
    for t in range(1, 4):
        g, _ = po.mlp_loss_grad(wr, X.astype(np.float64), Y.astype(np.float64))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        wr = wr - lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    print(f"[case] adam: rel L2 {rel_l2(w_gpu - w, wr - w):.2e} of the update")
    assert rel_l2(w_gpu, wr) < 1e-5
    assert rel_l2(w_gpu - w, wr - w) < 1e-3


def test_training_is_deterministic(wflib):
    w, X, Y = _data(3, 4000)
    out = []
    for _ in range(2):
        with _dom(wflib) as d:
            d.set_mlp_params(w)
            d.mlp_reset_adam()
            loss = d.mlp_fit(X, Y, 40, 2e-3)
            out.append((d.mlp_params(), loss))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_surrogate_saves_iterations_same_solution(wflib):
    """Coupled JC run from a yielding state: the plain return map and the
    surrogate-started one (trained online every 15 steps) reach the same
    fields; after training the surrogate needs fewer Newton iterations."""
    res = {}
    for sur in (0, 1):
        sc = randomize(scenarios.c3_johnson_cook(scale=0.2), seed=9, v_scale=0.5,
                       s_scale=250e6)
        sc.cfg.surrogate = sur
        sc.cfg.mlp_train_interval = 15
        with to_gpu(sc, wflib) as d:
            d.step(30)
            s0 = d.stats()
            d.step(30)
            s1 = d.stats()
            res[sur] = {"fields": {k: d.get(k) for k in ["u", "sigma", "pl_strain", "T"]},
                        "iters": s1["newton_iters"] - s0["newton_iters"],
                        "calls": s1["newton_calls"] - s0["newton_calls"],
                        "wave": s1["newton_iters_wave"] - s0["newton_iters_wave"],
                        "trainings": s1["mlp_trainings"], "loss": s1["mlp_loss"]}
    p, s = res[0], res[1]
    print(f"[case] plain: {p['iters']} iters / {p['calls']} calls; surrogate: {s['iters']} / "
          f"{s['calls']} after {s['trainings']} trainings (loss {s['loss']:.2e})")
    assert s["trainings"] >= 3 and p["trainings"] == 0
    assert p["calls"] > 0 and abs(s["calls"] - p["calls"]) <= 0.01 * p["calls"]
    assert s["iters"] / s["calls"] < p["iters"] / p["calls"]
    for k in ["u", "sigma", "pl_strain", "T"]:
        assert rel_l2(s["fields"][k], p["fields"][k]) < 1e-8, k
