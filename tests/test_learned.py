"""Learned cost model: the NumPy reference network is pinned by finite
differences; the GPU network (BF16 MFMA forward/backward, Adam) is checked
against it; training must reduce the loss; the learned order must not lose to
the cost model's order on held-out patterns."""
import os

import numpy as np
import pytest

import sge
import sge.workloads as W
from oracle import mlp_ref as R


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(64 * 128 + 128 + 128 * 128 + 128 + 128 + 1, np.float32)
    o = 0
    for n, fan in ((64 * 128, 64), (128, 0), (128 * 128, 128), (128, 0), (128, 128), (1, 0)):
        p[o:o + n] = rng.uniform(-1, 1, n) * (np.sqrt(6 / fan) if fan else 0.1)
        o += n
    return p


def test_reference_gradients_match_finite_differences():
    R.ROUND = False
    try:
        rng = np.random.default_rng(1)
        p = rand_params(1).astype(np.float64)
        X = rng.normal(size=(24, 64))
        t = rng.normal(size=24)
        g, loss = R.gradients(p, X, t)
        idx = list(rng.choice(len(p), 40, replace=False)) + [len(p) - 1, 64 * 128 + 3]
        for i in idx:
            e = np.zeros_like(p)
            e[i] = 1e-6
            lp = np.mean((R.forward(p + e, X)[0] - t) ** 2)
            lm = np.mean((R.forward(p - e, X)[0] - t) ** 2)
            num = (lp - lm) / 2e-6
            assert abs(num - g[i]) <= 1e-5 + 1e-4 * abs(num), i
    finally:
        R.ROUND = True


@pytest.mark.gpu
def test_gpu_forward_and_gradients_match_reference():
    rng = np.random.default_rng(2)
    m = sge.MLP(seed=3)
    p = rand_params(4)
    m.set_params(p)
    for N in (1, 17, 64, 300):
        X = rng.normal(size=(N, 64)).astype(np.float32)
        t = rng.normal(size=N).astype(np.float32)
        y = m.forward(X)
        yr, _ = R.forward(p, X)
        # BF16 operands, FP32 accumulation: compare at BF16-level tolerance
        assert np.allclose(y, yr, rtol=2e-2, atol=2e-2 * np.abs(yr).max()), N
        g, loss = m.gradients(X, t)
        gr, lr = R.gradients(p, X, t)
        assert abs(loss - lr) <= 2e-2 * max(1.0, lr)
        scale = np.abs(gr).max()
        assert np.abs(g - gr).max() <= 3e-2 * scale, N
        # per-block relative check
        o = 0
        for n in (64 * 128, 128, 128 * 128, 128, 128, 1):
            blk, rb = g[o:o + n], gr[o:o + n]
            assert np.linalg.norm(blk - rb) <= 5e-2 * np.linalg.norm(rb) + 1e-6, (N, o)
            o += n


@pytest.mark.gpu
def test_gpu_adam_steps_match_numpy_update():
    """Two Adam steps on the GPU (adam_kernel) against the textbook update in
    NumPy from the same gradients: m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps), with b1 = 0.9,
    b2 = 0.999, eps = 1e-8; the BF16 weight copies the forward pass reads are
    refreshed from the new master weights."""
    rng = np.random.default_rng(7)
    m = sge.MLP(seed=1)
    p = rand_params(8)
    m.set_params(p)
    b1, b2, eps, lr = 0.9, 0.999, 1e-8, 3e-3
    mom = np.zeros_like(p, np.float64)
    vel = np.zeros_like(p, np.float64)
    want = p.astype(np.float64)
    for step in (1, 2):
        X = rng.normal(size=(48, 64)).astype(np.float32)
        t = rng.normal(size=48).astype(np.float32)
        g, _ = m.gradients(X, t)  # the gradients train_step applies (same batch, same weights)
        g = g.astype(np.float64)
        m.train_step(X, t, lr)
        mom = b1 * mom + (1 - b1) * g
        vel = b2 * vel + (1 - b2) * g * g
        want = want - lr * (mom / (1 - b1 ** step)) / (np.sqrt(vel / (1 - b2 ** step)) + eps)
        got = m.params().astype(np.float64)
        assert np.abs(got - want).max() <= 1e-5 * max(1.0, np.abs(want).max()), step
        # the forward pass reads BF16 copies refreshed from the updated weights
        yr, _ = R.forward(got.astype(np.float32), X)
        y = m.forward(X)
        assert np.allclose(y, yr, rtol=2e-2, atol=2e-2 * np.abs(yr).max()), step


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    rng = np.random.default_rng(5)
    X = rng.normal(size=(512, 64)).astype(np.float32)
    w = rng.normal(size=64)
    t = (np.tanh(X @ w / 8) * 2 + 0.1 * rng.normal(size=512)).astype(np.float32)
    m = sge.MLP(seed=0)
    first = m.train_step(X, t, 1e-3)
    for _ in range(300):
        last = m.train_step(X, t, 1e-3)
    assert last < 0.2 * first


@pytest.mark.gpu
def test_learned_order_gain_over_cost_model():
    """The committed model (trained on (order, work) pairs of patterns disjoint
    from these, logged on the Chung-Lu graph of seed 1 and on R-MAT scale 12)
    chooses orders for the 16 six-vertex patterns over a Chung-Lu graph of the
    same parameters but seed 9, a graph no pair was logged on.  Measured by the
    same executor on the whole job (every changed choice gives the same count):
    the learned choices must be at least 1.15x faster in total (1.21x
    measured), and with the cost-model fallback, the sampled judge and the
    measured probe no pattern may get more than 10% slower (timing noise)."""
    from sge import learned as L

    path = os.path.join(os.path.dirname(sge.__file__), "models", "order_mlp.npz")
    assert os.path.exists(path), "trained order model missing: python -m sge.train_order_model"
    model = L.load(path)
    g = sge.Graph.chung_lu(7115, 100762, 2.1, 1100.0, 9)
    rows = L.evaluate(model, g, W.six_vertex_set(), sge.MODE_IEP, stride=1, reps=2)
    t_cost = sum(r["cost_ms"] for r in rows)
    t_learn = sum(r["learned_ms"] for r in rows)
    ratios = [r["learned_ms"] / max(r["cost_ms"], 1e-3) for r in rows]
    print(f"cost-model orders {t_cost:.2f} ms, learned orders {t_learn:.2f} ms ({t_cost / t_learn:.3f}x); "
          f"per-pattern learned/cost {[round(x, 3) for x in ratios]}")
    assert not any(r["capped"] for r in rows)
    assert t_cost >= 1.15 * t_learn
    assert max(ratios) <= 1.10


@pytest.mark.gpu
def test_learned_order_never_much_slower_on_rmat():
    """On R-MAT scale 8 (seed 7; 78 seven-vertex patterns, a graph family the
    C2 pairs do not cover and a graph no pair was logged on) the fallback,
    judge and probe keep every learned choice within 15% of the cost-model
    order's time (an earlier model without them made one pattern 11x slower)
    and the learned set no slower in total.  The probe accepts an order that
    does at least 13% less work on a sample; a whole job's time can still come
    out up to about 10% higher through how that work spreads over the waves
    (pattern 44: 5.50 -> 6.08 ms with the in-pass continuation queue)."""
    from sge import learned as L

    model = L.load(os.path.join(os.path.dirname(sge.__file__), "models", "order_mlp.npz"))
    g = W.graph_rmat(8, 7)
    rows = L.evaluate(model, g, W.seven_vertex_set(), sge.MODE_IEP, stride=1, reps=2)
    ratios = [r["learned_ms"] / max(r["cost_ms"], 1e-3) for r in rows]
    worst = int(np.argmax(ratios))
    print(f"total cost {sum(r['cost_ms'] for r in rows):.1f} ms learned {sum(r['learned_ms'] for r in rows):.1f} ms; "
          f"worst pattern {worst}: {ratios[worst]:.3f}")
    assert max(ratios) <= 1.15
    assert sum(r["learned_ms"] for r in rows) <= sum(r["cost_ms"] for r in rows)


@pytest.mark.gpu
def test_model_keeps_to_its_training_sizes():
    """Outside the pattern sizes it was trained on, the model returns the cost
    model's plan."""
    from sge import learned as L

    m = L.load(os.path.join(os.path.dirname(sge.__file__), "models", "order_mlp.npz"))
    assert m.sizes == [6, 7]
    g = W.graph_c1()
    house = W.named("house")
    fb = sge.build_plan(house, g, sge.MODE_IEP)
    plan, info = L.choose_plan(m, house, g, sge.MODE_IEP, fallback=fb)
    assert info["fallback"] and bytes(plan) == bytes(fb)
