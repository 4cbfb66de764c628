"""The learned planner: the cost-model network's kernels against the CPU
oracle (oracle/mlp.py), training, plan variants (same amplitudes, different
stage partitions) and the planner's choice with and without a model."""
import os

import numpy as np
import pytest

from oracle import mlp as ref
from qevedo.simulator import StateVector, library, lower_gates
from qevedo.simulator import _native, learned

pytestmark = pytest.mark.gpu


def random_params(inputs, hidden, seed):
    return learned.init_params(inputs, hidden, seed) + np.random.Generator(np.random.PCG64(seed)).normal(
        0, 0.05, ref.num_params(inputs, hidden)).astype(np.float32)


def test_forward_and_gradient_match_the_oracle():
    inputs, hidden, rows = _native.PLAN_FEATURES, 64, 300
    rng = np.random.Generator(np.random.PCG64(1))
    p = random_params(inputs, hidden, 2)
    x = rng.normal(0, 1, (rows, inputs)).astype(np.float32)
    t = rng.normal(0, 1, rows).astype(np.float32)
    net = _native.Mlp(inputs, hidden)
    assert net.num_params == ref.num_params(inputs, hidden)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    y = net.forward(x)
    y_ref, _, _ = ref.forward(p, x, inputs, hidden)
    assert np.max(np.abs(y - y_ref)) < 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    g, loss = net.gradient(x, t)
    g_ref, loss_ref = ref.gradient(p, x, t, inputs, hidden)
    assert abs(loss - loss_ref) < 1e-4 * max(1.0, loss_ref)
    assert np.max(np.abs(g - g_ref)) < 1e-4 * max(1.0, np.max(np.abs(g_ref)))
    net.close()


def test_training_fits_a_smooth_function():
    rng = np.random.Generator(np.random.PCG64(3))
    inputs = 6
    x = rng.normal(0, 1, (2048, inputs)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2).astype(np.float32)
    net = _native.Mlp(inputs, 64)
    net.set_params(learned.init_params(inputs, 64, 4))
    before = float(np.mean((net.forward(x) - t) ** 2))
    after = net.train(x, t, epochs=60, batch=128, lr=3e-3, seed=5)
    assert after < 0.1 * before and after < 0.05
    net.close()


CIRCUITS = [
    ("random", 18, lambda n: library.random_circuit(n, 12, 3)),
    ("qft", 17, lambda n: library.qft(n)),
    ("ising", 19, lambda n: library.kicked_ising(n, library.heavy_hex_edges(n), 4, 0.9)),
]


@pytest.mark.parametrize("name,n,make", CIRCUITS)
def test_every_variant_gives_the_same_state(name, n, make):
    ops = _native.OpArray(lower_gates(make(n)))
    base = StateVector(n, "double")
    base.apply_ops(ops)
    ref64 = StateVector(n, "single")
    ref64.apply_ops(ops)
    for v in range(_native.PLAN_VARIANTS):
        print(f"{name} {n}q variant {v}", flush=True)
        s = StateVector(n, "double")
        st = s.apply_ops(ops, variant=v)
        assert st["plan_variant"] == v
        e = s.compare(base)
        assert e["rel_l2"] < 1e-12
        s.close()
        s = StateVector(n, "single")
        s.apply_ops(ops, variant=v)
        assert s.compare(base)["rel_l2"] < 1e-5
        s.close()
    base.close()
    ref64.close()


IDENTITY_CIRCUITS = CIRCUITS + [
    ("hea", 20, lambda n: library.hardware_efficient_ansatz(n, 6, 5)),
]


@pytest.mark.parametrize("name,n,make", IDENTITY_CIRCUITS)
@pytest.mark.parametrize("precision", ["single", "double"])
def test_equivalent_variants_are_bit_identical(name, n, make, precision):
    """Every candidate of the learned planner (a variant whose plan image
    equals the heuristic's) gives the heuristic's amplitudes bit for bit."""
    ops = _native.OpArray(lower_gates(make(n)))
    eq = _native.equivalent_variants(n, ops, precision)
    assert 0 in eq
    base = StateVector(n, precision)
    base.apply_ops(ops)
    want = base.amplitudes()
    base.close()
    for v in eq:
        print(f"{name} {n}q {precision} variant {v}", flush=True)
        s = StateVector(n, precision)
        s.apply_ops(ops, variant=v)
        assert np.array_equal(s.amplitudes(), want)
        s.close()


def synthetic_model():
    """A model that predicts log time = sum of the features (all weights
    positive): it prefers the plan with the fewest, smallest stages."""
    inputs, hidden = _native.PLAN_FEATURES, 64
    net = _native.Mlp(inputs, hidden)
    W1 = np.full((hidden, inputs), 0.01, np.float32)
    p = np.concatenate([W1.ravel(), np.zeros(hidden), np.eye(hidden, dtype=np.float32).ravel(), np.zeros(hidden),
                        np.full(hidden, 1.0 / hidden, np.float32), [0.0]]).astype(np.float32)
    net.set_params(p)
    net.install(np.zeros(inputs), np.ones(inputs), -25.0, 1.0)
    return net


def test_learned_planner_falls_back_and_chooses_the_predicted_best():
    n = 20
    ops = _native.OpArray(lower_gates(library.random_circuit(n, 10, 9)))
    _native.clear_planner_model()
    s = StateVector(n, "single")
    for planner in (_native.PLANNER_LEARNED, _native.PLANNER_LEARNED_ANY):
        st = s.apply_ops(ops, planner=planner)
        assert st["plan_variant"] == 0  # no model: the heuristic
    s.set_zero()
    s.apply_ops(ops)
    heuristic = s.amplitudes()
    net = synthetic_model()
    scores = []
    for v in range(_native.PLAN_VARIANTS):
        f = _native.plan_features(n, ops, variant=v)
        scores.append(float(np.sum(np.exp(-25.0 + 0.01 * f.sum(axis=1)))))
    s.set_zero()
    st = s.apply_ops(ops, planner=_native.PLANNER_LEARNED_ANY)
    assert st["plan_variant"] == int(np.argmin(scores))
    # the identity-preserving planner chooses among the equivalent variants
    # only, and reproduces the heuristic's amplitudes
    eq = _native.equivalent_variants(n, ops)
    s.set_zero()
    st = s.apply_ops(ops, planner=_native.PLANNER_LEARNED)
    assert st["plan_variant"] in eq
    assert st["plan_variant"] == min(eq, key=lambda v: (scores[v], v))
    assert np.array_equal(s.amplitudes(), heuristic)
    _native.clear_planner_model()
    net.close()
    s.close()


@pytest.mark.skipif(not os.path.exists(learned.DEFAULT_MODEL), reason="no trained planner model in the tree")
def test_trained_model_regret_against_every_variant():
    """On a held-out circuit (26-qubit kicked Ising, a theta and depth not in
    the training set) every distinct plan variant is timed; the trained
    model's choice over all variants is within 10 % of the fastest, and the
    identity-preserving learned planner reproduces the heuristic's amplitudes
    bit for bit."""
    model = learned.load()
    net = learned.install(model)
    n = 26
    gates = library.kicked_ising(n, library.heavy_hex_edges(n), 7, 0.37)
    ops = _native.OpArray(lower_gates(gates))
    s = StateVector(n, "single")

    def timed(**kw):
        best = None
        for _ in range(3):
            s.set_zero()
            st = s.apply_ops(ops, timing=True, **kw)
            t = st["stage_kernel_seconds"] + st["other_kernel_seconds"]
            best = t if best is None else min(best, t)
        return best, st["plan_variant"]

    times, seen = {}, {}
    for v in range(_native.PLAN_VARIANTS):
        key = _native.plan_features(n, ops, variant=v).tobytes()
        if key in seen:
            times[v] = times[seen[key]]
            continue
        seen[key] = v
        times[v] = timed(variant=v)[0]
    fastest = min(times.values())
    t_any, v_any = timed(planner=_native.PLANNER_LEARNED_ANY)
    t_id, v_id = timed(planner=_native.PLANNER_LEARNED)
    print(f"fastest variant {min(times, key=times.get)} {fastest * 1e3:.2f} ms; heuristic {times[0] * 1e3:.2f} ms; "
          f"lookahead (24) {times[24] * 1e3:.2f} ms; learned over every variant: variant {v_any} "
          f"{t_any * 1e3:.2f} ms (regret {times[v_any] / fastest:.3f}); identity-preserving: variant {v_id}",
          flush=True)
    assert times[v_any] <= 1.10 * fastest
    assert v_id in _native.equivalent_variants(n, ops)
    s.set_zero()
    s.apply_ops(ops)
    heuristic = s.amplitudes()
    s.set_zero()
    s.apply_ops(ops, planner=_native.PLANNER_LEARNED)
    assert np.array_equal(s.amplitudes(), heuristic)
    _native.clear_planner_model()
    net.close()
    s.close()
