"""The learned planner: data collection, training and installation.

The engine's planner has PLAN_VARIANTS variants (lookahead capacity, greedy
choice of the next tile, register slots for the most-used qubits). With
PLANNER_LEARNED it builds every variant, describes each register-tile stage by
PLAN_FEATURES numbers (qvd_plan_features) and lets a 3-layer perceptron
predict the stage's kernel time; the variant with the smallest predicted sum
runs. Without an installed model it falls back to the heuristic (variant 0).

This module collects the training data from the engine itself (every stage
of random plan variants timed with HIP events around its launch,
qvd_sv_stage_seconds), trains the network with the library's own forward,
backward and Adam kernels (Mlp.train), and saves / loads / installs it.

    python -m qevedo.simulator.learned --out model.json   (on a GPU)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import _native, library
from ._native import PLAN_FEATURES, PLAN_VARIANTS, Mlp, StateVector, lower_gates, plan_features

HIDDEN = 64
DEFAULT_MODEL = os.path.join(os.path.dirname(__file__), "..", "..", "..", "share", "planner_model.json")


def training_circuits(qubits: list, seed: int, ising_qubits: list = ()):
    """Random, QFT and kicked-Ising circuits of the given sizes (the families
    the planner's choice is measured on), plus kicked-Ising circuits at the
    larger `ising_qubits` sizes, as (name, n, gates)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for n in ising_qubits:
        for steps in (3, 8, 20):
            theta = float(rng.uniform(0.1, 1.5))
            out.append((f"ising{n}s{steps}", n, library.kicked_ising(n, library.heavy_hex_edges(n), steps, theta)))
    for n in qubits:
        for d in (3, 5, 8, 12, 16, 20):
            for k in range(2):
                out.append((f"random{n}d{d}.{k}", n, library.random_circuit(n, d, int(rng.integers(1 << 30)))))
        out.append((f"qft{n}", n, library.qft(n)))
        for steps in (2, 4, 6, 9, 12):
            theta = float(rng.uniform(0.1, 1.5))
            out.append((f"ising{n}s{steps}", n, library.kicked_ising(n, library.heavy_hex_edges(n), steps, theta)))
        for layers in (2, 6):
            out.append((f"hea{n}l{layers}", n, library.hardware_efficient_ansatz(n, layers, int(rng.integers(1 << 30)))))
    return out


def collect(circuits, variants=None, repeats: int = 2, log=print) -> tuple:
    """Times every register-tile stage of each (circuit, variant) run, the
    fastest of `repeats` runs per stage after one untimed warm-up run of the
    circuit: returns features (rows x PLAN_FEATURES), targets log(seconds
    per amplitude) and row labels."""
    variants = list(range(PLAN_VARIANTS)) if variants is None else list(variants)
    X, y, labels = [], [], []
    states = {}
    for name, n, gates in circuits:
        ops = _native.OpArray(lower_gates(gates))
        if n not in states:
            for other in states.values():  # one state at a time: circuits come grouped by size
                other.close()
            states = {n: StateVector(n, "single")}
        sv = states[n]
        sv.set_zero()
        sv.apply_ops(ops, all_tiles=True)
        seen = set()
        for v in variants:
            f = plan_features(n, ops, "single", v)
            key = f.tobytes()
            if key in seen:  # the same plan as an earlier variant
                continue
            seen.add(key)
            best = None
            for _ in range(repeats):
                # every tile visited: the model prices whole stages, also a
                # first stage that starts from |0...0>
                sv.set_zero()
                sv.apply_ops(ops, timing=True, variant=v, all_tiles=True)
                t = sv.stage_seconds()
                if len(t) != len(f):
                    raise RuntimeError(f"{name} variant {v}: {len(t)} timed stages, {len(f)} feature rows")
                best = t if best is None else np.minimum(best, t)
            X.append(f)
            y.append(np.log(np.maximum(best, 1e-9) / float(1 << n)))
            labels += [f"{name}/v{v}"] * len(best)
        log(f"[learned] {name}: {len(seen)} distinct plans, {sum(len(a) for a in X)} rows so far")
    for sv in states.values():
        sv.close()
    return np.concatenate(X).astype(np.float32), np.concatenate(y).astype(np.float32), labels


def init_params(inputs: int, hidden: int, seed: int) -> np.ndarray:
    """He-normal weights, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W1 = rng.normal(0, math.sqrt(2.0 / inputs), (hidden, inputs))
    W2 = rng.normal(0, math.sqrt(2.0 / hidden), (hidden, hidden))
    W3 = rng.normal(0, math.sqrt(1.0 / hidden), hidden)
    return np.concatenate([W1.ravel(), np.zeros(hidden), W2.ravel(), np.zeros(hidden), W3, [0.0]]).astype(np.float32)


def train(X: np.ndarray, y: np.ndarray, epochs: int = 400, batch: int = 256, lr: float = 2e-3, seed: int = 0,
          holdout: float = 0.1) -> dict:
    """Standardises the data, trains the network on the device and returns
    the model (parameters and normalisation) with held-out metrics."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = rng.permutation(len(X))
    nval = int(len(X) * holdout)
    val, tr = idx[:nval], idx[nval:]
    mean = X[tr].mean(axis=0)
    scale = X[tr].std(axis=0)
    scale[scale < 1e-6] = 1.0
    ymean, yscale = float(y[tr].mean()), float(y[tr].std() or 1.0)
    Xn = ((X - mean) / scale).astype(np.float32)
    yn = ((y - ymean) / yscale).astype(np.float32)
    net = Mlp(PLAN_FEATURES, HIDDEN)
    net.set_params(init_params(PLAN_FEATURES, HIDDEN, seed))
    t0 = time.perf_counter()
    loss = net.train(Xn[tr], yn[tr], epochs, batch, lr, seed)
    seconds = time.perf_counter() - t0
    model = {"inputs": PLAN_FEATURES, "hidden": HIDDEN, "params": net.params().tolist(), "mean": mean.tolist(),
             "scale": scale.tolist(), "y_mean": ymean, "y_scale": yscale}
    pred = net.forward(Xn) * yscale + ymean
    rel = np.abs(np.exp(pred - y) - 1.0)
    model["metrics"] = {"train_rows": int(len(tr)), "holdout_rows": int(nval), "train_mse_normalised": loss,
                        "train_seconds": seconds, "epochs": epochs,
                        "holdout_median_rel_error": float(np.median(rel[val])) if nval else None,
                        "train_median_rel_error": float(np.median(rel[tr]))}
    net.close()
    return model


def save(model: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(model, f)


def load(path: str = DEFAULT_MODEL) -> dict:
    with open(path) as f:
        return json.load(f)


def install(model: dict) -> Mlp:
    """Installs `model` as the learned planner's cost model."""
    net = Mlp(model["inputs"], model["hidden"])
    net.set_params(np.asarray(model["params"], dtype=np.float32))
    net.install(np.asarray(model["mean"], dtype=np.float32), np.asarray(model["scale"], dtype=np.float32),
                model["y_mean"], model["y_scale"])
    return net


def main() -> int:
    ap = argparse.ArgumentParser(description="collect stage timings on the GPU and train the planner's cost model")
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--qubits", default="26,27,28,29,30")
    ap.add_argument("--ising-qubits", default="31,32,33", help="sizes of the extra kicked-Ising circuits")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--data", default="", help="also save the collected rows (.npz)")
    args = ap.parse_args()
    qubits = [int(q) for q in args.qubits.split(",")]
    ising_qubits = [int(q) for q in args.ising_qubits.split(",") if q]
    t = time.perf_counter()
    X, y, labels = collect(training_circuits(qubits, args.seed, ising_qubits),
                           log=lambda m: print(m, file=sys.stderr, flush=True))
    print(f"[learned] {len(X)} stage rows in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    if args.data:
        np.savez(args.data, X=X, y=y, labels=np.array(labels))
    model = train(X, y, epochs=args.epochs, seed=args.seed)
    model["data"] = {"qubits": qubits, "ising_qubits": ising_qubits, "rows": int(len(X)), "seed": args.seed,
                     "source": "per-stage kernel time, HIP events around each register-tile stage launch"}
    save(model, args.out)
    print(json.dumps(model["metrics"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
