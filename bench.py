#!/usr/bin/env python3
"""Benchmark of the MI355X state-vector engine on one GPU.

Workload (one step): the 30-qubit depth-20 random circuit (sqrt X / sqrt Y /
sqrt W on every qubit with the no-repeat rule, then CZ on the brick pattern;
890 gates) applied in complex64 to |0...0> held in HBM: reset to |0...0>,
plan and run every stage. Synthetic input,
generated from seed 42. The engine runs the circuit as register-tile stages
(2^14 amplitudes per workgroup tile, 32 per thread in VGPRs) plus full-pass
kernels for any op whose qubits cannot share a stage.

Printed as one JSON line (stdout):
  value          seconds per circuit (lower is better)
  roofline       the tile-stage kernel: algorithmic bytes per launch (one read
                 and one write of the 8 GiB state) over its average duration,
                 timed with HIP events on the engine's stream in the timed
                 steps, against the 8 TB/s HBM peak
  error          complex64 result against the same circuit in complex128 on
                 the GPU (max-abs, relative L2)
  cpu_baseline   the oracle's restatement of the reference CPU path (AVX2
                 blocks, fusion 4, 2^14-block regions, OpenMP) on this host's
                 cores, same circuit
Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "qevedo-simulator_amd", "python"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(n: int):
    """HBM bytes per stage launch of this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated against a copy pass
    of known size), or None when no such measurement exists for this size."""
    path = os.path.join(ROOT, "profiles", "k_vstage_random30_c64_traffic.json")
    if n != 30 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["hbm_bytes_per_launch"]


def run_batch(count: int) -> dict:
    """Batching workload: `count` hardware-efficient-ansatz circuits (6 layers
    of RY, RZ on every qubit plus a CZ ladder; qubit counts uniform in 12..22,
    angles uniform in [0, 2 pi)) in complex64, <Z_i> and <Z_i Z_i+1> of each:
    one graph-captured batch against the same circuits run one by one."""
    import numpy as np
    from qevedo.simulator import Batch, StateVector, library, lower_gates
    from qevedo.simulator import _native

    rng = np.random.Generator(np.random.PCG64(2024))
    sizes = [int(x) for x in rng.integers(12, 23, size=count)]
    circuits = [(n, _native.OpArray(lower_gates(library.hardware_efficient_ansatz(n, 6, 1000 + i))))
                for i, n in enumerate(sizes)]
    t = time.perf_counter()
    b = Batch(circuits, "single")
    build_s = time.perf_counter() - t
    b.run()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        out = b.run()
    batch_s = (time.perf_counter() - t) / reps
    stats = b.stats
    b.close()
    # the same circuits one by one: plan, run and expectation values per state
    t = time.perf_counter()
    worst = 0.0
    for i, (n, ops) in enumerate(circuits):
        sv = StateVector(n, "single")
        sv.apply_ops(ops)
        z, zz = sv.expect_z()
        sv.close()
        worst = max(worst, float(np.max(np.abs(z - out[i][0]))), float(np.max(np.abs(zz - out[i][1]))) if n > 1 else 0.0)
    single_s = time.perf_counter() - t
    log(f"batch of {count} HEA circuits: {batch_s * 1e3:.2f} ms per run ({stats['passes']} stage levels, "
        f"{stats['stages']} stage descriptors; build {build_s:.2f} s); one by one {single_s:.3f} s; "
        f"max |diff| {worst:.2e}")
    return {"circuits": count, "qubits": "12..22", "seconds": batch_s, "one_by_one_seconds": single_s,
            "speedup": single_s / batch_s, "stage_levels": stats["passes"], "build_seconds": build_s,
            "max_abs_diff_vs_one_by_one": worst}


def run_precision_form(n: int, depth: int, seed: int) -> dict:
    """Precision form: the random circuit fused to dense ops of at most 6
    qubits, each its own HBM pass (no stages), so every 5- and 6-qubit fused
    gate runs as a complex GEMM on FP32 matrix cores; the same plan in
    complex128 (VALU FP64) is the reference the complex64 result is compared
    with on the device."""
    from qevedo.simulator import StateVector, library, lower_gates
    from qevedo.simulator import _native

    ops = _native.OpArray(lower_gates(library.random_circuit(n, depth, seed)))
    sv = StateVector(n, "single")
    sv.apply_ops(ops, 6, None)
    sv.set_zero()
    t = time.perf_counter()
    st = sv.apply_ops(ops, 6, None, timing=True)
    c64_s = time.perf_counter() - t
    sv.set_zero()
    st_v = sv.apply_ops(ops, 6, None, timing=True, dense_valu=True)
    sv.set_zero()
    sv.apply_ops(ops, 6, None)
    ref = StateVector(n, "double")
    t = time.perf_counter()
    ref.apply_ops(ops, 6, None)
    c128_s = time.perf_counter() - t
    err = sv.compare(ref)
    ref.close()
    sv.close()
    log(f"precision form {n} q: {st['kernels']} fused ops, complex64 {c64_s:.3f} s (kernels "
        f"{st['other_kernel_seconds']:.3f} s; VALU dense kernels {st_v['other_kernel_seconds']:.3f} s), "
        f"complex128 {c128_s:.3f} s, max-abs {err['max_abs']:.2e}, rel L2 {err['rel_l2']:.2e}")
    return {"qubits": n, "depth": depth, "fused_ops": st["kernels"], "c64_seconds": c64_s,
            "c64_kernel_seconds": st["other_kernel_seconds"], "c64_valu_kernel_seconds": st_v["other_kernel_seconds"],
            "c128_seconds": c128_s, "max_abs": err["max_abs"], "rel_l2": err["rel_l2"]}


def run_learned_form(n: int, steps: int, seed: int, model_path: str, reps: int = 2) -> dict:
    """Learned form: a kicked-Ising Trotter circuit (RX(theta) on every qubit,
    theta uniform in [0.1, 1.5], then RZZ(-pi/2) on a heavy-hex-like graph),
    run with the heuristic planner and with the learned one; both final
    states are kept and compared on the device."""
    import numpy as np

    from qevedo.simulator import StateVector, learned, library, lower_gates
    from qevedo.simulator import _native

    theta = float(np.random.Generator(np.random.PCG64(seed)).uniform(0.1, 1.5))
    gates = library.kicked_ising(n, library.heavy_hex_edges(n), steps, theta)
    ops = _native.OpArray(lower_gates(gates))
    times = {}
    states = {}
    net = learned.install(learned.load(model_path))
    try:
        for name, planner in (("heuristic", _native.PLANNER_HEURISTIC), ("learned", _native.PLANNER_LEARNED)):
            sv = StateVector(n, "single")
            st = sv.apply_ops(ops, 4, 0, planner=planner)  # warm-up (and plan choice)
            t0 = time.perf_counter()
            for _ in range(reps):
                sv.set_zero()
                st = sv.apply_ops(ops, 4, 0, planner=planner)
            times[name] = (time.perf_counter() - t0) / reps
            times[name + "_stages"] = st["stages"]
            times[name + "_variant"] = st["plan_variant"]
            states[name] = sv
        diff = states["learned"].compare(states["heuristic"])
    finally:
        for sv in states.values():
            sv.close()
        _native.clear_planner_model()
        net.close()
    log(f"learned form {n} q kicked Ising ({len(gates)} gates, theta {theta:.3f}): heuristic {times['heuristic']:.3f} s "
        f"({times['heuristic_stages']} stages), learned {times['learned']:.3f} s (variant {times['learned_variant']}, "
        f"{times['learned_stages']} stages): {times['heuristic'] / times['learned']:.3f}x; "
        f"states differ by max-abs {diff['max_abs']:.2e}")
    return {"qubits": n, "steps": steps, "theta": theta, "gates": len(gates), "heuristic_seconds": times["heuristic"],
            "learned_seconds": times["learned"], "speedup": times["heuristic"] / times["learned"],
            "learned_variant": times["learned_variant"], "heuristic_stages": times["heuristic_stages"],
            "learned_stages": times["learned_stages"], "max_abs_diff": diff["max_abs"], "rel_l2_diff": diff["rel_l2"]}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--fusion", type=int, default=4)
    ap.add_argument("--tile-bits", type=int, default=0, help="0 = engine default")
    ap.add_argument("--shots", type=int, default=1_000_000)
    ap.add_argument("--no-check", action="store_true", help="skip the complex128 error check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores (capped at 16)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched hardware-efficient-ansatz workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned planner")
    ap.add_argument("--precision-form", type=int, default=0, metavar="QUBITS",
                    help="also run the precision form (fusion to 6-qubit GEMMs on matrix cores) at this size")
    ap.add_argument("--batch-circuits", type=int, default=512)
    ap.add_argument("--learned-form", type=int, default=33, metavar="QUBITS",
                    help="kicked-Ising circuit size for the learned-versus-heuristic comparison (0 = skip)")
    ap.add_argument("--calibration-passes", type=int, default=0,
                    help="extra in-place scale-by-1 passes (k_scale: one read and one write of the state), "
                         "the known-size kernel that hardware byte counters are scaled by")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the state-vector engine is single-device "
              "(multi-GPU sharding is a design, not part of this build)", file=sys.stderr)
        return 2

    import numpy as np
    from qevedo.simulator import StateVector, library, lower_gates
    from qevedo.simulator import _native

    n, depth = args.qubits, args.depth
    gates = library.random_circuit(n, depth, args.seed)
    ops = _native.OpArray(lower_gates(gates))
    log(f"{n} qubits, depth {depth}: {len(gates)} gates")

    sv = StateVector(n, "single")
    for i in range(args.warmup):
        sv.set_zero()
        st = sv.apply_ops(ops, args.fusion, args.tile_bits)
        log(f"warmup {i}: {st['gate_seconds']:.4f} s, {st['passes']} passes, {st['kernels']} fused ops")

    stage_s = other_s = plan_s = 0.0
    stage_launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sv.set_zero()
        st = sv.apply_ops(ops, args.fusion, args.tile_bits, timing=True)
        stage_s += st["stage_kernel_seconds"]
        other_s += st["other_kernel_seconds"]
        plan_s += st["plan_seconds"]
        stage_launches += st["stages"]
    elapsed = time.perf_counter() - t0
    sec = elapsed / args.steps
    log(f"timed {args.steps} steps: {sec:.4f} s/circuit (plan {plan_s / args.steps:.4f} s, "
        f"stage kernels {stage_s / args.steps:.4f} s, other kernels {other_s / args.steps:.4f} s)")

    state_bytes = (1 << n) * 8
    per_launch_bytes = 2 * state_bytes
    avg_stage = stage_s / max(stage_launches, 1)
    achieved = per_launch_bytes / avg_stage / 1e9 if stage_launches else 0.0

    result_extra = {"passes": st["passes"], "stages": st["stages"], "fused_ops": st["kernels"], "gates": st["gates"],
                    "stage_launch_ms": avg_stage * 1e3, "plan_seconds": plan_s / args.steps}

    for _ in range(args.calibration_passes):
        sv.scale(1.0)

    # The learned planner (a trained cost model in the tree): the same circuit
    # with the plan variant the model predicts fastest.
    learned_s = None
    model_path = os.path.join(ROOT, "qevedo-simulator_amd", "share", "planner_model.json")
    if os.path.exists(model_path) and not args.no_learned:
        from qevedo.simulator import learned
        net = learned.install(learned.load(model_path))
        sv.set_zero()
        st_l = sv.apply_ops(ops, args.fusion, args.tile_bits, planner=_native.PLANNER_LEARNED)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sv.set_zero()
            st_l = sv.apply_ops(ops, args.fusion, args.tile_bits, planner=_native.PLANNER_LEARNED)
        learned_s = (time.perf_counter() - t0) / args.steps
        result_extra["learned_variant"] = st_l["plan_variant"]
        result_extra["learned_plan_seconds"] = st_l["plan_seconds"]
        log(f"learned planner: variant {st_l['plan_variant']}, {learned_s:.4f} s/circuit "
            f"(heuristic {sec:.4f} s, plan {st_l['plan_seconds'] * 1e3:.1f} ms)")
        _native.clear_planner_model()
        net.close()

    # Sampling: 10^6 shots through the two-pass sampler; the first call also
    # sizes the handle's sampling workspace, the second is the steady state.
    if args.shots:
        u = np.random.Generator(np.random.PCG64(7)).random(args.shots)
        first = sv.sample(u)
        first_s = sv.last_sample_seconds()
        again = sv.sample(u)
        assert np.array_equal(first, again)
        result_extra["sample_seconds"] = sv.last_sample_seconds()
        result_extra["sample_seconds_first_call"] = first_s
        result_extra["shots"] = args.shots
        log(f"sampling {args.shots} shots: {result_extra['sample_seconds']:.4f} s (first call {first_s:.4f} s)")

    error = None
    if not args.no_check:
        ref = StateVector(n, "double")
        t = time.perf_counter()
        ref.apply_ops(ops, args.fusion, args.tile_bits)
        log(f"complex128 run: {time.perf_counter() - t:.3f} s")
        error = sv.compare(ref)
        ref.close()
        log(f"error vs complex128: max_abs {error['max_abs']:.3e} rel_l2 {error['rel_l2']:.3e}")

    # QFT on the same state size, fused to <= 4 qubits.
    qft = _native.OpArray(lower_gates(library.qft(n)))
    sv.set_zero()
    sv.apply_ops(qft, args.fusion, args.tile_bits)
    sv.set_zero()
    st_q = sv.apply_ops(qft, args.fusion, args.tile_bits)
    result_extra["qft_seconds"] = st_q["gate_seconds"]
    result_extra["qft_passes"] = st_q["passes"]
    log(f"QFT {n}: {st_q['gate_seconds']:.4f} s, {st_q['passes']} passes")
    sv.close()

    if not args.no_batch:
        result_extra["batch"] = run_batch(args.batch_circuits)
    if args.precision_form:
        result_extra["precision_form"] = run_precision_form(args.precision_form, args.depth, args.seed)
    if args.learned_form and os.path.exists(model_path) and not args.no_learned:
        result_extra["learned_form"] = run_learned_form(args.learned_form, 20, args.seed, model_path)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        oracle.set_threads(threads)
        o_ops = oracle.lower(gates)
        with oracle.OracleState(n, "single") as o:
            t = time.perf_counter()
            cst = o.apply_ops(o_ops, 4, 14)
            cpu_s = time.perf_counter() - t
        log(f"CPU baseline ({threads} threads): {cpu_s:.3f} s, {cst['passes']} passes")
        cpu = {"value": cpu_s, "unit": "s", "cores": threads, "kind": "port",
               "sample": f"the same {n}-qubit depth-{depth} circuit once, complex64, fusion 4, regions of 2^14 blocks"}

    # The headline is the learned planner's time when a trained model ships
    # in the tree (the heuristic's time is in "plans"), else the heuristic's.
    value = learned_s if learned_s is not None else sec
    out = {
        "metric": "30q random-circuit s (learned vs heuristic plan); HBM TB/s; err<=1e-5 vs FP64",
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"random circuit, {n} qubits, depth {depth} (sqrt X/Y/W with no repeats + CZ brick), "
                               f"complex64 from |0...0>", "qubits": n, "depth": depth, "gates": len(gates),
                   "max_fused_qubits": args.fusion, "tile_bits": args.tile_bits or "register tile, 2^14 amplitudes",
                   "planner": "learned" if learned_s is not None else "heuristic"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n)},
        "cpu_baseline": cpu,
        "error": error,
        "plans": {"heuristic": sec, "learned": learned_s},
    }
    out.update(result_extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
