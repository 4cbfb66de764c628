#!/usr/bin/env python3
"""Benchmark of the MI355X state-vector engine on one GPU.

Workload (one step): the 30-qubit depth-20 random circuit (sqrt X / sqrt Y /
sqrt W on every qubit with the no-repeat rule, then CZ on the brick pattern;
890 gates) applied in complex64 to |0...0> held in HBM: reset to |0...0>,
plan and run every stage. Synthetic input, generated from seed 42. The engine
runs the circuit as register-tile stages (2^14 amplitudes per workgroup tile,
16 per thread in VGPRs) plus full-pass kernels for any op whose qubits cannot
share a stage.

Plans timed on the same circuit and state:
  heuristic     the engine's heuristic (plan variant 0: region tiles of the
                circuit's coupling graph)
  learned       the learned planner over the plans that run the heuristic's
                arithmetic byte for byte (bit-identical amplitudes): the
                headline `value`
  learned_any   the learned planner over every plan variant (other op orders:
                rounding differs)
  lookahead     the tile choice of src/blocking.rs's lookahead alone (plan
                variant 24)

Printed as one JSON line (stdout):
  value          seconds per circuit of the learned plan (lower is better);
                 its stages skip the tiles known to be zero (qubits no
                 non-diagonal gate has touched yet since |0...0>: the first
                 stage computes one tile and the reset writes the rest);
                 all_tiles_seconds is the same plan visiting every tile
  roofline       the tile-stage kernel of that plan: its algorithmic bytes
                 (each stage reads and writes the tiles it visits) over the
                 stage kernels' summed duration, timed with HIP events on the
                 engine's stream in the timed steps, against the 8 TB/s HBM
                 peak
  error          complex64 result against the oracle's FP64 engine on the
                 host (and against the same circuit in complex128 on the GPU)
  cpu_baseline   the oracle's restatement of the reference CPU path (AVX2
                 blocks, fusion 4, 2^14-block regions, OpenMP) on this host's
                 cores, same circuit, best of 3 after a warm-up
Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import glob
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


def measured_traffic(n: int, launches: int, fixed_lanes: int):
    """HBM bytes per stage launch of this workload (the plan's measured
    stage traffic over its launches), from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (calibrated against a copy pass of known
    size), or None when no such measurement exists for this size, launch
    count and register-tile geometry."""
    path = os.path.join(ROOT, "profiles", "k_vstage_random30_c64_traffic.json")
    if n != 30 or not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("launches") != launches or rec.get("fixed_lanes") != fixed_lanes:
        return None
    return rec["hbm_bytes_per_launch"]


def host_info() -> dict:
    """CPU model, logical CPUs (all and the ones this process may use),
    physical cores and NUMA nodes of this host."""
    model, cores = "", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and not model:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
            "physical_cores": len(cores) or None,
            "numa_nodes": len(glob.glob("/sys/devices/system/node/node[0-9]*")) or None}


def time_plan(sv, ops, steps: int, warmup: int, **kw) -> dict:
    """Warm-up runs, then `steps` timed runs (reset + plan + stages), with HIP
    events around every kernel launch; seconds per circuit and per launch."""
    for _ in range(warmup):
        sv.set_zero()
        sv.apply_ops(ops, **kw)
    stage_s = other_s = plan_s = 0.0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        sv.set_zero()
        st = sv.apply_ops(ops, timing=True, **kw)
        stage_s += st["stage_kernel_seconds"]
        other_s += st["other_kernel_seconds"]
        plan_s += st["plan_seconds"]
        launches += st["stages"]
    sec = (time.perf_counter() - t0) / steps
    return {"seconds": sec, "stage_kernel_seconds": stage_s / steps, "other_kernel_seconds": other_s / steps,
            "plan_seconds": plan_s / steps, "stage_launch_seconds": stage_s / max(launches, 1), "stats": st,
            "stage_bytes": st["stage_bytes"]}


def run_batch(count: int, cpu_sample: int, threads: int) -> dict:
    """Batching workload: `count` hardware-efficient-ansatz circuits (6 layers
    of RY, RZ on every qubit plus a CZ ladder; qubit counts uniform in 12..22,
    angles uniform in [0, 2 pi)) in complex64, <Z_i> and <Z_i Z_i+1> of each:
    one graph-captured batch (every stage of every circuit in one launch of
    the stage-queue kernel, register-tile geometry autotuned per circuit
    class) against the same batch with the default geometry, the same
    circuits run one by one on the GPU, and the first `cpu_sample` of them
    one by one on the oracle's CPU path."""
    import numpy as np
    from qevedo.simulator import Batch, StateVector, library, lower_gates
    from qevedo.simulator import _native

    rng = np.random.Generator(np.random.PCG64(2024))
    sizes = [int(x) for x in rng.integers(12, 23, size=count)]
    gate_lists = [library.hardware_efficient_ansatz(n, 6, 1000 + i) for i, n in enumerate(sizes)]
    circuits = [(n, _native.OpArray(lower_gates(g))) for n, g in zip(sizes, gate_lists)]
    def timed(b, reps=5):
        b.run()
        t = time.perf_counter()
        for _ in range(reps):
            out = b.run()
        return (time.perf_counter() - t) / reps, out

    t = time.perf_counter()
    b = Batch(circuits, "single", autotune=True)
    build_s = time.perf_counter() - t
    batch_s, out = timed(b)
    stats = b.stats
    lanes, stage_counts = b.fixed_lanes(), b.stages()
    b.close()
    with Batch(circuits, "single") as d:
        default_s, _ = timed(d)
        default_stages = sum(d.stages())
    with Batch(circuits, "single", all_tiles=True) as a:
        all_tiles_s, _ = timed(a)
    geometry = {str(n): sorted({f for f, m in zip(lanes, sizes) if m == n}) for n in sorted(set(sizes))}
    # Algorithmic HBM bytes of one replay, as the engine counts them: each
    # stage reads and writes the tiles it visits (a first stage from |0...0>
    # only writes), the initialisation (when a first stage skips zero tiles)
    # writes every state and the expectation values read every state.
    passes_bytes = stats["stage_bytes"] + stats["other_bytes"]
    achieved = passes_bytes / batch_s / 1e9
    # the same circuits one by one on the GPU: plan, run and expectation values per state
    t = time.perf_counter()
    worst = 0.0
    for i, (n, ops) in enumerate(circuits):
        sv = StateVector(n, "single")
        sv.apply_ops(ops)
        z, zz = sv.expect_z()
        sv.close()
        worst = max(worst, float(np.max(np.abs(z - out[i][0]))), float(np.max(np.abs(zz - out[i][1]))) if n > 1 else 0.0)
    single_s = time.perf_counter() - t
    cpu = None
    if cpu_sample:
        from oracle import oracle
        oracle.set_threads(threads)
        t = time.perf_counter()
        cworst = 0.0
        for i in range(min(cpu_sample, count)):
            n = sizes[i]
            with oracle.OracleState(n, "single") as o:
                o.apply_ops(oracle.lower(gate_lists[i]), 4, 14)
                p = np.abs(o.amplitudes()) ** 2
            # marginals of neighbouring qubit pairs: p summed over every
            # other qubit, as a (b_{q+1}, b_q) table
            z, zz = np.zeros(n), np.zeros(max(n - 1, 0))
            for q in range(n - 1):
                m = p.reshape(1 << (n - q - 2), 2, 2, 1 << q).sum(axis=(0, 3))
                zz[q] = m[0, 0] + m[1, 1] - m[0, 1] - m[1, 0]
                z[q] = m[0, 0] + m[1, 0] - m[0, 1] - m[1, 1]
                if q == n - 2:
                    z[q + 1] = m[0, 0] + m[0, 1] - m[1, 0] - m[1, 1]
            cworst = max(cworst, float(np.max(np.abs(z - out[i][0]))),
                         float(np.max(np.abs(zz - out[i][1]))) if n > 1 else 0.0)
        cpu_s = time.perf_counter() - t
        k = min(cpu_sample, count)
        cpu = {"value": cpu_s * count / k, "unit": "s", "cores": threads, "kind": "port",
               "sample": f"the first {k} of the {count} circuits one by one (oracle apply + <Z>, <ZZ> from the "
                         f"probabilities), scaled to {count}", "max_abs_diff_vs_batch": cworst}
    log(f"batch of {count} HEA circuits: {batch_s * 1e3:.2f} ms per run, one stage launch ({stats['stages']} "
        f"stages, at most {stats['passes']} per circuit, {stats['skipped_tiles']} zero tiles skipped; {achieved:.0f} "
        f"GB/s algorithmic; build and autotune {build_s:.2f} s; default geometry {default_s * 1e3:.2f} ms, "
        f"{default_stages} stages; every tile visited {all_tiles_s * 1e3:.2f} ms); "
        f"one by one on the GPU {single_s:.3f} s; max |diff| {worst:.2e}"
        + (f"; CPU one by one {cpu['value']:.2f} s (sampled)" if cpu else ""))
    return {"circuits": count, "qubits": "12..22", "seconds": batch_s, "one_by_one_seconds": single_s,
            "speedup": single_s / batch_s, "stage_launches_per_replay": 1, "stages": stats["stages"],
            "max_stages_per_circuit": stats["passes"], "build_and_autotune_seconds": build_s,
            "fixed_lanes_by_qubits": geometry, "default_geometry_seconds": default_s,
            "skipped_tiles": stats["skipped_tiles"], "all_tiles_seconds": all_tiles_s,
            "default_geometry_stages": default_stages, "max_abs_diff_vs_one_by_one": worst,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_replay": passes_bytes},
            "cpu_baseline": cpu}


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
    norm = sv.norm_sqr()
    ref.close()
    sv.close()
    log(f"precision form {n} q: {st['kernels']} fused ops, complex64 {c64_s:.3f} s (kernels "
        f"{st['other_kernel_seconds']:.3f} s; VALU dense kernels {st_v['other_kernel_seconds']:.3f} s), "
        f"complex128 {c128_s:.3f} s, max-abs {err['max_abs']:.2e}, rel L2 {err['rel_l2']:.2e}")
    return {"qubits": n, "depth": depth, "fused_ops": st["kernels"], "c64_seconds": c64_s,
            "c64_kernel_seconds": st["other_kernel_seconds"], "c64_valu_kernel_seconds": st_v["other_kernel_seconds"],
            "c128_seconds": c128_s, "max_abs": err["max_abs"], "rel_l2": err["rel_l2"], "c64_norm_sqr": norm}


def run_learned_form(n: int, steps: int, seed: int, model_path: str, reps: int = 2) -> dict:
    """Learned form: a kicked-Ising Trotter circuit (RX(theta) on every qubit,
    theta uniform in [0.1, 1.5], then RZZ(-pi/2) on a heavy-hex-like graph),
    run with the heuristic planner, the learned one (bit-identical plans
    only), the learned one over every variant and the lookahead planner;
    the final states are compared on the device."""
    import numpy as np

    from qevedo.simulator import StateVector, learned, library, lower_gates
    from qevedo.simulator import _native

    theta = float(np.random.Generator(np.random.PCG64(seed)).uniform(0.1, 1.5))
    gates = library.kicked_ising(n, library.heavy_hex_edges(n), steps, theta)
    ops = _native.OpArray(lower_gates(gates))
    times = {}
    net = learned.install(learned.load(model_path))
    base = StateVector(n, "single")
    other = StateVector(n, "single")
    diffs = {}
    try:
        for name, kw in (("heuristic", {}), ("learned", {"planner": _native.PLANNER_LEARNED}),
                         ("learned_any", {"planner": _native.PLANNER_LEARNED_ANY}), ("lookahead", {"variant": 24})):
            sv = base if name == "heuristic" else other
            st = sv.apply_ops(ops, 4, 0, **kw)  # warm-up (and plan choice)
            t0 = time.perf_counter()
            for _ in range(reps):
                sv.set_zero()
                st = sv.apply_ops(ops, 4, 0, **kw)
            times[name] = {"seconds": (time.perf_counter() - t0) / reps, "stages": st["stages"],
                           "variant": st["plan_variant"]}
            if sv is other:
                diffs[name] = sv.compare(base)
    finally:
        base.close()
        other.close()
        _native.clear_planner_model()
        net.close()
    h, l = times["heuristic"]["seconds"], times["learned"]["seconds"]
    log(f"learned form {n} q kicked Ising ({len(gates)} gates, theta {theta:.3f}): "
        + ", ".join(f"{k} {v['seconds']:.3f} s (variant {v['variant']}, {v['stages']} stages)" for k, v in times.items())
        + f"; learned vs heuristic max-abs {diffs['learned']['max_abs']:.1e}")
    return {"qubits": n, "steps": steps, "theta": theta, "gates": len(gates), "heuristic_seconds": h,
            "learned_seconds": l, "speedup": h / l, "learned_variant": times["learned"]["variant"],
            "heuristic_stages": times["heuristic"]["stages"], "learned_stages": times["learned"]["stages"],
            "max_abs_diff": diffs["learned"]["max_abs"], "rel_l2_diff": diffs["learned"]["rel_l2"],
            "learned_any": {**times["learned_any"], "max_abs_diff": diffs["learned_any"]["max_abs"]},
            "lookahead": {**times["lookahead"], "max_abs_diff": diffs["lookahead"]["max_abs"]}}


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
    ap.add_argument("--no-check", action="store_true", help="skip the complex128 and FP64-oracle error checks")
    ap.add_argument("--no-oracle-check", action="store_true", help="skip the FP64 oracle on the host")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = this process's CPU share (OMP_NUM_THREADS if set, else every usable CPU)")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched hardware-efficient-ansatz workload")
    ap.add_argument("--batch-cpu-sample", type=int, default=64)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned planner")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default register-tile geometry instead of timing the candidates once per circuit")
    ap.add_argument("--fixed-lanes", type=int, default=0,
                    help="register-tile geometry to use instead of autotuning it (lane bits pinned to the lowest "
                         "qubits; 0 = autotune unless --no-autotune)")
    ap.add_argument("--precision-form", type=int, default=32, metavar="QUBITS",
                    help="size of the precision form (fusion to 6-qubit GEMMs on matrix cores); 0 = skip")
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

    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    n, depth = args.qubits, args.depth
    gates = library.random_circuit(n, depth, args.seed)
    ops = _native.OpArray(lower_gates(gates))
    log(f"{n} qubits, depth {depth}: {len(gates)} gates")
    # The register-tile geometry (lane bits pinned to the lowest qubits) is
    # autotuned: the first run of each plan times the candidates once on a
    # scratch state (in the warm-up), every later run uses the choice.
    kw = {"max_fused_qubits": args.fusion, "tile_bits": args.tile_bits,
          "autotune": not args.no_autotune and not args.fixed_lanes, "fixed_lanes": args.fixed_lanes}

    sv = StateVector(n, "single")
    plans = {}
    h = time_plan(sv, ops, args.steps, args.warmup, **kw)
    plans["heuristic"] = h
    log(f"heuristic: {h['seconds']:.4f} s/circuit ({h['stats']['stages']} stages, plan {h['plan_seconds'] * 1e3:.1f} ms, "
        f"stage kernels {h['stage_kernel_seconds']:.4f} s, other kernels {h['other_kernel_seconds']:.4f} s)")
    heuristic_amps = StateVector(n, "single")
    heuristic_amps.apply_ops(ops, **kw)
    hd = h
    if kw["autotune"]:
        hd = time_plan(sv, ops, args.steps, min(args.warmup, 1), **{**kw, "autotune": False, "fixed_lanes": 0})
        plans["heuristic_default_geometry"] = hd
        log(f"heuristic on the default geometry ({hd['stats']['fixed_lanes']} fixed lane bits): {hd['seconds']:.4f} "
            f"s/circuit ({hd['stats']['stages']} stages); autotuned {h['stats']['fixed_lanes']} fixed lane bits")

    la = time_plan(sv, ops, args.steps, min(args.warmup, 1), variant=24, **kw)
    plans["lookahead"] = la
    log(f"lookahead (variant 24): {la['seconds']:.4f} s/circuit ({la['stats']['stages']} stages)")
    at = time_plan(sv, ops, args.steps, min(args.warmup, 1), all_tiles=True, **kw)
    plans["heuristic_all_tiles"] = at
    log(f"heuristic, every tile visited: {at['seconds']:.4f} s/circuit (stage kernels {at['stage_kernel_seconds']:.4f} s; "
        f"with zero tiles skipped {h['stats']['skipped_tiles']} tiles are not visited)")

    # The learned planner (a trained cost model in the tree): over the plans
    # that reproduce the heuristic's amplitudes (the headline) and over every
    # variant.
    model_path = os.path.join(ROOT, "qevedo-simulator_amd", "share", "planner_model.json")
    have_model = os.path.exists(model_path) and not args.no_learned
    diffs = {}
    if have_model:
        from qevedo.simulator import learned
        net = learned.install(learned.load(model_path))
        for name, planner in (("learned", _native.PLANNER_LEARNED), ("learned_any", _native.PLANNER_LEARNED_ANY)):
            r = time_plan(sv, ops, args.steps, args.warmup, planner=planner, **kw)
            plans[name] = r
            diffs[name] = sv.compare(heuristic_amps)
            log(f"{name}: variant {r['stats']['plan_variant']}, {r['seconds']:.4f} s/circuit "
                f"({r['stats']['stages']} stages); against the heuristic's state max-abs {diffs[name]['max_abs']:.1e}")
        _native.clear_planner_model()
        net.close()
    heuristic_amps.close()
    head = plans["learned"] if have_model else h
    value = head["seconds"]
    st = head["stats"]
    achieved = head["stage_bytes"] / head["stage_kernel_seconds"] / 1e9

    result_extra = {"passes": st["passes"], "stages": st["stages"], "fused_ops": st["kernels"], "gates": st["gates"],
                    "stage_launch_ms": head["stage_launch_seconds"] * 1e3, "plan_seconds": head["plan_seconds"],
                    "stage_bytes_per_circuit": head["stage_bytes"], "skipped_tiles": st["skipped_tiles"],
                    "all_tiles_seconds": at["seconds"],
                    "fixed_lanes": st["fixed_lanes"], "default_geometry_seconds": hd["seconds"],
                    "all_tiles_roofline_achieved": at["stage_bytes"] / at["stage_kernel_seconds"] / 1e9}

    for _ in range(args.calibration_passes):
        sv.scale(1.0)

    # Sampling: 10^6 shots through the two-pass sampler; the first call also
    # sizes the handle's sampling workspace and second stream. The steady
    # state samples into an output array already in use (best of 3); a new
    # output array adds the host's first-touch page faults to the copy out.
    sv.set_zero()
    sv.apply_ops(ops, **kw)
    if args.shots:
        u = np.random.Generator(np.random.PCG64(7)).random(args.shots)
        first = sv.sample(u)
        first_s = sv.last_sample_seconds()
        fresh = sv.sample(u)
        fresh_s = sv.last_sample_seconds()
        steady = []
        for _ in range(3):
            again = sv.sample(u, out=fresh)
            steady.append(sv.last_sample_seconds())
        assert np.array_equal(first, again)
        result_extra["sample_seconds"] = min(steady)
        result_extra["sample_seconds_new_output_array"] = fresh_s
        result_extra["sample_seconds_first_call"] = first_s
        result_extra["shots"] = args.shots
        log(f"sampling {args.shots} shots: {result_extra['sample_seconds']:.4f} s (into a new output array "
            f"{fresh_s:.4f} s; first call {first_s:.4f} s)")

    error = None
    if not args.no_check:
        ref = StateVector(n, "double")
        t = time.perf_counter()
        ref.apply_ops(ops, **{**kw, "fixed_lanes": 0})  # complex128 has its own geometries
        log(f"complex128 run: {time.perf_counter() - t:.3f} s")
        vs_c128 = sv.compare(ref)
        log(f"complex64 vs GPU complex128: max_abs {vs_c128['max_abs']:.3e} rel_l2 {vs_c128['rel_l2']:.3e}")
        error = {"vs_gpu_complex128": vs_c128}
        if not args.no_oracle_check:
            from oracle import oracle
            oracle.set_threads(threads)
            t = time.perf_counter()
            with oracle.OracleState(n, "double") as o:
                o.apply_ops(oracle.lower(gates), 4, 14)
                want = o.amplitudes()
            fp64 = StateVector(n, "double")
            fp64.set_amplitudes(want)
            del want
            error["max_abs"], error["rel_l2"] = (lambda e: (e["max_abs"], e["rel_l2"]))(sv.compare(fp64))
            error["gpu_complex128_vs_oracle"] = ref.compare(fp64)
            error["reference"] = "the oracle's FP64 engine on the host (restated reference CPU path, complex128)"
            fp64.close()
            log(f"FP64 oracle on {threads} host threads: {time.perf_counter() - t:.1f} s; complex64 vs oracle "
                f"max_abs {error['max_abs']:.3e} rel_l2 {error['rel_l2']:.3e}; GPU complex128 vs oracle rel_l2 "
                f"{error['gpu_complex128_vs_oracle']['rel_l2']:.3e}")
        ref.close()

    # QFT on the same state size, fused to <= 4 qubits.
    qft = _native.OpArray(lower_gates(library.qft(n)))
    sv.set_zero()
    sv.apply_ops(qft, **kw)
    sv.set_zero()
    st_q = sv.apply_ops(qft, **kw)
    result_extra["qft_seconds"] = st_q["gate_seconds"]
    result_extra["qft_passes"] = st_q["passes"]
    log(f"QFT {n}: {st_q['gate_seconds']:.4f} s, {st_q['passes']} passes")
    sv.close()

    if not args.no_batch:
        result_extra["batch"] = run_batch(args.batch_circuits, 0 if args.no_cpu else args.batch_cpu_sample, threads)
    if args.precision_form:
        result_extra["precision_form"] = run_precision_form(args.precision_form, args.depth, args.seed)
    if args.learned_form and have_model:
        result_extra["learned_form"] = run_learned_form(args.learned_form, 20, args.seed, model_path)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        oracle.set_threads(threads)
        o_ops = oracle.lower(gates)
        runs = []
        with oracle.OracleState(n, "single") as o:
            cst = o.apply_ops(o_ops, 4, 14)  # warm-up: pages touched, caches and threads started
            for _ in range(args.cpu_repeats):
                t = time.perf_counter()
                cst = o.apply_ops(o_ops, 4, 14)
                runs.append(time.perf_counter() - t)
        cpu_s = min(runs)
        info = host_info()
        log(f"CPU baseline ({threads} threads; {info['cpu_model']}, {info['nproc']} logical CPUs, "
            f"{info['usable_cpus']} usable, {info['physical_cores']} cores, {info['numa_nodes']} NUMA nodes): "
            f"best {cpu_s:.3f} s of {', '.join(f'{x:.3f}' for x in runs)}, {cst['passes']} passes")
        cpu = {"value": cpu_s, "unit": "s", "cores": threads, "kind": "port",
               "sample": f"the same {n}-qubit depth-{depth} circuit, complex64, fusion 4, regions of 2^14 blocks: "
                         f"gate time, best of {len(runs)} after a warm-up run",
               "runs": runs, "host": info, "speedup_vs_gpu": cpu_s / value}

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
                   "fusion": "none inside register-tile stages (each stage program merges runs of 1-qubit gates "
                             f"per qubit); ops outside stages fused to at most {args.fusion} qubits",
                   "tile_bits": args.tile_bits or "register tile, 2^14 amplitudes",
                   "geometry": f"{st['fixed_lanes']} lane bits pinned to the lowest qubits"
                               + (" (autotuned once per circuit class)" if kw["autotune"] else
                                  " (given)" if args.fixed_lanes else " (the default)"),
                   "planner": "learned (bit-identical to the heuristic)" if have_model else "heuristic"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n, st["stages"], st["fixed_lanes"]),
                     "algorithmic_bytes_per_launch": head["stage_bytes"] / max(st["stages"], 1)},
        "cpu_baseline": cpu,
        "error": error,
        "plans": {k: {"seconds": v["seconds"], "variant": v["stats"]["plan_variant"], "stages": v["stats"]["stages"],
                      "fixed_lanes": v["stats"]["fixed_lanes"],
                      "stage_launch_ms": v["stage_launch_seconds"] * 1e3, "skipped_tiles": v["stats"]["skipped_tiles"],
                      **({"max_abs_diff_vs_heuristic": diffs[k]["max_abs"]} if k in diffs else {})}
                  for k, v in plans.items()},
    }
    out.update(result_extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
