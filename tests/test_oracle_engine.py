"""Pins the oracle's blocked CPU engine (the restated reference CPU path) the
way the reference's integration tests pin the real one (tests/kernels.rs,
tests/simulator.rs), against the naive FP64 applier and closed forms."""
import math

import numpy as np
import pytest

from helpers import (lower_with_oracle, max_distance, oracle_statevector, random_diagonal, random_qubits,
                     random_state, random_unitary, rng)
from oracle import oracle
from qevedo.simulator import library


def _kernel_sweep(precision, seed, tolerance, cases=400):
    """tests/kernels.rs:12-62: n in [1,11], k in [1, min(n,6)], up to two
    controls (value true with p = 0.7), 30% diagonal."""
    r = rng(seed)
    for case in range(cases):
        n = int(r.integers(1, 12))
        k = int(r.integers(1, min(n, 6) + 1))
        qubits = random_qubits(r, n, k)
        rest = [q for q in range(n) if q not in qubits]
        if not rest or k >= 6 or r.random() < 0.5:
            controls = []
        else:
            c = int(r.integers(1, min(len(rest), 2, 6 - k) + 1))
            controls = [(q, bool(r.random() < 0.7)) for q in rest[:c]]
        diagonal = r.random() < 0.3
        m = random_diagonal(r, k) if diagonal else random_unitary(r, k)
        initial = random_state(r, n)
        expected = initial.copy()
        state_bits = sum((1 if v else 0) << i for i, (_, v) in enumerate(controls))
        full = oracle.matrix_controlled(m, len(controls), state_bits) if controls else m
        oracle.reference_apply(expected, full, [q for q, _ in controls] + qubits)
        with oracle.OracleState(n, precision) as s:
            s.set_amplitudes(initial)
            s.apply(m, qubits, controls)
            d = max_distance(s.amplitudes(), expected)
        assert d < tolerance, f"case {case}: n={n} q={qubits} c={controls} diag={diagonal}: {d}"


def test_kernels_match_reference_f64():
    _kernel_sweep("double", 1, 1e-12)


def test_kernels_match_reference_f32():
    _kernel_sweep("single", 2, 2e-6)


def test_large_state_parallel_path():
    # tests/kernels.rs:74-92: 2^18 amplitudes, k = 1..5
    r = rng(3)
    n = 18
    initial = random_state(r, n)
    for k in range(1, 6):
        q = random_qubits(r, n, k)
        m = random_unitary(r, k)
        expected = initial.copy()
        oracle.reference_apply(expected, m, q)
        with oracle.OracleState(n, "double") as s:
            s.set_amplitudes(initial)
            s.apply(m, q)
            assert max_distance(s.amplitudes(), expected) < 1e-12


def test_circuits_every_fusion_width():
    # tests/simulator.rs:8-33 (no cache blocking)
    for seed in range(40):
        n = 1 + seed % 10
        gates = library.random_gate_circuit(n, 60, seed)
        expected = oracle_statevector(n, gates)
        ops = lower_with_oracle(gates)
        for width in range(1, 7):
            with oracle.OracleState(n, "double") as s:
                st = s.apply_ops(ops, width, None)
                assert max_distance(s.amplitudes(), expected) < 1e-11, (seed, width)
                assert st["kernels"] <= st["gates"]
            with oracle.OracleState(n, "single") as s:
                s.apply_ops(ops, width, None)
                assert max_distance(s.amplitudes(), expected) < 1e-5, (seed, width)


def test_library_circuits_default_options():
    # tests/simulator.rs:35-43
    for n, gates in [(9, library.ghz(9)), (9, library.qft(9)), (10, library.random_circuit(10, 12, 7))]:
        expected = oracle_statevector(n, gates)
        with oracle.OracleState(n, "double") as s:
            st = s.apply_ops(lower_with_oracle(gates), 4, 14)
            assert max_distance(s.amplitudes(), expected) < 1e-12
            assert st["kernels"] < st["gates"]


def test_closed_forms():
    # GHZ amplitudes 1/sqrt(2) (python/tests/test_simulator.py:77-81) and
    # QFT|0..0> = uniform 2^(-n/2) (every cp control is still |0>).
    for n in (3, 9, 12):
        a = oracle_statevector(n, library.ghz(n))
        want = np.zeros(1 << n, dtype=complex)
        want[0] = want[-1] = 1 / math.sqrt(2)
        assert max_distance(a, want) < 1e-12
        with oracle.OracleState(n, "single") as s:
            s.apply_ops(lower_with_oracle(library.qft(n)), 4, 14)
            assert max_distance(s.amplitudes(), np.full(1 << n, 2 ** (-n / 2))) < 1e-6


def test_fusion_cuts_kernel_calls():
    # tests/simulator.rs:45-57
    gates = library.random_circuit(20, 20, 1)
    with oracle.OracleState(20, "single") as s:
        st = s.apply_ops(lower_with_oracle(gates), 4, 14)
    assert st["gates"] > 500 and st["kernels"] * 4 < st["gates"]


def test_cache_blocking_matches_reference():
    # tests/simulator.rs:164-193: tiny regions force stages, global controls,
    # global diagonals and relabelings.
    for seed in range(30):
        n = 8 + seed % 7
        gates = library.random_gate_circuit(n, 120, 100 + seed)
        expected = oracle_statevector(n, gates)
        ops = lower_with_oracle(gates)
        for region in range(1, 5):
            for width in (1, 3, 4):
                with oracle.OracleState(n, "double") as s:
                    st = s.apply_ops(ops, width, region)
                    assert max_distance(s.amplitudes(), expected) < 1e-11, (seed, region, width)
                    assert st["passes"] > 0
                with oracle.OracleState(n, "single") as s:
                    s.apply_ops(ops, width, region)
                    assert max_distance(s.amplitudes(), expected) < 1e-5


def _run_terminal(n, gates, measures, shots, precision, seed, region=14, num_clbits=None):
    """simulator.rs:360-381: apply, sample with the uniforms, count."""
    with oracle.OracleState(n, precision) as s:
        s.apply_ops(lower_with_oracle(gates), 4, region)
        u = rng(seed).random(shots)
        out = s.sample(u)
    return oracle.count_outcomes(out, measures, num_clbits or n)


def test_sampling_after_relabeling():
    # tests/simulator.rs:195-214
    gates = [("x", (9,), ()), ("x", (7,), ()), ("h", (0,), ()), ("cx", (0, 8), ())]
    counts = _run_terminal(10, gates, [(q, q) for q in range(10)], 1000, "single", 1, region=1)
    assert sorted(counts) == ["1010000000", "1110000001"]


def test_layout_is_permuted_by_blocking():
    gates = [("x", (9,), ()), ("h", (0,), ()), ("cx", (0, 8), ())]
    with oracle.OracleState(10, "single") as s:
        st = s.apply_ops(lower_with_oracle(gates), 4, 1)
        assert st["swaps"] > 0
        assert s.layout() != list(range(10))
        want = oracle_statevector(10, gates)
        assert max_distance(s.amplitudes(), want) < 1e-6


def test_bell_sampling():
    # tests/simulator.rs:59-78: within 5 sigma
    counts = _run_terminal(2, [("h", (0,), ()), ("cx", (0, 1), ())], [(0, 0), (1, 1)], 100_000, "single", 5)
    assert set(counts) == {"00", "11"}
    assert abs(counts["00"] - 50_000) < 800


def test_sampling_matches_probabilities():
    # tests/simulator.rs:80-111
    gates = [("ry", (0,), (1.0,)), ("ry", (1,), (2.0,)), ("cx", (1, 2), ())]
    shots = 200_000
    counts = _run_terminal(3, gates, [(q, q) for q in range(3)], shots, "double", 9)
    p = np.abs(oracle_statevector(3, gates)) ** 2
    for i, pi in enumerate(p):
        obs = counts.get(format(i, "03b"), 0) / shots
        assert abs(obs - pi) <= 5 * math.sqrt(pi * (1 - pi) / shots) + 1e-9


def _walk(probs, uniforms, chunk):
    """measure.rs:104-174 from its definition, in pure Python floats, for
    tiny states (the chunked sequential sums, prefix, strict-< walk, skip of
    p == 0 and the last-nonzero fill)."""
    sums = []
    for c in range(0, len(probs), chunk):
        s = 0.0
        for p in probs[c:c + chunk]:
            s += p
        sums.append(s)
    prefix = [0.0]
    for s in sums:
        prefix.append(prefix[-1] + s)
    total = prefix[-1]
    targets = sorted(u * total for u in uniforms)
    out = []
    for ci in range(len(sums)):
        lo = sum(1 for t in targets if t < prefix[ci])
        hi = min(sum(1 for t in targets if t < prefix[ci + 1]), len(targets))
        if lo >= hi:
            continue
        cum, shot, last = prefix[ci], lo, ci * chunk
        found = []
        for j in range(ci * chunk, min((ci + 1) * chunk, len(probs))):
            p = probs[j]
            if p == 0.0:
                continue
            last = j
            cum += p
            while shot < hi and targets[shot] < cum:
                found.append(j)
                shot += 1
            if shot == hi:
                break
        found += [last] * (hi - lo - len(found))
        out += [min(f, len(probs) - 1) for f in found]
    return out


@pytest.mark.parametrize("precision,chunk", [("single", 8 * 4096), ("double", 4 * 4096)])
def test_sampler_walk_definition(precision, chunk):
    """Bit-exact agreement of the oracle's sampler with the walk's definition,
    on states spanning several chunks, with zero amplitudes and uniforms at
    chunk boundaries."""
    r = rng(21)
    n = 17
    with oracle.OracleState(n, precision) as s:
        a = random_state(r, n)
        a[r.integers(0, 1 << n, 5000)] = 0
        a[: chunk // 2] = 0  # a run of zeros at the start of chunk 0
        s.set_amplitudes(a)
        back = s.amplitudes()
        probs = [float(x.real) ** 2 + float(x.imag) ** 2 for x in back]
        u = r.random(3000)
        u[:3] = [0.0, 0.5, 0.999999999]
        got = list(map(int, s.sample(u)))
    assert got == _walk(probs, list(u), chunk)


def test_mid_circuit_measure_and_reset():
    # tests/simulator.rs:113-139 through the oracle's measure/reset with
    # uniforms from a seeded stream.
    r = rng(3)
    counts = {}
    for _ in range(2000):
        with oracle.OracleState(1, "double") as s:
            clbits = [0, 0, 0]
            s.apply(oracle.gate_matrix("x"), [0])
            clbits[0] = s.measure(0, r.random())
            s.reset(0, r.random())
            clbits[1] = s.measure(0, r.random())
            s.apply(oracle.gate_matrix("h"), [0])
            clbits[2] = s.measure(0, r.random())
        key = "".join(str(b) for b in reversed(clbits))
        counts[key] = counts.get(key, 0) + 1
    assert sorted(counts) == ["001", "101"]
    assert abs(counts["101"] - 1000) < 200


def test_probability_one_and_norm():
    r = rng(4)
    n = 16
    a = random_state(r, n)
    with oracle.OracleState(n, "double") as s:
        s.set_amplitudes(a)
        assert abs(s.norm_sqr() - 1) < 1e-12
        p = np.abs(a) ** 2
        idx = np.arange(1 << n)
        for q in (0, 1, 5, 15):
            assert abs(s.probability_one(q) - p[(idx >> q) & 1 == 1].sum()) < 1e-12
        assert max_distance(s.probabilities(), p) < 1e-15
