"""Parity of the HIP engine (through its C ABI) with the oracle.

Mirrors the reference's integration tests (tests/kernels.rs,
tests/simulator.rs) at the same shapes: random single ops against the FP64
applier, circuits at every fusion width, tiny tiles that force many stages,
non-tile controls and diagonals, sampling bit-exact against the restated
two-pass sampler on the same amplitudes and uniforms.
Tolerances: complex128 1e-12 per op / 1e-11 per circuit, complex64 2e-6 per
op / 1e-5 per circuit (max-abs), as the reference's tests use; circuits also
meet the relative L2 bounds of 1e-12 (complex128) and 1e-5 (complex64).
"""
import math
import sys

import numpy as np
import pytest

from helpers import (lower_with_oracle, max_distance, oracle_statevector, random_diagonal, random_qubits,
                     random_state, random_unitary, rel_l2, rng)
from oracle import oracle
from qevedo.simulator import NativeError, StateVector, library, lower_gates
from qevedo.simulator import _native

pytestmark = pytest.mark.gpu

PREC = [("single", 2e-6, 1e-5), ("double", 1e-12, 1e-11)]


def _note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("precision,tol,_", PREC)
def test_single_ops_match_reference(precision, tol, _):
    """tests/kernels.rs:12-62: 400 random ops, n in [1,11], k in [1,6],
    up to two controls, 30% diagonal."""
    _note(f"single ops {precision}")
    r = rng(1 if precision == "double" else 2)
    for case in range(400):
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
        bits = sum((1 if v else 0) << i for i, (_, v) in enumerate(controls))
        full = oracle.matrix_controlled(m, len(controls), bits) if controls else m
        oracle.reference_apply(expected, full, [q for q, _ in controls] + qubits)
        with StateVector(n, precision) as s:
            s.set_amplitudes(initial)
            s.apply(m, qubits, controls)
            d = max_distance(s.amplitudes(), expected)
        assert d < tol, f"case {case}: n={n} q={qubits} c={controls} diag={diagonal}: {d}"


def test_large_state_every_width():
    # tests/kernels.rs:74-92, extended to k = 6.
    _note("large state")
    r = rng(3)
    n = 18
    initial = random_state(r, n)
    for k in range(1, 7):
        q = random_qubits(r, n, k)
        m = random_unitary(r, k)
        expected = initial.copy()
        oracle.reference_apply(expected, m, q)
        with StateVector(n, "double") as s:
            s.set_amplitudes(initial)
            s.apply(m, q)
            assert max_distance(s.amplitudes(), expected) < 1e-12, k


@pytest.mark.parametrize("tile_bits", [None, 0])
def test_circuits_every_fusion_width(tile_bits):
    # tests/simulator.rs:8-33, with and without tiles.
    _note(f"fusion widths tile_bits={tile_bits}")
    for seed in range(40):
        n = 1 + seed % 10
        gates = library.random_gate_circuit(n, 60, seed)
        expected = oracle_statevector(n, gates)
        ops = lower_gates(gates)
        for width in range(1, 7):
            with StateVector(n, "double") as s:
                st = s.apply_ops(ops, width, tile_bits)
                got = s.amplitudes()
                assert max_distance(got, expected) < 1e-11, (seed, width)
                assert rel_l2(got, expected) <= 1e-12, (seed, width)
                assert st["kernels"] <= st["gates"]
            with StateVector(n, "single") as s:
                s.apply_ops(ops, width, tile_bits)
                assert max_distance(s.amplitudes(), expected) < 1e-5, (seed, width)


def test_tiny_tiles_force_stages():
    """tests/simulator.rs:164-193: tiles of 2..6 qubits on 8..14-qubit states
    force many stages, non-tile controls and diagonals, and full-state ops."""
    _note("tiny tiles")
    for seed in range(30):
        n = 8 + seed % 7
        gates = library.random_gate_circuit(n, 120, 100 + seed)
        expected = oracle_statevector(n, gates)
        ops = lower_gates(gates)
        for tile in (2, 3, 4, 6):
            for width in (1, 3, 4):
                with StateVector(n, "double") as s:
                    st = s.apply_ops(ops, width, tile)
                    got = s.amplitudes()
                    assert max_distance(got, expected) < 1e-11, (seed, tile, width)
                    assert rel_l2(got, expected) <= 1e-12, (seed, tile, width)
                    assert st["passes"] > 0
                with StateVector(n, "single") as s:
                    s.apply_ops(ops, width, tile)
                    assert max_distance(s.amplitudes(), expected) < 1e-5, (seed, tile, width)


def test_library_circuits():
    # tests/simulator.rs:35-57
    _note("library circuits")
    for n, gates in [(9, library.ghz(9)), (9, library.qft(9)), (10, library.random_circuit(10, 12, 7))]:
        expected = oracle_statevector(n, gates)
        with StateVector(n, "double") as s:
            st = s.apply_ops(lower_gates(gates), 4, 0)
            got = s.amplitudes()
            assert max_distance(got, expected) < 1e-12
            assert rel_l2(got, expected) <= 1e-12
            assert st["kernels"] < st["gates"]
    # Fusion cuts kernel calls on the fused-op (LDS tile) path; the register
    # tiles apply unfused gates, so this is checked with 13-qubit LDS tiles.
    gates = library.random_circuit(20, 20, 1)
    with StateVector(20, "single") as s:
        st = s.apply_ops(lower_gates(gates), 4, 13)
    assert st["gates"] > 500 and st["kernels"] * 4 < st["gates"]


@pytest.mark.parametrize("n", [20, 22])
def test_random_circuit_against_blocked_fp64(n):
    """The benchmark circuit family at 20-22 qubits: complex64 against the
    oracle's FP64 engine (itself pinned against the naive applier)."""
    _note(f"random circuit n={n}")
    gates = library.random_circuit(n, 20, 42)
    with oracle.OracleState(n, "double") as o:
        o.apply_ops(lower_with_oracle(gates), 4, 14)
        want = o.amplitudes()
    with StateVector(n, "single") as s:
        s.apply_ops(lower_gates(gates), 4, 0)
        got = s.amplitudes()
    assert rel_l2(got, want) <= 1e-5
    with StateVector(n, "double") as s:
        s.apply_ops(lower_gates(gates), 4, 0)
        got = s.amplitudes()
        assert max_distance(got, want) < 1e-11
        assert rel_l2(got, want) <= 1e-12


@pytest.mark.parametrize("precision,tol", [("single", 1e-5), ("double", 1e-11)])
def test_register_tiles_every_gate_kind(precision, tol):
    """States of at least one register tile (14 qubits in complex64, 13 in
    complex128) run the register-tile kernel: random circuits over the whole
    gate set exercise register-bit and lane-bit 1-qubit gates, controls on
    lane / wave / register / non-tile bits, diagonals, LDS exchanges and the
    in-LDS path for SWAP, CCX and CSWAP."""
    _note(f"register tiles {precision}")
    for seed in range(12):
        n = (14 if precision == "single" else 13) + seed % 4
        gates = library.random_gate_circuit(n, 150, 500 + seed)
        expected = oracle_statevector(n, gates)
        with StateVector(n, precision) as s:
            st = s.apply_ops(lower_gates(gates))
            got = s.amplitudes()
        assert max_distance(got, expected) < tol, (seed, n, max_distance(got, expected))
        assert rel_l2(got, expected) <= (1e-5 if precision == "single" else 1e-12), (seed, n)
        assert st["stages"] >= 1


@pytest.mark.parametrize("precision", ["single", "double"])
def test_register_tiles_structured_circuits(precision):
    """Kicked Ising on a heavy-hex graph, a hardware-efficient ansatz and the
    random circuit family on register tiles, against the FP64 oracle."""
    _note(f"structured circuits {precision}")
    n = 16
    # Layers of RX and RY on every qubit (off-diagonals purely imaginary or
    # real after the unit-form split) between CZ ladders, so the real and
    # imaginary unit forms run on every register and lane bit.
    rxry = []
    for layer in range(6):
        for q in range(n):
            rxry.append(("rx" if (q + layer) % 2 else "ry", (q,), (0.3 + 0.17 * q + 0.5 * layer,)))
        rxry += [("cz", (q, q + 1), ()) for q in range(layer % 2, n - 1, 2)]
    circuits = [
        rxry,
        library.kicked_ising(n, library.heavy_hex_edges(n), 5, 0.9),
        library.hardware_efficient_ansatz(n, 3, 4),
        library.random_circuit(n, 14, 11),
        library.qft(n),
    ]
    tol = 1e-5 if precision == "single" else 1e-11
    for gates in circuits:
        expected = oracle_statevector(n, gates)
        with StateVector(n, precision) as s:
            s.apply_ops(lower_gates(gates))
            got = s.amplitudes()
            assert max_distance(got, expected) < tol
            assert rel_l2(got, expected) <= (1e-5 if precision == "single" else 1e-12)


def test_qft_closed_form():
    _note("qft closed form")
    for n in (12, 20):
        with StateVector(n, "single") as s:
            s.apply_ops(lower_gates(library.qft(n)), 4, 0)
            assert max_distance(s.amplitudes(), np.full(1 << n, 2 ** (-n / 2))) < 1e-6


def test_tile_choice_and_determinism():
    """The same circuit twice gives bit-identical amplitudes; other tile
    sizes stay within the FP64 bound (the analogue of the thread-count
    reproducibility test, python/tests/test_simulator.py:167-176)."""
    _note("determinism")
    n = 16
    gates = library.random_circuit(n, 12, 3)
    ops = lower_gates(gates)
    runs = []
    for tile in (0, 0, 8, 11, 13):
        with StateVector(n, "single") as s:
            s.apply_ops(ops, 4, tile)
            runs.append(s.amplitudes())
    assert np.array_equal(runs[0], runs[1])
    want = oracle_statevector(n, gates)
    for a in runs:
        assert max_distance(a, want) < 1e-5


def _oracle_samples(amps, precision, uniforms):
    n = int(round(math.log2(amps.size)))
    with oracle.OracleState(n, precision) as o:
        o.set_amplitudes(amps)
        return o.sample(uniforms)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("n", [1, 3, 5, 15, 17, 19])
def test_sampling_bit_exact(precision, n):
    """measure.rs:104-174 on the same amplitudes and uniforms: identical
    outcome vectors (multi-chunk states, zero runs, boundary uniforms)."""
    _note(f"sampling {precision} n={n}")
    r = rng(100 + n)
    a = random_state(r, n)
    a[r.integers(0, 1 << n, max(1, (1 << n) // 16))] = 0
    a[: (1 << n) // 8] = 0
    with StateVector(n, precision) as s:
        s.set_amplitudes(a)
        stored = s.amplitudes()  # the values the GPU holds, in their precision
        u = r.random(20000)
        u[:4] = [0.0, 0.25, 0.5, 1 - 2 ** -53]
        got = s.sample(u)
    want = _oracle_samples(stored, precision, u)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sampling_bit_exact_peaked(precision):
    """A few amplitudes carry almost all the mass: runs of shots land on the
    same index, back to back, in several chunks (and in the last amplitude
    of a chunk), so a walk takes many hits in a row without moving on."""
    _note(f"sampling peaked {precision}")
    n = 17
    r = rng(61)
    a = random_state(r, n) * 1e-3
    chunk = 1 << (15 if precision == "single" else 14)
    for k, idx in enumerate([5, chunk - 1, chunk, 2 * chunk + 77, (1 << n) - 1]):
        a[idx] = 0.4 + 0.1 * k
    with StateVector(n, precision) as s:
        s.set_amplitudes(a / np.linalg.norm(a))
        stored = s.amplitudes()
        u = r.random(200_000)
        got = s.sample(u)
    want = _oracle_samples(stored, precision, u)
    assert np.array_equal(got, want)
    assert np.max(np.bincount(want.astype(np.int64))) > 10_000


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sampling_rounding_leftovers(precision):
    """measure.rs:167-168: chunk 1 holds probabilities of about 1e-20 each
    after a chunk of total 1. Its own sum is positive, so a target of
    exactly 1.0 falls in it, yet a running sum starting at 1.0 never moves:
    such shots go to the chunk's last nonzero index."""
    _note(f"sampling leftovers {precision}")
    n = 17
    chunk = 1 << (15 if precision == "single" else 14)
    a = np.zeros(1 << n, dtype=np.complex128)
    a[3] = 1.0
    a[chunk:2 * chunk] = 1e-10
    a[chunk + 100] = 0
    u = np.array([0.0, 0.3, 1 - 2 ** -52, 1 - 2 ** -53, 0.999999999999])
    with StateVector(n, precision) as s:
        s.set_amplitudes(a)
        stored = s.amplitudes()
        got = s.sample(u)
    want = _oracle_samples(stored, precision, u)
    assert np.array_equal(got, want)
    assert np.sum(want == 2 * chunk - 1) == 2


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sampling_one_shot(precision):
    """One uniform: one chunk walked by one thread."""
    n = 16
    r = rng(66)
    a = random_state(r, n)
    with StateVector(n, precision) as s:
        s.set_amplitudes(a)
        stored = s.amplitudes()
        for u in (0.0, 0.37, 1 - 2 ** -53):
            got = s.sample(np.array([u]))
            assert np.array_equal(got, _oracle_samples(stored, precision, np.array([u]))), u


def test_sampling_more_shots_than_one_grid():
    """5 * 10^6 shots: more than one thread per shot of the placement
    kernel's grid (16384 workgroups of 256), so its threads stride."""
    _note("sampling 5e6 shots")
    n = 14
    r = rng(71)
    with StateVector(n, "single") as s:
        s.set_amplitudes(random_state(r, n))
        stored = s.amplitudes()
        u = r.random(5_000_000)
        got = s.sample(u)
    assert np.array_equal(got, _oracle_samples(stored, "single", u))


def test_sampling_bit_exact_many_chunks():
    """27 qubits in complex64: 4096 chunks, so the prefix over chunk sums
    runs over more than one of the device's staging tiles."""
    _note("sampling 27 qubits")
    n = 27
    with StateVector(n, "single") as s:
        s.apply_ops(lower_gates(library.random_circuit(n, 3, 12)))
        u = rng(27).random(50_000)
        got = s.sample(u)
        amps = s.amplitudes()
    assert np.array_equal(got, _oracle_samples(amps, "single", u))


def test_sampling_after_circuit_matches_oracle():
    _note("sampling after circuit")
    n = 18
    gates = library.random_circuit(n, 10, 5)
    with StateVector(n, "single") as s:
        s.apply_ops(lower_gates(gates), 4, 0)
        amps = s.amplitudes()
        u = rng(9).random(100_000)
        got = s.sample(u)
    assert np.array_equal(got, _oracle_samples(amps, "single", u))


def test_sampling_into_a_given_array():
    _note("sampling into out")
    n = 16
    with StateVector(n, "single") as s:
        s.apply_ops(lower_gates(library.random_circuit(n, 6, 2)))
        u = rng(3).random(30_000)
        want = s.sample(u)
        buf = np.zeros(u.size, dtype=np.uint64)
        assert s.sample(u, out=buf) is buf
        assert np.array_equal(buf, want)
        for bad in [np.zeros(u.size, dtype=np.int64), np.zeros(u.size - 1, dtype=np.uint64),
                    np.zeros((u.size, 2), dtype=np.uint64)[:, 0]]:
            with pytest.raises(ValueError):
                s.sample(u, out=bad)


def test_bell_and_distribution_sampling():
    # tests/simulator.rs:59-111
    _note("bell sampling")
    with StateVector(2, "single") as s:
        s.apply_ops(lower_gates([("h", (0,), ()), ("cx", (0, 1), ())]))
        out = s.sample(rng(5).random(100_000))
    counts = np.bincount(out.astype(np.int64), minlength=4)
    assert counts[1] == 0 and counts[2] == 0 and abs(counts[0] - 50_000) < 800
    gates = [("ry", (0,), (1.0,)), ("ry", (1,), (2.0,)), ("cx", (1, 2), ())]
    shots = 200_000
    with StateVector(3, "double") as s:
        s.apply_ops(lower_gates(gates))
        out = s.sample(rng(9).random(shots))
    p = np.abs(oracle_statevector(3, gates)) ** 2
    obs = np.bincount(out.astype(np.int64), minlength=8) / shots
    for i in range(8):
        assert abs(obs[i] - p[i]) <= 5 * math.sqrt(p[i] * (1 - p[i]) / shots) + 1e-9


def test_reductions():
    _note("reductions")
    r = rng(4)
    for n, precision in [(1, "double"), (5, "double"), (16, "single"), (20, "double")]:
        a = random_state(r, n)
        with StateVector(n, precision) as s:
            s.set_amplitudes(a)
            stored = s.amplitudes()
            p = np.abs(stored) ** 2
            idx = np.arange(1 << n)
            assert abs(s.norm_sqr() - p.sum()) < 1e-12
            for q in range(n):
                assert abs(s.probability_one(q) - p[((idx >> q) & 1) == 1].sum()) < 1e-12
            z, zz = s.expect_z()
            for q in range(n):
                bit = (idx >> q) & 1
                assert abs(z[q] - np.sum(p * (1 - 2 * bit))) < 1e-12
                if q + 1 < n:
                    par = bit ^ ((idx >> (q + 1)) & 1)
                    assert abs(zz[q] - np.sum(p * (1 - 2 * par))) < 1e-12
            assert max_distance(s.probabilities(), p) < 1e-15
            s.scale(2.0)
            assert abs(s.norm_sqr() - 4 * p.sum()) < 1e-11


def test_measure_and_reset_follow_the_uniform():
    """measure.rs:68-87 with caller uniforms: outcome = u < P(1), collapse
    and renormalise; compared with the oracle step by step."""
    _note("measure/reset")
    r = rng(12)
    n = 10
    gates = library.random_circuit(n, 6, 2)
    for precision in ("single", "double"):
        with StateVector(n, precision) as s, oracle.OracleState(n, precision) as o:
            s.apply_ops(lower_gates(gates))
            o.set_amplitudes(s.amplitudes())
            for step in range(6):
                q = int(r.integers(0, n))
                u = float(r.random())
                p1 = o.probability_one(q)
                if abs(u - p1) < 1e-6:
                    continue
                if step % 2:
                    s.reset(q, u)
                    o.reset(q, u)
                else:
                    assert s.measure(q, u) == o.measure(q, u)
                tol = 2e-6 if precision == "single" else 1e-12
                assert max_distance(s.amplitudes(), o.amplitudes()) < tol
    # tests/simulator.rs:113-139: x; measure; reset; measure; h; measure
    counts = {}
    rr = rng(3)
    for _ in range(2000):
        with StateVector(1, "double") as s:
            c = [0, 0, 0]
            s.apply(_native.gate_matrix("x"), [0])
            c[0] = s.measure(0, rr.random())
            s.reset(0, rr.random())
            c[1] = s.measure(0, rr.random())
            s.apply(_native.gate_matrix("h"), [0])
            c[2] = s.measure(0, rr.random())
        key = "".join(str(b) for b in reversed(c))
        counts[key] = counts.get(key, 0) + 1
    assert sorted(counts) == ["001", "101"] and abs(counts["101"] - 1000) < 200


def test_errors_are_status_codes():
    _note("errors")
    with StateVector(4) as s:
        with pytest.raises(NativeError) as e:
            s.apply(np.eye(2), [4])
        assert e.value.code == _native.QVD_EINVAL and "out of range" in str(e.value)
        with pytest.raises(NativeError) as e:
            s.apply(np.eye(4), [1, 1])
        assert e.value.code == _native.QVD_EINVAL
        with pytest.raises(NativeError) as e:
            s.set_amplitudes(np.zeros(8))
        assert e.value.code == _native.QVD_EINVAL
    with StateVector(8) as s:
        with pytest.raises(NativeError) as e:
            s.apply(np.eye(128), list(range(7)))
        assert e.value.code == _native.QVD_EUNSUPPORTED
    with pytest.raises(NativeError) as e:
        StateVector(35)
    assert e.value.code == _native.QVD_EINVAL


def test_state_vector_fits_device_memory():
    """state_vector_fits (simulator.rs:254-259) against the device's HBM: a
    state fits when it takes at most 80 % of it, and run() refuses one that
    does not instead of falling back to another backend."""
    _note("fits")
    from qevedo.simulator import Circuit, choose_backend, run
    fits, sb, db = _native.state_vector_fits(30, "single")
    assert fits and sb == 8 << 30 and db > sb
    for n in range(20, 40):
        for prec, b in (("single", 8), ("double", 16)):
            fits, sb, db2 = _native.state_vector_fits(n, prec)
            assert db2 == db and sb == b << n
            assert fits == (sb <= 0.8 * db)
    big = Circuit(40)
    big.h(0)
    with pytest.raises(ValueError, match="does not fit"):
        choose_backend(big)
    with pytest.raises(ValueError, match="does not fit"):
        run(big, shots=1)
    assert choose_backend(Circuit(12)) == "statevector"
