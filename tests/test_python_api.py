"""The qevedo.simulator Python surface, mirroring the state-vector tests of
the reference's python/tests/test_simulator.py (:77-176). CPU tests cover
circuit building, OpenQASM and validation; GPU tests run circuits."""
import cmath
import math
import random

import numpy as np
import pytest

from helpers import oracle_statevector
from qevedo.simulator import Circuit, run, statevector


def test_qasm_round_trip_and_instructions():
    plain = Circuit(3, 3).h(0).cx(0, 1).rz(0.25, 2).measure(1, 2)
    again = Circuit.from_qasm(plain.to_qasm())
    assert again.instructions() == plain.instructions()
    assert plain.instructions()[2] == ("rz", [2], [0.25])
    assert plain.instructions()[3] == ("measure", [1], [2.0])
    with pytest.raises(ValueError, match="unitaries"):
        Circuit(2).unitary(np.eye(4), [0, 1]).to_qasm()


def test_qasm_expressions_and_broadcast():
    c = Circuit.from_qasm('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[3];\ncreg c[3];\n'
                          'h q;\nu3(pi/2, -pi/4, pi/4) q[1];\ncu1(pi/8) q[0],q[2];\nmeasure q -> c;\n')
    ins = c.instructions()
    assert [i[0] for i in ins[:3]] == ["h", "h", "h"]
    assert ins[3][0] == "u3" and ins[3][2] == pytest.approx([math.pi / 2, -math.pi / 4, math.pi / 4])
    assert ins[4] == ("cu1", [0, 2], [math.pi / 8])
    assert [i[0] for i in ins[5:]] == ["measure"] * 3


def test_errors():
    with pytest.raises(ValueError, match="out of range"):
        Circuit(2).h(2)
    with pytest.raises(ValueError, match="same qubit"):
        Circuit(2).cx(1, 1)
    with pytest.raises(ValueError, match="line 3"):
        Circuit.from_qasm('OPENQASM 2.0;\nqreg q[1];\nfoo q[0];')
    with pytest.raises(ValueError, match="unknown backend"):
        run(Circuit(1).measure_all(), backend="quantum")
    with pytest.raises(ValueError, match="unknown gate"):
        Circuit(1).gate("nope", [0])
    with pytest.raises(ValueError, match="not unitary"):
        Circuit(1).unitary([[1, 1], [0, 1]], [0])


def probabilities(circuit: Circuit) -> np.ndarray:
    gates = [(name, q, p) for kind, name, q, p in circuit._ins if kind == "gate"]
    return np.abs(oracle_statevector(circuit.num_qubits, gates)) ** 2


@pytest.mark.gpu
def test_statevector_matches_known_states():
    ghz = statevector(Circuit(3).h(0).cx(0, 1).cx(1, 2))
    r = 1 / math.sqrt(2)
    assert abs(ghz[0] - r) < 1e-12 and abs(ghz[7] - r) < 1e-12
    assert sum(abs(a) ** 2 for a in ghz) == pytest.approx(1.0)
    phased = statevector(Circuit(1).h(0).t(0))
    assert abs(phased[1] - r * cmath.exp(1j * math.pi / 4)) < 1e-12


@pytest.mark.gpu
def test_samples_the_exact_distribution():
    rng = random.Random(7)
    circuit = Circuit(5)
    for _ in range(40):
        a, b = rng.sample(range(5), 2)
        choice = rng.randrange(6)
        if choice == 0:
            circuit.h(a)
        elif choice == 1:
            circuit.t(a)
        elif choice == 2:
            circuit.rx(rng.uniform(0, 6.28), a)
        elif choice == 3:
            circuit.cx(a, b)
        elif choice == 4:
            circuit.cp(rng.uniform(0, 6.28), a, b)
        else:
            circuit.ccx(a, b, rng.choice([q for q in range(5) if q not in (a, b)]))
    p = probabilities(circuit)
    circuit.measure_all()
    shots = 20000
    result = run(circuit, shots=shots, backend="statevector", seed=5)
    assert result.backend == "statevector"
    for i, pi in enumerate(p):
        observed = result.counts.get(format(i, "05b"), 0) / shots
        assert abs(observed - pi) <= 5 * math.sqrt(pi * (1 - pi) / shots) + 3 / shots


@pytest.mark.gpu
def test_unitary_from_nested_lists():
    sx = [[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]
    a = statevector(Circuit(1).unitary(sx, [0]))
    b = statevector(Circuit(1).sx(0))
    assert all(abs(x - y) < 1e-12 for x, y in zip(a, b))


@pytest.mark.gpu
def test_reproducibility():
    circuit = Circuit(6)
    for q in range(6):
        circuit.h(q).t(q)
    for q in range(5):
        circuit.cx(q, q + 1)
    circuit.measure_all()
    a = run(circuit, shots=3000, seed=9, backend="statevector", threads=1)
    b = run(circuit, shots=3000, seed=9, backend="statevector", threads=4)
    assert a.counts == b.counts


@pytest.mark.gpu
def test_mid_circuit_measurement_reset_and_feedback():
    # tests/simulator.rs:113-139: x; measure; reset; measure; h; measure
    c = Circuit(1).x(0).measure(0, 0).reset(0).measure(0, 1).h(0).measure(0, 2)
    result = run(c, shots=400, seed=3)
    assert sorted(result.counts) == ["001", "101"]
    # feedback: measure a |+> qubit, flip it back to |0> when it read 1
    c = Circuit(1).h(0).measure(0, 0).feedback("X", 0, 0).measure(0, 1)
    result = run(c, shots=300, seed=4)
    assert set(result.counts) <= {"00", "01"} and len(result.counts) == 2


@pytest.mark.gpu
def test_sampling_after_relabeling_keys():
    """tests/simulator.rs:195-214: x(9) x(7) h(0) cx(0, 8), measured: the
    samples are reported in logical order, so the only keys are exactly
    1010000000 and 1110000001; also with tiles of 1-3 qubits, which force a
    stage or full-state pass per op."""
    from qevedo.simulator import StateVector, _count, lower_gates

    c = Circuit(10).x(9).x(7).h(0).cx(0, 8).measure_all()
    assert sorted(run(c, 1000, seed=1).counts) == ["1010000000", "1110000001"]
    gates = [("x", (9,), ()), ("x", (7,), ()), ("h", (0,), ()), ("cx", (0, 8), ())]
    u = np.random.Generator(np.random.PCG64(1)).random(1000)
    for tile in (1, 2, 3):
        with StateVector(10, "single") as s:
            s.apply_ops(lower_gates(gates), 4, tile)
            keys = sorted(_count(s.sample(u), [(q, q) for q in range(10)], 10))
        assert keys == ["1010000000", "1110000001"], tile


@pytest.mark.gpu
def test_teleportation_with_if_statements():
    """tests/qasm.rs:666-698 on the state vector: |+i> teleported from q[0]
    to q[2] with OpenQASM 3 `if`s (classical feedback), the preparation
    undone: q[2] reads 0 in every shot, and the four outcomes of the two
    measured qubits all occur."""
    program = ('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[3] q;\nbit[3] c;\n'
               "h q[0]; s q[0];\nh q[1]; cx q[1], q[2];\ncx q[0], q[1]; h q[0];\n"
               "c[0] = measure q[0];\nc[1] = measure q[1];\n"
               "if (c[1]) x q[2];\nif (c[0]) z q[2];\nsdg q[2]; h q[2];\nc[2] = measure q[2];\n")
    counts = run(Circuit.from_qasm(program), 400, seed=3).counts
    assert len(counts) == 4 and all(k.startswith("0") for k in counts), counts


@pytest.mark.gpu
def test_qasm_programs_with_modifiers_run_on_the_engine():
    """Programs whose gates become unitaries (modifiers, definitions, c3x,
    cu) run through the engine and match the oracle's FP64 applier."""
    from helpers import random_unitary, rng
    from oracle import oracle

    program = ('OPENQASM 3.0;\ninclude "stdgates.inc";\n'
               "gate g(a) x, y { rx(a) x; cz x, y; ry(a / 3) y; }\nqubit[5] q;\n"
               "h q; ctrl @ inv @ g(0.7) q[2], q[0], q[1];\nnegctrl(2) @ ry(1.1) q[3], q[4], q[0];\n"
               "c3x q[0], q[1], q[2], q[4];\ncu(0.3, 0.2, 0.1, 0.5) q[4], q[3];\npow(3) @ sx q[1];\n"
               "rccx q[1], q[3], q[2];\n")
    c = Circuit.from_qasm(program)
    want = np.zeros(32, dtype=np.complex128)
    want[0] = 1
    for kind, name, qubits, params in c._ins:
        m = oracle.gate_matrix(name, params) if kind == "gate" else name
        oracle.reference_apply(want, m, list(qubits))
    got = statevector(c)
    assert np.max(np.abs(got - want)) < 1e-12


def _dynamic_circuit(n, seed, steps=30):
    """Random gates with mid-circuit measures, resets and Pauli feedback."""
    r = random.Random(seed)
    c = Circuit(n, n)
    for _ in range(steps):
        k = r.randrange(8)
        a, b = r.sample(range(n), 2)
        if k == 0:
            c.measure(a, r.randrange(n))
        elif k == 1:
            c.reset(a)
        elif k == 2:
            c.feedback(r.choice("XYZ"), a, r.randrange(n))
        elif k == 3:
            c.cx(a, b)
        elif k == 4:
            c.rx(r.uniform(0, 6.28), a)
        elif k == 5:
            c.ry(r.uniform(0, 6.28), a)
        elif k == 6:
            c.h(a)
        else:
            c.cp(r.uniform(0, 6.28), a, b)
    return c.measure_all()


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed,shots,precision", [(3, 1, 500, "single"), (6, 2, 400, "double"),
                                                    (9, 3, 300, "single"), (15, 4, 200, "single")])
def test_branch_walk_counts_equal_the_shot_by_shot_loop(n, seed, shots, precision):
    """Shots sharing a measurement history share one state: the counts are
    exactly those of running every shot alone with the same uniforms."""
    from qevedo.simulator import _run_branches, _run_shot_by_shot

    c = _dynamic_circuit(n, seed)
    draws = sum(1 for k, *_ in c._ins if k in ("measure", "reset"))
    u = np.random.Generator(np.random.PCG64(seed)).random(shots * draws).reshape(shots, draws)
    branched = _run_branches(c, u, precision)
    serial = _run_shot_by_shot(c, u, precision)
    assert branched.counts == serial.counts
    assert sum(branched.counts.values()) == shots
    # the branch walk runs each gate segment once per branch, not per shot
    assert branched.gates < serial.gates
