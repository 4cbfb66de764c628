"""The OpenQASM 2 / 3 reader on the host, against circuits built directly
(the state-level cases of the reference's tests/qasm.rs:17-387). States are
computed by the oracle's FP64 naive applier (reference.rs:8-28) on the host."""
import math

import numpy as np
import pytest

from helpers import random_state, random_unitary, rng
from oracle import oracle
from qevedo.simulator import Circuit, library
from qevedo.simulator.qasm import QasmError, parse

HEAD2 = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'
HEAD3 = 'OPENQASM 3.0;\ninclude "stdgates.inc";\n'


def state_of(c: Circuit, psi=None) -> np.ndarray:
    n = c.num_qubits
    s = np.zeros(1 << n, dtype=np.complex128)
    if psi is None:
        s[0] = 1
    else:
        s[:] = psi
    for kind, name, qubits, params in c._ins:
        if kind == "gate":
            oracle.reference_apply(s, oracle.gate_matrix(name, params), list(qubits))
        elif kind == "unitary":
            oracle.reference_apply(s, name, list(qubits))
    return s


def test_round_trip_of_written_programs():
    for seed in range(30):
        n = 1 + seed % 6
        c = Circuit(n)
        for name, q, p in library.random_gate_circuit(n, 50, seed):
            c.gate(name, q, p)
        again = parse(c.to_qasm())
        assert len(again) == len(c)
        assert np.max(np.abs(state_of(again) - state_of(c))) < 1e-12, seed


CASES = [("id", "id", (), 1), ("h", "h", (), 1), ("x", "x", (), 1), ("y", "y", (), 1), ("z", "z", (), 1),
         ("s", "s", (), 1), ("sdg", "sdg", (), 1), ("t", "t", (), 1), ("tdg", "tdg", (), 1), ("sx", "sx", (), 1),
         ("sxdg", "sxdg", (), 1), ("rx(0.3)", "rx", (0.3,), 1), ("ry(-1.2)", "ry", (-1.2,), 1),
         ("rz(pi/3)", "rz", (math.pi / 3,), 1), ("p(0.7)", "p", (0.7,), 1), ("u1(0.7)", "p", (0.7,), 1),
         ("u2(0.1, 0.2)", "u", (math.pi / 2, 0.1, 0.2), 1), ("u3(0.1, 0.2, 0.3)", "u", (0.1, 0.2, 0.3), 1),
         ("U(0.1, 0.2, 0.3)", "u", (0.1, 0.2, 0.3), 1), ("u0(5)", "id", (), 1), ("cx", "cx", (), 2),
         ("CX", "cx", (), 2), ("cy", "cy", (), 2), ("cz", "cz", (), 2), ("ch", "ch", (), 2),
         ("cp(0.4)", "cp", (0.4,), 2), ("cu1(0.4)", "cp", (0.4,), 2), ("crx(0.5)", "crx", (0.5,), 2),
         ("cry(0.5)", "cry", (0.5,), 2), ("crz(0.5)", "crz", (0.5,), 2), ("swap", "swap", (), 2),
         ("rzz(0.6)", "rzz", (0.6,), 2), ("rxx(0.6)", "rxx", (0.6,), 2), ("ccx", "ccx", (), 3),
         ("cswap", "cswap", (), 3)]


def controlled(m, count):
    """m with `count` controls in state all-ones on the low index bits, built
    from the definition."""
    dim = m.shape[0]
    out = np.eye(dim << count, dtype=np.complex128)
    state = (1 << count) - 1
    rows = [(r << count) | state for r in range(dim)]
    out[np.ix_(rows, rows)] = m
    return out


def u3(t, p, l):
    return np.array([[math.cos(t / 2), -np.exp(1j * l) * math.sin(t / 2)],
                     [np.exp(1j * p) * math.sin(t / 2), np.exp(1j * (p + l)) * math.cos(t / 2)]])


SX = 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
UNITARY_CASES = [("csx", controlled(SX, 1), 2), ("cu3(0.1, 0.2, 0.3)", controlled(u3(0.1, 0.2, 0.3), 1), 2),
                 ("cu(0.1, 0.2, 0.3, 0.4)", controlled(np.exp(0.4j) * u3(0.1, 0.2, 0.3), 1), 2),
                 ("c3x", controlled(X, 3), 4), ("c3sqrtx", controlled(SX, 3), 4)]


def test_builtin_gates_match_direct_gates():
    """tests/qasm.rs:28-96: each gate applied after a random 4-qubit unitary,
    read from OpenQASM and applied directly, the same state."""
    r = rng(1)
    for call, name, params, arity in CASES:
        psi = random_state(r, 4)
        qubits = [int(q) for q in r.permutation(4)[:arity]]
        program = HEAD2 + "qreg q[4];\n" + f"{call} " + ", ".join(f"q[{q}]" for q in qubits) + ";"
        got = state_of(parse(program), psi)
        want = psi.copy()
        oracle.reference_apply(want, oracle.gate_matrix(name, params), qubits)
        assert np.max(np.abs(got - want)) < 1e-12, call
    for call, m, arity in UNITARY_CASES:
        psi = random_state(r, 4)
        qubits = [int(q) for q in r.permutation(4)[:arity]]
        program = HEAD2 + "qreg q[4];\n" + f"{call} " + ", ".join(f"q[{q}]" for q in qubits) + ";"
        got = state_of(parse(program), psi)
        want = psi.copy()
        oracle.reference_apply(want, m, qubits)
        assert np.max(np.abs(got - want)) < 1e-12, call
    c4x = parse(HEAD2 + "qreg q[5];\nx q[0]; x q[1]; x q[2]; x q[3];\nc4x q[0], q[1], q[2], q[3], q[4];")
    assert abs(state_of(c4x)[0b11111] - 1) < 1e-12


def test_openqasm_3_matches_openqasm_2():
    v2 = HEAD2 + """qreg a[2];
qreg b[1];
creg c[3];
h a[0];
cx a[0], a[1];
u3(0.1, 0.2, 0.3) b[0];
cu1(pi/4) a[1], b[0];
measure a -> c[0:1];
measure b[0] -> c[2];
"""
    v3 = HEAD3 + """qubit[2] a;
qubit b;
bit[3] c;
h a[0];
cx a[0], a[1];
U(0.1, 0.2, 0.3) b;
cp(π/4) a[1], b;
c[0:1] = measure a;
c[2] = measure b;
"""
    c2, c3 = parse(v2), parse(v3)
    assert (c2.num_qubits, c2.num_clbits) == (3, 3)
    assert len(c2) == len(c3)
    assert np.max(np.abs(state_of(c2) - state_of(c3))) < 1e-12
    assert [i for i in c2.instructions() if i[0] == "measure"] == [("measure", [0], [0.0]), ("measure", [1], [1.0]),
                                                                    ("measure", [2], [2.0])]


def test_gate_definitions_and_expressions():
    program = HEAD2 + """gate rot(theta, phi) q { rz(phi) q; ry(theta / 2 + 0.1) q; }
gate pair(a) x, y { rot(a, -a * 2) x; cx x, y; rot(sin(a)^2, cos(a)) y; }
qreg q[2];
pair(0.4) q[0], q[1];
pair(-1.0e-1) q[1], q[0];
"""
    direct = Circuit(2)
    for a, x, y in [(0.4, 0, 1), (-0.1, 1, 0)]:
        direct.rz(-a * 2, x).ry(a / 2 + 0.1, x).cx(x, y)
        direct.rz(math.cos(a), y).ry(math.sin(a) ** 2 / 2 + 0.1, y)
    assert np.max(np.abs(state_of(parse(program)) - state_of(direct))) < 1e-12


@pytest.mark.parametrize("with_,without", [
    ("ctrl @ x q[0], q[1];", "cx q[0], q[1];"),
    ("inv @ s q[1];", "sdg q[1];"),
    ("pow(2) @ s q[0];", "z q[0];"),
    ("pow(-1) @ t q[0];", "tdg q[0];"),
    ("ctrl(2) @ x q[0], q[1], q[2];", "ccx q[0], q[1], q[2];"),
    ("negctrl @ x q[2], q[0];", "x q[2]; cx q[2], q[0]; x q[2];"),
    ("ctrl @ inv @ rz(0.3) q[1], q[2];", "crz(-0.3) q[1], q[2];"),
    ("ctrl @ swap q[0], q[1], q[2];", "cswap q[0], q[1], q[2];"),
])
def test_modifiers(with_, without):
    """tests/qasm.rs:171-214."""
    r = rng(2)
    psi = random_state(r, 3)
    head = HEAD3 + "qubit[3] q;\n"
    assert np.max(np.abs(state_of(parse(head + with_), psi) - state_of(parse(head + without), psi))) < 1e-12


def test_modifiers_on_a_defined_gate():
    c = parse(HEAD3 + "gate g(a) x, y { rx(a) x; cz x, y; }\nqubit[3] q;\nctrl @ inv @ g(0.7) q[2], q[0], q[1];")
    (kind, m, qubits, _), = c._ins
    assert kind == "unitary" and qubits == (2, 0, 1)
    # g = cz . rx(0.7) on (x, y), x the low bit; expected: controlled(g^dagger)
    # with the control on the low bit
    rx = oracle.gate_matrix("rx", (0.7,))
    g = np.diag([1, 1, 1, -1]).astype(np.complex128) @ np.kron(np.eye(2), rx)
    assert np.max(np.abs(m - controlled(g.conj().T, 1))) < 1e-12


def test_broadcasting_ranges_resets_and_barriers():
    c = parse("OPENQASM 3;\nqubit[4] q;\nbit[4] c;\nh q;\ncx q[0:1], q[2:3];\nbarrier;\nreset q[0:2:2];\nc = measure q;")
    kinds = [(k, n, q, p) for k, n, q, p in c._ins]
    assert kinds[:4] == [("gate", "h", (q,), ()) for q in range(4)]
    assert kinds[4:6] == [("gate", "cx", (0, 2), ()), ("gate", "cx", (1, 3), ())]
    assert kinds[6][0] == "barrier"
    assert kinds[7:9] == [("reset", None, (0,), ()), ("reset", None, (2,), ())]
    assert kinds[9:] == [("measure", None, (q,), (q,)) for q in range(4)]


@pytest.mark.parametrize("program,line,message", [
    ("OPENQASM 2.0;\nqreg q[2];\nfoo q[0];", 3, "unknown gate `foo`"),
    ("OPENQASM 2.0;\nqreg q[2];\nh q[5];", 3, "out of range"),
    ("OPENQASM 2.0;\nqreg q[2];\ncx q[0];", 3, "acts on 2 qubits"),
    ("OPENQASM 2.0;\nqreg q[2];\nrx q[0];", 3, "takes 1 parameters"),
    ("OPENQASM 2.0;\nqreg q[2];\ncreg c[2];\nif (c == 1) x q[0];", 4, "a register of 2 bits, is not supported"),
    ("OPENQASM 3.0;\nqubit[2] q;\nbit[1] c;\nif (c[0]) h q[0];", 4, "`h` inside `if` is not supported"),
    ("OPENQASM 3.0;\nqubit[2] q;\nbit[1] c;\nif (c[0]) { x q[0];\n measure q[1] -> c[0]; }", 5,
     "`measure` inside `if` is not supported"),
    ("OPENQASM 3.0;\nqubit[2] q;\nbit[1] c;\nif (c[0] == 2) x q[0];", 4, "only be compared with 0, 1, true or false"),
    ("OPENQASM 3.0;\nqubit[2] q;\nbit[1] c;\nelse x q[0];", 4, "`else` is not supported"),
    ("OPENQASM 3.0;\nqubit[2] q;\nfor int i in [0:1] { h q[i]; }", 3, "`for` is not supported"),
    ("OPENQASM 2.0;\nqreg q[2];\ncx q[0], q[0];", 3, "same qubit twice"),
    ('OPENQASM 2.0;\ninclude "mylib.inc";', 2, "only qelib1.inc and stdgates.inc"),
    ("OPENQASM 2.0;\nqreg q[2];\nh r[0];", 3, "unknown qubit register `r`"),
    ("OPENQASM 2.0;\nqreg q[1];\nrz(theta) q[0];", 3, "unknown parameter `theta`"),
])
def test_errors_point_at_the_problem(program, line, message):
    """tests/qasm.rs:273-344."""
    with pytest.raises(QasmError) as e:
        parse(program)
    assert e.value.line == line, str(e.value)
    assert message in e.value.message, str(e.value)


def test_relative_phase_toffolis_follow_qelib1_and_can_be_redefined():
    c = parse(HEAD2 + "qreg q[3];\nrccx q[0], q[1], q[2];")
    u = np.stack([state_of(c, np.eye(8)[k]) for k in range(8)], axis=1)
    ccx = oracle.gate_matrix("ccx", ())
    # the Toffoli up to relative phases: same support, unit-modulus entries
    assert np.allclose(np.abs(u), np.abs(ccx), atol=1e-12)
    assert not np.allclose(u, ccx, atol=1e-6)
    c3 = parse(HEAD2 + "qreg q[4];\nrc3x q[0], q[1], q[2], q[3];")
    u3_ = np.stack([state_of(c3, np.eye(16)[k]) for k in range(16)], axis=1)
    assert np.allclose(np.abs(u3_), np.abs(controlled(X, 3)), atol=1e-12)
    redefined = parse(HEAD2 + "gate rccx a, b, c { ccx a, b, c; }\nqreg q[3];\nrccx q[0], q[1], q[2];")
    assert [i[1] for i in redefined._ins] == ["ccx"]


def test_conditional_paulis_become_feedback():
    """src/qasm.rs:636-684: the body for 0 runs always, the product of both
    bodies where the bit reads 1 runs as feedback."""
    head = HEAD3 + "qubit[2] q;\nbit[2] c;\n"
    c = parse(head + "if (c[1]) x q[0];")
    assert c._ins == [("feedback", "x", (0,), (1,))]
    c = parse(head + "if (!c[0]) { z q[0]; x q[1]; }")
    assert c._ins == [("gate", "z", (0,), ()), ("gate", "x", (1,), ()), ("feedback", "z", (0,), (0,)),
                      ("feedback", "x", (1,), (0,))]
    c = parse(head + "if (c[0] == 1) y q[1]; else x q[1];")
    assert c._ins == [("gate", "x", (1,), ()), ("feedback", "z", (1,), (0,))]
    c = parse(HEAD2 + "qreg q[1];\ncreg c[1];\nif (c == 1) x q[0];")
    assert c._ins == [("feedback", "x", (0,), (0,))]


def test_gphase_and_comments_are_ignored():
    c = parse(HEAD3 + "/* a block\ncomment */ qubit q; // a line comment\ngphase(0.3);\nh q;")
    assert [i[1] for i in c._ins] == ["h"]
