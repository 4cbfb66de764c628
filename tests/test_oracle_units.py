"""Pins the oracle's building blocks against the reference's own known answers
and against independent derivations (matrix exponentials, definitions)."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from helpers import controlled_full, full_operator, random_unitary, rng
from oracle import oracle

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1, -1]).astype(complex)
I2 = np.eye(2, dtype=complex)


def test_deposit_known_answers():
    # kernels.rs:541-546
    assert oracle.deposit(0b101, 0b1101_0000) == 0b1000_0000 | 0b0001_0000
    assert oracle.deposit(0b11, 0b1010) == 0b1010
    assert oracle.deposit(0, 0xFF) == 0


def _expected(name, p):
    """Independent closed forms: Qiskit's definitions via matrix exponentials."""
    rz = lambda t: expm(-0.5j * t * Z)
    ry = lambda t: expm(-0.5j * t * Y)
    table = {
        "id": lambda: I2,
        "h": lambda: np.array([[1, 1], [1, -1]]) / math.sqrt(2),
        "x": lambda: X, "y": lambda: Y, "z": lambda: Z,
        "s": lambda: np.diag([1, 1j]), "sdg": lambda: np.diag([1, -1j]),
        "t": lambda: np.diag([1, np.exp(1j * math.pi / 4)]),
        "tdg": lambda: np.diag([1, np.exp(-1j * math.pi / 4)]),
        "sx": lambda: np.exp(1j * math.pi / 4) * expm(-0.25j * math.pi * X),
        "sxdg": lambda: np.exp(-1j * math.pi / 4) * expm(0.25j * math.pi * X),
        "rx": lambda: expm(-0.5j * p[0] * X), "ry": lambda: ry(p[0]), "rz": lambda: rz(p[0]),
        "p": lambda: np.diag([1, np.exp(1j * p[0])]), "u1": lambda: np.diag([1, np.exp(1j * p[0])]),
        "u3": lambda: np.exp(0.5j * (p[1] + p[2])) * rz(p[1]) @ ry(p[0]) @ rz(p[2]),
        "u": lambda: np.exp(0.5j * (p[1] + p[2])) * rz(p[1]) @ ry(p[0]) @ rz(p[2]),
        # two-qubit: index bit 0 = first qubit, so the first factor of the
        # Kronecker product is the *second* qubit.
        "swap": lambda: np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
        "rzz": lambda: expm(-0.5j * p[0] * np.kron(Z, Z)),
        "rxx": lambda: expm(-0.5j * p[0] * np.kron(X, X)),
    }
    controlled = {"cx": ("x", 1), "cy": ("y", 1), "cz": ("z", 1), "ch": ("h", 1), "cp": ("p", 1),
                  "cu1": ("p", 1), "crx": ("rx", 1), "cry": ("ry", 1), "crz": ("rz", 1),
                  "ccx": ("x", 2), "cswap": ("swap", 1)}
    if name in controlled:
        base, c = controlled[name]
        return controlled_full(np.asarray(_expected(base, p), dtype=complex), [(q, True) for q in range(c)])
    return np.asarray(table[name](), dtype=complex)


GATES = [("id", ()), ("h", ()), ("x", ()), ("y", ()), ("z", ()), ("s", ()), ("sdg", ()), ("t", ()),
         ("tdg", ()), ("sx", ()), ("sxdg", ()), ("rx", (0.3,)), ("ry", (0.4,)), ("rz", (0.5,)),
         ("p", (0.6,)), ("u3", (0.1, 0.2, 0.3)), ("cx", ()), ("cy", ()), ("cz", ()), ("ch", ()),
         ("cp", (0.7,)), ("crx", (0.8,)), ("cry", (0.9,)), ("crz", (1.0,)), ("swap", ()),
         ("rzz", (1.1,)), ("rxx", (1.2,)), ("ccx", ()), ("cswap", ())]
DIAGONAL = {"id", "z", "s", "sdg", "t", "tdg", "rz", "p", "cz", "cp", "crz", "rzz"}


@pytest.mark.parametrize("name,params", GATES)
def test_gate_matrices_match_definitions(name, params):
    m = oracle.gate_matrix(name, params)
    assert np.max(np.abs(m - _expected(name, params))) < 1e-12
    # circuit.rs:765-791: unitary, and the diagonal flag by the exact test.
    assert np.max(np.abs(m @ m.conj().T - np.eye(m.shape[0]))) < 1e-12
    off = m - np.diag(np.diag(m))
    assert (not np.any(off != 0)) == (name in DIAGONAL)


def test_gate_known_answers():
    # circuit.rs:795-812: SX^2 = X, U(pi,0,pi) = X, U(pi/2,0,pi) = H.
    sx = oracle.gate_matrix("sx")
    assert np.max(np.abs(sx @ sx - X)) < 1e-12
    assert np.max(np.abs(oracle.gate_matrix("u3", (math.pi, 0, math.pi)) - X)) < 1e-12
    h = oracle.gate_matrix("h")
    assert np.max(np.abs(oracle.gate_matrix("u3", (math.pi / 2, 0, math.pi)) - h)) < 1e-12
    # sqrt(W) of the random circuit is U(pi/2,-pi/4,pi/4) and differs from SX.
    w = oracle.gate_matrix("u3", (math.pi / 2, -math.pi / 4, math.pi / 4))
    W = (X + Y) / math.sqrt(2)
    # w^2 is W up to a global phase: |tr(W^dag w^2)| / 2 == 1
    assert abs(abs(np.trace(W.conj().T @ w @ w)) / 2 - 1) < 1e-12


def test_embed_kronecker_order_and_cx():
    # matrix.rs:211-234
    big = oracle.matrix_embed(X, [1], [0, 1])
    for r in range(4):
        for c in range(4):
            assert big[r, c].real == (1.0 if r == c ^ 2 else 0.0)
    cx = oracle.matrix_controlled(X, 1, 1)
    assert cx[1, 3].real == 1 and cx[3, 1].real == 1 and cx[0, 0].real == 1 and cx[2, 2].real == 1
    assert cx[1, 1].real == 0


def test_embed_and_controlled_against_definitions():
    r = rng(5)
    for _ in range(20):
        k = int(r.integers(1, 4))
        m = random_unitary(r, k)
        qubits = list(r.permutation(5)[: k + 2])
        gate_qubits = [qubits[i] for i in r.permutation(len(qubits))[:k]]
        got = oracle.matrix_embed(m, gate_qubits, qubits)
        # same operator as acting on a 5-qubit register
        pos = [qubits.index(q) for q in gate_qubits]
        assert np.max(np.abs(got - full_operator(len(qubits), m, pos))) < 1e-14
        c = int(r.integers(1, 3))
        state = int(r.integers(0, 1 << c))
        ctrl = [(i, bool((state >> i) & 1)) for i in range(c)]
        assert np.max(np.abs(oracle.matrix_controlled(m, c, state) - controlled_full(m, ctrl))) == 0


def _op(name, qubits):
    return (list(qubits), [], oracle.gate_matrix(name))


def test_fusion_known_answers():
    # fusion.rs:488-503: h(0) h(5) cx(0,1) h(6) cx(1,2), width 3
    ops = [_op("h", [0]), _op("h", [5]), _op("cx", [0, 1]), _op("h", [6]), _op("cx", [1, 2])]
    fused = oracle.fuse(ops, 3)
    assert [f[0] for f in fused] == [[0, 1, 2], [5, 6]]
    # fusion.rs:505-517: cx(0,1) h(1) cx(1,2), width 2
    ops = [_op("cx", [0, 1]), _op("h", [1]), _op("cx", [1, 2])]
    fused = oracle.fuse(ops, 2)
    assert [f[0] for f in fused] == [[0, 1], [1, 2]]


def test_fused_matrix_is_the_product():
    """The fused matrix equals the ordered product of the members' operators."""
    r = rng(11)
    ops = []
    for _ in range(6):
        q = [int(x) for x in r.permutation(4)[:2]]
        ops.append((q, [], random_unitary(r, 2)))
    fused = oracle.fuse(ops, 4)
    assert len(fused) == 1
    targets, nc, m = fused[0]
    assert targets == [0, 1, 2, 3] and nc == 0
    U = np.eye(16, dtype=complex)
    for q, _, g in ops:
        U = full_operator(4, g, q) @ U
    assert np.max(np.abs(m - U)) < 1e-13


def test_reference_apply_brute_force():
    """reference.rs:8-28 against an explicit operator on tiny states."""
    r = rng(7)
    for _ in range(60):
        n = int(r.integers(1, 7))
        k = int(r.integers(1, min(n, 4) + 1))
        q = [int(x) for x in r.permutation(n)[:k]]
        m = random_unitary(r, k)
        psi = (r.random(1 << n) - 0.5) + 1j * (r.random(1 << n) - 0.5)
        want = full_operator(n, m, q) @ psi
        got = psi.copy()
        oracle.reference_apply(got, m, q)
        assert np.max(np.abs(got - want)) < 1e-13


def test_count_outcomes_bitstrings():
    # classical bit 0 on the right (simulator.rs:232-238)
    outcomes = np.array([0b01, 0b01, 0b10, 0b11], dtype=np.uint64)
    counts = oracle.count_outcomes(outcomes, [(0, 0), (1, 1)], 2)
    assert counts == {"01": 2, "10": 1, "11": 1}
    counts = oracle.count_outcomes(outcomes, [(0, 2)], 3)
    assert counts == {"000": 1, "100": 3}
