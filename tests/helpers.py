"""Shared test helpers: random inputs in the shapes of the reference's own test
helpers (tests/common/mod.rs:7-71), lowering of gate lists, and an
independent brute-force operator construction for tiny states."""
from __future__ import annotations

import numpy as np

from oracle import oracle


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_unitary(r: np.random.Generator, k: int) -> np.ndarray:
    """Gram-Schmidt on random complex columns (tests/common/mod.rs:12-37)."""
    dim = 1 << k
    cols = []
    while len(cols) < dim:
        v = (r.random(dim) - 0.5) + 1j * (r.random(dim) - 0.5)
        for u in cols:
            v = v - np.vdot(u, v) * u
        norm = np.sqrt(np.sum(np.abs(v) ** 2))
        if norm > 1e-6:
            cols.append(v / norm)
    return np.array(cols).T.copy()


def random_diagonal(r: np.random.Generator, k: int) -> np.ndarray:
    return np.diag(np.exp(1j * r.random(1 << k) * 2 * np.pi))


def random_state(r: np.random.Generator, n: int) -> np.ndarray:
    v = (r.random(1 << n) - 0.5) + 1j * (r.random(1 << n) - 0.5)
    return v / np.sqrt(np.sum(np.abs(v) ** 2))


def random_qubits(r: np.random.Generator, n: int, k: int) -> list:
    q = list(range(n))
    for i in range(k):
        j = int(r.integers(i, n))
        q[i], q[j] = q[j], q[i]
    return q[:k]


def max_distance(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.max(np.abs(a - b))) if a.size else 0.0


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def lower_with_oracle(gates) -> list:
    """gate_op (simulator.rs:106-119) over the oracle's gate table: controls
    stay separate, the matrix is the base gate's."""
    return oracle.lower(gates)


def full_operator(n: int, matrix: np.ndarray, qubits) -> np.ndarray:
    """The 2^n x 2^n operator of `matrix` on `qubits` (index bit j = qubits[j]),
    built entry by entry from the definition; independent of the oracle."""
    dim = 1 << n
    k = len(qubits)
    mask = sum(1 << q for q in qubits)
    U = np.zeros((dim, dim), dtype=np.complex128)
    for r in range(dim):
        for c in range(dim):
            if (r & ~mask) != (c & ~mask):
                continue
            lr = sum(((r >> qubits[j]) & 1) << j for j in range(k))
            lc = sum(((c >> qubits[j]) & 1) << j for j in range(k))
            U[r, c] = matrix[lr, lc]
    return U


def controlled_full(matrix: np.ndarray, controls) -> np.ndarray:
    """Matrix on [controls..., targets...] applying `matrix` when every
    control has its value (from the definition, independent of the oracle)."""
    c = len(controls)
    kt = int(round(np.log2(matrix.shape[0])))
    dim = 1 << (c + kt)
    state = sum((1 if v else 0) << i for i, (_, v) in enumerate(controls))
    out = np.eye(dim, dtype=np.complex128)
    mask = (1 << c) - 1
    for r in range(dim):
        for col in range(dim):
            if (r & mask) == state and (col & mask) == state:
                out[r, col] = matrix[r >> c, col >> c]
            elif (r & mask) == state or (col & mask) == state:
                out[r, col] = 0
    return out


def oracle_statevector(n: int, gates) -> np.ndarray:
    """reference::statevector (reference.rs:33-44) through the oracle's FP64
    naive applier."""
    state = np.zeros(1 << n, dtype=np.complex128)
    state[0] = 1.0
    for name, qubits, params in gates:
        oracle.reference_apply(state, oracle.gate_matrix(name, params), list(qubits))
    return state
