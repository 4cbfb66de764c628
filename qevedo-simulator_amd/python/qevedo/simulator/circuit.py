"""Circuits for the state-vector engine.

Mirrors the ``Circuit`` class of the reference's native module
(python/src/lib.rs:18-407) for what the state-vector backend runs: gates by
method or OpenQASM name, arbitrary unitaries, measurements, resets, barriers
and single-bit Pauli feedback. Validation messages follow the reference
(python/src/lib.rs:29-56, 164-190).
"""
from __future__ import annotations

import math

import numpy as np

from . import _native

# OpenQASM 2 qelib1 names written by to_qasm (circuit.rs:56-88).
_QASM_NAME = {"p": "u1", "u": "u3", "cp": "cu1", "i": "id", "cnot": "cx", "phase": "u1"}


class Circuit:
    """A quantum circuit on numbered qubits and classical bits (qubit 0 is the
    least significant bit, as in Qiskit)."""

    def __init__(self, num_qubits: int, num_clbits: int = 0):
        self._n = int(num_qubits)
        self._c = int(num_clbits)
        # (kind, name, qubits, params): kind is "gate", "unitary", "measure",
        # "reset", "barrier" or "feedback".
        self._ins: list = []

    # -- validation (python/src/lib.rs:28-56) -------------------------------
    def _check(self, qubits) -> None:
        for q in qubits:
            if q >= self._n or q < 0:
                raise ValueError(f"qubit {q} is out of range for {self._n} qubits")
        if len(set(qubits)) != len(qubits):
            raise ValueError("a gate cannot act on the same qubit twice")

    def _push(self, name: str, qubits, params=()):
        qubits = [int(q) for q in qubits]
        self._check(qubits)
        controls, base = _native.gate_lower(name, params)
        arity = controls + int(round(math.log2(base.shape[0])))
        if arity != len(qubits):
            raise ValueError(f"the gate acts on {arity} qubits, given {len(qubits)}")
        self._ins.append(("gate", name, tuple(qubits), tuple(float(p) for p in params)))
        return self

    # -- gates ---------------------------------------------------------------
    def gate(self, name: str, qubits, params=()):
        """Apply a gate by its OpenQASM name, e.g. circuit.gate("cu1", [0, 1], [0.5])."""
        try:
            return self._push(name, qubits, params)
        except _native.NativeError as e:
            raise ValueError(str(e)) from None

    def id(self, q): return self._push("id", [q])
    def h(self, q): return self._push("h", [q])
    def x(self, q): return self._push("x", [q])
    def y(self, q): return self._push("y", [q])
    def z(self, q): return self._push("z", [q])
    def s(self, q): return self._push("s", [q])
    def sdg(self, q): return self._push("sdg", [q])
    def t(self, q): return self._push("t", [q])
    def tdg(self, q): return self._push("tdg", [q])
    def sx(self, q): return self._push("sx", [q])
    def sxdg(self, q): return self._push("sxdg", [q])
    def rx(self, theta, q): return self._push("rx", [q], [theta])
    def ry(self, theta, q): return self._push("ry", [q], [theta])
    def rz(self, theta, q): return self._push("rz", [q], [theta])
    def p(self, lam, q): return self._push("p", [q], [lam])
    def u(self, theta, phi, lam, q): return self._push("u", [q], [theta, phi, lam])
    def cx(self, control, target): return self._push("cx", [control, target])
    def cy(self, control, target): return self._push("cy", [control, target])
    def cz(self, a, b): return self._push("cz", [a, b])
    def ch(self, control, target): return self._push("ch", [control, target])
    def swap(self, a, b): return self._push("swap", [a, b])
    def cp(self, lam, control, target): return self._push("cp", [control, target], [lam])
    def crx(self, theta, control, target): return self._push("crx", [control, target], [theta])
    def cry(self, theta, control, target): return self._push("cry", [control, target], [theta])
    def crz(self, theta, control, target): return self._push("crz", [control, target], [theta])
    def rzz(self, theta, a, b): return self._push("rzz", [a, b], [theta])
    def rxx(self, theta, a, b): return self._push("rxx", [a, b], [theta])
    def ccx(self, a, b, target): return self._push("ccx", [a, b, target])
    def cswap(self, control, a, b): return self._push("cswap", [control, a, b])

    def unitary(self, matrix, qubits):
        """A 2^k x 2^k unitary on k qubits; bit j of the matrix index is qubits[j]."""
        qubits = [int(q) for q in qubits]
        m = np.asarray(matrix, dtype=np.complex128)
        dim = 1 << len(qubits)
        if m.shape != (dim, dim):
            raise ValueError(f"a gate on {len(qubits)} qubits needs a {dim}x{dim} matrix")
        if np.max(np.abs(m @ m.conj().T - np.eye(dim))) > 1e-8:
            raise ValueError("the matrix is not unitary")
        self._check(qubits)
        self._ins.append(("unitary", m.copy(), tuple(qubits), ()))
        return self

    # -- non-unitary instructions -------------------------------------------
    def measure(self, qubit: int, clbit: int):
        self._check([qubit])
        self._c = max(self._c, clbit + 1)
        self._ins.append(("measure", None, (int(qubit),), (int(clbit),)))
        return self

    def measure_all(self):
        for q in range(self._n):
            self.measure(q, q)
        return self

    def reset(self, qubit: int):
        self._check([qubit])
        self._ins.append(("reset", None, (int(qubit),), ()))
        return self

    def barrier(self):
        self._ins.append(("barrier", None, tuple(range(self._n)), ()))
        return self

    def feedback(self, pauli: str, qubit: int, clbit: int):
        """The Pauli X, Y or Z on qubit if classical bit clbit reads 1 at this point."""
        if pauli not in ("X", "x", "Y", "y", "Z", "z"):
            raise ValueError(f"pauli must be X, Y or Z, not {pauli!r}")
        self._check([qubit])
        self._c = max(self._c, clbit + 1)
        self._ins.append(("feedback", pauli.lower(), (int(qubit),), (int(clbit),)))
        return self

    # -- properties ----------------------------------------------------------
    @property
    def num_qubits(self) -> int:
        return self._n

    @property
    def num_clbits(self) -> int:
        return self._c

    @property
    def gate_count(self) -> int:
        return sum(1 for i in self._ins if i[0] in ("gate", "unitary"))

    def has_noise(self) -> bool:
        return False

    def instructions(self) -> list:
        out = []
        for kind, name, qubits, params in self._ins:
            if kind == "gate":
                out.append((_QASM_NAME.get(name, name), list(qubits), list(params)))
            elif kind == "unitary":
                out.append(("unitary", list(qubits), []))
            elif kind == "feedback":
                out.append((f"feedback_{name}", list(qubits), [float(params[0])]))
            elif kind == "measure":
                out.append(("measure", list(qubits), [float(params[0])]))
            else:
                out.append((kind, list(qubits), []))
        return out

    def copy(self) -> "Circuit":
        c = Circuit(self._n, self._c)
        c._ins = list(self._ins)
        return c

    def __len__(self) -> int:
        return len(self._ins)

    def __repr__(self) -> str:
        return f"Circuit({self._n} qubits, {self._c} classical bits, {self.gate_count} gates)"

    # -- lowering ------------------------------------------------------------
    def ops(self, instructions=None) -> list:
        """gate_op (simulator.rs:106-119) of the unitary instructions:
        [(targets, [(control, True)], matrix)]."""
        ops, cache = [], {}
        for kind, name, qubits, params in (self._ins if instructions is None else instructions):
            if kind == "unitary":
                ops.append((list(qubits), [], name))
            elif kind == "gate":
                key = (name, params)
                if key not in cache:
                    cache[key] = _native.gate_lower(name, params)
                c, m = cache[key]
                ops.append((list(qubits[c:]), [(q, True) for q in qubits[:c]], m))
        return ops

    # -- OpenQASM ------------------------------------------------------------
    def to_qasm(self) -> str:
        """OpenQASM 2 with qelib1 names and 17-digit angles (circuit.rs:678-726)."""
        lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', f"qreg q[{self._n}];"]
        if self._c:
            lines.append(f"creg c[{self._c}];")
        for kind, name, qubits, params in self._ins:
            if kind == "gate":
                p = "(" + ",".join(f"{x:.17f}" for x in params) + ")" if params else ""
                lines.append(f"{_QASM_NAME.get(name, name)}{p} " + ",".join(f"q[{q}]" for q in qubits) + ";")
            elif kind == "measure":
                lines.append(f"measure q[{qubits[0]}] -> c[{params[0]}];")
            elif kind == "reset":
                lines.append(f"reset q[{qubits[0]}];")
            elif kind == "barrier":
                lines.append("barrier q;")
            elif kind == "unitary":
                raise ValueError("arbitrary unitaries cannot be written as OpenQASM 2")
            else:
                raise ValueError("feedback on one classical bit cannot be written as OpenQASM 2")
        return "\n".join(lines) + "\n"

    @staticmethod
    def from_qasm(source: str) -> "Circuit":
        """Read an OpenQASM 2 or 3 program (see qevedo.simulator.qasm):
        registers, qelib1 / stdgates gates, gate definitions, modifiers,
        measure, reset, barrier and Pauli feedback under `if`."""
        from .qasm import parse

        return parse(source)
