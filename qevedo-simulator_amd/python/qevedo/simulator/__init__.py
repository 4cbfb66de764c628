"""qevedo.simulator — the MI355X state-vector backend.

The state-vector surface of the reference's Python package
(python/python/qevedo/simulator/__init__.py): build a :class:`Circuit` (or
read OpenQASM 2 with :meth:`Circuit.from_qasm`), get its final state with
:func:`statevector` or sample it with :func:`run`. Every amplitude lives in
HBM and every gate, reduction and sample is computed by the HIP engine
``libqvd_hip.so`` through its C ABI.

    >>> from qevedo.simulator import Circuit, run
    >>> bell = Circuit(2).h(0).cx(0, 1).measure_all()
    >>> sorted(run(bell, shots=1000, seed=1).counts)
    ['00', '11']

Seeds drive NumPy's PCG64 generator (the reference's rand_pcg stream is not
reproduced): the same seed gives the same counts on this engine.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native, library
from ._native import (Batch, NativeError, StateVector, device_info, gate_lower, gate_matrix, lower_gates, plan_ops)
from .circuit import Circuit

__version__ = "0.1.18+mi355x"

__all__ = ["Batch", "Circuit", "NativeError", "RunResult", "StateVector", "choose_backend", "device_info", "expectation",
           "gate_lower", "gate_matrix", "library", "lower_gates", "plan_ops", "run", "statevector"]

BACKENDS = ("auto", "statevector")
_STATE_ALIASES = ("statevector", "state_vector")


class RunResult:
    """Counts and statistics of a run (python/src/lib.rs:410-478)."""

    def __init__(self, counts, gates, gate_seconds, measure_seconds, kernels=0, passes=0):
        self.counts = counts
        self.backend = "statevector"
        self.gates = gates
        self.gate_seconds = gate_seconds
        self.measure_seconds = measure_seconds
        self.kernels = kernels
        self.passes = passes
        self.fidelity = None
        self.max_bond_dimension = None
        self.peak_active_qubits = None

    def __repr__(self) -> str:
        return f"RunResult(backend='statevector', {len(self.counts)} outcomes)"


def _circuit(circuit) -> Circuit:
    if isinstance(circuit, str):
        return Circuit.from_qasm(circuit)
    if not isinstance(circuit, Circuit):
        raise TypeError("expected a qevedo.simulator.Circuit or OpenQASM text")
    return circuit


def _backend(name: str) -> str:
    if name in ("auto",) + _STATE_ALIASES:
        return "statevector"
    if name in ("stabilizer", "mps", "matrix_product_state", "near_clifford", "nearclifford"):
        raise ValueError(f"backend `{name}` is not part of this engine (state vector only)")
    raise ValueError(f"unknown backend `{name}` (auto, statevector, stabilizer, mps, near_clifford)")


def choose_backend(circuit, precision: str = "single") -> str:
    """choose_backend (simulator.rs:304-323) for this engine: the state vector
    when it takes at most 80 % of the device's HBM (state_vector_fits,
    simulator.rs:254-259). The reference's other backends (stabilizer,
    near-Clifford, matrix product state) are not part of it, so a state that
    does not fit is an error instead of a matrix product state."""
    c = _circuit(circuit)
    fits, state_bytes, device_bytes = _native.state_vector_fits(c.num_qubits, precision)
    if not fits:
        raise ValueError(f"a {c.num_qubits}-qubit {precision} state vector ({state_bytes} bytes) does not fit in 80 % "
                         f"of the device's {device_bytes} bytes of memory")
    return "statevector"


def _bitstring(key: int, num_clbits: int) -> str:
    return "".join("1" if (key >> c) & 1 else "0" for c in reversed(range(num_clbits)))


def _count(outcomes: np.ndarray, measures, num_clbits: int) -> dict:
    """count_outcomes (simulator.rs:187-238): a key per shot over the measured
    (qubit, clbit) pairs, counted, formatted with classical bit 0 on the
    right, in string order. Up to 128 classical bits the key is the OR of the
    measured bits (the reference's u128 key), kept as two 64-bit words; above
    128 each pair sets its classical bit, the last one written winning."""
    outcomes = np.asarray(outcomes, dtype=np.uint64)
    if num_clbits > 128:
        bits = np.zeros((outcomes.size, num_clbits), dtype=np.uint8)
        for q, c in measures:
            bits[:, c] = (outcomes >> np.uint64(q)) & np.uint64(1)
        uniq, counts = np.unique(bits, axis=0, return_counts=True)
        return dict(sorted(("".join("1" if b else "0" for b in row[::-1]), int(n)) for row, n in zip(uniq, counts)))
    words = np.zeros((outcomes.size, 2), dtype=np.uint64)
    for q, c in measures:
        words[:, c >> 6] |= ((outcomes >> np.uint64(q)) & np.uint64(1)) << np.uint64(c & 63)
    uniq, counts = np.unique(words, axis=0, return_counts=True)
    return dict(sorted((_bitstring(int(lo) | (int(hi) << 64), num_clbits), int(n)) for (lo, hi), n in zip(uniq, counts)))


def _terminal_split(ins) -> int | None:
    """terminal_measurements (simulator.rs:170-182)."""
    split = len(ins)
    for i in range(len(ins) - 1, -1, -1):
        if ins[i][0] in ("measure", "barrier"):
            split = i
        else:
            break
    if all(k[0] in ("gate", "unitary", "barrier") for k in ins[:split]):
        return split
    return None


def run(circuit, shots: int = 1024, *, backend: str = "auto", seed: int | None = None, precision: str = "single",
        max_bond_dimension: int = 256, truncation_threshold: float = 1e-16, threads=None) -> RunResult:
    """Sample shots of a circuit on the state vector (simulator.rs:291-427).

    Circuits whose measurements all come at the end are simulated once and
    sampled with the two-pass sampler; mid-circuit measurements, resets and
    feedback run shot by shot. ``max_bond_dimension``, ``truncation_threshold``
    and ``threads`` belong to other backends and CPU placement and are
    accepted for compatibility only.
    """
    c = _circuit(circuit)
    _backend(backend)
    if precision not in ("single", "double"):
        raise ValueError('precision must be "single" or "double"')
    choose_backend(c, precision)
    rng = np.random.Generator(np.random.PCG64(seed))
    ins = c._ins
    split = _terminal_split(ins)
    if split is not None:
        ops = c.ops(ins[:split])
        measures = [(q[0], p[0]) for k, _, q, p in ins[split:] if k == "measure"]
        with StateVector(c.num_qubits, precision) as sv:
            st = sv.apply_ops(ops) if ops else {"gates": 0, "gate_seconds": 0.0, "kernels": 0, "passes": 0}
            t0 = time.perf_counter()
            outcomes = sv.sample(rng.random(shots)) if shots else np.zeros(0, dtype=np.uint64)
            counts = _count(outcomes, measures, c.num_clbits)
            measure_seconds = time.perf_counter() - t0
        return RunResult(counts, st["gates"], st["gate_seconds"], measure_seconds, st["kernels"], st["passes"])
    counts: dict = {}
    gates = kernels = passes = 0
    gate_s = meas_s = 0.0
    with StateVector(c.num_qubits, precision) as sv:
        for _ in range(shots):
            sv.set_zero()
            clbits = [0] * c.num_clbits
            pending = []

            def flush():
                nonlocal gates, kernels, passes, gate_s
                if pending:
                    st = sv.apply_ops(c.ops(pending))
                    gates += st["gates"]
                    kernels += st["kernels"]
                    passes += st["passes"]
                    gate_s += st["gate_seconds"]
                    pending.clear()

            for ins_ in ins:
                kind, name, qubits, params = ins_
                if kind in ("gate", "unitary"):
                    pending.append(ins_)
                elif kind == "measure":
                    flush()
                    t0 = time.perf_counter()
                    clbits[params[0]] = sv.measure(qubits[0], rng.random())
                    meas_s += time.perf_counter() - t0
                elif kind == "reset":
                    flush()
                    sv.reset(qubits[0], rng.random())
                elif kind == "feedback":
                    if clbits[params[0]]:
                        pending.append(("gate", name, qubits, ()))
            flush()
            key = _bitstring(sum(b << i for i, b in enumerate(clbits)), c.num_clbits)
            counts[key] = counts.get(key, 0) + 1
    return RunResult(dict(sorted(counts.items())), gates, gate_s, meas_s, kernels, passes)


def statevector(circuit, *, threads=None) -> np.ndarray:
    """The final state of a circuit without measurements, in double precision
    (complex128, index bit q = qubit q), computed in HBM in complex128."""
    c = _circuit(circuit)
    if c.num_qubits > 34:
        raise ValueError("a state vector of more than 34 qubits does not fit in memory")
    for kind, *_ in c._ins:
        if kind not in ("gate", "unitary", "barrier"):
            raise ValueError(f"statevector() needs a unitary circuit, found {kind}")
    with StateVector(c.num_qubits, "double") as sv:
        ops = c.ops()
        if ops:
            sv.apply_ops(ops)
        return sv.amplitudes()


def expectation(circuit, observable, *, precision: str = "double") -> float:
    """<psi|O|psi> for the final state of a unitary circuit and an observable
    given as {Pauli label: coefficient} (Qiskit labels: the last character
    is qubit 0): the estimator's statevector path
    (python/python/qevedo/simulator/qiskit.py:584-589), with the state and
    every Pauli term evaluated in HBM (FP64 sums)."""
    c = _circuit(circuit)
    for kind, *_ in c._ins:
        if kind not in ("gate", "unitary", "barrier"):
            raise ValueError(f"expectation() needs a unitary circuit, found {kind}")
    with StateVector(c.num_qubits, precision) as sv:
        ops = c.ops()
        if ops:
            sv.apply_ops(ops)
        return sv.expectation_value(observable)
