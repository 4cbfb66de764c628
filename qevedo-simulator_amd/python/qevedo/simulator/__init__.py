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
    sampled with the two-pass sampler; with mid-circuit measurements, resets
    and feedback, shots that share a measurement history share one state
    (see _run_branches), with the counts of the reference's shot-by-shot
    loop. ``max_bond_dimension``, ``truncation_threshold``
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
    draws = sum(1 for k, *_ in ins if k in ("measure", "reset"))
    # One uniform per measure or reset, shot after shot: the order in which
    # the reference's shot loop draws them (simulator.rs:383-425).
    u = rng.random(shots * draws).reshape(shots, draws) if draws else np.zeros((shots, 0))
    _, state_bytes, device_bytes = _native.state_vector_fits(c.num_qubits, precision)
    if (draws + 2) * state_bytes <= 0.5 * device_bytes:
        return _run_branches(c, u, precision)
    return _run_shot_by_shot(c, u, precision)


class _Stats:
    def __init__(self):
        self.gates = self.kernels = self.passes = 0
        self.gate_s = self.meas_s = 0.0

    def add(self, st):
        self.gates += st["gates"]
        self.kernels += st["kernels"]
        self.passes += st["passes"]
        self.gate_s += st["gate_seconds"]


def _flush(c, sv, pending, stats):
    if pending:
        stats.add(sv.apply_ops(c.ops(pending)))
        pending.clear()


def _run_branches(c, u, precision) -> RunResult:
    """The shot loop with shots that share a measurement history sharing one
    state: a depth-first walk over the outcome tree. Each branch holds the
    shots whose outcomes so far agree; at a measure or reset the branch's
    P(1) is computed once, its shots split by their own uniform (outcome =
    u < P(1), as measure.rs:68-80 decides per shot), and a branch that
    splits continues in a device copy of its state for the shots with
    outcome 1. Every branch state is bit for bit the state each of its
    shots reaches alone, so the counts equal the shot-by-shot loop's; the
    gates run once per branch instead of once per shot."""
    ins = c._ins
    shots = u.shape[0]
    counts: dict = {}
    stats = _Stats()
    root = StateVector(c.num_qubits, precision)
    # (state, shot indices, next instruction, classical bits, next draw, pending gates)
    stack = [(root, np.arange(shots), 0, [0] * c.num_clbits, 0, [])]
    try:
        while stack:
            sv, idx, i, clbits, k, pending = stack.pop()
            while i < len(ins):
                kind, name, qubits, params = ins[i]
                i += 1
                if kind in ("gate", "unitary"):
                    pending.append(ins[i - 1])
                elif kind in ("measure", "reset"):
                    _flush(c, sv, pending, stats)
                    t0 = time.perf_counter()
                    q = qubits[0]
                    p1 = min(max(sv.probability_one(q), 0.0), 1.0)
                    one = u[idx, k] < p1
                    ones, zeros = idx[one], idx[~one]
                    k += 1
                    def collapse(state, outcome):
                        # a uniform of -1 is below any P(1) > 0 (some shot's u
                        # is below P(1) when outcome 1 occurs), 2 above any
                        f = state.measure if kind == "measure" else state.reset
                        f(q, -1.0 if outcome else 2.0)

                    if ones.size and zeros.size:
                        other = StateVector(c.num_qubits, precision)
                        other.copy_from(sv)
                        collapse(sv, 1)
                        bits = list(clbits)
                        if kind == "measure":
                            bits[params[0]] = 1
                        stack.append((sv, ones, i, bits, k, []))
                        sv, idx = other, zeros
                        collapse(sv, 0)
                        if kind == "measure":
                            clbits = list(clbits)
                            clbits[params[0]] = 0
                    else:
                        outcome = 1 if ones.size else 0
                        collapse(sv, outcome)
                        if kind == "measure":
                            clbits = list(clbits)
                            clbits[params[0]] = outcome
                    stats.meas_s += time.perf_counter() - t0
                elif kind == "feedback":
                    if clbits[params[0]]:
                        pending.append(("gate", name, qubits, ()))
            key = _bitstring(sum(b << j for j, b in enumerate(clbits)), c.num_clbits)
            counts[key] = counts.get(key, 0) + int(idx.size)
            sv.close()
    finally:
        for entry in stack:
            entry[0].close()
    return RunResult(dict(sorted(counts.items())), stats.gates, stats.gate_s, stats.meas_s, stats.kernels,
                     stats.passes)


def _run_shot_by_shot(c, u, precision) -> RunResult:
    """The reference's shot loop (simulator.rs:383-425), one shot after
    another on one state, with the same uniforms; used when the branch
    copies of _run_branches would not fit in device memory."""
    ins = c._ins
    counts: dict = {}
    stats = _Stats()
    with StateVector(c.num_qubits, precision) as sv:
        for s in range(u.shape[0]):
            sv.set_zero()
            clbits = [0] * c.num_clbits
            pending = []
            k = 0
            for kind, name, qubits, params in ins:
                if kind in ("gate", "unitary"):
                    pending.append((kind, name, qubits, params))
                elif kind == "measure":
                    _flush(c, sv, pending, stats)
                    t0 = time.perf_counter()
                    clbits[params[0]] = sv.measure(qubits[0], float(u[s, k]))
                    stats.meas_s += time.perf_counter() - t0
                    k += 1
                elif kind == "reset":
                    _flush(c, sv, pending, stats)
                    sv.reset(qubits[0], float(u[s, k]))
                    k += 1
                elif kind == "feedback":
                    if clbits[params[0]]:
                        pending.append(("gate", name, qubits, ()))
            key = _bitstring(sum(b << j for j, b in enumerate(clbits)), c.num_clbits)
            counts[key] = counts.get(key, 0) + 1
    return RunResult(dict(sorted(counts.items())), stats.gates, stats.gate_s, stats.meas_s, stats.kernels,
                     stats.passes)


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
