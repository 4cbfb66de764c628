"""ORACLE — test infrastructure only.

ctypes binding of ``oracle/build/liboracle_qvd.so``, the CPU restatement of the
reference's dense state-vector path (``src/state.rs``, ``src/kernels.rs``,
``src/fusion.rs``, ``src/blocking.rs``, ``src/measure.rs``,
``src/reference.rs``). Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this module, as the checker; the
product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle_qvd.so")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    cmd = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


class OrcOp(C.Structure):
    _fields_ = [
        ("targets", C.POINTER(C.c_uint32)),
        ("num_targets", C.c_uint32),
        ("controls", C.POINTER(C.c_uint32)),
        ("control_values", C.POINTER(C.c_uint8)),
        ("num_controls", C.c_uint32),
        ("matrix", C.POINTER(C.c_double)),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("gates", C.c_uint64),
        ("kernels", C.c_uint64),
        ("passes", C.c_uint64),
        ("swaps", C.c_uint64),
        ("gate_seconds", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, u32, u64, i32 = C.c_double, C.c_uint32, C.c_uint64, C.c_int32
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        L.orc_deposit.argtypes = [u64, u64]
        L.orc_deposit.restype = u64
        L.orc_gate_matrix.argtypes = [C.c_char_p, P(d), u32, P(d), u32, P(u32)]
        L.orc_gate_controls.argtypes = [C.c_char_p, P(u32)]
        L.orc_matrix_embed.argtypes = [P(d), u32, P(u32), P(u32), u32, P(d)]
        L.orc_matrix_controlled.argtypes = [P(d), u32, u32, u64, P(d)]
        L.orc_reference_apply.argtypes = [P(d), u32, P(d), P(u32), u32]
        L.orc_fuse.argtypes = [P(OrcOp), u64, u32, P(u32), P(u32), P(u32), P(u32), P(d), u64]
        L.orc_sv_create.argtypes = [u32, C.c_int]
        L.orc_sv_create.restype = C.c_void_p
        L.orc_sv_destroy.argtypes = [C.c_void_p]
        L.orc_sv_set.argtypes = [C.c_void_p, P(d), u64]
        L.orc_sv_get.argtypes = [C.c_void_p, P(d), u64]
        L.orc_sv_apply.argtypes = [C.c_void_p, P(d), P(u32), u32, P(u32), P(C.c_uint8), u32]
        L.orc_sv_apply_ops.argtypes = [C.c_void_p, P(OrcOp), u64, u32, i32, P(OrcStats)]
        L.orc_sv_norm_sqr.argtypes = [C.c_void_p, P(d)]
        L.orc_sv_scale.argtypes = [C.c_void_p, d]
        L.orc_sv_probability_one.argtypes = [C.c_void_p, u32, P(d)]
        L.orc_sv_probabilities.argtypes = [C.c_void_p, P(d), u64]
        L.orc_sv_measure.argtypes = [C.c_void_p, u32, d, P(C.c_uint8)]
        L.orc_sv_reset.argtypes = [C.c_void_p, u32, d]
        L.orc_sv_sample.argtypes = [C.c_void_p, P(d), u64, P(u64)]
        L.orc_sv_swap_physical.argtypes = [C.c_void_p, P(u32), u32]
        L.orc_sv_layout.argtypes = [C.c_void_p, P(u32)]
        L.orc_count_outcomes.argtypes = [P(u64), u64, P(u32), P(u32), u32, P(u64), P(u64), P(u64)]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return lib().orc_max_threads()


def deposit(value: int, mask: int) -> int:
    return int(lib().orc_deposit(value, mask))


def gate_matrix(name: str, params=()) -> np.ndarray:
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(2 * 1024, dtype=np.float64)  # up to 5 qubits (c4x)
    k = C.c_uint32()
    _check(lib().orc_gate_matrix(name.encode(), _p(p, C.c_double), len(p), _p(out, C.c_double), len(out) // 2, C.byref(k)))
    dim = 1 << k.value
    return out[: 2 * dim * dim].view(np.complex128).reshape(dim, dim).copy()


def gate_controls(name: str) -> int:
    c = C.c_uint32()
    _check(lib().orc_gate_controls(name.encode(), C.byref(c)))
    return c.value


def lower(gates) -> list:
    """gate_op (simulator.rs:106-119) over the oracle's gate table for a list
    of (name, qubits, params): controls stay separate, the matrix is the
    base gate's."""
    ops = []
    cache = {}
    for name, qubits, params in gates:
        key = (name, tuple(params))
        if key not in cache:
            full = gate_matrix(name, params)
            c = gate_controls(name)
            if c:
                mask = (1 << c) - 1
                idx = [(i << c) | mask for i in range(full.shape[0] >> c)]
                full = full[np.ix_(idx, idx)]
            cache[key] = (c, full)
        c, m = cache[key]
        ops.append((list(qubits[c:]), [(q, True) for q in qubits[:c]], m))
    return ops


def matrix_embed(m: np.ndarray, gate_qubits, qubits) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.complex128)
    g = np.ascontiguousarray(gate_qubits, dtype=np.uint32)
    q = np.ascontiguousarray(qubits, dtype=np.uint32)
    dim = 1 << len(q)
    out = np.zeros((dim, dim), dtype=np.complex128)
    _check(lib().orc_matrix_embed(_p(m.view(np.float64), C.c_double), len(g), _p(g, C.c_uint32), _p(q, C.c_uint32), len(q), _p(out.view(np.float64), C.c_double)))
    return out


def matrix_controlled(m: np.ndarray, controls: int, state: int) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.complex128)
    k = int(round(np.log2(m.shape[0])))
    dim = 1 << (k + controls)
    out = np.zeros((dim, dim), dtype=np.complex128)
    _check(lib().orc_matrix_controlled(_p(m.view(np.float64), C.c_double), k, controls, state, _p(out.view(np.float64), C.c_double)))
    return out


def reference_apply(state: np.ndarray, matrix: np.ndarray, qubits) -> None:
    """reference.rs:8-28 in place on a complex128 vector."""
    assert state.dtype == np.complex128 and state.flags.c_contiguous
    n = int(round(np.log2(state.size)))
    m = np.ascontiguousarray(matrix, dtype=np.complex128)
    q = np.ascontiguousarray(qubits, dtype=np.uint32)
    _check(lib().orc_reference_apply(_p(state.view(np.float64), C.c_double), n, _p(m.view(np.float64), C.c_double), _p(q, C.c_uint32), len(q)))


class _OpArray:
    """Keeps the numpy buffers behind an array of OrcOp alive."""

    def __init__(self, ops):
        self.keep = []
        self.arr = (OrcOp * max(len(ops), 1))()
        for i, (targets, controls, matrix) in enumerate(ops):
            t = np.ascontiguousarray(targets, dtype=np.uint32)
            cq = np.ascontiguousarray([c[0] for c in controls], dtype=np.uint32)
            cv = np.ascontiguousarray([1 if c[1] else 0 for c in controls], dtype=np.uint8)
            m = np.ascontiguousarray(matrix, dtype=np.complex128)
            assert m.shape == (1 << len(t), 1 << len(t))
            self.keep += [t, cq, cv, m]
            self.arr[i] = OrcOp(_p(t, C.c_uint32), len(t), _p(cq, C.c_uint32), _p(cv, C.c_uint8), len(cq), _p(m.view(np.float64), C.c_double))
        self.n = len(ops)


def fuse(ops, max_qubits: int):
    """fusion.rs fuse(): returns [(targets, num_controls, matrix)]."""
    a = _OpArray(ops)
    n = len(ops)
    count = C.c_uint32()
    targets = np.zeros(16 * max(n, 1), dtype=np.uint32)
    nt = np.zeros(max(n, 1), dtype=np.uint32)
    nc = np.zeros(max(n, 1), dtype=np.uint32)
    cap = 2 * sum(4 ** max(len(o[0]), max_qubits) for o in ops) + 16
    mats = np.zeros(cap, dtype=np.float64)
    _check(lib().orc_fuse(a.arr, n, max_qubits, C.byref(count), _p(targets, C.c_uint32), _p(nt, C.c_uint32), _p(nc, C.c_uint32), _p(mats, C.c_double), cap))
    out, mo = [], 0
    for i in range(count.value):
        k = int(nt[i])
        dim = 1 << k
        m = mats[mo: mo + 2 * dim * dim].view(np.complex128).reshape(dim, dim).copy()
        mo += 2 * dim * dim
        out.append((list(map(int, targets[16 * i: 16 * i + k])), int(nc[i]), m))
    return out


class OracleState:
    """The restated blocked CPU engine (state.rs / measure.rs / blocking.rs).

    precision: "single" (f32 blocks of 8) or "double" (f64 blocks of 4).
    """

    def __init__(self, num_qubits: int, precision: str = "single"):
        prec = {"single": 0, "double": 1}[precision]
        self.num_qubits = num_qubits
        self.precision = precision
        self.h = lib().orc_sv_create(num_qubits, prec)
        if not self.h:
            raise OracleError(lib().orc_last_error().decode())

    def close(self):
        if self.h:
            lib().orc_sv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_amplitudes(self, amps: np.ndarray) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        _check(lib().orc_sv_set(self.h, _p(a.view(np.float64), C.c_double), a.size))

    def amplitudes(self) -> np.ndarray:
        out = np.zeros(1 << self.num_qubits, dtype=np.complex128)
        _check(lib().orc_sv_get(self.h, _p(out.view(np.float64), C.c_double), out.size))
        return out

    def apply(self, matrix, targets, controls=()) -> None:
        m = np.ascontiguousarray(matrix, dtype=np.complex128)
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        cq = np.ascontiguousarray([c[0] for c in controls], dtype=np.uint32)
        cv = np.ascontiguousarray([1 if c[1] else 0 for c in controls], dtype=np.uint8)
        _check(lib().orc_sv_apply(self.h, _p(m.view(np.float64), C.c_double), _p(t, C.c_uint32), len(t), _p(cq, C.c_uint32), _p(cv, C.c_uint8), len(cq)))

    def apply_ops(self, ops, max_fused_qubits: int = 4, region_bits: int | None = 14) -> dict:
        a = _OpArray(ops)
        st = OrcStats()
        _check(lib().orc_sv_apply_ops(self.h, a.arr, a.n, max_fused_qubits, -1 if region_bits is None else region_bits, C.byref(st)))
        return {"gates": st.gates, "kernels": st.kernels, "passes": st.passes, "swaps": st.swaps, "gate_seconds": st.gate_seconds}

    def norm_sqr(self) -> float:
        v = C.c_double()
        _check(lib().orc_sv_norm_sqr(self.h, C.byref(v)))
        return v.value

    def scale(self, factor: float) -> None:
        _check(lib().orc_sv_scale(self.h, factor))

    def probability_one(self, qubit: int) -> float:
        v = C.c_double()
        _check(lib().orc_sv_probability_one(self.h, qubit, C.byref(v)))
        return v.value

    def probabilities(self) -> np.ndarray:
        out = np.zeros(1 << self.num_qubits, dtype=np.float64)
        _check(lib().orc_sv_probabilities(self.h, _p(out, C.c_double), out.size))
        return out

    def measure(self, qubit: int, uniform: float) -> int:
        o = C.c_uint8()
        _check(lib().orc_sv_measure(self.h, qubit, uniform, C.byref(o)))
        return o.value

    def reset(self, qubit: int, uniform: float) -> None:
        _check(lib().orc_sv_reset(self.h, qubit, uniform))

    def sample(self, uniforms: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.zeros(u.size, dtype=np.uint64)
        _check(lib().orc_sv_sample(self.h, _p(u, C.c_double), u.size, _p(out, C.c_uint64)))
        return out

    def swap_physical(self, pairs) -> None:
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1))
        _check(lib().orc_sv_swap_physical(self.h, _p(p, C.c_uint32), len(p) // 2))

    def layout(self) -> list:
        out = np.zeros(max(self.num_qubits, 1), dtype=np.uint32)
        _check(lib().orc_sv_layout(self.h, _p(out, C.c_uint32)))
        return list(map(int, out[: self.num_qubits]))


def count_outcomes(outcomes: np.ndarray, measures, num_clbits: int) -> dict:
    """simulator.rs:187-238: counts keyed by bitstring, classical bit 0 on the right."""
    o = np.ascontiguousarray(outcomes, dtype=np.uint64)
    mq = np.ascontiguousarray([m[0] for m in measures], dtype=np.uint32)
    mc = np.ascontiguousarray([m[1] for m in measures], dtype=np.uint32)
    keys = np.zeros(max(o.size, 1), dtype=np.uint64)
    counts = np.zeros(max(o.size, 1), dtype=np.uint64)
    d = C.c_uint64()
    _check(lib().orc_count_outcomes(_p(o, C.c_uint64), o.size, _p(mq, C.c_uint32), _p(mc, C.c_uint32), len(mq), _p(keys, C.c_uint64), _p(counts, C.c_uint64), C.byref(d)))
    out = {}
    for i in range(d.value):
        k = int(keys[i])
        out["".join("1" if (k >> c) & 1 else "0" for c in reversed(range(num_clbits)))] = int(counts[i])
    return out
