"""ctypes binding of ``libqvd_hip.so`` (the C ABI in ``include/qvd_hip.h``).

This module is the only way the Python package reaches the engine; every
computation runs in the HIP library on the GPU. If the library is missing or
no GPU is visible, calls raise ``NativeError`` instead of computing anything
another way.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG_ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "..", ".."))
LIB_PATH = os.environ.get("QVD_LIB", os.path.join(_PKG_ROOT, "lib", "libqvd_hip.so"))

QVD_OK, QVD_EINVAL, QVD_ENOMEM, QVD_EHIP, QVD_EUNSUPPORTED = 0, -1, -2, -3, -4
C64, C128 = 0, 1
PLANNER_HEURISTIC, PLANNER_LEARNED, PLANNER_LEARNED_ANY = 0, 1, 2
FLAG_TIMING, FLAG_DENSE_VALU, FLAG_AUTOTUNE, FLAG_ALL_TILES = 1, 2, 4, 8

# Every symbol include/qvd_hip.h declares.
EXPORTS = (
    "qvd_last_error", "qvd_device_info", "qvd_sv_create", "qvd_sv_destroy", "qvd_sv_num_qubits",
    "qvd_sv_precision", "qvd_sv_set_zero", "qvd_sv_set_amplitudes", "qvd_sv_get_amplitudes", "qvd_sv_apply",
    "qvd_sv_apply_ops", "qvd_plan_ops", "qvd_sv_norm_sqr", "qvd_sv_scale", "qvd_sv_probability_one", "qvd_sv_probabilities",
    "qvd_sv_expect_z", "qvd_sv_measure", "qvd_sv_reset", "qvd_sv_sample", "qvd_sv_compare", "qvd_sv_last_sample_seconds",
    "qvd_gate_matrix", "qvd_gate_lower", "qvd_batch_create", "qvd_batch_destroy", "qvd_batch_run",
    "qvd_batch_outputs", "qvd_batch_amplitudes", "qvd_batch_geometry", "qvd_plan_features", "qvd_sv_stage_seconds", "qvd_mlp_create", "qvd_mlp_destroy",
    "qvd_mlp_num_params", "qvd_mlp_set_params", "qvd_mlp_get_params", "qvd_mlp_forward", "qvd_mlp_gradient",
    "qvd_mlp_train", "qvd_planner_set_model", "qvd_state_vector_fits", "qvd_plan_equivalent_variants", "qvd_sv_sample_noalloc", "qvd_sv_expect_pauli", "qvd_sv_copy",
)
PLAN_VARIANTS, PLAN_FEATURES = 32, 24


class NativeError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Op(C.Structure):
    _fields_ = [
        ("targets", C.POINTER(C.c_uint32)),
        ("num_targets", C.c_uint32),
        ("controls", C.POINTER(C.c_uint32)),
        ("control_values", C.POINTER(C.c_uint8)),
        ("num_controls", C.c_uint32),
        ("matrix", C.POINTER(C.c_double)),
    ]


class Options(C.Structure):
    _fields_ = [
        ("max_fused_qubits", C.c_uint32),
        ("tile_bits", C.c_int32),
        ("planner", C.c_uint32),
        ("flags", C.c_uint32),
        ("variant", C.c_uint32),
        ("fixed_lanes", C.c_uint32),
    ]


class CircuitDesc(C.Structure):
    _fields_ = [
        ("num_qubits", C.c_uint32),
        ("ops", C.POINTER(Op)),
        ("num_ops", C.c_uint64),
        ("fixed_lanes", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("gates", C.c_uint64),
        ("kernels", C.c_uint64),
        ("passes", C.c_uint64),
        ("stages", C.c_uint64),
        ("gate_seconds", C.c_double),
        ("measure_seconds", C.c_double),
        ("plan_seconds", C.c_double),
        ("stage_kernel_seconds", C.c_double),
        ("other_kernel_seconds", C.c_double),
        ("plan_variant", C.c_uint64),
        ("stage_bytes", C.c_uint64),
        ("other_bytes", C.c_uint64),
        ("skipped_tiles", C.c_uint64),
        ("fixed_lanes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    """Load the engine; raises NativeError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(QVD_EHIP, f"{LIB_PATH} is not built (run `make -C qevedo-simulator_amd`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        u32, u64, i32, d, vp = C.c_uint32, C.c_uint64, C.c_int32, C.c_double, C.c_void_p
        L.qvd_last_error.restype = C.c_char_p
        L.qvd_device_info.argtypes = [P(C.c_int), C.c_char_p, u32]
        L.qvd_state_vector_fits.argtypes = [u32, C.c_int, P(C.c_int), P(u64), P(u64)]
        L.qvd_sv_create.argtypes = [u32, C.c_int, P(vp)]
        L.qvd_sv_destroy.argtypes = [vp]
        L.qvd_sv_num_qubits.argtypes = [vp, P(u32)]
        L.qvd_sv_precision.argtypes = [vp, P(C.c_int)]
        L.qvd_sv_set_zero.argtypes = [vp]
        L.qvd_sv_copy.argtypes = [vp, vp]
        L.qvd_sv_set_amplitudes.argtypes = [vp, P(d), u64]
        L.qvd_sv_get_amplitudes.argtypes = [vp, P(d), u64]
        L.qvd_sv_apply.argtypes = [vp, P(d), P(u32), u32, P(u32), P(C.c_uint8), u32]
        L.qvd_sv_apply_ops.argtypes = [vp, P(Op), u64, P(Options), P(Stats)]
        L.qvd_sv_norm_sqr.argtypes = [vp, P(d)]
        L.qvd_sv_scale.argtypes = [vp, d]
        L.qvd_sv_probability_one.argtypes = [vp, u32, P(d)]
        L.qvd_sv_probabilities.argtypes = [vp, P(d), u64]
        L.qvd_sv_expect_z.argtypes = [vp, P(d), P(d)]
        L.qvd_sv_expect_pauli.argtypes = [vp, P(u64), P(u64), u64, P(d)]
        L.qvd_sv_measure.argtypes = [vp, u32, d, P(C.c_uint8)]
        L.qvd_sv_reset.argtypes = [vp, u32, d]
        L.qvd_sv_sample.argtypes = [vp, P(d), u64, P(u64)]
        L.qvd_sv_sample_noalloc.argtypes = [vp, P(d), u64, P(u64)]
        L.qvd_plan_ops.argtypes = [u32, C.c_int, P(Op), u64, P(Options), P(Stats)]
        L.qvd_sv_compare.argtypes = [vp, vp, P(d), P(d), P(d)]
        L.qvd_sv_last_sample_seconds.argtypes = [vp, P(d)]
        L.qvd_gate_matrix.argtypes = [C.c_char_p, P(d), u32, P(d), u32, P(u32)]
        L.qvd_gate_lower.argtypes = [C.c_char_p, P(d), u32, P(u32), P(u32), P(d), u32]
        L.qvd_batch_create.argtypes = [P(CircuitDesc), u32, C.c_int, P(Options), P(vp)]
        L.qvd_batch_destroy.argtypes = [vp]
        L.qvd_batch_run.argtypes = [vp, P(d), P(d), P(Stats)]
        L.qvd_batch_outputs.argtypes = [vp, P(u64), P(u64)]
        L.qvd_batch_amplitudes.argtypes = [vp, C.c_uint32, vp, u64]
        L.qvd_batch_geometry.argtypes = [vp, P(u32), P(u32), u32]
        f = C.c_float
        L.qvd_plan_features.argtypes = [u32, C.c_int, P(Op), u64, P(Options), P(f), u64, P(u64)]
        L.qvd_plan_equivalent_variants.argtypes = [u32, C.c_int, P(Op), u64, P(Options), P(u32)]
        L.qvd_sv_stage_seconds.argtypes = [vp, P(d), u64, P(u64)]
        L.qvd_mlp_create.argtypes = [u32, u32, P(vp)]
        L.qvd_mlp_destroy.argtypes = [vp]
        L.qvd_mlp_num_params.argtypes = [vp, P(u64)]
        L.qvd_mlp_set_params.argtypes = [vp, P(f), u64]
        L.qvd_mlp_get_params.argtypes = [vp, P(f), u64]
        L.qvd_mlp_forward.argtypes = [vp, P(f), u64, P(f)]
        L.qvd_mlp_gradient.argtypes = [vp, P(f), P(f), u64, P(f), P(d)]
        L.qvd_mlp_train.argtypes = [vp, P(f), P(f), u64, u32, u32, f, u64, P(d)]
        L.qvd_planner_set_model.argtypes = [vp, P(f), P(f), f, f]
        for name in EXPORTS:
            getattr(L, name)  # fail early on a missing symbol
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != QVD_OK:
        raise NativeError(rc, lib().qvd_last_error().decode())


def _p(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def device_info() -> tuple:
    n = C.c_int()
    arch = C.create_string_buffer(64)
    _check(lib().qvd_device_info(C.byref(n), arch, 64))
    return n.value, arch.value.decode()


def state_vector_fits(num_qubits: int, precision: str = "single") -> tuple:
    """(fits, state bytes, device bytes): whether the state takes at most 80 %
    of the current device's HBM (simulator.rs:254-259)."""
    prec = {"single": C64, "double": C128}.get(precision)
    if prec is None:
        raise ValueError('precision must be "single" or "double"')
    fits, sb, db = C.c_int(), C.c_uint64(), C.c_uint64()
    _check(lib().qvd_state_vector_fits(num_qubits, prec, C.byref(fits), C.byref(sb), C.byref(db)))
    return bool(fits.value), sb.value, db.value


def gate_matrix(name: str, params=()) -> np.ndarray:
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(2 * 1024, dtype=np.float64)  # up to 5 qubits (c4x)
    k = C.c_uint32()
    _check(lib().qvd_gate_matrix(name.encode(), _p(p, C.c_double), p.size, _p(out, C.c_double), 1024, C.byref(k)))
    dim = 1 << k.value
    return out[: 2 * dim * dim].view(np.complex128).reshape(dim, dim).copy()


def gate_lower(name: str, params=()) -> tuple:
    """(number of leading controls, base matrix) of a named gate."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(2 * 64, dtype=np.float64)
    nc, nt = C.c_uint32(), C.c_uint32()
    _check(lib().qvd_gate_lower(name.encode(), _p(p, C.c_double), p.size, C.byref(nc), C.byref(nt), _p(out, C.c_double), 64))
    dim = 1 << nt.value
    return nc.value, out[: 2 * dim * dim].view(np.complex128).reshape(dim, dim).copy()


def lower_gates(gates) -> list:
    """gate_op over a list of (name, qubits, params): [(targets, controls, matrix)]."""
    cache = {}
    ops = []
    for name, qubits, params in gates:
        key = (name, tuple(params))
        if key not in cache:
            cache[key] = gate_lower(name, params)
        c, m = cache[key]
        qubits = list(qubits)
        ops.append((qubits[c:], [(q, True) for q in qubits[:c]], m))
    return ops


class OpArray:
    """A qvd_op array plus the NumPy buffers it points into."""

    def __init__(self, ops):
        self.keep = []
        self.n = len(ops)
        self.arr = (Op * max(self.n, 1))()
        for i, (targets, controls, matrix) in enumerate(ops):
            t = np.ascontiguousarray(targets, dtype=np.uint32)
            cq = np.ascontiguousarray([c[0] for c in controls], dtype=np.uint32)
            cv = np.ascontiguousarray([1 if c[1] else 0 for c in controls], dtype=np.uint8)
            m = np.ascontiguousarray(matrix, dtype=np.complex128)
            if m.shape != (1 << t.size, 1 << t.size):
                raise ValueError(f"a gate on {t.size} qubits needs a {1 << t.size}x{1 << t.size} matrix")
            self.keep += [t, cq, cv, m]
            self.arr[i] = Op(_p(t, C.c_uint32), t.size, _p(cq, C.c_uint32), _p(cv, C.c_uint8), cq.size, _p(m.view(np.float64), C.c_double))


def plan_features(num_qubits: int, ops, precision: str = "single", variant: int = 0,
                  max_fused_qubits: int = 4, fixed_lanes: int = 0) -> np.ndarray:
    """Per register-tile stage features (rows x PLAN_FEATURES) of a plan variant."""
    arr = ops if isinstance(ops, OpArray) else OpArray(ops)
    opt = Options(max_fused_qubits, 0, PLANNER_HEURISTIC, 0, variant, fixed_lanes)
    prec = {"single": C64, "double": C128}[precision]
    rows = C.c_uint64()
    _check(lib().qvd_plan_features(num_qubits, prec, arr.arr, arr.n, C.byref(opt), None, 0, C.byref(rows)))
    out = np.zeros((max(rows.value, 1), PLAN_FEATURES), dtype=np.float32)
    _check(lib().qvd_plan_features(num_qubits, prec, arr.arr, arr.n, C.byref(opt), _p(out, C.c_float), rows.value,
                                   C.byref(rows)))
    return out[: rows.value]


def equivalent_variants(num_qubits: int, ops, precision: str = "single", variant: int = 0,
                        max_fused_qubits: int = 4) -> list:
    """The plan variants whose programs equal variant `variant`'s byte for byte
    (bit-identical amplitudes): the candidates of PLANNER_LEARNED."""
    arr = ops if isinstance(ops, OpArray) else OpArray(ops)
    opt = Options(max_fused_qubits, 0, PLANNER_HEURISTIC, 0, variant)
    mask = C.c_uint32()
    _check(lib().qvd_plan_equivalent_variants(num_qubits, {"single": C64, "double": C128}[precision], arr.arr, arr.n,
                                              C.byref(opt), C.byref(mask)))
    return [v for v in range(PLAN_VARIANTS) if (mask.value >> v) & 1]


def plan_ops(num_qubits: int, ops, precision: str = "single", max_fused_qubits: int = 4, tile_bits: int | None = 0,
             planner: int = PLANNER_HEURISTIC, variant: int = 0, fixed_lanes: int = 0) -> dict:
    """Plan statistics (stages, passes, fused ops, geometry) of apply_ops, on the host only."""
    arr = ops if isinstance(ops, OpArray) else OpArray(ops)
    opt = Options(max_fused_qubits, -1 if tile_bits is None else tile_bits, planner, 0, variant, fixed_lanes)
    st = Stats()
    _check(lib().qvd_plan_ops(num_qubits, {"single": C64, "double": C128}[precision], arr.arr, arr.n, C.byref(opt), C.byref(st)))
    return st.as_dict()


def pauli_masks(label: str) -> tuple:
    """(x_mask, z_mask) of a Pauli label in Qiskit's order (the last
    character acts on qubit 0)."""
    x = z = 0
    for q, ch in enumerate(reversed(label.upper())):
        if ch in "XY":
            x |= 1 << q
        if ch in "ZY":
            z |= 1 << q
        if ch not in "IXYZ":
            raise ValueError(f"Pauli label {label!r}: {ch!r} is not one of I, X, Y, Z")
    return x, z


def pauli_terms(observable, num_qubits: int) -> tuple:
    """Labels and coefficients of an observable given as a {label:
    coefficient} mapping or a sequence of (label, coefficient) pairs."""
    items = list(observable.items()) if hasattr(observable, "items") else list(observable)
    labels, coeffs = [], []
    for label, c in items:
        if len(label) != num_qubits:
            raise ValueError(f"Pauli label {label!r} has {len(label)} qubits, the state {num_qubits}")
        labels.append(label)
        coeffs.append(complex(c))
    return labels, coeffs


class StateVector:
    """A state of `num_qubits` qubits in HBM (StateVector, src/state.rs).

    precision: "single" (complex64) or "double" (complex128).
    """

    def __init__(self, num_qubits: int, precision: str = "single"):
        prec = {"single": C64, "double": C128}.get(precision)
        if prec is None:
            raise ValueError('precision must be "single" or "double"')
        self.num_qubits = int(num_qubits)
        self.precision = precision
        h = C.c_void_p()
        _check(lib().qvd_sv_create(self.num_qubits, prec, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().qvd_sv_destroy(self.h)
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

    @property
    def size(self) -> int:
        return 1 << self.num_qubits

    def set_zero(self) -> None:
        _check(lib().qvd_sv_set_zero(self.h))

    def copy_from(self, other: "StateVector") -> None:
        """This state becomes a copy of `other` (same size and precision), on the device."""
        _check(lib().qvd_sv_copy(self.h, other.h))

    def set_amplitudes(self, amps) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        _check(lib().qvd_sv_set_amplitudes(self.h, _p(a.view(np.float64), C.c_double), a.size))

    def amplitudes(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.complex128)
        _check(lib().qvd_sv_get_amplitudes(self.h, _p(out.view(np.float64), C.c_double), out.size))
        return out

    def apply(self, matrix, targets, controls=()) -> None:
        m = np.ascontiguousarray(matrix, dtype=np.complex128)
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        cq = np.ascontiguousarray([c[0] for c in controls], dtype=np.uint32)
        cv = np.ascontiguousarray([1 if c[1] else 0 for c in controls], dtype=np.uint8)
        _check(lib().qvd_sv_apply(self.h, _p(m.view(np.float64), C.c_double), _p(t, C.c_uint32), t.size, _p(cq, C.c_uint32), _p(cv, C.c_uint8), cq.size))

    def apply_ops(self, ops, max_fused_qubits: int = 4, tile_bits: int | None = 0, planner: int = PLANNER_HEURISTIC,
                  timing: bool = False, variant: int = 0, dense_valu: bool = False, all_tiles: bool = False,
                  fixed_lanes: int = 0, autotune: bool = False) -> dict:
        """ops: [(targets, [(control, value)], matrix)]; tile_bits None = no blocking, 0 = default.

        planner PLANNER_LEARNED scores, with the installed cost model, the
        plan variants that run the heuristic's arithmetic byte for byte
        (bit-identical amplitudes); PLANNER_LEARNED_ANY scores every variant
        (other op orders round differently); both are the heuristic when no
        model is installed. variant picks a heuristic variant directly.
        all_tiles: visit the register tiles known to be zero too
        (QVD_FLAG_ALL_TILES). fixed_lanes: the register-tile geometry (lane
        bits pinned to the lowest qubits; 0 = the default); autotune: pick it
        by timing, once per circuit class (QVD_FLAG_AUTOTUNE). The stats'
        fixed_lanes is the geometry that ran."""
        arr = ops if isinstance(ops, OpArray) else OpArray(ops)
        flags = ((FLAG_TIMING if timing else 0) | (FLAG_DENSE_VALU if dense_valu else 0)
                 | (FLAG_ALL_TILES if all_tiles else 0) | (FLAG_AUTOTUNE if autotune else 0))
        opt = Options(max_fused_qubits, -1 if tile_bits is None else tile_bits, planner, flags, variant, fixed_lanes)
        st = Stats()
        _check(lib().qvd_sv_apply_ops(self.h, arr.arr, arr.n, C.byref(opt), C.byref(st)))
        return st.as_dict()

    def stage_seconds(self) -> np.ndarray:
        """Kernel time of each register-tile stage of the last timed apply_ops."""
        n = C.c_uint64()
        _check(lib().qvd_sv_stage_seconds(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.float64)
        _check(lib().qvd_sv_stage_seconds(self.h, _p(out, C.c_double), n.value, C.byref(n)))
        return out[: n.value]

    def norm_sqr(self) -> float:
        v = C.c_double()
        _check(lib().qvd_sv_norm_sqr(self.h, C.byref(v)))
        return v.value

    def scale(self, factor: float) -> None:
        _check(lib().qvd_sv_scale(self.h, factor))

    def probability_one(self, qubit: int) -> float:
        v = C.c_double()
        _check(lib().qvd_sv_probability_one(self.h, qubit, C.byref(v)))
        return v.value

    def probabilities(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.float64)
        _check(lib().qvd_sv_probabilities(self.h, _p(out, C.c_double), out.size))
        return out

    def expect_z(self) -> tuple:
        n = self.num_qubits
        z = np.zeros(max(n, 1), dtype=np.float64)
        zz = np.zeros(max(n - 1, 1), dtype=np.float64)
        _check(lib().qvd_sv_expect_z(self.h, _p(z, C.c_double), _p(zz, C.c_double)))
        return z[:n], zz[: max(n - 1, 0)]

    def expect_pauli(self, x_masks, z_masks) -> np.ndarray:
        """<P_t> in FP64 for Pauli strings given as bit masks (bit q = qubit
        q): X on x-only qubits, Z on z-only qubits, Y on qubits in both."""
        x = np.ascontiguousarray(x_masks, dtype=np.uint64)
        z = np.ascontiguousarray(z_masks, dtype=np.uint64)
        if x.shape != z.shape:
            raise ValueError("x_masks and z_masks must have the same length")
        out = np.zeros(max(x.size, 1), dtype=np.float64)
        _check(lib().qvd_sv_expect_pauli(self.h, _p(x, C.c_uint64), _p(z, C.c_uint64), x.size, _p(out, C.c_double)))
        return out[: x.size]

    def expectation_value(self, observable) -> float:
        """sum_k c_k <P_k> of an observable given as {label: coefficient} or
        [(label, coefficient)], with Qiskit's Pauli labels (characters I, X,
        Y, Z; the last character is qubit 0), as Statevector.expectation_value
        of a SparsePauliOp (python/python/qevedo/simulator/qiskit.py:584-589)."""
        labels, coeffs = pauli_terms(observable, self.num_qubits)
        x, z = zip(*(pauli_masks(label) for label in labels)) if labels else ((), ())
        values = self.expect_pauli(list(x), list(z))
        return float(np.real(np.sum(np.asarray(coeffs) * values)))

    def measure(self, qubit: int, uniform: float) -> int:
        o = C.c_uint8()
        _check(lib().qvd_sv_measure(self.h, qubit, uniform, C.byref(o)))
        return o.value

    def reset(self, qubit: int, uniform: float) -> None:
        _check(lib().qvd_sv_reset(self.h, qubit, uniform))

    def sample(self, uniforms, out=None) -> np.ndarray:
        """Outcomes for the uniforms (sample_unordered, measure.rs:104-174).
        `out`, a C-contiguous uint64 array of one element per uniform, is
        filled and returned instead of a new array: a caller sampling
        repeatedly saves the host's first-touch page faults of a new 8-byte-
        per-shot array, which the copy of the outcomes would otherwise pay."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        if out is None:
            out = np.empty(u.size, dtype=np.uint64)
        elif (not isinstance(out, np.ndarray) or out.dtype != np.uint64 or out.size != u.size
              or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError("out must be a writeable C-contiguous uint64 array with one element per uniform")
        _check(lib().qvd_sv_sample(self.h, _p(u, C.c_double), u.size, _p(out, C.c_uint64)))
        return out

    def compare(self, reference: "StateVector") -> dict:
        """max_abs and relative L2 distance to `reference`, computed on the GPU."""
        m, d, r = C.c_double(), C.c_double(), C.c_double()
        _check(lib().qvd_sv_compare(self.h, reference.h, C.byref(m), C.byref(d), C.byref(r)))
        return {"max_abs": m.value, "rel_l2": (d.value / r.value) ** 0.5 if r.value > 0 else float("nan")}

    def last_sample_seconds(self) -> float:
        v = C.c_double()
        _check(lib().qvd_sv_last_sample_seconds(self.h, C.byref(v)))
        return v.value


class Batch:
    """Many circuits, each on its own state, run together (qvd_batch_*).

    circuits: a list of (num_qubits, lowered ops) or (num_qubits, lowered ops,
    fixed_lanes), fixed_lanes being the register-tile geometry to build the
    circuit with (0 = default). autotune: time the candidate geometries per
    class of circuits (qvd_batch_create with QVD_FLAG_AUTOTUNE). run()
    returns, per circuit, (<Z_q> for every qubit, <Z_q Z_{q+1}> for
    neighbouring qubits).
    """

    def __init__(self, circuits, precision: str = "single", max_fused_qubits: int = 4, autotune: bool = False,
                 all_tiles: bool = False):
        self.h = None
        self.sizes = [int(c[0]) for c in circuits]
        self.arrays = [c[1] if isinstance(c[1], OpArray) else OpArray(c[1]) for c in circuits]
        cs = (CircuitDesc * len(self.sizes))()
        for i, (c, arr) in enumerate(zip(circuits, self.arrays)):
            cs[i] = CircuitDesc(int(c[0]), arr.arr, arr.n, int(c[2]) if len(c) > 2 else 0)
        opt = Options(max_fused_qubits, 0, PLANNER_HEURISTIC,
                      (FLAG_AUTOTUNE if autotune else 0) | (FLAG_ALL_TILES if all_tiles else 0), 0)
        h = C.c_void_p()
        _check(lib().qvd_batch_create(cs, len(self.sizes), {"single": C64, "double": C128}[precision], C.byref(opt),
                                      C.byref(h)))
        self.h = h
        nz, nzz = C.c_uint64(), C.c_uint64()
        _check(lib().qvd_batch_outputs(self.h, C.byref(nz), C.byref(nzz)))
        self.num_z, self.num_zz = nz.value, nzz.value
        # Each circuit's slice of the run's <Z> and <ZZ> arrays.
        self._z_off = np.cumsum([0] + self.sizes).tolist()
        self._zz_off = np.cumsum([0] + [n - 1 for n in self.sizes]).tolist()
        self.stats = {}
        self.precision = precision

    def fixed_lanes(self) -> list:
        """The register-tile geometry (fixed lane bits) of each circuit."""
        out = (C.c_uint32 * len(self.sizes))()
        _check(lib().qvd_batch_geometry(self.h, out, None, len(self.sizes)))
        return list(out)

    def stages(self) -> list:
        """The number of register-tile stages of each circuit."""
        out = (C.c_uint32 * len(self.sizes))()
        _check(lib().qvd_batch_geometry(self.h, None, out, len(self.sizes)))
        return list(out)

    def amplitudes(self, index: int) -> np.ndarray:
        """Circuit `index`'s state after the last run, in the batch's precision."""
        out = np.empty(1 << self.sizes[index], dtype=np.complex64 if self.precision == "single" else np.complex128)
        _check(lib().qvd_batch_amplitudes(self.h, index, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    def run(self) -> list:
        """Per circuit, (<Z_q> for every qubit, <Z_q Z_{q+1}>): views into
        arrays that this run allocates, so later runs leave them as they are."""
        z = np.zeros(max(self.num_z, 1), dtype=np.float64)
        zz = np.zeros(max(self.num_zz, 1), dtype=np.float64)
        st = Stats()
        _check(lib().qvd_batch_run(self.h, _p(z, C.c_double), _p(zz, C.c_double), C.byref(st)))
        self.stats = st.as_dict()
        oz, ozz = self._z_off, self._zz_off
        return [(z[oz[i]:oz[i + 1]], zz[ozz[i]:ozz[i + 1]]) for i in range(len(self.sizes))]

    def close(self) -> None:
        if self.h:
            lib().qvd_batch_destroy(self.h)
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


class Mlp:
    """The planner's cost model network on the device (qvd_mlp_*): inputs ->
    hidden -> hidden -> 1 with ReLU, FP32, trained by the library's kernels."""

    def __init__(self, inputs: int, hidden: int = 64):
        self.inputs, self.hidden = inputs, hidden
        h = C.c_void_p()
        _check(lib().qvd_mlp_create(inputs, hidden, C.byref(h)))
        self.h = h
        n = C.c_uint64()
        _check(lib().qvd_mlp_num_params(self.h, C.byref(n)))
        self.num_params = n.value

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().qvd_mlp_set_params(self.h, _p(p, C.c_float), p.size))

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, dtype=np.float32)
        _check(lib().qvd_mlp_get_params(self.h, _p(p, C.c_float), p.size))
        return p

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.inputs)
        y = np.zeros(max(len(x), 1), dtype=np.float32)
        _check(lib().qvd_mlp_forward(self.h, _p(x, C.c_float), len(x), _p(y, C.c_float)))
        return y[: len(x)]

    def gradient(self, x, t) -> tuple:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.inputs)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = C.c_double()
        _check(lib().qvd_mlp_gradient(self.h, _p(x, C.c_float), _p(t, C.c_float), len(x), _p(g, C.c_float),
                                      C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.inputs)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_double()
        _check(lib().qvd_mlp_train(self.h, _p(x, C.c_float), _p(t, C.c_float), len(x), epochs, batch, lr, seed,
                                   C.byref(loss)))
        return loss.value

    def install(self, mean, scale, y_mean: float, y_scale: float) -> None:
        """Make this network the learned planner's cost model."""
        mean = np.ascontiguousarray(mean, dtype=np.float32)
        scale = np.ascontiguousarray(scale, dtype=np.float32)
        _check(lib().qvd_planner_set_model(self.h, _p(mean, C.c_float), _p(scale, C.c_float), y_mean, y_scale))

    def close(self) -> None:
        if self.h:
            lib().qvd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clear_planner_model() -> None:
    """Remove the learned planner's cost model (it then falls back to the heuristic)."""
    _check(lib().qvd_planner_set_model(None, None, None, 0.0, 1.0))
