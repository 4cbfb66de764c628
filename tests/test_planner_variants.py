"""Plan variants and the learned planner's stage features, on the host (no
device): every variant plans every op, variant 0 is the heuristic, feature
rows match the register-tile stages, and out-of-range variants are errors."""
import numpy as np
import pytest

from qevedo.simulator import NativeError, library, lower_gates
from qevedo.simulator import _native

CIRCUITS = [
    ("random", 30, library.random_circuit(30, 20, 42)),
    ("qft", 30, library.qft(30)),
    ("ising", 33, library.kicked_ising(33, library.heavy_hex_edges(33), 6, 0.7)),
    ("hea", 22, library.hardware_efficient_ansatz(22, 6, 3)),
]


@pytest.mark.parametrize("name,n,gates", CIRCUITS)
def test_every_variant_plans_every_op(name, n, gates):
    ops = _native.OpArray(lower_gates(gates))
    base = _native.plan_ops(n, ops)
    assert base["plan_variant"] == 0
    again = _native.plan_ops(n, ops, variant=0)
    assert {k: v for k, v in base.items() if not k.endswith("seconds")} == \
        {k: v for k, v in again.items() if not k.endswith("seconds")}
    for v in range(_native.PLAN_VARIANTS):
        st = _native.plan_ops(n, ops, variant=v)
        assert st["gates"] == len(gates)
        assert st["kernels"] == len(gates)  # register-tile stages keep ops unfused
        assert st["plan_variant"] == v
        f = _native.plan_features(n, ops, variant=v)
        assert f.shape == (st["stages"], _native.PLAN_FEATURES)
        assert np.all(np.isfinite(f))
        assert np.allclose(f[:, 0], n / 32)
        # at least one instruction per stage, and the counts are log(1 + count)
        assert np.all(np.expm1(f[:, 23]) >= 1 - 1e-5)


def test_capacity_cut_never_plans_fewer_stages():
    # variants 24..27: the lookahead's tile alone (bit 4), register slots in
    # order of need (bit 3), capacity cut by 0..3 qubits (bits 0-1)
    n, gates = CIRCUITS[0][1], CIRCUITS[0][2]
    ops = _native.OpArray(lower_gates(gates))
    stages = [_native.plan_ops(n, ops, variant=v)["stages"] for v in range(24, 28)]
    assert stages == sorted(stages)


def test_region_tiles_plan_fewer_stages_than_the_lookahead_alone():
    # variant 0 (regions of the coupling graph) against variant 24 (the
    # lookahead's tile only, as src/blocking.rs:181-226 picks local qubits)
    for name, n, gates in CIRCUITS:
        ops = _native.OpArray(lower_gates(gates))
        assert _native.plan_ops(n, ops, variant=0)["stages"] <= _native.plan_ops(n, ops, variant=24)["stages"], name


@pytest.mark.parametrize("name,n,gates", CIRCUITS)
@pytest.mark.parametrize("precision", ["single", "double"])
def test_equivalent_variants_include_the_base_and_agree_both_ways(name, n, gates, precision):
    ops = _native.OpArray(lower_gates(gates))
    eq = _native.equivalent_variants(n, ops, precision)
    assert 0 in eq
    # equality of plan images is symmetric and transitive: every member sees
    # the same set
    for v in eq[1:3]:
        assert _native.equivalent_variants(n, ops, precision, variant=v) == eq
    outside = [v for v in range(_native.PLAN_VARIANTS) if v not in eq]
    if outside:
        assert 0 not in _native.equivalent_variants(n, ops, precision, variant=outside[0])


def test_gate_order_on_different_qubits_changes_float32_rounding():
    """Why the candidates of the learned planner must run the same programs:
    two 1-qubit gates on different qubits commute exactly, but applied in the
    other order their float32 arithmetic rounds differently, so a plan that
    reorders them cannot reproduce another plan's amplitudes bit for bit."""
    r = np.random.Generator(np.random.PCG64(11))
    psi = (r.standard_normal((2, 2, 2)) + 1j * r.standard_normal((2, 2, 2))).astype(np.complex64)

    def rot(theta, axis):
        c, s = np.cos(theta / 2), np.sin(theta / 2)
        m = np.array([[c, -1j * s], [-1j * s, c]]) if axis == "x" else np.array([[c, -s], [s, c]])
        return m.astype(np.complex64)

    def apply(state, m, axis):
        # float32 complex arithmetic, one gate at a time
        return np.moveaxis(np.tensordot(m, np.moveaxis(state, axis, 0), axes=1), 0, axis).astype(np.complex64)

    a, b = rot(0.7, "x"), rot(1.9, "y")
    ab = apply(apply(psi, a, 2), b, 1)
    ba = apply(apply(psi, b, 1), a, 2)
    assert np.allclose(ab, ba, atol=1e-6)
    assert not np.array_equal(ab, ba)


def test_variant_out_of_range_is_an_error():
    ops = _native.OpArray(lower_gates(library.qft(16)))
    with pytest.raises(NativeError, match="variant"):
        _native.plan_ops(16, ops, variant=_native.PLAN_VARIANTS)


@pytest.mark.parametrize("variant", [0, 4, 16, 24])
def test_op_wider_than_a_stage_gets_its_own_pass(variant):
    """An op on more targets than a stage applies (5 here) runs as a
    full-state pass between stages under every tile-choice rule."""
    import numpy as np
    from helpers import random_unitary, rng

    m = random_unitary(rng(3), 5)
    ops = [([0, 3, 6, 9, 12], [], m), ([1], [], np.eye(2)), ([0, 4, 8, 12, 15], [], m)]
    st = _native.plan_ops(16, ops, variant=variant)
    assert st["gates"] == 3 and st["passes"] >= 2 and st["passes"] - st["stages"] == 2


def test_geometry_of_host_plans():
    """Register-tile geometry on the host: the default and one fixed lane bit
    fewer or more plan (and are reported); others are refused; features
    follow the geometry (one row per register-tile stage)."""
    ops = _native.OpArray(lower_gates(library.random_circuit(30, 20, 42)))
    stages = {}
    for f, prec in ((3, "single"), (4, "single"), (5, "single"), (2, "double"), (3, "double"), (4, "double")):
        st = _native.plan_ops(30, ops, prec, fixed_lanes=f)
        assert st["fixed_lanes"] == f and st["stages"] > 0
        assert _native.plan_features(30, ops, prec, fixed_lanes=f).shape[0] == st["stages"]
        stages[(f, prec)] = st["stages"]
    assert _native.plan_ops(30, ops, "single")["fixed_lanes"] == 4
    assert _native.plan_ops(30, ops, "double")["fixed_lanes"] == 3
    for f, prec in ((2, "single"), (6, "single"), (1, "double"), (5, "double")):
        with pytest.raises(_native.NativeError):
            _native.plan_ops(30, ops, prec, fixed_lanes=f)
