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
    n, gates = CIRCUITS[0][1], CIRCUITS[0][2]
    ops = _native.OpArray(lower_gates(gates))
    stages = [_native.plan_ops(n, ops, variant=v)["stages"] for v in range(4)]
    assert stages == sorted(stages)


def test_variant_out_of_range_is_an_error():
    ops = _native.OpArray(lower_gates(library.qft(16)))
    with pytest.raises(NativeError, match="variant"):
        _native.plan_ops(16, ops, variant=_native.PLAN_VARIANTS)
