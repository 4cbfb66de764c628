"""Register-tile geometry of single states: the number F of lane bits pinned
to the lowest qubits (qvd_options.fixed_lanes), given or picked by timing
(QVD_FLAG_AUTOTUNE). Every geometry is checked against the oracle's FP64
state; the autotuner's result equals a run with the geometry it reports, bit
for bit, and its timing probes leave the caller's state alone."""
import numpy as np
import pytest

from helpers import oracle_statevector, rel_l2
from qevedo.simulator import StateVector, library, lower_gates
from qevedo.simulator import _native

pytestmark = pytest.mark.gpu

DEFAULT = {"single": 4, "double": 3}
TOL = {"single": 1e-5, "double": 1e-12}


@pytest.mark.parametrize("precision", ["single", "double"])
def test_every_geometry_matches_the_oracle(precision):
    n = 18
    gates = library.random_circuit(n, 10, 21)
    ops = _native.OpArray(lower_gates(gates))
    want = oracle_statevector(n, gates)
    d = DEFAULT[precision]
    for f in (d - 1, d, d + 1):
        print(f"geometry {precision} F={f}", flush=True)
        with StateVector(n, precision) as sv:
            st = sv.apply_ops(ops, fixed_lanes=f)
            got = sv.amplitudes()
        assert st["fixed_lanes"] == f and st["stages"] > 0
        assert rel_l2(got, want) <= TOL[precision], f


def test_geometry_out_of_range_is_an_error():
    with StateVector(16, "single") as sv:
        with pytest.raises(_native.NativeError):
            sv.apply_ops(_native.OpArray(lower_gates(library.random_circuit(16, 2, 1))), fixed_lanes=6)


def test_learned_plan_keeps_the_amplitudes_at_another_geometry():
    """The learned planner compares plan images built for the geometry that
    runs: at F = 5 its plan gives the heuristic's amplitudes bit for bit."""
    n = 20
    ops = _native.OpArray(lower_gates(library.random_circuit(n, 12, 8)))
    with StateVector(n, "single") as a, StateVector(n, "single") as b:
        a.apply_ops(ops, fixed_lanes=5)
        b.apply_ops(ops, fixed_lanes=5, planner=_native.PLANNER_LEARNED)
        assert np.array_equal(a.amplitudes(), b.amplitudes())


@pytest.mark.parametrize("from_zero", [True, False])
def test_autotuned_geometry_runs_as_reported(from_zero):
    """The autotuned run equals a run with the geometry it reports; a second
    call of the same class reuses the choice; a state that is not |0...0>
    (its probes start from the same known-zero qubits) comes out right."""
    n = 20
    first = library.random_circuit(n, 3, 30)
    # a class of its own per case (classes go by targets and controls)
    gates = library.random_circuit(n, 12 if from_zero else 13, 31)
    ops = _native.OpArray(lower_gates(gates))
    with StateVector(n, "single") as sv, StateVector(n, "single") as ref:
        if not from_zero:
            sv.apply_ops(_native.OpArray(lower_gates(first)))
            ref.apply_ops(_native.OpArray(lower_gates(first)))
        st = sv.apply_ops(ops, autotune=True)
        f = st["fixed_lanes"]
        assert f in (3, 4, 5)
        ref.apply_ops(ops, fixed_lanes=f)
        got = sv.amplitudes()
        assert np.array_equal(got, ref.amplitudes())
        want = oracle_statevector(n, gates if from_zero else first + gates)
        assert rel_l2(got, want) <= 1e-5
        sv.set_zero()
        again = sv.apply_ops(ops, autotune=True)
        assert again["fixed_lanes"] == f
