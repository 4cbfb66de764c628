"""|0...0> is written lazily: qvd_sv_create and qvd_sv_set_zero only mark the
state, the next apply_ops whose plan starts with a register-tile stage makes
that stage's tiles from |0...0> in registers, and every other use writes the
zeros first. These tests check that each path gives the state that writing
the zeros eagerly gives: bit for bit against a run from explicitly uploaded
zeros (the loading path), and against the oracle's FP64 state."""
import numpy as np
import pytest

from helpers import oracle_statevector
from qevedo.simulator import StateVector, library, lower_gates
from qevedo.simulator import _native

pytestmark = pytest.mark.gpu


def _zeros(n: int) -> np.ndarray:
    z = np.zeros(1 << n, dtype=np.complex128)
    z[0] = 1
    return z


@pytest.mark.parametrize("precision", ["single", "double"])
def test_first_stage_from_zero_equals_loaded_zeros(precision):
    """The same plan from a pending |0...0> (first stage made in registers)
    and from uploaded zeros (first stage loaded) gives identical amplitudes,
    within the oracle's tolerance of the FP64 state."""
    n = 18
    gates = library.random_circuit(n, 12, 77)
    ops = _native.OpArray(lower_gates(gates))
    with StateVector(n, precision) as a, StateVector(n, precision) as b:
        a.set_zero()
        a.apply_ops(ops)
        b.set_amplitudes(_zeros(n))
        b.apply_ops(ops)
        got_a, got_b = a.amplitudes(), b.amplitudes()
    assert np.array_equal(got_a, got_b)
    want = oracle_statevector(n, gates)
    tol = 1e-5 if precision == "single" else 1e-12
    assert np.linalg.norm(got_a - want) / np.linalg.norm(want) < tol


def test_every_use_of_a_pending_zero_sees_zero():
    """A fresh or reset state read, measured, sampled, copied, scaled or run
    through a plan that does not start with a register-tile stage is
    |0...0>; a second apply_ops continues from the first one's result."""
    n = 15
    with StateVector(n, "single") as sv:
        assert np.array_equal(sv.amplitudes(), _zeros(n))  # fresh handle
        sv.apply_ops(_native.OpArray(lower_gates([("h", (q,), ()) for q in range(n)])))
        sv.set_zero()
        assert sv.norm_sqr() == pytest.approx(1.0)
        assert sv.probabilities()[0] == 1.0
        sv.set_zero()
        z, zz = sv.expect_z()
        assert np.array_equal(z, np.ones(n)) and np.array_equal(zz, np.ones(n - 1))
        sv.set_zero()
        assert list(sv.sample(np.linspace(0.05, 0.95, 7))) == [0] * 7
        sv.set_zero()
        assert sv.measure(3, 0.5) == 0
        sv.set_zero()
        with StateVector(n, "single") as other:
            other.apply_ops(_native.OpArray(lower_gates([("x", (2,), ())])))
            other.copy_from(sv)  # a pending zero is copied
            assert np.array_equal(other.amplitudes(), _zeros(n))
        # an unblocked plan (no register-tile stage) writes the zeros first
        gates = [("h", (0,), ()), ("cx", (0, 5), ()), ("ry", (7,), (0.3,))]
        sv.set_zero()
        sv.apply_ops(_native.OpArray(lower_gates(gates)), tile_bits=None)
        assert np.max(np.abs(sv.amplitudes() - oracle_statevector(n, gates))) < 1e-6
        # two applies in a row: the second starts from the first's state
        first = library.random_circuit(n, 6, 5)
        second = library.random_circuit(n, 6, 6)
        sv.set_zero()
        sv.apply_ops(_native.OpArray(lower_gates(first)))
        sv.apply_ops(_native.OpArray(lower_gates(second)))
        want = oracle_statevector(n, first + second)
        assert np.linalg.norm(sv.amplitudes() - want) < 1e-5
        # an empty plan leaves the pending zero as it is
        sv.set_zero()
        sv.apply_ops(_native.OpArray([]))
        assert np.array_equal(sv.amplitudes(), _zeros(n))
