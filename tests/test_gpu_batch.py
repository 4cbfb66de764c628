"""Batches of circuits (qvd_batch_*): many circuits of uneven sizes run by one
graph-captured launch of the stage-queue kernel, checked circuit by circuit
against the oracle's FP64 state vector and against the same circuits run one
by one on single states (bit for bit where both run the same stage
programs)."""
import numpy as np
import pytest

from helpers import oracle_statevector
from qevedo.simulator import Batch, NativeError, StateVector, library, lower_gates
from qevedo.simulator import _native

pytestmark = pytest.mark.gpu


def z_from_state(psi: np.ndarray, n: int):
    """<Z_q> and <Z_q Z_{q+1}> of an FP64 state, from the definition."""
    p = np.abs(psi) ** 2
    idx = np.arange(p.size)
    bits = [(idx >> q) & 1 for q in range(n)]
    z = np.array([np.sum(p * (1 - 2 * b)) for b in bits])
    zz = np.array([np.sum(p * (1 - 2 * (bits[q] ^ bits[q + 1]))) for q in range(n - 1)])
    return z, zz


def hea(n: int, seed: int):
    return library.hardware_efficient_ansatz(n, 6, seed)


@pytest.mark.parametrize("precision,tol", [("single", 2e-5), ("double", 1e-12)])
def test_batch_matches_oracle(precision, tol):
    sizes = [1, 2, 5, 9, 12, 13, 14, 15, 16, 3, 17]
    circuits = [(n, lower_gates(hea(n, 100 + i))) for i, n in enumerate(sizes)]
    with Batch(circuits, precision) as b:
        out = b.run()
        assert b.stats["gates"] == sum(len(c[1]) for c in circuits)
        assert b.stats["passes"] >= 1
    for i, n in enumerate(sizes):
        print(f"circuit {i}: {n} qubits", flush=True)
        z_ref, zz_ref = z_from_state(oracle_statevector(n, hea(n, 100 + i)), n)
        z, zz = out[i]
        assert z.shape == (n,) and zz.shape == (n - 1,)
        assert np.max(np.abs(z - z_ref)) < tol
        if n > 1:
            assert np.max(np.abs(zz - zz_ref)) < tol


def test_batch_matches_one_by_one_and_replays():
    rng = np.random.Generator(np.random.PCG64(5))
    sizes = [int(x) for x in rng.integers(12, 23, size=24)]
    circuits = [(n, lower_gates(hea(n, 7 + i))) for i, n in enumerate(sizes)]
    with Batch(circuits, "single") as b:
        first = b.run()
        second = b.run()
    for (z1, zz1), (z2, zz2) in zip(first, second):
        assert np.array_equal(z1, z2) and np.array_equal(zz1, zz2)
    for i, (n, ops) in enumerate(circuits):
        sv = StateVector(n, "single")
        sv.apply_ops(_native.OpArray(ops))
        z, zz = sv.expect_z()
        sv.close()
        assert np.max(np.abs(first[i][0] - z)) < 1e-5
        assert np.max(np.abs(first[i][1] - zz)) < 1e-5


def test_batch_errors():
    ok = (4, lower_gates(hea(4, 1)))
    with pytest.raises(NativeError, match="out of range"):
        Batch([ok, (3, lower_gates([("h", (3,), ())]))])
    with pytest.raises(NativeError, match="fixed_lanes"):
        Batch([(15, lower_gates(hea(15, 1)), 6)])
    with pytest.raises(NativeError, match="fixed_lanes"):
        Batch([(15, lower_gates(hea(15, 1)), 5)], "double")
    # a 5-qubit dense op cannot join a register-tile stage
    m = np.eye(32, dtype=np.complex128)
    with pytest.raises(NativeError) as e:
        Batch([(16, [([0, 3, 6, 9, 12], [], m[::-1].copy())])])
    assert e.value.code == _native.QVD_EUNSUPPORTED


def _alone(n: int, ops, precision: str) -> np.ndarray:
    with StateVector(n, precision) as sv:
        sv.apply_ops(_native.OpArray(ops))
        return sv.amplitudes()


@pytest.mark.parametrize("precision", ["single", "double"])
def test_deep_circuit_alone_equals_single_state_bitwise(precision):
    """One 22-qubit circuit of many stages: every stage after the first waits
    in the queue for tiles that other workgroups (on other XCDs) wrote. The
    result equals the single-state run bit for bit over several replays."""
    n = 22
    ops = lower_gates(library.random_circuit(n, 20, 11))
    want = _alone(n, ops, precision)
    with Batch([(n, ops)], precision) as b:
        b.run()
        assert b.stats["passes"] >= 4
        for replay in range(3):
            if replay:
                b.run()
            got = b.amplitudes(0)
            print(f"replay {replay}", flush=True)
            assert np.array_equal(got.astype(np.complex128), want)


def test_mixed_batch_equals_single_states_bitwise():
    """Circuits of 14-22 qubits with 1-5 stages each (random and hardware-
    efficient), replayed: each state equals its single-state run bit for bit
    (both plan the same stages), and the first stage of each is made from
    |0...0> in registers."""
    rng = np.random.Generator(np.random.PCG64(12))
    circuits = []
    for i, n in enumerate(int(x) for x in rng.integers(14, 23, size=40)):
        g = library.random_circuit(n, 12, 300 + i) if i % 2 else hea(n, 300 + i)
        circuits.append((n, lower_gates(g)))
    with Batch(circuits, "single") as b:
        b.run()
        b.run()
        for i, (n, ops) in enumerate(circuits):
            print(f"circuit {i}: {n} qubits", flush=True)
            assert np.array_equal(b.amplitudes(i).astype(np.complex128), _alone(n, ops, "single"))


def test_batch_with_empty_circuits():
    """Circuits with no gates keep |0...0> (they have no stage, so the batch
    initialises their states), beside circuits that have stages."""
    circuits = [(3, []), (16, lower_gates(hea(16, 4))), (15, []), (1, [])]
    with Batch(circuits, "single") as b:
        out = b.run()
        out = b.run()
        for i in (0, 2, 3):
            n = circuits[i][0]
            z, zz = out[i]
            assert np.array_equal(z, np.ones(n)) and np.array_equal(zz, np.ones(n - 1))
            a = b.amplitudes(i)
            assert a[0] == 1 and not np.any(a[1:])
        z_ref, zz_ref = z_from_state(oracle_statevector(16, hea(16, 4)), 16)
        assert np.max(np.abs(out[1][0] - z_ref)) < 2e-5 and np.max(np.abs(out[1][1] - zz_ref)) < 2e-5


@pytest.mark.parametrize("precision,lanes,tol", [("single", (3, 4, 5), 2e-5), ("double", (2, 3, 4), 1e-12)])
def test_every_geometry_matches_the_oracle(precision, lanes, tol):
    """Each register-tile geometry the queue kernel accepts (fixed lane bits
    default - 1 .. default + 1, read per stage at run time) gives the oracle's
    FP64 state, for random and hardware-efficient circuits of 14-19 qubits,
    all in one batch."""
    cases = []
    for k, n in enumerate((14, 16, 19)):
        for f in lanes:
            g = library.random_circuit(n, 10, 40 + k) if k % 2 else hea(n, 40 + k)
            cases.append((n, g, f))
    with Batch([(n, lower_gates(g), f) for n, g, f in cases], precision) as b:
        b.run()
        assert b.fixed_lanes() == [f for _, _, f in cases]
        for i, (n, g, f) in enumerate(cases):
            print(f"circuit {i}: {n} qubits, {f} fixed lanes", flush=True)
            want = oracle_statevector(n, g)
            assert np.max(np.abs(b.amplitudes(i).astype(np.complex128) - want)) < tol


def test_autotuned_batch():
    """QVD_FLAG_AUTOTUNE: every circuit gets a geometry from the candidates,
    circuits of one class (size and gate layout) the same one, a second batch
    of the same classes reuses the choices, and the results agree with the
    default geometry's."""
    rng = np.random.Generator(np.random.PCG64(21))
    sizes = [int(x) for x in rng.integers(14, 23, size=48)]
    circuits = [(n, lower_gates(hea(n, 500 + i))) for i, n in enumerate(sizes)]
    with Batch(circuits, "single", autotune=True) as b:
        tuned = b.fixed_lanes()
        out = b.run()
    assert set(tuned) <= {3, 4, 5}
    for n in set(sizes):
        assert len({f for f, m in zip(tuned, sizes) if m == n}) == 1
    with Batch(circuits, "single", autotune=True) as b2:
        assert b2.fixed_lanes() == tuned
    with Batch(circuits, "single") as b3:
        assert b3.fixed_lanes() == [4] * len(circuits)
        ref = b3.run()
    for (z, zz), (rz, rzz) in zip(out, ref):
        assert np.max(np.abs(z - rz)) < 1e-5 and np.max(np.abs(zz - rzz)) < 1e-5
