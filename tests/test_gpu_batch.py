"""Batches of circuits (qvd_batch_*): many hardware-efficient-ansatz circuits
of uneven sizes in one graph-captured sequence of stage levels, checked
circuit by circuit against the oracle's FP64 state vector and against the
same circuits run one by one on single states."""
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
    # a 5-qubit dense op cannot join a register-tile stage
    m = np.eye(32, dtype=np.complex128)
    with pytest.raises(NativeError) as e:
        Batch([(16, [([0, 3, 6, 9, 12], [], m[::-1].copy())])])
    assert e.value.code == _native.QVD_EUNSUPPORTED
