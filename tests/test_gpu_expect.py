"""Expectation values on the device (qvd_sv_expect_pauli, the estimator's
statevector path, python/python/qevedo/simulator/qiskit.py:584-589) against
dense Pauli matrices built from Kronecker products on the host, and
probabilities computed on the device against |a|^2 of the amplitudes."""
import numpy as np
import pytest

from helpers import oracle_statevector, random_state, rng
from qevedo.simulator import Circuit, StateVector, expectation, library, lower_gates

pytestmark = pytest.mark.gpu

PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.array([[1, 0], [0, -1]])}


def dense(label):
    m = np.array([[1.0 + 0j]])
    for ch in label:  # first character: the highest qubit
        m = np.kron(m, PAULI[ch])
    return m


def random_labels(r, n, count):
    return ["".join("IXYZ"[int(k)] for k in r.integers(0, 4, n)) for _ in range(count)]


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-6)])
def test_pauli_strings_against_dense_matrices(precision, tol):
    r = rng(31)
    for n in (1, 3, 7, 10):
        psi = random_state(r, n)
        labels = random_labels(r, n, 40) + ["I" * n, "Z" * n, "Y" * n, "X" + "I" * (n - 1)]
        with StateVector(n, precision) as s:
            s.set_amplitudes(psi)
            stored = s.amplitudes()
            from qevedo.simulator._native import pauli_masks
            x, z = zip(*(pauli_masks(lb) for lb in labels))
            got = s.expect_pauli(list(x), list(z))
        want = [float(np.real(np.vdot(stored, dense(lb) @ stored))) for lb in labels]
        assert np.max(np.abs(got - np.array(want))) < tol, n


def test_observable_sums_and_z_agree_with_expect_z():
    r = rng(32)
    n = 16
    gates = library.random_circuit(n, 8, 4)
    psi = oracle_statevector(n, gates)
    with StateVector(n, "double") as s:
        s.apply_ops(lower_gates(gates))
        z, zz = s.expect_z()
        obs = {"I" * (n - 1 - q) + "Z" + "I" * q: 1.0 for q in range(n)}
        total = s.expectation_value(obs)
        assert abs(total - z.sum()) < 1e-12
        for q in range(n - 1):
            lb = "I" * (n - 2 - q) + "ZZ" + "I" * q
            assert abs(s.expectation_value({lb: 1.0}) - zz[q]) < 1e-12
        # mixed X / Y terms on a 16-qubit state against the oracle's FP64
        # amplitudes, checked term by term through the definition
        labels = random_labels(r, n, 6)
        coeffs = r.normal(size=6)
        got = s.expectation_value(list(zip(labels, coeffs)))
    idx = np.arange(1 << n)
    want = 0.0
    for lb, c in zip(labels, coeffs):
        from qevedo.simulator._native import pauli_masks
        x, zm = pauli_masks(lb)
        ny = bin(x & zm).count("1")
        sign = np.array([(-1) ** bin((j ^ x) & zm).count("1") for j in idx])
        v = (1j ** ny) * np.sum(sign * np.conj(psi) * psi[idx ^ x])
        want += c * np.real(v)
    assert abs(got - want) < 1e-11


def test_expectation_of_a_circuit():
    bell = Circuit(2).h(0).cx(0, 1)
    assert abs(expectation(bell, {"ZZ": 1.0}) - 1.0) < 1e-12
    assert abs(expectation(bell, {"XX": 1.0}) - 1.0) < 1e-12
    assert abs(expectation(bell, {"YY": 1.0}) + 1.0) < 1e-12
    assert abs(expectation(bell, {"ZI": 1.0, "IZ": 0.5})) < 1e-12
    with pytest.raises(ValueError, match="qubits"):
        expectation(bell, {"Z": 1.0})


@pytest.mark.parametrize("precision", ["single", "double"])
def test_probabilities_on_the_device(precision):
    r = rng(33)
    for n in (1, 9, 25):
        with StateVector(n, precision) as s:
            s.set_amplitudes(random_state(r, n))
            p = np.abs(s.amplitudes()) ** 2
            got = s.probabilities()
        assert np.max(np.abs(got - p)) < 1e-15
        assert abs(got.sum() - 1.0) < 1e-6


@pytest.mark.parametrize("precision,n", [("single", 23), ("double", 22)])
def test_large_state_reductions(precision, n):
    """States large enough for the reductions' batched loads (norm and
    probabilities over 8 strided loads per thread, the <Z> combine over 32
    segments at a time, Pauli strings over 4 iterations): norm_sqr,
    probability_one, <Z>, <ZZ> and Pauli strings against NumPy in FP64 on the
    stored amplitudes."""
    r = rng(33)
    psi = random_state(r, n)
    with StateVector(n, precision) as s:
        s.set_amplitudes(psi)
        a = s.amplitudes()
        p = np.abs(a) ** 2
        idx = np.arange(1 << n, dtype=np.uint64)
        print(f"large reductions n={n} {precision}", flush=True)
        assert abs(s.norm_sqr() - p.sum()) < 1e-12
        for q in (0, 5, 17, n - 1):
            bit = (idx >> np.uint64(q)) & np.uint64(1)
            assert abs(s.probability_one(q) - p[bit == 1].sum()) < 1e-12, q
        z, zz = s.expect_z()
        for q in range(n):
            bit = ((idx >> np.uint64(q)) & np.uint64(1)).astype(np.int64)
            assert abs(z[q] - np.sum(p * (1 - 2 * bit))) < 1e-12, q
            if q + 1 < n:
                par = bit ^ ((idx >> np.uint64(q + 1)) & np.uint64(1)).astype(np.int64)
                assert abs(zz[q] - np.sum(p * (1 - 2 * par))) < 1e-12, q
        xs = [0, 0b101, 1 << (n - 1), (1 << (n - 1)) | 3]
        zs = [0b11, 0b110, 1 << (n - 2), (1 << (n - 1)) | 1]
        got = s.expect_pauli(xs, zs)
    for t, (x, zm) in enumerate(zip(xs, zs)):
        k = idx ^ np.uint64(x)
        sign = 1 - 2 * (np.bitwise_count(k & np.uint64(zm)) & 1).astype(np.int64)
        v = np.sum(np.conj(a) * a[k.astype(np.int64)] * sign) * (1j ** int(bin(x & zm).count("1")))
        assert abs(got[t] - v.real) < 1e-12, t
