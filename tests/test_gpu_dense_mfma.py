"""Fused 5- and 6-qubit gates in complex64 on FP32 matrix cores
(csrc/dense_mfma.hip): against the oracle's FP64 naive applier
(reference.rs:8-28), against the VALU kernel, with controls, with targets that
do and do not include qubit 0, and on states with fewer columns than one
block."""
import numpy as np
import pytest

from helpers import controlled_full, random_state, random_unitary, rel_l2
from oracle import oracle
from qevedo.simulator import StateVector

pytestmark = pytest.mark.gpu


def run_op(n, psi, u, targets, controls=(), dense_valu=False):
    sv = StateVector(n, "single")
    sv.set_amplitudes(psi)
    if dense_valu:
        sv.apply_ops([(list(targets), list(controls), u)], tile_bits=None, dense_valu=True)
    else:
        sv.apply(u, targets, controls)
    out = sv.amplitudes()
    sv.close()
    return out


@pytest.mark.parametrize("k", [5, 6])
@pytest.mark.parametrize("n,case", [(12, 0), (14, 1), (16, 2), (7, 3), (9, 4)])
def test_mfma_dense_matches_fp64(k, n, case):
    if n < k + 1:
        pytest.skip("state too small")
    r = np.random.Generator(np.random.PCG64(100 * k + case))
    qubits = list(r.permutation(n)[:k])
    if case % 2 == 0 and 0 not in qubits:
        qubits[0] = 0  # lowest target is qubit 0: column-contiguous loads
    u = random_unitary(r, k)
    psi = random_state(r, n)
    print(f"k={k} n={n} targets={qubits}", flush=True)
    ref = psi.copy()
    oracle.reference_apply(ref, u, [int(q) for q in qubits])
    got = run_op(n, psi, u, [int(q) for q in qubits])
    assert rel_l2(got, ref) < 3e-6
    valu = run_op(n, psi, u, [int(q) for q in qubits], dense_valu=True)
    assert rel_l2(valu, ref) < 3e-6


@pytest.mark.parametrize("k", [5, 6])
def test_mfma_dense_with_controls(k):
    n = 13
    r = np.random.Generator(np.random.PCG64(7 + k))
    perm = [int(q) for q in r.permutation(n)]
    targets, ctrl = perm[:k], perm[k:k + 2]
    controls = [(ctrl[0], True), (ctrl[1], False)]
    u = random_unitary(r, k)
    psi = random_state(r, n)
    full = controlled_full(u, controls)  # on [controls..., targets...]
    ref = psi.copy()
    oracle.reference_apply(ref, full, [c for c, _ in controls] + targets)
    got = run_op(n, psi, u, targets, controls)
    assert rel_l2(got, ref) < 3e-6
