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


@pytest.mark.parametrize("n", [20, 22])
def test_precision_form_circuit_against_fp64(n):
    """The precision-form workload at 20-22 qubits: the depth-20 random
    circuit fused to dense ops of at most 6 qubits, each its own pass (5- and
    6-qubit ones on the FP32 matrix cores in complex64), against the oracle's
    FP64 engine: relative L2 within 1e-5 in complex64 and 1e-12 in
    complex128."""
    from helpers import lower_with_oracle
    from qevedo.simulator import library, lower_gates

    gates = library.random_circuit(n, 20, 42)
    print(f"precision form n={n}", flush=True)
    with oracle.OracleState(n, "double") as o:
        o.apply_ops(lower_with_oracle(gates), 4, 14)
        want = o.amplitudes()
    ops = lower_gates(gates)
    with StateVector(n, "single") as s:
        st = s.apply_ops(ops, 6, None)
        got = s.amplitudes()
    # fusion to at most 6 qubits: far fewer kernels than gates, each a pass
    assert st["kernels"] * 8 < st["gates"] and st["passes"] == st["kernels"] and st["stages"] == 0
    assert rel_l2(got, want) <= 1e-5
    with StateVector(n, "double") as s:
        s.apply_ops(ops, 6, None)
        got = s.amplitudes()
    assert rel_l2(got, want) <= 1e-12


def test_precision_form_at_32_qubits_crosses_2_to_31_amplitudes():
    """32 qubits (2^32 amplitudes: every 64-bit index path of the fused-gate
    kernels): the complex64 run on matrix cores against the same plan in
    complex128 on the device, relative L2 within 1e-5, and the norm kept."""
    from qevedo.simulator import library, lower_gates
    from qevedo.simulator import _native

    n = 32
    ops = _native.OpArray(lower_gates(library.random_circuit(n, 20, 42)))
    print("precision form 32 qubits: complex64", flush=True)
    c64 = StateVector(n, "single")
    st = c64.apply_ops(ops, 6, None)
    assert st["kernels"] < 60
    print("precision form 32 qubits: complex128", flush=True)
    c128 = StateVector(n, "double")
    c128.apply_ops(ops, 6, None)
    err = c64.compare(c128)
    assert err["rel_l2"] <= 1e-5, err
    assert abs(c64.norm_sqr() - 1.0) < 1e-6
    assert abs(c128.norm_sqr() - 1.0) < 1e-12
    c64.close()
    c128.close()
