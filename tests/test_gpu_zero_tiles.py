"""Register tiles known to be zero are skipped: a qubit that no non-diagonal
op has targeted since |0...0> is still |0>, so a stage whose tile leaves it
out neither reads nor writes the tiles with its bit set. Every case here
compares the result with the same plan visiting every tile
(QVD_FLAG_ALL_TILES; np.array_equal, under which -0.0 equals 0.0) and with
the oracle's FP64 state."""
import numpy as np
import pytest

from helpers import oracle_statevector
from qevedo.simulator import Batch, StateVector, library, lower_gates
from qevedo.simulator import _native

pytestmark = pytest.mark.gpu


def _run(n, gates, precision="single", all_tiles=False, before=None):
    with StateVector(n, precision) as sv:
        if before is not None:
            sv.apply_ops(_native.OpArray(lower_gates(before)), all_tiles=all_tiles)
        st = sv.apply_ops(_native.OpArray(lower_gates(gates)), all_tiles=all_tiles)
        return sv.amplitudes(), st


def _late(n):
    """Qubits brought in one at a time (a GHZ chain), with controls and
    diagonals on qubits still in |0> along the way."""
    g = [("h", (0,), ())]
    for q in range(1, n):
        g.append(("cx", (q - 1, q), ()))
        g.append(("rz", (n - 1,), (0.3 * q,)))             # diagonal on a qubit still in |0>
        g.append(("cx", (n - 1, q // 2), ()))                 # controlled by a qubit still in |0>
        g.append(("ry", (q,), (0.1 * q,)))
    return g


@pytest.mark.parametrize("precision,tol", [("single", 1e-5), ("double", 1e-12)])
@pytest.mark.parametrize("case", ["random", "late"])
def test_skipping_zero_tiles_changes_nothing(precision, tol, case):
    n = 20
    gates = library.random_circuit(n, 10, 5) if case == "random" else _late(n)
    got, st = _run(n, gates, precision)
    full, st_full = _run(n, gates, precision, all_tiles=True)
    assert st["skipped_tiles"] > 0 and st_full["skipped_tiles"] == 0
    assert st["stage_bytes"] < st_full["stage_bytes"]
    assert np.array_equal(got, full)
    want = oracle_statevector(n, gates)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


def test_second_apply_continues_from_the_tracked_qubits():
    """A first apply that touches qubits 0-15 of 22 leaves 16-21 in |0>: the
    second apply skips tiles of those that it has not touched yet."""
    n = 22
    first = library.random_circuit(16, 6, 8)  # qubits 0-15 only
    second = _late(n)
    got, st = _run(n, second, before=first)
    full, _ = _run(n, second, all_tiles=True, before=first)
    assert st["skipped_tiles"] > 0
    assert np.array_equal(got, full)
    want = oracle_statevector(n, first + second)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


def test_batch_skips_zero_tiles():
    """Batches start every circuit from |0...0>: their first stages compute
    only the tiles that can be nonzero (the states are written with zeros
    first), with the same results as visiting every tile."""
    rng = np.random.Generator(np.random.PCG64(31))
    sizes = [int(x) for x in rng.integers(14, 23, size=24)]
    circuits = []
    for i, n in enumerate(sizes):
        g = library.random_circuit(n, 8, 40 + i) if i % 3 else _late(n)
        circuits.append((n, lower_gates(g), g))
    with Batch([(n, ops) for n, ops, _ in circuits], "single") as b, \
            Batch([(n, ops) for n, ops, _ in circuits], "single", all_tiles=True) as a:
        out, ref = b.run(), a.run()
        assert b.stats["skipped_tiles"] > 0 and a.stats["skipped_tiles"] == 0
        for i, (n, _, g) in enumerate(circuits):
            print(f"circuit {i}: {n} qubits", flush=True)
            assert np.array_equal(b.amplitudes(i), a.amplitudes(i))
            assert np.array_equal(out[i][0], ref[i][0]) and np.array_equal(out[i][1], ref[i][1])
            want = oracle_statevector(n, g)
            got = b.amplitudes(i).astype(np.complex128)
            assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
