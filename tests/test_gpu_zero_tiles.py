"""Register tiles known to be zero are skipped: a qubit that no non-diagonal
op has targeted since |0...0> is still |0>, so a stage whose tile leaves it
out neither reads nor writes the tiles with its bit set. Those zeros are not
written either: a plan of register-tile stages makes |0...0> in its first
stage's registers and leaves the skipped qubits' amplitudes stale in memory;
later stages set them to zero instead of loading them, and any other use of
the state writes them first. Every case here compares the result with the
same plan visiting every tile (QVD_FLAG_ALL_TILES; np.array_equal, under
which -0.0 equals 0.0) and with the oracle's FP64 state; the cases of stale
memory fill it with nonzero values first."""
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
    only the tiles that can be nonzero (in registers; the states are not
    written first), with the same results as visiting every tile. The second
    run finds the first run's amplitudes in memory."""
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
        again = b.run()
        for i, (n, _, g) in enumerate(circuits):
            assert np.array_equal(again[i][0], out[i][0]) and np.array_equal(again[i][1], out[i][1])
            print(f"circuit {i}: {n} qubits", flush=True)
            assert np.array_equal(b.amplitudes(i), a.amplitudes(i))
            assert np.array_equal(out[i][0], ref[i][0]) and np.array_equal(out[i][1], ref[i][1])
            want = oracle_statevector(n, g)
            got = b.amplitudes(i).astype(np.complex128)
            assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


def _dirty_zero(sv, n, seed):
    """Nonzero values in every amplitude's memory, then |0...0> (pending, not
    written): memory that no stage may read as amplitudes."""
    r = np.random.default_rng(seed)
    a = r.standard_normal(1 << n) + 1j * r.standard_normal(1 << n)
    sv.set_amplitudes(a / np.linalg.norm(a))
    sv.set_zero()


def _partial(n):
    """Qubits 0..n-5 only: every stage skips the tiles of the top four, whose
    amplitudes stay stale after the apply."""
    return library.random_circuit(n - 4, 6, 11)


@pytest.mark.parametrize("precision,tol", [("single", 1e-5), ("double", 1e-12)])
def test_stale_amplitudes_read_as_zero(precision, tol):
    """An apply that leaves qubits stale writes no zeros; reading the state
    writes them."""
    n = 20
    gates = _partial(n)
    with StateVector(n, precision) as sv, StateVector(n, precision) as full:
        _dirty_zero(sv, n, 1)
        _dirty_zero(full, n, 2)
        st = sv.apply_ops(_native.OpArray(lower_gates(gates)))
        full.apply_ops(_native.OpArray(lower_gates(gates)), all_tiles=True)
        assert st["other_bytes"] == 0 and st["skipped_tiles"] > 0
        got = sv.amplitudes()
        assert np.array_equal(got, full.amplitudes())
        assert not np.any(got.reshape(16, -1)[1:])  # every amplitude with a top-four bit set
    want = oracle_statevector(n, gates)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


@pytest.mark.parametrize("precision,tol", [("single", 1e-5), ("double", 1e-12)])
def test_stages_over_stale_amplitudes(precision, tol):
    """A second apply whose stages bring the stale qubits in: their
    amplitudes are set to zero, not loaded, and the apply writes no zeros."""
    n = 20
    first, second = _partial(n), _late(n)
    with StateVector(n, precision) as sv, StateVector(n, precision) as full:
        _dirty_zero(sv, n, 3)
        _dirty_zero(full, n, 4)
        a = sv.apply_ops(_native.OpArray(lower_gates(first)))
        b = sv.apply_ops(_native.OpArray(lower_gates(second)))
        full.apply_ops(_native.OpArray(lower_gates(first)), all_tiles=True)
        c = full.apply_ops(_native.OpArray(lower_gates(second)), all_tiles=True)
        assert a["other_bytes"] == 0 and b["other_bytes"] == 0
        assert b["stage_bytes"] < c["stage_bytes"]
        got = sv.amplitudes()
        assert np.array_equal(got, full.amplitudes())
    want = oracle_statevector(n, first + second)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol


def test_every_tile_over_stale_amplitudes():
    """Visiting every tile after an apply that left qubits stale: the stale
    qubits' tiles are visited too, and read as zero."""
    n = 20
    first, second = _partial(n), _late(n)
    with StateVector(n) as sv, StateVector(n) as ref:
        _dirty_zero(sv, n, 5)
        sv.apply_ops(_native.OpArray(lower_gates(first)))
        st = sv.apply_ops(_native.OpArray(lower_gates(second)), all_tiles=True)
        assert st["other_bytes"] == 0 and st["skipped_tiles"] == 0
        ref.apply_ops(_native.OpArray(lower_gates(first)), all_tiles=True)
        ref.apply_ops(_native.OpArray(lower_gates(second)), all_tiles=True)
        assert np.array_equal(sv.amplitudes(), ref.amplitudes())


def test_other_plans_and_entry_points_write_stale_zeros():
    """After an apply that left qubits stale: a plan with a full-state pass
    (no blocking) writes the zeros first, and so does a single gate (an
    entry point of its own); a copy carries the stale qubits to its target."""
    n = 20
    first = _partial(n)
    h = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
    with StateVector(n) as sv, StateVector(n) as cp, StateVector(n) as ref:
        for k, case in enumerate(["unblocked", "gate", "copy"]):
            _dirty_zero(sv, n, 10 + k)
            _dirty_zero(cp, n, 20 + k)
            ref.set_zero()
            sv.apply_ops(_native.OpArray(lower_gates(first)))
            ref.apply_ops(_native.OpArray(lower_gates(first)), all_tiles=True)
            tgt = sv
            if case == "unblocked":
                st = sv.apply_ops(_native.OpArray(lower_gates(_late(n))), tile_bits=None)
                ref.apply_ops(_native.OpArray(lower_gates(_late(n))), tile_bits=None)
                assert st["other_bytes"] > 0
            elif case == "gate":
                sv.apply(h, [n - 1])
                ref.apply(h, [n - 1])
            else:
                cp.copy_from(sv)
                cp.apply_ops(_native.OpArray(lower_gates(_late(n))))
                ref.apply_ops(_native.OpArray(lower_gates(_late(n))), all_tiles=True)
                tgt = cp
            print(f"case {case}", flush=True)
            assert np.array_equal(tgt.amplitudes(), ref.amplitudes())


def test_sampling_a_state_with_stale_qubits():
    """Both sampling entry points on a state whose apply left qubits stale:
    the outcomes of the same amplitudes made visiting every tile, and the
    oracle's on those amplitudes (measure.rs:104-174)."""
    import ctypes as C
    from oracle import oracle

    n = 20
    gates = _partial(n)
    u = np.random.default_rng(7).random(100_000)
    u[:3] = [0.0, 0.5, 1 - 2 ** -53]
    with StateVector(n) as ref:
        ref.apply_ops(_native.OpArray(lower_gates(gates)), all_tiles=True)
        stored = ref.amplitudes()
        want = ref.sample(u)
    with oracle.OracleState(n, "single") as o:
        o.set_amplitudes(stored)
        assert np.array_equal(o.sample(u), want)
    for entry in ["sample", "noalloc"]:
        print(f"sampling through {entry}", flush=True)
        with StateVector(n) as sv:
            _dirty_zero(sv, n, 30)
            st = sv.apply_ops(_native.OpArray(lower_gates(gates)))
            assert st["other_bytes"] == 0
            if entry == "sample":
                got = sv.sample(u)
            else:
                got = np.empty(u.size, dtype=np.uint64)
                rc = _native.lib().qvd_sv_sample_noalloc(sv.h, _native._p(u, C.c_double), u.size,
                                                         _native._p(got, C.c_uint64))
                assert rc == 0, _native.lib().qvd_last_error()
            assert np.array_equal(got, want)
            assert np.array_equal(sv.amplitudes(), stored)


@pytest.mark.parametrize("precision,tol", [("single", 1e-5), ("double", 1e-12)])
@pytest.mark.parametrize("stale_end", [False, True])
def test_batch_lazy_and_written_states(precision, tol, stale_end):
    """A batch whose circuits all end visiting every tile writes no state
    before its stages; one with a circuit that ends with stale qubits (it
    touches only its lowest qubits) writes every state first. Either way the
    results equal every tile visited, on two runs in a row, and the
    oracle's."""
    circuits = []
    for i, n in enumerate([15, 17, 20, 22]):
        g = library.random_circuit(n, 6, 60 + i) if i % 2 else _late(n)
        circuits.append((n, g))
    if stale_end:
        circuits.append((20, _partial(20)))
    with Batch([(n, lower_gates(g)) for n, g in circuits], precision) as b, \
            Batch([(n, lower_gates(g)) for n, g in circuits], precision, all_tiles=True) as a:
        out, ref = b.run(), a.run()
        assert b.stats["skipped_tiles"] > 0
        state_bytes = sum((8 if precision == "single" else 16) << max(n, 14 if precision == "single" else 13)
                          for n, _ in circuits)
        assert b.stats["other_bytes"] == (2 if stale_end else 1) * state_bytes
        again = b.run()
        for i, (n, g) in enumerate(circuits):
            print(f"circuit {i}: {n} qubits", flush=True)
            assert np.array_equal(b.amplitudes(i), a.amplitudes(i))
            for r in (out, again):
                assert np.array_equal(r[i][0], ref[i][0]) and np.array_equal(r[i][1], ref[i][1])
            want = oracle_statevector(n, g)
            got = b.amplitudes(i).astype(np.complex128)
            assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
