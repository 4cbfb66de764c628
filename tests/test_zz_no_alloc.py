"""qvd_sv_sample_noalloc through the C ABI: the same outcomes as
qvd_sv_sample on two calls in a row (the second with fewer shots on a
smaller state, in the other precision), and a request beyond its fixed
workspace is an error after which the next call still works."""
import ctypes as C

import numpy as np
import pytest

from qevedo.simulator import StateVector, library, lower_gates
from qevedo.simulator import _native

pytestmark = pytest.mark.gpu


def sample_noalloc(sv, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(u.size, dtype=np.uint64)
    rc = _native.lib().qvd_sv_sample_noalloc(sv.h, _native._p(u, C.c_double), u.size, _native._p(out, C.c_uint64))
    return rc, out


def test_noalloc_sampling_matches_the_allocating_entry_point():
    r = np.random.Generator(np.random.PCG64(21))
    for n, precision, shots in [(22, "single", 200_000), (18, "double", 5_000)]:
        print(f"noalloc sampling n={n} {precision} shots={shots}", flush=True)
        with StateVector(n, precision) as sv:
            sv.apply_ops(lower_gates(library.random_circuit(n, 8, 3)))
            u = r.random(shots)
            u[:3] = [0.0, 0.5, 1 - 2 ** -53]
            rc, got = sample_noalloc(sv, u)
            assert rc == 0, _native.lib().qvd_last_error()
            assert np.array_equal(got, sv.sample(u))


def test_noalloc_rejects_requests_beyond_its_workspace():
    r = np.random.Generator(np.random.PCG64(22))
    with StateVector(16, "single") as sv:
        sv.apply_ops(lower_gates(library.random_circuit(16, 4, 1)))
        rc, _ = sample_noalloc(sv, r.random(1 << 23))  # 8M shots need about 300 MB
        assert rc == _native.QVD_EINVAL
        assert b"qvd_sv_sample" in _native.lib().qvd_last_error()
        u = r.random(1000)
        rc, got = sample_noalloc(sv, u)
        assert rc == 0 and np.array_equal(got, sv.sample(u))


def test_noalloc_with_coarser_marks_matches():
    """At 28 qubits in complex64 with 10^6 shots the walk's marks every 64
    amplitudes (32 MB) do not fit the fixed workspace beside the shots'
    buffers (32 MB and the sort's temporaries): the entry point marks every
    128 amplitudes instead, with the same outcomes."""
    r = np.random.Generator(np.random.PCG64(23))
    n = 28
    print(f"noalloc coarse marks n={n}", flush=True)
    with StateVector(n, "single") as sv:
        sv.apply_ops(lower_gates(library.random_circuit(n, 3, 4)))
        u = r.random(1_000_000)
        rc, got = sample_noalloc(sv, u)
        assert rc == 0, _native.lib().qvd_last_error()
        assert np.array_equal(got, sv.sample(u))
