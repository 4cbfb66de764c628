"""CPU checks of the product boundary: the C-ABI library loads, exports every
symbol include/qvd_hip.h declares, and its host-side paths (gate table,
argument validation) work without a GPU."""
import os
import re

import numpy as np
import pytest

from oracle import oracle
from qevedo.simulator import NativeError, StateVector, _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    header = open(os.path.join(ROOT, "include", "qvd_hip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(qvd_\w+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_native.EXPORTS)
    lib = _native.lib()
    for name in declared:
        assert hasattr(lib, name), name


GATES = [("id", ()), ("h", ()), ("x", ()), ("y", ()), ("z", ()), ("s", ()), ("sdg", ()), ("t", ()),
         ("tdg", ()), ("sx", ()), ("sxdg", ()), ("rx", (0.3,)), ("ry", (0.4,)), ("rz", (0.5,)),
         ("p", (0.6,)), ("u1", (0.6,)), ("u3", (0.1, 0.2, 0.3)), ("cx", ()), ("cy", ()), ("cz", ()),
         ("ch", ()), ("cp", (0.7,)), ("cu1", (0.7,)), ("crx", (0.8,)), ("cry", (0.9,)), ("crz", (1.0,)),
         ("swap", ()), ("rzz", (1.1,)), ("rxx", (1.2,)), ("ccx", ()), ("cswap", ()), ("U", (0.1, 0.2, 0.3)),
         ("CX", ()), ("u0", (0.5,)), ("u2", (0.1, 0.2)), ("csx", ()), ("cu3", (0.1, 0.2, 0.3)),
         ("cu", (0.1, 0.2, 0.3, 0.4)), ("c3x", ()), ("c4x", ()), ("c3sqrtx", ()), ("cnot", ()), ("toffoli", ()),
         ("fredkin", ()), ("cphase", (0.3,)), ("phase", (0.3,))]


@pytest.mark.parametrize("name,params", GATES)
def test_product_gate_table_matches_oracle(name, params):
    got = _native.gate_matrix(name, params)
    want = oracle.gate_matrix(name, params)
    assert got.shape == want.shape and np.array_equal(got, want)
    c, base = _native.gate_lower(name, params)
    assert c == oracle.gate_controls(name)


def test_unknown_gate_and_parameter_count():
    with pytest.raises(NativeError, match="unknown gate"):
        _native.gate_matrix("foo")
    with pytest.raises(NativeError, match="takes 1 parameter"):
        _native.gate_matrix("rx", ())


def test_create_without_gpu_or_over_cap(monkeypatch):
    # Argument checks come before any device call.
    with pytest.raises(NativeError) as e:
        StateVector(40)
    assert e.value.code == _native.QVD_EINVAL
    with pytest.raises(ValueError):
        StateVector(3, "half")
