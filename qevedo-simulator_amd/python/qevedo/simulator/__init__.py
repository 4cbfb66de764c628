"""qevedo.simulator — the MI355X state-vector backend.

The state-vector surface of the reference's Python package
(python/python/qevedo/simulator/__init__.py): circuits as gate lists, the
final state (``statevector``) and shot sampling (``run``), computed by the
HIP engine in ``libqvd_hip.so`` through its C ABI.
"""
from . import library
from ._native import NativeError, StateVector, device_info, gate_lower, gate_matrix, lower_gates

__all__ = ["NativeError", "StateVector", "device_info", "gate_lower", "gate_matrix", "library", "lower_gates"]
