"""Namespace package ``qevedo`` (the MI355X state-vector engine ships ``qevedo.simulator``)."""
