"""ORACLE — test infrastructure only: the CPU restatement of the reference's
state-vector path and its ctypes binding (``oracle.oracle``)."""
