"""CPU oracle (test infrastructure only): NumPy restatement of the reference's
V1-parity solver step and topology preparation.  Imported only by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py."""
