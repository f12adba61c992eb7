"""The oracle (test infrastructure only): a CPU restatement of the reference's
algorithm, loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
