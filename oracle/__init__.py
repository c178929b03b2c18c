"""TEST INFRASTRUCTURE ONLY: the CPU checker (restatement of the reference's cpu::snn) and
the wrapper of the reference's own CPU build. Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may use this package."""
