"""CPU oracle for the FDTD hot path -- test infrastructure only (the checker).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  See oracle.py and fdtd_oracle.cpp for what each part
restates and which reference file:line it follows.
"""
