"""Test infrastructure: the CPU oracle (FP64 restatement of the reference's
force routines) and the driver of the reference build in oracle/_ref.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product (mspms2_amd) never does.
"""
