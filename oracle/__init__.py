"""ORACLE — test infrastructure only.

CPU restatement of VirtualFluids' K17 LBM time step (FP64 and FP32), used as the
checker for the HIP path. Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg may import this package; the product never does.
"""
from .oracle import *  # noqa: F401,F403
