"""CPU FP64 oracle of holodeck's SAM GWB hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (as
the checker) and by bench.py's cpu_baseline leg.  The product package
(holodeck_amd) never imports it.
"""
from .pipeline import *  # noqa
