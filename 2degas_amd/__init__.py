"""2degas_amd: MI355X-native walker evaluation for the 2D electron gas.

The compute path is the HIP library lib2degas_hip.so (2degas_amd/csrc),
reached through its C-ABI (include/qmc_hip.h).  This Python package holds
only host-side tooling around it: the deck reader, the ctypes binding and
the batched drivers used by tests and the benchmark.
"""
from . import abi, deck  # noqa: F401
