"""plbfs: MI355X-native unstructured lattice Boltzmann flux solver (Python plumbing).

The solver itself is libplbfs.so (C++ host + HIP kernels); this package holds
the ctypes binding, the case description and the synthetic mesh generator.
"""
from . import case, meshgen  # noqa: F401
