"""fdfault_amd -- MI355X-native time-stepping core of the fdfault dynamic
rupture solver, behind a C ABI (include/fdfault_hip.h).

The drop-in for the reference's ``fdfault <file.in>`` executable is
``fdfault_amd/_lib/fdfault``; this package is its Python mirror for tests and
benchmarks: :class:`Solver` wraps one problem handle of the HIP library.
There is no CPU fallback: if the HIP library is missing, importing
:mod:`fdfault_amd.lib` raises.
"""
from .problems import Problem, Pert, Output, two_block_fault  # noqa: F401

__all__ = ["Problem", "Pert", "Output", "two_block_fault"]
