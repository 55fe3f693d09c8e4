"""fr_learn_amd -- MI355X-native flux-reconstruction residual engine.

The compute path is libfrhip.so (HIP kernels for gfx950 behind the C ABI in
include/fr.h).  This package holds the build recipe, the ctypes binding, the
Python mirror of the reference solver verbs, mesh generation and the
benchmark workloads' initial conditions.
"""
from . import cases, mesh  # noqa: F401
from ._lib import FrError  # noqa: F401


def Solver(*args, **kw):  # noqa: N802 - mirrors the reference class name
    from .solver import Solver as _S
    return _S(*args, **kw)


def Batch(*args, **kw):  # noqa: N802
    from .solver import Batch as _B
    return _B(*args, **kw)
