"""cvtamd -- host-side Python API of the MI355X CVT engine.

The compute lives in lib/libcvt_hip.so (HIP kernels for gfx950 behind the C
ABI of include/cvt.h); this package loads it with ctypes and provides the
seeded workload generators used by the tests and the benchmark.
"""
from ._lib import LIB_PATH, LABEL_NONE, CvtError, lib
from .engine import Batch, Engine
from . import compat, recover, weighting, workloads

__all__ = ["Batch", "Engine", "CvtError", "LABEL_NONE", "LIB_PATH", "lib", "compat", "recover", "weighting", "workloads"]
