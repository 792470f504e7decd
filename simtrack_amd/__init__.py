"""MI355X-native engine for SimTrack's tracking hot path.

The element-by-element FP64 tracking runs in hand-written HIP kernels for
gfx950 (csrc/), behind the C ABI of include/simtrack_amd.h.  This package
holds the build script, the ctypes binding of that ABI, the Element/Line
construction API mirroring the reference's C++ classes, and synthetic
workloads.
"""
from . import build  # noqa: F401
from ._capi import (ELEMENT_DTYPE, ST_DA_RADIAL, ST_DA_SINGLE, ST_DA_TRACK_DA, ST_DA_TRACK_DA1,  # noqa: F401
                    ST_DA_TRACK_DA2, ST_DA_TRACK_DA3, ST_DA_TRACK_DA_RADIAL, TRACK_RADIAL, TRACK_SPIN, TRACK_PERF,
                    NOALLOC_MAX_PARTICLES, NOALLOC_MAX_ELEMENTS, TYPES, Batch, BatchConfig, Context, Globals, Mlp,
                    STError, Table, load)
from .line import *  # noqa: F401,F403
