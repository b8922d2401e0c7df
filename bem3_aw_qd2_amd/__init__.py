"""MI355X-native implementation of the bem3_aw_qd2 boundary-integral hot path.

The product is two shared libraries built from csrc/ (see csrc/Makefile):
libbqd2_hip.so (HIP kernels for gfx950, C ABI in include/bqd2.h) and
libbem3_aw_qd2.so (the reference's bem3_aw_qd2.h C ABI on the host, calling the
HIP library).  This package only locates, builds and binds them.
"""
from ._native import (LIB_DIR, build, hip_lib, host_lib, QPD2, GfParams,
                      gf_params, BQD2Error, AssemblyStats, LargeStats, check)

__all__ = ["LIB_DIR", "build", "hip_lib", "host_lib", "QPD2", "GfParams",
           "gf_params", "BQD2Error", "AssemblyStats", "LargeStats", "check"]
