"""ed_amd -- MI355X-native symmetry-adapted Lanczos engine for QS3 (QS-Cube/ED).

The product is libqs3x.so (HIP kernels for gfx950 behind the C ABI of
include/qs3x.h) with Fortran ISO_C_BINDING shims in ed_amd/fortran and the
ctypes mirror in ed_amd/qs3.py.  Import ed_amd.qs3 to use it from Python.
"""
