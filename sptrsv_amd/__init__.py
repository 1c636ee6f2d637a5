"""MI355X (gfx950) sparse upper-triangular solver: U x = b and U X = B.

The solver is the HIP library libsptrsv_amd.so (kernels + C ABI + the reference's C++ entry
points).  `sptrsv_amd.capi` binds its C ABI (import it explicitly; it raises when the library
was not built); `sptrsv_amd.gen` builds the synthetic factors used by the benchmark and tests.
"""
