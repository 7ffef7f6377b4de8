"""hsmc_amd: MI355X-native hard-sphere Monte Carlo hot path.

The product is the C ABI library hsmc_amd/lib/libhsmc_gpu.so (HIP kernels for
gfx950, declared in include/hsmc_gpu.h) and the C host layer built on it
(include/hsmc.h).  This package adds the Python binding (``hsmc_amd.gpu``)
and the post-processing counterparts (``hsmc_amd.analysis``).
"""
__all__ = ["gpu"]
