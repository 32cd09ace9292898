"""MI355X-native LehrFEM++ assembly + PCG path.

The product is the HIP library liblfgpu.so behind the C ABI in
include/lfgpu.h; ``capi`` is its Python binding and ``include/lf_amd``
holds the C++ facade that mirrors LehrFEM++'s provider / assembler API.
"""
__all__ = ["capi", "build"]
