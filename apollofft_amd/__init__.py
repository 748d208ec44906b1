"""apollofft_amd -- MI355X (gfx950) Kaiser-Bessel NUFFT.

A HIP re-design of apollofft's fast KB NUFFT path (apollo-nufft): 1D and 3D
Type-1 / Type-2 with binned LDS spreading, a Stockham FFT with fused
pad/crop/deconvolution, and gradients through ``torch.autograd``.  The device
library is ``libapollo_nufft_hip.so`` (C ABI: include/apollo_nufft.h).
"""
from . import _native
from ._native import (CapacityExceeded, DeviceUnavailable, HipError, InvalidArgument,
                      NativeLibraryMissing, NufftError, ProfileMismatch, ShapeMismatch,
                      U32Overflow, Unsupported)
from .api import (DEFAULT_NUFFT_KERNEL_WIDTH, DEFAULT_NUFFT_OVERSAMPLING, nufft_type1_1d,
                  nufft_type1_1d_fast, nufft_type1_3d, nufft_type1_3d_fast, nufft_type2_1d,
                  nufft_type2_1d_fast, nufft_type2_3d, nufft_type2_3d_fast)
from .plan import NufftPlan1D, NufftPlan3D, autotune3d, execute_type1_batched

__version__ = "0.1.0"

__all__ = [
    "NufftPlan1D", "NufftPlan3D", "autotune3d", "execute_type1_batched",
    "nufft_type1_1d", "nufft_type2_1d", "nufft_type1_3d", "nufft_type2_3d",
    "nufft_type1_1d_fast", "nufft_type2_1d_fast", "nufft_type1_3d_fast", "nufft_type2_3d_fast",
    "DEFAULT_NUFFT_KERNEL_WIDTH", "DEFAULT_NUFFT_OVERSAMPLING",
    "NufftError", "InvalidArgument", "ShapeMismatch", "ProfileMismatch", "CapacityExceeded",
    "U32Overflow", "DeviceUnavailable", "HipError", "Unsupported", "NativeLibraryMissing",
]
