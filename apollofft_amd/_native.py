"""ctypes binding of libapollo_nufft_hip.so (C ABI in include/apollo_nufft.h).

The shared library is built in-tree by ``apollofft_amd.build``.  Loading never
falls back to another implementation: a missing library raises
``NativeLibraryMissing``.  When torch is importable it is imported first so
that torch and this library share one HIP runtime (both resolve the soname
libamdhip64.so.7 to the copy torch already loaded).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libapollo_nufft_hip.so")

# anufft_status
OK = 0
INVALID_ARG = 1
SHAPE_MISMATCH = 2
PROFILE_MISMATCH = 3
CAPACITY_EXCEEDED = 4
U32_OVERFLOW = 5
DEVICE_UNAVAILABLE = 6
HIP_ERROR = 7
UNSUPPORTED = 8

GRID_REFERENCE = 0
GRID_EXACT = 1

STORAGE_C128 = 0
STORAGE_C64 = 1
STORAGE_F16X2 = 2
PROFILE_HIGH_ACCURACY_F64 = 0
PROFILE_LOW_PRECISION_F32 = 1
PROFILE_MIXED_PRECISION_F16_F32 = 2

EXPORTS = [
    "anufft_default_opts", "anufft_plan1d_create", "anufft_plan3d_create",
    "anufft_plan_destroy", "anufft_plan_get_meta", "anufft_plan_get_deconv",
    "anufft_workspace_size", "anufft_execute_type1", "anufft_execute_type2",
    "anufft_execute_grad_points", "anufft_execute_type2_backward",
    "anufft_execute_type1_backward",
    "anufft_batch_workspace_size",
    "anufft_execute_type1_batched", "anufft_execute_type1_host",
    "anufft_execute_type2_host", "anufft_direct_type1", "anufft_direct_type2",
    "anufft_direct_type1_host", "anufft_direct_type2_host", "anufft_direct_type1_select",
    "anufft_last_error", "anufft_version",
    "anufft_debug_num_bins", "anufft_debug_sort", "anufft_debug_spread",
    "anufft_timing_enable", "anufft_timing_collect",
    "anufft_execute_type1_noalloc", "anufft_scratch_capacity", "anufft_autotune3d",
]

SPREAD_AUTO, SPREAD_VALU, SPREAD_MFMA = 0, 1, 2
STAGES = ["fill", "bin", "spread", "fft", "interp", "grad", "prep", "tile_max"]


class NufftError(RuntimeError):
    """Base class of errors reported by the device library."""

    status = -1


class InvalidArgument(NufftError, ValueError):
    status = INVALID_ARG


class ShapeMismatch(NufftError, ValueError):
    status = SHAPE_MISMATCH


class ProfileMismatch(NufftError, ValueError):
    status = PROFILE_MISMATCH


class CapacityExceeded(NufftError):
    status = CAPACITY_EXCEEDED


class U32Overflow(NufftError, ValueError):
    status = U32_OVERFLOW


class DeviceUnavailable(NufftError):
    status = DEVICE_UNAVAILABLE


class HipError(NufftError):
    status = HIP_ERROR


class Unsupported(NufftError):
    status = UNSUPPORTED


class NativeLibraryMissing(ImportError):
    pass


_ERRORS = {c.status: c for c in (InvalidArgument, ShapeMismatch, ProfileMismatch,
                                  CapacityExceeded, U32Overflow, DeviceUnavailable,
                                  HipError, Unsupported)}


class Opts(ctypes.Structure):
    _fields_ = [
        ("grid_policy", ctypes.c_int32),
        ("compat_reference_type2_3d_scaling", ctypes.c_int32),
        ("device_id", ctypes.c_int32),
        ("bin_size", ctypes.c_int32 * 3),
        ("max_subproblem", ctypes.c_int32),
        ("kernel_method", ctypes.c_int32),
        ("sub_work", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("fft_lines", ctypes.c_int32),
    ]


class Meta(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n", ctypes.c_int64 * 3),
        ("m", ctypes.c_int64 * 3),
        ("length", ctypes.c_double * 3),
        ("kernel_width", ctypes.c_int32),
        ("sigma", ctypes.c_int32),
        ("beta", ctypes.c_double),
        ("i0_beta", ctypes.c_double),
        ("grid_policy", ctypes.c_int32),
        ("compat_reference_type2_3d_scaling", ctypes.c_int32),
        ("bin_size", ctypes.c_int32 * 3),
        ("max_subproblem", ctypes.c_int32),
        ("interp_bin_size", ctypes.c_int32 * 3),
        ("interp_max_subproblem", ctypes.c_int32),
        ("poly_degree", ctypes.c_int32),
        ("poly_max_err", ctypes.c_double),
        ("kernel_method", ctypes.c_int32),
        ("sub_work", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("fft_lines", ctypes.c_int32),
    ]


_lib = None


def lib():
    """Load the library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing; build it with `python -m apollofft_amd.build`")
    try:  # share torch's HIP runtime when torch is present
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
    P = ctypes.POINTER
    sig = {
        "anufft_default_opts": (None, [P(Opts)]),
        "anufft_plan1d_create": (i32, [i64, ctypes.c_double, i32, i32, P(Opts), P(vp)]),
        "anufft_plan3d_create": (i32, [i64, i64, i64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, i32, i32, P(Opts), P(vp)]),
        "anufft_plan_destroy": (i32, [vp]),
        "anufft_plan_get_meta": (i32, [vp, P(Meta)]),
        "anufft_plan_get_deconv": (i32, [vp, i32, vp]),
        "anufft_workspace_size": (i32, [vp, i64, P(sz)]),
        "anufft_execute_type1": (i32, [vp, i64, P(vp), i64, vp, vp, vp, sz, vp]),
        "anufft_execute_type2": (i32, [vp, i64, P(vp), i64, vp, vp, vp, sz, vp]),
        "anufft_execute_type1_noalloc": (i32, [vp, i64, P(vp), i64, vp, vp, vp]),
        "anufft_scratch_capacity": (i32, [P(sz)]),
        "anufft_autotune3d": (i32, [i64, i64, i64, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, i32, i32, P(Opts), i64, P(vp), i64, i32,
                                    ctypes.c_char_p, P(Opts), P(ctypes.c_double)]),
        "anufft_execute_grad_points": (i32, [vp, i64, P(vp), i64, vp, vp, P(vp), vp, sz, vp]),
        "anufft_execute_type2_backward": (i32, [vp, i64, P(vp), i64, vp, vp, vp, P(vp), vp, sz,
                                                vp]),
        "anufft_execute_type1_backward": (i32, [vp, i64, P(vp), i64, vp, vp, vp, P(vp), vp, sz,
                                                vp]),
        "anufft_batch_workspace_size": (i32, [P(vp), i32, P(i64), P(sz)]),
        "anufft_execute_type1_batched": (i32, [P(vp), i32, P(i64), P(vp), i64, P(vp), P(vp),
                                               vp, sz, vp]),
        "anufft_execute_type1_host": (i32, [vp, i32, i32, i64, P(vp), i64, vp, vp]),
        "anufft_execute_type2_host": (i32, [vp, i32, i32, i64, P(vp), i64, vp, vp]),
        "anufft_direct_type1": (i32, [vp, i64, P(vp), i64, vp, vp, vp]),
        "anufft_direct_type2": (i32, [vp, i64, P(vp), i64, vp, vp, vp]),
        "anufft_direct_type1_host": (i32, [vp, i64, P(vp), i64, vp, vp]),
        "anufft_direct_type2_host": (i32, [vp, i64, P(vp), i64, vp, vp]),
        "anufft_direct_type1_select": (i32, [vp, i64, P(vp), i64, vp, i64, vp, vp, vp]),
        "anufft_debug_num_bins": (i32, [vp, P(i64)]),
        "anufft_debug_sort": (i32, [vp, i64, P(vp), i64, vp, vp, vp, sz, vp]),
        "anufft_debug_spread": (i32, [vp, i64, P(vp), i64, vp, vp, vp, sz, vp]),
        "anufft_timing_enable": (i32, [vp, i32]),
        "anufft_timing_collect": (i32, [vp, P(ctypes.c_double), P(i64)]),
        "anufft_last_error": (ctypes.c_char_p, []),
        "anufft_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status, what=""):
    """Raise the exception class of a non-zero anufft_status."""
    if status == OK:
        return
    msg = lib().anufft_last_error().decode(errors="replace")
    cls = _ERRORS.get(status, NufftError)
    raise cls(f"{what}: {msg}" if what else msg)


def ptr_array(ptrs, n=None):
    n = len(ptrs) if n is None else n
    arr = (ctypes.c_void_p * n)()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def default_opts():
    o = Opts()
    lib().anufft_default_opts(ctypes.byref(o))
    return o
