"""ctypes binding of ``libnect_amd.so`` (the C ABI declared in include/nect_amd.h).

The library is loaded after ``torch`` so that its HIP runtime dependency binds
to the runtime torch already holds (both carry the soname libamdhip64.so.7):
device pointers and streams from torch are then valid inside the library.
There is no fallback: if the library is missing or a tensor is not on the GPU,
the call raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libnect_amd.so"

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f32 = ctypes.c_float
c_vp = ctypes.c_void_p

MAX_LEVELS = 32


class NectGeometry(ctypes.Structure):
    _fields_ = [
        ("kind", c_i32),
        ("n_det", c_i32 * 2),
        ("d_det", c_f32 * 2),
        ("off_det", c_f32 * 2),
        ("rot_det", c_f32 * 3),
        ("cor", c_f32),
        ("dso", c_f32),
        ("dod", c_f32),
        ("s_voxel", c_f32 * 3),
        ("d_voxel", c_f32 * 3),
        ("radius", c_f32),
        ("height", c_f32),
        ("remove_top", c_f32),
        ("remove_bottom", c_f32),
    ]


class NectRay(ctypes.Structure):
    _fields_ = [
        ("start", ctypes.c_double * 3),
        ("step", ctypes.c_double * 3),
        ("count", c_i32),
        ("distance", c_f32),
        ("pad", c_i32 * 2),
    ]


RAY_BYTES = ctypes.sizeof(NectRay)
assert RAY_BYTES == 64


class NectAmpState(ctypes.Structure):
    _fields_ = [
        ("scale", c_f32),
        ("growth_tracker", c_i32),
        ("step", c_i32),
        ("found_inf", c_i32),
        ("skipped", c_i32),
        ("step_size", c_f32),
        ("inv_sqrt_bc2", c_f32),
        ("pad", c_i32),
    ]


AMP_STATE_BYTES = ctypes.sizeof(NectAmpState)
assert AMP_STATE_BYTES == 32


class NectGridDesc(ctypes.Structure):
    _fields_ = [
        ("n_dims", c_i32),
        ("n_levels", c_i32),
        ("n_features", c_i32),
        ("log2_hashmap_size", c_i32),
        ("base_resolution", c_i32),
        ("per_level_scale", c_f32),
        ("scale", c_f32 * MAX_LEVELS),
        ("resolution", ctypes.c_uint32 * MAX_LEVELS),
        ("level_size", ctypes.c_uint32 * MAX_LEVELS),
        ("offset", ctypes.c_uint32 * (MAX_LEVELS + 1)),
        ("n_params", c_i64),
    ]


class NectMlpDesc(ctypes.Structure):
    _fields_ = [
        ("n_input", c_i32),
        ("n_input_padded", c_i32),
        ("n_output", c_i32),
        ("n_output_padded", c_i32),
        ("width", c_i32),
        ("n_hidden_layers", c_i32),
        ("activation", c_i32),
        ("output_activation", c_i32),
        ("n_params", c_i64),
    ]


F32, F16, BF16 = 0, 1, 2
_DTYPE_CODE = {torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16}


def dtype_code(dt: torch.dtype) -> int:
    try:
        return _DTYPE_CODE[dt]
    except KeyError:
        raise ValueError(f"unsupported dtype {dt}; use float32, float16 or bfloat16") from None

# name -> (argtypes); every entry point returns int32 status.
_SIGNATURES: dict[str, list] = {
    "nect_sample": [ctypes.POINTER(NectGeometry), c_f32, c_vp, c_i64, c_i32, c_i32, c_i32, c_f32, c_i32,
                    c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "nect_ray_setup": [ctypes.POINTER(NectGeometry), c_f32, c_vp, c_i64, c_i32, c_i32, c_i32, c_f32, c_i32,
                       c_vp, c_vp, c_vp, c_vp],
    "nect_ray_tiles": [c_vp, c_i32, c_i64, c_vp, c_vp, c_vp, c_vp],
    "nect_ray_positions": [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
    "nect_ray_order": [ctypes.POINTER(NectGeometry), c_vp, c_i64, c_i32, c_i32, c_vp, c_vp, c_i64, c_vp],
    "nect_grid_init": [ctypes.POINTER(NectGridDesc), c_i32, c_i32, c_i32, c_i32, c_i32, c_f32],
    "nect_grid_index": [ctypes.POINTER(NectGridDesc), c_vp, c_i64, c_i64, c_vp, c_vp, c_vp],
    "nect_grid_encode": [ctypes.POINTER(NectGridDesc), c_i32, c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp],
    "nect_grid_backward": [ctypes.POINTER(NectGridDesc), c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp],
    "nect_grid_encode_parts": [c_i32, ctypes.POINTER(NectGridDesc), ctypes.POINTER(c_i64), ctypes.POINTER(c_i32),
                               ctypes.POINTER(c_i32), c_i32, c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp],
    "nect_grid_backward_parts": [c_i32, ctypes.POINTER(NectGridDesc), ctypes.POINTER(c_i64), ctypes.POINTER(c_i32),
                                 ctypes.POINTER(c_i32), c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp],
    "nect_mlp_init": [ctypes.POINTER(NectMlpDesc), c_i32, c_i32, c_i32, c_i32, c_i32, c_i32],
    "nect_mlp_forward": [ctypes.POINTER(NectMlpDesc), c_i32, c_vp, c_vp, c_i64, c_i64, ctypes.POINTER(NectGridDesc),
                         c_vp, c_vp, c_i64, c_vp],
    "nect_mlp_backward": [ctypes.POINTER(NectMlpDesc), c_i32, c_vp, c_vp, c_i64, c_i64, ctypes.POINTER(NectGridDesc),
                          c_vp, c_vp, c_i64, c_f32, c_vp, c_f32, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp],
    "nect_mlp_backward_noalloc": [ctypes.POINTER(NectMlpDesc), c_i32, c_vp, c_vp, c_i64, c_i64,
                                  ctypes.POINTER(NectGridDesc), c_vp, c_vp, c_i64, c_f32, c_vp, c_f32, c_vp, c_vp, c_i64,
                                  c_vp],
    "nect_grid_encode_rays": [ctypes.POINTER(NectGridDesc), c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_vp,
                              c_i64, c_i32, c_vp],
    "nect_grid_backward_rays": [ctypes.POINTER(NectGridDesc), c_vp, c_vp, c_vp, c_vp, c_i64, c_f32, c_vp, c_i64, c_vp,
                                c_vp],
    "nect_grid_backward_ray_groups": [ctypes.POINTER(NectGridDesc), c_vp, c_vp, c_i32, c_vp, c_i32, c_f32, c_vp,
                                      c_i64, c_vp, ctypes.POINTER(c_i32), c_i32, c_vp],
    "nect_grid_backward_ray_groups_cols": [ctypes.POINTER(NectGridDesc), ctypes.POINTER(c_i32), c_vp, c_vp, c_i32,
                                           c_vp, c_i32, c_f32, c_i32, c_vp, c_i64, c_vp, c_vp,
                                           ctypes.POINTER(c_i32), c_i32, c_vp],
    "nect_calibrate_atomics": [c_vp, c_i64, c_i32, c_i32, c_i64, ctypes.c_uint32, c_vp],
    "nect_wgrad": [c_i32, c_vp, c_i64, c_vp, c_i64, c_i64, c_i32, c_i32, c_f32, c_vp, c_i64, c_vp],
    "nect_gemm_act": [c_i32, c_vp, c_i64, c_vp, c_i64, c_i64, c_i32, c_i32, c_i32, c_f32, c_vp, c_i64, c_vp, c_i64,
                      c_vp],
    "nect_integral_loss": [c_vp, c_vp, c_i32, c_vp, c_vp, c_f32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "nect_adam": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i64, c_f32, c_f32, c_f32, c_f32, c_f32, c_i32, c_f32, c_vp],
    "nect_cast": [c_vp, c_vp, c_i32, c_i64, c_vp],
    "nect_amp_init": [c_vp, c_f32, c_vp],
    "nect_grad_check": [c_vp, c_i64, c_vp, c_vp],
    "nect_adam_amp": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i64, c_f32, c_f32, c_f32, c_f32, c_f32, c_vp, c_f32, c_f32,
                      c_i32, c_f32, c_vp],
    "nect_query": [ctypes.POINTER(NectGridDesc), ctypes.POINTER(NectMlpDesc), c_i32, c_vp, c_vp, c_vp, c_vp, c_i32,
                   c_i64, c_i32, c_vp],
}

# entry points that do not return an int32 status
_SIZE_QUERIES: dict[str, list] = {
    "nect_mlp_backward_workspace": [ctypes.POINTER(NectMlpDesc), c_i32, c_i64],
    "nect_ray_order_workspace": [ctypes.POINTER(NectGeometry)],
}

_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"nect_amd HIP library not found at {LIB_PATH}; build it with `python -m nect_amd.build`"
            )
        handle = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
        handle.nect_last_error.restype = ctypes.c_char_p
        handle.nect_last_error.argtypes = []
        handle.nect_abi_version.restype = c_i32
        for name, args in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = c_i32
            fn.argtypes = args
        for name, args in _SIZE_QUERIES.items():
            fn = getattr(handle, name)
            fn.restype = c_i64
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return ["nect_last_error", "nect_abi_version", *_SIGNATURES, *_SIZE_QUERIES]


def call(name: str, *args) -> None:
    """Calls an entry point and raises RuntimeError with the library's message on failure."""
    handle = lib()
    rc = getattr(handle, name)(*args)
    if rc != 0:
        msg = handle.nect_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed: {msg}")


def ptr(t: torch.Tensor | None) -> int | None:
    """Device pointer of a GPU tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("nect_amd kernels take GPU tensors only (got a CPU tensor)")
    if not t.is_contiguous():
        raise ValueError("nect_amd kernels take contiguous tensors")
    return t.data_ptr()


def stream_of(device: torch.device) -> int:
    """hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream
