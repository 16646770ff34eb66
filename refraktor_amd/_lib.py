"""ctypes bindings of `librefrakt_amd.so` (declared in include/refrakt.h).

The library must have been built (refraktor_amd.build / __graft_entry__.build);
importing this module raises if it is missing — there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librefrakt_amd.so"

RFK_OK = 0
RFK_ERR_INVALID = -1
RFK_ERR_PARSE = -2
RFK_ERR_COMPILE = -3
RFK_ERR_HIP = -4
RFK_ERR_NODEVICE = -5
RFK_ERR_OVERFLOW = -6

PIXEL_FLOAT4, PIXEL_HALF4, PIXEL_FLOAT3, PIXEL_HALF3, PIXEL_RGBA8, PIXEL_RGB8 = 0, 1, 2, 3, 4, 5
PIXEL_BYTES = {0: 16, 1: 8, 2: 12, 3: 6, 4: 4, 5: 3}
CONVERT_FLIP, CONVERT_SWAP = 1, 2


class ExecConfig(C.Structure):
    _fields_ = [("threads_per_block", C.c_uint32), ("atomic_width", C.c_uint32),
                ("bin_tile_log2", C.c_uint32), ("hist_bits", C.c_uint32), ("bin_queue", C.c_uint32)]


class Bailout(C.Structure):
    _fields_ = [("iters", C.c_uint64), ("millis", C.c_double), ("quality", C.c_double)]


class BinResult(C.Structure):
    _fields_ = [("quality", C.c_double), ("elapsed_ms", C.c_double), ("kernel_ms", C.c_double),
                ("total_passes", C.c_uint64), ("total_draws", C.c_uint64), ("total_bins", C.c_uint64),
                ("passes_per_thread", C.c_double), ("max_density", C.c_double),
                ("bad_values", C.c_uint64), ("hit_sum", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class StateInfo(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("walkers", C.c_uint32),
                ("threads_per_block", C.c_uint32), ("blocks", C.c_uint32), ("size_reals", C.c_uint32),
                ("param_stride", C.c_uint32), ("seed", C.c_uint32), ("passes_done", C.c_uint64),
                ("histogram", C.c_void_p), ("params", C.c_void_p), ("palettes", C.c_void_p),
                ("walker_x", C.c_void_p), ("walker_y", C.c_void_p), ("walker_c", C.c_void_p),
                ("stream", C.c_void_p), ("quality", C.c_double), ("hist_bits", C.c_uint32),
                ("walker_last", C.c_void_p)]


class BatchFrameInfo(C.Structure):
    _fields_ = [("histogram", C.c_void_p), ("quality", C.c_double), ("total_draws", C.c_uint64),
                ("total_passes", C.c_uint64), ("bad_values", C.c_uint64), ("passes_per_walker", C.c_uint32),
                ("walkers", C.c_uint32), ("hist_bits", C.c_uint32), ("params", C.c_void_p),
                ("palettes", C.c_void_p), ("param_stride", C.c_uint32), ("blocks", C.c_uint32), ("seed", C.c_uint32),
                ("stream_id", C.c_uint32), ("fuse", C.c_uint32)]


class TuneCandidate(C.Structure):
    _fields_ = [("threads_per_block", C.c_uint32), ("atomic_width", C.c_uint32), ("bin_tile_log2", C.c_uint32),
                ("walkers", C.c_uint32), ("bin_queue", C.c_uint32), ("iters_per_s", C.c_double),
                ("compile_ms", C.c_double), ("status", C.c_int)]


class TonemapArgs(C.Structure):
    _fields_ = [("quality", C.c_double), ("gamma", C.c_double), ("brightness", C.c_double),
                ("vibrancy", C.c_double), ("max_density", C.c_double)]


# (name, restype, argtypes)
_P = C.c_void_p
_SIGS = [
    ("rfk_last_error", C.c_char_p, []),
    ("rfk_version", C.c_char_p, []),
    ("rfk_db_load", C.c_int, [C.c_char_p, C.POINTER(_P), C.POINTER(C.c_int)]),
    ("rfk_db_free", C.c_int, [_P]),
    ("rfk_db_variation_count", C.c_int, [_P]),
    ("rfk_db_variation_name", C.c_char_p, [_P, C.c_int]),
    ("rfk_db_error", C.c_char_p, [_P, C.c_int]),
    ("rfk_db_hash", C.c_uint64, [_P]),
    ("rfk_flame_from_json", C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    ("rfk_flame_from_flam3", C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    ("rfk_flame_to_json", C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("rfk_flame_free", C.c_int, [_P]),
    ("rfk_flame_size_reals", C.c_int, [_P, C.POINTER(C.c_size_t)]),
    ("rfk_flame_affine_indices", C.c_int, [_P, C.POINTER(C.c_size_t), C.c_size_t, C.POINTER(C.c_size_t)]),
    ("rfk_flame_pack_samples", C.c_int, [_P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_double), C.c_size_t]),
    ("rfk_flame_xform_count", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("rfk_flames_align", C.c_int, [_P, C.POINTER(_P), C.c_int, C.POINTER(_P)]),
    ("rfk_compile", C.c_int, [_P, _P, C.c_int, C.POINTER(ExecConfig), C.POINTER(_P), C.c_char_p, C.c_size_t]),
    ("rfk_compile_check", C.c_int, [_P, _P, C.c_int, C.POINTER(ExecConfig), C.POINTER(C.c_size_t), C.c_char_p,
                                    C.c_size_t]),
    ("rfk_make_source", C.c_int, [_P, _P, C.c_int, C.POINTER(ExecConfig), C.c_char_p, C.c_size_t,
                                  C.POINTER(C.c_size_t)]),
    ("rfk_kernel_free", C.c_int, [_P]),
    ("rfk_kernel_size_reals", C.c_int, [_P, C.POINTER(C.c_size_t)]),
    ("rfk_kernel_threads_per_block", C.c_int, [_P, C.POINTER(C.c_uint32)]),
    ("rfk_kernel_compile_ms", C.c_double, [_P]),
    ("rfk_warmup", C.c_int, [_P, _P, C.POINTER(C.c_double), C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32,
                             C.c_uint32, C.c_uint32, C.c_int, C.POINTER(_P)]),
    ("rfk_warmup_ex", C.c_int, [_P, _P, C.POINTER(C.c_double), C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(_P)]),
    ("rfk_rewarm", C.c_int, [_P, _P, C.POINTER(C.c_double), C.c_size_t, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("rfk_state_free", C.c_int, [_P]),
    ("rfk_autotune", C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_double), C.c_size_t, C.c_uint32, C.c_uint32,
                               C.c_uint64, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("rfk_batch_create", C.c_int, [_P, _P, C.POINTER(C.c_double), C.c_size_t, C.c_uint32, C.POINTER(C.c_uint64),
                                   C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    ("rfk_batch_run", C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    ("rfk_batch_frame", C.c_int, [_P, C.c_uint32, C.c_void_p]),
    ("rfk_batch_info", C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("rfk_batch_free", C.c_int, [_P]),
    ("rfk_bin", C.c_int, [_P, _P, C.POINTER(Bailout), C.POINTER(BinResult)]),
    ("rfk_bin_async", C.c_int, [_P, _P, C.POINTER(Bailout)]),
    ("rfk_bin_wait", C.c_int, [_P, C.POINTER(BinResult)]),
    ("rfk_abort", C.c_int, [_P]),
    ("rfk_bin_points", C.c_int, [_P, _P, C.c_size_t]),
    ("rfk_trace", C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    ("rfk_state_get_info", C.c_int, [_P, C.POINTER(StateInfo)]),
    ("rfk_state_clear_histogram", C.c_int, [_P]),
    ("rfk_density_histogram", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    ("rfk_density_estimate", C.c_int, [_P, C.c_float, C.c_float, C.c_float, _P]),
    ("rfk_tonemap", C.c_int, [_P, _P, C.c_int, C.POINTER(TonemapArgs)]),
    ("rfk_tonemap_image", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_int, C.POINTER(TonemapArgs), _P]),
    ("rfk_density_hdr", C.c_int, [_P, C.c_int, C.c_size_t, C.POINTER(C.c_uint64), _P, _P]),
    ("rfk_convert_image", C.c_int, [_P, C.c_int, _P, C.c_int, C.c_uint32, C.c_uint32, C.c_int, _P]),
    ("rfk_jpeg_create", C.c_int, [C.POINTER(_P)]),
    ("rfk_jpeg_encode", C.c_int, [_P, _P, C.c_int, C.c_size_t, C.c_uint32, C.c_uint32, C.c_int, _P]),
    ("rfk_jpeg_retrieve", C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("rfk_jpeg_free", C.c_int, [_P]),
    ("rfk_catmull_coefficients", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    ("rfk_fix_rotation", C.c_int, [C.POINTER(C.c_double), C.c_size_t, C.c_size_t, C.POINTER(C.c_size_t),
                                   C.c_size_t, C.c_int]),
    ("rfk_denoiser_create", C.c_int, [C.c_uint64, C.POINTER(_P)]),
    ("rfk_denoiser_free", C.c_int, [_P]),
    ("rfk_denoiser_param_count", C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("rfk_denoiser_save", C.c_int, [_P, C.c_char_p]),
    ("rfk_denoiser_load", C.c_int, [_P, C.c_char_p]),
    ("rfk_denoiser_get_params", C.c_int, [_P, _P, _P]),
    ("rfk_denoiser_set_params", C.c_int, [_P, _P, _P]),
    ("rfk_denoiser_get_grads", C.c_int, [_P, _P, _P]),
    ("rfk_tonemap_features", C.c_int, [_P, C.POINTER(TonemapArgs), _P, _P]),
    ("rfk_denoiser_features", C.c_int, [_P, C.c_int, _P, C.c_size_t, _P, _P, _P]),
    ("rfk_denoiser_forward", C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
    ("rfk_denoiser_train_step", C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_float,
                                          C.POINTER(C.c_float), _P]),
    ("rfk_nn_conv3x3_forward", C.c_int, [_P, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int,
                                         C.c_int, _P, _P, _P]),
    ("rfk_nn_conv3x3_forward_ex", C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P,
                                            C.c_int, C.c_int, _P, _P, _P]),
    ("rfk_nn_conv3x3_forward_pool", C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P,
                                              C.c_int, C.c_int, _P, _P, _P]),
    ("rfk_nn_avgpool2", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P]),
    ("rfk_nn_conv3x3_pair_forward", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P,
                                              C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    ("rfk_nn_conv3x3_wgrad", C.c_int, [_P, C.c_int, _P, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P]),
    ("rfk_nn_conv3x3_dgrad", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, _P]),
    ("rfk_nn_gather_crops", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_int, C.c_int, _P,
                                      _P]),
    ("rfk_nn_inject_alloc_failures", C.c_int, [C.c_int]),
    ("rfk_nn_gather_crops_noalloc", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_int,
                                              C.c_int, _P, _P]),
    ("rfk_denoise", C.c_int, [_P, _P, C.c_int, _P, C.c_int, C.c_uint32, C.c_uint32, _P]),
    ("rfk_atomic_peak", C.c_int, [C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    ("rfk_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("rfk_malloc", C.c_int, [C.POINTER(_P), C.c_size_t]),
    ("rfk_free", C.c_int, [_P]),
    ("rfk_memcpy_htod", C.c_int, [_P, _P, C.c_size_t]),
    ("rfk_memcpy_dtoh", C.c_int, [_P, _P, C.c_size_t]),
    ("rfk_memcpy_dtod", C.c_int, [_P, _P, C.c_size_t]),
    ("rfk_memset", C.c_int, [_P, C.c_int, C.c_size_t]),
    ("rfk_synchronize", C.c_int, []),
]

EXPORTED_SYMBOLS = [name for name, _, _ in _SIGS]


class RfkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rfk {code}] {msg}")
        self.code = code


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is not built; run refraktor_amd.build.build_native() (or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(rc: int) -> None:
    if rc != RFK_OK:
        raise RfkError(rc, lib().rfk_last_error().decode(errors="replace"))
