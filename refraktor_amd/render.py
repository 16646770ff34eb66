"""Python mirror of librefrakt's render API over the C ABI.

    db = VariationDB()                          # rfkt::initialize
    k = db.compile(flame_dict)                  # flame_compiler::get_flame_kernel
    samples = k.pack_samples(width, height)     # flame::pack_samples
    st = k.warmup(samples, width, height, seed) # flame_kernel::warmup
    res = st.bin(iters=1 << 26)                 # flame_kernel::bin
    img = st.tonemap(gamma, brightness, vibrancy)   # tonemapper::run

Every call goes through librefrakt_amd.so; device buffers are owned by the
library or allocated with rfk_malloc.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import check

CONFIG_DIR = Path(__file__).resolve().parent / "config"


class DeviceBuffer:
    """A raw HBM allocation made by the library."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(_lib.lib().rfk_malloc(C.byref(p), max(self.nbytes, 16)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(_lib.lib().rfk_memcpy_htod(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(_lib.lib().rfk_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self) -> None:
        check(_lib.lib().rfk_memset(self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr:
            _lib.lib().rfk_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download(ptr: int, dtype, shape) -> np.ndarray:
    out = np.empty(shape, dtype=dtype)
    check(_lib.lib().rfk_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes))
    return out


def device_count() -> int:
    n = C.c_int(0)
    check(_lib.lib().rfk_device_count(C.byref(n)))
    return n.value


class VariationDB:
    def __init__(self, config_dir: str | Path | None = None):
        self._lib = _lib.lib()
        h = C.c_void_p()
        nerr = C.c_int(0)
        check(self._lib.rfk_db_load(str(config_dir or CONFIG_DIR).encode(), C.byref(h), C.byref(nerr)))
        self.handle = h.value
        self.errors = [self._lib.rfk_db_error(self.handle, i).decode() for i in range(nerr.value)]

    @property
    def variations(self) -> list[str]:
        n = self._lib.rfk_db_variation_count(self.handle)
        return [self._lib.rfk_db_variation_name(self.handle, i).decode() for i in range(n)]

    @property
    def hash(self) -> int:
        return self._lib.rfk_db_hash(self.handle)

    def flame(self, flame: dict | str) -> "Flame":
        return Flame(self, flame)

    def align(self, flames: list) -> list["Flame"]:
        """structurally aligned copies of keyframes (one kernel serves them all)"""
        fs = [f if isinstance(f, Flame) else self.flame(f) for f in flames]
        arr = (C.c_void_p * len(fs))(*[f.handle for f in fs])
        out = (C.c_void_p * len(fs))()
        check(self._lib.rfk_flames_align(self.handle, arr, len(fs), out))
        return [Flame(self, None, handle=out[i]) for i in range(len(fs))]

    def flame_from_flam3(self, xml: str) -> "Flame":
        h = C.c_void_p()
        check(self._lib.rfk_flame_from_flam3(self.handle, xml.encode(), C.byref(h)))
        return Flame(self, None, handle=h.value)

    def compile(self, flame, precision: int = 0, threads_per_block: int = 256, atomic_width: int = 32,
                bin_tile_log2: int = 0, hist_bits: int = 32, bin_queue: bool = False) -> "Kernel":
        f = flame if isinstance(flame, Flame) else self.flame(flame)
        return Kernel(self, f, precision, threads_per_block, atomic_width, bin_tile_log2, hist_bits, bin_queue)

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_db_free(self.handle)
        except Exception:
            pass


class Flame:
    def __init__(self, db: VariationDB, flame: dict | str | None, handle: int | None = None):
        self.db = db
        self._lib = db._lib
        if handle is None:
            text = flame if isinstance(flame, str) else json.dumps(flame)
            h = C.c_void_p()
            check(self._lib.rfk_flame_from_json(db.handle, text.encode(), C.byref(h)))
            handle = h.value
        self.handle = handle

    def size_reals(self) -> int:
        n = C.c_size_t()
        check(self._lib.rfk_flame_size_reals(self.handle, C.byref(n)))
        return n.value

    def affine_indices(self) -> list[int]:
        cnt = C.c_size_t()
        check(self._lib.rfk_flame_affine_indices(self.handle, None, 0, C.byref(cnt)))
        buf = (C.c_size_t * cnt.value)()
        check(self._lib.rfk_flame_affine_indices(self.handle, buf, cnt.value, C.byref(cnt)))
        return list(buf)

    def pack_samples(self, width: int, height: int, start: float = 0.0, offset: float = 0.0,
                     count: int = 4) -> np.ndarray:
        n = (self.size_reals() + 768) * count
        out = np.empty(n, dtype=np.float64)
        check(self._lib.rfk_flame_pack_samples(self.handle, start, offset, count, width, height,
                                               out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def to_json(self) -> dict:
        need = C.c_size_t()
        check(self._lib.rfk_flame_to_json(self.handle, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(self._lib.rfk_flame_to_json(self.handle, buf, need.value, C.byref(need)))
        return json.loads(buf.value.decode())

    def source(self, precision: int = 0, threads_per_block: int = 256) -> str:
        cfg = _lib.ExecConfig(threads_per_block, 32, 0, 0)
        need = C.c_size_t()
        check(self._lib.rfk_make_source(self.db.handle, self.handle, precision, C.byref(cfg), None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(self._lib.rfk_make_source(self.db.handle, self.handle, precision, C.byref(cfg), buf, need.value,
                                        C.byref(need)))
        return buf.value.decode()

    def compile_check(self, precision: int = 0, threads_per_block: int = 256, atomic_width: int = 32,
                      bin_tile_log2: int = 0, hist_bits: int = 32, bin_queue: bool = False) -> int:
        """hiprtc-compile the generated source without a device; returns code-object bytes"""
        cfg = _lib.ExecConfig(threads_per_block, atomic_width, bin_tile_log2, hist_bits, int(bin_queue))
        n = C.c_size_t()
        log = C.create_string_buffer(1 << 16)
        rc = self._lib.rfk_compile_check(self.db.handle, self.handle, precision, C.byref(cfg), C.byref(n), log,
                                         len(log))
        if rc != 0:
            raise _lib.RfkError(rc, self._lib.rfk_last_error().decode(errors="replace"))
        return n.value

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_flame_free(self.handle)
        except Exception:
            pass


class Kernel:
    def __init__(self, db: VariationDB, flame: Flame, precision: int, tpb: int, atomic_width: int,
                 bin_tile_log2: int = 0, hist_bits: int = 32, bin_queue: bool = False):
        self.db, self.flame, self._lib = db, flame, db._lib
        self.precision = precision
        self.hist_bits = hist_bits
        cfg = _lib.ExecConfig(tpb, atomic_width, bin_tile_log2, hist_bits, int(bin_queue))
        h = C.c_void_p()
        log = C.create_string_buffer(1 << 16)
        rc = self._lib.rfk_compile(db.handle, flame.handle, precision, C.byref(cfg), C.byref(h), log, len(log))
        if rc != 0:
            raise _lib.RfkError(rc, self._lib.rfk_last_error().decode(errors="replace"))
        self.handle = h.value
        self.compile_ms = self._lib.rfk_kernel_compile_ms(self.handle)
        self.threads_per_block = tpb
        self.chaos = "chaos" in flame.to_json()

    def pack_samples(self, width: int, height: int, start: float = 0.0, offset: float = 0.0, count: int = 4):
        return self.flame.pack_samples(width, height, start, offset, count)

    def warmup(self, samples: np.ndarray, width: int, height: int, seed: int, fuse: int = 20,
               walkers: int = 1 << 20, temporal_multiplier: int = 1, stream_id: int = 0) -> "State":
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        h = C.c_void_p()
        check(self._lib.rfk_warmup_ex(self.handle, None, samples.ctypes.data_as(C.POINTER(C.c_double)), samples.size,
                                      width, height, seed, stream_id, fuse, walkers, temporal_multiplier,
                                      C.byref(h)))
        return State(self, h.value)

    def batch(self, samples: np.ndarray, frames: int, iters, width: int, height: int, walkers: int = 1 << 18,
              seed: int = 1, fuse: int = 20) -> "Batch":
        """`frames` frames over the Catmull-Rom segments of `samples` (>= 4
        packed samples) in one persistent launch; iters: per-frame budgets"""
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        it = np.ascontiguousarray(np.broadcast_to(np.asarray(iters, dtype=np.uint64), (frames,)))
        h = C.c_void_p()
        check(self._lib.rfk_batch_create(self.handle, None, samples.ctypes.data_as(C.POINTER(C.c_double)),
                                         samples.size, frames, it.ctypes.data_as(C.POINTER(C.c_uint64)), width,
                                         height, walkers, seed, fuse, C.byref(h)))
        return Batch(self, h.value)

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_kernel_free(self.handle)
        except Exception:
            pass


class State:
    def __init__(self, kernel: Kernel, handle: int):
        self.kernel, self._lib, self.handle = kernel, kernel._lib, handle

    def rewarm(self, samples: np.ndarray, seed: int, fuse: int = 20, stream_id: int = 0) -> "State":
        """new frame on the same allocations (same kernel, size and walkers)"""
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        check(self._lib.rfk_rewarm(self.kernel.handle, self.handle, samples.ctypes.data_as(C.POINTER(C.c_double)),
                                   samples.size, seed, stream_id, fuse))
        return self

    def info(self) -> _lib.StateInfo:
        inf = _lib.StateInfo()
        check(self._lib.rfk_state_get_info(self.handle, C.byref(inf)))
        return inf

    def bin(self, iters: int = 0, millis: float = 0.0, quality: float = 0.0) -> dict:
        bo = _lib.Bailout(iters, millis, quality)
        res = _lib.BinResult()
        check(self._lib.rfk_bin(self.kernel.handle, self.handle, C.byref(bo), C.byref(res)))
        return res.as_dict()

    def bin_async(self, iters: int = 0, millis: float = 0.0, quality: float = 0.0) -> None:
        bo = _lib.Bailout(iters, millis, quality)
        check(self._lib.rfk_bin_async(self.kernel.handle, self.handle, C.byref(bo)))

    def bin_wait(self) -> dict:
        res = _lib.BinResult()
        check(self._lib.rfk_bin_wait(self.handle, C.byref(res)))
        return res.as_dict()

    def abort(self) -> None:
        check(self._lib.rfk_abort(self.handle))

    def histogram(self) -> np.ndarray:
        """(H, W, 4) R,G,B,W fixed-point bins: u32, or u64 for hist_bits=64 kernels"""
        inf = self.info()
        dt = np.uint64 if inf.hist_bits == 64 else np.uint32
        return download(inf.histogram, dt, (inf.height, inf.width, 4))

    def params(self) -> np.ndarray:
        inf = self.info()
        dt = np.float64 if self.kernel.precision == 1 else np.float32
        return download(inf.params, dt, (inf.blocks, inf.param_stride))

    def palettes(self) -> np.ndarray:
        inf = self.info()
        return download(inf.palettes, np.uint32, (inf.blocks, 256))

    def walkers(self) -> np.ndarray:
        inf = self.info()
        return np.stack([download(inf.walker_x, np.float32, (inf.walkers,)),
                         download(inf.walker_y, np.float32, (inf.walkers,)),
                         download(inf.walker_c, np.float32, (inf.walkers,))], axis=1)

    def walkers_last(self) -> np.ndarray:
        """each walker's last xform (xaos flames; 0xFFFFFFFF before the first pass)"""
        inf = self.info()
        if not inf.walker_last:
            raise ValueError("not a xaos flame")
        return download(inf.walker_last, np.uint32, (inf.walkers,))

    def trace(self, passes: int) -> tuple[np.ndarray, np.ndarray]:
        inf = self.info()
        tr = DeviceBuffer(passes * inf.walkers * 16)
        xi = DeviceBuffer(passes * inf.walkers * 4)
        check(self._lib.rfk_trace(self.kernel.handle, self.handle, passes, tr.ptr, xi.ptr))
        # xaos kernels record a choice per walker, the others one per wave
        per = inf.walkers if self.kernel.chaos else inf.walkers // 64
        return (tr.download(np.float32, (passes, inf.walkers, 4)), xi.download(np.uint32, (passes, per)))

    def bin_points(self, pts: np.ndarray) -> None:
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 4)
        buf = DeviceBuffer(pts.nbytes)
        buf.upload(pts)
        check(self._lib.rfk_bin_points(self.handle, buf.ptr, pts.shape[0]))

    def clear_histogram(self) -> None:
        check(self._lib.rfk_state_clear_histogram(self.handle))

    def density_histogram(self) -> tuple[np.ndarray, float]:
        bins = (C.c_uint64 * 97)()
        md = C.c_double()
        check(self._lib.rfk_density_histogram(self.handle, bins, C.byref(md)))
        return np.array(list(bins), dtype=np.uint64), md.value

    def density_estimate(self, max_radius: float, min_radius: float = 0.0, curve: float = 0.4) -> np.ndarray:
        inf = self.info()
        out = DeviceBuffer(inf.width * inf.height * 16)
        check(self._lib.rfk_density_estimate(self.handle, max_radius, min_radius, curve, out.ptr))
        return out.download(np.float32, (inf.height, inf.width, 4))

    def tonemap(self, gamma: float, brightness: float, vibrancy: float, quality: float = 0.0,
                fmt: int = _lib.PIXEL_FLOAT4) -> np.ndarray:
        inf = self.info()
        npx = inf.width * inf.height
        out = DeviceBuffer(npx * _lib.PIXEL_BYTES[fmt])
        args = _lib.TonemapArgs(quality, gamma, brightness, vibrancy, 0.0)
        check(self._lib.rfk_tonemap(self.handle, out.ptr, fmt, C.byref(args)))
        return _host_pixels(out, fmt, inf.height, inf.width)

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_state_free(self.handle)
        except Exception:
            pass


class Batch:
    """frames of an animation rendered by one persistent launch (optionally a hipGraph replay)"""

    def __init__(self, kernel: Kernel, handle: int):
        self.kernel, self._lib, self.handle = kernel, kernel._lib, handle
        n = [C.c_uint32() for _ in range(5)]
        check(self._lib.rfk_batch_info(self.handle, *[C.byref(v) for v in n]))
        self.frames, self.width, self.height, self.grid, self.items = (v.value for v in n)

    def run(self, graph: bool = True) -> float:
        ms = C.c_double()
        check(self._lib.rfk_batch_run(self.handle, 1 if graph else 0, C.byref(ms)))
        return ms.value

    def frame(self, f: int) -> dict:
        inf = _lib.BatchFrameInfo()
        check(self._lib.rfk_batch_frame(self.handle, f, C.byref(inf)))
        return {name: getattr(inf, name) for name, _ in inf._fields_}

    def histogram(self, f: int) -> np.ndarray:
        inf = self.frame(f)
        dt = np.uint64 if inf["hist_bits"] == 64 else np.uint32
        return download(inf["histogram"], dt, (self.height, self.width, 4))

    def params(self, f: int) -> np.ndarray:
        """frame f's parameter rows after precalc, one per temporal sample"""
        inf = self.frame(f)
        dt = np.float64 if self.kernel.precision == 1 else np.float32
        return download(inf["params"], dt, (inf["blocks"], inf["param_stride"]))

    def palettes(self, f: int) -> np.ndarray:
        inf = self.frame(f)
        return download(inf["palettes"], np.uint32, (inf["blocks"], 256))

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_batch_free(self.handle)
        except Exception:
            pass


def _host_pixels(buf: DeviceBuffer, fmt: int, h: int, w: int) -> np.ndarray:
    if fmt == _lib.PIXEL_FLOAT4:
        return buf.download(np.float32, (h, w, 4))
    if fmt == _lib.PIXEL_FLOAT3:
        return buf.download(np.float32, (h, w, 3))
    if fmt == _lib.PIXEL_HALF4:
        return buf.download(np.float16, (h, w, 4))
    if fmt == _lib.PIXEL_HALF3:
        return buf.download(np.float16, (h, w, 3))
    if fmt == _lib.PIXEL_RGB8:
        return buf.download(np.uint8, (h, w, 3))
    return buf.download(np.uint8, (h, w, 4))


def tonemap_image(img: np.ndarray, gamma: float, brightness: float, vibrancy: float, quality: float,
                  fmt: int = _lib.PIXEL_FLOAT4) -> np.ndarray:
    """tonemap a host float4 RGBA image in draw units on the device"""
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape[:2]
    src = DeviceBuffer(img.nbytes)
    src.upload(img)
    out = DeviceBuffer(h * w * _lib.PIXEL_BYTES[fmt])
    args = _lib.TonemapArgs(quality, gamma, brightness, vibrancy, 0.0)
    check(_lib.lib().rfk_tonemap_image(src.ptr, w, h, out.ptr, fmt, C.byref(args), None))
    return _host_pixels(out, fmt, h, w)


def density_hdr(img: np.ndarray, bins97: np.ndarray) -> np.ndarray:
    """density_hdr CDF remap of a host float4 tonemapped image on the device"""
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape[:2]
    src = DeviceBuffer(img.nbytes)
    src.upload(img)
    out = DeviceBuffer(h * w * 16)
    b = np.ascontiguousarray(bins97, dtype=np.uint64)
    check(_lib.lib().rfk_density_hdr(src.ptr, _lib.PIXEL_FLOAT4, h * w, b.ctypes.data_as(C.POINTER(C.c_uint64)),
                                     out.ptr, None))
    return out.download(np.float32, (h, w, 4))


def convert_image(img: np.ndarray, fmt: int, flip: bool = False, swap: bool = False) -> np.ndarray:
    """rfkt::converter on the device: a host half3 / half4 image (float16,
    h x w x 3|4) to `fmt` (PIXEL_FLOAT4 ... PIXEL_RGB8)"""
    img = np.ascontiguousarray(img, dtype=np.float16)
    h, w, c = img.shape
    if c not in (3, 4):
        raise ValueError("conversion reads half3 or half4 images")
    src = DeviceBuffer(img.nbytes)
    src.upload(img)
    out = DeviceBuffer(h * w * _lib.PIXEL_BYTES[fmt])
    flags = (_lib.CONVERT_FLIP if flip else 0) | (_lib.CONVERT_SWAP if swap else 0)
    check(_lib.lib().rfk_convert_image(src.ptr, _lib.PIXEL_HALF4 if c == 4 else _lib.PIXEL_HALF3, out.ptr, fmt, w, h,
                                       flags, None))
    return _host_pixels(out, fmt, h, w)


class JpegEncoder:
    """rfkt::jpeg_encoder on the device: 8-bit RGB / RGBA images (device
    buffers or host arrays) to baseline JFIF bytes"""

    def __init__(self):
        self._lib = _lib.lib()
        h = C.c_void_p()
        check(self._lib.rfk_jpeg_create(C.byref(h)))
        self.handle = h.value

    def encode_device(self, ptr: int, width: int, height: int, quality: int = 90, fmt: int = _lib.PIXEL_RGB8,
                      pitch: int = 0, stream=None) -> bytes:
        bpp = 3 if fmt == _lib.PIXEL_RGB8 else 4
        check(self._lib.rfk_jpeg_encode(self.handle, ptr, fmt, pitch or width * bpp, width, height, quality, stream))
        n = C.c_size_t()
        check(self._lib.rfk_jpeg_retrieve(self.handle, None, 0, C.byref(n)))
        out = C.create_string_buffer(n.value)
        check(self._lib.rfk_jpeg_retrieve(self.handle, out, n.value, C.byref(n)))
        return out.raw[:n.value]

    def encode(self, img: np.ndarray, quality: int = 90) -> bytes:
        """a host uint8 h x w x 3 (RGB) or x 4 (RGBA) image"""
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w, c = img.shape
        buf = DeviceBuffer(img.nbytes)
        buf.upload(img)
        return self.encode_device(buf.ptr, w, h, quality, _lib.PIXEL_RGB8 if c == 3 else _lib.PIXEL_RGBA8)

    def __del__(self):
        try:
            if self.handle:
                self._lib.rfk_jpeg_free(self.handle)
        except Exception:
            pass
