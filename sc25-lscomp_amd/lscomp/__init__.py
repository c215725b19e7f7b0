"""Python binding of the gfx950 lsCOMP uint16 codec (liblscomp.so, C ABI in include/lscomp.h).

Host-side mirror of the reference's compressor interface
(lsCOMP/include/cuLSZ_entry.h:10-11): ``compress`` / ``decompress`` take a
device volume of ``dims = (x, y, z)`` (x frames, z fastest), four
quantization ``bins`` (x <= y <= z <= w) and a ``pooling`` threshold in [0, 1].
Lossless is ``bins=(1, 1, 1, 1), pooling=1.0``.

Device memory and streams come from PyTorch (plumbing only); every byte of
compression and decompression is computed by the HIP kernels in the shared
library.  There is no CPU fallback: if the library is missing, importing this
module raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# LSCOMP_LIBRARY selects another build of the same library (e.g. a variant
# compiled with different flags); default: the in-tree build.
LIB_PATH = os.environ.get("LSCOMP_LIBRARY") or os.path.join(_HERE, "liblscomp.so")

OK = 0
ERR_INVALID_ARG = 1
ERR_CAPACITY = 2
ERR_CORRUPT = 3
ERR_HIP = 4
ERR_TOO_LARGE = 5
ERR_INTERNAL = 6
ERR_RANGE = 7

LOSSLESS_BINS = (1, 1, 1, 1)


class LscompError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _lib.lscomp_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


class Dim3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32)]


class ContainerItem(ctypes.Structure):
    """lscomp_container_item."""
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32),
                ("elem_bytes", ctypes.c_uint32), ("b0", ctypes.c_uint32), ("b1", ctypes.c_uint32),
                ("b2", ctypes.c_uint32), ("b3", ctypes.c_uint32), ("pooling_th", ctypes.c_float),
                ("reserved", ctypes.c_uint32), ("offset", ctypes.c_uint64), ("stream_bytes", ctypes.c_uint64),
                ("tag", ctypes.c_uint64)]


class Bins4(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32),
                ("w", ctypes.c_uint32)]


class ItemDesc(ctypes.Structure):
    """lscomp_item_u16."""
    _fields_ = [("dims", Dim3), ("bins", Bins4), ("pooling_th", ctypes.c_float),
                ("d_volume", ctypes.c_void_p), ("d_stream", ctypes.c_void_p), ("stream_bytes", ctypes.c_uint64)]


# Every symbol include/lscomp.h declares.
EXPORTS = (
    "lscomp_num_blocks_u16", "lscomp_header_bytes_u16", "lscomp_max_compressed_bytes_u16",
    "lscomp_compress_u16", "lscomp_compress_u16_async", "lscomp_compress_u16_noalloc_async", "lscomp_decompress_u16",
    "lscomp_decompress_u16_async", "lscomp_stream_status", "lscomp_reserve_u16",
    "lscomp_release_all", "lscomp_set_launch_shape", "lscomp_get_launch_shape", "lscomp_status_string",
    "lscomp_version", "lscomp_plan_create", "lscomp_plan_set_items_u16", "lscomp_plan_launch",
    "lscomp_plan_launch_graph", "lscomp_plan_device_sizes", "lscomp_plan_sizes", "lscomp_plan_status",
    "lscomp_plan_destroy", "lscomp_server_create", "lscomp_server_submit_u16", "lscomp_server_poll",
    "lscomp_server_release", "lscomp_server_drain", "lscomp_server_replay_u16", "lscomp_server_destroy",
    "lscomp_autotune_u16", "lscomp_set_decode_prepass", "lscomp_error_stats_u16", "lscomp_ssim_u16", "lscomp_copy_bytes",
    "lscomp_max_compressed_bytes_u32", "lscomp_compress_u32", "lscomp_compress_u32_async",
    "lscomp_decompress_u32", "lscomp_decompress_u32_async",
    "lscomp_container_bytes", "lscomp_container_pack", "lscomp_container_parse",
    "lscomp_set_launch_shape_ex", "lscomp_get_launch_shape_ex", "lscomp_server_create_device",
    "lscomp_server_stats", "lscomp_plan_set_shape", "lscomp_plan_get_shape", "lscomp_plan_autotune",
)

REPLAY_HOLD = 1


class ServerStats(ctypes.Structure):
    """lscomp_server_stats_t."""
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("device_loop", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class LaunchShape(ctypes.Structure):
    """lscomp_launch_shape."""
    _fields_ = [("tile", ctypes.c_uint32), ("wg_per_cu", ctypes.c_uint32), ("lb_window", ctypes.c_uint32),
                ("counters", ctypes.c_uint32)]


class ErrorStats(ctypes.Structure):
    """lscomp_error_stats."""
    _fields_ = [("n", ctypes.c_uint64), ("max_abs_err", ctypes.c_uint32), ("min_orig", ctypes.c_uint32),
                ("max_orig", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("sum_sq_err", ctypes.c_uint64),
                ("sum_sq_dec", ctypes.c_uint64), ("max_rel_err", ctypes.c_double), ("mse", ctypes.c_double),
                ("range", ctypes.c_double), ("psnr", ctypes.c_double), ("norm_err", ctypes.c_double),
                ("norm_err_norm", ctypes.c_double), ("nrmse", ctypes.c_double)]


class TuneEntry(ctypes.Structure):
    """lscomp_tune_entry."""
    _fields_ = [("direction", ctypes.c_uint32), ("tile", ctypes.c_uint32), ("wg_per_cu", ctypes.c_uint32),
                ("lb_window", ctypes.c_uint32), ("counters", ctypes.c_uint32), ("ms", ctypes.c_float)]

    def as_dict(self):
        return {"direction": self.direction, "tile": self.tile, "wg_per_cu": self.wg_per_cu,
                "lb_window": self.lb_window, "counters": self.counters, "ms": self.ms}


class Request(ctypes.Structure):
    """lscomp_request_u16."""
    _fields_ = [("dims", Dim3), ("bins", Bins4), ("pooling_th", ctypes.c_float),
                ("d_frame", ctypes.c_void_p), ("tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    """lscomp_completion."""
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int),
                ("reserved", ctypes.c_uint32), ("stream_bytes", ctypes.c_uint64), ("d_stream", ctypes.c_void_p),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_launch", ctypes.c_double),
                ("t_done", ctypes.c_double)]


def _load():
    # PyTorch ships its own HIP runtime.  If liblscomp.so pulled the system
    # runtime into the process first, torch's would bind to it and the process
    # could see no device; importing torch first makes one runtime serve both.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C sc25-lscomp_amd")
    lib = ctypes.CDLL(LIB_PATH)
    vp, sz, u64p = ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)
    for name in ("lscomp_num_blocks_u16", "lscomp_header_bytes_u16", "lscomp_max_compressed_bytes_u16"):
        getattr(lib, name).restype = sz
        getattr(lib, name).argtypes = [Dim3]
    lib.lscomp_max_compressed_bytes_u32.restype = sz
    lib.lscomp_max_compressed_bytes_u32.argtypes = [Dim3]
    lib.lscomp_compress_u32.restype = ctypes.c_int
    lib.lscomp_compress_u32.argtypes = [vp, vp, sz, ctypes.POINTER(sz), Dim3, Bins4, ctypes.c_float, vp]
    lib.lscomp_compress_u32_async.restype = ctypes.c_int
    lib.lscomp_compress_u32_async.argtypes = [vp, vp, sz, vp, Dim3, Bins4, ctypes.c_float, vp]
    lib.lscomp_decompress_u32.restype = ctypes.c_int
    lib.lscomp_decompress_u32.argtypes = [vp, vp, sz, Dim3, Bins4, vp]
    lib.lscomp_decompress_u32_async.restype = ctypes.c_int
    lib.lscomp_decompress_u32_async.argtypes = [vp, vp, sz, Dim3, Bins4, vp]
    lib.lscomp_container_bytes.restype = sz
    lib.lscomp_container_bytes.argtypes = [ctypes.POINTER(ContainerItem), ctypes.c_uint32]
    lib.lscomp_container_pack.restype = ctypes.c_int
    lib.lscomp_container_pack.argtypes = [ctypes.POINTER(ContainerItem), ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_void_p), vp, sz, ctypes.POINTER(sz)]
    lib.lscomp_container_parse.restype = ctypes.c_int
    lib.lscomp_container_parse.argtypes = [vp, sz, ctypes.POINTER(ContainerItem), ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint32)]
    lib.lscomp_compress_u16.restype = ctypes.c_int
    lib.lscomp_compress_u16.argtypes = [vp, vp, sz, ctypes.POINTER(sz), Dim3, Bins4, ctypes.c_float, vp]
    lib.lscomp_compress_u16_async.restype = ctypes.c_int
    lib.lscomp_compress_u16_async.argtypes = [vp, vp, sz, vp, Dim3, Bins4, ctypes.c_float, vp]
    lib.lscomp_compress_u16_noalloc_async.restype = ctypes.c_int
    lib.lscomp_compress_u16_noalloc_async.argtypes = [vp, vp, sz, vp, Dim3, Bins4, ctypes.c_float, vp]
    lib.lscomp_decompress_u16.restype = ctypes.c_int
    lib.lscomp_decompress_u16.argtypes = [vp, vp, sz, Dim3, Bins4, vp]
    lib.lscomp_decompress_u16_async.restype = ctypes.c_int
    lib.lscomp_decompress_u16_async.argtypes = [vp, vp, sz, Dim3, Bins4, vp]
    lib.lscomp_stream_status.restype = ctypes.c_int
    lib.lscomp_stream_status.argtypes = [vp]
    lib.lscomp_reserve_u16.restype = ctypes.c_int
    lib.lscomp_reserve_u16.argtypes = [Dim3, vp]
    lib.lscomp_release_all.restype = None
    lib.lscomp_release_all.argtypes = []
    lib.lscomp_set_launch_shape.restype = ctypes.c_int
    lib.lscomp_set_launch_shape.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
    lib.lscomp_get_launch_shape.restype = ctypes.c_int
    lib.lscomp_get_launch_shape.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_uint32)]
    lib.lscomp_set_launch_shape_ex.restype = ctypes.c_int
    lib.lscomp_set_launch_shape_ex.argtypes = [ctypes.c_int, ctypes.POINTER(LaunchShape)]
    lib.lscomp_get_launch_shape_ex.restype = ctypes.c_int
    lib.lscomp_get_launch_shape_ex.argtypes = [ctypes.c_int, ctypes.POINTER(LaunchShape)]
    lib.lscomp_plan_create.restype = ctypes.c_int
    lib.lscomp_plan_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(vp)]
    lib.lscomp_plan_set_items_u16.restype = ctypes.c_int
    lib.lscomp_plan_set_items_u16.argtypes = [vp, ctypes.POINTER(ItemDesc), ctypes.c_uint32, vp]
    for name in ("lscomp_plan_launch", "lscomp_plan_launch_graph", "lscomp_plan_status"):
        getattr(lib, name).restype = ctypes.c_int
        getattr(lib, name).argtypes = [vp, vp]
    lib.lscomp_plan_set_shape.restype = ctypes.c_int
    lib.lscomp_plan_set_shape.argtypes = [vp, ctypes.POINTER(LaunchShape), vp]
    lib.lscomp_plan_get_shape.restype = ctypes.c_int
    lib.lscomp_plan_get_shape.argtypes = [vp, ctypes.POINTER(LaunchShape)]
    lib.lscomp_plan_autotune.restype = ctypes.c_int
    lib.lscomp_plan_autotune.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(TuneEntry),
                                         ctypes.POINTER(ctypes.c_uint32)]
    lib.lscomp_plan_device_sizes.restype = vp
    lib.lscomp_plan_device_sizes.argtypes = [vp]
    lib.lscomp_plan_sizes.restype = ctypes.c_int
    lib.lscomp_plan_sizes.argtypes = [vp, u64p, vp]
    lib.lscomp_plan_destroy.restype = ctypes.c_int
    lib.lscomp_plan_destroy.argtypes = [vp]
    lib.lscomp_server_create.restype = ctypes.c_int
    lib.lscomp_server_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.POINTER(vp)]
    lib.lscomp_server_create_device.restype = ctypes.c_int
    lib.lscomp_server_create_device.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(vp)]
    lib.lscomp_server_stats.restype = ctypes.c_int
    lib.lscomp_server_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    lib.lscomp_server_submit_u16.restype = ctypes.c_int
    lib.lscomp_server_submit_u16.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    lib.lscomp_server_poll.restype = ctypes.c_int
    lib.lscomp_server_poll.argtypes = [vp, ctypes.POINTER(Completion), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32)]
    lib.lscomp_server_release.restype = ctypes.c_int
    lib.lscomp_server_release.argtypes = [vp, ctypes.c_uint64]
    lib.lscomp_server_drain.restype = ctypes.c_int
    lib.lscomp_server_drain.argtypes = [vp]
    lib.lscomp_server_replay_u16.restype = ctypes.c_int
    lib.lscomp_server_replay_u16.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double),
                                             ctypes.c_uint32, ctypes.POINTER(Completion), ctypes.c_uint32]
    lib.lscomp_server_destroy.restype = ctypes.c_int
    lib.lscomp_server_destroy.argtypes = [vp]
    lib.lscomp_error_stats_u16.restype = ctypes.c_int
    lib.lscomp_error_stats_u16.argtypes = [vp, vp, sz, ctypes.POINTER(ErrorStats), vp]
    lib.lscomp_ssim_u16.restype = ctypes.c_int
    lib.lscomp_ssim_u16.argtypes = [vp, vp, Dim3, ctypes.POINTER(ctypes.c_double), vp]
    lib.lscomp_copy_bytes.restype = ctypes.c_int
    lib.lscomp_copy_bytes.argtypes = [vp, vp, sz, vp]
    lib.lscomp_set_decode_prepass.restype = ctypes.c_int
    lib.lscomp_set_decode_prepass.argtypes = [ctypes.c_int]
    lib.lscomp_autotune_u16.restype = ctypes.c_int
    lib.lscomp_autotune_u16.argtypes = [vp, vp, sz, vp, Dim3, Bins4, ctypes.c_float, ctypes.c_int, vp,
                                        ctypes.POINTER(TuneEntry), ctypes.POINTER(ctypes.c_uint32)]
    lib.lscomp_status_string.restype = ctypes.c_char_p
    lib.lscomp_status_string.argtypes = [ctypes.c_int]
    lib.lscomp_version.restype = ctypes.c_int
    lib.lscomp_version.argtypes = []
    _ = u64p
    return lib


_lib = _load()


def lib():
    return _lib


def _dims(dims) -> Dim3:
    x, y, z = (int(v) for v in dims)
    return Dim3(x, y, z)


def _bins(bins) -> Bins4:
    b = tuple(int(v) for v in bins)
    if len(b) != 4:
        raise ValueError("bins must have 4 entries")
    return Bins4(*b)


def num_blocks(dims) -> int:
    return int(_lib.lscomp_num_blocks_u16(_dims(dims)))


def header_bytes(dims) -> int:
    return int(_lib.lscomp_header_bytes_u16(_dims(dims)))


def max_compressed_bytes(dims) -> int:
    return int(_lib.lscomp_max_compressed_bytes_u16(_dims(dims)))


def _stream_handle(stream):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def _check_volume(t, dims, esize: int = 2):
    x, y, z = (int(v) for v in dims)
    if not t.is_cuda:
        raise ValueError("volume must be a device tensor")
    if t.element_size() != esize or not t.is_contiguous() or t.numel() != x * y * z:
        raise ValueError(f"volume must be a contiguous {esize}-byte tensor of x*y*z elements")


def compress(vol, dims, bins=LOSSLESS_BINS, pooling: float = 1.0, out=None, stream=None):
    """Compress a device volume; returns (stream tensor uint8, stream bytes).

    ``out`` (uint8 device tensor) defaults to a worst-case sized buffer.
    Raises LscompError on failure (ERR_CAPACITY if ``out`` is too small).
    """
    import torch
    _check_volume(vol, dims)
    if out is None:
        out = torch.empty(max_compressed_bytes(dims), dtype=torch.uint8, device=vol.device)
    n = ctypes.c_size_t(0)
    st = _lib.lscomp_compress_u16(ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                  out.numel(), ctypes.byref(n), _dims(dims), _bins(bins),
                                  ctypes.c_float(pooling), _stream_handle(stream))
    if st != OK:
        err = LscompError(st, "lscomp_compress_u16")
        err.required = int(n.value)
        raise err
    return out, int(n.value)


def compress_async(vol, dims, bins, pooling, out, d_size, stream=None) -> None:
    """Stream-ordered compress; the stream size lands in ``d_size`` (uint64/int64 device tensor, 1 element)."""
    _check_volume(vol, dims)
    st = _lib.lscomp_compress_u16_async(ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                        out.numel(), ctypes.c_void_p(d_size.data_ptr()), _dims(dims),
                                        _bins(bins), ctypes.c_float(pooling), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_compress_u16_async")


def compress_noalloc_async(vol, dims, bins, pooling, out, d_size, stream=None) -> None:
    """compress_async through the entry point that allocates nothing (a fixed
    4 MiB device array per GPU holds its scratch; calls that overlap on one GPU
    race on it)."""
    _check_volume(vol, dims)
    st = _lib.lscomp_compress_u16_noalloc_async(ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                                out.numel(), ctypes.c_void_p(d_size.data_ptr()), _dims(dims),
                                                _bins(bins), ctypes.c_float(pooling), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_compress_u16_noalloc_async")


def decompress(stream_buf, nbytes: int, dims, bins=LOSSLESS_BINS, out=None, stream=None):
    """Decompress ``nbytes`` of a device stream into a (x, y, z) int16-typed volume (uint16 bits)."""
    import torch
    x, y, z = (int(v) for v in dims)
    if out is None:
        out = torch.empty((x, y, z), dtype=torch.int16, device=stream_buf.device)
    _check_volume(out, dims)
    st = _lib.lscomp_decompress_u16(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream_buf.data_ptr()),
                                    int(nbytes), _dims(dims), _bins(bins), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_decompress_u16")
    return out


def decompress_async(stream_buf, nbytes: int, dims, bins, out, stream=None) -> None:
    _check_volume(out, dims)
    st = _lib.lscomp_decompress_u16_async(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream_buf.data_ptr()),
                                          int(nbytes), _dims(dims), _bins(bins), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_decompress_u16_async")


# ---- uint32 volumes (int32-typed tensors carry the uint32 bits) -------------

def max_compressed_bytes_u32(dims) -> int:
    return int(_lib.lscomp_max_compressed_bytes_u32(_dims(dims)))


def compress_u32(vol, dims, bins=LOSSLESS_BINS, pooling: float = 1.0, out=None, stream=None):
    """Compress a device uint32 volume (4-byte tensor); returns (stream tensor uint8, stream bytes).
    Raises LscompError (ERR_RANGE when a block would need rate 32)."""
    import torch
    _check_volume(vol, dims, 4)
    if out is None:
        out = torch.empty(max_compressed_bytes_u32(dims), dtype=torch.uint8, device=vol.device)
    n = ctypes.c_size_t(0)
    st = _lib.lscomp_compress_u32(ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                  out.numel(), ctypes.byref(n), _dims(dims), _bins(bins),
                                  ctypes.c_float(pooling), _stream_handle(stream))
    if st != OK:
        err = LscompError(st, "lscomp_compress_u32")
        err.required = int(n.value)
        raise err
    return out, int(n.value)


def compress_u32_async(vol, dims, bins, pooling, out, d_size, stream=None) -> None:
    _check_volume(vol, dims, 4)
    st = _lib.lscomp_compress_u32_async(ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                        out.numel(), ctypes.c_void_p(d_size.data_ptr()), _dims(dims),
                                        _bins(bins), ctypes.c_float(pooling), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_compress_u32_async")


def decompress_u32(stream_buf, nbytes: int, dims, bins=LOSSLESS_BINS, out=None, stream=None):
    """Decompress into a (x, y, z) int32-typed volume (uint32 bits)."""
    import torch
    x, y, z = (int(v) for v in dims)
    if out is None:
        out = torch.empty((x, y, z), dtype=torch.int32, device=stream_buf.device)
    _check_volume(out, dims, 4)
    st = _lib.lscomp_decompress_u32(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream_buf.data_ptr()),
                                    int(nbytes), _dims(dims), _bins(bins), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_decompress_u32")
    return out


def decompress_u32_async(stream_buf, nbytes: int, dims, bins, out, stream=None) -> None:
    _check_volume(out, dims, 4)
    st = _lib.lscomp_decompress_u32_async(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream_buf.data_ptr()),
                                          int(nbytes), _dims(dims), _bins(bins), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_decompress_u32_async")


# ---- self-describing container (host memory) --------------------------------

def container_pack(entries) -> bytes:
    """Pack streams with their parameters into one container.

    ``entries``: dicts with ``stream`` (bytes-like, or a uint8 tensor whose first
    ``stream_bytes`` bytes are the stream), ``stream_bytes`` (optional for
    bytes-like), ``dims``, ``elem_bytes`` (2 or 4), ``bins``, ``pooling`` and
    optional ``tag``.  Device tensors are copied to the host."""
    import numpy as np
    n = len(entries)
    items = (ContainerItem * max(n, 1))()
    bufs = []
    for i, e in enumerate(entries):
        s = e["stream"]
        nbytes = int(e.get("stream_bytes", len(s) if isinstance(s, (bytes, bytearray, memoryview)) else 0))
        if hasattr(s, "cpu"):
            s = s[:nbytes].cpu().numpy()
        arr = np.ascontiguousarray(np.frombuffer(bytes(s[:nbytes]), np.uint8) if not isinstance(s, np.ndarray)
                                   else s[:nbytes], dtype=np.uint8)
        bufs.append(arr)
        x, y, z = (int(v) for v in e["dims"])
        b = [int(v) for v in e["bins"]]
        items[i] = ContainerItem(x, y, z, int(e["elem_bytes"]), b[0], b[1], b[2], b[3], float(e["pooling"]), 0, 0,
                                 arr.size, int(e.get("tag", i)))
    ptrs = (ctypes.c_void_p * max(n, 1))(*[a.ctypes.data for a in bufs])
    total = int(_lib.lscomp_container_bytes(items, n))
    out = np.zeros(max(total, 1), np.uint8)
    used = ctypes.c_size_t(0)
    st = _lib.lscomp_container_pack(items, n, ptrs, out.ctypes.data_as(ctypes.c_void_p), out.size,
                                    ctypes.byref(used))
    if st != OK:
        raise LscompError(st, "lscomp_container_pack")
    return out[: used.value].tobytes()


def container_parse(buf):
    """Index of a container: list of dicts (dims, elem_bytes, bins, pooling, tag,
    offset, stream_bytes, stream = memoryview into ``buf``)."""
    import numpy as np
    arr = np.frombuffer(bytes(buf), np.uint8)
    n = ctypes.c_uint32(0)
    st = _lib.lscomp_container_parse(arr.ctypes.data_as(ctypes.c_void_p), arr.size, None, 0, ctypes.byref(n))
    if st not in (OK, ERR_CAPACITY):
        raise LscompError(st, "lscomp_container_parse")
    items = (ContainerItem * max(n.value, 1))()
    st = _lib.lscomp_container_parse(arr.ctypes.data_as(ctypes.c_void_p), arr.size, items, n.value,
                                     ctypes.byref(n))
    if st != OK:
        raise LscompError(st, "lscomp_container_parse")
    out = []
    mv = memoryview(arr)
    for i in range(n.value):
        it = items[i]
        out.append({"dims": (it.x, it.y, it.z), "elem_bytes": it.elem_bytes, "bins": (it.b0, it.b1, it.b2, it.b3),
                    "pooling": it.pooling_th, "tag": it.tag, "offset": it.offset, "stream_bytes": it.stream_bytes,
                    "stream": mv[it.offset: it.offset + it.stream_bytes]})
    return out


def container_decompress(buf, device="cuda"):
    """Decompress every item of a container on the GPU; returns a list of
    device volumes (int16 tensors for uint16 items, int32 for uint32 items)."""
    import numpy as np
    import torch
    vols = []
    for e in container_parse(buf):
        s = torch.from_numpy(np.frombuffer(e["stream"], np.uint8).copy()).to(device)
        if e["elem_bytes"] == 2:
            vols.append(decompress(s, e["stream_bytes"], e["dims"], e["bins"]))
        else:
            vols.append(decompress_u32(s, e["stream_bytes"], e["dims"], e["bins"]))
    return vols


def stream_status(stream=None) -> int:
    return int(_lib.lscomp_stream_status(_stream_handle(stream)))


def reserve(dims, stream=None) -> None:
    st = _lib.lscomp_reserve_u16(_dims(dims), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_reserve_u16")


def release_all() -> None:
    _lib.lscomp_release_all()


COMPRESS, DECOMPRESS = 0, 1


def set_launch_shape(direction: int, tile: int = 0, wg_per_cu: int = 0) -> None:
    """Blocks per workgroup (128/256/512, 0 = default) and workgroups per CU (0 = occupancy limit)."""
    st = _lib.lscomp_set_launch_shape(direction, tile, wg_per_cu)
    if st != OK:
        raise LscompError(st, "lscomp_set_launch_shape")


def get_launch_shape(direction: int):
    t, w = ctypes.c_uint32(0), ctypes.c_uint32(0)
    st = _lib.lscomp_get_launch_shape(direction, ctypes.byref(t), ctypes.byref(w))
    if st != OK:
        raise LscompError(st, "lscomp_get_launch_shape")
    return int(t.value), int(w.value)


def set_launch_shape_ex(direction: int, tile: int = 0, wg_per_cu: int = 0, lb_window: int = 0,
                        counters: int = 0) -> None:
    """The full launch shape of a direction (0 = default for each field)."""
    sh = LaunchShape(tile, wg_per_cu, lb_window, counters)
    st = _lib.lscomp_set_launch_shape_ex(direction, ctypes.byref(sh))
    if st != OK:
        raise LscompError(st, "lscomp_set_launch_shape_ex")


def get_launch_shape_ex(direction: int) -> dict:
    sh = LaunchShape()
    st = _lib.lscomp_get_launch_shape_ex(direction, ctypes.byref(sh))
    if st != OK:
        raise LscompError(st, "lscomp_get_launch_shape_ex")
    return {"tile": sh.tile, "wg_per_cu": sh.wg_per_cu, "lb_window": sh.lb_window, "counters": sh.counters}


def autotune(vol, dims, bins=LOSSLESS_BINS, pooling: float = 1.0, reps: int = 3, stream=None):
    """Time every launch shape on this volume, apply the fastest per direction.

    Returns the measured table (list of dicts); empty when the shapes came from
    the tuner's cache for this size class and mode.
    """
    import torch
    _check_volume(vol, dims)
    work = torch.empty(max_compressed_bytes(dims), dtype=torch.uint8, device=vol.device)
    dec = torch.empty_like(vol)
    table = (TuneEntry * 128)()
    n = ctypes.c_uint32(128)
    st = _lib.lscomp_autotune_u16(ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(work.data_ptr()), work.numel(),
                                  ctypes.c_void_p(dec.data_ptr()), _dims(dims), _bins(bins), ctypes.c_float(pooling),
                                  reps, _stream_handle(stream), table, ctypes.byref(n))
    if st != OK:
        raise LscompError(st, "lscomp_autotune_u16")
    return [t.as_dict() for t in table[: n.value]]


def error_stats(orig, dec, stream=None) -> dict:
    """GPU error metrics between two device volumes of 2-byte elements (uint16 bits)."""
    if orig.numel() != dec.numel() or orig.element_size() != 2 or dec.element_size() != 2:
        raise ValueError("orig and dec must be 2-byte tensors of equal size")
    out = ErrorStats()
    st = _lib.lscomp_error_stats_u16(ctypes.c_void_p(orig.data_ptr()), ctypes.c_void_p(dec.data_ptr()),
                                     orig.numel(), ctypes.byref(out), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_error_stats_u16")
    return {name: getattr(out, name) for name, _ in ErrorStats._fields_ if name != "reserved"}


def ssim(orig, dec, dims, stream=None) -> float:
    """Mean SSIM of two device volumes of 2-byte elements (uint16 bits) of shape
    dims = (x, y, z), z fastest: the reference's computeSSIM
    (lsCOMP/src/cuLSZ_utility.cu:757-886) on the GPU."""
    x, y, z = map(int, dims)
    if orig.numel() != x * y * z or dec.numel() != x * y * z or orig.element_size() != 2 \
            or dec.element_size() != 2:
        raise ValueError("orig and dec must be 2-byte tensors of dims' size")
    out = ctypes.c_double(0.0)
    st = _lib.lscomp_ssim_u16(ctypes.c_void_p(orig.data_ptr()), ctypes.c_void_p(dec.data_ptr()), Dim3(x, y, z),
                              ctypes.byref(out), _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_ssim_u16")
    return out.value


def copy_bytes(dst, src, stream=None) -> None:
    """Plain device copy of src into dst (same byte size, 16-byte aligned): the
    reference copy for bandwidth and HBM counter calibration."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("dst and src must have the same size in bytes")
    st = _lib.lscomp_copy_bytes(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                                _stream_handle(stream))
    if st != OK:
        raise LscompError(st, "lscomp_copy_bytes")


DECODE_LOOKBACK, DECODE_PHASE, DECODE_PREPASS = 0, 1, 2


def set_decode_prepass(mode) -> int:
    """Single-volume decode: tile offsets from a first phase of the decode launch
    (DECODE_PHASE, default; True means this too), a separate pre-pass kernel
    (DECODE_PREPASS) or a look-back inside the decode (DECODE_LOOKBACK; False
    means this too).  Returns the old mode."""
    m = DECODE_PHASE if mode is True else DECODE_LOOKBACK if mode is False else int(mode)
    old = int(_lib.lscomp_set_decode_prepass(m))
    if old < 0:
        raise LscompError(-old, "lscomp_set_decode_prepass")
    return old


def version() -> int:
    return int(_lib.lscomp_version())


class Plan:
    """Batched form: many independent items (own streams) per persistent launch.

    ``items`` for ``set_items`` are tuples (volume tensor, dims, bins, pooling,
    stream tensor, stream_bytes): for compression the stream tensor is the output
    buffer and stream_bytes its capacity; for decompression the input stream and
    its length.  Volumes must stay alive until the launch finished.
    """

    def __init__(self, direction: int, max_items: int, max_blocks: int):
        h = ctypes.c_void_p()
        st = _lib.lscomp_plan_create(direction, max_items, max_blocks, ctypes.byref(h))
        if st != OK:
            raise LscompError(st, "lscomp_plan_create")
        self.h, self.direction, self.n = h, direction, 0

    def set_items(self, items, stream=None) -> None:
        arr = (ItemDesc * len(items))()
        for i, (vol, dims, bins, pooling, sbuf, sbytes) in enumerate(items):
            _check_volume(vol, dims)
            arr[i] = ItemDesc(_dims(dims), _bins(bins), float(pooling), vol.data_ptr(), sbuf.data_ptr(), int(sbytes))
        st = _lib.lscomp_plan_set_items_u16(self.h, arr, len(items), _stream_handle(stream))
        if st != OK:
            raise LscompError(st, "lscomp_plan_set_items_u16")
        self.n = len(items)

    def set_shape(self, tile: int = 0, wg_per_cu: int = 0, lb_window: int = 0, counters: int = 0,
                  stream=None) -> None:
        sh = LaunchShape(tile, wg_per_cu, lb_window, counters)
        st = _lib.lscomp_plan_set_shape(self.h, ctypes.byref(sh), _stream_handle(stream))
        if st != OK:
            raise LscompError(st, "lscomp_plan_set_shape")

    def get_shape(self) -> dict:
        sh = LaunchShape()
        st = _lib.lscomp_plan_get_shape(self.h, ctypes.byref(sh))
        if st != OK:
            raise LscompError(st, "lscomp_plan_get_shape")
        return {"tile": sh.tile, "wg_per_cu": sh.wg_per_cu, "lb_window": sh.lb_window, "counters": sh.counters}

    def autotune(self, reps: int = 3, stream=None):
        """Tune the plan's launch shape on its current batch; returns the measured entries."""
        table = (TuneEntry * 128)()
        n = ctypes.c_uint32(128)
        st = _lib.lscomp_plan_autotune(self.h, reps, _stream_handle(stream), table, ctypes.byref(n))
        if st != OK:
            raise LscompError(st, "lscomp_plan_autotune")
        return [t.as_dict() for t in table[: n.value]]

    def launch(self, stream=None, graph: bool = False) -> None:
        f = _lib.lscomp_plan_launch_graph if graph else _lib.lscomp_plan_launch
        st = f(self.h, _stream_handle(stream))
        if st != OK:
            raise LscompError(st, "lscomp_plan_launch")

    def sizes(self, stream=None):
        out = (ctypes.c_uint64 * max(self.n, 1))()
        st = _lib.lscomp_plan_sizes(self.h, out, _stream_handle(stream))
        if st != OK:
            raise LscompError(st, "lscomp_plan_sizes")
        return [int(v) for v in out[: self.n]]

    def device_sizes_ptr(self) -> int:
        return int(_lib.lscomp_plan_device_sizes(self.h) or 0)

    def status(self, stream=None) -> int:
        return int(_lib.lscomp_plan_status(self.h, _stream_handle(stream)))

    def close(self) -> None:
        if self.h:
            _lib.lscomp_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Served form: continuous batching of frame-compression requests (see include/lscomp.h).

    ``Server(slots, max_items, max_blocks, budget_bytes)`` is the host engine
    (host-scheduled batch launches); ``Server.device(max_blocks, budget_bytes,
    idle_us)`` the device-resident request loop."""

    def __init__(self, slots: int = 0, max_items: int = 0, max_blocks: int = 1 << 22, budget_bytes: int = 1 << 30,
                 device: bool = False, idle_us: int = 2000):
        h = ctypes.c_void_p()
        if device:
            st = _lib.lscomp_server_create_device(max_blocks, budget_bytes, idle_us, ctypes.byref(h))
        else:
            st = _lib.lscomp_server_create(slots, max_items, max_blocks, budget_bytes, ctypes.byref(h))
        if st != OK:
            raise LscompError(st, "lscomp_server_create_device" if device else "lscomp_server_create")
        self.h = h

    @classmethod
    def device(cls, max_blocks: int, budget_bytes: int, idle_us: int = 2000) -> "Server":
        return cls(max_blocks=max_blocks, budget_bytes=budget_bytes, device=True, idle_us=idle_us)

    def stats(self) -> dict:
        out = ServerStats()
        st = _lib.lscomp_server_stats(self.h, ctypes.byref(out))
        if st != OK:
            raise LscompError(st, "lscomp_server_stats")
        return {"submitted": out.submitted, "completed": out.completed, "launches": out.launches,
                "device_loop": bool(out.device_loop)}

    @staticmethod
    def request(frame, dims, bins=LOSSLESS_BINS, pooling: float = 1.0, tag: int = 0) -> Request:
        _check_volume(frame, dims)
        return Request(_dims(dims), _bins(bins), float(pooling), frame.data_ptr(), int(tag))

    def submit(self, req: Request) -> int:
        rid = ctypes.c_uint64(0)
        st = _lib.lscomp_server_submit_u16(self.h, ctypes.byref(req), ctypes.byref(rid))
        if st != OK:
            raise LscompError(st, "lscomp_server_submit_u16")
        return int(rid.value)

    def poll(self, max_n: int = 256):
        buf = (Completion * max_n)()
        n = ctypes.c_uint32(0)
        st = _lib.lscomp_server_poll(self.h, buf, max_n, ctypes.byref(n))
        if st != OK:
            raise LscompError(st, "lscomp_server_poll")
        return [buf[i] for i in range(n.value)]

    def release(self, rid: int) -> None:
        st = _lib.lscomp_server_release(self.h, rid)
        if st != OK:
            raise LscompError(st, "lscomp_server_release")

    def drain(self) -> int:
        return int(_lib.lscomp_server_drain(self.h))

    def replay(self, reqs, arrivals, hold: bool = False):
        """Submit reqs at the given arrival times (s); returns Completion per request, in order.
        With ``hold`` the outputs stay reserved (release them with ``release``)."""
        n = len(reqs)
        arr = (Request * n)(*reqs)
        times = (ctypes.c_double * n)(*[float(t) for t in arrivals])
        out = (Completion * n)()
        st = _lib.lscomp_server_replay_u16(self.h, arr, times, n, out, REPLAY_HOLD if hold else 0)
        if st != OK:
            raise LscompError(st, "lscomp_server_replay_u16")
        return [out[i] for i in range(n)]

    def close(self) -> None:
        if self.h:
            _lib.lscomp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
