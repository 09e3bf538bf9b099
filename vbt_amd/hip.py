"""Python host side of libvbt_hip.so (ctypes over the C-ABI in include/vbt_hip.h).

Device memory and streams come from the library itself (vbt_device_alloc,
vbt_stream_create), so no other GPU runtime is needed in the process.  Every
call raises VbtError with the library's message on failure; the library is
required -- there is no fallback decode.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvbt_hip.so")

VBT_OK = 0
CACHE_PACKED, CACHE_STRIDE576 = 0, 1
SIGN_NONE, SIGN_CLAMP_ZERO, SIGN_BACKGROUND = 0, 1, 2

c_void_p = ctypes.c_void_p
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int


class VbtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[vbt {code}] {msg}")
        self.code = code
        self.msg = msg


class FileHeader(ctypes.Structure):
    _fields_ = [(n, c_u32) for n in ("version", "width", "height", "depth", "frames", "leafSize",
                                       "coarseResolution", "maxCoarseKeep", "leafCount", "profileType",
                                       "coarseAcScaleCount")]


class AssetInfo(ctypes.Structure):
    _fields_ = [("header", FileHeader), ("activeLeaves", c_u32), ("badLeaves", c_u32),
                ("controlsPerFrame", c_u64), ("payloadBytes", c_u64), ("deviceBytes", c_u64),
                ("mode_counts", c_u64 * 4), ("fast_leaves", c_u64)]


class Compose(ctypes.Structure):
    _fields_ = [("signRule", c_int), ("background", ctypes.c_float), ("shellEmptyScale", ctypes.c_float),
                ("shellEmptyMax", ctypes.c_float)]


class MarchParams(ctypes.Structure):
    """vbt_march_params (the sampler's push constants)."""
    _fields_ = [("stepCount", c_u32), ("frameIndex", ctypes.c_int32), ("bboxMin", ctypes.c_float * 3),
                ("renderMode", c_u32), ("voxelSize", ctypes.c_float), ("backgroundValue", ctypes.c_float),
                ("surfaceIso", ctypes.c_float), ("surfaceEpsilonVoxels", ctypes.c_float),
                ("surfaceNormalStep", ctypes.c_float), ("maxStepVoxels", ctypes.c_float), ("surfaceModel", c_u32),
                ("floorHeight", ctypes.c_float), ("shadowSteps", c_u32), ("upAxis", c_u32)]


def march_params(**kw) -> MarchParams:
    """Defaults: 128 steps, frame 0, origin at 0, density mode, unit voxels,
    background 0, iso 0, epsilon 0.5 voxel, normal step 0.5, max step 4,
    no surface model, floor at 0, 64 shadow steps, y up."""
    p = MarchParams(stepCount=128, frameIndex=0, renderMode=0, voxelSize=1.0, backgroundValue=0.0, surfaceIso=0.0,
                    surfaceEpsilonVoxels=0.5, surfaceNormalStep=0.5, maxStepVoxels=4.0, surfaceModel=0,
                    floorHeight=0.0, shadowSteps=64, upAxis=1)
    for k, v in kw.items():
        if k == "bboxMin":
            p.bboxMin[:] = [float(x) for x in v]
        elif hasattr(p, k):
            setattr(p, k, v)
        else:
            raise TypeError(f"unknown march parameter {k}")
    return p


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbmBudgetBytes", c_u64), ("arenaBytes", c_u64), ("stagingBytes", c_u64),
                ("maxInflight", c_u32), ("blocksPerCU", c_u32), ("queryItem", c_u32), ("cacheItem", c_u32),
                ("idleUs", c_u32), ("leaseMs", c_u32), ("frameBuildsInLoop", c_u32), ("frameIndex", c_u32)]


class RequestResult(ctypes.Structure):
    _fields_ = [("id", c_u64), ("status", ctypes.c_int32), ("count", c_u32), ("d_values", c_void_p),
                ("d_status", c_void_p), ("latencyUs", ctypes.c_double), ("launch", c_u64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, c_u64) for n in ("submitted", "completed", "rejected", "launches", "values", "requeued")] + \
               [(n, ctypes.c_double) for n in ("p50Us", "p99Us", "meanUs", "maxUs")] + \
               [(n, c_u64) for n in ("maxQueueDepth", "maxInflight", "hbmInUse", "hbmBudget", "arenaInUse")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CacheJob(ctypes.Structure):
    _fields_ = [("asset", c_void_p), ("t0", c_int), ("nt", c_int), ("layout", c_int), ("d_out", c_void_p)]


class BatchShape(ctypes.Structure):
    _fields_ = [(n, c_int) for n in ("tile", "ranges", "chunk", "order", "queues", "wgPerCU", "persistent",
                                       "heavyTile")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def compose(sign_rule=SIGN_NONE, background=0.0, shell_empty_scale=1.0, shell_empty_max=math.nan) -> Compose:
    return Compose(sign_rule, background, shell_empty_scale, shell_empty_max)


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.vbt_last_error.restype = ctypes.c_char_p
        L.vbt_version.restype = ctypes.c_char_p
        L.vbt_status_message.restype = ctypes.c_char_p
        L.vbt_status_message.argtypes = [ctypes.c_int32]
        L.vbt_device_alloc.argtypes = [ctypes.POINTER(c_void_p), ctypes.c_size_t]
        L.vbt_device_free.argtypes = [c_void_p]
        L.vbt_memcpy_h2d.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]
        L.vbt_memcpy_d2h.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]
        L.vbt_memset_d.argtypes = [c_void_p, c_int, ctypes.c_size_t, c_void_p]
        L.vbt_device_copy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]
        L.vbt_device_fill.argtypes = [c_void_p, c_u32, ctypes.c_size_t, c_void_p]
        L.vbt_selftest_weights.argtypes = [c_u32, ctypes.POINTER(c_u64), c_void_p]
        L.vbt_stream_create.argtypes = [ctypes.POINTER(c_void_p)]
        L.vbt_stream_destroy.argtypes = [c_void_p]
        L.vbt_stream_synchronize.argtypes = [c_void_p]
        L.vbt_event_create.argtypes = [ctypes.POINTER(c_void_p)]
        L.vbt_event_destroy.argtypes = [c_void_p]
        L.vbt_event_record.argtypes = [c_void_p, c_void_p]
        L.vbt_event_elapsed_ms.argtypes = [c_void_p, c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.vbt_device_count.argtypes = [ctypes.POINTER(c_int)]
        L.vbt_asset_open_memory.argtypes = [c_void_p, c_u64, c_void_p, ctypes.POINTER(c_void_p)]
        L.vbt_asset_open_file.argtypes = [ctypes.c_char_p, c_void_p, ctypes.POINTER(c_void_p)]
        L.vbt_asset_open_memory_ex.argtypes = [c_void_p, c_u64, c_void_p, ctypes.c_uint32, ctypes.POINTER(c_void_p)]
        L.vbt_asset_open_file_ex.argtypes = [ctypes.c_char_p, c_void_p, ctypes.c_uint32, ctypes.POINTER(c_void_p)]
        L.vbt_asset_close.argtypes = [c_void_p]
        L.vbt_asset_close.restype = None
        L.vbt_asset_info_get.argtypes = [c_void_p, ctypes.POINTER(AssetInfo)]
        L.vbt_asset_page_table.argtypes = [c_void_p, c_void_p, c_void_p]
        L.vbt_asset_cache_offsets.argtypes = [c_void_p, c_void_p, c_void_p]
        L.vbt_asset_leaf_status.argtypes = [c_void_p, c_void_p, c_void_p]
        L.vbt_decode_frame_cache.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]
        L.vbt_decode_frame_dense.argtypes = [c_void_p, c_int, ctypes.POINTER(Compose), c_void_p, c_void_p]
        L.vbt_query_points.argtypes = [c_void_p, c_void_p, c_u32, ctypes.POINTER(Compose), c_void_p, c_void_p,
                                       c_void_p]
        L.vbt_query_points_plain.argtypes = L.vbt_query_points.argtypes
        L.vbt_sample_points.argtypes = [c_void_p, c_void_p, ctypes.POINTER(MarchParams), c_void_p, c_u32, c_int,
                                        c_void_p, c_void_p]
        L.vbt_march_rays.argtypes = [c_void_p, c_void_p, ctypes.POINTER(MarchParams), c_void_p, c_u32, c_void_p,
                                     c_void_p]
        L.vbt_cycles_blob_bytes.argtypes = [c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        L.vbt_asset_build_frame_index.argtypes = [c_void_p, ctypes.POINTER(ctypes.c_uint64), c_void_p]
        L.vbt_asset_drop_frame_index.argtypes = [c_void_p]
        L.vbt_cycles_pack_blob.argtypes = [c_void_p, c_int, c_void_p, ctypes.c_uint64, c_void_p]
        L.vbt_cycles_sample_density.argtypes = [c_void_p, ctypes.c_uint64, c_int, c_void_p, c_u32, c_void_p,
                                                c_void_p]
        L.vbt_decode_control_values.argtypes = [c_void_p, c_u32, ctypes.c_uint16, ctypes.c_uint8, ctypes.c_uint8,
                                                c_u32, c_u32, c_int, c_int, c_void_p, c_void_p, c_u32, c_void_p,
                                                c_void_p, c_void_p]
        L.vbt_decode_leaf_control_values.argtypes = [c_void_p, c_u32, c_int, c_int, c_int, c_void_p, c_void_p, c_u32,
                                                     c_void_p, c_void_p, c_void_p]
        L.vbt_decode_shell_voxel_values.argtypes = [c_void_p, c_u32, c_int, c_int, c_void_p, c_void_p, c_u32,
                                                    c_void_p, c_void_p, c_void_p]
        L.vbt_validate_file.argtypes = [c_void_p, c_u64, ctypes.POINTER(FileHeader)]
        L.vbt_cache_tile_count.restype = c_int
        L.vbt_cache_tile_shape.argtypes = [c_int, ctypes.POINTER(c_u32), ctypes.POINTER(c_u32)]
        L.vbt_asset_set_cache_tile.argtypes = [c_void_p, c_int]
        L.vbt_asset_get_cache_tile.argtypes = [c_void_p, ctypes.POINTER(c_int)]
        L.vbt_asset_autotune.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(c_int), c_void_p]
        L.vbt_tuning_key.argtypes = [c_void_p, c_int, ctypes.c_char_p, ctypes.c_size_t, c_void_p]
        L.vbt_asset_cache_ceiling.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]
        L.vbt_asset_autotune_table.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(ctypes.c_float),
                                               ctypes.POINTER(c_int), ctypes.POINTER(c_int), c_void_p]
        L.vbt_tuning_load.argtypes = [ctypes.c_char_p]
        L.vbt_tuning_save.argtypes = [ctypes.c_char_p]
        L.vbt_tuning_count.restype = c_int
        L.vbt_tuning_count.argtypes = []
        L.vbt_tuning_clear.argtypes = []
        L.vbt_server_default_config.argtypes = [ctypes.POINTER(ServerConfig)]
        L.vbt_server_default_config.restype = None
        L.vbt_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(c_void_p)]
        L.vbt_server_destroy.argtypes = [c_void_p]
        L.vbt_server_register.argtypes = [c_void_p, c_void_p, ctypes.POINTER(c_int)]
        L.vbt_server_submit_query.argtypes = [c_void_p, c_int, c_void_p, c_u32, ctypes.POINTER(Compose),
                                              ctypes.POINTER(c_u64)]
        L.vbt_server_submit_cache.argtypes = [c_void_p, c_int, c_int, ctypes.POINTER(c_u64)]
        L.vbt_server_poll.argtypes = [c_void_p, c_u64, c_int, ctypes.POINTER(RequestResult)]
        L.vbt_server_fetch.argtypes = [c_void_p, c_u64, c_void_p, c_void_p]
        L.vbt_server_release.argtypes = [c_void_p, c_u64]
        L.vbt_server_stats_get.argtypes = [c_void_p, c_int, ctypes.POINTER(ServerStats)]
        L.vbt_server_reset_stats.argtypes = [c_void_p]
        L.vbt_batch_default_shape.argtypes = [ctypes.POINTER(BatchShape)]
        L.vbt_batch_default_shape.restype = None
        L.vbt_cache_batch_create.argtypes = [ctypes.POINTER(CacheJob), c_int, ctypes.POINTER(BatchShape),
                                             ctypes.POINTER(c_void_p)]
        L.vbt_cache_batch_destroy.argtypes = [c_void_p]
        L.vbt_cache_batch_launch.argtypes = [c_void_p, c_void_p]
        L.vbt_cache_batch_launch_direct.argtypes = [c_void_p, c_void_p]
        L.vbt_cache_batch_set_shape.argtypes = [c_void_p, ctypes.POINTER(BatchShape)]
        L.vbt_cache_batch_get_shape.argtypes = [c_void_p, ctypes.POINTER(BatchShape)]
        L.vbt_cache_batch_plan_info.argtypes = [c_void_p, ctypes.POINTER(c_u64), ctypes.POINTER(c_u32)]
        L.vbt_cache_batch_autotune.argtypes = [c_void_p, c_int, ctypes.POINTER(BatchShape),
                                               ctypes.POINTER(ctypes.c_float), c_void_p]
        L.vbt_cache_batch_trial_count.argtypes = [c_void_p]
        L.vbt_cache_batch_trial.argtypes = [c_void_p, c_int, ctypes.POINTER(BatchShape),
                                            ctypes.POINTER(ctypes.c_float)]
        L.vbt_batch_noalloc_default_shape.argtypes = [ctypes.POINTER(BatchShape)]
        L.vbt_batch_noalloc_default_shape.restype = None
        L.vbt_cache_batch_run_noalloc.argtypes = [ctypes.POINTER(CacheJob), c_int, ctypes.POINTER(BatchShape),
                                                  c_void_p]
        L.vbt_parse_query_pattern.argtypes = [ctypes.c_char_p, ctypes.POINTER(c_int)]
        L.vbt_generate_queries.argtypes = [ctypes.POINTER(FileHeader), c_int, c_u32, c_u32, c_void_p]
        L.vbt_sort_queries_by_leaf.argtypes = [ctypes.POINTER(FileHeader), c_void_p, c_u32]
        L.vbt_half_bits_to_float.restype = ctypes.c_float
        L.vbt_half_bits_to_float.argtypes = [ctypes.c_uint16]
        L.vbt_float_to_half_bits.restype = ctypes.c_uint16
        L.vbt_float_to_half_bits.argtypes = [ctypes.c_float]
        _LIB = L
    return _LIB


def check(rc):
    if rc != VBT_OK:
        raise VbtError(rc, lib().vbt_last_error().decode())


def device_count() -> int:
    n = c_int(0)
    lib().vbt_device_count(ctypes.byref(n))
    return n.value


class Stream:
    def __init__(self):
        self.h = c_void_p()
        check(lib().vbt_stream_create(ctypes.byref(self.h)))

    def sync(self):
        check(lib().vbt_stream_synchronize(self.h))

    def close(self):
        if self.h:
            lib().vbt_stream_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation with a numpy dtype/shape (host copies on demand)."""

    def __init__(self, shape, dtype, stream: Stream | None = None):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = c_void_p()
        self.stream = stream
        check(lib().vbt_device_alloc(ctypes.byref(self.ptr), max(self.nbytes, 4)))

    @classmethod
    def from_host(cls, arr: np.ndarray, stream: Stream | None = None):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype, stream)
        check(lib().vbt_memcpy_h2d(d.ptr, arr.ctypes.data, arr.nbytes, stream.h if stream else None))
        # complete before any stream reads it (assets run on non-blocking streams)
        if stream:
            stream.sync()
        else:
            check(lib().vbt_stream_synchronize(None))
        return d

    def host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().vbt_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, self.stream.h if self.stream else None))
        return out

    def host_slice(self, start: int, count: int) -> np.ndarray:
        """Elements [start, start + count) of the flattened array."""
        n = int(np.prod(self.shape))
        if start < 0 or count < 0 or start + count > n:
            raise ValueError("slice outside the array")
        out = np.empty(count, self.dtype)
        src = c_void_p(self.ptr.value + start * self.dtype.itemsize)
        check(lib().vbt_memcpy_d2h(out.ctypes.data, src, count * self.dtype.itemsize,
                                   self.stream.h if self.stream else None))
        if self.stream:
            self.stream.sync()
        return out

    def fill_bytes(self, value: int):
        """Set every byte; an array without a stream is filled before this returns
        (work on non-blocking streams is not ordered after the null stream)."""
        check(lib().vbt_memset_d(self.ptr, value, self.nbytes, self.stream.h if self.stream else None))
        if not self.stream:
            check(lib().vbt_stream_synchronize(None))

    def free(self):
        if self.ptr:
            lib().vbt_device_free(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


CYCLES_KERNEL, CYCLES_WORLD, CYCLES_PLAIN = 1, 2, 4


def cycles_sample_into(blob: DeviceArray, xyz: DeviceArray, n: int, out: DeviceArray, flags: int = 0,
                       stream: Stream | None = None):
    check(lib().vbt_cycles_sample_density(blob.ptr, blob.nbytes, int(flags), xyz.ptr, n, out.ptr,
                                          (stream or blob.stream).h if (stream or blob.stream) else None))


def cycles_sample(blob: DeviceArray, xyz: np.ndarray, flags: int = 0) -> np.ndarray:
    """Density at n points (n x 3) read from a device Cycles blob:
    the standalone sampler, or with CYCLES_KERNEL the patched Cycles kernel's."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    n = len(xyz)
    if n == 0:
        return np.zeros(0, np.float32)
    d = DeviceArray.from_host(xyz, blob.stream)
    out = DeviceArray(n, np.float32, blob.stream)
    cycles_sample_into(blob, d, n, out, flags)
    return out.host()


class Event:
    def __init__(self):
        self.h = c_void_p()
        check(lib().vbt_event_create(ctypes.byref(self.h)))

    def record(self, stream: Stream):
        check(lib().vbt_event_record(self.h, stream.h))

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        check(lib().vbt_event_elapsed_ms(self.h, later.h, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                lib().vbt_event_destroy(self.h)
        except Exception:
            pass


OPEN_SCI_DENSE_BASIS = 1  # VBT_OPEN_SCI_DENSE_BASIS


class Asset:
    """A VBTPACK4 asset resident in HBM with its device page table.

    dense_basis: scientific dense subtypes 2 (tile) and 3 (patch) decode as
    the reference encoder reconstructs them instead of as its file decoder
    reads them (VBT_OPEN_SCI_DENSE_BASIS)."""

    def __init__(self, data: bytes, stream: Stream | None = None, dense_basis: bool = False):
        self.stream = stream or Stream()
        self._buf = np.frombuffer(bytearray(data), dtype=np.uint8)
        self.h = c_void_p()
        flags = OPEN_SCI_DENSE_BASIS if dense_basis else 0
        check(lib().vbt_asset_open_memory_ex(self._buf.ctypes.data, len(data), self.stream.h, flags,
                                             ctypes.byref(self.h)))
        self._buf = None
        self.info = AssetInfo()
        check(lib().vbt_asset_info_get(self.h, ctypes.byref(self.info)))

    @classmethod
    def open_file(cls, path: str, stream: Stream | None = None, dense_basis: bool = False):
        self = cls.__new__(cls)
        self.stream = stream or Stream()
        self.h = c_void_p()
        flags = OPEN_SCI_DENSE_BASIS if dense_basis else 0
        check(lib().vbt_asset_open_file_ex(path.encode(), self.stream.h, flags, ctypes.byref(self.h)))
        self.info = AssetInfo()
        check(lib().vbt_asset_info_get(self.h, ctypes.byref(self.info)))
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().vbt_asset_close(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def header(self):
        return self.info.header

    @property
    def dims(self):
        h = self.info.header
        return h.width, h.height, h.depth, h.frames

    def page_table(self) -> np.ndarray:
        d = DeviceArray(self.header.leafCount, np.uint32, self.stream)
        check(lib().vbt_asset_page_table(self.h, d.ptr, self.stream.h))
        return d.host()

    def cache_offsets(self) -> np.ndarray:
        d = DeviceArray(self.info.activeLeaves + 1, np.uint64, self.stream)
        check(lib().vbt_asset_cache_offsets(self.h, d.ptr, self.stream.h))
        return d.host()

    def leaf_status(self) -> np.ndarray:
        d = DeviceArray(self.header.leafCount, np.uint32, self.stream)
        check(lib().vbt_asset_leaf_status(self.h, d.ptr, self.stream.h))
        return d.host()

    # ---- frame-cache kernel tile shape (autotuned per asset) ----
    @staticmethod
    def tile_shapes():
        L = lib()
        out = []
        for i in range(L.vbt_cache_tile_count()):
            b, w = c_u32(), c_u32()
            check(L.vbt_cache_tile_shape(i, ctypes.byref(b), ctypes.byref(w)))
            out.append((b.value, w.value))
        return out

    @property
    def cache_tile(self) -> int:
        t = c_int()
        check(lib().vbt_asset_get_cache_tile(self.h, ctypes.byref(t)))
        return t.value

    @cache_tile.setter
    def cache_tile(self, tile: int):
        check(lib().vbt_asset_set_cache_tile(self.h, int(tile)))

    def autotune(self, nt: int | None = None, reps: int = 3):
        """Time every tile shape on this asset, keep the fastest; returns
        (best tile, [ms per tile])."""
        n = lib().vbt_cache_tile_count()
        ms = (ctypes.c_float * n)()
        best = c_int()
        check(lib().vbt_asset_autotune(self.h, int(nt or self.dims[3]), reps, ms, ctypes.byref(best), self.stream.h))
        return best.value, [round(float(x), 4) for x in ms]

    def cache_ceiling_into(self, out: DeviceArray, nt: int, tile: int, reads: bool = True,
                           stream: Stream | None = None):
        """The frame-cache build's memory traffic without its decode (the
        layout's ceiling on this asset); `out` gets placeholder values."""
        check(lib().vbt_asset_cache_ceiling(self.h, int(nt), int(tile), 1 if reads else 0, out.ptr,
                                            (stream or self.stream).h))

    def tuning_key(self, nt: int | None = None) -> str:
        """This asset's key in the tuning table for nt frames per launch."""
        buf = ctypes.create_string_buffer(256)
        check(lib().vbt_tuning_key(self.h, int(nt or self.dims[3]), buf, len(buf), self.stream.h))
        return buf.value.decode()

    def autotune_table(self, nt: int | None = None, reps: int = 3, times: bool = False):
        """The tile from the tuning table when this asset's key is in it, else
        a timed sweep whose pick joins the table; returns (tile, measured), or
        (tile, measured, [ms per tile] or None) with times=True."""
        n = lib().vbt_cache_tile_count()
        ms = (ctypes.c_float * n)()
        best, measured = c_int(), c_int()
        check(lib().vbt_asset_autotune_table(self.h, int(nt or self.dims[3]), reps, ms, ctypes.byref(best),
                                             ctypes.byref(measured), self.stream.h))
        if not times:
            return best.value, bool(measured.value)
        return best.value, bool(measured.value), [round(float(x), 4) for x in ms] if measured.value else None

    def frame_cache_into(self, out: DeviceArray, t0: int, nt: int, layout=CACHE_PACKED, stream: Stream | None = None):
        check(lib().vbt_decode_frame_cache(self.h, t0, nt, layout, out.ptr, (stream or self.stream).h))

    def frame_cache(self, t0: int, nt: int = 1, layout=CACHE_PACKED) -> np.ndarray:
        n = self.info.controlsPerFrame if layout == CACHE_PACKED else self.info.activeLeaves * 576
        out = DeviceArray((nt, n), np.float32, self.stream)
        self.frame_cache_into(out, t0, nt, layout)
        return out.host()

    def dense_into(self, out: DeviceArray, t: int, comp: Compose | None = None):
        comp = comp or compose()
        check(lib().vbt_decode_frame_dense(self.h, t, ctypes.byref(comp), out.ptr, self.stream.h))

    def dense(self, t: int, comp: Compose | None = None) -> np.ndarray:
        W, H, D, _ = self.dims
        out = DeviceArray((D, H, W), np.float32, self.stream)
        self.dense_into(out, t, comp)
        return out.host()

    def query_into(self, q: DeviceArray, n: int, out: DeviceArray, comp: Compose | None = None,
                   status: DeviceArray | None = None, plain=False):
        comp = comp or compose()
        fn = lib().vbt_query_points_plain if plain else lib().vbt_query_points
        check(fn(self.h, q.ptr, n, ctypes.byref(comp), out.ptr, status.ptr if status else None, self.stream.h))

    def query(self, q: np.ndarray, comp: Compose | None = None, plain=False, with_status=False):
        q = np.ascontiguousarray(q, dtype=np.uint32).reshape(-1, 4)
        n = len(q)
        dq = DeviceArray.from_host(q, self.stream)
        out = DeviceArray(n, np.float32, self.stream)
        st = DeviceArray(n, np.int32, self.stream)
        self.query_into(dq, n, out, comp, st, plain)
        vals = out.host()
        return (vals, st.host()) if with_status else vals


    def sample_points(self, cache576: DeviceArray, params: MarchParams, xyz: np.ndarray,
                      preserve_interior: bool = False) -> np.ndarray:
        """The viewer's sampleVbtValue at world points (n x 3) from a one-frame
        stride-576 cache."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        n = len(xyz)
        if n == 0:
            return np.zeros(0, np.float32)
        d = DeviceArray.from_host(xyz, self.stream)
        out = DeviceArray(n, np.float32, self.stream)
        check(lib().vbt_sample_points(self.h, cache576.ptr, ctypes.byref(params), d.ptr, n, int(preserve_interior),
                                      out.ptr, self.stream.h))
        return out.host()

    def march_into(self, cache576: DeviceArray, params: MarchParams, rays: DeviceArray, n: int, out: DeviceArray,
                   stream: Stream | None = None):
        check(lib().vbt_march_rays(self.h, cache576.ptr, ctypes.byref(params), rays.ptr, n, out.ptr,
                                   (stream or self.stream).h))

    def build_frame_index(self) -> int:
        """Builds the frame index that one- and two-frame packed-cache builds
        then read; returns its size in bytes."""
        n = ctypes.c_uint64()
        check(lib().vbt_asset_build_frame_index(self.h, ctypes.byref(n), self.stream.h))
        return n.value

    def drop_frame_index(self):
        check(lib().vbt_asset_drop_frame_index(self.h))

    def cycles_blob(self, frame_index: int) -> DeviceArray:
        """The Cycles resident blob (header, offsets, payload) for one frame,
        packed on the device (cycles_native_vbt/vbt_cycles_blob.cpp:22-87)."""
        n = ctypes.c_uint64()
        check(lib().vbt_cycles_blob_bytes(self.h, ctypes.byref(n)))
        blob = DeviceArray(n.value, np.uint8, self.stream)
        check(lib().vbt_cycles_pack_blob(self.h, int(frame_index), blob.ptr, n.value, self.stream.h))
        return blob

    def march(self, cache576: DeviceArray, params: MarchParams, rays: np.ndarray) -> np.ndarray:
        """The viewer's ray march (n x 8 rays: origin, tmin, direction, tmax)
        -> n x 8 results (primary, auxiliary)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        if n == 0:
            return np.zeros((0, 8), np.float32)
        d = DeviceArray.from_host(rays, self.stream)
        out = DeviceArray((n, 8), np.float32, self.stream)
        self.march_into(cache576, params, d, n, out)
        return out.host()


def tuning_load(path: str):
    """Merge a tuning table file into the process's table."""
    check(lib().vbt_tuning_load(os.fsencode(path)))


def tuning_save(path: str):
    check(lib().vbt_tuning_save(os.fsencode(path)))


def tuning_count() -> int:
    return lib().vbt_tuning_count()


def tuning_clear():
    check(lib().vbt_tuning_clear())


def status_message(code: int) -> str:
    return lib().vbt_status_message(int(code)).decode()


def decode_control_values(stream_bytes: bytes, codec, control_count, bins, keyframe_count, total_frames, ci, t,
                          align=1, stream_limit=None):
    """Batched decodeRenderTemporalControlValue on the GPU (host arrays in/out)."""
    s = Stream()
    buf = np.frombuffer(bytes(stream_bytes) + b"\0\0\0\0", dtype=np.uint8)
    d_stream = DeviceArray.from_host(buf, s)
    ci = np.ascontiguousarray(ci, np.uint16)
    t = np.ascontiguousarray(t, np.int32)
    d_ci = DeviceArray.from_host(ci, s)
    d_t = DeviceArray.from_host(t, s)
    out = DeviceArray(len(ci), np.float32, s)
    st = DeviceArray(len(ci), np.int32, s)
    limit = len(stream_bytes) if stream_limit is None else stream_limit
    check(lib().vbt_decode_control_values(d_stream.ptr, codec, control_count, bins, 0, keyframe_count, limit, align,
                                          total_frames, d_ci.ptr, d_t.ptr, len(ci), out.ptr, st.ptr, s.h))
    return out.host(), st.host()


def decode_leaf_values(leaf_bytes: bytes, total_frames, idx, t, use_fine=False, shell=False):
    """Batched decodeRenderTemporalLeafControlValue / ...ShellVoxelValue on the GPU."""
    s = Stream()
    buf = np.frombuffer(bytes(leaf_bytes) + b"\0\0\0\0", dtype=np.uint8)
    d_leaf = DeviceArray.from_host(buf, s)
    idx = np.ascontiguousarray(idx, np.uint16)
    t = np.ascontiguousarray(t, np.int32)
    d_i = DeviceArray.from_host(idx, s)
    d_t = DeviceArray.from_host(t, s)
    out = DeviceArray(len(idx), np.float32, s)
    st = DeviceArray(len(idx), np.int32, s)
    if shell:
        check(lib().vbt_decode_shell_voxel_values(d_leaf.ptr, len(leaf_bytes), 1, total_frames, d_i.ptr, d_t.ptr,
                                                  len(idx), out.ptr, st.ptr, s.h))
    else:
        check(lib().vbt_decode_leaf_control_values(d_leaf.ptr, len(leaf_bytes), int(use_fine), 1, total_frames,
                                                   d_i.ptr, d_t.ptr, len(idx), out.ptr, st.ptr, s.h))
    return out.host(), st.host()


REQ_QUERY, REQ_CACHE = 0, 1


QUERY_PATTERNS = {"random": 0, "same-t": 1, "same-xyz": 2, "coherent": 3}


def generate_queries(header: FileHeader, pattern, count: int, seed: int, sort_by_leaf: bool = False) -> np.ndarray:
    """The reference's request shapes (generateQueries, then sortQueriesByLeaf
    when sort_by_leaf): a (count, 4) u32 array of (x, y, z, t)."""
    if isinstance(pattern, str):
        p = c_int()
        check(lib().vbt_parse_query_pattern(pattern.encode(), ctypes.byref(p)))
        pattern = p.value
    out = np.zeros((count, 4), np.uint32)
    check(lib().vbt_generate_queries(ctypes.byref(header), int(pattern), count, seed, out.ctypes.data))
    if sort_by_leaf:
        check(lib().vbt_sort_queries_by_leaf(ctypes.byref(header), out.ctypes.data, count))
    return out


def default_batch_shape() -> BatchShape:
    s = BatchShape()
    lib().vbt_batch_default_shape(ctypes.byref(s))
    return s


def noalloc_batch_shape() -> BatchShape:
    s = BatchShape()
    lib().vbt_batch_noalloc_default_shape(ctypes.byref(s))
    return s


def cache_batch_run_noalloc(jobs, shape: BatchShape | None = None, stream: Stream | None = None):
    """The batch's frame-cache builds in one call that allocates no device
    memory (vbt_cache_batch_run_noalloc).  jobs: (asset, t0, nt, layout,
    DeviceArray out)."""
    jobs = list(jobs)
    arr = (CacheJob * max(1, len(jobs)))()
    for i, (asset, t0, nt, layout, out) in enumerate(jobs):
        arr[i] = CacheJob(asset.h, int(t0), int(nt), int(layout), out.ptr)
    check(lib().vbt_cache_batch_run_noalloc(arr, len(jobs), ctypes.byref(shape) if shape else None,
                                            stream.h if stream else None))


class CacheBatch:
    """Frame-cache builds of several assets as one graph replay
    (vbt_cache_batch_*).  jobs: (asset, t0, nt, layout, DeviceArray out)."""

    def __init__(self, jobs, shape: BatchShape | None = None, stream: Stream | None = None):
        self.stream = stream or Stream()
        self.jobs = list(jobs)  # keeps assets and outputs alive
        arr = (CacheJob * len(self.jobs))()
        for i, (asset, t0, nt, layout, out) in enumerate(self.jobs):
            arr[i] = CacheJob(asset.h, int(t0), int(nt), int(layout), out.ptr)
        self.h = c_void_p()
        check(lib().vbt_cache_batch_create(arr, len(self.jobs), ctypes.byref(shape) if shape else None,
                                           ctypes.byref(self.h)))

    def launch(self, graph=True):
        fn = lib().vbt_cache_batch_launch if graph else lib().vbt_cache_batch_launch_direct
        check(fn(self.h, self.stream.h))

    @property
    def shape(self) -> BatchShape:
        s = BatchShape()
        check(lib().vbt_cache_batch_get_shape(self.h, ctypes.byref(s)))
        return s

    @shape.setter
    def shape(self, s: BatchShape):
        check(lib().vbt_cache_batch_set_shape(self.h, ctypes.byref(s)))

    def plan_info(self):
        items, wg = c_u64(), c_u32()
        check(lib().vbt_cache_batch_plan_info(self.h, ctypes.byref(items), ctypes.byref(wg)))
        return items.value, wg.value

    def autotune(self, reps: int = 3):
        """Coordinate-descent over the queue shape; returns (best shape dict,
        best ms, [(shape dict, ms), ...])."""
        best, ms = BatchShape(), ctypes.c_float()
        check(lib().vbt_cache_batch_autotune(self.h, reps, ctypes.byref(best), ctypes.byref(ms), self.stream.h))
        trials = []
        for i in range(lib().vbt_cache_batch_trial_count(self.h)):
            s, t = BatchShape(), ctypes.c_float()
            check(lib().vbt_cache_batch_trial(self.h, i, ctypes.byref(s), ctypes.byref(t)))
            trials.append((s.as_dict(), round(float(t.value), 4)))
        return best.as_dict(), float(ms.value), trials

    def close(self):
        if getattr(self, "h", None):
            lib().vbt_cache_batch_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching decode server (vbt_server_* in the C-ABI)."""

    def __init__(self, **cfg):
        c = ServerConfig()
        lib().vbt_server_default_config(ctypes.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise TypeError(f"unknown server option {k}")
            setattr(c, k, v)
        self.config = c
        self.h = c_void_p()
        check(lib().vbt_server_create(ctypes.byref(c), ctypes.byref(self.h)))
        self._assets = []

    def register(self, asset: Asset) -> int:
        i = c_int(-1)
        check(lib().vbt_server_register(self.h, asset.h, ctypes.byref(i)))
        self._assets.append(asset)
        return i.value

    def submit_query(self, asset_id: int, q: np.ndarray, comp: Compose | None = None) -> int:
        q = np.ascontiguousarray(q, dtype=np.uint32).reshape(-1, 4)
        rid = c_u64(0)
        check(lib().vbt_server_submit_query(self.h, asset_id, q.ctypes.data, len(q),
                                            ctypes.byref(comp) if comp is not None else None, ctypes.byref(rid)))
        return rid.value

    def submit_cache(self, asset_id: int, t: int) -> int:
        rid = c_u64(0)
        check(lib().vbt_server_submit_cache(self.h, asset_id, t, ctypes.byref(rid)))
        return rid.value

    def poll(self, rid: int, timeout_ms: int = -1) -> RequestResult:
        r = RequestResult()
        check(lib().vbt_server_poll(self.h, rid, timeout_ms, ctypes.byref(r)))
        return r

    def fetch(self, rid: int, with_status=False, timeout_ms: int = 60000):
        """Values (and per-query statuses) of a completed request; releases it.
        Element failures do not raise: they show as NaN values and statuses."""
        r = self.poll(rid, timeout_ms)
        vals = np.empty(r.count, np.float32)
        st = np.zeros(r.count, np.int32)
        rc = lib().vbt_server_fetch(self.h, rid, vals.ctypes.data, st.ctypes.data)
        if rc not in (VBT_OK, 2):
            check(rc)
        return (vals, st, r) if with_status else vals

    def release(self, rid: int):
        check(lib().vbt_server_release(self.h, rid))

    def stats(self, request_type: int = -1) -> dict:
        s = ServerStats()
        check(lib().vbt_server_stats_get(self.h, request_type, ctypes.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        check(lib().vbt_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().vbt_server_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_copy(dst: DeviceArray, src: DeviceArray, nbytes: int, stream: Stream | None = None):
    """Plain device copy kernel (the byte-counter reference for profiling)."""
    check(lib().vbt_device_copy(dst.ptr, src.ptr, nbytes, stream.h if stream else None))


def device_fill(dst: DeviceArray, word: int, nbytes: int, stream: Stream | None = None):
    check(lib().vbt_device_fill(dst.ptr, word & 0xFFFFFFFF, nbytes, stream.h if stream else None))


def selftest_weights(max_gap: int = 65535, stream: Stream | None = None) -> int:
    """Mismatching lerp weights of the frame-cache sweep over every gap up to
    max_gap (0 expected)."""
    n = c_u64()
    check(lib().vbt_selftest_weights(max_gap, ctypes.byref(n), stream.h if stream else None))
    return n.value
