"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, when built, of
the reference decode compiled in oracle/_ref/libvbtref.so.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker / CPU baseline.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvbtref.so")

u8p = ctypes.POINTER(ctypes.c_uint8)


class Header(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint8) for n in ("mode", "coarseCodec", "fineCodec", "coarseResolution", "fineResolution")]


class StreamLayout(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("descriptorOffset", "descriptorBytes", "binIndexOffset",
                                                 "binIndexBytes", "keyframeOffset", "keyframeBytes", "totalBytes")]


class LeafLayout(ctypes.Structure):
    _fields_ = [("headerBytes", ctypes.c_uint32), ("prefixOffset", ctypes.c_uint32), ("prefixBytes", ctypes.c_uint32),
                ("coarseOffset", ctypes.c_uint32), ("coarse", StreamLayout), ("fineOffset", ctypes.c_uint32),
                ("fine", StreamLayout), ("totalBytes", ctypes.c_uint32)]


class LeafView(ctypes.Structure):
    _fields_ = [("header", Header), ("coarseKeyframeCount", ctypes.c_uint16), ("fineKeyframeCount", ctypes.c_uint16),
                ("layout", LeafLayout), ("shellActiveVoxelCount", ctypes.c_uint16),
                ("shellResidualOffset", ctypes.c_uint32), ("shellResidualBytes", ctypes.c_uint32)]


class Trace(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint32), ("hi", ctypes.c_uint32), ("f0", ctypes.c_int32), ("f1", ctypes.c_int32),
                ("bin", ctypes.c_uint32), ("keyStart", ctypes.c_uint32), ("keyCount", ctypes.c_uint32),
                ("kind", ctypes.c_uint32)]


class FileHeader(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("version", "width", "height", "depth", "frames", "leafSize",
                                                 "coarseResolution", "maxCoarseKeep", "leafCount", "profileType",
                                                 "coarseAcScaleCount")]


class VFile(ctypes.Structure):
    _fields_ = [("header", FileHeader), ("coarseAcScales", ctypes.c_void_p), ("offsetsWords", ctypes.c_void_p),
                ("payloadWords", ctypes.c_void_p), ("payloadWordCount", ctypes.c_uint32),
                ("sciDenseBasis", ctypes.c_uint32)]


class Compose(ctypes.Structure):
    _fields_ = [("signRule", ctypes.c_int), ("background", ctypes.c_float), ("shellEmptyScale", ctypes.c_float),
                ("shellEmptyMax", ctypes.c_float)]


_O = None
_R = None


def oracle():
    global _O
    if _O is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.vbto_last_error.restype = ctypes.c_char_p
        L.vbto_half_bits_to_float.restype = ctypes.c_float
        L.vbto_half_bits_to_float.argtypes = [ctypes.c_uint16]
        L.vbto_float_to_half_bits.restype = ctypes.c_uint16
        L.vbto_float_to_half_bits.argtypes = [ctypes.c_float]
        L.vbto_pack_header.restype = ctypes.c_uint32
        L.vbto_unpack_header.argtypes = [ctypes.c_uint32, ctypes.POINTER(Header)]
        L.vbto_compute_leaf_layout.argtypes = [ctypes.POINTER(Header), ctypes.c_uint16, ctypes.c_uint16, ctypes.c_int,
                                               ctypes.POINTER(LeafLayout)]
        L.vbto_parse_leaf.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(LeafView)]
        L.vbto_decode_control_value.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint16, ctypes.c_uint8,
                                                ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_uint16, ctypes.c_int,
                                                ctypes.POINTER(ctypes.c_float), ctypes.POINTER(Trace)]
        L.vbto_decode_leaf_control_value.argtypes = [ctypes.c_void_p, ctypes.POINTER(LeafView), ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int, ctypes.c_uint16, ctypes.c_int,
                                                     ctypes.POINTER(ctypes.c_float), ctypes.POINTER(Trace)]
        L.vbto_decode_shell_voxel_value.argtypes = [ctypes.c_void_p, ctypes.POINTER(LeafView), ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_uint16, ctypes.c_int,
                                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32)]
        L.vbto_load_file_memory.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(VFile)]
        L.vbto_leaf_index_for_voxel.restype = ctypes.c_uint32
        L.vbto_leaf_index_for_voxel.argtypes = [ctypes.POINTER(FileHeader), ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_uint32]
        L.vbto_decode_frame_dense.argtypes = [ctypes.POINTER(VFile), ctypes.c_int, ctypes.POINTER(Compose),
                                              ctypes.c_void_p, ctypes.c_int]
        L.vbto_query_points.argtypes = [ctypes.POINTER(VFile), ctypes.c_void_p, ctypes.c_uint64,
                                        ctypes.POINTER(Compose), ctypes.c_void_p, ctypes.c_int]
        L.vbto_page_table.restype = ctypes.c_uint32
        L.vbto_page_table.argtypes = [ctypes.POINTER(VFile), ctypes.c_void_p]
        L.vbto_leaf_cache_count.restype = ctypes.c_uint32
        L.vbto_leaf_cache_count.argtypes = [ctypes.POINTER(VFile), ctypes.c_uint32]
        L.vbto_frame_cache_576.argtypes = [ctypes.POINTER(VFile), ctypes.c_int, ctypes.c_void_p]
        L.vbto_frame_cache_packed.argtypes = [ctypes.POINTER(VFile), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.vbto_sample_points.argtypes = [ctypes.POINTER(VFile), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.vbto_march_rays.argtypes = [ctypes.POINTER(VFile), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int]
        L.vbto_dct_basis.restype = ctypes.c_float
        L.vbto_dct_basis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.vbto_sci_value_at.argtypes = [ctypes.POINTER(VFile)] + [ctypes.c_uint32] * 5 + [ctypes.POINTER(ctypes.c_float)]
        L.vbto_sci_leaf_frame.argtypes = [ctypes.POINTER(VFile), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        L.vbto_sci_basis_residual.argtypes = [ctypes.POINTER(VFile), ctypes.c_uint32, ctypes.c_uint32] + \
            [ctypes.c_uint32] * 4 + [ctypes.POINTER(ctypes.c_float)]
        L.vbto_sci_basis_extent.argtypes = [ctypes.POINTER(VFile), ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.vbto_decode_render_leaf_frame.argtypes = [ctypes.POINTER(VFile), ctypes.c_uint32, ctypes.c_int,
                                                    ctypes.POINTER(Compose), ctypes.c_void_p]
        L.vbto_cycles_pack_blob.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.vbto_cycles_sample.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int]
        L.vbto_cycles_error.restype = ctypes.c_char_p
        _O = L
    return _O


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _R
    if _R is None:
        L = ctypes.CDLL(REF_SO)
        L.vbtref_last_error.restype = ctypes.c_char_p
        L.vbtref_half_to_float.restype = ctypes.c_float
        L.vbtref_half_to_float.argtypes = [ctypes.c_uint16]
        L.vbtref_float_to_half.restype = ctypes.c_uint16
        L.vbtref_float_to_half.argtypes = [ctypes.c_float]
        L.vbtref_decode_control_value.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint16, ctypes.c_uint8,
                                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint16,
                                                  ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.vbtref_parse_leaf.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        L.vbtref_leaf_control_value.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_uint16, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.vbtref_shell_voxel_value.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint16,
                                               ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.vbtref_load_file.restype = ctypes.c_void_p
        L.vbtref_load_file.argtypes = [ctypes.c_char_p]
        L.vbtref_free_file.argtypes = [ctypes.c_void_p]
        L.vbtref_sci_value.argtypes = [ctypes.c_void_p] + [ctypes.c_uint32] * 4 + [ctypes.POINTER(ctypes.c_float)]
        L.vbtref_sci_frame.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.vbtref_cache_frame.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.vbtref_sci_values.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        L.vbtref_generate_queries.argtypes = [ctypes.c_uint32] * 4 + [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                                                      ctypes.c_int, ctypes.c_void_p]
        _R = L
    return _R


def last_error() -> str:
    return oracle().vbto_last_error().decode()


def compose(sign_rule=0, background=0.0, shell_empty_scale=1.0, shell_empty_max=math.nan) -> Compose:
    return Compose(sign_rule, background, shell_empty_scale, shell_empty_max)


class OracleFile:
    """A VBTPACK4 image validated by the oracle loader (keeps the bytes alive)."""

    def __init__(self, data: bytes, dense_basis: bool = False):
        self.buf = np.frombuffer(bytearray(data), dtype=np.uint8)  # 4-byte aligned numpy allocation
        self.vf = VFile()
        rc = oracle().vbto_load_file_memory(self.buf.ctypes.data, len(data), ctypes.byref(self.vf))
        if rc != 0:
            raise ValueError(last_error())
        self.vf.sciDenseBasis = 1 if dense_basis else 0
        h = self.vf.header
        self.header = h
        n_scale = h.coarseAcScaleCount
        base = 52 + 4 * n_scale
        self.offsets = self.buf[base: base + 4 * (h.leafCount + 1)].view(np.uint32)
        self.payload = self.buf[base + 4 * (h.leafCount + 1):]

    @property
    def dims(self):
        h = self.header
        return h.width, h.height, h.depth, h.frames

    def leaf_bytes(self, leaf: int) -> np.ndarray:
        a, b = int(self.offsets[leaf]), int(self.offsets[leaf + 1])
        return self.payload[4 * a: 4 * b]

    def frame(self, t: int, comp: Compose | None = None, threads: int = 1) -> np.ndarray:
        W, H, D, _ = self.dims
        out = np.empty((D, H, W), np.float32)
        comp = comp or compose()
        rc = oracle().vbto_decode_frame_dense(ctypes.byref(self.vf), t, ctypes.byref(comp), out.ctypes.data, threads)
        if rc != 0:
            raise ValueError(last_error())
        return out

    def query(self, q: np.ndarray, comp: Compose | None = None, threads: int = 1, check=True) -> np.ndarray:
        q = np.ascontiguousarray(q, dtype=np.uint32)
        out = np.empty(len(q), np.float32)
        comp = comp or compose()
        rc = oracle().vbto_query_points(ctypes.byref(self.vf), q.ctypes.data, len(q), ctypes.byref(comp),
                                        out.ctypes.data, threads)
        if rc != 0 and check:
            raise ValueError(last_error())
        return out

    def page_table(self):
        L = self.header.leafCount
        out = np.empty(L, np.uint32)
        n = oracle().vbto_page_table(ctypes.byref(self.vf), out.ctypes.data)
        return out, n

    def cache_counts(self) -> np.ndarray:
        return np.array([oracle().vbto_leaf_cache_count(ctypes.byref(self.vf), l)
                         for l in range(self.header.leafCount)], dtype=np.uint32)

    def packed_offsets(self) -> np.ndarray:
        """Packed-cache slot offset of every active leaf (+ the total)."""
        if getattr(self, "_offs", None) is None:
            compact, n = self.page_table()
            counts = self.cache_counts()
            active = counts[compact != 0xFFFFFFFF] if self.header.profileType != 2 else counts
            offs = np.zeros(len(active) + 1, np.uint64)
            offs[1:] = np.cumsum(active, dtype=np.uint64)
            self._offs = offs
        return self._offs

    def cache_packed(self, t: int):
        offs = self.packed_offsets()
        out = np.zeros(int(offs[-1]), np.float32)
        rc = oracle().vbto_frame_cache_packed(ctypes.byref(self.vf), t, offs.ctypes.data, out.ctypes.data)
        if rc != 0:
            raise ValueError(last_error())
        return out, offs

    def cache_packed_frames(self, ts, threads: int = 8) -> np.ndarray:
        """Packed caches of frames `ts` as one (len(ts), slots) array, one
        frame per host thread (the oracle call releases the GIL)."""
        from concurrent.futures import ThreadPoolExecutor

        offs = self.packed_offsets()
        ts = list(ts)
        out = np.zeros((len(ts), int(offs[-1])), np.float32)

        def one(i):
            row = out[i]
            rc = oracle().vbto_frame_cache_packed(ctypes.byref(self.vf), int(ts[i]), offs.ctypes.data,
                                                  row.ctypes.data)
            if rc != 0:
                raise ValueError(last_error())

        with ThreadPoolExecutor(max(1, threads)) as pool:
            list(pool.map(one, range(len(ts))))
        return out

    def cache_576(self, t: int):
        compact, n = self.page_table()
        out = np.zeros(int(n) * 576, np.float32)
        rc = oracle().vbto_frame_cache_576(ctypes.byref(self.vf), t, out.ctypes.data)
        if rc != 0:
            raise ValueError(last_error())
        return out.reshape(int(n), 576)

    def sample_points(self, cache576: np.ndarray, params, xyz: np.ndarray, preserve_interior=False,
                      threads: int = 8) -> np.ndarray:
        """sampleVbtValue (vbt_oracle_march.c); params is a vbt_march_params
        ctypes structure (same layout as the oracle's)."""
        cache = np.ascontiguousarray(cache576, np.float32)
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        out = np.empty(len(xyz), np.float32)
        rc = oracle().vbto_sample_points(ctypes.byref(self.vf), cache.ctypes.data, ctypes.byref(params),
                                         xyz.ctypes.data, len(xyz), int(preserve_interior), out.ctypes.data, threads)
        if rc != 0:
            raise ValueError(last_error())
        return out

    def march_rays(self, cache576: np.ndarray, params, rays: np.ndarray, threads: int = 8) -> np.ndarray:
        cache = np.ascontiguousarray(cache576, np.float32)
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.empty((len(rays), 8), np.float32)
        rc = oracle().vbto_march_rays(ctypes.byref(self.vf), cache.ctypes.data, ctypes.byref(params), rays.ctypes.data,
                                      len(rays), out.ctypes.data, threads)
        if rc != 0:
            raise ValueError(last_error())
        return out

    def parse(self, leaf: int) -> LeafView:
        b = self.leaf_bytes(leaf)
        v = LeafView()
        rc = oracle().vbto_parse_leaf(b.ctypes.data, len(b), 1, ctypes.byref(v))
        if rc != 0:
            raise ValueError(last_error())
        return v

    def modes(self) -> np.ndarray:
        return (self.payload.view(np.uint32)[self.offsets[:-1]] & 3).astype(np.uint8)


CYCLES_KERNEL, CYCLES_WORLD = 1, 2


def cycles_blob(data: bytes, frame: int) -> np.ndarray:
    """The Cycles resident blob of a VBTPACK4 image (load_pack4 + pack_blob)."""
    L = oracle()
    buf = np.frombuffer(data, np.uint8)
    n = ctypes.c_uint64()
    if L.vbto_cycles_pack_blob(buf.ctypes.data, len(data), frame, None, 0, ctypes.byref(n)) != 0:
        raise ValueError(L.vbto_cycles_error().decode())
    out = np.zeros(n.value, np.uint8)
    if L.vbto_cycles_pack_blob(buf.ctypes.data, len(data), frame, out.ctypes.data, n.value, ctypes.byref(n)) != 0:
        raise ValueError(L.vbto_cycles_error().decode())
    return out


def cycles_sample(blob: np.ndarray, xyz: np.ndarray, flags: int = 0, threads: int = 1) -> np.ndarray:
    """Density at n points read from a blob (standalone or kernel form)."""
    blob = np.ascontiguousarray(blob, np.uint8)
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    out = np.empty(len(xyz), np.float32)
    L = oracle()
    if L.vbto_cycles_sample(blob.ctypes.data, len(blob), flags, xyz.ctypes.data, len(xyz), out.ctypes.data,
                            threads) != 0:
        raise ValueError(L.vbto_cycles_error().decode())
    return out


class RefFile:
    """A VBTPACK4 file loaded by the reference's own loader (oracle/_ref/libvbtref.so)."""

    def __init__(self, data: bytes):
        import tempfile

        with tempfile.NamedTemporaryFile(suffix=".vbtp", delete=False) as fh:
            fh.write(data)
            path = fh.name
        try:
            self.h = ref().vbtref_load_file(path.encode())
        finally:
            os.unlink(path)
        if not self.h:
            raise ValueError(ref().vbtref_last_error().decode())

    def cache_frame_into(self, t: int, offs: np.ndarray, out: np.ndarray):
        """The packed cache of frame t through the reference's per-value decoders."""
        if ref().vbtref_cache_frame(self.h, int(t), offs.ctypes.data, out.ctypes.data) != 0:
            raise ValueError(ref().vbtref_last_error().decode())

    def sci_values(self, q: np.ndarray) -> np.ndarray:
        """decodeScientificValueAtCpu for each (x, y, z, t) row, serially."""
        q = np.ascontiguousarray(q, dtype=np.uint32)
        out = np.empty(len(q), np.float32)
        if ref().vbtref_sci_values(self.h, q.ctypes.data, len(q), out.ctypes.data) != 0:
            raise ValueError(ref().vbtref_last_error().decode())
        return out

    def close(self):
        if getattr(self, "h", None):
            ref().vbtref_free_file(self.h)
            self.h = None

    def __del__(self):
        self.close()
