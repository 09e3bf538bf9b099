"""Test helpers that build keyframe control streams and Mode2 leaves in the
VBTPACK4 byte layout, the way the reference's own unit tests build them
(tests/temporal_decode_tests.cpp:17-38 makeStream, :68-101 shellTest;
tests/render_temporal_payload_tests.cpp:30-61 makeValidMode2Leaf)."""
from __future__ import annotations

import math
import struct

import numpy as np


def half_round_trip(v: float) -> float:
    """float -> half -> float with IEEE round-to-nearest (numpy)."""
    return float(np.float32(np.float16(np.float32(v))))


def ref_half_bits(v: float) -> int:
    """renderTemporalFloatToHalfBits (round half up on the mantissa), restated for packing."""
    bits = struct.unpack("<I", struct.pack("<f", v))[0]
    sign = (bits >> 16) & 0x8000
    exp = ((bits >> 23) & 0xFF) - 127 + 15
    man = bits & 0x7FFFFF
    if exp <= 0:
        if exp < -10:
            return sign
        man |= 0x800000
        return sign | (((man >> (1 - exp)) + 0x1000) >> 13)
    if exp >= 31:
        return sign | 0x7C00
    rm = (man + 0x1000) >> 13
    if rm == 0x400:
        exp += 1
        return sign | 0x7C00 if exp >= 31 else sign | (exp << 10)
    return sign | (exp << 10) | (rm & 0x3FF)


def half_bits_to_float(h: int) -> float:
    return float(np.array([h], np.uint16).view(np.float16)[0].astype(np.float32))


class Stream:
    def __init__(self, codec=0, bins=8, max_bin_local_keys=255):
        self.codec = codec
        self.bins = bins
        self.max_bin_local_keys = max_bin_local_keys
        self.desc = []      # (keyStart, keyCount, reserved)
        self.bin_index = []  # flattened
        self.keys = []      # (frame, value)

    @property
    def count(self):
        return len(self.desc)

    def pack(self, align=True) -> bytes:
        out = bytearray()
        for ks, kc, r in self.desc:
            out += struct.pack("<HBB", ks, kc, r)
        b = bytes(self.bin_index)
        if align:
            b += b"\0" * ((-len(b)) % 4)
        out += b
        k = bytearray()
        for f, v in self.keys:
            if self.codec == 1:
                k += struct.pack("<HHf", f, 0, v)
            else:
                assert f <= 255
                k += struct.pack("<B", f) + struct.pack("<H", ref_half_bits(v))
        if align:
            k += b"\0" * ((-len(k)) % 4)
        out += k
        return bytes(out)


def make_stream(frames, bins, codec, dense) -> Stream:
    """makeStream of tests/temporal_decode_tests.cpp:17-38."""
    s = Stream(codec, bins, 255)
    for t in range(frames):
        if dense or t == 0 or t == frames - 1 or t % 13 == 7:
            value = float(np.float32(math.sin(float(np.float32(t) * np.float32(0.19)))))
            if codec == 0:
                value = half_bits_to_float(ref_half_bits(value))
            s.keys.append((t, value))
    s.desc.append((0, len(s.keys), 1))
    for b in range(bins):
        boundary = b * frames // bins
        idx = 0
        while idx < len(s.keys) and s.keys[idx][0] < boundary:
            idx += 1
        s.bin_index.append(0 if idx == 0 else idx - 1)
    return s


def lower_bound_reference(keys, frames, frame) -> float:
    """reference() of tests/temporal_decode_tests.cpp:40-50 (std::lower_bound lerp)."""
    frame = min(max(frame, 0), frames - 1)
    fs = [k[0] for k in keys]
    i = int(np.searchsorted(fs, frame, side="left"))
    if i == len(keys):
        return keys[-1][1]
    if i == 0 or keys[i][0] == frame:
        return keys[i][1]
    lo, hi = keys[i - 1], keys[i]
    alpha = np.float32(frame - lo[0]) / np.float32(hi[0] - lo[0])
    return float(np.float32(lo[1]) * (np.float32(1.0) - alpha) + np.float32(hi[1]) * alpha)


def replicate_controls(s: Stream, n: int) -> Stream:
    """Give a one-control stream n identical controls (as the shell test does)."""
    out = Stream(s.codec, s.bins, s.max_bin_local_keys)
    out.keys = list(s.keys)
    out.desc = [s.desc[0]] * n
    out.bin_index = list(s.bin_index[:s.bins]) * n
    return out


def occupancy_bytes(voxels) -> bytes:
    mask = [0] * 8
    for v in voxels:
        mask[v // 64] |= 1 << (v % 64)
    prefix = [sum(1 for v in voxels if v < g * 64) for g in range(8)]
    return struct.pack("<8Q", *mask) + struct.pack("<8H", *prefix)


def header_word(mode, coarse_codec=0, fine_codec=0, coarse_res=4, fine_res=0) -> int:
    return (mode & 3) | ((coarse_codec & 3) << 2) | ((fine_codec & 3) << 4) | ((min(coarse_res, 7) & 7) << 6) | \
        ((min(fine_res, 7) & 7) << 9)


def build_leaf(header: int, ckc: int, fkc: int, parts) -> bytes:
    out = bytearray(struct.pack("<IHH", header, ckc, fkc))
    for p in parts:
        out += p
        out += b"\0" * ((-len(out)) % 4)
    return bytes(out)


def shell_test_leaf():
    """shellTest of tests/temporal_decode_tests.cpp:68-101: Mode2 leaf with active
    voxels {0, 63, 64, 255, 511} over 101 frames; returns (leaf, residual stream)."""
    coarse = replicate_controls(make_stream(101, 8, 0, False), 64)
    residual = replicate_controls(make_stream(101, 4, 1, True), 5)
    voxels = [0, 63, 64, 255, 511]
    hdr = header_word(2, 0, residual.codec, 0, 0)  # header.coarseResolution defaults to 4 in the reference struct
    hdr = header_word(2, 0, residual.codec, 4, 0)
    leaf = build_leaf(hdr, len(coarse.keys), len(residual.keys), [coarse.pack(), occupancy_bytes(voxels),
                                                                    residual.pack()])
    return leaf, residual, voxels
