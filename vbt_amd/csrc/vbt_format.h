// vbt_format.h -- VBTPACK4 byte-layout arithmetic shared by the host format
// library and the CDNA4 kernels (__host__ __device__ inline).
//
// Everything here is integer work that must be bit-exact with the reference:
//   header bitfields      src/render_temporal_payload.cpp:87-107
//   stream layout         src/render_temporal_payload.cpp:43-78, 143-197
//   half -> float         src/render_temporal_decode.cpp:167-199
//   bracket + lerp        src/render_temporal_decode.cpp:79-131
//   scientific header     render/src/vbt_file.cpp:220-248
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VBT_HD __host__ __device__ __forceinline__

namespace vbtk {

enum : uint32_t {
    MODE_EMPTY = 0,
    MODE_GRID4 = 1,
    MODE_SHELL = 2, // TEMPORAL_FINE_COMPACT (Mode2)
    MODE_FINE6 = 3,
};

enum : uint32_t {
    CODEC_FP16 = 0,
    CODEC_FP32 = 1,
};

// per-leaf status codes (a leaf that fails validation poisons every request
// touching it; the messages are the reference's exception texts)
enum : uint32_t {
    LEAF_OK = 0,
    LEAF_ERR_TOO_SMALL = 1,        // "render temporal leaf bytes too small"
    LEAF_ERR_PREFIX_MISSING = 2,   // "render temporal leaf prefix missing"
    LEAF_ERR_SHORT = 3,            // "render temporal leaf shorter than computed layout"
    LEAF_ERR_OCCUPANCY_OOB = 4,    // "render temporal shell occupancy is out of bounds"
    LEAF_ERR_GROUP_PREFIX = 5,     // "render temporal shell group prefix is inconsistent"
    LEAF_ERR_NO_ACTIVE = 6,        // "render temporal shell mode contains no active voxels"
    LEAF_ERR_RESIDUAL_OOB = 7,     // "render temporal shell residual stream is out of bounds"
    LEAF_ERR_TRAILING = 8,         // "render temporal shell leaf contains trailing bytes"
    LEAF_ERR_KEY_RANGE = 9,        // "render temporal keyframe range out of bounds"
    LEAF_ERR_STREAM_OOB = 10,      // "render temporal stream decode out of bounds"
    LEAF_ERR_SCI_RANGE = 11,       // scientific payload read outside the payload pool
    LEAF_ERR_SCI_BASIS = 12,       // tile / patch basis block the encoder cannot have written
    LEAF_ERR_COUNT
};

VBT_HD uint32_t align_up4(uint32_t b) { return (b + 3u) & ~3u; }
VBT_HD uint32_t keyframe_bytes(uint32_t codec) { return codec == CODEC_FP32 ? 8u : 3u; }
VBT_HD uint32_t control_count(uint32_t res) { return res == 0 ? 0u : (uint32_t)(uint16_t)(res * res * res); }

struct Header {
    uint32_t mode, coarseCodec, fineCodec, coarseRes, fineRes;
};

VBT_HD Header unpack_header(uint32_t p)
{
    Header h;
    h.mode = p & 3u;
    h.coarseCodec = (p >> 2) & 3u;
    h.fineCodec = (p >> 4) & 3u;
    h.coarseRes = (p >> 6) & 7u;
    h.fineRes = (p >> 9) & 7u;
    return h;
}

// One keyframe stream as the decoder addresses it (offsets relative to the
// leaf's first byte).
struct StreamRef {
    uint32_t base;     // descriptor section
    uint32_t count;    // controls
    uint32_t bins;     // time bins per control
    uint32_t keys;     // keyframes in the stream
    uint32_t codec;
    uint32_t bytes;    // streamBytes limit used by the reference bounds check
    uint32_t keyBase;  // byte offset of keyframe section
};

VBT_HD uint32_t stream_total_bytes(uint32_t count, uint32_t bins, uint32_t keys, uint32_t codec)
{
    return count * 4u + align_up4(count * bins) + align_up4(keys * keyframe_bytes(codec));
}

// IEEE half bits -> float bits, the exact integer mapping of
// renderTemporalHalfBitsToFloat (subnormals normalised, NaN payload kept).
VBT_HD uint32_t half_to_float_bits(uint32_t h)
{
    const uint32_t s = (h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu;
    uint32_t m = h & 0x3FFu;
    if (e == 0u) {
        if (m == 0u) return s;
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t sh = (uint32_t)__clz(m) - 21u;
#else
        const uint32_t sh = (uint32_t)__builtin_clz(m) - 21u;
#endif
        m = (m << sh) & 0x3FFu;
        return s | ((113u - sh) << 23) | (m << 13);
    }
    if (e == 31u) return s | 0x7F800000u | (m << 13);
    return s | ((e + 112u) << 23) | (m << 13);
}

VBT_HD float half_to_float(uint32_t h)
{
    const uint32_t b = half_to_float_bits(h);
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(b);
#else
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
#endif
}

VBT_HD float bits_to_float(uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(b);
#else
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
#endif
}

// Byte reader over a leaf (global or LDS pointer).
struct Bytes {
    const uint8_t* p;
    VBT_HD uint32_t u8(uint32_t o) const { return p[o]; }
    VBT_HD uint32_t u16(uint32_t o) const { return (uint32_t)p[o] | ((uint32_t)p[o + 1] << 8); }
    VBT_HD uint32_t u32(uint32_t o) const
    {
        return (uint32_t)p[o] | ((uint32_t)p[o + 1] << 8) | ((uint32_t)p[o + 2] << 16) | ((uint32_t)p[o + 3] << 24);
    }
};

VBT_HD int key_frame(const Bytes& b, uint32_t keyBase, uint32_t codec, uint32_t idx)
{
    const uint32_t off = keyBase + idx * keyframe_bytes(codec);
    return codec == CODEC_FP32 ? (int)b.u16(off) : (int)b.u8(off);
}

VBT_HD float key_value(const Bytes& b, uint32_t keyBase, uint32_t codec, uint32_t idx)
{
    const uint32_t off = keyBase + idx * keyframe_bytes(codec);
    return codec == CODEC_FP32 ? bits_to_float(b.u32(off + 4)) : half_to_float(b.u16(off + 1));
}

VBT_HD int clamp_frame(int t, int F)
{
    const int fmax = F - 1 > 0 ? F - 1 : 0;
    int f = t < fmax ? t : fmax;
    return f < 0 ? 0 : f;
}

// decodeFromSequence (src/render_temporal_decode.cpp:79-131) exactly: bin hint,
// fallback, early outs, binary search, v0 + alpha * (v1 - v0).  Requires
// keyStart + keyCount <= keyframes (checked by the caller / validation).
VBT_HD float decode_sequence(const Bytes& b, uint32_t keyBase, uint32_t codec, uint32_t keyStart, uint32_t keyCount,
                             uint32_t binBase, uint32_t B, int F, int frameIndex)
{
    if (keyCount == 0u) return 0.0f;
    const int frame = clamp_frame(frameIndex, F);
    const uint32_t last = keyCount - 1u;
    uint32_t bin = 0u;
    if (B != 0u) {
        const int64_t q = ((int64_t)frame * (int64_t)B) / (int64_t)(F > 1 ? F : 1);
        bin = (uint32_t)q < B - 1u ? (uint32_t)q : B - 1u;
    }
    uint32_t lo = 0u;
    if (B != 0u) {
        const uint32_t h = b.u8(binBase + bin);
        lo = h < last ? h : last;
    }
    int f0 = key_frame(b, keyBase, codec, keyStart + lo);
    if (f0 > frame) {
        lo = 0u;
        f0 = key_frame(b, keyBase, codec, keyStart);
    }
    if (f0 >= frame || lo == last) return key_value(b, keyBase, codec, keyStart + lo);
    uint32_t hi = last;
    if (bin + 1u < B) {
        const uint32_t h = b.u8(binBase + bin + 1u) + 1u;
        hi = h < last ? h : last;
    }
    if (hi < lo + 1u) hi = lo + 1u;
    int f1 = key_frame(b, keyBase, codec, keyStart + hi);
    if (f1 < frame) {
        hi = last;
        f1 = key_frame(b, keyBase, codec, keyStart + hi);
    }
    if (f1 <= frame) return key_value(b, keyBase, codec, keyStart + hi);
    while (hi > lo + 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        const int fm = key_frame(b, keyBase, codec, keyStart + mid);
        if (fm < frame) {
            lo = mid;
            f0 = fm;
        } else {
            hi = mid;
            f1 = fm;
        }
    }
    const float v1 = key_value(b, keyBase, codec, keyStart + hi);
    if (f1 == frame) return v1;
    const float v0 = key_value(b, keyBase, codec, keyStart + lo);
    const float alpha = (float)(frame - f0) / (float)(f1 - f0);
    return v0 + alpha * (v1 - v0);
}

// One control of a stream (decodeRenderTemporalControlValue after its bounds
// checks, src/render_temporal_decode.cpp:316-356).
VBT_HD float decode_control(const Bytes& b, const StreamRef& s, uint32_t ci, int F, int t)
{
    const uint32_t d = s.base + ci * 4u;
    const uint32_t keyStart = b.u16(d);
    const uint32_t keyCount = b.u8(d + 2);
    return decode_sequence(b, s.keyBase, s.codec, keyStart, keyCount, s.base + s.count * 4u + ci * s.bins, s.bins, F, t);
}

// Per-leaf decode plan: the streams the composition reads, computed once per
// leaf from header + prefix (+ Mode2 occupancy).
struct LeafPlan {
    uint32_t mode;
    uint32_t coarseRes, fineRes;
    StreamRef coarse;
    StreamRef fine;      // the stream the composition's fine grid reads (hasFine)
    StreamRef shell;     // Mode2 shell residual stream
    uint32_t occOffset;  // Mode2 occupancy section (80 bytes)
    uint32_t hasFine;    // composition adds the fine grid (reference hasFine rule)
};

// computeRenderTemporalLeafLayout + parseRenderTemporalPackedLeaf checks
// (src/render_temporal_payload.cpp:143-311) plus descriptor range checks of
// decodeRenderTemporalControlValue.  Returns a LEAF_* status.
VBT_HD uint32_t plan_leaf(const Bytes& b, uint32_t leafBytes, LeafPlan& pl)
{
    pl.mode = 0;
    pl.coarseRes = pl.fineRes = 0;
    pl.hasFine = 0;
    pl.occOffset = 0;
    pl.coarse = StreamRef{0, 0, 8, 0, 0, 0, 0};
    pl.fine = StreamRef{0, 0, 8, 0, 0, 0, 0};
    pl.shell = StreamRef{0, 0, 4, 0, 0, 0, 0};
    if (leafBytes < 4u) return LEAF_ERR_TOO_SMALL;
    const Header h = unpack_header(b.u32(0));
    pl.mode = h.mode;
    pl.coarseRes = h.coarseRes;
    pl.fineRes = h.fineRes;
    if (h.mode == MODE_EMPTY) return LEAF_OK;
    if (leafBytes < 8u) return LEAF_ERR_PREFIX_MISSING;
    const uint32_t ckc = b.u16(4), fkc = b.u16(6);
    const uint32_t cc = control_count(h.coarseRes), fc = control_count(h.fineRes);
    pl.coarse.base = 8u;
    pl.coarse.count = cc;
    pl.coarse.bins = 8u;
    pl.coarse.keys = ckc;
    pl.coarse.codec = h.coarseCodec;
    pl.coarse.bytes = stream_total_bytes(cc, 8u, ckc, h.coarseCodec);
    pl.coarse.keyBase = 8u + cc * 4u + align_up4(cc * 8u);
    const uint32_t fineOffset = 8u + pl.coarse.bytes;
    uint32_t total = fineOffset;
    if (h.mode == MODE_FINE6) {
        pl.fine.base = fineOffset;
        pl.fine.count = fc;
        pl.fine.bins = 8u;
        pl.fine.keys = fkc;
        pl.fine.codec = h.fineCodec;
        pl.fine.bytes = stream_total_bytes(fc, 8u, fkc, h.fineCodec);
        pl.fine.keyBase = fineOffset + fc * 4u + align_up4(fc * 8u);
        total = fineOffset + pl.fine.bytes;
    } else if (h.mode == MODE_SHELL) {
        total = fineOffset + 80u + align_up4(fkc * keyframe_bytes(h.fineCodec));
    }
    if (total > leafBytes) return LEAF_ERR_SHORT;
    if (h.mode == MODE_SHELL) {
        pl.occOffset = fineOffset;
        if (fineOffset > leafBytes || 80u > leafBytes - fineOffset) return LEAF_ERR_OCCUPANCY_OOB;
        uint32_t expected = 0;
        for (uint32_t g = 0; g < 8u; ++g) {
            if (b.u16(fineOffset + 64u + 2u * g) != (expected & 0xFFFFu)) return LEAF_ERR_GROUP_PREFIX;
            const uint32_t lo = b.u32(fineOffset + 8u * g), hi = b.u32(fineOffset + 8u * g + 4u);
#ifdef __HIP_DEVICE_COMPILE__
            expected += (uint32_t)__popc(lo) + (uint32_t)__popc(hi);
#else
            expected += (uint32_t)__builtin_popcount(lo) + (uint32_t)__builtin_popcount(hi);
#endif
        }
        if (expected == 0u) return LEAF_ERR_NO_ACTIVE;
        const uint32_t resOff = fineOffset + 80u; // already 4-aligned
        const uint32_t resBytes = stream_total_bytes(expected, 4u, fkc, h.fineCodec);
        if (resOff > leafBytes || resBytes > leafBytes - resOff) return LEAF_ERR_RESIDUAL_OOB;
        if (resOff + resBytes != leafBytes) return LEAF_ERR_TRAILING;
        pl.shell.base = resOff;
        pl.shell.count = expected;
        pl.shell.bins = 4u;
        pl.shell.keys = fkc;
        pl.shell.codec = h.fineCodec;
        pl.shell.bytes = resBytes;
        pl.shell.keyBase = resOff + expected * 4u + align_up4(expected * 4u);
    }
    // composition rule of decodeLeafFrameValues (3D/render_temporal_vbt_to_vdb.cpp:211-214):
    // fine grid when fineResolution > 0, layout.fine.totalBytes > 0 and the
    // mode is not GRID4.  For a Mode2 leaf this reads a stream at the
    // occupancy offset (what decodeRenderTemporalLeafControlValue does with
    // useFineStream), bounded by the Mode2 layout's fine.totalBytes.
    if (h.fineRes > 0u && h.mode == MODE_SHELL) {
        const uint32_t limit = 80u + align_up4(fkc * keyframe_bytes(h.fineCodec));
        pl.fine.base = fineOffset;
        pl.fine.count = fc;
        pl.fine.bins = 8u;
        pl.fine.keys = fkc;
        pl.fine.codec = h.fineCodec;
        pl.fine.bytes = limit;
        pl.fine.keyBase = fineOffset + fc * 4u + align_up4(fc * 8u);
        if (stream_total_bytes(fc, 8u, fkc, h.fineCodec) > limit) return LEAF_ERR_STREAM_OOB;
        pl.hasFine = 1u;
    } else if (h.fineRes > 0u && h.mode == MODE_FINE6) {
        pl.hasFine = pl.fine.bytes > 0u ? 1u : 0u;
    }
    return LEAF_OK;
}

// Mode2 descriptor index of a voxel (decodeRenderTemporalLeafShellVoxelValue,
// src/render_temporal_decode.cpp:417-432); 0xffffffff when not occupied.
VBT_HD uint32_t shell_desc_index(const Bytes& b, uint32_t occOffset, uint32_t voxel)
{
    const uint32_t g = voxel >> 6, bit = voxel & 63u;
    const uint32_t lo = b.u32(occOffset + 8u * g), hi = b.u32(occOffset + 8u * g + 4u);
    const uint64_t mask = (uint64_t)lo | ((uint64_t)hi << 32);
    if (((mask >> bit) & 1ull) == 0ull) return 0xffffffffu;
    const uint64_t before = mask & ((1ull << bit) - 1ull);
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t pc = (uint32_t)__popcll(before);
#else
    const uint32_t pc = (uint32_t)__builtin_popcountll(before);
#endif
    return (b.u16(occOffset + 64u + 2u * g) + pc) & 0xFFFFu;
}

// locateCoord (3D/render_temporal_vbt_to_vdb.cpp:122-134) on the control
// lattice i * (7 / (res - 1)); res < 2 is defined as the single control.
VBT_HD void locate(float p, int extent, int res, int& i0, int& i1, float& w)
{
    if (res < 2) {
        i0 = i1 = 0;
        w = 0.0f;
        return;
    }
    const float scale = 7.0f / (float)(res - 1);
    const float maxP = (float)(extent - 1 > 0 ? extent - 1 : 0);
    if (p < 0.0f) p = 0.0f;
    if (p > maxP) p = maxP;
    int j = 1;
    while (j < res - 1 && (float)j * scale < p) ++j;
    i1 = j;
    i0 = j - 1 > 0 ? j - 1 : 0;
    const float c0 = (float)i0 * scale, c1 = (float)i1 * scale;
    w = (c1 > c0) ? ((p - c0) / (c1 - c0)) : 0.0f;
}

// trilinear in the reference's evaluation order; s[bit0 x | bit1 y | bit2 z]
VBT_HD float trilinear(const float s[8], float tx, float ty, float tz)
{
    const float c00 = s[0] * (1.0f - tx) + s[1] * tx;
    const float c01 = s[4] * (1.0f - tx) + s[5] * tx;
    const float c10 = s[2] * (1.0f - tx) + s[3] * tx;
    const float c11 = s[6] * (1.0f - tx) + s[7] * tx;
    const float c0 = c00 * (1.0f - ty) + c10 * ty;
    const float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}

// ---- scientific route (render/src/vbt_file.cpp:220-248) ----
struct SciHeader {
    uint32_t mode, keep, tier, fpb, count, bits, subtype, tk;
};

VBT_HD SciHeader sci_header(uint32_t p)
{
    SciHeader h;
    h.mode = p & 3u;
    h.keep = ((p >> 2) & 0x3Fu) + 1u;
    h.tier = (p >> 8) & 3u;
    h.fpb = ((p >> 10) & 0xFu) + 1u;
    h.count = (p >> 16) & 0x0FFFu;
    h.bits = (p >> 8) & 0xFu;
    h.subtype = (p >> 12) & 7u;
    h.tk = ((p >> 15) & 0xFu) + 1u;
    if (h.mode == 0u) { h.tier = h.fpb = h.count = h.bits = h.subtype = h.tk = 0; }
    if (h.mode == 1u) { h.bits = h.subtype = h.tk = 0; }
    if (h.mode >= 2u) { h.tier = h.fpb = h.count = 0; }
    return h;
}

// sampleCoarseAt's lattice (render/src/scientific_decode.cpp:99-123):
// f = (x / 7) * 3, i0 = clamp(floor(f), 0, 2)
VBT_HD void sci_coarse_coord(uint32_t x, int& i0, float& t)
{
    const float f = ((float)x / 7.0f) * 3.0f;
    int i = (int)floorf(f);
    i = i < 0 ? 0 : (i > 2 ? 2 : i);
    i0 = i;
    t = f - (float)i;
}

// dense lattice of decodeDenseValueAt (render/src/scientific_decode.cpp:141-152)
VBT_HD void sci_dense_coord(uint32_t x, int R, int& i0, float& t)
{
    const float f = ((float)x / 7.0f) * (float)(R - 1);
    int i = (int)floorf(f);
    if (i < 0) i = 0;
    if (i > R - 2) i = R - 2;
    i0 = i;
    t = f - (float)i;
}

// ---- dense subtypes 2 (tile) and 3 (patch) as the encoder reconstructs them
// The block after the coarse coefficients (appendPackedDenseTileTemporalBasis /
// appendPackedDensePatchTemporalBasis, src/spatial_first_encoder.cpp:426-458):
// tile = u64 active mask, u16 tile size, u16 active count; patch = u64 active
// mask, u16 patch size, u16 local resolution, u16 active count; then, when the
// value count (active count x cells) is positive, tk half scales and tk packed
// bases of (max(1, values) * bits + 7) / 8 bytes; tk and bits are the leaf
// header's.
struct SciBasis {
    uint64_t mask, values, basisBytes;
    uint32_t patch, size, res, count;
    uint32_t scalesOff, dataOff; // from the leaf's first byte
};

template <class B>
VBT_HD SciBasis sci_basis(const B& b, const SciHeader& h)
{
    SciBasis s;
    const uint32_t blk = 4u + 64u * (h.keep + 1u);
    s.patch = h.subtype == 3u ? 1u : 0u;
    s.mask = (uint64_t)b.u32(blk) | ((uint64_t)b.u32(blk + 4u) << 32);
    s.size = b.u16(blk + 8u);
    s.res = s.patch ? b.u16(blk + 10u) : 0u;
    s.count = b.u16(blk + (s.patch ? 12u : 10u));
    const uint64_t e = s.patch ? s.res : s.size;
    s.values = (uint64_t)s.count * e * e * e;
    s.basisBytes = ((s.values > 0u ? s.values : 1u) * h.bits + 7u) / 8u;
    s.scalesOff = blk + (s.patch ? 14u : 12u);
    s.dataOff = s.scalesOff + 2u * h.tk;
    return s;
}

// decodeBfp (render/src/scientific_decode.cpp:61-70)
VBT_HD float decode_bfp(uint32_t q, float scale, int bits)
{
    if (bits <= 2) {
        const float lv = q == 0u ? -1.0f : q == 1u ? -0.3333333f : q == 2u ? 0.3333333f : 1.0f;
        return lv * scale;
    }
    const int levels = (1 << bits) - 1;
    const float norm = ((float)q / (float)levels) * 2.0f - 1.0f;
    return norm * scale;
}

} // namespace vbtk
