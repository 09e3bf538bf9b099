// format.hip -- host side of the VBTPACK4 format library (pure functions of
// the C-ABI): half conversion, header/prefix/layout, leaf parse and build,
// shell section, control-stream pack/unpack/validate, container validation.
//
// These are the reference's format functions (src/render_temporal_payload.cpp,
// src/render_temporal_decode.cpp:135-302, render/src/vbt_file.cpp) as a C-ABI;
// they are the building blocks callers use to make and inspect payloads.  The
// decode itself runs on the GPU (decode.hip).
#include "vbt_format.h"
#include "vbt_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

using namespace vbtk;

namespace {

uint32_t rd16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
uint32_t rd32(const uint8_t* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

int popc64(uint64_t w) { return __builtin_popcountll(w); }

void stream_layout(uint32_t start, uint32_t cc, uint32_t bins, uint32_t keys, uint32_t codec, int align,
                   vbt_stream_layout* o)
{
    o->descriptorOffset = start;
    o->descriptorBytes = cc * 4u;
    o->binIndexOffset = start + o->descriptorBytes;
    o->binIndexBytes = align ? align_up4(cc * bins) : cc * bins;
    o->keyframeOffset = o->binIndexOffset + o->binIndexBytes;
    o->keyframeBytes = align ? align_up4(keys * keyframe_bytes(codec)) : keys * keyframe_bytes(codec);
    o->totalBytes = o->descriptorBytes + o->binIndexBytes + o->keyframeBytes;
}

} // namespace

extern "C" {

// renderTemporalFloatToHalfBits (src/render_temporal_decode.cpp:135-165):
// round half up on the mantissa (+0x1000), overflow to inf, flush below 2^-25.
uint16_t vbt_float_to_half_bits(float value)
{
    uint32_t bits;
    std::memcpy(&bits, &value, 4);
    const uint32_t sign = (bits >> 16) & 0x8000u;
    int32_t exponent = (int32_t)((bits >> 23) & 0xFFu) - 112;
    uint32_t mantissa = bits & 0x7FFFFFu;
    if (exponent <= 0) {
        if (exponent < -10) return (uint16_t)sign;
        mantissa |= 0x800000u;
        return (uint16_t)(sign | (((mantissa >> (uint32_t)(1 - exponent)) + 0x1000u) >> 13));
    }
    if (exponent >= 31) return (uint16_t)(sign | 0x7C00u);
    const uint32_t m = (mantissa + 0x1000u) >> 13;
    if (m == 0x400u) {
        ++exponent;
        return (uint16_t)(exponent >= 31 ? (sign | 0x7C00u) : (sign | ((uint32_t)exponent << 10)));
    }
    return (uint16_t)(sign | ((uint32_t)exponent << 10) | (m & 0x3FFu));
}

float vbt_half_bits_to_float(uint16_t bits) { return half_to_float(bits); }

uint32_t vbt_keyframe_bytes(uint32_t codec) { return keyframe_bytes(codec); }

uint32_t vbt_pack_header(const vbt_header_fields* f)
{
    const uint32_t cr = f->coarseResolution < 7 ? f->coarseResolution : 7u;
    const uint32_t fr = f->fineResolution < 7 ? f->fineResolution : 7u;
    return ((uint32_t)f->mode & 3u) | (((uint32_t)f->coarseCodec & 3u) << 2) | (((uint32_t)f->fineCodec & 3u) << 4) |
           ((cr & 7u) << 6) | ((fr & 7u) << 9);
}

void vbt_unpack_header(uint32_t p, vbt_header_fields* o)
{
    const Header h = unpack_header(p);
    o->mode = (uint8_t)h.mode;
    o->coarseCodec = (uint8_t)h.coarseCodec;
    o->fineCodec = (uint8_t)h.fineCodec;
    o->coarseResolution = (uint8_t)h.coarseRes;
    o->fineResolution = (uint8_t)h.fineRes;
}

int vbt_make_leaf_prefix(uint32_t mode, uint32_t ckc, uint32_t fkc, vbt_leaf_prefix* out)
{
    (void)mode;
    if (ckc > 0xFFFFu || fkc > 0xFFFFu)
        return vbt_fail(VBT_ERR_FORMAT, "render temporal prefix keyframe count exceeds uint16");
    out->coarseKeyframeCount = (uint16_t)ckc;
    out->fineKeyframeCount = (uint16_t)fkc;
    return VBT_OK;
}

// validateRenderTemporalLeafPrefix (src/render_temporal_payload.cpp:123-141): 1 valid, 0 invalid
int vbt_validate_leaf_prefix(const vbt_leaf_prefix* p, const vbt_header_fields* h, int timeBinCount)
{
    if (timeBinCount != 8) return 0;
    const uint32_t cc = control_count(h->coarseResolution), fc = control_count(h->fineResolution);
    if (cc == 0 && p->coarseKeyframeCount != 0) return 0;
    if ((h->mode == MODE_FINE6 || h->mode == MODE_SHELL) && fc == 0 && p->fineKeyframeCount != 0) return 0;
    return 1;
}

void vbt_compute_leaf_layout(const vbt_header_fields* h, const vbt_leaf_prefix* p, int align, vbt_leaf_layout* o)
{
    std::memset(o, 0, sizeof(*o));
    o->headerBytes = 4;
    o->prefixOffset = 4;
    o->prefixBytes = 4;
    o->coarseOffset = 8;
    stream_layout(8, control_count(h->coarseResolution), 8, p->coarseKeyframeCount, h->coarseCodec, align, &o->coarse);
    o->fineOffset = 8 + o->coarse.totalBytes;
    if (h->mode == MODE_FINE6) {
        stream_layout(o->fineOffset, control_count(h->fineResolution), 8, p->fineKeyframeCount, h->fineCodec, align,
                      &o->fine);
        o->totalBytes = o->fineOffset + o->fine.totalBytes;
    } else if (h->mode == MODE_SHELL) {
        o->fine.descriptorOffset = o->fineOffset;
        o->fine.descriptorBytes = 80;
        o->fine.binIndexOffset = o->fineOffset + 80;
        o->fine.keyframeOffset = o->fineOffset + 80;
        const uint32_t kb = p->fineKeyframeCount * keyframe_bytes(h->fineCodec);
        o->fine.keyframeBytes = align ? align_up4(kb) : kb;
        o->fine.totalBytes = 80 + o->fine.keyframeBytes;
        o->totalBytes = o->fineOffset + o->fine.totalBytes;
    } else {
        o->totalBytes = o->coarseOffset + o->coarse.totalBytes;
    }
    if (h->mode == MODE_EMPTY) {
        o->coarseOffset = 0;
        o->fineOffset = 0;
        std::memset(&o->coarse, 0, sizeof(o->coarse));
        std::memset(&o->fine, 0, sizeof(o->fine));
        o->totalBytes = o->headerBytes;
    }
}

static int copy_out(const std::vector<uint8_t>& v, uint8_t* out, size_t cap, size_t* len)
{
    if (len) *len = v.size();
    if (out) {
        if (cap < v.size()) return vbt_fail(VBT_ERR_INVALID, "output buffer too small");
        std::memcpy(out, v.data(), v.size());
    }
    return VBT_OK;
}

// buildRenderTemporalPackedLeafBytes 5-arg (src/render_temporal_payload.cpp:199-232)
int vbt_build_leaf_bytes(const vbt_header_fields* h, const vbt_leaf_prefix* p, const uint8_t* coarse, size_t cl,
                         const uint8_t* fine, size_t fl, int align, uint8_t* out, size_t cap, size_t* len)
{
    vbt_leaf_layout L;
    vbt_compute_leaf_layout(h, p, align, &L);
    std::vector<uint8_t> v;
    const uint32_t packed = vbt_pack_header(h);
    if (h->mode == MODE_EMPTY) {
        v.resize(4);
        std::memcpy(v.data(), &packed, 4);
        return copy_out(v, out, cap, len);
    }
    if (cl != L.coarse.totalBytes)
        return vbt_fail(VBT_ERR_FORMAT, "render temporal coarse stream bytes do not match computed layout");
    if (h->mode == MODE_FINE6 && fl != L.fine.totalBytes)
        return vbt_fail(VBT_ERR_FORMAT, "render temporal fine stream bytes do not match computed layout");
    v.assign(L.totalBytes, 0);
    std::memcpy(v.data(), &packed, 4);
    std::memcpy(v.data() + 4, p, 4);
    if (cl) std::memcpy(v.data() + 8, coarse, cl);
    if (fl) {
        // a fine stream longer than the layout (GRID4 / Mode2 through this
        // overload) would overrun the reference's buffer; here the leaf grows
        if (L.fineOffset + fl > v.size()) v.resize(L.fineOffset + fl, 0);
        std::memcpy(v.data() + L.fineOffset, fine, fl);
    }
    return copy_out(v, out, cap, len);
}

// buildRenderTemporalPackedLeafBytes 6-arg (src/render_temporal_payload.cpp:348-386)
int vbt_build_shell_leaf_bytes(const vbt_header_fields* h, const vbt_leaf_prefix* p, const uint8_t* coarse, size_t cl,
                               const uint8_t* occ, size_t ol, const uint8_t* res, size_t rl, int align, uint8_t* out,
                               size_t cap, size_t* len)
{
    std::vector<uint8_t> v(8, 0);
    const uint32_t packed = vbt_pack_header(h);
    std::memcpy(v.data(), &packed, 4);
    std::memcpy(v.data() + 4, p, 4);
    auto pad = [&] {
        if (align)
            while (v.size() % 4) v.push_back(0);
    };
    if (cl) { v.insert(v.end(), coarse, coarse + cl); pad(); }
    if (ol) { v.insert(v.end(), occ, occ + ol); pad(); }
    if (rl) { v.insert(v.end(), res, res + rl); pad(); }
    return copy_out(v, out, cap, len);
}

void vbt_pack_shell_section(const vbt_shell_section* s, uint8_t out[80])
{
    std::memcpy(out, s->shellMask, 64);
    std::memcpy(out + 64, s->groupPrefix, 16);
}

void vbt_unpack_shell_section(const uint8_t* d, uint64_t size, vbt_shell_section* s)
{
    std::memset(s, 0, sizeof(*s));
    if (d == nullptr || size < 80) return;
    std::memcpy(s->shellMask, d, 64);
    std::memcpy(s->groupPrefix, d + 64, 16);
}

uint16_t vbt_shell_active_count(const vbt_shell_section* s)
{
    uint32_t c = 0;
    for (int g = 0; g < 8; ++g) c += (uint32_t)popc64(s->shellMask[g]);
    return (uint16_t)c;
}

int vbt_parse_leaf(const uint8_t* leaf, size_t bytes, int align, vbt_leaf_view* o)
{
    std::memset(o, 0, sizeof(*o));
    if (leaf == nullptr || bytes < 4) return vbt_fail(VBT_ERR_FORMAT, "render temporal leaf bytes too small");
    vbt_unpack_header(rd32(leaf), &o->header);
    if (o->header.mode == MODE_EMPTY) {
        vbt_compute_leaf_layout(&o->header, &o->prefix, align, &o->layout);
        return VBT_OK;
    }
    if (bytes < 8) return vbt_fail(VBT_ERR_FORMAT, "render temporal leaf prefix missing");
    o->prefix.coarseKeyframeCount = (uint16_t)rd16(leaf + 4);
    o->prefix.fineKeyframeCount = (uint16_t)rd16(leaf + 6);
    vbt_compute_leaf_layout(&o->header, &o->prefix, align, &o->layout);
    if (o->layout.totalBytes > bytes) return vbt_fail(VBT_ERR_FORMAT, "render temporal leaf shorter than computed layout");
    if (o->header.mode == MODE_SHELL) {
        const size_t occOff = o->layout.fine.descriptorOffset, occBytes = o->layout.fine.descriptorBytes;
        if (occBytes != 80 || occOff > bytes || occBytes > bytes - occOff)
            return vbt_fail(VBT_ERR_FORMAT, "render temporal shell occupancy is out of bounds");
        vbt_shell_section s;
        vbt_unpack_shell_section(leaf + occOff, occBytes, &s);
        uint32_t expected = 0;
        for (int g = 0; g < 8; ++g) {
            if (s.groupPrefix[g] != (uint16_t)expected)
                return vbt_fail(VBT_ERR_FORMAT, "render temporal shell group prefix is inconsistent");
            expected = (uint16_t)(expected + (uint32_t)popc64(s.shellMask[g]));
        }
        if (expected == 0) return vbt_fail(VBT_ERR_FORMAT, "render temporal shell mode contains no active voxels");
        size_t resOff = occOff + occBytes;
        if (align) resOff = (resOff + 3u) & ~(size_t)3u;
        size_t binB = (size_t)expected * 4u, keyB = (size_t)o->prefix.fineKeyframeCount * keyframe_bytes(o->header.fineCodec);
        if (align) {
            binB = (binB + 3u) & ~(size_t)3u;
            keyB = (keyB + 3u) & ~(size_t)3u;
        }
        const size_t resBytes = (size_t)expected * 4u + binB + keyB;
        if (resOff > bytes || resBytes > bytes - resOff)
            return vbt_fail(VBT_ERR_FORMAT, "render temporal shell residual stream is out of bounds");
        if (resOff + resBytes != bytes) return vbt_fail(VBT_ERR_FORMAT, "render temporal shell leaf contains trailing bytes");
        o->shellActiveVoxelCount = (uint16_t)expected;
        o->shellResidualOffset = (uint32_t)resOff;
        o->shellResidualBytes = (uint32_t)resBytes;
    }
    return VBT_OK;
}

// validateRenderTemporalControlStreamData (src/render_temporal_decode.cpp:201-222)
int vbt_validate_control_stream(const vbt_stream_data* s)
{
    if (s->controlCount == 0 || s->timeBinCount == 0 || s->maxBinLocalKeys == 0) return 0;
    if (!s->descriptors || (!s->binIndex && s->timeBinCount)) return 0;
    for (uint32_t i = 0; i < s->controlCount; ++i) {
        const vbt_control_descriptor& d = s->descriptors[i];
        if ((size_t)d.keyStart + d.keyCount > s->keyframeCount) return 0;
        if ((uint32_t)d.keyCount > (uint32_t)s->timeBinCount * s->maxBinLocalKeys) return 0;
        if (d.reserved != 0 && d.reserved > s->maxBinLocalKeys) return 0;
    }
    return 1;
}

// packRenderTemporalControlStream (src/render_temporal_decode.cpp:224-264)
int vbt_pack_control_stream(const vbt_stream_data* s, int align, uint8_t* out, size_t cap, size_t* len)
{
    if (!vbt_validate_control_stream(s)) return vbt_fail(VBT_ERR_FORMAT, "invalid render temporal control stream");
    const uint32_t db = (uint32_t)s->controlCount * 4u;
    uint32_t bb = (uint32_t)s->controlCount * s->timeBinCount;
    uint32_t kb = s->keyframeCount * keyframe_bytes(s->codec);
    if (align) {
        bb = align_up4(bb);
        kb = align_up4(kb);
    }
    std::vector<uint8_t> v((size_t)db + bb + kb, 0);
    for (uint32_t i = 0; i < s->controlCount; ++i) {
        const vbt_control_descriptor& d = s->descriptors[i];
        v[4 * i] = (uint8_t)(d.keyStart & 0xFF);
        v[4 * i + 1] = (uint8_t)(d.keyStart >> 8);
        v[4 * i + 2] = d.keyCount;
        v[4 * i + 3] = d.reserved;
    }
    std::memcpy(v.data() + db, s->binIndex, (size_t)s->controlCount * s->timeBinCount);
    uint8_t* k = v.data() + db + bb;
    for (uint32_t i = 0; i < s->keyframeCount; ++i) {
        const vbt_control_keyframe& key = s->keyframes[i];
        const size_t off = (size_t)i * keyframe_bytes(s->codec);
        if (s->codec == CODEC_FP32) {
            k[off] = (uint8_t)(key.frame & 0xFF);
            k[off + 1] = (uint8_t)(key.frame >> 8);
            std::memcpy(k + off + 4, &key.value, 4);
        } else {
            if (key.frame > 255u)
                return vbt_fail(VBT_ERR_FORMAT, "render temporal fp16 keyframe frame exceeds uint8 range");
            const uint16_t hb = vbt_float_to_half_bits(key.value);
            k[off] = (uint8_t)key.frame;
            k[off + 1] = (uint8_t)(hb & 0xFF);
            k[off + 2] = (uint8_t)(hb >> 8);
        }
    }
    return copy_out(v, out, cap, len);
}

// unpackRenderTemporalControlStream (src/render_temporal_decode.cpp:266-302)
int vbt_unpack_control_stream(const uint8_t* base, uint32_t codec, uint16_t cc, uint8_t bins, uint8_t maxBin,
                              uint32_t keys, uint32_t streamBytes, int align, vbt_stream_data* out)
{
    const uint32_t db = (uint32_t)cc * 4u;
    uint32_t bb = (uint32_t)cc * bins, kb = keys * keyframe_bytes(codec);
    if (align) {
        bb = align_up4(bb);
        kb = align_up4(kb);
    }
    if (db + bb + kb > streamBytes) return vbt_fail(VBT_ERR_FORMAT, "render temporal stream bytes too small");
    out->codec = (uint8_t)codec;
    out->controlCount = cc;
    out->timeBinCount = bins;
    out->maxBinLocalKeys = maxBin;
    out->keyframeCount = keys;
    for (uint32_t i = 0; i < cc; ++i) {
        out->descriptors[i].keyStart = (uint16_t)rd16(base + 4 * i);
        out->descriptors[i].keyCount = base[4 * i + 2];
        out->descriptors[i].reserved = base[4 * i + 3];
    }
    std::memcpy(out->binIndex, base + db, (size_t)cc * bins);
    const Bytes b{base + db + bb};
    for (uint32_t i = 0; i < keys; ++i) {
        out->keyframes[i].frame = (uint16_t)key_frame(b, 0, codec, i);
        out->keyframes[i].value = key_value(b, 0, codec, i);
    }
    return VBT_OK;
}

// loadVbtFile validation (render/src/vbt_file.cpp:104-170, 174-275) over an image
int vbt_validate_file(const void* bytes_, uint64_t size, vbt_file_header* h)
{
    const uint8_t* bytes = (const uint8_t*)bytes_;
    std::memset(h, 0, sizeof(*h));
    if (size < 52) return vbt_fail(VBT_ERR_FORMAT, "Invalid VBTPACK4 file size.");
    uint32_t w[11];
    for (int i = 0; i < 11; ++i) w[i] = rd32(bytes + 8 + 4 * i);
    std::memcpy(h, w, sizeof(w));
    if (std::memcmp(bytes, "VBTPACK4", 8) != 0) return vbt_fail(VBT_ERR_FORMAT, "Unsupported magic. Expected VBTPACK4.");
    if (h->version != 4) {
        char m[96];
        std::snprintf(m, sizeof(m), "Unsupported VBTPACK4 version: %u", h->version);
        return vbt_fail(VBT_ERR_FORMAT, m);
    }
    if (!h->width || !h->height || !h->depth || !h->frames)
        return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 dimensions and frame count must be non-zero.");
    if (h->leafSize != 8) return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 v4 requires leafSize=8.");
    if (h->coarseResolution == 0 || h->coarseResolution > 8) return vbt_fail(VBT_ERR_FORMAT, "Invalid VBTPACK4 coarse resolution.");
    if (h->profileType > 2) return vbt_fail(VBT_ERR_FORMAT, "Invalid VBTPACK4 profile type.");
    if (h->maxCoarseKeep > h->frames || h->coarseAcScaleCount > 64 || h->coarseAcScaleCount > h->maxCoarseKeep)
        return vbt_fail(VBT_ERR_FORMAT, "Invalid VBTPACK4 coarse temporal metadata.");
    const uint64_t expected = (((uint64_t)h->width + 7) / 8) * (((uint64_t)h->height + 7) / 8) * (((uint64_t)h->depth + 7) / 8);
    if (expected > 0xffffffffull) return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 leaf count overflows the v4 header.");
    if (h->leafCount != expected) return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 leafCount does not match dimensions and leafSize.");
    const uint64_t minimum = 52 + (uint64_t)h->coarseAcScaleCount * 4 + ((uint64_t)h->leafCount + 1) * 4;
    if (minimum > size) return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 file is truncated before its offset table.");
    const uint8_t* offs = bytes + 52 + (uint64_t)h->coarseAcScaleCount * 4;
    if (rd32(offs) != 0) return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 offset table must begin at word zero.");
    uint32_t prev = 0;
    for (uint64_t i = 1; i <= h->leafCount; ++i) {
        const uint32_t cur = rd32(offs + 4 * i);
        if (prev >= cur) return vbt_fail(VBT_ERR_FORMAT, "VBTPACK4 offsets must be strictly increasing for every leaf.");
        prev = cur;
    }
    const uint64_t expectedBytes = minimum + (uint64_t)prev * 4;
    if (expectedBytes != size)
        return vbt_fail(VBT_ERR_FORMAT, expectedBytes > size ? "VBTPACK4 payload is truncated."
                                                             : "VBTPACK4 file contains trailing bytes.");
    return VBT_OK;
}

uint32_t vbt_leaf_index_for_voxel(const vbt_file_header* h, uint32_t x, uint32_t y, uint32_t z)
{
    if (!h->width || !h->height || !h->depth || !h->leafSize) return 0;
    const uint32_t s = h->leafSize;
    const uint32_t lx = (h->width + s - 1) / s, ly = (h->height + s - 1) / s, lz = (h->depth + s - 1) / s;
    const uint32_t bx = x / s < lx - 1 ? x / s : lx - 1, by = y / s < ly - 1 ? y / s : ly - 1,
                   bz = z / s < lz - 1 ? z / s : lz - 1;
    return (bz * ly + by) * lx + bx;
}

void vbt_decode_scientific_header(uint32_t p, vbt_scientific_header* o)
{
    std::memset(o, 0, sizeof(*o));
    o->mode = p & 3u;
    o->coarseKeepMinus1 = (p >> 2) & 0x3Fu;
    if (o->mode == 0u) {
        o->coarseQuantClass = (p >> 8) & 7u;
    } else if (o->mode == 1u) {
        o->tierId = (p >> 8) & 3u;
        o->framesPerBinMinus1 = (p >> 10) & 0xFu;
        o->eventQuantCode = (p >> 14) & 3u;
        o->packedEventCount = (p >> 16) & 0x0FFFu;
        o->sparseSubtype = (p >> 28) & 0xFu;
    } else {
        o->denseQuantBits = (p >> 8) & 0xFu;
        o->denseSubtype = (p >> 12) & 7u;
        o->denseTemporalKeepMinus1 = (p >> 15) & 0xFu;
    }
}

const char* vbt_status_message(int32_t s)
{
    switch (s) {
    case VBT_LEAF_OK: return "ok";
    case VBT_LEAF_ERR_TOO_SMALL: return "render temporal leaf bytes too small";
    case VBT_LEAF_ERR_PREFIX_MISSING: return "render temporal leaf prefix missing";
    case VBT_LEAF_ERR_SHORT: return "render temporal leaf shorter than computed layout";
    case VBT_LEAF_ERR_OCCUPANCY_OOB: return "render temporal shell occupancy is out of bounds";
    case VBT_LEAF_ERR_GROUP_PREFIX: return "render temporal shell group prefix is inconsistent";
    case VBT_LEAF_ERR_NO_ACTIVE: return "render temporal shell mode contains no active voxels";
    case VBT_LEAF_ERR_RESIDUAL_OOB: return "render temporal shell residual stream is out of bounds";
    case VBT_LEAF_ERR_TRAILING: return "render temporal shell leaf contains trailing bytes";
    case VBT_LEAF_ERR_KEY_RANGE: return "render temporal keyframe range out of bounds";
    case VBT_LEAF_ERR_STREAM_OOB: return "render temporal stream decode out of bounds";
    case VBT_LEAF_ERR_SCI_RANGE: return "scientific payload read outside the payload pool";
    case VBT_LEAF_ERR_SCI_BASIS: return "scientific dense basis block is inconsistent";
    case VBT_QUERY_ERR_OUTSIDE: return "query voxel outside the volume";
    case VBT_QUERY_ERR_FRAME: return "scientific query frame out of range";
    case VBT_ERR_CONTROL_INDEX: return "render temporal control index out of range";
    default: return "unknown status";
    }
}

} // extern "C"

// ------------------------------------------------------- query patterns --
// render/src/query_patterns.cpp:8-101.  The generator draws from the same
// standard engine and distributions in the same order as the reference, so
// a seed gives the reference's queries.
extern "C" {

int vbt_parse_query_pattern(const char* text, int* pattern)
{
    if (!text || !pattern) return vbt_fail(VBT_ERR_INVALID, "null query pattern text or output");
    static const char* names[] = {"random", "same-t", "same-xyz", "coherent"};
    for (int i = 0; i < 4; ++i)
        if (!std::strcmp(text, names[i])) {
            *pattern = i;
            return VBT_OK;
        }
    return vbt_fail(VBT_ERR_INVALID, std::string("unknown query pattern: ") + text);
}

int vbt_generate_queries(const vbt_file_header* h, int pattern, uint32_t count, uint32_t seed, uint32_t* out)
{
    if (!h || (!out && count)) return vbt_fail(VBT_ERR_INVALID, "null header or output");
    if (pattern < VBT_QUERY_RANDOM || pattern > VBT_QUERY_COHERENT)
        return vbt_fail(VBT_ERR_INVALID, "unknown query pattern");
    std::mt19937 rng(seed);
    auto axis = [](uint32_t n) { return std::uniform_int_distribution<uint32_t>(0u, std::max(1u, n) - 1u); };
    auto dx = axis(h->width), dy = axis(h->height), dz = axis(h->depth), dt = axis(h->frames);
    const uint32_t fx = dx(rng), fy = dy(rng), fz = dz(rng), ft = dt(rng);
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t* q = out + 4ull * i;
        switch (pattern) {
        case VBT_QUERY_RANDOM:
            q[0] = dx(rng);
            q[1] = dy(rng);
            q[2] = dz(rng);
            q[3] = dt(rng);
            break;
        case VBT_QUERY_SAME_T:
            q[0] = dx(rng);
            q[1] = dy(rng);
            q[2] = dz(rng);
            q[3] = ft;
            break;
        case VBT_QUERY_SAME_XYZ:
            q[0] = fx;
            q[1] = fy;
            q[2] = fz;
            q[3] = dt(rng);
            break;
        default: {
            const uint32_t bx = dx(rng) / 8u * 8u, by = dy(rng) / 8u * 8u, bz = dz(rng) / 8u * 8u;
            q[0] = std::min(h->width - 1u, bx + i % 8u);
            q[1] = std::min(h->height - 1u, by + (i / 8u) % 8u);
            q[2] = std::min(h->depth - 1u, bz + (i / 64u) % 8u);
            q[3] = ft;
        }
        }
    }
    return VBT_OK;
}

int vbt_sort_queries_by_leaf(const vbt_file_header* h, uint32_t* queries, uint32_t n)
{
    if (!h || (!queries && n)) return vbt_fail(VBT_ERR_INVALID, "null header or queries");
    struct Q {
        uint32_t leaf, x, y, z, t;
    };
    std::vector<Q> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* q = queries + 4ull * i;
        v[i] = Q{vbt_leaf_index_for_voxel(h, q[0], q[1], q[2]), q[0], q[1], q[2], q[3]};
    }
    std::stable_sort(v.begin(), v.end(), [](const Q& a, const Q& b) {
        if (a.leaf != b.leaf) return a.leaf < b.leaf;
        if (a.t != b.t) return a.t < b.t;
        if (a.z != b.z) return a.z < b.z;
        if (a.y != b.y) return a.y < b.y;
        return a.x < b.x;
    });
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t* q = queries + 4ull * i;
        q[0] = v[i].x;
        q[1] = v[i].y;
        q[2] = v[i].z;
        q[3] = v[i].t;
    }
    return VBT_OK;
}

} // extern "C"
