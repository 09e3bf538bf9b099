/*
 * vbt_oracle.c -- CPU restatement of the VBTPACK4 random-access decode path.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the CPU baseline).
 * Nothing in the product links or calls this file.
 *
 * Each function names the reference function it restates (file:line in the
 * reference repository).  How each part is pinned -- i.e. what independent
 * evidence says this restatement is right -- is:
 *
 *   half<->float                 exhaustive 65,536-value check against numpy's
 *                                IEEE float16 (tests/test_oracle_units.py) and
 *                                against the reference compiled in oracle/_ref.
 *   header/prefix/layout/parse   closed-form byte layouts + the reference's own
 *                                payload tests restated (bad prefix, truncated
 *                                residual, trailing bytes, no active voxels,
 *                                voxel stride) + oracle/_ref on random leaves.
 *   control decode (bracket)     std::lower_bound lerp semantics of the
 *                                reference's temporal_decode_tests over
 *                                frames 1..255, bins {4,8}, codecs {fp16,fp32},
 *                                dense/sparse keys, hints {-1,0,255}; the
 *                                integer trace is checked against a brute-force
 *                                linear scan; oracle/_ref on golden assets.
 *   shell voxel decode           the reference's 512-voxel Mode2 shell test
 *                                restated; popcount descIndex brute force.
 *   loader                       the reference's vbt_file_tests restated
 *                                (version, leafCount, monotonic, empty leaf,
 *                                trailing, truncated).
 *   composition (trilinear)      closed forms: lattice points reproduce the
 *                                controls, linear control fields are
 *                                reproduced at every voxel; an independent
 *                                float64 numpy evaluation (searchsorted-based)
 *                                agrees within 1e-6.
 *   composition (Fine6, Mode2,   the reference encoder's own sampled RMSE of
 *   sign rules)                  the leaves it saved (smoke64: Grid4 + Fine6,
 *                                liquid_mini: Grid4 + Mode2) reproduced to its
 *                                9 printed decimals; a Fine6 leaf of linear
 *                                grids decodes to their closed-form sum; a
 *                                Mode2 leaf follows a truth table of the three
 *                                sign rules written from the reference's
 *                                conditions, bit for bit
 *                                (tests/test_composition_pins.py).
 *   page table                   invariant: rank among active leaves.
 *   scientific DCT + residuals   oracle/_ref (reference scientific_decode.cpp
 *                                compiled here) on golden scientific assets
 *                                covering CoarseOnly/Sparse/DenseGrid3/4 and
 *                                both dense subtypes; closed-form DCT-II
 *                                inverse in float64 numpy.
 *
 * Build: gcc -std=c11 -O3 -ffp-contract=off (see oracle/Makefile).
 */
#include "vbt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static _Thread_local char g_err[256];

static int fail(const char* msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return VBTO_ERR;
}

const char* vbto_last_error(void) { return g_err; }

static uint32_t align_up4(uint32_t b) { return (b + 3u) & ~3u; }

static uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | ((uint16_t)p[1] << 8)); }
static uint32_t rd32(const uint8_t* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

static int popcount64(uint64_t w)
{
    int c = 0;
    for (; w != 0; w &= (w - 1)) ++c; /* same loop as render_temporal_payload.cpp:275-277 */
    return c;
}

/* ------------------------------------------------------------------------- */
/* half <-> float: src/render_temporal_decode.cpp:135-165 and :167-199        */
/* ------------------------------------------------------------------------- */
uint16_t vbto_float_to_half_bits(float value)
{
    uint32_t bits;
    memcpy(&bits, &value, 4);
    const uint32_t sign = (bits >> 16) & 0x8000u;
    int32_t exponent = (int32_t)((bits >> 23) & 0xFFu) - 127 + 15;
    uint32_t mantissa = bits & 0x7FFFFFu;
    if (exponent <= 0) {
        if (exponent < -10) return (uint16_t)sign;
        mantissa |= 0x800000u;
        const uint32_t shifted = mantissa >> (uint32_t)(1 - exponent);
        const uint32_t rounded = (shifted + 0x1000u) >> 13;
        return (uint16_t)(sign | rounded);
    }
    if (exponent >= 31) return (uint16_t)(sign | 0x7C00u);
    const uint32_t rm = (mantissa + 0x1000u) >> 13;
    if (rm == 0x400u) {
        exponent += 1;
        if (exponent >= 31) return (uint16_t)(sign | 0x7C00u);
        return (uint16_t)(sign | ((uint32_t)exponent << 10));
    }
    return (uint16_t)(sign | ((uint32_t)exponent << 10) | (rm & 0x3FFu));
}

float vbto_half_bits_to_float(uint16_t bits)
{
    const uint32_t sign = ((uint32_t)(bits & 0x8000u)) << 16;
    uint32_t exponent = (bits >> 10) & 0x1Fu;
    uint32_t mantissa = bits & 0x03FFu;
    uint32_t out = 0;
    if (exponent == 0) {
        if (mantissa == 0) {
            out = sign;
        } else {
            exponent = 1;
            while ((mantissa & 0x0400u) == 0) {
                mantissa <<= 1;
                --exponent;
            }
            mantissa &= 0x03FFu;
            out = sign | ((exponent + 127 - 15) << 23) | (mantissa << 13);
        }
    } else if (exponent == 31) {
        out = sign | 0x7F800000u | (mantissa << 13);
    } else {
        out = sign | ((exponent + 127 - 15) << 23) | (mantissa << 13);
    }
    float f;
    memcpy(&f, &out, 4);
    return f;
}

/* ------------------------------------------------------------------------- */
/* header / layout: src/render_temporal_payload.cpp:32-197                    */
/* ------------------------------------------------------------------------- */
uint32_t vbto_keyframe_bytes(uint32_t codec) { return codec == 1u ? 8u : 3u; }

uint32_t vbto_pack_header(const vbto_header* h)
{
    uint32_t p = 0;
    p |= (uint32_t)h->mode & 3u;
    p |= ((uint32_t)h->coarseCodec & 3u) << 2;
    p |= ((uint32_t)h->fineCodec & 3u) << 4;
    p |= ((h->coarseResolution < 7 ? h->coarseResolution : 7u) & 7u) << 6;
    p |= ((h->fineResolution < 7 ? h->fineResolution : 7u) & 7u) << 9;
    return p;
}

void vbto_unpack_header(uint32_t p, vbto_header* o)
{
    o->mode = (uint8_t)(p & 3u);
    o->coarseCodec = (uint8_t)((p >> 2) & 3u);
    o->fineCodec = (uint8_t)((p >> 4) & 3u);
    o->coarseResolution = (uint8_t)((p >> 6) & 7u);
    o->fineResolution = (uint8_t)((p >> 9) & 7u);
}

static uint32_t control_count_for_res(uint32_t r) { return r == 0 ? 0u : (uint16_t)(r * r * r); }

/* render_temporal_payload.cpp:62-78 */
static void stream_layout(uint32_t start, uint32_t controlCount, uint32_t bins, uint32_t keys,
                          uint32_t codec, int align, vbto_stream_layout* o)
{
    o->descriptorOffset = start;
    o->descriptorBytes = controlCount * 4u;
    o->binIndexOffset = o->descriptorOffset + o->descriptorBytes;
    o->binIndexBytes = controlCount * bins;
    if (align) o->binIndexBytes = align_up4(o->binIndexBytes);
    o->keyframeOffset = o->binIndexOffset + o->binIndexBytes;
    o->keyframeBytes = keys * vbto_keyframe_bytes(codec);
    if (align) o->keyframeBytes = align_up4(o->keyframeBytes);
    o->totalBytes = o->descriptorBytes + o->binIndexBytes + o->keyframeBytes;
}

/* render_temporal_payload.cpp:143-197 */
void vbto_compute_leaf_layout(const vbto_header* h, uint16_t ckc, uint16_t fkc, int align,
                              vbto_leaf_layout* o)
{
    memset(o, 0, sizeof(*o));
    o->headerBytes = 4;
    o->prefixOffset = 4;
    o->prefixBytes = 4;
    o->coarseOffset = o->prefixOffset + o->prefixBytes;
    const uint32_t cc = control_count_for_res(h->coarseResolution);
    const uint32_t fc = control_count_for_res(h->fineResolution);
    stream_layout(o->coarseOffset, cc, 8u, ckc, h->coarseCodec, align, &o->coarse);
    o->fineOffset = o->coarseOffset + o->coarse.totalBytes;
    if (h->mode == 3) {
        stream_layout(o->fineOffset, fc, 8u, fkc, h->fineCodec, align, &o->fine);
        o->totalBytes = o->fineOffset + o->fine.totalBytes;
    } else if (h->mode == 2) {
        memset(&o->fine, 0, sizeof(o->fine));
        o->fine.descriptorOffset = o->fineOffset;
        o->fine.descriptorBytes = 80u;
        o->fine.binIndexOffset = o->fineOffset + 80u;
        o->fine.binIndexBytes = 0u;
        o->fine.keyframeOffset = o->fineOffset + 80u;
        o->fine.keyframeBytes = fkc * vbto_keyframe_bytes(h->fineCodec);
        if (align) o->fine.keyframeBytes = align_up4(o->fine.keyframeBytes);
        o->fine.totalBytes = 80u + o->fine.keyframeBytes;
        o->totalBytes = o->fineOffset + o->fine.totalBytes;
    } else {
        o->totalBytes = o->coarseOffset + o->coarse.totalBytes;
    }
    if (h->mode == 0) {
        o->coarseOffset = 0;
        o->fineOffset = 0;
        memset(&o->coarse, 0, sizeof(o->coarse));
        memset(&o->fine, 0, sizeof(o->fine));
        o->totalBytes = o->headerBytes;
    }
}

/* render_temporal_payload.cpp:324-344 */
void vbto_unpack_shell_section(const uint8_t* d, uint64_t size, vbto_shell_section* s)
{
    memset(s, 0, sizeof(*s));
    if (d == NULL || size < 80) return;
    for (int g = 0; g < 8; ++g) s->shellMask[g] = rd64(d + 8 * g);
    for (int g = 0; g < 8; ++g) s->groupPrefix[g] = rd16(d + 64 + 2 * g);
}

uint16_t vbto_shell_active_count(const vbto_shell_section* s)
{
    uint16_t c = 0;
    for (int g = 0; g < 8; ++g) c = (uint16_t)(c + popcount64(s->shellMask[g]));
    return c;
}

/* render_temporal_payload.cpp:234-311 */
int vbto_parse_leaf(const uint8_t* leaf, size_t leafBytes, int align, vbto_leaf_view* o)
{
    memset(o, 0, sizeof(*o));
    if (leaf == NULL || leafBytes < 4) return fail("render temporal leaf bytes too small");
    vbto_unpack_header(rd32(leaf), &o->header);
    if (o->header.mode == 0) {
        vbto_compute_leaf_layout(&o->header, 0, 0, align, &o->layout);
        return VBTO_OK;
    }
    if (leafBytes < 8) return fail("render temporal leaf prefix missing");
    o->coarseKeyframeCount = rd16(leaf + 4);
    o->fineKeyframeCount = rd16(leaf + 6);
    vbto_compute_leaf_layout(&o->header, o->coarseKeyframeCount, o->fineKeyframeCount, align, &o->layout);
    if (o->layout.totalBytes > leafBytes) return fail("render temporal leaf shorter than computed layout");
    if (o->header.mode == 2) {
        const size_t occOff = o->layout.fine.descriptorOffset;
        const size_t occBytes = o->layout.fine.descriptorBytes;
        if (occBytes != 80 || occOff > leafBytes || occBytes > leafBytes - occOff)
            return fail("render temporal shell occupancy is out of bounds");
        vbto_shell_section s;
        vbto_unpack_shell_section(leaf + occOff, occBytes, &s);
        uint16_t expected = 0;
        for (int g = 0; g < 8; ++g) {
            if (s.groupPrefix[g] != expected) return fail("render temporal shell group prefix is inconsistent");
            expected = (uint16_t)(expected + popcount64(s.shellMask[g]));
        }
        if (expected == 0) return fail("render temporal shell mode contains no active voxels");
        size_t resOff = occOff + occBytes;
        if (align) resOff = (resOff + 3u) & ~(size_t)3u;
        const size_t descBytes = (size_t)expected * 4u;
        size_t binBytes = (size_t)expected * 4u;
        size_t keyBytes = (size_t)o->fineKeyframeCount * vbto_keyframe_bytes(o->header.fineCodec);
        if (align) {
            binBytes = (binBytes + 3u) & ~(size_t)3u;
            keyBytes = (keyBytes + 3u) & ~(size_t)3u;
        }
        const size_t resBytes = descBytes + binBytes + keyBytes;
        if (resOff > leafBytes || resBytes > leafBytes - resOff)
            return fail("render temporal shell residual stream is out of bounds");
        if (resOff + resBytes != leafBytes) return fail("render temporal shell leaf contains trailing bytes");
        o->shellActiveVoxelCount = expected;
        o->shellResidualOffset = (uint32_t)resOff;
        o->shellResidualBytes = (uint32_t)resBytes;
    }
    return VBTO_OK;
}

/* ------------------------------------------------------------------------- */
/* control decode: src/render_temporal_decode.cpp:69-131, 304-356             */
/* ------------------------------------------------------------------------- */
static int key_frame(const uint8_t* keyBase, uint32_t codec, uint32_t idx)
{
    const size_t off = (size_t)idx * vbto_keyframe_bytes(codec);
    return codec == 1u ? (int)rd16(keyBase + off) : (int)keyBase[off];
}

static float key_value(const uint8_t* keyBase, uint32_t codec, uint32_t idx)
{
    const size_t off = (size_t)idx * vbto_keyframe_bytes(codec);
    if (codec == 1u) {
        uint32_t b = rd32(keyBase + off + 4);
        float f;
        memcpy(&f, &b, 4);
        return f;
    }
    return vbto_half_bits_to_float(rd16(keyBase + off + 1));
}

/* decodeFromSequence, render_temporal_decode.cpp:79-131.  The caller guarantees (as
 * decodeRenderTemporalControlValue does) keyStart + keyCount <= keyframeCount,
 * so every read stays inside the keyframe section. */
static float decode_from_sequence(const uint8_t* keyBase, uint32_t codec, uint32_t keyStart,
                                  uint32_t keyCount, const uint8_t* bins, uint32_t B, int F,
                                  int frameIndex, vbto_trace* tr)
{
    if (keyCount == 0) {
        if (tr) { memset(tr, 0, sizeof(*tr)); tr->kind = 0; }
        return 0.0f;
    }
    int fmax = F - 1 > 0 ? F - 1 : 0;
    int frame = frameIndex < fmax ? frameIndex : fmax;
    if (frame < 0) frame = 0;
    const uint32_t last = keyCount - 1u;
    uint32_t bin = 0;
    if (B != 0) {
        const int64_t q = ((int64_t)frame * (int64_t)B) / (int64_t)(F > 1 ? F : 1);
        bin = (uint32_t)q < B - 1u ? (uint32_t)q : B - 1u;
    }
    uint32_t lo = 0;
    if (B != 0) lo = bins[bin] < last ? bins[bin] : last;
    int f0 = key_frame(keyBase, codec, keyStart + lo);
    if (f0 > frame) {
        lo = 0;
        f0 = key_frame(keyBase, codec, keyStart);
    }
    if (tr) { tr->bin = bin; tr->keyStart = keyStart; tr->keyCount = keyCount; }
    if (f0 >= frame || lo == last) {
        if (tr) { tr->lo = lo; tr->hi = lo; tr->f0 = f0; tr->f1 = f0; tr->kind = 1; }
        return key_value(keyBase, codec, keyStart + lo);
    }
    uint32_t hi = last;
    if (bin + 1u < B) {
        const uint32_t h = (uint32_t)bins[bin + 1u] + 1u;
        hi = h < last ? h : last;
    }
    if (hi < lo + 1u) hi = lo + 1u;
    int f1 = key_frame(keyBase, codec, keyStart + hi);
    if (f1 < frame) {
        hi = last;
        f1 = key_frame(keyBase, codec, keyStart + hi);
    }
    if (f1 <= frame) {
        if (tr) { tr->lo = lo; tr->hi = hi; tr->f0 = f0; tr->f1 = f1; tr->kind = 2; }
        return key_value(keyBase, codec, keyStart + hi);
    }
    while (hi > lo + 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        const int fm = key_frame(keyBase, codec, keyStart + mid);
        if (fm < frame) { lo = mid; f0 = fm; }
        else { hi = mid; f1 = fm; }
    }
    const float v1 = key_value(keyBase, codec, keyStart + hi);
    if (tr) { tr->lo = lo; tr->hi = hi; tr->f0 = f0; tr->f1 = f1; tr->kind = (f1 == frame) ? 2 : 3; }
    if (f1 == frame) return v1;
    const float v0 = key_value(keyBase, codec, keyStart + lo);
    const float alpha = (float)(frame - f0) / (float)(f1 - f0);
    return v0 + alpha * (v1 - v0);
}

int vbto_decode_control_value(const uint8_t* streamBase, uint32_t codec, uint16_t controlCount,
                              uint8_t timeBinCount, uint8_t maxBinLocalKeys, uint32_t keyframeCount,
                              uint32_t streamBytes, int align, int totalFrames, uint16_t controlIndex,
                              int frameIndex, float* out, vbto_trace* tr)
{
    (void)maxBinLocalKeys;
    if (controlIndex >= controlCount) return fail("render temporal control index out of range");
    const uint32_t descBytes = (uint32_t)controlCount * 4u;
    uint32_t binBytes = (uint32_t)controlCount * timeBinCount;
    uint32_t keyBytes = keyframeCount * vbto_keyframe_bytes(codec);
    if (align) {
        binBytes = align_up4(binBytes);
        keyBytes = align_up4(keyBytes);
    }
    const uint32_t total = descBytes + binBytes + keyBytes;
    if (total > streamBytes) return fail("render temporal stream decode out of bounds");
    if ((size_t)controlIndex * 4u + 4u > streamBytes) return fail("render temporal decode out of bounds");
    const uint8_t* d = streamBase + (size_t)controlIndex * 4u;
    const uint32_t keyStart = rd16(d);
    const uint32_t keyCount = d[2];
    if (keyCount == 0) {
        if (tr) memset(tr, 0, sizeof(*tr));
        *out = 0.0f;
        return VBTO_OK;
    }
    if (keyStart + keyCount > keyframeCount) return fail("render temporal keyframe range out of bounds");
    const uint8_t* binBase = streamBase + descBytes + (size_t)controlIndex * timeBinCount;
    const uint8_t* keyBase = streamBase + descBytes + binBytes;
    *out = decode_from_sequence(keyBase, codec, keyStart, keyCount, binBase, timeBinCount, totalFrames,
                                frameIndex, tr);
    return VBTO_OK;
}

/* render_temporal_decode.cpp:358-398 */
int vbto_decode_leaf_control_value(const uint8_t* leaf, const vbto_leaf_view* v, int useFine, int align,
                                   int totalFrames, uint16_t ci, int t, float* out, vbto_trace* tr)
{
    const vbto_header* h = &v->header;
    const uint16_t cc = (uint16_t)(h->coarseResolution * h->coarseResolution * h->coarseResolution);
    const uint16_t fc = h->fineResolution == 0
                            ? 0
                            : (uint16_t)(h->fineResolution * h->fineResolution * h->fineResolution);
    if (useFine)
        return vbto_decode_control_value(leaf + v->layout.fineOffset, h->fineCodec, fc, 8, 0,
                                         v->fineKeyframeCount, v->layout.fine.totalBytes, align,
                                         totalFrames, ci, t, out, tr);
    return vbto_decode_control_value(leaf + v->layout.coarseOffset, h->coarseCodec, cc, 8, 0,
                                     v->coarseKeyframeCount, v->layout.coarse.totalBytes, align, totalFrames,
                                     ci, t, out, tr);
}

/* render_temporal_decode.cpp:404-450 */
int vbto_decode_shell_voxel_value(const uint8_t* leaf, const vbto_leaf_view* v, int align, int totalFrames,
                                  uint16_t voxel, int t, float* out, uint32_t* descOut)
{
    *out = 0.0f;
    if (descOut) *descOut = 0xffffffffu;
    if (leaf == NULL || voxel >= 512) return VBTO_OK;
    if (v->header.mode != 2) return VBTO_OK;
    const uint32_t group = voxel >> 6, bit = voxel & 63u;
    vbto_shell_section s;
    vbto_unpack_shell_section(leaf + v->layout.fine.descriptorOffset, v->layout.fine.descriptorBytes, &s);
    const uint64_t mask = s.shellMask[group];
    if ((mask & (1ull << bit)) == 0ull) return VBTO_OK;
    const uint64_t before = mask & ((1ull << bit) - 1ull);
    const uint16_t desc = (uint16_t)(s.groupPrefix[group] + popcount64(before));
    if (descOut) *descOut = desc;
    if (v->shellActiveVoxelCount == 0 || v->fineKeyframeCount == 0) return VBTO_OK;
    return vbto_decode_control_value(leaf + v->shellResidualOffset, v->header.fineCodec,
                                     v->shellActiveVoxelCount, 4, 0, v->fineKeyframeCount,
                                     v->shellResidualBytes, align, totalFrames, desc, t, out, NULL);
}

/* ------------------------------------------------------------------------- */
/* container: render/src/vbt_file.cpp:104-170 (validateHeader), :174-275     */
/* ------------------------------------------------------------------------- */
int vbto_load_file_memory(const uint8_t* bytes, uint64_t size, vbto_file* o)
{
    memset(o, 0, sizeof(*o));
    const uint64_t kHeader = 52;
    if (size < kHeader) return fail("Invalid VBTPACK4 file size.");
    vbto_file_header* h = &o->header;
    uint32_t w[11];
    for (int i = 0; i < 11; ++i) w[i] = rd32(bytes + 8 + 4 * i);
    h->version = w[0]; h->width = w[1]; h->height = w[2]; h->depth = w[3]; h->frames = w[4];
    h->leafSize = w[5]; h->coarseResolution = w[6]; h->maxCoarseKeep = w[7]; h->leafCount = w[8];
    h->profileType = w[9]; h->coarseAcScaleCount = w[10];
    if (memcmp(bytes, "VBTPACK4", 8) != 0) return fail("Unsupported magic. Expected VBTPACK4.");
    if (h->version != 4) {
        snprintf(g_err, sizeof(g_err), "Unsupported VBTPACK4 version: %u", h->version);
        return VBTO_ERR;
    }
    if (!h->width || !h->height || !h->depth || !h->frames)
        return fail("VBTPACK4 dimensions and frame count must be non-zero.");
    if (h->leafSize != 8) return fail("VBTPACK4 v4 requires leafSize=8.");
    if (h->coarseResolution == 0 || h->coarseResolution > 8) return fail("Invalid VBTPACK4 coarse resolution.");
    if (h->profileType > 2) return fail("Invalid VBTPACK4 profile type.");
    if (h->maxCoarseKeep > h->frames || h->coarseAcScaleCount > 64 || h->coarseAcScaleCount > h->maxCoarseKeep)
        return fail("Invalid VBTPACK4 coarse temporal metadata.");
    const uint64_t lx = ((uint64_t)h->width + 7) / 8, ly = ((uint64_t)h->height + 7) / 8,
                   lz = ((uint64_t)h->depth + 7) / 8;
    const uint64_t expected = lx * ly * lz; /* cannot overflow 64 bits for 32-bit dims */
    if (expected > 0xffffffffull) return fail("VBTPACK4 leaf count overflows the v4 header.");
    if (h->leafCount != expected) return fail("VBTPACK4 leafCount does not match dimensions and leafSize.");
    const uint64_t minimum = kHeader + (uint64_t)h->coarseAcScaleCount * 4u + ((uint64_t)h->leafCount + 1u) * 4u;
    if (minimum > size) return fail("VBTPACK4 file is truncated before its offset table.");
    o->coarseAcScales = (const float*)(bytes + kHeader);
    o->offsetsWords = (const uint32_t*)(bytes + kHeader + (uint64_t)h->coarseAcScaleCount * 4u);
    if (o->offsetsWords[0] != 0) return fail("VBTPACK4 offset table must begin at word zero.");
    for (uint64_t i = 0; i < h->leafCount; ++i)
        if (o->offsetsWords[i] >= o->offsetsWords[i + 1])
            return fail("VBTPACK4 offsets must be strictly increasing for every leaf.");
    o->payloadWordCount = o->offsetsWords[h->leafCount];
    const uint64_t expectedBytes = minimum + (uint64_t)o->payloadWordCount * 4u;
    if (expectedBytes != size)
        return fail(expectedBytes > size ? "VBTPACK4 payload is truncated." : "VBTPACK4 file contains trailing bytes.");
    o->payloadWords = (const uint32_t*)(bytes + minimum);
    return VBTO_OK;
}

/* render/src/vbt_file.cpp:212-218 */
uint32_t vbto_read_packed_header_word(const vbto_file* f, uint32_t leaf)
{
    if (leaf >= f->header.leafCount) return 0;
    const uint32_t w = f->offsetsWords[leaf];
    if (w >= f->payloadWordCount) return 0;
    return f->payloadWords[w];
}

/* render/src/vbt_file.cpp:250-262 */
uint32_t vbto_leaf_index_for_voxel(const vbto_file_header* h, uint32_t x, uint32_t y, uint32_t z)
{
    if (!h->width || !h->height || !h->depth || !h->leafSize) return 0;
    const uint32_t s = h->leafSize;
    const uint32_t lcx = (h->width + s - 1) / s, lcy = (h->height + s - 1) / s, lcz = (h->depth + s - 1) / s;
    const uint32_t bx = (x / s) < lcx - 1 ? x / s : lcx - 1;
    const uint32_t by = (y / s) < lcy - 1 ? y / s : lcy - 1;
    const uint32_t bz = (z / s) < lcz - 1 ? z / s : lcz - 1;
    return (bz * lcy + by) * lcx + bx;
}

/* ------------------------------------------------------------------------- */
/* composition: 3D/render_temporal_vbt_to_vdb.cpp:101-289                    */
/* ------------------------------------------------------------------------- */
void vbto_compose_default(vbto_compose* c)
{
    c->signRule = VBTO_SIGN_NONE;
    c->background = 0.0f;
    c->shellEmptyScale = 1.0f;
    c->shellEmptyMax = NAN;
}

/* locateCoord, render_temporal_vbt_to_vdb.cpp:101-134.  res >= 2.  For
 * res < 2 the reference reads past its one-entry coordinate table (undefined);
 * this restatement defines it as the single control (i0 = i1 = 0, w = 0). */
static void locate(float p, int extent, int res, int* i0o, int* i1o, float* wo)
{
    if (res < 2) { *i0o = 0; *i1o = 0; *wo = 0.0f; return; }
    const float scale = 7.0f / (float)(res - 1);
    const float maxP = (float)(extent - 1 > 0 ? extent - 1 : 0);
    if (p < 0.0f) p = 0.0f;
    if (p > maxP) p = maxP;
    int i1 = 1;
    while (i1 < res - 1 && (float)i1 * scale < p) ++i1;
    const int i0 = i1 - 1 > 0 ? i1 - 1 : 0;
    const float c0 = (float)i0 * scale, c1 = (float)i1 * scale;
    *i0o = i0;
    *i1o = i1;
    *wo = (c1 > c0) ? ((p - c0) / (c1 - c0)) : 0.0f;
}

/* sampleControlGrid, render_temporal_vbt_to_vdb.cpp:136-167 */
static float sample_grid(const float* g, int res, int w, int h, int d, float fx, float fy, float fz)
{
    if (res < 1) return 0.0f; /* empty grid: the reference indexes an empty vector (undefined); defined as 0 */
    int x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    locate(fx, w, res, &x0, &x1, &tx);
    locate(fy, h, res, &y0, &y1, &ty);
    locate(fz, d, res, &z0, &z1, &tz);
#define AT(ix, iy, iz) g[((iz) * res + (iy)) * res + (ix)]
    const float c00 = AT(x0, y0, z0) * (1.0f - tx) + AT(x1, y0, z0) * tx;
    const float c01 = AT(x0, y0, z1) * (1.0f - tx) + AT(x1, y0, z1) * tx;
    const float c10 = AT(x0, y1, z0) * (1.0f - tx) + AT(x1, y1, z0) * tx;
    const float c11 = AT(x0, y1, z1) * (1.0f - tx) + AT(x1, y1, z1) * tx;
#undef AT
    const float c0 = c00 * (1.0f - ty) + c10 * ty;
    const float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}

static void leaf_extent(const vbto_file_header* h, uint32_t leaf, int* bx, int* by, int* bz, int* w, int* hh,
                        int* d)
{
    const uint32_t lcx = (h->width + 7) / 8, lcy = (h->height + 7) / 8;
    *bx = (int)(leaf % lcx);
    *by = (int)((leaf / lcx) % lcy);
    *bz = (int)(leaf / (lcx * lcy));
    const int W = (int)h->width, H = (int)h->height, D = (int)h->depth;
    *w = W - *bx * 8 < 8 ? W - *bx * 8 : 8;
    *hh = H - *by * 8 < 8 ? H - *by * 8 : 8;
    *d = D - *bz * 8 < 8 ? D - *bz * 8 : 8;
}

/* decodeLeafFrameValues (render_temporal_vbt_to_vdb.cpp:177-289) with the
 * sign-aware variants of render_smoke_vbt_bench_main.cpp:374-378 (CLAMP_ZERO)
 * and vbt_sample.comp:410-412 (BACKGROUND). */
int vbto_decode_render_leaf_frame(const vbto_file* f, uint32_t leaf, int t, const vbto_compose* c, float* out)
{
    int bx, by, bz, lw, lh, ld;
    leaf_extent(&f->header, leaf, &bx, &by, &bz, &lw, &lh, &ld);
    const uint32_t wb = f->offsetsWords[leaf], we = f->offsetsWords[leaf + 1];
    const uint8_t* base = (const uint8_t*)(f->payloadWords + wb);
    const size_t bytes = (size_t)(we - wb) * 4u;
    vbto_leaf_view v;
    if (vbto_parse_leaf(base, bytes, 1, &v) != VBTO_OK) return VBTO_ERR;
    const int n = lw * lh * ld;
    if (v.header.mode == 0) {
        for (int i = 0; i < n; ++i) out[i] = 0.0f;
        return VBTO_OK;
    }
    const int F = (int)f->header.frames;
    const int cr = v.header.coarseResolution, fr = v.header.fineResolution;
    const int cc = (uint16_t)(cr * cr * cr);
    float coarse[512];
    float fine[512];
    for (int i = 0; i < cc; ++i)
        if (vbto_decode_leaf_control_value(base, &v, 0, 1, F, (uint16_t)i, t, &coarse[i], NULL) != VBTO_OK)
            return VBTO_ERR;
    const int hasFine = fr > 0 && v.layout.fine.totalBytes > 0 && v.header.mode != 1;
    const int fc = hasFine ? (uint16_t)(fr * fr * fr) : 0;
    for (int i = 0; i < fc; ++i)
        if (vbto_decode_leaf_control_value(base, &v, 1, 1, F, (uint16_t)i, t, &fine[i], NULL) != VBTO_OK)
            return VBTO_ERR;
    const int hasShell = v.header.mode == 2;
    vbto_shell_section s;
    if (hasShell) vbto_unpack_shell_section(base + v.layout.fine.descriptorOffset, v.layout.fine.descriptorBytes, &s);
    for (int lz = 0; lz < ld; ++lz)
        for (int ly = 0; ly < lh; ++ly)
            for (int lx = 0; lx < lw; ++lx) {
                const uint16_t vox = (uint16_t)((lz * 8 + ly) * 8 + lx);
                float val = sample_grid(coarse, cr, lw, lh, ld, (float)lx, (float)ly, (float)lz);
                if (fc > 0) {
                    val += sample_grid(fine, fr, lw, lh, ld, (float)lx, (float)ly, (float)lz);
                } else if (hasShell) {
                    float r;
                    if (vbto_decode_shell_voxel_value(base, &v, 1, F, vox, t, &r, NULL) != VBTO_OK) return VBTO_ERR;
                    val += r;
                }
                if (hasShell) {
                    const int occ = (int)((s.shellMask[vox >> 6] >> (vox & 63)) & 1ull);
                    if (c->signRule == VBTO_SIGN_NONE) {
                        if (!isnan(c->shellEmptyMax) && c->shellEmptyScale < 1.0f && !occ && val > 0.0f &&
                            val <= c->shellEmptyMax)
                            val *= c->shellEmptyScale;
                    } else if (!occ && val < 0.0f) {
                        val = c->signRule == VBTO_SIGN_CLAMP_ZERO ? 0.0f : c->background;
                    }
                }
                out[(lz * lh + ly) * lw + lx] = val;
            }
    return VBTO_OK;
}

/* ------------------------------------------------------------------------- */
/* scientific route: render/src/vbt_file.cpp:220-248,                       */
/* render/src/scientific_decode.cpp:13-409                                   */
/* ------------------------------------------------------------------------- */
void vbto_decode_sci_header(uint32_t p, vbto_sci_header* d)
{
    memset(d, 0, sizeof(*d));
    d->mode = p & 3u;
    switch (d->mode) {
    case 0:
        d->coarseKeepMinus1 = (p >> 2) & 0x3Fu;
        d->coarseQuantClass = (p >> 8) & 7u;
        break;
    case 1:
        d->coarseKeepMinus1 = (p >> 2) & 0x3Fu;
        d->tierId = (p >> 8) & 3u;
        d->framesPerBinMinus1 = (p >> 10) & 0xFu;
        d->eventQuantCode = (p >> 14) & 3u;
        d->packedEventCount = (p >> 16) & 0x0FFFu;
        d->sparseSubtype = (p >> 28) & 0xFu;
        break;
    default:
        d->coarseKeepMinus1 = (p >> 2) & 0x3Fu;
        d->denseQuantBits = (p >> 8) & 0xFu;
        d->denseSubtype = (p >> 12) & 7u;
        d->denseTemporalKeepMinus1 = (p >> 15) & 0xFu;
        break;
    }
}

/* readPayloadByte/U16 (scientific_decode.cpp:42-59): bounds-checked against
 * the whole payload pool like std::vector::at. */
static int pbyte(const vbto_file* f, uint32_t off, uint32_t* out)
{
    if ((off >> 2) >= f->payloadWordCount) return fail("payload read out of range");
    *out = (f->payloadWords[off >> 2] >> ((off & 3u) * 8u)) & 0xFFu;
    return VBTO_OK;
}
static int pu16(const vbto_file* f, uint32_t off, uint32_t* out)
{
    uint32_t lo, hi;
    if (pbyte(f, off, &lo) || pbyte(f, off + 1u, &hi)) return VBTO_ERR;
    *out = (lo | (hi << 8)) & 0xFFFFu;
    return VBTO_OK;
}

/* scientific_decode.cpp:72-80 */
float vbto_dct_basis(int N, int index, int k)
{
    const double kPi = 3.14159265358979323846;
    const double invN = 1.0 / (double)N;
    const double alpha = (k == 0) ? sqrt(invN) : sqrt(2.0 * invN);
    return (float)(alpha * cos((kPi / (double)N) * ((double)index + 0.5) * (double)k));
}

/* decodeCoarseControlAt, scientific_decode.cpp:82-97 */
int vbto_sci_coarse_control(const vbto_file* f, uint32_t base, int keep, int ci, int t, float* out)
{
    const uint32_t cb = base + (uint32_t)(4 + ci * (keep + 1));
    uint32_t dc;
    if (pu16(f, cb, &dc)) return VBTO_ERR;
    double sum = (double)vbto_half_bits_to_float((uint16_t)dc) * (double)vbto_dct_basis((int)f->header.frames, t, 0);
    for (int k = 1; k < keep; ++k) {
        if ((uint32_t)(k - 1) >= f->header.coarseAcScaleCount) return fail("coarse AC scale index out of range");
        const float scale = f->coarseAcScales[k - 1];
        uint32_t b;
        if (pbyte(f, cb + 2u + (uint32_t)(k - 1), &b)) return VBTO_ERR;
        const float coeff = (float)(int8_t)(uint8_t)b * scale;
        const double term = (double)coeff * (double)vbto_dct_basis((int)f->header.frames, t, k);
        sum += term;
    }
    *out = (float)sum;
    return VBTO_OK;
}

/* sampleCoarseAt, scientific_decode.cpp:99-123 */
static float sci_sample_coarse(const float* ctrl, uint32_t x, uint32_t y, uint32_t z)
{
    const float fx = ((float)x / 7.0f) * 3.0f, fy = ((float)y / 7.0f) * 3.0f, fz = ((float)z / 7.0f) * 3.0f;
    int x0 = (int)floorf(fx), y0 = (int)floorf(fy), z0 = (int)floorf(fz);
    x0 = x0 < 0 ? 0 : (x0 > 2 ? 2 : x0);
    y0 = y0 < 0 ? 0 : (y0 > 2 ? 2 : y0);
    z0 = z0 < 0 ? 0 : (z0 > 2 ? 2 : z0);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const float tx = fx - (float)x0, ty = fy - (float)y0, tz = fz - (float)z0;
#define AT(cx, cy, cz) ctrl[((cz) * 4 + (cy)) * 4 + (cx)]
    const float c00 = AT(x0, y0, z0) * (1.0f - tx) + AT(x1, y0, z0) * tx;
    const float c01 = AT(x0, y0, z1) * (1.0f - tx) + AT(x1, y0, z1) * tx;
    const float c10 = AT(x0, y1, z0) * (1.0f - tx) + AT(x1, y1, z0) * tx;
    const float c11 = AT(x0, y1, z1) * (1.0f - tx) + AT(x1, y1, z1) * tx;
#undef AT
    const float c0 = c00 * (1.0f - ty) + c10 * ty;
    const float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}

/* decodeBfp, scientific_decode.cpp:61-70 */
static float decode_bfp(uint32_t q, float scale, int bits)
{
    if (bits <= 2) {
        static const float kLevels[4] = {-1.0f, -0.3333333f, 0.3333333f, 1.0f};
        return kLevels[q < 3 ? q : 3] * scale;
    }
    const int levels = (1 << bits) - 1;
    const float norm = ((float)q / (float)levels) * 2.0f - 1.0f;
    return norm * scale;
}

static int bfp_code(const vbto_file* f, uint32_t blockBase, int ci, int bits, uint32_t* q)
{
    const int bitOffset = ci * bits;
    const uint32_t byteIndex = blockBase + (uint32_t)(bitOffset / 8);
    const int shift = bitOffset % 8;
    uint32_t packed, hi;
    if (pbyte(f, byteIndex, &packed)) return VBTO_ERR;
    if (shift + bits > 8) {
        if (pbyte(f, byteIndex + 1u, &hi)) return VBTO_ERR;
        packed |= hi << 8;
    }
    *q = (packed >> shift) & ((1u << bits) - 1u);
    return VBTO_OK;
}

static void dense_coords(int R, uint32_t x, int* x0, float* tx)
{
    const float fx = ((float)x / 7.0f) * (float)(R - 1);
    int i = (int)floorf(fx);
    if (i < 0) i = 0;
    if (i > R - 2) i = R - 2;
    *x0 = i;
    *tx = fx - (float)i;
}

/* decodeDenseValueAt, scientific_decode.cpp:125-175 */
static int sci_dense_value(const vbto_file* f, uint32_t base, int keep, int R, int bits, uint32_t x, uint32_t y,
                           uint32_t z, uint32_t t, float* out)
{
    const int vpf = R * R * R;
    const int frameBytes = (vpf * bits + 7) / 8;
    const uint32_t scalesBase = base + (uint32_t)(4 + 64 * (keep + 1));
    const uint32_t dataBase = scalesBase + f->header.frames * 2u;
    uint32_t sb;
    if (pu16(f, scalesBase + t * 2u, &sb)) return VBTO_ERR;
    const float scale = vbto_half_bits_to_float((uint16_t)sb);
    int x0, y0, z0;
    float tx, ty, tz;
    dense_coords(R, x, &x0, &tx);
    dense_coords(R, y, &y0, &ty);
    dense_coords(R, z, &z0, &tz);
    const uint32_t frameBase = dataBase + (uint32_t)(t * (uint32_t)frameBytes);
    float s[8];
    for (int i = 0; i < 8; ++i) {
        const int cx = x0 + (i & 1), cy = y0 + ((i >> 1) & 1), cz = z0 + ((i >> 2) & 1);
        uint32_t q;
        if (bfp_code(f, frameBase, (cz * R + cy) * R + cx, bits, &q)) return VBTO_ERR;
        s[i] = decode_bfp(q, scale, bits);
    }
    /* s index: bit0 x, bit1 y, bit2 z */
    const float c00 = s[0] * (1.0f - tx) + s[1] * tx;
    const float c01 = s[4] * (1.0f - tx) + s[5] * tx;
    const float c10 = s[2] * (1.0f - tx) + s[3] * tx;
    const float c11 = s[6] * (1.0f - tx) + s[7] * tx;
    const float c0 = c00 * (1.0f - ty) + c10 * ty;
    const float c1 = c01 * (1.0f - ty) + c11 * ty;
    *out = c0 * (1.0f - tz) + c1 * tz;
    return VBTO_OK;
}

/* decodeDenseTemporalBasisAt, scientific_decode.cpp:177-233 */
static int sci_dense_basis(const vbto_file* f, uint32_t base, int keep, int R, int bits, int tk, uint32_t x,
                           uint32_t y, uint32_t z, uint32_t t, float* out)
{
    *out = 0.0f;
    if (tk <= 0 || bits <= 0) return VBTO_OK;
    const int vpb = R * R * R;
    const int basisBytes = (vpb * bits + 7) / 8;
    const uint32_t scalesBase = base + (uint32_t)(4 + 64 * (keep + 1));
    const uint32_t dataBase = scalesBase + (uint32_t)(tk * 2);
    int x0, y0, z0;
    float tx, ty, tz;
    dense_coords(R, x, &x0, &tx);
    dense_coords(R, y, &y0, &ty);
    dense_coords(R, z, &z0, &tz);
    float residual = 0.0f;
    for (int k = 0; k < tk; ++k) {
        uint32_t sb;
        if (pu16(f, scalesBase + (uint32_t)(k * 2), &sb)) return VBTO_ERR;
        const float scale = vbto_half_bits_to_float((uint16_t)sb);
        const uint32_t bb = dataBase + (uint32_t)(k * basisBytes);
        float s[8];
        for (int i = 0; i < 8; ++i) {
            const int cx = x0 + (i & 1), cy = y0 + ((i >> 1) & 1), cz = z0 + ((i >> 2) & 1);
            uint32_t q;
            if (bfp_code(f, bb, (cz * R + cy) * R + cx, bits, &q)) return VBTO_ERR;
            s[i] = decode_bfp(q, scale, bits);
        }
        const float c00 = s[0] * (1.0f - tx) + s[1] * tx;
        const float c01 = s[4] * (1.0f - tx) + s[5] * tx;
        const float c10 = s[2] * (1.0f - tx) + s[3] * tx;
        const float c11 = s[6] * (1.0f - tx) + s[7] * tx;
        const float c0 = c00 * (1.0f - ty) + c10 * ty;
        const float c1 = c01 * (1.0f - ty) + c11 * ty;
        const float tri = c0 * (1.0f - tz) + c1 * tz;
        residual += tri * vbto_dct_basis((int)f->header.frames, (int)t, k);
    }
    *out = residual;
    return VBTO_OK;
}

/* decodeSparseResidualAt, scientific_decode.cpp:235-295 */
static int sci_sparse(const vbto_file* f, uint32_t base, const vbto_sci_header* d, int keep, uint32_t idx,
                      uint32_t t, float* out)
{
    *out = 0.0f;
    const uint32_t count = d->packedEventCount;
    if (count == 0) return VBTO_OK;
    const uint32_t cap = d->tierId == 0 ? 64u : d->tierId == 1 ? 256u : d->tierId == 2 ? 768u : 1024u;
    const uint32_t fpb = d->framesPerBinMinus1 + 1u;
    const uint32_t bin = t / fpb < 15u ? t / fpb : 15u;
    const uint8_t localT = (uint8_t)(t - bin * fpb);
    const uint32_t sparseBase = base + 4u + (uint32_t)(64 * (keep + 1));
    if ((sparseBase >> 2) >= f->payloadWordCount) return fail("payload read out of range");
    float eventScale;
    memcpy(&eventScale, &f->payloadWords[sparseBase >> 2], 4);
    if (!(eventScale > 0.0f)) return VBTO_OK;
    const uint32_t maskOff = sparseBase + 4u + ((idx >> 5) * 4u);
    if ((maskOff >> 2) >= f->payloadWordCount) return fail("payload read out of range");
    if ((f->payloadWords[maskOff >> 2] & (1u << (idx & 31u))) == 0u) return VBTO_OK;
    const uint32_t binsBase = sparseBase + 68u;
    uint32_t start, end;
    if (pu16(f, binsBase + bin * 2u, &start)) return VBTO_ERR;
    if (bin + 1u < 16u) {
        if (pu16(f, binsBase + (bin + 1u) * 2u, &end)) return VBTO_ERR;
    } else {
        end = count & 0xFFFFu;
    }
    if (start >= end) return VBTO_OK;
    const uint32_t coordsBase = binsBase + 32u;
    const uint32_t residualBase = coordsBase + cap * 2u;
    const uint16_t target = (uint16_t)(((uint32_t)localT << 9) | (idx & 0x1FFu));
    int left = (int)start, right = (int)end - 1;
    while (left <= right) {
        const int mid = (left + right) >> 1;
        uint32_t key;
        if (pu16(f, coordsBase + (uint32_t)mid * 2u, &key)) return VBTO_ERR;
        if (key == target) {
            uint32_t byte;
            if (pbyte(f, residualBase + (uint32_t)(mid >> 1), &byte)) return VBTO_ERR;
            const uint32_t nib = (mid & 1) == 0 ? (byte & 0xFu) : ((byte >> 4) & 0xFu);
            int q = (int)nib;
            if (q & 8) q -= 16;
            *out = (float)q * eventScale;
            return VBTO_OK;
        }
        if (key < target) left = mid + 1;
        else right = mid - 1;
    }
    return VBTO_OK;
}

/* ---- dense subtypes 2 and 3: the encoder's tile / patch temporal bases ----
 * The reference's file decoder reads these subtypes as subtype 0; the encoder
 * reconstructs them from its in-memory encoding.  The encoding's fields come
 * from the block the encoder writes after the coarse coefficients
 * (appendPackedDenseTileTemporalBasis / appendPackedDensePatchTemporalBasis,
 * src/spatial_first_encoder.cpp:426-458): tile = u64 active mask, u16 tile
 * size, u16 active count; patch = u64 active mask, u16 patch size, u16 local
 * resolution, u16 active count; then, when the value count is positive,
 * temporalKeep half scales and temporalKeep packed bases of
 * basisBytes = (max(1, values) * bits + 7) / 8 bytes, values = count * cells
 * (encodeDenseTileTemporalBasis :1755-1773, encodeDensePatchTemporalBasis
 * :1850-1863).  temporalKeep and bits are the leaf header's (:225-248). */
typedef struct {
    int patch;
    uint64_t mask;
    uint32_t size, localRes, count, tk, bits;
    uint64_t values, basisBytes, scalesOff, dataOff, end; /* offsets: payload pool bytes */
} basis_block;

static int popcount_u64(uint64_t v)
{
    int n = 0;
    for (; v; v &= v - 1u) ++n;
    return n;
}

static int basis_block_read(const vbto_file* f, uint32_t base, const vbto_sci_header* dh, basis_block* b)
{
    const vbto_sci_header d = *dh;
    memset(b, 0, sizeof(*b));
    if (d.mode < 2 || d.denseSubtype < 2 || d.denseSubtype > 3) return fail("not a tile / patch basis leaf");
    const uint64_t pool = (uint64_t)f->payloadWordCount * 4u;
    const uint64_t blk = (uint64_t)base + 4u + 64u * (uint64_t)(d.coarseKeepMinus1 + 2u);
    b->patch = d.denseSubtype == 3;
    const uint64_t hdr = b->patch ? 14u : 12u;
    if (blk + hdr > pool) return fail("dense basis block out of range");
    uint32_t w[7];
    for (int i = 0; i < 7; ++i)
        if (pu16(f, (uint32_t)(blk + 2u * (uint64_t)i), &w[i])) return VBTO_ERR;
    b->mask = (uint64_t)w[0] | ((uint64_t)w[1] << 16) | ((uint64_t)w[2] << 32) | ((uint64_t)w[3] << 48);
    b->size = w[4];
    b->localRes = b->patch ? w[5] : 0u;
    b->count = b->patch ? w[6] : w[5];
    b->tk = d.denseTemporalKeepMinus1 + 1u;
    b->bits = d.denseQuantBits;
    if ((uint32_t)popcount_u64(b->mask) != b->count) return fail("dense basis block: active count does not match its mask");
    if (!b->patch && b->size > 8u) return fail("dense tile basis: tile size above the leaf size");
    if (b->patch && b->count && (b->size == 1u || b->localRes == 1u))
        return fail("dense patch basis: patch size or local resolution of 1");
    const uint64_t cells = b->patch ? (uint64_t)b->localRes * b->localRes * b->localRes
                                    : (uint64_t)b->size * b->size * b->size;
    b->values = (uint64_t)b->count * cells;
    b->basisBytes = ((b->values > 0u ? b->values : 1u) * b->bits + 7u) / 8u;
    b->scalesOff = blk + hdr;
    b->dataOff = b->scalesOff + 2u * b->tk;
    b->end = b->values > 0u ? b->dataOff + (uint64_t)b->tk * b->basisBytes : blk + hdr;
    if (b->end > pool) return fail("dense basis block out of range");
    return VBTO_OK;
}

int vbto_sci_basis_extent(const vbto_file* f, uint32_t base, uint32_t headerWord, uint64_t* end)
{
    basis_block b;
    vbto_sci_header d;
    vbto_decode_sci_header(headerWord, &d);
    if (basis_block_read(f, base, &d, &b)) return VBTO_ERR;
    *end = b.end - base;
    return VBTO_OK;
}

/* decodeDenseTileTemporalBasisValueAt (spatial_first_encoder.cpp:2078-2092)
 * and its patch twin (:2119-2133): the code keeps at most 8 bits (the mask is
 * a uint8_t) of a value read from at most two bytes */
static int basis_value(const vbto_file* f, const basis_block* b, uint64_t ci, uint32_t k, float* out)
{
    const uint64_t bitOffset = ci * b->bits;
    const uint64_t byteIndex = b->dataOff + (uint64_t)k * b->basisBytes + bitOffset / 8u;
    const uint32_t shift = (uint32_t)(bitOffset % 8u);
    uint32_t packed, hi, sb;
    if (pbyte(f, (uint32_t)byteIndex, &packed)) return VBTO_ERR;
    if (shift + b->bits > 8u) {
        if (pbyte(f, (uint32_t)byteIndex + 1u, &hi)) return VBTO_ERR;
        packed |= hi << 8;
    }
    const uint32_t q = (packed >> shift) & (((1u << b->bits) - 1u) & 0xFFu);
    if (pu16(f, (uint32_t)(b->scalesOff + 2u * k), &sb)) return VBTO_ERR;
    *out = decode_bfp(q, vbto_half_bits_to_float((uint16_t)sb), (int)b->bits);
    return VBTO_OK;
}

/* decodeDenseTileTemporalBasisAt, spatial_first_encoder.cpp:2094-2114 */
static int tile_basis(const vbto_file* f, const basis_block* b, uint32_t x, uint32_t y, uint32_t z, uint32_t t,
                      float* out)
{
    *out = 0.0f;
    if (b->count == 0u || b->size == 0u || b->values == 0u || b->basisBytes == 0u) return VBTO_OK;
    const uint32_t s = b->size, tpa = 8u / s > 1u ? 8u / s : 1u;
    const uint32_t tileIdx = ((z / s) * tpa + y / s) * tpa + x / s;
    const uint64_t bit = tileIdx < 64u ? (1ull << tileIdx) : 0ull;
    if (bit == 0u || (b->mask & bit) == 0u) return VBTO_OK;
    const uint64_t ci = (uint64_t)popcount_u64(b->mask & (bit - 1u)) * s * s * s + ((z % s) * s + y % s) * s + x % s;
    float residual = 0.0f;
    for (uint32_t k = 0; k < b->tk; ++k) {
        float v;
        if (basis_value(f, b, ci, k, &v)) return VBTO_ERR;
        residual += v * vbto_dct_basis((int)f->header.frames, (int)t, (int)k);
    }
    *out = residual;
    return VBTO_OK;
}

static void patch_coord(uint32_t local, uint32_t size, uint32_t res, int* i0, float* w)
{
    const float fx = ((float)local / (float)(size - 1u)) * (float)(res - 1u);
    int i = (int)floorf(fx);
    if (i < 0) i = 0;
    if (i > (int)res - 2) i = (int)res - 2;
    *i0 = i;
    *w = fx - (float)i;
}

/* decodeDensePatchTemporalBasisAt, spatial_first_encoder.cpp:2135-2176 */
static int patch_basis(const vbto_file* f, const basis_block* b, uint32_t x, uint32_t y, uint32_t z, uint32_t t,
                       float* out)
{
    *out = 0.0f;
    if (b->count == 0u || b->size == 0u || b->localRes == 0u || b->values == 0u || b->basisBytes == 0u)
        return VBTO_OK;
    const uint32_t s = b->size, R = b->localRes;
    const uint32_t patchIdx = ((z / s) * 2u + y / s) * 2u + x / s;
    const uint64_t bit = patchIdx < 64u ? (1ull << patchIdx) : 0ull;
    if (bit == 0u || (b->mask & bit) == 0u) return VBTO_OK;
    const uint64_t first = (uint64_t)popcount_u64(b->mask & (bit - 1u)) * R * R * R;
    int x0, y0, z0;
    float tx, ty, tz;
    patch_coord(x % s, s, R, &x0, &tx);
    patch_coord(y % s, s, R, &y0, &ty);
    patch_coord(z % s, s, R, &z0, &tz);
    float residual = 0.0f;
    for (uint32_t k = 0; k < b->tk; ++k) {
        float c[8]; /* bit0 x, bit1 y, bit2 z */
        for (int i = 0; i < 8; ++i) {
            const uint64_t cx = (uint64_t)(x0 + (i & 1)), cy = (uint64_t)(y0 + ((i >> 1) & 1)),
                           cz = (uint64_t)(z0 + ((i >> 2) & 1));
            if (basis_value(f, b, first + (cz * R + cy) * R + cx, k, &c[i])) return VBTO_ERR;
        }
        const float c00 = c[0] * (1.0f - tx) + c[1] * tx;
        const float c01 = c[4] * (1.0f - tx) + c[5] * tx;
        const float c10 = c[2] * (1.0f - tx) + c[3] * tx;
        const float c11 = c[6] * (1.0f - tx) + c[7] * tx;
        const float c0 = c00 * (1.0f - ty) + c10 * ty;
        const float c1 = c01 * (1.0f - ty) + c11 * ty;
        residual += (c0 * (1.0f - tz) + c1 * tz) * vbto_dct_basis((int)f->header.frames, (int)t, (int)k);
    }
    *out = residual;
    return VBTO_OK;
}

static int basis_residual(const vbto_file* f, uint32_t base, const vbto_sci_header* d, uint32_t lx, uint32_t ly,
                          uint32_t lz, uint32_t t, float* out)
{
    basis_block b;
    *out = 0.0f;
    if (basis_block_read(f, base, d, &b)) return VBTO_ERR;
    return b.patch ? patch_basis(f, &b, lx, ly, lz, t, out) : tile_basis(f, &b, lx, ly, lz, t, out);
}

int vbto_sci_basis_residual(const vbto_file* f, uint32_t base, uint32_t headerWord, uint32_t lx, uint32_t ly,
                            uint32_t lz, uint32_t t, float* out)
{
    vbto_sci_header d;
    vbto_decode_sci_header(headerWord, &d);
    return basis_residual(f, base, &d, lx, ly, lz, t, out);
}

/* decodeScientificResidualAt, scientific_decode.cpp:297-331; with
 * f->sciDenseBasis set, dense subtypes 2 and 3 go to vbto_sci_basis_residual */
static int sci_residual(const vbto_file* f, uint32_t base, const vbto_sci_header* d, int keep, uint32_t idx,
                        uint32_t lx, uint32_t ly, uint32_t lz, uint32_t t, float* out)
{
    switch (d->mode) {
    case 0: *out = 0.0f; return VBTO_OK;
    case 1: return sci_sparse(f, base, d, keep, idx, t, out);
    default: {
        if (f->sciDenseBasis && (d->denseSubtype == 2 || d->denseSubtype == 3))
            return basis_residual(f, base, d, lx, ly, lz, t, out);
        const int R = d->mode == 2 ? 3 : 4;
        if (d->denseSubtype == 1)
            return sci_dense_basis(f, base, keep, R, (int)d->denseQuantBits, (int)d->denseTemporalKeepMinus1 + 1,
                                   lx, ly, lz, t, out);
        return sci_dense_value(f, base, keep, R, (int)d->denseQuantBits, lx, ly, lz, t, out);
    }
    }
}

/* decodeScientificValueAtCpu, scientific_decode.cpp:335-359 */
int vbto_sci_value_at(const vbto_file* f, uint32_t x, uint32_t y, uint32_t z, uint32_t t, uint32_t leaf, float* out)
{
    vbto_sci_header d;
    vbto_decode_sci_header(vbto_read_packed_header_word(f, leaf), &d);
    if (leaf >= f->header.leafCount) return fail("leaf index out of range");
    const uint32_t base = f->offsetsWords[leaf] * 4u;
    const uint32_t lx = x % 8u, ly = y % 8u, lz = z % 8u;
    const uint32_t lidx = (lz * 8u + ly) * 8u + lx;
    const int keep = (int)d.coarseKeepMinus1 + 1;
    float ctrl[64];
    for (int i = 0; i < 64; ++i)
        if (vbto_sci_coarse_control(f, base, keep, i, (int)t, &ctrl[i])) return VBTO_ERR;
    const float coarse = sci_sample_coarse(ctrl, lx, ly, lz);
    float r;
    if (sci_residual(f, base, &d, keep, lidx, lx, ly, lz, t, &r)) return VBTO_ERR;
    *out = coarse + r;
    return VBTO_OK;
}

/* one leaf of reconstructScientificFrameCpu, render/src/scientific_decode.cpp:361-409 */
int vbto_sci_leaf_frame(const vbto_file* f, uint32_t leaf, uint32_t t, float* out)
{
    int bx, by, bz, lw, lh, ld;
    leaf_extent(&f->header, leaf, &bx, &by, &bz, &lw, &lh, &ld);
    vbto_sci_header d;
    vbto_decode_sci_header(vbto_read_packed_header_word(f, leaf), &d);
    const uint32_t base = f->offsetsWords[leaf] * 4u;
    const int keep = (int)d.coarseKeepMinus1 + 1;
    float ctrl[64];
    for (int i = 0; i < 64; ++i)
        if (vbto_sci_coarse_control(f, base, keep, i, (int)t, &ctrl[i])) return VBTO_ERR;
    for (int lz = 0; lz < ld; ++lz)
        for (int ly = 0; ly < lh; ++ly)
            for (int lx = 0; lx < lw; ++lx) {
                const uint32_t idx = (uint32_t)((lz * 8 + ly) * 8 + lx);
                float r;
                if (sci_residual(f, base, &d, keep, idx, (uint32_t)lx, (uint32_t)ly, (uint32_t)lz, t, &r))
                    return VBTO_ERR;
                out[(lz * lh + ly) * lw + lx] = sci_sample_coarse(ctrl, (uint32_t)lx, (uint32_t)ly, (uint32_t)lz) + r;
            }
    return VBTO_OK;
}

/* ------------------------------------------------------------------------- */
/* whole frame / queries                                                     */
/* ------------------------------------------------------------------------- */
static int leaf_frame_any(const vbto_file* f, uint32_t leaf, int t, const vbto_compose* c, float* out)
{
    if (f->header.profileType == 2) {
        if (t < 0) return fail("scientific frame index out of range");
        return vbto_sci_leaf_frame(f, leaf, (uint32_t)t, out);
    }
    return vbto_decode_render_leaf_frame(f, leaf, t, c, out);
}

/* decodeFrameToDense (render/src/render_smoke_vbt_bench_main.cpp:389-455) and
 * the leaf loop of render_temporal_vbt_to_vdb.cpp:334-382 / the serial loop of
 * reconstructScientificFrameCpu.  With nthreads > 1 the independent leaves are
 * split over OpenMP threads (the reference's only parallel decode pattern,
 * render_smoke_vbt_bench_main.cpp:655-667). */
int vbto_decode_frame_dense(const vbto_file* f, int t, const vbto_compose* c, float* out, int nthreads)
{
    const vbto_file_header* h = &f->header;
    const int64_t L = h->leafCount;
    int err = 0;
    char msg[256] = {0};
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 1 ? nthreads : 1)
#endif
    for (int64_t l = 0; l < L; ++l) {
        if (err) continue;
        float buf[512];
        int bx, by, bz, lw, lh, ld;
        leaf_extent(h, (uint32_t)l, &bx, &by, &bz, &lw, &lh, &ld);
        if (leaf_frame_any(f, (uint32_t)l, t, c, buf) != VBTO_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                if (!err) { err = 1; snprintf(msg, sizeof(msg), "%s", g_err); }
            }
            continue;
        }
        for (int lz = 0; lz < ld; ++lz)
            for (int ly = 0; ly < lh; ++ly) {
                const size_t row = ((size_t)(bz * 8 + lz) * h->height + (size_t)(by * 8 + ly)) * h->width + (size_t)bx * 8;
                memcpy(out + row, buf + (lz * lh + ly) * lw, sizeof(float) * (size_t)lw);
            }
    }
    if (err) return fail(msg);
    return VBTO_OK;
}

/* Render point value at an integer voxel: samplePackedLeafGridValue /
 * samplePackedValueAtWorld (3D/render_temporal_vbt_direct_preview.cpp:161-323)
 * at integer positions -- only the 8 coarse (+8 fine) controls or the 1 shell
 * control the voxel needs are decoded, and the arithmetic is the same as the
 * whole-frame decode, so the value equals the frame value at that voxel. */
static int point_grid(const uint8_t* base, const vbto_leaf_view* v, int useFine, int res, int F, int t, int lw,
                      int lh, int ld, int lx, int ly, int lz, float* out)
{
    int x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    if (res < 1) { *out = 0.0f; return VBTO_OK; }
    locate((float)lx, lw, res, &x0, &x1, &tx);
    locate((float)ly, lh, res, &y0, &y1, &ty);
    locate((float)lz, ld, res, &z0, &z1, &tz);
    float s[8];
    for (int i = 0; i < 8; ++i) {
        const int cx = (i & 1) ? x1 : x0, cy = (i & 2) ? y1 : y0, cz = (i & 4) ? z1 : z0;
        if (vbto_decode_leaf_control_value(base, v, useFine, 1, F, (uint16_t)((cz * res + cy) * res + cx), t, &s[i],
                                           NULL))
            return VBTO_ERR;
    }
    const float c00 = s[0] * (1.0f - tx) + s[1] * tx;
    const float c01 = s[4] * (1.0f - tx) + s[5] * tx;
    const float c10 = s[2] * (1.0f - tx) + s[3] * tx;
    const float c11 = s[6] * (1.0f - tx) + s[7] * tx;
    const float c0 = c00 * (1.0f - ty) + c10 * ty;
    const float c1 = c01 * (1.0f - ty) + c11 * ty;
    *out = c0 * (1.0f - tz) + c1 * tz;
    return VBTO_OK;
}

static int render_point(const vbto_file* f, uint32_t x, uint32_t y, uint32_t z, int t, const vbto_compose* c,
                        float* out)
{
    const uint32_t leaf = vbto_leaf_index_for_voxel(&f->header, x, y, z);
    int bx, by, bz, lw, lh, ld;
    leaf_extent(&f->header, leaf, &bx, &by, &bz, &lw, &lh, &ld);
    const uint32_t wb = f->offsetsWords[leaf], we = f->offsetsWords[leaf + 1];
    const uint8_t* base = (const uint8_t*)(f->payloadWords + wb);
    vbto_leaf_view v;
    if (vbto_parse_leaf(base, (size_t)(we - wb) * 4u, 1, &v) != VBTO_OK) return VBTO_ERR;
    if (v.header.mode == 0) { *out = 0.0f; return VBTO_OK; }
    const int F = (int)f->header.frames;
    const int lx = (int)x - bx * 8, ly = (int)y - by * 8, lz = (int)z - bz * 8;
    float val;
    if (point_grid(base, &v, 0, v.header.coarseResolution, F, t, lw, lh, ld, lx, ly, lz, &val)) return VBTO_ERR;
    const int fr = v.header.fineResolution;
    const int hasFine = fr > 0 && v.layout.fine.totalBytes > 0 && v.header.mode != 1;
    const int hasShell = v.header.mode == 2;
    const uint16_t vox = (uint16_t)((lz * 8 + ly) * 8 + lx);
    if (hasFine) {
        float fv;
        if (point_grid(base, &v, 1, fr, F, t, lw, lh, ld, lx, ly, lz, &fv)) return VBTO_ERR;
        val += fv;
    } else if (hasShell) {
        float r;
        if (vbto_decode_shell_voxel_value(base, &v, 1, F, vox, t, &r, NULL)) return VBTO_ERR;
        val += r;
    }
    if (hasShell) {
        vbto_shell_section s;
        vbto_unpack_shell_section(base + v.layout.fine.descriptorOffset, v.layout.fine.descriptorBytes, &s);
        const int occ = (int)((s.shellMask[vox >> 6] >> (vox & 63)) & 1ull);
        if (c->signRule == VBTO_SIGN_NONE) {
            if (!isnan(c->shellEmptyMax) && c->shellEmptyScale < 1.0f && !occ && val > 0.0f && val <= c->shellEmptyMax)
                val *= c->shellEmptyScale;
        } else if (!occ && val < 0.0f) {
            val = c->signRule == VBTO_SIGN_CLAMP_ZERO ? 0.0f : c->background;
        }
    }
    *out = val;
    return VBTO_OK;
}

int vbto_query_points(const vbto_file* f, const uint32_t* q, uint64_t n, const vbto_compose* c, float* out,
                      int nthreads)
{
    int err = 0;
    char msg[256] = {0};
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 1 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const uint32_t x = q[4 * i], y = q[4 * i + 1], z = q[4 * i + 2], t = q[4 * i + 3];
        int rc = VBTO_OK;
        float v = NAN;
        if (x >= f->header.width || y >= f->header.height || z >= f->header.depth) {
            rc = fail("query voxel outside the volume");
        } else if (f->header.profileType == 2) {
            if (t >= f->header.frames) rc = fail("scientific query frame out of range");
            else rc = vbto_sci_value_at(f, x, y, z, t, vbto_leaf_index_for_voxel(&f->header, x, y, z), &v);
        } else {
            rc = render_point(f, x, y, z, (int)(t > 0x7fffffffu ? 0x7fffffff : t), c, &v);
        }
        if (rc != VBTO_OK) {
            v = NAN;
#ifdef _OPENMP
#pragma omp critical
#endif
            {
                if (!err) { err = 1; snprintf(msg, sizeof(msg), "%s", g_err); }
            }
        }
        out[i] = v;
    }
    if (err) return fail(msg);
    return VBTO_OK;
}

/* ------------------------------------------------------------------------- */
/* page table + caches                                                        */
/* ------------------------------------------------------------------------- */
uint32_t vbto_page_table(const vbto_file* f, uint32_t* compact)
{
    uint32_t n = 0;
    for (uint32_t l = 0; l < f->header.leafCount; ++l) {
        /* scientific leaves have no EMPTY mode (mode 0 is CoarseOnly) */
        const uint32_t active = f->header.profileType == 2 || (f->payloadWords[f->offsetsWords[l]] & 3u) != 0u;
        compact[l] = active ? n++ : 0xffffffffu;
    }
    return n;
}

uint32_t vbto_leaf_cache_count(const vbto_file* f, uint32_t leaf)
{
    if (f->header.profileType == 2) {
        int bx, by, bz, lw, lh, ld;
        leaf_extent(&f->header, leaf, &bx, &by, &bz, &lw, &lh, &ld);
        return (uint32_t)(lw * lh * ld);
    }
    const uint32_t wb = f->offsetsWords[leaf], we = f->offsetsWords[leaf + 1];
    vbto_leaf_view v;
    if (vbto_parse_leaf((const uint8_t*)(f->payloadWords + wb), (size_t)(we - wb) * 4u, 1, &v) != VBTO_OK) return 0;
    if (v.header.mode == 0) return 0;
    const uint32_t cr = v.header.coarseResolution, fr = v.header.fineResolution;
    uint32_t n = (uint16_t)(cr * cr * cr);
    if (v.header.mode == 3) n += fr == 0 ? 0 : (uint16_t)(fr * fr * fr);
    if (v.header.mode == 2) n += v.shellActiveVoxelCount;
    return n;
}

/* cached control values of one active render leaf: coarse slots then fine6 or
 * shell slots (vbt_cache.comp:193-236 slot assignment, CPU arithmetic). */
static int render_leaf_controls(const vbto_file* f, uint32_t leaf, int t, float* slots, uint32_t* nOut)
{
    const uint32_t wb = f->offsetsWords[leaf], we = f->offsetsWords[leaf + 1];
    const uint8_t* base = (const uint8_t*)(f->payloadWords + wb);
    vbto_leaf_view v;
    if (vbto_parse_leaf(base, (size_t)(we - wb) * 4u, 1, &v) != VBTO_OK) return VBTO_ERR;
    const int F = (int)f->header.frames;
    const uint32_t cr = v.header.coarseResolution, fr = v.header.fineResolution;
    const uint32_t cc = (uint16_t)(cr * cr * cr);
    uint32_t n = 0;
    for (uint32_t i = 0; i < cc; ++i)
        if (vbto_decode_leaf_control_value(base, &v, 0, 1, F, (uint16_t)i, t, &slots[n++], NULL)) return VBTO_ERR;
    if (v.header.mode == 3) {
        const uint32_t fc = fr == 0 ? 0 : (uint16_t)(fr * fr * fr);
        for (uint32_t i = 0; i < fc; ++i)
            if (vbto_decode_leaf_control_value(base, &v, 1, 1, F, (uint16_t)i, t, &slots[n++], NULL)) return VBTO_ERR;
    } else if (v.header.mode == 2) {
        for (uint32_t i = 0; i < v.shellActiveVoxelCount; ++i)
            if (vbto_decode_control_value(base + v.shellResidualOffset, v.header.fineCodec, v.shellActiveVoxelCount, 4,
                                          0, v.fineKeyframeCount, v.shellResidualBytes, 1, F, (uint16_t)i, t,
                                          &slots[n++], NULL))
                return VBTO_ERR;
    }
    *nOut = n;
    return VBTO_OK;
}

int vbto_frame_cache_576(const vbto_file* f, int t, float* out)
{
    if (f->header.profileType == 2) return fail("stride-576 cache is defined for render-temporal assets only");
    uint32_t a = 0;
    float slots[1024];
    for (uint32_t l = 0; l < f->header.leafCount; ++l) {
        if ((f->payloadWords[f->offsetsWords[l]] & 3u) == 0u) continue;
        float* row = out + (size_t)a * 576u;
        ++a;
        for (int i = 0; i < 576; ++i) row[i] = 0.0f;
        uint32_t n;
        if (render_leaf_controls(f, l, t, slots, &n)) return VBTO_ERR;
        const uint32_t wb = f->offsetsWords[l];
        vbto_header h;
        vbto_unpack_header(f->payloadWords[wb], &h);
        const uint32_t cc = (uint16_t)(h.coarseResolution * h.coarseResolution * h.coarseResolution);
        for (uint32_t i = 0; i < cc && i < 64; ++i) row[i] = slots[i];
        for (uint32_t i = cc; i < n && (i - cc) < 512; ++i) row[64 + (i - cc)] = slots[i];
    }
    return VBTO_OK;
}

int vbto_frame_cache_packed(const vbto_file* f, int t, const uint64_t* ctrlOffset, float* out)
{
    uint32_t a = 0;
    float slots[1024];
    for (uint32_t l = 0; l < f->header.leafCount; ++l) {
        if (f->header.profileType == 2) {
            if (t < 0) return fail("scientific frame index out of range");
            if (vbto_sci_leaf_frame(f, l, (uint32_t)t, out + ctrlOffset[a])) return VBTO_ERR;
            ++a;
            continue;
        }
        if ((f->payloadWords[f->offsetsWords[l]] & 3u) == 0u) continue;
        uint32_t n;
        if (render_leaf_controls(f, l, t, slots, &n)) return VBTO_ERR;
        memcpy(out + ctrlOffset[a], slots, sizeof(float) * n);
        ++a;
    }
    return VBTO_OK;
}
