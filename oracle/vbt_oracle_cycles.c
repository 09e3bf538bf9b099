/*
 * vbt_oracle_cycles.c -- CPU restatement of the reference's Cycles consumer of
 * VBTPACK4: the resident blob (cycles_native_vbt/vbt_cycles_blob.{h,cpp}, the
 * file checks of cycles_native_vbt/vbt_cycles_loader.cpp) and the two density
 * samplers that read it straight from the packed payload:
 *   standalone  cycles_native_vbt/vbt_cycles_sampler.h:165-376, the form the
 *               reference's CPU probe (vbt_sampler_cpu_probe.cpp) runs;
 *   kernel      blender_patch/native_cycles_vbt.patch:688-929, the form the
 *               patched Cycles kernel runs on the device (voxel-corner
 *               trilinear over eight per-voxel leaf samples; Mode2 leaves and
 *               resolution <= 1 handled as that code does).
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the CPU baseline of the
 * device blob sampler).  Nothing in the product links or calls this file.
 *
 * Arithmetic: IEEE single precision, no contraction (-ffp-contract=off),
 * correctly rounded division; mix(a, b, t) = a * (1 - t) + b * t with each
 * operation rounded on its own, as both reference forms write it.
 *
 * Two places where the reference has no defined result are fixed here, and the
 * device kernels do the same:
 *   - every payload and offset read past the blob's sections reads zero bytes
 *     (the reference reads whatever lies there; no valid file reaches it);
 *   - a stream of resolution 0 contributes 0 in both forms (the standalone
 *     form's coordinate search would run for 2^32 steps and then read far past
 *     the leaf; the kernel form returns 0).
 *
 * How it is pinned (tests/test_cycles_oracle.py):
 *   standalone samples   bit for bit against the reference's own CPU probe,
 *                        compiled from its sources into oracle/_ref, on the
 *                        golden render assets (density, and the level set with
 *                        its Mode2 leaves), at random, voxel, leaf-boundary and
 *                        outside points and at several frames, also through
 *                        the committed fixture tests/golden/cycles_probe.npz;
 *   blob                 size equal to the probe's reported blob size; the
 *                        header fields equal the file header's, and the
 *                        offset and payload sections equal the file's bytes;
 *   kernel form          at integer voxel points equal to the probe's value
 *                        for leaves other than Mode2 (t = 0 in every mix),
 *                        zero in Mode2 leaves; between voxels the trilinear
 *                        mix of the eight corner values the probe returns.
 */
#include "vbt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#define BLOB_MAGIC 0x31544256u /* "VBT1" */
#define BLOB_HEADER_BYTES 76u  /* 19 four-byte fields */

static char g_cycles_err[160];

const char* vbto_cycles_error(void) { return g_cycles_err; }

static int cfail(const char* m)
{
    snprintf(g_cycles_err, sizeof g_cycles_err, "%s", m);
    return VBTO_ERR;
}

static uint32_t rd32(const uint8_t* p)
{
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

static void wr32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }

/* load_pack4 (vbt_cycles_loader.cpp:116-221) with validate_header (:36-114),
 * then pack_blob (vbt_cycles_blob.cpp:22-87). */
int vbto_cycles_pack_blob(const void* file, uint64_t fileBytes, int frameIndex, void* out, uint64_t capacity,
                          uint64_t* blobBytes)
{
    const uint8_t* b = (const uint8_t*)file;
    g_cycles_err[0] = 0;
    if (fileBytes < 52u) return cfail("Invalid VBTPACK4 file size.");
    uint32_t w[11];
    for (int i = 0; i < 11; ++i) w[i] = rd32(b + 8 + 4 * i);
    const uint32_t version = w[0], width = w[1], height = w[2], depth = w[3], frames = w[4], leafSize = w[5],
                   coarseRes = w[6], maxKeep = w[7], leafCount = w[8], profile = w[9], scaleCount = w[10];
    if (memcmp(b, "VBTPACK4", 8) != 0) return cfail("Unsupported VBT magic. Expected VBTPACK4.");
    if (version != 4u) return cfail("Unsupported VBTPACK4 version.");
    if (!width || !height || !depth || !frames) return cfail("VBTPACK4 dimensions and frame count must be non-zero.");
    if (leafSize != 8u) return cfail("VBTPACK4 v4 requires leaf_size=8.");
    if (coarseRes == 0u || coarseRes > 8u) return cfail("Invalid VBTPACK4 coarse resolution.");
    if (profile > 2u) return cfail("Invalid VBTPACK4 profile type.");
    if (maxKeep > frames || scaleCount > 64u || scaleCount > maxKeep)
        return cfail("Invalid VBTPACK4 coarse temporal metadata.");
    const uint64_t lcx = (width + 7ull) / 8u, lcy = (height + 7ull) / 8u, lcz = (depth + 7ull) / 8u;
    const uint64_t expect = lcx * lcy * lcz;
    if (expect > 0xFFFFFFFFull) return cfail("VBTPACK4 leaf count overflows the v4 header.");
    if (leafCount != expect) return cfail("VBTPACK4 leaf_count does not match dimensions and leaf_size.");
    const uint64_t offPos = 52ull + 4ull * scaleCount;
    if (offPos + 4ull * (leafCount + 1ull) > fileBytes) return cfail("VBTPACK4 file is truncated before its offset table.");
    const uint8_t* off = b + offPos;
    if (rd32(off) != 0u) return cfail("VBTPACK4 offset table must begin at word zero.");
    for (uint32_t i = 0; i < leafCount; ++i)
        if (rd32(off + 4ull * i) >= rd32(off + 4ull * (i + 1u)))
            return cfail("VBTPACK4 offsets must be strictly increasing for every leaf.");
    const uint32_t payloadWords = rd32(off + 4ull * leafCount);
    const uint64_t payPos = offPos + 4ull * (leafCount + 1ull);
    const uint64_t expectBytes = payPos + 4ull * payloadWords;
    if (expectBytes != fileBytes)
        return cfail(expectBytes > fileBytes ? "VBTPACK4 payload is truncated." : "VBTPACK4 file contains trailing bytes.");

    const uint64_t offsetsAt = BLOB_HEADER_BYTES; /* already a multiple of 4 */
    const uint64_t payloadAt = offsetsAt + 4ull * (leafCount + 1ull);
    const uint64_t total = payloadAt + 4ull * payloadWords;
    if (payloadAt > 0xFFFFFFFFull) return cfail("VBT blob section offset exceeds the 32-bit header range.");
    *blobBytes = total;
    if (!out) return VBTO_OK;
    if (capacity < total) return cfail("blob buffer too small");
    uint8_t* o = (uint8_t*)out;
    memset(o, 0, BLOB_HEADER_BYTES);
    const int last = (int)frames - 1;
    const uint32_t fi = (uint32_t)(frameIndex < 0 ? 0 : (frameIndex > last ? last : frameIndex));
    const uint32_t hdr[16] = {BLOB_MAGIC, 1u,         width,          height,       depth,
                              frames,     leafSize,   (uint32_t)lcx,  (uint32_t)lcy, (uint32_t)lcz,
                              leafCount,  fi,         leafCount + 1u, payloadWords, (uint32_t)offsetsAt,
                              (uint32_t)payloadAt};
    for (int i = 0; i < 16; ++i) wr32(o + 4 * i, hdr[i]); /* bbox_min stays 0.0f */
    memcpy(o + offsetsAt, off, 4ull * (leafCount + 1ull));
    memcpy(o + payloadAt, b + payPos, 4ull * payloadWords);
    return VBTO_OK;
}

/* ---- the sampler ---- */

typedef struct {
    uint32_t dimX, dimY, dimZ, frames, leafSize, lcx, lcy, lcz, offWords, payWords;
    int frameIndex;
    float bbox[3];
    const uint8_t* off;
    const uint8_t* pay;
} view_t;

static uint32_t byte_at(const view_t* v, uint32_t at)
{
    return (uint64_t)at < 4ull * v->payWords ? v->pay[at] : 0u;
}

static uint32_t u16_at(const view_t* v, uint32_t at) { return byte_at(v, at) | (byte_at(v, at + 1u) << 8); }

static uint32_t off_at(const view_t* v, uint32_t i) { return i < v->offWords ? rd32(v->off + 4ull * i) : 0u; }

static float f_bits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static float half_f(uint32_t h)
{
    /* sign, then a normalised, subnormal (shifted until the hidden bit shows)
     * or all-ones exponent carried over bit for bit */
    const uint32_t s = (h >> 15) << 31, e = (h >> 10) & 31u;
    uint32_t m = h & 0x3FFu;
    if (e == 0u) {
        if (m == 0u) return f_bits(s);
        int ex = 113;
        while (!(m & 0x400u)) { m <<= 1; --ex; }
        return f_bits(s | ((uint32_t)ex << 23) | ((m & 0x3FFu) << 13));
    }
    if (e == 31u) return f_bits(s | 0x7F800000u | (m << 13));
    return f_bits(s | ((e + 112u) << 23) | (m << 13));
}

static uint32_t key_frame_at(const view_t* v, uint32_t keys, uint32_t codec, uint32_t k)
{
    return codec == 1u ? u16_at(v, keys + 8u * k) : byte_at(v, keys + 3u * k);
}

static float key_value_at(const view_t* v, uint32_t keys, uint32_t codec, uint32_t k)
{
    if (codec == 1u) {
        const uint32_t a = keys + 8u * k + 4u;
        return f_bits(byte_at(v, a) | (byte_at(v, a + 1u) << 8) | (byte_at(v, a + 2u) << 16) |
                      (byte_at(v, a + 3u) << 24));
    }
    return half_f(u16_at(v, keys + 3u * k + 1u));
}

static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* decode_temporal_control_value, vbt_cycles_sampler.h:165-212 (the patch's
 * copy at native_cycles_vbt.patch:639-685 is the same): the descriptor's key
 * range, a bracket from the frame's bin hint and the next bin's, a binary
 * search inside it, and the lerp weight as an IEEE quotient of the unsigned
 * frame differences. */
static float control_value(const view_t* v, uint32_t stream, uint32_t codec, uint32_t nc, uint32_t nkeys,
                           uint32_t ci)
{
    const uint32_t descBytes = nc * 4u;
    const uint32_t keys = stream + descBytes + ((nc * 8u + 3u) & ~3u);
    const uint32_t start = u16_at(v, stream + ci * 4u);
    const uint32_t count = byte_at(v, stream + ci * 4u + 2u);
    if (count == 0u || start + count > nkeys) return 0.0f;
    const int F = (int)v->frames;
    const int hiF = F > 0 ? F - 1 : 0;
    const uint32_t t = (uint32_t)(v->frameIndex < 0 ? 0 : (v->frameIndex > hiF ? hiF : v->frameIndex));
    const uint32_t fd = v->frames > 1u ? v->frames : 1u;
    uint32_t bin = t * 8u / fd;
    if (bin > 7u) bin = 7u;
    const uint32_t bins = stream + descBytes + ci * 8u;
    uint32_t lo = byte_at(v, bins + bin);
    if (lo > count - 1u) lo = count - 1u;
    uint32_t f0 = key_frame_at(v, keys, codec, start + lo);
    if (f0 > t) {
        lo = 0u;
        f0 = key_frame_at(v, keys, codec, start);
    }
    if (f0 >= t || lo + 1u >= count) return key_value_at(v, keys, codec, start + lo);
    uint32_t hi = count - 1u;
    if (bin < 7u) {
        const uint32_t h = byte_at(v, bins + bin + 1u) + 1u;
        if (h < hi) hi = h;
    }
    if (hi < lo + 1u) hi = lo + 1u;
    uint32_t f1 = key_frame_at(v, keys, codec, start + hi);
    if (f1 < t) {
        hi = count - 1u;
        f1 = key_frame_at(v, keys, codec, start + hi);
    }
    if (f1 <= t) return key_value_at(v, keys, codec, start + hi);
    while (hi > lo + 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        const uint32_t fm = key_frame_at(v, keys, codec, start + mid);
        if (fm < t) { lo = mid; f0 = fm; }
        else { hi = mid; f1 = fm; }
    }
    const float next = key_value_at(v, keys, codec, start + hi);
    if (f1 == t) return next;
    const float prev = key_value_at(v, keys, codec, start + lo);
    const float alpha = (float)(uint32_t)(t - f0) / (float)(uint32_t)(f1 - f0);
    return mixf(prev, next, alpha);
}

static float ctrl_pos(uint32_t i, uint32_t res)
{
    return res <= 1u ? 0.0f : (float)i * (7.0f / (float)(res - 1u));
}

/* locate_coord, vbt_cycles_sampler.h:219-243; the kernel form
 * (native_cycles_vbt.patch:694-719) first sends resolution <= 1 to (0, 0, 0). */
static void locate(float p, uint32_t extent, uint32_t res, int kernelForm, uint32_t* i0, uint32_t* i1, float* w)
{
    if (kernelForm && res <= 1u) {
        *i0 = *i1 = 0u;
        *w = 0.0f;
        return;
    }
    const float maxP = (float)(extent > 0u ? extent - 1u : 0u);
    const float c = p < 0.0f ? 0.0f : (p > maxP ? maxP : p);
    uint32_t hi = 1u;
    while (hi < res - 1u && ctrl_pos(hi, res) < c) ++hi;
    const uint32_t lo = hi > 0u ? hi - 1u : 0u;
    const float c0 = ctrl_pos(lo, res), c1 = ctrl_pos(hi, res);
    *i0 = lo;
    *i1 = hi;
    *w = c1 > c0 ? (c - c0) / (c1 - c0) : 0.0f;
}

/* sample_control_stream, vbt_cycles_sampler.h:245-282 */
static float stream_sample(const view_t* v, uint32_t stream, uint32_t codec, uint32_t res, uint32_t nkeys,
                           const uint32_t ext[3], const float l[3], int kernelForm)
{
    if (res == 0u) return 0.0f;
    uint32_t i0[3], i1[3];
    float w[3];
    for (int a = 0; a < 3; ++a) locate(l[a], ext[a], res, kernelForm, &i0[a], &i1[a], &w[a]);
    const uint32_t nc = res * res * res;
    float c[8];
    for (int k = 0; k < 8; ++k) {
        const uint32_t x = (k & 1) ? i1[0] : i0[0], y = (k & 2) ? i1[1] : i0[1], z = (k & 4) ? i1[2] : i0[2];
        c[k] = control_value(v, stream, codec, nc, nkeys, (z * res + y) * res + x);
    }
    const float c00 = mixf(c[0], c[1], w[0]), c10 = mixf(c[2], c[3], w[0]);
    const float c01 = mixf(c[4], c[5], w[0]), c11 = mixf(c[6], c[7], w[0]);
    return mixf(mixf(c00, c10, w[1]), mixf(c01, c11, w[1]), w[2]);
}

/* sample_density_index, vbt_cycles_sampler.h:284-368; the kernel form's
 * vbt_sample_density_leaf_index (native_cycles_vbt.patch:788-874) differs in
 * Mode2 leaves (zero), leaf_count_z (from the header) and resolution <= 1. */
static float leaf_sample(const view_t* v, float x, float y, float z, int kernelForm)
{
    if (x < 0.0f || y < 0.0f || z < 0.0f || x > (float)(v->dimX - 1u) || y > (float)(v->dimY - 1u) ||
        z > (float)(v->dimZ - 1u))
        return 0.0f;
    uint32_t xi = (uint32_t)x, yi = (uint32_t)y, zi = (uint32_t)z;
    if (xi > v->dimX - 1u) xi = v->dimX - 1u;
    if (yi > v->dimY - 1u) yi = v->dimY - 1u;
    if (zi > v->dimZ - 1u) zi = v->dimZ - 1u;
    const uint32_t S = v->leafSize;
    const uint32_t lcz = kernelForm ? v->lcz : (v->dimZ + S - 1u) / S;
    uint32_t bx = xi / S, by = yi / S, bz = zi / S;
    if (bx > v->lcx - 1u) bx = v->lcx - 1u;
    if (by > v->lcy - 1u) by = v->lcy - 1u;
    if (bz > lcz - 1u) bz = lcz - 1u;
    const uint32_t leaf = (bz * v->lcy + by) * v->lcx + bx;
    const uint32_t wb = off_at(v, leaf), we = off_at(v, leaf + 1u);
    if (we <= wb) return 0.0f;
    const uint32_t base = wb * 4u;
    const uint32_t hdr = byte_at(v, base) | (byte_at(v, base + 1u) << 8) | (byte_at(v, base + 2u) << 16) |
                         (byte_at(v, base + 3u) << 24);
    const uint32_t mode = hdr & 3u;
    if (mode == 0u || (kernelForm && mode == 2u)) return 0.0f;
    const uint32_t cc = (hdr >> 2) & 3u, fc = (hdr >> 4) & 3u, cr = (hdr >> 6) & 7u, fr = (hdr >> 9) & 7u;
    const uint32_t ck = u16_at(v, base + 4u), fk = u16_at(v, base + 6u);
    const uint32_t cn = cr * cr * cr;
    const uint32_t cStream = base + 8u;
    const uint32_t fStream = cStream + cn * 4u + ((cn * 8u + 3u) & ~3u) + ((ck * (cc == 1u ? 8u : 3u) + 3u) & ~3u);
    const uint32_t ox = bx * S, oy = by * S, oz = bz * S;
    const uint32_t ext[3] = {S < v->dimX - ox ? S : v->dimX - ox, S < v->dimY - oy ? S : v->dimY - oy,
                             S < v->dimZ - oz ? S : v->dimZ - oz};
    const float l[3] = {x - (float)ox, y - (float)oy, z - (float)oz};
    float value = stream_sample(v, cStream, cc, cr, ck, ext, l, kernelForm);
    if (mode == 3u && fr > 0u) value += stream_sample(v, fStream, fc, fr, fk, ext, l, kernelForm);
    return value;
}

/* the kernel form's vbt_sample_density_index (native_cycles_vbt.patch:876-913) */
static float voxel_trilinear(const view_t* v, float x, float y, float z)
{
    if (x < 0.0f || y < 0.0f || z < 0.0f || x > (float)(v->dimX - 1u) || y > (float)(v->dimY - 1u) ||
        z > (float)(v->dimZ - 1u))
        return 0.0f;
    const float p[3] = {x, y, z};
    const uint32_t dim[3] = {v->dimX, v->dimY, v->dimZ};
    uint32_t a0[3], a1[3];
    float t[3];
    for (int a = 0; a < 3; ++a) {
        uint32_t i = (uint32_t)p[a];
        if (i > dim[a] - 1u) i = dim[a] - 1u;
        const uint32_t j = i + 1u < dim[a] - 1u ? i + 1u : dim[a] - 1u;
        a0[a] = i;
        a1[a] = j;
        t[a] = j > i ? p[a] - (float)i : 0.0f;
    }
    float c[8];
    for (int k = 0; k < 8; ++k)
        c[k] = leaf_sample(v, (float)((k & 1) ? a1[0] : a0[0]), (float)((k & 2) ? a1[1] : a0[1]),
                           (float)((k & 4) ? a1[2] : a0[2]), 1);
    const float c00 = mixf(c[0], c[1], t[0]), c10 = mixf(c[2], c[3], t[0]);
    const float c01 = mixf(c[4], c[5], t[0]), c11 = mixf(c[6], c[7], t[0]);
    return mixf(mixf(c00, c10, t[1]), mixf(c01, c11, t[1]), t[2]);
}

/* blob_payload_view's checks (vbt_cycles_blob.cpp:89-124) plus the kernel's
 * magic test; a blob failing them samples to 0 everywhere in the kernel form
 * and is an error in the standalone form. */
static int open_view(const uint8_t* blob, uint64_t bytes, view_t* v)
{
    if (bytes < BLOB_HEADER_BYTES) return 0;
    uint32_t h[19];
    for (int i = 0; i < 19; ++i) h[i] = rd32(blob + 4 * i);
    if (h[0] != BLOB_MAGIC || h[1] != 1u) return 0;
    if ((uint64_t)h[14] + 4ull * h[12] > bytes || (uint64_t)h[15] + 4ull * h[13] > bytes) return 0;
    if (h[14] % 4u || h[15] % 4u) return 0;
    v->dimX = h[2]; v->dimY = h[3]; v->dimZ = h[4]; v->frames = h[5]; v->leafSize = h[6];
    v->lcx = h[7]; v->lcy = h[8]; v->lcz = h[9]; v->frameIndex = (int)h[11];
    v->offWords = h[12]; v->payWords = h[13];
    v->off = blob + h[14];
    v->pay = blob + h[15];
    memcpy(v->bbox, blob + 64, 12);
    return 1;
}

int vbto_cycles_sample(const void* blob, uint64_t blobBytes, int flags, const float* xyz, uint64_t n, float* out,
                       int nthreads)
{
    view_t v;
    const int kernelForm = (flags & VBTO_CYCLES_KERNEL) != 0;
    if (!open_view((const uint8_t*)blob, blobBytes, &v)) {
        if (!kernelForm) return cfail("Invalid VBT blob header.");
        for (uint64_t i = 0; i < n; ++i) out[i] = 0.0f;
        return VBTO_OK;
    }
    if (v.leafSize == 0u) return cfail("blob leaf_size is zero");
    const int world = (flags & VBTO_CYCLES_WORLD) != 0;
    (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (world) {
            x -= v.bbox[0];
            y -= v.bbox[1];
            z -= v.bbox[2];
        }
        out[i] = kernelForm ? voxel_trilinear(&v, x, y, z) : leaf_sample(&v, x, y, z, 0);
    }
    return VBTO_OK;
}
