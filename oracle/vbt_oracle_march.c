/*
 * vbt_oracle_march.c -- CPU restatement of the reference viewer's sampling and
 * ray-march contract over a frame cache
 * (VBTStudio/frontend/shaders/vbt_sample.comp).
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the CPU baseline of the
 * device ray marcher).  Nothing in the product links or calls this file.
 *
 * GLSL leaves a few operations implementation-defined; this restatement fixes
 * them, and the device kernels do the same:
 *   mix(x, y, a)   x * (1 - a) + y * a, each operation rounded on its own
 *   round(x)       half away from zero
 *   min / max      min(x, y) = y < x ? y : x, max(x, y) = x < y ? y : x
 *   clamp(x,l,h)   min(max(x, l), h)
 *   length(v)      sqrt((x*x + y*y) + z*z);  normalize(v) = v / length(v)
 *   locateCoord    at resolution 0 the shader's loop would run for 2^32
 *                  steps; here it stops at once (no stored stream has it)
 * and every float operation is IEEE single precision with no contraction
 * (-ffp-contract=off), division and square root correctly rounded.
 *
 * How it is pinned (tests/test_march_oracle.py):
 *   sample at voxel centres  equals the oracle's dense frame decode (Mode2 with
 *                            the background sign rule) within 4 ulp of the
 *                            field range: the same controls and weights, only
 *                            mix's rounding differs from the exporter's lerp;
 *                            outside the volume and in empty leaves it is the
 *                            render mode's background; a bbox offset shifts
 *                            the grid exactly;
 *   density rays             the integral and the 8 depth bins equal sums of
 *                            point samples at the shader's step positions,
 *                            recomputed in float32 numpy;
 *   level-set tracing        hit t, normal, thickness and floor shadow equal
 *                            a float32 numpy restatement of the shader's
 *                            loops over this file's point samples, per ray.
 * Parity unpinned: decode_temporal (coarse grids above 64 controls, read from
 * the payload instead of the cache) -- no asset here has a coarse resolution
 * above 4; it restates decodeTemporalControlValue line by line.
 */
#include "vbt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { M_EMPTY = 0, M_GRID4 = 1, M_SHELL = 2, M_FINE6 = 3 };
#define TIME_BINS 8u
#define COARSE_CACHE 64u
#define STRIDE 576u

static float gmin(float x, float y) { return y < x ? y : x; }
static float gmax(float x, float y) { return x < y ? y : x; }
static float gclamp(float x, float l, float h) { return gmin(gmax(x, l), h); }
static float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

typedef struct {
    const vbto_file* f;
    const uint32_t* compact;
    const float* cache;
    const vbto_march_params* p;
    uint32_t dimX, dimY, dimZ, leafSize, lcx, lcy, lcz;
} ctx_t;

/* vbt_sample.comp:74-86 */
static uint32_t pbyte(const ctx_t* c, uint32_t baseWord, uint32_t off)
{
    const uint32_t w = c->f->payloadWords[baseWord + (off >> 2)];
    return (w >> ((off & 3u) * 8u)) & 0xFFu;
}
static uint32_t pu16(const ctx_t* c, uint32_t baseWord, uint32_t off)
{
    return pbyte(c, baseWord, off) | (pbyte(c, baseWord, off + 1u) << 8);
}
static uint32_t pu32(const ctx_t* c, uint32_t baseWord, uint32_t off)
{
    return pbyte(c, baseWord, off) | (pbyte(c, baseWord, off + 1u) << 8) | (pbyte(c, baseWord, off + 2u) << 16) |
           (pbyte(c, baseWord, off + 3u) << 24);
}
static float bits_f(uint32_t b)
{
    float f;
    memcpy(&f, &b, 4);
    return f;
}

/* halfToFloat, vbt_sample.comp:88-114 */
static float half_glsl(uint32_t h)
{
    const uint32_t sign = (h >> 15) & 1u, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t bits;
    if (e == 0u) {
        if (m == 0u) {
            bits = sign << 31;
        } else {
            int ex = -14;
            uint32_t mm = m;
            while ((mm & 0x400u) == 0u) {
                mm <<= 1;
                --ex;
            }
            mm &= 0x3FFu;
            bits = (sign << 31) | ((uint32_t)(ex + 127) << 23) | (mm << 13);
        }
    } else if (e == 31u) {
        bits = (sign << 31) | 0x7F800000u | (m << 13);
    } else {
        bits = (sign << 31) | ((e + 112u) << 23) | (m << 13);
    }
    return bits_f(bits);
}

/* readPackedKeyframeValue / readPackedKeyframeFrame, vbt_sample.comp:116-142 */
static float key_value(const ctx_t* c, uint32_t bw, uint32_t keyBase, uint32_t codec, uint32_t i)
{
    if (codec == 1u) return bits_f(pu32(c, bw, keyBase + i * 8u + 4u));
    return half_glsl(pu16(c, bw, keyBase + i * 3u + 1u));
}
static uint32_t key_frame(const ctx_t* c, uint32_t bw, uint32_t keyBase, uint32_t codec, uint32_t i)
{
    return codec == 1u ? pu16(c, bw, keyBase + i * 8u) : pbyte(c, bw, keyBase + i * 3u);
}

/* decodeTemporalControlValue, vbt_sample.comp:144-201 */
static float decode_temporal(const ctx_t* c, uint32_t bw, uint32_t streamBase, uint32_t codec, uint32_t count,
                             uint32_t keys, uint32_t ci)
{
    if (ci >= count) return 0.0f;
    const uint32_t descBytes = count * 4u;
    const uint32_t keyBase = streamBase + descBytes + ((count * TIME_BINS + 3u) & ~3u);
    const uint32_t desc = streamBase + ci * 4u;
    const uint32_t start = pu16(c, bw, desc), n = pbyte(c, bw, desc + 2u);
    if (n == 0u || start + n > keys) return 0.0f;
    const int F = (int)c->f->header.frames;
    const int fi = c->p->frameIndex;
    const int fmax = F - 1 > 0 ? F - 1 : 0;
    const uint32_t frame = (uint32_t)(fi < 0 ? 0 : (fi > fmax ? fmax : fi));
    const uint32_t bin = umin(TIME_BINS - 1u, frame * TIME_BINS / (uint32_t)(F > 1 ? F : 1));
    const uint32_t bins = streamBase + descBytes + ci * TIME_BINS;
    uint32_t lo = umin(pbyte(c, bw, bins + bin), n - 1u);
    uint32_t f0 = key_frame(c, bw, keyBase, codec, start + lo);
    if (f0 > frame) {
        lo = 0u;
        f0 = key_frame(c, bw, keyBase, codec, start);
    }
    if (f0 >= frame || lo + 1u >= n) return key_value(c, bw, keyBase, codec, start + lo);
    uint32_t hi = bin + 1u < TIME_BINS ? umin(n - 1u, pbyte(c, bw, bins + bin + 1u) + 1u) : n - 1u;
    hi = umax(hi, lo + 1u);
    uint32_t f1 = key_frame(c, bw, keyBase, codec, start + hi);
    if (f1 < frame) {
        hi = n - 1u;
        f1 = key_frame(c, bw, keyBase, codec, start + hi);
    }
    if (f1 <= frame) return key_value(c, bw, keyBase, codec, start + hi);
    while (hi > lo + 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        const uint32_t fm = key_frame(c, bw, keyBase, codec, start + mid);
        if (fm < frame) {
            lo = mid;
            f0 = fm;
        } else {
            hi = mid;
            f1 = fm;
        }
    }
    const float next = key_value(c, bw, keyBase, codec, start + hi);
    if (f1 == frame) return next;
    const float prev = key_value(c, bw, keyBase, codec, start + lo);
    return gmix(prev, next, (float)(frame - f0) / (float)(f1 - f0));
}

/* controlCoord / locateCoord, vbt_sample.comp:203-221 */
static float control_coord(uint32_t i, uint32_t res)
{
    if (res <= 1u) return 0.0f;
    return (float)i * (7.0f / (float)(res - 1u));
}
static void locate_coord(float p, uint32_t extent, uint32_t res, uint32_t* i0, uint32_t* i1, float* w)
{
    const float maxP = gmax((float)extent - 1.0f, 0.0f);
    const float cl = gclamp(p, 0.0f, maxP);
    uint32_t j = 1u;
    if (res >= 2u)
        while (j < res - 1u && control_coord(j, res) < cl) ++j;
    *i1 = j;
    *i0 = j >= 1u ? j - 1u : 0u;
    const float c0 = control_coord(*i0, res), c1 = control_coord(*i1, res);
    *w = c1 > c0 ? (cl - c0) / (c1 - c0) : 0.0f;
}

/* currentFrameControl, vbt_sample.comp:228-247 */
static float frame_control(const ctx_t* c, uint32_t cl, uint32_t slotBase, uint32_t bw, uint32_t streamBase,
                           uint32_t codec, uint32_t count, uint32_t keys, uint32_t ci)
{
    if (slotBase == 0u && count > COARSE_CACHE) return decode_temporal(c, bw, streamBase, codec, count, keys, ci);
    return c->cache[(size_t)cl * STRIDE + slotBase + ci];
}

/* sampleControlStream, vbt_sample.comp:249-289 */
static float sample_stream(const ctx_t* c, uint32_t cl, uint32_t slotBase, uint32_t bw, uint32_t streamBase,
                           uint32_t codec, uint32_t res, uint32_t keys, uint32_t lw, uint32_t lh, uint32_t ld, float lx,
                           float ly, float lz)
{
    uint32_t x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    locate_coord(lx, lw, res, &x0, &x1, &tx);
    locate_coord(ly, lh, res, &y0, &y1, &ty);
    locate_coord(lz, ld, res, &z0, &z1, &tz);
    const uint32_t n = res * res * res;
#define CI(x, y, z) (((z) * res + (y)) * res + (x))
    const float v000 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x0, y0, z0));
    const float v100 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x1, y0, z0));
    const float v010 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x0, y1, z0));
    const float v110 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x1, y1, z0));
    const float v001 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x0, y0, z1));
    const float v101 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x1, y0, z1));
    const float v011 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x0, y1, z1));
    const float v111 = frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, CI(x1, y1, z1));
#undef CI
    const float c00 = gmix(v000, v100, tx), c10 = gmix(v010, v110, tx);
    const float c01 = gmix(v001, v101, tx), c11 = gmix(v011, v111, tx);
    const float c0 = gmix(c00, c10, ty), c1 = gmix(c01, c11, ty);
    return gmix(c0, c1, tz);
}

/* sampleShellResidual, vbt_sample.comp:298-320 */
static float shell_residual(const ctx_t* c, uint32_t bw, uint32_t occBase, uint32_t cl, uint32_t vox, int* occupied)
{
    const uint32_t group = vox >> 6, bit = vox & 63u, wordInGroup = bit >> 5, localBit = bit & 31u;
    const uint32_t mask = pu32(c, bw, occBase + group * 8u + wordInGroup * 4u);
    *occupied = ((mask >> localBit) & 1u) != 0u;
    if (!*occupied) return 0.0f;
    const uint32_t prefix = pu16(c, bw, occBase + 64u + group * 2u);
    const uint32_t lowWord = pu32(c, bw, occBase + group * 8u);
    uint32_t di = prefix;
    if (wordInGroup != 0u) di += (uint32_t)__builtin_popcount(lowWord);
    const uint32_t before = localBit == 0u ? 0u : (0xffffffffu >> (32u - localBit));
    di += (uint32_t)__builtin_popcount(mask & before);
    return c->cache[(size_t)cl * STRIDE + COARSE_CACHE + di];
}

/* sampleVbtValue, vbt_sample.comp:322-415 */
static float sample_value(const ctx_t* c, float wx, float wy, float wz, int preserveInterior)
{
    const vbto_march_params* p = c->p;
    const float bg = p->renderMode == 1u ? p->backgroundValue : 0.0f;
    const float px = wx - p->bboxMin[0], py = wy - p->bboxMin[1], pz = wz - p->bboxMin[2];
    if (px < 0.0f || py < 0.0f || pz < 0.0f || px > (float)(c->dimX - 1u) || py > (float)(c->dimY - 1u) ||
        pz > (float)(c->dimZ - 1u))
        return bg;
    const uint32_t x = umin(c->dimX - 1u, (uint32_t)floorf(px));
    const uint32_t y = umin(c->dimY - 1u, (uint32_t)floorf(py));
    const uint32_t z = umin(c->dimZ - 1u, (uint32_t)floorf(pz));
    const uint32_t bx = umin(c->lcx - 1u, x / c->leafSize), by = umin(c->lcy - 1u, y / c->leafSize);
    const uint32_t bz = umin(c->lcz - 1u, z / c->leafSize);
    const uint32_t leaf = (bz * c->lcy + by) * c->lcx + bx;
    const uint32_t wb = c->f->offsetsWords[leaf], we = c->f->offsetsWords[leaf + 1u];
    if (we <= wb) return bg;
    const uint32_t hdr = c->f->payloadWords[wb];
    const uint32_t mode = hdr & 3u;
    if (mode == M_EMPTY) return bg;
    const uint32_t cl = c->compact[leaf];
    const uint32_t cc = (hdr >> 2) & 3u, fc = (hdr >> 4) & 3u, cr = (hdr >> 6) & 7u, fr = (hdr >> 9) & 7u;
    const uint32_t ckc = pu16(c, wb, 4u), fkc = pu16(c, wb, 6u);
    const uint32_t ccount = cr * cr * cr;
    const uint32_t coarseBase = 8u;
    const uint32_t fineBase = coarseBase + ccount * 4u + ((ccount * TIME_BINS + 3u) & ~3u) +
                              ((ckc * (cc == 1u ? 8u : 3u) + 3u) & ~3u);
    const uint32_t baseX = bx * c->leafSize, baseY = by * c->leafSize, baseZ = bz * c->leafSize;
    const uint32_t lw = umin(c->leafSize, c->dimX - baseX), lh = umin(c->leafSize, c->dimY - baseY);
    const uint32_t ld = umin(c->leafSize, c->dimZ - baseZ);
    const float lx = px - (float)baseX, ly = py - (float)baseY, lz = pz - (float)baseZ;
    float v = sample_stream(c, cl, 0u, wb, coarseBase, cc, cr, ckc, lw, lh, ld, lx, ly, lz);
    if (mode == M_FINE6 && fr > 0u) {
        v += sample_stream(c, cl, COARSE_CACHE, wb, fineBase, fc, fr, fkc, lw, lh, ld, lx, ly, lz);
    } else if (mode == M_SHELL) {
        const uint32_t ux = (uint32_t)gclamp(roundf(lx), 0.0f, 7.0f);
        const uint32_t uy = (uint32_t)gclamp(roundf(ly), 0.0f, 7.0f);
        const uint32_t uz = (uint32_t)gclamp(roundf(lz), 0.0f, 7.0f);
        int occ = 0;
        v += shell_residual(c, wb, fineBase, cl, (uz * 8u + uy) * 8u + ux, &occ);
        if (!preserveInterior && !occ && v < 0.0f) v = bg;
    }
    return v;
}

static void v3_at(const float* o, const float* d, float t, float* out)
{
    out[0] = o[0] + d[0] * t;
    out[1] = o[1] + d[1] * t;
    out[2] = o[2] + d[2] * t;
}
static float sample_at(const ctx_t* c, const float* o, const float* d, float t, int keep)
{
    float q[3];
    v3_at(o, d, t, q);
    return sample_value(c, q[0], q[1], q[2], keep);
}

/* intersectWorldBox, vbt_sample.comp:417-428 */
static int intersect_box(const ctx_t* c, const float* o, const float* d, float* nearT, float* farT)
{
    const float mn[3] = {c->p->bboxMin[0], c->p->bboxMin[1], c->p->bboxMin[2]};
    const float mx[3] = {mn[0] + (float)(c->dimX - 1u), mn[1] + (float)(c->dimY - 1u), mn[2] + (float)(c->dimZ - 1u)};
    float tmin[3], tmax[3];
    for (int k = 0; k < 3; ++k) {
        const float inv = 1.0f / d[k];
        const float a = (mn[k] - o[k]) * inv, b = (mx[k] - o[k]) * inv;
        tmin[k] = gmin(a, b);
        tmax[k] = gmax(a, b);
    }
    *nearT = gmax(0.0f, gmax(tmin[0], gmax(tmin[1], tmin[2])));
    *farT = gmin(tmax[0], gmin(tmax[1], tmax[2]));
    return *farT > *nearT;
}

/* traceShadow, vbt_sample.comp:430-452 */
static float trace_shadow(const ctx_t* c, const float* o, const float* d)
{
    const vbto_march_params* p = c->p;
    float nearT, farT;
    if (!intersect_box(c, o, d, &nearT, &farT)) return 0.0f;
    const float eps = gmax(p->surfaceEpsilonVoxels * p->voxelSize, 1e-5f);
    float t = nearT;
    float prev = sample_at(c, o, d, t, 0) - p->surfaceIso;
    for (uint32_t it = 0; it < p->shadowSteps; ++it) {
        const float step = gclamp(fabsf(prev) / gmax(p->voxelSize, 1e-6f) * 0.8f, 0.35f, 5.0f);
        const float nextT = gmin(t + step, farT);
        const float next = sample_at(c, o, d, nextT, 0) - p->surfaceIso;
        if ((prev < 0.0f) != (next < 0.0f)) {
            const int nextInside = next < 0.0f;
            const float neg = nextInside ? next : prev;
            const float supT = nextInside ? gmin(nextT + 0.35f, farT) : gmax(t - 0.35f, nearT);
            const float sup = sample_at(c, o, d, supT, 0) - p->surfaceIso;
            if (neg <= -eps * 0.5f || sup <= -eps * 0.25f) return 1.0f;
        }
        if (nextT >= farT) break;
        t = nextT;
        prev = fabsf(next) <= eps ? eps : next;
    }
    return 0.0f;
}

/* estimateThickness, vbt_sample.comp:454-474 */
static float thickness(const ctx_t* c, const float* o, const float* d, float tmax, float hitT)
{
    const vbto_march_params* p = c->p;
    const float eps = gmax(p->surfaceEpsilonVoxels * p->voxelSize, 1e-5f);
    float t = gmin(hitT + 0.35f, tmax);
    float prev = sample_at(c, o, d, t, 1) - p->surfaceIso;
    int entered = prev < -0.25f * eps;
    const uint32_t limit = umin(p->stepCount, 160u);
    for (uint32_t it = 0; it < limit; ++it) {
        const float step = gclamp(fabsf(prev) / gmax(p->voxelSize, 1e-6f) * 0.7f, 0.3f, 3.0f);
        const float nextT = gmin(t + step, tmax);
        const float next = sample_at(c, o, d, nextT, 1) - p->surfaceIso;
        entered = entered || next < -0.25f * eps;
        if (entered && prev < 0.0f && next >= 0.0f) return gmax((nextT - hitT) * p->voxelSize, p->voxelSize * 0.25f);
        if (nextT >= tmax) break;
        t = nextT;
        prev = next;
    }
    return p->voxelSize * 6.0f;
}

/* traceLevelSet, vbt_sample.comp:476-551; out = primary[4], auxiliary[4] */
static void trace_level_set(const ctx_t* c, const float* ray, float* out)
{
    const vbto_march_params* p = c->p;
    const float* o = ray;
    const float* d = ray + 4;
    const float tmax = ray[7];
    out[0] = out[1] = out[2] = 0.0f;
    out[3] = -1.0f;
    out[4] = out[5] = out[6] = out[7] = 0.0f;
    if (tmax <= 0.0f) return;
    const float eps = gmax(p->surfaceEpsilonVoxels * p->voxelSize, 1e-5f);
    float prevT = 0.0f;
    float prev = sample_value(c, o[0], o[1], o[2], 0) - p->surfaceIso;
    float hitT = -1.0f;
    for (uint32_t it = 0; it < p->stepCount; ++it) {
        const float step = gclamp(fabsf(prev) / gmax(p->voxelSize, 1e-6f) * 0.8f, 0.2f, gmax(p->maxStepVoxels, 0.2f));
        const float curT = gmin(prevT + step, tmax);
        const float cur = sample_at(c, o, d, curT, 0) - p->surfaceIso;
        if ((prev < 0.0f) != (cur < 0.0f)) {
            const int curInside = cur < 0.0f;
            const float neg = curInside ? cur : prev;
            const float supT = curInside ? gmin(curT + 0.35f, tmax) : gmax(prevT - 0.35f, 0.0f);
            const float sup = sample_at(c, o, d, supT, 0) - p->surfaceIso;
            if (!(neg <= -eps * 0.5f || sup <= -eps * 0.25f)) {
                if (curT >= tmax) break;
                prevT = curT;
                prev = gmax(cur, eps);
                continue;
            }
            float low = prevT, high = curT, lowPhi = prev;
            for (int r = 0; r < 5; ++r) {
                const float mid = 0.5f * (low + high);
                const float midPhi = sample_at(c, o, d, mid, 0) - p->surfaceIso;
                if ((lowPhi < 0.0f) != (midPhi < 0.0f)) {
                    high = mid;
                } else {
                    low = mid;
                    lowPhi = midPhi;
                }
            }
            hitT = 0.5f * (low + high);
            break;
        }
        if (curT >= tmax) break;
        prevT = curT;
        prev = cur;
    }
    if (hitT < 0.0f) return;
    float h[3];
    v3_at(o, d, hitT, h);
    const float ns = gmax(p->surfaceNormalStep, 0.25f);
    const float g0 = sample_value(c, h[0] + ns, h[1], h[2], 0) - sample_value(c, h[0] - ns, h[1], h[2], 0);
    const float g1 = sample_value(c, h[0], h[1] + ns, h[2], 0) - sample_value(c, h[0], h[1] - ns, h[2], 0);
    const float g2 = sample_value(c, h[0], h[1], h[2] + ns, 0) - sample_value(c, h[0], h[1], h[2] - ns, 0);
    const float len = sqrtf((g0 * g0 + g1 * g1) + g2 * g2);
    float n[3];
    if (len > 1e-6f) {
        n[0] = g0 / len;
        n[1] = g1 / len;
        n[2] = g2 / len;
    } else {
        n[0] = -d[0];
        n[1] = -d[1];
        n[2] = -d[2];
    }
    out[0] = n[0];
    out[1] = n[1];
    out[2] = n[2];
    out[3] = hitT;
    out[4] = p->surfaceModel != 0u ? thickness(c, o, d, tmax, hitT) : 0.0f;
    out[7] = fabsf((n[0] * -d[0] + n[1] * -d[1]) + n[2] * -d[2]);
}

/* main, vbt_sample.comp:553-622 */
static void march_ray(const ctx_t* c, const float* ray, float* out)
{
    const vbto_march_params* p = c->p;
    if (p->renderMode == 1u) {
        trace_level_set(c, ray, out);
        const int zUp = p->upAxis == 2u;
        const float fd = zUp ? ray[6] : ray[5];
        const float fo = zUp ? ray[2] : ray[1];
        if (p->surfaceModel != 0u && out[3] < 0.0f && fd < -1e-5f) {
            const float floorT = (p->floorHeight - fo) / fd;
            if (floorT > 0.0f) {
                float fp[3];
                v3_at(ray, ray + 4, floorT, fp);
                /* normalize(vec3(-0.55, 0.42, 0.72)) or (-0.55, 0.72, 0.42) */
                const float k0 = -0.55f, k1 = zUp ? 0.42f : 0.72f, k2 = zUp ? 0.72f : 0.42f;
                const float kl = sqrtf((k0 * k0 + k1 * k1) + k2 * k2);
                const float kd[3] = {k0 / kl, k1 / kl, k2 / kl};
                const float so[3] = {fp[0] + kd[0] * 0.5f, fp[1] + kd[1] * 0.5f, fp[2] + kd[2] * 0.5f};
                out[5] = 1.0f;
                out[6] = trace_shadow(c, so, kd);
            }
        }
        return;
    }
    const float dt = ray[7] / gmax((float)p->stepCount, 1.0f);
    float t = ray[3] + 0.5f * dt;
    float bins[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (p->surfaceModel != 0u) {
        for (uint32_t i = 0; i < p->stepCount; ++i) {
            const float v = sample_at(c, ray, ray + 4, t, 0);
            const uint32_t bin = umin((i * 8u) / umax(p->stepCount, 1u), 7u);
            bins[bin] += v;
            t += dt;
        }
        memcpy(out, bins, sizeof(bins));
        return;
    }
    float acc = 0.0f;
    for (uint32_t i = 0; i < p->stepCount; ++i) {
        acc += sample_at(c, ray, ray + 4, t, 0);
        t += dt;
    }
    memset(out, 0, 8 * sizeof(float));
    out[0] = acc;
}

static int make_ctx(const vbto_file* f, const float* cache576, const vbto_march_params* p, ctx_t* c, uint32_t** compact)
{
    if (f->header.profileType == 2u) return -1;
    *compact = (uint32_t*)malloc(sizeof(uint32_t) * (f->header.leafCount ? f->header.leafCount : 1u));
    if (!*compact) return -1;
    vbto_page_table(f, *compact);
    c->f = f;
    c->compact = *compact;
    c->cache = cache576;
    c->p = p;
    c->dimX = f->header.width;
    c->dimY = f->header.height;
    c->dimZ = f->header.depth;
    c->leafSize = f->header.leafSize;
    c->lcx = (c->dimX + c->leafSize - 1u) / c->leafSize;
    c->lcy = (c->dimY + c->leafSize - 1u) / c->leafSize;
    c->lcz = (c->dimZ + c->leafSize - 1u) / c->leafSize;
    return 0;
}

int vbto_sample_points(const vbto_file* f, const float* cache576, const vbto_march_params* p, const float* xyz,
                       uint64_t n, int preserveInterior, float* out, int nthreads)
{
    ctx_t c;
    uint32_t* compact = NULL;
    if (make_ctx(f, cache576, p, &c, &compact)) return VBTO_ERR;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        out[i] = sample_value(&c, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], preserveInterior);
    free(compact);
    return VBTO_OK;
}

int vbto_march_rays(const vbto_file* f, const float* cache576, const vbto_march_params* p, const float* rays,
                    uint64_t n, float* out, int nthreads)
{
    ctx_t c;
    uint32_t* compact = NULL;
    if (make_ctx(f, cache576, p, &c, &compact)) return VBTO_ERR;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t i = 0; i < (int64_t)n; ++i) march_ray(&c, rays + 8 * i, out + 8 * i);
    free(compact);
    return VBTO_OK;
}
