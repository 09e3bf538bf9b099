// march.hip -- the reference viewer's sampling and ray-march contract
// (VBTStudio/frontend/shaders/vbt_sample.comp) on a frame cache built by
// vbt_decode_frame_cache in the stride-576 layout: point samples at world
// positions and one ray per thread (density integration or level-set
// surface tracing with floor shadows).
//
// The shader leaves a few operations implementation-defined; these kernels
// fix them (and the CPU restatement used as the oracle fixes them the same
// way): mix(x, y, a) = x * (1 - a) + y * a with each operation rounded on its
// own, round() rounds half away from zero, min(x, y) = y < x ? y : x,
// max(x, y) = x < y ? y : x, clamp = min(max()), length = sqrt((x*x + y*y) +
// z*z), normalize divides by it.  Every float operation is IEEE single
// precision without contraction, division and square root correctly rounded,
// so the kernels and the oracle agree bit for bit.
#include "vbt_format.h"
#include "vbt_internal.h"
#include "vbt_kernels.h"

#include <hip/hip_runtime.h>

using namespace vbtk;

namespace {

constexpr uint32_t TIME_BINS = 8u, COARSE_CACHE = 64u, STRIDE = 576u;

__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float gclamp(float x, float l, float h) { return gmin(gmax(x, l), h); }
__device__ __forceinline__ float gmix(float x, float y, float a)
{
    const float p = x * (1.0f - a);
    const float q = y * a;
    return p + q;
}

struct MarchCtx {
    const uint32_t* offsets;
    const uint32_t* words;
    const uint32_t* compact;
    const float* cache;
    uint32_t dimX, dimY, dimZ, frames, leafSize, lcx, lcy, lcz;
    vbt_march_params p;
};

// readPayloadByte / U16 / U32 (vbt_sample.comp:74-86, 291-296)
__device__ __forceinline__ uint32_t pbyte(const MarchCtx& c, uint32_t bw, uint32_t off)
{
    return (c.words[bw + (off >> 2)] >> ((off & 3u) * 8u)) & 0xFFu;
}
__device__ __forceinline__ uint32_t pu16(const MarchCtx& c, uint32_t bw, uint32_t off)
{
    const uint32_t i = bw + (off >> 2);
    return __builtin_amdgcn_alignbyte(c.words[i + 1u], c.words[i], off & 3u) & 0xFFFFu;
}
__device__ __forceinline__ uint32_t pu32(const MarchCtx& c, uint32_t bw, uint32_t off)
{
    const uint32_t i = bw + (off >> 2);
    return __builtin_amdgcn_alignbyte(c.words[i + 1u], c.words[i], off & 3u);
}

// halfToFloat (vbt_sample.comp:88-114)
__device__ inline float half_glsl(uint32_t h)
{
    const uint32_t sign = (h >> 15) & 1u, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t bits;
    if (e == 0u) {
        if (m == 0u) {
            bits = sign << 31;
        } else {
            // normalise the subnormal: shift the leading one into bit 10
            const uint32_t lz = (uint32_t)__builtin_clz(m) - 21u; // m < 0x400: clz >= 22
            const uint32_t mm = (m << lz) & 0x3FFu;
            bits = (sign << 31) | ((uint32_t)(-14 - (int)lz + 127) << 23) | (mm << 13);
        }
    } else if (e == 31u) {
        bits = (sign << 31) | 0x7F800000u | (m << 13);
    } else {
        bits = (sign << 31) | ((e + 112u) << 23) | (m << 13);
    }
    return __uint_as_float(bits);
}

// readPackedKeyframeValue / Frame (vbt_sample.comp:116-142)
__device__ __forceinline__ float key_value(const MarchCtx& c, uint32_t bw, uint32_t kb, uint32_t codec, uint32_t i)
{
    if (codec == 1u) return __uint_as_float(pu32(c, bw, kb + i * 8u + 4u));
    return half_glsl(pu16(c, bw, kb + i * 3u + 1u));
}
__device__ __forceinline__ uint32_t key_frame(const MarchCtx& c, uint32_t bw, uint32_t kb, uint32_t codec, uint32_t i)
{
    return codec == 1u ? pu16(c, bw, kb + i * 8u) : pbyte(c, bw, kb + i * 3u);
}

// decodeTemporalControlValue (vbt_sample.comp:144-201): coarse grids with
// more than 64 controls are not in the cache and decode from the payload
__device__ float decode_temporal(const MarchCtx& c, uint32_t bw, uint32_t streamBase, uint32_t codec, uint32_t count,
                                 uint32_t keys, uint32_t ci)
{
    if (ci >= count) return 0.0f;
    const uint32_t descBytes = count * 4u;
    const uint32_t keyBase = streamBase + descBytes + ((count * TIME_BINS + 3u) & ~3u);
    const uint32_t desc = streamBase + ci * 4u;
    const uint32_t start = pu16(c, bw, desc), n = pbyte(c, bw, desc + 2u);
    if (n == 0u || start + n > keys) return 0.0f;
    const int F = (int)c.frames, fi = c.p.frameIndex, fmax = F - 1 > 0 ? F - 1 : 0;
    const uint32_t frame = (uint32_t)(fi < 0 ? 0 : (fi > fmax ? fmax : fi));
    const uint32_t bin = min(TIME_BINS - 1u, frame * TIME_BINS / (uint32_t)(F > 1 ? F : 1));
    const uint32_t bins = streamBase + descBytes + ci * TIME_BINS;
    uint32_t lo = min(pbyte(c, bw, bins + bin), n - 1u);
    uint32_t f0 = key_frame(c, bw, keyBase, codec, start + lo);
    if (f0 > frame) {
        lo = 0u;
        f0 = key_frame(c, bw, keyBase, codec, start);
    }
    if (f0 >= frame || lo + 1u >= n) return key_value(c, bw, keyBase, codec, start + lo);
    uint32_t hi = bin + 1u < TIME_BINS ? min(n - 1u, pbyte(c, bw, bins + bin + 1u) + 1u) : n - 1u;
    hi = max(hi, lo + 1u);
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

// controlCoord / locateCoord (vbt_sample.comp:203-221); at resolution 0 the
// shader's scan would run for 2^32 steps, here it stops at once
__device__ __forceinline__ float control_coord(uint32_t i, uint32_t res)
{
    if (res <= 1u) return 0.0f;
    return (float)i * (7.0f / (float)(res - 1u));
}
__device__ __forceinline__ void locate_coord(float p, uint32_t extent, uint32_t res, uint32_t& i0, uint32_t& i1,
                                             float& w)
{
    const float maxP = gmax((float)extent - 1.0f, 0.0f);
    const float cl = gclamp(p, 0.0f, maxP);
    uint32_t j = 1u;
    if (res >= 2u)
        while (j < res - 1u && control_coord(j, res) < cl) ++j;
    i1 = j;
    i0 = j - 1u;
    const float c0 = control_coord(i0, res), c1 = control_coord(i1, res);
    w = c1 > c0 ? (cl - c0) / (c1 - c0) : 0.0f;
}

// currentFrameControl (vbt_sample.comp:228-247)
__device__ __forceinline__ float frame_control(const MarchCtx& c, uint32_t cl, uint32_t slotBase, uint32_t bw,
                                               uint32_t streamBase, uint32_t codec, uint32_t count, uint32_t keys,
                                               uint32_t ci)
{
    if (slotBase == 0u && count > COARSE_CACHE) return decode_temporal(c, bw, streamBase, codec, count, keys, ci);
    return c.cache[(size_t)cl * STRIDE + slotBase + ci];
}

// sampleControlStream (vbt_sample.comp:249-289)
__device__ float sample_stream(const MarchCtx& c, uint32_t cl, uint32_t slotBase, uint32_t bw, uint32_t streamBase,
                               uint32_t codec, uint32_t res, uint32_t keys, uint32_t lw, uint32_t lh, uint32_t ld,
                               float lx, float ly, float lz)
{
    uint32_t x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    locate_coord(lx, lw, res, x0, x1, tx);
    locate_coord(ly, lh, res, y0, y1, ty);
    locate_coord(lz, ld, res, z0, z1, tz);
    const uint32_t n = res * res * res;
    auto at = [&](uint32_t x, uint32_t y, uint32_t z) {
        return frame_control(c, cl, slotBase, bw, streamBase, codec, n, keys, (z * res + y) * res + x);
    };
    const float v000 = at(x0, y0, z0), v100 = at(x1, y0, z0), v010 = at(x0, y1, z0), v110 = at(x1, y1, z0);
    const float v001 = at(x0, y0, z1), v101 = at(x1, y0, z1), v011 = at(x0, y1, z1), v111 = at(x1, y1, z1);
    const float c00 = gmix(v000, v100, tx), c10 = gmix(v010, v110, tx);
    const float c01 = gmix(v001, v101, tx), c11 = gmix(v011, v111, tx);
    const float c0 = gmix(c00, c10, ty), c1 = gmix(c01, c11, ty);
    return gmix(c0, c1, tz);
}

// sampleShellResidual (vbt_sample.comp:298-320)
__device__ __forceinline__ float shell_residual(const MarchCtx& c, uint32_t bw, uint32_t occBase, uint32_t cl,
                                                uint32_t vox, bool& occupied)
{
    const uint32_t group = vox >> 6, bit = vox & 63u, wordInGroup = bit >> 5, localBit = bit & 31u;
    const uint32_t mask = pu32(c, bw, occBase + group * 8u + wordInGroup * 4u);
    occupied = ((mask >> localBit) & 1u) != 0u;
    if (!occupied) return 0.0f;
    const uint32_t prefix = pu16(c, bw, occBase + 64u + group * 2u);
    const uint32_t lowWord = pu32(c, bw, occBase + group * 8u);
    uint32_t di = prefix;
    if (wordInGroup != 0u) di += (uint32_t)__popc(lowWord);
    const uint32_t before = localBit == 0u ? 0u : (0xffffffffu >> (32u - localBit));
    di += (uint32_t)__popc(mask & before);
    return c.cache[(size_t)cl * STRIDE + COARSE_CACHE + di];
}

// sampleVbtValue (vbt_sample.comp:322-415)
__device__ float sample_value(const MarchCtx& c, float wx, float wy, float wz, bool preserveInterior)
{
    const float bg = c.p.renderMode == 1u ? c.p.backgroundValue : 0.0f;
    const float px = wx - c.p.bboxMin[0], py = wy - c.p.bboxMin[1], pz = wz - c.p.bboxMin[2];
    if (px < 0.0f || py < 0.0f || pz < 0.0f || px > (float)(c.dimX - 1u) || py > (float)(c.dimY - 1u) ||
        pz > (float)(c.dimZ - 1u))
        return bg;
    const uint32_t x = min(c.dimX - 1u, (uint32_t)floorf(px));
    const uint32_t y = min(c.dimY - 1u, (uint32_t)floorf(py));
    const uint32_t z = min(c.dimZ - 1u, (uint32_t)floorf(pz));
    const uint32_t bx = min(c.lcx - 1u, x / c.leafSize), by = min(c.lcy - 1u, y / c.leafSize);
    const uint32_t bz = min(c.lcz - 1u, z / c.leafSize);
    const uint32_t leaf = (bz * c.lcy + by) * c.lcx + bx;
    const uint32_t wb = c.offsets[leaf], we = c.offsets[leaf + 1u];
    if (we <= wb) return bg;
    const uint32_t hdr = c.words[wb];
    const uint32_t mode = hdr & 3u;
    if (mode == MODE_EMPTY) return bg;
    const uint32_t cl = c.compact[leaf];
    const uint32_t cc = (hdr >> 2) & 3u, fc = (hdr >> 4) & 3u, cr = (hdr >> 6) & 7u, fr = (hdr >> 9) & 7u;
    const uint32_t prefix = c.words[wb + 1u];
    const uint32_t ckc = prefix & 0xFFFFu, fkc = prefix >> 16;
    const uint32_t ccount = cr * cr * cr;
    const uint32_t fineBase = 8u + ccount * 4u + ((ccount * TIME_BINS + 3u) & ~3u) +
                              ((ckc * (cc == 1u ? 8u : 3u) + 3u) & ~3u);
    const uint32_t baseX = bx * c.leafSize, baseY = by * c.leafSize, baseZ = bz * c.leafSize;
    const uint32_t lw = min(c.leafSize, c.dimX - baseX), lh = min(c.leafSize, c.dimY - baseY);
    const uint32_t ld = min(c.leafSize, c.dimZ - baseZ);
    const float lx = px - (float)baseX, ly = py - (float)baseY, lz = pz - (float)baseZ;
    float v = sample_stream(c, cl, 0u, wb, 8u, cc, cr, ckc, lw, lh, ld, lx, ly, lz);
    if (mode == MODE_FINE6 && fr > 0u) {
        v += sample_stream(c, cl, COARSE_CACHE, wb, fineBase, fc, fr, fkc, lw, lh, ld, lx, ly, lz);
    } else if (mode == MODE_SHELL) {
        const uint32_t ux = (uint32_t)gclamp(roundf(lx), 0.0f, 7.0f);
        const uint32_t uy = (uint32_t)gclamp(roundf(ly), 0.0f, 7.0f);
        const uint32_t uz = (uint32_t)gclamp(roundf(lz), 0.0f, 7.0f);
        bool occ = false;
        v += shell_residual(c, wb, fineBase, cl, (uz * 8u + uy) * 8u + ux, occ);
        if (!preserveInterior && !occ && v < 0.0f) v = bg;
    }
    return v;
}

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 ray_at(const V3& o, const V3& d, float t)
{
    const float ax = d.x * t, ay = d.y * t, az = d.z * t;
    return V3{o.x + ax, o.y + ay, o.z + az};
}
__device__ __forceinline__ float sample_at(const MarchCtx& c, const V3& o, const V3& d, float t, bool keep)
{
    const V3 q = ray_at(o, d, t);
    return sample_value(c, q.x, q.y, q.z, keep);
}

// intersectWorldBox (vbt_sample.comp:417-428)
__device__ bool intersect_box(const MarchCtx& c, const V3& o, const V3& d, float& nearT, float& farT)
{
    const float mn[3] = {c.p.bboxMin[0], c.p.bboxMin[1], c.p.bboxMin[2]};
    const float mx[3] = {mn[0] + (float)(c.dimX - 1u), mn[1] + (float)(c.dimY - 1u), mn[2] + (float)(c.dimZ - 1u)};
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float tmin[3], tmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float inv = 1.0f / dd[k];
        const float a = (mn[k] - oo[k]) * inv, b = (mx[k] - oo[k]) * inv;
        tmin[k] = gmin(a, b);
        tmax[k] = gmax(a, b);
    }
    nearT = gmax(0.0f, gmax(tmin[0], gmax(tmin[1], tmin[2])));
    farT = gmin(tmax[0], gmin(tmax[1], tmax[2]));
    return farT > nearT;
}

// traceShadow (vbt_sample.comp:430-452)
__device__ float trace_shadow(const MarchCtx& c, const V3& o, const V3& d)
{
    float nearT, farT;
    if (!intersect_box(c, o, d, nearT, farT)) return 0.0f;
    const float eps = gmax(c.p.surfaceEpsilonVoxels * c.p.voxelSize, 1e-5f);
    float t = nearT;
    float prev = sample_at(c, o, d, t, false) - c.p.surfaceIso;
    for (uint32_t it = 0; it < c.p.shadowSteps; ++it) {
        const float step = gclamp(fabsf(prev) / gmax(c.p.voxelSize, 1e-6f) * 0.8f, 0.35f, 5.0f);
        const float nextT = gmin(t + step, farT);
        const float next = sample_at(c, o, d, nextT, false) - c.p.surfaceIso;
        if ((prev < 0.0f) != (next < 0.0f)) {
            const bool nextInside = next < 0.0f;
            const float neg = nextInside ? next : prev;
            const float supT = nextInside ? gmin(nextT + 0.35f, farT) : gmax(t - 0.35f, nearT);
            const float sup = sample_at(c, o, d, supT, false) - c.p.surfaceIso;
            if (neg <= -eps * 0.5f || sup <= -eps * 0.25f) return 1.0f;
        }
        if (nextT >= farT) break;
        t = nextT;
        prev = fabsf(next) <= eps ? eps : next;
    }
    return 0.0f;
}

// estimateThickness (vbt_sample.comp:454-474)
__device__ float thickness(const MarchCtx& c, const V3& o, const V3& d, float tmax, float hitT)
{
    const float eps = gmax(c.p.surfaceEpsilonVoxels * c.p.voxelSize, 1e-5f);
    float t = gmin(hitT + 0.35f, tmax);
    float prev = sample_at(c, o, d, t, true) - c.p.surfaceIso;
    bool entered = prev < -0.25f * eps;
    const uint32_t limit = min(c.p.stepCount, 160u);
    for (uint32_t it = 0; it < limit; ++it) {
        const float step = gclamp(fabsf(prev) / gmax(c.p.voxelSize, 1e-6f) * 0.7f, 0.3f, 3.0f);
        const float nextT = gmin(t + step, tmax);
        const float next = sample_at(c, o, d, nextT, true) - c.p.surfaceIso;
        entered = entered || next < -0.25f * eps;
        if (entered && prev < 0.0f && next >= 0.0f)
            return gmax((nextT - hitT) * c.p.voxelSize, c.p.voxelSize * 0.25f);
        if (nextT >= tmax) break;
        t = nextT;
        prev = next;
    }
    return c.p.voxelSize * 6.0f;
}

// traceLevelSet (vbt_sample.comp:476-551): out = primary[4], auxiliary[4]
__device__ void trace_level_set(const MarchCtx& c, const V3& o, const V3& d, float tmax, float* out)
{
    out[0] = out[1] = out[2] = 0.0f;
    out[3] = -1.0f;
    out[4] = out[5] = out[6] = out[7] = 0.0f;
    if (tmax <= 0.0f) return;
    const float eps = gmax(c.p.surfaceEpsilonVoxels * c.p.voxelSize, 1e-5f);
    float prevT = 0.0f;
    float prev = sample_value(c, o.x, o.y, o.z, false) - c.p.surfaceIso;
    float hitT = -1.0f;
    for (uint32_t it = 0; it < c.p.stepCount; ++it) {
        const float step =
            gclamp(fabsf(prev) / gmax(c.p.voxelSize, 1e-6f) * 0.8f, 0.2f, gmax(c.p.maxStepVoxels, 0.2f));
        const float curT = gmin(prevT + step, tmax);
        const float cur = sample_at(c, o, d, curT, false) - c.p.surfaceIso;
        if ((prev < 0.0f) != (cur < 0.0f)) {
            const bool curInside = cur < 0.0f;
            const float neg = curInside ? cur : prev;
            const float supT = curInside ? gmin(curT + 0.35f, tmax) : gmax(prevT - 0.35f, 0.0f);
            const float sup = sample_at(c, o, d, supT, false) - c.p.surfaceIso;
            if (!(neg <= -eps * 0.5f || sup <= -eps * 0.25f)) {
                if (curT >= tmax) break;
                prevT = curT;
                prev = gmax(cur, eps);
                continue;
            }
            float low = prevT, high = curT, lowPhi = prev;
            for (int r = 0; r < 5; ++r) {
                const float mid = 0.5f * (low + high);
                const float midPhi = sample_at(c, o, d, mid, false) - c.p.surfaceIso;
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
    const V3 h = ray_at(o, d, hitT);
    const float ns = gmax(c.p.surfaceNormalStep, 0.25f);
    const float g0 = sample_value(c, h.x + ns, h.y, h.z, false) - sample_value(c, h.x - ns, h.y, h.z, false);
    const float g1 = sample_value(c, h.x, h.y + ns, h.z, false) - sample_value(c, h.x, h.y - ns, h.z, false);
    const float g2 = sample_value(c, h.x, h.y, h.z + ns, false) - sample_value(c, h.x, h.y, h.z - ns, false);
    const float s0 = g0 * g0, s1 = g1 * g1, s2 = g2 * g2;
    const float len = sqrtf((s0 + s1) + s2);
    V3 n;
    if (len > 1e-6f) {
        n = V3{g0 / len, g1 / len, g2 / len};
    } else {
        n = V3{-d.x, -d.y, -d.z};
    }
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
    out[3] = hitT;
    out[4] = c.p.surfaceModel != 0u ? thickness(c, o, d, tmax, hitT) : 0.0f;
    const float e0 = n.x * -d.x, e1 = n.y * -d.y, e2 = n.z * -d.z;
    out[7] = fabsf((e0 + e1) + e2);
}

// main (vbt_sample.comp:553-622): one thread per ray
__global__ void __launch_bounds__(64) k_march(MarchCtx c, const float4* rays, uint32_t n, float4* out)
{
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = rays[2u * i], r1 = rays[2u * i + 1u];
    const V3 o{r0.x, r0.y, r0.z}, d{r1.x, r1.y, r1.z};
    float res[8];
    if (c.p.renderMode == 1u) {
        trace_level_set(c, o, d, r1.w, res);
        const bool zUp = c.p.upAxis == 2u;
        const float fd = zUp ? d.z : d.y, fo = zUp ? o.z : o.y;
        if (c.p.surfaceModel != 0u && res[3] < 0.0f && fd < -1e-5f) {
            const float floorT = (c.p.floorHeight - fo) / fd;
            if (floorT > 0.0f) {
                const V3 fp = ray_at(o, d, floorT);
                const float k0 = -0.55f, k1 = zUp ? 0.42f : 0.72f, k2 = zUp ? 0.72f : 0.42f;
                const float q0 = k0 * k0, q1 = k1 * k1, q2 = k2 * k2;
                const float kl = sqrtf((q0 + q1) + q2);
                const V3 kd{k0 / kl, k1 / kl, k2 / kl};
                const V3 so = ray_at(fp, kd, 0.5f);
                res[5] = 1.0f;
                res[6] = trace_shadow(c, so, kd);
            }
        }
    } else {
        const float dt = r1.w / gmax((float)c.p.stepCount, 1.0f);
        float t = r0.w + 0.5f * dt;
#pragma unroll
        for (int k = 0; k < 8; ++k) res[k] = 0.0f;
        if (c.p.surfaceModel != 0u) {
            const uint32_t sc = max(c.p.stepCount, 1u);
            for (uint32_t s = 0; s < c.p.stepCount; ++s) {
                const float v = sample_at(c, o, d, t, false);
                const uint32_t bin = min((s * 8u) / sc, 7u);
                // the shader's if-chain over eight accumulators
#pragma unroll
                for (uint32_t k = 0; k < 8u; ++k)
                    if (bin == k) res[k] += v;
                t += dt;
            }
        } else {
            float acc = 0.0f;
            for (uint32_t s = 0; s < c.p.stepCount; ++s) {
                acc += sample_at(c, o, d, t, false);
                t += dt;
            }
            res[0] = acc;
        }
    }
    out[2u * i] = make_float4(res[0], res[1], res[2], res[3]);
    out[2u * i + 1u] = make_float4(res[4], res[5], res[6], res[7]);
}

__global__ void __launch_bounds__(256) k_sample(MarchCtx c, const float* xyz, uint32_t n, int keep, float* out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    out[i] = sample_value(c, xyz[3u * i], xyz[3u * i + 1u], xyz[3u * i + 2u], keep != 0);
}

int make_ctx(vbt_asset a, const float* d_cache, const vbt_march_params* p, MarchCtx& c)
{
    if (!a || !d_cache || !p) return vbt_fail(VBT_ERR_INVALID, "march: null argument");
    const vbt_file_header* h = vbt_internal_asset_header(a);
    if (h->profileType == 2u) return vbt_fail(VBT_ERR_INVALID, "march: render assets only (profile 1)");
    if (vbt_internal_asset_bad(a))
        return vbt_internal_report_bad(a, "march");
    const AssetDev ad = vbt_internal_asset_dev(a);
    c.offsets = ad.offsets;
    c.words = (const uint32_t*)ad.payload;
    c.compact = ad.compact;
    c.cache = d_cache;
    c.dimX = h->width;
    c.dimY = h->height;
    c.dimZ = h->depth;
    c.frames = h->frames;
    c.leafSize = h->leafSize;
    c.lcx = (c.dimX + c.leafSize - 1u) / c.leafSize;
    c.lcy = (c.dimY + c.leafSize - 1u) / c.leafSize;
    c.lcz = (c.dimZ + c.leafSize - 1u) / c.leafSize;
    c.p = *p;
    return VBT_OK;
}

} // namespace

extern "C" {

int vbt_sample_points(vbt_asset a, const float* d_cache576, const vbt_march_params* p, const float* d_xyz,
                      uint32_t n, int preserveInterior, float* d_out, void* stream)
{
    MarchCtx c;
    int rc = make_ctx(a, d_cache576, p, c);
    if (rc) return rc;
    if (n == 0u) return VBT_OK;
    if (!d_xyz || !d_out) return vbt_fail(VBT_ERR_INVALID, "vbt_sample_points: null buffer");
    if (hipSetDevice(vbt_internal_asset_device(a)) != hipSuccess) return vbt_fail(VBT_ERR_DEVICE, "hipSetDevice");
    k_sample<<<(n + 255u) / 256u, 256, 0, (hipStream_t)stream>>>(c, d_xyz, n, preserveInterior, d_out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? VBT_OK : vbt_fail(VBT_ERR_DEVICE, hipGetErrorString(e));
}

int vbt_march_rays(vbt_asset a, const float* d_cache576, const vbt_march_params* p, const float* d_rays, uint32_t n,
                   float* d_out, void* stream)
{
    MarchCtx c;
    int rc = make_ctx(a, d_cache576, p, c);
    if (rc) return rc;
    if (n == 0u) return VBT_OK;
    if (!d_rays || !d_out) return vbt_fail(VBT_ERR_INVALID, "vbt_march_rays: null buffer");
    if ((uintptr_t)d_rays % 16u || (uintptr_t)d_out % 16u)
        return vbt_fail(VBT_ERR_INVALID, "vbt_march_rays: rays and results must be 16-byte aligned");
    if (hipSetDevice(vbt_internal_asset_device(a)) != hipSuccess) return vbt_fail(VBT_ERR_DEVICE, "hipSetDevice");
    k_march<<<(n + 63u) / 64u, 64, 0, (hipStream_t)stream>>>(c, (const float4*)d_rays, n, (float4*)d_out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? VBT_OK : vbt_fail(VBT_ERR_DEVICE, hipGetErrorString(e));
}

} // extern "C"
