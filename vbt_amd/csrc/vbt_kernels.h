// vbt_kernels.h -- device-side decode building blocks shared by the one-shot
// kernels (decode.hip) and the serving work-queue kernel (server.hip).
#pragma once

#include "vbt_format.h"
#include "vbt_internal.h"

#include <hip/hip_runtime.h>

namespace vbtk {

constexpr uint32_t FLAG_SWEEP = 1u << 8;   // every control's keyframe frames strictly increase
constexpr uint32_t FLAG_NO_NAN = 1u << 9;  // no fp16 keyframe value of the leaf is a NaN pattern
constexpr uint32_t STATUS_MASK = 0xFFu;
constexpr uint32_t NO_LEAF = 0xffffffffu;
constexpr uint32_t CACHE_GROUP = 64;      // packed-cache slots per group-table entry

struct DevLeaf {
    uint32_t leaf, word, bytes, status;
    uint32_t ncache, cc;       // cache slots; coarse slots
    uint32_t lw, lh, ld;       // cropped extent
    uint32_t sciHeader;        // scientific header word
    LeafPlan plan;
};

// Everything a kernel needs to address one resident asset.
struct AssetDev {
    const uint32_t* offsets;
    const uint8_t* payload;
    const float* scales;
    const uint32_t* compact;
    const DevLeaf* leaves;
    const uint64_t* ctrl;
    const float* basis;
    const uint32_t* groupLeaf;  // first active leaf of each CACHE_GROUP-slot group
    uint32_t W, H, D, F, L, A, profile, scaleCount;
    uint64_t payloadBytes;
    uint64_t controlsPerFrame;
};

struct ComposeDev {
    int signRule;
    float background, shellEmptyScale, shellEmptyMax;
};

__host__ __device__ inline void leaf_coords(uint32_t leaf, uint32_t W, uint32_t H, uint32_t D, uint32_t& bx,
                                            uint32_t& by, uint32_t& bz, uint32_t& lw, uint32_t& lh, uint32_t& ld)
{
    const uint32_t lcx = (W + 7u) / 8u, lcy = (H + 7u) / 8u;
    bx = leaf % lcx;
    by = (leaf / lcx) % lcy;
    bz = leaf / (lcx * lcy);
    lw = W - bx * 8u < 8u ? W - bx * 8u : 8u;
    lh = H - by * 8u < 8u ? H - by * 8u : 8u;
    ld = D - bz * 8u < 8u ? D - bz * 8u : 8u;
}

__device__ inline uint32_t find_leaf(const uint64_t* ctrl, uint32_t A, uint64_t g)
{
    uint32_t lo = 0, hi = A; // ctrl[lo] <= g < ctrl[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ctrl[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------ scientific decode --
__device__ inline float sci_control(const AssetDev& a, const Bytes& b, const SciHeader& h, uint32_t ci, int t)
{
    const uint32_t cb = 4u + ci * (h.keep + 1u);
    const float* basis = a.basis + (size_t)t * 64u;
    double sum = (double)half_to_float(b.u16(cb)) * (double)basis[0];
    for (uint32_t k = 1; k < h.keep; ++k) {
        const float coeff = (float)(int8_t)(uint8_t)b.u8(cb + 2u + (k - 1u)) * a.scales[k - 1u];
        const double term = (double)coeff * (double)basis[k];
        sum += term;
    }
    return (float)sum;
}

__device__ inline uint32_t bfp_code(const Bytes& b, uint32_t blockBase, uint32_t ci, uint32_t bits)
{
    const uint32_t bit = ci * bits;
    const uint32_t byteIndex = blockBase + bit / 8u;
    const uint32_t shift = bit % 8u;
    uint32_t packed = b.u8(byteIndex);
    if (shift + bits > 8u) packed |= b.u8(byteIndex + 1u) << 8;
    return (packed >> shift) & ((1u << bits) - 1u);
}

__device__ inline float sci_dense_value(const AssetDev& a, const Bytes& b, const SciHeader& h, uint32_t lx, uint32_t ly,
                                        uint32_t lz, uint32_t t)
{
    const int R = h.mode == 2u ? 3 : 4;
    const uint32_t bits = h.bits;
    const uint32_t vals = (uint32_t)(R * R * R);
    const uint32_t blockBytes = (vals * bits + 7u) / 8u;
    const uint32_t scalesBase = 4u + 64u * (h.keep + 1u);
    int x0, y0, z0;
    float tx, ty, tz;
    sci_dense_coord(lx, R, x0, tx);
    sci_dense_coord(ly, R, y0, ty);
    sci_dense_coord(lz, R, z0, tz);
    if (h.subtype == 1u) {
        if (h.tk == 0u || bits == 0u) return 0.0f;
        const uint32_t dataBase = scalesBase + h.tk * 2u;
        float residual = 0.0f;
        for (uint32_t k = 0; k < h.tk; ++k) {
            const float scale = half_to_float(b.u16(scalesBase + k * 2u));
            const uint32_t bb = dataBase + k * blockBytes;
            float s[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int cx = x0 + (i & 1), cy = y0 + ((i >> 1) & 1), cz = z0 + ((i >> 2) & 1);
                s[i] = decode_bfp(bfp_code(b, bb, (uint32_t)((cz * R + cy) * R + cx), bits), scale, (int)bits);
            }
            residual += trilinear(s, tx, ty, tz) * a.basis[(size_t)t * 64u + k];
        }
        return residual;
    }
    const uint32_t dataBase = scalesBase + a.F * 2u;
    const float scale = half_to_float(b.u16(scalesBase + t * 2u));
    const uint32_t frameBase = dataBase + t * blockBytes;
    float s[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int cx = x0 + (i & 1), cy = y0 + ((i >> 1) & 1), cz = z0 + ((i >> 2) & 1);
        s[i] = decode_bfp(bfp_code(b, frameBase, (uint32_t)((cz * R + cy) * R + cx), bits), scale, (int)bits);
    }
    return trilinear(s, tx, ty, tz);
}

__device__ inline float sci_sparse(const Bytes& b, const SciHeader& h, uint32_t idx, uint32_t t)
{
    if (h.count == 0u) return 0.0f;
    const uint32_t cap = h.tier == 0 ? 64u : h.tier == 1 ? 256u : h.tier == 2 ? 768u : 1024u;
    const uint32_t bin = t / h.fpb < 15u ? t / h.fpb : 15u;
    const uint32_t localT = (t - bin * h.fpb) & 0xFFu;
    const uint32_t sparseBase = 4u + 64u * (h.keep + 1u);
    const float eventScale = bits_to_float(b.u32(sparseBase));
    if (!(eventScale > 0.0f)) return 0.0f;
    const uint32_t maskWord = b.u32(sparseBase + 4u + (idx >> 5) * 4u);
    if ((maskWord & (1u << (idx & 31u))) == 0u) return 0.0f;
    const uint32_t binsBase = sparseBase + 68u;
    const uint32_t start = b.u16(binsBase + bin * 2u);
    const uint32_t end = bin + 1u < 16u ? b.u16(binsBase + (bin + 1u) * 2u) : (h.count & 0xFFFFu);
    if (start >= end) return 0.0f;
    const uint32_t coordsBase = binsBase + 32u;
    const uint32_t residualBase = coordsBase + cap * 2u;
    const uint32_t target = ((localT << 9) | (idx & 0x1FFu)) & 0xFFFFu;
    int left = (int)start, right = (int)end - 1;
    while (left <= right) {
        const int mid = (left + right) >> 1;
        const uint32_t key = b.u16(coordsBase + (uint32_t)mid * 2u);
        if (key == target) {
            const uint32_t byte = b.u8(residualBase + (uint32_t)(mid >> 1));
            const uint32_t nib = (mid & 1) == 0 ? (byte & 0xFu) : ((byte >> 4) & 0xFu);
            int q = (int)nib;
            if (q & 8) q -= 16;
            return (float)q * eventScale;
        }
        if (key < target) left = mid + 1;
        else right = mid - 1;
    }
    return 0.0f;
}

__device__ inline float sci_residual(const AssetDev& a, const Bytes& b, const SciHeader& h, uint32_t lx, uint32_t ly,
                                     uint32_t lz, uint32_t t)
{
    if (h.mode == 0u) return 0.0f;
    if (h.mode == 1u) return sci_sparse(b, h, (lz * 8u + ly) * 8u + lx, t);
    return sci_dense_value(a, b, h, lx, ly, lz, t);
}

__device__ inline float sci_coarse_sample(const float* ctrl, uint32_t lx, uint32_t ly, uint32_t lz)
{
    int x0, y0, z0;
    float tx, ty, tz;
    sci_coarse_coord(lx, x0, tx);
    sci_coarse_coord(ly, y0, ty);
    sci_coarse_coord(lz, z0, tz);
    float s[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int cx = x0 + (i & 1), cy = y0 + ((i >> 1) & 1), cz = z0 + ((i >> 2) & 1);
        s[i] = ctrl[(cz * 4 + cy) * 4 + cx];
    }
    return trilinear(s, tx, ty, tz);
}

// ------------------------------------------------------ render voxels --
__device__ inline float apply_sign(const ComposeDev& cm, float v, bool occupied)
{
    if (cm.signRule == VBT_SIGN_NONE) {
        if (!isnan(cm.shellEmptyMax) && cm.shellEmptyScale < 1.0f && !occupied && v > 0.0f && v <= cm.shellEmptyMax)
            v *= cm.shellEmptyScale;
    } else if (!occupied && v < 0.0f) {
        v = cm.signRule == VBT_SIGN_CLAMP_ZERO ? 0.0f : cm.background;
    }
    return v;
}

// shell residual of a voxel (decodeRenderTemporalLeafShellVoxelValue semantics)
__device__ inline float shell_value(const Bytes& b, const DevLeaf& d, uint32_t vox, int F, int t, bool& occupied)
{
    const uint32_t di = shell_desc_index(b, d.plan.occOffset, vox);
    occupied = di != NO_LEAF;
    if (!occupied || d.plan.shell.keys == 0u) return 0.0f;
    return decode_control(b, d.plan.shell, di, F, t);
}

// render point value: 8 corners per grid decoded by the LANES lanes of a group
template <int LANES>
__device__ inline float render_point(const AssetDev& a, const DevLeaf& d, const ComposeDev& cm, uint32_t lx, uint32_t ly,
                                     uint32_t lz, int t, uint32_t sub)
{
    const Bytes b{a.payload + (size_t)d.word * 4u};
    const int F = (int)a.F;
    float v = 0.0f;
    for (int pass = 0; pass < 2; ++pass) {
        const int res = pass == 0 ? (int)d.plan.coarseRes : (int)d.plan.fineRes;
        if (pass == 1 && !d.plan.hasFine) break;
        if (res < 1) continue;
        const StreamRef& s = pass == 0 ? d.plan.coarse : d.plan.fine;
        int x0, x1, y0, y1, z0, z1;
        float tx, ty, tz;
        locate((float)lx, (int)d.lw, res, x0, x1, tx);
        locate((float)ly, (int)d.lh, res, y0, y1, ty);
        locate((float)lz, (int)d.ld, res, z0, z1, tz);
        float sv[8];
        if (LANES == 8) {
            const int cx = (sub & 1u) ? x1 : x0, cy = (sub & 2u) ? y1 : y0, cz = (sub & 4u) ? z1 : z0;
            const float mine = decode_control(b, s, (uint32_t)((cz * res + cy) * res + cx), F, t);
#pragma unroll
            for (int i = 0; i < 8; ++i) sv[i] = __shfl(mine, i, 8);
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int cx = (i & 1) ? x1 : x0, cy = (i & 2) ? y1 : y0, cz = (i & 4) ? z1 : z0;
                sv[i] = decode_control(b, s, (uint32_t)((cz * res + cy) * res + cx), F, t);
            }
        }
        const float g = trilinear(sv, tx, ty, tz);
        v = pass == 0 ? g : v + g;
    }
    if (d.plan.mode == MODE_SHELL) {
        const uint32_t vox = (lz * 8u + ly) * 8u + lx;
        bool occ = false;
        if (!d.plan.hasFine) {
            v += shell_value(b, d, vox, F, t, occ);
        } else {
            occ = shell_desc_index(b, d.plan.occOffset, vox) != NO_LEAF;
        }
        v = apply_sign(cm, v, occ);
    }
    return v;
}

template <int LANES>
__device__ inline float sci_point(const AssetDev& a, const DevLeaf& d, uint32_t lx, uint32_t ly, uint32_t lz, int t,
                                  uint32_t sub)
{
    const Bytes b{a.payload + (size_t)d.word * 4u};
    const SciHeader h = sci_header(d.sciHeader);
    int x0, y0, z0;
    float tx, ty, tz;
    sci_coarse_coord(lx, x0, tx);
    sci_coarse_coord(ly, y0, ty);
    sci_coarse_coord(lz, z0, tz);
    float sv[8];
    if (LANES == 8) {
        const int cx = x0 + (int)(sub & 1u), cy = y0 + (int)((sub >> 1) & 1u), cz = z0 + (int)((sub >> 2) & 1u);
        const float mine = sci_control(a, b, h, (uint32_t)((cz * 4 + cy) * 4 + cx), t);
#pragma unroll
        for (int i = 0; i < 8; ++i) sv[i] = __shfl(mine, i, 8);
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int cx = x0 + (i & 1), cy = y0 + ((i >> 1) & 1), cz = z0 + ((i >> 2) & 1);
            sv[i] = sci_control(a, b, h, (uint32_t)((cz * 4 + cy) * 4 + cx), t);
        }
    }
    return trilinear(sv, tx, ty, tz) + sci_residual(a, b, h, lx, ly, lz, (uint32_t)t);
}

// One point query (all LANES lanes of its group take part in the shuffles).
// Returns the value; st receives 0 or a status code.
template <int LANES>
__device__ inline float point_query(const AssetDev& a, const ComposeDev& cm, uint4 r, uint32_t sub, int32_t& st)
{
    st = 0;
    if (r.x >= a.W || r.y >= a.H || r.z >= a.D) {
        st = VBT_QUERY_ERR_OUTSIDE;
        return 0.0f;
    }
    if (a.profile == 2u && r.w >= a.F) {
        st = VBT_QUERY_ERR_FRAME;
        return 0.0f;
    }
    const uint32_t lcx = (a.W + 7u) / 8u, lcy = (a.H + 7u) / 8u;
    const uint32_t leaf = ((r.z / 8u) * lcy + r.y / 8u) * lcx + r.x / 8u;
    const uint32_t c = a.compact[leaf];
    const int t = r.w > 0x7fffffffu ? 0x7fffffff : (int)r.w;
    if (c == NO_LEAF) return 0.0f;
    const DevLeaf& d = a.leaves[c];
    if ((d.status & STATUS_MASK) != LEAF_OK) {
        st = (int32_t)(d.status & STATUS_MASK);
        return 0.0f;
    }
    if (a.profile == 2u) return sci_point<LANES>(a, d, r.x & 7u, r.y & 7u, r.z & 7u, t, sub);
    return render_point<LANES>(a, d, cm, r.x & 7u, r.y & 7u, r.z & 7u, t, sub);
}

// One packed-cache slot of a render asset at one frame (reference bracket
// search); returns NaN and sets st for a poisoned leaf.
__device__ inline float cache_slot_exact(const AssetDev& a, uint64_t g, int t, int32_t& st)
{
    uint32_t c = a.groupLeaf[g / CACHE_GROUP];
    while (c + 1u < a.A && a.ctrl[c + 1u] <= g) ++c;
    const DevLeaf& d = a.leaves[c];
    st = 0;
    if ((d.status & STATUS_MASK) != LEAF_OK) {
        st = (int32_t)(d.status & STATUS_MASK);
        return __builtin_nanf("");
    }
    const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
    const bool isCoarse = slot < d.cc;
    const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
    const Bytes b{a.payload + (size_t)d.word * 4u};
    return decode_control(b, s, isCoarse ? slot : slot - d.cc, (int)a.F, t);
}

// ------------------------------------------------- packed cache: render tile --
// Lerp weight of the sweep: alpha = (frame - f0) / (f1 - f0) for integer
// frames 0 <= frame - f0 < f1 - f0 < 2^16.  With r = RN64(1 / (f1 - f0)),
// RN32(RN64((frame - f0) * r)) equals the correctly rounded float quotient the
// reference computes: the double product is within 2^-52 (relative) of the
// true quotient, while a quotient a/b with b < 2^16 lies at least 2^-41
// (relative) away from every float rounding boundary, so both round alike.
__device__ inline float sweep_alpha(int num, double rinv) { return (float)((double)num * rinv); }

// One thread per (active leaf, cache slot); the thread produces that control
// for frames t0..t0+nt-1.  The workgroup first locates its leaf range in the
// page table (two binary searches per workgroup) and stages the leaves'
// payload bytes -- contiguous in the pool because active leaves are in leaf
// order -- into LDS with dword loads, so keyframe reads hit LDS.  Leaves
// whose every control has strictly increasing keyframe frames use a forward
// sweep over the keyframes: each frame's bracket is then the adjacent key pair
// around it, which is exactly the bracket the reference's hinted binary
// search returns for increasing keys.  Other leaves run the reference
// bracket search per frame.

// Branch-free form of half_to_float for the sweep: the hardware conversion is
// exact for zeros, subnormals, normals and infinities; NaNs take the
// reference's bit mapping (payload shifted, no quieting).  Leaves validated
// free of NaN patterns (FLAG_NO_NAN) skip the NaN mapping.
template <bool NANFIX>
__device__ __forceinline__ float sweep_half(uint32_t h)
{
    _Float16 x;
    const uint16_t hb = (uint16_t)h;
    __builtin_memcpy(&x, &hb, 2);
    const float f = (float)x;
    if (!NANFIX) return f;
    const uint32_t nanBits = ((h & 0x8000u) << 16) | 0x7F800000u | ((h & 0x3FFu) << 13);
    return ((h & 0x7C00u) == 0x7C00u && (h & 0x3FFu)) ? __uint_as_float(nanBits) : f;
}

template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void sweep_key(const uint8_t* p, uint32_t i, int& f, float& v)
{
    if (KB == 8u) {
        const uint8_t* q = p + i * 8u;
        f = (int)((uint32_t)q[0] | ((uint32_t)q[1] << 8));
        v = __uint_as_float((uint32_t)q[4] | ((uint32_t)q[5] << 8) | ((uint32_t)q[6] << 16) | ((uint32_t)q[7] << 24));
    } else {
        const uint8_t* q = p + i * 3u;
        f = (int)q[0];
        v = sweep_half<NANFIX>((uint32_t)q[1] | ((uint32_t)q[2] << 8));
    }
}

// The forward sweep of one control (strictly increasing key frames).  State:
// keys k, k+1 and k+2 (clamped to the last key) and RN64(1 / (f[k+1] - f[k])).
// Invariant after each frame: the next key's frame fn exceeds the frame just
// produced.  Frames advance by at most one per step, so a single predicated
// advance per step keeps the invariant (the key after the next is beyond fn
// because frames strictly increase); a catch-up establishes it for the first
// frame.  Each frame's bracket is then the adjacent key pair around it, the
// bracket the reference's hinted binary search returns for increasing keys.
// rtab[b] = RN64(1/b) for gaps b < 256.
struct SweepState {
    uint32_t k;
    int fk, fn, fp;
    float vk, vn, vp;
    double rinv;
};

template <uint32_t KB>
__device__ __forceinline__ double sweep_recip(int gap, const double* rtab)
{
    if (KB == 3u) return rtab[gap & 255];
    return gap < 256 ? rtab[gap] : 1.0 / (double)gap;
}

// State at the first requested frame, from the control's records in global
// memory (g points at key 0).
template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void sweep_start(const uint8_t* g, uint32_t last, int first, const double* rtab,
                                            SweepState& s)
{
    s.k = 0u;
    s.fn = 0;
    s.vn = 0.0f;
    sweep_key<KB, NANFIX>(g, 0u, s.fk, s.vk);
    if (last >= 1u) sweep_key<KB, NANFIX>(g, 1u, s.fn, s.vn);
    while (s.k < last && s.fn <= first) {
        ++s.k;
        s.fk = s.fn;
        s.vk = s.vn;
        if (s.k < last) sweep_key<KB, NANFIX>(g, s.k + 1u, s.fn, s.vn);
    }
    s.rinv = s.k < last ? sweep_recip<KB>(s.fn - s.fk, rtab) : 0.0;
    sweep_key<KB, NANFIX>(g, s.k + 2u <= last ? s.k + 2u : last, s.fp, s.vp);
}

// cn frames from tBegin on (out at the first; frames `stride` floats apart).
// rec holds the control's records r0..: record i at rec + (i - r0) * KB; the
// sweep reads records up to min(k + 2, last) with k <= r0 + cn.  The loop
// body is branch-free: the key after next is read every step and selected on
// advance.
template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void sweep_frames(const uint8_t* rec, uint32_t r0, uint32_t last, SweepState& s, int F,
                                             int tBegin, int cn, float* out, uint64_t stride, const double* rtab)
{
    const int fmax = F - 1 > 0 ? F - 1 : 0;
    uint32_t k = s.k;
    int fk = s.fk, fn = s.fn, fp = s.fp;
    float vk = s.vk, vn = s.vn, vp = s.vp;
    double rinv = s.rinv;
    for (int i = 0; i < cn; ++i) {
        const int tt = tBegin + i;
        const int frame = tt < fmax ? (tt < 0 ? 0 : tt) : fmax;
        const bool adv = (k < last) & (fn <= frame);
        const int gapNew = fp - fn;
        fk = adv ? fn : fk;
        vk = adv ? vn : vk;
        fn = adv ? fp : fn;
        vn = adv ? vp : vn;
        k += adv ? 1u : 0u;
        const double rNew = sweep_recip<KB>(gapNew > 0 ? gapNew : 1, rtab);
        rinv = adv ? rNew : rinv;
        sweep_key<KB, NANFIX>(rec, (k + 2u <= last ? k + 2u : last) - r0, fp, vp);
        const float alpha = sweep_alpha(frame - fk, rinv);
        const float lerp = vk + alpha * (vn - vk);
        out[(uint64_t)i * stride] = (fk >= frame || k == last) ? vk : lerp;
    }
    s.k = k;
    s.fk = fk;
    s.fn = fn;
    s.fp = fp;
    s.vk = vk;
    s.vn = vn;
    s.vp = vp;
    s.rinv = rinv;
}

template <uint32_t NW>
__device__ __forceinline__ uint32_t block_min(uint32_t v, uint32_t* sm, uint32_t lane, uint32_t wave)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor(v, off, 64));
    if (NW == 1) return v;
    if (lane == 0) sm[wave] = v;
    __syncthreads();
    v = sm[0];
#pragma unroll
    for (uint32_t w = 1; w < NW; ++w) v = min(v, sm[w]);
    __syncthreads();
    return v;
}

template <uint32_t NW>
__device__ __forceinline__ uint32_t block_max(uint32_t v, uint32_t* sm, uint32_t lane, uint32_t wave)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    if (NW == 1) return v;
    if (lane == 0) sm[wave] = v;
    __syncthreads();
    v = sm[0];
#pragma unroll
    for (uint32_t w = 1; w < NW; ++w) v = max(v, sm[w]);
    __syncthreads();
    return v;
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return v;
}

// One thread per (active leaf, cache slot): the thread produces that control
// for frames t0..t0+nt-1.  The workgroup reads its first leaf from the group
// table and each thread locates its leaf and stream.  Leaves whose every
// control has strictly increasing keyframe frames use the forward sweep;
// other leaves run the reference bracket search per frame from global memory.
// Keyframe records of consecutive controls are contiguous in the pool
// (consecutive controls of a stream store consecutive key ranges; active
// leaves are in leaf order), so the workgroup stages them into an LDS window
// of WINDOW bytes with coalesced dword loads, in as many passes as the window
// needs, and the lanes sweep from LDS.  A pass serves whole wavefronts whose
// records fit (so each such wavefront runs its frame loop once); a wavefront
// whose records exceed the window is served lane by lane over several passes.
// (Sweeping such wavefronts straight from global memory instead measured
// slower: the key read sits on the loop's dependency chain.)
// Tile: slots tile*BLOCK .. tile*BLOCK+BLOCK-1 of the packed cache; lds holds
// WINDOW bytes, sRed BLOCK/64 words; rtab[b] = RN64(1/b) (rtab[0] = 0) is
// filled before the first call.  BLOCK and WINDOW are the tile shape (the
// autotuner picks it for the one-shot kernel).
template <int LAYOUT, uint32_t BLOCK, uint32_t WINDOW>
__device__ __forceinline__ void cache_render_tile(const AssetDev& a, const uint32_t* groupLeaf, uint32_t tile, int t0,
                                                  int nt, float* out, uint64_t total, uint32_t* errFlag,
                                                  uint32_t* lds, uint32_t* sRed, const double* rtab)
{
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint64_t g = (uint64_t)tile * BLOCK + tid;
    const bool live = g < total;
    uint32_t c = groupLeaf[((uint64_t)tile * BLOCK) / CACHE_GROUP];
    if (live)
        while (c + 1u < a.A && a.ctrl[c + 1u] <= g) ++c;
    const DevLeaf& d = a.leaves[c];
    const uint32_t slot = live ? (uint32_t)(g - a.ctrl[c]) : 0u;
    const uint32_t status = d.status;
    const bool isCoarse = slot < d.cc;
    const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
    const uint32_t ci = isCoarse ? slot : slot - d.cc;
    const uint32_t leafByte = d.word * 4u;
    uint32_t keyStart = 0, keyCount = 0;
    const bool ok = live && (status & STATUS_MASK) == LEAF_OK;
    if (ok) {
        const uint32_t desc = *(const uint32_t*)(a.payload + leafByte + s.base + ci * 4u); // descriptors are 4-aligned
        keyStart = desc & 0xFFFFu;
        keyCount = (desc >> 16) & 0xFFu;
    }
    const uint32_t kb = keyframe_bytes(s.codec);
    const bool sweep = ok && keyCount > 0u && (status & FLAG_SWEEP);
    const uint32_t kLo = leafByte + s.keyBase + keyStart * kb;
    const uint32_t kHi = kLo + keyCount * kb;

    uint64_t stride = 0, dst = 0;
    bool writes = live;
    if (LAYOUT == VBT_CACHE_PACKED) {
        stride = total;
        dst = g;
    } else {
        const uint32_t s576 = isCoarse ? slot : 64u + (slot - d.cc);
        if ((isCoarse && slot >= 64u) || s576 >= 576u) writes = false;
        stride = (uint64_t)a.A * 576u;
        dst = (uint64_t)c * 576u + s576;
    }
    float* o = out + dst;
    const int F = (int)a.F;
    if (writes && !sweep) {
        if (!ok) {
            for (int i = 0; i < nt; ++i) o[(uint64_t)i * stride] = __builtin_nanf("");
            atomicOr(errFlag, 1u);
        } else if (keyCount == 0u) {
            for (int i = 0; i < nt; ++i) o[(uint64_t)i * stride] = 0.0f;
        } else {
            const Bytes b{a.payload + leafByte};
            const uint32_t binBase = s.base + s.count * 4u + ci * s.bins;
            for (int i = 0; i < nt; ++i)
                o[(uint64_t)i * stride] =
                    decode_sequence(b, s.keyBase, s.codec, keyStart, keyCount, binBase, s.bins, F, t0 + i);
        }
    }
    const bool sw = sweep && writes;
    const uint32_t mode = kb == 8u ? 0u : ((status & FLAG_NO_NAN) ? 1u : 2u);
    const uint32_t last = keyCount - 1u;
    const int first = clamp_frame(t0, F);
    auto run = [&](const uint8_t* rec) {
        SweepState st;
        if (mode == 0u) {
            sweep_start<8u, false>(rec, last, first, rtab, st);
            sweep_frames<8u, false>(rec, 0u, last, st, F, t0, nt, o, stride, rtab);
        } else if (mode == 1u) {
            sweep_start<3u, false>(rec, last, first, rtab, st);
            sweep_frames<3u, false>(rec, 0u, last, st, F, t0, nt, o, stride, rtab);
        } else {
            sweep_start<3u, true>(rec, last, first, rtab, st);
            sweep_frames<3u, true>(rec, 0u, last, st, F, t0, nt, o, stride, rtab);
        }
    };
    // a wavefront "fits" when its records fit in one window starting at the
    // dword holding its first record
    const uint32_t waveLo = wave_min(sw ? kLo : 0xFFFFFFFFu);
    const uint32_t waveHi = wave_max(sw ? kHi : 0u);
    const bool waveFits = waveLo != 0xFFFFFFFFu && waveHi - (waveLo & ~3u) <= WINDOW;
    bool pending = sw;
    for (;;) {
        const uint32_t lo = block_min<NW>(pending ? kLo : 0xFFFFFFFFu, sRed, lane, wave);
        if (lo == 0xFFFFFFFFu) break;
        const uint32_t w0 = lo / 4u;
        const uint32_t limit = w0 * 4u + WINDOW;
        // progress: the lowest pending record is always served (a single
        // control's records span at most 255 * 8 bytes < WINDOW)
        const bool mine = pending && ((waveFits ? waveHi <= limit : kHi <= limit) || kLo == lo);
        const uint32_t hi = block_max<NW>(mine ? kHi : 0u, sRed, lane, wave);
        const uint32_t nw = hi > w0 * 4u ? (hi + 3u) / 4u - w0 : 0u;
        // staging: eight dword loads in flight per lane before the LDS writes
        const uint32_t* pw = (const uint32_t*)a.payload + w0;
        for (uint32_t i0 = tid; i0 < nw; i0 += 8u * BLOCK) {
            uint32_t v[8];
#pragma unroll
            for (uint32_t j = 0; j < 8u; ++j) v[j] = i0 + j * BLOCK < nw ? pw[i0 + j * BLOCK] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < 8u; ++j)
                if (i0 + j * BLOCK < nw) lds[i0 + j * BLOCK] = v[j];
        }
        __syncthreads();
        if (mine) {
            run((const uint8_t*)lds + (kLo - w0 * 4u));
            pending = false;
        }
        __syncthreads();
    }
}

} // namespace vbtk
