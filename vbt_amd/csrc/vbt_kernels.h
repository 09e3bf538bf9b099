// vbt_kernels.h -- device-side decode building blocks shared by the one-shot
// kernels (decode.hip) and the serving work-queue kernel (server.hip).
#pragma once

#include "vbt_format.h"
#include "vbt_internal.h"

#include <hip/hip_runtime.h>

namespace vbtk {

constexpr uint32_t FLAG_SWEEP = 1u << 8;   // render: every control's keyframe frames strictly increase;
                                           // scientific: every event bin's keys strictly increase
constexpr uint32_t FLAG_NO_NAN = 1u << 9;  // no fp16 keyframe value of the leaf is a NaN pattern
constexpr uint32_t STATUS_MASK = 0xFFu;
constexpr uint32_t NO_LEAF = 0xffffffffu;
constexpr uint32_t CACHE_GROUP = 64;      // packed-cache slots per group-table entry

// Tile shapes the autotuner chooses from: (workgroup slots, LDS window bytes).
// A larger window serves more wavefronts in one pass, a smaller one leaves
// room for more resident waves.
struct CacheTile {
    uint32_t block, window;
};
constexpr CacheTile CACHE_TILES[] = {{64, 8192},  {64, 16384}, {128, 16384}, {128, 32768}, {256, 32768},
                                     {256, 49152}, {64, 4096},  {128, 8192},  {256, 16384}, {64, 12288}};
constexpr int CACHE_TILE_COUNT = (int)(sizeof(CACHE_TILES) / sizeof(CACHE_TILES[0]));
constexpr int CACHE_TILE_DEFAULT = 0;
// Up to this many frames per launch the exact per-slot kernel beats staging
// keyframe streams through LDS (a stream is read whole to be swept).
#ifndef VBT_CACHE_EXACT_MAX_FRAMES
#define VBT_CACHE_EXACT_MAX_FRAMES 2
#endif
constexpr int CACHE_EXACT_MAX_FRAMES = VBT_CACHE_EXACT_MAX_FRAMES;

struct DevLeaf {
    uint32_t leaf, word, bytes, status;
    uint32_t ncache, cc;       // cache slots; coarse slots
    uint32_t lw, lh, ld;       // cropped extent
    uint32_t sciHeader;        // scientific header word
    uint32_t sciEnd;           // scientific: bytes from the leaf start its decode may read (capped)
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
    const uint2* slotPlan;      // render: per packed-cache slot, where and how its records are swept
    const uint2* groupRange;    // render: per CACHE_GROUP-slot group, byte range of its swept records
    uint32_t W, H, D, F, L, A, profile, scaleCount;
    uint32_t flags;             // ASSET_* open flags
    uint64_t payloadBytes;
    uint64_t controlsPerFrame;
};

// scientific dense subtypes 2 and 3 decode as the encoder reconstructs them
// (the open flag VBT_OPEN_SCI_DENSE_BASIS)
constexpr uint32_t ASSET_SCI_DENSE_BASIS = 1u;

__host__ __device__ inline bool sci_basis_leaf(const AssetDev& a, const SciHeader& h)
{
    return (a.flags & ASSET_SCI_DENSE_BASIS) != 0u && h.mode >= 2u && (h.subtype == 2u || h.subtype == 3u);
}

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
template <class B>
__device__ inline float sci_control(const AssetDev& a, const B& b, const SciHeader& h, uint32_t ci, int t)
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

template <class B>
__device__ inline uint32_t bfp_code(const B& b, uint32_t blockBase, uint32_t ci, uint32_t bits)
{
    const uint32_t bit = ci * bits;
    const uint32_t byteIndex = blockBase + bit / 8u;
    const uint32_t shift = bit % 8u;
    uint32_t packed = b.u8(byteIndex);
    if (shift + bits > 8u) packed |= b.u8(byteIndex + 1u) << 8;
    return (packed >> shift) & ((1u << bits) - 1u);
}

// basis term k of value ci of a tile / patch block
// (decodeDenseTileTemporalBasisValueAt, src/spatial_first_encoder.cpp:2078-2092,
// and its patch twin :2119-2133): the code keeps at most 8 bits (the mask is
// a uint8_t) of a value read from at most two bytes
template <class B>
__device__ inline float sci_basis_value(const B& b, const SciBasis& s, uint32_t bits, uint64_t ci, uint32_t k)
{
    const uint64_t bit = ci * bits;
    const uint32_t byteIndex = (uint32_t)(s.dataOff + (uint64_t)k * s.basisBytes + bit / 8u);
    const uint32_t shift = (uint32_t)(bit % 8u);
    uint32_t packed = b.u8(byteIndex);
    if (shift + bits > 8u) packed |= b.u8(byteIndex + 1u) << 8;
    const uint32_t q = (packed >> shift) & (((1u << bits) - 1u) & 0xFFu);
    return decode_bfp(q, half_to_float(b.u16(s.scalesOff + 2u * k)), (int)bits);
}

// patch lattice of decodeDensePatchTemporalBasisAt
// (src/spatial_first_encoder.cpp:2147-2161)
__device__ inline void sci_patch_coord(uint32_t local, uint32_t size, uint32_t res, int& i0, float& w)
{
    const float f = ((float)local / (float)(size - 1u)) * (float)(res - 1u);
    int i = (int)floorf(f);
    if (i < 0) i = 0;
    if (i > (int)res - 2) i = (int)res - 2;
    i0 = i;
    w = f - (float)i;
}

// the residual of a tile (decodeDenseTileTemporalBasisAt,
// src/spatial_first_encoder.cpp:2094-2114) or patch block
// (decodeDensePatchTemporalBasisAt, :2135-2176).  Validation at open keeps
// every read inside the block: active count == popcount(mask), tile size <= 8,
// no patch size or resolution of 1 with active patches.
template <class B>
__device__ inline float sci_basis_residual(const AssetDev& a, const B& b, const SciHeader& h, uint32_t x,
                                           uint32_t y, uint32_t z, uint32_t t)
{
    const SciBasis s = sci_basis(b, h);
    if (s.count == 0u || s.size == 0u || s.values == 0u || s.basisBytes == 0u) return 0.0f;
    const float* basis = a.basis + (size_t)t * 64u;
    float residual = 0.0f;
    if (!s.patch) {
        const uint32_t n = s.size, tpa = 8u / n > 1u ? 8u / n : 1u;
        const uint32_t tile = ((z / n) * tpa + y / n) * tpa + x / n;
        if (tile >= 64u || ((s.mask >> tile) & 1u) == 0u) return 0.0f;
        const uint64_t ci = (uint64_t)__popcll(s.mask & ((1ull << tile) - 1u)) * n * n * n +
                            ((z % n) * n + y % n) * n + x % n;
        for (uint32_t k = 0; k < h.tk; ++k) residual += sci_basis_value(b, s, h.bits, ci, k) * basis[k];
        return residual;
    }
    if (s.res == 0u) return 0.0f;
    const uint32_t n = s.size, R = s.res;
    const uint32_t patch = ((z / n) * 2u + y / n) * 2u + x / n;
    if (patch >= 64u || ((s.mask >> patch) & 1u) == 0u) return 0.0f;
    const uint64_t first = (uint64_t)__popcll(s.mask & ((1ull << patch) - 1u)) * R * R * R;
    int x0, y0, z0;
    float tx, ty, tz;
    sci_patch_coord(x % n, n, R, x0, tx);
    sci_patch_coord(y % n, n, R, y0, ty);
    sci_patch_coord(z % n, n, R, z0, tz);
    for (uint32_t k = 0; k < h.tk; ++k) {
        float c[8];
        for (int i = 0; i < 8; ++i) {
            const uint64_t cx = (uint64_t)(x0 + (i & 1)), cy = (uint64_t)(y0 + ((i >> 1) & 1)),
                           cz = (uint64_t)(z0 + ((i >> 2) & 1));
            c[i] = sci_basis_value(b, s, h.bits, first + (cz * R + cy) * R + cx, k);
        }
        residual += trilinear(c, tx, ty, tz) * basis[k];
    }
    return residual;
}

template <class B>
__device__ inline float sci_dense_value(const AssetDev& a, const B& b, const SciHeader& h, uint32_t lx, uint32_t ly,
                                        uint32_t lz, uint32_t t)
{
    if (sci_basis_leaf(a, h)) return sci_basis_residual(a, b, h, lx, ly, lz, t);
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

// sparse-event section of a scientific leaf
// (decodeSparseResidualAt, render/src/scientific_decode.cpp:235-295)
__device__ inline uint32_t sci_event_cap(const SciHeader& h)
{
    return h.tier == 0 ? 64u : h.tier == 1 ? 256u : h.tier == 2 ? 768u : 1024u;
}

__device__ inline uint32_t sci_event_bin(const SciHeader& h, uint32_t t) { return t / h.fpb < 15u ? t / h.fpb : 15u; }

// the signed 4-bit residual code of event i
template <class B>
__device__ inline int sci_event_q(const B& b, uint32_t residualBase, uint32_t i)
{
    const uint32_t byte = b.u8(residualBase + (i >> 1));
    const uint32_t nib = (i & 1u) == 0u ? (byte & 0xFu) : ((byte >> 4) & 0xFu);
    int q = (int)nib;
    if (q & 8) q -= 16;
    return q;
}

template <class B>
__device__ inline float sci_sparse(const B& b, const SciHeader& h, uint32_t idx, uint32_t t)
{
    if (h.count == 0u) return 0.0f;
    const uint32_t cap = sci_event_cap(h);
    const uint32_t bin = sci_event_bin(h, t);
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
        if (key == target) return (float)sci_event_q(b, residualBase, (uint32_t)mid) * eventScale;
        if (key < target) left = mid + 1;
        else right = mid - 1;
    }
    return 0.0f;
}

template <class B>
__device__ inline float sci_residual(const AssetDev& a, const B& b, const SciHeader& h, uint32_t lx, uint32_t ly,
                                     uint32_t lz, uint32_t t)
{
    if (h.mode == 0u) return 0.0f;
    if (h.mode == 1u) return sci_sparse(b, h, (lz * 8u + ly) * 8u + lx, t);
    return sci_dense_value(a, b, h, lx, ly, lz, t);
}

// the coarse lattice cell and weights of one voxel (sampleCoarseAt,
// render/src/scientific_decode.cpp:99-123)
struct SciCoarseAt {
    int x0, y0, z0;
    float tx, ty, tz;
};

__device__ inline SciCoarseAt sci_coarse_at(uint32_t lx, uint32_t ly, uint32_t lz)
{
    SciCoarseAt p;
    sci_coarse_coord(lx, p.x0, p.tx);
    sci_coarse_coord(ly, p.y0, p.ty);
    sci_coarse_coord(lz, p.z0, p.tz);
    return p;
}

__device__ inline float sci_coarse_lerp(const float* ctrl, const SciCoarseAt& p)
{
    float s[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int cx = p.x0 + (i & 1), cy = p.y0 + ((i >> 1) & 1), cz = p.z0 + ((i >> 2) & 1);
        s[i] = ctrl[(cz * 4 + cy) * 4 + cx];
    }
    return trilinear(s, p.tx, p.ty, p.tz);
}

__device__ inline float sci_coarse_sample(const float* ctrl, uint32_t lx, uint32_t ly, uint32_t lz)
{
    return sci_coarse_lerp(ctrl, sci_coarse_at(lx, ly, lz));
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

// decode_control (decodeFromSequence, src/render_temporal_decode.cpp:79-131:
// bin hint, fallbacks, early outs, binary search, v0 + alpha * (v1 - v0))
// with aligned dword loads from the leaf base W: the descriptor and the two
// bin-hint bytes are read together, each keyframe record (frame and value)
// with one pair of dwords.  Same result as decode_control for every stream;
// reads stay inside the leaf's validated stream plus 4 bytes (the payload
// allocation is padded).
__device__ __forceinline__ uint32_t w_u32(const uint32_t* W, uint32_t byteOff)
{
    const uint32_t i = byteOff >> 2;
    return __builtin_amdgcn_alignbyte(W[i + 1u], W[i], byteOff & 3u);
}

__device__ __forceinline__ void w_rec(const uint32_t* W, uint32_t off, uint32_t codec, int& f, float& v)
{
    const uint32_t i = off >> 2, lo = W[i], hi = W[i + 1u];
    if (codec == CODEC_FP32) {
        f = (int)(lo & 0xFFFFu); // fp32 records are 4-byte aligned
        v = __uint_as_float(hi);
    } else {
        const uint32_t x = __builtin_amdgcn_alignbyte(hi, lo, off & 3u);
        f = (int)(x & 0xFFu);
        v = half_to_float((x >> 8) & 0xFFFFu);
    }
}

__device__ inline float decode_control_w(const uint32_t* W, const StreamRef& s, uint32_t ci, int F, int t)
{
    const int frame = clamp_frame(t, F);
    const uint32_t B = s.bins;
    const uint32_t q = ((uint32_t)frame * B) / (uint32_t)(F > 1 ? F : 1);
    const uint32_t bin = B != 0u ? (q < B - 1u ? q : B - 1u) : 0u;
    const uint32_t desc = W[(s.base + ci * 4u) >> 2]; // descriptors are 4-byte aligned
    const uint32_t hw = w_u32(W, s.base + s.count * 4u + ci * B + bin);
    const uint32_t keyStart = desc & 0xFFFFu, keyCount = (desc >> 16) & 0xFFu;
    if (keyCount == 0u) return 0.0f;
    const uint32_t last = keyCount - 1u, kb = keyframe_bytes(s.codec);
    const uint32_t k0 = s.keyBase + keyStart * kb;
    uint32_t lo = 0u;
    if (B != 0u) lo = (hw & 0xFFu) < last ? (hw & 0xFFu) : last;
    int f0, f1;
    float v0, v1;
    w_rec(W, k0 + lo * kb, s.codec, f0, v0);
    if (f0 > frame) {
        lo = 0u;
        w_rec(W, k0, s.codec, f0, v0);
    }
    if (f0 >= frame || lo == last) return v0;
    uint32_t hi = last;
    if (bin + 1u < B) {
        const uint32_t h = ((hw >> 8) & 0xFFu) + 1u;
        hi = h < last ? h : last;
    }
    if (hi < lo + 1u) hi = lo + 1u;
    w_rec(W, k0 + hi * kb, s.codec, f1, v1);
    if (f1 < frame) {
        hi = last;
        w_rec(W, k0 + hi * kb, s.codec, f1, v1);
    }
    if (f1 <= frame) return v1;
    while (hi > lo + 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        int fm;
        float vm;
        w_rec(W, k0 + mid * kb, s.codec, fm, vm);
        if (fm < frame) {
            lo = mid;
            f0 = fm;
            v0 = vm;
        } else {
            hi = mid;
            f1 = fm;
            v1 = vm;
        }
    }
    if (f1 == frame) return v1;
    const float alpha = (float)(frame - f0) / (float)(f1 - f0);
    return v0 + alpha * (v1 - v0);
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
    return decode_control_w((const uint32_t*)(a.payload + (size_t)d.word * 4u), s, isCoarse ? slot : slot - d.cc,
                            (int)a.F, t);
}

// ------------------------------------------------- packed cache: render tile --
// Lerp weight of the sweep: alpha = (frame - f0) / (f1 - f0) for integer
// frames 0 <= a = frame - f0 < b = f1 - f0 < 2^16, equal to the correctly
// rounded float quotient RN32(a / b) the reference computes.  Given y =
// RN32(1 / b), Markstein's correction (P. Markstein, IBM J. Res. Dev. 34(1),
// 1990) yields it in float arithmetic: q = RN32(a y) is within one ulp of
// a / b, the residual r = a - b q is exact in one FMA, and RN32(q + r y) =
// RN32(a / b).  b and a are exact floats (< 2^24).  Checked for every pair
// with y = RN32(1 / b) on the host (tests/native/alpha_identity.c), and on
// the device with y from sweep_recip, the bare hardware reciprocal, for every
// pair 0 <= a < b < 2^16 (vbt_selftest_weights).
__device__ __forceinline__ float sweep_alpha(float a, float b, float y)
{
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}

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

// The forward sweep of one control (strictly increasing key frames).  State:
// keys k and k+1 and 1 / (f[k+1] - f[k]); the loop also carries key k+2
// (clamped to the last key).
// Invariant after each frame: the next key's frame fn exceeds the frame just
// produced.  Frames advance by at most one per step, so a single predicated
// advance per step keeps the invariant (the key after the next is beyond fn
// because frames strictly increase); a catch-up establishes it for the first
// frame.  Each frame's bracket is then the adjacent key pair around it, the
// bracket the reference's hinted binary search returns for increasing keys.
struct SweepState {
    uint32_t k;
    int fk, fn;
    float vk, vn;
    float gap, y; // fn - fk and sweep_recip(fn - fk)
};

// Reciprocal of an integer key gap 1 <= b < 2^16 for sweep_alpha: the bare
// hardware float reciprocal (v_rcp_f32, within one ulp).  It differs from
// RN32(1 / b) for 6,450 of these gaps, yet with it Markstein's correction
// still returns RN32(a / b) for every numerator 0 <= a < b of every gap: the
// device self-test (vbt_selftest_weights) checks all 2^31 pairs, and the
// GPU test suite runs it.  (One Newton step, which makes y = RN32(1 / b)
// for every gap, cost two FMAs per frame and 3 % of the frame-cache launch.)
// A gap of 0 (never used for a weight) gives a non-finite value.
__device__ __forceinline__ float sweep_recip(float b) { return __builtin_amdgcn_rcpf(b); }

typedef __attribute__((address_space(3))) const uint8_t lds_byte;

// Record of a control staged in LDS at byte offset off (32-bit LDS addressing),
// read with aligned dword loads: an unaligned 16-bit LDS read stalls the LDS
// pipeline (the staged pool's 3-byte fp16 records sit at every byte offset),
// so the record is cut out of the two dwords covering it.  The second dword
// is only needed when the record crosses into it; the LDS arrays have
// padding dwords after the window (sweep_frames reads past a control's last
// record) so every read stays inside them.  fp32 records
// (8 bytes) are 4-byte aligned in the pool: frame in the low half of the
// first dword, value in the second.
// The two dwords holding the record at byte offset off, and the record
// decoded from them (split so a loop can issue the read well before it needs
// the record).
struct RawKey {
    uint32_t lo, hi, shift;
};

__device__ __forceinline__ RawKey lds_fetch(lds_byte* L, uint32_t off)
{
    typedef __attribute__((address_space(3))) const uint32_t lds_word;
    lds_word* W = (lds_word*)L;
    const uint32_t wi = off >> 2;
    return RawKey{W[wi], W[wi + 1u], off & 3u};
}

template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void key_decode(const RawKey& r, int& f, float& v)
{
    if (KB == 8u) {
        f = (int)(r.lo & 0xFFFFu);
        v = __uint_as_float(r.hi);
    } else {
        const uint32_t x = __builtin_amdgcn_alignbyte(r.hi, r.lo, r.shift); // bytes off .. off+3
        f = (int)(x & 0xFFu);
        v = sweep_half<NANFIX>((x >> 8) & 0xFFFFu);
    }
}

template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void lds_key(lds_byte* L, uint32_t off, int& f, float& v)
{
    key_decode<KB, NANFIX>(lds_fetch(L, off), f, v);
}

// State at the first requested frame from the control's records staged in
// LDS at byte offset off: the catch-up walk that establishes the invariant.
template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void sweep_start(lds_byte* L, uint32_t off, uint32_t last, int first, SweepState& s)
{
    s.k = 0u;
    s.fn = 0;
    s.vn = 0.0f;
    lds_key<KB, NANFIX>(L, off, s.fk, s.vk);
    if (last >= 1u) lds_key<KB, NANFIX>(L, off + KB, s.fn, s.vn);
    while (s.k < last && s.fn <= first) {
        ++s.k;
        s.fk = s.fn;
        s.vk = s.vn;
        if (s.k < last) lds_key<KB, NANFIX>(L, off + (s.k + 1u) * KB, s.fn, s.vn);
    }
    s.gap = (float)(s.fn - s.fk);
    s.y = sweep_recip(s.gap);
}

// Cache-policy bits of a buffer store on gfx950: 0 (write-back through L2)
// or NT (non-temporal: streamed, no cache allocation).
constexpr int STORE_WB = 0;
constexpr int STORE_NT = 2;

// Raw buffer descriptor of one frame of the output (base, size in bytes):
// stores address it with a scalar base and a 32-bit lane offset, and stores
// at or beyond `bytes` are dropped by the hardware range check.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_rsrc(char* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)bytes, 0x00020000);
}

// Frames tBegin .. tBegin+cn-1 (all inside [0, F)) of one control from state
// s at frame tBegin; its records are in LDS at byte offset off (record i at
// off + i * KB).  Frame i goes to out + i * strideBytes + dstBytes, with out
// the same for every lane: a buffer store with the frame's base in scalar
// registers and the lane's 32-bit offset (strideBytes < 2^32, dstBytes <
// strideBytes; the host checks the frame size).  The state update is
// branch-free: the key after next (record k + 2, tracked as a byte offset) is
// read every step and selected on advance.  Step i issues that LDS read first
// and produces frame i - 1's output (saved by step i - 1) while the read is
// in flight, then updates the state for frame i.  Once k + 2 passes the last
// record the read returns whatever follows it in LDS (the next control's
// records or padding); such a record only becomes key k + 1 when k reaches
// the last key, where the output is vk and never uses it.  The read reaches
// at most KB + 7 bytes past the window's staged records: record k + 2 <=
// last + 2 starts at most KB bytes past the control's end, which is inside
// the window, and the read covers two dwords from there.
// Record k + 2 of the sweep decoded with its frame as a float (frames are
// integers below 2^16, exact in float).  The alignbyte selector uses the
// offset's low two bits only, so the offset is passed as it is.
template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void key_decode_f(const RawKey& r, uint32_t off, float& f, float& v)
{
    if (KB == 8u) {
        f = (float)(r.lo & 0xFFFFu);
        v = __uint_as_float(r.hi);
    } else {
        const uint32_t x = __builtin_amdgcn_alignbyte(r.hi, r.lo, off); // bytes off .. off+3
        f = (float)(x & 0xFFu);
        v = sweep_half<NANFIX>((x >> 8) & 0xFFFFu);
    }
}

// Frames tBegin .. tBegin+cn-1 (all inside [0, F)) of one control from state
// s at frame tBegin; its records are in LDS at byte offset off (record i at
// off + i * KB).  Frame i goes to out + i * strideBytes + dstBytes, with out
// the same for every lane: a buffer store with the frame's base in scalar
// registers and the lane's 32-bit offset (strideBytes < 2^32, dstBytes <
// strideBytes; the host checks the frame size).
// The state update is branch-free.  Key frames are carried as floats (exact
// integers; the frame counter is a wave-uniform float), so the weight's
// numerator and the key gap are single float subtractions.  The key counter
// is replaced by the byte offset pP of record k + 2 (the record read ahead):
// k < last exactly when pP < e2 = off + (last + 2) * KB.  Record k + 2 is
// read every step and selected on advance.  Step i issues that LDS read
// first and produces frame i - 1's output while the read is in flight, then
// updates the state for frame i.  Once k + 2 passes the last record the read
// returns whatever follows it in LDS (the next control's records or
// padding); such a record only becomes key k + 1 when k reaches the last
// key, where the output is vk and never uses it.  The read reaches at most
// KB + 7 bytes past the window's staged records: record k + 2 <= last + 2
// starts at most KB bytes past the control's end, which is inside the
// window, and the read covers two dwords from there.
// The loop state of one control's sweep (the same fields for every record
// format, so lanes of different formats can share one loop around it) and its
// steps: sweep_begin sets it up at the first frame, sweep_update moves it to
// the current frame, sweep_value produces that frame's control.
struct SweepRegs {
    float fk, fn, vk, vn, gap, y;
    float fkS;   // the output's key frame: past every frame at the last key
    float frame; // the current frame
    uint32_t pP, e2;
    bool more;   // key k + 1 exists
    RawKey raw;  // record k + 2, read a step ahead
};

template <uint32_t KB>
__device__ __forceinline__ void sweep_begin(lds_byte* L, uint32_t off, uint32_t last, const SweepState& s, int tBegin,
                                            SweepRegs& r)
{
    r.fk = (float)s.fk;
    r.fn = (float)s.fn;
    r.vk = s.vk;
    r.vn = s.vn;
    r.gap = s.gap;
    r.y = s.y;
    r.pP = off + (s.k + 2u) * KB;
    r.e2 = off + (last + 2u) * KB;
    r.more = s.k < last;
    r.raw = lds_fetch(L, r.pP);
    // one comparison fkS >= frame selects vk both before the first key and
    // after the last
    r.fkS = r.more ? r.fk : __builtin_inff();
    r.frame = (float)tBegin;
}

template <uint32_t KB, bool NANFIX>
__device__ __forceinline__ void sweep_update(lds_byte* L, SweepRegs& r)
{
    float fp, vp;
    // the record read a step ago is consumed here on every lane: left to
    // itself the compiler re-reads it on advance only, behind a branch that
    // waits for the read
    asm volatile("" ::"v"(r.raw.lo), "v"(r.raw.hi));
    key_decode_f<KB, NANFIX>(r.raw, r.pP, fp, vp);
    asm volatile("" : "+v"(fp), "+v"(vp));
    const bool adv = r.more & (r.fn <= r.frame);
    const float gNew = fp - r.fn;
    const float yNew = sweep_recip(gNew);
    r.fk = adv ? r.fn : r.fk;
    r.vk = adv ? r.vn : r.vk;
    r.fn = adv ? fp : r.fn;
    r.vn = adv ? vp : r.vn;
    r.gap = adv ? gNew : r.gap;
    r.y = adv ? yNew : r.y;
    r.pP += adv ? KB : 0u;
    r.raw = lds_fetch(L, r.pP);
    r.more = r.pP < r.e2;
    r.fkS = r.more ? r.fk : __builtin_inff();
}

__device__ __forceinline__ float sweep_value(const SweepRegs& r)
{
    const float alpha = sweep_alpha(r.frame - r.fkS, r.gap, r.y); // frame - fkS: exact (integers < 2^16)
    const float lerp = r.vk + alpha * (r.vn - r.vk);
    return r.fkS >= r.frame ? r.vk : lerp;
}

// Frames tBegin .. tBegin+cn-1 (all inside [0, F)) of one control from state
// s at frame tBegin; its records are in LDS at byte offset off (record i at
// off + i * KB).  Frame i goes to out + i * strideBytes + dstBytes, with out
// the same for every lane: a buffer store (cache policy POLICY) with the
// frame's base in scalar registers and the lane's 32-bit offset (strideBytes
// < 2^32, dstBytes < strideBytes; the host checks the frame size).
// The state update is branch-free.  Key frames are carried as floats (exact
// integers; the frame counter is a wave-uniform float), so the weight's
// numerator and the key gap are single float subtractions.  The key counter
// is replaced by the byte offset pP of record k + 2 (the record read ahead):
// k < last exactly when pP < e2 = off + (last + 2) * KB.  Record k + 2 is
// read every step and selected on advance.  Step i issues that LDS read
// first and produces frame i - 1's output while the read is in flight, then
// updates the state for frame i.  Once k + 2 passes the last record the read
// returns whatever follows it in LDS (the next control's records or
// padding); such a record only becomes key k + 1 when k reaches the last
// key, where the output is vk and never uses it.  The read reaches at most
// KB + 7 bytes past the window's staged records: record k + 2 <= last + 2
// starts at most KB bytes past the control's end, which is inside the
// window, and the read covers two dwords from there.
template <uint32_t KB, bool NANFIX, int POLICY = STORE_NT>
__device__ __forceinline__ void sweep_frames(lds_byte* L, uint32_t off, uint32_t last, const SweepState& s,
                                             int tBegin, int cn, char* out, uint32_t dstBytes, uint64_t strideBytes)
{
    SweepRegs r;
    sweep_begin<KB>(L, off, last, s, tBegin, r);
    auto output = [&]() {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(sweep_value(r)), frame_rsrc(out, (uint32_t)strideBytes),
                                              (int)dstBytes, 0, POLICY);
        out += strideBytes;
    };
    // one basic block per step: the record read at the end of a step stays
    // there, and the next step's output stage runs while it is in flight
    sweep_update<KB, NANFIX>(L, r);
    for (int i = 1; i < cn; ++i) {
        output();
        r.frame += 1.0f;
        sweep_update<KB, NANFIX>(L, r);
    }
    output();
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
// WINDOW bytes, sRed BLOCK/64 words.  BLOCK and WINDOW are the tile shape (the
// autotuner picks it for the one-shot kernel).  STAGE is the number of dword
// loads each lane keeps in flight while staging a window: 32 brings a 32 KiB
// window of a 256-slot tile in one round trip (the one-shot kernel stays at
// 59 VGPRs); kernels with more live state use fewer to keep their occupancy.
template <int LAYOUT, uint32_t BLOCK, uint32_t WINDOW, uint32_t STAGE = 32u>
__device__ __forceinline__ void cache_render_tile(const AssetDev& a, const uint32_t* groupLeaf, uint32_t tile, int t0,
                                                  int nt, float* out, uint64_t total, uint32_t* errFlag,
                                                  uint32_t* lds, uint32_t* sRed)
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
    const int first = t0;
    // frames t0 .. t0+nt-1 lie inside [0, F) (checked by the host); slot byte
    // offsets within a frame fit 32 bits (checked by the host)
    char* const ob = (char*)out;
    const uint32_t dstBytes = (uint32_t)dst * 4u;
    const uint64_t strideBytes = stride * 4u;
    auto run = [&](uint32_t off) {
        lds_byte* L = (lds_byte*)lds;
        SweepState st;
        if (mode == 0u) {
            sweep_start<8u, false>(L, off, last, first, st);
            sweep_frames<8u, false>(L, off, last, st, t0, nt, ob, dstBytes, strideBytes);
        } else if (mode == 1u) {
            sweep_start<3u, false>(L, off, last, first, st);
            sweep_frames<3u, false>(L, off, last, st, t0, nt, ob, dstBytes, strideBytes);
        } else {
            sweep_start<3u, true>(L, off, last, first, st);
            sweep_frames<3u, true>(L, off, last, st, t0, nt, ob, dstBytes, strideBytes);
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
        // staging: STAGE dword loads in flight per lane before the LDS writes
        // (for the one-shot kernel 8 in flight measured 11 % slower per launch
        // than 32; 48 spends registers)
        constexpr uint32_t SU = STAGE;
        const uint32_t* pw = (const uint32_t*)a.payload + w0;
        for (uint32_t i0 = tid; i0 < nw; i0 += SU * BLOCK) {
            uint32_t v[SU];
#pragma unroll
            for (uint32_t j = 0; j < SU; ++j) v[j] = i0 + j * BLOCK < nw ? pw[i0 + j * BLOCK] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < SU; ++j)
                if (i0 + j * BLOCK < nw) lds[i0 + j * BLOCK] = v[j];
        }
        __syncthreads();
        if (mine) {
            run(kLo - w0 * 4u);
            pending = false;
        }
        __syncthreads();
    }
}

// ------------------------------------------------- slot plan of the packed cache --
// Per packed-cache slot, computed once per render asset (k_slot_plan): where
// the slot's keyframe records start and how the sweep reads them, so a tile's
// lanes start staging after one load instead of a walk over the control
// offsets, the leaf's plan and its descriptor.
//   x: byte offset in the payload of the slot's first keyframe record (swept
//      slots), or the slot's active leaf index (other slots)
//   y: keyCount | kind << 8
// PLAN_ZERO is a stream without keys (the control is 0 at every frame);
// PLAN_SLOW covers poisoned leaves and keys that do not strictly increase:
// those slots take the per-frame path of cache_slot_frames_slow.
// The group range of every CACHE_GROUP consecutive slots is the byte range
// [x, x + keyCount * record bytes) spanned by its swept slots (x = ~0 when it
// has none); a group holding a PLAN_SLOW slot gets the range [0, ~0), which
// no window holds, so its tiles take the general path.  A tile reads its
// groups' ranges to size its LDS window.
constexpr uint32_t PLAN_SLOW = 0u, PLAN_FP16 = 1u, PLAN_FP16_NAN = 2u, PLAN_FP32 = 3u, PLAN_ZERO = 4u;
__host__ __device__ inline bool plan_swept(uint32_t kind) { return kind >= PLAN_FP16 && kind <= PLAN_FP32; }
__host__ __device__ inline uint32_t plan_kind(uint2 p) { return (p.y >> 8) & 7u; }

// XCD-aware tile order of the one-shot packed build: workgroup b runs on XCD
// b % 8 (the dispatcher's round robin) and takes the tile
// ((j >> r) * 8 + b % 8) << r | (j & (2^r - 1)), j = b / 8, so each XCD works
// through runs of 2^r consecutive tiles.  Neighbouring tiles share the
// 128-byte lines at the ends of their row segments (a frame's row starts at
// any 4-byte offset); written from one XCD the two halves meet in its L2
// before the line goes to memory, from two XCDs each writes a partial line.
// The grid is rounded up to a multiple of 8 << r; tiles past the list are
// skipped.
__device__ __forceinline__ uint32_t xcd_run_tile(uint32_t b, uint32_t runLog2)
{
    const uint32_t j = b >> 3;
    return ((((j >> runLog2) << 3) + (b & 7u)) << runLog2) | (j & ((1u << runLog2) - 1u));
}

// r of the runs over n tiles (or work items): as long as every XCD keeps at
// least 64 runs, at most 16 tiles a run
__host__ __device__ inline uint32_t tile_run_log2(uint32_t n)
{
    uint32_t r = 0;
    while (r < 4u && (n >> (3u + r + 1u)) >= 64u) ++r;
    return r;
}

// the grid of n tiles in runs of 2^r: a multiple of 8 << r
__host__ __device__ inline uint32_t run_grid(uint32_t n, uint32_t runLog2)
{
    const uint32_t q = 8u << runLog2;
    return (uint32_t)(((uint64_t)n + q - 1u) / q * q);
}

// Wave-uniform loads from tables that no kernel writes while it runs (tile
// lists, group ranges, batch work items: written at asset open or before the
// launch).  Read through the constant address space they stay scalar loads
// even where the kernel has stored before them (a work-queue kernel's next
// item), which the compiler otherwise turns into vector loads.
__device__ __forceinline__ uint32_t invariant_u32(const uint32_t* p, uint64_t i)
{
    return ((const __attribute__((address_space(4))) uint32_t*)p)[i];
}
__device__ __forceinline__ uint2 invariant_u2(const uint2* p, uint64_t i)
{
    const uint64_t v = ((const __attribute__((address_space(4))) uint64_t*)p)[i];
    return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
}

__device__ inline uint2 slot_plan(const AssetDev& a, uint64_t g)
{
    uint32_t c = a.groupLeaf[g / CACHE_GROUP];
    while (c + 1u < a.A && a.ctrl[c + 1u] <= g) ++c;
    const DevLeaf& d = a.leaves[c];
    const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
    const uint32_t status = d.status;
    const bool isCoarse = slot < d.cc;
    const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
    const uint32_t ci = isCoarse ? slot : slot - d.cc;
    if ((status & STATUS_MASK) != LEAF_OK) return make_uint2(c, PLAN_SLOW << 8);
    const uint32_t desc = *(const uint32_t*)(a.payload + d.word * 4u + s.base + ci * 4u); // 4-aligned
    const uint32_t keyStart = desc & 0xFFFFu, keyCount = (desc >> 16) & 0xFFu;
    const uint32_t kb = keyframe_bytes(s.codec);
    if (keyCount == 0u) return make_uint2(c, PLAN_ZERO << 8);
    if (!(status & FLAG_SWEEP)) return make_uint2(c, keyCount | PLAN_SLOW << 8);
    const uint32_t kind = kb == 8u ? PLAN_FP32 : ((status & FLAG_NO_NAN) ? PLAN_FP16 : PLAN_FP16_NAN);
    return make_uint2(d.word * 4u + s.keyBase + keyStart * kb, keyCount | kind << 8);
}

// Per-frame path of a slot the sweep does not take (poisoned leaf: NaN and
// the error flag; no keys: 0; keys not strictly increasing: the reference's
// bracket search per frame from global memory).
__device__ __forceinline__ void cache_slot_frames_slow(const AssetDev& a, uint32_t c, uint64_t g, int t0, int nt,
                                                       float* o, uint64_t stride, uint32_t* errFlag)
{
    const DevLeaf& d = a.leaves[c];
    const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
    if ((d.status & STATUS_MASK) != LEAF_OK) {
        for (int i = 0; i < nt; ++i) o[(uint64_t)i * stride] = __builtin_nanf("");
        atomicOr(errFlag, 1u);
        return;
    }
    const bool isCoarse = slot < d.cc;
    const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
    const uint32_t ci = isCoarse ? slot : slot - d.cc;
    const Bytes b{a.payload + d.word * 4u};
    const uint32_t desc = b.u32(s.base + ci * 4u);
    const uint32_t keyStart = desc & 0xFFFFu, keyCount = (desc >> 16) & 0xFFu;
    if (keyCount == 0u) {
        for (int i = 0; i < nt; ++i) o[(uint64_t)i * stride] = 0.0f;
        return;
    }
    const uint32_t binBase = s.base + s.count * 4u + ci * s.bins;
    for (int i = 0; i < nt; ++i)
        o[(uint64_t)i * stride] = decode_sequence(b, s.keyBase, s.codec, keyStart, keyCount, binBase, s.bins,
                                                  (int)a.F, t0 + i);
}

// Copy payload dwords [w0, w0 + nw) into lds[0 .. nw) with STAGE loads in
// flight per lane.
template <uint32_t BLOCK, uint32_t STAGE>
__device__ __forceinline__ void stage_window(const uint8_t* payload, uint32_t w0, uint32_t nw, uint32_t* lds,
                                             uint32_t tid)
{
    const uint32_t* pw = (const uint32_t*)payload + w0;
    for (uint32_t i0 = tid; i0 < nw; i0 += STAGE * BLOCK) {
        uint32_t v[STAGE];
#pragma unroll
        for (uint32_t j = 0; j < STAGE; ++j) v[j] = i0 + j * BLOCK < nw ? pw[i0 + j * BLOCK] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < STAGE; ++j)
            if (i0 + j * BLOCK < nw) lds[i0 + j * BLOCK] = v[j];
    }
}

// One tile of the packed cache from the slot plan: the same results as
// cache_render_tile (VBT_CACHE_PACKED).  The tile is the CACHE_GROUP-slot
// groups [g0, g0 + ng) (ng <= BLOCK / CACHE_GROUP; the host cuts the packed
// cache into tiles whose records fit the window where a group allows it, see
// plan_tiles).  Its record range comes from its groups' ranges: a tile that
// fits stages its records in one pass with a single barrier, others take the
// multi-pass staging of cache_render_tile.  Each lane stores its control's
// frames (write-back stores: the tile's row segments share lines with their
// neighbours' and merge in L2).
template <uint32_t BLOCK, uint32_t WINDOW, uint32_t STAGE = 32u>
__device__ __forceinline__ void cache_plan_tile(const AssetDev& a, uint32_t g0, uint32_t ng, int t0, int nt,
                                                float* out, uint64_t total, uint32_t* errFlag, uint32_t* lds,
                                                uint32_t* sRed)
{
    constexpr uint32_t NW = BLOCK / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint64_t g = (uint64_t)g0 * CACHE_GROUP + tid;
    const bool live = tid < ng * CACHE_GROUP && g < total;
    uint32_t tlo = 0xFFFFFFFFu, thi = 0u;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
        if (w < ng) {
            const uint2 r = invariant_u2(a.groupRange, g0 + w);
            tlo = min(tlo, r.x);
            thi = max(thi, r.y);
        }
    }
    const uint2 p = live ? a.slotPlan[g] : make_uint2(0u, PLAN_ZERO << 8);
    const uint32_t kind = plan_kind(p), keyCount = p.y & 0xFFu;
    const bool sw = live && plan_swept(kind);
    const uint32_t kb = kind == PLAN_FP32 ? 8u : 3u;
    const uint32_t kLo = p.x, kHi = p.x + keyCount * kb;
    const uint32_t last = keyCount - 1u;
    char* const ob = (char*)out;
    const uint64_t strideBytes = total * 4u;
    const uint32_t dstBytes = (uint32_t)g * 4u;
    lds_byte* L = (lds_byte*)lds;
    if (live && !sw) cache_slot_frames_slow(a, p.x, g, t0, nt, out + g, total, errFlag);
    auto run = [&](uint32_t off) {
        SweepState st;
        if (kind == PLAN_FP32) {
            sweep_start<8u, false>(L, off, last, t0, st);
            sweep_frames<8u, false, STORE_WB>(L, off, last, st, t0, nt, ob, dstBytes, strideBytes);
        } else if (kind == PLAN_FP16) {
            sweep_start<3u, false>(L, off, last, t0, st);
            sweep_frames<3u, false, STORE_WB>(L, off, last, st, t0, nt, ob, dstBytes, strideBytes);
        } else {
            sweep_start<3u, true>(L, off, last, t0, st);
            sweep_frames<3u, true, STORE_WB>(L, off, last, st, t0, nt, ob, dstBytes, strideBytes);
        }
    };
    if (tlo == 0xFFFFFFFFu) return; // no swept slot in the tile
    if (thi - (tlo & ~3u) <= WINDOW) {
        const uint32_t w0 = tlo / 4u;
        stage_window<BLOCK, STAGE>(a.payload, w0, (thi + 3u) / 4u - w0, lds, tid);
        __syncthreads();
        if (sw) run(kLo - w0 * 4u);
        __syncthreads(); // the window is free again (work-queue kernels stage their next tile into it)
        return;
    }
    // multi-pass: as cache_render_tile
    const uint32_t waveLo = wave_min(sw ? kLo : 0xFFFFFFFFu);
    const uint32_t waveHi = wave_max(sw ? kHi : 0u);
    const bool waveFits = waveLo != 0xFFFFFFFFu && waveHi - (waveLo & ~3u) <= WINDOW;
    bool pending = sw;
    for (;;) {
        const uint32_t lo = block_min<NW>(pending ? kLo : 0xFFFFFFFFu, sRed, lane, wave);
        if (lo == 0xFFFFFFFFu) break;
        const uint32_t w0 = lo / 4u;
        const uint32_t limit = w0 * 4u + WINDOW;
        const bool mine = pending && ((waveFits ? waveHi <= limit : kHi <= limit) || kLo == lo);
        const uint32_t hi = block_max<NW>(mine ? kHi : 0u, sRed, lane, wave);
        const uint32_t nw = hi > w0 * 4u ? (hi + 3u) / 4u - w0 : 0u;
        stage_window<BLOCK, STAGE>(a.payload, w0, nw, lds, tid);
        __syncthreads();
        if (mine) {
            run(kLo - w0 * 4u);
            pending = false;
        }
        __syncthreads();
    }
}

} // namespace vbtk
