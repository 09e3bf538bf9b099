// decode.hip -- CDNA4 (gfx950) kernels of the VBTPACK4 random-access decode
// path and the asset / C-ABI layer around them.
//
// Data resident in HBM per asset (all uploaded once at open):
//   offsets  u32[L+1]   word offsets of the leaves (render/src/vbt_file.cpp layout)
//   payload  u32[P]     the leaf payload pool, byte-addressed by the kernels
//   scales   f32[S]     scientific AC scales
//   compact  u32[L]     page table: rank among active leaves (0xffffffff = EMPTY)
//   status   u32[L]     per-leaf status (LEAF_* | FLAG_SWEEP | FLAG_NO_NAN)
//   leaves   DevLeaf[A] per active leaf decode plan (compact order)
//   ctrl     u64[A+1]   packed-cache control offsets (prefix sum of C(l))
//   basis    f32[F*64]  scientific DCT basis table (dctBasis values, host-computed)
//
// Kernels (one launch each, stream-ordered):
//   k_scan_a / k_scan_b / k_scan_c   deterministic page table: wave64 ballot +
//                                    popcount ranks, LDS block scan, block-base scan
//   k_validate                       descriptor ranges + keyframe monotonicity per leaf
//   k_cache_render                   multi-frame control cache (sweep or exact bracket)
//   k_cache_sci                      scientific leaf voxels per frame (DCT in fp64)
//   k_dense_render / k_dense_sci     whole-frame voxel composition
//   k_query / k_query_plain          point queries (8 lanes per query / 1 thread per query)
//   k_control_values / ...           batched forms of the reference per-value decoders
#include "vbt_format.h"
#include "vbt_internal.h"
#include "vbt_kernels.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using namespace vbtk;

// ------------------------------------------------------------ error state --
static thread_local std::string g_err;

int vbt_fail(int code, const char* msg)
{
    g_err = msg;
    return code;
}
int vbt_fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return vbt_fail(e_ == hipErrorOutOfMemory ? VBT_ERR_NOMEM : VBT_ERR_DEVICE,                 \
                            std::string(#expr " failed: ") + hipGetErrorString(e_));                    \
    } while (0)

namespace {


inline hipStream_t S(void* s) { return (hipStream_t)s; }
inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

// ------------------------------------------------------------ wave helpers --
__device__ inline uint32_t wave_prefix_u32(uint32_t v, uint32_t lane, uint32_t& total)
{
    // inclusive scan across the 64-lane wavefront
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v; // exclusive
}

// per-leaf active flag and cache-slot count (0 for EMPTY and poisoned leaves)
__device__ inline void leaf_count(const AssetDev& a, uint32_t l, uint32_t& active, uint32_t& n, uint32_t& status,
                                  LeafPlan& pl)
{
    active = 0;
    n = 0;
    status = LEAF_OK;
    const uint32_t word = a.offsets[l];
    const uint32_t bytes = (a.offsets[l + 1] - word) * 4u;
    const Bytes b{a.payload + (size_t)word * 4u};
    if (a.profile == 2u) {
        uint32_t bx, by, bz, lw, lh, ld;
        leaf_coords(l, a.W, a.H, a.D, bx, by, bz, lw, lh, ld);
        active = 1;
        n = lw * lh * ld;
        pl.mode = b.u32(0) & 3u;
        return;
    }
    status = plan_leaf(b, bytes, pl);
    active = pl.mode != MODE_EMPTY ? 1u : 0u;
    if (active && status == LEAF_OK) {
        n = control_count(pl.coarseRes);
        if (pl.mode == MODE_FINE6) n += control_count(pl.fineRes);
        if (pl.mode == MODE_SHELL) n += pl.shell.count;
    }
}

struct Pair {
    uint32_t active;
    uint64_t n;
};

// pass A: per-block totals of (active leaves, cache slots)
__global__ void __launch_bounds__(256) k_scan_a(AssetDev a, Pair* blockSums)
{
    __shared__ uint32_t sA[4];
    __shared__ uint32_t sN[4];
    const uint32_t l = blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t act = 0, n = 0, st;
    LeafPlan pl;
    if (l < a.L) leaf_count(a, l, act, n, st, pl);
    const uint64_t m = __ballot(act);
    uint32_t tot;
    wave_prefix_u32(n, lane, tot);
    if (lane == 0) {
        sA[wave] = (uint32_t)__popcll(m);
        sN[wave] = tot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Pair p{0, 0};
        for (int w = 0; w < 4; ++w) {
            p.active += sA[w];
            p.n += sN[w];
        }
        blockSums[blockIdx.x] = p;
    }
}

// pass B: exclusive scan of block totals by one workgroup (deterministic)
__global__ void __launch_bounds__(1024) k_scan_b(Pair* blockSums, uint32_t nb, Pair* total)
{
    __shared__ uint32_t sA[16];
    __shared__ uint64_t sN[16];
    __shared__ Pair carry;
    if (threadIdx.x == 0) carry = Pair{0, 0};
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t base = 0; base < nb; base += 1024u) {
        const uint32_t i = base + threadIdx.x;
        Pair v = i < nb ? blockSums[i] : Pair{0, 0};
        // wave inclusive scans (64-bit for slot counts)
        uint32_t xa = v.active;
        uint64_t xn = v.n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t ya = __shfl_up(xa, d, 64);
            const uint64_t yn = __shfl_up(xn, d, 64);
            if (lane >= (uint32_t)d) {
                xa += ya;
                xn += yn;
            }
        }
        if (lane == 63u) {
            sA[wave] = xa;
            sN[wave] = xn;
        }
        __syncthreads();
        uint32_t wa = 0;
        uint64_t wn = 0;
        for (uint32_t w = 0; w < wave; ++w) {
            wa += sA[w];
            wn += sN[w];
        }
        const Pair c = carry;
        if (i < nb) blockSums[i] = Pair{c.active + wa + xa - v.active, c.n + wn + xn - v.n};
        __syncthreads();
        if (threadIdx.x == 1023u) carry = Pair{c.active + wa + xa, c.n + wn + xn};
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// pass C: ranks, page table, per-leaf plans
__global__ void __launch_bounds__(256) k_scan_c(AssetDev a, const Pair* blockBase, uint32_t* compact, uint32_t* status,
                                                DevLeaf* leaves, uint64_t* ctrl)
{
    __shared__ uint32_t sA[4];
    __shared__ uint32_t sN[4];
    const uint32_t l = blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t act = 0, n = 0, st = LEAF_OK;
    LeafPlan pl;
    if (l < a.L) leaf_count(a, l, act, n, st, pl);
    const uint64_t m = __ballot(act);
    const uint32_t rankInWave = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    uint32_t tot;
    const uint32_t nEx = wave_prefix_u32(n, lane, tot);
    if (lane == 0) {
        sA[wave] = (uint32_t)__popcll(m);
        sN[wave] = tot;
    }
    __syncthreads();
    uint32_t wa = 0, wn = 0;
    for (uint32_t w = 0; w < wave; ++w) {
        wa += sA[w];
        wn += sN[w];
    }
    if (l >= a.L) return;
    const Pair base = blockBase[blockIdx.x];
    status[l] = st;
    if (!act) {
        compact[l] = NO_LEAF;
        return;
    }
    const uint32_t c = base.active + wa + rankInWave;
    compact[l] = c;
    ctrl[c] = base.n + wn + nEx;
    DevLeaf d;
    d.leaf = l;
    d.word = a.offsets[l];
    d.bytes = (a.offsets[l + 1] - d.word) * 4u;
    d.status = st;
    d.ncache = n;
    d.cc = a.profile == 2u ? 0u : control_count(pl.coarseRes);
    uint32_t bx, by, bz;
    leaf_coords(l, a.W, a.H, a.D, bx, by, bz, d.lw, d.lh, d.ld);
    d.sciHeader = a.profile == 2u ? a.payload[(size_t)d.word * 4u] | ((uint32_t)a.payload[(size_t)d.word * 4u + 1] << 8) |
                                        ((uint32_t)a.payload[(size_t)d.word * 4u + 2] << 16) |
                                        ((uint32_t)a.payload[(size_t)d.word * 4u + 3] << 24)
                                  : 0u;
    d.sciEnd = 0xFFFFFFFFu; // set by validation
    d.plan = pl;
    leaves[c] = d;
}

// one past the last payload byte a scientific leaf decode reads (the
// reference reads through std::vector::at over the whole pool, so a read past
// the leaf but inside the pool is legal; past the pool it throws)
__device__ inline uint64_t sci_read_end(const AssetDev& a, const DevLeaf& d)
{
    const SciHeader h = sci_header(d.sciHeader);
    const Bytes b{a.payload};
    const uint64_t base = (uint64_t)d.word * 4u;
    uint64_t end = base + 4u + 64u * (h.keep + 1u);
    if (h.mode == 1u && h.count > 0u) {
        const uint32_t cap = h.tier == 0 ? 64u : h.tier == 1 ? 256u : h.tier == 2 ? 768u : 1024u;
        const uint64_t sparse = base + 4u + 64u * (h.keep + 1u);
        const uint64_t binsBase = sparse + 68u;
        if (binsBase + 32u > a.payloadBytes) return binsBase + 32u;
        // the binary search visits indices below max(bins[], count)
        uint32_t maxIdx = h.count;
        for (uint32_t i = 0; i < 16u; ++i) {
            const uint32_t v = b.u16((uint32_t)(binsBase + 2u * i));
            maxIdx = v > maxIdx ? v : maxIdx;
        }
        const uint64_t coordsEnd = binsBase + 32u + 2u * (uint64_t)maxIdx;
        const uint64_t nibEnd = binsBase + 32u + 2u * (uint64_t)cap + ((uint64_t)maxIdx + 1u) / 2u;
        end = coordsEnd > nibEnd ? coordsEnd : nibEnd;
        if (binsBase + 32u > end) end = binsBase + 32u;
    } else if (h.mode >= 2u && !(h.subtype == 1u && (h.tk == 0u || h.bits == 0u))) {
        const uint32_t R = h.mode == 2u ? 3u : 4u;
        const uint64_t vals = (uint64_t)R * R * R;
        const uint64_t blockBytes = (vals * h.bits + 7u) / 8u;
        const uint64_t scales = base + 4u + 64u * (h.keep + 1u);
        const uint64_t nb = h.subtype == 1u ? h.tk : a.F;
        const uint64_t lastBit = (vals - 1u) * h.bits;
        const uint64_t lastByte = lastBit / 8u + ((lastBit % 8u) + h.bits > 8u ? 2u : 1u);
        end = scales + 2u * nb + (nb - 1u) * blockBytes + lastByte;
        if (scales + 2u * nb > end) end = scales + 2u * nb;
    }
    return end;
}

// with ASSET_SCI_DENSE_BASIS, a subtype-2/3 leaf: one past the last payload
// byte its decode reads, and LEAF_ERR_SCI_BASIS for a block the encoder
// cannot have written (active count != popcount of the mask, tile size above
// 8, patch size or local resolution of 1 with active patches), whose decode
// would index outside the block
__device__ inline uint32_t sci_basis_check(const AssetDev& a, const DevLeaf& d, const SciHeader& h, uint64_t& end)
{
    const uint64_t base = (uint64_t)d.word * 4u;
    const uint64_t hdr = base + 4u + 64u * (h.keep + 1u) + (h.subtype == 3u ? 14u : 12u);
    end = hdr;
    if (hdr > a.payloadBytes) return LEAF_ERR_SCI_RANGE;
    const SciBasis s = sci_basis(Bytes{a.payload + base}, h);
    if (s.values > 0u) end = base + s.dataOff + (uint64_t)h.tk * s.basisBytes;
    if ((uint32_t)__popcll(s.mask) != s.count) return LEAF_ERR_SCI_BASIS;
    if (!s.patch && s.size > 8u) return LEAF_ERR_SCI_BASIS;
    if (s.patch && s.count != 0u && (s.size == 1u || s.res == 1u)) return LEAF_ERR_SCI_BASIS;
    if (end > a.payloadBytes || end - base >= 0xFFFFFFFFull) return LEAF_ERR_SCI_RANGE;
    return LEAF_OK;
}

// validation: one workgroup per active leaf
__global__ void __launch_bounds__(256) k_validate(AssetDev a, DevLeaf* leaves, uint32_t* status)
{
    __shared__ uint32_t sErr, sMono, sNoNan;
    const uint32_t c = blockIdx.x;
    DevLeaf& d = leaves[c];
    if (threadIdx.x == 0) {
        sErr = d.status & STATUS_MASK;
        sMono = 1u;
        sNoNan = 1u;
    }
    __syncthreads();
    if (a.profile == 2u) {
        const SciHeader h = sci_header(d.sciHeader);
        if (threadIdx.x == 0) {
            uint32_t st = LEAF_OK;
            if (h.keep > 1u && h.keep - 1u > a.scaleCount) st = LEAF_ERR_SCI_RANGE;
            uint64_t end;
            if (sci_basis_leaf(a, h)) {
                const uint32_t bs = sci_basis_check(a, d, h, end);
                if (st == LEAF_OK) st = bs;
            } else {
                end = sci_read_end(a, d);
                if (end > a.payloadBytes) st = LEAF_ERR_SCI_RANGE;
            }
            const uint64_t rel = end - (uint64_t)d.word * 4u;
            d.sciEnd = rel < 0xFFFFFFFFull ? (uint32_t)rel : 0xFFFFFFFFu;
            sErr = st;
        }
        __syncthreads();
        // events of one bin in strictly increasing key order: the binary
        // search of a frame then finds exactly the events a scatter of the
        // bin finds (the frame sweep's precondition); reads stay inside the
        // extent checked above
        if (sErr == LEAF_OK && h.mode == 1u && h.count > 0u) {
            const Bytes b{a.payload + (size_t)d.word * 4u};
            const uint32_t binsBase = 4u + 64u * (h.keep + 1u) + 68u, coordsBase = binsBase + 32u;
            for (uint32_t bin = 0; bin < 16u; ++bin) {
                const uint32_t start = b.u16(binsBase + bin * 2u);
                const uint32_t end = bin + 1u < 16u ? b.u16(binsBase + (bin + 1u) * 2u) : (h.count & 0xFFFFu);
                for (uint32_t i = start + 1u + threadIdx.x; i < end; i += blockDim.x)
                    if (b.u16(coordsBase + 2u * i) <= b.u16(coordsBase + 2u * (i - 1u))) sMono = 0u;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t st = sErr | (sErr == LEAF_OK && sMono ? FLAG_SWEEP : 0u);
            d.status = st;
            status[d.leaf] = st;
        }
        return;
    }
    if (sErr == LEAF_OK) {
        const Bytes b{a.payload + (size_t)d.word * 4u};
        const StreamRef* streams[3] = {&d.plan.coarse, d.plan.mode == MODE_FINE6 || d.plan.hasFine ? &d.plan.fine : nullptr,
                                       d.plan.mode == MODE_SHELL ? &d.plan.shell : nullptr};
        for (int s = 0; s < 3; ++s) {
            if (!streams[s]) continue;
            const StreamRef sr = *streams[s];
            if (sr.codec == CODEC_FP16)
                for (uint32_t i = threadIdx.x; i < sr.keys; i += blockDim.x) {
                    const uint32_t h = b.u16(sr.keyBase + 3u * i + 1u);
                    if ((h & 0x7C00u) == 0x7C00u && (h & 0x3FFu) != 0u) sNoNan = 0u;
                }
            for (uint32_t i = threadIdx.x; i < sr.count; i += blockDim.x) {
                const uint32_t ks = b.u16(sr.base + 4u * i), kc = b.u8(sr.base + 4u * i + 2u);
                if (kc == 0u) continue;
                if (ks + kc > sr.keys) {
                    atomicMax(&sErr, (uint32_t)LEAF_ERR_KEY_RANGE);
                    continue;
                }
                int prev = key_frame(b, sr.keyBase, sr.codec, ks);
                for (uint32_t j = 1; j < kc; ++j) {
                    const int f = key_frame(b, sr.keyBase, sr.codec, ks + j);
                    if (f <= prev) {
                        sMono = 0u;
                        break;
                    }
                    prev = f;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t st =
            sErr | (sErr == LEAF_OK && sMono ? FLAG_SWEEP : 0u) | (sErr == LEAF_OK && sNoNan ? FLAG_NO_NAN : 0u);
        d.status = st;
        status[d.leaf] = st;
    }
}

// ---------------------------------------------------------- cache: render --
// The packed cache is processed in tiles of BLOCK consecutive slots; the
// first active leaf of every CACHE_GROUP-slot group is tabulated at asset
// open so a workgroup of any size (a multiple of 64) finds its leaf with one
// load.

__global__ void k_group_leaf(const uint64_t* ctrl, uint32_t A, uint64_t total, uint32_t ng, uint32_t* groupLeaf)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > ng) return;
    const uint64_t g = (uint64_t)b * CACHE_GROUP;
    groupLeaf[b] = g < total ? find_leaf(ctrl, A, g) : (A ? A - 1u : 0u);
}

// The slot plan and group ranges of a render asset's packed cache (see
// slot_plan): one thread per slot, blocks of 256 = 4 whole groups.
__global__ void __launch_bounds__(256) k_slot_plan(AssetDev a, uint64_t total, uint2* plan, uint2* groupRange)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    if (g < total) {
        const uint2 p = slot_plan(a, g);
        plan[g] = p;
        const uint32_t kind = plan_kind(p), keyCount = p.y & 0xFFu;
        if (plan_swept(kind)) {
            lo = p.x;
            hi = p.x + keyCount * (kind == PLAN_FP32 ? 8u : 3u);
        } else if (kind == PLAN_SLOW) {
            lo = 0u;
            hi = 0xFFFFFFFFu;
        }
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if ((threadIdx.x & 63u) == 0u && g < total) groupRange[g / CACHE_GROUP] = make_uint2(lo, hi);
}

template <int LAYOUT, uint32_t BLOCK, uint32_t WINDOW>
__global__ void __launch_bounds__(BLOCK) k_cache_render(AssetDev a, const uint32_t* tileGroup, uint32_t nTiles,
                                                        uint32_t runLog2, int t0, int nt, float* out, uint64_t total,
                                                        uint32_t* errFlag)
{
    __shared__ uint32_t lds[WINDOW / 4 + 4]; // + 4: reads past the window, see sweep_frames
    __shared__ uint32_t sRed[BLOCK / 64];
    if (LAYOUT == VBT_CACHE_PACKED) {
        // tile i = groups [tileGroup[i], tileGroup[i + 1]) (plan_tiles), in XCD runs
        const uint32_t tile = xcd_run_tile(blockIdx.x, runLog2);
        if (tile >= nTiles) return;
        const uint32_t g0 = tileGroup[tile], g1 = tileGroup[tile + 1u];
        cache_plan_tile<BLOCK, WINDOW>(a, g0, g1 - g0, t0, nt, out, total, errFlag, lds, sRed);
    } else {
        cache_render_tile<LAYOUT, BLOCK, WINDOW>(a, a.groupLeaf, blockIdx.x, t0, nt, out, total, errFlag, lds, sRed);
    }
}

// The memory traffic of a packed frame-cache build without its decode: the
// same tiles, each tile's keyframe record span staged into LDS the same way
// (reads != 0; tiles holding a slot that takes the bracket search stage
// nothing) and the same write-back stores of every frame's row segment, with
// a value taken from the staged words so that the reads stay live.  It times
// the ceiling of the build for this asset and layout; what it writes is not a
// frame cache.
template <uint32_t BLOCK, uint32_t WINDOW>
__global__ void __launch_bounds__(BLOCK) k_cache_ceiling(AssetDev a, const uint32_t* tileGroup, uint32_t nTiles,
                                                         uint32_t runLog2, int nt, int reads, float* out,
                                                         uint64_t total)
{
    __shared__ uint32_t lds[WINDOW / 4];
    const uint32_t tid = threadIdx.x;
    const uint32_t tile = xcd_run_tile(blockIdx.x, runLog2);
    if (tile >= nTiles) return;
    const uint32_t g0 = tileGroup[tile], ng = tileGroup[tile + 1u] - g0;
    const uint64_t g = (uint64_t)g0 * CACHE_GROUP + tid;
    const bool live = tid < ng * CACHE_GROUP && g < total;
    uint32_t v = tid;
    if (reads) {
        uint32_t tlo = 0xFFFFFFFFu, thi = 0u;
#pragma unroll
        for (uint32_t w = 0; w < BLOCK / CACHE_GROUP; ++w) {
            if (w < ng) {
                const uint2 r = invariant_u2(a.groupRange, g0 + w);
                tlo = min(tlo, r.x);
                thi = max(thi, r.y);
            }
        }
        if (tlo != 0xFFFFFFFFu && thi != 0xFFFFFFFFu) {
            const uint32_t w0 = tlo / 4u, w1 = (thi + 3u) / 4u;
            uint32_t acc = 0u;
            for (uint32_t p = w0; p < w1; p += WINDOW / 4u) {
                const uint32_t nw = min(w1 - p, WINDOW / 4u);
                stage_window<BLOCK, 32u>(a.payload, p, nw, lds, tid);
                __syncthreads();
                acc += lds[tid % nw];
                __syncthreads();
            }
            v = acc;
        }
    }
    if (!live) return;
    char* ob = (char*)out;
    const uint64_t strideBytes = total * 4u;
    for (int f = 0; f < nt; ++f, ob += strideBytes)
        __builtin_amdgcn_raw_buffer_store_b32(v + (uint32_t)f, frame_rsrc(ob, (uint32_t)strideBytes), (int)(g * 4u), 0,
                                              STORE_WB);
}

// Few frames: one thread per cache slot, the reference bracket search per
// frame straight from the payload (about three keyframe records per frame
// instead of the slot's whole keyframe stream).
template <int LAYOUT>
__global__ void __launch_bounds__(256) k_cache_exact(AssetDev a, int t0, int nt, float* out, uint64_t total,
                                                     uint32_t* errFlag)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= total) return;
    uint32_t c = a.groupLeaf[g / CACHE_GROUP];
    while (c + 1u < a.A && a.ctrl[c + 1u] <= g) ++c;
    const DevLeaf& d = a.leaves[c];
    const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
    const bool isCoarse = slot < d.cc;
    uint64_t stride = total, dst = g;
    if (LAYOUT != VBT_CACHE_PACKED) {
        const uint32_t s576 = isCoarse ? slot : 64u + (slot - d.cc);
        if ((isCoarse && slot >= 64u) || s576 >= 576u) return;
        stride = (uint64_t)a.A * 576u;
        dst = (uint64_t)c * 576u + s576;
    }
    float* o = out + dst;
    if ((d.status & STATUS_MASK) != LEAF_OK) {
        for (int i = 0; i < nt; ++i) o[(uint64_t)i * stride] = __builtin_nanf("");
        atomicOr(errFlag, 1u);
        return;
    }
    const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
    const uint32_t ci = isCoarse ? slot : slot - d.cc;
    const uint32_t* W = (const uint32_t*)(a.payload + (size_t)d.word * 4u);
    for (int i = 0; i < nt; ++i) o[(uint64_t)i * stride] = decode_control_w(W, s, ci, (int)a.F, t0 + i);
}

// ------------------------------------------- frame index for few-frame builds --
// A copy of each slot's keyframe records regrouped by time bin: bin b holds,
// for every packed-cache slot in order, the slot's records from the last key
// at or before the bin's first frame to the first key at or after its last
// frame, as {frame, value bits}.  A one- or two-frame build then reads each
// slot's bracket from a short run that sits next to its neighbours' runs,
// instead of one line of the keyframe pool per slot.  Leaves whose keys do not
// strictly increase, poisoned leaves and empty streams keep no run and take
// the reference bracket search (cache_slot_exact).
// Bins of the frame range the index keeps runs for: a one-frame build reads
// its bin's run of every slot (about keys / bins + 2 records), the index holds
// about keys + 2 x bins records per slot.  32 bins against 8: liquid320
// one-frame builds 99 -> 63 us, smoke256 15 -> 13 us, for an index of 3.7 x
// the payload instead of 1.8 x.  At most one bin per frame.
constexpr uint32_t INDEX_BINS_MAX = 32u;
__host__ __device__ inline uint32_t index_bins(uint32_t F) { return F < INDEX_BINS_MAX ? (F ? F : 1u) : INDEX_BINS_MAX; }

__device__ __forceinline__ uint32_t index_bin(int t, uint32_t F, uint32_t B) { return min(B - 1u, (uint32_t)t * B / F); }

// the run [lo, hi] of records of one slot for bin b; false when the slot keeps none
__device__ __forceinline__ bool index_run(const AssetDev& a, uint64_t g, uint32_t b, uint32_t B, const uint32_t*& W,
                                          StreamRef& s, uint32_t& k0, uint32_t& lo, uint32_t& hi)
{
    uint32_t c = a.groupLeaf[g / CACHE_GROUP];
    while (c + 1u < a.A && a.ctrl[c + 1u] <= g) ++c;
    const DevLeaf& d = a.leaves[c];
    if ((d.status & STATUS_MASK) != LEAF_OK || !(d.status & FLAG_SWEEP)) return false;
    const uint32_t F = a.F;
    const uint32_t t0 = (b * F + B - 1u) / B;
    const uint32_t t1 = b + 1u < B ? ((b + 1u) * F + B - 1u) / B - 1u : F - 1u;
    if (t0 > t1) return false;
    const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
    const bool isCoarse = slot < d.cc;
    s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
    const uint32_t ci = isCoarse ? slot : slot - d.cc;
    W = (const uint32_t*)(a.payload + (size_t)d.word * 4u);
    const uint32_t desc = W[(s.base + ci * 4u) >> 2];
    const uint32_t keyStart = desc & 0xFFFFu, keyCount = (desc >> 16) & 0xFFu;
    if (keyCount == 0u) return false;
    const uint32_t kb = keyframe_bytes(s.codec);
    k0 = s.keyBase + keyStart * kb;
    // n0 = keys with frame <= t0, n1 = keys with frame < t1 (frames increase)
    uint32_t l = 0u, h = keyCount;
    while (l < h) {
        const uint32_t m = (l + h) / 2u;
        int f;
        float v;
        w_rec(W, k0 + m * kb, s.codec, f, v);
        if ((uint32_t)f <= t0) l = m + 1u;
        else h = m;
    }
    const uint32_t n0 = l;
    l = 0u;
    h = keyCount;
    while (l < h) {
        const uint32_t m = (l + h) / 2u;
        int f;
        float v;
        w_rec(W, k0 + m * kb, s.codec, f, v);
        if ((uint32_t)f < t1) l = m + 1u;
        else h = m;
    }
    lo = n0 > 0u ? n0 - 1u : 0u;
    hi = l < keyCount ? l : keyCount - 1u;
    return true;
}

// counts per bin and slot; *wide is set when a run holds fp32 records (the
// index then stores 8-byte records, else 4-byte {frame, half bits})
__global__ void __launch_bounds__(256) k_index_count(AssetDev a, uint32_t* cnt, uint64_t C, uint32_t B, uint32_t* wide)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g > C) return;
    bool fp32 = false;
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t* W;
        StreamRef s;
        uint32_t k0, lo, hi;
        const bool run = g < C && index_run(a, g, b, B, W, s, k0, lo, hi);
        cnt[b * (C + 1u) + g] = run ? hi - lo + 1u : 0u;
        fp32 = fp32 || (run && s.codec == CODEC_FP32);
    }
    if (fp32) atomicOr(wide, 1u);
}

// WIDE: {frame, value bits}; else {frame | half bits << 16} (fp16 records only)
template <bool WIDE>
__global__ void __launch_bounds__(256) k_index_scatter(AssetDev a, const uint32_t* off, void* rec, uint64_t C,
                                                       uint32_t B)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= C) return;
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t* W;
        StreamRef s;
        uint32_t k0, lo, hi;
        if (!index_run(a, g, b, B, W, s, k0, lo, hi)) continue;
        const uint32_t at = off[b * (C + 1u) + g];
        const uint32_t kb = keyframe_bytes(s.codec);
        for (uint32_t k = lo; k <= hi; ++k) {
            const uint32_t ro = k0 + k * kb;
            if constexpr (WIDE) {
                int f;
                float v;
                w_rec(W, ro, s.codec, f, v);
                ((uint2*)rec)[at + k - lo] = make_uint2((uint32_t)f, __float_as_uint(v));
            } else {
                const uint32_t x = __builtin_amdgcn_alignbyte(W[(ro >> 2) + 1u], W[ro >> 2], ro & 3u);
                ((uint32_t*)rec)[at + k - lo] = (x & 0xFFu) | (((x >> 8) & 0xFFFFu) << 16);
            }
        }
    }
}

template <bool WIDE>
__device__ __forceinline__ void index_rec(const void* rec, uint32_t j, int& f, float& v)
{
    if constexpr (WIDE) {
        const uint2 r = ((const uint2*)rec)[j];
        f = (int)r.x;
        v = __uint_as_float(r.y);
    } else {
        const uint32_t r = ((const uint32_t*)rec)[j];
        f = (int)(r & 0xFFFFu);
        v = half_to_float(r >> 16);
    }
}

// the sum of n counts, so the host can check it fits 32-bit record offsets
__global__ void __launch_bounds__(256) k_sum64(const uint32_t* cnt, uint64_t n, unsigned long long* sum)
{
    unsigned long long acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) acc += cnt[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63u) == 0u && acc) atomicAdd(sum, acc);
}

// Packed cache, nt <= 2 frames, from the frame index: the same values as
// k_cache_exact (for strictly increasing keys the reference bracket is the
// last key at or before the frame and the next one).
template <bool WIDE>
__global__ void __launch_bounds__(256) k_cache_indexed(AssetDev a, const uint32_t* off, const void* rec, uint32_t B,
                                                       int t0, int nt, float* out, uint64_t total, uint32_t* errFlag)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (g >= total) return;
    for (int i = 0; i < nt; ++i) {
        const int t = t0 + i;
        const uint32_t b = index_bin(t, a.F, B);
        const uint32_t o0 = off[b * (total + 1u) + g], o1 = off[b * (total + 1u) + g + 1u];
        float val;
        if (o0 == o1) {
            int32_t st;
            val = cache_slot_exact(a, g, t, st);
            if (st) atomicOr(errFlag, 1u);
        } else {
            uint32_t j = o0;
            int f0, f1 = 0;
            float v0, v1 = 0.0f;
            index_rec<WIDE>(rec, j, f0, v0);
            if (t < f0) {
                val = v0;
            } else {
                if (j + 1u < o1) index_rec<WIDE>(rec, j + 1u, f1, v1);
                while (j + 1u < o1 && f1 <= t) {
                    ++j;
                    f0 = f1;
                    v0 = v1;
                    if (j + 1u < o1) index_rec<WIDE>(rec, j + 1u, f1, v1);
                }
                if (f0 == t || j + 1u == o1) {
                    val = v0;
                } else {
                    const float alpha = (float)(t - f0) / (float)(f1 - f0);
                    val = v0 + alpha * (v1 - v0);
                }
            }
        }
        out[(uint64_t)i * total + g] = val;
    }
}

// Packed layout: grid = the shape's tile list (tileGroup, nTiles entries + 1)
// in XCD runs; 576 layout: one tile per BLOCK slots.
template <int LAYOUT>
void launch_cache_render(int tile, const AssetDev& ad, const uint32_t* tileGroup, uint32_t nTiles, int t0, int nt,
                         float* out, uint64_t total, uint32_t* err, hipStream_t s)
{
    const uint32_t r = tile_run_log2(nTiles);
    switch (tile) {
#define VBT_TILE(i, B, W)                                                                                   \
    case i:                                                                                                 \
        k_cache_render<LAYOUT, B, W><<<LAYOUT == VBT_CACHE_PACKED ? run_grid(nTiles, r) : blocks_for(total, B), B, \
                                       0, s>>>(ad, tileGroup, nTiles, r, t0, nt, out, total, err);          \
        break;
        VBT_TILE(0, 64, 8192)
        VBT_TILE(1, 64, 16384)
        VBT_TILE(2, 128, 16384)
        VBT_TILE(3, 128, 32768)
        VBT_TILE(4, 256, 32768)
        VBT_TILE(5, 256, 49152)
        VBT_TILE(6, 64, 4096)
        VBT_TILE(7, 128, 8192)
        VBT_TILE(8, 256, 16384)
        VBT_TILE(9, 64, 12288)
#undef VBT_TILE
    default:
        break;
    }
}

void launch_cache_ceiling(int tile, const AssetDev& ad, const uint32_t* tileGroup, uint32_t nTiles, int nt, int reads,
                          float* out, uint64_t total, hipStream_t s)
{
    const uint32_t r = tile_run_log2(nTiles);
    switch (tile) {
#define VBT_TILE(i, B, W)                                                                                  \
    case i:                                                                                                \
        k_cache_ceiling<B, W><<<run_grid(nTiles, r), B, 0, s>>>(ad, tileGroup, nTiles, r, nt, reads, out, total); \
        break;
        VBT_TILE(0, 64, 8192)
        VBT_TILE(1, 64, 16384)
        VBT_TILE(2, 128, 16384)
        VBT_TILE(3, 128, 32768)
        VBT_TILE(4, 256, 32768)
        VBT_TILE(5, 256, 49152)
        VBT_TILE(6, 64, 4096)
        VBT_TILE(7, 128, 8192)
        VBT_TILE(8, 256, 16384)
        VBT_TILE(9, 64, 12288)
#undef VBT_TILE
    default:
        break;
    }
}

// ------------------------------------------------------ scientific decode --
// scientific dense frame: one workgroup per active leaf, 512 lanes = voxels
__global__ void __launch_bounds__(512) k_sci_dense(AssetDev a, int t, float* out, uint32_t* errFlag)
{
    __shared__ float ctrl[64];
    const uint32_t c = blockIdx.x;
    const DevLeaf& d = a.leaves[c];
    const uint32_t tid = threadIdx.x;
    const uint32_t lx = tid & 7u, ly = (tid >> 3) & 7u, lz = tid >> 6;
    const bool inside = lx < d.lw && ly < d.lh && lz < d.ld;
    uint32_t bx, by, bz, lw, lh, ld;
    leaf_coords(d.leaf, a.W, a.H, a.D, bx, by, bz, lw, lh, ld);
    const uint64_t dst = ((uint64_t)(bz * 8u + lz) * a.H + (by * 8u + ly)) * a.W + bx * 8u + lx;
    if ((d.status & STATUS_MASK) != LEAF_OK) {
        if (inside) out[dst] = __builtin_nanf("");
        if (tid == 0) atomicOr(errFlag, 1u);
        return;
    }
    const SciHeader h = sci_header(d.sciHeader);
    const Bytes b{a.payload + (size_t)d.word * 4u};
    if (tid < 64u) ctrl[tid] = sci_control(a, b, h, tid, t);
    __syncthreads();
    if (!inside) return;
    out[dst] = sci_coarse_sample(ctrl, lx, ly, lz) + sci_residual(a, b, h, lx, ly, lz, (uint32_t)t);
}

// ---------------------------------------------- scientific frame sweep --
// Packed caches of a run of frames: one workgroup per (active leaf, run of
// frames).  A leaf whose decode reads at most SCI_STAGE_BYTES (every sparse
// and empty-residual leaf the encoder writes) is staged into LDS once, a
// larger one (a dense leaf's per-frame blocks) is read from the payload.  The
// 64 x keep coefficients (fp16 DC, int8 code x scale table entry) are
// expanded once; the run is then swept SCI_GROUP frames at a time.  One wave
// per frame evaluates the 64 controls (the FP64 basis sum of sci_control,
// term by term in the same order) and carries the coarse trilinear through
// its x and y passes with lane shuffles: trilinear() lerps x first, then y,
// then z, so the 4x4x8 x-pass and 4x8x8 y-pass values are exactly the
// intermediates every voxel of the lattice row would compute.  After one
// barrier each voxel's thread finishes its z lerp per frame, adds the
// residual and stores.
constexpr uint32_t SCI_STAGE_BYTES = 4096;
constexpr uint32_t SCI_GROUP = 8;      // frames per control pass, one wave each
constexpr uint32_t SCI_THREADS = 64u * SCI_GROUP; // also one thread per voxel
constexpr uint64_t SCI_TARGET_GROUPS = 65536; // split runs until the grid holds this many workgroups
constexpr uint32_t SCI_RUN_MAX = 64;  // frames per workgroup at most (their basis rows sit in LDS)

// a staged leaf: every byte its decode reads is in LDS
struct LdsBytes {
    const uint8_t* p;
    __device__ uint32_t u8(uint32_t o) const { return p[o]; }
    __device__ uint32_t u16(uint32_t o) const { return (uint32_t)p[o] | ((uint32_t)p[o + 1u] << 8); }
    __device__ uint32_t u32(uint32_t o) const { return u16(o) | (u16(o + 2u) << 16); }
    __device__ uint32_t u16a(uint32_t o) const { return *reinterpret_cast<const uint16_t*>(p + o); } // o even
};

// u16 at an even offset of a leaf (leaves are word-aligned)
__device__ inline uint32_t u16_even(const LdsBytes& b, uint32_t o) { return b.u16a(o); }
__device__ inline uint32_t u16_even(const Bytes& b, uint32_t o) { return *reinterpret_cast<const uint16_t*>(b.p + o); }

__device__ inline float lane_float(float v, int src) { return __shfl(v, src, 64); }

// coef: [k][control] for k < keep, then the run's basis rows [frame][k]
// (dynamic LDS sized by the asset's largest keep): the loop body issues no
// vector-memory load, whose wait would also wait for the stores before it.
// yp: per buffer and frame of a group, the y-pass values
// [cz][y][x] (256 floats).  evt: per voxel, one byte per frame of a group,
// 0x10 | event code where the frame has an event at the voxel.  yp and evt
// are double-buffered: a group's waves write one buffer before the barrier
// that the group's voxel pass reads it after; each voxel's thread clears its
// evt bytes as it reads them.
template <bool DENSE, class B>
__device__ inline void sci_sweep(const AssetDev& a, const B& b, const SciHeader& h, bool sorted, float* coef,
                                 float (*yp)[SCI_GROUP * 256u], uint64_t (*evt)[512], int t0, int tb, int te,
                                 float* out, uint64_t total, uint32_t dst, bool in)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t lx = tid & 7u, ly = (tid >> 3) & 7u, lz = tid >> 6;
    for (uint32_t i = tid; i < 64u * h.keep; i += SCI_THREADS) {
        const uint32_t ci = i & 63u, k = i >> 6;
        const uint32_t cb = 4u + ci * (h.keep + 1u);
        coef[i] = k == 0u ? half_to_float(b.u16(cb))
                          : (float)(int8_t)(uint8_t)b.u8(cb + 2u + (k - 1u)) * a.scales[k - 1u];
    }
    float* basisL = coef + 64u * h.keep;
    for (uint32_t i = tid; i < (uint32_t)(te - tb) * h.keep; i += SCI_THREADS)
        basisL[i] = a.basis[(size_t)(tb + (int)(i / h.keep)) * 64u + i % h.keep];
    // sparse residuals: sci_sparse is 0 unless the leaf has events, a positive
    // event scale and the voxel's mask bit, all fixed over the frames
    const uint32_t idx = (lz * 8u + ly) * 8u + lx; // == tid
    const uint32_t sparseBase = 4u + 64u * (h.keep + 1u);
    const uint32_t binsBase = sparseBase + 68u, coordsBase = binsBase + 32u;
    const uint32_t residualBase = DENSE ? 0u : coordsBase + sci_event_cap(h) * 2u;
    float eventScale = 0.0f;
    bool events = false, act = false;
    if (!DENSE && h.mode == 1u && h.count != 0u) {
        eventScale = bits_to_float(b.u32(sparseBase));
        events = eventScale > 0.0f;
        if (events) act = ((b.u32(sparseBase + 4u + (idx >> 5) * 4u) >> (idx & 31u)) & 1u) != 0u;
    }
    // a sorted leaf's events are scattered per frame group instead of each
    // voxel searching its frame's bin
    const bool scatter = events && sorted;
    if (scatter) {
        evt[0][tid] = 0u;
        evt[1][tid] = 0u;
    }
    const uint32_t j = tid >> 6, ci = tid & 63u;
    // the control wave's lane constants: x-pass entries ci and ci + 64
    // ([cz * 4 + cy][x], x = ci & 7), y-pass entries ci + 64 k ([cz = k][y][x],
    // y = ci >> 3); sci_coarse_coord gives each lattice cell and weight
    int xc, yc;
    float tx, ty;
    sci_coarse_coord(ci & 7u, xc, tx);
    sci_coarse_coord(ci >> 3, yc, ty);
    const int xs0 = (int)(ci >> 3) * 4 + xc;       // control lane of x-pass entry ci (row ci >> 3)
    const int ys0 = yc * 8 + (int)(ci & 7u);      // x-pass lane of y-pass entry ci, cz even
    // this voxel's z lerp
    int zc;
    float tz;
    sci_coarse_coord(lz, zc, tz);
    const uint32_t yo = (uint32_t)zc * 64u + ly * 8u + lx;
    __syncthreads();

    int g = 0;
    for (int tg = tb; tg < te; tg += (int)SCI_GROUP, g ^= 1) {
        const int t = __builtin_amdgcn_readfirstlane(tg + (int)j);
        if (t < te) {
            // lane k holds the frame's basis term k; the sum reads it back
            // lane by lane
            const int bl = ci < h.keep ? __builtin_bit_cast(int, basisL[(uint32_t)(t - tb) * h.keep + ci]) : 0;
            double sum = (double)coef[ci] * (double)__builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(bl));
            // a product of two floats is exact in double, so fma rounds as
            // sci_control's sum + term does
            uint32_t k = 1;
            for (; k + 4u <= h.keep; k += 4u) { // four LDS reads in flight per wait
                float cf[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) cf[u] = coef[(k + u) * 64u + ci];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    sum = __builtin_fma((double)cf[u],
                                        (double)__builtin_bit_cast(float, __builtin_amdgcn_readlane(bl, (int)(k + u))),
                                        sum);
            }
            for (; k < h.keep; ++k)
                sum = __builtin_fma((double)coef[k * 64u + ci],
                                    (double)__builtin_bit_cast(float, __builtin_amdgcn_readlane(bl, (int)k)), sum);
            const float cv = (float)sum; // control ci = (cz * 4 + cy) * 4 + cx
            // x pass (trilinear's c00/c01/c10/c11 for every lattice row)
            const float x0 = lane_float(cv, xs0) * (1.0f - tx) + lane_float(cv, xs0 + 1) * tx;
            const float x1 = lane_float(cv, xs0 + 32) * (1.0f - tx) + lane_float(cv, xs0 + 33) * tx;
            // y pass (its c0/c1): cz 0, 1 from x0's lanes, cz 2, 3 from x1's
            float* yw = yp[g] + j * 256u;
            yw[ci] = lane_float(x0, ys0) * (1.0f - ty) + lane_float(x0, ys0 + 8) * ty;
            yw[ci + 64u] = lane_float(x0, ys0 + 32) * (1.0f - ty) + lane_float(x0, ys0 + 40) * ty;
            yw[ci + 128u] = lane_float(x1, ys0) * (1.0f - ty) + lane_float(x1, ys0 + 8) * ty;
            yw[ci + 192u] = lane_float(x1, ys0 + 32) * (1.0f - ty) + lane_float(x1, ys0 + 40) * ty;
        }
        const int ng = te - tg < (int)SCI_GROUP ? te - tg : (int)SCI_GROUP;
        if (scatter && t < te) {
            // the frame's wave scatters the events of its bin
            uint8_t* eb = reinterpret_cast<uint8_t*>(evt[g]);
            const uint32_t bin = sci_event_bin(h, (uint32_t)t);
            const uint32_t lt = (((uint32_t)t - bin * h.fpb) & 0xFFu) & 0x7Fu; // the key keeps 7 bits of it
            // binsBase and coordsBase are even
            const uint32_t start = u16_even(b, binsBase + bin * 2u);
            const uint32_t end = bin + 1u < 16u ? u16_even(b, binsBase + (bin + 1u) * 2u) : (h.count & 0xFFFFu);
            for (uint32_t i = start + ci; i < end; i += 64u) {
                const uint32_t key = u16_even(b, coordsBase + 2u * i);
                if ((key >> 9) == lt)
                    eb[(key & 0x1FFu) * 8u + j] = (uint8_t)(0x10 | (sci_event_q(b, residualBase, i) & 0xF));
            }
        }
        __syncthreads(); // the buffers written two groups on are read by no wave still behind this barrier
        uint64_t row = 0;
        if (scatter) {
            row = evt[g][tid];
            evt[g][tid] = 0u;
        }
        if (!in) continue;
        const float* yr = yp[g] + yo;
        float* of = out + (uint64_t)(tg - t0) * total; // wave-uniform frame base
        if (DENSE) {
            for (int jj = 0; jj < ng; ++jj, of += total) {
                const float r = sci_dense_value(a, b, h, lx, ly, lz, (uint32_t)(tg + jj));
                const float v = yr[jj * 256] * (1.0f - tz) + yr[jj * 256 + 64] * tz;
                __builtin_nontemporal_store(v + r, of + dst);
            }
        } else {
            float r[SCI_GROUP];
#pragma unroll
            for (int jj = 0; jj < (int)SCI_GROUP; ++jj) r[jj] = 0.0f;
            if (scatter) {
                if (act && row != 0u) {
#pragma unroll
                    for (int jj = 0; jj < (int)SCI_GROUP; ++jj) {
                        const uint32_t e = (uint32_t)(row >> (8 * jj)) & 0xFFu;
                        int q = (int)(e & 0xFu);
                        if (q & 8) q -= 16;
                        r[jj] = e != 0u ? (float)q * eventScale : 0.0f;
                    }
                }
            } else if (act) {
                for (int jj = 0; jj < ng; ++jj) r[jj] = sci_sparse(b, h, idx, (uint32_t)(tg + jj));
            }
#pragma unroll
            for (int jj = 0; jj < (int)SCI_GROUP; ++jj, of += total) {
                if (jj >= ng) break;
                const float v = yr[jj * 256] * (1.0f - tz) + yr[jj * 256 + 64] * tz;
                __builtin_nontemporal_store(v + r[jj], of + dst);
            }
        }
    }
}

// DENSE: the leaves with a dense residual (mode 2/3) only, else every other
// leaf, poisoned ones included; the dense decode's registers would otherwise
// halve the occupancy of every leaf's sweep.  Dynamic LDS: 4 * (64 +
// SCI_RUN_MAX) * keepMax bytes of coefficients and basis rows.
template <bool DENSE>
__global__ void __launch_bounds__(SCI_THREADS) k_sci_frames(AssetDev a, int t0, int nt, int run, float* out,
                                                            uint64_t total, uint32_t* errFlag)
{
    __shared__ __align__(16) uint8_t stage[SCI_STAGE_BYTES];
    __shared__ float yp[2][SCI_GROUP * 256u];
    __shared__ uint64_t evt[2][512];
    extern __shared__ float coef[];
    const uint32_t c = blockIdx.x;
    const int tb = t0 + (int)blockIdx.y * run;
    const int te = min(t0 + nt, tb + run);
    const DevLeaf& d = a.leaves[c];
    const uint32_t tid = threadIdx.x;
    const uint32_t lx = tid & 7u, ly = (tid >> 3) & 7u, lz = tid >> 6;
    const bool in = lx < d.lw && ly < d.lh && lz < d.ld;
    // slot of this voxel in a frame: one frame is at most 4 GiB (checked at launch)
    const uint32_t dst = (uint32_t)(a.ctrl[c] + (lz * d.lh + ly) * d.lw + lx);
    const bool ok = (d.status & STATUS_MASK) == LEAF_OK;
    const SciHeader h = sci_header(d.sciHeader);
    if (DENSE ? !ok || h.mode < 2u : ok && h.mode >= 2u) return;
    if (!ok) {
        if (in)
            for (int t = tb; t < te; ++t) out[(uint64_t)(t - t0) * total + dst] = __builtin_nanf("");
        if (tid == 0) atomicOr(errFlag, 1u);
        return;
    }
    const uint8_t* leafG = a.payload + (size_t)d.word * 4u;
    const bool sorted = (d.status & FLAG_SWEEP) != 0u;
    if (d.sciEnd <= SCI_STAGE_BYTES) {
        // validation put the read extent inside the payload, whose size is whole words
        const uint32_t words = (d.sciEnd + 3u) / 4u;
        for (uint32_t i = tid; i < words; i += SCI_THREADS)
            reinterpret_cast<uint32_t*>(stage)[i] = reinterpret_cast<const uint32_t*>(leafG)[i];
        __syncthreads();
        sci_sweep<DENSE>(a, LdsBytes{stage}, h, sorted, coef, yp, evt, t0, tb, te, out, total, dst, in);
    } else {
        sci_sweep<DENSE>(a, Bytes{leafG}, h, sorted, coef, yp, evt, t0, tb, te, out, total, dst, in);
    }
}

static void launch_sci_frames(uint32_t A, uint32_t keepMax, bool dense, const AssetDev& ad, int t0, int nt,
                              float* d_out, uint64_t total, uint32_t* d_err, hipStream_t s)
{
    const size_t ldsBytes = 4u * (64u + SCI_RUN_MAX) * (size_t)keepMax;
    // at most 65535 runs of SCI_RUN_MAX frames per launch
    const uint64_t span = 65535ull * SCI_RUN_MAX;
    for (uint64_t f0 = 0; f0 < (uint64_t)nt; f0 += span) {
        const uint64_t n = std::min<uint64_t>(span, (uint64_t)nt - f0);
        // runs of whole frame groups: enough of them to fill the chip, none
        // longer than SCI_RUN_MAX frames
        uint64_t runs = std::min<uint64_t>((SCI_TARGET_GROUPS + A - 1u) / A, (n + SCI_GROUP - 1u) / SCI_GROUP);
        runs = std::max<uint64_t>(runs, (n + SCI_RUN_MAX - 1u) / SCI_RUN_MAX);
        const uint64_t run = std::min<uint64_t>(((n + runs - 1u) / runs + SCI_GROUP - 1u) / SCI_GROUP * SCI_GROUP,
                                                SCI_RUN_MAX);
        runs = (n + run - 1u) / run;
        float* o = d_out + f0 * total;
        const int ts = t0 + (int)f0;
        k_sci_frames<false><<<dim3(A, (uint32_t)runs), SCI_THREADS, ldsBytes, s>>>(ad, ts, (int)n, (int)run, o, total,
                                                                                    d_err);
        if (dense)
            k_sci_frames<true><<<dim3(A, (uint32_t)runs), SCI_THREADS, ldsBytes, s>>>(ad, ts, (int)n, (int)run, o,
                                                                                       total, d_err);
    }
}

// ------------------------------------------------------ render voxels --
__device__ inline float sample_grid(const float* g, int res, int lw, int lh, int ld, int lx, int ly, int lz)
{
    if (res < 1) return 0.0f;
    int x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    locate((float)lx, lw, res, x0, x1, tx);
    locate((float)ly, lh, res, y0, y1, ty);
    locate((float)lz, ld, res, z0, z1, tz);
    float s[8];
    s[0] = g[(z0 * res + y0) * res + x0];
    s[1] = g[(z0 * res + y0) * res + x1];
    s[2] = g[(z0 * res + y1) * res + x0];
    s[3] = g[(z0 * res + y1) * res + x1];
    s[4] = g[(z1 * res + y0) * res + x0];
    s[5] = g[(z1 * res + y0) * res + x1];
    s[6] = g[(z1 * res + y1) * res + x0];
    s[7] = g[(z1 * res + y1) * res + x1];
    return trilinear(s, tx, ty, tz);
}

// whole frame, render route: one workgroup per leaf, 512 lanes = voxels
__global__ void __launch_bounds__(512) k_dense_render(AssetDev a, int t, ComposeDev cm, float* out, uint32_t* errFlag)
{
    __shared__ float coarse[343];
    __shared__ float fine[343];
    const uint32_t leaf = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    uint32_t bx, by, bz, lw, lh, ld;
    leaf_coords(leaf, a.W, a.H, a.D, bx, by, bz, lw, lh, ld);
    const uint32_t lx = tid & 7u, ly = (tid >> 3) & 7u, lz = tid >> 6;
    const bool inside = lx < lw && ly < lh && lz < ld;
    const uint64_t dst = ((uint64_t)(bz * 8u + lz) * a.H + (by * 8u + ly)) * a.W + bx * 8u + lx;
    const uint32_t c = a.compact[leaf];
    if (c == NO_LEAF) {
        if (inside) out[dst] = 0.0f;
        return;
    }
    const DevLeaf& d = a.leaves[c];
    if ((d.status & STATUS_MASK) != LEAF_OK) {
        if (inside) out[dst] = __builtin_nanf("");
        if (tid == 0) atomicOr(errFlag, 1u);
        return;
    }
    const Bytes b{a.payload + (size_t)d.word * 4u};
    const int F = (int)a.F;
    const uint32_t cc = control_count(d.plan.coarseRes);
    const uint32_t fc = d.plan.hasFine ? control_count(d.plan.fineRes) : 0u;
    for (uint32_t i = tid; i < cc; i += blockDim.x) coarse[i] = decode_control(b, d.plan.coarse, i, F, t);
    for (uint32_t i = tid; i < fc; i += blockDim.x) fine[i] = decode_control(b, d.plan.fine, i, F, t);
    __syncthreads();
    if (!inside) return;
    float v = sample_grid(coarse, (int)d.plan.coarseRes, (int)lw, (int)lh, (int)ld, (int)lx, (int)ly, (int)lz);
    const uint32_t vox = (lz * 8u + ly) * 8u + lx;
    bool occupied = false;
    if (fc > 0u) {
        v += sample_grid(fine, (int)d.plan.fineRes, (int)lw, (int)lh, (int)ld, (int)lx, (int)ly, (int)lz);
        if (d.plan.mode == MODE_SHELL) occupied = shell_desc_index(b, d.plan.occOffset, vox) != NO_LEAF;
    } else if (d.plan.mode == MODE_SHELL) {
        v += shell_value(b, d, vox, F, t, occupied);
    }
    if (d.plan.mode == MODE_SHELL) v = apply_sign(cm, v, occupied);
    out[dst] = v;
}

// ----------------------------------------------------------- point queries --
template <int LANES>
__global__ void __launch_bounds__(256) k_query(AssetDev a, const uint4* q, uint32_t n, ComposeDev cm, float* out,
                                               int32_t* status)
{
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t qi = gid / LANES;
    const uint32_t sub = (uint32_t)(gid % LANES);
    // groups past the end still take part in the lane shuffles, with a clamped query
    const bool valid = qi < n;
    const uint4 r = q[valid ? qi : (n - 1u)];
    int32_t st;
    const float v = point_query<LANES>(a, cm, r, sub, st);
    if (valid && sub == 0u) {
        out[qi] = st ? __builtin_nanf("") : v;
        if (status) status[qi] = st;
    }
}

// ------------------------------------------- per-value decode (batched) --
__device__ inline int32_t stream_checks(uint32_t count, uint32_t bins, uint32_t keys, uint32_t codec, uint32_t bytes,
                                        int align, uint32_t ci, uint32_t& keyBase)
{
    if (ci >= count) return VBT_ERR_CONTROL_INDEX;
    const uint32_t db = count * 4u;
    uint32_t bb = count * bins, kb = keys * keyframe_bytes(codec);
    if (align) {
        bb = align_up4(bb);
        kb = align_up4(kb);
    }
    if (db + bb + kb > bytes) return LEAF_ERR_STREAM_OOB;
    keyBase = db + bb;
    return 0;
}

__device__ inline float value_of_stream(const uint8_t* base, uint32_t codec, uint32_t count, uint32_t bins,
                                        uint32_t keys, uint32_t bytes, int align, int F, uint32_t ci, int t,
                                        int32_t& st)
{
    uint32_t keyBase = 0;
    st = stream_checks(count, bins, keys, codec, bytes, align, ci, keyBase);
    if (st) return __builtin_nanf("");
    const Bytes b{base};
    const uint32_t ks = b.u16(ci * 4u), kc = b.u8(ci * 4u + 2u);
    if (kc == 0u) return 0.0f;
    if (ks + kc > keys) {
        st = LEAF_ERR_KEY_RANGE;
        return __builtin_nanf("");
    }
    return decode_sequence(b, keyBase, codec, ks, kc, count * 4u + ci * bins, bins, F, t);
}

__global__ void k_control_values(const uint8_t* stream, uint32_t codec, uint32_t count, uint32_t bins, uint32_t keys,
                                 uint32_t bytes, int align, int F, const uint16_t* ci, const int32_t* t, uint32_t n,
                                 float* out, int32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t st;
    out[i] = value_of_stream(stream, codec, count, bins, keys, bytes, align, F, ci[i], t[i], st);
    if (status) status[i] = st;
}

// decodeRenderTemporalLeafControlValue after parseRenderTemporalPackedLeaf:
// the stream is chosen from the parse layout (EMPTY leaves have a zero layout)
__global__ void k_leaf_control_values(const uint8_t* leaf, uint32_t leafBytes, int useFine, int F, const uint16_t* ci,
                                      const int32_t* t, uint32_t n, float* out, int32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Bytes b{leaf};
    LeafPlan pl;
    int32_t st = (int32_t)plan_leaf(b, leafBytes, pl);
    // parse-level failures only (descriptor ranges are checked per control below)
    if (st == LEAF_ERR_STREAM_OOB) st = 0;
    float v = __builtin_nanf("");
    if (st == 0) {
        const Header h = unpack_header(b.u32(0));
        const uint32_t cc = control_count(h.coarseRes), fc = control_count(h.fineRes);
        uint32_t off = 0, bytes = 0, keys = 0, count = 0, codec = 0;
        if (h.mode != MODE_EMPTY) {
            const uint32_t ckc = b.u16(4), fkc = b.u16(6);
            const uint32_t coarseBytes = stream_total_bytes(cc, 8u, ckc, h.coarseCodec);
            if (useFine) {
                off = 8u + coarseBytes;
                count = fc;
                keys = fkc;
                codec = h.fineCodec;
                bytes = h.mode == MODE_FINE6   ? stream_total_bytes(fc, 8u, fkc, h.fineCodec)
                        : h.mode == MODE_SHELL ? 80u + align_up4(fkc * keyframe_bytes(h.fineCodec))
                                               : 0u;
            } else {
                off = 8u;
                count = cc;
                keys = ckc;
                codec = h.coarseCodec;
                bytes = coarseBytes;
            }
        } else {
            count = useFine ? fc : cc;
            codec = useFine ? h.fineCodec : h.coarseCodec;
        }
        v = value_of_stream(leaf + off, codec, count, 8u, keys, bytes, 1, F, ci[i], t[i], st);
    }
    out[i] = v;
    if (status) status[i] = st;
}

__global__ void k_shell_values(const uint8_t* leaf, uint32_t leafBytes, int F, const uint16_t* vox, const int32_t* t,
                               uint32_t n, float* out, int32_t* status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Bytes b{leaf};
    LeafPlan pl;
    int32_t st = (int32_t)plan_leaf(b, leafBytes, pl);
    if (st == LEAF_ERR_STREAM_OOB) st = 0;
    float v = 0.0f;
    if (st == 0 && pl.mode == MODE_SHELL && vox[i] < 512u) {
        const uint32_t di = shell_desc_index(b, pl.occOffset, vox[i]);
        if (di != NO_LEAF && pl.shell.count != 0u && pl.shell.keys != 0u)
            v = value_of_stream(leaf + pl.shell.base, pl.shell.codec, pl.shell.count, 4u, pl.shell.keys, pl.shell.bytes,
                                1, F, di, t[i], st);
    } else if (st != 0) {
        v = __builtin_nanf("");
    }
    out[i] = v;
    if (status) status[i] = st;
}

// Plain device copy (16 B per lane per step): the reference point that scales
// the HBM byte counters of the decode kernels.
__global__ void __launch_bounds__(256) k_copy(const uint4* src, uint4* dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_fill16(uint4* dst, uint4 v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) dst[i] = v;
}

// Every lerp weight the sweep can form: block g-1 checks each weight a / g,
// 0 <= a < g, of gap g (formed from the sweep's reciprocal of g) against
// RN32(a / g).  The expected values are double quotients rounded to float:
// a quotient a / g with g < 2^16 that is not itself a float lies at least
// 2^-41 (relative) from every float rounding boundary (a boundary is a 25-bit
// odd multiple of a power of two, and |a / g - m 2^e| >= 2^e / g), far beyond
// the double rounding error, so rounding twice gives the correctly rounded
// float quotient.
__global__ void __launch_bounds__(256) k_selftest_weights(uint32_t maxGap, unsigned long long* mismatches)
{
    const uint32_t g = blockIdx.x + 1u;
    if (g > maxGap) return;
    const float b = (float)g;
    const float y = sweep_recip(b);
    uint32_t bad = 0u;
    for (uint32_t a = threadIdx.x; a < g; a += 256u) {
        const float got = sweep_alpha((float)a, b, y);
        const float exp = (float)((double)a / (double)g);
        bad += __float_as_uint(got) != __float_as_uint(exp) ? 1u : 0u;
    }
    if (bad) atomicAdd(mismatches, (unsigned long long)bad);
}

__global__ void k_fill(float* p, uint64_t n, float v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}


} // namespace

// Cuts the packed cache into tiles of at most maxGroups CACHE_GROUP-slot
// groups whose swept records span at most `window` bytes (greedy, in slot
// order; a group that alone exceeds the window is a tile of its own and takes
// the multi-pass staging).  Appends the tiles' first groups and one past the
// last group to `out`; returns the tile count.
static uint32_t plan_tiles(const std::vector<uint2>& gr, uint32_t maxGroups, uint32_t window, std::vector<uint32_t>& out)
{
    const uint32_t ng = (uint32_t)gr.size();
    uint32_t tiles = 0, i = 0;
    while (i < ng) {
        out.push_back(i);
        ++tiles;
        uint32_t lo = gr[i].x, hi = gr[i].y, j = i + 1;
        while (j < ng && j - i < maxGroups) {
            const uint32_t nlo = std::min(lo, gr[j].x), nhi = std::max(hi, gr[j].y);
            if (nlo != 0xFFFFFFFFu && nhi - (nlo & ~3u) > window) break;
            lo = nlo;
            hi = nhi;
            ++j;
        }
        i = j;
    }
    out.push_back(ng);
    return tiles;
}

// ============================================================== assets ==
struct vbt_asset_s {
    vbt_file_header h{};
    int device = 0;
    uint32_t* d_offsets = nullptr;
    uint32_t* d_payload = nullptr;
    float* d_scales = nullptr;
    uint32_t* d_compact = nullptr;
    uint32_t* d_status = nullptr;
    DevLeaf* d_leaves = nullptr;
    uint64_t* d_ctrl = nullptr;
    float* d_basis = nullptr;
    uint32_t* d_err = nullptr;   // error flag word for launches
    uint32_t* d_blockLeaf = nullptr; // first active leaf of each 64-slot group of the packed cache
    uint2* d_slotPlan = nullptr;     // render: slot plan of the packed cache (k_slot_plan)
    uint2* d_groupRange = nullptr;   // render: swept-record byte range of each 64-slot group
    // render: the packed cache cut into tiles per tile shape (plan_tiles):
    // shape i's tile j is groups [d_tiles[tileOff[i] + j], d_tiles[tileOff[i] + j + 1])
    uint32_t* d_tiles = nullptr;
    std::vector<uint32_t> hTiles; // host copy of d_tiles
    uint32_t tileOff[CACHE_TILE_COUNT] = {}, tileCount[CACHE_TILE_COUNT] = {};
    int cacheTile = CACHE_TILE_DEFAULT;
    Pair* d_blocks = nullptr;
    VbtFrameIndex index; // the asset's own frame index (vbt_asset_build_frame_index)
    // one event per stream that launched a build reading the index, recorded
    // after the launch: dropping the index waits for them only
    std::vector<std::pair<hipStream_t, hipEvent_t>> indexUse;
    uint32_t A = 0, bad = 0;
    uint64_t controlsPerFrame = 0, payloadBytes = 0, deviceBytes = 0, fast = 0;
    uint64_t modeCounts[4] = {0, 0, 0, 0};
    uint32_t sciKeepMax = 1;        // scientific: largest DCT term count of a leaf
    uint32_t flags = 0;             // ASSET_* (the VBT_OPEN_* flags it was opened with)
    std::string badMessage; // first poisoned leaf and its reference exception text

    AssetDev dev() const
    {
        AssetDev a;
        a.offsets = d_offsets;
        a.payload = (const uint8_t*)d_payload;
        a.scales = d_scales;
        a.compact = d_compact;
        a.leaves = d_leaves;
        a.ctrl = d_ctrl;
        a.basis = d_basis;
        a.W = h.width;
        a.H = h.height;
        a.D = h.depth;
        a.F = h.frames;
        a.L = h.leafCount;
        a.A = A;
        a.profile = h.profileType;
        a.scaleCount = h.coarseAcScaleCount;
        a.payloadBytes = payloadBytes;
        a.groupLeaf = d_blockLeaf;
        a.slotPlan = d_slotPlan;
        a.groupRange = d_groupRange;
        a.controlsPerFrame = controlsPerFrame;
        a.flags = flags;
        return a;
    }
};

namespace {

template <class T> int dalloc(vbt_asset a, T** p, size_t count)
{
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 4;
    HIPCHK(hipMalloc((void**)p, bytes));
    a->deviceBytes += bytes;
    return VBT_OK;
}

void free_asset(vbt_asset a)
{
    if (!a) return;
    vbt_asset_drop_frame_index(a);
    void* ptrs[] = {a->d_offsets, a->d_payload, a->d_scales, a->d_compact, a->d_status,
                    a->d_leaves,  a->d_ctrl,    a->d_basis,  a->d_err,     a->d_blocks, a->d_blockLeaf,
                    a->d_slotPlan, a->d_groupRange, a->d_tiles};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete a;
}

// Whole-frame requests decode every leaf, so an asset with poisoned leaves
// fails them (as the reference's decode throws); the valid leaves' outputs are
// still written and the poisoned ones are NaN or absent.
int report_bad(vbt_asset a, const char* what)
{
    return vbt_fail(VBT_ERR_FORMAT, std::string(what) + ": " + std::to_string(a->bad) +
                                        " leaves failed validation; " + a->badMessage);
}

ComposeDev to_dev(const vbt_compose* c)
{
    ComposeDev d{VBT_SIGN_NONE, 0.0f, 1.0f, NAN};
    if (c) {
        d.signRule = c->signRule;
        d.background = c->background;
        d.shellEmptyScale = c->shellEmptyScale;
        d.shellEmptyMax = c->shellEmptyMax;
    }
    return d;
}

} // namespace

namespace {

void launch_indexed(vbt_asset a, const VbtFrameIndex& ix, int t0, int nt, float* d_out, hipStream_t s)
{
    const AssetDev ad = a->dev();
    if (ix.wide)
        k_cache_indexed<true><<<blocks_for(a->controlsPerFrame, 256), 256, 0, s>>>(ad, ix.off, ix.rec, ix.bins, t0, nt, d_out,
                                                                                  a->controlsPerFrame, a->d_err);
    else
        k_cache_indexed<false><<<blocks_for(a->controlsPerFrame, 256), 256, 0, s>>>(ad, ix.off, ix.rec, ix.bins, t0, nt, d_out,
                                                                                   a->controlsPerFrame, a->d_err);
}

// Records that a build on stream s reads the asset's own index.
int note_index_use(vbt_asset a, hipStream_t s)
{
    for (auto& u : a->indexUse)
        if (u.first == s) {
            HIPCHK(hipEventRecord(u.second, s));
            return VBT_OK;
        }
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    a->indexUse.push_back({s, ev});
    HIPCHK(hipEventRecord(ev, s));
    return VBT_OK;
}

} // namespace

int vbt_internal_frame_cache_indexed(vbt_asset a, const VbtFrameIndex& ix, int t0, int nt, float* d_out, void* stream)
{
    if (!a || !d_out || nt <= 0 || nt > CACHE_EXACT_MAX_FRAMES || !ix.off)
        return vbt_fail(VBT_ERR_INVALID, "invalid indexed frame cache request");
    if (t0 < 0 || (uint64_t)t0 + (uint64_t)nt > a->h.frames) return vbt_fail(VBT_ERR_INVALID, "frame range outside the asset");
    if (a->A == 0) return a->bad ? report_bad(a, "frame cache") : VBT_OK;
    launch_indexed(a, ix, t0, nt, d_out, S(stream));
    HIPCHK(hipGetLastError());
    if (a->bad) return report_bad(a, "frame cache");
    return VBT_OK;
}

void vbt_internal_frame_index_free(VbtFrameIndex& ix)
{
    if (ix.off) (void)hipFree(ix.off);
    if (ix.rec) (void)hipFree(ix.rec);
    ix = VbtFrameIndex{};
}

int vbt_internal_frame_index_build(vbt_asset a, VbtFrameIndex& ix, void* stream)
{
    if (!a) return vbt_fail(VBT_ERR_INVALID, "frame index: null asset");
    if (a->h.profileType == 2) return vbt_fail(VBT_ERR_INVALID, "frame index: render assets only");
    if (ix.off || ix.rec) return vbt_fail(VBT_ERR_INVALID, "frame index: already built");
    const uint32_t B = index_bins(a->h.frames);
    const uint64_t C = a->controlsPerFrame, N = (uint64_t)B * (C + 1u);
    if (a->A == 0 || C == 0) return VBT_OK;
    if (N > 0x7FFFFFFFull) return vbt_fail(VBT_ERR_INVALID, "frame index: too many controls");
    HIPCHK(hipSetDevice(a->device));
    hipStream_t s = S(stream);
    const AssetDev ad = a->dev();
    uint32_t* cnt = nullptr;
    void* tmp = nullptr;
    size_t tmpBytes = 0;
    uint32_t last = 0;
    auto fail = [&](int code, const char* msg) {
        (void)hipStreamSynchronize(s);
        if (cnt) (void)hipFree(cnt);
        if (tmp) (void)hipFree(tmp);
        vbt_internal_frame_index_free(ix);
        return vbt_fail(code, msg);
    };
    if (hipMalloc((void**)&cnt, 4ull * N) != hipSuccess || hipMalloc((void**)&ix.off, 4ull * N) != hipSuccess)
        return fail(VBT_ERR_NOMEM, "frame index: out of device memory");
    // d_err[1]: the fp32 flag, d_err[2..3]: the 64-bit record total
    if (hipMemsetAsync(a->d_err + 1, 0, 12, s) != hipSuccess) return fail(VBT_ERR_DEVICE, "frame index: clear failed");
    k_index_count<<<blocks_for(C + 1u, 256), 256, 0, s>>>(ad, cnt, C, B, a->d_err + 1);
    // the record total must fit the 32-bit offsets (each prefix is at most it)
    unsigned long long sum = 0;
    uint32_t wide = 0;
    k_sum64<<<1024, 256, 0, s>>>(cnt, N, (unsigned long long*)(a->d_err + 2));
    if (hipMemcpyAsync(&sum, a->d_err + 2, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&wide, a->d_err + 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(VBT_ERR_DEVICE, "frame index: count failed");
    if (sum > 0xFFFFFFFFull) return fail(VBT_ERR_INVALID, "frame index: more than 2^32 records");
    if (sum == 0) { // nothing indexable: builds keep the bracket search
        (void)hipFree(cnt);
        vbt_internal_frame_index_free(ix);
        return VBT_OK;
    }
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmpBytes, cnt, ix.off, (int)N, s) != hipSuccess ||
        hipMalloc(&tmp, tmpBytes ? tmpBytes : 4) != hipSuccess ||
        hipcub::DeviceScan::ExclusiveSum(tmp, tmpBytes, cnt, ix.off, (int)N, s) != hipSuccess ||
        hipMemcpyAsync(&last, ix.off + (N - 1u), 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(VBT_ERR_DEVICE, "frame index: scan failed");
    // the slot count after each bin's last slot is 0, so `last` is the total
    const uint64_t total = last;
    const uint64_t recBytes = wide ? 8u : 4u;
    if (hipMalloc(&ix.rec, total * recBytes) != hipSuccess) return fail(VBT_ERR_NOMEM, "frame index: out of device memory");
    if (wide) k_index_scatter<true><<<blocks_for(C, 256), 256, 0, s>>>(ad, ix.off, ix.rec, C, B);
    else k_index_scatter<false><<<blocks_for(C, 256), 256, 0, s>>>(ad, ix.off, ix.rec, C, B);
    const hipError_t e = hipGetLastError();
    const hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(cnt);
    (void)hipFree(tmp);
    cnt = nullptr;
    tmp = nullptr;
    if (e != hipSuccess || e2 != hipSuccess) return fail(VBT_ERR_DEVICE, hipGetErrorString(e != hipSuccess ? e : e2));
    ix.records = total;
    ix.bins = B;
    ix.wide = wide != 0;
    ix.bytes = 4ull * N + recBytes * total;
    return VBT_OK;
}

// accessors for the serving layer (server.hip)
vbtk::AssetDev vbt_internal_asset_dev(vbt_asset a) { return a->dev(); }
uint64_t vbt_internal_asset_bytes(vbt_asset a) { return a->deviceBytes; }
int vbt_internal_asset_device(vbt_asset a) { return a->device; }
uint32_t vbt_internal_asset_bad(vbt_asset a) { return a->bad; }
uint32_t* vbt_internal_asset_err(vbt_asset a) { return a->d_err; }
const uint32_t* vbt_internal_asset_tiles(vbt_asset a, int shape, uint32_t* count, const uint32_t** host)
{
    if (!a->d_tiles || shape < 0 || shape >= CACHE_TILE_COUNT) {
        *count = 0;
        *host = nullptr;
        return nullptr;
    }
    *count = a->tileCount[shape];
    *host = a->hTiles.data() + a->tileOff[shape];
    return a->d_tiles + a->tileOff[shape];
}
int vbt_internal_report_bad(vbt_asset a, const char* what) { return report_bad(a, what); }
const vbt_file_header* vbt_internal_asset_header(vbt_asset a) { return &a->h; }

extern "C" {

const char* vbt_last_error(void) { return g_err.c_str(); }

const char* vbt_version(void) { return "vbt_hip 1.0 (gfx950)"; }

int vbt_device_count(int* n)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return VBT_OK;
}

int vbt_asset_open_memory(const void* bytes, uint64_t size, void* stream, vbt_asset* out)
{
    return vbt_asset_open_memory_ex(bytes, size, stream, 0u, out);
}

int vbt_asset_open_memory_ex(const void* bytes, uint64_t size, void* stream, uint32_t flags, vbt_asset* out)
{
    *out = nullptr;
    if (flags & ~(uint32_t)VBT_OPEN_SCI_DENSE_BASIS) return vbt_fail(VBT_ERR_INVALID, "unknown asset open flags");
    vbt_file_header h;
    int rc = vbt_validate_file(bytes, size, &h);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return vbt_fail(VBT_ERR_NO_DEVICE, "no HIP device");
    hipStream_t s = S(stream);
    vbt_asset a = new vbt_asset_s();
    a->h = h;
    a->flags = (flags & VBT_OPEN_SCI_DENSE_BASIS) ? ASSET_SCI_DENSE_BASIS : 0u;
    HIPCHK(hipGetDevice(&a->device));
    const uint8_t* base = (const uint8_t*)bytes;
    const uint64_t scaleOff = 52, offOff = 52 + 4ull * h.coarseAcScaleCount;
    const uint64_t payOff = offOff + 4ull * (h.leafCount + 1ull);
    a->payloadBytes = size - payOff;
    const uint32_t L = h.leafCount;
    const uint32_t nb = blocks_for(L, 256);
    if ((rc = dalloc(a, &a->d_offsets, (size_t)L + 1)) || (rc = dalloc(a, &a->d_payload, (a->payloadBytes + 3) / 4 + 1)) ||
        (rc = dalloc(a, &a->d_scales, h.coarseAcScaleCount + 1)) || (rc = dalloc(a, &a->d_compact, L)) ||
        (rc = dalloc(a, &a->d_status, L)) || (rc = dalloc(a, &a->d_err, 4)) || (rc = dalloc(a, &a->d_blocks, nb + 1))) {
        free_asset(a);
        return rc;
    }
    auto fail = [&](int code) {
        free_asset(a);
        return code;
    };
#define TRY(x)                                                                                          \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            vbt_fail(e_ == hipErrorOutOfMemory ? VBT_ERR_NOMEM : VBT_ERR_DEVICE,                        \
                     std::string(#x " failed: ") + hipGetErrorString(e_));                              \
            return fail(e_ == hipErrorOutOfMemory ? VBT_ERR_NOMEM : VBT_ERR_DEVICE);                    \
        }                                                                                               \
    } while (0)
    TRY(hipMemcpyAsync(a->d_offsets, base + offOff, 4ull * (L + 1ull), hipMemcpyHostToDevice, s));
    if (a->payloadBytes) TRY(hipMemcpyAsync(a->d_payload, base + payOff, a->payloadBytes, hipMemcpyHostToDevice, s));
    if (h.coarseAcScaleCount)
        TRY(hipMemcpyAsync(a->d_scales, base + scaleOff, 4ull * h.coarseAcScaleCount, hipMemcpyHostToDevice, s));
    if (h.profileType == 2) {
        // dctBasis table (render/src/scientific_decode.cpp:72-80): computed in double, stored as float
        std::vector<float> basis((size_t)h.frames * 64);
        for (uint32_t t = 0; t < h.frames; ++t)
            for (int k = 0; k < 64; ++k) {
                const double invN = 1.0 / (double)h.frames;
                const double alpha = (k == 0) ? std::sqrt(invN) : std::sqrt(2.0 * invN);
                basis[(size_t)t * 64 + k] = (float)(alpha * std::cos((3.14159265358979323846 / (double)h.frames) *
                                                                     ((double)t + 0.5) * (double)k));
            }
        if ((rc = dalloc(a, &a->d_basis, basis.size()))) return fail(rc);
        TRY(hipMemcpyAsync(a->d_basis, basis.data(), basis.size() * 4, hipMemcpyHostToDevice, s));
        TRY(hipStreamSynchronize(s));
    }
    AssetDev ad = a->dev();
    Pair* d_total = a->d_blocks + nb;
    k_scan_a<<<nb, 256, 0, s>>>(ad, a->d_blocks);
    k_scan_b<<<1, 1024, 0, s>>>(a->d_blocks, nb, d_total);
    TRY(hipGetLastError());
    Pair total;
    TRY(hipMemcpyAsync(&total, d_total, sizeof(Pair), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    a->A = total.active;
    a->controlsPerFrame = total.n;
    if ((rc = dalloc(a, &a->d_leaves, a->A + 1)) || (rc = dalloc(a, &a->d_ctrl, (size_t)a->A + 1))) return fail(rc);
    ad = a->dev();
    k_scan_c<<<nb, 256, 0, s>>>(ad, a->d_blocks, a->d_compact, a->d_status, a->d_leaves, a->d_ctrl);
    TRY(hipGetLastError());
    TRY(hipMemcpyAsync(a->d_ctrl + a->A, &a->controlsPerFrame, 8, hipMemcpyHostToDevice, s));
    if (a->A) k_validate<<<a->A, 256, 0, s>>>(ad, a->d_leaves, a->d_status);
    TRY(hipGetLastError());
    {
        const uint32_t ng = blocks_for(a->controlsPerFrame, CACHE_GROUP);
        if ((rc = dalloc(a, &a->d_blockLeaf, (size_t)ng + 1))) return fail(rc);
        k_group_leaf<<<blocks_for(ng + 1ull, 256), 256, 0, s>>>(a->d_ctrl, a->A, a->controlsPerFrame, ng,
                                                                 a->d_blockLeaf);
        TRY(hipGetLastError());
        if (h.profileType != 2 && a->controlsPerFrame) {
            if ((rc = dalloc(a, &a->d_slotPlan, (size_t)a->controlsPerFrame)) ||
                (rc = dalloc(a, &a->d_groupRange, (size_t)ng)))
                return fail(rc);
            ad = a->dev();
            k_slot_plan<<<blocks_for(a->controlsPerFrame, 256), 256, 0, s>>>(ad, a->controlsPerFrame, a->d_slotPlan,
                                                                            a->d_groupRange);
            TRY(hipGetLastError());
            std::vector<uint2> gr(ng);
            TRY(hipMemcpyAsync(gr.data(), a->d_groupRange, 8ull * ng, hipMemcpyDeviceToHost, s));
            TRY(hipStreamSynchronize(s));
            std::vector<uint32_t> all;
            for (int i = 0; i < CACHE_TILE_COUNT; ++i) {
                a->tileOff[i] = (uint32_t)all.size();
                a->tileCount[i] = plan_tiles(gr, CACHE_TILES[i].block / CACHE_GROUP, CACHE_TILES[i].window, all);
            }
            if ((rc = dalloc(a, &a->d_tiles, all.size()))) return fail(rc);
            TRY(hipMemcpyAsync(a->d_tiles, all.data(), 4ull * all.size(), hipMemcpyHostToDevice, s));
            TRY(hipStreamSynchronize(s));
            a->hTiles = std::move(all);
        }
    }
    // host-side statistics of the page table
    std::vector<uint32_t> st(L);
    TRY(hipMemcpyAsync(st.data(), a->d_status, 4ull * L, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    const uint32_t* offs = (const uint32_t*)(base + offOff);
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t w = offs[l];
        uint32_t hw;
        std::memcpy(&hw, base + payOff + 4ull * w, 4);
        a->modeCounts[hw & 3u]++;
        if (a->h.profileType == 2) a->sciKeepMax = std::max(a->sciKeepMax, ((hw >> 2) & 0x3Fu) + 1u);
        if ((st[l] & STATUS_MASK) != LEAF_OK) {
            if (a->bad == 0)
                a->badMessage = "first: leaf " + std::to_string(l) + ": " + vbt_status_message((int32_t)(st[l] & STATUS_MASK));
            a->bad++;
        }
        if (st[l] & FLAG_SWEEP) a->fast++;
    }
#undef TRY
    *out = a;
    return VBT_OK;
}

int vbt_asset_open_file(const char* path, void* stream, vbt_asset* out)
{
    return vbt_asset_open_file_ex(path, stream, 0u, out);
}

int vbt_asset_open_file_ex(const char* path, void* stream, uint32_t flags, vbt_asset* out)
{
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return vbt_fail(VBT_ERR_INVALID, std::string("Failed to open file: ") + path);
    const std::streamsize n = in.tellg();
    std::vector<uint32_t> buf(((size_t)n + 3) / 4);
    in.seekg(0);
    in.read((char*)buf.data(), n);
    if (!in) return vbt_fail(VBT_ERR_INVALID, std::string("Failed to read file: ") + path);
    return vbt_asset_open_memory_ex(buf.data(), (uint64_t)n, stream, flags, out);
}

void vbt_asset_close(vbt_asset a) { free_asset(a); }

int vbt_asset_info_get(vbt_asset a, vbt_asset_info* o)
{
    if (!a || !o) return vbt_fail(VBT_ERR_INVALID, "null asset");
    std::memset(o, 0, sizeof(*o));
    o->header = a->h;
    o->activeLeaves = a->A;
    o->badLeaves = a->bad;
    o->controlsPerFrame = a->controlsPerFrame;
    o->payloadBytes = a->payloadBytes;
    o->deviceBytes = a->deviceBytes;
    for (int i = 0; i < 4; ++i) o->mode_counts[i] = a->modeCounts[i];
    o->fast_leaves = a->fast;
    return VBT_OK;
}

int vbt_asset_page_table(vbt_asset a, uint32_t* d_compact, void* stream)
{
    HIPCHK(hipMemcpyAsync(d_compact, a->d_compact, 4ull * a->h.leafCount, hipMemcpyDeviceToDevice, S(stream)));
    return VBT_OK;
}

int vbt_asset_cache_offsets(vbt_asset a, uint64_t* d_offsets, void* stream)
{
    HIPCHK(hipMemcpyAsync(d_offsets, a->d_ctrl, 8ull * (a->A + 1ull), hipMemcpyDeviceToDevice, S(stream)));
    return VBT_OK;
}

int vbt_asset_leaf_status(vbt_asset a, uint32_t* d_status, void* stream)
{
    HIPCHK(hipMemcpyAsync(d_status, a->d_status, 4ull * a->h.leafCount, hipMemcpyDeviceToDevice, S(stream)));
    return VBT_OK;
}

static int frame_cache(vbt_asset a, int t0, int nt, int layout, float* d_out, void* stream, bool useIndex);

int vbt_decode_frame_cache(vbt_asset a, int t0, int nt, int layout, float* d_out, void* stream)
{
    return frame_cache(a, t0, nt, layout, d_out, stream, true);
}

static int frame_cache(vbt_asset a, int t0, int nt, int layout, float* d_out, void* stream, bool useIndex)
{
    if (!a || !d_out || nt <= 0) return vbt_fail(VBT_ERR_INVALID, "invalid frame cache request");
    if (t0 < 0 || (uint64_t)t0 + (uint64_t)nt > a->h.frames)
        return vbt_fail(VBT_ERR_INVALID, "frame range outside the asset");
    if (layout != VBT_CACHE_PACKED && layout != VBT_CACHE_STRIDE576)
        return vbt_fail(VBT_ERR_INVALID, "unknown cache layout");
    if (a->h.profileType == 2 && layout != VBT_CACHE_PACKED)
        return vbt_fail(VBT_ERR_INVALID, "scientific assets support the packed cache only");
    if (4ull * (layout == VBT_CACHE_PACKED ? a->controlsPerFrame : (uint64_t)a->A * 576u) > 0xFFFFFFFFull)
        return vbt_fail(VBT_ERR_INVALID, "frame cache: one frame exceeds 4 GiB");
    hipStream_t s = S(stream);
    const AssetDev ad = a->dev();
    if (a->A == 0) return a->bad ? report_bad(a, "frame cache") : VBT_OK;
    if (a->h.profileType == 2) {
        launch_sci_frames(a->A, a->sciKeepMax, a->modeCounts[2] + a->modeCounts[3] > 0, ad, t0, nt, d_out,
                          a->controlsPerFrame, a->d_err, s);
    } else if (layout == VBT_CACHE_PACKED) {
        if (nt <= CACHE_EXACT_MAX_FRAMES && useIndex && a->index.off) {
            launch_indexed(a, a->index, t0, nt, d_out, s);
            HIPCHK(hipGetLastError());
            int rc = note_index_use(a, s);
            if (rc) return rc;
        } else if (nt <= CACHE_EXACT_MAX_FRAMES)
            k_cache_exact<VBT_CACHE_PACKED><<<blocks_for(a->controlsPerFrame, 256), 256, 0, s>>>(
                ad, t0, nt, d_out, a->controlsPerFrame, a->d_err);
        else
            launch_cache_render<VBT_CACHE_PACKED>(a->cacheTile, ad, a->d_tiles + a->tileOff[a->cacheTile],
                                                  a->tileCount[a->cacheTile], t0, nt, d_out, a->controlsPerFrame,
                                                  a->d_err, s);
    } else {
        const uint64_t n = (uint64_t)a->A * 576u * (uint64_t)nt;
        k_fill<<<(uint32_t)std::min<uint64_t>(blocks_for(n, 256), 65535), 256, 0, s>>>(d_out, n, 0.0f);
        if (nt <= CACHE_EXACT_MAX_FRAMES)
            k_cache_exact<VBT_CACHE_STRIDE576><<<blocks_for(a->controlsPerFrame, 256), 256, 0, s>>>(
                ad, t0, nt, d_out, a->controlsPerFrame, a->d_err);
        else
            launch_cache_render<VBT_CACHE_STRIDE576>(a->cacheTile, ad, nullptr, 0u, t0, nt, d_out,
                                                     a->controlsPerFrame, a->d_err, s);
    }
    HIPCHK(hipGetLastError());
    if (a->bad) return report_bad(a, "frame cache");
    return VBT_OK;
}

int vbt_asset_cache_ceiling(vbt_asset a, int nt, int tile, int reads, float* d_out, void* stream)
{
    if (!a || !d_out || nt <= 0 || (uint64_t)nt > a->h.frames)
        return vbt_fail(VBT_ERR_INVALID, "invalid frame-cache ceiling request");
    if (tile < 0 || tile >= CACHE_TILE_COUNT) return vbt_fail(VBT_ERR_INVALID, "unknown tile shape");
    if (a->h.profileType == 2 || !a->d_tiles) return vbt_fail(VBT_ERR_INVALID, "the ceiling covers render assets");
    if (4ull * a->controlsPerFrame > 0xFFFFFFFFull) return vbt_fail(VBT_ERR_INVALID, "frame cache: one frame exceeds 4 GiB");
    if (a->tileCount[tile] == 0) return VBT_OK;
    launch_cache_ceiling(tile, a->dev(), a->d_tiles + a->tileOff[tile], a->tileCount[tile], nt, reads, d_out,
                         a->controlsPerFrame, S(stream));
    HIPCHK(hipGetLastError());
    return VBT_OK;
}

int vbt_decode_frame_dense(vbt_asset a, int t, const vbt_compose* c, float* d_out, void* stream)
{
    if (!a || !d_out) return vbt_fail(VBT_ERR_INVALID, "invalid dense frame request");
    if (a->h.profileType == 2 && (t < 0 || (uint32_t)t >= a->h.frames))
        return vbt_fail(VBT_ERR_INVALID, "scientific frame index out of range");
    hipStream_t s = S(stream);
    const AssetDev ad = a->dev();
    if (a->h.profileType == 2) k_sci_dense<<<a->A, 512, 0, s>>>(ad, t, d_out, a->d_err);
    else k_dense_render<<<a->h.leafCount, 512, 0, s>>>(ad, t, to_dev(c), d_out, a->d_err);
    HIPCHK(hipGetLastError());
    if (a->bad) return report_bad(a, "dense frame");
    return VBT_OK;
}

static int launch_query(vbt_asset a, const uint32_t* d_q, uint32_t n, const vbt_compose* c, float* d_out,
                        int32_t* d_status, void* stream, bool plain)
{
    if (!a || !d_q || !d_out) return vbt_fail(VBT_ERR_INVALID, "invalid query request");
    if (n == 0) return VBT_OK;
    hipStream_t s = S(stream);
    const AssetDev ad = a->dev();
    if (plain) k_query<1><<<blocks_for(n, 256), 256, 0, s>>>(ad, (const uint4*)d_q, n, to_dev(c), d_out, d_status);
    else k_query<8><<<blocks_for(8ull * n, 256), 256, 0, s>>>(ad, (const uint4*)d_q, n, to_dev(c), d_out, d_status);
    HIPCHK(hipGetLastError());
    return VBT_OK;
}

int vbt_query_points(vbt_asset a, const uint32_t* d_q, uint32_t n, const vbt_compose* c, float* d_out,
                     int32_t* d_status, void* stream)
{
    return launch_query(a, d_q, n, c, d_out, d_status, stream, false);
}

int vbt_query_points_plain(vbt_asset a, const uint32_t* d_q, uint32_t n, const vbt_compose* c, float* d_out,
                           int32_t* d_status, void* stream)
{
    return launch_query(a, d_q, n, c, d_out, d_status, stream, true);
}

int vbt_decode_control_values(const uint8_t* d_stream, uint32_t codec, uint16_t controlCount, uint8_t timeBinCount,
                              uint8_t maxBinLocalKeys, uint32_t keyframeCount, uint32_t streamBytes, int align,
                              int totalFrames, const uint16_t* d_ci, const int32_t* d_t, uint32_t n, float* d_out,
                              int32_t* d_status, void* stream)
{
    (void)maxBinLocalKeys;
    if (n == 0) return VBT_OK;
    k_control_values<<<blocks_for(n, 256), 256, 0, S(stream)>>>(d_stream, codec, controlCount, timeBinCount,
                                                                keyframeCount, streamBytes, align, totalFrames, d_ci,
                                                                d_t, n, d_out, d_status);
    HIPCHK(hipGetLastError());
    return VBT_OK;
}

int vbt_decode_leaf_control_values(const uint8_t* d_leaf, uint32_t leafBytes, int useFine, int align, int totalFrames,
                                   const uint16_t* d_ci, const int32_t* d_t, uint32_t n, float* d_out,
                                   int32_t* d_status, void* stream)
{
    if (!align) return vbt_fail(VBT_ERR_INVALID, "the device leaf decoders read 4-byte aligned leaves");
    if (n == 0) return VBT_OK;
    k_leaf_control_values<<<blocks_for(n, 256), 256, 0, S(stream)>>>(d_leaf, leafBytes, useFine, totalFrames, d_ci,
                                                                     d_t, n, d_out, d_status);
    HIPCHK(hipGetLastError());
    return VBT_OK;
}

int vbt_decode_shell_voxel_values(const uint8_t* d_leaf, uint32_t leafBytes, int align, int totalFrames,
                                  const uint16_t* d_vox, const int32_t* d_t, uint32_t n, float* d_out,
                                  int32_t* d_status, void* stream)
{
    if (!align) return vbt_fail(VBT_ERR_INVALID, "the device leaf decoders read 4-byte aligned leaves");
    if (n == 0) return VBT_OK;
    k_shell_values<<<blocks_for(n, 256), 256, 0, S(stream)>>>(d_leaf, leafBytes, totalFrames, d_vox, d_t, n, d_out,
                                                              d_status);
    HIPCHK(hipGetLastError());
    return VBT_OK;
}

// -------------------------------------------------------------- autotune --
int vbt_cache_tile_count(void) { return CACHE_TILE_COUNT; }

int vbt_cache_tile_shape(int tile, uint32_t* block, uint32_t* window)
{
    if (tile < 0 || tile >= CACHE_TILE_COUNT) return vbt_fail(VBT_ERR_INVALID, "tile index out of range");
    *block = CACHE_TILES[tile].block;
    *window = CACHE_TILES[tile].window;
    return VBT_OK;
}

int vbt_asset_drop_frame_index(vbt_asset a)
{
    if (!a) return vbt_fail(VBT_ERR_INVALID, "vbt_asset_drop_frame_index: null asset");
    int rc = VBT_OK;
    if (a->index.off || a->index.rec) {
        // builds that read the index may still run: wait for each stream's last one
        for (auto& u : a->indexUse) {
            const hipError_t e = hipEventSynchronize(u.second);
            if (e != hipSuccess && rc == VBT_OK) rc = vbt_fail(VBT_ERR_DEVICE, "frame index: wait for its readers");
        }
        a->deviceBytes -= a->index.bytes;
        vbt_internal_frame_index_free(a->index);
    }
    for (auto& u : a->indexUse) (void)hipEventDestroy(u.second);
    a->indexUse.clear();
    return rc;
}

int vbt_asset_build_frame_index(vbt_asset a, uint64_t* bytes, void* stream)
{
    if (!a) return vbt_fail(VBT_ERR_INVALID, "vbt_asset_build_frame_index: null asset");
    if (a->h.profileType == 2) return vbt_fail(VBT_ERR_INVALID, "frame index: render assets only");
    int rc = vbt_asset_drop_frame_index(a);
    if (rc) return rc;
    if (bytes) *bytes = 0;
    if ((rc = vbt_internal_frame_index_build(a, a->index, stream))) return rc;
    a->deviceBytes += a->index.bytes; // counted in the asset's device memory (and so in server budgets)
    if (bytes) *bytes = a->index.bytes;
    return VBT_OK;
}

int vbt_asset_set_cache_tile(vbt_asset a, int tile)
{
    if (!a || tile < 0 || tile >= CACHE_TILE_COUNT) return vbt_fail(VBT_ERR_INVALID, "tile index out of range");
    a->cacheTile = tile;
    return VBT_OK;
}

int vbt_asset_get_cache_tile(vbt_asset a, int* tile)
{
    if (!a) return vbt_fail(VBT_ERR_INVALID, "null asset");
    *tile = a->cacheTile;
    return VBT_OK;
}

int vbt_asset_autotune(vbt_asset a, int nt, int reps, float* msPerTile, int* best, void* stream)
{
    if (!a || nt <= 0 || reps <= 0) return vbt_fail(VBT_ERR_INVALID, "invalid autotune request");
    if (a->h.profileType == 2 || a->A == 0) {
        if (best) *best = a->cacheTile;
        return VBT_OK; // scientific caches have a single kernel shape
    }
    if ((uint64_t)nt > a->h.frames) nt = (int)a->h.frames;
    hipStream_t s = S(stream);
    float* scratch = nullptr;
    HIPCHK(hipMalloc((void**)&scratch, 4ull * a->controlsPerFrame * (uint64_t)nt));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const AssetDev ad = a->dev();
    int bestTile = a->cacheTile;
    float bestMs = 1e30f;
    for (int tile = 0; tile < CACHE_TILE_COUNT; ++tile) {
        launch_cache_render<VBT_CACHE_PACKED>(tile, ad, a->d_tiles + a->tileOff[tile], a->tileCount[tile], 0, nt,
                                              scratch, a->controlsPerFrame, a->d_err, s);
        HIPCHK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r)
            launch_cache_render<VBT_CACHE_PACKED>(tile, ad, a->d_tiles + a->tileOff[tile], a->tileCount[tile], 0,
                                                  nt, scratch, a->controlsPerFrame, a->d_err, s);
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)reps;
        if (msPerTile) msPerTile[tile] = ms;
        if (ms < bestMs) {
            bestMs = ms;
            bestTile = tile;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    a->cacheTile = bestTile;
    if (best) *best = bestTile;
    return VBT_OK;
}

// ----------------------------------------------------------- device utils --
int vbt_device_alloc(void** p, size_t bytes)
{
    HIPCHK(hipMalloc(p, bytes ? bytes : 4));
    return VBT_OK;
}
int vbt_device_free(void* p)
{
    if (p) HIPCHK(hipFree(p));
    return VBT_OK;
}
int vbt_memcpy_h2d(void* d, const void* s, size_t n, void* stream)
{
    HIPCHK(hipMemcpyAsync(d, s, n, hipMemcpyHostToDevice, S(stream)));
    return VBT_OK;
}
int vbt_memcpy_d2h(void* d, const void* s, size_t n, void* stream)
{
    HIPCHK(hipMemcpyAsync(d, s, n, hipMemcpyDeviceToHost, S(stream)));
    HIPCHK(hipStreamSynchronize(S(stream)));
    return VBT_OK;
}
int vbt_device_copy(void* d_dst, const void* d_src, size_t bytes, void* stream)
{
    if (bytes % 16u || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16u)
        return vbt_fail(VBT_ERR_INVALID, "vbt_device_copy needs 16-byte aligned pointers and sizes");
    const uint64_t n = bytes / 16u;
    if (n) {
        k_copy<<<(uint32_t)std::min<uint64_t>(blocks_for(n, 256), 8192), 256, 0, S(stream)>>>((const uint4*)d_src,
                                                                                            (uint4*)d_dst, n);
        HIPCHK(hipGetLastError());
    }
    return VBT_OK;
}
int vbt_device_fill(void* d_dst, uint32_t word, size_t bytes, void* stream)
{
    if (bytes % 16u || (uintptr_t)d_dst % 16u)
        return vbt_fail(VBT_ERR_INVALID, "vbt_device_fill needs a 16-byte aligned pointer and size");
    const uint64_t n = bytes / 16u;
    if (n) {
        k_fill16<<<(uint32_t)std::min<uint64_t>(blocks_for(n, 256), 16384), 256, 0, S(stream)>>>(
            (uint4*)d_dst, make_uint4(word, word, word, word), n);
        HIPCHK(hipGetLastError());
    }
    return VBT_OK;
}
int vbt_selftest_weights(uint32_t maxGap, uint64_t* mismatches, void* stream)
{
    if (!mismatches || maxGap == 0u || maxGap > 65535u)
        return vbt_fail(VBT_ERR_INVALID, "vbt_selftest_weights: maxGap must be 1..65535");
    unsigned long long* d = nullptr;
    HIPCHK(hipMalloc((void**)&d, sizeof(*d)));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(*d), S(stream));
    if (e == hipSuccess) {
        k_selftest_weights<<<maxGap, 256, 0, S(stream)>>>(maxGap, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));
    (void)hipFree(d);
    HIPCHK(e);
    *mismatches = h;
    return VBT_OK;
}
int vbt_memset_d(void* d, int v, size_t n, void* stream)
{
    HIPCHK(hipMemsetAsync(d, v, n, S(stream)));
    return VBT_OK;
}
int vbt_stream_create(void** s)
{
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = st;
    return VBT_OK;
}
int vbt_stream_destroy(void* s)
{
    if (s) HIPCHK(hipStreamDestroy(S(s)));
    return VBT_OK;
}
int vbt_stream_synchronize(void* s)
{
    HIPCHK(hipStreamSynchronize(S(s)));
    return VBT_OK;
}
int vbt_event_create(void** e)
{
    hipEvent_t ev;
    HIPCHK(hipEventCreate(&ev));
    *e = ev;
    return VBT_OK;
}
int vbt_event_destroy(void* e)
{
    if (e) HIPCHK(hipEventDestroy((hipEvent_t)e));
    return VBT_OK;
}
int vbt_event_record(void* e, void* s)
{
    HIPCHK(hipEventRecord((hipEvent_t)e, S(s)));
    return VBT_OK;
}
int vbt_event_elapsed_ms(void* a, void* b, float* ms)
{
    HIPCHK(hipEventSynchronize((hipEvent_t)b));
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return VBT_OK;
}

} // extern "C"

// the build without the frame index (no event records: batch graphs capture it)
int vbt_internal_frame_cache_plain(vbt_asset a, int t0, int nt, int layout, float* d_out, void* stream)
{
    return frame_cache(a, t0, nt, layout, d_out, stream, false);
}

