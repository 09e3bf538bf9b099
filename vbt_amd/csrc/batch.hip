// batch.hip -- batched frame-cache builds: one persistent kernel per batch
// over a work queue of (job, leaf tile, frame range) items, replayed from a
// hipGraph, with the queue shape picked by a timing autotuner.
//
// A job is one asset's frame range [t0, t0+nt) into one output buffer (the
// layout vbt_decode_frame_cache writes).  The unit of work of a render job is
// a leaf tile: BLOCK consecutive packed-cache slots (any mix of Grid4, Fine6
// and Mode2 leaves, one lane per slot).  Each job's frames are split into
// `ranges` frame ranges; an item is one (tile, range) pair.  The ranges of a
// tile are adjacent in the queue, so a tile's keyframe records are re-read
// from L2 rather than HBM.  Tiles are ordered by job and leaf, or heaviest
// first (keyframe bytes staged plus bytes written), and split into 1 or 8
// queues.  With 8 queues, workgroup i starts on queue i % 8 -- workgroups are
// dispatched to the 8 XCDs round-robin, so a queue's tiles stay in one XCD's
// L2 -- and steals from the next queues when its own is empty.  A queue is a
// counter that workgroups advance by `chunk` items; a workgroup takes its next
// ticket before processing the items it holds, so the atomic's round trip
// overlaps the work.  A kernel node clears the counters at the start of every
// replay (no memset node).  The static form of the same queue launches one
// workgroup per `chunk` items in queue order and leaves the balancing to the
// hardware dispatcher; the autotuner times both forms.
//
// Scientific jobs (a DCT per leaf control, then voxels) need three times the
// registers of the render tile; folded into the persistent kernel they would
// cut its occupancy for every job.  They run as their own kernel nodes on a
// forked branch of the graph, concurrently with the persistent kernel.
#include "vbt_format.h"
#include "vbt_internal.h"
#include "vbt_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

using namespace vbtk;

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return vbt_fail(e_ == hipErrorOutOfMemory ? VBT_ERR_NOMEM : VBT_ERR_DEVICE,                 \
                            std::string(#expr " failed: ") + hipGetErrorString(e_));                    \
    } while (0)

namespace {

enum : uint32_t { JOB_PACKED = 0, JOB_576 = 1, JOB_SCI = 2 };

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t MAX_QUEUES = 8;
constexpr uint32_t COUNTER_STRIDE = 32; // counters 128 bytes apart
constexpr int MAX_RANGES = 64;
constexpr int MAX_CHUNK = 64;

struct BatchJob {
    AssetDev a;
    float* out;
    uint32_t* err;
    const uint32_t* tiles; // packed jobs: the asset's tile list of the batch's tile shape
    int t0, nt, per;       // per: frames per range
    uint32_t kind;
};

struct BatchArgs {
    const BatchJob* jobs;
    const uint2* entries;   // (job, tile) in queue order
    const uint32_t* bounds; // items of queue q: [bounds[2q], bounds[2q+1])
    uint32_t* counters;     // queue q's counter at counters[q * COUNTER_STRIDE]
    uint32_t queues, ranges, chunk;
    uint32_t items; // all queues' items
};

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void k_fill_f32(float* p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// Keyframe-record bytes (plus descriptor and bins) of each 64-slot group of
// the packed cache: the cost model of heaviest-first ordering.
__global__ void __launch_bounds__(256) k_group_cost(AssetDev a, uint32_t* cost)
{
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint32_t bytes = 0;
    if (g < a.controlsPerFrame) {
        uint32_t c = a.groupLeaf[g / CACHE_GROUP];
        while (c + 1u < a.A && a.ctrl[c + 1u] <= g) ++c;
        const DevLeaf& d = a.leaves[c];
        if ((d.status & STATUS_MASK) == LEAF_OK) {
            const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
            const bool isCoarse = slot < d.cc;
            const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
            const uint32_t ci = isCoarse ? slot : slot - d.cc;
            const uint32_t desc = *(const uint32_t*)(a.payload + (size_t)d.word * 4u + s.base + ci * 4u);
            bytes = ((desc >> 16) & 0xFFu) * keyframe_bytes(s.codec) + 4u + s.bins;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) bytes += (uint32_t)__shfl_xor(bytes, off, 64);
    if ((threadIdx.x & 63u) == 0u && g < a.controlsPerFrame) cost[g / CACHE_GROUP] = bytes;
}

// Few frames: the reference bracket search per slot and frame (no staging);
// the tile is the groups [g0, g0 + ng).
__device__ __forceinline__ void exact_tile(const AssetDev& a, uint32_t g0, uint32_t ng, int f0, int n, float* out,
                                           uint64_t total, uint32_t* err)
{
    const uint64_t g = (uint64_t)g0 * CACHE_GROUP + threadIdx.x;
    if (threadIdx.x >= ng * CACHE_GROUP || g >= total) return;
    bool bad = false;
    for (int i = 0; i < n; ++i) {
        int32_t st;
        out[(uint64_t)i * total + g] = cache_slot_exact(a, g, f0 + i, st);
        bad |= st != 0;
    }
    if (bad) atomicOr(err, 1u);
}

// staging depth of the batch kernels: their item loop keeps more state live
// than the one-shot kernel, and deeper staging would cost them occupancy
constexpr uint32_t BATCH_STAGE = 8u;

template <uint32_t BLOCK, uint32_t WINDOW>
__device__ __forceinline__ void batch_item(const BatchArgs& A, uint32_t it, uint32_t* lds, uint32_t* sRed)
{
    const uint2 e = invariant_u2(A.entries, it / A.ranges);
    const uint32_t r = it % A.ranges;
    BatchJob J; // the job table is invariant while the kernel runs: scalar loads
    {
        static_assert(sizeof(BatchJob) % 4 == 0, "job table words");
        const uint32_t* src = (const uint32_t*)(A.jobs + e.x);
        uint32_t* dst = (uint32_t*)&J;
#pragma unroll
        for (uint32_t i = 0; i < sizeof(BatchJob) / 4; ++i) dst[i] = invariant_u32(src, i);
    }
    const AssetDev a = J.a;
    const int f0 = J.t0 + (int)r * J.per;
    const int n = min(J.per, J.t0 + J.nt - f0);
    if (n <= 0) return;
    const uint64_t total = a.controlsPerFrame;
    if (J.kind == JOB_PACKED) {
        float* out = J.out + (uint64_t)(f0 - J.t0) * total;
        const uint32_t g0 = invariant_u32(J.tiles, e.y), ng = invariant_u32(J.tiles, e.y + 1u) - g0;
        if (n <= CACHE_EXACT_MAX_FRAMES) exact_tile(a, g0, ng, f0, n, out, total, J.err);
        else cache_plan_tile<BLOCK, WINDOW, BATCH_STAGE>(a, g0, ng, f0, n, out, total, J.err, lds, sRed);
    } else {
        float* out = J.out + (uint64_t)(f0 - J.t0) * a.A * 576u;
        cache_render_tile<VBT_CACHE_STRIDE576, BLOCK, WINDOW, BATCH_STAGE>(a, a.groupLeaf, e.y, f0, n, out, total, J.err, lds,
                                                              sRed);
    }
}

// Thread 0: turn a ticket taken on queue q into an item range; when q is
// exhausted move on to the next queue.  Returns NONE once every queue is.
__device__ inline uint32_t resolve(const BatchArgs& A, uint32_t& q, uint32_t& tried, uint32_t ticket, uint32_t& end)
{
    for (;;) {
        const uint32_t b0 = A.bounds[2u * q], b1 = A.bounds[2u * q + 1u];
        if (ticket < b1 - b0) {
            end = min(b0 + ticket + A.chunk, b1);
            return b0 + ticket;
        }
        if (++tried >= A.queues) return NONE;
        q = q + 1u == A.queues ? 0u : q + 1u;
        ticket = atomicAdd(A.counters + q * COUNTER_STRIDE, A.chunk);
    }
}

// Every workgroup leaves the loop once all queues are exhausted: each claim
// advances a counter, and a counter past its queue's end sends the workgroup
// to the next queue (at most `queues` moves).
template <uint32_t BLOCK, uint32_t WINDOW>
__global__ void __launch_bounds__(BLOCK) k_cache_batch(BatchArgs A)
{
    __shared__ uint32_t lds[WINDOW / 4 + 4]; // + 4: reads past the window, see sweep_frames
    __shared__ uint32_t sRed[BLOCK / 64];
    __shared__ uint32_t sItem, sEnd;
    const uint32_t tid = threadIdx.x;
    uint32_t q = blockIdx.x % A.queues, tried = 0, ticket = 0;
    if (tid == 0) {
        uint32_t end = 0;
        ticket = atomicAdd(A.counters + q * COUNTER_STRIDE, A.chunk);
        sItem = resolve(A, q, tried, ticket, end);
        sEnd = end;
    }
    __syncthreads();
    for (;;) {
        uint32_t it = __builtin_amdgcn_readfirstlane(sItem);
        const uint32_t end = __builtin_amdgcn_readfirstlane(sEnd);
        __syncthreads();
        if (it == NONE) break;
        if (tid == 0) ticket = atomicAdd(A.counters + q * COUNTER_STRIDE, A.chunk);
        for (; it < end; ++it) batch_item<BLOCK, WINDOW>(A, it, lds, sRed);
        if (tid == 0) {
            uint32_t e2 = 0;
            sItem = resolve(A, q, tried, ticket, e2);
            sEnd = e2;
        }
        __syncthreads();
    }
}

// Static form of the same queue: the claim c = xcd_run_tile(i) of workgroup
// i takes items c*chunk .. c*chunk+chunk-1 of the queue order (the hardware
// dispatcher balances; each XCD works through runs of consecutive claims, so
// neighbouring tiles' shared row-end lines meet in one L2).
template <uint32_t BLOCK, uint32_t WINDOW>
__global__ void __launch_bounds__(BLOCK) k_cache_batch_static(BatchArgs A)
{
    __shared__ uint32_t lds[WINDOW / 4 + 4]; // + 4: reads past the window, see sweep_frames
    __shared__ uint32_t sRed[BLOCK / 64];
    const uint32_t claims = (A.items + A.chunk - 1u) / A.chunk;
    const uint32_t c = xcd_run_tile(blockIdx.x, tile_run_log2(claims));
    if (c >= claims) return;
    const uint32_t first = c * A.chunk;
    const uint32_t end = min(first + A.chunk, A.items);
    for (uint32_t it = first; it < end; ++it) batch_item<BLOCK, WINDOW>(A, it, lds, sRed);
}

// kernel instance of a tile shape (the CACHE_TILES table)
const void* batch_kernel(int tile, bool persistent)
{
    switch (tile) {
#define VBT_BATCH(i, B, W)                                                                                    \
    case i: return persistent ? (const void*)k_cache_batch<B, W> : (const void*)k_cache_batch_static<B, W>;
        VBT_BATCH(0, 64, 8192)
        VBT_BATCH(1, 64, 16384)
        VBT_BATCH(2, 128, 16384)
        VBT_BATCH(3, 128, 32768)
        VBT_BATCH(4, 256, 32768)
        VBT_BATCH(5, 256, 49152)
        VBT_BATCH(6, 64, 4096)
        VBT_BATCH(7, 128, 8192)
        VBT_BATCH(8, 256, 16384)
        VBT_BATCH(9, 64, 12288)
#undef VBT_BATCH
    default: return nullptr;
    }
}

void launch_batch(int tile, bool persistent, uint32_t grid, const BatchArgs& A, hipStream_t s)
{
    switch (tile) {
#define VBT_BATCH(i, B, W)                                                                                    \
    case i:                                                                                                   \
        if (persistent) k_cache_batch<B, W><<<grid, B, 0, s>>>(A);                                            \
        else k_cache_batch_static<B, W><<<run_grid(grid, tile_run_log2(grid)), B, 0, s>>>(A);                                                \
        break;
        VBT_BATCH(0, 64, 8192)
        VBT_BATCH(1, 64, 16384)
        VBT_BATCH(2, 128, 16384)
        VBT_BATCH(3, 128, 32768)
        VBT_BATCH(4, 256, 32768)
        VBT_BATCH(5, 256, 49152)
        VBT_BATCH(6, 64, 4096)
        VBT_BATCH(7, 128, 8192)
        VBT_BATCH(8, 256, 16384)
        VBT_BATCH(9, 64, 12288)
#undef VBT_BATCH
    default: break;
    }
}

static_assert(CACHE_TILE_COUNT == 10, "batch_kernel / launch_batch list every tile shape");

} // namespace

struct vbt_cache_batch_s {
    int device = 0;
    uint32_t numCU = 0;
    std::vector<vbt_cache_job> jobs;
    std::vector<uint32_t> kinds;
    std::vector<std::vector<uint32_t>> groupCost; // render jobs: cost per 64-slot group
    vbt_batch_shape shape{};
    BatchJob* d_jobs = nullptr;
    uint2* d_entries = nullptr;
    size_t entriesCap = 0;
    uint32_t* d_bounds = nullptr;   // 2 * MAX_QUEUES
    uint32_t* d_counters = nullptr; // MAX_QUEUES * COUNTER_STRIDE
    uint64_t items = 0;
    uint32_t grid = 0;
    hipStream_t cap = nullptr;  // capture stream
    hipStream_t side = nullptr; // branch of the scientific jobs
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    hipStream_t side2 = nullptr; // branch of the heavy render tiles
    hipEvent_t evFork2 = nullptr, evJoin2 = nullptr;
    uint32_t heavyBase = 0, heavyGrid = 0;
    uint64_t heavyItems = 0;
    uint32_t sciJobs = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<vbt_batch_shape, float>> trials;
};

namespace {

int check_shape(const vbt_batch_shape& s)
{
    if (s.tile < 0 || s.tile >= CACHE_TILE_COUNT) return vbt_fail(VBT_ERR_INVALID, "batch shape: tile out of range");
    if (s.ranges < 1 || s.ranges > MAX_RANGES) return vbt_fail(VBT_ERR_INVALID, "batch shape: ranges out of range");
    if (s.chunk < 1 || s.chunk > MAX_CHUNK) return vbt_fail(VBT_ERR_INVALID, "batch shape: chunk out of range");
    if (s.order != 0 && s.order != 1) return vbt_fail(VBT_ERR_INVALID, "batch shape: order must be 0 or 1");
    if (s.queues < 1 || s.queues > (int)MAX_QUEUES) return vbt_fail(VBT_ERR_INVALID, "batch shape: queues out of range");
    if (s.wgPerCU < 0 || s.wgPerCU > 64) return vbt_fail(VBT_ERR_INVALID, "batch shape: wgPerCU out of range");
    if (s.persistent < 0 || s.persistent > 2) return vbt_fail(VBT_ERR_INVALID, "batch shape: persistent must be 0, 1 or 2");
    if (s.persistent == 2 && s.heavyTile >= 0)
        return vbt_fail(VBT_ERR_INVALID, "batch shape: the per-job form has no heavy-tile kernel");
    if (s.heavyTile < -1 || s.heavyTile >= CACHE_TILE_COUNT)
        return vbt_fail(VBT_ERR_INVALID, "batch shape: heavyTile out of range");
    if (s.heavyTile >= 0 && (s.persistent || CACHE_TILES[s.heavyTile].block != CACHE_TILES[s.tile].block ||
                             CACHE_TILES[s.heavyTile].window <= CACHE_TILES[s.tile].window))
        return vbt_fail(VBT_ERR_INVALID,
                        "batch shape: heavyTile needs the static form, the tile's slots and a larger window");
    return VBT_OK;
}

void release_graph(vbt_cache_batch b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Build the queue of `shape` and upload it (stream-ordered after earlier work
// on the null stream: the copies are synchronous).
int plan(vbt_cache_batch b, const vbt_batch_shape& shape)
{
    int rc = check_shape(shape);
    if (rc) return rc;
    const uint32_t BLOCK = CACHE_TILES[shape.tile].block, NW = BLOCK / CACHE_GROUP;
    const uint32_t R = (uint32_t)shape.ranges;
    struct Entry {
        uint64_t cost;
        uint32_t job, tile;
    };
    std::vector<Entry> entries, heavy;
    const uint64_t lightSpan = CACHE_TILES[shape.tile].window;
    std::vector<BatchJob> hj(b->jobs.size());
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const vbt_cache_job& job = b->jobs[j];
        const AssetDev a = vbt_internal_asset_dev(job.asset);
        BatchJob& J = hj[j];
        J.a = a;
        J.out = job.d_out;
        J.err = vbt_internal_asset_err(job.asset);
        J.t0 = job.t0;
        J.nt = job.nt;
        J.per = (job.nt + (int)R - 1) / (int)R;
        J.kind = b->kinds[j];
        J.tiles = nullptr;
        if (a.A == 0 || J.kind == JOB_SCI) continue;
        {
            // packed jobs: the asset's tiles of this shape; 576 jobs: BLOCK slots per tile
            uint32_t nPlan = 0;
            const uint32_t* hPlan = nullptr;
            if (J.kind == JOB_PACKED) J.tiles = vbt_internal_asset_tiles(job.asset, shape.tile, &nPlan, &hPlan);
            const uint64_t tiles = J.tiles ? nPlan : (a.controlsPerFrame + BLOCK - 1u) / BLOCK;
            if (tiles > 0xffffffffull) return vbt_fail(VBT_ERR_INVALID, "batch: too many tiles");
            const std::vector<uint32_t>& gc = b->groupCost[j];
            for (uint64_t t = 0; t < tiles; ++t) {
                const uint64_t ga = J.tiles ? hPlan[t] : t * NW;
                const uint64_t gb = std::min<uint64_t>(J.tiles ? hPlan[t + 1] : (t + 1) * NW, gc.size());
                uint64_t c = 0; // keyframe records + descriptors + bins: about the tile's byte span
                for (uint64_t g = ga; g < gb; ++g) c += gc[g];
                const Entry e{c + 4ull * CACHE_GROUP * (gb - ga) * (uint64_t)job.nt, (uint32_t)j, (uint32_t)t};
                if (shape.heavyTile >= 0 && c > lightSpan) heavy.push_back(e);
                else entries.push_back(e);
            }
        }
    }
    if ((uint64_t)(entries.size() + heavy.size()) * R + (uint64_t)MAX_CHUNK * 65536u >= 0xffffffffull)
        return vbt_fail(VBT_ERR_INVALID, "batch: too many work items");
    const uint32_t Q = (uint32_t)shape.queues;
    std::vector<std::vector<uint2>> seg(Q);
    if (shape.order == 1) {
        std::stable_sort(entries.begin(), entries.end(), [](const Entry& x, const Entry& y) { return x.cost > y.cost; });
        for (size_t i = 0; i < entries.size(); ++i) seg[i % Q].push_back(make_uint2(entries[i].job, entries[i].tile));
    } else {
        uint64_t sum = 0;
        for (const Entry& e : entries) sum += e.cost;
        uint64_t acc = 0;
        for (const Entry& e : entries) {
            const uint32_t q = sum ? (uint32_t)std::min<uint64_t>(Q - 1u, (acc * Q) / sum) : 0u;
            seg[q].push_back(make_uint2(e.job, e.tile));
            acc += e.cost;
        }
    }
    std::vector<uint2> flat;
    flat.reserve(entries.size() + heavy.size());
    uint32_t bounds[2 * MAX_QUEUES] = {};
    for (uint32_t q = 0; q < Q; ++q) {
        bounds[2 * q] = (uint32_t)(flat.size() * R);
        flat.insert(flat.end(), seg[q].begin(), seg[q].end());
        bounds[2 * q + 1] = (uint32_t)(flat.size() * R);
    }
    // tiles whose records exceed the light window: their own static kernel
    // with the larger window, after the light tiles in the entry array
    const size_t heavyBase = flat.size();
    if (shape.order == 1)
        std::stable_sort(heavy.begin(), heavy.end(), [](const Entry& x, const Entry& y) { return x.cost > y.cost; });
    for (const Entry& e : heavy) flat.push_back(make_uint2(e.job, e.tile));
    if (flat.size() > b->entriesCap) {
        if (b->d_entries) (void)hipFree(b->d_entries);
        b->d_entries = nullptr;
        b->entriesCap = 0;
        HIPCHK(hipMalloc((void**)&b->d_entries, flat.size() * sizeof(uint2)));
        b->entriesCap = flat.size();
    }
    if (!flat.empty()) HIPCHK(hipMemcpy(b->d_entries, flat.data(), flat.size() * sizeof(uint2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_bounds, bounds, sizeof(bounds), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_jobs, hj.data(), hj.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
    int occ = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel(shape.tile, shape.persistent != 0),
                                                        (int)BLOCK, 0));
    if (occ < 1) occ = 1;
    const uint32_t perCU = shape.wgPerCU ? std::min<uint32_t>((uint32_t)shape.wgPerCU, (uint32_t)occ) : (uint32_t)occ;
    b->items = (uint64_t)heavyBase * R;
    const uint64_t claims = (b->items + (uint64_t)shape.chunk - 1u) / (uint64_t)shape.chunk;
    b->grid = shape.persistent ? (uint32_t)std::min<uint64_t>((uint64_t)b->numCU * perCU, claims) : (uint32_t)claims;
    if (shape.persistent == 2) b->grid = 0; // per-job one-shot kernels instead of the work queue
    b->heavyBase = (uint32_t)heavyBase;
    b->heavyItems = (uint64_t)heavy.size() * R;
    b->heavyGrid = (uint32_t)((b->heavyItems + (uint64_t)shape.chunk - 1u) / (uint64_t)shape.chunk);
    b->shape = shape;
    return VBT_OK;
}

// Per-job form: each render job's own one-shot build (its asset's tile shape,
// XCD runs), the jobs alternating between the stream and a second branch so
// that one job's tail overlaps the next job's start; scientific jobs on a
// third branch as in the queue form.
int enqueue_per_job(vbt_cache_batch b, hipStream_t s)
{
    // branches only where there is work for them (a branch costs a fork and a join)
    const uint32_t render = (uint32_t)b->jobs.size() - b->sciJobs;
    const bool useSide = b->sciJobs > 0, useSide2 = render > 1;
    if (useSide || useSide2) HIPCHK(hipEventRecord(b->evFork, s));
    if (useSide) HIPCHK(hipStreamWaitEvent(b->side, b->evFork, 0));
    if (useSide2) HIPCHK(hipStreamWaitEvent(b->side2, b->evFork, 0));
    int k = 0;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const vbt_cache_job& job = b->jobs[j];
        const bool sci = b->kinds[j] == JOB_SCI;
        hipStream_t st = sci ? b->side : ((k++ & 1) ? b->side2 : s);
        const int rc = vbt_internal_frame_cache_plain(job.asset, job.t0, job.nt,
                                                      sci ? VBT_CACHE_PACKED : job.layout, job.d_out, st);
        if (rc != VBT_OK && rc != VBT_ERR_FORMAT) return rc; // poisoned leaves are reported after the launch
    }
    if (useSide) {
        HIPCHK(hipEventRecord(b->evJoin, b->side));
        HIPCHK(hipStreamWaitEvent(s, b->evJoin, 0));
    }
    if (useSide2) {
        HIPCHK(hipEventRecord(b->evJoin2, b->side2));
        HIPCHK(hipStreamWaitEvent(s, b->evJoin2, 0));
    }
    return VBT_OK;
}

int enqueue(vbt_cache_batch b, hipStream_t s)
{
    if (b->shape.persistent == 2) return enqueue_per_job(b, s);
    k_zero_u32<<<1, 256, 0, s>>>(b->d_counters, MAX_QUEUES * COUNTER_STRIDE);
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        if (b->kinds[j] != JOB_576) continue;
        const AssetDev a = vbt_internal_asset_dev(b->jobs[j].asset);
        const uint64_t n = (uint64_t)a.A * 576u * (uint64_t)b->jobs[j].nt;
        if (n) k_fill_f32<<<(uint32_t)std::min<uint64_t>((n + 255u) / 256u, 65535u), 256, 0, s>>>(b->jobs[j].d_out, n);
    }
    const bool fork = b->sciJobs > 0;
    if (fork) {
        // scientific jobs on the side branch, concurrent with the queue kernel
        HIPCHK(hipEventRecord(b->evFork, s));
        HIPCHK(hipStreamWaitEvent(b->side, b->evFork, 0));
        for (size_t j = 0; j < b->jobs.size(); ++j) {
            if (b->kinds[j] != JOB_SCI) continue;
            const vbt_cache_job& job = b->jobs[j];
            const int rc = vbt_decode_frame_cache(job.asset, job.t0, job.nt, VBT_CACHE_PACKED, job.d_out, b->side);
            if (rc != VBT_OK && rc != VBT_ERR_FORMAT) return rc; // poisoned leaves are reported after the launch
        }
        HIPCHK(hipEventRecord(b->evJoin, b->side));
    }
    BatchArgs A;
    A.jobs = b->d_jobs;
    A.entries = b->d_entries;
    A.bounds = b->d_bounds;
    A.counters = b->d_counters;
    A.queues = (uint32_t)b->shape.queues;
    A.ranges = (uint32_t)b->shape.ranges;
    A.chunk = (uint32_t)b->shape.chunk;
    A.items = (uint32_t)b->items;
    const bool heavy = b->heavyGrid > 0;
    if (heavy) {
        // heavy tiles (larger LDS window) on a second branch, concurrent
        HIPCHK(hipEventRecord(b->evFork2, s));
        HIPCHK(hipStreamWaitEvent(b->side2, b->evFork2, 0));
        BatchArgs H = A;
        H.entries = b->d_entries + b->heavyBase;
        H.items = (uint32_t)b->heavyItems;
        launch_batch(b->shape.heavyTile, false, b->heavyGrid, H, b->side2);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b->evJoin2, b->side2));
    }
    if (b->grid) launch_batch(b->shape.tile, b->shape.persistent != 0, b->grid, A, s);
    HIPCHK(hipGetLastError());
    if (fork) HIPCHK(hipStreamWaitEvent(s, b->evJoin, 0));
    if (heavy) HIPCHK(hipStreamWaitEvent(s, b->evJoin2, 0));
    return VBT_OK;
}

int capture(vbt_cache_batch b)
{
    release_graph(b);
    HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b, b->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIPCHK(e);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return VBT_OK;
}

// One job of a batch: an asset on device `dev`, a frame range inside it and
// a layout it supports.  Sets the job's kind.
int check_job(const vbt_cache_job& job, int dev, uint32_t& kind)
{
    if (!job.asset || !job.d_out) return vbt_fail(VBT_ERR_INVALID, "cache batch job without asset or output");
    if (vbt_internal_asset_device(job.asset) != dev)
        return vbt_fail(VBT_ERR_INVALID, "cache batch asset lives on another device");
    const vbt_file_header* h = vbt_internal_asset_header(job.asset);
    if (job.nt <= 0 || job.t0 < 0 || (uint64_t)job.t0 + (uint64_t)job.nt > h->frames)
        return vbt_fail(VBT_ERR_INVALID, "frame range outside the asset");
    if (job.layout != VBT_CACHE_PACKED && job.layout != VBT_CACHE_STRIDE576)
        return vbt_fail(VBT_ERR_INVALID, "unknown cache layout");
    if (h->profileType == 2 && job.layout != VBT_CACHE_PACKED)
        return vbt_fail(VBT_ERR_INVALID, "scientific assets support the packed cache only");
    const AssetDev ad = vbt_internal_asset_dev(job.asset);
    if (4ull * (job.layout == VBT_CACHE_PACKED ? ad.controlsPerFrame : (uint64_t)ad.A * 576u) > 0xFFFFFFFFull)
        return vbt_fail(VBT_ERR_INVALID, "frame cache: one frame exceeds 4 GiB");
    kind = h->profileType == 2 ? JOB_SCI : (job.layout == VBT_CACHE_PACKED ? JOB_PACKED : JOB_576);
    return VBT_OK;
}

int report(vbt_cache_batch b)
{
    for (const vbt_cache_job& j : b->jobs)
        if (vbt_internal_asset_bad(j.asset)) return vbt_internal_report_bad(j.asset, "frame cache");
    return VBT_OK;
}

} // namespace

extern "C" {

void vbt_batch_default_shape(vbt_batch_shape* s)
{
    if (!s) return;
    s->tile = CACHE_TILE_DEFAULT;
    s->ranges = 1;
    s->chunk = 1;
    s->order = 1;
    s->queues = 8;
    s->wgPerCU = 0;
    s->persistent = 1;
    s->heavyTile = -1;
}

int vbt_cache_batch_create(const vbt_cache_job* jobs, int nJobs, const vbt_batch_shape* shape, vbt_cache_batch* out)
{
    if (!out || nJobs <= 0 || !jobs) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    *out = nullptr;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    vbt_cache_batch b = new vbt_cache_batch_s();
    b->device = dev;
    auto fail = [&](int rc) {
        vbt_cache_batch_destroy(b);
        return rc;
    };
    for (int j = 0; j < nJobs; ++j) {
        uint32_t kind = 0;
        const int rc = check_job(jobs[j], dev, kind);
        if (rc) return fail(rc);
        b->jobs.push_back(jobs[j]);
        b->kinds.push_back(kind);
    }
    hipDeviceProp_t prop;
    {
        const hipError_t e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return fail(vbt_fail(VBT_ERR_DEVICE, "hipGetDeviceProperties failed"));
    }
    b->numCU = (uint32_t)prop.multiProcessorCount;
    if (hipMalloc((void**)&b->d_jobs, b->jobs.size() * sizeof(BatchJob)) != hipSuccess ||
        hipMalloc((void**)&b->d_bounds, 2 * MAX_QUEUES * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_counters, MAX_QUEUES * COUNTER_STRIDE * sizeof(uint32_t)) != hipSuccess)
        return fail(vbt_fail(VBT_ERR_NOMEM, "cache batch: device allocation failed"));
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->evFork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->evJoin, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->evFork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->evJoin2, hipEventDisableTiming) != hipSuccess)
        return fail(vbt_fail(VBT_ERR_DEVICE, "cache batch: stream creation failed"));
    for (uint32_t k : b->kinds) b->sciJobs += k == JOB_SCI;
    // cost of every 64-slot group of the render jobs (for heaviest-first order)
    b->groupCost.resize(b->jobs.size());
    uint32_t* d_cost = nullptr;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        if (b->kinds[j] == JOB_SCI) continue;
        const AssetDev a = vbt_internal_asset_dev(b->jobs[j].asset);
        if (a.A == 0 || a.controlsPerFrame == 0) continue;
        const uint64_t ng = (a.controlsPerFrame + CACHE_GROUP - 1u) / CACHE_GROUP;
        if (hipMalloc((void**)&d_cost, ng * 4u) != hipSuccess) return fail(vbt_fail(VBT_ERR_NOMEM, "cache batch: cost table"));
        k_group_cost<<<(uint32_t)((a.controlsPerFrame + 255u) / 256u), 256, 0, b->cap>>>(a, d_cost);
        b->groupCost[j].resize(ng);
        const hipError_t e1 = hipMemcpyAsync(b->groupCost[j].data(), d_cost, ng * 4u, hipMemcpyDeviceToHost, b->cap);
        const hipError_t e2 = hipStreamSynchronize(b->cap);
        (void)hipFree(d_cost);
        d_cost = nullptr;
        if (e1 != hipSuccess || e2 != hipSuccess) return fail(vbt_fail(VBT_ERR_DEVICE, "cache batch: cost table copy"));
    }
    vbt_batch_shape s;
    vbt_batch_default_shape(&s);
    if (shape) s = *shape;
    int rc = plan(b, s);
    if (!rc) rc = capture(b);
    if (rc) return fail(rc);
    *out = b;
    return VBT_OK;
}

int vbt_cache_batch_destroy(vbt_cache_batch b)
{
    if (!b) return VBT_OK;
    release_graph(b);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->evFork) (void)hipEventDestroy(b->evFork);
    if (b->evJoin) (void)hipEventDestroy(b->evJoin);
    if (b->side2) (void)hipStreamDestroy(b->side2);
    if (b->evFork2) (void)hipEventDestroy(b->evFork2);
    if (b->evJoin2) (void)hipEventDestroy(b->evJoin2);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_entries) (void)hipFree(b->d_entries);
    if (b->d_bounds) (void)hipFree(b->d_bounds);
    if (b->d_counters) (void)hipFree(b->d_counters);
    delete b;
    return VBT_OK;
}

int vbt_cache_batch_launch(vbt_cache_batch b, void* stream)
{
    if (!b || !b->exec) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    HIPCHK(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return report(b);
}

int vbt_cache_batch_launch_direct(vbt_cache_batch b, void* stream)
{
    if (!b) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    const int rc = enqueue(b, (hipStream_t)stream);
    return rc ? rc : report(b);
}

int vbt_cache_batch_set_shape(vbt_cache_batch b, const vbt_batch_shape* shape)
{
    if (!b || !shape) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    HIPCHK(hipDeviceSynchronize()); // no replay may read the queue being replaced
    int rc = plan(b, *shape);
    if (!rc) rc = capture(b);
    return rc;
}

int vbt_cache_batch_get_shape(vbt_cache_batch b, vbt_batch_shape* shape)
{
    if (!b || !shape) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    *shape = b->shape;
    return VBT_OK;
}

int vbt_cache_batch_plan_info(vbt_cache_batch b, uint64_t* items, uint32_t* workgroups)
{
    if (!b) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    if (items) *items = b->items + b->heavyItems;
    if (workgroups) *workgroups = b->grid + b->heavyGrid;
    return VBT_OK;
}

int vbt_cache_batch_autotune(vbt_cache_batch b, int reps, vbt_batch_shape* best, float* bestMs, void* stream)
{
    if (!b || reps <= 0) return vbt_fail(VBT_ERR_INVALID, "invalid autotune request");
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    b->trials.clear();
    vbt_batch_shape cur = b->shape, top = b->shape;
    float topMs = 1e30f;
    int rc = VBT_OK;
    auto timeShape = [&](const vbt_batch_shape& sh) -> int {
        for (const auto& t : b->trials)
            if (!std::memcmp(&t.first, &sh, sizeof(sh))) return VBT_OK;
        int r = plan(b, sh);
        if (r) return r;
        if ((r = enqueue(b, s))) return r;
        if (hipEventRecord(e0, s) != hipSuccess) return vbt_fail(VBT_ERR_DEVICE, "hipEventRecord failed");
        for (int i = 0; i < reps; ++i)
            if ((r = enqueue(b, s))) return r;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return vbt_fail(VBT_ERR_DEVICE, "hipEventSynchronize failed");
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return vbt_fail(VBT_ERR_DEVICE, "hipEventElapsedTime failed");
        ms /= (float)reps;
        b->trials.push_back({sh, ms});
        if (ms < topMs) {
            topMs = ms;
            top = sh;
        }
        return VBT_OK;
    };
    auto msOf = [&](const vbt_batch_shape& sh) {
        for (const auto& t : b->trials)
            if (!std::memcmp(&t.first, &sh, sizeof(sh))) return t.second;
        return 1e30f;
    };
    // coordinate descent: kernel form x tile shape, then -- from the best tile
    // of each form, since the two forms often tie on tiles alone -- frame
    // ranges, chunk, order, queues (persistent form) or a heavy-tile kernel
    // (static form)
    for (int p : {1, 0})
        for (int t = 0; t < CACHE_TILE_COUNT && !rc; ++t) {
            vbt_batch_shape sh = cur;
            sh.persistent = p;
            sh.tile = t;
            rc = timeShape(sh);
        }
    if (!rc) { // each job's own one-shot build, jobs on two branches
        vbt_batch_shape sh = cur;
        sh.persistent = 2;
        sh.heavyTile = -1;
        rc = timeShape(sh);
    }
    for (int form : {1, 0}) {
        if (rc) break;
        vbt_batch_shape lt = top;
        float lms = 1e30f;
        for (const auto& t : b->trials)
            if (t.first.persistent == form && t.second < lms) {
                lms = t.second;
                lt = t.first;
            }
        auto step = [&](const vbt_batch_shape& sh) {
            if (rc) return;
            rc = timeShape(sh);
            const float ms = msOf(sh);
            if (!rc && ms < lms) {
                lms = ms;
                lt = sh;
            }
        };
        vbt_batch_shape cur2 = lt;
        for (int r : {1, 2, 4, 8}) {
            vbt_batch_shape sh = cur2;
            sh.ranges = r;
            step(sh);
        }
        cur2 = lt;
        for (int c : {1, 2, 4}) {
            vbt_batch_shape sh = cur2;
            sh.chunk = c;
            step(sh);
        }
        cur2 = lt;
        for (int o : {0, 1}) {
            vbt_batch_shape sh = cur2;
            sh.order = o;
            step(sh);
        }
        cur2 = lt;
        for (int q : {1, 8}) {
            if (!cur2.persistent) break;
            vbt_batch_shape sh = cur2;
            sh.queues = q;
            step(sh);
        }
        // static form: tiles whose records exceed the window go to a second
        // kernel with a larger window (same slots per workgroup)
        cur2 = lt;
        for (int h = 0; h < CACHE_TILE_COUNT && !cur2.persistent; ++h) {
            if (CACHE_TILES[h].block != CACHE_TILES[cur2.tile].block ||
                CACHE_TILES[h].window <= CACHE_TILES[cur2.tile].window)
                continue;
            vbt_batch_shape sh = cur2;
            sh.heavyTile = h;
            step(sh);
        }
        // with a heavy kernel, a smaller light window may now pay
        cur2 = lt;
        for (int t = 0; t < CACHE_TILE_COUNT && cur2.heavyTile >= 0; ++t) {
            if (CACHE_TILES[t].block != CACHE_TILES[cur2.tile].block ||
                CACHE_TILES[t].window >= CACHE_TILES[cur2.heavyTile].window)
                continue;
            vbt_batch_shape sh = cur2;
            sh.tile = t;
            step(sh);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
    if ((rc = plan(b, top)) || (rc = capture(b))) return rc;
    if (best) *best = top;
    if (bestMs) *bestMs = topMs;
    return VBT_OK;
}

int vbt_cache_batch_trial_count(vbt_cache_batch b) { return b ? (int)b->trials.size() : 0; }

int vbt_cache_batch_trial(vbt_cache_batch b, int i, vbt_batch_shape* shape, float* ms)
{
    if (!b || i < 0 || i >= (int)b->trials.size()) return vbt_fail(VBT_ERR_INVALID, "trial index out of range");
    if (shape) *shape = b->trials[(size_t)i].first;
    if (ms) *ms = b->trials[(size_t)i].second;
    return VBT_OK;
}

} // extern "C"

// ----------------------------------------------- batch run, no allocation --
// vbt_cache_batch_run_noalloc: the batch's work queue (job table, entries,
// queue bounds, queue counters) lives in the file-scope device array
// `scratch` instead of memory allocated per batch, and is written by two
// small kernels from kernel arguments, so a run calls no hipMalloc / hipFree
// and no host-to-device copy.  Entries are in job / leaf order (order 0) and
// the queues split them evenly by tile count; the decode kernels are the
// batch's own.
namespace {

constexpr uint32_t NA_MAX_JOBS = VBT_NOALLOC_MAX_JOBS;
constexpr uint32_t NA_MAX_TILES = VBT_NOALLOC_MAX_TILES;
constexpr size_t na_round(size_t n) { return (n + 255u) & ~(size_t)255u; }
constexpr size_t NA_JOBS_OFF = 0;
constexpr size_t NA_BOUNDS_OFF = NA_JOBS_OFF + na_round(NA_MAX_JOBS * sizeof(BatchJob));
constexpr size_t NA_COUNTERS_OFF = NA_BOUNDS_OFF + na_round(2 * MAX_QUEUES * sizeof(uint32_t));
constexpr size_t NA_ENTRIES_OFF = NA_COUNTERS_OFF + na_round(MAX_QUEUES * COUNTER_STRIDE * sizeof(uint32_t));
constexpr size_t NA_SCRATCH_BYTES = NA_ENTRIES_OFF + (size_t)NA_MAX_TILES * sizeof(uint2);
static_assert(NA_SCRATCH_BYTES <= VBT_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the stated maximum");

// One copy per device, held until the process ends.  Host code never uses
// this name as an address: it takes the device address from
// hipGetSymbolAddress (na_scratch below).
static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

constexpr uint32_t NA_CHUNK = 8; // job records per setup launch (kernel arguments)

struct NaJobChunk {
    BatchJob jobs[NA_CHUNK];
    uint32_t first, count;
};

struct NaPlan {
    uint32_t prefix[NA_MAX_JOBS + 1]; // first tile of each job in the entry order
    uint32_t bounds[2 * MAX_QUEUES];
    uint32_t nJobs, tiles;
};

__global__ void k_na_jobs(NaJobChunk c, BatchJob* dst)
{
    if (threadIdx.x != 0) return;
#pragma unroll
    for (uint32_t j = 0; j < NA_CHUNK; ++j)
        if (j < c.count) dst[c.first + j] = c.jobs[j];
}

// Entries (job, tile) in job / leaf order, the queue bounds, and the queue
// counters cleared (every run: the array keeps the previous run's values).
__global__ void __launch_bounds__(256) k_na_entries(NaPlan p, uint2* entries, uint32_t* bounds, uint32_t* counters)
{
    if (blockIdx.x == 0) {
        for (uint32_t k = threadIdx.x; k < MAX_QUEUES * COUNTER_STRIDE; k += 256u) counters[k] = 0u;
        if (threadIdx.x < 2u * MAX_QUEUES) bounds[threadIdx.x] = p.bounds[threadIdx.x];
    }
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.tiles) return;
    uint32_t j = 0;
    while (j + 1u < p.nJobs && p.prefix[j + 1u] <= i) ++j;
    entries[i] = make_uint2(j, i - p.prefix[j]);
}

struct NaDevice {
    uint8_t* ptr = nullptr; // device address of `scratch` on this device
    uint32_t numCU = 0;
};

std::mutex naMu;
NaDevice naDev[64];

// The device address of `scratch` on the current device, looked up once per
// device with that device current (each device holds its own copy).
int na_scratch(int dev, NaDevice& out)
{
    if (dev < 0 || dev >= 64) return vbt_fail(VBT_ERR_INVALID, "no-alloc batch: device index above 63");
    std::lock_guard<std::mutex> g(naMu);
    NaDevice& d = naDev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return vbt_fail(VBT_ERR_DEVICE, std::string("no-alloc batch: hipGetSymbolAddress failed: ") +
                                                hipGetErrorString(e));
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            return vbt_fail(VBT_ERR_DEVICE, "no-alloc batch: compute unit count");
        d.numCU = (uint32_t)cu;
        d.ptr = (uint8_t*)p;
    }
    out = d;
    return VBT_OK;
}

} // namespace

extern "C" {

void vbt_batch_noalloc_default_shape(vbt_batch_shape* s)
{
    if (!s) return;
    vbt_batch_default_shape(s);
    s->order = 0;
}

int vbt_cache_batch_run_noalloc(const vbt_cache_job* jobs, int nJobs, const vbt_batch_shape* shape, void* stream)
{
    if (!jobs || nJobs <= 0) return vbt_fail(VBT_ERR_INVALID, "invalid cache batch");
    if (nJobs > (int)NA_MAX_JOBS)
        return vbt_fail(VBT_ERR_INVALID, "no-alloc batch: more than " + std::to_string(NA_MAX_JOBS) + " jobs");
    vbt_batch_shape sh;
    vbt_batch_noalloc_default_shape(&sh);
    if (shape) sh = *shape;
    int rc = check_shape(sh);
    if (rc) return rc;
    if (sh.order != 0 || sh.heavyTile != -1)
        return vbt_fail(VBT_ERR_INVALID, "no-alloc batch: shape needs order 0 (job / leaf order) and heavyTile -1");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    NaDevice nd;
    if ((rc = na_scratch(dev, nd))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (sh.persistent == 2) { // each job's own one-shot build: no work queue, nothing in the scratch array
        uint32_t kind;
        for (int j = 0; j < nJobs; ++j)
            if ((rc = check_job(jobs[j], dev, kind))) return rc;
        bool bad = false;
        for (int j = 0; j < nJobs; ++j) {
            const vbt_cache_job& job = jobs[j];
            rc = vbt_internal_frame_cache_plain(job.asset, job.t0, job.nt, job.layout, job.d_out, s);
            if (rc == VBT_ERR_FORMAT) bad = true;
            else if (rc) return rc;
        }
        return bad ? VBT_ERR_FORMAT : VBT_OK;
    }
    const uint32_t BLOCK = CACHE_TILES[sh.tile].block, R = (uint32_t)sh.ranges;
    BatchJob hj[NA_MAX_JOBS];
    uint32_t kinds[NA_MAX_JOBS];
    NaPlan plan;
    std::memset(&plan, 0, sizeof(plan));
    uint64_t tiles = 0;
    for (int j = 0; j < nJobs; ++j) {
        const vbt_cache_job& job = jobs[j];
        if ((rc = check_job(job, dev, kinds[j]))) return rc;
        const AssetDev a = vbt_internal_asset_dev(job.asset);
        BatchJob& J = hj[j];
        J.a = a;
        J.out = job.d_out;
        J.err = vbt_internal_asset_err(job.asset);
        J.t0 = job.t0;
        J.nt = job.nt;
        J.per = (job.nt + (int)R - 1) / (int)R;
        J.kind = kinds[j];
        J.tiles = nullptr;
        plan.prefix[j] = (uint32_t)std::min<uint64_t>(tiles, NA_MAX_TILES);
        if (a.A != 0 && J.kind == JOB_PACKED) {
            uint32_t nPlan = 0;
            const uint32_t* hPlan = nullptr;
            J.tiles = vbt_internal_asset_tiles(job.asset, sh.tile, &nPlan, &hPlan);
            tiles += nPlan;
        } else if (a.A != 0 && J.kind != JOB_SCI) {
            tiles += (a.controlsPerFrame + BLOCK - 1u) / BLOCK;
        }
    }
    if (tiles > NA_MAX_TILES)
        return vbt_fail(VBT_ERR_INVALID, "no-alloc batch: " + std::to_string(tiles) + " leaf tiles of " +
                                             std::to_string(BLOCK) + " slots exceed the scratch maximum of " +
                                             std::to_string(NA_MAX_TILES));
    plan.prefix[nJobs] = (uint32_t)tiles;
    plan.nJobs = (uint32_t)nJobs;
    plan.tiles = (uint32_t)tiles;
    const uint32_t Q = (uint32_t)sh.queues;
    for (uint32_t q = 0; q < Q; ++q) {
        plan.bounds[2 * q] = (uint32_t)(tiles * q / Q) * R;
        plan.bounds[2 * q + 1] = (uint32_t)(tiles * (q + 1u) / Q) * R;
    }
    const uint64_t items = tiles * R;
    BatchJob* dJobs = (BatchJob*)(nd.ptr + NA_JOBS_OFF);
    uint32_t* dBounds = (uint32_t*)(nd.ptr + NA_BOUNDS_OFF);
    uint32_t* dCounters = (uint32_t*)(nd.ptr + NA_COUNTERS_OFF);
    uint2* dEntries = (uint2*)(nd.ptr + NA_ENTRIES_OFF);
    for (uint32_t f = 0; f < (uint32_t)nJobs; f += NA_CHUNK) {
        NaJobChunk c;
        std::memset(&c, 0, sizeof(c));
        c.first = f;
        c.count = std::min<uint32_t>(NA_CHUNK, (uint32_t)nJobs - f);
        for (uint32_t i = 0; i < c.count; ++i) c.jobs[i] = hj[f + i];
        k_na_jobs<<<1, 64, 0, s>>>(c, dJobs);
    }
    k_na_entries<<<(uint32_t)std::max<uint64_t>(1u, (tiles + 255u) / 256u), 256, 0, s>>>(plan, dEntries, dBounds,
                                                                                        dCounters);
    HIPCHK(hipGetLastError());
    for (int j = 0; j < nJobs; ++j) {
        if (kinds[j] != JOB_576) continue;
        const uint64_t n = (uint64_t)hj[j].a.A * 576u * (uint64_t)jobs[j].nt;
        if (n) k_fill_f32<<<(uint32_t)std::min<uint64_t>((n + 255u) / 256u, 65535u), 256, 0, s>>>(jobs[j].d_out, n);
    }
    for (int j = 0; j < nJobs; ++j) {
        if (kinds[j] != JOB_SCI) continue;
        rc = vbt_decode_frame_cache(jobs[j].asset, jobs[j].t0, jobs[j].nt, VBT_CACHE_PACKED, jobs[j].d_out, stream);
        if (rc != VBT_OK && rc != VBT_ERR_FORMAT) return rc; // poisoned leaves are reported after the launch
    }
    if (items) {
        int occ = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel(sh.tile, sh.persistent != 0),
                                                            (int)BLOCK, 0));
        if (occ < 1) occ = 1;
        const uint32_t perCU = sh.wgPerCU ? std::min<uint32_t>((uint32_t)sh.wgPerCU, (uint32_t)occ) : (uint32_t)occ;
        const uint64_t claims = (items + (uint64_t)sh.chunk - 1u) / (uint64_t)sh.chunk;
        const uint32_t grid =
            sh.persistent ? (uint32_t)std::min<uint64_t>((uint64_t)nd.numCU * perCU, claims) : (uint32_t)claims;
        BatchArgs A;
        A.jobs = dJobs;
        A.entries = dEntries;
        A.bounds = dBounds;
        A.counters = dCounters;
        A.queues = Q;
        A.ranges = R;
        A.chunk = (uint32_t)sh.chunk;
        A.items = (uint32_t)items;
        launch_batch(sh.tile, sh.persistent != 0, grid, A, s);
        HIPCHK(hipGetLastError());
    }
    for (int j = 0; j < nJobs; ++j)
        if (vbt_internal_asset_bad(jobs[j].asset)) return vbt_internal_report_bad(jobs[j].asset, "frame cache");
    return VBT_OK;
}

} // extern "C"
