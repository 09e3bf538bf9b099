// server.hip -- continuous-batching decode server with a device-resident
// request loop.
//
// Host side: submitters copy query inputs into pinned staging and queue the
// request; a dispatcher thread admits queued requests in arrival order
// (device arena + HBM budget + in-flight limit), writes them into a request
// ring in host-coherent pinned memory and publishes the ring tail.  It polls
// per-request completion flags that the device writes into host memory, and
// (re)launches the device loop when work is waiting and no loop runs.
//
// Device side (k_serve; one launch serves at most RING requests):
//   block 0, wave 0   scheduler: polls the host ring tail, copies new ring
//                     entries into a device-memory mirror and appends them to
//                     their class list (point queries / frame builds),
//                     publishes them.
//   other blocks      workers: claim items of the oldest request with
//                     unclaimed items in a class (a per-request atomic
//                     counter; the class cursor only moves forward), query
//                     batches first in 7 of 8 workgroups and frame builds
//                     first in the 8th, decode the item, and count it done;
//                     the workgroup that completes a request's last item
//                     flushes and writes its host flag.
// Ring entries are never reused within a launch, so every mirror entry and
// counter is written once before it is published and needs no generation tag.
// Every wait is bounded: the scheduler leaves on the host stop flag, on ring
// exhaustion, after idleUs without new entries, or after leaseMs; workers
// leave once the scheduler has left and no published item remains, or after
// a hard bound.  The host requeues whatever a finished launch did not
// complete (entries the scheduler never consumed) into the next launch.
//
// Work items (uneven cost; mixed assets, routes and request kinds interleave
// on the CUs):
//   query item   queryItem point queries of one request, 8 lanes per query
//   cache item   cacheItem packed-cache slots of one render frame
//   sci item     four scientific leaves (their voxels) of one frame
#include "vbt_format.h"
#include "vbt_internal.h"
#include "vbt_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace vbtk;

namespace {

constexpr uint32_t RING = 65536;  // requests per device-loop launch
constexpr uint32_t MAX_ASSETS = 32;
#ifndef VBT_SERVE_BLOCK
#define VBT_SERVE_BLOCK 256
#endif
constexpr uint32_t SERVE_BLOCK = VBT_SERVE_BLOCK;  // lanes per workgroup
constexpr uint32_t QUERY_PASS = SERVE_BLOCK / 8u;  // queries per pass of a workgroup
constexpr uint32_t QUERY_CHUNK = 512;              // queries staged in LDS at once
constexpr uint32_t SCI_LEAVES = SERVE_BLOCK / 64u; // scientific leaves per item (one control per lane)
constexpr uint32_t SBUF_WORDS = QUERY_CHUNK * 4u;  // >= SCI_LEAVES * 64 controls

enum : uint32_t { REQ_QUERY = 0, REQ_CACHE = 1 };

// Host -> device request record (ring entry), 64 bytes.
struct RingEntry {
    uint32_t type, asset;
    int32_t t;
    uint32_t n;        // queries, or packed slots per frame
    uint64_t q;        // device-mapped host array of queries (x,y,z,t u32 each)
    uint64_t out;      // device output values
    uint64_t status;   // device per-query statuses (queries)
    uint32_t itemCount, pad;
    ComposeDev cm;
};
static_assert(sizeof(RingEntry) == 64, "ring entry layout");

// Device mirror of a ring entry plus its counters.
struct DevReq {
    RingEntry e;
    uint32_t claim;    // items handed out
    uint32_t done;     // items finished
    uint32_t err;      // an element failed
    uint32_t pad;
};
static_assert(sizeof(DevReq) == 80, "mirror entry layout");

// Host-coherent control words, each on its own 128-byte line.
struct HostShared {
    uint32_t tail;     // host: ring entries published in this launch
    uint32_t pad0[31];
    uint32_t stop;     // host: leave the loop
    uint32_t pad1[31];
    uint32_t head;     // device: ring entries consumed
    uint32_t pad2[31];
    uint32_t exited;   // device: the scheduler has left
    uint32_t pad3[31];
};

// Device-memory loop state, cleared before every launch.
// Requests are claimed in two classes, each in arrival order: point-query
// batches (latency first) and frame builds.  The scheduler appends each new
// request's mirror index to its class list.
struct DevState {
    uint32_t pub[2];     // per class: list entries visible to workers
    uint32_t cur[2];     // per class: lowest list position that may still have unclaimed items
    uint32_t exit;       // the scheduler has left
    uint32_t pad[3];
};

struct ServeArgs {
    HostShared* hs;
    const RingEntry* ring;
    uint32_t* doneHost;  // per ring entry: 1 | err << 1 once complete
    DevState* st;
    DevReq* mirror;
    uint32_t* lists;     // 2 x RING mirror indices, one list per class
    const AssetDev* assets;
    uint64_t idleTicks, leaseTicks, workerTicks;
    uint32_t queryItem, cacheItem;
};

template <typename T>
__device__ __forceinline__ T ld_sys(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ __forceinline__ T ld_dev(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Release-only fences: write back (L2) and wait, never invalidate.  No load in
// the loop is an acquire: everything a worker reads that another agent or
// workgroup wrote during the launch (ring entries, mirror entries, counters,
// query inputs) is read with coherent scoped loads, and the asset data is
// immutable while the loop runs, so cached lines never need dropping.
__device__ __forceinline__ void release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void release_system() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, ""); }

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Scheduler: wave 0 of block 0.
__device__ void serve_scheduler(const ServeArgs& A)
{
    const uint32_t lane = threadIdx.x;
    const uint64_t start = wall_clock64();
    uint64_t last = start;
    uint32_t head = 0, n[2] = {0u, 0u};
    const uint64_t below = (1ull << lane) - 1ull;
    for (;;) {
        const uint32_t tail0 = ld_sys(&A.hs->tail);
        const uint32_t stop = ld_sys(&A.hs->stop);
        const uint32_t tail = tail0 < RING ? tail0 : RING;
        if (tail > head) {
            for (uint32_t j0 = head; j0 < tail; j0 += 64u) {
                const uint32_t j = j0 + lane;
                const bool act = j < tail;
                uint32_t type = 0u;
                if (act) {
                    const uint32_t* src = (const uint32_t*)(A.ring + j);
                    uint32_t* dst = (uint32_t*)(A.mirror + j);
#pragma unroll
                    for (int w = 0; w < 16; ++w) dst[w] = ld_sys(src + w);
                    dst[16] = 0u;
                    dst[17] = 0u;
                    dst[18] = 0u;
                    dst[19] = 0u;
                    type = dst[0];
                }
                // append to the class lists in arrival order (wave ballot + popcount)
                const uint64_t mq = __ballot(act && type == REQ_QUERY);
                const uint64_t mc = __ballot(act && type != REQ_QUERY);
                if (act) {
                    const uint32_t cls = type == REQ_QUERY ? 0u : 1u;
                    const uint32_t pos = n[cls] + (uint32_t)__popcll((cls ? mc : mq) & below);
                    A.lists[cls * RING + pos] = j;
                }
                n[0] += (uint32_t)__popcll(mq);
                n[1] += (uint32_t)__popcll(mc);
            }
            release_agent();  // mirror entries and lists reach device-coherent memory before the publish
            if (lane == 0) {
                __hip_atomic_store(&A.st->pub[0], n[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&A.st->pub[1], n[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&A.hs->head, tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            head = tail;
            last = wall_clock64();
        }
        const uint64_t t = wall_clock64();
        if (stop || head >= RING || t - last > A.idleTicks || t - start > A.leaseTicks) break;
        __builtin_amdgcn_s_sleep(8);
    }
    release_agent();  // the last publish completes before the exit flag
    if (lane == 0) {
        __hip_atomic_store(&A.st->exit, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&A.hs->exited, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Waves per SIMD the register allocation is sized for (4: 128 VGPRs, every
// item path without spills).
#ifndef VBT_SERVE_WAVES
#define VBT_SERVE_WAVES 4
#endif
__global__ void __launch_bounds__(SERVE_BLOCK) __attribute__((amdgpu_waves_per_eu(VBT_SERVE_WAVES, VBT_SERVE_WAVES)))
k_serve(ServeArgs A)
{
    if (blockIdx.x == 0) {
        if (threadIdx.x < 64u) serve_scheduler(A);
        return;
    }
    __shared__ uint32_t sCmd, sR, sK, sErr;
    __shared__ RingEntry sE;
    __shared__ uint32_t sBuf[SBUF_WORDS];  // query chunk or scientific controls
    uint32_t* const sQ = sBuf;
    float* const ctrl = (float*)sBuf;
    const uint32_t tid = threadIdx.x;
    uint32_t haveR = 0xFFFFFFFFu;  // request whose entry sits in sE (thread 0's view)
    uint64_t waitStart = 0;
    for (;;) {
        if (tid == 0) {
            uint32_t cmd = 2u;  // 0 work, 1 leave, 2 poll again
            // one workgroup in eight looks at frame builds first, the rest at
            // point queries first; either falls through to the other class
            const uint32_t firstCls = (blockIdx.x & 7u) == 0u ? 1u : 0u;
            for (uint32_t i = 0; i < 2u && cmd == 2u; ++i) {
                const uint32_t cls = firstCls ^ i;
                uint32_t pos = ld_dev(&A.st->cur[cls]);
                const uint32_t p = ld_dev(&A.st->pub[cls]);
                while (pos < p) {  // pos only grows, up to p
                    const uint32_t r = ld_dev(A.lists + cls * RING + pos);
                    DevReq* R = A.mirror + r;
                    const uint32_t cnt = ld_dev(&R->e.itemCount);
                    const uint32_t k =
                        __hip_atomic_fetch_add(&R->claim, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (k < cnt) {
                        if (haveR != r) {
                            const uint32_t* src = (const uint32_t*)&R->e;
                            uint32_t* dst = (uint32_t*)&sE;
#pragma unroll
                            for (int w = 0; w < 16; ++w) dst[w] = ld_dev(src + w);
                            haveR = r;
                        }
                        sR = r;
                        sK = k;
                        sErr = 0u;
                        cmd = 0u;
                        waitStart = 0;
                        break;
                    }
                    __hip_atomic_fetch_max(&A.st->cur[cls], pos + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ++pos;
                }
            }
            if (cmd == 2u) {
                // the counts are read only after exit is seen set (the branch
                // waits for that load), so they are final
                const uint32_t ex = ld_dev(&A.st->exit);
                const uint64_t t = wall_clock64();
                if (ex && ld_dev(&A.st->cur[0]) >= ld_dev(&A.st->pub[0]) &&
                    ld_dev(&A.st->cur[1]) >= ld_dev(&A.st->pub[1])) {
                    cmd = 1u;
                } else if (waitStart == 0) {
                    waitStart = t;
                } else if (t - waitStart > A.workerTicks) {
                    cmd = 1u;
                }
                if (cmd == 2u) __builtin_amdgcn_s_sleep(16);
            }
            sCmd = cmd;
        }
        __syncthreads();
        const uint32_t cmd = sCmd;
        if (cmd == 1u) break;
        if (cmd == 2u) {
            __syncthreads();
            continue;
        }
        // the item's fields are wave-uniform: keep them (and so the asset
        // descriptor's address) in scalar registers
        const uint32_t r = uni(sR), k = uni(sK);
        RingEntry E;
        {
            const uint32_t* src = (const uint32_t*)&sE;
            uint32_t* dst = (uint32_t*)&E;
#pragma unroll
            for (int w = 0; w < 16; ++w) dst[w] = uni(src[w]);
        }
        const AssetDev& a = A.assets[E.asset];
        bool err = false;
        if (E.type == REQ_QUERY) {
            // queries come from host memory: one round trip per chunk into LDS
            const uint32_t* q = (const uint32_t*)E.q;
            float* out = (float*)E.out;
            int32_t* status = (int32_t*)E.status;
            const uint32_t first = k * A.queryItem;
            const uint32_t endQ = min(E.n, first + A.queryItem);
            for (uint32_t c0 = first; c0 < endQ; c0 += QUERY_CHUNK) {
                const uint32_t c1 = min(endQ, c0 + QUERY_CHUNK);
                const uint32_t nd = (c1 - c0) * 4u;
                uint32_t tmp[QUERY_CHUNK * 4u / SERVE_BLOCK];
#pragma unroll
                for (uint32_t i = 0; i < QUERY_CHUNK * 4u / SERVE_BLOCK; ++i) {
                    const uint32_t w = tid + i * SERVE_BLOCK;
                    tmp[i] = w < nd ? ld_sys(q + (size_t)c0 * 4u + w) : 0u;
                }
#pragma unroll
                for (uint32_t i = 0; i < QUERY_CHUNK * 4u / SERVE_BLOCK; ++i) {
                    const uint32_t w = tid + i * SERVE_BLOCK;
                    if (w < nd) sQ[w] = tmp[i];
                }
                __syncthreads();
                for (uint32_t base = c0; base < c1; base += QUERY_PASS) {
                    const uint32_t qi = base + tid / 8u, sub = tid & 7u;
                    const bool valid = qi < c1;
                    const uint32_t o = ((valid ? qi : c1 - 1u) - c0) * 4u;
                    const uint4 rq = make_uint4(sQ[o], sQ[o + 1u], sQ[o + 2u], sQ[o + 3u]);
                    int32_t stq;
                    const float v = point_query<8>(a, E.cm, rq, sub, stq);
                    if (valid && sub == 0u) {
                        out[qi] = stq ? __builtin_nanf("") : v;
                        status[qi] = stq;
                        err |= stq != 0;
                    }
                }
                __syncthreads();
            }
        } else if (a.profile != 2u) {
            // cacheItem packed-cache slots of one frame, each by the reference
            // bracket search (one frame needs ~3 keyframe records per slot,
            // far fewer bytes than staging the slot's whole keyframe stream);
            // four independent slots per lane at a time for memory parallelism
            float* out = (float*)E.out;
            const uint64_t first = (uint64_t)k * A.cacheItem;
            const uint64_t endS = min<uint64_t>(a.controlsPerFrame, first + A.cacheItem);
            for (uint64_t g0 = first + tid; g0 < endS; g0 += 4u * SERVE_BLOCK) {
                float v[4];
                int32_t st[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint64_t g = g0 + (uint64_t)j * SERVE_BLOCK;
                    v[j] = g < endS ? cache_slot_exact(a, g, E.t, st[j]) : 0.0f;
                    if (g >= endS) st[j] = 0;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint64_t g = g0 + (uint64_t)j * SERVE_BLOCK;
                    if (g < endS) out[g] = v[j];
                    err |= st[j] != 0;
                }
            }
        } else {
            // SCI_LEAVES scientific leaves: one DCT control per lane into LDS,
            // then the leaves' voxels across the workgroup
            float* out = (float*)E.out;
            const uint32_t c0 = k * SCI_LEAVES;
            const uint32_t cl = c0 + tid / 64u;
            if (cl < a.A) {
                const DevLeaf& d = a.leaves[cl];
                const Bytes b{a.payload + (size_t)d.word * 4u};
                ctrl[tid] = (d.status & STATUS_MASK) == LEAF_OK ? sci_control(a, b, sci_header(d.sciHeader), tid & 63u, E.t)
                                                                : 0.0f;
            }
            __syncthreads();
            for (uint32_t v = tid; v < SCI_LEAVES * 512u; v += SERVE_BLOCK) {
                const uint32_t c = c0 + v / 512u, vox = v & 511u;
                if (c >= a.A) break;
                const DevLeaf& d = a.leaves[c];
                const uint32_t lx = vox & 7u, ly = (vox >> 3) & 7u, lz = vox >> 6;
                if (lx < d.lw && ly < d.lh && lz < d.ld) {
                    const bool bad = (d.status & STATUS_MASK) != LEAF_OK;
                    const uint64_t dst = a.ctrl[c] + (lz * d.lh + ly) * d.lw + lx;
                    const Bytes b{a.payload + (size_t)d.word * 4u};
                    out[dst] = bad ? __builtin_nanf("")
                                   : sci_coarse_sample(ctrl + (v / 512u) * 64u, lx, ly, lz) +
                                         sci_residual(a, b, sci_header(d.sciHeader), lx, ly, lz, (uint32_t)E.t);
                    err |= bad;
                }
            }
        }
        if (err) atomicOr(&sErr, 1u);
        release_agent();  // this lane's outputs reach memory before the count
        __syncthreads();
        if (tid == 0) {
            DevReq* R = A.mirror + r;
            if (sErr) __hip_atomic_fetch_or(&R->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t old = __hip_atomic_fetch_add(&R->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1u == E.itemCount) {
                release_system();
                const uint32_t e = ld_dev(&R->err);
                __hip_atomic_store(&A.doneHost[r], 1u | (e << 1), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        __syncthreads();
    }
}

__global__ void k_clear_state(DevState* st)
{
    if (threadIdx.x == 0) *st = DevState{{0u, 0u}, {0u, 0u}, 0u, {0u, 0u, 0u}};
}

using Clock = std::chrono::steady_clock;

struct Req {
    uint64_t id = 0;
    uint32_t type = 0, asset = 0, n = 0, items = 0;
    int32_t t = 0;
    ComposeDev cm{};
    uint64_t stageOff = 0, stageBytes = 0;  // pinned staging (query inputs)
    uint64_t devOff = 0, devBytes = 0;      // device arena (outputs)
    bool hasDev = false, hasStage = false;
    Clock::time_point submitted, finished;
    int state = 0;                          // 0 queued, 1 admitted, 2 done
    int32_t status = 0;
    uint32_t ring = 0;
    uint64_t launch = 0;
};

// First-fit allocator with a coalescing free list (offsets into one block).
struct Arena {
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> freeList;  // offset -> size
    static uint64_t round(uint64_t b) { return (b + 255u) & ~uint64_t(255u); }
    void init(uint64_t n)
    {
        size = n;
        used = 0;
        freeList.clear();
        if (n) freeList[0] = n;
    }
    bool alloc(uint64_t bytes, uint64_t& off)
    {
        bytes = round(bytes);
        for (auto it = freeList.begin(); it != freeList.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const uint64_t rest = it->second - bytes;
                freeList.erase(it);
                if (rest) freeList[off + bytes] = rest;
                used += bytes;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = round(bytes);
        used -= bytes;
        auto it = freeList.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != freeList.end() && it->first + it->second == next->first) {
            it->second += next->second;
            freeList.erase(next);
        }
        if (it != freeList.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                freeList.erase(it);
            }
        }
    }
};

uint64_t out_bytes(const Req& r)
{
    const uint64_t v = Arena::round(4ull * r.n);
    return r.type == REQ_QUERY ? 2u * v : v;
}

int hfail(hipError_t e, const char* what)
{
    return vbt_fail(e == hipErrorOutOfMemory ? VBT_ERR_NOMEM : VBT_ERR_DEVICE,
                    std::string(what) + " failed: " + hipGetErrorString(e));
}

#define SCHK(expr)                                                                                      \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return hfail(e_, #expr);                                                  \
    } while (0)

} // namespace

struct vbt_server_s {
    vbt_server_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr, fetchStream = nullptr;
    hipEvent_t loopDone = nullptr;
    uint32_t grid = 0;
    double ticksPerUs = 100.0;

    // resident assets
    std::vector<vbt_asset> assets;
    std::vector<uint64_t> assetCpf;
    std::vector<uint32_t> assetProfile, assetActive, assetFrames, assetBad;
    std::vector<VbtFrameIndex> assetIndex; // the server's frame index of each asset (empty: none)
    uint64_t assetBytes = 0;
    AssetDev* dAssets = nullptr;

    // memory
    char* dArena = nullptr;
    Arena arena;
    uint8_t* hStage = nullptr;  // pinned, device-mapped
    uint8_t* dStage = nullptr;
    Arena stage;
    HostShared* hs = nullptr;
    HostShared* dHs = nullptr;
    RingEntry* ring = nullptr;
    RingEntry* dRing = nullptr;
    uint32_t* doneHost = nullptr;
    uint32_t* dDoneHost = nullptr;
    DevState* dState = nullptr;
    DevReq* dMirror = nullptr;
    uint32_t* dLists = nullptr;

    // dispatcher state (guarded by mu)
    std::mutex mu;
    std::condition_variable cv;       // dispatcher wake-up
    std::condition_variable doneCv;   // pollers, registration
    std::condition_variable stageCv;  // submitters waiting for staging space
    std::deque<std::shared_ptr<Req>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
    std::vector<std::shared_ptr<Req>> inflight;  // admitted, not complete
    bool loopRunning = false;
    uint32_t ringTail = 0;
    // whole-frame builds launched as one-shot kernels beside the loop
    hipStream_t frameStream = nullptr;
    std::vector<hipEvent_t> eventPool;
    struct FrameBuild {
        std::shared_ptr<Req> r;
        hipEvent_t done;
    };
    std::vector<FrameBuild> frames;
    uint64_t launchSeq = 0;
    uint64_t nextId = 1;
    bool stopping = false, pauseLoop = false, paused = false;
    std::thread dispatcher;
    std::string fatal;
    std::mutex fetchMu;

    // statistics
    uint64_t submitted = 0, completed = 0, rejected = 0, launches = 0, values = 0, requeued = 0, maxDepth = 0,
             maxInflightSeen = 0;
    std::vector<double> lat[2];

    void run();
    void complete(const std::shared_ptr<Req>& r, int32_t status, Clock::time_point now);
    bool launch_loop();
    void end_of_launch();
    void admit();
    bool launch_frame(const std::shared_ptr<Req>& r);
    void poll_frames(Clock::time_point now);
};

// A whole-frame build as the one-shot frame-cache kernel on the frame stream
// (it runs beside the device loop on the CUs the loop leaves free); its event
// marks completion.  False when the launch itself failed.
bool vbt_server_s::launch_frame(const std::shared_ptr<Req>& r)
{
    hipEvent_t ev = nullptr;
    if (!eventPool.empty()) {
        ev = eventPool.back();
        eventPool.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        return false;
    }
    float* out = reinterpret_cast<float*>(dArena + r->devOff);
    const VbtFrameIndex& ix = assetIndex[r->asset];
    const int rc = ix.off ? vbt_internal_frame_cache_indexed(assets[r->asset], ix, r->t, 1, out, frameStream)
                          : vbt_decode_frame_cache(assets[r->asset], r->t, 1, VBT_CACHE_PACKED, out, frameStream);
    // VBT_ERR_FORMAT only reports poisoned leaves after launching (complete()
    // turns it into the request's status)
    if ((rc != VBT_OK && rc != VBT_ERR_FORMAT) || hipEventRecord(ev, frameStream) != hipSuccess) {
        eventPool.push_back(ev);
        return false;
    }
    frames.push_back(FrameBuild{r, ev});
    return true;
}

void vbt_server_s::poll_frames(Clock::time_point now)
{
    size_t w = 0;
    bool any = false;
    for (size_t i = 0; i < frames.size(); ++i) {
        const hipError_t e = hipEventQuery(frames[i].done);
        if (e == hipErrorNotReady) {
            frames[w++] = frames[i];
            continue;
        }
        complete(frames[i].r, e == hipSuccess ? VBT_OK : VBT_ERR_DEVICE, now);
        eventPool.push_back(frames[i].done);
        any = true;
    }
    frames.resize(w);
    if (any) doneCv.notify_all();
}

void vbt_server_s::complete(const std::shared_ptr<Req>& r, int32_t status, Clock::time_point now)
{
    // a whole-frame request on an asset with poisoned leaves fails as a whole
    // (those leaves own no packed-cache slots), as vbt_decode_frame_cache does
    if (status == VBT_OK && r->type == REQ_CACHE && assetBad[r->asset]) status = VBT_ERR_FORMAT;
    r->status = status;
    r->finished = now;
    r->state = 2;
    if (r->hasStage) {
        stage.release(r->stageOff, r->stageBytes);
        r->hasStage = false;
        stageCv.notify_all();
    }
    lat[r->type].push_back(std::chrono::duration<double, std::micro>(now - r->submitted).count());
    ++completed;
    values += r->n;
}

bool vbt_server_s::launch_loop()
{
    ServeArgs A;
    A.hs = dHs;
    A.ring = dRing;
    A.doneHost = dDoneHost;
    A.st = dState;
    A.mirror = dMirror;
    A.lists = dLists;
    A.assets = dAssets;
    A.idleTicks = (uint64_t)(cfg.idleUs * ticksPerUs);
    A.leaseTicks = (uint64_t)(cfg.leaseMs * 1000.0 * ticksPerUs);
    A.workerTicks = A.leaseTicks + A.idleTicks + (uint64_t)(1e6 * ticksPerUs);
    A.queryItem = cfg.queryItem;
    A.cacheItem = cfg.cacheItem;
    k_clear_state<<<1, 64, 0, stream>>>(dState);
    k_serve<<<grid, SERVE_BLOCK, 0, stream>>>(A);
    if (hipGetLastError() != hipSuccess) return false;
    if (hipEventRecord(loopDone, stream) != hipSuccess) return false;
    loopRunning = true;
    ++launchSeq;
    ++launches;
    return true;
}

// A launch has ended: complete what it finished and requeue the rest (in ring
// order, ahead of everything still queued).
void vbt_server_s::end_of_launch()
{
    const auto now = Clock::now();
    std::vector<std::shared_ptr<Req>> again;
    for (auto& r : inflight) {
        const uint32_t f = __atomic_load_n(&doneHost[r->ring], __ATOMIC_ACQUIRE);
        if (f & 1u) complete(r, (f >> 1) ? VBT_ERR_FORMAT : VBT_OK, now);
        else again.push_back(r);
    }
    std::sort(again.begin(), again.end(), [](const auto& x, const auto& y) { return x->ring < y->ring; });
    for (auto it = again.rbegin(); it != again.rend(); ++it) {
        (*it)->state = 0;
        queue.push_front(*it);
        ++requeued;
    }
    inflight.clear();
    loopRunning = false;
    ringTail = 0;
    doneCv.notify_all();
}

// Admission in arrival order while the budget, the arena, the ring and the
// in-flight limit allow; launches the device loop when none runs.
void vbt_server_s::admit()
{
    const uint32_t first = ringTail;
    while (!queue.empty() && ringTail < RING && inflight.size() + frames.size() < cfg.maxInflight) {
        auto r = queue.front();
        const uint64_t need = out_bytes(*r);
        if (!r->hasDev && need) {
            if (assetBytes + arena.used + need > cfg.hbmBudgetBytes) break;
            if (!arena.alloc(need, r->devOff)) break;
            r->hasDev = true;
            r->devBytes = need;
        }
        queue.pop_front();
        if (r->items == 0) {
            complete(r, VBT_OK, Clock::now());
            doneCv.notify_all();
            continue;
        }
        if (r->type == REQ_CACHE && !cfg.frameBuildsInLoop) {
            r->state = 1;
            r->launch = 0;
            if (!launch_frame(r)) {
                complete(r, VBT_ERR_DEVICE, Clock::now());
                doneCv.notify_all();
            }
            continue;
        }
        RingEntry& e = ring[ringTail];
        e.type = r->type;
        e.asset = r->asset;
        e.t = r->t;
        e.n = r->n;
        e.q = r->type == REQ_QUERY ? (uint64_t)(dStage + r->stageOff) : 0u;
        e.out = (uint64_t)(dArena + r->devOff);
        e.status = r->type == REQ_QUERY ? e.out + Arena::round(4ull * r->n) : 0u;
        e.itemCount = r->items;
        e.pad = 0;
        e.cm = r->cm;
        doneHost[ringTail] = 0u;
        r->ring = ringTail++;
        r->state = 1;
        r->launch = loopRunning ? launchSeq : launchSeq + 1u;
        inflight.push_back(r);
    }
    maxInflightSeen = std::max<uint64_t>(maxInflightSeen, inflight.size() + frames.size());
    if (ringTail == first) return;
    if (!loopRunning) {
        __atomic_store_n(&hs->stop, 0u, __ATOMIC_RELAXED);
        __atomic_store_n(&hs->head, 0u, __ATOMIC_RELAXED);
        __atomic_store_n(&hs->exited, 0u, __ATOMIC_RELAXED);
    }
    __atomic_store_n(&hs->tail, ringTail, __ATOMIC_RELEASE);
    if (!loopRunning && !launch_loop()) {
        fatal = "device loop launch failed";
        const auto now = Clock::now();
        for (auto& r : inflight) complete(r, VBT_ERR_DEVICE, now);
        inflight.clear();
        ringTail = 0;
        doneCv.notify_all();
    }
}

void vbt_server_s::run()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // a finished launch
        if (loopRunning && hipEventQuery(loopDone) == hipSuccess) end_of_launch();
        if (!frames.empty()) poll_frames(Clock::now());
        // completions
        if (!inflight.empty()) {
            const auto now = Clock::now();
            size_t w = 0;
            for (size_t i = 0; i < inflight.size(); ++i) {
                auto& r = inflight[i];
                const uint32_t f = __atomic_load_n(&doneHost[r->ring], __ATOMIC_ACQUIRE);
                if (f & 1u) complete(r, (f >> 1) ? VBT_ERR_FORMAT : VBT_OK, now);
                else inflight[w++] = r;
            }
            if (w != inflight.size()) {
                inflight.resize(w);
                doneCv.notify_all();
            }
        }
        if (stopping) {
            if (!loopRunning && frames.empty()) break;
            __atomic_store_n(&hs->stop, 1u, __ATOMIC_RELEASE);
        } else if (pauseLoop) {
            // asset registration: end the loop, then wait until resumed
            if (loopRunning) {
                __atomic_store_n(&hs->stop, 1u, __ATOMIC_RELEASE);
            } else {
                paused = true;
                doneCv.notify_all();
                cv.wait(lk, [&] { return !pauseLoop || stopping; });
                paused = false;
                continue;
            }
        } else if (fatal.empty() && !(loopRunning && __atomic_load_n(&hs->exited, __ATOMIC_ACQUIRE))) {
            admit();
        }
        if (!loopRunning && frames.empty() && (queue.empty() || !fatal.empty()) && !stopping && !pauseLoop) {
            cv.wait(lk);
        } else {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
}

extern "C" {

void vbt_server_default_config(vbt_server_config* c)
{
    if (!c) return;
    c->hbmBudgetBytes = 80ull << 30;
    c->arenaBytes = 4ull << 30;
    c->stagingBytes = 256ull << 20;
    c->maxInflight = 4096;
    c->blocksPerCU = 2;
    c->queryItem = 256; // mixed six-asset trace at saturation: 5.5 x the per-launch decode vs 5.3 x with 128
    c->cacheItem = 1024;
    c->idleUs = 2000;
    c->leaseMs = 250;
    c->frameBuildsInLoop = 0;
    c->frameIndex = 1;
}

int vbt_server_destroy(vbt_server s);

int vbt_server_create(const vbt_server_config* c, vbt_server* out)
{
    if (!out) return vbt_fail(VBT_ERR_INVALID, "null output");
    *out = nullptr;
    vbt_server_config cfg;
    vbt_server_default_config(&cfg);
    if (c) cfg = *c;
    if (cfg.maxInflight == 0 || cfg.maxInflight > RING) return vbt_fail(VBT_ERR_INVALID, "maxInflight must be 1..65536");
    if (cfg.queryItem == 0 || cfg.queryItem % QUERY_PASS || cfg.queryItem > 65536)
        return vbt_fail(VBT_ERR_INVALID, "queryItem must be a positive multiple of " + std::to_string(QUERY_PASS) +
                                             ", at most 65536");
    if (cfg.cacheItem == 0 || cfg.cacheItem % SERVE_BLOCK || cfg.cacheItem > (1u << 24))
        return vbt_fail(VBT_ERR_INVALID, "cacheItem must be a positive multiple of " + std::to_string(SERVE_BLOCK) +
                                             ", at most 2^24");
    if (cfg.blocksPerCU == 0 || cfg.blocksPerCU > 8) return vbt_fail(VBT_ERR_INVALID, "blocksPerCU must be 1..8");
    if (cfg.idleUs == 0 || cfg.leaseMs == 0 || cfg.leaseMs > 60000 || cfg.idleUs > 10000000)
        return vbt_fail(VBT_ERR_INVALID, "idleUs (at most 10 s) and leaseMs (at most 60000) must be positive");
    if (cfg.arenaBytes > cfg.hbmBudgetBytes) cfg.arenaBytes = cfg.hbmBudgetBytes;
    if (cfg.arenaBytes < 4096 || cfg.stagingBytes < 4096) return vbt_fail(VBT_ERR_INVALID, "arena or staging too small");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return vbt_fail(VBT_ERR_NO_DEVICE, "no HIP device");

    auto destroy = [](vbt_server_s* p) { vbt_server_destroy(p); };
    std::unique_ptr<vbt_server_s, decltype(destroy)> s(new vbt_server_s(), destroy);
    s->cfg = cfg;
    SCHK(hipGetDevice(&s->device));
    int cus = 256, rateKHz = 100000;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device);
    if (hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, s->device) != hipSuccess || rateKHz <= 0)
        rateKHz = 100000;
    s->ticksPerUs = rateKHz / 1000.0;
    int occ = 0;
    SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_serve, SERVE_BLOCK, 0));
    // workers wait only on a scheduler that is running: keep the grid co-resident
    const uint32_t per = std::min<uint32_t>(cfg.blocksPerCU, occ > 0 ? (uint32_t)occ : 1u);
    s->grid = 1u + per * (uint32_t)cus;
    if (occ > 0 && s->grid > (uint32_t)occ * (uint32_t)cus) s->grid = (uint32_t)occ * (uint32_t)cus;

    SCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    SCHK(hipStreamCreateWithFlags(&s->fetchStream, hipStreamNonBlocking));
    SCHK(hipStreamCreateWithFlags(&s->frameStream, hipStreamNonBlocking));
    SCHK(hipEventCreateWithFlags(&s->loopDone, hipEventDisableTiming));
    SCHK(hipMalloc((void**)&s->dAssets, sizeof(AssetDev) * MAX_ASSETS));
    SCHK(hipMalloc((void**)&s->dArena, cfg.arenaBytes));
    s->arena.init(cfg.arenaBytes);
    SCHK(hipMalloc((void**)&s->dState, sizeof(DevState)));
    SCHK(hipMalloc((void**)&s->dMirror, sizeof(DevReq) * RING));
    SCHK(hipMalloc((void**)&s->dLists, sizeof(uint32_t) * 2u * RING));
    const unsigned hostFlags = hipHostMallocMapped | hipHostMallocCoherent;
    SCHK(hipHostMalloc((void**)&s->hStage, cfg.stagingBytes, hostFlags));
    SCHK(hipHostGetDevicePointer((void**)&s->dStage, s->hStage, 0));
    s->stage.init(cfg.stagingBytes);
    SCHK(hipHostMalloc((void**)&s->hs, sizeof(HostShared), hostFlags));
    SCHK(hipHostGetDevicePointer((void**)&s->dHs, s->hs, 0));
    SCHK(hipHostMalloc((void**)&s->ring, sizeof(RingEntry) * RING, hostFlags));
    SCHK(hipHostGetDevicePointer((void**)&s->dRing, s->ring, 0));
    SCHK(hipHostMalloc((void**)&s->doneHost, sizeof(uint32_t) * RING, hostFlags));
    SCHK(hipHostGetDevicePointer((void**)&s->dDoneHost, s->doneHost, 0));
    std::memset(s->hs, 0, sizeof(HostShared));
    std::memset(s->doneHost, 0, sizeof(uint32_t) * RING);
    vbt_server_s* raw = s.get();
    raw->dispatcher = std::thread([raw] { raw->run(); });
    *out = s.release();
    return VBT_OK;
}

int vbt_server_destroy(vbt_server s)
{
    if (!s) return VBT_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
    }
    s->cv.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->fetchStream) (void)hipStreamSynchronize(s->fetchStream);
    if (s->frameStream) (void)hipStreamSynchronize(s->frameStream);
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (auto& kv : s->live)
            if (kv.second->state != 2) {
                kv.second->state = 2;
                kv.second->status = VBT_ERR_DEVICE;
            }
    }
    s->doneCv.notify_all();
    if (s->loopDone) (void)hipEventDestroy(s->loopDone);
    if (s->hStage) (void)hipHostFree(s->hStage);
    if (s->hs) (void)hipHostFree(s->hs);
    if (s->ring) (void)hipHostFree(s->ring);
    if (s->doneHost) (void)hipHostFree(s->doneHost);
    if (s->dArena) (void)hipFree(s->dArena);
    if (s->dAssets) (void)hipFree(s->dAssets);
    if (s->dState) (void)hipFree(s->dState);
    if (s->dMirror) (void)hipFree(s->dMirror);
    if (s->dLists) (void)hipFree(s->dLists);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->fetchStream) (void)hipStreamDestroy(s->fetchStream);
    if (s->frameStream) (void)hipStreamDestroy(s->frameStream);
    for (auto& f : s->frames) (void)hipEventDestroy(f.done);
    for (hipEvent_t e : s->eventPool) (void)hipEventDestroy(e);
    for (auto& ix : s->assetIndex) vbt_internal_frame_index_free(ix); // every stream that read them has drained
    delete s;
    return VBT_OK;
}

int vbt_server_register(vbt_server s, vbt_asset a, int* assetId)
{
    if (!s || !a || !assetId) return vbt_fail(VBT_ERR_INVALID, "null server, asset or output");
    if (vbt_internal_asset_device(a) != s->device) return vbt_fail(VBT_ERR_INVALID, "asset lives on another device");
    vbt_asset_info info;
    int rc = vbt_asset_info_get(a, &info);
    if (rc) return rc;
    uint64_t bytes = vbt_internal_asset_bytes(a);
    const AssetDev ad = vbt_internal_asset_dev(a);
    // the server's own frame index of the asset (built before the loop pauses:
    // it only reads the asset)
    VbtFrameIndex ix;
    if (s->cfg.frameIndex && info.header.profileType != 2 && !vbt_internal_asset_bad(a)) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (s->assets.size() >= MAX_ASSETS) return vbt_fail(VBT_ERR_INVALID, "too many assets (at most 32)");
            if (s->assetBytes + bytes + s->arena.used > s->cfg.hbmBudgetBytes)
                return vbt_fail(VBT_ERR_BUDGET, "asset does not fit in the HBM budget");
        }
        if ((rc = vbt_internal_frame_index_build(a, ix, s->fetchStream))) return rc;
        bytes += ix.bytes;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto drop = [&](int code) {
        vbt_internal_frame_index_free(ix);
        return code;
    };
    if (s->assets.size() >= MAX_ASSETS) return drop(vbt_fail(VBT_ERR_INVALID, "too many assets (at most 32)"));
    if (s->assetBytes + bytes + s->arena.used > s->cfg.hbmBudgetBytes)
        return drop(vbt_fail(VBT_ERR_BUDGET, "asset does not fit in the HBM budget"));
    // the asset table changes only while no device loop runs
    s->pauseLoop = true;
    s->cv.notify_all();
    s->doneCv.wait(lk, [&] { return s->paused || s->stopping; });
    const int id = (int)s->assets.size();
    hipError_t e = hipMemcpyAsync(s->dAssets + id, &ad, sizeof(ad), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) {
        s->assets.push_back(a);
        s->assetCpf.push_back(info.controlsPerFrame);
        s->assetProfile.push_back(info.header.profileType);
        s->assetActive.push_back(info.activeLeaves);
        s->assetFrames.push_back(info.header.frames);
        s->assetBad.push_back(vbt_internal_asset_bad(a));
        s->assetIndex.push_back(ix);
        ix = VbtFrameIndex{};
        s->assetBytes += bytes;
        *assetId = id;
    }
    s->pauseLoop = false;
    s->cv.notify_all();
    if (e != hipSuccess) return drop(hfail(e, "asset table upload"));
    return VBT_OK;
}

static int enqueue(vbt_server s, const std::shared_ptr<Req>& r, const uint32_t* queries, uint64_t* id)
{
    r->submitted = Clock::now();  // latency covers the staging copy too
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->fatal.empty()) return vbt_fail(VBT_ERR_DEVICE, s->fatal);
    if (s->stopping) return vbt_fail(VBT_ERR_INVALID, "server is shutting down");
    const uint64_t need = out_bytes(*r);
    if (need > s->arena.size || s->assetBytes + need > s->cfg.hbmBudgetBytes || r->stageBytes > s->stage.size) {
        ++s->rejected;
        return vbt_fail(VBT_ERR_BUDGET, "request buffers exceed the HBM budget, the arena or the staging area");
    }
    if (r->stageBytes) {
        const auto deadline = Clock::now() + std::chrono::seconds(1);
        while (!s->stage.alloc(r->stageBytes, r->stageOff)) {
            if (s->stageCv.wait_until(lk, deadline) == std::cv_status::timeout) {
                if (s->stage.alloc(r->stageBytes, r->stageOff)) break;
                ++s->rejected;
                return vbt_fail(VBT_ERR_AGAIN, "query staging area is full");
            }
        }
        r->hasStage = true;
        lk.unlock();
        std::memcpy(s->hStage + r->stageOff, queries, r->stageBytes);
        lk.lock();
        if (s->stopping) {
            s->stage.release(r->stageOff, r->stageBytes);
            return vbt_fail(VBT_ERR_INVALID, "server is shutting down");
        }
    }
    r->id = s->nextId++;
    s->queue.push_back(r);
    s->live[r->id] = r;
    ++s->submitted;
    s->maxDepth = std::max<uint64_t>(s->maxDepth, s->queue.size());
    *id = r->id;
    s->cv.notify_all();
    return VBT_OK;
}

int vbt_server_submit_query(vbt_server s, int assetId, const uint32_t* queries, uint32_t n, const vbt_compose* c,
                            uint64_t* requestId)
{
    if (!s || !queries || !requestId || n == 0 || assetId < 0) return vbt_fail(VBT_ERR_INVALID, "invalid query request");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if ((size_t)assetId >= s->assets.size()) return vbt_fail(VBT_ERR_INVALID, "unknown asset id");
    }
    auto r = std::make_shared<Req>();
    r->type = REQ_QUERY;
    r->asset = (uint32_t)assetId;
    r->n = n;
    r->cm = ComposeDev{VBT_SIGN_NONE, 0.0f, 1.0f, NAN};
    if (c) r->cm = ComposeDev{c->signRule, c->background, c->shellEmptyScale, c->shellEmptyMax};
    r->items = (uint32_t)(((uint64_t)n + s->cfg.queryItem - 1u) / s->cfg.queryItem);
    r->stageBytes = 16ull * n;
    return enqueue(s, r, queries, requestId);
}

int vbt_server_submit_cache(vbt_server s, int assetId, int t, uint64_t* requestId)
{
    if (!s || !requestId || assetId < 0) return vbt_fail(VBT_ERR_INVALID, "invalid cache request");
    auto r = std::make_shared<Req>();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if ((size_t)assetId >= s->assets.size()) return vbt_fail(VBT_ERR_INVALID, "unknown asset id");
        if (t < 0 || (uint32_t)t >= s->assetFrames[assetId]) return vbt_fail(VBT_ERR_INVALID, "frame out of range");
        const uint64_t cpf = s->assetCpf[assetId];
        if (cpf > 0xFFFFFFFFull) return vbt_fail(VBT_ERR_INVALID, "frame cache too large for one request");
        r->n = (uint32_t)cpf;
        r->items = s->assetProfile[assetId] == 2 ? (s->assetActive[assetId] + SCI_LEAVES - 1u) / SCI_LEAVES
                                                 : (uint32_t)((cpf + s->cfg.cacheItem - 1u) / s->cfg.cacheItem);
    }
    r->type = REQ_CACHE;
    r->asset = (uint32_t)assetId;
    r->t = t;
    r->cm = ComposeDev{VBT_SIGN_NONE, 0.0f, 1.0f, NAN};
    return enqueue(s, r, nullptr, requestId);
}

int vbt_server_poll(vbt_server s, uint64_t id, int timeoutMs, vbt_request_result* out)
{
    if (!s || !out) return vbt_fail(VBT_ERR_INVALID, "null server or output");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) return vbt_fail(VBT_ERR_INVALID, "unknown or released request");
    std::shared_ptr<Req> r = it->second;
    auto ready = [&] { return r->state == 2; };
    if (timeoutMs < 0) s->doneCv.wait(lk, ready);
    else if (!s->doneCv.wait_for(lk, std::chrono::milliseconds(timeoutMs), ready))
        return vbt_fail(VBT_ERR_AGAIN, "request not complete");
    out->id = r->id;
    out->status = r->status;
    out->count = r->n;
    const char* base = s->dArena + r->devOff;
    out->d_values = r->hasDev ? (const float*)base : nullptr;
    out->d_status = r->hasDev && r->type == REQ_QUERY ? (const int32_t*)(base + Arena::round(4ull * r->n)) : nullptr;
    out->latencyUs = std::chrono::duration<double, std::micro>(r->finished - r->submitted).count();
    out->launch = r->launch;
    return VBT_OK;
}

int vbt_server_release(vbt_server s, uint64_t id)
{
    if (!s) return vbt_fail(VBT_ERR_INVALID, "null server");
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end()) return vbt_fail(VBT_ERR_INVALID, "unknown or released request");
        if (it->second->state != 2) return vbt_fail(VBT_ERR_AGAIN, "request not complete");
        if (it->second->hasDev) s->arena.release(it->second->devOff, it->second->devBytes);
        it->second->hasDev = false;
        s->live.erase(it);
    }
    s->cv.notify_all();  // arena space may unblock admission
    return VBT_OK;
}

int vbt_server_fetch(vbt_server s, uint64_t id, float* values, int32_t* status)
{
    vbt_request_result r;
    int rc = vbt_server_poll(s, id, -1, &r);
    if (rc) return rc;
    if (r.status == VBT_ERR_DEVICE) {
        (void)vbt_server_release(s, id);
        return vbt_fail(VBT_ERR_DEVICE, "request failed on the device");
    }
    if (r.d_values) {
        std::lock_guard<std::mutex> g(s->fetchMu);
        hipError_t e = hipSuccess;
        if (values) e = hipMemcpyAsync(values, r.d_values, 4ull * r.count, hipMemcpyDeviceToHost, s->fetchStream);
        if (e == hipSuccess && status && r.d_status)
            e = hipMemcpyAsync(status, r.d_status, 4ull * r.count, hipMemcpyDeviceToHost, s->fetchStream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->fetchStream);
        if (e != hipSuccess) return hfail(e, "result copy");
    } else if (status) {
        std::memset(status, 0, 4ull * r.count);
    }
    rc = vbt_server_release(s, id);
    if (rc) return rc;
    if (r.status) return vbt_fail(r.status, "an element of the request failed (see the statuses)");
    return VBT_OK;
}

static double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

int vbt_server_stats_get(vbt_server s, int type, vbt_server_stats* o)
{
    if (!s || !o) return vbt_fail(VBT_ERR_INVALID, "null server or output");
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(o, 0, sizeof(*o));
    std::vector<double> v;
    if (type < 0 || type == 0) v.insert(v.end(), s->lat[0].begin(), s->lat[0].end());
    if (type < 0 || type == 1) v.insert(v.end(), s->lat[1].begin(), s->lat[1].end());
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->rejected = s->rejected;
    o->launches = s->launches;
    o->values = s->values;
    o->requeued = s->requeued;
    o->p50Us = percentile(v, 0.50);
    o->p99Us = percentile(v, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double x : v) {
        sum += x;
        mx = std::max(mx, x);
    }
    o->meanUs = v.empty() ? 0.0 : sum / (double)v.size();
    o->maxUs = mx;
    o->maxQueueDepth = s->maxDepth;
    o->maxInflight = s->maxInflightSeen;
    o->hbmInUse = s->assetBytes + s->arena.used;
    o->hbmBudget = s->cfg.hbmBudgetBytes;
    o->arenaInUse = s->arena.used;
    return VBT_OK;
}

int vbt_server_reset_stats(vbt_server s)
{
    if (!s) return vbt_fail(VBT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    s->submitted = s->completed = s->rejected = s->launches = s->values = s->requeued = s->maxDepth =
        s->maxInflightSeen = 0;
    s->lat[0].clear();
    s->lat[1].clear();
    return VBT_OK;
}

} // extern "C"
