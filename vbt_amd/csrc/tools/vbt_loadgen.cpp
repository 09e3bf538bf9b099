// vbt_loadgen.cpp -- open-loop load generator for the decode server and for
// the one-request-per-launch path, over the same seeded request trace.
//
// Trace: requests arrive as a Poisson process (exponential gaps; rate 0 puts
// every arrival at t = 0, a saturation run).  Each request is a whole-frame
// packed cache build with probability cacheFraction, else a point-query batch
// of log-uniform size in 1..maxQueries with uniform (x, y, z, t); the asset is
// uniform among those given.  Latency is measured from the request's arrival
// to the completion the serving side observes (results in HBM).
//
// Modes: 0 submits to a vbt_server; 1 runs each request as its own launch
// (vbt_query_points_plain, one thread per query; vbt_decode_frame_cache for
// frames) on `workers` streams in arrival order; 2 is mode 1 with the batched
// query kernel (vbt_query_points).
#include "../../../include/vbt_hip.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

extern "C" {

typedef struct {
    uint32_t nAssets;
    const int32_t* serverIds;  // mode 0: server asset ids
    const vbt_asset* assets;   // asset handles (dims; modes 1 and 2 decode with them)
    double rate;               // mean arrivals per second; 0: all arrive at once
    uint32_t requests;
    double cacheFraction;
    uint32_t maxQueries;
    uint64_t seed;
    int32_t mode;
    uint32_t workers;          // modes 1 and 2: concurrent streams
} vbtl_config;

typedef struct {
    double seconds;            // first arrival -> last completion
    uint64_t requests, queryRequests, cacheRequests, queries, cacheValues, failed;
    double p50Us, p99Us, meanUs, maxUs;  // all requests
    double qp50Us, qp99Us;               // query batches
    double cp50Us, cp99Us;               // frame-cache builds
    double queriesPerSec, requestsPerSec, offeredPerSec;
} vbtl_result;

int vbtl_run(vbt_server srv, const vbtl_config* c, vbtl_result* out);
}

namespace {

using Clock = std::chrono::steady_clock;

struct Item {
    double arrivalUs;
    int kind;  // 0 queries, 1 cache
    uint32_t asset, n;
    int t;
    uint64_t qoff;  // in queries
};

double pct(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

double u01(std::mt19937_64& g) { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

int vbtl_run(vbt_server srv, const vbtl_config* c, vbtl_result* out)
{
    if (!c || !out || c->nAssets == 0 || !c->assets || c->requests == 0 || c->maxQueries == 0) return VBT_ERR_INVALID;
    if (c->mode == 0 && (!srv || !c->serverIds)) return VBT_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    std::vector<vbt_asset_info> info(c->nAssets);
    for (uint32_t a = 0; a < c->nAssets; ++a)
        if (int rc = vbt_asset_info_get(c->assets[a], &info[a])) return rc;

    // the trace
    std::mt19937_64 g(c->seed);
    std::vector<Item> items(c->requests);
    double at = 0.0;
    uint64_t totalQ = 0;
    for (auto& it : items) {
        if (c->rate > 0.0) at += -std::log(1.0 - u01(g)) / c->rate * 1e6;
        it.arrivalUs = at;
        it.asset = (uint32_t)(g() % c->nAssets);
        const auto& h = info[it.asset].header;
        it.kind = u01(g) < c->cacheFraction ? 1 : 0;
        it.t = (int)(g() % h.frames);
        it.n = it.kind ? 0u : std::min<uint32_t>(c->maxQueries,
                                                  std::max<uint32_t>(1u, (uint32_t)std::exp(u01(g) * std::log((double)c->maxQueries + 1.0))));
        it.qoff = totalQ;
        totalQ += it.n;
    }
    uint32_t* qbuf = nullptr;
    if (hipHostMalloc((void**)&qbuf, std::max<uint64_t>(16, 16 * totalQ), hipHostMallocDefault) != hipSuccess)
        return VBT_ERR_NOMEM;
    for (auto& it : items) {
        const auto& h = info[it.asset].header;
        for (uint32_t i = 0; i < it.n; ++i) {
            uint32_t* q = qbuf + 4 * (it.qoff + i);
            q[0] = (uint32_t)(g() % h.width);
            q[1] = (uint32_t)(g() % h.height);
            q[2] = (uint32_t)(g() % h.depth);
            q[3] = (uint32_t)(g() % h.frames);
        }
    }

    const uint32_t R = c->requests;
    std::vector<double> lat(R, 0.0), doneAt(R, 0.0);
    std::vector<char> failed(R, 0);
    const auto t0 = Clock::now() + std::chrono::milliseconds(2);
    auto usSince = [&](Clock::time_point t) { return std::chrono::duration<double, std::micro>(t - t0).count(); };
    auto waitArrival = [&](const Item& it) {
        const auto due = t0 + std::chrono::nanoseconds((int64_t)(it.arrivalUs * 1000.0));
        while (Clock::now() < due) {
        }
    };
    int rcAll = VBT_OK;

    if (c->mode == 0) {
        std::vector<std::atomic<uint64_t>> ids(R);
        std::vector<double> submitAt(R, 0.0);
        for (auto& x : ids) x.store(0);
        std::thread reaper([&] {
            for (uint32_t i = 0; i < R; ++i) {
                uint64_t id;
                while ((id = ids[i].load(std::memory_order_acquire)) == 0) std::this_thread::yield();
                if (id == ~uint64_t(0)) continue;
                vbt_request_result res;
                if (vbt_server_poll(srv, id, -1, &res) != VBT_OK) {
                    failed[i] = 1;
                    continue;
                }
                failed[i] = res.status != VBT_OK;
                doneAt[i] = submitAt[i] + res.latencyUs;
                lat[i] = doneAt[i] - items[i].arrivalUs;
                vbt_server_release(srv, id);
            }
        });
        for (uint32_t i = 0; i < R; ++i) {
            const Item& it = items[i];
            waitArrival(it);
            submitAt[i] = usSince(Clock::now());
            uint64_t id = 0;
            const int rc = it.kind ? vbt_server_submit_cache(srv, c->serverIds[it.asset], it.t, &id)
                                   : vbt_server_submit_query(srv, c->serverIds[it.asset], qbuf + 4 * it.qoff, it.n,
                                                             nullptr, &id);
            ids[i].store(rc == VBT_OK ? id : ~uint64_t(0), std::memory_order_release);
            if (rc != VBT_OK) failed[i] = 1;
        }
        reaper.join();
    } else {
        // one request per launch, in arrival order, on `workers` streams
        std::mutex mu;
        std::condition_variable cv;
        std::deque<uint32_t> fifo;
        bool closed = false;
        uint64_t maxCpf = 1;
        for (auto& in : info) maxCpf = std::max<uint64_t>(maxCpf, in.controlsPerFrame);
        const uint32_t W = std::max<uint32_t>(1u, c->workers);
        std::vector<std::thread> th;
        std::atomic<int> err{VBT_OK};
        for (uint32_t w = 0; w < W; ++w) {
            th.emplace_back([&] {
                hipStream_t st;
                void *dq = nullptr, *dout = nullptr, *dst = nullptr, *dcache = nullptr;
                if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                    hipMalloc(&dq, 16ull * c->maxQueries) != hipSuccess ||
                    hipMalloc(&dout, 4ull * c->maxQueries) != hipSuccess ||
                    hipMalloc(&dst, 4ull * c->maxQueries) != hipSuccess ||
                    hipMalloc(&dcache, 4ull * maxCpf) != hipSuccess) {
                    err = VBT_ERR_NOMEM;
                }
                for (;;) {
                    uint32_t i;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [&] { return closed || !fifo.empty(); });
                        if (fifo.empty()) break;
                        i = fifo.front();
                        fifo.pop_front();
                    }
                    const Item& it = items[i];
                    int rc = err.load();
                    if (rc == VBT_OK) {
                        if (it.kind) {
                            rc = vbt_decode_frame_cache(c->assets[it.asset], it.t, 1, VBT_CACHE_PACKED, (float*)dcache,
                                                        st);
                        } else {
                            rc = hipMemcpyAsync(dq, qbuf + 4 * it.qoff, 16ull * it.n, hipMemcpyHostToDevice, st) ==
                                         hipSuccess
                                     ? VBT_OK
                                     : VBT_ERR_DEVICE;
                            if (rc == VBT_OK)
                                rc = (c->mode == 1 ? vbt_query_points_plain : vbt_query_points)(
                                    c->assets[it.asset], (const uint32_t*)dq, it.n, nullptr, (float*)dout,
                                    (int32_t*)dst, st);
                        }
                        if (rc == VBT_OK && hipStreamSynchronize(st) != hipSuccess) rc = VBT_ERR_DEVICE;
                    }
                    doneAt[i] = usSince(Clock::now());
                    lat[i] = doneAt[i] - it.arrivalUs;
                    failed[i] = rc != VBT_OK;
                }
                if (dq) (void)hipFree(dq);
                if (dout) (void)hipFree(dout);
                if (dst) (void)hipFree(dst);
                if (dcache) (void)hipFree(dcache);
                (void)hipStreamDestroy(st);
            });
        }
        for (uint32_t i = 0; i < R; ++i) {
            waitArrival(items[i]);
            {
                std::lock_guard<std::mutex> lk(mu);
                fifo.push_back(i);
            }
            cv.notify_one();
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            closed = true;
        }
        cv.notify_all();
        for (auto& t : th) t.join();
        rcAll = err.load();
    }
    (void)hipHostFree(qbuf);

    std::vector<double> all, ql, cl;
    double end = 0.0;
    for (uint32_t i = 0; i < R; ++i) {
        const Item& it = items[i];
        out->failed += failed[i] ? 1u : 0u;
        end = std::max(end, doneAt[i]);
        all.push_back(lat[i]);
        if (it.kind) {
            ++out->cacheRequests;
            out->cacheValues += info[it.asset].controlsPerFrame;
            cl.push_back(lat[i]);
        } else {
            ++out->queryRequests;
            out->queries += it.n;
            ql.push_back(lat[i]);
        }
    }
    out->requests = R;
    out->seconds = end * 1e-6;
    out->p50Us = pct(all, 0.5);
    out->p99Us = pct(all, 0.99);
    double s = 0.0, mx = 0.0;
    for (double x : all) {
        s += x;
        mx = std::max(mx, x);
    }
    out->meanUs = s / (double)R;
    out->maxUs = mx;
    out->qp50Us = pct(ql, 0.5);
    out->qp99Us = pct(ql, 0.99);
    out->cp50Us = pct(cl, 0.5);
    out->cp99Us = pct(cl, 0.99);
    out->queriesPerSec = out->seconds > 0 ? (double)out->queries / out->seconds : 0.0;
    out->requestsPerSec = out->seconds > 0 ? (double)R / out->seconds : 0.0;
    out->offeredPerSec = items.back().arrivalUs > 0 ? (double)R / (items.back().arrivalUs * 1e-6) : 0.0;
    return rcAll;
}
