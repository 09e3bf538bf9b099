// tuning.hip -- the frame-cache autotuner's persistent table: tile-shape
// picks keyed by the statistics the best shape depends on (route mix,
// keyframe-count histogram, slots per frame, frames per launch), kept per
// process and loaded from / saved to a JSON file, so a run on a known kind of
// asset skips the timing sweep.
#include "vbt_internal.h"
#include "vbt_kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

using namespace vbtk;

namespace {

constexpr int KEY_BUCKETS = 9; // keyframes per control: 0, 1, 2-3, 4-7, ..., 128-255
constexpr const char* TABLE_FORMAT = "vbt-tuning-1";

struct Entry {
    int tile;
    float ms;
};

std::mutex g_mu;
std::map<std::string, Entry> g_table;

// keyframes per packed-cache slot, bucketed by powers of two (one thread per slot)
__global__ void __launch_bounds__(256) k_key_hist(AssetDev a, unsigned long long* hist)
{
    __shared__ unsigned int local[KEY_BUCKETS];
    if (threadIdx.x < KEY_BUCKETS) local[threadIdx.x] = 0u;
    __syncthreads();
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.controlsPerFrame) {
        const uint32_t c = find_leaf(a.ctrl, a.A, g);
        const DevLeaf& d = a.leaves[c];
        const uint32_t slot = (uint32_t)(g - a.ctrl[c]);
        uint32_t kc = 0;
        if ((d.status & STATUS_MASK) == LEAF_OK) {
            const bool isCoarse = slot < d.cc;
            const StreamRef s = isCoarse ? d.plan.coarse : (d.plan.mode == MODE_FINE6 ? d.plan.fine : d.plan.shell);
            const uint32_t ci = isCoarse ? slot : slot - d.cc;
            kc = (*(const uint32_t*)(a.payload + (size_t)d.word * 4u + s.base + ci * 4u) >> 16) & 0xFFu;
        }
        atomicAdd(&local[kc == 0u ? 0 : 32 - __clz(kc)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < KEY_BUCKETS && local[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)local[threadIdx.x]);
}

int tenths(uint64_t part, uint64_t whole) { return whole ? (int)std::lround(10.0 * (double)part / (double)whole) : 0; }

int asset_key(vbt_asset a, int nt, std::string& key, hipStream_t s)
{
    const vbt_file_header* h = vbt_internal_asset_header(a);
    vbt_asset_info info;
    int rc = vbt_asset_info_get(a, &info);
    if (rc) return rc;
    if ((uint64_t)nt > h->frames) nt = (int)h->frames;
    std::ostringstream k;
    k << "p" << h->profileType << " nt" << nt << " s"
      << (info.controlsPerFrame ? (int)std::lround(std::log2((double)info.controlsPerFrame)) : 0);
    const uint64_t A = info.activeLeaves;
    k << " r" << tenths(info.mode_counts[1], A) << "." << tenths(info.mode_counts[3], A) << "."
      << tenths(info.mode_counts[2], A);
    unsigned long long hist[KEY_BUCKETS] = {};
    if (h->profileType != 2 && info.controlsPerFrame) {
        unsigned long long* d = nullptr;
        if (hipMalloc((void**)&d, sizeof(hist)) != hipSuccess) return vbt_fail(VBT_ERR_DEVICE, "tuning key: alloc failed");
        const AssetDev ad = vbt_internal_asset_dev(a);
        const uint32_t blocks = (uint32_t)((info.controlsPerFrame + 255u) / 256u);
        bool ok = hipMemsetAsync(d, 0, sizeof(hist), s) == hipSuccess;
        if (ok) {
            k_key_hist<<<blocks, 256, 0, s>>>(ad, d);
            ok = hipGetLastError() == hipSuccess && hipMemcpyAsync(hist, d, sizeof(hist), hipMemcpyDeviceToHost, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess;
        }
        (void)hipFree(d);
        if (!ok) return vbt_fail(VBT_ERR_DEVICE, "tuning key: histogram failed");
    }
    k << " k";
    for (int b = 0; b < KEY_BUCKETS; ++b) k << (b ? "." : "") << tenths(hist[b], info.controlsPerFrame);
    key = k.str();
    return VBT_OK;
}

// the table file: {"format": ..., "entries": [{"key": "...", "tile": t, "ms": x}, ...]}
std::string table_text()
{
    std::ostringstream o;
    o << "{\n \"format\": \"" << TABLE_FORMAT << "\",\n \"arch\": \"gfx950\",\n \"entries\": [";
    bool first = true;
    for (const auto& [k, e] : g_table) {
        char ms[32];
        std::snprintf(ms, sizeof(ms), "%.4f", (double)e.ms);
        o << (first ? "\n" : ",\n") << "  {\"key\": \"" << k << "\", \"tile\": " << e.tile << ", \"ms\": " << ms << "}";
        first = false;
    }
    o << "\n ]\n}\n";
    return o.str();
}

} // namespace

extern "C" {

int vbt_tuning_key(vbt_asset a, int nt, char* buf, size_t len, void* stream)
{
    if (!a || !buf || len == 0 || nt <= 0) return vbt_fail(VBT_ERR_INVALID, "invalid tuning key request");
    std::string key;
    const int rc = asset_key(a, nt, key, (hipStream_t)stream);
    if (rc) return rc;
    if (key.size() + 1 > len) return vbt_fail(VBT_ERR_INVALID, "tuning key buffer too small");
    std::memcpy(buf, key.c_str(), key.size() + 1);
    return VBT_OK;
}

int vbt_tuning_count(void)
{
    std::lock_guard<std::mutex> g(g_mu);
    return (int)g_table.size();
}

int vbt_tuning_clear(void)
{
    std::lock_guard<std::mutex> g(g_mu);
    g_table.clear();
    return VBT_OK;
}

int vbt_tuning_load(const char* path)
{
    if (!path) return vbt_fail(VBT_ERR_INVALID, "invalid tuning table path");
    std::ifstream in(path);
    if (!in) return vbt_fail(VBT_ERR_INVALID, std::string("cannot read tuning table: ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string t = ss.str();
    if (t.find(std::string("\"format\": \"") + TABLE_FORMAT + "\"") == std::string::npos)
        return vbt_fail(VBT_ERR_FORMAT, std::string("not a tuning table: ") + path);
    // entries: objects with "key" (string), "tile" and "ms" (numbers), any whitespace
    auto value_at = [&](const char* name, size_t from, size_t to) -> size_t {
        const std::string tag = std::string("\"") + name + "\"";
        size_t p = t.find(tag, from);
        if (p == std::string::npos || p >= to) return std::string::npos;
        p = t.find(':', p + tag.size());
        if (p == std::string::npos || p >= to) return std::string::npos;
        return t.find_first_not_of(" \t\r\n", p + 1);
    };
    std::map<std::string, Entry> got;
    for (size_t p = t.find('{', t.find("\"entries\"")); p != std::string::npos; p = t.find('{', p + 1)) {
        const size_t end = t.find('}', p);
        if (end == std::string::npos) return vbt_fail(VBT_ERR_FORMAT, "malformed tuning table entry");
        const size_t k = value_at("key", p, end), ti = value_at("tile", p, end), mi = value_at("ms", p, end);
        if (k == std::string::npos || ti == std::string::npos || mi == std::string::npos || t[k] != '"')
            return vbt_fail(VBT_ERR_FORMAT, "malformed tuning table entry");
        const size_t q = t.find('"', k + 1);
        if (q == std::string::npos || q > end) return vbt_fail(VBT_ERR_FORMAT, "malformed tuning table entry");
        char* stop = nullptr;
        const long tile = std::strtol(t.c_str() + ti, &stop, 10);
        if (stop == t.c_str() + ti) return vbt_fail(VBT_ERR_FORMAT, "malformed tuning table entry");
        const double ms = std::strtod(t.c_str() + mi, &stop);
        if (tile < 0 || tile >= CACHE_TILE_COUNT) return vbt_fail(VBT_ERR_FORMAT, "tuning table tile out of range");
        got[t.substr(k + 1, q - k - 1)] = Entry{(int)tile, (float)ms};
        p = end;
    }
    std::lock_guard<std::mutex> g(g_mu);
    for (const auto& [k, e] : got) g_table[k] = e;
    return VBT_OK;
}

int vbt_tuning_save(const char* path)
{
    if (!path) return vbt_fail(VBT_ERR_INVALID, "invalid tuning table path");
    std::string text;
    {
        std::lock_guard<std::mutex> g(g_mu);
        text = table_text();
    }
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream out(tmp, std::ios::trunc);
        if (!out || !(out << text) || !out.flush())
            return vbt_fail(VBT_ERR_INVALID, std::string("cannot write tuning table: ") + path);
    }
    if (std::rename(tmp.c_str(), path) != 0)
        return vbt_fail(VBT_ERR_INVALID, std::string("cannot write tuning table: ") + path);
    return VBT_OK;
}

int vbt_asset_autotune_table(vbt_asset a, int nt, int reps, float* msPerTile, int* best, int* measured,
                             void* stream)
{
    if (!a || nt <= 0 || reps <= 0) return vbt_fail(VBT_ERR_INVALID, "invalid autotune request");
    std::string key;
    int rc = asset_key(a, nt, key, (hipStream_t)stream);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(g_mu);
        const auto it = g_table.find(key);
        if (it != g_table.end()) {
            rc = vbt_asset_set_cache_tile(a, it->second.tile);
            if (rc) return rc;
            if (best) *best = it->second.tile;
            if (measured) *measured = 0;
            return VBT_OK;
        }
    }
    float ms[CACHE_TILE_COUNT];
    int tile = 0;
    rc = vbt_asset_autotune(a, nt, reps, ms, &tile, stream);
    if (rc) return rc;
    if (msPerTile) std::memcpy(msPerTile, ms, sizeof(ms));
    if (best) *best = tile;
    if (measured) *measured = 1;
    if (vbt_internal_asset_header(a)->profileType != 2) {
        std::lock_guard<std::mutex> g(g_mu);
        g_table[key] = Entry{tile, ms[tile]};
    }
    return VBT_OK;
}

} // extern "C"
