// C++ owners for the batched entry points of libvbt_hip.so: a resident asset
// (vbt_asset_*), its frame caches, dense frames and point queries, and the
// decode server (vbt_server_*).  No counterpart in the reference: these are
// the bulk forms its per-value API is called in a loop for.  Errors throw
// std::runtime_error with the library's message.
#pragma once

#include "vbt_file.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace vbt::hip {

inline void check(int rc)
{
    if (rc != VBT_OK) throw std::runtime_error(vbt_last_error());
}

// Device memory owned by the library allocator.
template <class T>
class DeviceBuffer {
public:
    explicit DeviceBuffer(size_t n = 0) { resize(n); }
    ~DeviceBuffer() { vbt_device_free(p_); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void resize(size_t n)
    {
        if (n == n_) return;
        vbt_device_free(p_);
        p_ = nullptr;
        n_ = 0;
        if (n) {
            void* q = nullptr;
            check(vbt_device_alloc(&q, n * sizeof(T)));
            p_ = static_cast<T*>(q);
            n_ = n;
        }
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T* src, size_t n, void* stream = nullptr) { check(vbt_memcpy_h2d(p_, src, n * sizeof(T), stream)); }
    std::vector<T> download(void* stream = nullptr) const
    {
        std::vector<T> out(n_);
        check(vbt_memcpy_d2h(out.data(), p_, n_ * sizeof(T), stream));
        check(vbt_stream_synchronize(stream));
        return out;
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// A VBTPACK4 asset resident in HBM, with its device page table.  flags:
// VBT_OPEN_* (VBT_OPEN_SCI_DENSE_BASIS: scientific dense subtypes 2 / 3 as
// the encoder reconstructs them).
class Asset {
public:
    Asset(const void* bytes, uint64_t size, void* stream = nullptr, uint32_t flags = 0) : stream_(stream)
    {
        check(vbt_asset_open_memory_ex(bytes, size, stream, flags, &a_));
        check(vbt_asset_info_get(a_, &info_));
    }
    explicit Asset(const vbt::render::VbtFile& f, void* stream = nullptr, uint32_t flags = 0) : stream_(stream)
    {
        const std::vector<uint8_t> img = vbt::render::hip_detail::image(f);
        check(vbt_asset_open_memory_ex(img.data(), img.size(), stream, flags, &a_));
        check(vbt_asset_info_get(a_, &info_));
    }
    explicit Asset(const std::string& path, void* stream = nullptr, uint32_t flags = 0) : stream_(stream)
    {
        check(vbt_asset_open_file_ex(path.c_str(), stream, flags, &a_));
        check(vbt_asset_info_get(a_, &info_));
    }
    ~Asset() { vbt_asset_close(a_); }
    Asset(const Asset&) = delete;
    Asset& operator=(const Asset&) = delete;

    vbt_asset handle() const { return a_; }
    const vbt_asset_info& info() const { return info_; }
    void* stream() const { return stream_; }

    // Frames [t0, t0+nt) into the packed compact cache (device pointer).
    void frameCache(int t0, int nt, float* d_out, int layout = VBT_CACHE_PACKED) const
    {
        check(vbt_decode_frame_cache(a_, t0, nt, layout, d_out, stream_));
    }
    // Whole frame t, row-major (z, y, x) floats.
    void denseFrame(int t, float* d_out, const vbt_compose* c = nullptr) const
    {
        check(vbt_decode_frame_dense(a_, t, c, d_out, stream_));
    }
    // n queries (x, y, z, t as 4 x u32, device memory).
    void query(const uint32_t* d_queries, uint32_t n, float* d_out, int32_t* d_status = nullptr,
               const vbt_compose* c = nullptr) const
    {
        check(vbt_query_points(a_, d_queries, n, c, d_out, d_status, stream_));
    }
    // Host convenience: queries in, values out (synchronous).
    std::vector<float> query(const std::vector<uint32_t>& queries, const vbt_compose* c = nullptr) const
    {
        const uint32_t n = static_cast<uint32_t>(queries.size() / 4u);
        DeviceBuffer<uint32_t> q(queries.size());
        DeviceBuffer<float> out(n);
        q.upload(queries.data(), queries.size(), stream_);
        query(q.data(), n, out.data(), nullptr, c);
        return out.download(stream_);
    }

private:
    vbt_asset a_ = nullptr;
    vbt_asset_info info_{};
    void* stream_ = nullptr;
};

// The continuous-batching decode server (see vbt_server_* in vbt_hip.h).
class Server {
public:
    explicit Server(const vbt_server_config* c = nullptr) { check(vbt_server_create(c, &s_)); }
    ~Server() { vbt_server_destroy(s_); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    int add(const Asset& a)
    {
        int id = -1;
        check(vbt_server_register(s_, a.handle(), &id));
        return id;
    }
    uint64_t submitQuery(int asset, const uint32_t* queries, uint32_t n, const vbt_compose* c = nullptr)
    {
        uint64_t id = 0;
        check(vbt_server_submit_query(s_, asset, queries, n, c, &id));
        return id;
    }
    uint64_t submitCache(int asset, int t)
    {
        uint64_t id = 0;
        check(vbt_server_submit_cache(s_, asset, t, &id));
        return id;
    }
    vbt_request_result wait(uint64_t id, int timeoutMs = -1)
    {
        vbt_request_result r;
        check(vbt_server_poll(s_, id, timeoutMs, &r));
        return r;
    }
    // Values (and statuses) to host memory; releases the request.  Returns
    // the request status (VBT_OK, or VBT_ERR_FORMAT when an element failed).
    int fetch(uint64_t id, float* values, int32_t* status = nullptr)
    {
        const int rc = vbt_server_fetch(s_, id, values, status);
        if (rc != VBT_OK && rc != VBT_ERR_FORMAT) check(rc);
        return rc;
    }
    void release(uint64_t id) { check(vbt_server_release(s_, id)); }
    vbt_server_stats stats(int requestType = -1)
    {
        vbt_server_stats s;
        check(vbt_server_stats_get(s_, requestType, &s));
        return s;
    }
    vbt_server handle() const { return s_; }

private:
    vbt_server s_ = nullptr;
};

} // namespace vbt::hip
