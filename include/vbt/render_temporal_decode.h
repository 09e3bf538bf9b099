// Drop-in for src/render_temporal_decode.h: the stream model, half
// conversion, stream (de)serialisation and the three per-value decoders,
// with the reference's names and signatures, backed by libvbt_hip.so.
//
// The per-value decoders run on the GPU (one request each: the stream bytes
// are copied to the device, decoded by the batched stream kernel, and the
// value copied back).  That keeps every decoded value on the HIP path, but a launch per
// value is slow: bulk callers use the batched entry points
// (vbt_decode_control_values / vbt_decode_leaf_control_values /
// vbt_decode_shell_voxel_values, or an asset's frame cache and point queries).
// Errors throw std::runtime_error with the reference's message text.
#pragma once

#include "render_temporal_payload.h"

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace vbt {

struct RenderTemporalControlDescriptor {  // render_temporal_decode.h:11-15
    uint16_t keyStart = 0;
    uint8_t keyCount = 0;
    uint8_t reserved = 0;
};

struct RenderTemporalControlKeyframe {  // render_temporal_decode.h:17-20
    uint16_t frame = 0;
    float value = 0.0f;
};

static_assert(sizeof(RenderTemporalControlDescriptor) == 4, "RenderTemporalControlDescriptor must stay 4 bytes");

struct RenderTemporalControlStreamData {  // render_temporal_decode.h:25-33
    RenderTemporalSequenceCodec codec = RenderTemporalSequenceCodec::DP_KEYFRAME_FP16;
    uint16_t controlCount = 0;
    uint8_t timeBinCount = 0;
    uint8_t maxBinLocalKeys = 0;
    std::vector<RenderTemporalControlDescriptor> descriptors;
    std::vector<uint8_t> binIndex;
    std::vector<RenderTemporalControlKeyframe> keyframes;
};

namespace hip_detail {
inline vbt_stream_data view(const RenderTemporalControlStreamData& s, std::vector<vbt_control_descriptor>& d,
                            std::vector<vbt_control_keyframe>& k)
{
    d.resize(s.descriptors.size());
    for (size_t i = 0; i < d.size(); ++i)
        d[i] = vbt_control_descriptor{s.descriptors[i].keyStart, s.descriptors[i].keyCount, s.descriptors[i].reserved};
    k.resize(s.keyframes.size());
    for (size_t i = 0; i < k.size(); ++i) k[i] = vbt_control_keyframe{s.keyframes[i].frame, s.keyframes[i].value};
    vbt_stream_data v;
    v.codec = static_cast<uint8_t>(s.codec);
    v.controlCount = s.controlCount;
    v.timeBinCount = s.timeBinCount;
    v.maxBinLocalKeys = s.maxBinLocalKeys;
    v.descriptors = d.data();
    v.binIndex = const_cast<uint8_t*>(s.binIndex.data());
    v.keyframes = k.data();
    v.keyframeCount = static_cast<uint32_t>(k.size());
    return v;
}

// One decode request on the device: the bytes, one index, one frame.
struct OneValue {
    void* d = nullptr;
    explicit OneValue(size_t bytes)
    {
        check(vbt_device_alloc(&d, bytes + 64));
    }
    ~OneValue() { vbt_device_free(d); }
    OneValue(const OneValue&) = delete;
    OneValue& operator=(const OneValue&) = delete;
    uint8_t* bytes() const { return static_cast<uint8_t*>(d); }
    template <class Launch>
    float run(const void* src, size_t bytes, uint16_t index, int frame, Launch launch)
    {
        uint8_t* base = static_cast<uint8_t*>(d);
        const size_t off = (bytes + 15u) & ~size_t(15);
        auto* dIndex = reinterpret_cast<uint16_t*>(base + off);
        auto* dFrame = reinterpret_cast<int32_t*>(base + off + 16);
        auto* dOut = reinterpret_cast<float*>(base + off + 32);
        auto* dStatus = reinterpret_cast<int32_t*>(base + off + 48);
        if (bytes) check(vbt_memcpy_h2d(base, src, bytes, nullptr));
        check(vbt_memcpy_h2d(dIndex, &index, sizeof(index), nullptr));
        check(vbt_memcpy_h2d(dFrame, &frame, sizeof(frame), nullptr));
        check(launch(base, dIndex, dFrame, dOut, dStatus));
        float value = 0.0f;
        int32_t status = 0;
        check(vbt_memcpy_d2h(&value, dOut, sizeof(value), nullptr));
        check(vbt_memcpy_d2h(&status, dStatus, sizeof(status), nullptr));
        check(vbt_stream_synchronize(nullptr));
        if (status != 0) throw std::runtime_error(vbt_status_message(status));
        return value;
    }
};
} // namespace hip_detail

inline uint16_t renderTemporalFloatToHalfBits(float value) { return vbt_float_to_half_bits(value); }
inline float renderTemporalHalfBitsToFloat(uint16_t bits) { return vbt_half_bits_to_float(bits); }

inline bool validateRenderTemporalControlStreamData(const RenderTemporalControlStreamData& stream)
{
    std::vector<vbt_control_descriptor> d;
    std::vector<vbt_control_keyframe> k;
    const vbt_stream_data v = hip_detail::view(stream, d, k);
    return vbt_validate_control_stream(&v) != 0;
}

inline std::vector<uint8_t> packRenderTemporalControlStream(const RenderTemporalControlStreamData& stream,
                                                            bool alignTo4Bytes)
{
    std::vector<vbt_control_descriptor> d;
    std::vector<vbt_control_keyframe> k;
    const vbt_stream_data v = hip_detail::view(stream, d, k);
    size_t len = 0;
    hip_detail::check(vbt_pack_control_stream(&v, alignTo4Bytes ? 1 : 0, nullptr, 0, &len));
    std::vector<uint8_t> out(len);
    hip_detail::check(vbt_pack_control_stream(&v, alignTo4Bytes ? 1 : 0, out.data(), out.size(), &len));
    out.resize(len);
    return out;
}

inline RenderTemporalControlStreamData unpackRenderTemporalControlStream(const uint8_t* streamBase,
                                                                        RenderTemporalSequenceCodec codec,
                                                                        uint16_t controlCount, uint8_t timeBinCount,
                                                                        uint8_t maxBinLocalKeys,
                                                                        uint32_t keyframeCount, uint32_t streamBytes,
                                                                        bool alignTo4Bytes)
{
    std::vector<vbt_control_descriptor> d(controlCount);
    std::vector<uint8_t> bins(static_cast<size_t>(controlCount) * timeBinCount);
    std::vector<vbt_control_keyframe> k(keyframeCount);
    vbt_stream_data v{};
    v.descriptors = d.data();
    v.binIndex = bins.data();
    v.keyframes = k.data();
    v.keyframeCount = keyframeCount;
    hip_detail::check(vbt_unpack_control_stream(streamBase, static_cast<uint32_t>(codec), controlCount, timeBinCount,
                                                maxBinLocalKeys, keyframeCount, streamBytes, alignTo4Bytes ? 1 : 0,
                                                &v));
    RenderTemporalControlStreamData o;
    o.codec = static_cast<RenderTemporalSequenceCodec>(v.codec);
    o.controlCount = v.controlCount;
    o.timeBinCount = v.timeBinCount;
    o.maxBinLocalKeys = v.maxBinLocalKeys;
    o.descriptors.resize(controlCount);
    for (size_t i = 0; i < d.size(); ++i) o.descriptors[i] = {d[i].keyStart, d[i].keyCount, d[i].reserved};
    o.binIndex = bins;
    o.keyframes.resize(keyframeCount);
    for (size_t i = 0; i < k.size(); ++i) o.keyframes[i] = {k[i].frame, k[i].value};
    return o;
}

inline float decodeRenderTemporalControlValue(const uint8_t* streamBase, RenderTemporalSequenceCodec codec,
                                              uint16_t controlCount, uint8_t timeBinCount, uint8_t maxBinLocalKeys,
                                              uint32_t keyframeCount, uint32_t streamBytes, bool alignTo4Bytes,
                                              int totalFrames, uint16_t controlIndex, int frameIndex)
{
    hip_detail::OneValue one(streamBytes);
    return one.run(streamBase, streamBytes, controlIndex, frameIndex,
                   [&](const uint8_t* dBytes, const uint16_t* dIndex, const int32_t* dFrame, float* dOut,
                       int32_t* dStatus) {
                       return vbt_decode_control_values(dBytes, static_cast<uint32_t>(codec), controlCount,
                                                        timeBinCount, maxBinLocalKeys, keyframeCount, streamBytes,
                                                        alignTo4Bytes ? 1 : 0, totalFrames, dIndex, dFrame, 1, dOut,
                                                        dStatus, nullptr);
                   });
}

// As the reference: the caller's header, prefix and layout locate the coarse
// or fine stream inside the leaf, and that stream's value is decoded.
inline float decodeRenderTemporalLeafControlValue(const uint8_t* leafBase,
                                                  const RenderTemporalPackedHeaderFields& header,
                                                  const RenderTemporalPackedLeafPrefix& prefix,
                                                  const RenderTemporalPackedLeafLayout& layout, bool useFineStream,
                                                  bool alignTo4Bytes, int totalFrames, uint16_t controlIndex,
                                                  int frameIndex)
{
    const uint16_t coarseCount = static_cast<uint16_t>(header.coarseResolution * header.coarseResolution *
                                                       header.coarseResolution);
    const uint16_t fineCount = header.fineResolution == 0u
                                   ? 0u
                                   : static_cast<uint16_t>(header.fineResolution * header.fineResolution *
                                                           header.fineResolution);
    if (useFineStream)
        return decodeRenderTemporalControlValue(leafBase + layout.fineOffset, header.fineCodec, fineCount,
                                                kRenderTemporalPackedTimeBinCount, 0u, prefix.fineKeyframeCount,
                                                layout.fine.totalBytes, alignTo4Bytes, totalFrames, controlIndex,
                                                frameIndex);
    return decodeRenderTemporalControlValue(leafBase + layout.coarseOffset, header.coarseCodec, coarseCount,
                                            kRenderTemporalPackedTimeBinCount, 0u, prefix.coarseKeyframeCount,
                                            layout.coarse.totalBytes, alignTo4Bytes, totalFrames, controlIndex,
                                            frameIndex);
}

// As the reference: the occupancy section at the view's fine descriptor
// offset maps the voxel to its shell control (group prefix + set bits below
// it); that control of the shell residual stream is decoded.  Voxels outside
// the shell, and leaves that are not Mode2, give 0.
inline float decodeRenderTemporalLeafShellVoxelValue(const uint8_t* leafBase,
                                                     const RenderTemporalPackedLeafView& leafView,
                                                     bool alignTo4Bytes, int totalFrames, std::uint16_t voxelIndex,
                                                     int frameIndex)
{
    if (leafBase == nullptr || voxelIndex >= 512) return 0.0f;
    if (leafView.header.mode != TemporalFirstPackedMode::TEMPORAL_FINE_COMPACT) return 0.0f;
    vbt_shell_section sec;
    vbt_unpack_shell_section(leafBase + leafView.layout.fine.descriptorOffset, leafView.layout.fine.descriptorBytes,
                             &sec);
    const uint32_t group = voxelIndex >> 6, bit = voxelIndex & 63u;
    const uint64_t mask = sec.shellMask[group];
    if ((mask & (1ull << bit)) == 0ull) return 0.0f;
    uint16_t descIndex = sec.groupPrefix[group];
    for (uint64_t w = mask & ((1ull << bit) - 1ull); w != 0; w &= w - 1) ++descIndex;
    if (leafView.shellActiveVoxelCount == 0 || leafView.prefix.fineKeyframeCount == 0) return 0.0f;
    return decodeRenderTemporalControlValue(leafBase + leafView.shellResidualOffset, leafView.header.fineCodec,
                                            leafView.shellActiveVoxelCount, kRenderTemporalShellPackedTimeBinCount,
                                            0u, leafView.prefix.fineKeyframeCount, leafView.shellResidualBytes,
                                            alignTo4Bytes, totalFrames, descIndex, frameIndex);
}

} // namespace vbt
