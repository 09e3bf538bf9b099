// Drop-in for src/render_temporal_decode.h: the stream model, half
// conversion, stream (de)serialisation and the three per-value decoders,
// with the reference's names and signatures; defined in libvbt_hip_cxx.so
// over libvbt_hip.so.
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

} // namespace hip_detail

uint16_t renderTemporalFloatToHalfBits(float value);
float renderTemporalHalfBitsToFloat(uint16_t bits);

bool validateRenderTemporalControlStreamData(const RenderTemporalControlStreamData& stream);

std::vector<uint8_t> packRenderTemporalControlStream(const RenderTemporalControlStreamData& stream,
                                                            bool alignTo4Bytes);

RenderTemporalControlStreamData unpackRenderTemporalControlStream(const uint8_t* streamBase,
                                                                        RenderTemporalSequenceCodec codec,
                                                                        uint16_t controlCount, uint8_t timeBinCount,
                                                                        uint8_t maxBinLocalKeys,
                                                                        uint32_t keyframeCount, uint32_t streamBytes,
                                                                        bool alignTo4Bytes);

float decodeRenderTemporalControlValue(const uint8_t* streamBase, RenderTemporalSequenceCodec codec,
                                              uint16_t controlCount, uint8_t timeBinCount, uint8_t maxBinLocalKeys,
                                              uint32_t keyframeCount, uint32_t streamBytes, bool alignTo4Bytes,
                                              int totalFrames, uint16_t controlIndex, int frameIndex);

// As the reference: the caller's header, prefix and layout locate the coarse
// or fine stream inside the leaf, and that stream's value is decoded.
float decodeRenderTemporalLeafControlValue(const uint8_t* leafBase,
                                                  const RenderTemporalPackedHeaderFields& header,
                                                  const RenderTemporalPackedLeafPrefix& prefix,
                                                  const RenderTemporalPackedLeafLayout& layout, bool useFineStream,
                                                  bool alignTo4Bytes, int totalFrames, uint16_t controlIndex,
                                                  int frameIndex);

// As the reference: the occupancy section at the view's fine descriptor
// offset maps the voxel to its shell control (group prefix + set bits below
// it); that control of the shell residual stream is decoded.  Voxels outside
// the shell, and leaves that are not Mode2, give 0.
float decodeRenderTemporalLeafShellVoxelValue(const uint8_t* leafBase,
                                                     const RenderTemporalPackedLeafView& leafView,
                                                     bool alignTo4Bytes, int totalFrames, std::uint16_t voxelIndex,
                                                     int frameIndex);

} // namespace vbt
