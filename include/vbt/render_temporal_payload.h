// Drop-in for src/render_temporal_payload.h (the reference's leaf format API):
// same namespace, type names, field order and function signatures.  The
// functions are defined in libvbt_hip_cxx.so over libvbt_hip.so's C-ABI (link
// -lvbt_hip_cxx -lvbt_hip); validateRenderTemporalLeafPrefix, which reads the
// encoder's option struct, is defined inline here.  Format errors throw
// std::runtime_error with the reference's message text.
#pragma once

#include "temporal_first_encoder.h"
#include "../vbt_hip.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace vbt {

enum class RenderTemporalSequenceCodec : uint8_t {
    DP_KEYFRAME_FP16 = 0,
    DP_KEYFRAME_FP32 = 1,
};

struct RenderTemporalPackedHeaderFields {  // render_temporal_payload.h:21-27
    TemporalFirstPackedMode mode = TemporalFirstPackedMode::EMPTY;
    RenderTemporalSequenceCodec coarseCodec = RenderTemporalSequenceCodec::DP_KEYFRAME_FP16;
    RenderTemporalSequenceCodec fineCodec = RenderTemporalSequenceCodec::DP_KEYFRAME_FP16;
    uint8_t coarseResolution = 4;
    uint8_t fineResolution = 0;
};

struct RenderTemporalPackedLeafPrefix {  // render_temporal_payload.h:32-35
    uint16_t coarseKeyframeCount = 0;
    uint16_t fineKeyframeCount = 0;
};
static_assert(sizeof(RenderTemporalPackedLeafPrefix) == 4, "RenderTemporalPackedLeafPrefix must stay 4 bytes");

struct RenderTemporalStreamLayout {  // render_temporal_payload.h:40-48
    uint32_t descriptorOffset = 0;
    uint32_t descriptorBytes = 0;
    uint32_t binIndexOffset = 0;
    uint32_t binIndexBytes = 0;
    uint32_t keyframeOffset = 0;
    uint32_t keyframeBytes = 0;
    uint32_t totalBytes = 0;
};

struct RenderTemporalPackedLeafLayout {  // render_temporal_payload.h:50-59
    uint32_t headerBytes = 4;
    uint32_t prefixOffset = 4;
    uint32_t prefixBytes = sizeof(RenderTemporalPackedLeafPrefix);
    uint32_t coarseOffset = 0;
    RenderTemporalStreamLayout coarse;
    uint32_t fineOffset = 0;
    RenderTemporalStreamLayout fine;
    uint32_t totalBytes = 0;
};

struct RenderTemporalPackedLeafView {  // render_temporal_payload.h:61-68
    RenderTemporalPackedHeaderFields header;
    RenderTemporalPackedLeafPrefix prefix;
    RenderTemporalPackedLeafLayout layout;
    uint16_t shellActiveVoxelCount = 0;
    uint32_t shellResidualOffset = 0;
    uint32_t shellResidualBytes = 0;
};

constexpr uint32_t kRenderTemporalHeaderBytes = 4;
constexpr uint32_t kRenderTemporalControlDescriptorBytes = 4;
constexpr uint8_t kRenderTemporalPackedTimeBinCount = 8;
constexpr uint8_t kRenderTemporalShellPackedTimeBinCount = 4;
constexpr uint16_t kRenderTemporalLeafSize = 8;

constexpr uint16_t renderTemporalLeafVoxelIndex(uint16_t lx, uint16_t ly, uint16_t lz,
                                                uint16_t leafSize = kRenderTemporalLeafSize)
{
    return static_cast<uint16_t>((lz * leafSize + ly) * leafSize + lx);
}

struct RenderTemporalShellOccupancySection {  // render_temporal_payload.h:87-90
    std::array<std::uint64_t, 8> shellMask{};
    std::array<std::uint16_t, 8> groupPrefix{};
};
static_assert(sizeof(RenderTemporalShellOccupancySection) == 80,
              "RenderTemporalShellOccupancySection must stay 80 bytes");

struct RenderTemporalShellSectionLayout {
    uint32_t occupancyOffset = 0;
    uint32_t occupancyBytes = 0;
    uint32_t residualOffset = 0;
    uint32_t residualBytes = 0;
    uint32_t totalBytes = 0;
};

namespace hip_detail {
[[noreturn]] inline void raise() { throw std::runtime_error(vbt_last_error()); }
inline void check(int rc)
{
    if (rc != VBT_OK) raise();
}
inline vbt_header_fields to_c(const RenderTemporalPackedHeaderFields& h)
{
    return vbt_header_fields{static_cast<uint8_t>(h.mode), static_cast<uint8_t>(h.coarseCodec),
                             static_cast<uint8_t>(h.fineCodec), h.coarseResolution, h.fineResolution};
}
inline RenderTemporalPackedHeaderFields from_c(const vbt_header_fields& h)
{
    RenderTemporalPackedHeaderFields o;
    o.mode = static_cast<TemporalFirstPackedMode>(h.mode);
    o.coarseCodec = static_cast<RenderTemporalSequenceCodec>(h.coarseCodec);
    o.fineCodec = static_cast<RenderTemporalSequenceCodec>(h.fineCodec);
    o.coarseResolution = h.coarseResolution;
    o.fineResolution = h.fineResolution;
    return o;
}
inline vbt_leaf_prefix to_c(const RenderTemporalPackedLeafPrefix& p) { return {p.coarseKeyframeCount, p.fineKeyframeCount}; }
inline RenderTemporalStreamLayout from_c(const vbt_stream_layout& s)
{
    RenderTemporalStreamLayout o;
    o.descriptorOffset = s.descriptorOffset;
    o.descriptorBytes = s.descriptorBytes;
    o.binIndexOffset = s.binIndexOffset;
    o.binIndexBytes = s.binIndexBytes;
    o.keyframeOffset = s.keyframeOffset;
    o.keyframeBytes = s.keyframeBytes;
    o.totalBytes = s.totalBytes;
    return o;
}
inline RenderTemporalPackedLeafLayout from_c(const vbt_leaf_layout& l)
{
    RenderTemporalPackedLeafLayout o;
    o.headerBytes = l.headerBytes;
    o.prefixOffset = l.prefixOffset;
    o.prefixBytes = l.prefixBytes;
    o.coarseOffset = l.coarseOffset;
    o.coarse = from_c(l.coarse);
    o.fineOffset = l.fineOffset;
    o.fine = from_c(l.fine);
    o.totalBytes = l.totalBytes;
    return o;
}
inline std::vector<uint8_t> sized(int rc, std::vector<uint8_t>& buf, size_t len)
{
    check(rc);
    buf.resize(len);
    return buf;
}
} // namespace hip_detail

uint32_t renderTemporalKeyframeBytes(RenderTemporalSequenceCodec codec);

uint32_t packRenderTemporalHeader(const RenderTemporalPackedHeaderFields& fields);

RenderTemporalPackedHeaderFields unpackRenderTemporalHeader(uint32_t packedHeader);

RenderTemporalPackedLeafPrefix makeRenderTemporalLeafPrefix(TemporalFirstPackedMode mode,
                                                                   uint32_t coarseKeyframeCount,
                                                                   uint32_t fineKeyframeCount);

inline bool validateRenderTemporalLeafPrefix(const RenderTemporalPackedLeafPrefix& prefix,
                                             const RenderTemporalPackedHeaderFields& header,
                                             const TemporalFirstEncoderOptions& options)
{
    const vbt_leaf_prefix p = hip_detail::to_c(prefix);
    const vbt_header_fields h = hip_detail::to_c(header);
    return vbt_validate_leaf_prefix(&p, &h, options.timeBinCount) != 0;
}

RenderTemporalPackedLeafLayout computeRenderTemporalLeafLayout(const RenderTemporalPackedHeaderFields& header,
                                                                     const RenderTemporalPackedLeafPrefix& prefix,
                                                                     bool alignTo4Bytes);

std::vector<uint8_t> buildRenderTemporalPackedLeafBytes(const RenderTemporalPackedHeaderFields& header,
                                                               const RenderTemporalPackedLeafPrefix& prefix,
                                                               const std::vector<uint8_t>& coarseStreamBytes,
                                                               const std::vector<uint8_t>& fineStreamBytes,
                                                               bool alignTo4Bytes);

std::vector<uint8_t> buildRenderTemporalPackedLeafBytes(const RenderTemporalPackedHeaderFields& header,
                                                               const RenderTemporalPackedLeafPrefix& prefix,
                                                               const std::vector<uint8_t>& coarseStreamBytes,
                                                               const std::vector<uint8_t>& shellOccupancyBytes,
                                                               const std::vector<uint8_t>& shellResidualStreamBytes,
                                                               bool alignTo4Bytes);

std::vector<uint8_t> packRenderTemporalShellOccupancySection(const RenderTemporalShellOccupancySection& section);

RenderTemporalShellOccupancySection unpackRenderTemporalShellOccupancySection(const uint8_t* data,
                                                                                     std::uint64_t size);

std::uint16_t renderTemporalShellActiveVoxelCount(const RenderTemporalShellOccupancySection& section);

RenderTemporalPackedLeafView parseRenderTemporalPackedLeaf(const uint8_t* leafBase, size_t leafBytes,
                                                                  bool alignTo4Bytes);

} // namespace vbt
