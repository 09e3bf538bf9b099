// vbt_cxx.cpp -- libvbt_hip_cxx.so: the reference's C++ decode API
// (src/render_temporal_decode.h, src/render_temporal_payload.h,
// render/src/vbt_file.h, render/src/scientific_decode.h) as compiled,
// exported functions over the C-ABI of libvbt_hip.so.  A reference target
// drops the four CPU .cpp files and links this library instead; the
// reference's own headers and the ones in include/vbt/ declare the same
// functions with the same types, so either set compiles against it.
// (validateRenderTemporalLeafPrefix is the exception: it takes the encoder's
// option struct, so it stays an inline function of include/vbt/.)
#include "../../include/vbt/query_patterns.h"
#include "../../include/vbt/render_temporal_decode.h"
#include "../../include/vbt/scientific_decode.h"
#include "../../include/vbt/vbt_file.h"

#include <algorithm>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <system_error>
#include <vector>

namespace vbt::hip_detail {
namespace {
// One decode request on the device: the stream bytes, one index and one
// frame go over in a single copy into a per-thread device buffer (grown on
// demand, kept for the thread's lifetime), the value and status come back in
// one copy.  A launch per value is what the reference's per-value API implies;
// bulk callers use the batched C entry points.
struct OneValue {
    void* d = nullptr;
    size_t cap = 0;
    std::vector<uint8_t> host;

    template <class Launch>
    float run(const void* src, size_t bytes, uint16_t index, int frame, Launch launch)
    {
        const size_t off = (bytes + 15u) & ~size_t(15);
        const size_t need = off + 64u;
        if (need > cap) {
            if (d) vbt_device_free(d);
            d = nullptr;
            cap = 0;
            check(vbt_device_alloc(&d, need * 2u));
            cap = need * 2u;
        }
        host.assign(off + 16u, 0u);
        if (bytes) std::memcpy(host.data(), src, bytes);
        std::memcpy(host.data() + off, &index, sizeof(index));
        std::memcpy(host.data() + off + 4u, &frame, sizeof(frame));
        uint8_t* base = static_cast<uint8_t*>(d);
        check(vbt_memcpy_h2d(base, host.data(), host.size(), nullptr));
        auto* dIndex = reinterpret_cast<uint16_t*>(base + off);
        auto* dFrame = reinterpret_cast<int32_t*>(base + off + 4u);
        auto* dOut = reinterpret_cast<float*>(base + off + 16u);
        auto* dStatus = reinterpret_cast<int32_t*>(base + off + 20u);
        check(launch(base, dIndex, dFrame, dOut, dStatus));
        uint32_t res[2] = {0u, 0u};
        check(vbt_memcpy_d2h(res, dOut, sizeof(res), nullptr));
        check(vbt_stream_synchronize(nullptr));
        const int32_t status = static_cast<int32_t>(res[1]);
        if (status != 0) throw std::runtime_error(vbt_status_message(status));
        float value;
        std::memcpy(&value, &res[0], sizeof(value));
        return value;
    }
};

OneValue& one_value()
{
    // never freed: the device buffer outlives static destruction order issues
    // of the HIP runtime at process exit
    thread_local OneValue* v = new OneValue();
    return *v;
}
} // namespace
} // namespace vbt::hip_detail

// ---- render_temporal_payload.h
namespace vbt {

uint32_t renderTemporalKeyframeBytes(RenderTemporalSequenceCodec codec)
{
    return vbt_keyframe_bytes(static_cast<uint32_t>(codec));
}

uint32_t packRenderTemporalHeader(const RenderTemporalPackedHeaderFields& fields)
{
    const vbt_header_fields h = hip_detail::to_c(fields);
    return vbt_pack_header(&h);
}

RenderTemporalPackedHeaderFields unpackRenderTemporalHeader(uint32_t packedHeader)
{
    vbt_header_fields h;
    vbt_unpack_header(packedHeader, &h);
    return hip_detail::from_c(h);
}

RenderTemporalPackedLeafPrefix makeRenderTemporalLeafPrefix(TemporalFirstPackedMode mode,
                                                                   uint32_t coarseKeyframeCount,
                                                                   uint32_t fineKeyframeCount)
{
    vbt_leaf_prefix p;
    hip_detail::check(vbt_make_leaf_prefix(static_cast<uint32_t>(mode), coarseKeyframeCount, fineKeyframeCount, &p));
    return RenderTemporalPackedLeafPrefix{p.coarseKeyframeCount, p.fineKeyframeCount};
}

RenderTemporalPackedLeafLayout computeRenderTemporalLeafLayout(const RenderTemporalPackedHeaderFields& header,
                                                                     const RenderTemporalPackedLeafPrefix& prefix,
                                                                     bool alignTo4Bytes)
{
    const vbt_leaf_prefix p = hip_detail::to_c(prefix);
    const vbt_header_fields h = hip_detail::to_c(header);
    vbt_leaf_layout l;
    vbt_compute_leaf_layout(&h, &p, alignTo4Bytes ? 1 : 0, &l);
    return hip_detail::from_c(l);
}

std::vector<uint8_t> buildRenderTemporalPackedLeafBytes(const RenderTemporalPackedHeaderFields& header,
                                                               const RenderTemporalPackedLeafPrefix& prefix,
                                                               const std::vector<uint8_t>& coarseStreamBytes,
                                                               const std::vector<uint8_t>& fineStreamBytes,
                                                               bool alignTo4Bytes)
{
    const vbt_leaf_prefix p = hip_detail::to_c(prefix);
    const vbt_header_fields h = hip_detail::to_c(header);
    size_t len = 0;
    hip_detail::check(vbt_build_leaf_bytes(&h, &p, coarseStreamBytes.data(), coarseStreamBytes.size(),
                                           fineStreamBytes.data(), fineStreamBytes.size(), alignTo4Bytes ? 1 : 0,
                                           nullptr, 0, &len));
    std::vector<uint8_t> out(len);
    return hip_detail::sized(vbt_build_leaf_bytes(&h, &p, coarseStreamBytes.data(), coarseStreamBytes.size(),
                                                  fineStreamBytes.data(), fineStreamBytes.size(),
                                                  alignTo4Bytes ? 1 : 0, out.data(), out.size(), &len),
                             out, len);
}

std::vector<uint8_t> buildRenderTemporalPackedLeafBytes(const RenderTemporalPackedHeaderFields& header,
                                                               const RenderTemporalPackedLeafPrefix& prefix,
                                                               const std::vector<uint8_t>& coarseStreamBytes,
                                                               const std::vector<uint8_t>& shellOccupancyBytes,
                                                               const std::vector<uint8_t>& shellResidualStreamBytes,
                                                               bool alignTo4Bytes)
{
    const vbt_leaf_prefix p = hip_detail::to_c(prefix);
    const vbt_header_fields h = hip_detail::to_c(header);
    size_t len = 0;
    auto call = [&](uint8_t* out, size_t cap) {
        return vbt_build_shell_leaf_bytes(&h, &p, coarseStreamBytes.data(), coarseStreamBytes.size(),
                                          shellOccupancyBytes.data(), shellOccupancyBytes.size(),
                                          shellResidualStreamBytes.data(), shellResidualStreamBytes.size(),
                                          alignTo4Bytes ? 1 : 0, out, cap, &len);
    };
    hip_detail::check(call(nullptr, 0));
    std::vector<uint8_t> out(len);
    return hip_detail::sized(call(out.data(), out.size()), out, len);
}

std::vector<uint8_t> packRenderTemporalShellOccupancySection(const RenderTemporalShellOccupancySection& section)
{
    vbt_shell_section s;
    for (int i = 0; i < 8; ++i) {
        s.shellMask[i] = section.shellMask[i];
        s.groupPrefix[i] = section.groupPrefix[i];
    }
    std::vector<uint8_t> out(80);
    vbt_pack_shell_section(&s, out.data());
    return out;
}

RenderTemporalShellOccupancySection unpackRenderTemporalShellOccupancySection(const uint8_t* data,
                                                                                     std::uint64_t size)
{
    vbt_shell_section s;
    vbt_unpack_shell_section(data, size, &s);
    RenderTemporalShellOccupancySection o;
    for (int i = 0; i < 8; ++i) {
        o.shellMask[i] = s.shellMask[i];
        o.groupPrefix[i] = s.groupPrefix[i];
    }
    return o;
}

std::uint16_t renderTemporalShellActiveVoxelCount(const RenderTemporalShellOccupancySection& section)
{
    vbt_shell_section s;
    for (int i = 0; i < 8; ++i) {
        s.shellMask[i] = section.shellMask[i];
        s.groupPrefix[i] = section.groupPrefix[i];
    }
    return vbt_shell_active_count(&s);
}

RenderTemporalPackedLeafView parseRenderTemporalPackedLeaf(const uint8_t* leafBase, size_t leafBytes,
                                                                  bool alignTo4Bytes)
{
    vbt_leaf_view v;
    hip_detail::check(vbt_parse_leaf(leafBase, leafBytes, alignTo4Bytes ? 1 : 0, &v));
    RenderTemporalPackedLeafView o;
    o.header = hip_detail::from_c(v.header);
    o.prefix = RenderTemporalPackedLeafPrefix{v.prefix.coarseKeyframeCount, v.prefix.fineKeyframeCount};
    o.layout = hip_detail::from_c(v.layout);
    o.shellActiveVoxelCount = v.shellActiveVoxelCount;
    o.shellResidualOffset = v.shellResidualOffset;
    o.shellResidualBytes = v.shellResidualBytes;
    return o;
}

} // namespace vbt

// ---- render_temporal_decode.h
namespace vbt {

uint16_t renderTemporalFloatToHalfBits(float value)
{ return vbt_float_to_half_bits(value); }

float renderTemporalHalfBitsToFloat(uint16_t bits)
{ return vbt_half_bits_to_float(bits); }

bool validateRenderTemporalControlStreamData(const RenderTemporalControlStreamData& stream)
{
    std::vector<vbt_control_descriptor> d;
    std::vector<vbt_control_keyframe> k;
    const vbt_stream_data v = hip_detail::view(stream, d, k);
    return vbt_validate_control_stream(&v) != 0;
}

std::vector<uint8_t> packRenderTemporalControlStream(const RenderTemporalControlStreamData& stream,
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

RenderTemporalControlStreamData unpackRenderTemporalControlStream(const uint8_t* streamBase,
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

float decodeRenderTemporalControlValue(const uint8_t* streamBase, RenderTemporalSequenceCodec codec,
                                              uint16_t controlCount, uint8_t timeBinCount, uint8_t maxBinLocalKeys,
                                              uint32_t keyframeCount, uint32_t streamBytes, bool alignTo4Bytes,
                                              int totalFrames, uint16_t controlIndex, int frameIndex)
{
    return hip_detail::one_value().run(streamBase, streamBytes, controlIndex, frameIndex,
                   [&](const uint8_t* dBytes, const uint16_t* dIndex, const int32_t* dFrame, float* dOut,
                       int32_t* dStatus) {
                       return vbt_decode_control_values(dBytes, static_cast<uint32_t>(codec), controlCount,
                                                        timeBinCount, maxBinLocalKeys, keyframeCount, streamBytes,
                                                        alignTo4Bytes ? 1 : 0, totalFrames, dIndex, dFrame, 1, dOut,
                                                        dStatus, nullptr);
                   });
}

float decodeRenderTemporalLeafControlValue(const uint8_t* leafBase,
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

float decodeRenderTemporalLeafShellVoxelValue(const uint8_t* leafBase,
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

// ---- vbt_file.h
namespace vbt::render {

bool loadVbtFile(const std::filesystem::path& path, VbtFile& outFile, std::string& error)
{
    error.clear();
    std::error_code ec;
    const uintmax_t size = std::filesystem::file_size(path, ec);
    if (ec) {
        error = "Failed to query file size: " + path.string();
        return false;
    }
    if (size < 52u) {
        error = "Invalid VBTPACK4 file size.";
        return false;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        error = "Failed to open file: " + path.string();
        return false;
    }
    std::vector<uint8_t> bytes(static_cast<size_t>(size));
    in.read(reinterpret_cast<char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    if (!in) {
        error = "Failed to read VBTPACK header.";
        return false;
    }
    vbt_file_header h;
    if (vbt_validate_file(bytes.data(), bytes.size(), &h) != VBT_OK) {
        error = vbt_last_error();
        return false;
    }
    VbtFile f;
    std::memcpy(f.header.magic, bytes.data(), 8);
    f.header.version = h.version;
    f.header.width = h.width;
    f.header.height = h.height;
    f.header.depth = h.depth;
    f.header.frames = h.frames;
    f.header.leafSize = h.leafSize;
    f.header.coarseResolution = h.coarseResolution;
    f.header.maxCoarseKeep = h.maxCoarseKeep;
    f.header.leafCount = h.leafCount;
    f.header.profileType = h.profileType;
    f.header.coarseAcScaleCount = h.coarseAcScaleCount;
    size_t o = 52;
    f.coarseAcScales.resize(h.coarseAcScaleCount);
    std::memcpy(f.coarseAcScales.data(), bytes.data() + o, 4u * h.coarseAcScaleCount);
    o += 4u * h.coarseAcScaleCount;
    f.offsetsWords.resize(static_cast<size_t>(h.leafCount) + 1u);
    std::memcpy(f.offsetsWords.data(), bytes.data() + o, 4u * f.offsetsWords.size());
    o += 4u * f.offsetsWords.size();
    f.payloadWords.resize((bytes.size() - o) / 4u);
    std::memcpy(f.payloadWords.data(), bytes.data() + o, 4u * f.payloadWords.size());
    outFile = std::move(f);
    return true;
}

uint32_t readPackedHeaderWord(const VbtFile& file, uint32_t leafIndex)
{
    if (leafIndex >= file.header.leafCount) return 0u;
    const uint32_t w = file.offsetsWords[leafIndex];
    return w < file.payloadWords.size() ? file.payloadWords[w] : 0u;
}

ScientificHeaderV2 decodeScientificHeaderV2(uint32_t packedHeader)
{
    vbt_scientific_header s;
    vbt_decode_scientific_header(packedHeader, &s);
    ScientificHeaderV2 o;
    o.mode = static_cast<ScientificMode>(s.mode);
    o.coarseKeepMinus1 = s.coarseKeepMinus1;
    o.coarseQuantClass = s.coarseQuantClass;
    o.tierId = s.tierId;
    o.framesPerBinMinus1 = s.framesPerBinMinus1;
    o.eventQuantCode = s.eventQuantCode;
    o.sparseSubtype = s.sparseSubtype;
    o.packedEventCount = s.packedEventCount;
    o.denseQuantBits = s.denseQuantBits;
    o.denseSubtype = s.denseSubtype;
    o.denseTemporalKeepMinus1 = s.denseTemporalKeepMinus1;
    return o;
}

uint32_t leafIndexForVoxel(const VbtFileHeader& header, uint32_t x, uint32_t y, uint32_t z)
{
    const vbt_file_header h = hip_detail::to_c(header);
    return vbt_leaf_index_for_voxel(&h, x, y, z);
}

} // namespace vbt::render

// ---- scientific_decode.h
namespace vbt::render {

float decodeScientificValueAtCpu(const VbtFile& file, const Query4D& query)
{
    vbt::hip::Asset asset(file);
    const std::vector<uint32_t> q = {query.x, query.y, query.z, query.t};
    vbt::hip::DeviceBuffer<uint32_t> dq(4);
    vbt::hip::DeviceBuffer<float> out(1);
    vbt::hip::DeviceBuffer<int32_t> st(1);
    dq.upload(q.data(), 4);
    asset.query(dq.data(), 1, out.data(), st.data());
    const int32_t status = st.download()[0];
    if (status != 0) throw std::runtime_error(vbt_status_message(status));
    return out.download()[0];
}

std::vector<float> reconstructScientificFrameCpu(const VbtFile& file, uint32_t frameIndex)
{
    vbt::hip::Asset asset(file);
    const auto& h = asset.info().header;
    vbt::hip::DeviceBuffer<float> out(static_cast<size_t>(h.width) * h.height * h.depth);
    asset.denseFrame(static_cast<int>(frameIndex), out.data());
    return out.download();
}

} // namespace vbt::render

// ---- query_patterns.h
namespace vbt::render {

bool parseQueryPattern(const std::string& text, QueryPattern& outPattern)
{
    int p = 0;
    if (vbt_parse_query_pattern(text.c_str(), &p) != VBT_OK) return false;
    outPattern = static_cast<QueryPattern>(p);
    return true;
}

std::vector<Query4D> generateQueries(const VbtFileHeader& header, QueryPattern pattern, uint32_t count, uint32_t seed)
{
    const vbt_file_header h = hip_detail::to_c(header);
    std::vector<uint32_t> rows(4ull * count);
    if (vbt_generate_queries(&h, static_cast<int>(pattern), count, seed, rows.data()) != VBT_OK)
        throw std::runtime_error(vbt_last_error());
    std::vector<Query4D> out(count);
    for (uint32_t i = 0; i < count; ++i) {
        out[i].x = rows[4ull * i];
        out[i].y = rows[4ull * i + 1];
        out[i].z = rows[4ull * i + 2];
        out[i].t = rows[4ull * i + 3];
    }
    return out;
}

void fillLeafIndices(const VbtFileHeader& header, std::vector<Query4D>& queries)
{
    for (auto& q : queries) q.leafIndex = leafIndexForVoxel(header, q.x, q.y, q.z);
}

// the key order of vbt_sort_queries_by_leaf, on the record's own leafIndex
void sortQueriesByLeaf(std::vector<Query4D>& queries)
{
    std::stable_sort(queries.begin(), queries.end(), [](const Query4D& a, const Query4D& b) {
        if (a.leafIndex != b.leafIndex) return a.leafIndex < b.leafIndex;
        if (a.t != b.t) return a.t < b.t;
        if (a.z != b.z) return a.z < b.z;
        if (a.y != b.y) return a.y < b.y;
        return a.x < b.x;
    });
}

} // namespace vbt::render
