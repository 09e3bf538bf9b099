// Drop-in for render/src/vbt_file.h: the VBTPACK4 container types and loader
// with the reference's names and signatures (namespace vbt::render), backed
// by libvbt_hip.so.  loadVbtFile runs the library's container validation
// (the reference's checks and error texts) and keeps the file's arrays on the
// host, as the reference does; vbt::hip::Asset (hip_asset.h) puts one in HBM.
#pragma once

#include "../vbt_hip.h"

#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <system_error>
#include <vector>

namespace vbt::render {

struct VbtFileHeader {  // render/src/vbt_file.h:10-23
    char magic[8]{};
    uint32_t version = 0;
    uint32_t width = 0;
    uint32_t height = 0;
    uint32_t depth = 0;
    uint32_t frames = 0;
    uint32_t leafSize = 0;
    uint32_t coarseResolution = 0;
    uint32_t maxCoarseKeep = 0;
    uint32_t leafCount = 0;
    uint32_t profileType = 0;
    uint32_t coarseAcScaleCount = 0;
};

struct VbtFile {  // render/src/vbt_file.h:25-30
    VbtFileHeader header;
    std::vector<float> coarseAcScales;
    std::vector<uint32_t> offsetsWords;
    std::vector<uint32_t> payloadWords;
};

enum class ScientificMode : uint32_t {
    CoarseOnly = 0,
    SparseEvents = 1,
    DenseGrid3 = 2,
    DenseGrid4 = 3,
};

struct ScientificHeaderV2 {  // render/src/vbt_file.h:39-51
    ScientificMode mode = ScientificMode::CoarseOnly;
    uint32_t coarseKeepMinus1 = 0;
    uint32_t coarseQuantClass = 0;
    uint32_t tierId = 0;
    uint32_t framesPerBinMinus1 = 0;
    uint32_t eventQuantCode = 0;
    uint32_t sparseSubtype = 0;
    uint32_t packedEventCount = 0;
    uint32_t denseQuantBits = 0;
    uint32_t denseSubtype = 0;
    uint32_t denseTemporalKeepMinus1 = 0;
};

enum : uint32_t {
    kScientificDenseSubtypeLegacyPerFrame = 0,
    kScientificDenseSubtypeTemporalBasis = 1,
    kScientificDenseSubtypeTileTemporalBasis = 2,
    kScientificDenseSubtypePatchTemporalBasis = 3,
};

namespace hip_detail {
inline vbt_file_header to_c(const VbtFileHeader& h)
{
    return vbt_file_header{h.version, h.width,   h.height,    h.depth,       h.frames,           h.leafSize,
                           h.coarseResolution, h.maxCoarseKeep, h.leafCount, h.profileType, h.coarseAcScaleCount};
}

// The file image (header, scales, offsets, payload) of a loaded VbtFile.
inline std::vector<uint8_t> image(const VbtFile& f)
{
    std::vector<uint8_t> out(52 + 4 * (f.coarseAcScales.size() + f.offsetsWords.size() + f.payloadWords.size()));
    const vbt_file_header h = to_c(f.header);
    std::memcpy(out.data(), "VBTPACK4", 8);
    std::memcpy(out.data() + 8, &h, 44);
    size_t o = 52;
    std::memcpy(out.data() + o, f.coarseAcScales.data(), 4 * f.coarseAcScales.size());
    o += 4 * f.coarseAcScales.size();
    std::memcpy(out.data() + o, f.offsetsWords.data(), 4 * f.offsetsWords.size());
    o += 4 * f.offsetsWords.size();
    std::memcpy(out.data() + o, f.payloadWords.data(), 4 * f.payloadWords.size());
    return out;
}
} // namespace hip_detail

// loadVbtFile (render/src/vbt_file.cpp:109-210): false + error text on failure.
inline bool loadVbtFile(const std::filesystem::path& path, VbtFile& outFile, std::string& error)
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

// readPackedHeaderWord (render/src/vbt_file.cpp:212-218)
inline uint32_t readPackedHeaderWord(const VbtFile& file, uint32_t leafIndex)
{
    if (leafIndex >= file.header.leafCount) return 0u;
    const uint32_t w = file.offsetsWords[leafIndex];
    return w < file.payloadWords.size() ? file.payloadWords[w] : 0u;
}

inline ScientificHeaderV2 decodeScientificHeaderV2(uint32_t packedHeader)
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

inline uint32_t leafIndexForVoxel(const VbtFileHeader& header, uint32_t x, uint32_t y, uint32_t z)
{
    const vbt_file_header h = hip_detail::to_c(header);
    return vbt_leaf_index_for_voxel(&h, x, y, z);
}

} // namespace vbt::render
