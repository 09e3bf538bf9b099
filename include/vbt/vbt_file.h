// Drop-in for render/src/vbt_file.h: the VBTPACK4 container types and loader
// with the reference's names and signatures (namespace vbt::render), defined
// in libvbt_hip_cxx.so over libvbt_hip.so.  loadVbtFile runs the library's
// container validation (the reference's checks and error texts) and keeps the
// file's arrays on the host, as the reference does; vbt::hip::Asset
// (hip_asset.h) puts one in HBM.
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
bool loadVbtFile(const std::filesystem::path& path, VbtFile& outFile, std::string& error);

// readPackedHeaderWord (render/src/vbt_file.cpp:212-218)
uint32_t readPackedHeaderWord(const VbtFile& file, uint32_t leafIndex);

ScientificHeaderV2 decodeScientificHeaderV2(uint32_t packedHeader);

uint32_t leafIndexForVoxel(const VbtFileHeader& header, uint32_t x, uint32_t y, uint32_t z);

} // namespace vbt::render
