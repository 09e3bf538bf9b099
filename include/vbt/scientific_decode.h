// Drop-in for render/src/scientific_decode.h and the Query4D record of
// render/src/query_patterns.h:11-17, with the reference's names and
// signatures, decoding on the GPU (defined in libvbt_hip_cxx.so).
//
// Each call uploads the file to HBM, decodes, and frees it again: correct
// for a drop-in, but callers that decode more than once keep a
// vbt::hip::Asset (hip_asset.h) and use its batched queries / frames.
#pragma once

#include "hip_asset.h"
#include "vbt_file.h"

#include <cstdint>
#include <stdexcept>
#include <vector>

// Inside the reference's tree, query_patterns.h (render/src) is on the include
// path and declares Query4D and the query generators (its .cpp stays in the
// build: it is not decode code).  Standalone, the record is defined here.
#if __has_include("query_patterns.h")
#include "query_patterns.h"
#define VBT_HIP_HAVE_QUERY_PATTERNS 1
#endif

namespace vbt::render {

#ifndef VBT_HIP_HAVE_QUERY_PATTERNS
struct Query4D {  // render/src/query_patterns.h:11-17
    uint32_t x = 0;
    uint32_t y = 0;
    uint32_t z = 0;
    uint32_t t = 0;
    uint32_t leafIndex = 0;
};

// fillLeafIndices (render/src/query_patterns.h:31)
inline void fillLeafIndices(const VbtFileHeader& header, std::vector<Query4D>& queries)
{
    for (auto& q : queries) q.leafIndex = leafIndexForVoxel(header, q.x, q.y, q.z);
}
#endif

// decodeScientificValueAtCpu (render/src/scientific_decode.h:11).  A query
// outside the volume or at t >= frames throws.
float decodeScientificValueAtCpu(const VbtFile& file, const Query4D& query);

// reconstructScientificFrameCpu (render/src/scientific_decode.h:12): frame
// t as row-major (z, y, x) floats.
std::vector<float> reconstructScientificFrameCpu(const VbtFile& file, uint32_t frameIndex);

} // namespace vbt::render
