// Drop-in for render/src/scientific_decode.h, with the reference's names and
// signatures, decoding on the GPU (defined in libvbt_hip_cxx.so); the query
// record is query_patterns.h's.
//
// Each call uploads the file to HBM, decodes, and frees it again: correct
// for a drop-in, but callers that decode more than once keep a
// vbt::hip::Asset (hip_asset.h) and use its batched queries / frames.
#pragma once

#include "hip_asset.h"
#include "query_patterns.h"
#include "vbt_file.h"

#include <cstdint>
#include <stdexcept>
#include <vector>

namespace vbt::render {

// decodeScientificValueAtCpu (render/src/scientific_decode.h:11).  A query
// outside the volume or at t >= frames throws.
float decodeScientificValueAtCpu(const VbtFile& file, const Query4D& query);

// reconstructScientificFrameCpu (render/src/scientific_decode.h:12): frame
// t as row-major (z, y, x) floats.
std::vector<float> reconstructScientificFrameCpu(const VbtFile& file, uint32_t frameIndex);

} // namespace vbt::render
