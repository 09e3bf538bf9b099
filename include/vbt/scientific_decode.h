// Drop-in for render/src/scientific_decode.h and the Query4D record of
// render/src/query_patterns.h:11-17, with the reference's names and
// signatures, decoding on the GPU through libvbt_hip.so.
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

namespace vbt::render {

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

// decodeScientificValueAtCpu (render/src/scientific_decode.h:11).  A query
// outside the volume or at t >= frames throws.
inline float decodeScientificValueAtCpu(const VbtFile& file, const Query4D& query)
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

// reconstructScientificFrameCpu (render/src/scientific_decode.h:12): frame
// t as row-major (z, y, x) floats.
inline std::vector<float> reconstructScientificFrameCpu(const VbtFile& file, uint32_t frameIndex)
{
    vbt::hip::Asset asset(file);
    const auto& h = asset.info().header;
    vbt::hip::DeviceBuffer<float> out(static_cast<size_t>(h.width) * h.height * h.depth);
    asset.denseFrame(static_cast<int>(frameIndex), out.data());
    return out.download();
}

} // namespace vbt::render
