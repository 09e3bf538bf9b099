// ref_capi.cpp -- C-ABI wrapper over the reference's own decode sources, built
// only into oracle/_ref/libvbtref.so (test infrastructure, never shipped).
//
// This file contains no decode logic of its own: every entry point forwards
// to the reference functions named in its comment (one loops over leaves and
// controls the way the reference's CPU frame decode does), catches their
// exceptions and returns the message.  It lets the tests check the oracle restatement against the
// real reference implementation on the same bytes.
#include "query_patterns.h"
#include "render_temporal_decode.h"
#include "render_temporal_payload.h"
#include "scientific_decode.h"
#include "vbt_file.h"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

namespace {
thread_local std::string g_err;
int fail(const std::exception& e)
{
    g_err = e.what();
    return 1;
}
} // namespace

extern "C" {

const char* vbtref_last_error() { return g_err.c_str(); }

// src/render_temporal_decode.cpp:167
float vbtref_half_to_float(uint16_t b) { return vbt::renderTemporalHalfBitsToFloat(b); }
// src/render_temporal_decode.cpp:135
uint16_t vbtref_float_to_half(float v) { return vbt::renderTemporalFloatToHalfBits(v); }

// src/render_temporal_decode.cpp:304
int vbtref_decode_control_value(const uint8_t* stream, uint32_t codec, uint16_t controlCount, uint8_t bins,
                                uint32_t keyframeCount, uint32_t streamBytes, int totalFrames,
                                uint16_t controlIndex, int frameIndex, float* out)
{
    try {
        *out = vbt::decodeRenderTemporalControlValue(stream, static_cast<vbt::RenderTemporalSequenceCodec>(codec),
                                                     controlCount, bins, 0, keyframeCount, streamBytes, true,
                                                     totalFrames, controlIndex, frameIndex);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// src/render_temporal_payload.cpp:234 -- fields flattened into 24 u32
int vbtref_parse_leaf(const uint8_t* leaf, uint64_t bytes, uint32_t* o)
{
    try {
        const auto v = vbt::parseRenderTemporalPackedLeaf(leaf, bytes, true);
        o[0] = static_cast<uint32_t>(v.header.mode);
        o[1] = static_cast<uint32_t>(v.header.coarseCodec);
        o[2] = static_cast<uint32_t>(v.header.fineCodec);
        o[3] = v.header.coarseResolution;
        o[4] = v.header.fineResolution;
        o[5] = v.prefix.coarseKeyframeCount;
        o[6] = v.prefix.fineKeyframeCount;
        o[7] = v.layout.coarseOffset;
        o[8] = v.layout.coarse.totalBytes;
        o[9] = v.layout.fineOffset;
        o[10] = v.layout.fine.descriptorOffset;
        o[11] = v.layout.fine.descriptorBytes;
        o[12] = v.layout.fine.totalBytes;
        o[13] = v.layout.totalBytes;
        o[14] = v.shellActiveVoxelCount;
        o[15] = v.shellResidualOffset;
        o[16] = v.shellResidualBytes;
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// src/render_temporal_decode.cpp:358 (after parse, as every caller does)
int vbtref_leaf_control_value(const uint8_t* leaf, uint64_t bytes, int useFine, int totalFrames,
                              uint16_t controlIndex, int frameIndex, float* out)
{
    try {
        const auto v = vbt::parseRenderTemporalPackedLeaf(leaf, bytes, true);
        *out = vbt::decodeRenderTemporalLeafControlValue(leaf, v.header, v.prefix, v.layout, useFine != 0, true,
                                                         totalFrames, controlIndex, frameIndex);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// src/render_temporal_decode.cpp:404
int vbtref_shell_voxel_value(const uint8_t* leaf, uint64_t bytes, int totalFrames, uint16_t voxel, int frameIndex,
                             float* out)
{
    try {
        const auto v = vbt::parseRenderTemporalPackedLeaf(leaf, bytes, true);
        *out = vbt::decodeRenderTemporalLeafShellVoxelValue(leaf, v, true, totalFrames, voxel, frameIndex);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// render/src/vbt_file.cpp:174 -- returns an opaque handle or null
void* vbtref_load_file(const char* path)
{
    auto* f = new vbt::render::VbtFile();
    std::string err;
    if (!vbt::render::loadVbtFile(path, *f, err)) {
        g_err = err;
        delete f;
        return nullptr;
    }
    return f;
}

void vbtref_free_file(void* h) { delete static_cast<vbt::render::VbtFile*>(h); }

// render/src/scientific_decode.cpp:335
int vbtref_sci_value(void* h, uint32_t x, uint32_t y, uint32_t z, uint32_t t, float* out)
{
    try {
        const auto& f = *static_cast<vbt::render::VbtFile*>(h);
        vbt::render::Query4D q{x, y, z, t, vbt::render::leafIndexForVoxel(f.header, x, y, z)};
        *out = vbt::render::decodeScientificValueAtCpu(f, q);
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// render/src/scientific_decode.cpp:335 over a batch of (x, y, z, t) queries,
// one after the other (the serial query loop of
// render/src/cpu_query_bench_main.cpp:161-170)
int vbtref_sci_values(void* h, const uint32_t* q, uint64_t n, float* out)
{
    try {
        const auto& f = *static_cast<vbt::render::VbtFile*>(h);
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t* p = q + 4u * i;
            vbt::render::Query4D qq{p[0], p[1], p[2], p[3], vbt::render::leafIndexForVoxel(f.header, p[0], p[1], p[2])};
            out[i] = vbt::render::decodeScientificValueAtCpu(f, qq);
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// render/src/scientific_decode.cpp:361
int vbtref_sci_frame(void* h, uint32_t t, float* out)
{
    try {
        const auto& f = *static_cast<vbt::render::VbtFile*>(h);
        const auto frame = vbt::render::reconstructScientificFrameCpu(f, t);
        std::memcpy(out, frame.data(), frame.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// The packed frame cache of frame t through the reference's per-value
// decoders, in the order of the reference's CPU frame decode
// (render/src/render_smoke_vbt_bench_main.cpp:286-330: parse each leaf once,
// then decodeRenderTemporalLeafControlValue for every coarse and fine
// control; Mode2 leaves: decodeRenderTemporalLeafShellVoxelValue for every
// occupied voxel, which in voxel order are the shell stream's controls).
// Active leaves (header mode != EMPTY) are written at ctrlOffset[rank].
int vbtref_cache_frame(void* h, int t, const uint64_t* ctrlOffset, float* out)
{
    try {
        const auto& f = *static_cast<vbt::render::VbtFile*>(h);
        const uint8_t* payload = reinterpret_cast<const uint8_t*>(f.payloadWords.data());
        const int frames = static_cast<int>(f.header.frames);
        uint32_t rank = 0;
        for (uint32_t l = 0; l < f.header.leafCount; ++l) {
            if ((vbt::render::readPackedHeaderWord(f, l) & 3u) == 0u) continue;
            const uint32_t wb = f.offsetsWords[l], we = f.offsetsWords[l + 1u];
            const uint8_t* leaf = payload + static_cast<size_t>(wb) * 4u;
            const auto v = vbt::parseRenderTemporalPackedLeaf(leaf, static_cast<size_t>(we - wb) * 4u, true);
            float* o = out + ctrlOffset[rank++];
            const int cr = v.header.coarseResolution;
            const uint16_t cc = static_cast<uint16_t>(cr * cr * cr);
            for (uint16_t i = 0; i < cc; ++i)
                *o++ = vbt::decodeRenderTemporalLeafControlValue(leaf, v.header, v.prefix, v.layout, false, true,
                                                                 frames, i, t);
            if (v.header.mode == vbt::TemporalFirstPackedMode::TEMPORAL_FINE6) {
                const int fr = v.header.fineResolution;
                const uint16_t fc = static_cast<uint16_t>(fr * fr * fr);
                for (uint16_t i = 0; i < fc; ++i)
                    *o++ = vbt::decodeRenderTemporalLeafControlValue(leaf, v.header, v.prefix, v.layout, true, true,
                                                                     frames, i, t);
            } else if (v.header.mode == vbt::TemporalFirstPackedMode::TEMPORAL_FINE_COMPACT) {
                const auto occ = vbt::unpackRenderTemporalShellOccupancySection(leaf + v.layout.fine.descriptorOffset,
                                                                                v.layout.fine.descriptorBytes);
                for (uint16_t vox = 0; vox < 512; ++vox)
                    if ((occ.shellMask[vox >> 6] >> (vox & 63u)) & 1u)
                        *o++ = vbt::decodeRenderTemporalLeafShellVoxelValue(leaf, v, true, frames, vox, t);
            }
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// render/src/query_patterns.cpp:29-101: generateQueries for a header of
// these dims, then optionally sortQueriesByLeaf (after fillLeafIndices);
// out: (x, y, z, t) rows
int vbtref_generate_queries(uint32_t width, uint32_t height, uint32_t depth, uint32_t frames, int pattern,
                            uint32_t count, uint32_t seed, int sortByLeaf, uint32_t* out)
{
    try {
        vbt::render::VbtFileHeader h{};
        h.width = width;
        h.height = height;
        h.depth = depth;
        h.frames = frames;
        h.leafSize = 8;
        auto q = vbt::render::generateQueries(h, static_cast<vbt::render::QueryPattern>(pattern), count, seed);
        if (sortByLeaf) {
            vbt::render::fillLeafIndices(h, q);
            vbt::render::sortQueriesByLeaf(q);
        }
        for (size_t i = 0; i < q.size(); ++i) {
            out[4 * i] = q[i].x;
            out[4 * i + 1] = q[i].y;
            out[4 * i + 2] = q[i].z;
            out[4 * i + 3] = q[i].t;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

} // extern "C"
