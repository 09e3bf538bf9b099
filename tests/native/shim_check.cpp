// Consumer-side check of the C++ drop-in headers (include/vbt/*.h): written
// the way a reference consumer calls the reference API, it prints what it
// computes as JSON lines; tests/test_cpp_shim.py compares them with the
// oracle.
//
//   shim_check host <file.vbtp>      container + leaf format (no GPU)
//   shim_check gpu  <file.vbtp> <requests.txt>
//        per-value decodes, one request per line:
//          c <leaf> <useFine> <control> <frame>   decodeRenderTemporalLeafControlValue
//          s <leaf> <voxel> <frame>               decodeRenderTemporalLeafShellVoxelValue
//          q <x> <y> <z> <t>                      decodeScientificValueAtCpu
//          f <t>                                  reconstructScientificFrameCpu (checksum)
#include "../../include/vbt/query_patterns.h"
#include "../../include/vbt/render_temporal_decode.h"
#include "../../include/vbt/scientific_decode.h"
#include "../../include/vbt/vbt_file.h"

#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

using namespace vbt;
using namespace vbt::render;

static const uint8_t* leafBase(const VbtFile& f, uint32_t leaf, size_t& bytes)
{
    const uint32_t w0 = f.offsetsWords[leaf], w1 = f.offsetsWords[leaf + 1];
    bytes = 4ull * (w1 - w0);
    return reinterpret_cast<const uint8_t*>(f.payloadWords.data() + w0);
}

static uint32_t bits(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
}

static int host(const VbtFile& f)
{
    // half conversion over every non-NaN pattern: checksum of the float bits
    uint64_t sum = 0;
    for (uint32_t h = 0; h < 65536u; ++h)
        if ((h & 0x7C00u) != 0x7C00u || (h & 0x3FFu) == 0u)
            sum = sum * 1000003u + bits(renderTemporalHalfBitsToFloat((uint16_t)h));
    std::printf("{\"half_checksum\": %" PRIu64 "}\n", sum);
    const auto& h = f.header;
    std::printf("{\"header\": [%u, %u, %u, %u, %u, %u, %u]}\n", h.version, h.width, h.height, h.depth, h.frames,
                h.leafCount, h.profileType);
    const uint32_t vox[6][3] = {{0, 0, 0}, {h.width - 1, 0, 0}, {0, h.height - 1, h.depth - 1},
                                {h.width, h.height, h.depth}, {7, 8, 9}, {h.width / 2, h.height / 3, h.depth / 4}};
    for (auto& v : vox) std::printf("{\"leaf_of\": [%u, %u, %u, %u]}\n", v[0], v[1], v[2], leafIndexForVoxel(h, v[0], v[1], v[2]));
    // request shapes as the reference's query tools make them: generate, then
    // leaf indices and the leaf-major order
    for (const char* name : {"random", "same-t", "same-xyz", "coherent"}) {
        QueryPattern p;
        if (!parseQueryPattern(name, p)) std::printf("{\"error\": \"pattern %s\"}\n", name);
        auto q = generateQueries(h, p, 1000, 77u);
        fillLeafIndices(h, q);
        sortQueriesByLeaf(q);
        uint64_t qs = 0;
        for (const auto& r : q) qs = ((((qs * 1000003u + r.x) * 1000003u + r.y) * 1000003u + r.z) * 1000003u + r.t) * 1000003u + r.leafIndex;
        std::printf("{\"queries\": [\"%s\", %" PRIu64 "]}\n", name, qs);
    }
    for (uint32_t l = 0; l < h.leafCount; ++l) {
        const uint32_t word = readPackedHeaderWord(f, l);
        if (h.profileType == 2) {
            const ScientificHeaderV2 s = decodeScientificHeaderV2(word);
            std::printf("{\"sci\": [%u, %u, %u, %u, %u, %u, %u, %u, %u, %u, %u, %u]}\n", l, (uint32_t)s.mode,
                        s.coarseKeepMinus1, s.coarseQuantClass, s.tierId, s.framesPerBinMinus1, s.eventQuantCode,
                        s.sparseSubtype, s.packedEventCount, s.denseQuantBits, s.denseSubtype,
                        s.denseTemporalKeepMinus1);
            continue;
        }
        const RenderTemporalPackedHeaderFields hf = unpackRenderTemporalHeader(word);
        if (packRenderTemporalHeader(hf) != word && hf.mode != TemporalFirstPackedMode::EMPTY) {
            std::printf("{\"error\": \"header round trip leaf %u\"}\n", l);
            return 1;
        }
        if (hf.mode == TemporalFirstPackedMode::EMPTY) continue;
        size_t bytes = 0;
        const uint8_t* b = leafBase(f, l, bytes);
        const RenderTemporalPackedLeafView v = parseRenderTemporalPackedLeaf(b, bytes, true);
        const RenderTemporalPackedLeafLayout lay = computeRenderTemporalLeafLayout(v.header, v.prefix, true);
        std::printf("{\"leaf\": [%u, %u, %u, %u, %u, %u, %u, %u, %u, %u, %u]}\n", l, (uint32_t)v.header.mode,
                    v.prefix.coarseKeyframeCount, v.prefix.fineKeyframeCount, v.layout.coarse.totalBytes,
                    v.layout.fine.totalBytes, v.layout.totalBytes, (uint32_t)v.shellActiveVoxelCount,
                    v.shellResidualOffset, v.shellResidualBytes, lay.totalBytes);
        // the coarse stream through the stream model and back
        const RenderTemporalControlStreamData s = unpackRenderTemporalControlStream(
            b + v.layout.coarseOffset, v.header.coarseCodec, (uint16_t)(v.header.coarseResolution * v.header.coarseResolution * v.header.coarseResolution),
            kRenderTemporalPackedTimeBinCount, 255, v.prefix.coarseKeyframeCount, v.layout.coarse.totalBytes, true);
        const std::vector<uint8_t> re = packRenderTemporalControlStream(s, true);
        if (!validateRenderTemporalControlStreamData(s) || re.size() != v.layout.coarse.totalBytes ||
            std::memcmp(re.data(), b + v.layout.coarseOffset, re.size()) != 0) {
            std::printf("{\"error\": \"coarse stream round trip leaf %u\"}\n", l);
            return 1;
        }
    }
    // a reference error text through the shim
    try {
        (void)makeRenderTemporalLeafPrefix(TemporalFirstPackedMode::TEMPORAL_GRID4, 70000u, 0u);
        std::printf("{\"error\": \"prefix overflow accepted\"}\n");
        return 1;
    } catch (const std::runtime_error& e) {
        std::printf("{\"prefix_error\": \"%s\"}\n", e.what());
    }
    return 0;
}

static int gpu(const VbtFile& f, const char* requests)
{
    std::ifstream in(requests);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        char kind;
        ss >> kind;
        try {
            if (kind == 'c' || kind == 's') {
                uint32_t leaf, a, b2 = 0;
                int t;
                if (kind == 'c') ss >> leaf >> a >> b2 >> t;
                else ss >> leaf >> a >> t;
                size_t bytes = 0;
                const uint8_t* b = leafBase(f, leaf, bytes);
                const RenderTemporalPackedLeafView v = parseRenderTemporalPackedLeaf(b, bytes, true);
                const float val =
                    kind == 'c'
                        ? decodeRenderTemporalLeafControlValue(b, v.header, v.prefix, v.layout, a != 0, true,
                                                               (int)f.header.frames, (uint16_t)b2, t)
                        : decodeRenderTemporalLeafShellVoxelValue(b, v, true, (int)f.header.frames, (uint16_t)a, t);
                std::printf("{\"v\": %u}\n", bits(val));
            } else if (kind == 'q') {
                Query4D q;
                ss >> q.x >> q.y >> q.z >> q.t;
                std::printf("{\"v\": %u}\n", bits(decodeScientificValueAtCpu(f, q)));
            } else if (kind == 'f') {
                uint32_t t;
                ss >> t;
                const std::vector<float> fr = reconstructScientificFrameCpu(f, t);
                uint64_t sum = 0;
                for (float x : fr) sum = sum * 1000003u + bits(x);
                std::printf("{\"frame\": %" PRIu64 ", \"n\": %zu}\n", sum, fr.size());
            }
        } catch (const std::runtime_error& e) {
            std::printf("{\"error\": \"%s\"}\n", e.what());
        }
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: shim_check host|gpu file.vbtp [requests]\n");
        return 2;
    }
    VbtFile f;
    std::string err;
    if (!loadVbtFile(argv[2], f, err)) {
        std::printf("{\"load_error\": \"%s\"}\n", err.c_str());
        return 0;
    }
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    if (std::strcmp(argv[1], "host") == 0) {
        try {
            return host(f);
        } catch (const std::exception& e) {
            std::printf("{\"error\": \"%s\"}\n", e.what());
            return 1;
        }
    }
    if (std::strcmp(argv[1], "gpu") == 0 && argc >= 4) return gpu(f, argv[3]);
    return 2;
}
