// Drop-in for render/src/query_patterns.h: the query record and the request
// shapes of the reference's query tools, with the reference's names and
// signatures (namespace vbt::render), defined in libvbt_hip_cxx.so.
// generateQueries gives the reference's queries for the same header, pattern,
// count and seed (it is vbt_generate_queries of the C-ABI).
#pragma once

#include "vbt_file.h"

#include <cstdint>
#include <string>
#include <vector>

namespace vbt::render {

struct Query4D {  // render/src/query_patterns.h:11-17
    uint32_t x = 0;
    uint32_t y = 0;
    uint32_t z = 0;
    uint32_t t = 0;
    uint32_t leafIndex = 0;
};

enum class QueryPattern {  // render/src/query_patterns.h:19-24
    RandomFull,
    SameT,
    SameXYZ,
    CoherentTiles,
};

// parseQueryPattern (query_patterns.h:26): "random", "same-t", "same-xyz", "coherent".
bool parseQueryPattern(const std::string& text, QueryPattern& outPattern);
// generateQueries (query_patterns.h:27-30); leafIndex is left 0, as the reference leaves it.
std::vector<Query4D> generateQueries(const VbtFileHeader& header, QueryPattern pattern, uint32_t count,
                                     uint32_t seed);
// fillLeafIndices (query_patterns.h:31)
void fillLeafIndices(const VbtFileHeader& header, std::vector<Query4D>& queries);
// sortQueriesByLeaf (query_patterns.h:32): stable, by (leafIndex, t, z, y, x).
void sortQueriesByLeaf(std::vector<Query4D>& queries);

} // namespace vbt::render
