// ref_sf_basis.cpp -- C-ABI over the reference encoder's own reconstruction of
// scientific dense subtypes 2 (tile temporal basis) and 3 (patch temporal
// basis), built only into oracle/_ref/libvbtref_sf.so (test infrastructure,
// never shipped).
//
// The reference's file decoder has no reader for these subtypes; its encoder
// reconstructs them from its in-memory encodings with functions local to
// src/spatial_first_encoder.cpp.  This file includes that source so that it
// can call them:
//   decodeDenseTileTemporalBasisAt   src/spatial_first_encoder.cpp:2094
//   decodeDensePatchTemporalBasisAt  src/spatial_first_encoder.cpp:2135
//   appendPackedDenseTileTemporalBasis / appendPackedDensePatchTemporalBasis
//                                    src/spatial_first_encoder.cpp:426-458
// It fills the encoding struct from the bytes of a leaf's block, checks that
// the encoder's own append function writes exactly those bytes back (so the
// fields were read where the encoder puts them), then evaluates the encoder's
// decode function at every voxel of the leaf for one frame.
#include "spatial_first_encoder.cpp"

#include "render_temporal_decode.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace {

uint64_t le(const uint8_t* p, int n)
{
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}

template <class E>
int fill_common(E& e, const uint8_t* block, uint64_t avail, uint64_t hdr, int cells, int tk, int bits,
                uint64_t* used)
{
    e.temporalKeep = tk;
    e.bitsPerValue = bits;
    e.valuesPerBasis = e.valuesPerBasis * cells;
    e.basisBytes = (std::max(1, e.valuesPerBasis) * e.bitsPerValue + 7) / 8;
    uint64_t n = hdr;
    if (e.temporalKeep > 0 && e.valuesPerBasis > 0) {
        n += 2ull * (uint64_t)tk + (uint64_t)tk * (uint64_t)e.basisBytes;
        if (n > avail) return 1;
        for (int k = 0; k < tk; ++k)
            e.basisScales.push_back(vbt::renderTemporalHalfBitsToFloat((uint16_t)le(block + hdr + 2 * k, 2)));
        const uint8_t* data = block + hdr + 2ull * (uint64_t)tk;
        e.packedBasisValues.assign(data, data + (size_t)tk * (size_t)e.basisBytes);
    }
    *used = n;
    return 0;
}

} // namespace

extern "C" {

// block: the leaf's bytes from its dense block on (avail of them readable);
// headerWord: the leaf's packed header word; out: 512 floats, voxel
// (z * 8 + y) * 8 + x.  Returns 0, 1 when the block reaches past avail, 2
// when the encoder's append function does not write the block back, 3 when
// the leaf is not a subtype-2/3 dense leaf.
int vbtref_sf_basis_leaf(const uint8_t* block, uint64_t avail, uint32_t headerWord, int frames, int t, float* out,
                         uint64_t* blockBytes)
{
    const uint32_t mode = headerWord & 3u, subtype = (headerWord >> 12) & 7u;
    if (mode < 2u || (subtype != 2u && subtype != 3u)) return 3;
    const int bits = (int)((headerWord >> 8) & 0xFu), tk = (int)((headerWord >> 15) & 0xFu) + 1;
    std::vector<uint8_t> back;
    uint64_t used = 0;
    if (subtype == 2u) {
        if (avail < 12) return 1;
        vbt::DenseTileTemporalBasisEncoding e;
        e.activeTileMask = le(block, 8);
        e.tileSize = (int)le(block + 8, 2);
        e.activeTileCount = (int)le(block + 10, 2);
        e.valuesPerBasis = e.activeTileCount;
        if (fill_common(e, block, avail, 12, e.tileSize * e.tileSize * e.tileSize, tk, bits, &used)) return 1;
        vbt::appendPackedDenseTileTemporalBasis(back, e);
        for (int z = 0; z < 8; ++z)
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x)
                    out[(z * 8 + y) * 8 + x] = vbt::decodeDenseTileTemporalBasisAt(e, x, y, z, t, frames);
    } else {
        if (avail < 14) return 1;
        vbt::DensePatchTemporalBasisEncoding e;
        e.activePatchMask = le(block, 8);
        e.patchSize = (int)le(block + 8, 2);
        e.localResolution = (int)le(block + 10, 2);
        e.activePatchCount = (int)le(block + 12, 2);
        e.valuesPerBasis = e.activePatchCount;
        const int r = e.localResolution;
        if (fill_common(e, block, avail, 14, r * r * r, tk, bits, &used)) return 1;
        vbt::appendPackedDensePatchTemporalBasis(back, e);
        for (int z = 0; z < 8; ++z)
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x)
                    out[(z * 8 + y) * 8 + x] = vbt::decodeDensePatchTemporalBasisAt(e, x, y, z, t, frames);
    }
    *blockBytes = used;
    if (back.size() != used || std::memcmp(back.data(), block, used) != 0) return 2;
    return 0;
}

} // extern "C"
