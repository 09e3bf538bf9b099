// Drop-in for the two encoder-side types the decode headers name:
// TemporalFirstPackedMode (src/temporal_first_encoder.h:15-20) and the fields
// of TemporalFirstEncoderOptions (src/temporal_first_encoder.h:25-44) that the
// leaf-prefix check reads.  Decode consumers never need the encoder itself.
#pragma once

#include <cstdint>

namespace vbt {

enum class TemporalFirstPackedMode : uint8_t {
    EMPTY = 0,
    TEMPORAL_GRID4 = 1,
    TEMPORAL_FINE_COMPACT = 2,
    TEMPORAL_FINE6 = 3,
};

// Decode-side subset: validateRenderTemporalLeafPrefix reads timeBinCount only.
struct TemporalFirstEncoderOptions {
    float cutoff = 0.0f;
    float cutoffBand = 0.0f;
    int timeBinCount = 8;
    int maxBinLocalKeys = 8;
    bool useGlobalOffsetTable = true;
    bool alignPayloadTo4Bytes = true;
};

} // namespace vbt
