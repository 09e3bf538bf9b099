// vbt_synth.h -- synthetic 4D fields and a VBTPACK4 encoder for making test
// and benchmark assets.  Tooling, not the decode path: it is what produces
// the large assets on a machine that has no reference encoder.  The streams it
// writes follow the VBTPACK4 layout that the decoder reads (header word,
// prefix, descriptor/bin/keyframe streams, Mode2 occupancy, scientific
// DCT + sparse/dense residual blocks); they are format-valid but not
// byte-identical to the reference encoder's output.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VBTS_DENSITY = 0,     // advected Gaussian blobs, zero-clamped
    VBTS_FLAMES = 1,      // narrow band of the density
    VBTS_TEMPERATURE = 2, // 300 + 1500 * band(density smoothed by sigma 2)
    VBTS_LEVELSET = 3,    // smooth-min of falling sphere SDFs
    VBTS_SCIENTIFIC = 4,  // sum of sines + sparse spikes
};

typedef struct {
    int kind;
    uint32_t width, height, depth, frames;
    uint32_t count;      // blobs / spheres / sine terms
    float sizeMin, sizeMax;  // sigma range (blobs) or radius range (spheres)
    float peakMin, peakMax;  // blob peaks
    float noise;         // relative uniform noise inside blob support (0.01 = 1%)
    float spikeFraction; // scientific: fraction of voxel-frames carrying a spike
    float spikeGain;     // scientific: spike amplitude relative to the field amplitude
    float motion;        // path amplitude scale
    uint64_t seed;
    float support;       // blob support radius in sigmas (values outside are zero)
} vbts_params;

void vbts_default_params(int kind, uint32_t w, uint32_t h, uint32_t d, uint32_t f, uint64_t seed, vbts_params* p);

void* vbts_field_create(const vbts_params* p);
void vbts_field_free(void* field);
// Rasterise frame t at integer voxels, row-major (z, y, x).
int vbts_field_frame(void* field, uint32_t t, float* out, int threads);
float vbts_field_eval(void* field, float x, float y, float z, uint32_t t);

// Render fields are encoded the way the reference's render encoder does it
// (src/render_temporal_kf_probe_main.cpp): keyframes by the adaptive
// piecewise-linear split of its density detector with cutoff protection, and
// each non-empty leaf routed to Fine6 or Grid4 by its sampled coarse error.
typedef struct {
    int codec;            // 0 fp16 keyframes, 1 fp32
    float epsAbs;         // keyframe tolerance at frame t: epsAbs + epsRel * max(|v_t|, stddev),
    float epsRel;         //   divided by 1 + gammaDelta * |v_t - v_t-1| / (90th percentile step)
    float gammaDelta;
    float cutoff;         // render cutoff: frames in [cutoff - band, cutoff + band] or crossing it keep keys
    float cutoffBand;
    float routeCoarseRmse; // Fine6 when the coarse error (relative to the field's range) exceeds either
    float routeCoarsePeak; //   of these, the leaf is visible enough and the 6^3 residual gains enough
    float routeVisible;   // fraction of samples >= cutoff - band
    float routeBand;      // fraction of samples inside the cutoff band
    float routeFineGain;  // (coarse RMSE - Fine6 RMSE) / coarse RMSE
    float shellBand;      // Mode2 surface band: shell voxels come within it of phi = 0 (guard band: half of it, at least 0.5)
    float burstFraction;  // fraction of controls that receive extra burst keyframes
    uint32_t burstMax;    // burst keyframes per control 1..burstMax-1 extra
    int sciKeep;          // scientific: DCT terms kept per control
    int threads;
} vbts_encode_opts;

void vbts_default_encode_opts(vbts_encode_opts* o);

// Encode the whole field into a VBTPACK4 image (allocated with malloc; free
// with vbts_free).  Scientific fields are encoded with profileType 2, others
// with profileType 1.  Returns 0 on success.
int vbts_encode(void* field, const vbts_encode_opts* o, uint8_t** outBytes, uint64_t* outSize);
void vbts_free(void* p);
const char* vbts_last_error(void);

#ifdef __cplusplus
}
#endif
