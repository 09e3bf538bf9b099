/*
 * vbt_oracle.h -- CPU restatement of the VBTPACK4 random-access decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle that the HIP
 * product path is checked against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline -- never as the thing measured or shipped.
 *
 * Every function restates one reference function; the citation next to each
 * declaration names the reference file:line it follows.  Pinning status of
 * each function (which golden vector / closed form / brute force checks it)
 * is listed in the header of vbt_oracle.c.
 *
 * Plain C11, no dependencies.  Built with -O2 -ffp-contract=off so that FP32
 * expression order is exactly the reference's.
 */
#ifndef VBT_ORACLE_H
#define VBT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws std::runtime_error; we return) ---- */
enum {
    VBTO_OK = 0,
    VBTO_ERR = 1, /* message via vbto_last_error() */
};

const char* vbto_last_error(void);

/* ---- half <-> float  (src/render_temporal_decode.cpp:135-199) ---- */
uint16_t vbto_float_to_half_bits(float value);
float vbto_half_bits_to_float(uint16_t bits);

/* ---- packed header  (src/render_temporal_payload.cpp:87-107) ---- */
typedef struct {
    uint8_t mode;        /* 0 EMPTY, 1 GRID4, 2 FINE_COMPACT (Mode2 shell), 3 FINE6 */
    uint8_t coarseCodec; /* 0 fp16 keyframes, 1 fp32 keyframes */
    uint8_t fineCodec;
    uint8_t coarseResolution;
    uint8_t fineResolution;
} vbto_header;

uint32_t vbto_pack_header(const vbto_header* h);
void vbto_unpack_header(uint32_t packed, vbto_header* out);
uint32_t vbto_keyframe_bytes(uint32_t codec); /* render_temporal_payload.cpp:32-41 */

/* ---- layout  (src/render_temporal_payload.cpp:43-78, 143-197) ---- */
typedef struct {
    uint32_t descriptorOffset, descriptorBytes;
    uint32_t binIndexOffset, binIndexBytes;
    uint32_t keyframeOffset, keyframeBytes;
    uint32_t totalBytes;
} vbto_stream_layout;

typedef struct {
    uint32_t headerBytes, prefixOffset, prefixBytes;
    uint32_t coarseOffset;
    vbto_stream_layout coarse;
    uint32_t fineOffset;
    vbto_stream_layout fine;
    uint32_t totalBytes;
} vbto_leaf_layout;

void vbto_compute_leaf_layout(const vbto_header* h, uint16_t coarseKeyframeCount,
                              uint16_t fineKeyframeCount, int alignTo4Bytes,
                              vbto_leaf_layout* out);

typedef struct {
    vbto_header header;
    uint16_t coarseKeyframeCount, fineKeyframeCount;
    vbto_leaf_layout layout;
    uint16_t shellActiveVoxelCount;
    uint32_t shellResidualOffset;
    uint32_t shellResidualBytes;
} vbto_leaf_view;

/* src/render_temporal_payload.cpp:234-311 */
int vbto_parse_leaf(const uint8_t* leaf, size_t leafBytes, int alignTo4Bytes, vbto_leaf_view* out);

/* ---- shell occupancy  (src/render_temporal_payload.cpp:316-344) ---- */
typedef struct {
    uint64_t shellMask[8];
    uint16_t groupPrefix[8];
} vbto_shell_section;
void vbto_unpack_shell_section(const uint8_t* data, uint64_t size, vbto_shell_section* out);
uint16_t vbto_shell_active_count(const vbto_shell_section* s);

/* ---- control decode  (src/render_temporal_decode.cpp:79-131, 304-450) ---- */
typedef struct {
    uint32_t lo, hi;   /* keyframe indices (relative to keyStart) of the bracket used */
    int32_t f0, f1;    /* bracket frames */
    uint32_t bin;      /* time bin consulted */
    uint32_t keyStart, keyCount;
    uint32_t kind;     /* 0 zero (keyCount 0), 1 value(lo), 2 value(hi), 3 lerp */
} vbto_trace;

int vbto_decode_control_value(const uint8_t* streamBase, uint32_t codec, uint16_t controlCount,
                              uint8_t timeBinCount, uint8_t maxBinLocalKeys, uint32_t keyframeCount,
                              uint32_t streamBytes, int alignTo4Bytes, int totalFrames,
                              uint16_t controlIndex, int frameIndex, float* out, vbto_trace* trace);

int vbto_decode_leaf_control_value(const uint8_t* leaf, const vbto_leaf_view* view, int useFine,
                                   int alignTo4Bytes, int totalFrames, uint16_t controlIndex,
                                   int frameIndex, float* out, vbto_trace* trace);

int vbto_decode_shell_voxel_value(const uint8_t* leaf, const vbto_leaf_view* view, int alignTo4Bytes,
                                  int totalFrames, uint16_t voxelIndex, int frameIndex, float* out,
                                  uint32_t* descIndexOut);

/* ---- container  (render/src/vbt_file.cpp:104-262) ---- */
typedef struct {
    uint32_t version, width, height, depth, frames, leafSize, coarseResolution, maxCoarseKeep,
        leafCount, profileType, coarseAcScaleCount;
} vbto_file_header;

typedef struct {
    vbto_file_header header;
    const float* coarseAcScales;   /* points into the caller's memory image */
    const uint32_t* offsetsWords;  /* leafCount + 1 */
    const uint32_t* payloadWords;
    uint32_t payloadWordCount;
    /* 0 (as loaded): dense subtypes 2 and 3 decode as the reference's file
     * decoder reads them (as subtype 0).  1: as the encoder reconstructs its
     * tile / patch temporal-basis blocks (vbto_sci_basis_residual). */
    uint32_t sciDenseBasis;
} vbto_file;

/* Validates a whole file image exactly like loadVbtFile (same messages). The
 * image must be 4-byte aligned; vbto_file points into it. */
int vbto_load_file_memory(const uint8_t* bytes, uint64_t size, vbto_file* out);
uint32_t vbto_read_packed_header_word(const vbto_file* f, uint32_t leafIndex);
uint32_t vbto_leaf_index_for_voxel(const vbto_file_header* h, uint32_t x, uint32_t y, uint32_t z);

/* ---- voxel composition  (3D/render_temporal_vbt_to_vdb.cpp:101-289,
 *      render/src/render_smoke_vbt_bench_main.cpp:374-378,
 *      VBTStudio/frontend/shaders/vbt_sample.comp:399-413) ---- */
enum {
    VBTO_SIGN_NONE = 0,        /* exporter default: no clamp (optionally shellEmptyScale) */
    VBTO_SIGN_CLAMP_ZERO = 1,  /* Mode2: !occupied && v<0 -> 0 */
    VBTO_SIGN_BACKGROUND = 2,  /* Mode2: !occupied && v<0 -> background */
};

typedef struct {
    int signRule;
    float background;
    float shellEmptyScale; /* 1 disables */
    float shellEmptyMax;   /* NaN disables */
} vbto_compose;

void vbto_compose_default(vbto_compose* c);

/* Decode one render-temporal leaf at one frame into its cropped voxel block
 * (leafWidth*leafHeight*leafDepth floats, x fastest). */
int vbto_decode_render_leaf_frame(const vbto_file* f, uint32_t leafIndex, int frameIndex,
                                  const vbto_compose* c, float* outLeaf);

/* Whole frame, row-major (z, y, x).  Profile 1 -> render route, profile 2 ->
 * scientific route.  nthreads<=1 is the serial reference loop. */
int vbto_decode_frame_dense(const vbto_file* f, int frameIndex, const vbto_compose* c,
                            float* out, int nthreads);

/* Point queries at integer voxels: q is n x 4 u32 {x,y,z,t}; the leaf is
 * leafIndexForVoxel(x,y,z) and the value equals the whole-frame decode at that
 * voxel.  Out-of-volume voxels (and t >= frames on the scientific route)
 * produce NaN and an error return after the whole batch is written. */
int vbto_query_points(const vbto_file* f, const uint32_t* q, uint64_t n, const vbto_compose* c,
                      float* out, int nthreads);

/* ---- page table (rank order) and frame caches ---- */
/* compact[l] = rank of l among active leaves (header mode != 0), else 0xffffffff.
 * Active predicate: VBTStudio/frontend/shaders/vbt_active_index.comp:33-38. */
uint32_t vbto_page_table(const vbto_file* f, uint32_t* compact);

/* Meaningful control count of a leaf: coarse + (fine6 | shell active), or 512
 * voxels for scientific leaves (cropped voxel count at the boundary). */
uint32_t vbto_leaf_cache_count(const vbto_file* f, uint32_t leafIndex);

/* Stride-576 cache of VBTStudio/frontend/shaders/vbt_cache.comp:168-236 for one
 * frame: out[compact*576 + slot]; unused slots 0. */
int vbto_frame_cache_576(const vbto_file* f, int frameIndex, float* out);

/* Packed cache: out[ctrlOffset[compact] + slot] with the meaningful counts only. */
int vbto_frame_cache_packed(const vbto_file* f, int frameIndex, const uint64_t* ctrlOffset, float* out);

/* ---- viewer sampling and ray marching over a stride-576 frame cache
 *      (VBTStudio/frontend/shaders/vbt_sample.comp, restated in
 *      vbt_oracle_march.c; render assets only) ---- */
typedef struct {
    uint32_t stepCount;  /* march steps per ray */
    int32_t frameIndex;  /* frame the cache holds (direct decode of coarse grids > 64 controls) */
    float bboxMin[3];    /* world position of voxel (0, 0, 0) */
    uint32_t renderMode; /* 0 density integration, 1 level-set surface */
    float voxelSize, backgroundValue, surfaceIso, surfaceEpsilonVoxels, surfaceNormalStep, maxStepVoxels;
    uint32_t surfaceModel; /* density: 8 depth bins; level set: thickness and floor shadow */
    float floorHeight;
    uint32_t shadowSteps;
    uint32_t upAxis;     /* 2: z up, else y up (the shader's pc.padding) */
} vbto_march_params;

/* sampleVbtValue at n world points xyz (3 floats each). */
int vbto_sample_points(const vbto_file* f, const float* cache576, const vbto_march_params* p, const float* xyz,
                       uint64_t n, int preserveInterior, float* out, int nthreads);
/* The shader's main over n rays (8 floats each: origin, tmin, direction,
 * tmax); out gets 8 floats per ray (primary, auxiliary). */
int vbto_march_rays(const vbto_file* f, const float* cache576, const vbto_march_params* p, const float* rays,
                    uint64_t n, float* out, int nthreads);

/* ---- Cycles resident blob and its samplers  (cycles_native_vbt/, blender_patch/native_cycles_vbt.patch) ---- */
#define VBTO_CYCLES_KERNEL 1 /* the patched Cycles kernel's form; else the standalone header's */
#define VBTO_CYCLES_WORLD 2  /* subtract the blob's bbox_min first (sample_density_world) */
/* load_pack4's checks and pack_blob: writes the blob to out (when non-null and
 * capacity allows) and its size to *blobBytes. */
int vbto_cycles_pack_blob(const void* file, uint64_t fileBytes, int frameIndex, void* out, uint64_t capacity,
                          uint64_t* blobBytes);
/* density at n points xyz (3 floats each) read from a blob */
int vbto_cycles_sample(const void* blob, uint64_t blobBytes, int flags, const float* xyz, uint64_t n, float* out,
                       int nthreads);
const char* vbto_cycles_error(void);

/* ---- scientific route  (render/src/vbt_file.cpp:220-248, render/src/scientific_decode.cpp) ---- */
typedef struct {
    uint32_t mode, coarseKeepMinus1, coarseQuantClass, tierId, framesPerBinMinus1, eventQuantCode,
        sparseSubtype, packedEventCount, denseQuantBits, denseSubtype, denseTemporalKeepMinus1;
} vbto_sci_header;

void vbto_decode_sci_header(uint32_t packed, vbto_sci_header* out);
float vbto_dct_basis(int totalLength, int index, int k); /* scientific_decode.cpp:72-80 */
int vbto_sci_coarse_control(const vbto_file* f, uint32_t payloadByteBase, int keep, int controlIndex,
                            int timeIndex, float* out);
int vbto_sci_value_at(const vbto_file* f, uint32_t x, uint32_t y, uint32_t z, uint32_t t,
                      uint32_t leafIndex, float* out);
int vbto_sci_leaf_frame(const vbto_file* f, uint32_t leafIndex, uint32_t frameIndex, float* outLeaf);
/* The residual of a dense subtype-2 (tile) or subtype-3 (patch) temporal-basis
 * block at local voxel (lx, ly, lz) and frame t, read as the encoder's
 * reconstruction reads its in-memory encoding
 * (src/spatial_first_encoder.cpp:2078-2176); the leaf's block starts at
 * payload byte payloadByteBase + 4 + 64 * (keep + 1).  VBTO_ERR for a block
 * that reaches past the pool or that the encoder cannot have written
 * (active count != popcount of the mask, tile size above 8, patch size or
 * local resolution of 1). */
int vbto_sci_basis_residual(const vbto_file* f, uint32_t payloadByteBase, uint32_t headerWord, uint32_t lx,
                            uint32_t ly, uint32_t lz, uint32_t t, float* out);
/* one past the last payload byte (relative to the leaf's first byte) that
 * vbto_sci_basis_residual reads of the leaf, or VBTO_ERR as it would fail */
int vbto_sci_basis_extent(const vbto_file* f, uint32_t payloadByteBase, uint32_t headerWord, uint64_t* end);

#ifdef __cplusplus
}
#endif
#endif
