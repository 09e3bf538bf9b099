/*
 * vbt_hip.h -- C-ABI of libvbt_hip.so, the MI355X (gfx950) VBTPACK4 decoder.
 *
 * Plain C types only (pointers, sizes, PODs).  Device pointers are HIP device
 * allocations (vbt_device_alloc, or any hipMalloc'd memory); work is ordered on
 * the caller's hipStream_t passed as void* (NULL = the null stream).
 * Every entry point returns VBT_OK (0) or an error code; the message of the
 * last error on the calling thread is vbt_last_error().  Error texts of format
 * failures are the reference's exception texts.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * the reference repository).  The C++ drop-in headers in include/vbt/ wrap
 * these with the reference's own names and signatures.
 */
#ifndef VBT_HIP_H
#define VBT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status -- */
enum {
    VBT_OK = 0,
    VBT_ERR_INVALID = 1,   /* bad argument */
    VBT_ERR_FORMAT = 2,    /* VBTPACK4 / leaf / stream format error (reference exception text) */
    VBT_ERR_DEVICE = 3,    /* HIP runtime error */
    VBT_ERR_NOMEM = 4,     /* device allocation failed */
    VBT_ERR_BUDGET = 5,    /* server admission: request exceeds the HBM budget */
    VBT_ERR_AGAIN = 6,     /* server: queue full / result not ready */
    VBT_ERR_NO_DEVICE = 7, /* no HIP device present */
};

const char* vbt_last_error(void);
/* Library version string and the offload target it was built for. */
const char* vbt_version(void);

/* ------------------------------------------------- format (host, pure) -- */
/* Mirrors of src/render_temporal_payload.h and src/render_temporal_decode.h
 * value types. */
typedef struct {
    uint8_t mode;        /* 0 EMPTY, 1 TEMPORAL_GRID4, 2 TEMPORAL_FINE_COMPACT (Mode2), 3 TEMPORAL_FINE6 */
    uint8_t coarseCodec; /* 0 DP_KEYFRAME_FP16, 1 DP_KEYFRAME_FP32 */
    uint8_t fineCodec;
    uint8_t coarseResolution;
    uint8_t fineResolution;
} vbt_header_fields;         /* RenderTemporalPackedHeaderFields, render_temporal_payload.h:21-27 */

typedef struct {
    uint16_t coarseKeyframeCount;
    uint16_t fineKeyframeCount;
} vbt_leaf_prefix;           /* RenderTemporalPackedLeafPrefix, render_temporal_payload.h:32-35 */

typedef struct {
    uint32_t descriptorOffset, descriptorBytes, binIndexOffset, binIndexBytes, keyframeOffset, keyframeBytes,
        totalBytes;
} vbt_stream_layout;         /* RenderTemporalStreamLayout, render_temporal_payload.h:40-48 */

typedef struct {
    uint32_t headerBytes, prefixOffset, prefixBytes, coarseOffset;
    vbt_stream_layout coarse;
    uint32_t fineOffset;
    vbt_stream_layout fine;
    uint32_t totalBytes;
} vbt_leaf_layout;           /* RenderTemporalPackedLeafLayout, render_temporal_payload.h:50-59 */

typedef struct {
    vbt_header_fields header;
    vbt_leaf_prefix prefix;
    vbt_leaf_layout layout;
    uint16_t shellActiveVoxelCount;
    uint32_t shellResidualOffset;
    uint32_t shellResidualBytes;
} vbt_leaf_view;             /* RenderTemporalPackedLeafView, render_temporal_payload.h:61-68 */

typedef struct {
    uint64_t shellMask[8];
    uint16_t groupPrefix[8];
} vbt_shell_section;         /* RenderTemporalShellOccupancySection, render_temporal_payload.h:87-90 */

typedef struct {
    uint16_t keyStart;
    uint8_t keyCount;
    uint8_t reserved;
} vbt_control_descriptor;    /* RenderTemporalControlDescriptor, render_temporal_decode.h:11-15 */

typedef struct {
    uint16_t frame;
    float value;
} vbt_control_keyframe;      /* RenderTemporalControlKeyframe, render_temporal_decode.h:17-20 */

/* RenderTemporalControlStreamData (render_temporal_decode.h:25-33) with
 * caller-owned arrays. */
typedef struct {
    uint8_t codec;
    uint16_t controlCount;
    uint8_t timeBinCount;
    uint8_t maxBinLocalKeys;
    vbt_control_descriptor* descriptors; /* controlCount */
    uint8_t* binIndex;                   /* controlCount * timeBinCount */
    vbt_control_keyframe* keyframes;     /* keyframeCount */
    uint32_t keyframeCount;
} vbt_stream_data;

uint16_t vbt_float_to_half_bits(float value);   /* renderTemporalFloatToHalfBits, render_temporal_decode.h:35 */
float vbt_half_bits_to_float(uint16_t bits);    /* renderTemporalHalfBitsToFloat, render_temporal_decode.h:36 */
uint32_t vbt_keyframe_bytes(uint32_t codec);    /* renderTemporalKeyframeBytes, render_temporal_payload.h:103 */
uint32_t vbt_pack_header(const vbt_header_fields* fields);                /* render_temporal_payload.h:105 */
void vbt_unpack_header(uint32_t packedHeader, vbt_header_fields* out);   /* render_temporal_payload.h:106 */
/* makeRenderTemporalLeafPrefix, render_temporal_payload.h:108 (throws above u16 -> VBT_ERR_FORMAT) */
int vbt_make_leaf_prefix(uint32_t mode, uint32_t coarseKeyframeCount, uint32_t fineKeyframeCount,
                         vbt_leaf_prefix* out);
/* validateRenderTemporalLeafPrefix, render_temporal_payload.h:112 (options reduced to timeBinCount) */
int vbt_validate_leaf_prefix(const vbt_leaf_prefix* prefix, const vbt_header_fields* header, int timeBinCount);
/* computeRenderTemporalLeafLayout, render_temporal_payload.h:116 */
void vbt_compute_leaf_layout(const vbt_header_fields* header, const vbt_leaf_prefix* prefix, int alignTo4Bytes,
                             vbt_leaf_layout* out);
/* buildRenderTemporalPackedLeafBytes (5-arg), render_temporal_payload.h:120.  Writes at most outCap bytes;
 * *outLen receives the full size (call with out = NULL to size). */
int vbt_build_leaf_bytes(const vbt_header_fields* header, const vbt_leaf_prefix* prefix, const uint8_t* coarse,
                         size_t coarseLen, const uint8_t* fine, size_t fineLen, int alignTo4Bytes, uint8_t* out,
                         size_t outCap, size_t* outLen);
/* buildRenderTemporalPackedLeafBytes (6-arg, Mode2), render_temporal_payload.h:128 */
int vbt_build_shell_leaf_bytes(const vbt_header_fields* header, const vbt_leaf_prefix* prefix, const uint8_t* coarse,
                               size_t coarseLen, const uint8_t* occupancy, size_t occupancyLen,
                               const uint8_t* residual, size_t residualLen, int alignTo4Bytes, uint8_t* out,
                               size_t outCap, size_t* outLen);
void vbt_pack_shell_section(const vbt_shell_section* s, uint8_t out[80]);        /* render_temporal_payload.h:135 */
void vbt_unpack_shell_section(const uint8_t* data, uint64_t size, vbt_shell_section* out); /* :136 */
uint16_t vbt_shell_active_count(const vbt_shell_section* s);                      /* :137 */
/* parseRenderTemporalPackedLeaf, render_temporal_payload.h:139 */
int vbt_parse_leaf(const uint8_t* leafBase, size_t leafBytes, int alignTo4Bytes, vbt_leaf_view* out);
/* validateRenderTemporalControlStreamData, render_temporal_decode.h:38 -> 1 valid, 0 invalid */
int vbt_validate_control_stream(const vbt_stream_data* stream);
/* packRenderTemporalControlStream, render_temporal_decode.h:39 */
int vbt_pack_control_stream(const vbt_stream_data* stream, int alignTo4Bytes, uint8_t* out, size_t outCap,
                            size_t* outLen);
/* unpackRenderTemporalControlStream, render_temporal_decode.h:42; out's arrays are caller-allocated
 * (controlCount descriptors, controlCount*timeBinCount bins, keyframeCount keyframes). */
int vbt_unpack_control_stream(const uint8_t* streamBase, uint32_t codec, uint16_t controlCount, uint8_t timeBinCount,
                              uint8_t maxBinLocalKeys, uint32_t keyframeCount, uint32_t streamBytes, int alignTo4Bytes,
                              vbt_stream_data* out);

/* --------------------------------------------- per-value decode (GPU) -- */
/* Batched forms of the reference's per-value decoders: request i decodes
 * (controlIndex[i], frameIndex[i]) and writes out[i]; status[i] = 0 or a
 * VBT_LEAF_* code (the value is then NaN).  All pointers are device memory. */
/* decodeRenderTemporalControlValue, render_temporal_decode.h:51 */
int vbt_decode_control_values(const uint8_t* d_stream, uint32_t codec, uint16_t controlCount, uint8_t timeBinCount,
                              uint8_t maxBinLocalKeys, uint32_t keyframeCount, uint32_t streamBytes, int alignTo4Bytes,
                              int totalFrames, const uint16_t* d_controlIndex, const int32_t* d_frameIndex, uint32_t n,
                              float* d_out, int32_t* d_status, void* stream);
/* decodeRenderTemporalLeafControlValue, render_temporal_decode.h:63 (leaf parsed on the device) */
int vbt_decode_leaf_control_values(const uint8_t* d_leaf, uint32_t leafBytes, int useFineStream, int alignTo4Bytes,
                                   int totalFrames, const uint16_t* d_controlIndex, const int32_t* d_frameIndex,
                                   uint32_t n, float* d_out, int32_t* d_status, void* stream);
/* decodeRenderTemporalLeafShellVoxelValue, render_temporal_decode.h:76 */
int vbt_decode_shell_voxel_values(const uint8_t* d_leaf, uint32_t leafBytes, int alignTo4Bytes, int totalFrames,
                                  const uint16_t* d_voxelIndex, const int32_t* d_frameIndex, uint32_t n, float* d_out,
                                  int32_t* d_status, void* stream);

/* Leaf status codes written to d_status / reported per leaf. */
enum {
    VBT_LEAF_OK = 0,
    VBT_LEAF_ERR_TOO_SMALL = 1,
    VBT_LEAF_ERR_PREFIX_MISSING = 2,
    VBT_LEAF_ERR_SHORT = 3,
    VBT_LEAF_ERR_OCCUPANCY_OOB = 4,
    VBT_LEAF_ERR_GROUP_PREFIX = 5,
    VBT_LEAF_ERR_NO_ACTIVE = 6,
    VBT_LEAF_ERR_RESIDUAL_OOB = 7,
    VBT_LEAF_ERR_TRAILING = 8,
    VBT_LEAF_ERR_KEY_RANGE = 9,
    VBT_LEAF_ERR_STREAM_OOB = 10,
    VBT_LEAF_ERR_SCI_RANGE = 11,
    VBT_LEAF_ERR_SCI_BASIS = 12,     /* with VBT_OPEN_SCI_DENSE_BASIS: a tile / patch basis block whose active
                                        count is not its mask's popcount, tile size above 8, or patch size or
                                        local resolution of 1 */
    VBT_QUERY_ERR_OUTSIDE = 32,      /* point query voxel outside the volume */
    VBT_QUERY_ERR_FRAME = 33,        /* scientific query frame >= frames */
    VBT_ERR_CONTROL_INDEX = 34,      /* control index >= controlCount */
};
/* The reference's exception text for a status code. */
const char* vbt_status_message(int32_t status);

/* ------------------------------------------------- assets (container) -- */
typedef struct {
    uint32_t version, width, height, depth, frames, leafSize, coarseResolution, maxCoarseKeep, leafCount,
        profileType, coarseAcScaleCount;
} vbt_file_header;           /* VbtFileHeader, render/src/vbt_file.h:10-23 */

/* loadVbtFile's validation (render/src/vbt_file.cpp:39-105, 109-210) over an in-memory
 * image: same checks, same messages. */
int vbt_validate_file(const void* bytes, uint64_t size, vbt_file_header* out);
/* leafIndexForVoxel, render/src/vbt_file.h:63 */
uint32_t vbt_leaf_index_for_voxel(const vbt_file_header* h, uint32_t x, uint32_t y, uint32_t z);

typedef struct {
    uint32_t mode; /* 0 CoarseOnly, 1 SparseEvents, 2 DenseGrid3, 3 DenseGrid4 */
    uint32_t coarseKeepMinus1, coarseQuantClass, tierId, framesPerBinMinus1, eventQuantCode, sparseSubtype,
        packedEventCount, denseQuantBits, denseSubtype, denseTemporalKeepMinus1;
} vbt_scientific_header;     /* ScientificHeaderV2, render/src/vbt_file.h:39-51 */
/* decodeScientificHeaderV2, render/src/vbt_file.h:62 */
void vbt_decode_scientific_header(uint32_t packedHeader, vbt_scientific_header* out);

/* Query request shapes, render/src/query_patterns.h:19-33 (host).  Queries
 * are (x, y, z, t) rows of four u32, the layout vbt_query_points reads. */
enum {
    VBT_QUERY_RANDOM = 0,   /* QueryPattern::RandomFull: uniform x, y, z, t */
    VBT_QUERY_SAME_T = 1,   /* SameT: uniform x, y, z at one frame */
    VBT_QUERY_SAME_XYZ = 2, /* SameXYZ: one voxel at uniform frames */
    VBT_QUERY_COHERENT = 3, /* CoherentTiles: 8x8x8 tiles of one random leaf-aligned corner per query, one frame */
};
/* parseQueryPattern, query_patterns.h:28 ("random", "same-t", "same-xyz",
 * "coherent"); VBT_ERR_INVALID for other text. */
int vbt_parse_query_pattern(const char* text, int* pattern);
/* generateQueries, query_patterns.h:29-32: the same queries as the reference
 * for the same header, pattern, count and seed (std::mt19937 seeded with
 * `seed`, std::uniform_int_distribution per axis).  out: 4 * count u32. */
int vbt_generate_queries(const vbt_file_header* h, int pattern, uint32_t count, uint32_t seed, uint32_t* out);
/* sortQueriesByLeaf, query_patterns.h:34: stable sort by (leaf, t, z, y, x),
 * the leaf from leafIndexForVoxel. */
int vbt_sort_queries_by_leaf(const vbt_file_header* h, uint32_t* queries, uint32_t n);

typedef struct vbt_asset_s* vbt_asset;

/* Validate (as loadVbtFile), upload the offset table, payload pool and AC
 * scales to HBM, then build the page table on the device: the compact index
 * of every leaf (rank among active leaves in leaf order; 0xffffffff when
 * EMPTY -- the deterministic form of vbt_active_index.comp), the per-leaf
 * decode plan, the packed-cache control offsets and the per-leaf status.
 * Runs on the given stream and synchronises it once at the end. */
int vbt_asset_open_memory(const void* bytes, uint64_t size, void* stream, vbt_asset* out);
int vbt_asset_open_file(const char* path, void* stream, vbt_asset* out);
/* Open flags.  VBT_OPEN_SCI_DENSE_BASIS: scientific leaves with dense
 * subtype 2 (tile temporal basis) or 3 (patch temporal basis) decode as the
 * reference encoder reconstructs its encodings
 * (decodeDenseTileTemporalBasisAt / decodeDensePatchTemporalBasisAt,
 * src/spatial_first_encoder.cpp:2094-2176) from the block it writes
 * (src/spatial_first_encoder.cpp:426-458): value = coarse + that residual.
 * Without it these subtypes decode as the reference's file decoder reads them
 * (render/src/scientific_decode.cpp:297-331 takes every subtype but 1 for
 * per-frame BFP data).  No counterpart in the reference's file decoder: for
 * callers that read what the encoder measured. */
enum { VBT_OPEN_SCI_DENSE_BASIS = 1 };
int vbt_asset_open_memory_ex(const void* bytes, uint64_t size, void* stream, uint32_t flags, vbt_asset* out);
int vbt_asset_open_file_ex(const char* path, void* stream, uint32_t flags, vbt_asset* out);
void vbt_asset_close(vbt_asset a);

typedef struct {
    vbt_file_header header;
    uint32_t activeLeaves;       /* A */
    uint32_t badLeaves;          /* leaves whose status != VBT_LEAF_OK */
    uint64_t controlsPerFrame;   /* packed cache floats per frame (meaningful controls only) */
    uint64_t payloadBytes;       /* compressed payload resident in HBM */
    uint64_t deviceBytes;        /* all device memory owned by the asset */
    uint64_t mode_counts[4];     /* leaves per header mode (render) / scientific mode */
    uint64_t fast_leaves;        /* active leaves eligible for the multi-frame sweep */
} vbt_asset_info;
int vbt_asset_info_get(vbt_asset a, vbt_asset_info* out);

/* Device-side page table copies (device pointers; sizes from the info). */
int vbt_asset_page_table(vbt_asset a, uint32_t* d_compact /* leafCount */, void* stream);
int vbt_asset_cache_offsets(vbt_asset a, uint64_t* d_offsets /* activeLeaves + 1 */, void* stream);
int vbt_asset_leaf_status(vbt_asset a, uint32_t* d_status /* leafCount */, void* stream);

/* ------------------------------------------------------ frame caches -- */
enum {
    VBT_CACHE_PACKED = 0,    /* out[(t-t0)*controlsPerFrame + offset[a] + slot], meaningful controls only */
    VBT_CACHE_STRIDE576 = 1, /* out[(t-t0)*A*576 + a*576 + slot], vbt_cache.comp slot layout, unused slots 0 */
};
/* Decode frames [t0, t0+nt) of every active leaf into the compact active-leaf
 * cache.  Render leaves: coarse controls then Fine6 or Mode2 shell controls
 * (vbt_cache.comp:193-236 slot order, CPU arithmetic).  Scientific leaves: the
 * leaf's decoded voxels (cropped, x fastest).  Returns VBT_ERR_FORMAT when any
 * active leaf is poisoned (its slots are NaN, the others are valid). */
int vbt_decode_frame_cache(vbt_asset a, int t0, int nt, int layout, float* d_out, void* stream);

/* Autotuning of the frame-cache kernel's tile shape (workgroup slots x LDS
 * keyframe window).  vbt_asset_autotune times every shape on this asset for
 * nt frames (reps launches each, into library scratch memory), keeps the
 * fastest and reports the per-shape times (msPerTile has vbt_cache_tile_count()
 * entries; may be NULL).  The choice is per asset and can be set directly
 * from a stored tuning table.  No counterpart in the reference (its cache
 * shader has one fixed shape, vbt_cache.comp:3). */
int vbt_cache_tile_count(void);
int vbt_cache_tile_shape(int tile, uint32_t* block, uint32_t* window);
int vbt_asset_set_cache_tile(vbt_asset a, int tile);
int vbt_asset_get_cache_tile(vbt_asset a, int* tile);
int vbt_asset_autotune(vbt_asset a, int nt, int reps, float* msPerTile, int* best, void* stream);
/* The memory traffic of a packed frame-cache build of nt frames without its
 * decode, for timing the layout's ceiling on this render asset: the same
 * tiles of shape `tile`, the same keyframe record bytes staged into LDS
 * (reads != 0; none when 0) and the same output stores.  d_out (nt x
 * controlsPerFrame floats) receives placeholder values, not a frame cache.
 * No counterpart in the reference (a measurement aid, like vbt_device_copy). */
int vbt_asset_cache_ceiling(vbt_asset a, int nt, int tile, int reads, float* d_out, void* stream);

/* The tuning table: tile-shape picks keyed by the asset statistics the best
 * shape depends on, one table per process, kept in a JSON file between runs.
 * The key (vbt_tuning_key, at most 256 bytes) is "p<profile> nt<frames per
 * launch> s<round(log2 slots per frame)> r<Grid4>.<Fine6>.<Mode2> k<h0>...<h8>":
 * route mix in tenths of the active leaves, and the keyframe-count histogram
 * (0, 1, 2-3, 4-7, ..., 128-255 keyframes per slot) in tenths of the slots;
 * it runs one kernel on `stream` and waits for it.
 * vbt_asset_autotune_table sets the tile from the table when the key is
 * there (*measured = 0, msPerTile untouched), else times every shape as
 * vbt_asset_autotune does (msPerTile as there) and records the pick
 * (*measured = 1).  vbt_tuning_load merges a file's
 * entries into the table; vbt_tuning_save writes the table (atomically, via a
 * temporary file).  Scientific assets have one kernel shape and no entries.
 * No counterpart in the reference. */
int vbt_tuning_key(vbt_asset a, int nt, char* buf, size_t len, void* stream);
int vbt_asset_autotune_table(vbt_asset a, int nt, int reps, float* msPerTile, int* best, int* measured,
                             void* stream);
int vbt_tuning_load(const char* path);
int vbt_tuning_save(const char* path);
int vbt_tuning_count(void);
int vbt_tuning_clear(void);

/* ------------------------------------------------ batched frame caches -- */
/* Frame-cache builds of several assets as one hipGraph replay: the render
 * jobs (any mix of Grid4 / Fine6 / Mode2 leaves) as ONE persistent kernel over
 * a work queue of (job, leaf tile, frame range) items; scientific jobs as
 * kernel nodes on a concurrent branch of the same graph.  Each job's output is
 * exactly what vbt_decode_frame_cache(asset, t0, nt, layout, d_out) writes.
 * No counterpart in the reference (it decodes one frame of one asset per
 * dispatch, vbt_cache.comp:168-236); the work-queue shape is chosen by
 * vbt_cache_batch_autotune. */
typedef struct {
    vbt_asset asset;
    int t0, nt;   /* frames [t0, t0+nt) */
    int layout;   /* VBT_CACHE_PACKED or VBT_CACHE_STRIDE576 (render assets only) */
    float* d_out; /* device memory, nt * frame size floats */
} vbt_cache_job;

typedef struct {
    int tile;    /* leaf-tile shape, 0 .. vbt_cache_tile_count()-1 (slots per workgroup x LDS window) */
    int ranges;  /* frame ranges per leaf tile: each job's frames split into this many items per tile */
    int chunk;   /* items a workgroup claims per queue access */
    int order;   /* 0: job / leaf order; 1: heaviest tiles (keyframe bytes) first */
    int queues;  /* 1: one queue; 8: one queue per XCD (workgroup i starts on queue i % 8, then steals) */
    int wgPerCU; /* persistent workgroups per CU; 0: as many as fit */
    int persistent; /* 1: persistent workgroups claim items from the queues; 0: one workgroup per `chunk`
                       items in queue order (static); 2: no queue, each job's own one-shot build
                       (vbt_decode_frame_cache's kernels at its asset's tile shape), the jobs alternating
                       between two graph branches */
    int heavyTile;  /* static form only: tiles whose keyframe bytes exceed the tile's LDS window run in a
                       second, concurrent kernel with this tile shape (same slots, larger window); -1: none */
} vbt_batch_shape;

typedef struct vbt_cache_batch_s* vbt_cache_batch;

void vbt_batch_default_shape(vbt_batch_shape* shape);
/* Plans the queue for `shape` (NULL: default) and captures the graph.  The
 * assets and outputs must stay alive and unchanged in size while the batch
 * exists; every asset must live on the current device. */
int vbt_cache_batch_create(const vbt_cache_job* jobs, int nJobs, const vbt_batch_shape* shape,
                           vbt_cache_batch* out);
int vbt_cache_batch_destroy(vbt_cache_batch b);
/* One replay of the captured graph on `stream` (all jobs).  Returns
 * VBT_ERR_FORMAT (after launching) when a job's asset has poisoned leaves, as
 * vbt_decode_frame_cache does. */
int vbt_cache_batch_launch(vbt_cache_batch b, void* stream);
/* The same work as plain launches (no graph), for comparison and tuning. */
int vbt_cache_batch_launch_direct(vbt_cache_batch b, void* stream);
int vbt_cache_batch_set_shape(vbt_cache_batch b, const vbt_batch_shape* shape);
int vbt_cache_batch_get_shape(vbt_cache_batch b, vbt_batch_shape* shape);
/* Work items of the current plan and persistent workgroups of its launch. */
int vbt_cache_batch_plan_info(vbt_cache_batch b, uint64_t* items, uint32_t* workgroups);
/* Times candidate shapes (reps direct launches each, coordinate descent over
 * kernel form x tile, then ranges, chunk, order, queues, heavy tile), keeps
 * the fastest and re-captures the graph.  Trials can be read back with vbt_cache_batch_trial. */
int vbt_cache_batch_autotune(vbt_cache_batch b, int reps, vbt_batch_shape* best, float* bestMs, void* stream);
int vbt_cache_batch_trial_count(vbt_cache_batch b);
int vbt_cache_batch_trial(vbt_cache_batch b, int i, vbt_batch_shape* shape, float* ms);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the work of vbt_cache_batch_create + vbt_cache_batch_launch_direct
 * in one call that allocates nothing.  The batch's work queue (job table,
 * entries, queue bounds and counters) lives in a file-scope device array of
 * VBT_NOALLOC_SCRATCH_BYTES that each device the library runs on holds until
 * the process ends; it is written by kernels from kernel arguments (no
 * memset, no host-to-device copy) and its counters are cleared on every run.
 * Limits: at most VBT_NOALLOC_MAX_JOBS jobs and VBT_NOALLOC_MAX_TILES render
 * leaf tiles (slots / tile slots, summed over the jobs); larger requests
 * return VBT_ERR_INVALID.  The shape (NULL: vbt_batch_noalloc_default_shape)
 * must have order 0 (job / leaf order, queues split evenly by tile count) and
 * heavyTile -1; scientific jobs run on `stream` before the queue kernel.
 * Outputs equal vbt_cache_batch_launch's bit for bit.  Every call on a device
 * shares that device's array: calls that overlap on one device (other
 * streams, other threads) race and must be serialised by the caller. */
#define VBT_NOALLOC_MAX_JOBS 32
#define VBT_NOALLOC_MAX_TILES (1u << 19)
#define VBT_NOALLOC_SCRATCH_BYTES (4u * 1024u * 1024u + 8u * 1024u)
void vbt_batch_noalloc_default_shape(vbt_batch_shape* shape);
int vbt_cache_batch_run_noalloc(const vbt_cache_job* jobs, int nJobs, const vbt_batch_shape* shape, void* stream);

/* --------------------------------------------------- voxels & queries -- */
enum {
    VBT_SIGN_NONE = 0,        /* render_temporal_vbt_to_vdb default (optional shellEmptyScale) */
    VBT_SIGN_CLAMP_ZERO = 1,  /* render_smoke_vbt_bench_main.cpp:374-378: Mode2 !occupied && v<0 -> 0 */
    VBT_SIGN_BACKGROUND = 2,  /* vbt_sample.comp:410-412: Mode2 !occupied && v<0 -> background */
};
typedef struct {
    int signRule;
    float background;
    float shellEmptyScale; /* 1 disables */
    float shellEmptyMax;   /* NaN disables */
} vbt_compose;

/* decodeFrameToDense (render/src/render_smoke_vbt_bench_main.cpp:389-455) /
 * reconstructScientificFrameCpu (render/src/scientific_decode.h:12): frame t,
 * row-major (z, y, x) floats. */
int vbt_decode_frame_dense(vbt_asset a, int t, const vbt_compose* c, float* d_out, void* stream);

/* Point queries (Query4D x,y,z,t as 4 x u32, render/src/query_patterns.h:11-17;
 * the leaf is leafIndexForVoxel(x,y,z)): value = the whole-frame decode at
 * that voxel.  d_status may be NULL.  decodeScientificValueAtCpu,
 * render/src/scientific_decode.h:11, and the render point sampler. */
int vbt_query_points(vbt_asset a, const uint32_t* d_queries, uint32_t n, const vbt_compose* c, float* d_out,
                     int32_t* d_status, void* stream);
/* The plain GPU decode used as the comparison point: one thread per query,
 * one launch per request, no grouping. */
int vbt_query_points_plain(vbt_asset a, const uint32_t* d_queries, uint32_t n, const vbt_compose* c, float* d_out,
                           int32_t* d_status, void* stream);

/* ----------------------------------------------------------- serving -- */
/* A continuous-batching decode server over resident assets (no counterpart in
 * the reference, whose decode functions are called one value at a time from
 * OpenMP loops).  Requests -- point-query batches and whole-frame packed cache
 * builds, on any registered asset -- are submitted from any thread (MPSC).
 * A host dispatcher admits them in arrival order while the device arena, the
 * HBM budget (resident assets + request outputs) and the in-flight limit
 * allow, and appends them to a request ring in host-coherent pinned memory.
 * On the device a persistent loop runs: one scheduler wave polls the ring and
 * publishes new requests to device memory; every other workgroup claims work
 * items (a chunk of queries, a range of cache slots, four scientific leaves)
 * from the oldest request of a class with unclaimed items -- point-query
 * batches first in seven of eight workgroups, frame builds first in the
 * eighth -- so requests that arrive while others run start on the CUs as soon
 * as workgroups free up.  The workgroup
 * finishing a request's last item writes its completion flag to host memory.
 * The loop exits after idleUs without new requests or after leaseMs, and the
 * dispatcher relaunches it when work arrives; every device wait is bounded by
 * the device clock.  Results stay in HBM until fetched or released.
 * Registering an asset stops the device loop first (the asset table changes
 * only between launches). */
typedef struct vbt_server_s* vbt_server;

typedef struct {
    uint64_t hbmBudgetBytes; /* admission: resident assets + request outputs (default 80 GiB) */
    uint64_t arenaBytes;     /* device arena for request outputs (default 4 GiB, clipped to the budget) */
    uint64_t stagingBytes;   /* pinned host staging for query inputs (default 256 MiB) */
    uint32_t maxInflight;    /* admitted, unfinished requests at once (default 4096) */
    uint32_t blocksPerCU;    /* persistent workgroups per CU, 256 lanes each (default 2) */
    uint32_t queryItem;      /* point queries per work item, multiple of 32 (default 256) */
    uint32_t cacheItem;      /* packed-cache slots per work item, multiple of 256 (default 1024) */
    uint32_t idleUs;         /* device loop exits after this long without new requests (default 2000) */
    uint32_t leaseMs;        /* device loop exits (and is relaunched) at least this often (default 250) */
    uint32_t frameBuildsInLoop; /* 0 (default): whole-frame builds run as one-shot frame-cache kernels on a
                                   stream beside the device loop; 1: as cacheItem-slot work items of the loop */
    uint32_t frameIndex;        /* 1 (default): registering a render asset builds a frame index owned by the
                                   server (its bytes count as resident), and one-shot frame builds read it
                                   (same values); 0: frame builds search the keyframe pool */
} vbt_server_config;

typedef struct {
    uint64_t id;
    int32_t status;          /* VBT_OK, or VBT_ERR_FORMAT when an element failed (see d_status) */
    uint32_t count;          /* values */
    const float* d_values;   /* device pointer, valid until fetch/release */
    const int32_t* d_status; /* per-query statuses (queries) or NULL */
    double latencyUs;        /* submit -> completion observed by the dispatcher */
    uint64_t launch;         /* device-loop launch it ran in (0: a frame build beside the loop) */
} vbt_request_result;

typedef struct {
    uint64_t submitted, completed, rejected, launches, values, requeued;
    double p50Us, p99Us, meanUs, maxUs;
    uint64_t maxQueueDepth, maxInflight, hbmInUse, hbmBudget, arenaInUse;
} vbt_server_stats;

void vbt_server_default_config(vbt_server_config* c);
int vbt_server_create(const vbt_server_config* c, vbt_server* out);
int vbt_server_destroy(vbt_server s);
/* Register a resident asset (counts against the HBM budget). */
int vbt_server_register(vbt_server s, vbt_asset a, int* assetId);
/* Queue a point-query batch (host array of n x {x,y,z,t}; copied to pinned
 * staging before returning).  VBT_ERR_BUDGET when its buffers can never fit;
 * VBT_ERR_AGAIN when the staging area stays full for a second. */
int vbt_server_submit_query(vbt_server s, int assetId, const uint32_t* queries, uint32_t n, const vbt_compose* c,
                            uint64_t* requestId);
/* Queue a whole-frame packed cache build of frame t (layout VBT_CACHE_PACKED,
 * same values as vbt_decode_frame_cache). */
int vbt_server_submit_cache(vbt_server s, int assetId, int t, uint64_t* requestId);
/* Wait up to timeoutMs (-1 = forever) for completion; VBT_ERR_AGAIN on timeout. */
int vbt_server_poll(vbt_server s, uint64_t requestId, int timeoutMs, vbt_request_result* out);
/* Copy a completed request's values (and per-query statuses, may be NULL) to
 * host memory and release its device buffers. */
int vbt_server_fetch(vbt_server s, uint64_t requestId, float* values, int32_t* status);
int vbt_server_release(vbt_server s, uint64_t requestId);
/* Latency statistics over completed requests of a type (-1 all, 0 queries, 1 cache builds). */
int vbt_server_stats_get(vbt_server s, int requestType, vbt_server_stats* out);
int vbt_server_reset_stats(vbt_server s);

/* ------------------------------------------- device / stream helpers -- */
int vbt_device_count(int* n);
int vbt_device_alloc(void** d_ptr, size_t bytes);
int vbt_device_free(void* d_ptr);
int vbt_memcpy_h2d(void* d_dst, const void* src, size_t bytes, void* stream);
int vbt_memcpy_d2h(void* dst, const void* d_src, size_t bytes, void* stream);
int vbt_memset_d(void* d_dst, int value, size_t bytes, void* stream);
/* Device-to-device copy by a plain copy kernel (k_copy; 16-byte aligned
 * pointers and size): the known-size reference used to scale HBM byte
 * counters when profiling the decode kernels. */
int vbt_device_copy(void* d_dst, const void* d_src, size_t bytes, void* stream);
/* Write-only reference: fills bytes (16-byte aligned) with a 32-bit word by a
 * plain store kernel; the write-bandwidth ceiling next to vbt_device_copy's
 * copy ceiling, and a clear that a captured graph may contain. */
int vbt_device_fill(void* d_dst, uint32_t word, size_t bytes, void* stream);
/* Diagnostic without a counterpart in the reference: counts the lerp weights
 * (frame - f0) / (f1 - f0), 0 <= frame - f0 < f1 - f0 <= maxGap, for which the
 * frame-cache sweep's reciprocal-based weight differs from the correctly
 * rounded float quotient the reference computes (all 2^31 of them for
 * maxGap = 65535); expected 0.  Synchronises `stream`. */
int vbt_selftest_weights(uint32_t maxGap, uint64_t* mismatches, void* stream);
/* ---- sampling and ray marching over a frame cache --------------------------
 * The reference viewer's sampler (VBTStudio/frontend/shaders/vbt_sample.comp:
 * sampleVbtValue :322-415, intersectWorldBox :417-428, traceShadow :430-452,
 * estimateThickness :454-474, traceLevelSet :476-551, main :553-622) reading
 * a one-frame cache that vbt_decode_frame_cache built in the
 * VBT_CACHE_STRIDE576 layout (activeLeaves * 576 floats).  Render assets
 * (profile 1) whose leaves all validate.  Operations the shader leaves
 * implementation-defined are fixed: mix(x, y, a) = x*(1-a) + y*a, round()
 * half away from zero, min(x, y) = y < x ? y : x, max(x, y) = x < y ? y : x,
 * length = sqrt((x*x + y*y) + z*z); IEEE float without contraction.  The
 * parameter block mirrors the shader's push constants (vbt_sample.comp:
 * 38-65) that are not image or volume sizes. */
typedef struct {
    uint32_t stepCount;   /* march steps per ray (pc.stepCount) */
    int32_t frameIndex;   /* frame the cache holds; coarse grids above 64 controls decode from the payload at it */
    float bboxMin[3];     /* world position of voxel (0, 0, 0) */
    uint32_t renderMode;  /* 0 density integration, 1 level-set surface */
    float voxelSize, backgroundValue, surfaceIso, surfaceEpsilonVoxels, surfaceNormalStep, maxStepVoxels;
    uint32_t surfaceModel; /* density: 8 depth bins; level set: thickness and floor shadow */
    float floorHeight;
    uint32_t shadowSteps;
    uint32_t upAxis;      /* 2: z up, else y up (pc.padding) */
} vbt_march_params;

/* sampleVbtValue at n world points (3 floats each, device memory). */
int vbt_sample_points(vbt_asset a, const float* d_cache576, const vbt_march_params* p, const float* d_xyz,
                      uint32_t n, int preserveInterior, float* d_out, void* stream);
/* The shader's main for n rays: d_rays holds 8 floats per ray (origin xyz,
 * tmin, direction xyz, tmax), d_out 8 per ray (primary, auxiliary: density
 * sum or 8 depth bins; level set: normal, hit t, thickness, floor flag,
 * shadow, |n.-d|); both 16-byte aligned. */
int vbt_march_rays(vbt_asset a, const float* d_cache576, const vbt_march_params* p, const float* d_rays, uint32_t n,
                   float* d_out, void* stream);

/* Frame index for repeated one- and two-frame packed-cache builds (serving,
 * scrubbing): each slot's keyframe records regrouped by bin of the frame range
 * (min(32, frames) bins), so such a build reads short adjacent runs instead
 * of one keyframe line per slot.  No counterpart in the reference (its
 * decoder reads the payload per value); it changes no result: with the index,
 * frame builds of one or two frames in the packed layout read it, other
 * builds and leaves whose keys do not strictly increase keep the bracket
 * search.  Costs 4 * bins * (controlsPerFrame + 1) bytes of offsets plus 4
 * bytes per indexed record (8 when a run holds fp32 records) (*bytes, also
 * added to the asset's deviceBytes); synchronizes the stream.  Render assets
 * only.  Build and drop change the asset: call them while no other call on
 * it runs. */
int vbt_asset_build_frame_index(vbt_asset a, uint64_t* bytes, void* stream);
/* Frees the frame index, after waiting for the last build that read it on
 * each stream. */
int vbt_asset_drop_frame_index(vbt_asset a);

/* ---- Cycles resident blob (cycles_native_vbt/, blender_patch/native_cycles_vbt.patch) ----
 * The blob is the reference's BlobHeader (cycles_native_vbt/vbt_cycles_blob.h:14-34:
 * 19 four-byte fields, magic "VBT1", version 1, bbox_min 0) followed by the
 * leaf offset table and the payload pool, the layout the patched Cycles
 * kernel reads (native_cycles_vbt.patch:456-476, 551-563). */
#define VBT_CYCLES_KERNEL 1 /* the patched Cycles kernel's sampler (native_cycles_vbt.patch:876-929);
                               else the standalone header's (vbt_cycles_sampler.h:284-376) */
#define VBT_CYCLES_WORLD 2  /* subtract the blob's bbox_min first (vbt_cycles_sampler.h:370-376) */
#define VBT_CYCLES_PLAIN 4  /* one lane per sample, the reference's structure (for comparison) */

/* Size of the blob pack_blob makes for this asset (vbt_cycles_blob.cpp:50-52). */
int vbt_cycles_blob_bytes(vbt_asset a, uint64_t* bytes);
/* pack_blob (vbt_cycles_blob.cpp:22-87) on the device from the asset's resident
 * offsets and payload: frameIndex is clamped to [0, frames - 1]; d_blob is
 * 4-byte aligned with room for vbt_cycles_blob_bytes.  Stream-ordered. */
int vbt_cycles_pack_blob(vbt_asset a, int frameIndex, void* d_blob, uint64_t capacity, void* stream);
/* Density at n points (d_xyz: 3 floats each, index space unless
 * VBT_CYCLES_WORLD) read from a device blob: sample_density_index
 * (vbt_cycles_sampler.h:284-368) or, with VBT_CYCLES_KERNEL,
 * kernel_image_interp_vbt_density (native_cycles_vbt.patch:915-929).  A blob
 * whose header fails blob_payload_view's checks (vbt_cycles_blob.cpp:89-124)
 * samples to 0, as the Cycles kernel does for a bad magic.  Reads past the
 * blob's sections read zero bytes. */
int vbt_cycles_sample_density(const void* d_blob, uint64_t blobBytes, int flags, const float* d_xyz, uint32_t n,
                              float* d_out, void* stream);

int vbt_stream_create(void** stream);
int vbt_stream_destroy(void* stream);
int vbt_stream_synchronize(void* stream);
int vbt_event_create(void** ev);
int vbt_event_destroy(void* ev);
int vbt_event_record(void* ev, void* stream);
int vbt_event_elapsed_ms(void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif
#endif
