// vbt_internal.h -- shared declarations of libvbt_hip.so (not installed).
#pragma once

#include "../../include/vbt_hip.h"

#include <string>

// Record the calling thread's last error and return the code.
int vbt_fail(int code, const char* msg);
int vbt_fail(int code, const std::string& msg);

// Accessors of libvbt_hip.so's asset object for the other translation units.
namespace vbtk {
struct AssetDev;
}
vbtk::AssetDev vbt_internal_asset_dev(vbt_asset a);
uint64_t vbt_internal_asset_bytes(vbt_asset a);
int vbt_internal_asset_device(vbt_asset a);
uint32_t vbt_internal_asset_bad(vbt_asset a);
uint32_t* vbt_internal_asset_err(vbt_asset a);
int vbt_internal_report_bad(vbt_asset a, const char* what);
const vbt_file_header* vbt_internal_asset_header(vbt_asset a);
// The packed-cache tile list of tile shape `shape` (render assets): tile j is
// the CACHE_GROUP-slot groups [tiles[j], tiles[j + 1]); *count tiles, the
// device list returned, the host copy in *host (count + 1 entries).  Scientific
// assets have none (count 0, nullptr).
const uint32_t* vbt_internal_asset_tiles(vbt_asset a, int shape, uint32_t* count, const uint32_t** host);
// vbt_decode_frame_cache without the frame index (no event records, so a
// stream capture can hold it)
int vbt_internal_frame_cache_plain(vbt_asset a, int t0, int nt, int layout, float* d_out, void* stream);

// A frame index of one asset: per bin of the frame range, each packed-cache
// slot's run of keyframe records (offsets bins x (slots + 1), then the
// records); one- and two-frame packed builds read it instead of the keyframe
// pool.  An asset can own one (vbt_asset_build_frame_index) and a server one
// per registered asset.
struct VbtFrameIndex {
    uint32_t* off = nullptr;
    void* rec = nullptr;
    uint64_t records = 0;
    uint64_t bytes = 0; // device bytes of off and rec
    uint32_t bins = 0;
    bool wide = false;  // 8-byte {frame, value bits} records (else 4-byte {frame, half bits})
};
// Builds ix for asset a on `stream` (synchronises it); ix must be empty.  An
// asset with nothing indexable leaves ix empty and returns VBT_OK.
int vbt_internal_frame_index_build(vbt_asset a, VbtFrameIndex& ix, void* stream);
void vbt_internal_frame_index_free(VbtFrameIndex& ix);
// Packed cache of frames [t0, t0 + nt), nt <= 2, from an index of asset a.
int vbt_internal_frame_cache_indexed(vbt_asset a, const VbtFrameIndex& ix, int t0, int nt, float* d_out, void* stream);
