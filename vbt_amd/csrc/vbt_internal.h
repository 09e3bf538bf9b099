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
