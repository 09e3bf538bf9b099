// vbt_internal.h -- shared declarations of libvbt_hip.so (not installed).
#pragma once

#include "../../include/vbt_hip.h"

#include <string>

// Record the calling thread's last error and return the code.
int vbt_fail(int code, const char* msg);
int vbt_fail(int code, const std::string& msg);
