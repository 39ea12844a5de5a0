// sw_engine_internal.h -- hooks of the engine used by the request server
// (sw_server.cpp); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ipuma_hip.h"

namespace ipuma {
// 256-entry character -> residue code table of the engine's data type.
const uint8_t *engine_code_table(const ipuma_engine *e);
// The engine's HIP stream.
hipStream_t engine_stream(const ipuma_engine *e);
// Error bits the kernels of the last batch raised (after its completion).
int engine_check_error(ipuma_engine *e);
// Sets the calling thread's last-error message; returns `code`.
int set_last_error(int code, const char *fmt, ...);
}  // namespace ipuma
