// sw_engine_internal.h -- hooks of the engine used by the request server
// (sw_server.cpp); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <exception>
#include <new>

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
// Runs an entry point's body and turns a C++ exception escaping it (a failed
// host allocation, a thread that could not start) into an error code, so no
// exception crosses the C ABI.
template <class F>
int no_throw(F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return set_last_error(IPUMA_ENOMEM, "host allocation failed");
  } catch (const std::exception &x) {
    return set_last_error(IPUMA_EDEVICE, "internal: %s", x.what());
  } catch (...) {
    return set_last_error(IPUMA_EDEVICE, "internal: unknown exception");
  }
}
}  // namespace ipuma
