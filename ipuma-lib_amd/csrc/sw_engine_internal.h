// sw_engine_internal.h -- hooks of the engine used by the request server
// (sw_server.cpp); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <exception>
#include <new>
#include <system_error>
#include <thread>
#include <utility>

#include "../../include/ipuma_hip.h"

namespace ipuma {
// 256-entry character -> residue code table of the engine's data type.
const uint8_t *engine_code_table(const ipuma_engine *e);
// The engine's HIP stream.
hipStream_t engine_stream(const ipuma_engine *e);
// Device addresses of the error words of the engine's last device batch (the
// forward chain's and, with start positions, the reverse chain's; w1 may be
// NULL).  A caller copies them on the batch's stream right after the batch,
// before a later batch clears them.
int engine_error_words(ipuma_engine *e, const int32_t **w0, const int32_t **w1);
// The status (IPUMA_OK or an error code with its message set) for error bits.
int engine_error_status(int32_t bits);
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
// Host threads the library may still start; -1 (the default) is unlimited.
// Fault injection for the tests of the thread-start failure paths: a start
// past the limit throws std::system_error, as a failed std::thread start does.
extern std::atomic<int> g_thread_start_budget;
// Starts a host thread of the library (std::thread under the limit above).
template <class F>
std::thread start_thread(F &&f) {
  int b = g_thread_start_budget.load(std::memory_order_relaxed);
  while (b >= 0) {
    if (b == 0)
      throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again),
                              "thread start refused");
    if (g_thread_start_budget.compare_exchange_weak(b, b - 1, std::memory_order_relaxed)) break;
  }
  return std::thread(std::forward<F>(f));
}
}  // namespace ipuma
