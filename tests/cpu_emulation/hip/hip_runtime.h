// CPU emulation of the HIP subset the engine uses -- TEST INFRASTRUCTURE.
//
// Lets tests compile the engine's own kernel and host sources
// (ipuma-lib_amd/csrc/*.hip, *.cpp) for the host: every launch runs each block
// as blockDim.x std::threads; the 64 lanes of a wave exchange values for DPP,
// shuffles, ballots and readfirstlane through a per-wave barrier, so the
// kernels run in lockstep exactly as written.  Device memory is host memory
// pre-filled with a garbage pattern (0xA5) so reliance on zeroed memory shows.
// Built with -fsanitize=address this bounds-checks every kernel access.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
#define __shared__ static
#define __HIPCC__ 1

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint2 {
  uint32_t x, y;
};
struct int4 {
  int x, y, z, w;
};
inline uint2 make_uint2(uint32_t a, uint32_t b) { return uint2{a, b}; }

namespace emu {
struct Wave {
  std::barrier<> bar{64};
  uint64_t x[64];
};
struct Ctx {
  dim3 tid, bid, bdim, gdim;
  int lane = 0;
  Wave *wave = nullptr;
  std::barrier<> *block = nullptr;
};
inline thread_local Ctx ctx;
extern int sm_count;
extern int fill_byte;

template <class T>
T exchange_get(T v, int src) {
  Wave &w = *ctx.wave;
  uint64_t raw = 0;
  std::memcpy(&raw, &v, sizeof(T));
  w.x[ctx.lane] = raw;
  w.bar.arrive_and_wait();
  uint64_t r = w.x[src & 63];
  w.bar.arrive_and_wait();
  T out;
  std::memcpy(&out, &r, sizeof(T));
  return out;
}

inline int update_dpp(int old, int src, int ctrl, int, int, bool bound_ctrl) {
  Wave &w = *ctx.wave;
  const int lane = ctx.lane;
  w.x[lane] = (uint32_t)src;
  w.bar.arrive_and_wait();
  int srcl = -1;
  if (ctrl == 0x111) srcl = (lane % 16) ? lane - 1 : -1;  // row_shr:1
  else if (ctrl == 0x138) srcl = lane ? lane - 1 : -1;    // wave_shr:1
  else { std::fprintf(stderr, "emu: unsupported dpp ctrl %#x\n", ctrl); std::abort(); }
  const int r = srcl >= 0 ? (int)(uint32_t)w.x[srcl] : (bound_ctrl ? 0 : old);
  w.bar.arrive_and_wait();
  return r;
}

inline uint32_t perm(uint32_t s0, uint32_t s1, uint32_t sel) {
  const uint64_t v = ((uint64_t)s0 << 32) | s1;
  uint32_t out = 0;
  for (int i = 0; i < 4; ++i) {
    const uint32_t b = (sel >> (8 * i)) & 0xff;
    uint32_t byte;
    if (b < 8) byte = (uint32_t)((v >> (8 * b)) & 0xff);
    else if (b == 12) byte = 0;
    else if (b > 12) byte = 0xff;
    else { std::fprintf(stderr, "emu: perm sign selector\n"); std::abort(); }
    out |= byte << (8 * i);
  }
  return out;
}

template <class K, class... Args>
void launch(const char *name, K kernel, dim3 grid, dim3 block, Args... args) {
  const unsigned nthreads = block.x;
  if (std::getenv("EMU_TRACE")) std::fprintf(stderr, "emu launch %s grid %u block %u\n", name, grid.x, block.x);
  if (nthreads % 64) { std::fprintf(stderr, "emu: block size %u not a wave multiple\n", nthreads); std::abort(); }
  // One thread per lane for the whole grid: the blocks run one after another
  // on the same threads, separated by a barrier outside the kernel.
  std::barrier<> blockbar(nthreads), endbar(nthreads);
  std::vector<std::unique_ptr<Wave>> waves;
  for (unsigned w = 0; w < nthreads / 64; ++w) waves.emplace_back(new Wave());
  std::vector<std::thread> th;
  th.reserve(nthreads);
  for (unsigned t = 0; t < nthreads; ++t)
    th.emplace_back([&, t]() {
      ctx.tid = dim3(t);
      ctx.bdim = block;
      ctx.gdim = grid;
      ctx.lane = (int)(t & 63);
      ctx.wave = waves[t / 64].get();
      ctx.block = &blockbar;
      for (unsigned b = 0; b < grid.x; ++b) {
        ctx.bid = dim3(b);
        kernel(args...);
        // a thread that returns keeps arriving at its wave / block barriers
        // nowhere: kernels here only return wave-uniformly after their last
        // exchange, or never exchange at all, so the barriers' phases agree
        // when the next block starts.
        endbar.arrive_and_wait();
      }
    });
  for (auto &x : th) x.join();
  if (std::getenv("EMU_TRACE")) std::fprintf(stderr, "emu done %s\n", name);
}
}  // namespace emu

#define threadIdx (emu::ctx.tid)
#define blockIdx (emu::ctx.bid)
#define blockDim (emu::ctx.bdim)
#define gridDim (emu::ctx.gdim)
#define __syncthreads() (emu::ctx.block->arrive_and_wait())
#define __builtin_amdgcn_update_dpp(o, s, c, r, b, bc) emu::update_dpp((o), (s), (c), (r), (b), (bc))
#define __builtin_amdgcn_mov_dpp(s, c, r, b, bc) emu::update_dpp(0, (s), (c), (r), (b), (bc))
#define __builtin_amdgcn_perm(a, b, s) emu::perm((a), (b), (s))
// v_alignbit_b32: the low 32 bits of ({a, b} >> s)
#define __builtin_amdgcn_alignbit(a, b, s) \
  ((uint32_t)(((((uint64_t)(uint32_t)(a)) << 32) | (uint32_t)(b)) >> ((s) & 31)))
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_readfirstlane(v) emu::exchange_get((v), 0)
// v_readlane: the lane index is uniform over the wave (an SGPR operand)
#define __builtin_amdgcn_readlane(v, l) emu::exchange_get((v), (l) & 63)
#define __builtin_amdgcn_wave_barrier() ((void)emu::exchange_get(0, 0))

template <class T> inline T __shfl_xor(T v, int m) { return emu::exchange_get(v, emu::ctx.lane ^ m); }
template <class T> inline T __shfl(T v, int src) { return emu::exchange_get(v, src); }
template <class T> inline T __shfl_up(T v, unsigned d, int width = 64) {
  const int lane = emu::ctx.lane;
  const int src = (lane % width) >= (int)d ? lane - (int)d : lane;
  return emu::exchange_get(v, src);
}
inline unsigned long long __ballot(int pred) {
  emu::Wave &w = *emu::ctx.wave;
  w.x[emu::ctx.lane] = pred ? 1 : 0;
  w.bar.arrive_and_wait();
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i) m |= (w.x[i] ? 1ull : 0ull) << i;
  w.bar.arrive_and_wait();
  return m;
}
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline float __log2f(float x) { return std::log2(x); }

inline int atomicAdd(int *p, int v) { return std::atomic_ref<int>(*p).fetch_add(v); }
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  return std::atomic_ref<unsigned long long>(*p).fetch_add(v);
}
inline int atomicOr(int *p, int v) { return std::atomic_ref<int>(*p).fetch_or(v); }
inline int atomicCAS(int *p, int cmp, int v) {
  std::atomic_ref<int>(*p).compare_exchange_strong(cmp, v);
  return cmp;  // the value found
}
inline int atomicMax(int *p, int v) {
  std::atomic_ref<int> a(*p);
  int cur = a.load();
  while (cur < v && !a.compare_exchange_weak(cur, v)) {}
  return cur;
}

// scoped atomics (every scope is the whole emulated device)
namespace emu {
template <class T> inline T hload(T *p) { return std::atomic_ref<T>(*p).load(); }
template <class T, class V> inline void hstore(T *p, V v) { std::atomic_ref<T>(*p).store((T)v); }
template <class T, class V> inline bool hcas(T *p, T *e, V d) {
  return std::atomic_ref<T>(*p).compare_exchange_strong(*e, (T)d);
}
template <class T, class V> inline T hadd(T *p, V v) { return std::atomic_ref<T>(*p).fetch_add((T)v); }
template <class T, class V> inline T hor(T *p, V v) { return std::atomic_ref<T>(*p).fetch_or((T)v); }
template <class T, class V> inline T hand(T *p, V v) { return std::atomic_ref<T>(*p).fetch_and((T)v); }
template <class T, class V> inline T hmax(T *p, V v) {
  std::atomic_ref<T> a(*p);
  T cur = a.load();
  while (cur < (T)v && !a.compare_exchange_weak(cur, (T)v)) {}
  return cur;
}
}  // namespace emu
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __hip_atomic_load(p, o, s) emu::hload(p)
#define __hip_atomic_store(p, v, o, s) emu::hstore((p), (v))
#define __hip_atomic_compare_exchange_strong(p, e, d, o1, o2, s) emu::hcas((p), (e), (d))
#define __hip_atomic_fetch_add(p, v, o, s) emu::hadd((p), (v))
#define __hip_atomic_fetch_or(p, v, o, s) emu::hor((p), (v))
#define __hip_atomic_fetch_and(p, v, o, s) emu::hand((p), (v))
#define __hip_atomic_fetch_max(p, v, o, s) emu::hmax((p), (v))
#define __builtin_amdgcn_s_sleep(n) std::this_thread::yield()
#define __builtin_amdgcn_s_setprio(n) ((void)0)
// HW_ID: the wave's SIMD in bits 5:4 (slot % 4, as the hardware spreads a block)
#define __builtin_amdgcn_s_getreg(r) ((int)(((emu::ctx.tid.x >> 6) & 3) << 4))

inline int max(int a, int b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }

// ---- runtime API -----------------------------------------------------------
typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1 };
struct hipDeviceProp_t {
  int multiProcessorCount;
  int clockRate;  // kHz
};
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated error"; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  p->multiProcessorCount = emu::sm_count;
  p->clockRate = 2400000;
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (void *)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipMalloc(void **p, size_t n) {
  *p = std::malloc(n);
  if (!*p) return hipErrorOutOfMemory;
  std::memset(*p, emu::fill_byte, n);
  return hipSuccess;
}
template <class T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
// a __device__ variable is a host variable here: its address is the symbol's
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, int) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
// an event holds the host time of its last record (work runs at launch)
namespace emu {
struct Event {
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
};
}  // namespace emu
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu::Event(); return hipSuccess; }
enum { hipEventDisableTiming = 2, hipHostMallocDefault = 0 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }  // work runs at launch
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) {
  *p = std::malloc(n);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <class T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned f) {
  return hipHostMalloc((void **)p, n, f);
}
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete static_cast<emu::Event *>(e); return hipSuccess; }
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *s) {
  *s = hipStreamCaptureStatusNone;
  return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  static_cast<emu::Event *>(e)->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(static_cast<emu::Event *>(b)->t -
                                                 static_cast<emu::Event *>(a)->t).count();
  return hipSuccess;
}
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, K, int, size_t) {
  *n = 1;
  return hipSuccess;
}
#ifdef EMU_DEFINE_GLOBALS
namespace emu {
int sm_count = 2;
int fill_byte = 0xA5;
}  // namespace emu
#endif
#define hipLaunchKernelGGL(k, grid, block, shm, st, ...) emu::launch(#k, k, dim3(grid), dim3(block), __VA_ARGS__)
