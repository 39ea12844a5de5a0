// sw_calib.hip -- throughput of the integer VALU instructions the DP kernels
// issue, measured on the current device: the calibrated peak against which
// the packed kernel's roofline fraction is reported beside the nominal one.
//
// k_valu_pk16: per lane, kChains independent chains alternating a saturating
// packed int16 add (v_pk_add_i16 clamp) and a packed int16 max (v_pk_max_i16),
// the two instructions of the DP's H / E / F updates; 2 int16 ops each.
// k_valu_i32: the same with 32-bit adds and maxes (1 op each).
// Enough waves are launched that every SIMD always has work to issue.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/ipuma_hip.h"

namespace {

typedef short s16x2 __attribute__((ext_vector_type(2)));

constexpr int kChains = 8;
constexpr int kUnroll = 16;
constexpr int kCalibThreads = 256;

__global__ __launch_bounds__(kCalibThreads) void k_valu_pk16(uint32_t *out, int iters,
                                                             uint32_t seed) {
  s16x2 x[kChains];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int c = 0; c < kChains; ++c)
    x[c] = __builtin_bit_cast(s16x2, (t * 2654435761u + seed * (c + 1)) & 0x0FFF0FFFu);
  // step values from a kernel argument: not foldable by the compiler
  const s16x2 inc = __builtin_bit_cast(s16x2, seed | (seed << 16));
  const s16x2 lim = __builtin_bit_cast(s16x2, (seed + 7u) | ((seed + 7u) << 16));
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        x[c] = __builtin_elementwise_add_sat(x[c], inc);
        x[c] = __builtin_elementwise_max(x[c], lim);
      }
    }
  }
  uint32_t acc = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) acc ^= __builtin_bit_cast(uint32_t, x[c]);
  out[t] = acc;
}

__global__ __launch_bounds__(kCalibThreads) void k_valu_i32(uint32_t *out, int iters,
                                                            uint32_t seed) {
  int x[kChains];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = (int)((t * 2654435761u + seed * (c + 1)) & 0xFFFFu);
  const int inc = (int)seed, lim = (int)seed + 7;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        x[c] = x[c] + inc;
        x[c] = max(x[c], lim);
      }
    }
  }
  uint32_t acc = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) acc ^= (uint32_t)x[c];
  out[t] = acc;
}

}  // namespace

extern "C" int ipuma_valu_peak(double *pk_int16_tops, double *int32_tops) {
  if (!pk_int16_tops && !int32_tops) return IPUMA_EINVAL;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return IPUMA_EDEVICE;
  // 32 waves per CU (8 per SIMD)
  const int blocks = prop.multiProcessorCount * 8;
  const int threads = blocks * kCalibThreads;
  const int iters = 2048;  // about 40 ms per launch at the nominal rate
  uint32_t *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = IPUMA_OK;
  if (hipMalloc(&out, (size_t)threads * 4) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = IPUMA_EDEVICE;
  // lane-instructions per launch x ops per instruction
  const double instr = (double)threads * iters * kUnroll * kChains * 2.0;
  for (int kind = 0; kind < 2 && rc == IPUMA_OK; ++kind) {
    double *res = kind == 0 ? pk_int16_tops : int32_tops;
    if (!res) continue;
    float best = 1e30f;
    for (int rep = 0; rep < 4 && rc == IPUMA_OK; ++rep) {  // the first launch warms up
      if (hipEventRecord(e0, st) != hipSuccess) rc = IPUMA_EDEVICE;
      if (kind == 0)
        hipLaunchKernelGGL(k_valu_pk16, dim3(blocks), dim3(kCalibThreads), 0, st, out, iters,
                           3u + rep);
      else
        hipLaunchKernelGGL(k_valu_i32, dim3(blocks), dim3(kCalibThreads), 0, st, out, iters,
                           3u + rep);
      float ms = 0.f;
      if (rc == IPUMA_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                             hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = IPUMA_EDEVICE;
      if (rep > 0) best = std::min(best, ms);
    }
    if (rc == IPUMA_OK) *res = instr * (kind == 0 ? 2.0 : 1.0) / (best * 1e-3) / 1e12;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (out) (void)hipFree(out);
  return rc;
}
