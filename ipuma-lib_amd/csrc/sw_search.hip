// sw_search.hip -- one-vs-many search around the batched engine: the pair
// descriptors of every (query, database sequence) comparison, and the top-k
// selection per query.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw_device.h"

namespace ipuma {

// Comparison p = q * ndb + d aligns database sequence d (A, the columns)
// against query q (B, the rows), or with query_cols the query as A and the
// database sequence as B (the same score for a symmetric matrix): the pairs
// of one query then share A, and the packed kernel's groups of one query read
// both comparisons' scores of a row with one load.
__global__ void k_search_pairs(int32_t nq, int32_t ndb, const int64_t *q_off,
                               const int32_t *q_len, const int64_t *db_off,
                               const int32_t *db_len, int64_t *a_off, int32_t *a_len,
                               int64_t *b_off, int32_t *b_len, int32_t query_cols) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)nq * ndb) return;
  const int q = (int)(p / ndb), d = (int)(p - (int64_t)q * ndb);
  if (query_cols) {
    a_off[p] = q_off[q];
    a_len[p] = q_len[q];
    b_off[p] = db_off[d];
    b_len[p] = db_len[d];
  } else {
    a_off[p] = db_off[d];
    a_len[p] = db_len[d];
    b_off[p] = q_off[q];
    b_len[p] = q_len[q];
  }
}

// Top k of one query per block: entries ordered by score (descending), then
// database index (ascending).  A radix select over the 24-bit score finds the
// k-th score T; every entry above T and the first entries equal to T in index
// order are gathered, then sorted in LDS.
constexpr int kTopThreads = 1024;
constexpr int kTopMax = 1024;  // largest k

__device__ __forceinline__ uint32_t score_key(int32_t v) {
  return (uint32_t)min(max(v, 0), (1 << 24) - 1);
}

__global__ __launch_bounds__(kTopThreads) void k_topk(int32_t ndb, int32_t k,
                                                      const int32_t *scores, int32_t *top_idx,
                                                      int32_t *top_score) {
  __shared__ int32_t hist[256];
  __shared__ int32_t scan[kTopThreads];
  __shared__ uint32_t sel_key[kTopMax];
  __shared__ int32_t sel_idx[kTopMax];
  __shared__ uint32_t sh_prefix, sh_mask;
  __shared__ int32_t sh_need, sh_ngt, sh_neq;
  const int q = blockIdx.x, tid = threadIdx.x;
  const int32_t *s = scores + (size_t)q * ndb;
  const int kk = min(k, ndb);
  if (tid == 0) {
    sh_prefix = 0;
    sh_mask = 0;
    sh_need = kk;
  }
  __syncthreads();
  // radix select of the kk-th largest key, 8 bits at a time
  for (int shift = 16; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += kTopThreads) hist[i] = 0;
    __syncthreads();
    const uint32_t prefix = sh_prefix, mask = sh_mask;
    for (int i = tid; i < ndb; i += kTopThreads) {
      const uint32_t v = score_key(s[i]);
      if ((v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255], 1);
    }
    __syncthreads();
    if (tid == 0 && kk > 0) {
      int need = sh_need, above = 0, b = 255;
      for (; b > 0; --b) {
        if (above + hist[b] >= need) break;
        above += hist[b];
      }
      sh_prefix = prefix | ((uint32_t)b << shift);
      sh_mask = mask | (255u << shift);
      sh_need = need - above;
    }
    __syncthreads();
  }
  const uint32_t T = sh_prefix;
  const int need_eq = sh_need;  // entries equal to T to take, the lowest indices first
  if (tid == 0) {
    sh_ngt = 0;
    sh_neq = 0;
  }
  __syncthreads();
  // gather: all keys > T, and the first need_eq keys == T in index order
  for (int base = 0; base < ndb; base += kTopThreads) {
    const int i = base + tid;
    const uint32_t v = i < ndb ? score_key(s[i]) : 0u;
    const int is_eq = (i < ndb && v == T) ? 1 : 0;
    const int is_gt = (i < ndb && v > T) ? 1 : 0;
    // inclusive scan of is_eq over the chunk (index order)
    scan[tid] = is_eq;
    __syncthreads();
    for (int off = 1; off < kTopThreads; off <<= 1) {
      const int x = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += x;
      __syncthreads();
    }
    const int eq_before = sh_neq;
    const int eq_rank = eq_before + scan[tid] - is_eq;  // rank among equal keys
    if (is_gt) {
      const int slot = atomicAdd(&sh_ngt, 1);  // < kk - need_eq by the selection
      sel_key[slot] = v;
      sel_idx[slot] = i;
    }
    __syncthreads();
    if (is_eq && eq_rank < need_eq) {
      const int slot = (kk - need_eq) + eq_rank;
      sel_key[slot] = v;
      sel_idx[slot] = i;
    }
    if (tid == kTopThreads - 1) sh_neq = eq_before + scan[tid];
    __syncthreads();
  }
  __syncthreads();
  // sort the kk selected entries: score descending, index ascending (bitonic,
  // padded to a power of two with entries that sort last)
  int np2 = 1;
  while (np2 < kk) np2 <<= 1;
  for (int i = kk + tid; i < np2; i += kTopThreads) {
    sel_key[i] = 0;
    sel_idx[i] = 0x7fffffff;
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < np2; i += kTopThreads) {
        const int j = i ^ stride;
        if (j > i) {
          // "before" = higher score, then lower index
          const bool i_first = sel_key[i] > sel_key[j] ||
                               (sel_key[i] == sel_key[j] && sel_idx[i] < sel_idx[j]);
          const bool up = (i & size) == 0;
          if (i_first != up) {
            const uint32_t tk = sel_key[i];
            sel_key[i] = sel_key[j];
            sel_key[j] = tk;
            const int32_t ti = sel_idx[i];
            sel_idx[i] = sel_idx[j];
            sel_idx[j] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int r = tid; r < k; r += kTopThreads) {
    top_idx[(size_t)q * k + r] = r < kk ? sel_idx[r] : -1;
    top_score[(size_t)q * k + r] = r < kk ? (int32_t)sel_key[r] : 0;
  }
}

static inline int nblocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

hipError_t launch_search_pairs(int32_t nq, int32_t ndb, const int64_t *q_off,
                               const int32_t *q_len, const int64_t *db_off,
                               const int32_t *db_len, int64_t *a_off, int32_t *a_len,
                               int64_t *b_off, int32_t *b_len, int32_t query_cols,
                               hipStream_t st) {
  const int64_t n = (int64_t)nq * ndb;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_search_pairs, dim3(nblocks(n, 256)), dim3(256), 0, st, nq, ndb, q_off,
                     q_len, db_off, db_len, a_off, a_len, b_off, b_len, query_cols);
  return hipGetLastError();
}

hipError_t launch_topk(int32_t nq, int32_t ndb, int32_t k, const int32_t *scores,
                       int32_t *top_idx, int32_t *top_score, hipStream_t st) {
  if (nq <= 0 || k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_topk, dim3(nq), dim3(kTopThreads), 0, st, ndb, k, scores, top_idx,
                     top_score);
  return hipGetLastError();
}

}  // namespace ipuma
