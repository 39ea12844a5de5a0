// sw_kernels.hip -- gfx950 kernels of the batched Smith-Waterman affine-gap engine.
//
// Semantics (the contract every kernel here keeps bit-exact) follow the
// reference's SWAffine vertex, src/codelets/codelet_batch_affine.cpp:37-124:
// rows i run over B, columns j over A, H = max(diag + S[a_j][b_i], E, F, 0),
// E = max(H_left + gI + gE, E + gE), F = max(H_up + gI + gE, F + gE), and the
// end is the first cell in B-major order (i, then j) holding the maximum.
//
// Kernel families
//   k_fast<KP, STOP>
//                  packed int16, two comparisons per lane (low / high halves),
//                  L lanes per comparison pipelined over the B rows (L = 1 is
//                  the inter-sequence kernel, L = 64 the intra-sequence
//                  anti-diagonal kernel), persistent over the list of groups
//                  built by the device bucketing below.  STOP: the start
//                  pass's instance, which ends a group once every pair has
//                  reached its forward score.
//   k_wave32       exact int32 re-statement of the vertex loop, one wave per
//                  comparison (a 64-lane pipeline over its rows); re-runs
//                  int16-saturated comparisons and serves gap settings outside
//                  the packed kernel's domain.
//   k_classify ... device-side length bucketing (counting sort by lanes-per-
//                  pair class, super-strips and |A|) and longest-first order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "sw_device.h"

namespace ipuma {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 U2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ s16x2 S2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t W(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t W(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// The larger of three packed 16-bit values per half, in one v_pk_maximum3_f16
// (the compiler fuses this nesting of fp16 maximum).  The DP's values are
// biased by kBias16 so that every H, E and F is a positive normal fp16 bit
// pattern (0x0400 .. 0x7BFF), and the diagonal candidate is one too, a
// positive subnormal (below zero by at most 128: the scores are 8-bit), or a
// negative finite pattern (the pad score's -32768 added): on those the fp16
// order is the integer order, and a pattern at 0x7C00 or above (an overflow,
// Inf or NaN) yields one at 0x7C00 or above, which flags the comparison.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f16x2 F2(uint32_t v) { return __builtin_bit_cast(f16x2, v); }
__device__ __forceinline__ uint32_t max3h(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_bit_cast(uint32_t,
                            __builtin_elementwise_maximum(F2(a), __builtin_elementwise_maximum(F2(b), F2(c))));
}
// kBias16 in both halves: H = 0 (and the floor of E and F)
constexpr uint32_t kFloor2 = (uint32_t)kBias16 * 0x10001u;

// Pipeline modes of the packed kernel: one lane per comparison (L = 1, no
// cross-lane moves), 2..16 lanes (DPP row_shr:1 inside rows of 16 lanes), and
// the whole wave (L = 64, DPP wave_shr:1).
constexpr int kModeSolo = 0;
// Longest padded |A| of a whole-wave group with row-pair keys: its columns
// run to mpad + 62 and must fit 15 bits.
constexpr int kPairKeyMaxM = 32768 - 64;
constexpr int kModeRow = 1;
// Largest H a row-tag key holds: H * 16 + 15 fits 16 bits.
constexpr int kTagKeyMaxH = 4095;
constexpr int kModeWave = 2;

// Value of lane k-1 of the same pipeline (k = lane % L); the first lane of a
// pipeline (k == 0) takes `own`, selected explicitly.
template <int MODE>
__device__ __forceinline__ uint32_t from_prev(uint32_t own, uint32_t v, bool first) {
  if constexpr (MODE == kModeSolo) {
    (void)v; (void)first;
    return own;
  } else if constexpr (MODE == kModeWave) {
    // DPP wave_shr:1: lane i takes lane i-1 across the whole wave; lane 0 has
    // no source and keeps `old` (bound_ctrl off), which is `own`
    (void)first;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)own, (int)v, 0x138, 0xf, 0xf, false);
  } else {
    // DPP row_shr:1 inside rows of 16 lanes; a lane without a source (a row
    // start) reads 0 (bound_ctrl), and every first lane of a pipeline (row
    // starts included) selects `own`
    const uint32_t r = (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xf, 0xf, true);
    return first ? own : r;
  }
}

// ---------------------------------------------------------------------------
// Packed int16 kernel.
//
// A wave runs one group: 64/L pipelines of L lanes; pipeline g holds two
// comparisons, slot g (low halves) and slot 64/L + g (high halves), of equal
// padded shape.  A super-strip is L*kRows rows of B; lane k of a pipeline owns
// rows [ (ss*L + k)*kRows, +kRows ) and at step t computes column j = t - k of
// them, receiving H/F of the row above (and the column's A code) from lane k-1
// (from_prev).  Lane 0 reads the row above the super-strip from the wave's
// scratch, lane L-1 writes the last row back in place.  L is a power of two
// read from the group descriptor; the three modes compile the lane moves.
//
// Arithmetic is the reference recurrence with E and F clamped at 0:
// H = max(diag + s, E', F') with E' = max(E, 0), F' = max(F, 0) is unchanged
// because H is clamped at 0, and E'/F' follow E' <- max(H - o, E' - x, 0) when
// o = -(gI + gE) >= 0 and x = -gE >= 0.  Every value is held plus kBias16
// (1024): H = max3(diag + s, E', F') and E' = max3(H - o, E' - x, 1024) are
// one v_pk_maximum3_f16 each after the unsigned saturating subtractions (see
// max3h), and a maximum of at least kSat16 (30720: a biased H at 0x7C00) is
// re-run exactly in int32.
//
// Padding (columns past |A|, rows past |B|, pipeline fill and drain) scores
// the pad code at -32768, so padded cells only carry gap chains of real cells
// and stay strictly below, or tie later in B-major order than, the real cell
// they came from: they never move the maximum or its first position.
//
// Substitution scores come from a row-pair table in LDS: entry (a, b1, b2)
// holds S[a][b1] in its low and S[a][b2] in its high half, so one 32-bit load
// serves rows 2q and 2q+1 of one comparison, and two loads plus two byte
// permutes give both halves' scores of a row pair.  The table (KP^3 words,
// 70 KB for the 26-letter protein alphabet) is not replicated: its loads hit
// random banks, which costs less than the address arithmetic a per-row
// lookup needs.  Where both comparisons of every pipeline read the same A (a
// search's query), word (a, b_lo, b_hi) is a row's pair of scores itself: one
// load per row, no permutes.
//
// Position tracking, row-tag keys: while every H of a super-strip stays below
// 4096 (the best H above it plus the largest score times its rows, checked
// for the whole wave before the super-strip), each row's packed H becomes a
// 16-bit key H * 16 + (15 - r) (one v_pk_mad_u16 for both comparisons), a
// packed max over the lane's rows gives (max H, first row) of the step, and
// one 32-bit key per comparison, that << 16 | (0xFFFF - j), merges the steps:
// the max key is the lane's first maximum in B-major order.  Otherwise, per
// row, key = H << 16 | (0xFFFF - j); pipelines of up to 16 lanes share one
// key between rows 2q and 2q+1: H << 16 | (row 2q ? 1 : 0) << 15 |
// (0x7FFF - j), which orders (max H, min row, min column) inside the pair, and
// rows merge in order with a strict >.  Lanes merge at the end by (max H, min
// row).
// ---------------------------------------------------------------------------
// Issue priority among the waves of one SIMD.  By default the oldest wave
// wins every issue conflict, so the youngest wave on a SIMD barely advances
// the group it took at launch and finishes it alone, after the queue has run
// dry.  Instead each wave publishes the steps left in its group and ranks
// itself among the waves of its SIMD: the most work left issues first, the
// groups on a SIMD finish together, and the queue's tail stays short.
struct SimdShare {
  int *left;   // LDS, per wave slot of the block: steps left (-1: none)
  int self;    // the wave's slot
  int p1, p2;  // the other slots on the same SIMD (a slot that stays -1 if fewer)
};

__device__ __forceinline__ void share_simd(const SimdShare &S, int left, int lane) {
  if (lane == 0)
    __hip_atomic_store(&S.left[S.self], left, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const int o1 = __hip_atomic_load(&S.left[S.p1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const int o2 = __hip_atomic_load(&S.left[S.p2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const int below = __builtin_amdgcn_readfirstlane((o1 < left ? 1 : 0) + (o2 < left ? 1 : 0));
  if (below == 0) __builtin_amdgcn_s_setprio(0);
  else if (below == 1) __builtin_amdgcn_s_setprio(1);
  else __builtin_amdgcn_s_setprio(2);
}

template <int MODE, int KP, bool PAIRKEY, bool STOP>
__device__ __forceinline__ void run_group(const FastArgs &A, const uint32_t *__restrict__ tab,
                                          const GroupDesc gd, uint32_t *__restrict__ wscr,
                                          int lane, int lg_rt, const SimdShare &share) {
  constexpr uint32_t PADC = KP - 1;
  // pipeline depth L = 2^lg: fixed for the solo and whole-wave modes
  const int lg = MODE == kModeSolo ? 0 : MODE == kModeWave ? 6 : lg_rt;
  // Row-pair keys hold 15-bit columns: every column a lane computes (up to
  // mpad + L - 2) must stay below 32768.  Pipelines of up to 16 lanes have
  // |A| <= 16384 (the wave scratch holds mpad * 64/L <= 65536 words); the
  // whole-wave pipeline takes them up to mpad = kPairKeyMaxM and keeps one
  // key per row with 16-bit columns beyond (|A| up to 65535).
  const int L = 1 << lg;
  const int NG = 64 >> lg;  // pipelines per wave
  const int k = lane & (L - 1);
  const int g = lane >> lg;
  const bool first = (k == 0);
  const bool last = (k == L - 1);
  const int mpad = gd.mpad;
  const int nss = gd.nss;

  const int p_lo = A.sorted[gd.base + g];
  const int p_hi = A.sorted[gd.base + NG + g];
  int m_lo = 0, n_lo = 0, m_hi = 0, n_hi = 0;
  const uint8_t *a_lo = A.codes, *b_lo = A.codes, *a_hi = A.codes, *b_hi = A.codes;
  if (p_lo >= 0) {
    m_lo = A.a_len[p_lo];
    n_lo = A.b_len[p_lo];
    a_lo = A.codes + A.a_off[p_lo];
    b_lo = A.codes + A.b_off[p_lo];
  }
  if (p_hi >= 0) {
    m_hi = A.a_len[p_hi];
    n_hi = A.b_len[p_hi];
    a_hi = A.codes + A.a_off[p_hi];
    b_hi = A.codes + A.b_off[p_hi];
  }
  const uint32_t *a32_lo = (const uint32_t *)a_lo;
  const uint32_t *a32_hi = (const uint32_t *)a_hi;
  const uint32_t gap_o = A.gap_o, gap_x = A.gap_x;
  // Shared A: both comparisons of every pipeline align the same A (a search's
  // query), so the row-pair table word (a, b_lo, b_hi) holds a row's scores
  // for both halves and one load replaces two loads and two permutes.  A
  // pipeline with no comparison, or only a low one, shares trivially.
  bool shared = false;
  if constexpr (MODE == kModeRow) {
    const bool same =
        (p_lo < 0 && p_hi < 0) || (p_lo >= 0 && (p_hi < 0 || (a_lo == a_hi && m_lo == m_hi)));
    shared = __ballot(!same) == 0;
  }

  int bh_lo = 0, bi_lo = 0, bj_lo = 0, bh_hi = 0, bi_hi = 0, bj_hi = 0;

  // One super-strip with R rows per lane: kRows, or for the last super-strip
  // of a group the 8 or 12 that cover its rows (lane k then owns rows
  // ss * L * kRows + k * R .. + R).
  // weight of the group's last super-strip (its rows per lane)
  const int wlast =
      step_weight32(PAIRKEY ? last_strip_rows(gd.last_rows, MODE == kModeWave) : kRows);

  // Column table of the group, behind the boundary row in the wave's scratch:
  // entry (t, g) packs the LDS table offsets (in words) of the A codes at
  // column t of pipeline g's two comparisons (the pad code past each |A|,
  // codes outside the matrix clamped to it), for every step a lane 0 takes
  // (t < round4(mpad + L - 1), plus one prefetched).  Built once per group by
  // the pipeline's lanes, it serves every super-strip: lane 0 then loads one
  // word per step instead of decoding, clamping and scaling two codes.
  // One-lane pipelines decode their own codes instead: every lane is a lane 0
  // there, and the table would add a load per lane and step to the boundary
  // row's.
  constexpr uint32_t KK = KP * KP;
  const int ncols = ((mpad + L - 1 + 3) & ~3) + 1;
  uint32_t *const ctab = wscr + (size_t)mpad * NG * 2;
  if constexpr (MODE != kModeSolo) {
    for (int w = k; 4 * w < ncols; w += L) {
      const int t = 4 * w;
      const uint32_t wl = t < m_lo ? a32_lo[w] : 0u;
      const uint32_t wh = t < m_hi ? a32_hi[w] : 0u;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (t + u < ncols) {
          const uint32_t cl = t + u < m_lo ? min((wl >> (8 * u)) & 0xffu, PADC) : PADC;
          const uint32_t ch = t + u < m_hi ? min((wh >> (8 * u)) & 0xffu, PADC) : PADC;
          ctab[(t + u) * NG + g] = cl * KK | (ch * KK) << 16;
        }
      }
    }
    // the entries are read by lane 0 of each pipeline: complete the stores
    // and drop stale L1 lines
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  }
  // TAG: row-tag keys (the super-strip's H values stay below kTagKeyMaxH + 1):
  // per step, a packed 16-bit key H * 16 + (15 - r) per row, a max over the
  // rows (both comparisons at once), then one 32-bit key per comparison.
  auto strip = [&](auto rows_per_lane, auto tag_keys, auto shared_a, const int ss) {
    constexpr int R = decltype(rows_per_lane)::value;
    constexpr bool TAG = decltype(tag_keys)::value;
    constexpr bool SH = decltype(shared_a)::value;
    constexpr bool PK = PAIRKEY && !TAG;
    static_assert(!TAG || (PAIRKEY && R <= 16), "row tags: 4 bits, 16-bit columns");
    constexpr int KN = TAG ? 1 : PAIRKEY ? R / 2 : R;
    // The boundary row written by lane L-1 during the previous super-strip is
    // read back by lane 0: complete those stores and drop stale L1 lines.
    if (ss > 0) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    const int row0 = (ss << lg) * kRows + k * R;
    uint32_t rp[R / 2];  // row-pair table offsets: low half low comparison, high half high comparison
    {
      uint32_t cl[R], ch[R];
      if constexpr (R % 4 == 0) {  // row0 is a multiple of 4: word loads
        uint32_t wl[R / 4], wh[R / 4];
        const uint32_t *bl = (const uint32_t *)(b_lo + row0);
        const uint32_t *bh = (const uint32_t *)(b_hi + row0);
#pragma unroll
        for (int q = 0; q < R / 4; ++q) {
          wl[q] = (row0 + 4 * q < n_lo) ? bl[q] : 0u;
          wh[q] = (row0 + 4 * q < n_hi) ? bh[q] : 0u;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          cl[r] = (row0 + r < n_lo) ? min((wl[r >> 2] >> (8 * (r & 3))) & 0xffu, PADC) : PADC;
          ch[r] = (row0 + r < n_hi) ? min((wh[r >> 2] >> (8 * (r & 3))) & 0xffu, PADC) : PADC;
        }
      } else {  // 10 or 14 rows per lane: row0 = k * R is not word-aligned
#pragma unroll
        for (int r = 0; r < R; ++r) {
          cl[r] = (row0 + r < n_lo) ? min((uint32_t)b_lo[row0 + r], PADC) : PADC;
          ch[r] = (row0 + r < n_hi) ? min((uint32_t)b_hi[row0 + r], PADC) : PADC;
        }
      }
#pragma unroll
      for (int q = 0; q < R / 2; ++q) {
        if constexpr (SH)  // rows 2q and 2q+1, each (low, high) comparison
          rp[q] = ((cl[2 * q] * KP + ch[2 * q]) * 4u) |
                  (((cl[2 * q + 1] * KP + ch[2 * q + 1]) * 4u) << 16);
        else  // low comparison (rows 2q, 2q+1), high comparison
          rp[q] = ((cl[2 * q] * KP + cl[2 * q + 1]) * 4u) |
                  (((ch[2 * q] * KP + ch[2 * q + 1]) * 4u) << 16);
      }
    }
    uint32_t H[R], E[R], KL[KN], KH[KN];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      H[r] = kFloor2;
      E[r] = kFloor2;
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      KL[q] = 0u;
      KH[q] = 0u;
    }
    uint32_t out_h = kFloor2, out_f = kFloor2, hdiag = kFloor2;
    uint32_t colw = PADC * KK * 0x10001u;  // the column's table offsets: low | high << 16
    const bool load_bnd = ss > 0;
    const bool store_bnd = ss + 1 < nss;

    uint2 nb = make_uint2(kFloor2, kFloor2);
    if (first && load_bnd && mpad > 0) nb = *(const uint2 *)(wscr + (size_t)g * 2);
    uint32_t ncw = 0u;
    uint32_t na_lo = 0u, na_hi = 0u;  // one-lane pipelines: the next four A codes
    if constexpr (MODE == kModeSolo) {
      na_lo = (0 < m_lo) ? a32_lo[0] : 0u;
      na_hi = (0 < m_hi) ? a32_hi[0] : 0u;
    } else {
      if (first) ncw = ctab[g];
    }
    const int T = mpad + L - 1;
    for (int t0 = 0; t0 < T; t0 += 4) {
      if ((t0 & 7) == 0)  // work left, in 32nds of a kRows step
        share_simd(share, ((nss - 1 - ss) * 32 - 32 + wlast) * T + (T - t0) * step_weight32(R) +
                              (ss + 1 == nss ? (32 - wlast) * T : 0), lane);
      const uint32_t wa_lo = na_lo, wa_hi = na_hi;
      if constexpr (MODE == kModeSolo) {
        const int tn = t0 + 4;
        na_lo = (tn < m_lo) ? a32_lo[tn >> 2] : 0u;
        na_hi = (tn < m_hi) ? a32_hi[tn >> 2] : 0u;
      }
#pragma unroll 1
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u;
        const uint2 cb = nb;
        const uint32_t ccw = ncw;
        // t + 1 < ncols: the table covers every step of the loop plus one
        if (MODE != kModeSolo && first) ncw = ctab[(t + 1) * NG + g];
        // byte offsets into the wave scratch fit 32 bits (it holds 769 KiB)
        if (first && load_bnd && t + 1 < mpad)
          nb = *(const uint2 *)((const char *)wscr + (uint32_t)(((t + 1) * NG + g) * 8));
        else
          nb = make_uint2(kFloor2, kFloor2);
        const uint32_t rh = from_prev<MODE>(cb.x, out_h, first);
        const uint32_t rf = from_prev<MODE>(cb.y, out_f, first);
        uint32_t col_lo, col_hi;  // byte offsets of the column's A-code rows of the table
        if constexpr (MODE == kModeSolo) {
          col_lo = ((t < m_lo) ? min((wa_lo >> (8 * u)) & 0xffu, PADC) : PADC) * (KK * 4);
          col_hi = ((t < m_hi) ? min((wa_hi >> (8 * u)) & 0xffu, PADC) : PADC) * (KK * 4);
        } else {
          colw = from_prev<MODE>(ccw, colw, first);
          col_lo = (colw & 0xFFFFu) * 4u;
          col_hi = (colw >> 16) * 4u;
        }
        const int j = t - k;
        // key low parts: 16-bit 0xFFFF - j, or with row pairs the preferred
        // (even) row's bit 15 over a 15-bit 0x7FFF - j
        const uint32_t jinv = PK ? ((0x7FFFu - (uint32_t)j) & 0x7FFFu) | 0x8000u
                                 : 0xFFFFu - (uint32_t)j;
        const char *tl = (const char *)tab + col_lo;
        const char *th = (const char *)tab + col_hi;
        // keep the row-pair offsets packed two per register across the loop
        // (pinned once per step, ahead of the loads, so the R loads can issue
        // back to back and their latency overlaps)
#pragma unroll
        for (int q = 0; q < R / 2; ++q) asm volatile("" : "+v"(rp[q]));
        uint32_t sv[R];
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
          if constexpr (SH) {
            sv[2 * q] = *(const uint32_t *)(tl + (rp[q] & 0xFFFFu));
            sv[2 * q + 1] = *(const uint32_t *)(tl + (rp[q] >> 16));
          } else {
            const uint32_t wlo = *(const uint32_t *)(tl + (rp[q] & 0xFFFFu));
            const uint32_t whi = *(const uint32_t *)(th + (rp[q] >> 16));
            sv[2 * q] = __builtin_amdgcn_perm(whi, wlo, 0x05040100u);
            sv[2 * q + 1] = __builtin_amdgcn_perm(whi, wlo, 0x07060302u);
          }
        }
        // t2 of row r is diag + s; the old H of row r is consumed (as the
        // diagonal of row r + 1) before the new H of row r is written.
        s16x2 t2 = __builtin_elementwise_add_sat(S2(hdiag), S2(sv[0]));
        hdiag = rh;
        uint32_t f = rf;
        uint32_t kl_even = 0u, kh_even = 0u;
        // the two row-pair key tails, computed once per step and pinned in
        // VGPRs (otherwise they are re-derived per row, or a uniform tail
        // lands in an SGPR where v_and_or_b32 cannot take it with its mask)
        uint32_t jr_even = jinv, jr_odd = jinv & 0x7FFFu;
        if constexpr (PK) asm volatile("" : "+v"(jr_even), "+v"(jr_odd));
        u16x2 tk2[2] = {(u16x2)0, (u16x2)0};  // row-tag keys: running maxima over the rows
        uint32_t tag_mul = 0x00100010u;  // 16 in both halves
        if constexpr (TAG) asm volatile("" : "+v"(tag_mul));
#pragma unroll
        for (int r = 0; r < R; ++r) {
          s16x2 t2n = t2;
          if (r + 1 < R) t2n = __builtin_elementwise_add_sat(S2(H[r]), S2(sv[r + 1]));
          const uint32_t hn = max3h(W(t2), E[r], f);
          H[r] = hn;
          const uint32_t ho = W(__builtin_elementwise_sub_sat(U2(hn), U2(gap_o)));
          E[r] = max3h(ho, W(__builtin_elementwise_sub_sat(U2(E[r]), U2(gap_x))), kFloor2);
          f = max3h(ho, W(__builtin_elementwise_sub_sat(U2(f), U2(gap_x))), kFloor2);
          if constexpr (TAG) {
            // H < 4096 in both halves: H * 16 + (15 - r) in 16 bits orders
            // (max H, min row) within the step; from the biased H that is
            // H' * 16 + (15 - r - 16 * kBias16) modulo 2^16.  One
            // v_pk_mad_u16 per row: the multiplier is opaque to the compiler
            // (a shift and an or would take two instructions, and the tag
            // constants registers)
            const u16x2 kr = U2(hn) * U2(tag_mul) + (unsigned short)(15 - r - 16 * kBias16);
            // two running maxima (even and odd rows) keep the chain short
            if (r < 2) tk2[r] = kr;
            else tk2[r & 1] = __builtin_elementwise_max(tk2[r & 1], kr);
          } else if constexpr (PAIRKEY) {
            // H << 16 | row bit << 15 | 15-bit column; one max3 per row pair
            const uint32_t jr = (r & 1) ? jr_odd : jr_even;
            const uint32_t kl = (hn << 16) | jr;
            const uint32_t kh = (hn & 0xFFFF0000u) | jr;
            if (r & 1) {
              KL[r >> 1] = max(max(KL[r >> 1], kl_even), kl);
              KH[r >> 1] = max(max(KH[r >> 1], kh_even), kh);
            } else {
              kl_even = kl;
              kh_even = kh;
            }
          } else {
            KL[r] = max(KL[r], __builtin_amdgcn_perm(hn, jinv, 0x05040100u));
            KH[r] = max(KH[r], __builtin_amdgcn_perm(hn, jinv, 0x07060100u));
          }
          t2 = t2n;
        }
        if constexpr (TAG) {
          // step key over 16-bit columns: (max H, min row) << 16 | 0xFFFF - j
          const u16x2 tk = __builtin_elementwise_max(tk2[0], tk2[1]);
          KL[0] = max(KL[0], __builtin_amdgcn_perm(W(tk), jinv, 0x05040100u));
          KH[0] = max(KH[0], __builtin_amdgcn_perm(W(tk), jinv, 0x07060100u));
        }
        out_h = H[R - 1];
        out_f = f;
        // the last lane of every pipeline is at column t - (L - 1): a uniform
        // (scalar) range test, then the lane mask
        const int jl = t - (L - 1);
        if (store_bnd && (uint32_t)jl < (uint32_t)mpad && last)
          *(uint2 *)((char *)wscr + (uint32_t)((jl * NG + g) * 8)) = make_uint2(out_h, out_f);
      }
    }
    // rows (or row pairs) in order, strict >: the first maximum in B order
    if constexpr (TAG) {
      const int hl = (int)(KL[0] >> 20);
      if (hl > bh_lo) {
        bh_lo = hl;
        bi_lo = row0 + 15 - (int)((KL[0] >> 16) & 15u);
        bj_lo = 0xFFFF - (int)(KL[0] & 0xFFFFu);
      }
      const int hh = (int)(KH[0] >> 20);
      if (hh > bh_hi) {
        bh_hi = hh;
        bi_hi = row0 + 15 - (int)((KH[0] >> 16) & 15u);
        bj_hi = 0xFFFF - (int)(KH[0] & 0xFFFFu);
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int hl = (int)(KL[q] >> 16) - kBias16;
      if (hl > bh_lo) {
        bh_lo = hl;
        if constexpr (PAIRKEY) {
          bi_lo = row0 + 2 * q + ((KL[q] & 0x8000u) ? 0 : 1);
          bj_lo = 0x7FFF - (int)(KL[q] & 0x7FFFu);
        } else {
          bi_lo = row0 + q;
          bj_lo = 0xFFFF - (int)(KL[q] & 0xFFFFu);
        }
      }
      const int hh = (int)(KH[q] >> 16) - kBias16;
      if (hh > bh_hi) {
        bh_hi = hh;
        if constexpr (PAIRKEY) {
          bi_hi = row0 + 2 * q + ((KH[q] & 0x8000u) ? 0 : 1);
          bj_hi = 0x7FFF - (int)(KH[q] & 0x7FFFu);
        } else {
          bi_hi = row0 + q;
          bj_hi = 0xFFFF - (int)(KH[q] & 0xFFFFu);
        }
      }
    }
  };
  // Start pass: stop after the super-strip by which every pair of the wave has
  // reached its target (a pipeline has it when any of its lanes does).
  // (compiled into the start pass's kernel only: the check's registers would
  // push the forward kernel's step loops into spilling)
  const int32_t *const target = A.target;
  using Full = std::integral_constant<int, kRows>;
  // Key form of a super-strip, uniform over the wave.  A path into the
  // super-strip gains at most the largest score per row it adds (gaps cost),
  // so every H in it is at most the best H above it (a lane's bh bounds its
  // pipeline's best only together with the other lanes: the ballot covers
  // every lane) plus A.smax times its rows.  Row-tag keys hold when that stays
  // within kTagKeyMaxH for every pair of the wave.
  //
  // The last super-strip of a group (nothing after it reads its boundary row,
  // so it can run again) takes row-tag keys even when the bound does not hold,
  // as long as the best H above it stays within kTagKeyMaxH - A.smax, and
  // checks afterwards: an H above kTagKeyMaxH is reached from below that by
  // steps of at most A.smax up the diagonals (gaps only lower H), so some cell
  // of the super-strip then holds an H within A.smax of kTagKeyMaxH, whose key
  // is exact.  When no lane's maximum came that close, every key was exact;
  // otherwise the super-strip runs again with 32-bit keys.
  auto run_strip = [&](auto rows, const int ss) {
    constexpr int R = decltype(rows)::value;
    if constexpr (PAIRKEY) {
      const int lim = kTagKeyMaxH - A.smax * (R << lg);
      const int near = kTagKeyMaxH - A.smax;
      const bool bound = __ballot(max(bh_lo, bh_hi) > lim) == 0;
      if (bound || (ss + 1 == nss && __ballot(max(bh_lo, bh_hi) > near) == 0)) {
        const int s_hl = bh_lo, s_il = bi_lo, s_jl = bj_lo, s_hh = bh_hi, s_ih = bi_hi, s_jh = bj_hi;
        if (MODE == kModeRow && shared) strip(rows, std::true_type{}, std::true_type{}, ss);
        else strip(rows, std::true_type{}, std::false_type{}, ss);
        if (bound || __ballot(max(bh_lo, bh_hi) > near) == 0) return;
        bh_lo = s_hl, bi_lo = s_il, bj_lo = s_jl, bh_hi = s_hh, bi_hi = s_ih, bj_hi = s_jh;
      }
    }
    strip(rows, std::false_type{}, std::false_type{}, ss);
  };
  bool reached = false;
  for (int ss = 0; ss + 1 < nss; ++ss) {
    run_strip(Full{}, ss);
    if constexpr (STOP) {
      const int tg_lo = p_lo >= 0 ? target[p_lo] : INT32_MIN;
      const int tg_hi = p_hi >= 0 ? target[p_hi] : INT32_MIN;
      const uint64_t mlo = __ballot(bh_lo >= tg_lo), mhi = __ballot(bh_hi >= tg_hi);
      const uint64_t win = L == 64 ? ~0ull : ((1ull << L) - 1ull);
      const bool ok = !first || ((((mlo >> lane) & win) != 0) && (((mhi >> lane) & win) != 0));
      if (__ballot(!ok) == 0) {
        reached = true;
        break;
      }
    }
  }
  if (nss > 0 && !reached) {
    if constexpr (PAIRKEY) {
      const int rl = last_strip_rows(gd.last_rows, MODE == kModeWave);
      if (rl == 8) run_strip(std::integral_constant<int, 8>{}, nss - 1);
      else if (rl == 12) run_strip(std::integral_constant<int, 12>{}, nss - 1);
      else if (MODE == kModeWave && rl == 10) run_strip(std::integral_constant<int, 10>{}, nss - 1);
      else if (MODE == kModeWave && rl == 14) run_strip(std::integral_constant<int, 14>{}, nss - 1);
      else run_strip(Full{}, nss - 1);
    } else {
      run_strip(Full{}, nss - 1);
    }
  }

  if constexpr (MODE != kModeSolo) {
    for (int off = 1; off < L; off <<= 1) {
      const int oh = __shfl_xor(bh_lo, off), oi = __shfl_xor(bi_lo, off), oj = __shfl_xor(bj_lo, off);
      if (oh > bh_lo || (oh == bh_lo && oi < bi_lo)) {
        bh_lo = oh;
        bi_lo = oi;
        bj_lo = oj;
      }
      const int qh = __shfl_xor(bh_hi, off), qi = __shfl_xor(bi_hi, off), qj = __shfl_xor(bj_hi, off);
      if (qh > bh_hi || (qh == bh_hi && qi < bi_hi)) {
        bh_hi = qh;
        bi_hi = qi;
        bj_hi = qj;
      }
    }
  }
  if (first) {
    if (p_lo >= 0) {
      if (bh_lo >= kSat16) {
        A.ovf_list[atomicAdd(A.ovf_count, 1)] = p_lo;
      } else {
        A.scores[p_lo] = bh_lo;
        A.a_range[p_lo] = (int32_t)((uint32_t)(bj_lo & 0xFFFF) << 16);
        A.b_range[p_lo] = (int32_t)((uint32_t)(bi_lo & 0xFFFF) << 16);
      }
    }
    if (p_hi >= 0) {
      if (bh_hi >= kSat16) {
        A.ovf_list[atomicAdd(A.ovf_count, 1)] = p_hi;
      } else {
        A.scores[p_hi] = bh_hi;
        A.a_range[p_hi] = (int32_t)((uint32_t)(bj_hi & 0xFFFF) << 16);
        A.b_range[p_hi] = (int32_t)((uint32_t)(bi_hi & 0xFFFF) << 16);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Band groups: whole-wave pipeline groups (L = 64) with two or more
// super-strips, shared by the waves of one block.  Super-strip s needs the
// boundary row that super-strip s-1 writes; the group's boundary buffer is
// updated in place (strip s reads column j at its step j - 1 and overwrites it
// at step j + 63), and the writer of strip s-1 publishes how many columns it
// has written in LDS (prog[s-1]), which strip s waits for before reading.
// Strips are handed out in order, so a strip only waits for one handed out
// earlier, which itself only waits for an earlier one: every wait ends.  A
// block holds at most one band group per wave at a time (one slot each, each
// slot with its own boundary buffer); the wave that finds the current group's
// strips exhausted loads the next band group into a free slot.
//
// Ticket, CAS and counter updates run on every lane with lane 0 carrying the
// update and the other lanes a no-op (CAS with an impossible expected value,
// add of 0), read back with readfirstlane: no lane-conditional branch
// surrounds them (see next_group).
// ---------------------------------------------------------------------------
struct BandSlot {
  GroupDesc gd;
  int32_t ticket;  // seq << 16 | nss << 8 | next strip; next == nss + 1: loading the next group
  int32_t done;    // strips finished
  int32_t prog[kMaxStrips];      // boundary columns each strip has written
  unsigned long long best[2];    // H << 32 | (0xFFFF - row) << 16 | column (low, high)
};
struct BandState {
  BandSlot slot[kBandSlots + 1];  // [kBandSlots]: the empty group before the first
  int32_t cur;                    // seq << 4 | slot of the group handing out strips, -1: none left
  int32_t free_mask;              // slots not holding a group
};
static_assert(kBandSlots < 16 && kMaxStrips < 255, "band ticket layout");
// Steps by which super-strip s + 1 of a band group trails super-strip s: it
// reads boundary column j at its step j - 1, which strip s writes at its step
// j + 63 and publishes at most 16 steps later.
constexpr int kBandLag = 76;

// Loads of LDS words other waves update are read back through readfirstlane:
// the wave's control flow depends on them, so it must see one value.
__device__ __forceinline__ int lds_load_acq(int *p) {
  return __builtin_amdgcn_readfirstlane(
      __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ __forceinline__ void lds_store_rel(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Compare-and-swap carried by lane 0; the wave reads lane 0's result.
__device__ __forceinline__ int lds_cas_lane0(int *p, int expected, int desired, int lane) {
  int e = lane == 0 ? expected : -1;  // tickets are never -1: the other lanes never write
  const int d = lane == 0 ? desired : -1;
  __hip_atomic_compare_exchange_strong(p, &e, d, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
  return __builtin_amdgcn_readfirstlane(e);
}

// Waits until the producer of the boundary row has written `need` columns.
__device__ __forceinline__ int band_wait(int *prog, int need) {
  int v = lds_load_acq(prog);
  while (v < need) {
    __builtin_amdgcn_s_sleep(1);
    v = lds_load_acq(prog);
  }
  return v;
}

// Super-strip s of the band group in slot `sl`: the whole-wave pipeline body
// of run_group, with the boundary row shared through the slot's buffer.
template <int KP, bool PAIRKEY, bool TAG>
__device__ __forceinline__ void band_strip(const FastArgs &A, const uint32_t *__restrict__ tab, BandState &B,
                           int sl, int s, int lane, const SimdShare &share) {
  constexpr int R = kRows;
  constexpr uint32_t PADC = KP - 1;
  constexpr uint32_t COLSTRIDE = KP * KP * 4;
  BandSlot &S = B.slot[sl];
  const int k = lane;
  const bool first = (k == 0);
  const bool last = (k == 63);
  const int mpad = __builtin_amdgcn_readfirstlane(S.gd.mpad);
  const int nss = __builtin_amdgcn_readfirstlane(S.gd.nss);
  const int gbase = __builtin_amdgcn_readfirstlane(S.gd.base);
  uint32_t *__restrict__ bscr =
      A.band_scratch + ((int64_t)blockIdx.x * kBandSlots + sl) * A.band_words;

  const int p_lo = A.sorted[gbase];
  const int p_hi = A.sorted[gbase + 1];
  int m_lo = 0, n_lo = 0, m_hi = 0, n_hi = 0;
  const uint8_t *a_lo = A.codes, *b_lo = A.codes, *a_hi = A.codes, *b_hi = A.codes;
  if (p_lo >= 0) {
    m_lo = A.a_len[p_lo];
    n_lo = A.b_len[p_lo];
    a_lo = A.codes + A.a_off[p_lo];
    b_lo = A.codes + A.b_off[p_lo];
  }
  if (p_hi >= 0) {
    m_hi = A.a_len[p_hi];
    n_hi = A.b_len[p_hi];
    a_hi = A.codes + A.a_off[p_hi];
    b_hi = A.codes + A.b_off[p_hi];
  }
  const uint32_t *a32_lo = (const uint32_t *)a_lo;
  const uint32_t *a32_hi = (const uint32_t *)a_hi;
  const uint32_t gap_o = A.gap_o, gap_x = A.gap_x;

  const int row0 = ((s << 6) + k) * R;
  uint32_t rp[R / 2];  // row-pair table offsets: low half low comparison, high half high comparison
  {
    uint32_t wl[R / 4], wh[R / 4];
    const uint32_t *bl = (const uint32_t *)(b_lo + row0);
    const uint32_t *bh = (const uint32_t *)(b_hi + row0);
#pragma unroll
    for (int q = 0; q < R / 4; ++q) {
      wl[q] = (row0 + 4 * q < n_lo) ? bl[q] : 0u;
      wh[q] = (row0 + 4 * q < n_hi) ? bh[q] : 0u;
    }
    uint32_t cl[R], ch[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cl[r] = (row0 + r < n_lo) ? min((wl[r >> 2] >> (8 * (r & 3))) & 0xffu, PADC) : PADC;
      ch[r] = (row0 + r < n_hi) ? min((wh[r >> 2] >> (8 * (r & 3))) & 0xffu, PADC) : PADC;
    }
#pragma unroll
    for (int q = 0; q < R / 2; ++q)
      rp[q] = ((cl[2 * q] * KP + cl[2 * q + 1]) * 4u) | (((ch[2 * q] * KP + ch[2 * q + 1]) * 4u) << 16);
  }
  // keys as in run_group: one per row pair (15-bit columns) up to
  // mpad = kPairKeyMaxM, one per row (16-bit columns) beyond, or with TAG
  // (every H of the strip below 4,096) row-tag keys over 16-bit columns
  constexpr int KN = TAG ? 1 : PAIRKEY ? R / 2 : R;
  uint32_t H[R], E[R], KL[KN], KH[KN];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    H[r] = kFloor2;
    E[r] = kFloor2;
  }
#pragma unroll
  for (int q = 0; q < KN; ++q) {
    KL[q] = 0u;
    KH[q] = 0u;
  }
  uint32_t out_h = kFloor2, out_f = kFloor2, hdiag = kFloor2;
  uint32_t col_lo = PADC * COLSTRIDE, col_hi = PADC * COLSTRIDE;
  const bool load_bnd = s > 0;
  const bool store_bnd = s + 1 < nss;
  int *const prev_prog = &S.prog[s > 0 ? s - 1 : 0];
  int *const my_prog = &S.prog[s];
  int avail = 0;

  uint2 nb = make_uint2(kFloor2, kFloor2);
  if (load_bnd && mpad > 0) {
    avail = band_wait(prev_prog, 1);
    if (first) nb = *(const uint2 *)bscr;
  }
  uint32_t na_lo = 0u, na_hi = 0u;
  if (first) {
    na_lo = (0 < m_lo) ? a32_lo[0] : 0u;
    na_hi = (0 < m_hi) ? a32_hi[0] : 0u;
  }
  const int T = mpad + 63;
  for (int t0 = 0; t0 < T; t0 += 4) {
    if ((t0 & 7) == 0) {
      // The group's remaining critical path through this strip, in 32nds of
      // a kRows step like the other groups' work left: this strip's steps,
      // the lag of each strip after it, and one more strip's length while
      // some of the group's strips wait for a wave.
      const int tk = lds_load_acq(&S.ticket);
      const bool waiting = (tk & 0xFF) < nss;
      share_simd(share, ((T - t0) + (nss - 1 - s) * kBandLag + (waiting ? T : 0)) * 32, lane);
    }
    // columns t0 + 1 .. t0 + 4 of the boundary row are read below
    if (load_bnd) {
      const int need = min(mpad, t0 + 5);
      if (avail < need) avail = band_wait(prev_prog, need);
    }
    const uint32_t wa_lo = na_lo, wa_hi = na_hi;
    if (first) {
      const int tn = t0 + 4;
      na_lo = (tn < m_lo) ? a32_lo[tn >> 2] : 0u;
      na_hi = (tn < m_hi) ? a32_hi[tn >> 2] : 0u;
    }
#pragma unroll 1
    for (int u = 0; u < 4; ++u) {
      const int t = t0 + u;
      const uint2 cb = nb;
      if (first && load_bnd && t + 1 < mpad)
        nb = *(const uint2 *)(bscr + (size_t)(t + 1) * 2);
      else
        nb = make_uint2(kFloor2, kFloor2);
      const uint32_t ac_lo = (t < m_lo) ? min((wa_lo >> (8 * u)) & 0xffu, PADC) : PADC;
      const uint32_t ac_hi = (t < m_hi) ? min((wa_hi >> (8 * u)) & 0xffu, PADC) : PADC;
      const uint32_t rh = from_prev<kModeWave>(cb.x, out_h, first);
      const uint32_t rf = from_prev<kModeWave>(cb.y, out_f, first);
      col_lo = from_prev<kModeWave>(ac_lo * COLSTRIDE, col_lo, first);
      col_hi = from_prev<kModeWave>(ac_hi * COLSTRIDE, col_hi, first);
      const int j = t - k;
      const uint32_t jinv = PAIRKEY && !TAG ? ((0x7FFFu - (uint32_t)j) & 0x7FFFu) | 0x8000u
                                            : 0xFFFFu - (uint32_t)j;
      const char *tl = (const char *)tab + col_lo;
      const char *th = (const char *)tab + col_hi;
#pragma unroll
      for (int q = 0; q < R / 2; ++q) asm volatile("" : "+v"(rp[q]));
      uint32_t sv[R];
#pragma unroll
      for (int q = 0; q < R / 2; ++q) {
        const uint32_t wlo = *(const uint32_t *)(tl + (rp[q] & 0xFFFFu));
        const uint32_t whi = *(const uint32_t *)(th + (rp[q] >> 16));
        sv[2 * q] = __builtin_amdgcn_perm(whi, wlo, 0x05040100u);
        sv[2 * q + 1] = __builtin_amdgcn_perm(whi, wlo, 0x07060302u);
      }
      s16x2 t2 = __builtin_elementwise_add_sat(S2(hdiag), S2(sv[0]));
      hdiag = rh;
      uint32_t f = rf;
      uint32_t kl_even = 0u, kh_even = 0u;
      uint32_t jr_even = jinv, jr_odd = jinv & 0x7FFFu;
      if constexpr (PAIRKEY && !TAG) asm volatile("" : "+v"(jr_even), "+v"(jr_odd));
      u16x2 tk2[2] = {(u16x2)0, (u16x2)0};  // row-tag keys: running maxima over the rows
      uint32_t tag_mul = 0x00100010u;       // 16 in both halves, opaque (one v_pk_mad_u16 per row)
      if constexpr (TAG) asm volatile("" : "+v"(tag_mul));
#pragma unroll
      for (int r = 0; r < R; ++r) {
        s16x2 t2n = t2;
        if (r + 1 < R) t2n = __builtin_elementwise_add_sat(S2(H[r]), S2(sv[r + 1]));
        const uint32_t hn = max3h(W(t2), E[r], f);
        H[r] = hn;
        const uint32_t ho = W(__builtin_elementwise_sub_sat(U2(hn), U2(gap_o)));
        E[r] = max3h(ho, W(__builtin_elementwise_sub_sat(U2(E[r]), U2(gap_x))), kFloor2);
        f = max3h(ho, W(__builtin_elementwise_sub_sat(U2(f), U2(gap_x))), kFloor2);
        if constexpr (TAG) {
          // H * 16 + (15 - r) from the biased H, as in run_group
          const u16x2 kr = U2(hn) * U2(tag_mul) + (unsigned short)(15 - r - 16 * kBias16);
          if (r < 2) tk2[r] = kr;
          else tk2[r & 1] = __builtin_elementwise_max(tk2[r & 1], kr);
        } else if constexpr (PAIRKEY) {
          const uint32_t jr = (r & 1) ? jr_odd : jr_even;
          const uint32_t kl = (hn << 16) | jr;
          const uint32_t kh = (hn & 0xFFFF0000u) | jr;
          if (r & 1) {
            KL[r >> 1] = max(max(KL[r >> 1], kl_even), kl);
            KH[r >> 1] = max(max(KH[r >> 1], kh_even), kh);
          } else {
            kl_even = kl;
            kh_even = kh;
          }
        } else {
          KL[r] = max(KL[r], __builtin_amdgcn_perm(hn, jinv, 0x05040100u));
          KH[r] = max(KH[r], __builtin_amdgcn_perm(hn, jinv, 0x07060100u));
        }
        t2 = t2n;
      }
      if constexpr (TAG) {
        const u16x2 tk = __builtin_elementwise_max(tk2[0], tk2[1]);
        KL[0] = max(KL[0], __builtin_amdgcn_perm(W(tk), jinv, 0x05040100u));
        KH[0] = max(KH[0], __builtin_amdgcn_perm(W(tk), jinv, 0x07060100u));
      }
      out_h = H[R - 1];
      out_f = f;
      // lane 63 is at column t - 63: a uniform range test, then the lane mask
      if (store_bnd && (uint32_t)(t - 63) < (uint32_t)mpad && last)
        *(uint2 *)(bscr + (size_t)(t - 63) * 2) = make_uint2(out_h, out_f);
    }
    // publish the boundary columns written so far (lane 63 wrote j = t - 63)
    if (store_bnd && (t0 & 15) == 12) {
      __builtin_amdgcn_wave_barrier();
      lds_store_rel(my_prog, max(0, min(mpad, t0 - 59)));
    }
  }
  if (store_bnd) {
    __builtin_amdgcn_wave_barrier();
    lds_store_rel(my_prog, mpad);
  }

  int bh_lo = 0, bi_lo = 0, bj_lo = 0, bh_hi = 0, bi_hi = 0, bj_hi = 0;
  if constexpr (TAG) {
    const int hl = (int)(KL[0] >> 20);
    if (hl > 0) {
      bh_lo = hl;
      bi_lo = row0 + 15 - (int)((KL[0] >> 16) & 15u);
      bj_lo = 0xFFFF - (int)(KL[0] & 0xFFFFu);
    }
    const int hh = (int)(KH[0] >> 20);
    if (hh > 0) {
      bh_hi = hh;
      bi_hi = row0 + 15 - (int)((KH[0] >> 16) & 15u);
      bj_hi = 0xFFFF - (int)(KH[0] & 0xFFFFu);
    }
  }
#pragma unroll
  for (int q = 0; q < (TAG ? 0 : KN); ++q) {
    const int hl = (int)(KL[q] >> 16) - kBias16;
    if (hl > bh_lo) {
      bh_lo = hl;
      if constexpr (PAIRKEY) {
        bi_lo = row0 + 2 * q + ((KL[q] & 0x8000u) ? 0 : 1);
        bj_lo = 0x7FFF - (int)(KL[q] & 0x7FFFu);
      } else {
        bi_lo = row0 + q;
        bj_lo = 0xFFFF - (int)(KL[q] & 0xFFFFu);
      }
    }
    const int hh = (int)(KH[q] >> 16) - kBias16;
    if (hh > bh_hi) {
      bh_hi = hh;
      if constexpr (PAIRKEY) {
        bi_hi = row0 + 2 * q + ((KH[q] & 0x8000u) ? 0 : 1);
        bj_hi = 0x7FFF - (int)(KH[q] & 0x7FFFu);
      } else {
        bi_hi = row0 + q;
        bj_hi = 0xFFFF - (int)(KH[q] & 0xFFFFu);
      }
    }
  }
  for (int off = 1; off < 64; off <<= 1) {
    const int oh = __shfl_xor(bh_lo, off), oi = __shfl_xor(bi_lo, off), oj = __shfl_xor(bj_lo, off);
    if (oh > bh_lo || (oh == bh_lo && oi < bi_lo)) {
      bh_lo = oh;
      bi_lo = oi;
      bj_lo = oj;
    }
    const int qh = __shfl_xor(bh_hi, off), qi = __shfl_xor(bi_hi, off), qj = __shfl_xor(bj_hi, off);
    if (qh > bh_hi || (qh == bh_hi && qi < bi_hi)) {
      bh_hi = qh;
      bi_hi = qi;
      bj_hi = qj;
    }
  }
  // strips hold disjoint rows: the group's first maximum is (max H, min row)
  const unsigned long long key_lo = ((unsigned long long)(uint32_t)bh_lo << 32) |
                                    ((unsigned long long)(0xFFFF - (bi_lo & 0xFFFF)) << 16) |
                                    (unsigned long long)(bj_lo & 0xFFFF);
  const unsigned long long key_hi = ((unsigned long long)(uint32_t)bh_hi << 32) |
                                    ((unsigned long long)(0xFFFF - (bi_hi & 0xFFFF)) << 16) |
                                    (unsigned long long)(bj_hi & 0xFFFF);
  __hip_atomic_fetch_max(&S.best[0], key_lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_max(&S.best[1], key_hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const int d = __builtin_amdgcn_readfirstlane(__hip_atomic_fetch_add(
      &S.done, lane == 0 ? 1 : 0, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP));
  if (d != nss - 1) return;
  // the last strip of the group: write its results and free the slot
  const unsigned long long bl =
      __hip_atomic_load(&S.best[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
  const unsigned long long bhk =
      __hip_atomic_load(&S.best[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (first) {
    const int sl_h = (int)(bl >> 32), sl_i = 0xFFFF - (int)((bl >> 16) & 0xFFFF),
              sl_j = (int)(bl & 0xFFFF);
    const int sh_h = (int)(bhk >> 32), sh_i = 0xFFFF - (int)((bhk >> 16) & 0xFFFF),
              sh_j = (int)(bhk & 0xFFFF);
    if (p_lo >= 0) {
      if (sl_h >= kSat16) {
        A.ovf_list[atomicAdd(A.ovf_count, 1)] = p_lo;
      } else {
        A.scores[p_lo] = sl_h;
        A.a_range[p_lo] = (int32_t)((uint32_t)sl_j << 16);
        A.b_range[p_lo] = (int32_t)((uint32_t)sl_i << 16);
      }
    }
    if (p_hi >= 0) {
      if (sh_h >= kSat16) {
        A.ovf_list[atomicAdd(A.ovf_count, 1)] = p_hi;
      } else {
        A.scores[p_hi] = sh_h;
        A.a_range[p_hi] = (int32_t)((uint32_t)sh_j << 16);
        A.b_range[p_hi] = (int32_t)((uint32_t)sh_i << 16);
      }
    }
  }
  __hip_atomic_fetch_or(&B.free_mask, 1 << sl, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Loads the next band group into a free slot and makes it current (called by
// the wave that exhausted group `cur`), or marks the band phase finished.
__device__ __forceinline__ void band_load(const FastArgs &A, BandState &B, int cur, int lane, int nband) {
  const int seq = ((cur >> 4) + 1) & 0xFFFF;
  for (;;) {
    const int tk = __builtin_amdgcn_readfirstlane(atomicAdd(&A.ctl->band_queue, lane == 0 ? 1 : 0));
    if (tk >= nband) break;
    const GroupDesc gd = A.groups[tk];
    const bool ok = gd.cls == kNumClasses - 1 && gd.base >= 0 && gd.mpad > 0 &&
                    gd.mpad <= kMaxSeqLen && (int64_t)gd.mpad * 2 <= A.band_words &&
                    gd.nss >= 2 && gd.nss <= kMaxStrips;
    if (!ok) {
      atomicOr(&A.ctl->error, 4);
      continue;
    }
    const int mask = lds_load_acq(&B.free_mask);
    if (mask == 0) {  // cannot happen: at most kBandSlots - 1 other groups are unfinished
      atomicOr(&A.ctl->error, 8);
      break;
    }
    const int sl = __builtin_ctz((unsigned)mask);
    __hip_atomic_fetch_and(&B.free_mask, ~(1 << sl), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    BandSlot &S = B.slot[sl];
    S.gd = gd;
    S.done = 0;
    for (int i = lane; i < kMaxStrips; i += 64) S.prog[i] = 0;
    S.best[0] = 0ull;
    S.best[1] = 0ull;
    __builtin_amdgcn_wave_barrier();
    // a free slot's ticket is not updated by other waves (they only take
    // strips of the current group), nor is cur (only its loader moves it)
    lds_cas_lane0(&S.ticket, lds_load_acq(&S.ticket), (seq << 16) | (gd.nss << 8), lane);
    lds_cas_lane0(&B.cur, cur, (seq << 4) | sl, lane);
    return;
  }
  lds_cas_lane0(&B.cur, cur, -1, lane);
}

// The band phase of a block: its waves take strips of the current band group
// until no band group is left.
template <int KP>
__device__ __forceinline__ void band_phase(const FastArgs &A, const uint32_t *__restrict__ tab, BandState &B,
                           int lane, int nband, const SimdShare &share) {
  for (;;) {
    const int cur = lds_load_acq(&B.cur);
    if (cur < 0) return;
    const int sl = cur & 15, seq = (cur >> 4) & 0xFFFF;
    BandSlot &S = B.slot[sl];
    const int t = lds_load_acq(&S.ticket);
    if (((t >> 16) & 0xFFFF) != seq) {  // the slot is being reloaded: re-read cur
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    const int nss = (t >> 8) & 0xFF, nx = t & 0xFF;
    if (nx <= nss) {
      if (lds_cas_lane0(&S.ticket, t, t + 1, lane) != t) continue;  // another wave took it
      if (nx < nss) {
        // The strip ranks against the other waves on its SIMD by the work
        // left on its path (the longest path issues first, the others fill
        // its gaps), as every group of the main queue does.
        // Row-tag keys where no H of the strip can pass 4,095 (its rows end
        // at (nx + 1) * 64 * kRows, and H <= smax * rows): the first strips of
        // the small alphabet's band groups (DNA match scores are small)
        bool tag = false;
        if constexpr (KP == 18) tag = (int64_t)A.smax * (nx + 1) * 64 * kRows <= kTagKeyMaxH;
        if (tag)
          band_strip<KP, true, true>(A, tab, B, sl, nx, lane, share);
        else if (__builtin_amdgcn_readfirstlane(S.gd.mpad) <= kPairKeyMaxM)
          band_strip<KP, true, false>(A, tab, B, sl, nx, lane, share);
        else
          band_strip<KP, false, false>(A, tab, B, sl, nx, lane, share);
        share_simd(share, -1, lane);
      } else {
        band_load(A, B, cur, lane, nband);
      }
      continue;
    }
    while (lds_load_acq(&B.cur) == cur) __builtin_amdgcn_s_sleep(2);
  }
}

// Next group for the calling wave, or -1 when the queue is empty.  Every
// lane executes the same atomic (lane 0 adds 1, the others 0), so no
// lane-conditional branch surrounds it; the wave takes lane 0's ticket.  (A
// fetch written as `if (lane == 0) atomicAdd` followed by readfirstlane is
// unsafe here: for the 64-lane pipeline that branch has the same predicate as
// the first lane's result write, the compiler may thread one into the other
// across the loop back-edge, and lanes 1..63 then loop without lane 0.)
__device__ __forceinline__ int next_group(int32_t *queue, int first, int ng, int lane) {
  const int t = first + __builtin_amdgcn_readfirstlane(atomicAdd(queue, lane == 0 ? 1 : 0));
  return t < ng ? t : -1;
}

// One persistent kernel for the whole batch: waves take groups from a queue
// ordered longest first over all pipeline depths, so deep and shallow
// pipelines run side by side and the tail is short.
template <int KP, bool STOP>
__global__ __launch_bounds__(kFastThreads) void k_fast(FastArgs A) {
  constexpr int WPB = kFastThreads / 64;
  __shared__ uint32_t tab[KP * KP * KP];
  __shared__ BandState band;
  __shared__ int left[WPB + 1];
  __shared__ int simd_of[WPB];
  const int lane = threadIdx.x & 63;
  const int ng = min(A.ctl->ngroups, A.max_groups);
  if (ng <= 0) return;  // uniform over the grid
  const int nband = min(A.ctl->nband, ng);
  if (threadIdx.x == 0) {
    for (int i = 0; i < kBandSlots; ++i) band.slot[i].ticket = 0;
    band.slot[kBandSlots].ticket = 0;  // seq 0, no strips: the first wave loads a group
    band.cur = kBandSlots;
    band.free_mask = (1 << kBandSlots) - 1;
  }
  if (threadIdx.x <= WPB) left[threadIdx.x] = -1;
  if (lane == 0)  // SIMD of the wave: HW_ID bits 5:4
    simd_of[threadIdx.x >> 6] = (__builtin_amdgcn_s_getreg((31 << 11) | 4) >> 4) & 3;
  // row-pair table: word (a * KP + b1) * KP + b2 = S[a][b1] | S[a][b2] << 16
  for (int i = threadIdx.x; i < KP * KP * KP; i += blockDim.x) {
    const int a = i / (KP * KP), b1 = (i / KP) % KP, b2 = i % KP;
    tab[i] = (uint32_t)(uint16_t)A.table[a * KP + b1] |
             ((uint32_t)(uint16_t)A.table[a * KP + b2] << 16);
  }
  __syncthreads();
  const int wslot = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int wave = (int)blockIdx.x * WPB + wslot;
  uint32_t *wscr = A.scratch + (int64_t)wave * A.wave_words;
  SimdShare share{left, wslot, WPB, WPB};
  for (int i = 0; i < WPB; ++i)
    if (i != wslot && simd_of[i] == simd_of[wslot]) {
      if (share.p1 == WPB) share.p1 = i;
      else if (share.p2 == WPB) share.p2 = i;
    }
  share.p1 = __builtin_amdgcn_readfirstlane(share.p1);
  share.p2 = __builtin_amdgcn_readfirstlane(share.p2);
  // Band groups first (they are the longest), one block each, then the head
  // groups (the longest of the rest), then the rest.  When there are band
  // groups, the first wave of each SIMD takes one head group before the band
  // phase instead: behind it the head groups would end the launch, and this
  // way they start spread one per SIMD.
  const int nhead = nband > 0 ? min(A.ctl->nhead, ng - nband) : 0;
  bool head_left = nhead > 0;
  int gi = -1;
  if (head_left && wslot < share.p1 && wslot < share.p2)
    gi = next_group(&A.ctl->head_queue, nband, nband + nhead, lane);
  if (gi < 0 && nband > 0) band_phase<KP>(A, tab, band, lane, nband, share);
  for (;; gi = -1) {
    if (gi < 0 && head_left) {
      gi = next_group(&A.ctl->head_queue, nband, nband + nhead, lane);
      head_left = gi >= 0;
    }
    if (gi < 0) gi = next_group(&A.ctl->queue, nband + nhead, ng, lane);
    if (gi < 0) break;
    const GroupDesc gd = A.groups[gi];
    // a descriptor outside what the bucketing produces (class, boundary row
    // and column table within the wave's scratch, rows <= kMaxSeqLen) is
    // reported and skipped
    const int cls = gd.cls;
    const int lg = cls < kNumClasses - 1 ? cls : 6;
    const bool ok = cls >= 0 && cls < kNumClasses && gd.base >= 0 && gd.mpad >= 0 &&
                    gd.mpad <= kMaxSeqLen && wave_scratch_words(gd.mpad, lg) <= A.wave_words &&
                    gd.nss >= 0 && ((int64_t)(gd.nss - 1) << lg) * kRows < kMaxSeqLen &&
                    gd.last_rows >= 1 && gd.last_rows <= kRows;
    if (!ok) {
      atomicOr(&A.ctl->error, 4);  // every lane: no lane-conditional branch here either
    } else if (cls == 0) {
      run_group<kModeSolo, KP, true, STOP>(A, tab, gd, wscr, lane, 0, share);
    } else if (cls == kNumClasses - 1) {
      if (gd.mpad <= kPairKeyMaxM)
        run_group<kModeWave, KP, true, STOP>(A, tab, gd, wscr, lane, 6, share);
      else
        run_group<kModeWave, KP, false, STOP>(A, tab, gd, wscr, lane, 6, share);
    } else {
      run_group<kModeRow, KP, true, STOP>(A, tab, gd, wscr, lane, lg, share);
    }
  }
  if (lane == 0)
    __hip_atomic_store(&left[wslot], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __builtin_amdgcn_s_setprio(0);
}

template __global__ void k_fast<18, false>(FastArgs);
template __global__ void k_fast<26, false>(FastArgs);
template __global__ void k_fast<18, true>(FastArgs);
template __global__ void k_fast<26, true>(FastArgs);

// ---------------------------------------------------------------------------
// Exact int32 kernel: the vertex loop of codelet_batch_affine.cpp:62-87 with
// the reference's own (unclamped) recurrence, initial values and bounds, one
// wave per comparison.  The wave is one 64-lane pipeline over the B rows, as
// the packed kernel's whole-wave pipeline: lane k owns kW32Rows rows of each
// super-strip of 64 * kW32Rows rows and at step t computes column j = t - k
// of them, receiving H and F of the row above (and the column's A code) from
// lane k - 1 through DPP; lane 0 reads the row above the super-strip from the
// wave's work area, where lane 63 wrote it during the previous super-strip.
// Per row i (B) and column j (A), with H(i, -1) = 0, E(i, -1) = gapInit,
// H(-1, j) = F(-1, j) = 0:
//   E(i,j) = max(H(i,j-1) + gI + gE, E(i,j-1) + gE)
//   F(i,j) = max(H(i-1,j) + gI + gE, F(i-1,j) + gE)
//   H(i,j) = max(H(i-1,j-1) + S[a_j][b_i], E(i,j), F(i,j), 0)
// and the end is the first (i, then j) cell holding the maximum, by a strict >.
// Re-runs int16-saturated comparisons and serves gap settings outside the
// packed kernel's domain (and the int32 precision policy).
// ---------------------------------------------------------------------------
constexpr int kW32Rows = 16;
static_assert(kW32Rows <= 16, "k_wave32 row keys hold the row in 4 bits");
// Waves of a block: they share one comparison at a time, super-strip s on
// wave s % kW32Waves (three waves fill a SIMD's issue slots where one alone
// waits on its own dependency chains).
constexpr int kW32Waves = 3;
// Sleeps after which a wait for the boundary row is taken as lost (about a
// second; every wait ends long before in a working kernel): the kernel then
// reports error bit 16 and stops waiting.
constexpr int kW32SpinLimit = 1 << 24;

__global__ __launch_bounds__(64 * kW32Waves) void k_wave32(GenericArgs G) {
  constexpr int R = kW32Rows;
  constexpr int NW = kW32Waves;
  __shared__ int32_t tab[26 * 26];
  // per wave: (super-strip << 16) + boundary columns its current super-strip
  // has written (monotone over a comparison)
  __shared__ int prog[NW];
  __shared__ int res[NW][3];  // per wave: its best (H, row, column) of the comparison
  __shared__ int lost;        // a wait timed out: stop waiting
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int kp = G.kp;
  const int total = G.list ? min(*G.count, G.n) : G.n;
  if ((int)blockIdx.x >= total) return;  // uniform over the block: no comparison for it
  for (int i = threadIdx.x; i < kp * kp; i += blockDim.x) tab[i] = G.table[i];
  if (threadIdx.x == 0) lost = 0;
  __syncthreads();
  // boundary row of the block's comparison: (H, F) per column, updated in
  // place (super-strip s reads column c a block of 64 columns before its step
  // c and overwrites it at step c + 63)
  int32_t *const bnd = G.work + (int64_t)blockIdx.x * G.work_stride;
  const int gI = G.gap_init, gE = G.gap_extend, gIE = gI + gE;
  const bool pen = gIE <= 0 && gE <= 0;
  const bool first = lane == 0, last = lane == 63;
  // Waits until the previous super-strip's wave has published `need`; a lost
  // wait (a kernel fault, never a working schedule) is reported, not hung on.
  auto wait_prog = [&](int *pr, int need) {
    int v = lds_load_acq(pr);
    for (int spins = 0; v < need; ++spins) {
      if (spins >= kW32SpinLimit || lds_load_acq(&lost)) {
        // every lane: lane 0 carries the error bit (no lane-conditional branch)
        atomicOr(G.error, lane == 0 ? 16 : 0);
        __hip_atomic_store(&lost, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
      }
      __builtin_amdgcn_s_sleep(2);
      v = lds_load_acq(pr);
    }
  };
  for (int idx = blockIdx.x; idx < total; idx += gridDim.x) {
    const int p = G.list ? G.list[idx] : idx;
    const int m = G.a_len[p], n = G.b_len[p];
    if (m < 0 || n < 0 || m > G.max_a || m > kMaxSeqLen || n > kMaxSeqLen ||
        (int64_t)2 * m > G.work_stride) {
      if (threadIdx.x == 0) {
        atomicOr(G.error, 1);
        G.scores[p] = -1;
        G.a_range[p] = 0;
        G.b_range[p] = 0;
      }
      continue;
    }
    if (m == 0 || n == 0) {
      if (threadIdx.x == 0) {
        G.scores[p] = 0;
        G.a_range[p] = 0;
        G.b_range[p] = 0;
      }
      continue;
    }
    const uint8_t *a = G.codes + G.a_off[p];
    const uint8_t *b = G.codes + G.b_off[p];
    int best = 0, bi = 0, bj = 0;
    const int nss = (n + 64 * R - 1) / (64 * R);
    if (threadIdx.x < NW) prog[threadIdx.x] = -1;
    __syncthreads();
    // Super-strip ss runs on wave ss % NW after the wave's earlier ones; it
    // waits only for super-strip ss - 1 (on another wave), which waits only
    // for earlier ones: every wait ends.  Its publications grow over the
    // comparison, so a later super-strip's value also satisfies a wait.
    for (int ss = wv; ss < nss; ss += NW) {
      int *const prev = &prog[(ss + NW - 1) % NW];
      int *const mine = &prog[wv];
      const int pbase = (ss - 1) << 16;
      const int i0 = (ss * 64 + lane) * R;
      const bool load_bnd = ss > 0, store_bnd = ss + 1 < nss;
      // Gaps that cost (gI + gE <= 0, gE <= 0) bound every H by |A| * 127 <
      // 2^23: each row's H then becomes a key H * 16 + (15 - r), and a max over
      // the rows gives the step's (max H, first row).  Other gap settings keep
      // a maximum and its column per row.
      auto run = [&](auto pen_c) {
        constexpr bool PEN = decltype(pen_c)::value;
        int boff[R], H[R], E[R], rmax[PEN ? 1 : R], rcol[PEN ? 1 : R], toff[PEN ? R : 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          boff[r] = i0 + r < n ? min((int)b[i0 + r], kp - 1) : 0;
          H[r] = 0;  // H(i, -1)
          if constexpr (PEN) {
            // E of the row at the lane's next column: E(i, 0) from H(i, -1) = 0
            // and E(i, -1) = gI; a row past |B| keys below every real row
            E[r] = max(gIE, gI + gE);
            toff[r] = i0 + r < n ? 15 - r : -(1 << 30);
          } else {
            E[r] = gI;  // E(i, -1)
            // a row past |B| never takes a maximum (nothing exceeds INT_MAX):
            // the step needs no row check, the merge below skips such rows
            rmax[r] = i0 + r < n ? 0 : INT32_MAX;
            rcol[r] = 0;
          }
        }
        int bkey = INT32_MIN, bcol = 0;  // the lane's best key in this super-strip, its column
        int out_h = 0, out_f = 0, col = 0;
        int hdiag = 0;  // H(i0 - 1, j - 1)
        // Lane 0's inputs at step t (column t: its A code and the boundary row's
        // H and F) come from blocks of 64 columns that every lane loads one
        // column of, a block ahead; the wave reads lane t % 64's values with
        // v_readlane (a uniform lane index: no LDS round trip).  Lane 63
        // overwrites column c of the boundary at step c + 63, after the load of
        // c's block (at step c - 64 or earlier).
        int a_cur, h_cur = 0, f_cur = 0, a_nxt, h_nxt = 0, f_nxt = 0;
        auto load_block = [&](int c0, int &av, int &hv, int &fv) {
          const int c = c0 + lane;
          av = c < m ? min((int)a[c], kp - 1) * kp : 0;
          if (load_bnd) {
            // the 64 columns from c0 written by the previous super-strip
            if (c0 < m) wait_prog(prev, pbase + min(m, c0 + 64));
            if (c < m) {
              hv = bnd[2 * c];
              fv = bnd[2 * c + 1];
            }
          }
        };
        load_block(0, a_cur, h_cur, f_cur);
        load_block(64, a_nxt, h_nxt, f_nxt);
        const int T = m + 63;
        for (int t = 0; t < T; ++t) {
          const int j = t - lane;
          if (t > 0 && (t & 63) == 0) {
            a_cur = a_nxt;
            h_cur = h_nxt;
            f_cur = f_nxt;
            load_block(t + 64, a_nxt, h_nxt, f_nxt);
          }
          const int sl = t & 63;
          const int ac = __builtin_amdgcn_readlane(a_cur, sl);
          const int bh = load_bnd ? __builtin_amdgcn_readlane(h_cur, sl) : 0;
          const int bf = load_bnd ? __builtin_amdgcn_readlane(f_cur, sl) : 0;
          const int up_h = (int)from_prev<kModeWave>((uint32_t)bh, (uint32_t)out_h, first);
          const int up_f = (int)from_prev<kModeWave>((uint32_t)bf, (uint32_t)out_f, first);
          col = (int)from_prev<kModeWave>((uint32_t)ac, (uint32_t)col, first);
          if (j >= 0 && j < m) {
            int hd = hdiag, hup = up_h, fup = up_f;
            if constexpr (PEN) {
              // H(i - 1, j) + gI + gE carried down the rows with F
              int hoo = up_h + gIE, kev = 0, tk = INT32_MIN;
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const int sc = tab[col + boff[r]];
                const int e = E[r];  // E(i, j)
                const int f = max(hoo, fup + gE);
                const int h = max(max(hd + sc, max(e, f)), 0);
                hd = H[r];
                H[r] = h;
                const int ho = h + gIE;
                E[r] = max(ho, e + gE);  // E(i, j + 1)
                hoo = ho;
                fup = f;
                const int kr = (h << 4) + toff[r];
                if (r & 1) tk = max(tk, max(kev, kr));
                else kev = kr;
              }
              hup = H[R - 1];
              if (tk > bkey) {  // strict: the first column of the lane's best key
                bkey = tk;
                bcol = j;
              }
            } else {
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const int sc = tab[col + boff[r]];
                const int e = max(H[r] + gIE, E[r] + gE);
                const int f = max(hup + gIE, fup + gE);
                const int h = max(max(hd + sc, e), max(f, 0));
                hd = H[r];
                H[r] = h;
                E[r] = e;
                hup = h;
                fup = f;
                if (h > rmax[r]) {
                  rmax[r] = h;
                  rcol[r] = j;
                }
              }
            }
            out_h = hup;
            out_f = fup;
            if (last && store_bnd) {
              bnd[2 * j] = hup;
              bnd[2 * j + 1] = fup;
            }
            hdiag = up_h;
          }
          // publish the boundary columns written so far (lane 63 wrote
          // column t - 63 this step)
          if (store_bnd && (t & 15) == 15) {
            __builtin_amdgcn_wave_barrier();
            lds_store_rel(mine, (ss << 16) + max(0, min(m, t - 62)));
          }
        }
        if (store_bnd) {
          __builtin_amdgcn_wave_barrier();
          lds_store_rel(mine, (ss << 16) + m);
        }
        // rows in order, strict >: the lane's first maximum in B order so far
        if constexpr (PEN) {
          if (bkey >= 0 && (bkey >> 4) > best) {
            best = bkey >> 4;
            bi = i0 + 15 - (bkey & 15);
            bj = bcol;
          }
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (i0 + r < n && rmax[r] > best) {
              best = rmax[r];
              bi = i0 + r;
              bj = rcol[r];
            }
        }
      };
      if (pen) run(std::true_type{});
      else run(std::false_type{});
    }
    // lanes and waves hold disjoint rows: (max H, min row)
    for (int off = 1; off < 64; off <<= 1) {
      const int oh = __shfl_xor(best, off), oi = __shfl_xor(bi, off), oj = __shfl_xor(bj, off);
      if (oh > best || (oh == best && oi < bi)) {
        best = oh;
        bi = oi;
        bj = oj;
      }
    }
    if (first) {
      res[wv][0] = best;
      res[wv][1] = bi;
      res[wv][2] = bj;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < NW; ++w)
        if (res[w][0] > best || (res[w][0] == best && res[w][1] < bi)) {
          best = res[w][0];
          bi = res[w][1];
          bj = res[w][2];
        }
      G.scores[p] = best;
      G.a_range[p] = (int32_t)((uint32_t)(bj & 0xFFFF) << 16);
      G.b_range[p] = (int32_t)((uint32_t)(bi & 0xFFFF) << 16);
    }
    __syncthreads();  // prog and res serve the next comparison
  }
}

// ---------------------------------------------------------------------------
// Small utility kernels.
// ---------------------------------------------------------------------------
// The bucketing's buffers cleared in one launch: up to four (pointer, count,
// value) ranges over one grid-stride index.
struct FillRanges {
  int32_t *p[4];
  int64_t n[4];
  int32_t v[4];
};
__global__ void k_fill_ranges(FillRanges f) {
  const int64_t total = f.n[0] + f.n[1] + f.n[2] + f.n[3];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = i;
    int s = 0;
    while (s < 3 && j >= f.n[s]) j -= f.n[s++];
    f.p[s][j] = f.v[s];
  }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_encode(const uint8_t *chars, int64_t len, uint8_t *codes, const uint8_t *table) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * blockDim.x)
    codes[i] = table[chars[i]];
}

// ---------------------------------------------------------------------------
// Device-side length bucketing (replaces partition::fillFirst / roundRobin /
// greedy, src/partition.cpp:197-462).
//
// Class: lanes per comparison L = max(L_scratch, min(L_share, L_rows)) with
//   L_scratch = ceil(|A| * 64 / wave_cols)   (boundary of one wave <= its scratch:
//                                    512 KiB, ceil(|A| / 1024), in the default layout),
//   L_share   = ceil(|A||B| / share) (cells per pipeline near the fair share
//                                    of the launch, so big pairs get more lanes),
//   L_rows    = ceil(|B| / kRows)  (no more lanes than row strips),
// each rounded up to a power of two in {1, 2, 4, 8, 16, 64}.
// Bin: (class, super-strips, |A| bucket); pairs scatter into a class-aligned
// sorted array; groups of 128/L consecutive slots share one padded shape.
// ---------------------------------------------------------------------------


__device__ __forceinline__ int pow2_class(int64_t x) {
  // smallest class whose L >= x
  if (x <= 1) return 0;
  if (x <= 2) return 1;
  if (x <= 4) return 2;
  if (x <= 8) return 3;
  if (x <= 16) return 4;
  return 5;
}

// Super-strip bin: exact below 32, then four bins per octave (a group pads
// to its longest B, so neighbours in a bin differ by at most 19%).
__device__ __forceinline__ int nss_bin_of(int nss) {
  if (nss < 32) return nss;
  return min(kNssBins - 1, 32 + (int)(4.0f * __log2f((float)nss * (1.0f / 32.0f))));
}

// |A| bin inside a run: 16 columns wide up to 2048, 256 wide up to 12288,
// then one bin (a group pads to its longest A).
__device__ __forceinline__ int mbin_of(int m) {
  return m <= 2048 ? (m + 15) / 16 : min(kMBinsPerRows - 1, 128 + (m - 2048 + 255) / 256);
}
// |A| bin of a whole-wave pair of one super-strip (five row counts share the
// run's bins): 16 columns wide up to 1024, 64 wide up to 3328, then one bin.
__device__ __forceinline__ int mbin_wave_of(int m) {
  return m <= 1024 ? (m + 15) / 16 : min(kMBinsPerRowsWave - 1, 64 + (m - 1024 + 63) / 64);
}

// Device-scope atomics on one address serialize across the XCDs, so the
// batch-wide sums below are reduced per block first: one atomic per block.
// Batch totals (cells, longest A and B): kSumBlocks blocks at most.
constexpr int kSumThreads = 1024;
constexpr int kSumBlocks = 64;
__global__ __launch_bounds__(kSumThreads) void k_sum_cells(int32_t n, const int32_t *a_len,
                                                           const int32_t *b_len, BucketCtl *ctl) {
  __shared__ unsigned long long s_c[kSumThreads / 64];
  __shared__ int s_m[kSumThreads / 64], s_n[kSumThreads / 64];
  const int w = threadIdx.x >> 6;
  unsigned long long c = 0;
  int m = 0, nb = 0;
  // grid-stride: a few blocks, so that few device atomics hit the three
  // counters (atomics on one address serialize across the chip)
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int pm = a_len[p], pn = b_len[p];
    if (pm > 0 && pn > 0 && pm <= kMaxSeqLen && pn <= kMaxSeqLen) {
      c += (unsigned long long)pm * pn;
      m = max(m, pm);
      nb = max(nb, pn);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_xor(c, off);
    m = max(m, __shfl_xor(m, off));
    nb = max(nb, __shfl_xor(nb, off));
  }
  if ((threadIdx.x & 63) == 0) {
    s_c[w] = c;
    s_m[w] = m;
    s_n[w] = nb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      c += s_c[i];
      m = max(m, s_m[i]);
      nb = max(nb, s_n[i]);
    }
    if (c) atomicAdd(&ctl->total_cells, c);
    atomicMax(&ctl->maxm, m);
    atomicMax(&ctl->maxn, nb);
  }
}

// Per-pair bin and rank.  Device-scope atomics run at about 2 G/s over the
// whole chip, so each block first counts its pairs per bin in an LDS hash
// table (the pair's rank inside the block comes back from the LDS atomic),
// then adds one count per distinct bin to the global histogram; the value that
// atomic returns is the block's first slot in the bin.  rank[p] is then the
// pair's slot inside its bin, and k_scatter places pairs without atomics.
constexpr int kClsThreads = 1024;
constexpr int kClsHash = 2048;  // at most kClsThreads keys: load factor <= 1/2
__global__ __launch_bounds__(kClsThreads) void k_classify(int32_t n, const int32_t *a_len,
                                                          const int32_t *b_len,
                                                          int64_t target_slots,
                                                          int32_t force_class, int32_t wave_cols,
                                                          int32_t max_a, int32_t *bin_of,
                                                          int32_t *rank, int32_t *hist,
                                                          BucketCtl *ctl, int32_t *scores,
                                                          int32_t *a_range, int32_t *b_range,
                                                          const int32_t *target) {
  __shared__ int s_count[kNumClasses];
  __shared__ int32_t s_key[kClsHash];
  __shared__ int32_t s_cnt[kClsHash];
  const int tid = threadIdx.x;
  if (tid < kNumClasses) s_count[tid] = 0;
  for (int i = tid; i < kClsHash; i += blockDim.x) {
    s_key[i] = -1;
    s_cnt[i] = 0;
  }
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + tid;
  if (p == 0) {
    const int64_t cap = max((int64_t)1, (int64_t)(ctl->total_cells /
                                                  (unsigned long long)(kRows * target_slots)));
    ctl->head_cost = (int32_t)min(cap / kHeadCapDiv, (int64_t)0x7FFFFFFF);
  }
  int bin = -1;
  if (p < n) {
    const int m = a_len[p], nb = b_len[p];
    if (m < 0 || nb < 0 || m > max_a || m > kMaxSeqLen || nb > kMaxSeqLen) {
      atomicOr(&ctl->error, 1);
      bin_of[p] = -1;
      scores[p] = -1;
      a_range[p] = 0;
      b_range[p] = 0;
    } else if (m == 0 || nb == 0) {
      bin_of[p] = -1;
      scores[p] = 0;
      a_range[p] = 0;
      b_range[p] = 0;
    } else {
      int cls;
      if (force_class >= 0) {
        // never below the class the wave scratch needs (mpad * 64 / L <= wave_cols)
        cls = max(force_class, pow2_class((m * 64 + wave_cols - 1) / wave_cols));
      } else {
        // The shallowest pipeline whose group takes at most `cap` steps, the
        // batch's cells spread evenly over the resident half-lanes
        // (cells / (kRows * slots)); the wave scratch sets a floor
        // (mpad * 64/L <= wave_cols).  Groups of similar length then keep the
        // longest-first queue's tail short.
        //
        // Two refinements: a comparison that spans more than one row of 16
        // never runs alone in a lane (the one-lane pipeline reads and writes
        // the boundary row on every lane and step; two lanes halve that),
        // and one that a band group could take (|B| > 1024) must fit a third
        // of the cap on one wave: otherwise it joins the band groups, which
        // spread it over the waves of a block, instead of running long after
        // the band groups as the launch's tail.
        const int64_t cap = max((int64_t)1, (int64_t)(ctl->total_cells /
                                                      (unsigned long long)(kRows * target_slots)));
        const int64_t capc = nb > 64 * kRows ? max((int64_t)1, cap / 3) : cap;
        cls = max(nb > kRows ? 1 : 0, pow2_class((m * 64 + wave_cols - 1) / wave_cols));
        while (cls < kNumClasses - 1) {
          const int Lc = kClassL[cls];
          if ((int64_t)((nb + Lc * kRows - 1) / (Lc * kRows)) * (m + Lc - 1) <= capc) break;
          ++cls;
        }
      }
      const int L = kClassL[cls];
      // start pass: bin by the rows the pair is expected to need (the class,
      // and the group's super-strips, still follow |B|)
      const int nb_key = target ? min(nb, 2 * max(target[p], 0) + 16) : nb;
      const int nss = (nb_key + L * kRows - 1) / (L * kRows);
      // rows per lane the last super-strip of |B| needs (a group runs it with
      // 8, 12 or 16, or a whole-wave group of one super-strip with an even
      // count from 8, by its longest B): pairs that need fewer share groups
      const int nss_b = (nb + L * kRows - 1) / (L * kRows);
      const int last = (nb - (nss_b - 1) * L * kRows + L - 1) / L;
      const bool wave1 = cls == kNumClasses - 1 && nss_b == 1;
      const int rlc = rows_class(last, wave1);
      bin = (cls * kNssBins + nss_bin_of(nss)) * kMBins +
            (wave1 ? rlc * kMBinsPerRowsWave + mbin_wave_of(m) : rlc * kMBinsPerRows + mbin_of(m));
      bin_of[p] = bin;
      atomicAdd(&s_count[cls], 1);
    }
  }
  int slot = 0, local = 0;
  if (bin >= 0) {
    slot = (int)(((uint32_t)bin * 2654435761u) >> 21);  // 11 bits: kClsHash
    for (;;) {
      const int old = atomicCAS(&s_key[slot], -1, bin);
      if (old == -1 || old == bin) break;
      slot = (slot + 1) & (kClsHash - 1);
    }
    local = atomicAdd(&s_cnt[slot], 1);
  }
  __syncthreads();
  for (int i = tid; i < kClsHash; i += blockDim.x) {
    const int k = s_key[i];
    if (k >= 0) s_cnt[i] = atomicAdd(&hist[k], s_cnt[i]);  // count -> the block's base
  }
  __syncthreads();
  if (bin >= 0) rank[p] = s_cnt[slot] + local;
  if (tid < kNumClasses && s_count[tid]) atomicAdd(&ctl->class_count[tid], s_count[tid]);
}

// Scan of the bins in two kernels.  k_scan_run: one wave per run (class,
// super-strip bin) of kMBins consecutive bins, eight bins per lane, turns each
// bin's count into its offset inside the run and leaves the run's total in
// run_start.  k_scan_runs: one thread per run gives every run its start (a
// multiple of kRunAlign slots) and its first group.  k_scatter adds the two.
static_assert(kMBins == 512 && kRuns % 64 == 0 && kRuns <= 1024, "scan layout");
constexpr int kScanRunBlocks = 16;  // x 4 waves: each wave scans every 64th run
__global__ __launch_bounds__(256) void k_scan_run(int32_t *hist, BucketCtl *ctl) {
  const int lane = threadIdx.x & 63;
  const int nw = (int)gridDim.x * (int)(blockDim.x >> 6);
  for (int r = (int)blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6); r < kRuns;
       r += nw) {
    int4 *h4 = reinterpret_cast<int4 *>(hist + r * kMBins) + 2 * lane;
    const int4 c0 = h4[0], c1 = h4[1];
    const int t0 = c0.x + c0.y + c0.z + c0.w;
    const int tot = t0 + c1.x + c1.y + c1.z + c1.w;
    if (__ballot(tot != 0) == 0) {  // an empty run: its offsets are already 0
      if (lane == 63) ctl->run_start[r] = 0;
      continue;
    }
    int inc = tot;
    for (int off = 1; off < 64; off <<= 1) {
      const int u = __shfl_up(inc, off);
      if (lane >= off) inc += u;
    }
    const int base = inc - tot;
    int4 o0, o1;
    o0.x = base;
    o0.y = o0.x + c0.x;
    o0.z = o0.y + c0.y;
    o0.w = o0.z + c0.z;
    o1.x = base + t0;
    o1.y = o1.x + c1.x;
    o1.z = o1.y + c1.y;
    o1.w = o1.z + c1.z;
    h4[0] = o0;
    h4[1] = o1;
    if (lane == 63) ctl->run_start[r] = inc;
  }
}

__global__ __launch_bounds__(kRuns) void k_scan_runs(BucketCtl *ctl) {
  __shared__ int32_t wsum[2][kRuns / 64];
  const int tid = threadIdx.x;
  // the run's aligned size and group count, scanned over the runs (wave
  // scans, then the wave totals)
  const int cnt = ctl->run_start[tid];
  const int gs = 128 / kClassL[tid / kNssBins];
  const int ng = (cnt + gs - 1) / gs;
  const int al = (cnt + kRunAlign - 1) / kRunAlign * kRunAlign;
  int inc_al = al, inc_ng = ng;
  for (int off = 1; off < 64; off <<= 1) {
    const int ua = __shfl_up(inc_al, off), un = __shfl_up(inc_ng, off);
    if ((tid & 63) >= off) {
      inc_al += ua;
      inc_ng += un;
    }
  }
  if ((tid & 63) == 63) {
    wsum[0][tid >> 6] = inc_al;
    wsum[1][tid >> 6] = inc_ng;
  }
  __syncthreads();
  int pa = 0, pn = 0;
  for (int w = 0; w < (tid >> 6); ++w) {
    pa += wsum[0][w];
    pn += wsum[1][w];
  }
  const int start = pa + inc_al - al, gbase = pn + inc_ng - ng;
  ctl->run_start[tid] = start;
  ctl->run_group_base[tid] = gbase;
  if (tid == kRuns - 1) {
    ctl->run_group_base[kRuns] = gbase + ng;
    ctl->ngroups = gbase + ng;
  }
}

__global__ void k_scatter(int32_t n, const int32_t *bin_of, const int32_t *rank,
                          const int32_t *offsets, const BucketCtl *ctl, int32_t *sorted) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int b = bin_of[p];
  if (b < 0) return;
  sorted[ctl->run_start[b / kMBins] + offsets[b] + rank[p]] = p;
}

__device__ __forceinline__ int cost_bin(int64_t cost) {
  // 8 bins per octave of the group's step count
  if (cost <= 1) return 0;
  const float l2 = __log2f((float)cost);
  return min(kCostBinsPerClass - 1, (int)(l2 * 8.0f));
}

// One wave per group, grid-stride over the groups: its run (binary search
// over the run group bases), its slots, and the padded shape over the pairs it
// holds.  The cost bins are counted per block in LDS (one global atomic per
// bin and block, whose result is the block's base in the bin), so
// k_order_groups places groups without atomics.
constexpr int kGroupThreads = 1024;
constexpr int kGroupBlocks = 128;
__global__ __launch_bounds__(kGroupThreads) void k_groups(int32_t max_groups, const int32_t *sorted,
                                                          const int32_t *a_len,
                                                          const int32_t *b_len, BucketCtl *ctl,
                                                          GroupDesc *gtmp, int32_t *gbin,
                                                          int32_t *grank, int32_t *costhist) {
  __shared__ int32_t s_cost[kCostBins];
  __shared__ unsigned long long s_pad[kGroupThreads / 64];
  __shared__ int32_t s_nband, s_nhead;
  static_assert(kCostBins >= 2 * kCostBinsPerClass, "band and other cost bins");
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) {
    s_nband = 0;
    s_nhead = 0;
  }
  // groups in this cost bin or above that are not band groups are head groups
  const int head_bin = cost_bin(ctl->head_cost);
  for (int i = tid; i < kCostBins; i += blockDim.x) s_cost[i] = 0;
  __syncthreads();
  const int ng = min(max_groups, ctl->ngroups);
  const int nw = (int)(blockDim.x >> 6);
  const int stride = (int)gridDim.x * nw;
  unsigned long long pad = 0;
  for (int gi = (int)blockIdx.x * nw + w; gi < ng; gi += stride) {
    int lo = 0, hi = kRuns - 1;  // last run whose first group is <= gi
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ctl->run_group_base[mid] <= gi) lo = mid; else hi = mid - 1;
    }
    const int c = lo / kNssBins;
    const int L = kClassL[c];
    const int gs = 128 / L;
    const int base = ctl->run_start[lo] + (gi - ctl->run_group_base[lo]) * gs;
    int mpad = 0, nbmax = 0;
    for (int s = lane; s < gs; s += 64) {
      const int p = sorted[base + s];
      if (p >= 0) {
        mpad = max(mpad, a_len[p]);
        nbmax = max(nbmax, b_len[p]);
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      mpad = max(mpad, __shfl_xor(mpad, off));
      nbmax = max(nbmax, __shfl_xor(nbmax, off));
    }
    if (lane == 0) {
      GroupDesc d;
      d.cls = c;
      d.mpad = mpad;
      d.nss = (nbmax + L * kRows - 1) / (L * kRows);
      d.base = base;
      // the last super-strip's rows per lane (its lanes cover the rest of B)
      d.last_rows = d.nss > 0 ? (nbmax - (d.nss - 1) * L * kRows + L - 1) / L : kRows;
      gtmp[gi] = d;
      const bool band_g = c == kNumClasses - 1 && d.nss >= 2;  // band strips run kRows rows
      const int rows_last =
          band_g || d.nss == 0 ? kRows : last_strip_rows(d.last_rows, c == kNumClasses - 1);
      pad += (unsigned long long)128 * (unsigned long long)mpad *
             (unsigned long long)(d.nss > 0 ? (d.nss - 1) * kRows + rows_last : 0);
      // longest groups first over the whole batch (a wave's time is about its
      // super-strips times its steps per super-strip)
      // band groups (whole-wave pipelines over two or more super-strips,
      // shared by a block) come before all others
      const bool band = c == kNumClasses - 1 && d.nss >= 2;
      const int64_t strips32 =
          d.nss > 0 ? (int64_t)(d.nss - 1) * 32 + step_weight32(rows_last) : 0;
      const int cbin = cost_bin(strips32 * (mpad + L - 1) / 32);
      const int cb = (band ? 0 : kCostBinsPerClass) + kCostBinsPerClass - 1 - cbin;
      gbin[gi] = cb;
      grank[gi] = atomicAdd(&s_cost[cb], 1);
      if (band) atomicAdd(&s_nband, 1);
      else if (cbin >= head_bin) atomicAdd(&s_nhead, 1);
    }
  }
  if (lane == 0) s_pad[w] = pad;
  __syncthreads();
  for (int i = tid; i < kCostBins; i += blockDim.x) {
    const int cnt = s_cost[i];
    if (cnt) s_cost[i] = atomicAdd(&costhist[i], cnt);  // count -> the block's base
  }
  if (tid == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < nw; ++i) t += s_pad[i];
    if (t) atomicAdd(&ctl->padded_cells, t);
    if (s_nband) atomicAdd(&ctl->nband, s_nband);
    if (s_nhead) atomicAdd(&ctl->nhead, s_nhead);
  }
  __syncthreads();
  for (int gi = (int)blockIdx.x * nw + w; gi < ng; gi += stride)
    if (lane == 0) grank[gi] += s_cost[gbin[gi]];
}

// Single block: exclusive scan of the cost bins.
__global__ __launch_bounds__(1024) void k_scan_cost(int32_t *costhist) {
  __shared__ int32_t part[1024];
  constexpr int PER = kCostBins / 1024;
  static_assert(kCostBins % 1024 == 0, "cost bins");
  const int tid = threadIdx.x;
  int sum = 0;
  for (int i = 0; i < PER; ++i) sum += costhist[tid * PER + i];
  part[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int i = 0; i < PER; ++i) {
    const int b = tid * PER + i;
    const int h = costhist[b];
    costhist[b] = run;
    run += h;
  }
}

__global__ void k_order_groups(int32_t max_groups, const BucketCtl *ctl, const GroupDesc *gtmp,
                               const int32_t *gbin, const int32_t *grank, const int32_t *coff,
                               GroupDesc *groups) {
  const int ng = min(max_groups, ctl->ngroups);
  for (int gi = blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += gridDim.x * blockDim.x)
    groups[coff[gbin[gi]] + grank[gi]] = gtmp[gi];
}

// ---------------------------------------------------------------------------
// Start positions: the reference's reverse pass (codelet_batch_affine.cpp:89-115)
// walks i = Bend..0, j = Aend..0 with the forward recurrence; that is the
// forward pass on the reversed prefixes A[0..Aend], B[0..Bend], whose first
// maximum (i', j') maps back to Bstart = Bend - i', Astart = Aend - j'.
// ---------------------------------------------------------------------------
// One wave per pair: the lanes copy the reversed prefixes byte by byte, the
// reads of a wave falling in one span of the sequence and the writes
// contiguous.
__global__ __launch_bounds__(256) void k_reverse_prefix(
    int32_t n, const uint8_t *codes, const int64_t *a_off, const int32_t *a_len,
    const int64_t *b_off, const int32_t *b_len, const int32_t *scores, const int32_t *a_range,
    const int32_t *b_range, uint8_t *rcodes, int32_t *ra_len, int32_t *rb_len, int32_t *error) {
  const int64_t p64 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (p64 >= n) return;  // uniform over the wave
  const int p = (int)p64;
  int la = 0, lb = 0;
  if (scores[p] > 0) {
    la = (int)(((uint32_t)a_range[p] >> 16) & 0xFFFF) + 1;
    lb = (int)(((uint32_t)b_range[p] >> 16) & 0xFFFF) + 1;
    // an end outside its sequence would write past the reversed copy
    if (la > a_len[p] || lb > b_len[p]) {
      if (lane == 0) atomicOr(error, 2);
      la = lb = 0;
    }
  }
  if (la > 0) {
    const uint8_t *a = codes + a_off[p];
    const uint8_t *b = codes + b_off[p];
    uint8_t *ra = rcodes + a_off[p];
    uint8_t *rb = rcodes + b_off[p];
    for (int t = lane; t < la; t += 64) ra[t] = a[la - 1 - t];
    for (int t = lane; t < lb; t += 64) rb[t] = b[lb - 1 - t];
  }
  if (lane == 0) {
    ra_len[p] = la;
    rb_len[p] = lb;
  }
}

__global__ void k_finish_starts(int32_t n, const int32_t *scores, const int32_t *ra_range,
                                const int32_t *rb_range, int32_t *a_range, int32_t *b_range) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (scores[p] <= 0) return;
  const uint32_t aend = ((uint32_t)a_range[p] >> 16) & 0xFFFF;
  const uint32_t bend = ((uint32_t)b_range[p] >> 16) & 0xFFFF;
  const uint32_t ra = ((uint32_t)ra_range[p] >> 16) & 0xFFFF;
  const uint32_t rb = ((uint32_t)rb_range[p] >> 16) & 0xFFFF;
  a_range[p] = (int32_t)((aend << 16) | ((aend - ra) & 0xFFFF));
  b_range[p] = (int32_t)((bend << 16) | ((bend - rb) & 0xFFFF));
}

// ---------------------------------------------------------------------------
// Plain padded batches (the baselines the bucketing is measured against):
// pairs in input order, one pair per lane unless the longest A of the batch
// needs a deeper pipeline for the wave scratch (mpad * 64/L <= wave_cols); group g holds
// pairs [g*gs, (g+1)*gs).  Mode 1 pads every group to the longest A and B of
// the whole batch, mode 2 pads each group to its own longest pair.
// ---------------------------------------------------------------------------
__global__ void k_plain_setup(int32_t n, const int32_t *bin_of, int32_t *sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sorted[i] = bin_of[i] >= 0 ? i : -1;
}

// One wave per group.
__global__ __launch_bounds__(256) void k_plain_groups(int32_t n, int32_t per_group,
                                                      int32_t wave_cols, const int32_t *sorted,
                                                      const int32_t *a_len,
                                                      const int32_t *b_len, BucketCtl *ctl,
                                                      GroupDesc *groups) {
  const int cls = pow2_class(((int64_t)ctl->maxm * 64 + wave_cols - 1) / wave_cols);
  const int L = kClassL[cls];
  const int gs = 128 / L;
  const int g = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int ng = (n + gs - 1) / gs;
  if (g == 0 && lane == 0) ctl->ngroups = ng;
  if (g >= ng) return;
  int mpad = ctl->maxm, nbmax = ctl->maxn;
  if (per_group) {
    mpad = 0;
    nbmax = 0;
    for (int s = lane; s < gs; s += 64) {
      const int i = g * gs + s;
      const int p = i < n ? sorted[i] : -1;
      if (p >= 0) {
        mpad = max(mpad, a_len[p]);
        nbmax = max(nbmax, b_len[p]);
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      mpad = max(mpad, __shfl_xor(mpad, off));
      nbmax = max(nbmax, __shfl_xor(nbmax, off));
    }
  }
  if (lane != 0) return;
  GroupDesc d;
  d.cls = cls;
  d.mpad = mpad;
  d.nss = (nbmax + L * kRows - 1) / (L * kRows);
  d.base = g * gs;
  d.last_rows = kRows;  // the padded baselines keep full super-strips
  groups[g] = d;
  atomicAdd(&ctl->padded_cells,
            (unsigned long long)128 * d.nss * kRows * (unsigned long long)d.mpad);
}

// ---------------------------------------------------------------------------
// Traceback: the CIGAR of each comparison between the start and end positions
// a compare with starts reported.  A[as..ae] and B[bs..be] are aligned
// globally with affine gaps, rows over A and columns over B; the optimum
// equals the reported local score (checked), and the traceback prefers the
// diagonal when it is at least both gap states, then the A-gap state when it
// is strictly larger, and opens a gap only when opening is strictly better
// than extending (the rules of SSW's banded_sw, without its band).  One
// comparison per lane, int64 arithmetic; a direction byte per cell (bits 0-1
// H source 0 diag / 1 A gap / 2 B gap, bit 2 A gap opened, bit 3 B gap opened).
// Runs (length << 4 | op, op 0 M, 1 I = A only, 2 D = B only) are written end
// to start.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_traceback(TraceArgs T) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= T.n) return;
  const int s = T.scores[p];
  const uint32_t ar = (uint32_t)T.a_range[p], br = (uint32_t)T.b_range[p];
  const int as = (int)(ar & 0xFFFF), ae = (int)(ar >> 16);
  const int bs = (int)(br & 0xFFFF), be = (int)(br >> 16);
  if (s <= 0 || ae < as || be < bs) {
    T.cig_len[p] = 0;
    return;
  }
  const int m = ae - as + 1, n = be - bs + 1;
  const uint8_t *a = T.codes + T.a_off[p] + as;
  const uint8_t *b = T.codes + T.b_off[p] + bs;
  uint8_t *dir = T.dir + T.dir_off[p];
  int64_t *hrow = T.work + T.work_off[p];
  int64_t *erow = hrow + (n + 1);
  const int64_t go = T.gap_o, ge = T.gap_x;
  const int64_t NEG = -((int64_t)1 << 50);
  hrow[0] = 0;
  erow[0] = NEG;
  for (int j = 1; j <= n; ++j) {
    hrow[j] = -(go + (int64_t)(j - 1) * ge);
    erow[j] = NEG;
  }
  for (int i = 1; i <= m; ++i) {
    int64_t hdiag = hrow[0];
    const int64_t h0 = -(go + (int64_t)(i - 1) * ge);
    hrow[0] = h0;
    int64_t hleft = h0, f = NEG;
    const int16_t *row = T.table + (int)a[i - 1] * T.kp;
    uint8_t *drow = dir + (int64_t)(i - 1) * n;
    for (int j = 1; j <= n; ++j) {
      const int64_t hup = hrow[j], eup = erow[j];
      const bool eo = hup - go > eup - ge;
      const int64_t e = eo ? hup - go : eup - ge;
      const bool fo = hleft - go > f - ge;
      f = fo ? hleft - go : f - ge;
      const int64_t d = hdiag + row[b[j - 1]];
      int64_t h = d;
      uint32_t src = 0;
      if (e > h || f > h) {
        src = e > f ? 1u : 2u;
        h = e > f ? e : f;
      }
      drow[j - 1] = (uint8_t)(src | (eo ? 4u : 0u) | (fo ? 8u : 0u));
      hdiag = hup;
      hrow[j] = h;
      erow[j] = e;
      hleft = h;
    }
  }
  if (hrow[n] != (int64_t)s) {
    T.cig_len[p] = -2;
    return;
  }
  uint32_t *out = T.cigar + T.cig_off[p];
  int runs = 0, run_op = -1, run_len = 0;
  int i = m, j = n, st = 0;
  while (i > 0 || j > 0) {
    int op;
    if (i == 0) {
      op = 2;
      --j;
    } else if (j == 0) {
      op = 1;
      --i;
    } else {
      const uint32_t dv = dir[(int64_t)(i - 1) * n + (j - 1)];
      if (st == 0) st = (int)(dv & 3u);
      if (st == 0) {
        op = 0;
        --i;
        --j;
      } else if (st == 1) {
        op = 1;
        st = (dv & 4u) ? 0 : 1;
        --i;
      } else {
        op = 2;
        st = (dv & 8u) ? 0 : 2;
        --j;
      }
    }
    if (op == run_op) {
      ++run_len;
    } else {
      if (run_len) out[runs++] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
      run_op = op;
      run_len = 1;
    }
  }
  if (run_len) out[runs++] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
  T.cig_len[p] = runs;
}

// The same traceback on one wave per comparison, in int32 (the host takes it
// when every value of the call's regions is bounded well inside int32: the
// boundary value kTrNeg and its differences with real values then stay in
// range, and every comparison decides as the int64 kernel's).  The region's
// rows (A) run in strips of 64 * R (R = trace_rows(|A region|): 1 to 8 rows
// per lane, the fewest that cover a region of up to 512 rows in one strip);
// lane k owns rows k * R .. + R - 1 of a strip and at step t computes
// column t - k of them,
// receiving H and E of the row above (and the column's B code) from lane
// k - 1 through DPP; lane 0 reads the row above the strip from the
// comparison's work area, where lane 63 wrote it during the previous strip.
// Four direction bits per cell, one 32-bit word per lane and step, stored
// coalesced at dir[(strip * (n + 63) + t) * 64 + lane]: row r's nibble at
// bit 4 * (R - 1 - r), its bits 3 (H not from the diagonal), 2 (E above F),
// 1 (E opened) and 0 (F opened).  The wave then walks back as k_traceback does.
constexpr int kTrNeg = -(1 << 30);

template <int R>
__device__ __forceinline__ void trace_wave(const TraceArgs &T, const int32_t *tab, int p, int s, int as, int bs,
                           int m, int n) {
  const int lane = threadIdx.x;
  const int kp = T.kp;
  const uint8_t *a = T.codes + T.a_off[p] + as;
  const uint8_t *b = T.codes + T.b_off[p] + bs;
  uint32_t *const dirw = reinterpret_cast<uint32_t *>(T.dir + T.dir_off[p]);
  int32_t *const bnd = reinterpret_cast<int32_t *>(T.work + T.work_off[p]);
  const int go = T.gap_o, ge = T.gap_x;
  const int nss = (m + 64 * R - 1) / (64 * R);
  const int TT = n + 63;  // steps per strip
  const bool first = lane == 0, last = lane == 63;
  int Hc[R];
  for (int ss = 0; ss < nss; ++ss) {
    // the lane's rows: 0-based A index i0 + r, DP row i0 + r + 1
    const int i0 = (ss * 64 + lane) * R;
    int arow[R], F[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = i0 + r;
      arow[r] = (i < m ? min((int)a[i], kp - 1) : 0) * kp;
      Hc[r] = -(go + i * ge);  // H(i + 1, 0)
      F[r] = kTrNeg;           // F(i + 1, 0)
    }
    int hdiag = i0 == 0 ? 0 : -(go + (i0 - 1) * ge);  // H(i0, 0)
    // lane 0's inputs at step t (column t's B code, and H and E of the row
    // above the strip) come from blocks of 64 columns that every lane loads
    // one column of, a block ahead, read with v_readlane; lane 63 overwrites
    // column c of the work area at step c + 63, after c's block was loaded
    auto load_block = [&](int c0, int &bv, int &hv, int &ev) {
      const int c = c0 + lane;
      bv = c < n ? min((int)b[c], kp - 1) : 0;
      if (ss == 0) {
        hv = -(go + c * ge);  // H(0, c + 1)
        ev = kTrNeg;          // E(0, c + 1)
      } else if (c < n) {
        hv = bnd[2 * c];
        ev = bnd[2 * c + 1];
      }
    };
    int b_cur, h_cur = 0, e_cur = 0, b_nxt, h_nxt = 0, e_nxt = 0;
    load_block(0, b_cur, h_cur, e_cur);
    load_block(64, b_nxt, h_nxt, e_nxt);
    int out_h = 0, out_e = 0, col = 0;
    uint32_t *const dws = dirw + (int64_t)ss * TT * 64 + lane;
    const bool store_bnd = last && ss + 1 < nss;
    for (int t = 0; t < TT; ++t) {
      const int j = t - lane;  // 0-based B index: DP column j + 1
      if (t > 0 && (t & 63) == 0) {
        b_cur = b_nxt;
        h_cur = h_nxt;
        e_cur = e_nxt;
        load_block(t + 64, b_nxt, h_nxt, e_nxt);
      }
      const int sl = t & 63;
      const int bc = __builtin_amdgcn_readlane(b_cur, sl);
      const int bh = __builtin_amdgcn_readlane(h_cur, sl);
      const int bee = __builtin_amdgcn_readlane(e_cur, sl);
      const int up_h = (int)from_prev<kModeWave>((uint32_t)bh, (uint32_t)out_h, first);
      const int up_e = (int)from_prev<kModeWave>((uint32_t)bee, (uint32_t)out_e, first);
      col = (int)from_prev<kModeWave>((uint32_t)bc, (uint32_t)col, first);
      if (j >= 0 && j < n) {
        int hd = hdiag, hup = up_h, eup = up_e;
        uint32_t w = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int sc = tab[arow[r] + col];
          const int eop = hup - go, eex = eup - ge;
          const int e = max(eop, eex);
          const int fop = Hc[r] - go, fex = F[r] - ge;
          const int f = max(fop, fex);
          const int d = hd + sc;
          const int h = max(d, max(e, f));
          // each bit is the sign of a difference, shifted in from the right
          w = __builtin_amdgcn_alignbit(w, (uint32_t)(d - h), 31);      // d < h
          w = __builtin_amdgcn_alignbit(w, (uint32_t)(f - e), 31);      // e > f
          w = __builtin_amdgcn_alignbit(w, (uint32_t)(eex - eop), 31);  // E opened
          w = __builtin_amdgcn_alignbit(w, (uint32_t)(fex - fop), 31);  // F opened
          hd = Hc[r];
          Hc[r] = h;
          F[r] = f;
          hup = h;
          eup = e;
        }
        out_h = hup;
        out_e = eup;
        dws[(int64_t)t * 64] = w;
        if (store_bnd) {
          bnd[2 * j] = hup;
          bnd[2 * j + 1] = eup;
        }
        hdiag = up_h;
      }
    }
    __syncthreads();  // the work area and the direction words, for the next strip and the walk
  }
  // H(m, n): row m - 1 (0-based) of the last strip, after its last column
  const int rm = (m - 1) % R, lm = ((m - 1) % (64 * R)) / R;
  int hv = Hc[0];
#pragma unroll
  for (int r = 1; r < R; ++r)
    if (r == rm) hv = Hc[r];
  const int hend = __shfl(hv, lm);
  if (hend != s) {
    if (first) T.cig_len[p] = -2;
    return;
  }
  // The walk runs on every lane (uniform values); the direction nibbles come
  // 64 at a time: lane q loads the cell q steps from the walk's position along
  // its current move (diagonal, up, left), so that one round trip to memory
  // serves a run of up to 64 moves in the same state.
  uint32_t *out = T.cigar + T.cig_off[p];
  int runs = 0, run_op = -1, run_len = 0;
  int i = m, j = n, st = 0;
  int pi = -1, pj = -1, pst = -1;  // origin and state of the prefetched line
  uint32_t pn = 0;                 // lane q: the nibble of the line's cell q
  while (i > 0 || j > 0) {
    int op;
    if (i == 0) {
      op = 2;
      --j;
    } else if (j == 0) {
      op = 1;
      --i;
    } else {
      const int di = pi - i, dj = pj - j;
      int q = -1;
      if (pst == 0 && di == dj && di >= 0 && di < 64) q = di;
      else if (pst == 1 && dj == 0 && di >= 0 && di < 64) q = di;
      else if (pst == 2 && di == 0 && dj >= 0 && dj < 64) q = dj;
      if (q < 0) {
        pi = i;
        pj = j;
        pst = st;
        const int ci = i - (st != 2 ? lane : 0), cj = j - (st != 1 ? lane : 0);
        pn = 0;
        if (ci >= 1 && cj >= 1) {
          const int ri = ci - 1, li = (ri % (64 * R)) / R;
          const int64_t idx = ((int64_t)(ri / (64 * R)) * TT + (cj - 1) + li) * 64 + li;
          pn = (dirw[idx] >> (4 * (R - 1 - ri % R))) & 15u;
        }
        q = 0;
      }
      const uint32_t nib = (uint32_t)__builtin_amdgcn_readlane((int)pn, q);
      if (st == 0) st = (nib & 8u) ? ((nib & 4u) ? 1 : 2) : 0;
      if (st == 0) {
        op = 0;
        --i;
        --j;
      } else if (st == 1) {
        op = 1;
        st = (nib & 2u) ? 0 : 1;
        --i;
      } else {
        op = 2;
        st = (nib & 1u) ? 0 : 2;
        --j;
      }
    }
    if (op == run_op) {
      ++run_len;
    } else {
      if (run_len && first) out[runs] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
      runs += run_len ? 1 : 0;
      run_op = op;
      run_len = 1;
    }
  }
  if (run_len && first) out[runs] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
  runs += run_len ? 1 : 0;
  if (!first) return;
  T.cig_len[p] = runs;
}

__global__ __launch_bounds__(64) void k_trace_wave(TraceArgs T) {
  __shared__ int32_t tab[26 * 26];
  const int lane = threadIdx.x;
  const int kp = T.kp;
  for (int i = lane; i < kp * kp; i += 64) tab[i] = T.table[i];
  __syncthreads();
  const int p = T.order ? T.order[blockIdx.x] : (int)blockIdx.x;
  const int s = T.scores[p];
  const uint32_t ar = (uint32_t)T.a_range[p], br = (uint32_t)T.b_range[p];
  const int as = (int)(ar & 0xFFFF), ae = (int)(ar >> 16);
  const int bs = (int)(br & 0xFFFF), be = (int)(br >> 16);
  if (s <= 0 || ae < as || be < bs) {
    if (lane == 0) T.cig_len[p] = 0;
    return;
  }
  const int m = ae - as + 1, n = be - bs + 1;
  // The longest comparisons set the launch's end: a wave issues ahead of the
  // other waves of its SIMD by its steps times rows (the lanes' cells)
  const int rows = trace_rows(m);
  const int64_t work = (int64_t)((m + 64 * rows - 1) / (64 * rows)) * (n + 63) * rows;
  if (work > 12000) __builtin_amdgcn_s_setprio(3);
  else if (work > 6000) __builtin_amdgcn_s_setprio(2);
  else if (work > 3000) __builtin_amdgcn_s_setprio(1);
  switch (rows) {
    case 1: trace_wave<1>(T, tab, p, s, as, bs, m, n); break;
    case 2: trace_wave<2>(T, tab, p, s, as, bs, m, n); break;
    case 3: trace_wave<3>(T, tab, p, s, as, bs, m, n); break;
    case 4: trace_wave<4>(T, tab, p, s, as, bs, m, n); break;
    case 5: trace_wave<5>(T, tab, p, s, as, bs, m, n); break;
    case 6: trace_wave<6>(T, tab, p, s, as, bs, m, n); break;
    case 7: trace_wave<7>(T, tab, p, s, as, bs, m, n); break;
    default: trace_wave<kTrRows>(T, tab, p, s, as, bs, m, n); break;
  }
}

// ---------------------------------------------------------------------------
// Prepared input buffers (the reference's wire format): the residues of each
// slot's A and B are copied from their tile regions to 4-aligned places of a
// codes arena, codes outside the scoring matrix mapped by `remap` (as the
// character path maps them).  Sequence q (A of slot q, or B of slot q - slots)
// is copied by one wave.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wire_gather(int64_t nseq, const uint8_t *wire,
                                                     const int64_t *src, const int64_t *dst,
                                                     const int32_t *len, const uint8_t *remap,
                                                     uint8_t *arena) {
  const int lane = threadIdx.x & 63;
  const int64_t step = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < nseq; q += step) {
    const int l = len[q];
    const uint8_t *s = wire + src[q];
    uint8_t *d = arena + dst[q];
    for (int i = lane; i < l; i += 64) d[i] = remap[s[i]];
  }
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side).
// ---------------------------------------------------------------------------
static inline int nblocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

hipError_t launch_fill_ranges(int32_t *const p[4], const int64_t n[4], const int32_t v[4],
                              hipStream_t st) {
  FillRanges f;
  int64_t total = 0;
  for (int i = 0; i < 4; ++i) {
    f.p[i] = p[i];
    f.n[i] = std::max<int64_t>(n[i], 0);
    f.v[i] = v[i];
    total += f.n[i];
  }
  if (total <= 0) return hipSuccess;
  const int blocks = (int)std::min<int64_t>(nblocks(total, 256), 128);
  hipLaunchKernelGGL(k_fill_ranges, dim3(blocks), dim3(256), 0, st, f);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int blocks = (int)std::min<int64_t>(nblocks(n, 256), 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_encode(const uint8_t *chars, int64_t len, uint8_t *codes, const uint8_t *table,
                         hipStream_t st) {
  if (len <= 0) return hipSuccess;
  const int blocks = (int)std::min<int64_t>(nblocks(len, 256), 8192);
  hipLaunchKernelGGL(k_encode, dim3(blocks), dim3(256), 0, st, chars, len, codes, table);
  return hipGetLastError();
}

hipError_t launch_bucketing(const BucketArgs &B, hipStream_t st) {
  const int n = B.n;
  const int nb = nblocks(std::max(n, 1), 256);  // one thread per pair, 256 per block
  hipLaunchKernelGGL(k_sum_cells, dim3(std::min(nblocks(std::max(n, 1), kSumThreads), kSumBlocks)),
                     dim3(kSumThreads), 0, st, n, B.a_len, B.b_len, B.ctl);
  hipLaunchKernelGGL(k_classify, dim3(nblocks(std::max(n, 1), kClsThreads)), dim3(kClsThreads), 0,
                     st, n, B.a_len, B.b_len, B.target_slots, B.plain ? 0 : B.force_class,
                     B.wave_cols, B.max_a, B.bin_of, B.rank, B.hist, B.ctl, B.scores, B.a_range, B.b_range,
                     B.plain ? nullptr : B.target);
  if (B.plain) {
    hipLaunchKernelGGL(k_plain_setup, dim3(nb), dim3(256), 0, st, n, B.bin_of, B.sorted);
    hipLaunchKernelGGL(k_plain_groups, dim3(nblocks((int64_t)n * 64 + 64, 256)), dim3(256), 0,
                       st, n, B.plain == 2 ? 1 : 0, B.wave_cols, B.sorted, B.a_len, B.b_len,
                       B.ctl, B.groups);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_run, dim3(kScanRunBlocks), dim3(256), 0, st, B.hist, B.ctl);
  hipLaunchKernelGGL(k_scan_runs, dim3(1), dim3(kRuns), 0, st, B.ctl);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, st, n, B.bin_of, B.rank, B.hist,
                     B.ctl, B.sorted);
  const int gb = std::min(nblocks(B.max_groups, kGroupThreads / 64), kGroupBlocks);
  hipLaunchKernelGGL(k_groups, dim3(gb), dim3(kGroupThreads), 0, st, B.max_groups, B.sorted,
                     B.a_len, B.b_len, B.ctl, B.gtmp, B.gbin, B.grank, B.costhist);
  hipLaunchKernelGGL(k_scan_cost, dim3(1), dim3(1024), 0, st, B.costhist);
  hipLaunchKernelGGL(k_order_groups, dim3(std::min(nblocks(B.max_groups, 256), 512)), dim3(256), 0,
                     st, B.max_groups, B.ctl, B.gtmp, B.gbin, B.grank, B.costhist, B.groups);
  return hipGetLastError();
}

hipError_t launch_traceback(const TraceArgs &T, hipStream_t st) {
  if (T.n <= 0) return hipSuccess;
  if (T.wave)
    hipLaunchKernelGGL(k_trace_wave, dim3(T.n), dim3(64), 0, st, T);
  else
    hipLaunchKernelGGL(k_traceback, dim3(nblocks(T.n, 64)), dim3(64), 0, st, T);
  return hipGetLastError();
}

hipError_t launch_wire_gather(int64_t nseq, const uint8_t *wire, const int64_t *src,
                              const int64_t *dst, const int32_t *len, const uint8_t *remap,
                              uint8_t *arena, hipStream_t st) {
  if (nseq <= 0) return hipSuccess;
  const int blocks = (int)std::min<int64_t>(nblocks(nseq * 64, 256), 8192);
  hipLaunchKernelGGL(k_wire_gather, dim3(blocks), dim3(256), 0, st, nseq, wire, src, dst, len,
                     remap, arena);
  return hipGetLastError();
}

hipError_t launch_fast(int kp, const FastArgs &A, int blocks, hipStream_t st) {
  // the start pass (targets set) runs the kernel with the early stop compiled in
  if (A.target != nullptr) {
    if (kp == 18)
      hipLaunchKernelGGL((k_fast<18, true>), dim3(blocks), dim3(kFastThreads), 0, st, A);
    else
      hipLaunchKernelGGL((k_fast<26, true>), dim3(blocks), dim3(kFastThreads), 0, st, A);
  } else if (kp == 18) {
    hipLaunchKernelGGL((k_fast<18, false>), dim3(blocks), dim3(kFastThreads), 0, st, A);
  } else {
    hipLaunchKernelGGL((k_fast<26, false>), dim3(blocks), dim3(kFastThreads), 0, st, A);
  }
  return hipGetLastError();
}

int fast_blocks_per_cu(int kp) {
  // resident blocks of both instances (forward, and start pass with its stop)
  int nb = 0, ns = 0;
  hipError_t e = kp == 18
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fast<18, false>, kFastThreads, 0)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fast<26, false>, kFastThreads, 0);
  hipError_t e2 = kp == 18
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&ns, k_fast<18, true>, kFastThreads, 0)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&ns, k_fast<26, true>, kFastThreads, 0);
  if (e != hipSuccess || e2 != hipSuccess) return 1;
  const int m = std::min(nb, ns);
  return m > 0 ? m : 1;
}

hipError_t launch_generic(const GenericArgs &G, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_wave32, dim3(blocks), dim3(64 * kW32Waves), 0, st, G);
  return hipGetLastError();
}

hipError_t launch_reverse_prefix(int32_t n, const uint8_t *codes, const int64_t *a_off,
                                 const int32_t *a_len, const int64_t *b_off,
                                 const int32_t *b_len, const int32_t *scores,
                                 const int32_t *a_range, const int32_t *b_range, uint8_t *rcodes,
                                 int32_t *ra_len, int32_t *rb_len, int32_t *error,
                                 hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reverse_prefix, dim3(nblocks((int64_t)n * 64, 256)), dim3(256), 0, st, n,
                     codes, a_off, a_len, b_off, b_len, scores, a_range, b_range, rcodes, ra_len,
                     rb_len, error);
  return hipGetLastError();
}

hipError_t launch_finish_starts(int32_t n, const int32_t *scores, const int32_t *ra_range,
                                const int32_t *rb_range, int32_t *a_range, int32_t *b_range,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finish_starts, dim3(nblocks(n, 256)), dim3(256), 0, st, n, scores,
                     ra_range, rb_range, a_range, b_range);
  return hipGetLastError();
}

}  // namespace ipuma
