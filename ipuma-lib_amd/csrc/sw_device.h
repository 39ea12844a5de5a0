// sw_device.h -- shared host/device definitions of the MI355X Smith-Waterman engine.
#pragma once
#include <stdint.h>

namespace ipuma {

// Rows of one lane's strip held in registers by the packed kernel.
constexpr int kRows = 16;
// Threads per block of the packed kernel (one block per CU: the row-pair
// scoring table takes most of the LDS).
constexpr int kFastThreads = 768;
// Lanes-per-pair classes of the packed kernel (pipeline depth inside a wave);
// class c < 5 has L = 2^c, the last class L = 64.
constexpr int kNumClasses = 6;
constexpr int kClassL[kNumClasses] = {1, 2, 4, 8, 16, 64};
// Scratch of one wave, in 32-bit words: a group's boundary row (two words per
// pipeline column, mpad * 64/L columns) and its column table (one word per
// pipeline step and pipeline: at most (mpad + L + 4) * 64/L).
__host__ __device__ constexpr int64_t wave_scratch_words(int64_t mpad, int lg) {
  return (3 * mpad + (1 << lg) + 4) << (6 - lg);
}
// Boundary columns (mpad * 64/L) that a wave scratch of `words` words holds
// for every pipeline depth: (3 * mpad + L + 4) * 64/L <= 3 * mpad * 64/L + 320.
__host__ __device__ constexpr int64_t wave_cols_of(int64_t words) { return (words - 320) / 3; }
// The class rules keep mpad * (64/L) <= 65536: 769 KiB per wave.
constexpr int64_t kWaveScratchWords = 3 * 65536 + 320;
// Boundary row of one band slot (a 64-lane group: two words per column).
constexpr int64_t kBandScratchWords = 2 * 65536;
// Longest sequence the engine accepts (positions are packed as uint16).
constexpr int32_t kMaxSeqLen = 65535;
// Rows per lane of k_trace_wave: a strip is 64 * R rows, and a lane's
// direction bits (4 per cell) take one 32-bit word per step.  R is the fewest
// (1..8) whose strip holds the region's |A| rows, else kTrRows.
constexpr int kTrRows = 8;
__host__ __device__ constexpr int trace_rows(int64_t m) {
  return m <= 64 * kTrRows ? (int)(m > 64 ? (m + 63) / 64 : 1) : kTrRows;
}
// Counting-sort bins of the length bucketing: class x super-strip bin x
// (rows of the last super-strip, m-bin).  A run is one (class, super-strip
// bin); groups never straddle runs, and each run starts at a multiple of
// kRunAlign slots of the sorted pair array.  Inside a run, pairs whose last
// super-strip needs 8, 12 or 16 rows per lane are kept apart (a group runs
// its last super-strip with the rows its longest B needs), then ordered by
// |A| (kMBinsPerRows bins of each).
constexpr int kNssBins = 64;
constexpr int kMBins = 512;
constexpr int kMBinsPerRows = 170;
static_assert(3 * kMBinsPerRows <= kMBins, "m-bins of the three last-strip row counts");
// Whole-wave groups of one super-strip: five row counts, m-bins of their own.
constexpr int kMBinsPerRowsWave = 102;
static_assert(5 * kMBinsPerRowsWave <= kMBins, "m-bins of the five whole-wave row counts");
constexpr int kRuns = kNumClasses * kNssBins;
constexpr int kNumBins = kRuns * kMBins;
constexpr int kRunAlign = 128;
constexpr int kSortedPad = kRuns * kRunAlign;  // empty slots the run alignment can add
// Group-cost bins of the processing order (longest first).
constexpr int kCostBinsPerClass = 256;
constexpr int kCostBins = 2048;  // >= kCostBinsPerClass, multiple of 1024
static_assert(kCostBins >= kCostBinsPerClass, "cost bins");
// Band groups: a whole-wave pipeline group (L = 64) with at least two
// super-strips is shared by the waves of one block, one super-strip per wave
// at a time; a block keeps up to one band group per wave in flight, each with
// its own boundary row of kBandScratchWords words.
// Groups outside the band groups whose steps reach a quarter of the step cap
// are head groups: they start with the band groups instead of after them.
constexpr int kHeadCapDiv = 4;
constexpr int kBandSlots = kFastThreads / 64;
constexpr int kMaxStrips = 64;  // super-strips of 1,024 rows cover 65,535 rows
// Saturation of the packed int16 path: the packed kernel holds its DP values
// plus kBias16, as positive normal fp16 bit patterns below 0x7C00, and a
// comparison whose maximum reaches kSat16 is re-run in int32.
constexpr int32_t kBias16 = 1024;
constexpr int32_t kSat16 = 0x7C00 - kBias16;

// Group descriptor: one wave's work in the packed kernel.
struct GroupDesc {
  int32_t cls;        // class index into kClassL
  int32_t mpad;       // columns (|A|) padded for the group
  int32_t nss;        // super-strips (rows / (L * kRows), rounded up)
  int32_t base;       // first slot in the sorted pair array
  int32_t last_rows;  // rows per lane the last super-strip needs (1..kRows)
};

// Rows per lane the packed kernel runs the last super-strip of a group with:
// the smallest of 8, 12 and kRows that covers `need`; whole-wave groups (one
// super-strip of 64 lanes, where four rows per lane are 256 rows of B) take
// the smallest even count from 8 up.
__host__ __device__ inline int last_strip_rows(int need, bool wave = false) {
  if (wave) return need <= 8 ? 8 : need >= kRows ? kRows : (need + 1) & ~1;
  return need <= 8 ? 8 : need <= 12 ? 12 : kRows;
}
// Rows class of a comparison's last super-strip in the bucketing (pairs of
// one class share groups): 8, 12, 16 rows per lane, or for whole-wave groups
// of one super-strip 8, 10, 12, 14, 16.
__host__ __device__ inline int rows_class(int need, bool wave) {
  if (wave) return need <= 8 ? 0 : (min(need, kRows) - 7) / 2;
  return need <= 8 ? 0 : need <= 12 ? 1 : 2;
}
// Cost of one step with `rows` rows per lane, in 32nds of a kRows step (the
// step bodies take about 135, 187 and 240 VALU instructions for 8, 12, 16).
__host__ __device__ inline int step_weight32(int rows) {
  return rows <= 8 ? 18 : rows <= 10 ? 21 : rows <= 12 ? 25 : rows <= 14 ? 28 : 32;
}

// Control block of the device bucketing (one per batch, zeroed before use).
struct BucketCtl {
  int32_t class_count[kNumClasses];
  int32_t run_start[kRuns];           // aligned start of each run in `sorted`
  int32_t run_group_base[kRuns + 1];  // first group of each run; [kRuns] = total
  int32_t ngroups;
  int32_t error;       // bit 0: a length outside [0, kMaxSeqLen]; bit 1: an end
                       // position outside its sequence; bit 2: a bad group; bit 3: no
                       // free band slot; bit 4: an int32 boundary-row wait timed out
  int32_t ovf_count;   // int16-saturated comparisons listed for int32
  int32_t maxm;        // longest A of the batch
  int32_t maxn;        // longest B of the batch
  int32_t queue;       // next per-wave group of the packed kernel's work queue
  int32_t nband;       // band groups: the first nband entries of the group order
  int32_t band_queue;  // next band group (taken by a whole block)
  int32_t nhead;       // head groups: the nhead entries after the band groups
  int32_t head_queue;  // next head group
  int32_t head_cost;   // steps from which a group that is not a band group is a head group
  unsigned long long total_cells;
  unsigned long long padded_cells;
};
static_assert(sizeof(BucketCtl) % 8 == 0, "BucketCtl is cleared as whole int32 words");

// Arguments of the packed kernel.
struct FastArgs {
  const uint8_t *codes;
  const int64_t *a_off;
  const int32_t *a_len;
  const int64_t *b_off;
  const int32_t *b_len;
  const int16_t *table;        // KP x KP, row = A code, col = B code
  const int32_t *sorted;       // pair ids in bucket order (-1 = empty slot)
  const GroupDesc *groups;     // in processing order
  BucketCtl *ctl;              // group ranges per class, error bits
  uint32_t *scratch;           // wave_words per wave
  uint32_t *band_scratch;      // kBandSlots * band_words per block
  int64_t wave_words;          // scratch of one wave, in words (kWaveScratchWords, or less
                               // for a smaller scratch): boundary row + column table
  int64_t band_words;          // boundary row of one band slot, in words
  int32_t *scores;
  int32_t *a_range;
  int32_t *b_range;
  int32_t *ovf_list;           // pairs that saturated int16
  int32_t *ovf_count;
  int32_t max_groups;          // bound of the group count
  // Start pass only: the forward score of each pair.  The maximum over the
  // reversed prefixes equals it, so a group stops after the first
  // super-strip by which every pair has reached it (the first maximum in
  // B-major order lies in that super-strip or an earlier one); nullptr: run
  // every super-strip.
  const int32_t *target;
  uint32_t gap_o;              // -(gapInit + gapExtend), packed twice
  uint32_t gap_x;              // -gapExtend, packed twice
  int32_t smax;                // largest substitution score, at least 0
};

// Arguments of the exact int32 kernel.
struct GenericArgs {
  const uint8_t *codes;
  const int64_t *a_off;
  const int32_t *a_len;
  const int64_t *b_off;
  const int32_t *b_len;
  const int16_t *table;
  int32_t kp;
  const int32_t *list;   // pair ids, or nullptr for 0..n-1
  const int32_t *count;  // device count of `list`, or nullptr
  int32_t n;             // pair count (bound of *count when list != nullptr)
  int32_t *work;         // work_stride ints per wave (its boundary row: 2 per column)
  int64_t work_stride;   // >= 2 * longest A the kernel accepts
  int32_t max_a;         // longest A accepted; a longer one raises error bit 0
  int32_t *error;        // set to 1 when a length exceeds kMaxSeqLen
  int32_t gap_init;
  int32_t gap_extend;
  int32_t *scores;
  int32_t *a_range;
  int32_t *b_range;
};

// Arguments of the traceback kernel (one comparison per lane).
struct TraceArgs {
  const uint8_t *codes;
  const int64_t *a_off;     // [n] start of A in `codes`
  const int64_t *b_off;     // [n]
  const int32_t *a_range;   // [n] end << 16 | start
  const int32_t *b_range;   // [n]
  const int32_t *scores;    // [n]
  const int64_t *dir_off;   // [n] byte offset of the comparison's direction cells
  const int64_t *work_off;  // [n] int64 offset of its row buffers (2 x (|B region| + 1))
  const int64_t *cig_off;   // [n] offset of its runs in `cigar`
  int32_t n;
  int32_t gap_o;            // -(gapInit + gapExtend): the first residue of a gap
  int32_t gap_x;            // -gapExtend: each further residue
  int32_t kp;               // table side
  const int16_t *table;     // kp x kp, row = A code
  uint8_t *dir;
  int64_t *work;
  uint32_t *cigar;          // runs, end to start
  int32_t *cig_len;         // [n] runs, or -2 when the region's optimum differs from the score
  const int32_t *order;     // [n] comparison of each wave (k_trace_wave), or null
  int32_t wave;             // 1: k_trace_wave (int32, a wave per comparison), 0: k_traceback
};

// Arguments of the bucketing pipeline.
struct BucketArgs {
  int32_t n;
  const int32_t *a_len;
  const int32_t *b_len;
  int64_t target_slots;   // resident half-lanes x rounds: step cap = cells / (kRows * this)
  int32_t force_class;    // -1: choose per pair
  int32_t wave_cols;      // boundary columns a wave's scratch holds: a pair's pipeline is
                          // deep enough that mpad * 64/L <= wave_cols (wave_cols_of(words))
  int32_t max_a;          // longest A accepted (kMaxSeqLen, or less with a smaller scratch);
                          // a longer one is an error like a length past kMaxSeqLen
  // Start pass only: the forward scores.  A pair's super-strip bin then counts
  // the rows it is expected to need before reaching its score (2 * score + 16,
  // at most |B|), so pairs that stop early share groups; nullptr otherwise.
  const int32_t *target;
  int32_t plain;          // 0: bucketed; 1: input order, batch-max padding; 2: input order,
                          // each group padded to its own longest pair
  int32_t max_groups;
  int32_t *bin_of;
  int32_t *rank;          // n: slot of the pair inside its bin
  int32_t *hist;          // kNumBins
  int32_t *costhist;      // kCostBins
  BucketCtl *ctl;
  int32_t *sorted;        // n + kSortedPad, pre-filled with -1
  GroupDesc *gtmp;
  int32_t *gbin;
  int32_t *grank;         // max_groups: slot of the group inside its cost bin
  GroupDesc *groups;
  int32_t *scores;
  int32_t *a_range;
  int32_t *b_range;
};

}  // namespace ipuma

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
namespace ipuma {
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t st);
// Up to four ranges filled by one launch (a range with count 0 is skipped).
hipError_t launch_fill_ranges(int32_t *const p[4], const int64_t n[4], const int32_t v[4],
                              hipStream_t st);
hipError_t launch_encode(const uint8_t *chars, int64_t len, uint8_t *codes, const uint8_t *table,
                         hipStream_t st);
hipError_t launch_bucketing(const BucketArgs &B, hipStream_t st);
hipError_t launch_traceback(const TraceArgs &T, hipStream_t st);
hipError_t launch_wire_gather(int64_t nseq, const uint8_t *wire, const int64_t *src,
                              const int64_t *dst, const int32_t *len, const uint8_t *remap,
                              uint8_t *arena, hipStream_t st);
hipError_t launch_fast(int kp, const FastArgs &A, int blocks, hipStream_t st);
int fast_blocks_per_cu(int kp);
hipError_t launch_generic(const GenericArgs &G, int blocks, hipStream_t st);
hipError_t launch_reverse_prefix(int32_t n, const uint8_t *codes, const int64_t *a_off,
                                 const int32_t *a_len, const int64_t *b_off,
                                 const int32_t *b_len, const int32_t *scores,
                                 const int32_t *a_range, const int32_t *b_range, uint8_t *rcodes,
                                 int32_t *ra_len, int32_t *rb_len, int32_t *error,
                                 hipStream_t st);
hipError_t launch_search_pairs(int32_t nq, int32_t ndb, const int64_t *q_off,
                               const int32_t *q_len, const int64_t *db_off,
                               const int32_t *db_len, int64_t *a_off, int32_t *a_len,
                               int64_t *b_off, int32_t *b_len, int32_t query_cols,
                               hipStream_t st);
hipError_t launch_topk(int32_t nq, int32_t ndb, int32_t k, const int32_t *scores,
                       int32_t *top_idx, int32_t *top_score, hipStream_t st);
hipError_t launch_finish_starts(int32_t n, const int32_t *scores, const int32_t *ra_range,
                                const int32_t *rb_range, int32_t *a_range, int32_t *b_range,
                                hipStream_t st);
}  // namespace ipuma
#endif
