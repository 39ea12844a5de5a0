// sw_engine.cpp -- host side of the MI355X Smith-Waterman engine and its C ABI.
//
// One engine = one HIP device + one stream + the scoring table in HBM + grow-
// only scratch.  A device batch is a fixed chain of kernels on one stream,
// with no host round trip:
//   bucketing (k_sum_cells, k_classify, k_scan_run, k_scan_runs, k_scatter,
//   k_groups, k_scan_cost, k_order_groups)  ->  k_fast (packed int16, persistent)  ->
//   k_wave32 (exact int32 re-run of saturated comparisons)  ->
//   [k_reverse_prefix -> same chain on the reversed prefixes -> k_finish_starts]
// The last block only runs without forward_only, for start positions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ipuma_hip.h"
#include "sw_device.h"
#include "sw_engine_internal.h"
#include "sw_tables.h"

using namespace ipuma;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

double since_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(IPUMA_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

// Work area of the int32 kernel per wave: the boundary row (H, F) of an A of
// up to kMaxSeqLen columns.
constexpr int64_t kWorkStride = 2 * 65536;

// Grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return IPUMA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&p, want) != hipSuccess) {
      p = nullptr;
      return fail(IPUMA_ENOMEM, "hipMalloc(%zu) failed", want);
    }
    cap = want;
    return IPUMA_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// Device state of one bucketing + kernel chain (the forward pass and the
// reverse pass each own one, so they can be in flight together).
struct ChainBufs {
  DevBuf bin_of, rank, hist, costhist, ctl, sorted, gtmp, gbin, grank, groups, ovf;
  void release() {
    for (DevBuf *b : {&bin_of, &rank, &hist, &costhist, &ctl, &sorted, &gtmp, &gbin, &grank,
                      &groups, &ovf})
      b->release();
  }
};

// A prepared batch queued for the engine's executor thread.
struct PreparedJob {
  const int32_t *in = nullptr;
  int64_t in_words = 0;
  int32_t *res = nullptr;
  int64_t res_words = 0;
  int rc = 0;
  std::string err;
  bool done = false;
  ipuma_job_times times{};  // copy in, gather + device batch (HIP events); executor wall time
};

}  // namespace

struct ipuma_engine {
  ipuma_sw_config sw{};
  ipuma_algo_config algo{};
  int device = 0;
  int num_cu = 0;
  double clock_hz = 0.0;  // peak shader clock
  int kp = 0;   // padded alphabet (matrix side + pad code)
  int dim = 0;  // matrix side
  int smax = 0; // largest entry of the score table, at least 0
  bool symmetric = false;  // the score table equals its transpose
  hipStream_t stream = nullptr;
  std::vector<int16_t> h_table;
  uint8_t h_codes[256];
  int16_t *d_table = nullptr;
  uint8_t *d_codetab = nullptr;
  int batching = IPUMA_BATCH_BUCKETED;
  int precision = IPUMA_PREC_AUTO;
  int forward_only = 0;
  int force_class = -1;
  bool packed_ok = false;  // gap settings inside the packed kernel's domain
  int fast_blocks = 0;
  int generic_blocks = 512;
  int64_t target_slots = 0;
  int rounds = 1;  // groups per resident wave the bucketing aims for
  int cap_milli = 1000;  // step cap scale, in thousandths
  DevBuf scratch, band_scratch, work;
  ChainBufs fwd, rev;
  DevBuf rcodes, rlen, rres;
  DevBuf s_pairs, s_res;  // one-vs-many search: pair descriptors and results
  // host staging of compare_local / compare_mn_local
  DevBuf in_chars, in_codes, in_meta, out_res;
  char *h_pin = nullptr;  // pinned staging of host characters (grown as needed)
  size_t h_pin_cap = 0;
  char *h_meta = nullptr;  // pinned staging of descriptors in, results out
  size_t h_meta_cap = 0;
  std::vector<int64_t> h_aoff, h_boff;
  std::vector<int32_t> h_alen, h_blen;
  ipuma_batch_stats stats{};
  BucketCtl *last_ctl = nullptr;  // device ctl of the last forward chain
  BucketCtl *na_ctl = nullptr;    // device ctl of the allocation-free entry, when the last
                                  // device batch was one of its calls
  int32_t last_n = 0;
  int timing = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // batch begin, fast begin/end, batch end
  // IPUMA_LOG set: one JSON line per device batch on stderr (tag IPUSWLOG)
  bool log_batches = false;
  bool log_pending = false;  // a device batch ran since the last line
  bool ev_recorded = false;
  // Host calls are serialized; device work of calls on different streams is
  // ordered through order_ev (the engine's buffers are shared).
  std::mutex mu;
  hipEvent_t order_ev = nullptr;
  // prepared input buffers (the reference's wire format)
  uint8_t *d_wiretab = nullptr;  // code -> code inside the matrix
  DevBuf w_wire, w_desc, w_arena, w_res;
  hipEvent_t w_ev[3] = {nullptr, nullptr, nullptr};  // prepared batch: begin, input copied, batch done
  double w_h2d_ms = 0.0, w_device_ms = 0.0;          // their intervals for the last prepared batch
  // traceback
  DevBuf t_desc, t_dir, t_work, t_cig;
  hipEvent_t t_ev[2] = {nullptr, nullptr};  // around each traceback launch (timing on)
  int64_t t_cells = 0;                       // region cells of the last traceback call
  double t_kernel_ms = 0.0;                  // its kernels' time (timing on)
  std::vector<int64_t> h_src, h_dst;
  std::vector<int32_t> h_len;
  // executor thread of asynchronously submitted prepared batches
  std::mutex pj_mu;
  std::condition_variable pj_cv;
  std::deque<int64_t> pj_queue;
  std::unordered_map<int64_t, PreparedJob> pj_jobs;
  int64_t pj_next = 1;
  std::thread pj_thread;
  bool pj_stop = false;
};

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry (ipuma_compare_device_noalloc): the
// bucketing buffers, the boundary rows and the int32 work area of one call,
// in one static device array per GPU, sized for at most kNaPairs comparisons
// with |A| <= kNaMaxA.  Every call on a GPU uses that GPU's copy.
// ---------------------------------------------------------------------------
namespace {
constexpr int32_t kNaPairs = 65536;  // comparisons per call
constexpr int32_t kNaMaxA = 4096;    // longest A: a boundary row of mpad * 64/L <= 4096 columns
// per wave: that boundary row and the group's column table (49.25 KiB)
constexpr int64_t kNaWaveWords = 3 * (int64_t)kNaMaxA + 320;
static_assert(wave_cols_of(kNaWaveWords) == kNaMaxA, "wave scratch of the allocation-free entry");
constexpr int64_t kNaBandWords = 2 * (int64_t)kNaMaxA;  // per band slot: its boundary row (32 KiB)
constexpr int kNaBlocks = 256;         // persistent blocks of the packed kernel (one per CU)
constexpr int kNaGenericBlocks = 256;  // waves of the int32 kernel
constexpr int64_t kNaWorkStride = 2 * (int64_t)kNaMaxA;  // its boundary row per wave (32 KiB)
constexpr int64_t kNaGroups = (int64_t)kNaPairs + kRuns;
constexpr int64_t na_align(int64_t x) { return (x + 255) / 256 * 256; }
// byte offsets of the parts, in this order
enum NaPart { kNaCtl, kNaHist, kNaCost, kNaBin, kNaRank, kNaOvf, kNaGbin, kNaGrank, kNaSorted,
              kNaGtmp, kNaGroupsPart, kNaWave, kNaBand, kNaWork, kNaParts };
constexpr int64_t na_size(int part) {
  switch (part) {
    case kNaCtl: return (int64_t)sizeof(BucketCtl);
    case kNaHist: return (int64_t)kNumBins * 4;
    case kNaCost: return (int64_t)kCostBins * 4;
    case kNaBin: case kNaRank: case kNaOvf: return (int64_t)kNaPairs * 4;
    case kNaGbin: case kNaGrank: return kNaGroups * 4;
    case kNaSorted: return ((int64_t)kNaPairs + kSortedPad) * 4;
    case kNaGtmp: case kNaGroupsPart: return kNaGroups * (int64_t)sizeof(GroupDesc);
    case kNaWave: return (int64_t)kNaBlocks * (kFastThreads / 64) * kNaWaveWords * 4;
    case kNaBand: return (int64_t)kNaBlocks * kBandSlots * kNaBandWords * 4;
    case kNaWork: return (int64_t)kNaGenericBlocks * kNaWorkStride * 4;
    default: return 0;
  }
}
constexpr int64_t na_offset(int part) {
  int64_t o = 0;
  for (int i = 0; i < part; ++i) o += na_align(na_size(i));
  return o;
}
constexpr int64_t kNaBytes = na_offset(kNaParts);  // about 260 MiB
static_assert(kNaBytes < ((int64_t)288 << 20), "static scratch stays below 288 MiB");

// The array itself.  Host code never uses its name as an address: each GPU
// holds its own copy, whose device address hipGetSymbolAddress gives per
// device at engine set-up (NaDevice::ptr).
static __device__ __align__(256) uint8_t scratch[kNaBytes];

// Per device: the array's address there and an event that orders the calls
// on that device (they share the array).
constexpr int kNaMaxDevices = 64;
struct NaDevice {
  bool ready = false;
  int rc = IPUMA_OK;  // set-up failure, reported by the entry point
  std::string err;
  uint8_t *ptr = nullptr;
  hipEvent_t order = nullptr;
  std::mutex mu;      // serializes the calls' host side on this device
};
NaDevice g_na[kNaMaxDevices];
std::mutex g_na_setup_mu;

// Once per device, with that device current (called from engine set-up).
void na_setup(int dev) {
  if (dev < 0 || dev >= kNaMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_na_setup_mu);
  NaDevice &d = g_na[dev];
  if (d.ready || d.rc != IPUMA_OK) return;
  void *p = nullptr;
  hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (err == hipSuccess && p != nullptr)
    err = hipEventCreateWithFlags(&d.order, hipEventDisableTiming);
  if (err != hipSuccess || p == nullptr) {
    d.rc = IPUMA_EDEVICE;
    d.err = std::string("static scratch of device ") + std::to_string(dev) + ": " +
            (err != hipSuccess ? hipGetErrorString(err) : "no address");
    return;
  }
  d.ptr = static_cast<uint8_t *>(p);
  d.ready = true;
}
}  // namespace

extern "C" {
static void log_batch(ipuma_engine *e, const char *entry, double host_ms, int status);
}

// Resident half-lanes x rounds, divided by the step cap scale: the bucketing's
// cap on a group's steps is cells / (kRows * target_slots).
static void update_target_slots(ipuma_engine *e) {
  const int64_t base = (int64_t)e->fast_blocks * (kFastThreads / 64) * 128 * e->rounds;
  e->target_slots = std::max<int64_t>(1, base * 1000 / std::max(1, e->cap_milli));
}

namespace {

int init_engine(ipuma_engine *e) {
  HIP_TRY(hipGetDevice(&e->device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, e->device));
  e->num_cu = prop.multiProcessorCount;
  e->clock_hz = prop.clockRate > 0 ? prop.clockRate * 1e3 : 2.4e9;  // clockRate is in kHz
  HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  std::vector<int8_t> mat(25 * 25);
  const int dim = select_matrix(e->sw.similarity, e->sw.match_value, e->sw.mismatch_value,
                                e->sw.ambiguity_value, mat.data());
  if (dim < 0) return fail(IPUMA_ECONFIG, "No matrix for simple penalty available.");
  e->dim = dim;
  e->kp = dim + 1;
  e->h_table.resize((size_t)e->kp * e->kp);
  build_pad_table(mat.data(), dim, e->h_table.data());
  e->smax = std::max<int>(0, *std::max_element(e->h_table.begin(), e->h_table.end()));
  e->symmetric = true;
  for (int i = 0; i < e->kp; ++i)
    for (int j = 0; j < i; ++j)
      if (e->h_table[(size_t)i * e->kp + j] != e->h_table[(size_t)j * e->kp + i]) e->symmetric = false;
  build_code_table(e->sw.datatype, e->h_codes);
  // codes outside the matrix (e.g. a nucleic-acid encoding with an amino-acid
  // matrix) would index past the table: clamp them to the last real code.
  for (int i = 0; i < 256; ++i)
    if (e->h_codes[i] >= dim) e->h_codes[i] = (uint8_t)(dim - 1);
  HIP_TRY(hipMalloc(&e->d_table, e->h_table.size() * sizeof(int16_t)));
  HIP_TRY(hipMemcpy(e->d_table, e->h_table.data(), e->h_table.size() * sizeof(int16_t),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&e->d_codetab, 256));
  HIP_TRY(hipMemcpy(e->d_codetab, e->h_codes, 256, hipMemcpyHostToDevice));
  uint8_t wiretab[256];
  for (int i = 0; i < 256; ++i) wiretab[i] = (uint8_t)(i < dim ? i : dim - 1);
  HIP_TRY(hipMalloc(&e->d_wiretab, 256));
  HIP_TRY(hipMemcpy(e->d_wiretab, wiretab, 256, hipMemcpyHostToDevice));
  HIP_TRY(hipEventCreateWithFlags(&e->order_ev, hipEventDisableTiming));
  na_setup(e->device);
  const int o = -(e->sw.gap_init + e->sw.gap_extend), x = -e->sw.gap_extend;
  e->packed_ok = o >= 0 && x >= 0 && o <= 32767 && x <= 32767;
  e->forward_only = e->algo.forward_only ? 1 : 0;
  {
    const char *lg = std::getenv("IPUMA_LOG");
    e->log_batches = lg && *lg && std::strcmp(lg, "0") != 0;
  }
  if (e->log_batches) {
    for (auto &ev : e->ev) HIP_TRY(hipEventCreate(&ev));
    e->timing = 1;
  }
  e->fast_blocks = e->num_cu * fast_blocks_per_cu(e->kp);
  e->generic_blocks = std::max(8, 4 * e->num_cu);  // one wave each
  update_target_slots(e);
  return IPUMA_OK;
}

int ensure_chain(ChainBufs &c, int32_t n) {
  int rc;
  const size_t nn = (size_t)std::max(n, 1);
  const size_t max_groups = nn + kRuns;  // at most one part-filled group per run
  if ((rc = c.bin_of.ensure(nn * 4)) || (rc = c.rank.ensure(nn * 4)) ||
      (rc = c.grank.ensure(max_groups * 4)) || (rc = c.hist.ensure((size_t)kNumBins * 4)) ||
      (rc = c.costhist.ensure((size_t)kCostBins * 4)) || (rc = c.ctl.ensure(sizeof(BucketCtl))) ||
      (rc = c.sorted.ensure((nn + kSortedPad) * 4)) || (rc = c.gtmp.ensure(max_groups * sizeof(GroupDesc))) ||
      (rc = c.gbin.ensure(max_groups * 4)) || (rc = c.groups.ensure(max_groups * sizeof(GroupDesc))) ||
      (rc = c.ovf.ensure(nn * 4)))
    return rc;
  return IPUMA_OK;
}

// Device buffers of one forward chain, wherever they live (the engine's
// grow-only buffers, or the static array of the allocation-free entry).
struct ChainPtrs {
  int32_t *bin_of, *rank, *hist, *costhist, *sorted, *gbin, *grank, *ovf;
  BucketCtl *ctl;
  GroupDesc *gtmp, *groups;
  uint32_t *wave, *band;  // wave_words per wave / band_words per band slot
  int32_t *work;          // int32 kernel: work_stride ints per wave
  int64_t work_stride;
  int64_t wave_words, band_words;
  int32_t max_a;          // longest A accepted
  int fast_blocks, generic_blocks;
  int64_t target_slots;
};

// One forward chain on given buffers: bucketing -> packed kernel -> int32 re-run.
int run_chain_on(ipuma_engine *e, const ChainPtrs &c, int32_t n, const uint8_t *codes,
                 const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
                 const int32_t *b_len, int32_t *scores, int32_t *a_range, int32_t *b_range,
                 hipStream_t st, const int32_t *target, bool time_fast) {
  const bool packed = e->packed_ok && e->precision == IPUMA_PREC_AUTO;
  BucketCtl *ctl = c.ctl;
  if (!packed) {
    HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t *>(ctl), sizeof(BucketCtl) / 4, 0, st));
    // every comparison through the exact int32 kernel
    GenericArgs G{};
    G.codes = codes; G.a_off = a_off; G.a_len = a_len; G.b_off = b_off; G.b_len = b_len;
    G.table = e->d_table; G.kp = e->kp; G.list = nullptr; G.count = nullptr; G.n = n;
    G.work = c.work; G.work_stride = c.work_stride; G.max_a = c.max_a;
    G.gap_init = e->sw.gap_init; G.gap_extend = e->sw.gap_extend;
    G.scores = scores; G.a_range = a_range; G.b_range = b_range; G.error = &ctl->error;
    HIP_TRY(launch_generic(G, c.generic_blocks, st));
    return IPUMA_OK;
  }
  {  // control block, bin and cost histograms cleared, sorted slots empty
    int32_t *const fp[4] = {reinterpret_cast<int32_t *>(ctl), c.hist, c.costhist, c.sorted};
    const int64_t fn[4] = {(int64_t)(sizeof(BucketCtl) / 4), kNumBins, kCostBins,
                           (int64_t)n + kSortedPad};
    const int32_t fv[4] = {0, 0, 0, -1};
    HIP_TRY(launch_fill_ranges(fp, fn, fv, st));
  }
  BucketArgs B{};
  B.n = n; B.a_len = a_len; B.b_len = b_len; B.target_slots = c.target_slots;
  B.force_class = e->force_class;
  B.wave_cols = (int32_t)wave_cols_of(c.wave_words); B.max_a = c.max_a;
  B.plain = e->batching == IPUMA_BATCH_PLAIN ? 1 : e->batching == IPUMA_BATCH_PLAIN_TILE ? 2 : 0;
  B.max_groups = n + kRuns; B.bin_of = c.bin_of; B.rank = c.rank;
  B.hist = c.hist; B.grank = c.grank;
  B.costhist = c.costhist; B.ctl = ctl; B.sorted = c.sorted;
  B.gtmp = c.gtmp; B.gbin = c.gbin; B.groups = c.groups;
  B.scores = scores; B.a_range = a_range; B.b_range = b_range; B.target = target;
  HIP_TRY(launch_bucketing(B, st));
  FastArgs A{};
  A.codes = codes; A.a_off = a_off; A.a_len = a_len; A.b_off = b_off; A.b_len = b_len;
  A.table = e->d_table; A.sorted = c.sorted; A.groups = c.groups;
  A.ctl = ctl; A.scratch = c.wave; A.band_scratch = c.band; A.wave_words = c.wave_words;
  A.band_words = c.band_words;
  A.scores = scores; A.a_range = a_range; A.b_range = b_range;
  A.ovf_list = c.ovf; A.ovf_count = &ctl->ovf_count;
  A.max_groups = n + kRuns;
  A.target = target;
  const uint32_t o = (uint32_t)(-(e->sw.gap_init + e->sw.gap_extend)) & 0xFFFFu;
  const uint32_t x = (uint32_t)(-e->sw.gap_extend) & 0xFFFFu;
  A.gap_o = o | (o << 16);
  A.gap_x = x | (x << 16);
  A.smax = e->smax;
  if (time_fast) HIP_TRY(hipEventRecord(e->ev[1], st));
  HIP_TRY(launch_fast(e->kp, A, c.fast_blocks, st));
  if (time_fast) HIP_TRY(hipEventRecord(e->ev[2], st));
  GenericArgs G{};
  G.codes = codes; G.a_off = a_off; G.a_len = a_len; G.b_off = b_off; G.b_len = b_len;
  G.table = e->d_table; G.kp = e->kp; G.list = c.ovf; G.count = &ctl->ovf_count;
  G.n = n; G.work = c.work; G.work_stride = c.work_stride; G.max_a = c.max_a;
  G.gap_init = e->sw.gap_init;
  G.gap_extend = e->sw.gap_extend; G.scores = scores; G.a_range = a_range; G.b_range = b_range;
  G.error = &ctl->error;
  HIP_TRY(launch_generic(G, c.generic_blocks, st));
  return IPUMA_OK;
}

// One forward chain on the engine's grow-only buffers.
int run_chain(ipuma_engine *e, ChainBufs &c, int32_t n, const uint8_t *codes,
              const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
              const int32_t *b_len, int32_t *scores, int32_t *a_range, int32_t *b_range,
              hipStream_t st, const int32_t *target) {
  int rc = ensure_chain(c, n);
  if (rc) return rc;
  if ((rc = e->work.ensure((size_t)e->generic_blocks * kWorkStride * 4))) return rc;
  const bool packed = e->packed_ok && e->precision == IPUMA_PREC_AUTO;
  if (packed) {
    const size_t scratch_bytes =
        (size_t)e->fast_blocks * (kFastThreads / 64) * kWaveScratchWords * 4;
    if ((rc = e->scratch.ensure(scratch_bytes))) return rc;
    if ((rc = e->band_scratch.ensure((size_t)e->fast_blocks * kBandSlots * kBandScratchWords * 4)))
      return rc;
  }
  ChainPtrs p{};
  p.bin_of = c.bin_of.as<int32_t>(); p.rank = c.rank.as<int32_t>(); p.hist = c.hist.as<int32_t>();
  p.costhist = c.costhist.as<int32_t>(); p.sorted = c.sorted.as<int32_t>();
  p.gbin = c.gbin.as<int32_t>(); p.grank = c.grank.as<int32_t>(); p.ovf = c.ovf.as<int32_t>();
  p.ctl = c.ctl.as<BucketCtl>(); p.gtmp = c.gtmp.as<GroupDesc>(); p.groups = c.groups.as<GroupDesc>();
  p.wave = e->scratch.as<uint32_t>(); p.band = e->band_scratch.as<uint32_t>();
  p.work = e->work.as<int32_t>(); p.work_stride = kWorkStride;
  p.wave_words = kWaveScratchWords; p.band_words = kBandScratchWords; p.max_a = kMaxSeqLen;
  p.fast_blocks = e->fast_blocks; p.generic_blocks = e->generic_blocks;
  p.target_slots = e->target_slots;
  return run_chain_on(e, p, n, codes, a_off, a_len, b_off, b_len, scores, a_range, b_range, st,
                      target, e->timing && &c == &e->fwd);
}

// Device work on `st` after the engine's previous work on any stream (its
// buffers are shared); skipped while `st` is being captured into a graph.
bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}
int begin_work(ipuma_engine *e, hipStream_t st) {
  if (!capturing(st)) HIP_TRY(hipStreamWaitEvent(st, e->order_ev, 0));
  return IPUMA_OK;
}
int end_work(ipuma_engine *e, hipStream_t st) {
  if (!capturing(st)) HIP_TRY(hipEventRecord(e->order_ev, st));
  return IPUMA_OK;
}

int compare_device_chain(ipuma_engine *e, int32_t n, const uint8_t *codes, int64_t codes_len,
                         const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
                         const int32_t *b_len, int32_t *scores, int32_t *a_range,
                         int32_t *b_range, hipStream_t st);

int compare_device_impl(ipuma_engine *e, int32_t n, const uint8_t *codes, int64_t codes_len,
                        const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
                        const int32_t *b_len, int32_t *scores, int32_t *a_range,
                        int32_t *b_range, hipStream_t st) {
  int rc = begin_work(e, st);
  if (rc) return rc;
  rc = compare_device_chain(e, n, codes, codes_len, a_off, a_len, b_off, b_len, scores, a_range,
                            b_range, st);
  if (rc) return rc;
  return end_work(e, st);
}

int compare_device_chain(ipuma_engine *e, int32_t n, const uint8_t *codes, int64_t codes_len,
                         const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
                         const int32_t *b_len, int32_t *scores, int32_t *a_range,
                         int32_t *b_range, hipStream_t st) {
  if (e->timing) HIP_TRY(hipEventRecord(e->ev[0], st));
  e->ev_recorded = false;
  e->log_pending = e->log_batches;
  e->na_ctl = nullptr;  // errors of this batch are in the chains' control blocks
  int rc = run_chain(e, e->fwd, n, codes, a_off, a_len, b_off, b_len, scores, a_range, b_range, st,
                     nullptr);
  if (rc) return rc;
  e->last_ctl = e->fwd.ctl.as<BucketCtl>();
  e->last_n = n;
  if (e->forward_only || n == 0) {
    if (e->timing) {
      HIP_TRY(hipEventRecord(e->ev[3], st));
      e->ev_recorded = e->packed_ok && e->precision == IPUMA_PREC_AUTO;
    }
    return IPUMA_OK;
  }
  // start positions: forward pass over the reversed prefixes
  if ((rc = e->rcodes.ensure((size_t)codes_len + 16)) || (rc = e->rlen.ensure((size_t)n * 8)) ||
      (rc = e->rres.ensure((size_t)n * 12)))
    return rc;
  int32_t *ra_len = e->rlen.as<int32_t>(), *rb_len = ra_len + n;
  int32_t *rs = e->rres.as<int32_t>(), *rar = rs + n, *rbr = rs + 2 * (size_t)n;
  HIP_TRY(launch_reverse_prefix(n, codes, a_off, a_len, b_off, b_len, scores, a_range, b_range,
                                e->rcodes.as<uint8_t>(), ra_len, rb_len,
                                &e->fwd.ctl.as<BucketCtl>()->error, st));
  // the forward scores are the targets: a group stops once every pair reached its own
  rc = run_chain(e, e->rev, n, e->rcodes.as<uint8_t>(), a_off, ra_len, b_off, rb_len, rs, rar, rbr,
                 st, scores);
  if (rc) return rc;
  HIP_TRY(launch_finish_starts(n, scores, rar, rbr, a_range, b_range, st));
  if (e->timing) {
    HIP_TRY(hipEventRecord(e->ev[3], st));
    e->ev_recorded = e->packed_ok && e->precision == IPUMA_PREC_AUTO;
  }
  return IPUMA_OK;
}

// Host strings -> 4-aligned device arena of codes (encoded on the device).
// Grows a pinned host buffer to hold `need` bytes (in 1 MB steps); its old
// contents are not kept.  The caller has synchronized the stream that reads it.
int ensure_pinned(char *&p, size_t &cap, size_t need) {
  if (need <= cap) return IPUMA_OK;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  const size_t c = (need + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
  HIP_TRY(hipHostMalloc(&p, c, hipHostMallocDefault));
  cap = c;
  return IPUMA_OK;
}

int stage_arena(ipuma_engine *e, const std::vector<const char *> &ptrs,
                const std::vector<int64_t> &lens, std::vector<int64_t> &offs, int64_t &total) {
  offs.resize(ptrs.size());
  int64_t pos = 0;
  for (size_t i = 0; i < ptrs.size(); ++i) {
    offs[i] = pos;
    pos += (lens[i] + 3) / 4 * 4;
  }
  total = pos + 16;
  int rc;
  if ((rc = e->in_chars.ensure((size_t)total)) || (rc = e->in_codes.ensure((size_t)total))) return rc;
  // The characters go through a pinned staging buffer (one DMA at full link
  // speed instead of the runtime's pageable path); the previous call's copy
  // out of it has finished before it is overwritten.
  HIP_TRY(hipStreamSynchronize(e->stream));
  if ((rc = ensure_pinned(e->h_pin, e->h_pin_cap, (size_t)total))) return rc;
  char *const dst = e->h_pin;
  // sequences at 4-byte aligned offsets, the bytes between them and the tail zero
  auto copy_range = [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; ++i) {
      if (lens[i]) std::memcpy(dst + offs[i], ptrs[i], (size_t)lens[i]);
      const int64_t end = (i + 1 < ptrs.size()) ? offs[i + 1] : pos;
      if (end > offs[i] + lens[i]) std::memset(dst + offs[i] + lens[i], 0, (size_t)(end - offs[i] - lens[i]));
    }
  };
  const size_t np = ptrs.size();
  std::memset(dst + pos, 0, 16);
  if (total <= ((int64_t)8 << 20)) {
    copy_range(0, np);
    HIP_TRY(hipMemcpyAsync(e->in_chars.p, dst, (size_t)total, hipMemcpyHostToDevice, e->stream));
  } else {
    // large batches: eight slices, each gathered by up to eight threads (this one
    // among them) and sent as soon as all have finished it, while the threads go
    // on to the next slice
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nth = std::max<size_t>(1, std::min<size_t>(std::min(8u, hw), np));
    constexpr size_t kSlices = 8;
    std::atomic<size_t> done[kSlices];
    for (auto &d : done) d.store(0, std::memory_order_relaxed);
    auto part = [&](size_t t, size_t sl) {
      const size_t i0 = np * sl / kSlices, i1 = np * (sl + 1) / kSlices;
      copy_range(i0 + (i1 - i0) * t / nth, i0 + (i1 - i0) * (t + 1) / nth);
      done[sl].fetch_add(1, std::memory_order_release);
    };
    // A worker that cannot start (std::system_error) leaves its parts to this
    // thread; the started ones are always joined before returning.
    std::vector<std::thread> th;
    th.reserve(nth);
    size_t started = 1;
    try {
      for (size_t t = 1; t < nth; ++t, ++started)
        th.push_back(start_thread([&part, t] {
          for (size_t sl = 0; sl < kSlices; ++sl) part(t, sl);
        }));
    } catch (const std::system_error &) {
    }
    hipError_t err = hipSuccess;
    for (size_t sl = 0; sl < kSlices; ++sl) {
      part(0, sl);
      for (size_t t = started; t < nth; ++t) part(t, sl);
      while (done[sl].load(std::memory_order_acquire) < nth) std::this_thread::yield();
      const size_t i0 = np * sl / kSlices, i1 = np * (sl + 1) / kSlices;
      if (i0 == i1 || err != hipSuccess) continue;
      const int64_t b0 = offs[i0];
      const int64_t b1 = sl + 1 == kSlices ? total : offs[i1];
      err = hipMemcpyAsync(e->in_chars.as<char>() + b0, dst + b0, (size_t)(b1 - b0),
                           hipMemcpyHostToDevice, e->stream);
    }
    for (auto &x : th) x.join();
    HIP_TRY(err);
  }
  HIP_TRY(launch_encode(e->in_chars.as<uint8_t>(), total, e->in_codes.as<uint8_t>(), e->d_codetab,
                        e->stream));
  return IPUMA_OK;
}

int error_from_bits(int32_t err);

// Error bits the kernels raise in the chains' control blocks.
int check_device_error(ipuma_engine *e) {
  int32_t err = 0;
  // the control blocks of the last device batch: the allocation-free entry's,
  // or the forward and reverse chains'
  BucketCtl *const chain_ctl[2] = {e->fwd.ctl.as<BucketCtl>(), e->rev.ctl.as<BucketCtl>()};
  BucketCtl *const na_ctl[2] = {e->na_ctl, nullptr};
  BucketCtl *const *const ctls = e->na_ctl ? na_ctl : chain_ctl;
  for (int i = 0; i < 2; ++i) {
    BucketCtl *const c = ctls[i];
    if (!c) continue;
    int32_t v = 0;
    HIP_TRY(hipMemcpy(&v, &c->error, 4, hipMemcpyDeviceToHost));
    err |= v;
  }
  return error_from_bits(err);
}

// The status an entry reports for the kernels' error bits.
int error_from_bits(int32_t err) {
  if (err & 1)
    return fail(IPUMA_ETOOLONG,
                "a sequence is longer than %d (|A| longer than %d through "
                "ipuma_compare_device_noalloc)", kMaxSeqLen, kNaMaxA);
  if (err & 2) return fail(IPUMA_EDEVICE, "internal: an end position outside its sequence");
  if (err & 4) return fail(IPUMA_EDEVICE, "internal: a work-group descriptor out of range");
  if (err & 8) return fail(IPUMA_EDEVICE, "internal: no free band slot");
  if (err & 16) return fail(IPUMA_EDEVICE, "internal: an int32 boundary-row wait timed out");
  return IPUMA_OK;
}

int run_host_batch(ipuma_engine *e, int32_t n, int64_t codes_len, int32_t *scores,
                   int32_t *a_range, int32_t *b_range) {
  int rc;
  const size_t meta_bytes = (size_t)n * (8 + 4 + 8 + 4);
  if ((rc = e->in_meta.ensure(meta_bytes)) || (rc = e->out_res.ensure((size_t)n * 12 + 16)))
    return rc;
  char *m = e->in_meta.as<char>();
  int64_t *d_aoff = reinterpret_cast<int64_t *>(m);
  int64_t *d_boff = d_aoff + n;
  int32_t *d_alen = reinterpret_cast<int32_t *>(d_boff + n);
  int32_t *d_blen = d_alen + n;
  // descriptors in and results out go through one pinned buffer (one DMA each
  // way); the stream was synchronized when the previous call returned
  const size_t res_bytes = (size_t)n * 12;
  if ((rc = ensure_pinned(e->h_meta, e->h_meta_cap, std::max(meta_bytes, res_bytes) + 16)))
    return rc;
  if (n > 0) {
    char *h = e->h_meta;
    std::memcpy(h, e->h_aoff.data(), (size_t)n * 8);
    std::memcpy(h + (size_t)n * 8, e->h_boff.data(), (size_t)n * 8);
    std::memcpy(h + (size_t)n * 16, e->h_alen.data(), (size_t)n * 4);
    std::memcpy(h + (size_t)n * 20, e->h_blen.data(), (size_t)n * 4);
    HIP_TRY(hipMemcpyAsync(m, h, meta_bytes, hipMemcpyHostToDevice, e->stream));
  }
  int32_t *d_s = e->out_res.as<int32_t>(), *d_ar = d_s + n, *d_br = d_s + 2 * (size_t)n;
  rc = compare_device_impl(e, n, e->in_codes.as<uint8_t>(), codes_len, d_aoff, d_alen, d_boff,
                           d_blen, d_s, d_ar, d_br, e->stream);
  if (rc) {
    (void)hipStreamSynchronize(e->stream);  // the descriptor copy may still read h_meta
    return rc;
  }
  if (n > 0) HIP_TRY(hipMemcpyAsync(e->h_meta, d_s, res_bytes, hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (n > 0) {
    const int32_t *r = reinterpret_cast<const int32_t *>(e->h_meta);
    std::memcpy(scores, r, (size_t)n * 4);
    std::memcpy(a_range, r + n, (size_t)n * 4);
    std::memcpy(b_range, r + 2 * (size_t)n, (size_t)n * 4);
  }
  return check_device_error(e);
}

int check_sizes(ipuma_engine *e, int32_t n, const int64_t *a_lens_or_null,
                const std::vector<int64_t> &a_lens) {
  (void)a_lens_or_null;
  const int64_t cap = (int64_t)e->algo.tiles_used * e->algo.max_batches;
  if (cap > 0 && n > cap)
    return fail(IPUMA_EINVAL,
                "Input sequence (A) is over the max length: %d comparisons > tilesUsed*maxBatches = %lld",
                n, (long long)cap);
  for (int64_t l : a_lens)
    if (e->algo.max_ab > 0 && l > e->algo.max_ab)
      return fail(IPUMA_ETOOLONG, "Sequence size in seq %lld > maxAB %d", (long long)l,
                  e->algo.max_ab);
  return IPUMA_OK;
}

// A prepared input buffer (the reference's wire format, laid out for the
// engine's algo config and layout lot): the metadata is checked on the host,
// the buffer is copied to HBM as is, the residues of every slot are gathered
// to 4-aligned places, and the slots run as one device batch (empty slots
// score 0).  The results buffer receives the first result_words words of
// [scores | a ranges | b ranges], one word per slot.
int prepared_impl(ipuma_engine *e, const int32_t *in, int64_t in_words, int32_t *res,
                  int64_t res_words, int lot) {
  const auto t_entry = std::chrono::steady_clock::now();
  const ipuma_algo_config &c = e->algo;
  const int lots = c.transmission_programs;
  if (c.tiles_used < 0 || c.max_batches < 0 || c.bufsize < 0 || lots < 1)
    return fail(IPUMA_ECONFIG, "tiles_used, max_batches, bufsize >= 0 and transmission_programs >= 1 required");
  if (lot < 0 || lot >= lots) return fail(IPUMA_EINVAL, "layout lot %d outside 0..%d", lot, lots - 1);
  if (in_words < 0 || res_words < 0 || (in_words > 0 && !in) || (res_words > 0 && !res))
    return fail(IPUMA_EINVAL, "bad buffers");
  auto scale = [&](int64_t x) { return (x / lots + (x % lots == 0 ? 0 : 1)) * (int64_t)(lot + 1); };
  const int64_t ibuf = ((int64_t)c.bufsize + 3) / 4;
  const int64_t tile_bytes = scale(4 * ibuf);
  const int64_t meta_off = scale(ibuf) * c.tiles_used;
  const int64_t slots = (int64_t)c.tiles_used * c.max_batches;
  if (slots > INT32_MAX - kSortedPad) return fail(IPUMA_EINVAL, "too many slots");
  if (in_words < meta_off + 4 * slots)
    return fail(IPUMA_EINVAL, "input buffer of %lld words is shorter than the layout's %lld",
                (long long)in_words, (long long)(meta_off + 4 * slots));
  if (res_words > 3 * slots)
    return fail(IPUMA_EINVAL, "results buffer of %lld words exceeds 3 x %lld slots",
                (long long)res_words, (long long)slots);
  const int32_t *meta = in + meta_off;
  e->h_src.resize(2 * slots);
  e->h_dst.resize(2 * slots);
  e->h_len.resize(2 * slots);
  int64_t pos = 0;
  for (int64_t q = 0; q < slots; ++q) {
    const int64_t tile = q / std::max(c.max_batches, 1);
    const int32_t la = meta[4 * q], oa = meta[4 * q + 1], lb = meta[4 * q + 2], ob = meta[4 * q + 3];
    const bool empty = la == 0 || lb == 0;
    if (!empty) {
      if (la < 0 || lb < 0) return fail(IPUMA_EINVAL, "slot %lld: negative length", (long long)q);
      if (la > kMaxSeqLen || lb > kMaxSeqLen)
        return fail(IPUMA_ETOOLONG, "slot %lld: a sequence longer than %d", (long long)q, kMaxSeqLen);
      if (oa < 0 || ob < 0 || (int64_t)oa + la > tile_bytes || (int64_t)ob + lb > tile_bytes)
        return fail(IPUMA_EINVAL, "slot %lld: a sequence outside its tile's %lld bytes",
                    (long long)q, (long long)tile_bytes);
    }
    const int32_t na = empty ? 0 : la, nb = empty ? 0 : lb;
    e->h_len[q] = na;
    e->h_len[slots + q] = nb;
    e->h_src[q] = tile * tile_bytes + (empty ? 0 : oa);
    e->h_src[slots + q] = tile * tile_bytes + (empty ? 0 : ob);
    e->h_dst[q] = pos;
    pos += (na + 3) / 4 * 4;
    e->h_dst[slots + q] = pos;
    pos += (nb + 3) / 4 * 4;
  }
  const int64_t arena = pos + 16;
  const size_t n2 = (size_t)(2 * slots);
  int rc;
  if ((rc = e->w_wire.ensure((size_t)in_words * 4)) || (rc = e->w_desc.ensure(n2 * 20)) ||
      (rc = e->w_arena.ensure((size_t)arena)) || (rc = e->w_res.ensure((size_t)slots * 12 + 16)))
    return rc;
  int64_t *d_src = e->w_desc.as<int64_t>(), *d_dst = d_src + n2;
  int32_t *d_len = reinterpret_cast<int32_t *>(d_dst + n2);
  int32_t *d_res = e->w_res.as<int32_t>();
  hipStream_t st = e->stream;
  if (!e->w_ev[0])
    for (auto &ev : e->w_ev) HIP_TRY(hipEventCreate(&ev));
  if ((rc = begin_work(e, st))) return rc;
  HIP_TRY(hipEventRecord(e->w_ev[0], st));
  if (slots > 0) {
    HIP_TRY(hipMemcpyAsync(e->w_wire.p, in, (size_t)in_words * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_src, e->h_src.data(), n2 * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_dst, e->h_dst.data(), n2 * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_len, e->h_len.data(), n2 * 4, hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipEventRecord(e->w_ev[1], st));
  if (slots > 0) {
    HIP_TRY(launch_wire_gather((int64_t)n2, e->w_wire.as<uint8_t>(), d_src, d_dst, d_len,
                               e->d_wiretab, e->w_arena.as<uint8_t>(), st));
  }
  rc = compare_device_chain(e, (int32_t)slots, e->w_arena.as<uint8_t>(), arena, d_dst, d_len,
                            d_dst + slots, d_len + slots, d_res, d_res + slots, d_res + 2 * slots,
                            st);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(e->w_ev[2], st));
  if ((rc = end_work(e, st))) return rc;
  if (res_words > 0)
    HIP_TRY(hipMemcpyAsync(res, d_res, (size_t)res_words * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  float h2d = 0.f, dev = 0.f;
  HIP_TRY(hipEventElapsedTime(&h2d, e->w_ev[0], e->w_ev[1]));
  HIP_TRY(hipEventElapsedTime(&dev, e->w_ev[1], e->w_ev[2]));
  e->w_h2d_ms = h2d;
  e->w_device_ms = dev;
  rc = check_device_error(e);
  log_batch(e, "prepared", since_ms(t_entry), rc);
  return rc;
}

// Executor of asynchronously submitted prepared batches, in submission order.
void executor_loop(ipuma_engine *e) {
  for (;;) {
    int64_t id;
    PreparedJob job;
    {
      std::unique_lock<std::mutex> lk(e->pj_mu);
      e->pj_cv.wait(lk, [&] { return e->pj_stop || !e->pj_queue.empty(); });
      if (e->pj_queue.empty()) return;  // stopping, nothing left
      id = e->pj_queue.front();
      e->pj_queue.pop_front();
      job = e->pj_jobs[id];
    }
    int rc;
    double run_ms = 0.0;
    {
      std::lock_guard<std::mutex> g(e->mu);
      const auto t0 = std::chrono::steady_clock::now();
      rc = prepared_impl(e, job.in, job.in_words, job.res, job.res_words, 0);
      run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    std::lock_guard<std::mutex> lk(e->pj_mu);
    PreparedJob &j = e->pj_jobs[id];
    j.rc = rc;
    if (rc) j.err = g_last_error;
    j.times.h2d_ms = e->w_h2d_ms;
    j.times.device_ms = e->w_device_ms;
    j.times.run_ms = run_ms;
    j.done = true;
    e->pj_cv.notify_all();
  }
}

}  // namespace

namespace ipuma {
std::atomic<int> g_thread_start_budget{-1};
const uint8_t *engine_code_table(const ipuma_engine *e) { return e->h_codes; }
hipStream_t engine_stream(const ipuma_engine *e) { return e->stream; }
int engine_error_status(int32_t bits) { return error_from_bits(bits); }
int engine_error_words(ipuma_engine *e, const int32_t **w0, const int32_t **w1) {
  std::lock_guard<std::mutex> lk(e->mu);
  if (e->na_ctl) {
    *w0 = &e->na_ctl->error;
    *w1 = nullptr;
  } else {
    *w0 = e->fwd.ctl.p ? &e->fwd.ctl.as<BucketCtl>()->error : nullptr;
    *w1 = (!e->forward_only && e->rev.ctl.p) ? &e->rev.ctl.as<BucketCtl>()->error : nullptr;
  }
  return IPUMA_OK;
}
int set_last_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
}  // namespace ipuma

extern "C" {

void ipuma_default_sw_config(ipuma_sw_config *c) {
  if (!c) return;
  c->gap_init = 0;
  c->gap_extend = -1;
  c->match_value = 1;
  c->mismatch_value = -1;
  c->ambiguity_value = -1;
  c->similarity = IPUMA_SIM_NA;
  c->datatype = IPUMA_DT_NA;
}

void ipuma_default_algo_config(ipuma_algo_config *c) {
  if (!c) return;
  c->tiles_used = 1;
  c->max_ab = 300;
  c->max_batches = 20;
  c->bufsize = 3000;
  c->vtype = IPUMA_VT_CPP;
  c->fill_algo = IPUMA_FILL_FIRST;
  c->forward_only = 0;
  c->use_remote_buffer = 0;
  c->transmission_programs = 1;
  c->io_tiles = 0;
}

const char *ipuma_last_error(void) { return g_last_error.c_str(); }

static int ipuma_engine_create_body(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                        ipuma_engine **out) {
  if (!out) return fail(IPUMA_EINVAL, "out is NULL");
  *out = nullptr;
  auto *e = new (std::nothrow) ipuma_engine();
  if (!e) return fail(IPUMA_ENOMEM, "host allocation failed");
  if (sw) e->sw = *sw; else ipuma_default_sw_config(&e->sw);
  if (algo) e->algo = *algo; else ipuma_default_algo_config(&e->algo);
  if (e->sw.similarity < 0 || e->sw.similarity > 3 || e->sw.datatype < 0 || e->sw.datatype > 2) {
    delete e;
    return fail(IPUMA_EINVAL, "invalid similarity/datatype");
  }
  const int rc = init_engine(e);
  if (rc) {
    ipuma_engine_destroy(e);
    return rc;
  }
  *out = e;
  return IPUMA_OK;
}

void ipuma_engine_destroy(ipuma_engine *e) {
  if (!e) return;
  if (e->pj_thread.joinable()) {  // queued prepared batches still run
    {
      std::lock_guard<std::mutex> lk(e->pj_mu);
      e->pj_stop = true;
    }
    e->pj_cv.notify_all();
    e->pj_thread.join();
  }
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->h_pin) (void)hipHostFree(e->h_pin);
  e->h_pin = nullptr;
  e->h_pin_cap = 0;
  if (e->h_meta) (void)hipHostFree(e->h_meta);
  e->h_meta = nullptr;
  e->h_meta_cap = 0;
  e->scratch.release();
  e->band_scratch.release();
  e->work.release();
  e->fwd.release();
  e->rev.release();
  e->rcodes.release();
  e->rlen.release();
  e->rres.release();
  e->s_pairs.release();
  e->s_res.release();
  e->in_chars.release();
  e->in_codes.release();
  e->in_meta.release();
  e->out_res.release();
  e->w_wire.release();
  e->w_desc.release();
  e->w_arena.release();
  e->w_res.release();
  e->t_desc.release();
  e->t_dir.release();
  e->t_work.release();
  e->t_cig.release();
  if (e->d_wiretab) (void)hipFree(e->d_wiretab);
  if (e->order_ev) (void)hipEventDestroy(e->order_ev);
  for (auto &ev : e->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : e->w_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : e->t_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->d_table) (void)hipFree(e->d_table);
  if (e->d_codetab) (void)hipFree(e->d_codetab);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

static int ipuma_compare_local_body(ipuma_engine *e, int32_t n, const char *a_chars, const int64_t *a_offsets,
                        const char *b_chars, const int64_t *b_offsets, int32_t *scores,
                        int32_t *a_range, int32_t *b_range) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  const auto t_entry = std::chrono::steady_clock::now();
  if (n < 0) return fail(IPUMA_EINVAL, "n < 0");
  if (n > 0 && (!a_offsets || !b_offsets || !scores || !a_range || !b_range))
    return fail(IPUMA_EINVAL, "NULL buffer");
  std::vector<const char *> ptrs(2 * (size_t)n);
  std::vector<int64_t> lens(2 * (size_t)n), alens((size_t)n);
  for (int32_t k = 0; k < n; ++k) {
    const int64_t la = a_offsets[k + 1] - a_offsets[k], lb = b_offsets[k + 1] - b_offsets[k];
    if (la < 0 || lb < 0) return fail(IPUMA_EINVAL, "offsets not ascending at %d", k);
    if ((la > 0 && !a_chars) || (lb > 0 && !b_chars))
      return fail(IPUMA_EINVAL, "character buffer is NULL");
    if (la > kMaxSeqLen || lb > kMaxSeqLen)
      return fail(IPUMA_ETOOLONG, "sequence %d longer than %d", k, kMaxSeqLen);
    ptrs[2 * k] = a_chars + a_offsets[k];
    ptrs[2 * k + 1] = b_chars + b_offsets[k];
    lens[2 * k] = la;
    lens[2 * k + 1] = lb;
    alens[k] = la;
  }
  int rc = check_sizes(e, n, nullptr, alens);
  if (rc) return rc;
  std::vector<int64_t> offs;
  int64_t total = 0;
  if ((rc = stage_arena(e, ptrs, lens, offs, total))) return rc;
  e->h_aoff.resize(n); e->h_boff.resize(n); e->h_alen.resize(n); e->h_blen.resize(n);
  for (int32_t k = 0; k < n; ++k) {
    e->h_aoff[k] = offs[2 * k];
    e->h_boff[k] = offs[2 * k + 1];
    e->h_alen[k] = (int32_t)lens[2 * k];
    e->h_blen[k] = (int32_t)lens[2 * k + 1];
  }
  rc = run_host_batch(e, n, total, scores, a_range, b_range);
  log_batch(e, "compare_local", since_ms(t_entry), rc);
  return rc;
}

static int ipuma_compare_mn_local_body(ipuma_engine *e, int32_t nseqs, const char *chars,
                           const int64_t *offsets, int32_t ncmps, const int32_t *cmps,
                           int32_t *scores, int32_t *a_range, int32_t *b_range) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  const auto t_entry = std::chrono::steady_clock::now();
  if (nseqs < 0 || ncmps < 0) return fail(IPUMA_EINVAL, "negative count");
  if (ncmps > 0 && (!cmps || !scores || !a_range || !b_range))
    return fail(IPUMA_EINVAL, "NULL buffer");
  if (nseqs > 0 && !offsets) return fail(IPUMA_EINVAL, "offsets is NULL");
  std::vector<const char *> ptrs((size_t)nseqs);
  std::vector<int64_t> lens((size_t)nseqs);
  for (int32_t s = 0; s < nseqs; ++s) {
    lens[s] = offsets[s + 1] - offsets[s];
    if (lens[s] < 0) return fail(IPUMA_EINVAL, "offsets not ascending at %d", s);
    if (lens[s] > 0 && !chars) return fail(IPUMA_EINVAL, "character buffer is NULL");
    if (lens[s] > kMaxSeqLen) return fail(IPUMA_ETOOLONG, "sequence %d longer than %d", s, kMaxSeqLen);
    ptrs[s] = chars + offsets[s];
  }
  std::vector<int64_t> alens((size_t)ncmps);
  for (int32_t k = 0; k < ncmps; ++k) {
    const int32_t ia = cmps[2 * k], ib = cmps[2 * k + 1];
    if (ia < 0 || ia >= nseqs || ib < 0 || ib >= nseqs)
      return fail(IPUMA_EINVAL, "comparison %d indexes outside the sequences", k);
    alens[k] = lens[ia];
  }
  int rc = check_sizes(e, ncmps, nullptr, alens);
  if (rc) return rc;
  std::vector<int64_t> offs;
  int64_t total = 0;
  if ((rc = stage_arena(e, ptrs, lens, offs, total))) return rc;
  e->h_aoff.resize(ncmps); e->h_boff.resize(ncmps); e->h_alen.resize(ncmps); e->h_blen.resize(ncmps);
  for (int32_t k = 0; k < ncmps; ++k) {
    const int32_t ia = cmps[2 * k], ib = cmps[2 * k + 1];
    e->h_aoff[k] = offs[ia];
    e->h_boff[k] = offs[ib];
    e->h_alen[k] = (int32_t)lens[ia];
    e->h_blen[k] = (int32_t)lens[ib];
  }
  rc = run_host_batch(e, ncmps, total, scores, a_range, b_range);
  log_batch(e, "compare_mn_local", since_ms(t_entry), rc);
  return rc;
}

static int ipuma_compare_device_body(ipuma_engine *e, int32_t n, const uint8_t *codes, int64_t codes_len,
                         const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
                         const int32_t *b_len, int32_t *scores, int32_t *a_range,
                         int32_t *b_range, void *stream) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (n < 0) return fail(IPUMA_EINVAL, "n < 0");
  if (n > 0 && (!codes || !a_off || !a_len || !b_off || !b_len || !scores || !a_range || !b_range))
    return fail(IPUMA_EINVAL, "NULL buffer");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : e->stream;
  return compare_device_impl(e, n, codes, codes_len, a_off, a_len, b_off, b_len, scores, a_range,
                             b_range, st);
}

// ipuma_compare_device without allocation: the forward chain on the static
// scratch of the engine's device.  Calls on one device are ordered through
// that device's event (outside graph capture); they are issued one at a time.
static int ipuma_compare_device_noalloc_body(ipuma_engine *e, int32_t n, const uint8_t *codes,
                                             int64_t codes_len, const int64_t *a_off,
                                             const int32_t *a_len, const int64_t *b_off,
                                             const int32_t *b_len, int32_t *scores,
                                             int32_t *a_range, int32_t *b_range, void *stream) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (n < 0) return fail(IPUMA_EINVAL, "n < 0");
  if (n > kNaPairs)
    return fail(IPUMA_EINVAL, "%d comparisons: the allocation-free entry takes at most %d", n,
                kNaPairs);
  if (n > 0 && (!codes || !a_off || !a_len || !b_off || !b_len || !scores || !a_range || !b_range))
    return fail(IPUMA_EINVAL, "NULL buffer");
  if (!e->forward_only)
    return fail(IPUMA_EINVAL, "the allocation-free entry computes end positions only: the "
                              "engine must be forward-only (start positions need "
                              "ipuma_compare_device)");
  if (e->device < 0 || e->device >= kNaMaxDevices)
    return fail(IPUMA_EDEVICE, "device %d has no static scratch", e->device);
  NaDevice &d = g_na[e->device];
  if (!d.ready) return fail(d.rc ? d.rc : IPUMA_EDEVICE, "%s", d.err.c_str());
  (void)codes_len;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : e->stream;
  std::lock_guard<std::mutex> dlk(d.mu);
  int rc = begin_work(e, st);
  if (rc) return rc;
  const bool cap = capturing(st);
  if (!cap) HIP_TRY(hipStreamWaitEvent(st, d.order, 0));
  uint8_t *const base = d.ptr;
  auto at = [&](int part) { return base + na_offset(part); };
  ChainPtrs p{};
  p.ctl = reinterpret_cast<BucketCtl *>(at(kNaCtl));
  p.hist = reinterpret_cast<int32_t *>(at(kNaHist));
  p.costhist = reinterpret_cast<int32_t *>(at(kNaCost));
  p.bin_of = reinterpret_cast<int32_t *>(at(kNaBin));
  p.rank = reinterpret_cast<int32_t *>(at(kNaRank));
  p.ovf = reinterpret_cast<int32_t *>(at(kNaOvf));
  p.gbin = reinterpret_cast<int32_t *>(at(kNaGbin));
  p.grank = reinterpret_cast<int32_t *>(at(kNaGrank));
  p.sorted = reinterpret_cast<int32_t *>(at(kNaSorted));
  p.gtmp = reinterpret_cast<GroupDesc *>(at(kNaGtmp));
  p.groups = reinterpret_cast<GroupDesc *>(at(kNaGroupsPart));
  p.wave = reinterpret_cast<uint32_t *>(at(kNaWave));
  p.band = reinterpret_cast<uint32_t *>(at(kNaBand));
  p.work = reinterpret_cast<int32_t *>(at(kNaWork));
  p.work_stride = kNaWorkStride;
  p.wave_words = kNaWaveWords;
  p.band_words = kNaBandWords;
  p.max_a = kNaMaxA;
  p.fast_blocks = std::min(e->fast_blocks, kNaBlocks);
  p.generic_blocks = kNaGenericBlocks;
  // the step cap follows the resident waves of this launch
  p.target_slots = std::max<int64_t>(1, e->target_slots * p.fast_blocks / std::max(1, e->fast_blocks));
  if (e->timing) HIP_TRY(hipEventRecord(e->ev[0], st));
  e->ev_recorded = false;
  e->log_pending = e->log_batches;
  rc = run_chain_on(e, p, n, codes, a_off, a_len, b_off, b_len, scores, a_range, b_range, st,
                    nullptr, e->timing != 0);
  if (rc) return rc;
  e->na_ctl = p.ctl;
  e->last_ctl = p.ctl;
  e->last_n = n;
  if (e->timing) {
    HIP_TRY(hipEventRecord(e->ev[3], st));
    e->ev_recorded = e->packed_ok && e->precision == IPUMA_PREC_AUTO;
  }
  if (!cap) HIP_TRY(hipEventRecord(d.order, st));
  return end_work(e, st);
}

static int ipuma_search_device_body(ipuma_engine *e, const uint8_t *codes, int64_t codes_len, int32_t nq,
                        const int64_t *q_off, const int32_t *q_len, int32_t ndb,
                        const int64_t *db_off, const int32_t *db_len, int32_t k,
                        int32_t *scores, int32_t *top_idx, int32_t *top_score, void *stream) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (nq < 0 || ndb < 0 || k < 0 || k > 1024)
    return fail(IPUMA_EINVAL, "nq, ndb >= 0 and 0 <= k <= 1024 required");
  const int64_t n = (int64_t)nq * ndb;
  if (n > INT32_MAX - kSortedPad) return fail(IPUMA_EINVAL, "nq * ndb too large");
  if (n > 0 && (!codes || !q_off || !q_len || !db_off || !db_len))
    return fail(IPUMA_EINVAL, "NULL buffer");
  if (k > 0 && (!top_idx || !top_score)) return fail(IPUMA_EINVAL, "NULL top-k buffer");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : e->stream;
  int rc;
  if ((rc = e->s_pairs.ensure((size_t)std::max<int64_t>(n, 1) * 24)) ||
      (rc = e->s_res.ensure((size_t)std::max<int64_t>(n, 1) * 12)))
    return rc;
  int64_t *a_off = e->s_pairs.as<int64_t>(), *b_off = a_off + n;
  int32_t *a_len = reinterpret_cast<int32_t *>(b_off + n), *b_len = a_len + n;
  int32_t *res = e->s_res.as<int32_t>();
  int32_t *sc = scores ? scores : res + 2 * n;
  if ((rc = begin_work(e, st))) return rc;
  // a symmetric table scores (query, sequence) as (sequence, query): the query
  // then runs over the columns, shared by both comparisons of a lane
  HIP_TRY(launch_search_pairs(nq, ndb, q_off, q_len, db_off, db_len, a_off, a_len, b_off, b_len,
                              e->symmetric ? 1 : 0, st));
  // scores only: the forward pass (its end positions are discarded)
  const int saved = e->forward_only;
  e->forward_only = 1;
  rc = compare_device_impl(e, (int32_t)n, codes, codes_len, a_off, a_len, b_off, b_len, sc, res,
                           res + n, st);
  e->forward_only = saved;
  if (rc) return rc;
  HIP_TRY(launch_topk(nq, ndb, k, sc, top_idx, top_score, st));
  return end_work(e, st);
}

int ipuma_set_batching(ipuma_engine *e, int32_t mode) {
  if (!e || (mode != IPUMA_BATCH_BUCKETED && mode != IPUMA_BATCH_PLAIN &&
             mode != IPUMA_BATCH_PLAIN_TILE))
    return fail(IPUMA_EINVAL, "invalid batching mode");
  std::lock_guard<std::mutex> lock_(e->mu);
  e->batching = mode;
  return IPUMA_OK;
}

int ipuma_set_precision(ipuma_engine *e, int32_t mode) {
  if (!e || (mode != IPUMA_PREC_AUTO && mode != IPUMA_PREC_INT32))
    return fail(IPUMA_EINVAL, "invalid precision mode");
  std::lock_guard<std::mutex> lock_(e->mu);
  e->precision = mode;
  return IPUMA_OK;
}

int ipuma_set_forward_only(ipuma_engine *e, int32_t forward_only) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  e->forward_only = forward_only ? 1 : 0;
  return IPUMA_OK;
}

int ipuma_last_traceback_stats(ipuma_engine *e, int64_t *region_cells, double *kernel_ms) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (region_cells) *region_cells = e->t_cells;
  if (kernel_ms) *kernel_ms = e->t_kernel_ms;
  return IPUMA_OK;
}

int ipuma_set_timing(ipuma_engine *e, int32_t on) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (on && !e->ev[0])
    for (auto &ev : e->ev) HIP_TRY(hipEventCreate(&ev));
  e->timing = on || e->log_batches ? 1 : 0;
  return IPUMA_OK;
}

int ipuma_set_work_rounds(ipuma_engine *e, int32_t rounds) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (rounds < 1 || rounds > 64) return fail(IPUMA_EINVAL, "rounds must be in 1..64");
  e->rounds = rounds;
  update_target_slots(e);
  return IPUMA_OK;
}

int ipuma_set_step_cap_scale(ipuma_engine *e, float scale) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (!(scale >= 0.25f && scale <= 4.0f)) return fail(IPUMA_EINVAL, "scale must be in 0.25..4");
  e->cap_milli = (int)(scale * 1000.0f + 0.5f);
  update_target_slots(e);
  return IPUMA_OK;
}

int ipuma_set_lanes_per_pair(ipuma_engine *e, int32_t lanes) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  if (lanes == 0) {
    e->force_class = -1;
    return IPUMA_OK;
  }
  for (int c = 0; c < kNumClasses; ++c)
    if (kClassL[c] == lanes) {
      e->force_class = c;
      return IPUMA_OK;
    }
  return fail(IPUMA_EINVAL, "lanes per pair must be 0, 1, 2, 4, 8, 16 or 64");
}

static int ipuma_encode_device_body(ipuma_engine *e, const char *d_chars, int64_t len, uint8_t *d_codes,
                        void *stream) {
  if (!e || len < 0 || (len > 0 && (!d_chars || !d_codes))) return fail(IPUMA_EINVAL, "bad args");
  std::lock_guard<std::mutex> lock_(e->mu);
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : e->stream;
  HIP_TRY(launch_encode(reinterpret_cast<const uint8_t *>(d_chars), len, d_codes, e->d_codetab, st));
  return IPUMA_OK;
}

static int stats_locked(ipuma_engine *e, ipuma_batch_stats *out);

static int ipuma_last_batch_stats_body(ipuma_engine *e, ipuma_batch_stats *out) {
  if (!e || !out) return fail(IPUMA_EINVAL, "bad args");
  std::lock_guard<std::mutex> lock_(e->mu);
  return stats_locked(e, out);
}

// Statistics of the last device batch (the engine mutex is held).
static int stats_locked(ipuma_engine *e, ipuma_batch_stats *out) {
  std::memset(out, 0, sizeof(*out));
  if (!e->last_ctl) return IPUMA_OK;
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipDeviceSynchronize());
  BucketCtl c;
  HIP_TRY(hipMemcpy(&c, e->last_ctl, sizeof(c), hipMemcpyDeviceToHost));
  out->pairs = e->last_n;
  out->cells = (int64_t)c.total_cells;
  out->padded_cells = (int64_t)c.padded_cells;
  out->groups = c.ngroups;
  out->rerun_pairs = c.ovf_count;
  int lmax = 1;
  for (int k = 0; k < kNumClasses; ++k)
    if (c.class_count[k] > 0) lmax = kClassL[k];
  out->lanes_per_pair_max = lmax;
  out->rows_per_lane = kRows;
  out->band_groups = c.nband;
  out->head_groups = c.nband > 0 ? std::min(c.nhead, c.ngroups - c.nband) : 0;
  if (e->ev_recorded) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev[1], e->ev[2]));
    out->fast_kernel_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev[0], e->ev[3]));
    out->batch_ms = ms;
  }
  return IPUMA_OK;
}

static int ipuma_prepared_compare_body(ipuma_engine *e, const int32_t *inputs, int64_t input_words,
                           int32_t *results, int64_t result_words, int32_t lot) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  return prepared_impl(e, inputs, input_words, results, result_words, lot);
}

static int ipuma_prepared_submit_body(ipuma_engine *e, const int32_t *inputs, int64_t input_words,
                          int32_t *results, int64_t result_words, int64_t *job) {
  if (!e || !job) return fail(IPUMA_EINVAL, "engine or job is NULL");
  if (input_words < 0 || result_words < 0 || (input_words > 0 && !inputs) ||
      (result_words > 0 && !results))
    return fail(IPUMA_EINVAL, "bad buffers");
  std::lock_guard<std::mutex> lk(e->pj_mu);
  if (e->pj_stop) return fail(IPUMA_ECLOSED, "engine is closing");
  if (!e->pj_thread.joinable()) e->pj_thread = start_thread([e] { executor_loop(e); });
  const int64_t id = e->pj_next++;
  PreparedJob &j = e->pj_jobs[id];
  j.in = inputs;
  j.in_words = input_words;
  j.res = results;
  j.res_words = result_words;
  e->pj_queue.push_back(id);
  e->pj_cv.notify_all();
  *job = id;
  return IPUMA_OK;
}

static int ipuma_prepared_join_body(ipuma_engine *e, int64_t job, ipuma_job_times *times) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::unique_lock<std::mutex> lk(e->pj_mu);
  auto it = e->pj_jobs.find(job);
  if (it == e->pj_jobs.end()) return fail(IPUMA_EINVAL, "unknown or already joined job %lld", (long long)job);
  e->pj_cv.wait(lk, [&] {
    auto f = e->pj_jobs.find(job);
    return f == e->pj_jobs.end() || f->second.done;
  });
  it = e->pj_jobs.find(job);
  if (it == e->pj_jobs.end()) return fail(IPUMA_EINVAL, "job %lld joined twice", (long long)job);
  const PreparedJob j = it->second;
  e->pj_jobs.erase(it);
  if (j.rc) return fail(j.rc, "%s", j.err.c_str());
  if (times) *times = j.times;
  return IPUMA_OK;
}

static int ipuma_traceback_local_body(ipuma_engine *e, int32_t n, const char *a_chars,
                          const int64_t *a_offsets, const char *b_chars, const int64_t *b_offsets,
                          const int32_t *scores, const int32_t *a_range, const int32_t *b_range,
                          uint32_t *cigar, const int64_t *cigar_offsets, int32_t *cigar_len) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  if (n < 0) return fail(IPUMA_EINVAL, "n < 0");
  if (n > 0 && (!a_offsets || !b_offsets || !scores || !a_range || !b_range || !cigar ||
                !cigar_offsets || !cigar_len))
    return fail(IPUMA_EINVAL, "NULL buffer");
  std::lock_guard<std::mutex> lock_(e->mu);
  // k_trace_wave (int32) when every region's values are bounded by 2^28: a
  // path of the region has at most m + n steps, each worth at most a gap
  // open, an extension and a score in magnitude (one pad cell adds 2^15)
  const int64_t go = -(int64_t)(e->sw.gap_init + e->sw.gap_extend), ge = -(int64_t)e->sw.gap_extend;
  int64_t sabs = 0;
  for (int x = 0; x + 1 < e->kp; ++x)
    for (int y = 0; y + 1 < e->kp; ++y)
      sabs = std::max<int64_t>(sabs, std::abs((int64_t)e->h_table[(size_t)x * e->kp + y]));
  const int64_t per_step = std::abs(go) + std::abs(ge) + sabs;
  bool wave = true;
  // regions, capacities, scratch offsets
  std::vector<const char *> ptrs(2 * (size_t)n);
  std::vector<int64_t> lens(2 * (size_t)n);
  std::vector<int64_t> dir_off(n), work_off(n), cig_off(n);
  int64_t need_cig = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int64_t la = a_offsets[k + 1] - a_offsets[k], lb = b_offsets[k + 1] - b_offsets[k];
    if (la < 0 || lb < 0) return fail(IPUMA_EINVAL, "offsets not ascending at %d", k);
    if ((la > 0 && !a_chars) || (lb > 0 && !b_chars))
      return fail(IPUMA_EINVAL, "character buffer is NULL");
    if (la > kMaxSeqLen || lb > kMaxSeqLen)
      return fail(IPUMA_ETOOLONG, "sequence %d longer than %d", k, kMaxSeqLen);
    ptrs[2 * k] = a_chars + a_offsets[k];
    ptrs[2 * k + 1] = b_chars + b_offsets[k];
    lens[2 * k] = la;
    lens[2 * k + 1] = lb;
    const uint32_t ar = (uint32_t)a_range[k], br = (uint32_t)b_range[k];
    const int64_t as = ar & 0xFFFF, ae = ar >> 16, bs = br & 0xFFFF, be = br >> 16;
    int64_t m = 0, nb = 0;
    if (scores[k] > 0) {
      if (ae < as || be < bs || ae >= la || be >= lb)
        return fail(IPUMA_EINVAL, "comparison %d: ranges outside its sequences", k);
      m = ae - as + 1;
      nb = be - bs + 1;
    }
    const int64_t cap = cigar_offsets[k + 1] - cigar_offsets[k];
    if (cap < m + nb)
      return fail(IPUMA_EINVAL, "comparison %d: CIGAR capacity %lld < %lld", k, (long long)cap,
                  (long long)(m + nb));
    if ((m + nb + 2) * per_step + 65536 > ((int64_t)1 << 28)) wave = false;
    dir_off[k] = m * nb;   // sizes for now; turned into offsets per chunk below
    work_off[k] = 2 * (nb + 1);
    cig_off[k] = m + nb;
    need_cig += m + nb;
  }
  (void)need_cig;
  e->t_cells = 0;
  e->t_kernel_ms = 0.0;
  for (int32_t k = 0; k < n; ++k) e->t_cells += dir_off[k];  // region cells (sizes not yet replaced)
  if (e->timing && !e->t_ev[0])
    for (auto &ev : e->t_ev) HIP_TRY(hipEventCreate(&ev));
  if (wave)  // a 32-bit word per lane and step: strips of 64 * R A rows x (|B region| + 63) steps x 256 B
    for (int32_t k = 0; k < n; ++k) {
      if (!dir_off[k]) continue;
      const int64_t cols = work_off[k] / 2 - 1;  // |B region|
      const int64_t rows = dir_off[k] / cols, strip = 64 * trace_rows(rows);
      dir_off[k] = (rows + strip - 1) / strip * (cols + 63) * 256;
    }
  std::vector<int64_t> offs;
  int64_t total = 0;
  int rc;
  if ((rc = stage_arena(e, ptrs, lens, offs, total))) return rc;
  // chunks of at most 8 GiB of direction storage (one pair may exceed it
  // alone): one launch for most calls, so that the longest regions, which
  // start first, overlap every other comparison of the call
  const int64_t kChunkDir = (int64_t)8 << 30;
  std::vector<int64_t> h_desc;
  for (int32_t c0 = 0; c0 < n;) {
    int32_t c1 = c0;
    int64_t dsum = 0, wsum = 0, csum = 0;
    while (c1 < n && (c1 == c0 || dsum + dir_off[c1] <= kChunkDir)) {
      dsum += dir_off[c1];
      wsum += work_off[c1];
      csum += cig_off[c1];
      ++c1;
    }
    const int32_t cn = c1 - c0;
    // descriptors: a_off, b_off, dir_off, work_off, cig_off (int64) + ranges, scores (int32)
    h_desc.assign((size_t)cn * 5, 0);
    int64_t dpos = 0, wpos = 0, cpos = 0;
    for (int32_t k = 0; k < cn; ++k) {
      h_desc[k] = offs[2 * (c0 + k)];
      h_desc[cn + k] = offs[2 * (c0 + k) + 1];
      h_desc[2 * cn + k] = dpos;
      h_desc[3 * cn + k] = wpos;
      h_desc[4 * cn + k] = cpos;
      dpos += dir_off[c0 + k];
      wpos += work_off[c0 + k];
      cpos += cig_off[c0 + k];
    }
    // k_trace_wave: largest regions first (one wave each)
    std::vector<int32_t> order(cn);
    for (int32_t k = 0; k < cn; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
      return dir_off[c0 + x] > dir_off[c0 + y];
    });
    if ((rc = e->t_desc.ensure((size_t)cn * (5 * 8 + 4 * 4 + 4))) ||
        (rc = e->t_dir.ensure((size_t)std::max<int64_t>(dsum, 1))) ||
        (rc = e->t_work.ensure((size_t)std::max<int64_t>(wsum, 1) * 8)) ||
        (rc = e->t_cig.ensure((size_t)std::max<int64_t>(csum, 1) * 4)))
      return rc;
    int64_t *d64 = e->t_desc.as<int64_t>();
    int32_t *d32 = reinterpret_cast<int32_t *>(d64 + 5 * (size_t)cn);
    HIP_TRY(hipMemcpyAsync(d64, h_desc.data(), (size_t)cn * 40, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(d32, a_range + c0, (size_t)cn * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(d32 + cn, b_range + c0, (size_t)cn * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(d32 + 2 * (size_t)cn, scores + c0, (size_t)cn * 4, hipMemcpyHostToDevice,
                           e->stream));
    HIP_TRY(hipMemcpyAsync(d32 + 4 * (size_t)cn, order.data(), (size_t)cn * 4, hipMemcpyHostToDevice,
                           e->stream));
    TraceArgs T{};
    T.codes = e->in_codes.as<uint8_t>();
    T.a_off = d64; T.b_off = d64 + cn; T.dir_off = d64 + 2 * (size_t)cn;
    T.work_off = d64 + 3 * (size_t)cn; T.cig_off = d64 + 4 * (size_t)cn;
    T.a_range = d32; T.b_range = d32 + cn; T.scores = d32 + 2 * (size_t)cn;
    T.cig_len = d32 + 3 * (size_t)cn;
    T.order = d32 + 4 * (size_t)cn;
    T.wave = wave ? 1 : 0;
    T.n = cn;
    T.gap_o = -(e->sw.gap_init + e->sw.gap_extend);
    T.gap_x = -e->sw.gap_extend;
    T.kp = e->kp; T.table = e->d_table;
    T.dir = e->t_dir.as<uint8_t>(); T.work = e->t_work.as<int64_t>();
    T.cigar = e->t_cig.as<uint32_t>();
    if ((rc = begin_work(e, e->stream))) return rc;
    if (e->timing) HIP_TRY(hipEventRecord(e->t_ev[0], e->stream));
    HIP_TRY(launch_traceback(T, e->stream));
    if (e->timing) HIP_TRY(hipEventRecord(e->t_ev[1], e->stream));
    if ((rc = end_work(e, e->stream))) return rc;
    // runs and counts back through the pinned staging buffer of results (one
    // DMA each at full link rate; the runs take each comparison's whole capacity)
    if ((rc = ensure_pinned(e->h_meta, e->h_meta_cap, (size_t)(csum + cn) * 4 + 16))) return rc;
    uint32_t *const h_cig = reinterpret_cast<uint32_t *>(e->h_meta);
    int32_t *const h_len = reinterpret_cast<int32_t *>(h_cig + csum);
    HIP_TRY(hipMemcpyAsync(h_cig, T.cigar, (size_t)csum * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(h_len, T.cig_len, (size_t)cn * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->timing) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, e->t_ev[0], e->t_ev[1]));
      e->t_kernel_ms += ms;
    }
    for (int32_t k = 0; k < cn; ++k) {
      const int32_t r = h_len[k];
      cigar_len[c0 + k] = r;
      const uint32_t *src = h_cig + h_desc[4 * cn + k];
      uint32_t *dst = cigar + cigar_offsets[c0 + k];
      for (int32_t q = 0; q < r; ++q) dst[q] = src[r - 1 - q];  // written end to start
    }
    c0 = c1;
  }
  for (int32_t k = 0; k < n; ++k)
    if (cigar_len[k] == -2)
      return fail(IPUMA_EINVAL, "comparison %d: the ranges do not give the reported score", k);
  return IPUMA_OK;
}

// One line per device batch when IPUMA_LOG is set: the batch_perf record of
// the reference's batch driver (ipusw/ipumultidriver.hpp:103-121) with this
// engine's counters -- comparisons, cells, padded cells, groups, int32
// re-runs, the packed kernel's and the whole device batch's times (HIP
// events), and the entry point's host time when it has one -- written to
// stderr behind the tag IPUSWLOG (src/types.h:6).  The engine mutex is held.
static void log_batch(ipuma_engine *e, const char *entry, double host_ms, int status) {
  if (!e->log_pending) return;
  e->log_pending = false;
  ipuma_batch_stats st;
  if (stats_locked(e, &st) != IPUMA_OK) return;
  char line[768];
  const double cells = (double)st.cells;
  int k = std::snprintf(line, sizeof(line),
                        "IPUSWLOG{\"tag\":\"batch_perf\",\"entry\":\"%s\",\"status\":%d,"
                        "\"pairs\":%lld,\"cells\":%lld,\"padded_cells\":%lld,\"groups\":%lld,"
                        "\"rerun_pairs\":%lld,\"band_groups\":%d,\"head_groups\":%d,\"kernel_ms\":%.4f,"
                        "\"batch_ms\":%.4f,\"gcups_inner\":%.2f,\"gcups_batch\":%.2f",
                        entry, status, (long long)st.pairs, (long long)st.cells,
                        (long long)st.padded_cells, (long long)st.groups,
                        (long long)st.rerun_pairs, st.band_groups, st.head_groups, st.fast_kernel_ms,
                        st.batch_ms,
                        st.fast_kernel_ms > 0 ? cells / st.fast_kernel_ms / 1e6 : 0.0,
                        st.batch_ms > 0 ? cells / st.batch_ms / 1e6 : 0.0);
  if (host_ms > 0 && k > 0 && k < (int)sizeof(line))
    k += std::snprintf(line + k, sizeof(line) - k, ",\"host_ms\":%.4f,\"gcups_outer\":%.2f",
                       host_ms, cells / host_ms / 1e6);
  std::fprintf(stderr, "%s}\n", line);
  std::fflush(stderr);
}

int ipuma_synchronize(ipuma_engine *e) {
  if (!e) return fail(IPUMA_EINVAL, "engine is NULL");
  std::lock_guard<std::mutex> lock_(e->mu);
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipDeviceSynchronize());
  const int rc = check_device_error(e);
  log_batch(e, "device", 0.0, rc);
  return rc;
}

// Exported entry points: the bodies above, with no exception crossing the ABI.
int ipuma_engine_create(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                        ipuma_engine **out) {
  return ipuma::no_throw([&] { return ipuma_engine_create_body(sw, algo, out); });
}

int ipuma_compare_local(ipuma_engine *e, int32_t n, const char *a_chars, const int64_t *a_offsets,
                        const char *b_chars, const int64_t *b_offsets, int32_t *scores,
                        int32_t *a_range, int32_t *b_range) {
  return ipuma::no_throw([&] { return ipuma_compare_local_body(e, n, a_chars, a_offsets, b_chars, b_offsets, scores, a_range, b_range); });
}

int ipuma_compare_mn_local(ipuma_engine *e, int32_t nseqs, const char *chars,
                           const int64_t *offsets, int32_t ncmps, const int32_t *cmps,
                           int32_t *scores, int32_t *a_range, int32_t *b_range) {
  return ipuma::no_throw([&] { return ipuma_compare_mn_local_body(e, nseqs, chars, offsets, ncmps, cmps, scores, a_range, b_range); });
}

int ipuma_compare_device(ipuma_engine *e, int32_t n, const uint8_t *codes, int64_t codes_len,
                         const int64_t *a_off, const int32_t *a_len, const int64_t *b_off,
                         const int32_t *b_len, int32_t *scores, int32_t *a_range,
                         int32_t *b_range, void *stream) {
  return ipuma::no_throw([&] { return ipuma_compare_device_body(e, n, codes, codes_len, a_off, a_len, b_off, b_len, scores, a_range, b_range, stream); });
}

int ipuma_compare_device_noalloc(ipuma_engine *e, int32_t n, const uint8_t *codes,
                                 int64_t codes_len, const int64_t *a_off, const int32_t *a_len,
                                 const int64_t *b_off, const int32_t *b_len, int32_t *scores,
                                 int32_t *a_range, int32_t *b_range, void *stream) {
  return ipuma::no_throw([&] {
    return ipuma_compare_device_noalloc_body(e, n, codes, codes_len, a_off, a_len, b_off, b_len,
                                             scores, a_range, b_range, stream);
  });
}

int ipuma_search_device(ipuma_engine *e, const uint8_t *codes, int64_t codes_len, int32_t nq,
                        const int64_t *q_off, const int32_t *q_len, int32_t ndb,
                        const int64_t *db_off, const int32_t *db_len, int32_t k,
                        int32_t *scores, int32_t *top_idx, int32_t *top_score, void *stream) {
  return ipuma::no_throw([&] { return ipuma_search_device_body(e, codes, codes_len, nq, q_off, q_len, ndb, db_off, db_len, k, scores, top_idx, top_score, stream); });
}

int ipuma_prepared_compare(ipuma_engine *e, const int32_t *inputs, int64_t input_words,
                           int32_t *results, int64_t result_words, int32_t lot) {
  return ipuma::no_throw([&] { return ipuma_prepared_compare_body(e, inputs, input_words, results, result_words, lot); });
}

int ipuma_prepared_submit(ipuma_engine *e, const int32_t *inputs, int64_t input_words,
                          int32_t *results, int64_t result_words, int64_t *job) {
  return ipuma::no_throw([&] { return ipuma_prepared_submit_body(e, inputs, input_words, results, result_words, job); });
}

int ipuma_prepared_join(ipuma_engine *e, int64_t job) {
  return ipuma::no_throw([&] { return ipuma_prepared_join_body(e, job, nullptr); });
}

int ipuma_prepared_join_timed(ipuma_engine *e, int64_t job, ipuma_job_times *times) {
  return ipuma::no_throw([&] { return ipuma_prepared_join_body(e, job, times); });
}

int ipuma_device_clock_hz(ipuma_engine *e, double *hz) {
  if (!e || !hz) return fail(IPUMA_EINVAL, "engine or hz is NULL");
  *hz = e->clock_hz;
  return IPUMA_OK;
}

int ipuma_traceback_local(ipuma_engine *e, int32_t n, const char *a_chars,
                          const int64_t *a_offsets, const char *b_chars, const int64_t *b_offsets,
                          const int32_t *scores, const int32_t *a_range, const int32_t *b_range,
                          uint32_t *cigar, const int64_t *cigar_offsets, int32_t *cigar_len) {
  return ipuma::no_throw([&] { return ipuma_traceback_local_body(e, n, a_chars, a_offsets, b_chars, b_offsets, scores, a_range, b_range, cigar, cigar_offsets, cigar_len); });
}

int ipuma_last_batch_stats(ipuma_engine *e, ipuma_batch_stats *out) {
  return ipuma::no_throw([&] { return ipuma_last_batch_stats_body(e, out); });
}

int ipuma_encode_device(ipuma_engine *e, const char *d_chars, int64_t len, uint8_t *d_codes,
                        void *stream) {
  return ipuma::no_throw([&] { return ipuma_encode_device_body(e, d_chars, len, d_codes, stream); });
}

}  // extern "C"
