// ipu_batch_affine.h -- the reference's public class, ipu::batchaffine::SWAlgorithm
// (src/ipu_batch_affine.h:22-118), on top of the MI355X engine's C ABI
// (include/ipuma_hip.h).  Same names, argument meaning and result layout:
// get_result() returns vectors of tilesUsed*maxBatches entries of which the
// first N are written by the last compare call; ranges pack end<<16 | start.
//
// Errors: the reference throws std::runtime_error for too many comparisons
// and exit(1)s for |A| > maxAB (src/ipu_batch_affine.cpp:410-423); here both
// throw std::runtime_error, as does any device failure.
#ifndef IPUMA_AMD_IPU_BATCH_AFFINE_H
#define IPUMA_AMD_IPU_BATCH_AFFINE_H

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include <tuple>

#include "ipu_config.h"
#include "ipuma_hip.h"
#include "partition.h"
#include "timing.hpp"
#include "types.h"

namespace ipu {
namespace batchaffine {

using slotToken = long;

struct BlockAlignmentResults {
  std::vector<int32_t> scores;
  std::vector<int32_t> a_range_result;
  std::vector<int32_t> b_range_result;
};

// The buffers and timers of a submitted prepared batch
// (src/ipu_batch_affine.h:28-38).  cyclesH2D / cyclesInner point at the
// owning Job's counters.  The reference's completion channel has no
// counterpart here (the engine's executor signals the join).
struct SubmittedBatch {
  int32_t *inputs_begin = nullptr;
  int32_t *inputs_end = nullptr;
  int32_t *results_begin = nullptr;
  int32_t *results_end = nullptr;
  slotToken slot = 0;
  void *signal_done = nullptr;
  uint64_t *cyclesH2D = nullptr;
  uint64_t *cyclesInner = nullptr;
  swatlib::TickTock runTick;  // the executor's run of the batch (accumulated time)
};

// A prepared batch submitted with async_submit_prepared_remote_compare
// (src/ipu_batch_affine.h:40-48).  join() waits for its results; the input
// and result buffers must stay valid until then.  The caller deletes it.
// After join: h2dCycles / innerCycles are the input copy and the device
// batch, from HIP events, in cycles of getTileClockFrequency() (so
// cycles / getTileClockFrequency() are seconds, as in the reference);
// tick spans submit to join, sb.runTick holds the executor's run of it.
struct Job {
  void join();
  uint64_t h2dCycles = 0;
  uint64_t innerCycles = 0;
  swatlib::TickTock tick;
  SubmittedBatch sb;
  ipuma_engine *engine = nullptr;
  int64_t id = 0;
  double clock_hz = 0.0;
  bool joined = false;
};

class SWAlgorithm {
 public:
  IPUAlgoConfig algoconfig;

  SWAlgorithm(SWConfig config, IPUAlgoConfig algoconfig);
  // thread_id / slotCap / ipuCount / runExecutor are accepted for source
  // compatibility; the engine binds the current HIP device.
  SWAlgorithm(SWConfig config, IPUAlgoConfig algoconfig, int thread_id, size_t slotCap = 1,
              size_t ipuCount = 1, bool runExecutor = true);
  ~SWAlgorithm();
  SWAlgorithm(const SWAlgorithm &) = delete;
  SWAlgorithm &operator=(const SWAlgorithm &) = delete;
  SWAlgorithm(SWAlgorithm &&other) noexcept;

  static void checkSequenceSizes(const IPUAlgoConfig &algoconfig, const std::vector<int> &SeqSizes);
  static void checkSequenceSizes(const IPUAlgoConfig &algoconfig, const RawSequences &Seqs);

  BlockAlignmentResults get_result();

  void compare_local(const std::vector<std::string> &A, const std::vector<std::string> &B,
                     bool errcheck = false);
  void compare_mn_local(const std::vector<std::string> &Seqs, const Comparisons &Cmps,
                        bool errcheck = false);

  // --- prepared input buffers (src/ipu_batch_affine.h:81-116) -------------
  // Layout of a prepared input buffer (int32 words), the reference's wire
  // format: tilesUsed tile regions of residue codes (one byte each, region
  // stride scale(ceil(bufsize/4)*4) bytes), then at word
  // scale(ceil(bufsize/4))*tilesUsed the metadata, maxBatches x
  // {|A|, offset of A, |B|, offset of B} per tile (offsets inside the tile's
  // region).  scale(x) = ceil(x / transmissionPrograms) * (lot + 1), lot =
  // calculate_slot_region_index(algoconfig, largest tile fill).  Comparison k
  // lands in slot mapping[k] = tile * maxBatches + position; results come back
  // as [scores | a ranges | b ranges], one word per slot.
  static int calculate_slot_region_index(const IPUAlgoConfig &algoconfig, int max_buffer_size);
  int calculate_slot_region_index(int max_buffer_size);
  std::tuple<int, slotToken> unpack_slot(slotToken);
  bool slot_available(int max_buffer_size);
  slotToken queue_slot(int max_buffer_size);

  static void fillBuckets(Algorithm algo, partition::BucketMap &map, const RawSequences &A,
                          const RawSequences &B, int offset = 0);
  static void fillMNBuckets(Algorithm algo, partition::BucketMap &map, const RawSequences &Seqs,
                            const Comparisons &Cmps, int offset = 0);
  // Writes the tiles and metadata of `map` into [inputs_begin, inputs_end)
  // and each comparison's slot into mapping.  The A/B form clears the buffer
  // first; the comparison-list form writes only what the map covers (as the
  // reference).  Throws std::runtime_error if the buffer is too small.
  static void fill_input_buffer(const partition::BucketMap &map, const swatlib::DataType dtype,
                                const IPUAlgoConfig &algoconfig, const RawSequences &Seqs,
                                const Comparisons &Cmps, int32_t *inputs_begin,
                                int32_t *inputs_end, int32_t *mapping);
  static void fill_input_buffer(const partition::BucketMap &map, const swatlib::DataType dtype,
                                const IPUAlgoConfig &algoconfig, const RawSequences &A,
                                const RawSequences &B, int32_t *inputs_begin, int32_t *inputs_end,
                                int32_t *mapping);
  // checkSequenceSizes(A) + fillBuckets(fillAlgo) + fill_input_buffer; returns
  // the largest tile fill in bytes.
  static int prepare_remote(const SWConfig &swconfig, const IPUAlgoConfig &algoconfig,
                            const std::vector<std::string> &A, const std::vector<std::string> &B,
                            int32_t *inputs_begin, int32_t *inputs_end, int *deviceMapping);
  // Result words of comparison k = slot mapping[k] of each third of the
  // results buffer; throws if the buffer is not a multiple of 3 words.
  static void transferResults(int32_t *results_begin, int32_t *results_end, int *mapping_begin,
                              int *mapping_end, int32_t *scores_begin, int32_t *scores_end,
                              int32_t *arange_begin, int32_t *arange_end, int32_t *brange_begin,
                              int32_t *brange_end);
  static void transferResults(int32_t *results_begin, int32_t *results_end, int *mapping_begin,
                              int *mapping_end, int32_t *scores_begin, int32_t *scores_end,
                              int32_t *arange_begin, int32_t *arange_end, int32_t *brange_begin,
                              int32_t *brange_end, int numComparisons);

  // Runs a prepared buffer; the results buffer receives its first
  // (results_end - results_begin) words of [scores | a ranges | b ranges].
  // The slot token's lot selects the layout's lot (transmissionPrograms > 1).
  void prepared_remote_compare(int32_t *inputs_begin, int32_t *inputs_end,
                               int32_t *results_begin, int32_t *results_end,
                               slotToken slot_token = 0);
  Job *async_submit_prepared_remote_compare(int32_t *inputs_begin, int32_t *inputs_end,
                                            int32_t *results_begin, int32_t *results_end);
  void blocking_join_prepared_remote_compare(Job &job);
  // The device loop of the reference runs per engine here and starts on
  // demand; run_executor() is accepted and does nothing.
  void run_executor() {}
  // Remote buffers are a Poplar feature the reference leaves unimplemented
  // (src/ipu_batch_affine.cpp:466-487, 665-672): these throw.
  void upload(int32_t *inputs_begin, int32_t *inputs_end, slotToken slot);
  void refetch();
  std::string printTensors();

  // Clock (Hz) that converts a Job's cycle counts to seconds
  // (IPUAlgorithm::getTileClockFrequency, src/ipu_base.h:50): the device's
  // peak shader clock.
  double getTileClockFrequency();

  // The engine handle, for the device-resident path (ipuma_compare_device).
  ipuma_engine *engine() const { return engine_; }
  const SWConfig &config() const { return config_; }

 private:
  void store(size_t n);
  SWConfig config_;
  ipuma_engine *engine_ = nullptr;
  std::vector<int32_t> scores, a_range_result, b_range_result;
  std::vector<int32_t> tmp_s_, tmp_a_, tmp_b_;
  void release_slot(slotToken i);
  int slot_size_ = 1;
  std::vector<int> slot_avail_;
  std::vector<std::vector<bool>> slots_;
};

}  // namespace batchaffine
}  // namespace ipu
#endif
