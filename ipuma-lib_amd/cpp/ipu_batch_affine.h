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

#include "ipu_config.h"
#include "ipuma_hip.h"
#include "types.h"

namespace ipu {
namespace batchaffine {

struct BlockAlignmentResults {
  std::vector<int32_t> scores;
  std::vector<int32_t> a_range_result;
  std::vector<int32_t> b_range_result;
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

  // The engine handle, for the device-resident path (ipuma_compare_device).
  ipuma_engine *engine() const { return engine_; }
  const SWConfig &config() const { return config_; }

 private:
  void store(size_t n);
  SWConfig config_;
  ipuma_engine *engine_ = nullptr;
  std::vector<int32_t> scores, a_range_result, b_range_result;
  std::vector<int32_t> tmp_s_, tmp_a_, tmp_b_;
};

}  // namespace batchaffine
}  // namespace ipu
#endif
