// ipu_batch_affine.cpp -- SWAlgorithm facade over the C ABI.
#include "ipu_batch_affine.h"

#include <stdexcept>

namespace ipu {
namespace batchaffine {
namespace {
void check(int rc) {
  if (rc != IPUMA_OK) throw std::runtime_error(ipuma_last_error());
}
ipuma_sw_config to_c(const SWConfig &c) {
  ipuma_sw_config s;
  s.gap_init = c.gapInit;
  s.gap_extend = c.gapExtend;
  s.match_value = c.matchValue;
  s.mismatch_value = c.mismatchValue;
  s.ambiguity_value = c.ambiguityValue;
  s.similarity = static_cast<int32_t>(c.similarity);
  s.datatype = static_cast<int32_t>(c.datatype);
  return s;
}
ipuma_algo_config to_c(const IPUAlgoConfig &c) {
  ipuma_algo_config a;
  a.tiles_used = c.tilesUsed;
  a.max_ab = c.maxAB;
  a.max_batches = c.maxBatches;
  a.bufsize = c.bufsize;
  a.vtype = static_cast<int32_t>(c.vtype);
  a.fill_algo = static_cast<int32_t>(c.fillAlgo);
  a.forward_only = c.forwardOnly ? 1 : 0;
  a.use_remote_buffer = c.useRemoteBuffer ? 1 : 0;
  a.transmission_programs = c.transmissionPrograms;
  a.io_tiles = c.ioTiles;
  return a;
}
void pack(const std::vector<std::string> &v, std::string &chars, std::vector<int64_t> &offs) {
  offs.assign(v.size() + 1, 0);
  size_t total = 0;
  for (size_t i = 0; i < v.size(); ++i) total += v[i].size();
  chars.clear();
  chars.reserve(total + 1);
  for (size_t i = 0; i < v.size(); ++i) {
    offs[i] = (int64_t)chars.size();
    chars += v[i];
  }
  offs[v.size()] = (int64_t)chars.size();
  chars.push_back('\0');
}
}  // namespace

SWAlgorithm::SWAlgorithm(SWConfig config, IPUAlgoConfig algoconfig)
    : SWAlgorithm(config, algoconfig, 0) {}

SWAlgorithm::SWAlgorithm(SWConfig config, IPUAlgoConfig algoconfig, int, size_t slotCap, size_t,
                         bool)
    : algoconfig(algoconfig), config_(config) {
  const ipuma_sw_config s = to_c(config);
  const ipuma_algo_config a = to_c(algoconfig);
  check(ipuma_engine_create(&s, &a, &engine_));
  const size_t total = (size_t)algoconfig.getTotalNumberOfComparisons();
  scores.assign(total, 0);
  a_range_result.assign(total, 0);
  b_range_result.assign(total, 0);
  // slot bookkeeping of the remote-buffer API (src/ipu_batch_affine.cpp:373-381)
  slot_size_ = slotCap > 0 ? (int)slotCap : 1;
  const int lots = algoconfig.transmissionPrograms > 0 ? algoconfig.transmissionPrograms : 1;
  slots_.assign(lots, std::vector<bool>(slot_size_, false));
  slot_avail_.assign(lots, slot_size_);
}

SWAlgorithm::SWAlgorithm(SWAlgorithm &&o) noexcept
    : algoconfig(o.algoconfig), config_(o.config_), engine_(o.engine_), scores(std::move(o.scores)),
      a_range_result(std::move(o.a_range_result)), b_range_result(std::move(o.b_range_result)),
      slot_size_(o.slot_size_), slot_avail_(std::move(o.slot_avail_)), slots_(std::move(o.slots_)) {
  o.engine_ = nullptr;
}

SWAlgorithm::~SWAlgorithm() { ipuma_engine_destroy(engine_); }

void SWAlgorithm::checkSequenceSizes(const IPUAlgoConfig &algoconfig,
                                     const std::vector<int> &SeqSizes) {
  if (SeqSizes.size() > (size_t)algoconfig.maxBatches * algoconfig.tilesUsed)
    throw std::runtime_error("Input sequence (A) is over the max length.");
  for (int s : SeqSizes)
    if (s > algoconfig.maxAB)
      throw std::runtime_error("Sequence size " + std::to_string(s) + " > maxAB " +
                               std::to_string(algoconfig.maxAB));
}

void SWAlgorithm::checkSequenceSizes(const IPUAlgoConfig &algoconfig, const RawSequences &Seqs) {
  std::vector<int> sizes;
  sizes.reserve(Seqs.size());
  for (const auto &s : Seqs) sizes.push_back((int)s.size());
  checkSequenceSizes(algoconfig, sizes);
}

BlockAlignmentResults SWAlgorithm::get_result() {
  return {scores, a_range_result, b_range_result};
}

void SWAlgorithm::store(size_t n) {
  if (scores.size() < n) {
    scores.resize(n);
    a_range_result.resize(n);
    b_range_result.resize(n);
  }
  for (size_t i = 0; i < n; ++i) {
    scores[i] = tmp_s_[i];
    a_range_result[i] = tmp_a_[i];
    b_range_result[i] = tmp_b_[i];
  }
}

void SWAlgorithm::compare_local(const std::vector<std::string> &A,
                                const std::vector<std::string> &B, bool) {
  if (A.size() != B.size()) throw std::runtime_error("A and B differ in size");
  std::string ca, cb;
  std::vector<int64_t> oa, ob;
  pack(A, ca, oa);
  pack(B, cb, ob);
  const size_t n = A.size();
  tmp_s_.assign(n, 0);
  tmp_a_.assign(n, 0);
  tmp_b_.assign(n, 0);
  check(ipuma_compare_local(engine_, (int32_t)n, ca.data(), oa.data(), cb.data(), ob.data(),
                            tmp_s_.data(), tmp_a_.data(), tmp_b_.data()));
  store(n);
}

void SWAlgorithm::compare_mn_local(const std::vector<std::string> &Seqs, const Comparisons &Cmps,
                                   bool) {
  std::string cs;
  std::vector<int64_t> os;
  pack(Seqs, cs, os);
  std::vector<int32_t> flat(2 * Cmps.size());
  for (size_t k = 0; k < Cmps.size(); ++k) {
    flat[2 * k] = Cmps[k].indexA;
    flat[2 * k + 1] = Cmps[k].indexB;
  }
  const size_t n = Cmps.size();
  tmp_s_.assign(n, 0);
  tmp_a_.assign(n, 0);
  tmp_b_.assign(n, 0);
  check(ipuma_compare_mn_local(engine_, (int32_t)Seqs.size(), cs.data(), os.data(), (int32_t)n,
                               flat.data(), tmp_s_.data(), tmp_a_.data(), tmp_b_.data()));
  store(n);
}

}  // namespace batchaffine
}  // namespace ipu
