// wire.cpp -- prepared input buffers: the reference's wire format
// (src/ipu_batch_affine.cpp:674-950, src/ipu_config.cpp:6-14), bucket fills,
// the slot bookkeeping of the remote-buffer API, and their C ABI.
//
// The byte layout is the reference's, so buffers prepared here and buffers
// prepared by the reference are interchangeable:
//   ibuf      = ceil(bufsize / 4)                        words per tile
//   scale(x)  = ceil(x / transmissionPrograms) * (lot + 1)
//   tile t's codes at byte t * scale(4 * ibuf), one byte per residue
//   metadata at word scale(ibuf) * tilesUsed: per tile maxBatches x
//     {|A|, A's byte offset in the tile, |B|, B's byte offset in the tile}
//   mapping[k] = tile * maxBatches + position of comparison k in the tile
// with lot = calculate_slot_region_index(largest tile fill) (0 whenever the
// fill fits bufsize with one transmission program).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "ipu_batch_affine.h"
#include "sw_tables.h"

namespace ipuma {
int set_last_error(int code, const char *fmt, ...);  // the C ABI's last-error slot
}

namespace ipu {
namespace batchaffine {
namespace {

void check(int rc) {
  if (rc != IPUMA_OK) throw std::runtime_error(ipuma_last_error());
}

struct Layout {
  size_t tile_bytes;   // byte stride of the tile regions
  size_t meta_offset;  // word offset of the metadata
};

Layout layout_for(const IPUAlgoConfig &c, int max_bucket) {
  const int lots = c.transmissionPrograms > 0 ? c.transmissionPrograms : 1;
  const int lot = SWAlgorithm::calculate_slot_region_index(c, max_bucket);
  auto scale = [&](size_t x) {
    return (x / lots + (x % lots == 0 ? 0 : 1)) * (size_t)(lot + 1);
  };
  const size_t ibuf = (size_t)c.getBufsize32b();
  return {scale(ibuf * 4), scale(ibuf) * (size_t)c.tilesUsed};
}

int max_fill(const partition::BucketMap &map) {
  int m = 0;
  for (const auto &b : map.buckets) m = std::max(m, b.seqSize);
  return m;
}

// Bounds of what writing `map` touches, checked before any write.
void check_fits(const partition::BucketMap &map, const IPUAlgoConfig &c, const Layout &L,
                const int32_t *begin, const int32_t *end) {
  const size_t words = (size_t)(end - begin);
  for (const auto &b : map.buckets) {
    if (b.bucketIndex < 0) throw std::runtime_error("negative bucket index");
    const size_t seq_end = (size_t)b.bucketIndex * L.tile_bytes + (size_t)b.seqSize;
    const size_t meta_end =
        L.meta_offset + ((size_t)b.bucketIndex * c.maxBatches + b.cmps.size()) * 4;
    if ((b.seqSize > 0 && seq_end > words * 4) || (!b.cmps.empty() && meta_end > words))
      throw std::runtime_error("Input buffer too small for the prepared layout.");
  }
}

template <class SeqOf>
void write_map(const partition::BucketMap &map, swatlib::DataType dtype, const IPUAlgoConfig &c,
               int32_t *begin, int32_t *end, int32_t *mapping, const SeqOf &seq_of) {
  const Layout L = layout_for(c, max_fill(map));
  check_fits(map, c, L, begin, end);
  uint8_t table[256];
  ipuma::build_code_table((int)dtype, table);
  int8_t *seqs = reinterpret_cast<int8_t *>(begin);
  int32_t *meta = begin + L.meta_offset;
  for (const auto &b : map.buckets) {
    int8_t *tile = seqs + (size_t)b.bucketIndex * L.tile_bytes;
    int32_t *tmeta = meta + (size_t)b.bucketIndex * c.maxBatches * 4;
    for (const auto &sm : b.seqs) {
      const std::string &s = seq_of(sm);
      for (size_t j = 0; j < s.size(); ++j)
        tile[sm.offset + j] = (int8_t)table[(unsigned char)s[j]];
    }
    for (size_t i = 0; i < b.cmps.size(); ++i) {
      const auto &cm = b.cmps[i];
      tmeta[i * 4] = cm.sizeA;
      tmeta[i * 4 + 1] = cm.offsetA;
      tmeta[i * 4 + 2] = cm.sizeB;
      tmeta[i * 4 + 3] = cm.offsetB;
      mapping[cm.comparisonIndex] = map.cmpCapacity * b.bucketIndex + (int)i;
    }
  }
}

IPUAlgoConfig from_c(const ipuma_algo_config &a) {
  IPUAlgoConfig c;
  c.tilesUsed = a.tiles_used;
  c.maxAB = a.max_ab;
  c.maxBatches = a.max_batches;
  c.bufsize = a.bufsize;
  c.vtype = static_cast<VertexType>(a.vtype);
  c.fillAlgo = static_cast<Algorithm>(a.fill_algo);
  c.forwardOnly = a.forward_only != 0;
  c.useRemoteBuffer = a.use_remote_buffer != 0;
  c.transmissionPrograms = a.transmission_programs;
  c.ioTiles = a.io_tiles;
  return c;
}

std::vector<std::string> strings(int32_t n, const char *chars, const int64_t *offs) {
  std::vector<std::string> v((size_t)n);
  for (int32_t k = 0; k < n; ++k) {
    const int64_t len = offs[k + 1] - offs[k];
    if (len < 0) throw std::invalid_argument("offsets not ascending at " + std::to_string(k));
    v[k].assign(chars + offs[k], (size_t)len);
  }
  return v;
}

}  // namespace

int SWAlgorithm::calculate_slot_region_index(const IPUAlgoConfig &c, int max_buffer_size) {
  return (int)std::ceil((max_buffer_size / (double)c.bufsize) *
                        static_cast<double>(c.transmissionPrograms)) -
         1;
}

int SWAlgorithm::calculate_slot_region_index(int max_buffer_size) {
  return calculate_slot_region_index(algoconfig, max_buffer_size);
}

std::tuple<int, slotToken> SWAlgorithm::unpack_slot(slotToken s) {
  return {(int)(s / slot_size_), s % slot_size_};
}

bool SWAlgorithm::slot_available(int max_buffer_size) {
  const int lot = calculate_slot_region_index(max_buffer_size);
  return lot >= 0 && lot < (int)slot_avail_.size() && slot_avail_[lot] > 0;
}

slotToken SWAlgorithm::queue_slot(int max_buffer_size) {
  const int lot = calculate_slot_region_index(max_buffer_size);
  if (lot < 0 || lot >= (int)slots_.size()) throw std::runtime_error("no slot region for this size");
  for (int i = 0; i < slot_size_; ++i) {
    if (!slots_[lot][i]) {
      slots_[lot][i] = true;
      --slot_avail_[lot];
      return i + (slotToken)lot * slot_size_;
    }
  }
  throw std::runtime_error("no free slot");
}

void SWAlgorithm::release_slot(slotToken token) {
  const auto [lot, sid] = unpack_slot(token);
  if (lot < 0 || lot >= (int)slots_.size() || sid < 0 || sid >= slot_size_) return;
  if (slots_[lot][sid]) {  // tokens that were never queued (e.g. 0) release nothing
    slots_[lot][sid] = false;
    ++slot_avail_[lot];
  }
}

void SWAlgorithm::fillBuckets(Algorithm algo, partition::BucketMap &map, const RawSequences &A,
                              const RawSequences &B, int offset) {
  switch (algo) {
    case Algorithm::fillFirst: partition::fillFirst(map, A, B, offset); break;
    case Algorithm::roundRobin: partition::roundRobin(map, A, B, offset); break;
    case Algorithm::greedy: partition::greedy(map, A, B, offset); break;
  }
}

void SWAlgorithm::fillMNBuckets(Algorithm algo, partition::BucketMap &map,
                                const RawSequences &Seqs, const Comparisons &Cmps, int offset) {
  switch (algo) {
    case Algorithm::fillFirst: partition::fillFirst(map, Seqs, Cmps, offset); break;
    case Algorithm::roundRobin: partition::roundRobin(map, Seqs, Cmps, offset); break;
    case Algorithm::greedy: partition::greedy(map, Seqs, Cmps, offset); break;
  }
}

void SWAlgorithm::fill_input_buffer(const partition::BucketMap &map, const swatlib::DataType dtype,
                                    const IPUAlgoConfig &c, const RawSequences &Seqs,
                                    const Comparisons &, int32_t *begin, int32_t *end,
                                    int32_t *mapping) {
  write_map(map, dtype, c, begin, end, mapping, [&](const partition::SequenceMapping &sm)
                -> const std::string & {
    if (sm.origin != partition::SequenceOrigin::unordered)
      throw std::runtime_error("Using A/B mapping with unordered comparison.");
    return Seqs[sm.index];
  });
}

void SWAlgorithm::fill_input_buffer(const partition::BucketMap &map, const swatlib::DataType dtype,
                                    const IPUAlgoConfig &c, const RawSequences &A,
                                    const RawSequences &B, int32_t *begin, int32_t *end,
                                    int32_t *mapping) {
  std::memset(begin, 0, (size_t)(end - begin) * sizeof(int32_t));
  write_map(map, dtype, c, begin, end, mapping, [&](const partition::SequenceMapping &sm)
                -> const std::string & {
    switch (sm.origin) {
      case partition::SequenceOrigin::A: return A[sm.index];
      case partition::SequenceOrigin::B: return B[sm.index];
      default: throw std::runtime_error("Using unordered mapping with A/B comparison.");
    }
  });
}

int SWAlgorithm::prepare_remote(const SWConfig &sw, const IPUAlgoConfig &c,
                                const std::vector<std::string> &A,
                                const std::vector<std::string> &B, int32_t *begin, int32_t *end,
                                int *mapping) {
  if (A.size() != B.size()) throw std::runtime_error("A and B differ in size");
  checkSequenceSizes(c, A);
  partition::BucketMap map(c.tilesUsed, c.maxBatches, c.bufsize);
  fillBuckets(c.fillAlgo, map, A, B, 0);
  fill_input_buffer(map, sw.datatype, c, A, B, begin, end, mapping);
  return max_fill(map);
}

void SWAlgorithm::transferResults(int32_t *rb, int32_t *re, int *mb, int *me, int32_t *sb,
                                  int32_t *se, int32_t *ab, int32_t *ae, int32_t *bb,
                                  int32_t *be) {
  transferResults(rb, re, mb, me, sb, se, ab, ae, bb, be, (int)(me - mb));
}

void SWAlgorithm::transferResults(int32_t *rb, int32_t *re, int *mb, int *, int32_t *sb,
                                  int32_t *, int32_t *ab, int32_t *, int32_t *bb, int32_t *,
                                  int n) {
  const size_t size = (size_t)(re - rb);
  if (size % 3 != 0) throw std::runtime_error("Results buffer not 3 aligned.");
  const size_t part = size / 3;
  for (int i = 0; i < n; ++i) {
    const int m = mb[i];
    if (m < 0 || (size_t)m >= part) throw std::runtime_error("mapping outside the results buffer");
    sb[i] = rb[m];
    ab[i] = rb[part + m];
    bb[i] = rb[2 * part + m];
  }
}

void SWAlgorithm::prepared_remote_compare(int32_t *ib, int32_t *ie, int32_t *rb, int32_t *re,
                                          slotToken slot_token) {
  release_slot(slot_token);
  const int lot = std::get<0>(unpack_slot(slot_token));
  check(ipuma_prepared_compare(engine_, ib, (int64_t)(ie - ib), rb, (int64_t)(re - rb), lot));
}

Job *SWAlgorithm::async_submit_prepared_remote_compare(int32_t *ib, int32_t *ie, int32_t *rb,
                                                       int32_t *re) {
  Job *job = new Job();
  job->engine = engine_;
  job->clock_hz = getTileClockFrequency();
  job->sb.inputs_begin = ib;
  job->sb.inputs_end = ie;
  job->sb.results_begin = rb;
  job->sb.results_end = re;
  job->sb.cyclesH2D = &job->h2dCycles;
  job->sb.cyclesInner = &job->innerCycles;
  job->tick.tick();
  int64_t id = 0;
  const int rc = ipuma_prepared_submit(engine_, ib, (int64_t)(ie - ib), rb, (int64_t)(re - rb), &id);
  if (rc != IPUMA_OK) {
    delete job;
    check(rc);
  }
  job->id = id;
  job->sb.slot = (slotToken)id;
  return job;
}

void SWAlgorithm::blocking_join_prepared_remote_compare(Job &job) { job.join(); }

double SWAlgorithm::getTileClockFrequency() {
  double hz = 0.0;
  check(ipuma_device_clock_hz(engine_, &hz));
  return hz;
}

void Job::join() {
  if (joined) return;
  joined = true;
  ipuma_job_times t{};
  check(ipuma_prepared_join_timed(engine, id, &t));
  tick.tock();
  sb.runTick.add_microseconds((int64_t)(t.run_ms * 1e3 + 0.5));
  h2dCycles = (uint64_t)(t.h2d_ms * 1e-3 * clock_hz + 0.5);
  innerCycles = (uint64_t)(t.device_ms * 1e-3 * clock_hz + 0.5);
}

void SWAlgorithm::upload(int32_t *, int32_t *, slotToken) {
  throw std::runtime_error("Not implemented");
}
void SWAlgorithm::refetch() { throw std::runtime_error("Not implemented"); }
std::string SWAlgorithm::printTensors() { throw std::runtime_error("Not implemented"); }

}  // namespace batchaffine
}  // namespace ipu

// ---- C ABI -----------------------------------------------------------------
namespace {
using ipu::batchaffine::SWAlgorithm;

int fail_from(const std::exception &e) {
  const std::string m = e.what();
  const int code = m.find("maxAB") != std::string::npos ? IPUMA_ETOOLONG : IPUMA_EINVAL;
  return ipuma::set_last_error(code, "%s", m.c_str());
}
}  // namespace

extern "C" {

int64_t ipuma_input_buffer_words(const ipuma_algo_config *algo) {
  if (!algo) return ipuma::set_last_error(IPUMA_EINVAL, "algo is NULL");
  return (int64_t)ipu::batchaffine::from_c(*algo).getInputBufferSize32b();
}

int ipuma_prepare_remote(const ipuma_sw_config *sw, const ipuma_algo_config *algo, int32_t n,
                         const char *a_chars, const int64_t *a_offsets, const char *b_chars,
                         const int64_t *b_offsets, int32_t *inputs, int64_t input_words,
                         int32_t *mapping, int32_t *max_bucket) {
  if (!sw || !algo || n < 0 || input_words < 0 || (input_words > 0 && !inputs))
    return ipuma::set_last_error(IPUMA_EINVAL, "bad arguments");
  if (n > 0 && (!a_chars || !a_offsets || !b_chars || !b_offsets || !mapping))
    return ipuma::set_last_error(IPUMA_EINVAL, "NULL buffer");
  try {
    ipu::SWConfig s;
    s.datatype = static_cast<swatlib::DataType>(sw->datatype);
    const auto A = ipu::batchaffine::strings(n, a_chars, a_offsets);
    const auto B = ipu::batchaffine::strings(n, b_chars, b_offsets);
    const int mb = SWAlgorithm::prepare_remote(s, ipu::batchaffine::from_c(*algo), A, B, inputs,
                                               inputs + input_words, mapping);
    if (max_bucket) *max_bucket = mb;
    return IPUMA_OK;
  } catch (const std::exception &e) {
    return fail_from(e);
  }
}

int ipuma_prepare_remote_mn(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                            int32_t nseqs, const char *chars, const int64_t *offsets,
                            int32_t ncmps, const int32_t *cmps, int32_t *inputs,
                            int64_t input_words, int32_t *mapping, int32_t *max_bucket) {
  if (!sw || !algo || nseqs < 0 || ncmps < 0 || input_words < 0 || (input_words > 0 && !inputs))
    return ipuma::set_last_error(IPUMA_EINVAL, "bad arguments");
  if ((nseqs > 0 && (!chars || !offsets)) || (ncmps > 0 && (!cmps || !mapping)))
    return ipuma::set_last_error(IPUMA_EINVAL, "NULL buffer");
  try {
    const auto seqs = ipu::batchaffine::strings(nseqs, chars, offsets);
    ipu::Comparisons cs((size_t)ncmps);
    for (int32_t k = 0; k < ncmps; ++k) {
      cs[k].indexA = cmps[2 * k];
      cs[k].indexB = cmps[2 * k + 1];
      if (cs[k].indexA < 0 || cs[k].indexA >= nseqs || cs[k].indexB < 0 || cs[k].indexB >= nseqs)
        return ipuma::set_last_error(IPUMA_EINVAL, "comparison %d indexes outside the sequences", k);
    }
    const ipu::IPUAlgoConfig c = ipu::batchaffine::from_c(*algo);
    if (input_words > 0) std::memset(inputs, 0, (size_t)input_words * sizeof(int32_t));
    ipu::partition::BucketMap map(c.tilesUsed, c.maxBatches, c.bufsize);
    SWAlgorithm::fillMNBuckets(c.fillAlgo, map, seqs, cs, 0);
    SWAlgorithm::fill_input_buffer(map, static_cast<swatlib::DataType>(sw->datatype), c, seqs, cs,
                                   inputs, inputs + input_words, mapping);
    if (max_bucket) *max_bucket = ipu::batchaffine::max_fill(map);
    return IPUMA_OK;
  } catch (const std::exception &e) {
    return fail_from(e);
  }
}

int ipuma_transfer_results(const int32_t *results, int64_t result_words, const int32_t *mapping,
                           int32_t n, int32_t *scores, int32_t *a_range, int32_t *b_range) {
  if (result_words < 0 || n < 0 || (n > 0 && (!results || !mapping || !scores || !a_range ||
                                                !b_range)))
    return ipuma::set_last_error(IPUMA_EINVAL, "bad arguments");
  try {
    int32_t *r = const_cast<int32_t *>(results);
    int *m = const_cast<int32_t *>(mapping);
    SWAlgorithm::transferResults(r, r + result_words, m, m + n, scores, scores + n, a_range,
                                 a_range + n, b_range, b_range + n, n);
    return IPUMA_OK;
  } catch (const std::exception &e) {
    return fail_from(e);
  }
}

}  // extern "C"
