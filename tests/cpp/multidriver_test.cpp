// multidriver_test.cpp -- a caller written in the shape of the reference's
// batch driver (ipusw/ipumultidriver.hpp:10-145, ipusw/run_comparison.hpp:67-172,
// ipusw/ipuworker.hpp:8-66, ipusw/ipuswconfig.hpp:10-46), compiled against the
// facade (ipuma-lib_amd/cpp/) and nlohmann::json: a run configuration that
// embeds SWConfig / IPUAlgoConfig through their to_json / from_json, the
// driver's clock frequency, prepared buffers filled by worker threads,
// asynchronous jobs submitted from one thread and joined by receiver threads,
// per-batch performance from Job::h2dCycles / innerCycles / tick / sb, and
// results checked against the expectations of test/simple-test.cpp:25-45.
// TEST INFRASTRUCTURE.  Exit code 0 when every check passes.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <nlohmann/json.hpp>

#include "ipu_batch_affine.h"

using json = nlohmann::json;

namespace {
std::atomic<int> failures{0};
std::mutex print_mu;
void expect(bool ok, const std::string &what) {
  if (!ok) {
    std::lock_guard<std::mutex> lk(print_mu);
    std::printf("FAIL %s\n", what.c_str());
    ++failures;
  }
}

struct Row {
  const char *ref, *query;
  int qstart, qend, rstart, rend, score;
};
// test/simple-test.cpp:25-45 (query = A, ref = B)
const Row kRows[] = {
    {"acgt", "acgt", 0, 3, 0, 3, 4},
    {"ACGT", "ACGT", 0, 3, 0, 3, 4},
    {"AACGT", "ACGT", 0, 3, 1, 4, 4},
    {"ACGTT", "ACGT", 0, 3, 0, 3, 4},
    {"ACGT", "TACGT", 1, 4, 0, 3, 4},
    {"ACGT", "TTACGT", 2, 5, 0, 3, 4},
    {"ACGT", "ACGTTT", 0, 3, 0, 3, 4},
    {"ACGT", "TACGTT", 1, 4, 0, 3, 4},
    {"ACGT", "TTACGTT", 2, 5, 0, 3, 4},
    {"ACGT", "TACGTTT", 1, 4, 0, 3, 4},
    {"ACGT", "TTACGTTT", 2, 5, 0, 3, 4},
    {"AAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGG", 0, 15, 0, 15, 16},
    {"AAAATTTTCCCCGGGG", "AAAATTTTACCCGGGG", 0, 15, 0, 15, 14},
    {"AAAATTTTCCCCGGGG", "AAAATTTTACCCCGGGG", 0, 16, 0, 15, 15},
    {"AAAATTTTCCCCGGGG", "AAAATTTCCCCGGGG", 0, 14, 0, 15, 14},
    {"AAAATTTTCCCCGGGG", "GCTAGCTAGCTAGCTA", 3, 3, 0, 0, 1},
    {"AAAATTTTCCCCGGGG", "GCTAAAATTTTCCCCGGGG", 3, 18, 0, 15, 16},
    {"AAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGGACT", 0, 15, 0, 15, 16},
};
constexpr int kNumRows = sizeof(kRows) / sizeof(kRows[0]);

// The run configuration of the batch driver, serialized with the facade's
// conversions of its members.
struct RunConfig {
  ipu::SWConfig swconfig;
  ipu::IPUAlgoConfig ipuconfig;
  int numDevices = 1;
  int numThreads = 4;
  int repeat = 1;
};
void to_json(json &j, const RunConfig &c) {
  j = json{{"sw", c.swconfig}, {"ipu", c.ipuconfig}, {"numDevices", c.numDevices},
           {"numThreads", c.numThreads}, {"repeat", c.repeat}};
}
void from_json(const json &j, RunConfig &c) {
  j.at("sw").get_to(c.swconfig);
  j.at("ipu").get_to(c.ipuconfig);
  j.at("numDevices").get_to(c.numDevices);
  j.at("numThreads").get_to(c.numThreads);
  j.at("repeat").get_to(c.repeat);
}

struct Batch {
  int first = 0, count = 0;
  std::vector<int32_t> input, mapping, result;
  ipu::batchaffine::Job *job = nullptr;
};

class MultiDriver {
 public:
  explicit MultiDriver(RunConfig c) : config(c) {}
  ~MultiDriver() { delete driver_; }
  void init() {
    driver_ = new ipu::batchaffine::SWAlgorithm(config.swconfig, config.ipuconfig, 0, 0,
                                                config.numDevices, false);
    clock_ = driver_->getTileClockFrequency();
  }
  void run() { driver_->run_executor(); }
  double clock() const { return clock_; }
  void fill(const ipu::RawSequences &a, const ipu::RawSequences &b, Batch &batch) {
    batch.input.assign(config.ipuconfig.getInputBufferSize32b(), 0);
    batch.mapping.assign(config.ipuconfig.getTotalNumberOfComparisons(), 0);
    ipu::batchaffine::SWAlgorithm::prepare_remote(config.swconfig, config.ipuconfig, a, b,
                                                  batch.input.data(),
                                                  batch.input.data() + batch.input.size(),
                                                  batch.mapping.data());
  }
  ipu::batchaffine::Job *submit(Batch &batch) {
    return driver_->async_submit_prepared_remote_compare(
        batch.input.data(), batch.input.data() + batch.input.size(), batch.result.data(),
        batch.result.data() + batch.result.size());
  }
  // waits for the job and logs its batch_perf line (the metrics of
  // ipusw/ipumultidriver.hpp:54-121)
  json wait(ipu::batchaffine::Job *job) {
    job->join();
    const double timeJob = (double)job->tick.accumulate_microseconds() / 1e6;
    const double timeBatch = (double)job->sb.runTick.accumulate_microseconds() / 1e6;
    const double timeOuter = (double)(job->h2dCycles + job->innerCycles) / clock_;
    const double timeInner = (double)job->innerCycles / clock_;
    const int32_t *meta = job->sb.inputs_begin + config.ipuconfig.getTotalBufsize32b();
    uint64_t cells = 0, data = 0;
    for (int i = 0; i < config.ipuconfig.getTotalNumberOfComparisons(); ++i) {
      cells += (uint64_t)meta[4 * i] * (uint64_t)meta[4 * i + 2];
      data += (uint64_t)meta[4 * i] + (uint64_t)meta[4 * i + 2];
    }
    json log = {{"tag", "batch_perf"},
                {"cycle_inner", job->innerCycles},
                {"cycle_outer", job->h2dCycles + job->innerCycles},
                {"time_inner", timeInner},
                {"time_outer", timeOuter},
                {"time_job", timeJob},
                {"time_batch", timeBatch},
                {"cell_count", cells},
                {"data_count", data},
                {"gcups_inner", (double)cells / timeInner / 1e9},
                {"gcups_outer", (double)cells / timeOuter / 1e9}};
    return log;
  }
  void results(Batch &batch, ipu::batchaffine::BlockAlignmentResults &r) {
    const int n = batch.count;
    r.scores.resize(n);
    r.a_range_result.resize(n);
    r.b_range_result.resize(n);
    ipu::batchaffine::SWAlgorithm::transferResults(
        batch.result.data(), batch.result.data() + batch.result.size(), batch.mapping.data(),
        batch.mapping.data() + batch.mapping.size(), r.scores.data(), r.scores.data() + n,
        r.a_range_result.data(), r.a_range_result.data() + n, r.b_range_result.data(),
        r.b_range_result.data() + n, n);
  }
  RunConfig config;

 private:
  ipu::batchaffine::SWAlgorithm *driver_ = nullptr;
  double clock_ = 0.0;
};
}  // namespace

int main() {
  RunConfig cfg;
  cfg.swconfig.gapInit = 0;
  cfg.swconfig.gapExtend = -1;
  cfg.ipuconfig.tilesUsed = 4;
  cfg.ipuconfig.maxAB = 300;
  cfg.ipuconfig.maxBatches = 16;
  cfg.ipuconfig.bufsize = 800;
  cfg.ipuconfig.vtype = ipu::VertexType::multiasm;
  cfg.ipuconfig.fillAlgo = ipu::Algorithm::greedy;
  cfg.ipuconfig.forwardOnly = false;

  // configuration through JSON and back (the reference logs it this way,
  // ipusw/run_comparison.hpp:77-84)
  const json j = cfg;
  expect(j.at("sw").at("similarity") == "na", "json sw.similarity is its config string");
  expect(j.at("sw").at("datatype") == "na", "json sw.datatype");
  expect(j.at("ipu").at("vtype") == "multiasm", "json ipu.vtype");
  expect(j.at("ipu").at("fillAlgo") == "greedy", "json ipu.fillAlgo");
  expect(j.at("ipu").at("maxBatches") == 16, "json ipu.maxBatches");
  RunConfig back = json::parse(j.dump()).get<RunConfig>();
  expect(back.ipuconfig.bufsize == 800 && back.ipuconfig.tilesUsed == 4, "json round trip ipu");
  expect(back.ipuconfig.fillAlgo == ipu::Algorithm::greedy, "json round trip fillAlgo");
  expect(back.ipuconfig.vtype == ipu::VertexType::multiasm, "json round trip vtype");
  expect(back.swconfig.gapExtend == -1 && back.swconfig.matchValue == 1, "json round trip sw");
  {
    ipu::SWConfig p;
    p.similarity = swatlib::Similarity::blosum62;
    p.datatype = swatlib::DataType::aminoAcid;
    p.mismatchValue = -3;
    json pj = p;
    expect(pj.at("similarity") == "blosum62" && pj.at("mismatchValue") == -3, "json SWConfig");
    expect(pj.get<ipu::SWConfig>().datatype == swatlib::DataType::aminoAcid, "json datatype back");
    json ej = swatlib::Similarity::blosum50;
    expect(ej == "blosum50", "json Similarity");
  }

  // 20 copies of the golden rows, 360 comparisons, cut into batches of at
  // most 60 comparisons (the limits of ipusw/run_comparison.hpp:93-95)
  ipu::RawSequences A, B;
  std::vector<int> row_of;
  for (int rep = 0; rep < 20; ++rep)
    for (int r = 0; r < kNumRows; ++r) {
      const int k = (r * 7 + rep) % kNumRows;
      A.push_back(kRows[k].query);
      B.push_back(kRows[k].ref);
      row_of.push_back(k);
    }
  const int per_batch = 60;
  std::vector<Batch> batches;
  for (int first = 0; first < (int)A.size(); first += per_batch) {
    Batch b;
    b.first = first;
    b.count = std::min(per_batch, (int)A.size() - first);
    b.result.assign(3 * cfg.ipuconfig.getTotalNumberOfComparisons(), -7);
    batches.push_back(std::move(b));
  }

  MultiDriver driver(back);
  std::thread init([&] { driver.init(); });
  // fill workers (ipusw/ipuworker.hpp:38-66)
  std::vector<std::thread> fillers;
  for (int w = 0; w < 2; ++w)
    fillers.emplace_back([&, w] {
      for (size_t i = w; i < batches.size(); i += 2) {
        Batch &b = batches[i];
        ipu::RawSequences a(A.begin() + b.first, A.begin() + b.first + b.count);
        ipu::RawSequences bb(B.begin() + b.first, B.begin() + b.first + b.count);
        driver.fill(a, bb, b);
      }
    });
  for (auto &t : fillers) t.join();
  init.join();
  driver.run();
  expect(driver.clock() > 1e8 && driver.clock() < 1e10, "clock frequency in Hz");

  // submit everything from this thread, receive on two (run_comparison.hpp:131-145)
  std::atomic<int> received{0};
  std::vector<std::string> logs(batches.size());
  for (auto &b : batches) b.job = driver.submit(b);
  std::vector<std::thread> receivers;
  for (int w = 0; w < 2; ++w)
    receivers.emplace_back([&, w] {
      for (size_t i = w; i < batches.size(); i += 2) {
        Batch &b = batches[i];
        const json log = driver.wait(b.job);
        logs[i] = log.dump();
        expect(log.at("cycle_inner").get<uint64_t>() > 0, "inner cycles counted");
        const double ti = log.at("time_inner").get<double>();
        expect(ti > 0.0 && ti < 30.0, "inner time is positive seconds");
        expect(log.at("time_outer").get<double>() >= ti, "outer time includes the inner time");
        expect(log.at("time_job").get<double>() >= log.at("time_batch").get<double>() - 2e-6,
               "job time (submit to join) spans the executor's run");
        expect(log.at("time_batch").get<double>() > 0.0, "run time measured");
        uint64_t cells = 0;
        for (int k = 0; k < b.count; ++k) cells += A[b.first + k].size() * B[b.first + k].size();
        expect(log.at("cell_count").get<uint64_t>() == cells, "cell count from the buffer's metadata");
        ipu::batchaffine::BlockAlignmentResults r;
        driver.results(b, r);
        for (int k = 0; k < b.count; ++k) {
          const Row &row = kRows[row_of[b.first + k]];
          const uint32_t a = (uint32_t)r.a_range_result[k], bb = (uint32_t)r.b_range_result[k];
          const bool ok = r.scores[k] == row.score && (int16_t)(a & 0xffff) == row.qstart &&
                          (int16_t)(a >> 16) == row.qend && (int16_t)(bb & 0xffff) == row.rstart &&
                          (int16_t)(bb >> 16) == row.rend;
          expect(ok, "batch " + std::to_string(i) + " comparison " + std::to_string(k));
        }
        delete b.job;
        b.job = nullptr;
        ++received;
      }
    });
  for (auto &t : receivers) t.join();
  expect(received == (int)batches.size(), "every batch received");
  for (const auto &l : logs) std::printf("%s%s\n", IPU_JSON_LOG_TAG, l.c_str());
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures.load());
  return failures ? 1 : 0;
}
