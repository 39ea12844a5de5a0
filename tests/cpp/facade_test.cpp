// facade_test.cpp -- the reference's gtest cases for the public API, re-hosted
// as a plain C++ program over ipu::batchaffine::SWAlgorithm (the facade of the
// MI355X engine).  Rows and expectations: test/simple-test.cpp:25-45, 105-171;
// test/mn_cmp-test.cpp:29-57.  Exit code 0 when every check passes.
#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <cstdint>
#include <string>
#include <vector>

#include "driver.hpp"
#include "ipu_batch_affine.h"

namespace {
int failures = 0;
#define EXPECT_EQ(a, b, what)                                                              \
  do {                                                                                     \
    if ((a) != (b)) {                                                                      \
      std::printf("FAIL %s: got %lld want %lld\n", what, (long long)(a), (long long)(b)); \
      ++failures;                                                                          \
    }                                                                                      \
  } while (0)

struct Row {
  std::string ref, query;
  int qstart, qend, rstart, rend, score;
};

const std::vector<Row> kRows = {
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
    {"ACGT", "TTACGTTT", 2, 5, 0, 3, 4},
    {"AAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGG", 0, 15, 0, 15, 16},
    {"AAAATTTTCCCCGGGG", "AAAATTTTACCCGGGG", 0, 15, 0, 15, 14},
    {"AAAATTTTCCCCGGGG", "AAAATTTTACCCCGGGG", 0, 16, 0, 15, 15},
    {"AAAATTTTCCCCGGGG", "AAAATTTCCCCGGGG", 0, 14, 0, 15, 14},
    {"AAAATTTTCCCCGGGG", "GCTAGCTAGCTAGCTA", 3, 3, 0, 0, 1},
    {"AAAATTTTCCCCGGGG", "GCTAAAATTTTCCCCGGGG", 3, 18, 0, 15, 16},
    {"AAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGGACT", 0, 15, 0, 15, 16},
};

void check_rows(const ipu::batchaffine::BlockAlignmentResults &r, const char *label) {
  for (size_t i = 0; i < kRows.size(); ++i) {
    const uint32_t a = (uint32_t)r.a_range_result[i], b = (uint32_t)r.b_range_result[i];
    char what[96];
    std::snprintf(what, sizeof(what), "%s row %zu score", label, i);
    EXPECT_EQ(r.scores[i], kRows[i].score, what);
    std::snprintf(what, sizeof(what), "%s row %zu query begin", label, i);
    EXPECT_EQ((int16_t)(a & 0xffff), kRows[i].qstart, what);
    std::snprintf(what, sizeof(what), "%s row %zu query end", label, i);
    EXPECT_EQ((int16_t)(a >> 16), kRows[i].qend, what);
    std::snprintf(what, sizeof(what), "%s row %zu ref begin", label, i);
    EXPECT_EQ((int16_t)(b & 0xffff), kRows[i].rstart, what);
    std::snprintf(what, sizeof(what), "%s row %zu ref end", label, i);
    EXPECT_EQ((int16_t)(b >> 16), kRows[i].rend, what);
  }
}
}  // namespace

int main() {
  std::vector<std::string> queries, refs;
  for (const auto &r : kRows) {
    queries.push_back(r.query);
    refs.push_back(r.ref);
  }
  {  // SimpleCorrectnessTest.UseCppVertex / UseCppMultiVertex / UseAsmMultiVertex
    ipu::IPUAlgoConfig algo;
    algo.tilesUsed = 2;
    algo.maxAB = 300;
    algo.maxBatches = 20;
    algo.bufsize = 3000;
    algo.fillAlgo = ipu::Algorithm::roundRobin;
    for (auto vt : {ipu::VertexType::cpp, ipu::VertexType::multi, ipu::VertexType::multiasm}) {
      algo.vtype = vt;
      ipu::batchaffine::SWAlgorithm driver({}, algo);
      driver.compare_local(queries, refs);
      auto res = driver.get_result();
      EXPECT_EQ(res.scores.size(), (size_t)40, "result vector size (tilesUsed*maxBatches)");
      check_rows(res, ipu::vertexTypeToConfigString(vt).c_str());
    }
  }
  {  // SimpleCorrectnessTest.useMNComparisons
    ipu::IPUAlgoConfig algo;
    algo.tilesUsed = 4;
    algo.maxAB = 300;
    algo.maxBatches = 5;
    ipu::batchaffine::SWAlgorithm driver({}, algo);
    std::vector<std::string> seqs;
    ipu::Comparisons cmps;
    for (int i = 0; i < (int)queries.size(); ++i) {
      seqs.push_back(queries[i]);
      seqs.push_back(refs[i]);
      cmps.push_back({2 * i, 2 * i + 1});
    }
    driver.compare_mn_local(seqs, cmps);
    check_rows(driver.get_result(), "mn");
  }
  {  // MN_Test.CompareAll
    std::vector<std::string> seqs = {"AAAAAA", "TTTTTT", "TTTTAT", "AAAAAA"};
    ipu::Comparisons cmps = {{0, 2}, {0, 3}, {1, 2}, {1, 3}};
    ipu::batchaffine::SWAlgorithm driver({0, -1, 1, -1, -1, swatlib::Similarity::nucleicAcid,
                                          swatlib::DataType::nucleicAcid},
                                         {1, 20, 30, 1000, ipu::VertexType::assembly});
    driver.compare_mn_local(seqs, cmps);
    auto r = driver.get_result();
    EXPECT_EQ(r.scores[0], 1, "mn 0");
    EXPECT_EQ(r.scores[1], 6, "mn 1");
    EXPECT_EQ(r.scores[2], 4, "mn 2");
    EXPECT_EQ(r.scores[3], 0, "mn 3");
  }
  {  // errors throw instead of exit(1)
    ipu::IPUAlgoConfig algo;
    algo.tilesUsed = 1;
    algo.maxBatches = 2;
    algo.maxAB = 10;
    ipu::batchaffine::SWAlgorithm driver({}, algo);
    bool threw = false;
    try {
      driver.compare_local({"A", "A", "A"}, {"A", "A", "A"});
    } catch (const std::runtime_error &) {
      threw = true;
    }
    EXPECT_EQ(threw, true, "too many comparisons throws");
    threw = false;
    try {
      driver.compare_local({std::string(11, 'A')}, {"A"});
    } catch (const std::runtime_error &) {
      threw = true;
    }
    EXPECT_EQ(threw, true, "|A| > maxAB throws");
  }
  {  // prepared buffers: prepare_remote, prepared_remote_compare, transferResults (the
     // flow of the reference's batch driver, ipusw/batch.hpp:34-63, ipusw/ipuworker.hpp:8-66),
     // then the same buffers through the asynchronous job API
    ipu::IPUAlgoConfig algo;
    algo.tilesUsed = 3;
    algo.maxAB = 300;
    algo.maxBatches = 10;
    algo.bufsize = 1000;
    for (auto fill : {ipu::Algorithm::fillFirst, ipu::Algorithm::roundRobin, ipu::Algorithm::greedy}) {
      algo.fillAlgo = fill;
      ipu::batchaffine::SWAlgorithm driver({}, algo);
      std::vector<int32_t> inputs(algo.getInputBufferSize32b());
      std::vector<int> mapping(queries.size());
      ipu::batchaffine::SWAlgorithm::prepare_remote({}, algo, queries, refs, inputs.data(),
                                                    inputs.data() + inputs.size(), mapping.data());
      std::vector<int32_t> results(3 * algo.getTotalNumberOfComparisons());
      ipu::batchaffine::BlockAlignmentResults r;
      const size_t n = queries.size();
      r.scores.resize(n);
      r.a_range_result.resize(n);
      r.b_range_result.resize(n);
      for (int pass = 0; pass < 2; ++pass) {
        std::fill(results.begin(), results.end(), -7);
        if (pass == 0) {
          driver.prepared_remote_compare(inputs.data(), inputs.data() + inputs.size(),
                                         results.data(), results.data() + results.size());
        } else {
          auto *job = driver.async_submit_prepared_remote_compare(
              inputs.data(), inputs.data() + inputs.size(), results.data(),
              results.data() + results.size());
          driver.blocking_join_prepared_remote_compare(*job);
          delete job;
        }
        ipu::batchaffine::SWAlgorithm::transferResults(
            results.data(), results.data() + results.size(), mapping.data(),
            mapping.data() + mapping.size(), r.scores.data(), r.scores.data() + n,
            r.a_range_result.data(), r.a_range_result.data() + n, r.b_range_result.data(),
            r.b_range_result.data() + n);
        check_rows(r, pass == 0 ? "prepared" : "prepared async");
      }
    }
  }
  {  // driver.hpp singleton with the compile-time configuration
    init_single_ipu(SW_CONFIGURATION, ALGO_CONFIGURATION);
    auto *d = getDriver();
    EXPECT_EQ(d != nullptr, true, "getDriver");
    d->compare_local({"AAAATTTTCCCCGGGG"}, {"AAAATTTTCCCCGGGG"});
    EXPECT_EQ(d->get_result().scores[0], 16, "driver score");
  }
  {  // config JSON round trip (src/ipu_config.cpp:16-80 keys)
    ipu::SWConfig c;
    c.gapInit = -2;
    c.mismatchValue = -3;
    c.similarity = swatlib::Similarity::blosum62;
    c.datatype = swatlib::DataType::aminoAcid;
    auto back = ipu::sw_config_from_json(ipu::to_json(c));
    EXPECT_EQ(back.gapInit, -2, "json gapInit");
    EXPECT_EQ(back.mismatchValue, -3, "json mismatchValue");
    EXPECT_EQ((int)back.similarity, (int)swatlib::Similarity::blosum62, "json similarity");
    ipu::IPUAlgoConfig a;
    a.forwardOnly = true;
    a.fillAlgo = ipu::Algorithm::greedy;
    auto ab = ipu::algo_config_from_json(ipu::to_json(a));
    EXPECT_EQ(ab.forwardOnly, true, "json forwardOnly");
    EXPECT_EQ((int)ab.fillAlgo, (int)ipu::Algorithm::greedy, "json fillAlgo");
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
