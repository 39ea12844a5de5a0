// host_threads.cpp -- the engine's threaded host layer under a sanitizer.
// TEST INFRASTRUCTURE: built against the CPU HIP emulation (tests/cpu_emulation/hip)
// together with the engine, server and kernel sources and the oracle
// (oracle/sw_oracle.c), once with -fsanitize=thread and once with
// -fsanitize=undefined, and run as a plain program.  It drives, through the C ABI:
//  1. the request server from four submitting threads, with an arena small
//     enough that admission blocks and the ring wraps;
//  2. prepared-buffer jobs submitted and joined from two threads at once
//     (the engine's executor thread);
//  3. a host batch above 8 MB, gathered by the staging threads, with and
//     without thread starts refused;
// and checks every result against the oracle.  Exit code 0 when all pass.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "ipuma_hip.h"

namespace ipuma {
extern std::atomic<int> g_thread_start_budget;
}

extern "C" {
int oracle_select_matrix(int similarity, int match, int mismatch, int ambiguity, int8_t *out);
void oracle_encode(int datatype, const char *s, int64_t len, uint8_t *out);
void oracle_sw_pair(const uint8_t *a, int m, const uint8_t *b, int n, const int8_t *mat, int dim,
                    int gI, int gE, int forward_only, int *work, int32_t *score, int32_t *arange,
                    int32_t *brange);
}

namespace {
std::atomic<int> failures{0};
void expect(bool ok, const std::string &what) {
  if (!ok) {
    std::printf("FAIL %s\n", what.c_str());
    ++failures;
  }
}

struct Batch {
  std::string a_chars, b_chars;
  std::vector<int64_t> a_off{0}, b_off{0};
  int n() const { return (int)a_off.size() - 1; }
  void add(const std::string &a, const std::string &b) {
    a_chars += a;
    b_chars += b;
    a_off.push_back((int64_t)a_chars.size());
    b_off.push_back((int64_t)b_chars.size());
  }
};

std::string random_dna(std::mt19937 &g, int len) {
  static const char kL[] = "ACGT";
  std::string s(len, 'A');
  for (auto &c : s) c = kL[g() % 4];
  return s;
}

Batch make_batch(uint32_t seed, int n, int max_len) {
  std::mt19937 g(seed);
  Batch b;
  for (int k = 0; k < n; ++k) {
    const std::string a = random_dna(g, 1 + (int)(g() % max_len));
    b.add(a, random_dna(g, (int)(g() % 20)) + a.substr(0, a.size() / 2) + random_dna(g, (int)(g() % 20)));
  }
  return b;
}

// the oracle's results for a batch (DNA, match 1 / mismatch -1, gap init / extend)
void oracle_batch(const Batch &b, int gI, int gE, bool fwd, std::vector<int32_t> &s,
                  std::vector<int32_t> &ar, std::vector<int32_t> &br) {
  int8_t mat[25 * 25];
  const int dim = oracle_select_matrix(0, 1, -1, -1, mat);
  const int n = b.n();
  s.assign(n, 0);
  ar.assign(n, 0);
  br.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    const int la = (int)(b.a_off[k + 1] - b.a_off[k]), lb = (int)(b.b_off[k + 1] - b.b_off[k]);
    std::vector<uint8_t> ca(la + 1), cb(lb + 1);
    oracle_encode(0, b.a_chars.data() + b.a_off[k], la, ca.data());
    oracle_encode(0, b.b_chars.data() + b.b_off[k], lb, cb.data());
    std::vector<int> work(2 * (la + 1));
    oracle_sw_pair(ca.data(), la, cb.data(), lb, mat, dim, gI, gE, fwd ? 1 : 0, work.data(), &s[k],
                   &ar[k], &br[k]);
  }
}

bool same(const std::vector<int32_t> &x, const int32_t *y, int n) {
  return std::memcmp(x.data(), y, (size_t)n * 4) == 0;
}

void test_server() {
  ipuma_sw_config sw;
  ipuma_default_sw_config(&sw);
  sw.gap_init = -2;
  ipuma_algo_config al;
  ipuma_default_algo_config(&al);
  al.forward_only = 0;
  ipuma_server_config sc;
  ipuma_default_server_config(&sc);
  sc.hbm_budget_bytes = 4096;  // a few requests at a time: admission blocks, the ring wraps
  sc.max_batch_pairs = 16;
  sc.max_wait_us = 300;
  ipuma_server *srv = nullptr;
  expect(ipuma_server_create(&sw, &al, &sc, &srv) == IPUMA_OK, "server create");
  if (!srv) return;
  std::vector<Batch> reqs;
  for (int r = 0; r < 8; ++r) reqs.push_back(make_batch(100 + r, 1 + r % 5, 60));
  std::vector<std::thread> th;
  for (int w = 0; w < 4; ++w)
    th.emplace_back([&, w] {
      for (int r = w; r < (int)reqs.size(); r += 4) {
        const Batch &b = reqs[r];
        int64_t t = 0;
        if (ipuma_server_submit(srv, b.n(), b.a_chars.data(), b.a_off.data(), b.b_chars.data(),
                                b.b_off.data(), &t) != IPUMA_OK) {
          expect(false, "server submit");
          continue;
        }
        std::vector<int32_t> s(b.n()), ar(b.n()), br(b.n());
        double ms = 0;
        expect(ipuma_server_wait(srv, t, s.data(), ar.data(), br.data(), &ms) == IPUMA_OK,
               "server wait");
        std::vector<int32_t> es, ea, eb;
        oracle_batch(b, -2, -1, false, es, ea, eb);
        expect(same(es, s.data(), b.n()) && same(ea, ar.data(), b.n()) &&
                   same(eb, br.data(), b.n()),
               "server request " + std::to_string(r));
      }
    });
  for (auto &t : th) t.join();
  ipuma_server_stats st;
  ipuma_server_get_stats(srv, &st);
  expect(st.requests == (int64_t)reqs.size() && st.arena_used_bytes == 0, "server stats");
  ipuma_server_destroy(srv);
}

void test_prepared_jobs() {
  ipuma_sw_config sw;
  ipuma_default_sw_config(&sw);
  ipuma_algo_config al;
  ipuma_default_algo_config(&al);
  al.tiles_used = 3;
  al.max_batches = 8;
  al.bufsize = 400;
  al.max_ab = 100;
  al.fill_algo = IPUMA_GREEDY;
  ipuma_engine *e = nullptr;
  expect(ipuma_engine_create(&sw, &al, &e) == IPUMA_OK, "engine create");
  if (!e) return;
  const int64_t words = ipuma_input_buffer_words(&al);
  const int64_t rwords = 3 * (int64_t)al.tiles_used * al.max_batches;
  std::vector<std::thread> th;
  for (int w = 0; w < 2; ++w)
    th.emplace_back([&, w] {
      for (int j = 0; j < 2; ++j) {
        const Batch b = make_batch(200 + 10 * w + j, 8, 40);
        std::vector<int32_t> in(words), res(rwords, -7), map(b.n());
        int32_t maxb = 0;
        expect(ipuma_prepare_remote(&sw, &al, b.n(), b.a_chars.data(), b.a_off.data(),
                                    b.b_chars.data(), b.b_off.data(), in.data(), words, map.data(),
                                    &maxb) == IPUMA_OK,
               "prepare_remote");
        int64_t job = 0;
        expect(ipuma_prepared_submit(e, in.data(), words, res.data(), rwords, &job) == IPUMA_OK,
               "prepared submit");
        ipuma_job_times times{};
        expect(ipuma_prepared_join_timed(e, job, &times) == IPUMA_OK, "prepared join");
        std::vector<int32_t> s(b.n()), ar(b.n()), br(b.n());
        expect(ipuma_transfer_results(res.data(), rwords, map.data(), b.n(), s.data(), ar.data(),
                                      br.data()) == IPUMA_OK,
               "transfer results");
        std::vector<int32_t> es, ea, eb;
        oracle_batch(b, 0, -1, false, es, ea, eb);
        expect(same(es, s.data(), b.n()) && same(ea, ar.data(), b.n()) &&
                   same(eb, br.data(), b.n()),
               "prepared job " + std::to_string(w) + "/" + std::to_string(j));
      }
    });
  for (auto &t : th) t.join();
  ipuma_engine_destroy(e);
}

void test_staging() {
  ipuma_sw_config sw;
  ipuma_default_sw_config(&sw);
  sw.gap_init = -2;
  ipuma_algo_config al;
  ipuma_default_algo_config(&al);
  al.tiles_used = 1 << 20;
  al.max_batches = 1;
  al.max_ab = 65535;
  al.forward_only = 1;
  ipuma_engine *e = nullptr;
  expect(ipuma_engine_create(&sw, &al, &e) == IPUMA_OK, "engine create");
  if (!e) return;
  // 9 MB of characters in long A's against empty B's (no DP cells),
  // interleaved with short pairs that land in every staging slice
  std::mt19937 g(7);
  const std::string big = random_dna(g, 65535);
  Batch b, shorts;
  for (int k = 0; k < 136; ++k) {
    b.add(big, "");
    const std::string a = random_dna(g, 20 + (int)(g() % 40));
    const std::string bb = random_dna(g, (int)(g() % 20)) + a.substr(3, 20);
    b.add(a, bb);
    shorts.add(a, bb);
  }
  std::vector<int32_t> es, ea, eb;
  oracle_batch(shorts, -2, -1, true, es, ea, eb);
  for (int cap : {3, 0}) {  // some worker threads refused, then all
    ipuma::g_thread_start_budget = cap;
    std::vector<int32_t> s(b.n()), ar(b.n()), br(b.n());
    expect(ipuma_compare_local(e, b.n(), b.a_chars.data(), b.a_off.data(), b.b_chars.data(),
                               b.b_off.data(), s.data(), ar.data(), br.data()) == IPUMA_OK,
           "staged compare");
    ipuma::g_thread_start_budget = -1;
    bool ok = true;
    for (int k = 0; k < shorts.n(); ++k)
      ok = ok && s[2 * k] == 0 && s[2 * k + 1] == es[k] && ar[2 * k + 1] == ea[k] &&
           br[2 * k + 1] == eb[k];
    expect(ok, "staged batch with thread budget " + std::to_string(cap));
  }
  ipuma_engine_destroy(e);
}
}  // namespace

int main() {
  test_server();
  std::printf("server done\n");
  std::fflush(stdout);
  test_prepared_jobs();
  std::printf("prepared jobs done\n");
  std::fflush(stdout);
  test_staging();
  std::printf("staging done\n");
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures.load());
  return failures ? 1 : 0;
}
