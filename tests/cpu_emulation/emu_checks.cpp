// emu_checks.cpp -- CPU (emulated HIP) regression checks of kernel input
// validation.  TEST INFRASTRUCTURE: compiled with tests/cpu_emulation/hip.
//  1. k_reverse_prefix with an end position past its sequence must raise
//     error bit 2 and write nothing (under ASan: no write past the buffer).
//  2. k_fast given a corrupted group descriptor (super-strip count far past
//     any sequence) must raise error bit 4 and skip it instead of running it.
//  3. k_fast given a group whose boundary row exceeds the wave scratch it was
//     given must raise error bit 4 and skip it.
#include <cstdio>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "sw_device.h"

using namespace ipuma;

static int failures = 0;
#define CHECK(c, what)                           \
  do {                                           \
    if (!(c)) {                                  \
      std::printf("FAIL: %s\n", what);           \
      ++failures;                                \
    }                                            \
  } while (0)

int main() {
  {  // 1. reverse prefix with an out-of-range end
    const int n = 2;
    std::vector<uint8_t> codes(32, 1), rcodes(32 + 16, 0);
    int64_t a_off[2] = {0, 16}, b_off[2] = {8, 24};
    int32_t a_len[2] = {5, 5}, b_len[2] = {5, 5};
    int32_t scores[2] = {3, 3};
    int32_t a_range[2] = {(int32_t)(0xFFFFu << 16), 2 << 16};  // pair 0: Aend 65535
    int32_t b_range[2] = {1 << 16, 1 << 16};
    int32_t ra_len[2] = {-7, -7}, rb_len[2] = {-7, -7};
    int32_t error = 0;
    launch_reverse_prefix(n, codes.data(), a_off, a_len, b_off, b_len, scores, a_range, b_range,
                          rcodes.data(), ra_len, rb_len, &error, nullptr);
    CHECK(error & 2, "bad end raises error bit 2");
    CHECK(ra_len[0] == 0 && rb_len[0] == 0, "bad pair left empty");
    CHECK(ra_len[1] == 3 && rb_len[1] == 2, "good pair reversed");
  }
  {  // 2. corrupted group descriptor
    const int kp = 18;
    std::vector<int16_t> table(kp * kp, 1);
    std::vector<uint8_t> codes(64, 1);
    int64_t a_off[1] = {0}, b_off[1] = {32};
    int32_t a_len[1] = {8}, b_len[1] = {8};
    std::vector<int32_t> sorted(256, -1);
    sorted[0] = 0;
    GroupDesc gd[2];
    gd[0] = {0, 8, 0x01020304, 0, 16};  // nss from stray residue bytes
    gd[1] = {0, 8, 1, 0, 8};            // the real group (8 rows in its one super-strip)
    BucketCtl ctl;
    std::memset(&ctl, 0, sizeof(ctl));
    ctl.ngroups = 2;
    int32_t ovf = 0;
    std::vector<uint32_t> scratch((size_t)kWaveScratchWords * (kFastThreads / 64), 0);
    int32_t score = -5, ar = 0, br = 0, ovf_list[4];
    FastArgs A{};
    A.codes = codes.data(); A.a_off = a_off; A.a_len = a_len; A.b_off = b_off; A.b_len = b_len;
    A.table = table.data(); A.sorted = sorted.data(); A.groups = gd; A.ctl = &ctl;
    A.scratch = scratch.data(); A.scores = &score; A.a_range = &ar;
    A.b_range = &br; A.ovf_list = ovf_list; A.ovf_count = &ovf;
    A.max_groups = 2; A.gap_o = 0x00010001u; A.gap_x = 0x00010001u;
    A.wave_words = kWaveScratchWords;
    A.band_words = kBandScratchWords;
    launch_fast(kp, A, 1, nullptr);
    CHECK(ctl.error & 4, "corrupted descriptor raises error bit 4");
    CHECK(score == 8, "the valid group still runs (8 matches)");
    // 3. a group whose boundary row and column table would not fit a smaller
    // wave scratch: one-lane pipelines of mpad 16 need (3 * 16 + 5) * 64 =
    // 3392 words, the real group (mpad 8) (3 * 8 + 5) * 64 = 1856
    std::memset(&ctl, 0, sizeof(ctl));
    ctl.ngroups = 2;
    gd[0] = {0, 16, 2, 0, 16};
    score = -5;
    A.wave_words = 2048;
    launch_fast(kp, A, 1, nullptr);
    CHECK(ctl.error & 4, "a boundary row past the wave scratch raises error bit 4");
    CHECK(score == 8, "the group that fits still runs");
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
