// ssw_ref_driver.cpp -- C entry points over the reference's CPU comparator.
//
// TEST INFRASTRUCTURE ONLY.  The reference checks its IPU results against the
// Striped Smith-Waterman library it vendors (extern/include/ssw/, used by
// test/parity-test.cpp:70-87 and cpusw/aligners.hpp:128-152).  oracle/Makefile
// compiles that library straight from the reference tree together with this
// file into oracle/_ref/libssw_ref.so; nothing from the reference is copied.
// tests/ use it to pin the oracle's DP and to freeze golden fixtures, and
// bench.py's CPU baseline times it (ssw_ref_batch_codes) as the reference's
// own CPU path.
//
// Orientation follows the reference's parity test: the reference's A is SSW's
// "query" and its B is SSW's "ref" (test/parity-test.cpp:80, 90-93).
#include <cstdint>
#include <cstring>

#include "ssw.hpp"
#include "ssw_core.hpp"

namespace {
void fill(const StripedSmithWaterman::Alignment &aln, int32_t out[5]) {
  out[0] = aln.sw_score;
  out[1] = aln.query_begin;
  out[2] = aln.query_end;
  out[3] = aln.ref_begin;
  out[4] = aln.ref_end;
}
void cigar_out(const StripedSmithWaterman::Alignment &aln, char *cigar, int cap) {
  if (!cigar || cap <= 0) return;
  std::strncpy(cigar, aln.cigar_string.c_str(), (size_t)cap - 1);
  cigar[cap - 1] = 0;
}
}  // namespace

extern "C" {

// DNA aligner over SSW's built-in ACGTN alphabet (ssw.cpp:289-308).
// out = {score, query(A) begin, query(A) end, ref(B) begin, ref(B) end}.
int ssw_ref_align_dna(const char *query, int qlen, const char *ref, int rlen, int match,
                      int mismatch, int gap_open, int gap_extend, int ambiguity,
                      int32_t out[5]) {
  StripedSmithWaterman::Aligner aligner(match, mismatch, gap_open, gap_extend, ambiguity);
  StripedSmithWaterman::Filter filter;
  StripedSmithWaterman::Alignment aln;
  int mask = qlen / 2 > 15 ? qlen / 2 : 15;
  if (!aligner.Align(query, qlen, ref, rlen, filter, &aln, mask)) return -1;
  fill(aln, out);
  return 0;
}

// Matrix aligner: `matrix` is dim x dim int8 (row = query code), `translate`
// maps 128 chars to codes (ssw.cpp:310-329).
int ssw_ref_align_matrix(const char *query, int qlen, const char *ref, int rlen,
                         const int8_t *matrix, int dim, const int8_t *translate,
                         int gap_open, int gap_extend, int32_t out[5]) {
  StripedSmithWaterman::Aligner aligner(matrix, dim, translate, 128);
  aligner.SetGapPenalty(gap_open, gap_extend);
  StripedSmithWaterman::Filter filter;
  StripedSmithWaterman::Alignment aln;
  int mask = qlen / 2 > 15 ? qlen / 2 : 15;
  if (!aligner.Align(query, qlen, ref, rlen, filter, &aln, mask)) return -1;
  fill(aln, out);
  return 0;
}

// As ssw_ref_align_dna / ssw_ref_align_matrix, plus SSW's CIGAR string (soft
// clips, M, I, D; query = A, ref = B) written to `cigar` (cap bytes).
int ssw_ref_align_dna_cigar(const char *query, int qlen, const char *ref, int rlen, int match,
                            int mismatch, int gap_open, int gap_extend, int ambiguity,
                            int32_t out[5], char *cigar, int cap) {
  StripedSmithWaterman::Aligner aligner(match, mismatch, gap_open, gap_extend, ambiguity);
  StripedSmithWaterman::Filter filter;
  StripedSmithWaterman::Alignment aln;
  int mask = qlen / 2 > 15 ? qlen / 2 : 15;
  if (!aligner.Align(query, qlen, ref, rlen, filter, &aln, mask)) return -1;
  fill(aln, out);
  cigar_out(aln, cigar, cap);
  return 0;
}

int ssw_ref_align_matrix_cigar(const char *query, int qlen, const char *ref, int rlen,
                               const int8_t *matrix, int dim, const int8_t *translate,
                               int gap_open, int gap_extend, int32_t out[5], char *cigar,
                               int cap) {
  StripedSmithWaterman::Aligner aligner(matrix, dim, translate, 128);
  aligner.SetGapPenalty(gap_open, gap_extend);
  StripedSmithWaterman::Filter filter;
  StripedSmithWaterman::Alignment aln;
  int mask = qlen / 2 > 15 ? qlen / 2 : 15;
  if (!aligner.Align(query, qlen, ref, rlen, filter, &aln, mask)) return -1;
  fill(aln, out);
  cigar_out(aln, cigar, cap);
  return 0;
}

// The reference's CPU comparator over a whole batch, as its benchmark driver
// times it (cpusw/aligners.hpp:128-152): per pair a query profile of A
// (ssw_init, 8- and 16-bit) and ssw_align against B, end positions only
// (flag 0: the forward-only benchmark), mask max(|A| / 2, 15), pairs spread
// over `threads` OpenMP threads.  `codes` are residue codes and `mat` is
// dim x dim in the same code space; gap_open / gap_extend are SSW's positive
// penalties (first gap residue, each further one).  Scores go to `scores`
// (-1 where SSW returns no alignment).
void ssw_ref_batch_codes(int32_t n, const uint8_t *codes, const int64_t *a_off,
                         const int32_t *a_len, const int64_t *b_off, const int32_t *b_len,
                         const int8_t *mat, int32_t dim, int32_t gap_open, int32_t gap_extend,
                         int32_t threads, int32_t *scores) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
  for (int32_t p = 0; p < n; ++p) {
    const int32_t la = a_len[p], lb = b_len[p];
    if (la <= 0 || lb <= 0) {
      scores[p] = 0;
      continue;
    }
    s_profile *prof = ssw_init(reinterpret_cast<const int8_t *>(codes + a_off[p]), la, mat, dim, 2);
    const int32_t mask = la / 2 > 15 ? la / 2 : 15;
    s_align *al = ssw_align(prof, reinterpret_cast<const int8_t *>(codes + b_off[p]), lb,
                            (uint8_t)gap_open, (uint8_t)gap_extend, 0, 0, 0, mask);
    scores[p] = al ? al->score1 : -1;
    if (al) align_destroy(al);
    init_destroy(prof);
  }
}
}
