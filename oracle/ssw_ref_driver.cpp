// ssw_ref_driver.cpp -- C entry points over the reference's CPU comparator.
//
// TEST INFRASTRUCTURE ONLY.  The reference checks its IPU results against the
// Striped Smith-Waterman library it vendors (extern/include/ssw/, used by
// test/parity-test.cpp:70-87 and cpusw/aligners.hpp:128-152).  oracle/Makefile
// compiles that library straight from the reference tree together with this
// file into oracle/_ref/libssw_ref.so; nothing from the reference is copied.
// tests/ use it to pin the oracle's DP and to freeze golden fixtures.
//
// Orientation follows the reference's parity test: the reference's A is SSW's
// "query" and its B is SSW's "ref" (test/parity-test.cpp:80, 90-93).
#include <cstdint>
#include <cstring>

#include "ssw.hpp"

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
}
