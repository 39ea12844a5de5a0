// sw_tables.h -- scoring matrices and residue encodings (host side).
#pragma once
#include <stdint.h>

namespace ipuma {
// Side of the matrix for a similarity (0 na, 1 blosum50, 2 blosum62), -1 for `simple`.
int matrix_dim(int similarity);
// selectMatrix: dim x dim int8, row = A code; returns dim or -1.
int select_matrix(int similarity, int match, int mismatch, int ambiguity, int8_t *out);
// (dim+1) x (dim+1) int16 with the pad code dim scored -32768.
void build_pad_table(const int8_t *mat, int dim, int16_t *out);
// 256-entry char -> code table (0 na, 1 aa, 2 string).
void build_code_table(int datatype, uint8_t table[256]);
}  // namespace ipuma
