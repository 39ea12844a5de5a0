/*
 * sw_oracle.c -- CPU restatement of the reference's batched Smith-Waterman
 * affine-gap path (xiamaz/ipuma-lib).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker / the timed CPU baseline.  The product library
 * (ipuma-lib_amd) never links or calls it.
 *
 * What it restates (file:line into the reference):
 *   - scoring tables       src/swatlib/similarity.h:17-94
 *   - selectMatrix/scale   src/swatlib/similarity.cpp:19-33, 89-104
 *   - char encodings       src/swatlib/encoding.h:27-73, encoding.cpp:40-48
 *   - forward DP + first-max end tracking
 *                          src/codelets/codelet_batch_affine.cpp:44-87
 *   - reverse DP for start positions
 *                          src/codelets/codelet_batch_affine.cpp:89-115
 *   - range packing (end<<16 | start, uint16 each)
 *                          src/codelets/codelet_batch_affine.cpp:117-123
 *   - CIGAR between start and end (no counterpart in the IPU path; the
 *     traceback rules of SSW's banded_sw, ssw_core.cpp:854-1053, without
 *     its band): oracle_cigar below
 *
 * Definitions where the reference is undefined (documented, deliberate):
 *   - an empty A or B gives score 0 and ranges 0 (the reference vertex
 *     `break`s out of its slot loop, codelet_batch_affine.cpp:60);
 *   - the nucleic-acid "unknown" code 16 lies outside the reference's 16x16
 *     matrix (an out-of-bounds read there); here row/column 16 score as the
 *     ambiguity value, like 'N'.
 *   - characters outside 0..126 encode like '\0' (unknown).
 *
 * Pinning: tests/test_oracle.py checks every function here against the
 * reference's golden rows (test/simple-test.cpp:25-45), its MN expected
 * scores (test/mn_cmp-test.cpp:36-41), closed forms, an exhaustive
 * brute-force enumerator on tiny inputs, and SSW (the reference's CPU
 * comparator, built from the reference sources into oracle/_ref).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Nucleic-acid 0/1 similarity, rows/cols N A T C G U R Y K M S W B D H V
 * (similarity.h:17-35). */
static const int8_t NA_BIN[16 * 16] = {
    1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
    1, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 1, 1, 1,
    1, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1, 1, 1, 1, 0,
    1, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1,
    1, 0, 0, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 1, 0, 1,
    1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    1, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 1, 1, 1, 1,
    1, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 1, 1,
    1, 0, 1, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0,
    1, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1,
    1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 1, 0, 0, 1,
    1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 1, 0,
    1, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 0, 1, 0, 0, 0,
    1, 1, 1, 0, 1, 0, 1, 1, 1, 0, 0, 1, 0, 1, 0, 0,
    1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 0, 1, 0, 0, 1, 0,
    1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 0, 0, 0, 1,
};

/* BLOSUM50, rows/cols A R N D C Q E G H I L K M F P S T W Y V B J Z X *
 * (similarity.h:38-64). */
static const int8_t BL50[25 * 25] = {
     5,-2,-1,-2,-1,-1,-1, 0,-2,-1,-2,-1,-1,-3,-1, 1, 0,-3,-2, 0,-2,-2,-1,-1,-5,
    -2, 7,-1,-2,-4, 1, 0,-3, 0,-4,-3, 3,-2,-3,-3,-1,-1,-3,-1,-3,-1,-3, 0,-1,-5,
    -1,-1, 7, 2,-2, 0, 0, 0, 1,-3,-4, 0,-2,-4,-2, 1, 0,-4,-2,-3, 5,-4, 0,-1,-5,
    -2,-2, 2, 8,-4, 0, 2,-1,-1,-4,-4,-1,-4,-5,-1, 0,-1,-5,-3,-4, 6,-4, 1,-1,-5,
    -1,-4,-2,-4,13,-3,-3,-3,-3,-2,-2,-3,-2,-2,-4,-1,-1,-5,-3,-1,-3,-2,-3,-1,-5,
    -1, 1, 0, 0,-3, 7, 2,-2, 1,-3,-2, 2, 0,-4,-1, 0,-1,-1,-1,-3, 0,-3, 4,-1,-5,
    -1, 0, 0, 2,-3, 2, 6,-3, 0,-4,-3, 1,-2,-3,-1,-1,-1,-3,-2,-3, 1,-3, 5,-1,-5,
     0,-3, 0,-1,-3,-2,-3, 8,-2,-4,-4,-2,-3,-4,-2, 0,-2,-3,-3,-4,-1,-4,-2,-1,-5,
    -2, 0, 1,-1,-3, 1, 0,-2,10,-4,-3, 0,-1,-1,-2,-1,-2,-3, 2,-4, 0,-3, 0,-1,-5,
    -1,-4,-3,-4,-2,-3,-4,-4,-4, 5, 2,-3, 2, 0,-3,-3,-1,-3,-1, 4,-4, 4,-3,-1,-5,
    -2,-3,-4,-4,-2,-2,-3,-4,-3, 2, 5,-3, 3, 1,-4,-3,-1,-2,-1, 1,-4, 4,-3,-1,-5,
    -1, 3, 0,-1,-3, 2, 1,-2, 0,-3,-3, 6,-2,-4,-1, 0,-1,-3,-2,-3, 0,-3, 1,-1,-5,
    -1,-2,-2,-4,-2, 0,-2,-3,-1, 2, 3,-2, 7, 0,-3,-2,-1,-1, 0, 1,-3, 2,-1,-1,-5,
    -3,-3,-4,-5,-2,-4,-3,-4,-1, 0, 1,-4, 0, 8,-4,-3,-2, 1, 4,-1,-4, 1,-4,-1,-5,
    -1,-3,-2,-1,-4,-1,-1,-2,-2,-3,-4,-1,-3,-4,10,-1,-1,-4,-3,-3,-2,-3,-1,-1,-5,
     1,-1, 1, 0,-1, 0,-1, 0,-1,-3,-3, 0,-2,-3,-1, 5, 2,-4,-2,-2, 0,-3, 0,-1,-5,
     0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 2, 5,-3,-2, 0, 0,-1,-1,-1,-5,
    -3,-3,-4,-5,-5,-1,-3,-3,-3,-3,-2,-3,-1, 1,-4,-4,-3,15, 2,-3,-5,-2,-2,-1,-5,
    -2,-1,-2,-3,-3,-1,-2,-3, 2,-1,-1,-2, 0, 4,-3,-2,-2, 2, 8,-1,-3,-1,-2,-1,-5,
     0,-3,-3,-4,-1,-3,-3,-4,-4, 4, 1,-3, 1,-1,-3,-2, 0,-3,-1, 5,-3, 2,-3,-1,-5,
    -2,-1, 5, 6,-3, 0, 1,-1, 0,-4,-4, 0,-3,-4,-2, 0, 0,-5,-3,-3, 6,-4, 1,-1,-5,
    -2,-3,-4,-4,-2,-3,-3,-4,-3, 4, 4,-3, 2, 1,-3,-3,-1,-2,-1, 2,-4, 4,-3,-1,-5,
    -1, 0, 0, 1,-3, 4, 5,-2, 0,-3,-3, 1,-1,-4,-1, 0,-1,-2,-2,-3, 1,-3, 5,-1,-5,
    -1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-5,
    -5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5,-5, 1,
};

/* BLOSUM62, same row/col labels (similarity.h:67-93). */
static const int8_t BL62[25 * 25] = {
     4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0,-2,-1,-1,-1,-4,
    -1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3,-1,-2, 0,-1,-4,
    -2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3, 4,-3, 0,-1,-4,
    -2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3, 4,-3, 1,-1,-4,
     0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1,-3,-1,-3,-1,-4,
    -1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2, 0,-2, 4,-1,-4,
    -1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2, 1,-3, 4,-1,-4,
     0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3,-1,-4,-2,-1,-4,
    -2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3, 0,-3, 0,-1,-4,
    -1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3,-3, 3,-3,-1,-4,
    -1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4, 3,-3,-1,-4,
    -1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2, 0,-3, 1,-1,-4,
    -1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1,-3, 2,-1,-1,-4,
    -2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1,-3, 0,-3,-1,-4,
    -1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2,-2,-3,-1,-1,-4,
     1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2, 0,-2, 0,-1,-4,
     0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0,-1,-1,-1,-1,-4,
    -3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3,-4,-2,-2,-1,-4,
    -2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1,-3,-1,-2,-1,-4,
     0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4,-3, 2,-2,-1,-4,
    -2,-1, 4, 4,-3, 0, 1,-1, 0,-3,-4, 0,-3,-3,-2, 0,-1,-4,-3,-3, 4,-3, 0,-1,-4,
    -1,-2,-3,-3,-1,-2,-3,-4,-3, 3, 3,-3, 2, 0,-3,-2,-1,-2,-1, 2,-3, 3,-3,-1,-4,
    -1, 0, 0, 1,-3, 4, 4,-2, 0,-3,-3, 1,-1,-3,-1, 0,-1,-2,-2,-2, 0,-3, 4,-1,-4,
    -1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-1,-4,
    -4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4, 1,
};

enum { SIM_NA = 0, SIM_BLOSUM50 = 1, SIM_BLOSUM62 = 2, SIM_SIMPLE = 3 };
enum { DT_NA = 0, DT_AA = 1, DT_STRING = 2 };

/* Matrix side length the oracle uses for each similarity (17 for NA: the
 * reference's 16 plus the unknown code 16, see header). */
int oracle_matrix_dim(int similarity) {
    switch (similarity) {
    case SIM_NA: return 17;
    case SIM_BLOSUM50:
    case SIM_BLOSUM62: return 25;
    default: return -1;
    }
}

/* selectMatrix (similarity.cpp:89-104) with scaleMatrix (:19-33): NA entries
 * become match/mismatch, then row 0 and column 0 ('N') become ambiguity.
 * Writes dim*dim int8 entries row-major (row = A code, col = B code).
 * Returns dim, or -1 for `simple` (the reference throws there). */
int oracle_select_matrix(int similarity, int match, int mismatch, int ambiguity,
                         int8_t *out) {
    int dim = oracle_matrix_dim(similarity);
    if (dim < 0) return -1;
    if (similarity == SIM_NA) {
        for (int r = 0; r < 17; ++r)
            for (int c = 0; c < 17; ++c) {
                int8_t v;
                if (r == 16 || c == 16) v = (int8_t)ambiguity;
                else v = NA_BIN[r * 16 + c] ? (int8_t)match : (int8_t)mismatch;
                if (r == 0 || c == 0) v = (int8_t)ambiguity;
                out[r * 17 + c] = v;
            }
        return 17;
    }
    memcpy(out, similarity == SIM_BLOSUM50 ? BL50 : BL62, 25 * 25);
    return 25;
}

/* Character -> code tables (encoding.h:27-73; Encoding ctor encoding.cpp:40-48:
 * 127 entries defaulting to the code of '\0', upper and lower case keyed). */
static void build_code_table(int datatype, uint8_t table[256]) {
    static const char NA_KEYS[] = "NATCGURYKMSWBDHV";
    static const char AA_KEYS[] = "ARNDCQEGHILKMFPSTWYVBZX*";
    if (datatype == DT_STRING) {
        for (int i = 0; i < 256; ++i) table[i] = (uint8_t)(i < 127 ? i : 0);
        return;
    }
    const char *keys = datatype == DT_NA ? NA_KEYS : AA_KEYS;
    uint8_t unknown = datatype == DT_NA ? 16 : 24;
    for (int i = 0; i < 256; ++i) table[i] = unknown;
    for (int k = 0; keys[k]; ++k) {
        unsigned char c = (unsigned char)keys[k];
        table[c] = (uint8_t)k;
        if (c >= 'A' && c <= 'Z') table[c - 'A' + 'a'] = (uint8_t)k;
    }
}

void oracle_encode(int datatype, const char *s, int64_t len, uint8_t *out) {
    uint8_t table[256];
    build_code_table(datatype, table);
    for (int64_t i = 0; i < len; ++i) out[i] = table[(unsigned char)s[i]];
}

static inline int imax(int a, int b) { return a > b ? a : b; }

/*
 * One comparison, exactly as SWAffine::compute does it for one slot
 * (codelet_batch_affine.cpp:44-124): row loop over B outside, column loop
 * over A inside, C[] / bG[] indexed by the A position, first strict
 * improvement wins, positions held as uint16.
 * `work` must hold 2*m ints.
 */
void oracle_sw_pair(const uint8_t *a, int m, const uint8_t *b, int n,
                    const int8_t *mat, int dim, int gI, int gE, int forward_only,
                    int *work, int32_t *score, int32_t *arange, int32_t *brange) {
    uint16_t Astart = 0, Bstart = 0, Aend = 0, Bend = 0;
    int s = 0;
    if (m == 0 || n == 0) {
        *score = 0;
        *arange = 0;
        *brange = 0;
        return;
    }
    int *C = work, *bG = work + m;
    memset(C, 0, sizeof(int) * (size_t)m);
    memset(bG, 0, sizeof(int) * (size_t)m);
    for (int i = 0; i < n; ++i) { /* :66-85 */
        int lastNoGap = 0, prevNoGap = 0, aGap = gI;
        const int8_t *row_b = mat + b[i]; /* column b[i] of row a[j] below */
        for (int j = 0; j < m; ++j) {
            aGap = imax(lastNoGap + gI + gE, aGap + gE);
            bG[j] = imax(C[j] + gI + gE, bG[j] + gE);
            lastNoGap = imax(prevNoGap + row_b[a[j] * dim], aGap);
            lastNoGap = imax(lastNoGap, bG[j]);
            lastNoGap = imax(lastNoGap, 0);
            prevNoGap = C[j];
            C[j] = lastNoGap;
            if (lastNoGap > s) {
                Aend = (uint16_t)j;
                Bend = (uint16_t)i;
                s = lastNoGap;
            }
        }
    }
    *score = s;
    if (!forward_only) { /* :89-115 */
        int s2 = 0;
        /* The reference re-zeroes maxAB entries; only [0, Aend] are read. */
        memset(C, 0, sizeof(int) * (size_t)m);
        memset(bG, 0, sizeof(int) * (size_t)m);
        for (int i = Bend; i >= 0; --i) {
            int lastNoGap = 0, prevNoGap = 0, aGap = gI;
            const int8_t *row_b = mat + b[i];
            for (int j = Aend; j >= 0; --j) {
                aGap = imax(lastNoGap + gI + gE, aGap + gE);
                bG[j] = imax(C[j] + gI + gE, bG[j] + gE);
                lastNoGap = imax(prevNoGap + row_b[a[j] * dim], aGap);
                lastNoGap = imax(lastNoGap, bG[j]);
                lastNoGap = imax(lastNoGap, 0);
                prevNoGap = C[j];
                C[j] = lastNoGap;
                if (lastNoGap > s2) {
                    Astart = (uint16_t)j;
                    Bstart = (uint16_t)i;
                    s2 = lastNoGap;
                }
            }
        }
    }
    *arange = (int32_t)(((uint32_t)Aend << 16) | (uint32_t)Astart); /* :117-120 */
    *brange = (int32_t)(((uint32_t)Bend << 16) | (uint32_t)Bstart); /* :121-123 */
}

/*
 * A batch of comparisons over one encoded arena.  Pair k aligns
 * codes[a_off[k] .. +a_len[k]) (A) against codes[b_off[k] .. +b_len[k]) (B).
 * nthreads <= 0 uses every OpenMP thread.  Returns 0, or -1 on bad args.
 */
int oracle_sw_batch(int64_t npairs, const uint8_t *codes,
                    const int64_t *a_off, const int32_t *a_len,
                    const int64_t *b_off, const int32_t *b_len,
                    const int8_t *mat, int dim, int gI, int gE, int forward_only,
                    int nthreads, int32_t *scores, int32_t *aranges, int32_t *branges) {
    if (npairs < 0 || dim <= 0) return -1;
    int32_t maxm = 0;
    for (int64_t k = 0; k < npairs; ++k) maxm = a_len[k] > maxm ? a_len[k] : maxm;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int rc = 0;
#pragma omp parallel
    {
        int *work = (int *)malloc(sizeof(int) * 2 * (size_t)(maxm > 0 ? maxm : 1));
        if (!work) {
#pragma omp atomic write
            rc = -1;
        } else {
#pragma omp for schedule(dynamic, 16)
            for (int64_t k = 0; k < npairs; ++k)
                oracle_sw_pair(codes + a_off[k], a_len[k], codes + b_off[k], b_len[k],
                               mat, dim, gI, gE, forward_only, work,
                               &scores[k], &aranges[k], &branges[k]);
            free(work);
        }
    }
    return rc;
}

/*
 * CIGAR of one comparison's alignment, between the start and end positions a
 * compare with starts reported.  The reference's IPU path never produces
 * one; this is the definition the device traceback is checked against:
 *
 *   A[as..ae] (query) and B[bs..be] (ref) are aligned globally with affine
 *   gaps (a gap of k residues costs go + (k-1)*ge, go = -(gI + gE), ge = -gE;
 *   src/driver.hpp:47-49), and the optimal score equals the local score the
 *   compare reported (any path between the two cells is a local alignment
 *   and the reported one reaches the score).  Rows run over A, columns over
 *   B (SSW's orientation, extern/include/ssw/ssw_core.cpp:854-1053) and the
 *   traceback takes, like SSW's: the diagonal when it is at least both gap
 *   states, else the A-gap state when it is strictly larger, else the B-gap
 *   state; a gap state opens (back to H) only when opening is strictly
 *   better than extending.
 *
 * Output: runs (length << 4 | op), op 0 = M (A and B), 1 = I (A only),
 * 2 = D (B only), start to end.  Returns the number of runs, -1 if `cap` is
 * too small or an allocation fails, -2 if the region's optimum differs from
 * `score`.
 */
int oracle_cigar(const uint8_t *a, const uint8_t *b, int as, int ae, int bs, int be,
                 const int8_t *mat, int dim, int gI, int gE, int score, uint32_t *out,
                 int cap) {
    const int m = ae - as + 1, n = be - bs + 1;
    if (m <= 0 || n <= 0) return 0;
    const int go = -(gI + gE), ge = -gE;
    const long NEG = -(1L << 40);
    long *H = (long *)malloc(sizeof(long) * (size_t)(m + 1) * (n + 1));
    long *E = (long *)malloc(sizeof(long) * (size_t)(m + 1) * (n + 1));
    long *F = (long *)malloc(sizeof(long) * (size_t)(m + 1) * (n + 1));
    uint8_t *ops = (uint8_t *)malloc((size_t)(m + n) + 1);
    if (!H || !E || !F || !ops) {
        free(H); free(E); free(F); free(ops);
        return -1;
    }
#define IX(i, j) ((size_t)(i) * (n + 1) + (j)) /* i, j shifted by one: row 0 / col 0 = before */
    H[IX(0, 0)] = 0;
    E[IX(0, 0)] = F[IX(0, 0)] = NEG;
    for (int i = 1; i <= m; ++i) {
        H[IX(i, 0)] = -(long)(go + (long)(i - 1) * ge);
        E[IX(i, 0)] = H[IX(i, 0)];
        F[IX(i, 0)] = NEG;
    }
    for (int j = 1; j <= n; ++j) {
        H[IX(0, j)] = -(long)(go + (long)(j - 1) * ge);
        F[IX(0, j)] = H[IX(0, j)];
        E[IX(0, j)] = NEG;
    }
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const long e = E[IX(i - 1, j)] - ge, eo = H[IX(i - 1, j)] - go;
            const long f = F[IX(i, j - 1)] - ge, fo = H[IX(i, j - 1)] - go;
            E[IX(i, j)] = eo > e ? eo : e;
            F[IX(i, j)] = fo > f ? fo : f;
            const long d = H[IX(i - 1, j - 1)] + mat[a[as + i - 1] * dim + b[bs + j - 1]];
            long h = d;
            if (E[IX(i, j)] > h) h = E[IX(i, j)];
            if (F[IX(i, j)] > h) h = F[IX(i, j)];
            H[IX(i, j)] = h;
        }
    int rc = 0;
    if (H[IX(m, n)] != score) {
        rc = -2;
    } else {
        int i = m, j = n, st = 0, len = 0; /* st: 0 H, 1 E (A gap), 2 F (B gap) */
        while (i > 0 || j > 0) {
            if (i == 0) { ops[len++] = 2; --j; continue; }
            if (j == 0) { ops[len++] = 1; --i; continue; }
            if (st == 0) {
                const long d = H[IX(i - 1, j - 1)] + mat[a[as + i - 1] * dim + b[bs + j - 1]];
                const long e = E[IX(i, j)], f = F[IX(i, j)];
                if (d >= e && d >= f) { ops[len++] = 0; --i; --j; continue; }
                st = e > f ? 1 : 2;
            }
            if (st == 1) {
                ops[len++] = 1;
                st = (H[IX(i - 1, j)] - go > E[IX(i - 1, j)] - ge) ? 0 : 1;
                --i;
            } else {
                ops[len++] = 2;
                st = (H[IX(i, j - 1)] - go > F[IX(i, j - 1)] - ge) ? 0 : 2;
                --j;
            }
        }
        /* runs, start to end */
        int runs = 0;
        for (int k = len - 1; k >= 0;) {
            int l = k;
            while (l >= 0 && ops[l] == ops[k]) --l;
            if (runs >= cap) { rc = -1; break; }
            out[runs++] = ((uint32_t)(k - l) << 4) | ops[k];
            k = l;
        }
        if (rc == 0) rc = runs;
    }
#undef IX
    free(H); free(E); free(F); free(ops);
    return rc;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
