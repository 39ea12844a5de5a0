"""The traceback oracle (oracle_cigar): CIGARs that score the reported local
score and span exactly the reported region, optimal by exhaustive enumeration
on tiny pairs, and equal to the reference's CPU comparator (SSW, built from the
reference sources into oracle/_ref) wherever the optimal alignment is unique.
Host only; the device traceback is checked against this oracle in the GPU and
emulated tests."""
import os
import re
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

CASES = [(O.SIM_NA, O.DT_NA, "ACGT", -2, -1), (O.SIM_NA, O.DT_NA, "ACGT", 0, -1),
         (O.SIM_BLOSUM62, O.DT_AA, "ARNDCQEGHILKMFPSTWYV", -10, -1),
         (O.SIM_BLOSUM50, O.DT_AA, "ARNDCQEGHILKMFPSTWYV", -2, -1)]


def _pairs(rng, letters, n, lo, hi):
    A = ["".join(rng.choice(list(letters), int(rng.integers(lo, hi)))) for _ in range(n)]
    B = []
    for a in A:
        b = list(a)
        for _ in range(max(1, len(b) // 8)):  # substitutions and indels
            k = int(rng.integers(0, len(b) + 1))
            r = rng.random()
            if r < 0.4 and k < len(b):
                b[k] = str(rng.choice(list(letters)))
            elif r < 0.7:
                b.insert(k, str(rng.choice(list(letters))))
            elif k < len(b):
                del b[k]
        B.append("".join(rng.choice(list(letters), int(rng.integers(0, 6)))) + "".join(b))
    return A, B


@pytest.mark.parametrize("case", range(len(CASES)))
def test_cigar_scores_and_spans(case):
    sim, dt, letters, gi, ge = CASES[case]
    rng = np.random.default_rng(100 + case)
    A, B = _pairs(rng, letters, 60, 5, 90)
    s, ar, br = O.sw_batch(A, B, similarity=sim, datatype=dt, gap_init=gi, gap_extend=ge)
    mat = O.select_matrix(sim)
    for k in range(len(A)):
        a, b = O.encode(dt, A[k]), O.encode(dt, B[k])
        runs = O.cigar(a, b, int(ar[k]), int(br[k]), int(s[k]), mat, gi, ge)
        if s[k] <= 0:
            assert len(runs) == 0
            continue
        as_, ae = int(ar[k]) & 0xFFFF, (int(ar[k]) >> 16) & 0xFFFF
        bs, be = int(br[k]) & 0xFFFF, (int(br[k]) >> 16) & 0xFFFF
        score, ai, bj = O.cigar_score(a, b, as_, bs, runs, mat, gi, ge)
        assert (score, ai, bj) == (int(s[k]), ae, be)
        ops = [int(r) & 15 for r in runs]
        assert all(x != y for x, y in zip(ops, ops[1:]))  # runs are maximal
        assert ops[0] == 0 and ops[-1] == 0  # a local alignment starts and ends matched


def test_cigar_optimal_by_enumeration():
    # tiny pairs: the region's best global alignment (every alignment
    # enumerated) is the local score, and the CIGAR reaches it
    rng = np.random.default_rng(7)
    mat = O.select_matrix(O.SIM_NA)
    for _ in range(60):
        a = rng.integers(1, 5, int(rng.integers(1, 7))).astype(np.uint8)
        b = rng.integers(1, 5, int(rng.integers(1, 7))).astype(np.uint8)
        s, ar, br = O.sw_batch_codes(np.concatenate([a, b]), np.array([0]), np.array([len(a)], np.int32),
                                     np.array([len(a)]), np.array([len(b)], np.int32), mat, -2, -1,
                                     False)
        s, ar, br = int(s[0]), int(ar[0]), int(br[0])
        if s <= 0:
            continue
        as_, ae, bs, be = ar & 0xFFFF, ar >> 16, br & 0xFFFF, br >> 16
        assert O._best_global(list(a[as_:ae + 1]), list(b[bs:be + 1]), mat, 3, 1) == s
        runs = O.cigar(a, b, ar, br, s, mat, -2, -1)
        assert O.cigar_score(a, b, as_, bs, runs, mat, -2, -1)[0] == s


def _optimal_paths(a, b, mat, go, ge):
    """Number of optimal global alignments of a and b (affine gaps)."""
    m, n = len(a), len(b)
    NEG = -10 ** 9
    H = np.full((m + 1, n + 1), NEG, np.int64)
    E = np.full((m + 1, n + 1), NEG, np.int64)
    F = np.full((m + 1, n + 1), NEG, np.int64)
    cH = np.zeros((m + 1, n + 1), object)
    cE = np.zeros((m + 1, n + 1), object)
    cF = np.zeros((m + 1, n + 1), object)
    H[0, 0], cH[0, 0] = 0, 1
    for i in range(m + 1):
        for j in range(n + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                o, x = H[i - 1, j] - go, E[i - 1, j] - ge
                E[i, j] = max(o, x)
                cE[i, j] = (cH[i - 1, j] if o == E[i, j] else 0) + (cE[i - 1, j] if x == E[i, j] else 0)
            if j > 0:
                o, x = H[i, j - 1] - go, F[i, j - 1] - ge
                F[i, j] = max(o, x)
                cF[i, j] = (cH[i, j - 1] if o == F[i, j] else 0) + (cF[i, j - 1] if x == F[i, j] else 0)
            d = H[i - 1, j - 1] + int(mat[a[i - 1], b[j - 1]]) if i > 0 and j > 0 else NEG
            h = max(d, E[i, j], F[i, j])
            H[i, j] = h
            cH[i, j] = ((cH[i - 1, j - 1] if i > 0 and j > 0 and d == h else 0)
                        + (cE[i, j] if E[i, j] == h else 0) + (cF[i, j] if F[i, j] == h else 0))
    return H[m, n], cH[m, n]


def _as_m(cigar: str) -> str:
    """SSW reports matches and mismatches as '=' / 'X' (ssw.cpp:129-254): merge them into M."""
    out = []
    for ln, op in re.findall(r"(\d+)([MIDX=])", cigar):
        op = "M" if op in "=X" else op
        if out and out[-1][1] == op:
            out[-1][0] += int(ln)
        else:
            out.append([int(ln), op])
    return "".join(f"{n}{o}" for n, o in out)


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref (SSW from the reference) not built")
def test_cigar_matches_ssw_when_unique():
    # SSW's CIGAR (soft clips removed) equals the oracle's whenever SSW reports
    # the same region and the region has exactly one optimal alignment
    rng = np.random.default_rng(9)
    mat = O.select_matrix(O.SIM_BLOSUM62)
    tr = O.translate_table(O.DT_AA)
    A, B = _pairs(rng, "ARNDCQEGHILKMFPSTWYV", 80, 20, 120)
    s, ar, br = O.sw_batch(A, B, similarity=O.SIM_BLOSUM62, datatype=O.DT_AA, gap_init=-10,
                           gap_extend=-1)
    compared = 0
    for k in range(len(A)):
        if s[k] <= 0:
            continue
        (ss, qb, qe, rb, re_), cig = O.ssw_matrix_cigar(A[k], B[k], mat, tr, 11, 1)
        as_, ae = int(ar[k]) & 0xFFFF, (int(ar[k]) >> 16) & 0xFFFF
        bs, be = int(br[k]) & 0xFFFF, (int(br[k]) >> 16) & 0xFFFF
        if (ss, qb, qe, rb, re_) != (int(s[k]), as_, ae, bs, be):
            continue
        a, b = O.encode(O.DT_AA, A[k]), O.encode(O.DT_AA, B[k])
        best, count = _optimal_paths(list(a[as_:ae + 1]), list(b[bs:be + 1]), mat, 11, 1)
        assert best == s[k]
        if count != 1:
            continue
        core = _as_m(re.sub(r"\d+S", "", cig))
        assert O.cigar_string(O.cigar(a, b, int(ar[k]), int(br[k]), int(s[k]), mat, -10, -1)) == core
        compared += 1
    assert compared >= 20
