"""Pins the CPU oracle (oracle/sw_oracle.c) before it is trusted as the checker.

Each check is independent of the oracle's own recurrence: the reference's
golden rows, its MN expectations, closed forms, an exhaustive enumerator on
tiny inputs, and SSW (the reference's CPU comparator) through frozen fixtures.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_simple_golden_rows():
    d = load("simple_test.json")
    A = [r[1] for r in d["rows"]]
    B = [r[0] for r in d["rows"]]
    s, ar, br = O.sw_batch(A, B)  # defaults = SWConfig{} (src/ipu_config.h:11-19)
    for k, (ref, query, qs, qe, rs, re_, score) in enumerate(d["rows"]):
        assert int(s[k]) == score, k
        assert O.unpack_range(ar[k]) == (qs, qe), k
        assert O.unpack_range(br[k]) == (rs, re_), k


def test_mn_expected_scores():
    d = load("mn_test.json")
    seqs = d["seqs"]
    A = [seqs[a] for a, _ in d["comparisons"]]
    B = [seqs[b] for _, b in d["comparisons"]]
    s, _, _ = O.sw_batch(A, B)
    assert list(map(int, s)) == d["expected_scores"]


@pytest.mark.parametrize("L", [1, 7, 150, 600])
def test_closed_forms(L):
    # 'A'^L vs 'A'^L -> L ending at (L-1, L-1); 'A'^L vs 'T'^L -> 0
    # (test/perf/ipuma-perf.cpp:73-144)
    s, ar, br = O.sw_batch(["A" * L, "T" * L], ["A" * L, "A" * L], forward_only=True)
    assert int(s[0]) == L and int(s[1]) == 0
    assert O.unpack_range(ar[0])[1] == L - 1 and O.unpack_range(br[0])[1] == L - 1
    assert int(ar[1]) == 0 and int(br[1]) == 0


def test_identical_protein_self_score():
    # every off-diagonal BLOSUM62 entry of a standard residue is below its
    # diagonal entries, so identity scores sum S[x][x] and ends at the corner
    rng = np.random.default_rng(3)
    mat = O.select_matrix(O.SIM_BLOSUM62)
    letters = "ARNDCQEGHILKMFPSTWYV"
    for L in (5, 64, 333):
        seq = "".join(rng.choice(list(letters), L))
        s, ar, br = O.sw_batch([seq], [seq], similarity=O.SIM_BLOSUM62, datatype=O.DT_AA,
                               gap_init=-10, gap_extend=-1)
        codes = O.encode(O.DT_AA, seq)
        assert int(s[0]) == int(sum(int(mat[c, c]) for c in codes))
        assert O.unpack_range(ar[0]) == (0, L - 1) and O.unpack_range(br[0]) == (0, L - 1)


def test_matrix_and_encoding_tables():
    na = O.select_matrix(O.SIM_NA, 2, -3, -7)
    # scaleMatrix: row/col 0 ('N') get the ambiguity value (similarity.cpp:19-33)
    assert (na[0, :] == -7).all() and (na[:, 0] == -7).all()
    assert na[1, 1] == 2 and na[1, 2] == -3  # A/A match, A/T mismatch
    assert na[1, 6] == 2  # A vs R (purine) is compatible in the IUPAC table
    for sim in (O.SIM_BLOSUM50, O.SIM_BLOSUM62):
        m = O.select_matrix(sim)
        assert (m == m.T).all()
    b62 = O.select_matrix(O.SIM_BLOSUM62)
    assert b62[0, 0] == 4 and b62[17, 17] == 11 and b62[4, 4] == 9 and b62[24, 24] == 1
    # codes (encoding.h:27-73), lower case like upper, unknown -> 16 / 24
    assert list(O.encode(O.DT_NA, "NATCGnatcgX")) == [0, 1, 2, 3, 4, 0, 1, 2, 3, 4, 16]
    assert list(O.encode(O.DT_AA, "ARV*xz?")) == [0, 1, 19, 23, 22, 21, 24]
    with pytest.raises(RuntimeError):
        O.select_matrix(O.SIM_SIMPLE)


def test_blosum50_against_reference_test_table():
    # test/parity-test.cpp:259-297 holds an independent BLOSUM50 (SSW layout,
    # own alphabet order); both tables must agree on the standard residues.
    d = load("aa_parity.json")
    mat = np.array(d["ssw_blosum50"], np.int8).reshape(24, 24)
    trans = np.array(d["ssw_aa_table"], np.int64)
    ours = O.select_matrix(O.SIM_BLOSUM50)
    for x in "ARNDCQEGHILKMFPSTWYV":
        for y in "ARNDCQEGHILKMFPSTWYV":
            a, b = O.encode(O.DT_AA, x)[0], O.encode(O.DT_AA, y)[0]
            assert ours[a, b] == mat[trans[ord(x)], trans[ord(y)]], (x, y)


def test_bruteforce_tiny():
    mat = O.select_matrix(O.SIM_NA, 1, -1, -1)
    rng = np.random.default_rng(5)
    for t in range(45):
        m, n = rng.integers(1, 5, 2)
        a = rng.choice([1, 2, 3, 4], m).astype(np.uint8)
        b = rng.choice([1, 2, 3, 4], n).astype(np.uint8)
        gI, gE = [(0, -1), (-2, -1), (-1, -2)][t % 3]
        exp = O.bruteforce_pair(a, b, mat, gI, gE)
        s, ar, br = O.sw_batch_codes(np.concatenate([a, b]), [0], [m], [m], [n], mat, gI, gE)
        assert (int(s[0]), int(ar[0]), int(br[0])) == tuple(int(x) for x in exp), (a, b, gI, gE)


def test_bruteforce_catches_planted_mistakes():
    # Plant a plausible slip in the oracle -- the gap-open term dropped, so a
    # gap costs only its extensions -- and check the enumerator rejects it.
    mat = O.select_matrix(O.SIM_NA, 1, -1, -1)
    a = np.array([4, 3, 4, 1, 1], np.uint8)
    b = np.array([4, 3, 1, 1], np.uint8)
    codes = np.concatenate([a, b])
    truth = tuple(int(x) for x in O.bruteforce_pair(a, b, mat, -2, -1))
    good = O.sw_batch_codes(codes, [0], [5], [5], [4], mat, -2, -1)
    slip = O.sw_batch_codes(codes, [0], [5], [5], [4], mat, 0, -1)  # open term dropped
    assert tuple(int(x[0]) for x in good) == truth
    assert tuple(int(x[0]) for x in slip) != truth


def test_ssw_frozen_fixtures():
    d = load("ssw_fixtures.json")
    for case in d["cases"]:
        c = case["config"]
        A = [p[0] for p in case["pairs"]]
        B = [p[1] for p in case["pairs"]]
        s, ar, br = O.sw_batch(A, B, similarity=c["similarity"], datatype=c["datatype"],
                               match=c["match"], mismatch=c["mismatch"], ambiguity=c["ambiguity"],
                               gap_init=c["gap_init"], gap_extend=c["gap_extend"])
        for k, exp in enumerate(case["expected"]):
            score, qb, qe, rb, re_ = exp
            assert int(s[k]) == score, (case["name"], k)
            if score > 0:
                assert O.unpack_range(ar[k])[1] == qe, (case["name"], k)
                assert O.unpack_range(br[k])[1] == re_, (case["name"], k)
                if case.get("check_begin", True):
                    assert O.unpack_range(ar[k])[0] == qb, (case["name"], k)
                    assert O.unpack_range(br[k])[0] == rb, (case["name"], k)


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref (SSW) not built")
def test_live_ssw_random_dna():
    rng = np.random.default_rng(11)
    for t in range(150):
        m, n = rng.integers(1, 120, 2)
        a = "".join(rng.choice(list("ACGT"), m))
        b = "".join(rng.choice(list("ACGT"), n))
        if t % 2:
            b = b[: n // 2] + a + b[n // 2:]
        sc, qb, qe, rb, re_ = O.ssw_dna(a, b, 1, 1, 3, 1, 1)
        s, ar, br = O.sw_batch([a], [b], gap_init=-2, gap_extend=-1)
        assert int(s[0]) == sc
        if sc > 0:
            assert O.unpack_range(ar[0]) == (qb, qe) and O.unpack_range(br[0]) == (rb, re_)


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (needs the reference tree)")
def test_ssw_batch_scores_equal_oracle():
    """The reference's SSW comparator over a coded batch (the CPU baseline bench.py
    times) gives the oracle's scores: protein and DNA, gap settings mapped as SSW's
    positive open (first gap residue) and extend penalties."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                                    "ipuma-lib_amd"))
    from ipuma_amd import workloads as W
    for b, mat in ((W.config1(n=400, seed=3), O.select_matrix(O.SIM_BLOSUM62)),
                   (W.config0(seed=4), O.select_matrix(O.SIM_NA, 1, -1, -1))):
        sw = b["sw"]
        s = O.ssw_batch_codes(b["codes"], b["a_off"], b["a_len"], b["b_off"], b["b_len"], mat,
                              -(sw["gapInit"] + sw["gapExtend"]), -sw["gapExtend"], 4)
        e, _, _ = O.sw_batch_codes(b["codes"], b["a_off"], b["a_len"], b["b_off"], b["b_len"],
                                   mat, sw["gapInit"], sw["gapExtend"], True)
        assert np.array_equal(s, e)
