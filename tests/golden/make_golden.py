"""Freeze SSW results (the reference's CPU comparator) into tests/golden/ssw_fixtures.json.

Runs in the build container only: it needs oracle/_ref/libssw_ref.so, which
oracle/Makefile compiles from the reference's vendored SSW sources.  Inputs:
the reference's parity pairs (parity_inputs.json, test/parity-test.cpp:20-59,
242-253) and seeded random DNA / BLOSUM62 pairs at the benchmark gap settings.
Orientation as in the reference's parity test: A = SSW query, B = SSW ref.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def dna_case(name, pairs, gap_open, gap_extend):
    exp = [list(O.ssw_dna(a, b, 1, 1, gap_open, gap_extend, 1)) for a, b in pairs]
    return dict(name=name, pairs=[list(p) for p in pairs], expected=exp,
                config=dict(similarity=O.SIM_NA, datatype=O.DT_NA, match=1, mismatch=-1,
                            ambiguity=-1, gap_init=-(gap_open - gap_extend),
                            gap_extend=-gap_extend))


def matrix_case(name, pairs, similarity, gap_open, gap_extend, matrix, translate):
    exp = [list(O.ssw_matrix(a, b, matrix, translate, gap_open, gap_extend)) for a, b in pairs]
    return dict(name=name, pairs=[list(p) for p in pairs], expected=exp,
                config=dict(similarity=similarity, datatype=O.DT_AA, match=1, mismatch=-1,
                            ambiguity=-1, gap_init=-(gap_open - gap_extend),
                            gap_extend=-gap_extend))


def main():
    O.build()
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libssw_ref.so is not built (needs the reference tree)")
    inp = json.load(open(os.path.join(HERE, "parity_inputs.json")))
    # test_aligns_ipu(alns, refs, queries) -> A = refs, B = queries (parity-test.cpp:170, 90-93)
    dna = list(zip(inp["dna_refs"], inp["dna_queries"]))
    aa = list(zip(inp["aa_refs"], inp["aa_queries"]))
    cases = [dna_case("parity_dna_open1_ext1", dna, 1, 1),
             dna_case("parity_dna_open3_ext1", dna, 3, 1)]
    aa_tr = O.translate_table(O.DT_AA)
    cases.append(matrix_case("parity_aa_blosum50_open3_ext1", aa, O.SIM_BLOSUM50, 3, 1,
                             O.select_matrix(O.SIM_BLOSUM50), aa_tr))
    rng = np.random.default_rng(20261203)
    letters = list("ARNDCQEGHILKMFPSTWYV")
    pairs = []
    for t in range(60):
        m, n = rng.integers(20, 300, 2)
        a = "".join(rng.choice(letters, m))
        b = "".join(rng.choice(letters, n))
        if t % 2:
            cut = int(rng.integers(0, n))
            b = b[:cut] + a[int(m // 4): int(3 * m // 4)] + b[cut:]
        pairs.append((a, b))
    cases.append(matrix_case("random_blosum62_open11_ext1", pairs, O.SIM_BLOSUM62, 11, 1,
                             O.select_matrix(O.SIM_BLOSUM62), aa_tr))
    pairs = []
    for t in range(60):
        m, n = rng.integers(20, 300, 2)
        a = "".join(rng.choice(list("ACGT"), m))
        b = "".join(rng.choice(list("ACGT"), n))
        if t % 2:
            cut = int(rng.integers(0, n))
            b = b[:cut] + a + b[cut:]
        pairs.append((a, b))
    cases.append(dna_case("random_dna_open3_ext1", pairs, 3, 1))
    out = dict(source="SSW (extern/include/ssw, the reference's comparator) via oracle/_ref; "
                      "expected = [score, A begin, A end, B begin, B end]", cases=cases)
    with open(os.path.join(HERE, "ssw_fixtures.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", sum(len(c["pairs"]) for c in cases), "pairs")


if __name__ == "__main__":
    main()
