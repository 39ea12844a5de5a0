"""Write the dataset-file fixtures of tests/golden/datasets/: sequences in the
two formats the reference's batch driver reads (ipusw/run_comparison.hpp:22-61):
references as FASTA (multi-line records wrapped at 60 columns, as downloaded
benchmark files are) and queries as .txt, one sequence per line.  Seeded
synthetic data: DNA pairs (B embeds a mutated part of A, with lower-case runs
and IUPAC N) and BLOSUM62 protein pairs."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "datasets")


def mutate(rng, s, alphabet, p_sub, p_indel):
    out = []
    for c in s:
        r = rng.random()
        if r < p_indel / 2:
            continue  # deletion
        if r < p_indel:
            out.append(alphabet[rng.integers(len(alphabet))])  # insertion
        out.append(alphabet[rng.integers(len(alphabet))] if rng.random() < p_sub else c)
    return "".join(out)


def pairs(rng, n, alphabet, la, lb_extra, p_sub, p_indel):
    refs, queries = [], []
    for _ in range(n):
        a = "".join(alphabet[i] for i in rng.integers(0, len(alphabet), int(rng.integers(*la))))
        i0 = int(rng.integers(0, max(1, len(a) // 3)))
        core = mutate(rng, a[i0:i0 + int(rng.integers(len(a) // 3, len(a) + 1))], alphabet,
                      p_sub, p_indel)
        flank = lambda k: "".join(alphabet[i] for i in rng.integers(0, len(alphabet), k))
        b = flank(int(rng.integers(0, lb_extra))) + core + flank(int(rng.integers(0, lb_extra)))
        refs.append(a)
        queries.append(b)
    return refs, queries


def write_fasta(path, seqs, tag):
    with open(path, "w") as f:
        for k, s in enumerate(seqs):
            f.write(f">{tag}_{k} synthetic\n")
            for i in range(0, max(len(s), 1), 60):
                f.write(s[i:i + 60] + "\n")


def write_txt(path, seqs):
    with open(path, "w") as f:
        for s in seqs:
            f.write(s + "\n")


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(2024)
    refs, queries = pairs(rng, 300, "ACGT", (40, 400), 60, 0.1, 0.02)
    # lower-case runs and N in some references (the encoder folds case; N is
    # the ambiguity code)
    for k in range(0, 300, 7):
        s = list(refs[k])
        for i in range(0, len(s), 11):
            s[i] = s[i].lower()
        if len(s) > 10:
            s[5] = "N"
        refs[k] = "".join(s)
    write_fasta(os.path.join(OUT, "dna_refs.fa"), refs, "dna")
    write_txt(os.path.join(OUT, "dna_queries.txt"), queries)
    aa = "ARNDCQEGHILKMFPSTWYV"
    refs, queries = pairs(rng, 150, aa, (30, 350), 40, 0.45, 0.05)
    write_fasta(os.path.join(OUT, "aa_refs.fa"), refs, "aa")
    write_txt(os.path.join(OUT, "aa_queries.txt"), queries)


if __name__ == "__main__":
    main()
