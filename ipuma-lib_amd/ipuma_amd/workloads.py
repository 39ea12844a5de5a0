"""Seeded synthetic workloads of the benchmark configurations.

Every generator works in residue-code space (the reference's encodings,
src/swatlib/encoding.h:27-73) and returns an arena: one uint8 code buffer with
every sequence at a 4-byte aligned offset, plus per-pair offsets and lengths.
The same arena feeds the engine's device path and the CPU oracle.

Decisions the configurations leave open, fixed here:
  * gap convention of the reference (src/driver.hpp:47-49): a gap of length k
    costs open + (k-1)*extend, so "open o, extend e" is gapInit = -(o - e),
    gapExtend = -e;
  * indels are single residues: an insertion of a random residue or a deletion,
    with equal probability;
  * substitutions never keep the residue; protein substitutions draw the new
    residue with weight q_b * 2^(S62[a][b] / 2) (BLOSUM62 is in half bits);
  * protein residues follow the BLOSUM62 background frequencies (Henikoff &
    Henikoff 1992);
  * the long-pair length law of the skewed batch is pdf ~ x^-alpha on [lo, hi].
"""
from __future__ import annotations

import numpy as np

# NA codes A=1 T=2 C=3 G=4 (src/swatlib/encoding.h:27-45)
DNA_CODES = np.array([1, 2, 3, 4], np.uint8)
NA_LETTERS = "NATCGURYKMSWBDHV"
AA_LETTERS = "ARNDCQEGHILKMFPSTWYVBZX*"

# BLOSUM62 background frequencies, residues A R N D C Q E G H I L K M F P S T W Y V
# (codes 0..19 of src/swatlib/encoding.h:47-67).
BG62 = np.array([0.074, 0.052, 0.045, 0.054, 0.025, 0.034, 0.054, 0.074, 0.026, 0.068,
                 0.099, 0.058, 0.025, 0.047, 0.039, 0.057, 0.051, 0.013, 0.032, 0.073])
BG62 = BG62 / BG62.sum()

# BLOSUM62 restricted to the 20 standard residues (src/swatlib/similarity.h:67-88).
_B62_20 = np.array([
    [4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0],
    [-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3],
    [-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3],
    [-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3],
    [0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1],
    [-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2],
    [-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2],
    [0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3],
    [-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3],
    [-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3],
    [-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1],
    [-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2],
    [-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1],
    [-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1],
    [-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2],
    [1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2],
    [0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0],
    [-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3],
    [-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1],
    [0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4]])


def _subst_cdf():
    w = BG62[None, :] * np.power(2.0, _B62_20 / 2.0)
    np.fill_diagonal(w, 0.0)
    w /= w.sum(axis=1, keepdims=True)
    return np.cumsum(w, axis=1)


_SUBST_CDF = _subst_cdf()


class Arena:
    """Sequences packed into one code buffer (4-byte aligned offsets)."""

    def __init__(self):
        self.parts = []
        self.pos = 0

    def add(self, codes: np.ndarray) -> int:
        off = self.pos
        codes = np.asarray(codes, np.uint8)
        pad = (-len(codes)) % 4
        self.parts.append(codes)
        if pad:
            self.parts.append(np.zeros(pad, np.uint8))
        self.pos += len(codes) + pad
        return off

    def finish(self) -> np.ndarray:
        self.parts.append(np.zeros(16, np.uint8))
        return np.concatenate(self.parts)


def make_batch(pairs, name: str, sw: dict, **meta) -> dict:
    ar = Arena()
    a_off, a_len, b_off, b_len = [], [], [], []
    for a, b in pairs:
        a_off.append(ar.add(a)); a_len.append(len(a))
        b_off.append(ar.add(b)); b_len.append(len(b))
    codes = ar.finish()
    a_len = np.array(a_len, np.int32)
    b_len = np.array(b_len, np.int32)
    return dict(name=name, sw=sw, codes=codes, a_off=np.array(a_off, np.int64), a_len=a_len,
                b_off=np.array(b_off, np.int64), b_len=b_len,
                cells=int(np.dot(a_len.astype(np.int64), b_len.astype(np.int64))), **meta)


def _mutate(rng, seq, p_sub, p_indel, alphabet_draw, subst):
    """Single-residue substitutions and indels, vectorised over the sequence."""
    n = len(seq)
    u = rng.random(n)
    indel = u < p_indel
    sub = (~indel) & (u < p_indel + p_sub)
    out = seq.copy()
    if sub.any():
        out[sub] = subst(rng, seq[sub])
    ins = indel & (rng.random(n) < 0.5)
    dele = indel & ~ins
    keep = ~dele
    if ins.any():
        # insert a random residue before each chosen position
        idx = np.nonzero(ins)[0]
        ins_codes = alphabet_draw(rng, len(idx))
        res = np.insert(out, idx, ins_codes)
        keep = np.insert(keep, idx, True)
        return res[keep]
    return out[keep]


def _dna_draw(rng, k):
    return DNA_CODES[rng.integers(0, 4, k)]


def _dna_subst(rng, x):
    # a different base: shift the base index by 1..3
    idx = (x.astype(np.int64) - 1 + rng.integers(1, 4, len(x))) % 4
    return DNA_CODES[idx]


def _aa_draw(rng, k):
    return rng.choice(20, size=k, p=BG62).astype(np.uint8)


def _aa_subst(rng, x):
    u = rng.random(len(x))
    out = np.empty(len(x), np.uint8)
    for c in np.unique(x):
        sel = x == c
        out[sel] = np.searchsorted(_SUBST_CDF[int(c)], u[sel]).clip(0, 19)
    return out


DNA_SW = dict(gapInit=-2, gapExtend=-1, matchValue=1, mismatchValue=-1, ambiguityValue=-1,
              similarity="na", datatype="na")
PROTEIN_SW = dict(gapInit=-10, gapExtend=-1, matchValue=1, mismatchValue=-1, ambiguityValue=-1,
                  similarity="blosum62", datatype="aa")


def dna_embedded_pair(rng, la, lb):
    """B = flank + mutate(A; 10% substitutions, 2% indels) + flank, |B| = lb."""
    a = _dna_draw(rng, la)
    m = _mutate(rng, a, 0.10, 0.02, _dna_draw, _dna_subst)
    lb = max(lb, len(m))
    left = int(rng.integers(0, lb - len(m) + 1))
    b = np.concatenate([_dna_draw(rng, left), m, _dna_draw(rng, lb - len(m) - left)])
    return a, b


def config0(seed: int = 0, n: int = 1024) -> dict:
    """Plain DNA: |A| ~ U[100,150], |B| ~ U[150,300], B embeds a mutated A."""
    rng = np.random.default_rng(seed)
    pairs = [dna_embedded_pair(rng, int(rng.integers(100, 151)), int(rng.integers(150, 301)))
             for _ in range(n)]
    return make_batch(pairs, "dna_plain_1024", DNA_SW, seed=seed)


def _lognormal_len(rng, k, median, sigma, lo, hi):
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, k)), lo, hi).astype(np.int64)


def config1(seed: int = 1, n: int = 100_000) -> dict:
    """Protein: lognormal(300, 0.5) lengths clipped to [50, 1024]; half related
    (B = A at 40% identity via BLOSUM62-weighted substitutions and 5% indels)."""
    rng = np.random.default_rng(seed)
    la = _lognormal_len(rng, n, 300, 0.5, 50, 1024)
    lb = _lognormal_len(rng, n, 300, 0.5, 50, 1024)
    related = rng.random(n) < 0.5
    pairs = []
    for k in range(n):
        a = _aa_draw(rng, int(la[k]))
        if related[k]:
            b = _mutate(rng, a, 0.60, 0.05, _aa_draw, _aa_subst)
            b = b[:1024] if len(b) > 1024 else b
            if len(b) < 50:
                b = np.concatenate([b, _aa_draw(rng, 50 - len(b))])
        else:
            b = _aa_draw(rng, int(lb[k]))
        pairs.append((a, b))
    return make_batch(pairs, "protein_blosum62_100k", PROTEIN_SW, seed=seed)


def _powerlaw_len(rng, k, alpha, lo, hi):
    # inverse CDF of pdf ~ x^-alpha on [lo, hi]
    u = rng.random(k)
    a1 = 1.0 - alpha
    x = (lo ** a1 + u * (hi ** a1 - lo ** a1)) ** (1.0 / a1)
    return np.round(x).astype(np.int64)


def config2(seed: int = 2, n_short: int = 1_000_000, n_long: int = 5_000) -> dict:
    """Skewed DNA: short 150 x U[150,400] embedded pairs plus power-law long
    pairs (alpha 1.5 over 1-16 kb) at 85% identity."""
    rng = np.random.default_rng(seed)
    pairs = [dna_embedded_pair(rng, 150, int(rng.integers(150, 401))) for _ in range(n_short)]
    ll = _powerlaw_len(rng, n_long, 1.5, 1000, 16000)
    for k in range(n_long):
        a = _dna_draw(rng, int(ll[k]))
        b = _mutate(rng, a, 0.13, 0.02, _dna_draw, _dna_subst)
        pairs.append((a, b))
    order = rng.permutation(len(pairs))
    pairs = [pairs[i] for i in order]
    return make_batch(pairs, "dna_skewed_1M_plus_5k_long", DNA_SW, seed=seed)


def random_pairs(seed, n, la_range, lb_range, protein=False, related=0.5) -> dict:
    """Small random mix used by the parity tests."""
    rng = np.random.default_rng(seed)
    pairs = []
    for _ in range(n):
        la = int(rng.integers(la_range[0], la_range[1] + 1))
        lb = int(rng.integers(lb_range[0], lb_range[1] + 1))
        if protein:
            a = _aa_draw(rng, la)
            b = _mutate(rng, a, 0.6, 0.05, _aa_draw, _aa_subst) if rng.random() < related \
                else _aa_draw(rng, lb)
        else:
            if rng.random() < related and lb >= 1:
                a, b = dna_embedded_pair(rng, la, lb)
            else:
                a, b = _dna_draw(rng, la), _dna_draw(rng, lb)
        pairs.append((a, b))
    return make_batch(pairs, f"random_{'aa' if protein else 'na'}_{n}",
                      PROTEIN_SW if protein else DNA_SW, seed=seed)


def decode(codes: np.ndarray, protein: bool) -> str:
    letters = AA_LETTERS if protein else NA_LETTERS
    return "".join(letters[c] for c in codes)
