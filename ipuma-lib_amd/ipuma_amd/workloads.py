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


def arena_from_flat(seqs_flat: np.ndarray, lens: np.ndarray):
    """Pack sequences given back to back in `seqs_flat` (lengths `lens`) into an
    arena with 4-byte aligned offsets; returns (codes, offsets)."""
    lens = np.asarray(lens, np.int64)
    padded = (lens + 3) // 4 * 4
    offs = np.zeros(len(lens), np.int64)
    if len(lens) > 1:
        offs[1:] = np.cumsum(padded)[:-1]
    total = int(padded.sum()) + 16
    codes = np.zeros(total, np.uint8)
    src_start = np.zeros(len(lens), np.int64)
    if len(lens) > 1:
        src_start[1:] = np.cumsum(lens)[:-1]
    # destination index of every residue
    rep = np.repeat(offs - src_start, lens)
    codes[np.arange(len(seqs_flat), dtype=np.int64) + rep] = seqs_flat
    return codes, offs


def _embedded_pairs_vec(rng, n, la, lb_lo, lb_hi, p_sub, p_indel):
    """Vectorised dna_embedded_pair for n pairs with |A| = la, |B| ~ U[lb_lo, lb_hi]."""
    A = DNA_CODES[rng.integers(0, 4, (n, la))]
    u = rng.random((n, la))
    indel = u < p_indel
    sub = (~indel) & (u < p_indel + p_sub)
    M = A.copy()
    shift = rng.integers(1, 4, (n, la))
    M[sub] = DNA_CODES[(A[sub].astype(np.int64) - 1 + shift[sub]) % 4]
    ins = indel & (rng.random((n, la)) < 0.5)
    dele = indel & ~ins
    # per position: deleted -> 0 residues, inserted -> random residue then the
    # original, else the (substituted) residue
    count = np.where(dele, 0, np.where(ins, 2, 1))
    mlen = count.sum(axis=1)
    lb = np.maximum(rng.integers(lb_lo, lb_hi + 1, n), mlen)
    left = (rng.random(n) * (lb - mlen + 1)).astype(np.int64)
    right = lb - mlen - left
    # flat B: left flank, mutated copy, right flank
    cnt = count.ravel()
    flat_m = np.repeat(M.ravel(), cnt)
    # an inserted position emits a random residue, then its own residue
    first_out = np.cumsum(cnt) - cnt
    ins_at = first_out[cnt == 2]
    flat_m[ins_at] = DNA_CODES[rng.integers(0, 4, len(ins_at))]
    m_start = np.zeros(n, np.int64)
    m_start[1:] = np.cumsum(mlen)[:-1]
    B = np.empty(int(lb.sum()), np.uint8)
    b_start = np.zeros(n, np.int64)
    b_start[1:] = np.cumsum(lb)[:-1]
    B[:] = DNA_CODES[rng.integers(0, 4, len(B))]  # flanks
    dst = np.repeat(b_start + left - m_start, mlen) + np.arange(int(mlen.sum()), dtype=np.int64)
    B[dst] = flat_m
    return A, B, lb


def config2(seed: int = 2, n_short: int = 1_000_000, n_long: int = 5_000) -> dict:
    """Skewed DNA: short 150 x U[150,400] embedded pairs (B = flank + mutate(A;
    10% substitutions, 2% indels) + flank, as in config 0) plus power-law long
    pairs (pdf ~ x^-1.5 over 1-16 kb) with B = A at 85% identity (13%
    substitutions, 2% indels); pairs shuffled."""
    rng = np.random.default_rng(seed)
    A, Bflat, lb = _embedded_pairs_vec(rng, n_short, 150, 150, 400, 0.10, 0.02)
    ll = _powerlaw_len(rng, n_long, 1.5, 1000, 16000)
    longs = []
    for k in range(n_long):
        a = _dna_draw(rng, int(ll[k]))
        b = _mutate(rng, a, 0.13, 0.02, _dna_draw, _dna_subst)
        longs.append((a, b))
    n = n_short + n_long
    order = rng.permutation(n)  # position of pair k in the batch
    a_lens = np.concatenate([np.full(n_short, 150, np.int64), [len(a) for a, _ in longs]])
    b_lens = np.concatenate([lb, [len(b) for _, b in longs]])
    a_flat = np.concatenate([A.ravel()] + [a for a, _ in longs])
    b_flat = np.concatenate([Bflat] + [b for _, b in longs])
    # sequences in batch order: A_k then B_k for pair k = order^-1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    a_src = np.zeros(n, np.int64); a_src[1:] = np.cumsum(a_lens)[:-1]
    b_src = np.zeros(n, np.int64); b_src[1:] = np.cumsum(b_lens)[:-1]
    pa = inv  # batch slot -> source pair
    la_b, lb_b = a_lens[pa], b_lens[pa]
    lens = np.empty(2 * n, np.int64)
    lens[0::2], lens[1::2] = la_b, lb_b
    src = np.empty(2 * n, np.int64)
    src[0::2], src[1::2] = a_src[pa], b_src[pa] + len(a_flat)
    both = np.concatenate([a_flat, b_flat])
    gather = np.repeat(src - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + \
        np.arange(int(lens.sum()), dtype=np.int64)
    codes, offs = arena_from_flat(both[gather], lens)
    a_len = la_b.astype(np.int32); b_len = lb_b.astype(np.int32)
    return dict(name="dna_skewed_1M_plus_5k_long", sw=DNA_SW, codes=codes, a_off=offs[0::2],
                a_len=a_len, b_off=offs[1::2], b_len=b_len,
                cells=int(np.dot(a_len.astype(np.int64), b_len.astype(np.int64))), seed=seed,
                n_long=n_long)


CONFIG4_QUERY_LENGTHS = (128, 256, 512, 768, 1024, 1536, 2048, 4096)


def config4(seed: int = 4, ndb: int = 200_000, qlens=CONFIG4_QUERY_LENGTHS,
            homolog_frac: float = 0.001) -> dict:
    """One-vs-many protein search: queries of the given lengths against a
    database with lognormal(350, 0.5) lengths clipped to [30, 5000]; a fraction
    of the database are planted homologs of a random query at 30-60% identity
    (BLOSUM62-weighted substitutions, 5% indels).  All residues follow the
    BLOSUM62 background.  Returns one arena with the queries first."""
    rng = np.random.default_rng(seed)
    queries = [_aa_draw(rng, int(L)) for L in qlens]
    dl = _lognormal_len(rng, ndb, 350, 0.5, 30, 5000)
    nh = int(round(ndb * homolog_frac))
    hidx = rng.choice(ndb, nh, replace=False)
    planted = {}
    for i in hidx:
        q = queries[int(rng.integers(0, len(queries)))]
        ident = rng.uniform(0.3, 0.6)
        h = _mutate(rng, q, 1.0 - ident, 0.05, _aa_draw, _aa_subst)[:5000]
        if len(h) < 30:
            h = np.concatenate([h, _aa_draw(rng, 30 - len(h))])
        planted[int(i)] = h
        dl[i] = len(h)
    flat = _aa_draw(rng, int(dl.sum()))
    starts = np.concatenate([[0], np.cumsum(dl)[:-1]])
    for i, h in planted.items():
        flat[starts[i]:starts[i] + len(h)] = h
    qflat = np.concatenate(queries)
    lens = np.concatenate([np.array(qlens, np.int64), dl])
    codes, offs = arena_from_flat(np.concatenate([qflat, flat]), lens)
    nq = len(qlens)
    q_len = np.array(qlens, np.int32)
    db_len = dl.astype(np.int32)
    return dict(name="protein_search_8q_x_200k", sw=PROTEIN_SW, codes=codes,
                q_off=offs[:nq], q_len=q_len, db_off=offs[nq:], db_len=db_len,
                cells=int(q_len.astype(np.int64).sum() * db_len.astype(np.int64).sum()),
                homologs=np.sort(hidx), seed=seed)


def saturating_pairs(seed: int = 5, n: int = 1000, la=(2000, 4001)) -> dict:
    """DNA pairs of 2-4 kb at 95% identity, match +20 / mismatch -20, gap open
    20 / extend 2: optimal scores above 32,767, so the packed int16 kernel
    saturates on every pair and the exact int32 kernel re-runs it."""
    rng = np.random.default_rng(seed)
    pairs = []
    for _ in range(n):
        m = int(rng.integers(*la))
        a = DNA_CODES[rng.integers(0, 4, m)]
        b = a.copy()
        sub = rng.random(m) < 0.05
        b[sub] = DNA_CODES[rng.integers(0, 4, int(sub.sum()))]
        pairs.append((a, b))
    sw = dict(gapInit=-18, gapExtend=-2, matchValue=20, mismatchValue=-20, ambiguityValue=-20,
              similarity="na", datatype="na")
    return make_batch(pairs, f"dna_saturating_{n}", sw, seed=seed)


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
