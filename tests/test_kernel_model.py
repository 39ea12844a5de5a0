"""CPU model of the packed kernel's schedule, checked against the oracle.

It replays, for one comparison, what k_fast does lane by lane: super-strips of
L*R rows, lane k computing column j = t - k at step t, H/F/code handed to lane
k+1 each step, the row above each super-strip read from / written to a
boundary row, every value held plus 1,024 (so that H, E and F are positive
normal fp16 bit patterns and H = max3(diag + s, E, F), E = max3(H - o, E - x,
1024) are fp16 maxima of three, NaN-propagating, on those patterns), E and F
clamped at 0 with unsigned saturating subtraction, the substitution add
saturating at 32767 (a biased H at 0x7C00 or above flags the comparison for the
int32 re-run), pad code scores of -32768, per-row keys
H << 16 | (0xFFFF - j) -- or, for pipelines of up to 16 lanes, one key per
row pair, H << 16 | (even row) << 15 | (0x7FFF - j) -- rows merged with a
strict > and lanes by (max H, min row).  With tagkey, a super-strip whose H
values are bounded below 4096 (the best H of the lanes so far plus the
largest score times the super-strip's rows) takes row-tag keys instead: per
step and lane the max over rows of H * 16 + (15 - r), merged over the steps as
that << 16 | (0xFFFF - j); the model asserts the bound on every cell.  Agreement with the oracle on random inputs -- including pipeline fill
and drain, padded rows and columns, and the saturation flag -- checks the
algebra of the kernel before it ever runs on a device.
"""
import numpy as np
import pytest

import oracle as O

BIAS = 1024
SAT = 0x7C00 - BIAS  # the re-run threshold of a comparison's maximum


def max3(*vals):
    """v_pk_maximum3_f16 on one half: IEEE maximum of the 16-bit patterns read
    as fp16 (a NaN operand gives the quiet NaN 0x7E00)."""
    f = np.array([v & 0xFFFF for v in vals], np.uint16).view(np.float16)
    if np.isnan(f).any():
        return 0x7E00
    return int(np.array([f.max()], np.float16).view(np.uint16)[0])


def model(a, b, mat, gI, gE, L, R, mpad=None, pairkey=False, tagkey=False, force_tag=False):
    m, n = len(a), len(b)
    o, x = -(gI + gE), -gE
    dim = mat.shape[0]
    pad = dim
    table = np.full((dim + 1, dim + 1), -32768, np.int64)
    table[:dim, :dim] = mat
    mpad = m if mpad is None else mpad
    nss = -(-n // (L * R))
    bnd_h = np.full(mpad, BIAS, np.int64)
    bnd_f = np.full(mpad, BIAS, np.int64)
    best = [(0, 0, 0)] * L  # per lane (h, i, j)
    smax = max(0, int(np.max(mat)))
    tags_used = 0

    def subsat(v, c):
        return max(v - c, 0)

    for ss in range(nss):
        rows = [[(ss * L + k) * R + r for r in range(R)] for k in range(L)]
        bcode = [[b[i] if i < n else pad for i in rows[k]] for k in range(L)]
        H = np.full((L, R), BIAS, np.int64)
        E = np.full((L, R), BIAS, np.int64)
        key = np.zeros((L, R), np.int64)
        pkey = np.zeros((L, (R + 1) // 2), np.int64)
        use_tag = force_tag or (tagkey and R <= 16 and
                                max(bh for bh, _, _ in best) + smax * R * L <= 4095)
        tags_used += use_tag
        tkey = [0] * L
        out_h = [BIAS] * L
        out_f = [BIAS] * L
        col = [pad] * L
        hdiag = [BIAS] * L
        new_bh = bnd_h.copy()
        new_bf = bnd_f.copy()
        T = mpad + L - 1
        T = -(-T // 4) * 4
        for t in range(T):
            prev_h, prev_f, prev_c = out_h[:], out_f[:], col[:]
            for k in range(L):
                if k == 0:
                    rh = bnd_h[t] if (ss > 0 and t < mpad) else BIAS
                    rf = bnd_f[t] if (ss > 0 and t < mpad) else BIAS
                    col[k] = a[t] if t < m else pad
                else:
                    rh, rf = prev_h[k - 1], prev_f[k - 1]
                    col[k] = prev_c[k - 1]
                j = t - k
                jinv = (0xFFFF - j) & 0xFFFF
                j15 = (0x7FFF - j) & 0x7FFF
                d = hdiag[k]
                hdiag[k] = rh
                f = rf
                k16 = 0
                for r in range(R):
                    s = table[col[k], bcode[k][r]]
                    # saturating int16 add of the biased diagonal, as a bit pattern
                    t2 = (min(max(int(np.int16(d)) + s, -32768), 32767)) & 0xFFFF
                    hn = max3(t2, E[k, r], f)
                    ho = subsat(hn, o)
                    E[k, r] = max3(ho, subsat(E[k, r], x), BIAS)
                    f = max3(ho, subsat(f, x), BIAS)
                    key[k, r] = max(key[k, r], (hn << 16) | jinv)
                    pkey[k, r // 2] = max(pkey[k, r // 2],
                                          (hn << 16) | ((r % 2 == 0) << 15) | j15)
                    if use_tag:
                        assert force_tag or hn - BIAS <= 4095, (hn, ss)  # the bound
                        k16 = max(k16, (hn * 16 + 15 - r - 16 * BIAS) & 0xFFFF)
                    d = H[k, r]
                    H[k, r] = hn
                if use_tag:
                    tkey[k] = max(tkey[k], (k16 << 16) | jinv)
                out_h[k] = int(H[k, R - 1])
                out_f[k] = int(f)
                if k == L - 1 and 0 <= j < mpad:
                    new_bh[j] = out_h[k]
                    new_bf[j] = out_f[k]
        bnd_h, bnd_f = new_bh, new_bf
        for k in range(L):
            if use_tag:
                h = tkey[k] >> 20
                if h > best[k][0]:
                    best[k] = (h, rows[k][15 - ((tkey[k] >> 16) & 15)],
                               0xFFFF - (tkey[k] & 0xFFFF))
                continue
            if pairkey:
                for q in range((R + 1) // 2):
                    h = int(pkey[k, q] >> 16) - BIAS
                    if h > best[k][0]:
                        r = 2 * q + (0 if pkey[k, q] & 0x8000 else 1)
                        best[k] = (h, rows[k][r], 0x7FFF - int(pkey[k, q] & 0x7FFF))
                continue
            for r in range(R):
                h = int(key[k, r] >> 16) - BIAS
                if h > best[k][0]:
                    best[k] = (h, rows[k][r], 0xFFFF - int(key[k, r] & 0xFFFF))
    h, i, j = 0, 0, 0
    for bh, bi, bj in best:
        if bh > h or (bh == h and bi < i):
            h, i, j = bh, bi, bj
    model.tags_used = tags_used
    return h, i, j


@pytest.mark.parametrize("L,R,pairkey", [(1, 4, False), (2, 4, False), (4, 2, False),
                                          (3, 3, False), (8, 1, False), (1, 4, True),
                                          (2, 6, True), (4, 2, True)])
def test_model_matches_oracle_dna(L, R, pairkey):
    rng = np.random.default_rng(L * 10 + R)
    mat = O.select_matrix(O.SIM_NA, 1, -1, -1)
    for trial in range(25):
        m, n = rng.integers(1, 30, 2)
        a = rng.choice([1, 2, 3, 4], m).astype(np.int64)
        b = rng.choice([1, 2, 3, 4], n).astype(np.int64)
        if trial % 2:
            b = np.concatenate([b[: n // 2], a, b[n // 2:]])
        gI, gE = [(0, -1), (-2, -1), (-4, -2)][trial % 3]
        mpad = m + int(rng.integers(0, 9))  # group padding
        h, i, j = model(a, b, mat, gI, gE, L, R, mpad, pairkey)
        codes = np.concatenate([a, b]).astype(np.uint8)
        s, ar, br = O.sw_batch_codes(codes, [0], [len(a)], [len(a)], [len(b)], mat, gI, gE, True)
        assert h == int(s[0])
        assert j == O.unpack_range(ar[0])[1] if h else True
        assert i == O.unpack_range(br[0])[1] if h else True


def test_model_matches_oracle_protein():
    rng = np.random.default_rng(7)
    mat = O.select_matrix(O.SIM_BLOSUM62)
    for trial in range(20):
        m, n = rng.integers(1, 40, 2)
        a = rng.integers(0, 20, m)
        b = rng.integers(0, 20, n)
        if trial % 2:
            b = np.concatenate([b[:3], a[2:], b[3:]])
        h, i, j = model(a, b, mat, -10, -1, L=4, R=3, mpad=m + 3, pairkey=bool(trial % 2))
        codes = np.concatenate([a, b]).astype(np.uint8)
        s, ar, br = O.sw_batch_codes(codes, [0], [len(a)], [len(a)], [len(b)], mat, -10, -1, True)
        assert (h, j, i) == (int(s[0]), O.unpack_range(ar[0])[1] if h else 0,
                             O.unpack_range(br[0])[1] if h else 0)


@pytest.mark.parametrize("n", [242, 260, 300])
def test_model_flags_saturation(n):
    # an exact score at or above SAT must surface at or above SAT (the int32
    # re-run flag); 241 * 127 = 30607 stays below it and is exact
    mat = np.full((25, 25), -1, np.int64)
    np.fill_diagonal(mat, 127)
    a = np.zeros(n, np.int64)  # n * 127 >= 30734 > SAT
    h, _, _ = model(a, a, mat, -10, -1, L=2, R=8)
    assert h >= SAT
    b = np.zeros(241, np.int64)
    h, i, j = model(b, b, mat, -10, -1, L=2, R=8)
    assert (h, i, j) == (241 * 127, 240, 240)


@pytest.mark.parametrize("L,R", [(1, 16), (2, 8), (4, 4), (3, 12)])
def test_model_row_tag_keys(L, R):
    """Row-tag keys where the super-strip bound allows them, 32-bit row-pair
    keys elsewhere: short pairs (ties everywhere), and pairs scoring above
    4,095 whose later super-strips switch forms (match 7); the model asserts
    that no H of a row-tag super-strip passes 4,095."""
    rng = np.random.default_rng(100 + L * 16 + R)
    for trial in range(14):
        if trial < 2:
            match, m = 7, int(rng.integers(600, 700))
            a = rng.choice([1, 2, 3, 4], m).astype(np.int64)
            b = np.concatenate([rng.choice([1, 2, 3, 4], 5), a]).astype(np.int64)
        else:
            match = int(rng.integers(1, 4))
            m, n = rng.integers(1, 40, 2)
            a = rng.choice([1, 2, 3, 4], m).astype(np.int64)
            b = rng.choice([1, 2, 3, 4], n).astype(np.int64)
            if trial % 3 == 0:  # A = X + Y, B = Y + X: equal maxima, earlier row at the later column
                x, y = a[: m // 2], a[m // 2:]
                a, b = np.concatenate([x, y]), np.concatenate([y, x])
        mat = O.select_matrix(O.SIM_NA, match, -1, -1)
        gI, gE = -2, -1
        h, i, j = model(a, b, mat, gI, gE, L, R, len(a) + int(rng.integers(0, 5)), pairkey=True,
                        tagkey=True)
        if trial < 2:
            assert model.tags_used >= 1 and h > 4095 - match * R * L
        codes = np.concatenate([a, b]).astype(np.uint8)
        s, ar, br = O.sw_batch_codes(codes, [0], [len(a)], [len(a)], [len(b)], mat, gI, gE, True)
        assert (h, j, i) == (int(s[0]), O.unpack_range(ar[0])[1] if h else 0,
                             O.unpack_range(br[0])[1] if h else 0), (trial, match)


def model_single_strip(a, b, mat, gI, gE, L, R, mpad):
    """The kernel's rule for a group of one super-strip: row-tag keys when the
    bound holds; otherwise row-tag keys anyway, kept unless the maximum came
    within smax of 4,095, else a second run with 32-bit keys.  Returns the
    result and the forced run's maximum (None when the bound held)."""
    smax = max(0, int(np.max(mat)))
    res = model(a, b, mat, gI, gE, L, R, mpad, pairkey=True, tagkey=True)
    if model.tags_used:
        return res, None
    forced = model(a, b, mat, gI, gE, L, R, mpad, pairkey=True, force_tag=True)
    if forced[0] <= 4095 - smax:
        return forced, forced[0]
    return model(a, b, mat, gI, gE, L, R, mpad, pairkey=True), forced[0]


def test_model_single_strip_optimistic_tags():
    """Single super-strips (|B| <= L * R) whose rows' bound fails: DNA at match
    20 on 16 lanes x 16 rows.  Below the limit the forced row-tag keys are
    exact; above it (scores up to 20 x 256) the forced run's maximum comes
    within smax of 4,095, so the kernel falls back to 32-bit keys.  Either way
    the result equals the oracle's."""
    rng = np.random.default_rng(9)
    mat = O.select_matrix(O.SIM_NA, 20, -3, -1)
    fell_back = kept = 0
    for trial in range(10):
        n = int(rng.integers(100, 257))
        a = rng.choice([1, 2, 3, 4], n + 40).astype(np.int64)
        b = a[20: 20 + n] if trial % 2 == 0 else np.concatenate([rng.choice([1, 2, 3, 4], 7), a[: n - 40]])
        (h, i, j), fmax = model_single_strip(a, b, mat, -2, -1, 16, 16, len(a) + 2)
        codes = np.concatenate([a, b]).astype(np.uint8)
        s, ar, br = O.sw_batch_codes(codes, [0], [len(a)], [len(a)], [len(b)], mat, -2, -1, True)
        assert (h, j, i) == (int(s[0]), O.unpack_range(ar[0])[1], O.unpack_range(br[0])[1])
        assert fmax is not None  # 20 x 256 > 4,095: the bound never holds here
        if int(s[0]) > 4095:
            assert fmax > 4095 - 20  # the overflow is always detected
            fell_back += 1
        elif fmax <= 4095 - 20:
            kept += 1
    assert fell_back >= 2 and kept >= 2, (fell_back, kept)
