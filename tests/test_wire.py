"""Prepared input buffers (the reference's wire format), host side: the
product's ipuma_prepare_remote / ipuma_prepare_remote_mn / ipuma_transfer_results
against the oracle restatement (oracle/wire.py), the oracle pinned by the
reference's own expectations (test/partition-test.cpp:36-132,
test/mn_cmp-test.cpp:15-27, test/simple-test.cpp:195-227), and the greedy fill
by properties.  No GPU: these calls are host-only."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "ipuma-lib_amd"))
import oracle as O  # noqa: E402
import wire as OW  # noqa: E402
import ipuma_amd as I  # noqa: E402

NA, AA = 0, 1


def _algo(tiles, max_batches, bufsize, max_ab=1 << 16, fill="fillfirst", programs=1):
    return I.IPUAlgoConfig(tilesUsed=tiles, maxAB=max_ab, maxBatches=max_batches,
                           bufsize=bufsize, fillAlgo=fill, transmissionPrograms=programs)


def _sw(dt):
    return I.SWConfig(datatype="na" if dt == NA else "aa",
                      similarity="na" if dt == NA else "blosum62")


def _rand_seqs(rng, n, lo, hi, dt):
    letters = list("ACGTNacgtRYX") if dt == NA else list("ARNDCQEGHILKMFPSTWYVBZX*Jo")
    return ["".join(rng.choice(letters, int(rng.integers(lo, hi + 1)))) for _ in range(n)]


def _seq_sizes(buf, tiles, max_batches, bufsize, programs=1):
    """Per tile: (comparisons, bytes used) read back from a buffer's metadata."""
    words = OW.input_words(tiles, max_batches, bufsize)
    out = []
    fill_guess = bufsize  # the metadata offset depends on the largest fill; try lot 0 first
    _, meta = OW.layout(tiles, bufsize, programs, fill_guess)
    assert meta + tiles * max_batches * 4 <= words
    for t in range(tiles):
        m = buf[meta + t * max_batches * 4: meta + (t + 1) * max_batches * 4].reshape(-1, 4)
        used = m[(m[:, 0] > 0) | (m[:, 2] > 0)]
        end = int(max(np.max(used[:, 1] + used[:, 0]), np.max(used[:, 3] + used[:, 2]))) \
            if len(used) else 0
        out.append((len(used), end))
    return out


# --- the oracle, pinned by the reference's expectations ----------------------
def test_oracle_partition_reference_cases():
    # test/partition-test.cpp:26-31, 36-56: 10 tiles x 2 slots, 10-residue pairs
    A = ["A" * 10] * 20
    for fill in (OW.fill_first, OW.round_robin):
        m = fill(10, 2, 40, A, A)
        assert [(b.max_len, b.seq_size) for b in m] == [(10, 40)] * 10
    # :70-104 partial fills
    a, b = ["AAAAAA", "AA"], ["AAAAAA", "AAA"]
    m = OW.fill_first(10, 2, 40, a, b)
    assert (m[0].max_len, m[0].seq_size) == (6, 17)
    assert all((x.max_len, x.seq_size) == (0, 0) for x in m[1:])
    m = OW.round_robin(10, 2, 40, a, b)
    assert [(x.max_len, x.seq_size) for x in m[:2]] == [(6, 12), (3, 5)]
    assert all((x.max_len, x.seq_size) == (0, 0) for x in m[2:])
    # :106-132 compression of repeated sequences (one tile, 5 slots, 100 bytes)
    seqs, cmps = ["AAAA"] * 3, [(0, 1), (0, 2), (0, 2)]
    for fill in (OW.fill_first_mn, OW.round_robin_mn):
        assert fill(1, 5, 100, seqs, cmps)[0].seq_size == 12
    # test/mn_cmp-test.cpp:8-27: {2, 30, 10, 1000}, one comparison {0, 1}
    m = OW.fill_first_mn(2, 10, 1000, ["AAAA", "TTTT", "GGGG", "CCCC"], [(0, 1)])
    assert len(m) == 2 and len(m[0].cmps) == 1 and len(m[1].cmps) == 0
    assert m[0].cmps[0][0] == 0 and m[0].cmps[0][2] == 0 and m[0].cmps[0][4] == 4


def test_oracle_layout_closed_forms():
    # getInputBufferSize32b (src/ipu_config.cpp:6-14): 10 tiles x ceil(40/4) + 10 x 2 x 4
    assert OW.input_words(10, 2, 40) == 180
    assert OW.input_words(3, 5, 1001) == 3 * 251 + 60
    # one program: tile stride ceil(bufsize/4)*4 bytes, metadata after all tiles
    assert OW.layout(3, 1001, 1, 900) == (1004, 753)
    # two programs, fill above half of bufsize: lot 1, the full layout again
    assert OW.region_index(1000, 2, 600) == 1 and OW.layout(4, 1000, 2, 600) == (1000, 1000)
    # two programs, fill at most half: lot 0, half-size tile regions
    assert OW.region_index(1000, 2, 500) == 0 and OW.layout(4, 1000, 2, 500) == (500, 500)


# --- product against the oracle ----------------------------------------------
@pytest.mark.parametrize("fill", ["fillfirst", "roundrobin"])
@pytest.mark.parametrize("dt", [NA, AA])
@pytest.mark.parametrize("programs", [1, 2])
def test_prepare_matches_oracle(fill, dt, programs):
    rng = np.random.default_rng(hash((fill, dt, programs)) % 2**32)
    tiles, mb, bufsize = 7, 5, 1001
    n = 28
    A = _rand_seqs(rng, n, 0, 60, dt)
    B = _rand_seqs(rng, n, 1, 90, dt)
    algo = _algo(tiles, mb, bufsize, fill=fill, programs=programs)
    buf, mapping, maxb = I.prepare_remote(_sw(dt), algo, A, B)
    ofill = OW.fill_first if fill == "fillfirst" else OW.round_robin
    m = ofill(tiles, mb, bufsize, A, B)
    words = OW.input_words(tiles, mb, bufsize)
    ebuf, emap, emax = OW.write_buffer(
        m, words, mb, bufsize, programs, lambda s: O.encode(dt, s),
        lambda idx, origin: A[idx] if origin == "A" else B[idx])
    assert maxb == emax
    assert np.array_equal(mapping, emap)
    assert np.array_equal(buf, ebuf)


@pytest.mark.parametrize("fill", ["fillfirst", "roundrobin"])
def test_prepare_mn_matches_oracle(fill):
    rng = np.random.default_rng(5)
    seqs = _rand_seqs(rng, 12, 1, 50, NA)
    cmps = [(int(rng.integers(0, 12)), int(rng.integers(0, 12))) for _ in range(40)]
    cmps += [(3, 3), (3, 4), (4, 3)]  # self-comparison and both orders
    tiles, mb, bufsize = 6, 9, 240
    algo = _algo(tiles, mb, bufsize, fill=fill)
    buf, mapping, maxb = I.prepare_remote_mn(_sw(NA), algo, seqs, cmps)
    ofill = OW.fill_first_mn if fill == "fillfirst" else OW.round_robin_mn
    m = ofill(tiles, mb, bufsize, seqs, cmps)
    ebuf, emap, emax = OW.write_buffer(m, OW.input_words(tiles, mb, bufsize), mb, bufsize, 1,
                                       lambda s: O.encode(NA, s), lambda idx, o: seqs[idx])
    assert maxb == emax and np.array_equal(mapping, emap) and np.array_equal(buf, ebuf)
    # dedup happened: some tile holds fewer bytes than its comparisons' sequences
    assert sum(b.seq_size for b in m) < sum(len(seqs[a]) + len(seqs[b]) for a, b in cmps)


def _decode_check(buf, mapping, A, B, tiles, mb, bufsize, dt):
    """Every comparison's metadata points at its own encoded A and B."""
    maxfill = 0
    stride, meta = OW.layout(tiles, bufsize, 1, bufsize)
    raw = buf.view(np.uint8)
    seen = set()
    for k, slot in enumerate(mapping):
        assert slot not in seen
        seen.add(int(slot))
        t = slot // mb
        la, oa, lb, ob = buf[meta + slot * 4: meta + slot * 4 + 4]
        assert (la, lb) == (len(A[k]), len(B[k]))
        assert oa + la <= bufsize and ob + lb <= bufsize
        assert np.array_equal(raw[t * stride + oa: t * stride + oa + la], O.encode(dt, A[k]))
        assert np.array_equal(raw[t * stride + ob: t * stride + ob + lb], O.encode(dt, B[k]))
        maxfill = max(maxfill, oa + la, ob + lb)
    return maxfill


def test_greedy_properties():
    # test/partition-test.cpp:58-68: full buckets, every tile 40 bytes
    A = ["A" * 10] * 20
    algo = _algo(10, 2, 40, max_ab=10, fill="greedy")
    buf, mapping, maxb = I.prepare_remote(_sw(NA), algo, A, A)
    assert maxb == 40
    assert _seq_sizes(buf, 10, 2, 40) == [(2, 40)] * 10
    # random batch: one slot each, capacity kept, metadata decodes to the pair
    rng = np.random.default_rng(11)
    A = _rand_seqs(rng, 60, 1, 80, AA)
    B = _rand_seqs(rng, 60, 1, 120, AA)
    algo = _algo(9, 8, 1400, fill="greedy")
    buf, mapping, maxb = I.prepare_remote(_sw(AA), algo, A, B)
    assert maxb == _decode_check(buf, mapping, A, B, 9, 8, 1400, AA)
    # the largest comparison is placed first, into tile 0 (all tiles equally light)
    big = int(np.argmax([len(a) * len(b) for a, b in zip(A, B)]))
    assert mapping[big] == 0


def test_greedy_mn_properties():
    rng = np.random.default_rng(12)
    seqs = _rand_seqs(rng, 10, 1, 60, NA)
    cmps = [(int(rng.integers(0, 10)), int(rng.integers(0, 10))) for _ in range(30)]
    algo = _algo(5, 8, 400, fill="greedy")
    buf, mapping, maxb = I.prepare_remote_mn(_sw(NA), algo, seqs, cmps)
    A = [seqs[a] for a, _ in cmps]
    B = [seqs[b] for _, b in cmps]
    assert maxb >= _decode_check(buf, mapping, A, B, 5, 8, 400, NA)
    # test/partition-test.cpp:124-132
    buf, _, maxb = I.prepare_remote_mn(_sw(NA), _algo(1, 5, 100, fill="greedy"), ["AAAA"] * 3,
                                       [(0, 1), (0, 2), (0, 2)])
    assert maxb == 12


def test_canaries():
    # test/simple-test.cpp:195-227: prepare into [begin + 1, end - 1) of a larger buffer
    tiles, mb, maxab = 10, 2, 10
    algo = _algo(tiles, mb, mb * maxab * 2, max_ab=maxab)
    A = ["A" * maxab] * (tiles * mb)
    words = I.input_buffer_words(algo)
    assert words == OW.input_words(tiles, mb, mb * maxab * 2)
    big = np.full(words + 2, 0x5EADBEEF, np.int32)
    inner = big[1:-1]
    I.prepare_remote(_sw(NA), algo, A, A, inputs=inner)
    assert big[0] == 0x5EADBEEF and big[-1] == 0x5EADBEEF


def test_transfer_results_matches_oracle():
    rng = np.random.default_rng(3)
    res = rng.integers(-1000, 1000, 3 * 40).astype(np.int32)
    mapping = rng.permutation(40)[:25].astype(np.int32)
    got = I.transfer_results(res, mapping)
    for g, e in zip(got, OW.transfer(res, mapping)):
        assert np.array_equal(g, e)
    with pytest.raises(I.IpumaError, match="3 aligned"):
        I.transfer_results(res[:-1], mapping)


def test_prepare_errors():
    algo = _algo(2, 2, 40, max_ab=10)
    with pytest.raises(I.IpumaError) as e:  # |A| > maxAB (the reference exits)
        I.prepare_remote(_sw(NA), algo, ["A" * 11], ["A"])
    assert e.value.code == I.IPUMA_ETOOLONG
    with pytest.raises(I.IpumaError) as e:  # more comparisons than tiles x slots
        I.prepare_remote(_sw(NA), algo, ["A"] * 5, ["A"] * 5)
    assert e.value.code == I.IPUMA_EINVAL
    with pytest.raises(I.IpumaError, match="Out of buckets"):
        I.prepare_remote(_sw(NA), algo, ["A" * 10] * 3, ["A" * 25] * 3)
    with pytest.raises(I.IpumaError, match="too small"):
        I.prepare_remote(_sw(NA), algo, ["A"], ["A"], inputs=np.zeros(10, np.int32))
