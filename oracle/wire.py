"""Oracle of the prepared-buffer wire format (bucket fills and buffer layout).

TEST INFRASTRUCTURE ONLY: only tests/ may import this module, as the checker
of the product's ipuma_prepare_remote / ipuma_prepare_remote_mn /
ipuma_transfer_results.

Plain-Python restatement of the reference's host-side preparation:

* ``fill_first`` / ``round_robin`` on A/B lists -- src/partition.cpp:197-224
  (fillFirst) and :244-269 (roundRobin): search from the current bucket,
  wrapping, for room for one more comparison (cmps + 1 <= maxBatches) and its
  bytes (seqSize + |A| + |B| <= bufsize); A then B appended at seqSize;
  round robin advances the current bucket by one after each placement.
* ``fill_first_mn`` / ``round_robin_mn`` on comparison lists -- getNextBucket
  and addCmpToBucket (src/partition.cpp:105-155), fillFirst (:157-171) and
  roundRobin (:226-242): a sequence already in the bucket adds no bytes and
  is referenced at its stored offset (the last matching entry).
* ``layout`` / ``write_buffer`` -- fill_input_buffer (src/ipu_batch_affine.cpp:
  752-889) with getBufsize32b / getTotalBufsize32b (src/ipu_config.cpp:6-14)
  and calculate_slot_region_index (:976-980).
* ``transfer`` -- transferResults (src/ipu_batch_affine.cpp:674-695).

The greedy fill (src/partition.cpp:287-462) breaks ties through std::sort and
std::priority_queue; it is not restated here, and the tests check the
product's greedy maps by properties instead (capacity, one slot per
comparison, metadata that decodes to the comparison's own sequences, and the
reference's own expectations, test/partition-test.cpp:58-68).

Pins: test/partition-test.cpp:36-132, test/mn_cmp-test.cpp:15-27 and the
canaries of test/simple-test.cpp:195-227 (see tests/test_wire.py).
"""
from __future__ import annotations

import math

import numpy as np


class Bucket:
    def __init__(self, index: int):
        self.index = index
        self.max_len = 0
        self.seq_size = 0
        self.seqs: list[tuple[int, int, str]] = []   # (index, offset, origin "A"|"B"|"U")
        self.cmps: list[tuple[int, int, int, int, int]] = []  # (cmp, |A|, offA, |B|, offB)


def new_map(tiles: int):
    return [Bucket(i) for i in range(tiles)]


def _find(buckets, cur, ok):
    nb = len(buckets)
    for off in range(nb):
        w = (cur + off) % nb
        if ok(buckets[w]):
            return w
    raise RuntimeError("Out of buckets.")


def _place_ab(b: Bucket, ci: int, la: int, lb: int):
    b.max_len = max(b.max_len, la, lb)
    oa, ob = b.seq_size, b.seq_size + la
    b.seqs += [(ci, oa, "A"), (ci, ob, "B")]
    b.cmps.append((ci, la, oa, lb, ob))
    b.seq_size += la + lb


def _fill_ab(buckets, A, B, max_batches, bufsize, advance):
    cur = 0
    for i, (a, b) in enumerate(zip(A, B)):
        la, lb = len(a), len(b)
        w = _find(buckets, cur,
                  lambda x: len(x.cmps) + 1 <= max_batches and x.seq_size + la + lb <= bufsize)
        _place_ab(buckets[w], i, la, lb)
        cur = w + 1 if advance else w
    return buckets


def fill_first(tiles, max_batches, bufsize, A, B):
    return _fill_ab(new_map(tiles), A, B, max_batches, bufsize, advance=False)


def round_robin(tiles, max_batches, bufsize, A, B):
    return _fill_ab(new_map(tiles), A, B, max_batches, bufsize, advance=True)


def _fill_mn(buckets, seqs, cmps, max_batches, bufsize, advance):
    cur = 0
    for i, (ia, ib) in enumerate(cmps):
        la, lb = len(seqs[ia]), len(seqs[ib])
        need = {}

        def ok(x):
            ea, eb, oa, ob = la, lb, 0, 0
            for (idx, off, _) in x.seqs:
                if idx == ia:
                    ea, oa = 0, off
                if idx == ib:
                    eb, ob = 0, off
            need[x.index] = (ea, eb, oa, ob)
            return len(x.cmps) + 1 <= max_batches and x.seq_size + ea + eb <= bufsize

        w = _find(buckets, cur, ok)
        b = buckets[w]
        ea, eb, oa, ob = need[w]
        b.max_len = max(b.max_len, la, lb)
        if ea:
            oa = b.seq_size
            b.seqs.append((ia, oa, "U"))
        if eb:
            ob = b.seq_size + ea
            b.seqs.append((ib, ob, "U"))
        b.cmps.append((i, la, oa, lb, ob))
        b.seq_size += ea + eb
        cur = w + 1 if advance else w
    return buckets


def fill_first_mn(tiles, max_batches, bufsize, seqs, cmps):
    return _fill_mn(new_map(tiles), seqs, cmps, max_batches, bufsize, advance=False)


def round_robin_mn(tiles, max_batches, bufsize, seqs, cmps):
    return _fill_mn(new_map(tiles), seqs, cmps, max_batches, bufsize, advance=True)


def input_words(tiles, max_batches, bufsize):
    """getInputBufferSize32b: tiles * ceil(bufsize / 4) + tiles * maxBatches * 4."""
    return tiles * math.ceil(bufsize / 4) + tiles * max_batches * 4


def region_index(bufsize, programs, max_fill):
    return math.ceil(max_fill / bufsize * programs) - 1


def layout(tiles, bufsize, programs, max_fill):
    """(byte stride of the tile regions, word offset of the metadata)."""
    lot = region_index(bufsize, programs, max_fill)

    def scale(x):
        return (x // programs + (0 if x % programs == 0 else 1)) * (lot + 1)

    ibuf = math.ceil(bufsize / 4)
    return scale(ibuf * 4), scale(ibuf) * tiles


def write_buffer(buckets, words, max_batches, bufsize, programs, encode, seq_of):
    """The input buffer (int32 words, cleared first) and each comparison's slot."""
    buf = np.zeros(words, np.int32)
    raw = buf.view(np.int8)
    max_fill = max((b.seq_size for b in buckets), default=0)
    stride, meta = layout(len(buckets), bufsize, programs, max_fill)
    mapping = {}
    for b in buckets:
        base = b.index * stride
        for (idx, off, origin) in b.seqs:
            codes = encode(seq_of(idx, origin))
            raw[base + off: base + off + len(codes)] = codes.view(np.int8)
        for i, (ci, la, oa, lb, ob) in enumerate(b.cmps):
            m = meta + (b.index * max_batches + i) * 4
            buf[m: m + 4] = (la, oa, lb, ob)
            mapping[ci] = max_batches * b.index + i
    return buf, np.array([mapping[k] for k in sorted(mapping)], np.int32), max_fill


def transfer(results, mapping):
    results = np.asarray(results)
    if len(results) % 3:
        raise ValueError("Results buffer not 3 aligned.")
    part = len(results) // 3
    m = np.asarray(mapping)
    return results[m], results[part + m], results[2 * part + m]
