"""The allocation-free entry (ipuma_compare_device_noalloc) through the C ABI
against the allocating one (ipuma_compare_device) on the same device inputs:
two calls in a row, the second smaller, bit-exact scores and ranges, plus a
sample against the oracle and the entry's limits."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
import ipuma_amd as I
from ipuma_amd import workloads as W

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _protein_batch(seed, n, long_pairs):
    """n protein pairs of the 100k-batch generator plus long pairs (|A| up to
    the entry's 4096, |B| past one 1,024-row super-strip, so band groups run)."""
    b = W.config1(seed=seed, n=n)
    if not long_pairs:
        return b
    rng = np.random.default_rng(seed)
    seqs = []
    for k in range(n):
        seqs.append(b["codes"][b["a_off"][k]:b["a_off"][k] + b["a_len"][k]])
        seqs.append(b["codes"][b["b_off"][k]:b["b_off"][k] + b["b_len"][k]])
    for m, nb in long_pairs:
        a = rng.integers(0, 20, m).astype(np.uint8)
        bb = rng.integers(0, 20, nb).astype(np.uint8)
        bb[: min(m, nb) // 2] = a[: min(m, nb) // 2]
        seqs += [a, bb]
    lens = np.array([len(x) for x in seqs], np.int64)
    codes, offs = W.arena_from_flat(np.concatenate(seqs), lens)
    return dict(sw=b["sw"], codes=codes, a_off=offs[0::2].copy(),
                a_len=lens[0::2].astype(np.int32), b_off=offs[1::2].copy(),
                b_len=lens[1::2].astype(np.int32))


def _run(eng, fn, batch):
    import torch
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(batch[k])).to(dev)
         for k in ("codes", "a_off", "a_len", "b_off", "b_len")}
    n = len(batch["a_len"])
    out = [torch.full((n,), -7, dtype=torch.int32, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    fn(eng._h, n, t["codes"].data_ptr(), t["codes"].numel(), t["a_off"].data_ptr(),
       t["a_len"].data_ptr(), t["b_off"].data_ptr(), t["b_len"].data_ptr(),
       *[o.data_ptr() for o in out], None)
    return out


def test_noalloc_matches_compare_device(gpu):
    L = I.lib()
    batch1 = _protein_batch(71, 60_000, [(4096, 2500), (3000, 5000), (1500, 1400)])
    batch2 = _protein_batch(72, 3_000, [(800, 1200)])
    eng = I.SWAlgorithm(I.SWConfig(**batch1["sw"]),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=True))
    mat = O.select_matrix(O.SIM_BLOSUM62)
    for label, batch in (("first call, 60,003 pairs", batch1), ("second call, 3,001 pairs", batch2)):
        note(f"noalloc {label}")
        got = _run(eng, L.ipuma_compare_device_noalloc, batch)
        I._check(L.ipuma_synchronize(eng._h))
        st = eng.stats()
        assert st["pairs"] == len(batch["a_len"])
        got = [g.cpu().numpy() for g in got]
        note(f"compare_device {label}")
        ref = _run(eng, L.ipuma_compare_device, batch)
        I._check(L.ipuma_synchronize(eng._h))
        ref = [r.cpu().numpy() for r in ref]
        for g, r, name in zip(got, ref, ("scores", "a_range", "b_range")):
            bad = np.nonzero(g != r)[0]
            assert bad.size == 0, f"{label}: {name} differ at {bad.size} pairs, first {bad[:5]}"
        # a sample and every long pair against the oracle
        n = len(batch["a_len"])
        idx = np.concatenate([np.arange(min(1500, n)), np.arange(n - 4, n)])
        idx = np.unique(idx[idx >= 0])
        exp = O.sw_batch_codes(batch["codes"], batch["a_off"][idx], batch["a_len"][idx],
                               batch["b_off"][idx], batch["b_len"][idx], mat,
                               batch["sw"]["gapInit"], batch["sw"]["gapExtend"], True)
        for g, e in zip(got, exp):
            assert np.array_equal(g[idx], e)
    eng.close()


def test_noalloc_limits(gpu):
    note("noalloc limits")
    L = I.lib()
    eng = I.SWAlgorithm(I.SWConfig(similarity="blosum62", datatype="aa", gapInit=-10),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=True))
    # more comparisons than the static scratch holds: refused before any launch
    rc = L.ipuma_compare_device_noalloc(eng._h, 65537, 1, 16, 1, 1, 1, 1, 1, 1, 1, None)
    assert rc == I.IPUMA_EINVAL and b"65536" in L.ipuma_last_error()
    # an A longer than 4096: the bucketing flags it, nothing is run for it
    batch = _protein_batch(73, 100, [(4100, 300)])
    out = _run(eng, L.ipuma_compare_device_noalloc, batch)
    assert L.ipuma_synchronize(eng._h) == I.IPUMA_ETOOLONG
    s = out[0].cpu().numpy()
    assert s[-1] == -1
    ref = _run(eng, L.ipuma_compare_device, dict(batch, a_len=batch["a_len"][:-1],
                                                 a_off=batch["a_off"][:-1],
                                                 b_len=batch["b_len"][:-1],
                                                 b_off=batch["b_off"][:-1]))
    I._check(L.ipuma_synchronize(eng._h))
    assert np.array_equal(s[:-1], ref[0].cpu().numpy())
    eng.close()
    # an engine that computes start positions
    eng2 = I.SWAlgorithm(I.SWConfig(), I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16,
                                                       maxBatches=1))
    rc = L.ipuma_compare_device_noalloc(eng2._h, 0, None, 0, None, None, None, None, None, None,
                                        None, None)
    assert rc == I.IPUMA_EINVAL and b"forward-only" in L.ipuma_last_error()
    eng2.close()
