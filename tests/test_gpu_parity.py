"""Parity of the HIP engine (through the C ABI) with the oracle: bit-exact
scores and packed A/B ranges on the reference's fixtures and on seeded batches."""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
import ipuma_amd as I
from ipuma_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
BIG = I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1)
SIM = {"na": O.SIM_NA, "blosum50": O.SIM_BLOSUM50, "blosum62": O.SIM_BLOSUM62}
DT = {"na": O.DT_NA, "aa": O.DT_AA}


def note(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def oracle_batch(batch, forward_only):
    sw = batch["sw"]
    mat = O.select_matrix(SIM[sw["similarity"]], sw["matchValue"], sw["mismatchValue"],
                          sw["ambiguityValue"])
    return O.sw_batch_codes(batch["codes"], batch["a_off"], batch["a_len"], batch["b_off"],
                            batch["b_len"], mat, sw["gapInit"], sw["gapExtend"], forward_only)


def engine_for(sw: dict, forward_only=False):
    return I.SWAlgorithm(I.SWConfig(**sw), I.IPUAlgoConfig(
        tilesUsed=BIG.tilesUsed, maxAB=BIG.maxAB, maxBatches=1, forwardOnly=forward_only))


def device_run(eng, batch):
    import torch
    dev = torch.device("cuda:0")
    codes = torch.from_numpy(batch["codes"]).to(dev)
    a_off = torch.from_numpy(batch["a_off"]).to(dev)
    b_off = torch.from_numpy(batch["b_off"]).to(dev)
    a_len = torch.from_numpy(batch["a_len"]).to(dev)
    b_len = torch.from_numpy(batch["b_len"]).to(dev)
    n = len(batch["a_len"])
    s = torch.empty(n, dtype=torch.int32, device=dev)
    ar = torch.empty(n, dtype=torch.int32, device=dev)
    br = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    eng.compare_device(codes, codes.numel(), a_off, a_len, b_off, b_len, s, ar, br)
    eng.synchronize()
    return s.cpu().numpy(), ar.cpu().numpy(), br.cpu().numpy()


def assert_same(got, exp, label):
    names = ("scores", "a_range", "b_range")
    for g, e, nm in zip(got, exp, names):
        g = np.asarray(g)[: len(e)]
        bad = np.nonzero(g != e)[0]
        assert bad.size == 0, (f"{label}: {nm} differ at {bad.size} pairs, first {bad[:5]}: "
                               f"got {g[bad[:5]]} want {e[bad[:5]]}")


def test_simple_golden_rows(gpu):
    note("simple-test golden rows via compare_local")
    d = load("simple_test.json")
    A = [r[1] for r in d["rows"]]
    B = [r[0] for r in d["rows"]]
    eng = I.SWAlgorithm(I.SWConfig(), I.IPUAlgoConfig(tilesUsed=2, maxAB=300, maxBatches=20))
    eng.compare_local(A, B)
    res = eng.get_result()
    assert len(res.scores) == 40  # tilesUsed * maxBatches, like the reference
    qb, qe = I.unpack_range(res.a_range_result)
    rb, re_ = I.unpack_range(res.b_range_result)
    for k, (_, _, qs, qend, rs, rend, score) in enumerate(d["rows"]):
        assert (int(res.scores[k]), int(qb[k]), int(qe[k]), int(rb[k]), int(re_[k])) == \
            (score, qs, qend, rs, rend), k


def test_mn_comparisons(gpu):
    note("mn_cmp-test expected scores via compare_mn_local")
    d = load("mn_test.json")
    eng = I.SWAlgorithm(I.SWConfig(), I.IPUAlgoConfig(tilesUsed=1, maxAB=20, maxBatches=30))
    eng.compare_mn_local(d["seqs"], d["comparisons"])
    assert list(map(int, eng.get_result().scores[:4])) == d["expected_scores"]


def test_simple_rows_as_mn(gpu):
    # test/simple-test.cpp:150-171: the same rows through the MN path
    note("simple rows through compare_mn_local")
    d = load("simple_test.json")
    seqs, cmps = [], []
    for i, r in enumerate(d["rows"]):
        seqs += [r[1], r[0]]
        cmps.append((2 * i, 2 * i + 1))
    eng = I.SWAlgorithm(I.SWConfig(), I.IPUAlgoConfig(tilesUsed=4, maxAB=300, maxBatches=5))
    eng.compare_mn_local(seqs, cmps)
    res = eng.get_result()
    for k, r in enumerate(d["rows"]):
        assert int(res.scores[k]) == r[6]
        qb, qe = I.unpack_range(res.a_range_result[k:k + 1])
        assert (int(qb[0]), int(qe[0])) == (r[2], r[3])


SSW_CASES = [c["name"] for c in load("ssw_fixtures.json")["cases"]]
# random_blosum62_open11_ext1 once hung the GPU.  The cause, shown in the
# kernel's machine code, was the 64-lane class's next-group fetch: one lane's
# atomic read back with readfirstlane, which the compiler threaded into the
# first lane's result write, after which lanes 1..63 looped on group 0
# without lane 0.  Every lane now executes the fetch's atomic (lane 0 adds 1,
# the others 0) with no lane-conditional branch around it (next_group in
# sw_kernels.hip); this case is the regression test of that fix.


@pytest.mark.parametrize("name", SSW_CASES)
def test_ssw_fixtures(gpu, name):
    d = load("ssw_fixtures.json")
    names = {0: "na", 1: "blosum50", 2: "blosum62"}
    case = next(c for c in d["cases"] if c["name"] == name)
    note(f"ssw fixture {case['name']}")
    c = case["config"]
    sw = dict(gapInit=c["gap_init"], gapExtend=c["gap_extend"], matchValue=c["match"],
              mismatchValue=c["mismatch"], ambiguityValue=c["ambiguity"],
              similarity=names[c["similarity"]], datatype="na" if c["datatype"] == 0 else "aa")
    eng = engine_for(sw)
    A = [p[0] for p in case["pairs"]]
    B = [p[1] for p in case["pairs"]]
    s, ar, br = eng.compare(A, B)
    qb, qe = I.unpack_range(ar)
    rb, re_ = I.unpack_range(br)
    for k, (score, eqb, eqe, erb, ere) in enumerate(case["expected"]):
        assert int(s[k]) == score, (case["name"], k)
        if score > 0:
            assert (int(qb[k]), int(qe[k]), int(rb[k]), int(re_[k])) == (eqb, eqe, erb, ere), \
                (case["name"], k)


@pytest.mark.parametrize("protein", [False, True])
@pytest.mark.parametrize("mode", [I.BATCH_BUCKETED, I.BATCH_PLAIN, I.BATCH_PLAIN_TILE])
@pytest.mark.parametrize("forward_only", [True, False])
def test_random_batches(gpu, protein, mode, forward_only):
    note(f"random batch protein={protein} mode={mode} forward_only={forward_only}")
    batch = W.random_pairs(100 + protein * 7 + mode * 3, 700, (1, 400), (1, 400), protein=protein)
    eng = engine_for(batch["sw"], forward_only)
    eng.set_batching(mode)
    got = device_run(eng, batch)
    assert_same(got, oracle_batch(batch, forward_only), "random")


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 64])
def test_every_pipeline_depth(gpu, lanes):
    note(f"lanes per pair {lanes}")
    batch = W.random_pairs(300 + lanes, 300, (1, 700), (1, 1500), protein=lanes % 2 == 0)
    eng = engine_for(batch["sw"], False)
    eng.set_lanes_per_pair(lanes)
    got = device_run(eng, batch)
    assert_same(got, oracle_batch(batch, False), f"lanes={lanes}")


def test_config0_bucketed_and_plain(gpu):
    note("dna plain config (1024 pairs) bucketed vs plain vs oracle")
    batch = W.config0()
    exp = oracle_batch(batch, False)
    for mode in (I.BATCH_BUCKETED, I.BATCH_PLAIN, I.BATCH_PLAIN_TILE):
        eng = engine_for(batch["sw"], False)
        eng.set_batching(mode)
        assert_same(device_run(eng, batch), exp, f"config0 mode={mode}")


def test_skewed_slice_all_modes(gpu):
    note("skewed DNA slice (3000 short + 12 long pairs) in every batching mode")
    batch = W.config2(seed=5, n_short=3000, n_long=12)
    exp = oracle_batch(batch, True)
    for mode in (I.BATCH_BUCKETED, I.BATCH_PLAIN_TILE, I.BATCH_PLAIN):
        eng = engine_for(batch["sw"], True)
        eng.set_batching(mode)
        assert_same(device_run(eng, batch), exp, f"skewed mode={mode}")


def subset(batch, idx):
    return dict(batch, a_off=batch["a_off"][idx], a_len=batch["a_len"][idx],
                b_off=batch["b_off"][idx], b_len=batch["b_len"][idx])


def test_protein_full_size_with_starts(gpu):
    # the whole 100,000-pair protein batch (1.29e10 cells), scores and both
    # ranges of every pair against the oracle
    note("protein batch at full size (100,000 pairs) with starts vs oracle")
    batch = W.config1()
    assert len(batch["a_len"]) == 100_000
    eng = engine_for(batch["sw"], False)
    assert_same(device_run(eng, batch), oracle_batch(batch, False), "protein full size")
    eng.close()


def test_skewed_full_size(gpu):
    # the whole skewed DNA batch (1,000,000 short + 5,000 long pairs): the
    # bucketed and the plain launch agree on every pair, and the 40 longest
    # pairs plus 3,000 random short ones agree with the oracle
    note("skewed DNA batch at full size (1,005,000 pairs): bucketed vs plain vs oracle sample")
    batch = W.config2()
    n = len(batch["a_len"])
    assert n == 1_005_000
    eng = engine_for(batch["sw"], True)
    got = device_run(eng, batch)
    eng.set_batching(I.BATCH_PLAIN)
    assert_same(device_run(eng, batch), got, "skewed plain vs bucketed")
    eng.close()
    cells = batch["a_len"].astype(np.int64) * batch["b_len"].astype(np.int64)
    longest = np.argsort(cells)[-40:]
    short = np.random.default_rng(41).choice(np.nonzero(batch["a_len"] == 150)[0], 3000,
                                             replace=False)
    idx = np.concatenate([longest, short])
    exp = oracle_batch(subset(batch, idx), True)
    assert_same([np.asarray(x)[idx] for x in got], exp, "skewed sample vs oracle")


def test_protein_slice_forward(gpu):
    note("protein config slice (4000 pairs) forward only")
    batch = W.config1(n=4000)
    eng = engine_for(batch["sw"], True)
    assert_same(device_run(eng, batch), oracle_batch(batch, True), "protein slice")


def test_host_staging_large(gpu):
    # host-buffer entry on a batch above the 8 MB line where the pinned staging
    # copy is gathered in slices by several threads, each slice sent while the
    # next is gathered: same results as the device-resident entry, and the first
    # 1500 pairs against the oracle
    note("host-buffer staging: protein batch of 30000 pairs (about 20 MB)")
    batch = W.config1(n=30000)
    letters = np.frombuffer(b"ARNDCQEGHILKMFPSTWYVBZX*", np.uint8)
    dec = np.zeros(256, np.uint8)
    dec[: len(letters)] = letters
    codes = batch["codes"]
    A = [dec[codes[o:o + l]].tobytes() for o, l in zip(batch["a_off"], batch["a_len"])]
    B = [dec[codes[o:o + l]].tobytes() for o, l in zip(batch["b_off"], batch["b_len"])]
    assert sum(map(len, A)) + sum(map(len, B)) > 8 << 20
    eng = engine_for(batch["sw"], False)
    got = eng.compare(A, B)
    assert_same(got, device_run(eng, batch), "host staging vs device entry")
    sw = batch["sw"]
    exp = O.sw_batch([a.decode() for a in A[:1500]], [b.decode() for b in B[:1500]],
                     similarity=SIM[sw["similarity"]], datatype=DT[sw["datatype"]],
                     gap_init=sw["gapInit"], gap_extend=sw["gapExtend"])
    assert_same([np.asarray(x)[:1500] for x in got], exp, "host staging vs oracle")
    eng.close()


def test_long_pairs(gpu):
    note("long DNA pairs (2k-6k) with starts")
    rng = np.random.default_rng(9)
    pairs = []
    for L in (2000, 3500, 6000):
        a = W._dna_draw(rng, L)
        pairs.append((a, W._mutate(rng, a, 0.13, 0.02, W._dna_draw, W._dna_subst)))
    batch = W.make_batch(pairs, "long", W.DNA_SW)
    eng = engine_for(batch["sw"], False)
    assert_same(device_run(eng, batch), oracle_batch(batch, False), "long")


def test_band_groups(gpu):
    # whole-wave groups of two or more super-strips are shared by the waves of
    # a block: thousands of them (several per block, so the block's group
    # slots are reused) and a few very long pairs, with starts
    note("band groups: 8000 pairs |B| 1-2 kb + 40 pairs 1.5-12 kb, with starts")
    rng = np.random.default_rng(17)
    pairs = []
    for _ in range(8000):
        a = W._dna_draw(rng, int(rng.integers(40, 120)))
        left = W._dna_draw(rng, int(rng.integers(900, 2000)))
        pairs.append((a, np.concatenate([left, W._mutate(rng, a, 0.1, 0.02, W._dna_draw,
                                                         W._dna_subst)])))
    for _ in range(40):
        a = W._dna_draw(rng, int(rng.integers(1500, 12000)))
        pairs.append((a, W._mutate(rng, a, 0.15, 0.02, W._dna_draw, W._dna_subst)))
    batch = W.make_batch(pairs, "band", W.DNA_SW)
    exp = oracle_batch(batch, False)
    for lanes in (0, 64):  # the classifier's choice, then every pair on 64 lanes
        eng = engine_for(batch["sw"], False)
        if lanes:
            eng.set_lanes_per_pair(lanes)
        got = device_run(eng, batch)
        assert_same(got, exp, f"band groups, lanes={lanes}")
        assert eng.stats()["band_groups"] >= (3500 if lanes else 256), eng.stats()


def test_wide_a_key_layouts(gpu):
    # whole-wave groups keep row-pair keys (15-bit columns) while |A| <= 32704
    # and one key per row (16-bit columns) beyond; the best cell sits near the
    # end of A, so its column uses the top bits of either layout.  |B| = 200
    # runs one super-strip per wave, |B| = 1,500 a band group.
    note("wide A: |A| 32704 / 32705 / 40000 x |B| 200 / 1500, with starts")
    rng = np.random.default_rng(23)
    pairs = []
    for m in (32704, 32705, 40000):
        for nb in (200, 1500):
            a = W._dna_draw(rng, m)
            b = W._mutate(rng, a[m - nb - 5: m - 5], 0.08, 0.01, W._dna_draw, W._dna_subst)
            pairs.append((a, b))
    batch = W.make_batch(pairs, "wide", W.DNA_SW)
    eng = engine_for(batch["sw"], False)
    got = device_run(eng, batch)
    assert_same(got, oracle_batch(batch, False), "wide A")
    assert eng.stats()["band_groups"] >= 1, eng.stats()
    assert min(got[1].astype(np.uint32) >> 16) > 30000  # end columns in the top bits


def test_prepared_buffers(gpu):
    # the reference's wire format end to end: prepare_remote with each fill
    # algorithm, prepared_remote_compare, transferResults; a comparison-list
    # buffer; asynchronous jobs from two threads
    import threading
    note("prepared buffers: 5000 DNA + 2000 protein pairs, three fills, MN, async")
    rng = np.random.default_rng(23)

    def strings(n, letters, lo, hi):
        return ["".join(rng.choice(letters, int(rng.integers(lo, hi)))) for _ in range(n)]

    dna = list("ACGT")
    A = strings(5000, dna, 80, 200)
    B = [strings(1, dna, 20, 80)[0] + a + strings(1, dna, 0, 60)[0] for a in A]
    aa = list("ARNDCQEGHILKMFPSTWYV")
    PA = strings(2000, aa, 50, 400)
    PB = [a[: len(a) // 2] + strings(1, aa, 10, 200)[0] for a in PA]
    cases = [(I.SWConfig(gapInit=-2, gapExtend=-1), A, B, O.SIM_NA, O.DT_NA),
             (I.SWConfig(gapInit=-10, gapExtend=-1, similarity="blosum62", datatype="aa"), PA,
              PB, O.SIM_BLOSUM62, O.DT_AA)]
    for sw, a, b, sim, dt in cases:
        exp = O.sw_batch(a, b, similarity=sim, datatype=dt, gap_init=sw.gapInit,
                         gap_extend=sw.gapExtend)
        for fill in ("fillfirst", "roundrobin", "greedy"):
            algo = I.IPUAlgoConfig(tilesUsed=1472, maxAB=1000, maxBatches=4, bufsize=4000,
                                   fillAlgo=fill)
            eng = I.SWAlgorithm(sw, algo)
            inputs, mapping, _ = I.prepare_remote(sw, algo, a, b)
            got = I.transfer_results(eng.prepared_remote_compare(inputs), mapping)
            assert_same(got, exp, f"prepared {sw.datatype} {fill}")
            eng.close()
    # comparison list: every A against two B's
    seqs = A[:300] + B[:300]
    cmps = [(i, 300 + i) for i in range(300)] + [(i, 300 + (i + 7) % 300) for i in range(300)]
    sw = I.SWConfig(gapInit=-2, gapExtend=-1)
    algo = I.IPUAlgoConfig(tilesUsed=200, maxAB=1000, maxBatches=6, bufsize=2500)
    eng = I.SWAlgorithm(sw, algo)
    inputs, mapping, _ = I.prepare_remote_mn(sw, algo, seqs, cmps)
    got = I.transfer_results(eng.prepared_remote_compare(inputs), mapping)
    assert_same(got, O.sw_batch([seqs[x] for x, _ in cmps], [seqs[y] for _, y in cmps],
                                gap_init=-2, gap_extend=-1), "prepared mn")
    # asynchronous jobs
    algo = I.IPUAlgoConfig(tilesUsed=600, maxAB=1000, maxBatches=4, bufsize=4000,
                           forwardOnly=True)
    eng = I.SWAlgorithm(sw, algo)
    parts = [(A[i::4], B[i::4]) for i in range(4)]
    prepared = [I.prepare_remote(sw, algo, x, y) for x, y in parts]
    jobs = [None] * 4

    def submit(ix):
        for i in ix:
            jobs[i] = eng.async_submit_prepared_remote_compare(prepared[i][0])

    th = [threading.Thread(target=submit, args=(ix,)) for ix in ([0, 2], [1, 3])]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in (3, 1, 0, 2):
        got = I.transfer_results(eng.blocking_join_prepared_remote_compare(jobs[i]),
                                 prepared[i][1])
        assert_same(got, O.sw_batch(parts[i][0], parts[i][1], gap_init=-2, gap_extend=-1,
                                    forward_only=True), f"async job {i}")
    with pytest.raises(I.IpumaError):
        eng.prepared_remote_compare(prepared[0][0][:100])  # shorter than the layout
    eng.close()


def test_traceback(gpu):
    # CIGARs of a compare with starts, bit-exact against the oracle: DNA and
    # BLOSUM62 batches and a few long pairs (one wave per comparison, regions
    # of one to ten strips, 2 to 8 rows per lane)
    note("traceback: 3000 DNA + 1500 protein pairs + long pairs, CIGAR vs oracle")
    rng = np.random.default_rng(29)

    def strings(n, letters, lo, hi):
        return ["".join(rng.choice(letters, int(rng.integers(lo, hi)))) for _ in range(n)]

    dna, aa = list("ACGT"), list("ARNDCQEGHILKMFPSTWYV")
    A = strings(3000, dna, 30, 250)
    B = [strings(1, dna, 0, 50)[0] + a[5:] + strings(1, dna, 0, 30)[0] for a in A]
    for L in (3000, 5000):
        a = strings(1, dna, L, L + 1)[0]
        A.append(a)
        B.append(a[:L // 3] + "ACGTTA" + a[L // 3:2 * L // 3] + a[2 * L // 3 + 7:])
    PA = strings(1500, aa, 40, 400)
    PB = [a[: len(a) // 2] + strings(1, aa, 0, 5)[0] + a[len(a) // 2:] for a in PA]
    for sw, a, b, sim, dt in ((I.SWConfig(gapInit=-2, gapExtend=-1), A, B, O.SIM_NA, O.DT_NA),
                              (I.SWConfig(gapInit=-10, gapExtend=-1, similarity="blosum62",
                                          datatype="aa"), PA, PB, O.SIM_BLOSUM62, O.DT_AA)):
        eng = I.SWAlgorithm(sw, I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
        s, ar, br = eng.compare(a, b)
        eng.set_timing(True)
        got = eng.traceback(a, b, s, ar, br)
        mat = O.select_matrix(sim)
        for k in range(len(a)):
            exp = O.cigar(O.encode(dt, a[k]), O.encode(dt, b[k]), int(ar[k]), int(br[k]),
                          int(s[k]), mat, sw.gapInit, sw.gapExtend)
            assert np.array_equal(got[k], exp), (sw.datatype, k, I.cigar_string(got[k]),
                                                 O.cigar_string(exp))
        # the region cells and the kernel time of the call
        u = ar.astype(np.int64) & 0xFFFFFFFF
        v = br.astype(np.int64) & 0xFFFFFFFF
        cells = int(np.sum(np.where(s > 0, ((u >> 16) - (u & 0xFFFF) + 1) *
                                    ((v >> 16) - (v & 0xFFFF) + 1), 0)))
        ts = eng.traceback_stats()
        assert ts["region_cells"] == cells and ts["kernel_ms"] > 0, ts
        eng.close()


def test_traceback_int64_gaps(gpu):
    # a gap open of 20,000,000 puts the regions past the int32 bound of the
    # wave traceback: the one-lane int64 kernel takes the call
    note("traceback: gap open 20,000,000 (int64 kernel), CIGAR vs oracle")
    rng = np.random.default_rng(31)
    A = ["".join(rng.choice(list("ACGT"), int(rng.integers(20, 200)))) for _ in range(64)]
    B = [a[:len(a) // 2] + "T" + a[len(a) // 2:] if k % 2 else a[3:] for k, a in enumerate(A)]
    sw = I.SWConfig(gapInit=-20_000_000, gapExtend=-1)
    eng = I.SWAlgorithm(sw, I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    s, ar, br = eng.compare(A, B)
    got = eng.traceback(A, B, s, ar, br)
    mat = O.select_matrix(O.SIM_NA)
    for k in range(len(A)):
        exp = O.cigar(O.encode(O.DT_NA, A[k]), O.encode(O.DT_NA, B[k]), int(ar[k]), int(br[k]),
                      int(s[k]), mat, sw.gapInit, sw.gapExtend)
        assert np.array_equal(got[k], exp), (k, I.cigar_string(got[k]), O.cigar_string(exp))
    eng.close()


def test_int16_saturation_rerun(gpu):
    # W^L vs W^L under BLOSUM62 scores 11*L: 45056 at L = 4096 overflows int16
    note("int16 saturation -> int32 re-run (W^4096)")
    batch = W.make_batch([(np.full(4096, 17, np.uint8), np.full(4096, 17, np.uint8)),
                          (np.full(3000, 17, np.uint8), np.full(2990, 17, np.uint8)),
                          (np.full(100, 17, np.uint8), np.full(100, 17, np.uint8))],
                         "sat", W.PROTEIN_SW)
    eng = engine_for(batch["sw"], False)
    got = device_run(eng, batch)
    assert int(got[0][0]) == 11 * 4096
    assert eng.stats()["rerun_pairs"] >= 1
    assert_same(got, oracle_batch(batch, False), "saturation")


def test_saturation_threshold(gpu):
    # the packed kernel holds H + 1024 below 0x7C00: W^2792 (30,712) stays
    # exact in the packed kernel, W^2793 (30,723) is re-run in int32
    note("saturation threshold: W^2792 packed, W^2793 re-run")
    batch = W.make_batch([(np.full(2792, 17, np.uint8), np.full(2792, 17, np.uint8)),
                          (np.full(2793, 17, np.uint8), np.full(2793, 17, np.uint8))],
                         "thr", W.PROTEIN_SW)
    eng = engine_for(batch["sw"], False)
    got = device_run(eng, batch)
    assert [int(x) for x in got[0]] == [11 * 2792, 11 * 2793]
    assert eng.stats()["rerun_pairs"] == 1
    assert_same(got, oracle_batch(batch, False), "threshold")


@pytest.mark.parametrize("lanes", [16, 64])
def test_high_scores_near_tag_limit(gpu, lanes):
    # match 16 over |A| <= 250: maxima up to 4,000, just below the largest H a
    # row-tag key holds, on 16- and 64-lane pipelines (whose super-strips take
    # 32-bit keys: 16 x their rows exceeds it); with starts
    note(f"high scores near the row-tag limit, lanes={lanes}")
    rng = np.random.default_rng(31 + lanes)
    pairs = []
    for _ in range(120):
        a = W._dna_draw(rng, int(rng.integers(200, 251)))
        pairs.append((a, W._mutate(rng, a, 0.02, 0.01, W._dna_draw, W._dna_subst)))
        pairs.append((a, np.concatenate([W._dna_draw(rng, int(rng.integers(0, 300))), a])))
    batch = W.make_batch(pairs, "colbound", dict(W.DNA_SW, matchValue=16, mismatchValue=-3))
    eng = engine_for(batch["sw"], False)
    eng.set_lanes_per_pair(lanes)
    got = device_run(eng, batch)
    assert_same(got, oracle_batch(batch, False), f"high scores, lanes={lanes}")
    assert int(got[0].max()) >= 16 * 240


def test_single_strip_tag_fallback(gpu):
    # groups of one 64-lane super-strip take row-tag keys without the bound
    # and run again with 32-bit keys when a maximum comes near 4,095: match 20
    # over 120-300 matching residues scores 2,400-6,000, on both sides
    note("single super-strips: row-tag keys kept below 4,075, re-run above")
    rng = np.random.default_rng(41)
    pairs = []
    for _ in range(160):
        n = int(rng.integers(120, 301))
        a = W._dna_draw(rng, n + 60)
        pairs.append((a, a[30: 30 + n]))
        pairs.append((a, W._mutate(rng, a[: n], 0.05, 0.01, W._dna_draw, W._dna_subst)))
    batch = W.make_batch(pairs, "tagfall", dict(W.DNA_SW, matchValue=20, mismatchValue=-3))
    eng = engine_for(batch["sw"], False)
    eng.set_lanes_per_pair(64)
    got = device_run(eng, batch)
    assert_same(got, oracle_batch(batch, False), "single strip tag fallback")
    assert int(got[0].min()) < 4000 < 4200 < int(got[0].max())


@pytest.mark.parametrize("lanes", [8, 16])
def test_last_strip_tag_fallback(gpu, lanes):
    # the last super-strip of a group takes row-tag keys without the bound and
    # runs again with 32-bit keys when a maximum comes near 4,095: B = random
    # rows (one or more super-strips) + a 150-260-residue copy of A, match 20,
    # with starts (scores about 20 x the copy: 3,000-5,200)
    note(f"last super-strips: row-tag keys kept or re-run, lanes={lanes}")
    rng = np.random.default_rng(43 + lanes)
    pairs = []
    for _ in range(200):
        a = W._dna_draw(rng, 300)
        n = int(rng.integers(150, 261))
        pre = W._dna_draw(rng, int(rng.integers(lanes * 16, 3 * lanes * 16)))
        pairs.append((a, np.concatenate([pre, a[20: 20 + n]])))
    # mismatches and gaps cost enough that the random rows add nothing
    batch = W.make_batch(pairs, "lastfall", dict(W.DNA_SW, matchValue=20, mismatchValue=-20,
                                                 ambiguityValue=-20, gapInit=-20, gapExtend=-5))
    eng = engine_for(batch["sw"], False)
    eng.set_lanes_per_pair(lanes)
    got = device_run(eng, batch)
    assert_same(got, oracle_batch(batch, False), f"last strip tag fallback, lanes={lanes}")
    assert int(got[0].min()) < 3500 and int(got[0].max()) > 4500


@pytest.mark.parametrize("forward_only", [True, False])
def test_int32_rerun_long_pairs(gpu, forward_only):
    """Long saturating pairs (2-4 kb, every score above 32,767) through the
    exact int32 kernel, one wave per pair: all re-run, 40 against the oracle,
    with and without starts."""
    note(f"int32 re-run of 300 saturating 2-4 kb pairs, forward_only={forward_only}")
    batch = W.saturating_pairs(seed=9, n=300)
    eng = engine_for(batch["sw"], forward_only)
    got = device_run(eng, batch)
    assert eng.stats()["rerun_pairs"] == 300
    assert int(got[0].min()) > 32767
    idx = np.arange(0, 300, 300 // 40)[:40]
    sub = dict(batch, a_off=batch["a_off"][idx], a_len=batch["a_len"][idx],
               b_off=batch["b_off"][idx], b_len=batch["b_len"][idx])
    assert_same([g[idx] for g in got], oracle_batch(sub, forward_only), "int32 re-run")


def test_int32_mode_matches(gpu):
    note("int32 precision mode")
    batch = W.random_pairs(77, 200, (1, 300), (1, 300), protein=True)
    eng = engine_for(batch["sw"], False)
    eng.set_precision(I.PREC_INT32)
    assert_same(device_run(eng, batch), oracle_batch(batch, False), "int32 mode")


def test_empty_and_tiny(gpu):
    note("empty and tiny sequences")
    eng = engine_for(W.DNA_SW, False)
    A = ["", "A", "", "ACGT", "T", "G" * 5]
    B = ["ACGT", "", "", "A", "T", "G"]
    s, ar, br = eng.compare(A, B)
    exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1)
    assert_same((s, ar, br), exp, "empty")


def test_batch_log_lines(gpu, capfd, monkeypatch):
    # IPUMA_LOG at engine creation: one IPUSWLOG batch_perf JSON line per
    # device batch, with the batch's counts and its HIP-event times
    note("IPUMA_LOG batch_perf lines")
    batch = W.random_pairs(57, 2000, (50, 400), (50, 400))
    cells = batch["cells"]
    monkeypatch.setenv("IPUMA_LOG", "1")
    eng = engine_for(W.DNA_SW, True)
    monkeypatch.delenv("IPUMA_LOG")
    capfd.readouterr()
    got = device_run(eng, batch)
    host = eng.compare(["ACGT" * 30], ["ACGT" * 20])
    eng.close()
    assert_same(got, oracle_batch(batch, True), "logged batch")
    assert int(host[0][0]) == 80
    lines = [json.loads(ln[len("IPUSWLOG"):]) for ln in capfd.readouterr().err.splitlines()
             if ln.startswith("IPUSWLOG")]
    assert [ln["entry"] for ln in lines] == ["device", "compare_local"]
    dev, loc = lines
    assert dev["tag"] == "batch_perf" and dev["status"] == 0
    assert dev["pairs"] == 2000 and dev["cells"] == cells
    assert dev["padded_cells"] >= cells and dev["groups"] > 0
    assert 0 < dev["kernel_ms"] <= dev["batch_ms"] and dev["gcups_inner"] > 0
    assert loc["pairs"] == 1 and loc["cells"] == 120 * 80 and loc["host_ms"] > 0


def test_gap_settings_outside_packed_domain(gpu):
    # a positive gap extension is outside the packed kernel; the exact int32
    # kernel takes over and still matches the reference recurrence
    note("positive gap extension -> exact int32 path")
    sw = dict(W.DNA_SW, gapInit=-3, gapExtend=1)
    batch = W.random_pairs(55, 100, (1, 60), (1, 60))
    batch["sw"] = sw
    eng = engine_for(sw, False)
    assert_same(device_run(eng, batch), oracle_batch(batch, False), "gap domain")


def test_errors(gpu):
    note("error paths")
    eng = I.SWAlgorithm(I.SWConfig(), I.IPUAlgoConfig(tilesUsed=1, maxAB=10, maxBatches=2))
    with pytest.raises(I.IpumaError):
        eng.compare_local(["A"] * 3, ["A"] * 3)  # more than tilesUsed*maxBatches
    with pytest.raises(I.IpumaError):
        eng.compare_local(["A" * 11], ["A"])  # |A| > maxAB
    with pytest.raises(I.IpumaError):
        I.SWAlgorithm(I.SWConfig(similarity="simple"), I.IPUAlgoConfig())
    for bad in (0.0, 0.2, 4.5, float("nan")):
        with pytest.raises(I.IpumaError):
            eng.set_step_cap_scale(bad)
    for bad in (0, 65):
        with pytest.raises(I.IpumaError):
            eng.set_work_rounds(bad)
    eng.set_step_cap_scale(0.25)
    eng.set_step_cap_scale(4.0)


def test_cpp_facade(gpu):
    # ipu::batchaffine::SWAlgorithm (C++) over the C ABI: the reference's gtest
    # cases re-hosted in tests/cpp/facade_test.cpp
    import subprocess
    note("C++ facade test binary")
    exe = os.path.join(os.path.dirname(I.LIB_PATH), "facade_test")
    assert os.path.exists(exe), "build with `make -C ipuma-lib_amd`"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]


def test_cpp_multidriver_shape(gpu):
    # a caller in the shape of the reference's batch driver (ipusw/ipumultidriver.hpp):
    # JSON run configuration, getTileClockFrequency, prepared buffers, asynchronous jobs
    # joined by receiver threads, batch_perf metrics from the Job's cycle counts
    import subprocess
    exe = os.path.join(os.path.dirname(I.LIB_PATH), "multidriver_test")
    if not os.path.exists(exe):
        pytest.skip("multidriver_test is built only where nlohmann/json.hpp is available")
    note("C++ batch-driver-shaped caller")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert r.stdout.count("IPUSWLOG") == 6  # 360 comparisons / 60


@pytest.mark.parametrize("kind", ["dna", "aa"])
def test_dataset_files(gpu, kind):
    """Dataset files as the reference's batch driver reads them
    (ipusw/run_comparison.hpp:22-61: references from FASTA, queries one per
    line, A = reference, B = query), through compare_local, against the
    oracle; the C++ loader (ipu::load_data) reads the same sequences."""
    import subprocess
    from ipuma_amd import io as IO
    note(f"dataset files ({kind}) through compare_local")
    ref_path = os.path.join(GOLD, "datasets", f"{kind}_refs.fa")
    qry_path = os.path.join(GOLD, "datasets", f"{kind}_queries.txt")
    refs, queries = IO.load_data(ref_path), IO.load_data(qry_path)
    assert len(refs) == len(queries) and len(refs) > 100
    io_bin = os.path.join(os.path.dirname(I.LIB_PATH), "io_test")
    if os.path.exists(io_bin):
        for path, seqs in ((ref_path, refs), (qry_path, queries)):
            out = subprocess.run([io_bin, path, "0"], capture_output=True, check=True).stdout
            lines = out.decode("latin-1").split("\n")
            assert int(lines[0]) == len(seqs)
            assert [ln[1:-1] for ln in lines[1:1 + len(seqs)]] == seqs
    if kind == "dna":
        sw = dict(gapInit=-2, gapExtend=-1)
        osw = dict(similarity=O.SIM_NA, datatype=O.DT_NA, gap_init=-2, gap_extend=-1)
    else:
        sw = dict(gapInit=-10, gapExtend=-1, similarity="blosum62", datatype="aa")
        osw = dict(similarity=O.SIM_BLOSUM62, datatype=O.DT_AA, gap_init=-10, gap_extend=-1)
    n = len(refs)
    eng = I.SWAlgorithm(I.SWConfig(**sw), I.IPUAlgoConfig(
        tilesUsed=4, maxAB=max(len(x) for x in refs), maxBatches=(n + 3) // 4))
    eng.compare_local(refs, queries)
    r = eng.get_result()
    exp = O.sw_batch(refs, queries, **osw)
    assert_same((r.scores[:n], r.a_range_result[:n], r.b_range_result[:n]), exp,
                f"dataset {kind}")
    eng.close()


def test_request_server(gpu):
    note("request server: 4 threads x 6 DNA requests of 1-300 pairs vs the oracle")
    import threading
    rng = np.random.default_rng(21)
    reqs = []
    for r in range(24):
        b = W.config0(seed=100 + r, n=int(rng.integers(1, 301)))
        A = [W.decode(b["codes"][o:o + l], False) for o, l in zip(b["a_off"], b["a_len"])]
        B = [W.decode(b["codes"][o:o + l], False) for o, l in zip(b["b_off"], b["b_len"])]
        reqs.append((A, B))
    srv = I.Server(I.SWConfig(**W.DNA_SW), I.IPUAlgoConfig(forwardOnly=False),
                   hbm_budget_bytes=1 << 20, max_batch_pairs=4096, max_wait_us=1000)
    out = [None] * len(reqs)
    errors = []

    def worker(idx):
        try:
            for i in idx:
                out[i] = srv.wait(srv.submit(*reqs[i]), len(reqs[i][0]))
        except Exception as e:
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=120)
    assert not errors, errors
    for (A, B), got in zip(reqs, out):
        exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1)
        for g, e in zip(got[:3], exp):
            assert np.array_equal(g, e)
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["arena_used_bytes"] == 0
    srv.close()


@pytest.mark.parametrize("forward_only", [True, False])
def test_serving_load(gpu, forward_only):
    """The serving workload through the request server: DNA (plain DNA
    generator) and protein (protein batch generator) requests of uniform
    1..4,096 comparisons, submitted from six threads at once to one server per
    scoring on the GPU (as the load generator runs them), each request's results
    against the oracle.  Reference loop: ipusw/run_comparison.hpp:131-145,
    ipusw/ipuworker.hpp:8-36."""
    import threading
    from ipuma_amd import loadgen as LG
    note(f"serving load: 36 DNA/protein requests of 1-4096 pairs, 6 threads, "
         f"forward_only={forward_only}")
    pools = LG.make_pools(pairs=12_000, seed=31)
    rng = np.random.default_rng(32)
    m = 36
    kinds = ["dna" if rng.random() < 0.5 else "protein" for _ in range(m)]
    sizes = rng.integers(1, 4097, m)
    sizes[:2] = (4096, 1)  # both ends of the range
    starts = [int(rng.integers(0, pools[k][0].n - int(z) + 1)) for k, z in zip(kinds, sizes)]
    servers = {k: I.Server(I.SWConfig(**sw), I.IPUAlgoConfig(forwardOnly=forward_only),
                           hbm_budget_bytes=256 << 20, max_batch_pairs=1 << 16, max_wait_us=1500)
               for k, (_, sw) in pools.items()}
    out = [None] * m
    errors = []

    def worker(idx):
        try:
            tickets = []
            for i in idx:  # submit all of this thread's requests, then wait for them
                pool = pools[kinds[i]][0]
                n, ap, ao, bp, bo, _ = pool.window(starts[i], int(sizes[i]))
                tickets.append((i, servers[kinds[i]].submit_raw(n, ap, ao, bp, bo)))
            for i, t in tickets:
                out[i] = servers[kinds[i]].wait(t, int(sizes[i]))
        except Exception as e:  # surfaced below
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(list(range(j, m, 6)),)) for j in range(6)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=240)
    assert not errors, errors
    for i in range(m):
        b = pools[kinds[i]][0].batch
        sl = slice(starts[i], starts[i] + int(sizes[i]))
        sw = b["sw"]
        mat = O.select_matrix(SIM[sw["similarity"]], sw["matchValue"], sw["mismatchValue"],
                              sw["ambiguityValue"])
        exp = O.sw_batch_codes(b["codes"], b["a_off"][sl], b["a_len"][sl], b["b_off"][sl],
                               b["b_len"][sl], mat, sw["gapInit"], sw["gapExtend"], forward_only)
        assert_same(out[i][:3], exp, f"request {i} ({kinds[i]}, {int(sizes[i])} pairs)")
    for k, srv in servers.items():
        st = srv.stats()
        assert st["requests"] == kinds.count(k) and st["arena_used_bytes"] == 0, (k, st)
        srv.close()


def test_server_admission(gpu):
    """Admission against the HBM budget on the GPU: with blocking on, requests
    whose total is many times the arena all complete (each waits for space);
    with it off, a request larger than the arena is refused with IPUMA_EBUSY
    and the server keeps serving."""
    import threading
    note("server admission: 8 KiB arena, blocking and refusing")
    rng = np.random.default_rng(41)
    reqs = []
    for r in range(16):
        b = W.config0(seed=200 + r, n=int(rng.integers(5, 12)))
        A = [W.decode(b["codes"][o:o + l], False) for o, l in zip(b["a_off"], b["a_len"])]
        B = [W.decode(b["codes"][o:o + l], False) for o, l in zip(b["b_off"], b["b_len"])]
        reqs.append((A, B))
    srv = I.Server(I.SWConfig(**W.DNA_SW), I.IPUAlgoConfig(forwardOnly=True),
                   hbm_budget_bytes=8192, max_batch_pairs=64, max_wait_us=200)
    out = [None] * len(reqs)
    errors = []

    def worker(idx):
        try:
            for i in idx:
                out[i] = srv.wait(srv.submit(*reqs[i]), len(reqs[i][0]))
        except Exception as e:
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 4)),)) for j in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=120)
    assert not errors, errors
    total = sum(len(a) + len(b) for A, B in reqs for a, b in zip(A, B))
    assert total > 4 * 8192  # the requests did not fit the arena at once
    for (A, B), got in zip(reqs, out):
        exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=True)
        assert_same(got[:3], exp, "admission (blocking)")
    assert srv.stats()["arena_used_bytes"] == 0
    srv.close()
    srv = I.Server(I.SWConfig(**W.DNA_SW), I.IPUAlgoConfig(forwardOnly=True),
                   hbm_budget_bytes=8192, max_batch_pairs=64, block_on_full=False)
    with pytest.raises(I.IpumaError) as ei:
        srv.submit(["A" * 5000], ["C" * 4000])
    assert ei.value.code == I.IPUMA_EBUSY
    A, B = reqs[0]
    got = srv.wait(srv.submit(A, B), len(A))
    assert_same(got[:3], O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=True),
                "admission (after a refusal)")
    srv.close()


def test_search_top100(gpu):
    note("one-vs-many search: 4 protein queries x 2500 database sequences, top 100")
    import torch
    b = W.config4(seed=9, ndb=2500, qlens=(50, 120, 300, 700), homolog_frac=0.01)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(b[k])).to(dev)
         for k in ("codes", "q_off", "q_len", "db_off", "db_len")}
    nq, ndb, k = 4, 2500, 100
    scores = torch.empty(nq * ndb, dtype=torch.int32, device=dev)
    ti = torch.empty(nq * k, dtype=torch.int32, device=dev)
    ts = torch.empty(nq * k, dtype=torch.int32, device=dev)
    eng = I.SWAlgorithm(I.SWConfig(**W.PROTEIN_SW),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    eng.search_device(t["codes"], t["codes"].numel(), t["q_off"], t["q_len"], t["db_off"],
                      t["db_len"], k, scores, ti, ts)
    eng.synchronize()
    got = scores.cpu().numpy().reshape(nq, ndb)
    mat = O.select_matrix(O.SIM_BLOSUM62)
    for q in range(nq):
        exp, _, _ = O.sw_batch_codes(b["codes"], b["db_off"], b["db_len"],
                                     np.full(ndb, b["q_off"][q]), np.full(ndb, b["q_len"][q],
                                                                          np.int32),
                                     mat, -10, -1, True)
        assert np.array_equal(got[q], exp), q
        order = np.lexsort((np.arange(ndb), -exp.astype(np.int64)))[:k]
        assert np.array_equal(ti.cpu().numpy().reshape(nq, k)[q], order), q
        assert np.array_equal(ts.cpu().numpy().reshape(nq, k)[q], exp[order]), q
    eng.close()


def test_search_full_size(gpu):
    # the whole one-vs-many search (8 queries x 200,000 database sequences):
    # every query's top 100 is the (score desc, index asc) order of its own
    # 200,000 scores; 1,000 random database sequences per query and every
    # top-100 entry agree with the oracle
    note("one-vs-many search at full size: 8 queries x 200,000 sequences, top 100")
    import torch
    b = W.config4()
    nq, ndb, k = len(b["q_len"]), len(b["db_len"]), 100
    assert (nq, ndb) == (8, 200_000)
    dev = torch.device("cuda:0")
    t = {x: torch.from_numpy(np.ascontiguousarray(b[x])).to(dev)
         for x in ("codes", "q_off", "q_len", "db_off", "db_len")}
    scores = torch.empty(nq * ndb, dtype=torch.int32, device=dev)
    ti = torch.empty(nq * k, dtype=torch.int32, device=dev)
    ts = torch.empty(nq * k, dtype=torch.int32, device=dev)
    eng = I.SWAlgorithm(I.SWConfig(**W.PROTEIN_SW),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    eng.search_device(t["codes"], t["codes"].numel(), t["q_off"], t["q_len"], t["db_off"],
                      t["db_len"], k, scores, ti, ts)
    eng.synchronize()
    got = scores.cpu().numpy().reshape(nq, ndb)
    top_i = ti.cpu().numpy().reshape(nq, k)
    top_s = ts.cpu().numpy().reshape(nq, k)
    eng.close()
    mat = O.select_matrix(O.SIM_BLOSUM62)
    rng = np.random.default_rng(44)
    for q in range(nq):
        order = np.lexsort((np.arange(ndb), -got[q].astype(np.int64)))[:k]
        assert np.array_equal(top_i[q], order), q
        assert np.array_equal(top_s[q], got[q][order]), q
        idx = np.concatenate([rng.choice(ndb, 1000, replace=False), order])
        exp, _, _ = O.sw_batch_codes(b["codes"], b["db_off"][idx], b["db_len"][idx],
                                     np.full(idx.size, b["q_off"][q]),
                                     np.full(idx.size, b["q_len"][q], np.int32),
                                     mat, -10, -1, True)
        assert np.array_equal(got[q][idx], exp), q
