"""The engine's own kernel and host sources, compiled for the CPU against the
HIP emulation in tests/cpu_emulation (lockstep 64-lane waves, DPP and
shuffles through barriers, device memory pre-filled with garbage).

* emu_checks: input validation of the kernels under AddressSanitizer;
* the reference's SSW parity pairs through the emulated engine (C ABI),
  bit-exact against the frozen SSW results, with device memory filled with a
  pattern that reads as large positive integers.
"""
import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "cpu_emulation")
CSRC = os.path.join(ROOT, "ipuma-lib_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"
BUILD = os.path.join(ROOT, "_build_emu")

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="host clang++ not found")


def _compile(src, out, extra=(), lang="c++"):
    cmd = [CLANG, "-std=c++20", "-O1", "-g", "-fPIC", "-pthread", "-x", lang, *extra,
           "-I" + EMU, "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), "-c", src, "-o", out]
    subprocess.run(cmd, check=True)


@pytest.fixture(scope="module")
def build_dir():
    os.makedirs(BUILD, exist_ok=True)
    yield BUILD


def test_kernel_input_validation_asan(build_dir):
    san = ["-fsanitize=address", "-fno-omit-frame-pointer"]
    objs = []
    for src, define in (("sw_kernels.hip", []), ("sw_tables.cpp", [])):
        o = os.path.join(build_dir, src + ".asan.o")
        _compile(os.path.join(CSRC, src), o, san + define)
        objs.append(o)
    main_o = os.path.join(build_dir, "emu_checks.o")
    _compile(os.path.join(EMU, "emu_checks.cpp"), main_o, san + ["-DEMU_DEFINE_GLOBALS"])
    exe = os.path.join(build_dir, "emu_checks")
    subprocess.run([CLANG, "-fsanitize=address", "-pthread", main_o, *objs, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.fixture(scope="module")
def emu_lib(build_dir):
    """The engine, server, kernels and tables built against the emulation, loaded
    as the binding's library for the duration of the module."""
    objs = []
    for src, extra in (("sw_engine.cpp", ["-DEMU_DEFINE_GLOBALS"]), ("sw_server.cpp", []),
                       ("sw_kernels.hip", []), ("sw_search.hip", []), ("sw_tables.cpp", [])):
        o = os.path.join(build_dir, src + ".o")
        _compile(os.path.join(CSRC, src), o, extra)
        objs.append(o)
    cpp = os.path.join(ROOT, "ipuma-lib_amd", "cpp")
    for src in sorted(f for f in os.listdir(cpp) if f.endswith(".cpp")):
        o = os.path.join(build_dir, "cpp_" + src + ".o")
        _compile(os.path.join(cpp, src), o, ["-I" + cpp])
        objs.append(o)
    lib_path = os.path.join(build_dir, "libipuma_emu.so")
    subprocess.run([CLANG, "-shared", "-pthread", *objs, "-o", lib_path], check=True)
    lib = ctypes.CDLL(lib_path)
    ctypes.c_int.in_dll(lib, "_ZN3emu9fill_byteE").value = 0x3F
    import ipuma_amd as I
    I.lib.cache_clear()
    saved = I.LIB_PATH
    I.LIB_PATH = lib_path
    yield I
    I.LIB_PATH = saved
    I.lib.cache_clear()


SSW_CASES = [c["name"] for c in json.load(open(os.path.join(HERE, "golden", "ssw_fixtures.json")))["cases"]]


@pytest.mark.parametrize("name", SSW_CASES)
def test_ssw_parity_pairs_emulated(emu_lib, name):
    """Every SSW-frozen fixture (the reference's parity-test pairs and the seeded
    BLOSUM62 11/1 and DNA 3/1 batches) through the emulated engine."""
    I = emu_lib
    d = json.load(open(os.path.join(HERE, "golden", "ssw_fixtures.json")))
    case = next(c for c in d["cases"] if c["name"] == name)
    c = case["config"]
    names = {0: "na", 1: "blosum50", 2: "blosum62"}
    sw = I.SWConfig(gapInit=c["gap_init"], gapExtend=c["gap_extend"], matchValue=c["match"],
                    mismatchValue=c["mismatch"], ambiguityValue=c["ambiguity"],
                    similarity=names[c["similarity"]],
                    datatype="na" if c["datatype"] == 0 else "aa")
    eng = I.SWAlgorithm(sw, I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    s, ar, br = eng.compare([p[0] for p in case["pairs"]], [p[1] for p in case["pairs"]])
    qb, qe = I.unpack_range(ar)
    rb, re_ = I.unpack_range(br)
    for k, (score, eqb, eqe, erb, ere) in enumerate(case["expected"]):
        assert (int(s[k]), int(qb[k]), int(qe[k]), int(rb[k]), int(re_[k])) == \
            (score, eqb, eqe, erb, ere), (name, k)
    eng.close()


def test_shape_knobs_emulated(emu_lib):
    """The autotuner's knobs through the C ABI: ranges enforced, and a batch
    under the extreme settings still matches the oracle."""
    import oracle as O
    I = emu_lib
    eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=True))
    for bad in (0.0, 0.2, 4.5, float("nan")):
        with pytest.raises(I.IpumaError):
            eng.set_step_cap_scale(bad)
    with pytest.raises(I.IpumaError):
        eng.set_work_rounds(0)
    rng = np.random.default_rng(41)
    A = ["".join(rng.choice(list("ACGT"), int(rng.integers(20, 200)))) for _ in range(12)]
    B = ["".join(rng.choice(list("ACGT"), int(rng.integers(20, 300)))) for _ in range(12)]
    exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=True)
    for rounds, scale in ((1, 0.25), (4, 4.0)):
        eng.set_work_rounds(rounds)
        eng.set_step_cap_scale(scale)
        for g, e in zip(eng.compare(A, B), exp):
            assert np.array_equal(np.asarray(g), np.asarray(e)), (rounds, scale)
    eng.close()


def test_last_strip_rows_emulated(emu_lib):
    """Pipelines of 2, 4 and 16 lanes whose last super-strip needs 1-16 rows per
    lane (run with 8, 12 or 16), and whole-wave groups of one super-strip
    needing 7-16 rows per lane (run with 8, 10, 12, 14 or 16), forward and with
    starts (whose pass stops early), against the oracle."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(43)
    for lanes in (2, 4, 16, 64):
        A, B = [], []
        for need in ((1, 3, 8, 9, 12, 13, 16) if lanes < 64 else (7, 9, 10, 11, 13, 14, 15, 16)):
            # one full super-strip, then `need` rows per lane (whole-wave: one super-strip)
            nb = (16 * lanes if lanes < 64 else 0) + lanes * (need - 1) + 1
            a = "".join(rng.choice(list("ACGT"), int(rng.integers(20, 60))))
            b = "".join(rng.choice(list("ACGT"), nb - len(a) // 2)) + a[: len(a) // 2]
            A.append(a)
            B.append(b[:nb])
        for fwd in (True, False):
            eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                                I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                                forwardOnly=fwd))
            eng.set_lanes_per_pair(lanes)
            exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=fwd)
            for g, e in zip(eng.compare(A, B), exp):
                assert np.array_equal(np.asarray(g), np.asarray(e)), (lanes, fwd)
            eng.close()


def test_request_server_emulated(emu_lib):
    """Three submitting threads, a 2 KiB arena (so admission blocks and the ring
    wraps), results checked against the oracle; then a request larger than the
    arena is refused with IPUMA_EBUSY when blocking is off."""
    import threading
    import numpy as np
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(4)
    reqs = []
    for r in range(6):
        k = int(rng.integers(1, 7))
        A = ["".join(rng.choice(list("ACGT"), int(rng.integers(1, 60)))) for _ in range(k)]
        B = ["".join(rng.choice(list("ACGT"), int(rng.integers(1, 60)))) + a[: len(a) // 2]
             for a in A]
        reqs.append((A, B))
    srv = I.Server(I.SWConfig(gapInit=-2, gapExtend=-1), I.IPUAlgoConfig(forwardOnly=True),
                   hbm_budget_bytes=2048, max_batch_pairs=16, max_wait_us=500)
    out = [None] * len(reqs)
    errors = []

    def worker(idx):
        try:
            for i in idx:
                t = srv.submit(*reqs[i])
                out[i] = srv.wait(t, len(reqs[i][0]))
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(list(range(j, len(reqs), 3)),))
               for j in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=600)
    assert not errors, errors
    for (A, B), got in zip(reqs, out):
        exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=True)
        for g, e in zip(got[:3], exp):
            assert np.array_equal(g, e)
        assert got[3] >= 0.0
    st = srv.stats()
    assert st["requests"] == len(reqs) and st["pairs"] == sum(len(a) for a, _ in reqs)
    assert st["arena_used_bytes"] == 0 and st["pending_pairs"] == 0
    srv.close()

    srv = I.Server(I.SWConfig(), I.IPUAlgoConfig(), hbm_budget_bytes=2048, block_on_full=False)
    with pytest.raises(I.IpumaError) as ei:
        srv.submit(["A" * 3000], ["C" * 3000])
    assert ei.value.code == I.IPUMA_EBUSY
    srv.close()


def _topk_reference(scores, k):
    """Top k of every row: score descending, then index ascending."""
    out_i, out_s = [], []
    for row in scores:
        order = np.lexsort((np.arange(len(row)), -row.astype(np.int64)))[:k]
        out_i.append(np.concatenate([order, -np.ones(k - len(order), np.int64)]))
        out_s.append(np.concatenate([row[order], np.zeros(k - len(order), np.int64)]))
    return np.array(out_i), np.array(out_s)


def test_search_topk_emulated(emu_lib):
    """Every query against every database sequence, with the per-query top k:
    short DNA (many tied scores, so the index tie-break matters), a k that is
    not a power of two, and a k larger than the database."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(8)
    qs = [rng.choice([1, 2, 3, 4], int(rng.integers(5, 40))).astype(np.uint8) for _ in range(3)]
    db = [rng.choice([1, 2, 3, 4], int(rng.integers(1, 30))).astype(np.uint8) for _ in range(120)]
    from ipuma_amd.workloads import arena_from_flat
    lens = np.array([len(x) for x in qs + db], np.int64)
    codes, offs = arena_from_flat(np.concatenate(qs + db), lens)
    q_off, db_off = offs[:3].copy(), offs[3:].copy()
    q_len = lens[:3].astype(np.int32)
    db_len = lens[3:].astype(np.int32)
    eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    mat = O.select_matrix(O.SIM_NA, 1, -1, -1)
    exp = np.zeros((3, 120), np.int64)
    for qi in range(3):
        s, _, _ = O.sw_batch_codes(codes, db_off, db_len, np.full(120, q_off[qi]),
                                   np.full(120, q_len[qi], np.int32), mat, -2, -1, True)
        exp[qi] = s
    for k in (17, 150):
        scores = np.zeros(3 * 120, np.int32)
        ti = np.zeros(3 * k, np.int32)
        ts = np.zeros(3 * k, np.int32)
        # device memory is host memory in the emulation: raw pointers
        eng.search_device(codes.ctypes.data, len(codes), q_off.ctypes.data, q_len.ctypes.data,
                          db_off.ctypes.data, db_len.ctypes.data, k, scores.ctypes.data,
                          ti.ctypes.data, ts.ctypes.data, nq=3, ndb=120)
        assert np.array_equal(scores.reshape(3, 120), exp)
        ri, rs = _topk_reference(exp, k)
        assert np.array_equal(ti.reshape(3, k), ri), k
        assert np.array_equal(ts.reshape(3, k), rs), k
    eng.close()


def test_search_shared_query_emulated(emu_lib):
    """BLOSUM62 is symmetric, so the search puts each query on the columns and
    the packed kernel's groups of one query read both comparisons' scores of a
    row with one load: queries of 20-300 residues (two of equal length, whose
    groups can mix them), database sequences with planted homologs, pipelines
    of 4 lanes, every score against the oracle."""
    import oracle as O
    from ipuma_amd.workloads import arena_from_flat
    I = emu_lib
    rng = np.random.default_rng(12)
    qlens = (20, 75, 75, 150, 300)
    qs = [rng.integers(0, 20, n).astype(np.uint8) for n in qlens]
    db = []
    for i in range(90):
        if i % 9 == 0:  # a homolog of query i % 5 at about 60% identity
            q = qs[i % 5]
            x = q.copy()
            mut = rng.random(len(x)) < 0.4
            x[mut] = rng.integers(0, 20, int(mut.sum()))
            db.append(np.concatenate([rng.integers(0, 20, 5), x]).astype(np.uint8))
        else:
            db.append(rng.integers(0, 20, int(rng.integers(10, 200))).astype(np.uint8))
    nq, ndb = len(qs), len(db)
    lens = np.array([len(x) for x in qs + db], np.int64)
    codes, offs = arena_from_flat(np.concatenate(qs + db), lens)
    q_off, db_off = offs[:nq].copy(), offs[nq:].copy()
    q_len = lens[:nq].astype(np.int32)
    db_len = lens[nq:].astype(np.int32)
    mat = O.select_matrix(O.SIM_BLOSUM62)
    exp = np.zeros((nq, ndb), np.int64)
    for qi in range(nq):
        s, _, _ = O.sw_batch_codes(codes, db_off, db_len, np.full(ndb, q_off[qi]),
                                   np.full(ndb, q_len[qi], np.int32), mat, -10, -1, True)
        exp[qi] = s
    assert exp.max() > 200  # the planted homologs
    for lanes in (4,):
        eng = I.SWAlgorithm(I.SWConfig(gapInit=-10, gapExtend=-1, similarity="blosum62",
                                       datatype="aa"),
                            I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
        eng.set_lanes_per_pair(lanes)
        k = 10
        scores = np.zeros(nq * ndb, np.int32)
        ti = np.zeros(nq * k, np.int32)
        ts = np.zeros(nq * k, np.int32)
        eng.search_device(codes.ctypes.data, len(codes), q_off.ctypes.data, q_len.ctypes.data,
                          db_off.ctypes.data, db_len.ctypes.data, k, scores.ctypes.data,
                          ti.ctypes.data, ts.ctypes.data, nq=nq, ndb=ndb)
        assert np.array_equal(scores.reshape(nq, ndb), exp), lanes
        ri, rs = _topk_reference(exp, k)
        assert np.array_equal(ti.reshape(nq, k), ri), lanes
        eng.close()


def test_band_groups_emulated(emu_lib):
    """Whole-wave groups of two or more super-strips (|B| > 1024) shared by the
    waves of a block, at the classifier's pipeline depth and with every pair
    on 64 lanes: long DNA pairs among short ones, scores and both ranges (the
    start pass runs the band groups again on the reversed prefixes), against
    the oracle."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(21)
    A, B = [], []
    for i in range(8):
        m = int(rng.integers(200, 700))
        a = "".join(rng.choice(list("ACGT"), m))
        mut = "".join(c if rng.random() > 0.1 else str(rng.choice(list("ACGT"))) for c in a)
        left = "".join(rng.choice(list("ACGT"), int(rng.integers(400, 1500))))
        right = "".join(rng.choice(list("ACGT"), int(rng.integers(100, 600))))
        A.append(a)
        B.append(left + mut + right)
    for i in range(24):
        A.append("".join(rng.choice(list("ACGT"), int(rng.integers(5, 60)))))
        B.append("".join(rng.choice(list("ACGT"), int(rng.integers(5, 60)))))
    exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=False)
    for lanes in (0, 64):  # the classifier's choice, then every pair on 64 lanes
        eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                            I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
        if lanes:
            eng.set_lanes_per_pair(lanes)
        s, ar, br = eng.compare(A, B)
        st = eng.stats()
        assert st["band_groups"] >= 2, (lanes, st)
        # the longest other groups start with the band groups
        assert st["head_groups"] >= 1, (lanes, st)
        for got, want in zip((s, ar, br), exp):
            assert np.array_equal(np.asarray(got), np.asarray(want)), lanes
        eng.close()


def test_last_strip_optimistic_tags_emulated(emu_lib):
    """The last super-strip of a group takes row-tag keys without the a-priori
    bound and runs again with 32-bit keys when a maximum comes within smax of
    4,095: 16-lane pipelines (super-strips of 256 rows), B = 260 random rows +
    a 190- or 240-residue copy of A, match 20 (scores near 3,800 keep the keys,
    near 4,800 re-run), one pair on a single super-strip too; with starts,
    against the oracle."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(77)
    A, B = [], []
    for k in range(6):
        a = "".join(rng.choice(list("ACGT"), 260))
        n = 190 if k % 2 == 0 else 240
        B.append("".join(rng.choice(list("ACGT"), 260)) + a[10: 10 + n])
        A.append(a)
    A.append(A[0])
    B.append(A[0][:230])  # one super-strip, 4,600
    # mismatches and gaps cost enough that the random rows add nothing
    exp = O.sw_batch(A, B, match=20, mismatch=-20, ambiguity=-20, gap_init=-20, gap_extend=-5,
                     forward_only=False)
    assert max(exp[0]) > 4095 > 4075 > min(exp[0])
    eng = I.SWAlgorithm(I.SWConfig(gapInit=-20, gapExtend=-5, matchValue=20, mismatchValue=-20,
                                   ambiguityValue=-20),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    eng.set_lanes_per_pair(16)
    got = eng.compare(A, B)
    for g, e, nm in zip(got, exp, ("scores", "a_range", "b_range")):
        assert np.array_equal(np.asarray(g), np.asarray(e)), nm
    eng.close()


def test_many_pairs_emulated(emu_lib):
    """3,000 short pairs (more than one block of every bucketing kernel: the
    batch totals, classification, scatter), each result against the oracle."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(88)
    A = ["".join(rng.choice(list("ACGT"), int(rng.integers(3, 24)))) for _ in range(3000)]
    B = ["".join(rng.choice(list("ACGT"), int(rng.integers(3, 24)))) for _ in range(3000)]
    exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1, forward_only=True)
    eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=True))
    got = eng.compare(A, B)
    for g, e, nm in zip(got, exp, ("scores", "a_range", "b_range")):
        assert np.array_equal(np.asarray(g), np.asarray(e)), nm
    eng.close()


def _tie_pair(rng, n):
    """A = X + Y, B = Y + X with |X| = |Y| = n: two equal best alignments, the
    one on the earlier B row ending at the later A column."""
    x = "".join(rng.choice(list("ACGT"), n))
    y = "".join(rng.choice(list("ACGT"), n))
    return x + y, y + x


def test_row_tag_keys_emulated(emu_lib):
    """Super-strips whose H values stay within 12 bits take 16-bit row-tag keys
    (H * 16 + 15 - row, a max over the step's rows); the rest the 32-bit keys.
    A super-strip of R rows can raise H by at most R times the largest score,
    so a nearly identical pair scoring above 4,095 crosses from one form to the
    other mid-pair: match 7 at depths 1-16, match 3 at depth 64 (whose first
    super-strip, 1,024 rows, already passes 4,095 - 3 * 1,024).  Tie pairs put
    two equal maxima in one lane's rows with the earlier row at the later
    column; short random pairs tie everywhere.  Scores and both ranges, against
    the oracle."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(77)

    def batch(m):
        a = "".join(rng.choice(list("ACGT"), m))
        mut = "".join(c if rng.random() > 0.01 else str(rng.choice(list("ACGT"))) for c in a)
        A, B = [a], ["".join(rng.choice(list("ACGT"), 7)) + mut]
        for n in (3, 5, 7, 8, 12, 30):
            a, b = _tie_pair(rng, n)
            A.append(a)
            B.append(b)
        for i in range(12):
            A.append("".join(rng.choice(list("ACGT"), int(rng.integers(1, 40)))))
            B.append("".join(rng.choice(list("ACGT"), int(rng.integers(1, 90)))))
        return A, B

    # (configuration, batch, pipeline depths, forward only)
    cases = ((dict(gapInit=-6, gapExtend=-2, matchValue=7, mismatchValue=-4), batch(600),
              (1, 16), False),
             (dict(gapInit=-4, gapExtend=-1, matchValue=3, mismatchValue=-2), batch(1040), (64,),
              True))
    for cfg, (A, B), depths, fwd in cases:
        exp = O.sw_batch(A, B, gap_init=cfg["gapInit"], gap_extend=cfg["gapExtend"],
                         match=cfg["matchValue"], mismatch=cfg["mismatchValue"],
                         forward_only=fwd)
        # the long pair passes the row-tag bound of its deepest super-strips
        assert int(np.max(exp[0])) > 4095 - cfg["matchValue"] * 16 * max(depths)
        for lanes in depths:
            eng = I.SWAlgorithm(I.SWConfig(**cfg),
                                I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                                forwardOnly=fwd))
            eng.set_lanes_per_pair(lanes)
            got = eng.compare(A, B)
            for g, e, what in zip(got, exp, ("scores", "a_range", "b_range")):
                g = np.asarray(g)
                bad = np.nonzero(g != np.asarray(e))[0]
                assert bad.size == 0, (cfg, lanes, what, bad[:5], g[bad[:5]],
                                       np.asarray(e)[bad[:5]])
            eng.close()


def test_prepared_buffers_emulated(emu_lib):
    """The wire-format path through the emulated engine: prepare_remote (fill
    first and greedy), prepared_remote_compare, transferResults, against the
    oracle; a comparison-list buffer; and three buffers submitted
    asynchronously from two threads, joined in another order."""
    import threading
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(31)
    sw = I.SWConfig(gapInit=-2, gapExtend=-1)
    A = ["".join(rng.choice(list("ACGTN"), int(rng.integers(1, 70)))) for _ in range(30)]
    B = ["".join(rng.choice(list("ACGT"), int(rng.integers(1, 90)))) + A[k][: len(A[k]) // 2]
         for k in range(30)]
    exp = O.sw_batch(A, B, gap_init=-2, gap_extend=-1)
    for fill in ("fillfirst", "greedy"):
        algo = I.IPUAlgoConfig(tilesUsed=8, maxAB=200, maxBatches=5, bufsize=700, fillAlgo=fill)
        eng = I.SWAlgorithm(sw, algo)
        inputs, mapping, _ = I.prepare_remote(sw, algo, A, B)
        res = eng.prepared_remote_compare(inputs)
        got = I.transfer_results(res, mapping)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e), fill
        # unused slots score 0
        used = np.zeros(8 * 5, bool)
        used[mapping] = True
        assert not res[:40][~used].any()
        eng.close()
    seqs = A[:10] + B[:10]
    cmps = [(i, 10 + (i * 3) % 10) for i in range(10)] + [(2, 2), (5, 12), (5, 13)]
    algo = I.IPUAlgoConfig(tilesUsed=3, maxAB=200, maxBatches=6, bufsize=800)
    eng = I.SWAlgorithm(sw, algo)
    inputs, mapping, _ = I.prepare_remote_mn(sw, algo, seqs, cmps)
    got = I.transfer_results(eng.prepared_remote_compare(inputs), mapping)
    expm = O.sw_batch([seqs[a] for a, _ in cmps], [seqs[b] for _, b in cmps], gap_init=-2,
                      gap_extend=-1)
    for g, e in zip(got, expm):
        assert np.array_equal(g, e)
    eng.close()
    algo = I.IPUAlgoConfig(tilesUsed=4, maxAB=200, maxBatches=5, bufsize=700, forwardOnly=True)
    eng = I.SWAlgorithm(sw, algo)
    parts = [(A[i::3], B[i::3]) for i in range(3)]
    prepared = [I.prepare_remote(sw, algo, a, b) for a, b in parts]
    jobs = [None] * 3

    def submit(idx):
        for i in idx:
            jobs[i] = eng.async_submit_prepared_remote_compare(prepared[i][0])

    th = [threading.Thread(target=submit, args=(ix,)) for ix in ([0, 2], [1])]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in (2, 0, 1):
        res = eng.blocking_join_prepared_remote_compare(jobs[i])
        s, ar, br = I.transfer_results(res, prepared[i][1])
        es, ea, eb = O.sw_batch(parts[i][0], parts[i][1], gap_init=-2, gap_extend=-1,
                                forward_only=True)
        assert np.array_equal(s, es) and np.array_equal(ar, ea) and np.array_equal(br, eb)
    with pytest.raises(I.IpumaError):
        jobs[0].joined = False
        jobs[0].join()  # joined twice
    eng.close()


def test_traceback_emulated(emu_lib):
    """CIGARs from the emulated traceback kernel equal the oracle's, on DNA and
    BLOSUM62 pairs with starts, including zero-score pairs."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(41)
    for sim, dt, letters, gi, ge in ((O.SIM_NA, O.DT_NA, "ACGT", -2, -1),
                                     (O.SIM_BLOSUM62, O.DT_AA, "ARNDCQEGHILKMFPSTWYV", -10, -1)):
        A = ["".join(rng.choice(list(letters), int(rng.integers(1, 120)))) for _ in range(20)]
        B = [("".join(rng.choice(list(letters), int(rng.integers(0, 40)))) + a[3:]
              + "".join(rng.choice(list(letters), int(rng.integers(0, 40))))) or "A" for a in A]
        A[0], B[0] = "A", "C" * 5  # score 0 under DNA scoring
        sw = I.SWConfig(gapInit=gi, gapExtend=ge, similarity="na" if dt == O.DT_NA else "blosum62",
                        datatype="na" if dt == O.DT_NA else "aa")
        eng = I.SWAlgorithm(sw, I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
        s, ar, br = eng.compare(A, B)
        got = eng.traceback(A, B, s, ar, br)
        mat = O.select_matrix(sim)
        for k in range(len(A)):
            a, b = O.encode(dt, A[k]), O.encode(dt, B[k])
            exp = O.cigar(a, b, int(ar[k]), int(br[k]), int(s[k]), mat, gi, ge)
            assert np.array_equal(got[k], exp), (k, I.cigar_string(got[k]), O.cigar_string(exp))
        eng.close()


def test_traceback_strips_emulated(emu_lib):
    """The wave traceback over regions of several strips of 512 rows (the row
    above each strip handed over through the work area) and of one strip of
    3-7 rows per lane, and the int64 kernel that gap settings beyond the int32
    bound take: CIGARs equal the oracle's."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(43)

    def mutate(a, letters):
        out = []
        for c in a:
            u = rng.random()
            if u < 0.08:
                out.append(rng.choice(list(letters)))
            elif u < 0.10:
                continue
            elif u < 0.12:
                out.extend([c, rng.choice(list(letters))])
            else:
                out.append(c)
        return "".join(out)

    dna, aa = "ACGT", "ARNDCQEGHILKMFPSTWYV"
    cases = []
    # regions of 2-3 strips of 512 rows, and of one strip of 3-7 rows per lane
    A = ["".join(rng.choice(list(dna), n)) for n in (1300, 530, 1030, 200, 350, 170, 420)]
    cases.append((O.SIM_NA, O.DT_NA, -2, -1, A, [mutate(a, dna) for a in A]))
    P = ["".join(rng.choice(list(aa), n)) for n in (700, 513, 300, 240)]
    cases.append((O.SIM_BLOSUM62, O.DT_AA, -10, -1, P, [mutate(a, aa) for a in P]))
    # a gap open of 20,000,000: every region here (at least 12 cells along its
    # two sides) is past the int32 bound (2^28), so the int64 kernel runs
    D = ["".join(rng.choice(list(dna), int(rng.integers(20, 90)))) for _ in range(6)]
    cases.append((O.SIM_NA, O.DT_NA, -20_000_000, -1, D, [mutate(a, dna) for a in D]))
    for sim, dt, gi, ge, A, B in cases:
        sw = I.SWConfig(gapInit=gi, gapExtend=ge, similarity="na" if dt == O.DT_NA else "blosum62",
                        datatype="na" if dt == O.DT_NA else "aa")
        eng = I.SWAlgorithm(sw, I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
        s, ar, br = eng.compare(A, B)
        got = eng.traceback(A, B, s, ar, br)
        mat = O.select_matrix(sim)
        for k in range(len(A)):
            exp = O.cigar(O.encode(dt, A[k]), O.encode(dt, B[k]), int(ar[k]), int(br[k]), int(s[k]),
                          mat, gi, ge)
            assert np.array_equal(got[k], exp), (gi, k, I.cigar_string(got[k]), O.cigar_string(exp))
        if gi == -2:  # the long DNA pairs span two and three strips of 512 rows
            rows = [((int(ar[k]) >> 16) - (int(ar[k]) & 0xFFFF) + 1) for k in range(len(A))]
            assert max(rows) > 1024, rows
        eng.close()


def _thread_budget(I):
    import ctypes
    return ctypes.c_int.in_dll(I.lib(), "_ZN5ipuma21g_thread_start_budgetE")


def test_thread_start_failures_emulated(emu_lib):
    """Host threads that cannot start: the staging gather of a batch above
    8 MB does the parts of the workers that never started itself (results
    unchanged), and a server whose dispatcher or completer cannot start
    reports IPUMA_EDEVICE with nothing left running."""
    import oracle as O
    I = emu_lib
    budget = _thread_budget(I)
    rng = np.random.default_rng(17)
    # 9 MB of characters: 136 sequences of 65,535 residues against empty B
    # (no DP cells), interleaved with short pairs whose results show that every
    # slice was gathered to the right place
    A, B = [], []
    big = "".join(rng.choice(list("ACGT"), 65535))
    for i in range(136):
        A.append(big)
        B.append("")
        a = "".join(rng.choice(list("ACGT"), int(rng.integers(20, 80))))
        A.append(a)
        B.append("".join(rng.choice(list("ACGT"), int(rng.integers(5, 40)))) + a[5:50])
    exp = O.sw_batch(A[1::2], B[1::2], gap_init=-2, gap_extend=-1, forward_only=True)
    eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=True))
    try:
        for cap in (3,):  # three workers start, the caller gathers the other slices
            budget.value = cap
            s, ar, br = eng.compare(A, B)
            budget.value = -1
            assert not s[0::2].any()
            for g, e in zip((s[1::2], ar[1::2], br[1::2]), exp):
                assert np.array_equal(g, e), cap
        for cap in (0, 1):
            budget.value = cap
            with pytest.raises(I.IpumaError) as ei:
                I.Server(I.SWConfig(), I.IPUAlgoConfig(forwardOnly=True), hbm_budget_bytes=4096)
            budget.value = -1
            assert ei.value.code == I.IPUMA_EDEVICE and "thread start" in str(ei.value)
        # and the library still starts threads afterwards
        srv = I.Server(I.SWConfig(), I.IPUAlgoConfig(forwardOnly=True), hbm_budget_bytes=4096)
        t = srv.submit(["ACGT"], ["ACGT"])
        assert int(srv.wait(t, 1)[0][0]) == 4
        srv.close()
    finally:
        budget.value = -1
        eng.close()


def test_null_character_buffers_emulated(emu_lib):
    """compare_local / compare_mn_local / traceback_local refuse a NULL
    character buffer or offsets with IPUMA_EINVAL instead of reading them."""
    import ctypes
    I = emu_lib
    L = I.lib()
    eng = I.SWAlgorithm(I.SWConfig(), I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16,
                                                      maxBatches=1))
    offs = np.array([0, 4], np.int64)
    out = [np.zeros(1, np.int32) for _ in range(3)]
    o = [x.ctypes.data for x in out]
    assert L.ipuma_compare_local(eng._h, 1, None, offs.ctypes.data, b"ACGT",
                                 offs.ctypes.data, *o) == I.IPUMA_EINVAL
    assert b"NULL" in L.ipuma_last_error()
    assert L.ipuma_compare_local(eng._h, 1, b"ACGT", offs.ctypes.data, None,
                                 offs.ctypes.data, *o) == I.IPUMA_EINVAL
    cmps = np.array([0, 0], np.int32)
    assert L.ipuma_compare_mn_local(eng._h, 1, None, offs.ctypes.data, 1, cmps.ctypes.data,
                                    *o) == I.IPUMA_EINVAL
    assert L.ipuma_compare_mn_local(eng._h, 1, b"ACGT", None, 0, None, None, None,
                                    None) == I.IPUMA_EINVAL
    cig = np.zeros(16, np.uint32)
    coff = np.array([0, 16], np.int64)
    clen = np.zeros(1, np.int32)
    assert L.ipuma_traceback_local(eng._h, 1, None, offs.ctypes.data, b"ACGT", offs.ctypes.data,
                                   *o, cig.ctypes.data, coff.ctypes.data,
                                   clen.ctypes.data) == I.IPUMA_EINVAL
    # empty sequences need no characters
    zero = np.array([0, 0], np.int64)
    assert L.ipuma_compare_local(eng._h, 1, None, zero.ctypes.data, None, zero.ctypes.data,
                                 *o) == 0
    assert int(out[0][0]) == 0
    eng.close()


def test_batch_log_lines_emulated(emu_lib, capfd, monkeypatch):
    """IPUMA_LOG set at engine creation: one IPUSWLOG batch_perf JSON line on
    stderr per device batch, carrying the batch's comparisons and cells (the
    sum of |A| x |B|), its status and the entry point that ran it; an engine
    created without it writes nothing."""
    I = emu_lib
    A = ["ACGTACGT", "GATTACA", "", "TTTT"]
    B = ["ACGTTACGT", "GATACA", "ACGT", "TTATT"]
    cells = sum(len(a) * len(b) for a, b in zip(A, B))
    cfg = I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1)
    monkeypatch.setenv("IPUMA_LOG", "1")
    eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1), cfg)
    monkeypatch.delenv("IPUMA_LOG")
    eng.compare(A, B)
    eng.compare(A[:2], B[:2])
    eng.close()
    err = capfd.readouterr().err
    lines = [json.loads(ln[len("IPUSWLOG"):]) for ln in err.splitlines()
             if ln.startswith("IPUSWLOG")]
    assert len(lines) == 2, err
    first, second = lines
    assert first["tag"] == "batch_perf" and first["entry"] == "compare_local"
    assert first["status"] == 0 and first["pairs"] == 4 and first["cells"] == cells
    assert first["padded_cells"] >= cells and first["host_ms"] > 0
    assert second["pairs"] == 2 and second["cells"] == 8 * 9 + 7 * 6
    eng2 = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1), cfg)
    eng2.compare(A, B)
    eng2.close()
    assert "IPUSWLOG" not in capfd.readouterr().err


def test_noalloc_entry_emulated(emu_lib):
    """The allocation-free entry on its static scratch: two calls in a row (the
    second smaller, with long and short pairs so band groups and deeper
    pipelines run) equal the allocating entry's results and the oracle's; its
    limits are refused."""
    import oracle as O
    from ipuma_amd.workloads import arena_from_flat
    I = emu_lib
    rng = np.random.default_rng(29)

    def batch(npairs, long_pairs):
        seqs = []
        for i in range(npairs):
            m = int(rng.integers(1, 300))
            a = rng.integers(1, 5, m).astype(np.uint8)
            b = np.concatenate([rng.integers(1, 5, int(rng.integers(0, 200))).astype(np.uint8),
                                a[: m // 2]])
            seqs += [a, b]
        for m, nb in long_pairs:  # |A| up to the 4096 limit, |B| past one super-strip
            a = rng.integers(1, 5, m).astype(np.uint8)
            seqs += [a, np.concatenate([a[:nb // 2], rng.integers(1, 5, nb - nb // 2).astype(np.uint8)])]
        lens = np.array([len(x) for x in seqs], np.int64)
        codes, offs = arena_from_flat(np.concatenate(seqs), lens)
        return (codes, offs[0::2].copy(), lens[0::2].astype(np.int32), offs[1::2].copy(),
                lens[1::2].astype(np.int32))

    eng = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=True))
    mat = O.select_matrix(O.SIM_NA, 1, -1, -1)

    def run(fn, b):
        codes, a_off, a_len, b_off, b_len = b
        n = len(a_len)
        out = [np.full(n, -7, np.int32) for _ in range(3)]
        fn(codes.ctypes.data, len(codes), a_off.ctypes.data, a_len.ctypes.data,
           b_off.ctypes.data, b_len.ctypes.data, *[o.ctypes.data for o in out], n=n)
        eng.synchronize()
        return out

    for b in (batch(40, [(4096, 1030)]), batch(20, [(700, 1100)])):
        got = run(eng.compare_device_noalloc, b)
        ref = run(eng.compare_device, b)
        exp = O.sw_batch_codes(*b, mat, -2, -1, True)
        for g, r, e in zip(got, ref, exp):
            assert np.array_equal(g, r)
            assert np.array_equal(g, e)
    assert eng.stats()["pairs"] == 21  # the second call's pairs
    # limits: more than 65536 comparisons, a forward-and-start engine, |A| > 4096
    z = np.zeros(65537, np.int32)
    with pytest.raises(I.IpumaError) as ei:
        eng.compare_device_noalloc(1, 4, 1, z.ctypes.data, 1, z.ctypes.data, 1, 1, 1, n=65537)
    assert ei.value.code == I.IPUMA_EINVAL and "65536" in str(ei.value)
    long_a = batch(0, [(4100, 20)])
    with pytest.raises(I.IpumaError) as ei:
        run(eng.compare_device_noalloc, long_a)
    assert ei.value.code == I.IPUMA_ETOOLONG
    # the error belongs to that call: later batches of either entry report their own
    run(eng.compare_device, batch(3, []))
    run(eng.compare_device_noalloc, batch(3, []))
    eng.close()
    eng2 = I.SWAlgorithm(I.SWConfig(gapInit=-2, gapExtend=-1),
                         I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    with pytest.raises(I.IpumaError) as ei:
        run(eng2.compare_device_noalloc, batch(3, []))
    assert ei.value.code == I.IPUMA_EINVAL and "forward-only" in str(ei.value)
    eng2.close()


NLOHMANN = "/root/reference/extern/include"  # the reference tree's copy of nlohmann::json


@pytest.mark.skipif(not os.path.exists(os.path.join(NLOHMANN, "nlohmann", "json.hpp")),
                    reason="nlohmann/json.hpp not available")
def test_multidriver_shape_emulated(emu_lib, build_dir):
    """A caller in the shape of the reference's batch driver (JSON run
    configuration, getTileClockFrequency, prepared buffers from worker
    threads, asynchronous jobs joined by receiver threads, batch_perf metrics
    from the Job's cycle counts) against the emulated engine."""
    I = emu_lib
    exe = os.path.join(build_dir, "multidriver_emu")
    cpp = os.path.join(ROOT, "ipuma-lib_amd", "cpp")
    subprocess.run([CLANG, "-std=c++17", "-O1", "-pthread", "-I" + cpp,
                    "-I" + os.path.join(ROOT, "include"), "-isystem", NLOHMANN,
                    os.path.join(HERE, "cpp", "multidriver_test.cpp"), I.LIB_PATH,
                    "-Wl,-rpath," + os.path.dirname(I.LIB_PATH), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "PASSED" in r.stdout and r.stdout.count("IPUSWLOG") == 6  # 360 comparisons / 60


def test_int32_wave_kernel_emulated(emu_lib):
    """The exact int32 kernel (one wave per comparison, a 64-lane pipeline of
    16 rows per lane, the boundary row between super-strips in the wave's work
    area): the int32 policy on pairs spanning up to two super-strips, gap
    settings outside the packed kernel's domain (a gap that extends for free,
    a positive gap-open value), and an int16-saturated pair over three
    super-strips re-run from the packed kernel's list, against the oracle."""
    import oracle as O
    I = emu_lib
    rng = np.random.default_rng(61)
    A, B = [], []
    for i in range(6):
        m = int(rng.integers(1, 120))
        a = "".join(rng.choice(list("ACGT"), m))
        nb = int(rng.integers(1030, 1300)) if i % 3 == 0 else int(rng.integers(1, 120))
        b = "".join(rng.choice(list("ACGT"), max(0, nb - m // 2))) + a[: m // 2]
        A.append(a)
        B.append(b[:nb])
    A += ["", "ACGT"]
    B += ["ACGT", ""]
    cases = [((-2, -1, 1, -1), I.PREC_INT32, (True, False)), ((0, 0, 1, -1), I.PREC_AUTO, (True,)),
             ((3, -1, 1, -1), I.PREC_AUTO, (False,)), ((-2, -1, 127, -127), I.PREC_AUTO, (True,))]
    for (gi, ge, ma, mi), prec, fwds in cases:
        for fwd in fwds:
            eng = I.SWAlgorithm(I.SWConfig(gapInit=gi, gapExtend=ge, matchValue=ma,
                                           mismatchValue=mi, ambiguityValue=mi),
                                I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                                forwardOnly=fwd))
            eng.set_precision(prec)
            if ma == 127:  # 300 x 127 > 32767, B over three super-strips: saturates
                A, B = A + ["A" * 300], B + ["A" * 2060]
            exp = O.sw_batch(A, B, match=ma, mismatch=mi, ambiguity=mi, gap_init=gi,
                             gap_extend=ge, forward_only=fwd)
            got = eng.compare(A, B)
            for g, e, nm in zip(got, exp, ("scores", "a_range", "b_range")):
                bad = np.nonzero(np.asarray(g) != np.asarray(e))[0]
                assert bad.size == 0, ((gi, ge, ma, mi), prec, fwd, nm, bad[:5],
                                       np.asarray(g)[bad[:5]], np.asarray(e)[bad[:5]])
            if ma == 127:
                assert eng.stats()["rerun_pairs"] > 0
            eng.close()


def _build_host_threads(build_dir, tag, san):
    """tests/cpu_emulation/host_threads.cpp linked with the engine, server,
    kernel, facade and oracle sources, all built against the emulation with
    the given sanitizer flags; returns the executable's path."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(build_dir, tag)
    os.makedirs(out, exist_ok=True)
    cpp = os.path.join(ROOT, "ipuma-lib_amd", "cpp")
    jobs = [(os.path.join(CSRC, "sw_engine.cpp"), ["-DEMU_DEFINE_GLOBALS"], "c++")]
    jobs += [(os.path.join(CSRC, f), [], "c++") for f in
             ("sw_server.cpp", "sw_kernels.hip", "sw_search.hip", "sw_tables.cpp")]
    jobs += [(os.path.join(cpp, f), ["-I" + cpp], "c++") for f in sorted(os.listdir(cpp))
             if f.endswith(".cpp")]
    jobs += [(os.path.join(EMU, "host_threads.cpp"), [], "c++")]

    def one(job):
        src, extra, lang = job
        o = os.path.join(out, os.path.basename(src) + ".o")
        _compile(src, o, san + extra, lang)
        return o

    with ThreadPoolExecutor(8) as ex:
        objs = list(ex.map(one, jobs))
    orc = os.path.join(out, "sw_oracle.o")
    subprocess.run(["/opt/rocm/llvm/bin/clang", "-O1", "-g", "-std=c11", *san, "-c",
                    os.path.join(ROOT, "oracle", "sw_oracle.c"), "-o", orc], check=True)
    exe = os.path.join(out, "host_threads")
    subprocess.run([CLANG, *san, "-pthread", *objs, orc, "-o", exe], check=True)
    return exe


@pytest.mark.skipif(os.environ.get("IPUMA_SANITIZER_TESTS") != "1",
                    reason="minutes under the sanitizers: set IPUMA_SANITIZER_TESTS=1")
@pytest.mark.parametrize("sanitizer", ["thread", "undefined"])
def test_host_layer_sanitized(build_dir, sanitizer):
    """The request server (four submitting threads, a blocking 4 KiB arena),
    prepared jobs from two threads on the engine's executor, and the staging
    threads of a 9 MB host batch (with thread starts refused too), under
    ThreadSanitizer and under UndefinedBehaviorSanitizer, against the oracle."""
    san = [f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer"]
    if sanitizer == "undefined":
        san.append("-fno-sanitize-recover=undefined")
    exe = _build_host_threads(build_dir, f"san_{sanitizer}", san)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=1500, env=env)
    assert r.returncode == 0 and "PASSED" in r.stdout, (r.stdout[-3000:] + r.stderr[-6000:])
