#!/usr/bin/env python3
"""Benchmark of the batched Smith-Waterman engine on one MI355X.

The headline (``value``) is the protein batch: 100,000 BLOSUM62 pairs,
lognormal lengths (median 300, sigma 0.5, clipped to 50..1024), half of them
related at 40% identity, gap open 11 / extend 1, compared with the plain
kernel padded to the batch maximum.  One step = one device batch over all
pairs (device-side bucketing, the packed int16 kernel, the exact int32 re-run
of saturated pairs), scores plus A/B end positions (forward pass, as the
reference's benchmark runs with --forwardOnly); inputs are resident in HBM
before the timed region.

The same default run also measures, each as a sub-record under ``configs``
with its own parity flags and roofline (packed-kernel time from HIP events):
  dna1024   1,024 DNA pairs (|A| 100-150, |B| 150-300, B embeds a mutated A),
            open 3 / extend 1, against the plain padded kernel;
  skewed    1,000,000 DNA pairs 150 x U[150,400] plus 5,000 power-law
            (alpha 1.5) pairs of 1-16 kb at 85% identity, against the plain
            kernel padding each wave's group of input-order pairs to its own
            longest pair (padding every group to the batch maximum would take
            minutes per step);
  search    8 protein queries (128..4096) x 200,000 database sequences, every
            score plus the top 100 per query;
  int32_rerun  1,000 DNA pairs of 2-4 kb at 95% identity scored +20 / -20,
            gap open 20 / extend 2, whose scores pass the int16 range, so
            every pair is re-run exactly in int32;
  traceback  CIGARs of 20,000 protein pairs from a compare with starts
            (k_trace_wave, a wave per comparison, int32);
and ``serving``: Poisson arrivals at 50, 200 and 1,000 req/s (shortened to
--serve-seconds each), merged batches and one launch per request.
``valu_peak`` is the measured packed-int16 / int32 VALU throughput the
roofline fractions are also reported against.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ipuma-lib_amd"))

OPS_PER_CELL = 11          # E: 2 add + max, F: 2 add + max, H: add + 3 max, best: max
PEAK_INT16_TOPS = 78.6     # 256 CU x 4 SIMD x 16 lanes x 2.4 GHz x 2 (packed int16)
PEAK_INT32_TOPS = 39.3     # the same, one op per lane
METRIC = "GCUPS on 1 MI355X (bucketed vs plain, bit-exact); p50/p99 latency at fixed req/s"
ALL_SUBCONFIGS = ("dna1024", "skewed", "search", "int32_rerun", "traceback")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=100_000)
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="steps of the plain padded-batch kernel for the comparison (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the baseline sample (0 = skip)")
    ap.add_argument("--with-starts", action="store_true", help="time the reverse pass too")
    ap.add_argument("--rounds", type=int, default=0,
                    help="groups per resident wave for the bucketing (0 = engine default)")
    ap.add_argument("--cap-scale", type=float, default=0.0,
                    help="step cap scale of the bucketing of the --config workload (0 = engine "
                         "default or the autotuner's choice; set: no autotuning of that workload; "
                         "the sub-records keep their autotuning)")
    ap.add_argument("--autotune", choices=["on", "off"], default="on",
                    help="time a small grid of bucketing shapes (rounds x step cap scale) on "
                         "each batch before its timed region and keep the fastest (off with "
                         "--rounds)")
    ap.add_argument("--serve-rates", default="50,200,1000",
                    help="comma-separated request rates (req/s) of the serving measurement")
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="seconds of Poisson arrivals per rate (0 = skip serving)")
    ap.add_argument("--serve-unbatched", choices=["on", "off"], default="on",
                    help="also run the one-launch-per-request baseline at every rate")
    ap.add_argument("--host-buffers", action="store_true",
                    help="also time the host-buffer entry (ipuma_compare_local: characters in "
                         "host memory, results back to host) on the protein batch; reported "
                         "beside value, never as it")
    ap.add_argument("--config", choices=["protein", "skewed", "search", "dna1024", "int32_rerun"],
                    default="protein",
                    help="the headline workload (value); the default protein run also measures "
                         "the other configurations (--subconfigs)")
    ap.add_argument("--subconfigs", default=",".join(ALL_SUBCONFIGS),
                    help="comma-separated configurations measured beside the protein headline "
                         "('' = none)")
    ap.add_argument("--sub-steps", type=int, default=5,
                    help="timed steps of each sub-configuration")
    ap.add_argument("--calibrate", choices=["on", "off"], default="on",
                    help="measure the packed-int16 / int32 VALU peak (ipuma_valu_peak)")
    args = ap.parse_args()
    unknown = [c for c in args.subconfigs.split(",") if c and c not in ALL_SUBCONFIGS]
    if unknown:
        ap.error(f"--subconfigs: unknown {', '.join(unknown)} (choose from "
                 f"{', '.join(ALL_SUBCONFIGS)})")
    return args


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (else the platform's machine name)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def matrix_of(O, sw):
    if sw.get("similarity", "na") == "blosum62":
        return O.select_matrix(O.SIM_BLOSUM62)
    return O.select_matrix(O.SIM_NA, sw["matchValue"], sw["mismatchValue"], sw["ambiguityValue"])


def cpu_baseline(batch, target_s, gpu_scores=None):
    """The reference's CPU path (its SSW comparator, compiled from the reference's sources
    into oracle/_ref) on a bounded leading sample of the batch, OpenMP over pairs; where that
    library is absent, the oracle (C restatement of the IPU vertex) instead."""
    import numpy as np
    O = oracle()
    O.build()
    sw = batch["sw"]
    mat = matrix_of(O, sw)
    threads = O.max_threads()
    cells = batch["a_len"].astype(np.int64) * batch["b_len"].astype(np.int64)
    csum = np.cumsum(cells)
    reference = O.have_ref()
    out = {}

    def run(k, nthreads=threads):
        t = time.perf_counter()
        if reference:
            out["s"] = O.ssw_batch_codes(batch["codes"], batch["a_off"][:k], batch["a_len"][:k],
                                         batch["b_off"][:k], batch["b_len"][:k], mat,
                                         -(sw["gapInit"] + sw["gapExtend"]), -sw["gapExtend"],
                                         nthreads)
        else:
            out["s"] = O.sw_batch_codes(batch["codes"], batch["a_off"][:k], batch["a_len"][:k],
                                        batch["b_off"][:k], batch["b_len"][:k], mat,
                                        sw["gapInit"], sw["gapExtend"], True, nthreads)[0]
        return time.perf_counter() - t

    k0 = int(np.searchsorted(csum, 2e8 * threads / 16)) + 1
    k0 = min(k0, len(cells))
    t0 = run(k0)
    rate = csum[k0 - 1] / max(t0, 1e-6)
    k = int(np.searchsorted(csum, rate * target_s)) + 1
    k = min(max(k, k0), len(cells))
    t = run(k)
    g = csum[k - 1] / t / 1e9
    what = ("the reference's SSW comparator (cpusw/aligners.hpp: query profile + ssw_align, "
            "end positions only)" if reference else "int32 oracle")
    res = {"value": round(float(g), 3), "unit": "GCUPS", "cores": int(threads),
           "kind": "reference" if reference else "port",
           "sample": f"first {k} of {len(cells)} pairs ({int(csum[k - 1])} cells), forward "
                     f"pass, {what}, {t:.1f} s on {threads} threads "
                     f"of {os.cpu_count()} visible ({cpu_model()})"}
    if gpu_scores is not None:
        res["scores_equal_gpu"] = bool(np.array_equal(np.asarray(out["s"])[:k],
                                                      np.asarray(gpu_scores)[:k]))
    # one core on a shorter leading sample (about a tenth of the time above)
    k1 = int(np.searchsorted(csum, g * 1e9 / threads * max(target_s / 10, 0.5))) + 1
    k1 = min(max(k1, 1), len(cells))
    t1 = run(k1, 1)
    res["value_1core"] = round(float(csum[k1 - 1] / t1 / 1e9), 3)
    res["sample_1core"] = f"first {k1} pairs ({int(csum[k1 - 1])} cells), {t1:.1f} s on 1 thread"
    return res


def measured_traffic():
    """HBM bytes per launch of the packed kernel on the protein batch, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by a 1 GiB copy's counts)."""
    p = os.path.join(ROOT, "profiles", "k_fast_protein_100k_traffic.json")
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def valu_peak():
    """Measured packed-int16 and int32 VALU throughput of this GPU (ipuma_valu_peak)."""
    import ipuma_amd as I
    L = I.lib()
    L.ipuma_valu_peak.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2
    p16, p32 = ctypes.c_double(0), ctypes.c_double(0)
    rc = L.ipuma_valu_peak(ctypes.byref(p16), ctypes.byref(p32))
    if rc != 0:
        log(f"valu peak measurement failed ({rc})")
        return None
    return {"pk_int16_tops": round(p16.value, 2), "int32_tops": round(p32.value, 2),
            "nominal_pk_int16_tops": PEAK_INT16_TOPS, "nominal_int32_tops": PEAK_INT32_TOPS,
            "what": "k_valu_pk16 / k_valu_i32: 8 independent chains of saturating add + max "
                    "per lane, 32 waves per CU, best of 3 launches"}


def roofline(cells, kernel_ms, kernel, peak_cal=None, traffic=None, precision="int16"):
    peak = PEAK_INT16_TOPS if precision == "int16" else PEAK_INT32_TOPS
    achieved = cells * OPS_PER_CELL / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    r = {"bound": "valu", "achieved": round(achieved, 2), "peak": peak,
         "unit": f"TOP/s ({precision})", "frac": round(achieved / peak, 4),
         "traffic": traffic["traffic_bytes"] if traffic else None,
         "kernel": kernel, "kernel_ms": round(kernel_ms, 3), "ops_per_cell": OPS_PER_CELL}
    if traffic:
        r["traffic_source"] = traffic["source"]
        r["hbm_gbs"] = round(traffic["traffic_bytes"] / (kernel_ms * 1e6), 1)
    if peak_cal:
        r["peak_calibrated"] = peak_cal
        r["frac_calibrated"] = round(achieved / peak_cal, 4)
    return r


class DeviceBatch:
    """A batch resident in HBM plus its result tensors and an engine to run it."""

    def __init__(self, batch, forward_only=True, rounds=0):
        import torch
        import ipuma_amd as I
        self.batch = batch
        dev = torch.device("cuda:0")
        self.t = {k: torch.from_numpy(batch[k]).to(dev)
                  for k in ("codes", "a_off", "a_len", "b_off", "b_len")}
        self.n = len(batch["a_len"])
        self.out = [torch.empty(self.n, dtype=torch.int32, device=dev) for _ in range(3)]
        self.eng = I.SWAlgorithm(I.SWConfig(**batch["sw"]),
                                 I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                                 forwardOnly=forward_only))
        if rounds:
            self.eng.set_work_rounds(rounds)
        self.sptr = torch.cuda.current_stream().cuda_stream

    def step(self):
        t = self.t
        self.eng.compare_device(t["codes"], t["codes"].numel(), t["a_off"], t["a_len"],
                                t["b_off"], t["b_len"], *self.out, stream=self.sptr)

    def tune(self, alphabet, reps=3):
        import torch
        from ipuma_amd import autotune as AT
        key = AT.signature(self.batch["a_len"], self.batch["b_len"], alphabet)
        tuned = AT.tune(self.eng, self.step, key=key, reps=reps, sync=torch.cuda.synchronize)
        tab = ", ".join("%dx%.2f: %.2f ms" % (r["rounds"], r["scale"], r["ms"])
                        for r in tuned["table"])
        log(f"autotune {key}: rounds {tuned['rounds']} scale {tuned['scale']} ({tab})")
        return tuned

    def timed(self, steps, warmup):
        """(ms per step over `steps` back-to-back steps, mean packed-kernel ms from the
        engine's HIP events over `steps` more, the last batch's stats)."""
        import numpy as np
        import torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        self.eng.set_timing(True)
        fast = []
        st = {}
        for _ in range(steps):
            self.step()
            st = self.eng.stats()  # synchronizes; reads this step's events
            fast.append(st["fast_kernel_ms"])
        self.eng.set_timing(False)
        return ms, float(np.mean(fast)), st

    def plain(self, mode, steps):
        import torch
        import ipuma_amd as I
        self.eng.set_batching(mode)
        self.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        pe = (time.perf_counter() - t0) / steps
        st = self.eng.stats()
        self.eng.set_batching(I.BATCH_BUCKETED)
        return pe * 1e3, st

    def results(self):
        return [o.cpu().numpy() for o in self.out]

    def close(self):
        self.eng.close()


def oracle_equal(batch, idx, got, forward_only=True):
    """Scores and ranges of pairs `idx` against the oracle."""
    import numpy as np
    O = oracle()
    sw = batch["sw"]
    idx = np.asarray(idx)
    exp = O.sw_batch_codes(batch["codes"], batch["a_off"][idx], batch["a_len"][idx],
                           batch["b_off"][idx], batch["b_len"][idx], matrix_of(O, sw),
                           sw["gapInit"], sw["gapExtend"], forward_only)
    return all(bool(np.array_equal(g[idx], e)) for g, e in zip(got, exp))


def bench_dna1024(args, peak_cal):
    """1,024 DNA pairs, open 3 / extend 1, bucketed vs the plain padded kernel; every pair
    against the oracle."""
    import ipuma_amd as I
    from ipuma_amd import workloads as W
    b = W.config0()
    db = DeviceBatch(b)
    steps = max(args.sub_steps * 4, 10)
    ms, kms, st = db.timed(steps, 3)
    got = db.results()
    parity = oracle_equal(b, range(db.n), got)
    pms, pst = db.plain(I.BATCH_PLAIN, steps)
    plain_equal = all(bool((g == h).all()) for g, h in zip(got, db.results()))
    db.close()
    cells = b["cells"]
    rec = {"workload": "1,024 DNA pairs, |A| U[100,150], |B| U[150,300], B embeds a mutated A "
                       "(10% substitutions, 2% indels), match 1 mismatch -1 gap open 3 extend 1, "
                       "scores + end positions",
           "pairs": db.n, "cells": cells, "steps": steps,
           "ms_per_step": round(ms, 3), "gcups": round(cells / ms / 1e6, 1),
           "plain": {"kind": "plain kernel padded to the batch maximum",
                     "ms_per_step": round(pms, 3), "gcups": round(cells / pms / 1e6, 1),
                     "padded_cells": pst["padded_cells"]},
           "bucketed_vs_plain": round(pms / ms, 2),
           "parity": {"all_pairs_oracle": parity, "plain_equal": plain_equal},
           "roofline": roofline(cells, kms, "k_fast<18, false>", peak_cal),
           "stats": {k: st[k] for k in ("groups", "padded_cells", "lanes_per_pair_max")}}
    log(f"dna1024: {rec['ms_per_step']} ms/step, {rec['gcups']} GCUPS, plain "
        f"{rec['plain']['ms_per_step']} ms, parity {rec['parity']}")
    return rec


def bench_skewed(args, peak_cal):
    """1,000,000 short + 5,000 long DNA pairs: bucketed persistent kernel vs the plain
    per-wave-padded kernel (target: >= 2x)."""
    import numpy as np
    import ipuma_amd as I
    from ipuma_amd import workloads as W
    t = time.time()
    b = W.config2()
    log(f"skewed: generated {len(b['a_len'])} pairs, {b['cells']:.3e} cells in "
        f"{time.time() - t:.1f}s")
    db = DeviceBatch(b)
    fixed = args.config == "skewed" and args.cap_scale
    if fixed:
        db.eng.set_step_cap_scale(args.cap_scale)
    tuned = db.tune("dna", reps=2) if args.autotune == "on" and not fixed else None
    ms, kms, st = db.timed(args.sub_steps, 1)
    got = db.results()
    pms, pst = db.plain(I.BATCH_PLAIN_TILE, 1)
    plain_got = db.results()
    db.close()
    plain_equal = all(bool((g == h).all()) for g, h in zip(got, plain_got))
    cells = b["cells"]
    long_idx = np.argsort(b["a_len"].astype(np.int64) * b["b_len"])[-40:]
    rng = np.random.default_rng(0)
    short_idx = rng.choice(len(b["a_len"]) - 5000, 2000, replace=False)
    parity_long = oracle_equal(b, long_idx, got)
    parity_short = oracle_equal(b, short_idx, got)
    rec = {"workload": "skewed DNA batch: 1,000,000 pairs 150 x U[150,400] (B embeds a mutated "
                       "A) + 5,000 pairs power-law(alpha 1.5) 1-16 kb at 85% identity, match 1 "
                       "mismatch -1 gap open 3 extend 1, scores + end positions",
           "pairs": int(len(b["a_len"])), "cells": cells, "steps": args.sub_steps,
           "ms_per_step": round(ms, 3), "gcups": round(cells / ms / 1e6, 1),
           "plain": {"kind": "plain kernel, input order, each wave padded to its longest pair",
                     "ms_per_step": round(pms, 2), "gcups": round(cells / pms / 1e6, 1),
                     "padded_cells": pst["padded_cells"]},
           "bucketed_vs_plain": round(pms / ms, 2), "target_bucketed_vs_plain": 2.0,
           "parity": {"bucketed_equals_plain_all_pairs": plain_equal,
                      "oracle_40_longest": parity_long, "oracle_2000_short": parity_short},
           "roofline": roofline(cells, kms, "k_fast<18, false>", peak_cal),
           "autotune": ({"rounds": tuned["rounds"], "step_cap_scale": tuned["scale"]}
                        if tuned else None),
           "step_cap_scale": args.cap_scale if fixed else None,
           "stats": {k: st[k] for k in ("groups", "padded_cells", "band_groups")}}
    log(f"skewed: {rec['ms_per_step']} ms/step, {rec['gcups']} GCUPS, plain "
        f"{rec['plain']['ms_per_step']} ms ({rec['bucketed_vs_plain']}x), parity {rec['parity']}")
    return rec


def bench_search(args, peak_cal, cpu_seconds=0.0):
    """One-vs-many search: 8 protein queries (128..4096) against 200,000 database
    sequences, every score and the top 100 per query."""
    import numpy as np
    import torch
    import ipuma_amd as I
    from ipuma_amd import workloads as W
    O = oracle()
    t = time.time()
    b = W.config4()
    nq, ndb, k = len(b["q_len"]), len(b["db_len"]), 100
    cells = b["cells"]
    log(f"search: generated {nq} queries x {ndb} database sequences, {cells:.3e} cells in "
        f"{time.time() - t:.1f}s")
    dev = torch.device("cuda:0")
    d = {x: torch.from_numpy(np.ascontiguousarray(b[x])).to(dev)
         for x in ("codes", "q_off", "q_len", "db_off", "db_len")}
    scores = torch.empty(nq * ndb, dtype=torch.int32, device=dev)
    ti = torch.empty(nq * k, dtype=torch.int32, device=dev)
    ts = torch.empty(nq * k, dtype=torch.int32, device=dev)
    eng = I.SWAlgorithm(I.SWConfig(**b["sw"]),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    sptr = torch.cuda.current_stream().cuda_stream

    def step():
        eng.search_device(d["codes"], d["codes"].numel(), d["q_off"], d["q_len"], d["db_off"],
                          d["db_len"], k, scores, ti, ts, stream=sptr)

    tuned = None
    fixed = args.config == "search" and args.cap_scale
    if fixed:
        eng.set_step_cap_scale(args.cap_scale)
    if args.autotune == "on" and not args.rounds and not fixed:
        from ipuma_amd import autotune as AT
        dbs = b["db_len"][:4096]  # pairs of every query with a database sample
        key = AT.signature(np.repeat(b["q_len"], len(dbs)), np.tile(dbs, nq), "search-blosum62")
        tuned = AT.tune(eng, step, key=key, reps=2, sync=torch.cuda.synchronize)
        tab = ", ".join("%dx%.2f: %.1f ms" % (r["rounds"], r["scale"], r["ms"])
                        for r in tuned["table"])
        log(f"autotune {key}: rounds {tuned['rounds']} scale {tuned['scale']} ({tab})")
    steps = args.sub_steps
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / steps * 1e3
    gcups = cells / (elapsed / steps) / 1e9
    eng.set_timing(True)
    fast_ms = []
    for _ in range(steps):
        step()
        fast_ms.append(eng.stats()["fast_kernel_ms"])
    eng.set_timing(False)
    st = eng.stats()
    fast_avg = float(np.mean(fast_ms))
    # parity: oracle scores of every query against a 2000-sequence subset, the top 100
    # recomputed on the host from the device's full score array, and the oracle's scores
    # of every top-100 entry
    got = scores.cpu().numpy().reshape(nq, ndb)
    top_i = ti.cpu().numpy().reshape(nq, k)
    top_s = ts.cpu().numpy().reshape(nq, k)
    mat = O.select_matrix(O.SIM_BLOSUM62)
    sub = 2000
    ok_subset = ok_top = ok_top_oracle = True
    for q in range(nq):
        exp, _, _ = O.sw_batch_codes(b["codes"], b["db_off"][:sub], b["db_len"][:sub],
                                     np.full(sub, b["q_off"][q]),
                                     np.full(sub, b["q_len"][q], np.int32), mat,
                                     b["sw"]["gapInit"], b["sw"]["gapExtend"], True)
        ok_subset &= bool(np.array_equal(got[q, :sub], exp))
        order = np.lexsort((np.arange(ndb), -got[q].astype(np.int64)))[:k]
        ok_top &= bool(np.array_equal(top_i[q], order) and np.array_equal(top_s[q], got[q, order]))
        e2, _, _ = O.sw_batch_codes(b["codes"], b["db_off"][top_i[q]], b["db_len"][top_i[q]],
                                    np.full(k, b["q_off"][q]), np.full(k, b["q_len"][q], np.int32),
                                    mat, b["sw"]["gapInit"], b["sw"]["gapExtend"], True)
        ok_top_oracle &= bool(np.array_equal(e2, top_s[q]))
    planted = set(int(x) for x in b["homologs"])
    found = len(planted & set(int(x) for x in top_i.ravel()))
    eng.close()
    cpu = None
    if cpu_seconds > 0:
        sample = {"codes": b["codes"], "sw": b["sw"],
                  "a_off": b["db_off"][:20000], "a_len": b["db_len"][:20000],
                  "b_off": np.full(20000, b["q_off"][4]),
                  "b_len": np.full(20000, b["q_len"][4], np.int32)}
        cpu = cpu_baseline(sample, cpu_seconds)
        cpu["sample"] = "query 5 (1024 residues) vs " + cpu["sample"]
    rec = {"workload": "one-vs-many search: 8 protein queries (128..4096) x 200,000 database "
                       "sequences lognormal(350, 0.5) clipped 30-5000, BLOSUM62 open 11 extend 1, "
                       "every score + top 100 per query",
           "queries": nq, "database": ndb, "cells": cells, "steps": steps,
           "ms_per_step": round(ms_step, 3), "gcups": round(gcups, 1),
           "parity": {"subset_2000_scores": ok_subset, "top100_order": ok_top,
                      "top100_oracle_scores": ok_top_oracle,
                      "planted_homologs_found": f"{found}/{len(planted)}"},
           "roofline": roofline(cells, fast_avg, "k_fast<26, false>", peak_cal),
           "autotune": ({"rounds": tuned["rounds"], "step_cap_scale": tuned["scale"]}
                        if tuned else None),
           "step_cap_scale": args.cap_scale if fixed else None,
           "cpu_baseline": cpu,
           "stats": {kk: st[kk] for kk in ("groups", "padded_cells", "band_groups")}}
    log(f"search: {rec['ms_per_step']} ms/step, {rec['gcups']} GCUPS, parity {rec['parity']}")
    return rec


def bench_rerun32(args, peak_cal):
    """1,000 saturating 2-4 kb pairs: the exact int32 re-run on many lanes."""
    from ipuma_amd import workloads as W
    b = W.saturating_pairs()
    db = DeviceBatch(b)
    steps = max(2, args.sub_steps)
    ms, kms, st = db.timed(steps, 1)
    got = db.results()
    import numpy as np
    rng = np.random.default_rng(1)
    idx = rng.choice(db.n, 60, replace=False)
    parity = oracle_equal(b, idx, got)
    db.close()
    cells = b["cells"]
    rec = {"workload": "1,000 DNA pairs 2-4 kb at 95% identity, match 20 mismatch -20 gap open "
                       "20 extend 2 (optimal scores above 32,767: every pair saturates int16 "
                       "and is re-run exactly in int32), scores + end positions",
           "pairs": db.n, "cells": cells, "steps": steps,
           "ms_per_step": round(ms, 3), "gcups": round(cells / ms / 1e6, 1),
           "rerun_pairs": st["rerun_pairs"], "min_score": int(got[0].min()),
           "parity": {"oracle_60_pairs": parity},
           "roofline": roofline(cells, st["batch_ms"] - st["fast_kernel_ms"],
                                "k_wave32 (device batch minus the packed kernel, HIP events)",
                                getattr(args, "pk32", None), precision="int32")}
    log(f"int32 re-run: {rec['ms_per_step']} ms/step, {rec['gcups']} GCUPS, "
        f"{st['rerun_pairs']} re-run pairs, parity {rec['parity']}")
    return rec


def bench_traceback(args, peak_cal):
    """CIGARs of 20,000 protein pairs (the protein batch's generator) from a compare with
    starts: host characters in, runs out (ipuma_traceback_local), k_trace_wave timed with
    HIP events; 300 CIGARs against the oracle."""
    import numpy as np
    import ipuma_amd as I
    from ipuma_amd import workloads as W
    b = W.config1(n=20_000)
    letters = np.frombuffer(b"ARNDCQEGHILKMFPSTWYVBZX*", np.uint8)
    codes = b["codes"]
    A = [letters[codes[o:o + n]].tobytes() for o, n in zip(b["a_off"], b["a_len"])]
    B = [letters[codes[o:o + n]].tobytes() for o, n in zip(b["b_off"], b["b_len"])]
    eng = I.SWAlgorithm(I.SWConfig(**b["sw"]),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1))
    s, ar, br = eng.compare(A, B)
    n = len(A)
    abuf, aoff = I._pack(A)
    bbuf, boff = I._pack(B)
    u = ar.astype(np.int64) & 0xFFFFFFFF
    v = br.astype(np.int64) & 0xFFFFFFFF
    cap = np.where(s > 0, ((u >> 16) - (u & 0xFFFF) + 1) + ((v >> 16) - (v & 0xFFFF) + 1), 0)
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum(cap)
    cig = np.zeros(max(int(offs[-1]), 1), np.uint32)
    lens = np.zeros(n, np.int32)
    L = I.lib()

    def call():
        I._check(L.ipuma_traceback_local(eng._h, n, abuf, aoff.ctypes.data, bbuf, boff.ctypes.data,
                                         s.ctypes.data, ar.ctypes.data, br.ctypes.data,
                                         cig.ctypes.data, offs.ctypes.data, lens.ctypes.data))

    eng.set_timing(True)
    call()
    steps = max(3, args.sub_steps)
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
        kms.append(eng.traceback_stats()["kernel_ms"])
    ms = (time.perf_counter() - t0) / steps * 1e3
    cells = eng.traceback_stats()["region_cells"]
    O = oracle()
    mat = matrix_of(O, b["sw"])
    rng = np.random.default_rng(3)
    ok = True
    for k in rng.choice(n, 300, replace=False):
        a = codes[b["a_off"][k]:b["a_off"][k] + b["a_len"][k]]
        bb = codes[b["b_off"][k]:b["b_off"][k] + b["b_len"][k]]
        exp = O.cigar(a, bb, int(ar[k]), int(br[k]), int(s[k]), mat, b["sw"]["gapInit"],
                      b["sw"]["gapExtend"])
        ok &= bool(np.array_equal(cig[offs[k]:offs[k] + lens[k]], exp))
    eng.close()
    kbest = min(kms)
    rec = {"workload": "traceback (CIGAR) of 20,000 protein pairs of the protein batch's "
                       "generator from a compare with starts, BLOSUM62, gap open 11 extend 1",
           "pairs": n, "region_cells": cells, "steps": steps,
           "ms_per_call": round(ms, 3),
           "gcups_call": round(cells / ms / 1e6, 1),
           "kernel_ms": round(float(np.median(kms)), 3), "kernel_ms_min": round(kbest, 3),
           "gcups_kernel": round(cells / float(np.median(kms)) / 1e6, 1),
           "what": "ms_per_call: ipuma_traceback_local with host characters in and runs out "
                   "(staging, kernel, copy back); kernel: k_trace_wave, HIP events",
           "parity": {"oracle_cigar_300_pairs": ok},
           "roofline": roofline(cells, float(np.median(kms)), "k_trace_wave (HIP events)",
                                getattr(args, "pk32", None), precision="int32")}
    log(f"traceback: {rec['ms_per_call']} ms/call, kernel {rec['kernel_ms']} ms, "
        f"{rec['gcups_kernel']} GCUPS (kernel), parity {rec['parity']}")
    return rec


def bench_serving(args):
    from ipuma_amd import loadgen
    log("serving: generating request pools ...")
    pools = loadgen.make_pools()
    out = []
    for rate in [float(r) for r in args.serve_rates.split(",") if r]:
        for merge in ([True, False] if args.serve_unbatched == "on" else [True]):
            res = loadgen.run(pools, rate, args.serve_seconds, merge=merge)
            log(f"serving {res}")
            out.append(res)
    return out


def headline_record(value, ms_step, args, data, config, dtype):
    return {"metric": METRIC, "value": value, "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data,
            "config": config}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(pairs are independent: a batch shards by cumulative cells, one engine per GPU, "
              "with a gather of the results by pair index)",
              file=sys.stderr)
        return 2
    import numpy as np
    import torch
    import ipuma_amd as I
    from ipuma_amd import workloads as W

    peak = valu_peak() if args.calibrate == "on" else None
    if peak:
        log(f"valu peak: {peak}")
    pk16 = peak["pk_int16_tops"] if peak else None
    args.pk32 = peak["int32_tops"] if peak else None

    if args.config != "protein":
        fn = {"skewed": bench_skewed, "search": bench_search, "dna1024": bench_dna1024,
              "int32_rerun": bench_rerun32}[args.config]
        args.sub_steps = args.steps
        rec = (fn(args, pk16, args.cpu_seconds) if args.config == "search" else fn(args, pk16))
        out = headline_record(rec["gcups"], rec["ms_per_step"], args, "synthetic (seeded)",
                              {"workload": rec["workload"]},
                              "int32" if args.config == "int32_rerun" else
                              "int16x2 (int32 re-run on saturation)")
        out.update({k: v for k, v in rec.items() if k not in ("gcups", "ms_per_step", "workload")})
        out["valu_peak"] = peak
        print(json.dumps(out), flush=True)
        return 0

    t = time.time()
    batch = W.config1(n=args.pairs)
    log(f"generated {len(batch['a_len'])} pairs, {batch['cells']:.3e} cells in {time.time()-t:.1f}s")
    db = DeviceBatch(batch, forward_only=not args.with_starts, rounds=args.rounds)
    if args.cap_scale:
        db.eng.set_step_cap_scale(args.cap_scale)
    tuned = (db.tune("blosum62")
             if (args.autotune == "on" and not args.rounds and not args.cap_scale) else None)
    ms_step, fast_avg, st = db.timed(args.steps, args.warmup)
    cells = batch["cells"]
    gcups = cells / ms_step / 1e6
    log(f"bucketed: {ms_step:.2f} ms/step, {gcups:.1f} GCUPS end-to-end, packed kernel "
        f"{fast_avg:.2f} ms ({cells / fast_avg / 1e6:.1f} GCUPS), stats {st}")
    n = db.n
    scores_np = db.out[0].cpu().numpy()
    # correctness guard on the measured output: a 2000-pair slice against the oracle
    k = min(2000, n)
    got = db.results()
    parity = oracle_equal(batch, np.arange(k), got, forward_only=not args.with_starts)
    log(f"parity on {k} pairs: {parity}")

    plain = None
    if args.plain_steps > 0:
        pms, pst = db.plain(I.BATCH_PLAIN, args.plain_steps)
        plain = {"kind": "plain kernel padded to the batch maximum", "ms_per_step": round(pms, 2),
                 "gcups": round(cells / pms / 1e6, 1), "padded_cells": pst["padded_cells"]}
        log(f"plain padded batch: {plain}")

    cpu = None
    if args.cpu_seconds > 0:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(batch, args.cpu_seconds, gpu_scores=scores_np)
        log(f"cpu baseline: {cpu}")

    host_buffers = None
    if args.host_buffers:
        log("timing the host-buffer entry ...")
        letters = np.frombuffer(b"ARNDCQEGHILKMFPSTWYVBZX*", np.uint8)
        dec = np.zeros(256, np.uint8)
        dec[: len(letters)] = letters
        codes = batch["codes"]
        abuf, aoff = I._pack([dec[codes[o:o + l]].tobytes()
                              for o, l in zip(batch["a_off"], batch["a_len"])])
        bbuf, boff = I._pack([dec[codes[o:o + l]].tobytes()
                              for o, l in zip(batch["b_off"], batch["b_len"])])
        hs, ha, hb = (np.zeros(n, np.int32) for _ in range(3))

        def host_call():
            I._check(I.lib().ipuma_compare_local(db.eng._h, n, abuf, aoff.ctypes.data, bbuf,
                                                 boff.ctypes.data, hs.ctypes.data,
                                                 ha.ctypes.data, hb.ctypes.data))

        for _ in range(max(1, args.warmup)):
            host_call()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            host_call()
        hms = (time.perf_counter() - t0) / args.steps * 1e3
        dev_res = db.results()
        host_buffers = {"ms_per_step": round(hms, 3), "gcups": round(cells / hms / 1e6, 1),
                        "results_equal_device": all(bool(np.array_equal(h, d)) for h, d in
                                                    zip((hs, ha, hb), dev_res)),
                        "what": "ipuma_compare_local, host characters in, host results out "
                                "(PCIe-inclusive)"}
        log(f"host buffers: {host_buffers}")
    db.close()
    del db
    torch.cuda.empty_cache()

    configs = {}
    for name in [c for c in args.subconfigs.split(",") if c]:
        fn = {"dna1024": bench_dna1024, "skewed": bench_skewed, "search": bench_search,
              "int32_rerun": bench_rerun32, "traceback": bench_traceback}[name]
        try:
            configs[name] = fn(args, pk16)
        except Exception as e:  # reported in the line; the headline stands
            log(f"{name} failed: {e!r}")
            configs[name] = {"error": repr(e)}
        torch.cuda.empty_cache()

    serving = bench_serving(args) if args.serve_seconds > 0 else None

    tr = measured_traffic() if (args.pairs == 100_000 and not args.with_starts) else None
    out = headline_record(
        round(gcups, 1), round(ms_step, 3), args,
        "synthetic (seeded generator, BLOSUM62 background residues)",
        {"workload": "protein batch: 100k BLOSUM62 pairs, lognormal(300, 0.5) lengths clipped "
                     "50-1024, half related at 40% identity, gap open 11 extend 1, scores + end "
                     "positions",
         "pairs": n, "cells": cells, "starts": bool(args.with_starts),
         "work_rounds": args.rounds or "default",
         "step_cap_scale": args.cap_scale or None,
         "autotune": ({"rounds": tuned["rounds"], "step_cap_scale": tuned["scale"],
                       "candidates": len(tuned["table"])} if tuned else None)},
        "int16x2 (int32 re-run on saturation)")
    out.update({
        "bucketed_vs_plain": (round(gcups / plain["gcups"], 2) if plain else None),
        "plain": plain,
        "parity_2000_pairs": parity,
        "roofline": roofline(cells, fast_avg, "k_fast<26, false>", pk16, tr),
        "cpu_baseline": cpu,
        "host_buffers": host_buffers,
        "valu_peak": peak,
        "configs": configs,
        "serving": serving,
        "serving_note": (f"{args.serve_seconds:g} s of Poisson arrivals per rate (the "
                         "configuration asks for 60 s)" if serving else None),
    })
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
