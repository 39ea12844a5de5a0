#!/usr/bin/env python3
"""Benchmark of the batched Smith-Waterman engine on one MI355X.

Workloads (synthetic, seeded, resident in HBM before the timed region):
  protein (default, the headline): 100,000 BLOSUM62 pairs, lognormal lengths
    (median 300, sigma 0.5, clipped to 50..1024), half of them related at 40%
    identity, gap open 11 / extend 1; compared with the plain kernel padded to
    the batch maximum.
  skewed: 1,000,000 DNA pairs of 150 x U[150,400] (B embeds a mutated A) plus
    5,000 pairs with power-law lengths (alpha 1.5) over 1-16 kb at 85%
    identity, open 3 / extend 1; compared with the plain kernel that pads each
    wave's group of input-order pairs to its own longest pair (padding every
    group to the batch maximum would take minutes per step here).
One step = one device batch over all pairs: device-side bucketing, the packed
int16 kernel, the exact int32 re-run of saturated pairs; scores plus A/B end
positions (forward pass, as the reference's benchmark runs with --forwardOnly).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ipuma-lib_amd"))

OPS_PER_CELL = 11          # E: 2 add + max, F: 2 add + max, H: add + 3 max, best: max
PEAK_INT16_TOPS = 78.6     # 256 CU x 4 SIMD x 16 lanes x 2.4 GHz x 2 (packed int16)


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
    ap.add_argument("--autotune", choices=["on", "off"], default="on",
                    help="time a small grid of bucketing shapes (rounds x step cap scale) on "
                         "the batch before the timed region and keep the fastest (off with "
                         "--rounds)")
    ap.add_argument("--serve-rates", default="200",
                    help="comma-separated request rates (req/s) of the serving measurement")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="seconds of Poisson arrivals per rate (0 = skip serving)")
    ap.add_argument("--serve-unbatched", action="store_true",
                    help="also run the one-launch-per-request baseline at every rate")
    ap.add_argument("--host-buffers", action="store_true",
                    help="also time the host-buffer entry (ipuma_compare_local: characters in "
                         "host memory, results back to host) on the protein batch; reported "
                         "beside value, never as it")
    ap.add_argument("--config", choices=["protein", "skewed", "search"], default="protein",
                    help="protein: 100k BLOSUM62 pairs (the headline); skewed: 1M short + 5k "
                         "long DNA pairs, bucketed vs the plain padded batch")
    return ap.parse_args()


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


def cpu_baseline(batch, target_s, gpu_scores=None):
    """The reference's CPU path (its SSW comparator, compiled from the reference's sources
    into oracle/_ref) on a bounded leading sample of the batch, OpenMP over pairs; where that
    library is absent, the oracle (C restatement of the IPU vertex) instead."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    O.build()
    sw = batch["sw"]
    mat = (O.select_matrix(O.SIM_BLOSUM62) if sw["similarity"] == "blosum62"
           else O.select_matrix(O.SIM_NA, sw["matchValue"], sw["mismatchValue"],
                                sw["ambiguityValue"]))
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
    """HBM bytes per launch of the packed kernel on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by a 1 GiB copy's counts)."""
    p = os.path.join(ROOT, "profiles", "k_fast_protein_100k_traffic.json")
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def run_search(args):
    """One-vs-many search: 8 protein queries (128..4096) against 200,000 database
    sequences, every score and the top 100 per query."""
    import numpy as np
    import torch
    import ipuma_amd as I
    from ipuma_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    t = time.time()
    b = W.config4()
    nq, ndb, k = len(b["q_len"]), len(b["db_len"]), 100
    cells = b["cells"]
    log(f"generated {nq} queries x {ndb} database sequences, {cells:.3e} cells in "
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
    if args.autotune == "on" and not args.rounds:
        # bucketing shape chosen on this search, outside the timed region
        from ipuma_amd import autotune as AT
        dbs = b["db_len"][:4096]  # pairs of every query with a database sample
        key = AT.signature(np.repeat(b["q_len"], len(dbs)), np.tile(dbs, nq), "search-blosum62")
        tuned = AT.tune(eng, step, key=key, reps=2, sync=torch.cuda.synchronize)
        tab = ", ".join("%dx%.2f: %.1f ms" % (r["rounds"], r["scale"], r["ms"])
                        for r in tuned["table"])
        log(f"autotune {key}: rounds {tuned['rounds']} scale {tuned['scale']} ({tab})")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    gcups = cells / (elapsed / args.steps) / 1e9
    eng.set_timing(True)
    fast_ms = []
    for _ in range(args.steps):
        step()
        fast_ms.append(eng.stats()["fast_kernel_ms"])
    eng.set_timing(False)
    st = eng.stats()
    fast_avg = float(np.mean(fast_ms))
    achieved = cells * OPS_PER_CELL / (fast_avg * 1e-3) / 1e12
    log(f"search: {ms_step:.2f} ms/step, {gcups:.1f} GCUPS end-to-end, packed kernel "
        f"{fast_avg:.2f} ms ({cells / fast_avg / 1e6:.1f} GCUPS), stats {st}")
    # parity: oracle scores of every query against a 2000-sequence subset, the
    # top 100 recomputed on the host from the device's full score array, and the
    # oracle's scores of every top-100 entry
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
    log(f"parity: subset {ok_subset}, top-100 order {ok_top}, top-100 oracle scores "
        f"{ok_top_oracle}; planted homologs in the top lists {found}/{len(planted)}")
    cpu = None
    if args.cpu_seconds > 0:
        sample = {"codes": b["codes"], "sw": b["sw"],
                  "a_off": np.tile(b["db_off"][:20000], 1), "a_len": b["db_len"][:20000],
                  "b_off": np.full(20000, b["q_off"][4]), "b_len": np.full(20000, b["q_len"][4],
                                                                           np.int32)}
        cpu = cpu_baseline(sample, args.cpu_seconds)
        cpu["sample"] = "query 5 (1024 residues) vs " + cpu["sample"]
    return {
        "metric": "GCUPS on 1 MI355X (bucketed vs plain, bit-exact); p50/p99 latency at fixed req/s",
        "value": round(gcups, 1), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int16x2 (int32 re-run on saturation)",
        "data": "synthetic (seeded generator, BLOSUM62 background residues, 0.1% planted homologs)",
        "config": {"workload": "one-vs-many search: 8 protein queries (128..4096) x 200,000 database "
                               "sequences lognormal(350, 0.5) clipped 30-5000, BLOSUM62 open 11 "
                               "extend 1, every score + top 100 per query",
                   "queries": nq, "database": ndb, "cells": cells,
                   "autotune": ({"rounds": tuned["rounds"], "step_cap_scale": tuned["scale"],
                                 "candidates": len(tuned["table"])} if tuned else None)},
        "parity": {"subset_2000_scores": ok_subset, "top100_order": ok_top,
                   "top100_oracle_scores": ok_top_oracle,
                   "planted_homologs_found": f"{found}/{len(planted)}"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_INT16_TOPS,
                     "unit": "TOP/s (int16)", "frac": round(achieved / PEAK_INT16_TOPS, 4),
                     "traffic": None, "kernel": "k_fast<26, false>", "kernel_ms": round(fast_avg, 3),
                     "ops_per_cell": OPS_PER_CELL},
        "cpu_baseline": cpu,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(pairs are independent: a batch shards by cumulative cells, one engine per GPU, "
              "with a gather of the results by pair index)",
              file=sys.stderr)
        return 2
    if args.config == "search":
        print(json.dumps(run_search(args)), flush=True)
        return 0
    import numpy as np
    import torch
    import ipuma_amd as I
    from ipuma_amd import workloads as W

    t = time.time()
    if args.config == "protein":
        batch = W.config1(n=args.pairs)
        plain_mode, plain_name = I.BATCH_PLAIN, "plain kernel padded to the batch maximum"
        workload = ("protein batch: 100k BLOSUM62 pairs, lognormal(300, 0.5) lengths clipped "
                    "50-1024, half related at 40% identity, gap open 11 extend 1, scores + end "
                    "positions")
        data = "synthetic (seeded generator, BLOSUM62 background residues)"
    else:
        batch = W.config2()
        plain_mode, plain_name = (I.BATCH_PLAIN_TILE,
                                  "plain kernel, input order, each wave padded to its longest pair")
        workload = ("skewed DNA batch: 1,000,000 pairs 150 x U[150,400] (B embeds a mutated A) "
                    "+ 5,000 pairs power-law(alpha 1.5) 1-16 kb at 85% identity, match 1 "
                    "mismatch -1 gap open 3 extend 1, scores + end positions")
        data = "synthetic (seeded generator, uniform ACGT)"
    log(f"generated {len(batch['a_len'])} pairs, {batch['cells']:.3e} cells in {time.time()-t:.1f}s")
    dev = torch.device("cuda:0")
    codes = torch.from_numpy(batch["codes"]).to(dev)
    a_off = torch.from_numpy(batch["a_off"]).to(dev)
    b_off = torch.from_numpy(batch["b_off"]).to(dev)
    a_len = torch.from_numpy(batch["a_len"]).to(dev)
    b_len = torch.from_numpy(batch["b_len"]).to(dev)
    n = len(batch["a_len"])
    scores = torch.empty(n, dtype=torch.int32, device=dev)
    ar = torch.empty(n, dtype=torch.int32, device=dev)
    br = torch.empty(n, dtype=torch.int32, device=dev)
    eng = I.SWAlgorithm(I.SWConfig(**batch["sw"]),
                        I.IPUAlgoConfig(tilesUsed=1 << 20, maxAB=1 << 16, maxBatches=1,
                                        forwardOnly=not args.with_starts))
    if args.rounds:
        eng.set_work_rounds(args.rounds)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def step():
        eng.compare_device(codes, codes.numel(), a_off, a_len, b_off, b_len, scores, ar, br,
                           stream=sptr)

    tuned = None
    if args.autotune == "on" and not args.rounds:
        # bucketing shape chosen on this batch, outside the timed region
        from ipuma_amd import autotune as AT
        key = AT.signature(batch["a_len"], batch["b_len"], batch["sw"].get("similarity", "dna"))
        tuned = AT.tune(eng, step, key=key, sync=torch.cuda.synchronize)
        tab = ", ".join("%dx%.2f: %.2f ms" % (r["rounds"], r["scale"], r["ms"])
                        for r in tuned["table"])
        log(f"autotune {key}: rounds {tuned['rounds']} scale {tuned['scale']} ({tab})")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region: K steps back to back on one stream
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / args.steps * 1e3
    cells = batch["cells"]
    gcups = cells / (elapsed / args.steps) / 1e9
    # packed-kernel duration: HIP events the engine records around its launch
    # on the batch's stream, read after each of K more steps
    eng.set_timing(True)
    fast_ms = []
    for _ in range(args.steps):
        step()
        st = eng.stats()  # synchronizes; reads the events of this step
        fast_ms.append(st["fast_kernel_ms"])
    eng.set_timing(False)
    fast_avg = float(np.mean(fast_ms))
    achieved = cells * OPS_PER_CELL / (fast_avg * 1e-3) / 1e12
    log(f"bucketed: {ms_step:.2f} ms/step, {gcups:.1f} GCUPS end-to-end, packed kernel "
        f"{fast_avg:.2f} ms ({cells / fast_avg / 1e6:.1f} GCUPS), stats {st}")

    # correctness guard on the measured output: a 2000-pair slice against the oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    k = min(2000, n)
    sw = batch["sw"]
    mat = (O.select_matrix(O.SIM_BLOSUM62) if sw["similarity"] == "blosum62"
           else O.select_matrix(O.SIM_NA, sw["matchValue"], sw["mismatchValue"],
                                sw["ambiguityValue"]))
    exp = O.sw_batch_codes(batch["codes"], batch["a_off"][:k], batch["a_len"][:k],
                           batch["b_off"][:k], batch["b_len"][:k], mat,
                           sw["gapInit"], sw["gapExtend"], not args.with_starts)
    got = (scores[:k].cpu().numpy(), ar[:k].cpu().numpy(), br[:k].cpu().numpy())
    parity = all(bool((g == e).all()) for g, e in zip(got, exp))
    log(f"parity on {k} pairs: {parity}")

    plain = None
    if args.plain_steps > 0:
        eng.set_batching(plain_mode)
        step()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(args.plain_steps):
            step()
        torch.cuda.synchronize()
        pe = (time.perf_counter() - tp) / args.plain_steps
        pst = eng.stats()
        plain = {"kind": plain_name, "ms_per_step": round(pe * 1e3, 2),
                 "gcups": round(cells / pe / 1e9, 1), "padded_cells": pst["padded_cells"]}
        log(f"plain padded batch: {plain}")
        eng.set_batching(I.BATCH_BUCKETED)

    cpu = None
    if args.cpu_seconds > 0:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(batch, args.cpu_seconds, gpu_scores=scores.cpu().numpy())
        log(f"cpu baseline: {cpu}")

    host_buffers = None
    if args.host_buffers and args.config == "protein":
        import ipuma_amd as I
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
            I._check(I.lib().ipuma_compare_local(eng._h, n, abuf, aoff.ctypes.data, bbuf,
                                                 boff.ctypes.data, hs.ctypes.data,
                                                 ha.ctypes.data, hb.ctypes.data))

        for _ in range(max(1, args.warmup)):
            host_call()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            host_call()
        hms = (time.perf_counter() - t0) / args.steps * 1e3
        host_buffers = {"ms_per_step": round(hms, 3), "gcups": round(cells / hms / 1e6, 1),
                        "scores_equal_device": bool(np.array_equal(hs, scores.cpu().numpy())),
                        "what": "ipuma_compare_local, host characters in, host results out "
                                "(PCIe-inclusive)"}
        log(f"host buffers: {host_buffers}")

    serving = None
    if args.serve_seconds > 0:
        from ipuma_amd import loadgen
        del eng  # the measured engine's buffers are not needed any more
        torch.cuda.empty_cache()
        log("serving: generating request pools ...")
        pools = loadgen.make_pools()
        serving = []
        for rate in [float(r) for r in args.serve_rates.split(",") if r]:
            for merge in ([True, False] if args.serve_unbatched else [True]):
                res = loadgen.run(pools, rate, args.serve_seconds, merge=merge)
                log(f"serving {res}")
                serving.append(res)

    tr = measured_traffic() if (args.config == "protein" and args.pairs == 100_000) else None
    out = {
        "metric": "GCUPS on 1 MI355X (bucketed vs plain, bit-exact); p50/p99 latency at fixed req/s",
        "value": round(gcups, 1),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16x2 (int32 re-run on saturation)",
        "data": data,
        "config": {"workload": workload,
                   "pairs": n, "cells": cells,
                   "starts": bool(args.with_starts), "work_rounds": args.rounds or "default",
                   "autotune": ({"rounds": tuned["rounds"], "step_cap_scale": tuned["scale"],
                                 "candidates": len(tuned["table"])} if tuned else None)},
        "bucketed_vs_plain": (round(gcups / plain["gcups"], 2) if plain else None),
        "plain": plain,
        "parity_2000_pairs": parity,
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_INT16_TOPS,
                     "unit": "TOP/s (int16)", "frac": round(achieved / PEAK_INT16_TOPS, 4),
                     "traffic": tr["traffic_bytes"] if tr else None,
                     "traffic_source": tr["source"] if tr else None,
                     "hbm_gbs": (round(tr["traffic_bytes"] / (fast_avg * 1e6), 1) if tr else None),
                     "kernel": f"k_fast<{26 if args.config == 'protein' else 18}, false>",
                     "kernel_ms": round(fast_avg, 3),
                     "ops_per_cell": OPS_PER_CELL},
        "cpu_baseline": cpu,
        "host_buffers": host_buffers,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
