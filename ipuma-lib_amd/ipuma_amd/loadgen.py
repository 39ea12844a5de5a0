"""Open-loop load generator for the request server (the serving configuration).

Requests arrive as a Poisson process at a fixed rate; each carries a uniform
1..max_pairs comparisons, DNA (the plain DNA generator: |A| ~ U[100,150],
|B| ~ U[150,300], B embeds a mutated A) or protein (the protein batch
generator) with equal probability, drawn as windows of pre-generated pools
(the sequences are sent as characters, so every request pays its encoding and
upload).  DNA and protein requests go to two servers on the same GPU (the
scoring is per engine), each with half of the HBM budget.

Latency is end to end: from the request's scheduled arrival to its results in
host memory (encoding, upload, queueing, the device batch, copy-back).  The
unbatched baseline is the same load with one device batch per request.
"""
from __future__ import annotations

import queue
import threading
import time

import numpy as np

from . import Server, SWConfig, IPUAlgoConfig
from . import workloads as W

NA_LETTERS = np.frombuffer(W.NA_LETTERS.encode(), np.uint8)
AA_LETTERS = np.frombuffer(W.AA_LETTERS.encode(), np.uint8)


class Pool:
    """A batch's sequences as characters: concatenated A's and B's with n + 1 offsets."""

    def __init__(self, batch: dict, letters: np.ndarray):
        self.batch = batch  # the encoded batch the characters come from
        self.n = len(batch["a_len"])
        self.cells = batch["a_len"].astype(np.int64) * batch["b_len"].astype(np.int64)
        self.ccells = np.concatenate([[0], np.cumsum(self.cells)])
        for side in ("a", "b"):
            lens = batch[f"{side}_len"].astype(np.int64)
            offs = np.zeros(self.n + 1, np.int64)
            offs[1:] = np.cumsum(lens)
            src = np.repeat(batch[f"{side}_off"] - offs[:-1], lens) + np.arange(int(offs[-1]))
            chars = letters[batch["codes"][src]]
            # one byte per residue plus a terminating 0 (np.append with a Python
            # int would promote the array to int64)
            setattr(self, f"{side}_chars",
                    np.concatenate([chars, np.zeros(1, np.uint8)]).astype(np.uint8))
            setattr(self, f"{side}_offs", offs)

    def window(self, start: int, n: int):
        a, b = self.a_chars, self.b_chars
        return (n, a.ctypes.data, self.a_offs.ctypes.data + 8 * start, b.ctypes.data,
                self.b_offs.ctypes.data + 8 * start, int(self.ccells[start + n] - self.ccells[start]))


def make_pools(pairs: int = 40_000, seed: int = 11):
    dna = W.config0(seed=seed, n=pairs)
    prot = W.config1(seed=seed + 1, n=pairs)
    return {"dna": (Pool(dna, NA_LETTERS), dna["sw"]), "protein": (Pool(prot, AA_LETTERS), prot["sw"])}


def run(pools, rate: float, duration: float, max_pairs: int = 4096, merge: bool = True,
        hbm_budget: int = 64 << 30, max_wait_us: int = 2000, seed: int = 7,
        submitters: int = 8, waiters: int = 32) -> dict:
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * duration * 1.5) + 16))
    arrivals = arrivals[arrivals < duration]
    m = len(arrivals)
    kinds = rng.random(m) < 0.5                      # True: DNA
    sizes = rng.integers(1, max_pairs + 1, m)
    servers = {}
    for kind, (pool, sw) in pools.items():
        servers[kind] = Server(SWConfig(**sw), IPUAlgoConfig(forwardOnly=True),
                               hbm_budget_bytes=hbm_budget // 2, max_batch_pairs=1 << 16,
                               max_wait_us=max_wait_us, merge_requests=merge)
    starts = [int(rng.integers(0, pools["dna" if k else "protein"][0].n - s + 1))
              for k, s in zip(kinds, sizes)]
    lat = np.zeros(m)
    cells = np.zeros(m, np.int64)
    done_t = np.zeros(m)
    errors = []
    nxt = [0]
    lock = threading.Lock()
    tickets: "queue.Queue" = queue.Queue()
    t0 = time.perf_counter() + 0.05

    def submitter():
        while True:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= m:
                return
            kind = "dna" if kinds[i] else "protein"
            pool = pools[kind][0]
            n, ap, ao, bp, bo, c = pool.window(starts[i], int(sizes[i]))
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            try:
                t = servers[kind].submit_raw(n, ap, ao, bp, bo)
            except Exception as e:  # reported in the result
                errors.append(repr(e))
                continue
            cells[i] = c
            tickets.put((i, kind, t, n))

    def waiter():
        while True:
            item = tickets.get()
            if item is None:
                return
            i, kind, t, n = item
            try:
                servers[kind].wait(t, n)
            except Exception as e:
                errors.append(repr(e))
            done_t[i] = time.perf_counter()
            lat[i] = done_t[i] - (t0 + arrivals[i])

    sub = [threading.Thread(target=submitter) for _ in range(submitters)]
    wt = [threading.Thread(target=waiter) for _ in range(waiters)]
    for th in sub + wt:
        th.start()
    for th in sub:
        th.join()
    for _ in wt:
        tickets.put(None)
    for th in wt:
        th.join()
    st = {k: s.stats() for k, s in servers.items()}
    for s in servers.values():
        s.close()
    ok = lat > 0
    span = (done_t[ok].max() - t0) if ok.any() else 0.0
    batches = sum(v["batches"] for v in st.values())
    pairs = sum(v["pairs"] for v in st.values())
    return {
        "offered_req_s": rate, "duration_s": duration, "requests": int(m),
        "completed": int(ok.sum()), "errors": errors[:3],
        "achieved_req_s": round(float(ok.sum() / max(span, 1e-9)), 1),
        "p50_ms": round(float(np.percentile(lat[ok], 50) * 1e3), 3) if ok.any() else None,
        "p99_ms": round(float(np.percentile(lat[ok], 99) * 1e3), 3) if ok.any() else None,
        "gcups": round(float(cells[ok].sum() / max(span, 1e-9) / 1e9), 2),
        "offered_gcups": round(float(cells.sum() / duration / 1e9), 2),
        "batches": int(batches), "mean_batch_pairs": round(pairs / max(batches, 1), 1),
        "merged": bool(merge),
    }
