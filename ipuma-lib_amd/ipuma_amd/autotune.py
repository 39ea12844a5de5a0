"""Autotuner of the device bucketing's group shapes, with a result cache.

The packed kernel's groups come from the device length bucketing: each
comparison gets the shallowest pipeline (lanes per comparison) whose group
fits a cap on its steps, cells / (16 * resident half-lanes * rounds) times the
step cap scale.  The cap trades row padding (a high cap keeps pipelines
shallow) against the length of the longest-first queue's tail (a low cap makes
more, shorter groups).  Where the balance lies depends on the batch's length
distribution, so it is measured rather than modelled:

* ``tune`` times a grid of (rounds, step cap scale) on the caller's own
  batch (one untimed pass, then the median of ``reps`` timed passes each),
  then refines below the fastest point in steps of 0.01, applies the fastest
  and returns the table.  The time is a sawtooth in the scale: the queue runs
  in rounds of about one group per wave, a larger cap makes every group
  longer until the groups fit one round fewer, where the time drops (on the
  protein batch 2.95 ms at scale 0.77 and 0.98, 3.10-3.33 ms just below
  them).  The fastest point of a tooth lies just above its drop, so the
  refinement walks down from the best grid point while the time improves;
* ``TuneCache`` keeps the choice in a JSON file keyed by the device, the
  alphabet and a coarse signature of the batch (pair count, cells, length
  quantiles), so a later batch of the same kind reuses it without timing.

The reference has no counterpart: its tile fill (src/partition.cpp:197-462)
runs with the capacities the caller configures.
"""
from __future__ import annotations

import json
import math
import os
import time
from typing import Callable, Iterable, Sequence

import numpy as np

# (rounds, step cap scale): the default first, then a grid of caps (the cap
# depends on scale / rounds only, so one round covers them)
CANDIDATES: tuple = ((1, 1.0),) + tuple((1, round(0.6 + 0.05 * i, 2)) for i in range(13) if i != 8)
REFINE_STEP = 0.01

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "ipuma_amd", "autotune.json")


def _lbin(x: float, per_octave: int = 4) -> int:
    return int(round(per_octave * math.log2(max(float(x), 1.0))))


def signature(a_len: Sequence[int], b_len: Sequence[int], alphabet: str,
              device: str | None = None) -> str:
    """Cache key of a batch: device, alphabet, and pair count, total cells and
    the 10/50/90% quantiles of |A| and |B| on a log scale (four bins per
    octave; pair count and cells two per octave)."""
    a = np.asarray(a_len, dtype=np.int64)
    b = np.asarray(b_len, dtype=np.int64)
    if device is None:
        device = device_name()
    cells = float((a * b).sum()) if a.size else 0.0
    qa = [_lbin(v) for v in np.percentile(a, (10, 50, 90))] if a.size else [0, 0, 0]
    qb = [_lbin(v) for v in np.percentile(b, (10, 50, 90))] if b.size else [0, 0, 0]
    return "|".join([device, alphabet, f"n{_lbin(a.size, 2)}", f"c{_lbin(cells, 2)}",
                     "a" + ",".join(map(str, qa)), "b" + ",".join(map(str, qb))])


def device_name() -> str:
    """The GPU's architecture name (gfx target), or "unknown"."""
    try:
        import torch
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(0)
            return str(getattr(p, "gcnArchName", p.name)).split(":")[0]
    except Exception:  # pragma: no cover - no torch / no device
        pass
    return "unknown"


class TuneCache:
    """JSON file of tuned settings: {key: {"rounds", "scale", "ms", "table"}}."""

    def __init__(self, path: str | None = None):
        self.path = path or os.environ.get("IPUMA_AUTOTUNE_CACHE", DEFAULT_CACHE)
        self.entries: dict = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                data = json.load(f)
            if not isinstance(data, dict):
                raise ValueError(f"{self.path}: not an autotune cache")
            self.entries = data

    def get(self, key: str):
        e = self.entries.get(key)
        if e is None:
            return None
        r, s = int(e["rounds"]), float(e["scale"])
        if not (1 <= r <= 64 and 0.25 <= s <= 4.0):
            raise ValueError(f"{self.path}: bad entry for {key!r}: {e}")
        return e

    def put(self, key: str, entry: dict):
        self.entries[key] = entry
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.entries, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def apply(eng, rounds: int, scale: float):
    eng.set_work_rounds(int(rounds))
    eng.set_step_cap_scale(float(scale))


def tune(eng, run: Callable[[], None], key: str | None = None, cache: TuneCache | None = None,
         candidates: Iterable = CANDIDATES, reps: int = 3,
         sync: Callable[[], None] | None = None, refine_step: float = REFINE_STEP,
         refine_max: int = 4) -> dict:
    """Choose (rounds, step cap scale) for ``eng`` by timing ``run`` (one device
    batch through ``eng``) under each candidate, then at up to ``refine_max``
    scales below the fastest, ``refine_step`` apart, while the time improves
    (``refine_step`` 0: no refinement); apply the fastest and return
    {"rounds", "scale", "ms", "table", "cached"}.  With a cache and a key
    already in it, the cached choice is applied without timing."""
    if cache is not None and key is not None:
        hit = cache.get(key)
        if hit is not None:
            apply(eng, hit["rounds"], hit["scale"])
            return dict(hit, cached=True)
    sync = sync or eng.synchronize
    table = []

    def measure(rounds, scale):
        apply(eng, rounds, scale)
        run()  # untimed: first pass of a shape
        sync()
        ms = []
        for _ in range(max(1, reps)):
            t0 = time.perf_counter()
            run()
            sync()
            ms.append((time.perf_counter() - t0) * 1e3)
        table.append({"rounds": int(rounds), "scale": float(scale),
                      "ms": float(np.median(ms))})
        return table[-1]

    for rounds, scale in candidates:
        measure(rounds, scale)
    best = min(table, key=lambda r: r["ms"])
    if refine_step > 0:
        tried = {(r["rounds"], round(r["scale"], 4)) for r in table}
        cur = best
        for i in range(1, refine_max + 1):
            s = round(best["scale"] - i * refine_step, 4)
            if s < 0.25 or (best["rounds"], s) in tried:
                break
            r = measure(best["rounds"], s)
            if r["ms"] >= cur["ms"]:
                break
            cur = r
        best = min(table, key=lambda r: r["ms"])
    apply(eng, best["rounds"], best["scale"])
    entry = {"rounds": best["rounds"], "scale": best["scale"], "ms": best["ms"], "table": table}
    if cache is not None and key is not None:
        cache.put(key, entry)
    return dict(entry, cached=False)
