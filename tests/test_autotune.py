"""The bucketing autotuner: cache keys, the JSON cache, and the choice logic
against a stand-in engine whose pass time is a known function of the knobs
(CPU); on the GPU, a tuned engine stays bit-exact with the oracle."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "ipuma-lib_amd"))
from ipuma_amd import autotune as AT  # noqa: E402


class FakeEngine:
    """Records the knobs; a pass 'takes' cost[(rounds, scale)] ms."""

    def __init__(self, cost):
        self.cost = cost
        self.rounds, self.scale = 1, 1.0
        self.passes = 0
        self.clock = 0.0

    def set_work_rounds(self, r):
        assert 1 <= r <= 64
        self.rounds = r

    def set_step_cap_scale(self, s):
        assert 0.25 <= s <= 4.0
        self.scale = s

    def synchronize(self):
        pass

    def run(self):
        self.passes += 1
        c = self.cost
        self.clock += c(self.rounds, self.scale) if callable(c) else c[(self.rounds, self.scale)]


def test_signature_stable_and_coarse():
    rng = np.random.default_rng(0)
    a = rng.integers(50, 1024, 10000)
    b = rng.integers(50, 1024, 10000)
    k1 = AT.signature(a, b, "blosum62", device="gfx950")
    assert k1 == AT.signature(a.copy(), b.copy(), "blosum62", device="gfx950")
    # a few pairs more or a little longer: same kind of batch, same key
    assert AT.signature(a[:-5], b[:-5], "blosum62", device="gfx950") == k1
    # another alphabet, device, or a batch 4x larger: another key
    assert AT.signature(a, b, "dna", device="gfx950") != k1
    assert AT.signature(a, b, "blosum62", device="gfx942") != k1
    assert AT.signature(np.tile(a, 4), np.tile(b, 4), "blosum62", device="gfx950") != k1
    assert AT.signature([], [], "dna", device="x").startswith("x|dna|")


def test_tune_picks_the_fastest_and_caches(tmp_path, monkeypatch):
    cost = {c: 10.0 + i for i, c in enumerate(AT.CANDIDATES)}
    best = AT.CANDIDATES[3]
    cost[best] = 5.0
    eng = FakeEngine(cost)
    # the timer reads the fake clock, so the choice depends only on `cost`
    monkeypatch.setattr(AT.time, "perf_counter", lambda: eng.clock / 1e3)
    cache = AT.TuneCache(str(tmp_path / "tune.json"))
    r = AT.tune(eng, eng.run, key="k", cache=cache, reps=3, refine_step=0)
    assert (r["rounds"], r["scale"]) == best and not r["cached"]
    assert (eng.rounds, eng.scale) == best  # applied
    assert eng.passes == len(AT.CANDIDATES) * 4  # one untimed + three timed each
    assert np.allclose([t["ms"] for t in r["table"]], [cost[c] for c in AT.CANDIDATES])
    # a new process: the file gives the same choice without timing
    eng2 = FakeEngine(cost)
    r2 = AT.tune(eng2, eng2.run, key="k", cache=AT.TuneCache(str(tmp_path / "tune.json")))
    assert r2["cached"] and (eng2.rounds, eng2.scale) == best and eng2.passes == 0
    # another key is timed again
    r3 = AT.tune(eng2, eng2.run, key="other", cache=cache, refine_step=0)
    assert not r3["cached"] and eng2.passes > 0
    data = json.load(open(tmp_path / "tune.json"))
    assert set(data) == {"k", "other"}


def test_tune_refines_down_to_the_drop(monkeypatch):
    """A sawtooth like the measured one (time rises with the scale within a
    tooth, drops where the groups fit one round fewer): the grid lands inside
    a tooth, the refinement walks down to the point just above the drop and
    stops at the first step that does not improve."""
    def cost(rounds, scale):  # one tooth over the grid, dropping at 0.98
        base = 3.0 + 2.0 * (scale - 0.98) if scale >= 0.98 - 1e-9 else 3.5 + 2.0 * (scale - 0.6)
        return base + (0.5 if rounds > 1 else 0.0)
    eng = FakeEngine(cost)
    monkeypatch.setattr(AT.time, "perf_counter", lambda: eng.clock / 1e3)
    r = AT.tune(eng, eng.run, reps=1)
    assert r["rounds"] == 1 and abs(r["scale"] - 0.98) < 1e-9, r["table"]
    assert (eng.rounds, eng.scale) == (1, r["scale"])
    refined = r["table"][len(AT.CANDIDATES):]
    # from the best grid point 1.0 down: 0.99, 0.98 improve, 0.97 (below the drop) does not
    assert [round(t["scale"], 2) for t in refined] == [0.99, 0.98, 0.97]


def test_cache_rejects_bad_entries(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text(json.dumps({"k": {"rounds": 0, "scale": 1.0, "ms": 1.0, "table": []}}))
    with pytest.raises(ValueError):
        AT.TuneCache(str(p)).get("k")
    p.write_text(json.dumps([1, 2]))
    with pytest.raises(ValueError):
        AT.TuneCache(str(p))


@pytest.mark.gpu
def test_tuned_engine_stays_exact(gpu):
    import torch
    import ipuma_amd as I
    from ipuma_amd import workloads as W
    sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
    import test_gpu_parity as T
    T.note("autotune: protein 3000 pairs, tuned, then parity")
    batch = W.config1(n=3000, seed=11)
    eng = T.engine_for(batch["sw"], True)
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(batch[k]).to(dev) for k in ("codes", "a_off", "b_off", "a_len", "b_len")}
    out = [torch.empty(3000, dtype=torch.int32, device=dev) for _ in range(3)]

    def run():
        eng.compare_device(t["codes"], t["codes"].numel(), t["a_off"], t["a_len"], t["b_off"],
                           t["b_len"], *out)

    r = AT.tune(eng, run, reps=1)
    assert any(r["rounds"] == c[0] and abs(r["scale"] - c[1]) <= 4 * AT.REFINE_STEP + 1e-9
               for c in AT.CANDIDATES)
    for rounds, scale in ((r["rounds"], r["scale"]), (1, 0.6), (2, 1.0)):
        AT.apply(eng, rounds, scale)
        got = T.device_run(eng, batch)
        T.assert_same(got, T.oracle_batch(batch, True), f"autotune {rounds}x{scale}")
