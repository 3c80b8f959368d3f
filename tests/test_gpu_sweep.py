"""A seeded sweep over the request space against the oracle: n (4..17, so 4-bit, 5-bit and held-card
piles), mode (full, half, canonical), rule, detector (Brent, the alternation and pair-run checkpoint
detectors, the streak test), kernel (persistent or plain), schedule (flow kernel or staged pipeline,
tile, waves per CU, first stage boundary), histograms, cycle walks (distinct, weights, shape, with
and without the cycle-state table) and step caps, drawn with numpy PCG64 from a fixed seed; each range
small enough for the oracle. Every output equals the oracle's (tolerance 0)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DRAWS = 600
KEYS = ("cyc", "done", "maxfin", "turns", "undetermined", "hist", "lenhist")
WALK_KEYS = ("oncycle", "distinct_minkey", "dhist", "shape_nonstd", "shape_unchk")


def total(n, mode):
    return math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)


def draw(rng):
    """One valid request: (n, mode, lo, hi, engine keywords, oracle keywords)."""
    n = int(rng.integers(4, 18))
    modes = ["full", "canon"] + (["half"] if n % 2 == 0 else [])
    mode = modes[int(rng.integers(len(modes)))]
    rule = int(rng.integers(2))
    t = total(n, mode)
    length = min(t, int(10 ** rng.uniform(1, 6)))
    lo = int(rng.integers(0, t - length + 1))
    kernel = "plain" if rng.random() < 0.2 else "persistent"
    det = "brent"
    if n % 2 == 1 and rule == 0:
        u = rng.random()
        if u < 0.1:
            det, kernel = "streak", "plain"
        elif u < 0.5 and kernel == "persistent":
            det = "alt"
    elif n % 4 == 2 and rule == 0 and kernel == "persistent" and rng.random() < 0.3:
        det = "pair"
    kw = {"rule": rule, "kernel": kernel, "detector": det}
    okw = {"rule": rule, "detector": det}
    if det != "streak":
        kw["hbins"] = okw["hbins"] = int(rng.choice([0, 64, 256]))
    kw["lbins"] = okw["lbins"] = int(rng.choice([0, 128, 512]))
    walks = mode == "canon" and n <= 15 and det != "streak" and rng.random() < 0.5
    if walks:
        if rng.random() < 0.6:
            kw["distinct"] = okw["distinct"] = True
        if rng.random() < 0.6:
            kw["dbins"] = okw["dbins"] = 64
        if rng.random() < 0.6 or not (kw.get("distinct") or kw.get("dbins")):
            kw["shape_blk"] = okw["shape_blk"] = n & -n
        if rule == 0 and n <= 14:
            kw["cycle_table"] = [None, True, False][int(rng.integers(3))]
    flow = True
    if kernel == "persistent" and rng.random() < 0.5:
        tun = {"ranks_per_refill": 64 * int(rng.integers(1, 129)),
               "waves_per_cu": int(rng.choice([1, 2, 4, 7, 8, 16, 28, 32])),
               "brent_cap": int(rng.choice([7, 15, 31, 63, 127]))}
        if det != "pair":
            tun["stage_kind"] = str(rng.choice(["flow", "flow", "refill", "lockstep"]))
            flow = tun["stage_kind"] == "flow"
        kw["tuning"] = tun
    if not walks and det != "streak" and rng.random() < 0.15:
        if kernel == "plain":
            kw["step_cap"] = okw["step_cap"] = int(rng.integers(8, 200))
        elif flow:
            kw["step_cap"] = okw["step_cap"] = 32 * int(rng.integers(1, 6))
    return n, mode, lo, lo + length, kw, okw


def test_sweep_against_oracle(eng, orc):
    rng = np.random.Generator(np.random.PCG64(20261204))
    for i in range(DRAWS):
        n, mode, lo, hi, kw, okw = draw(rng)
        print("draw", i, n, mode, lo, hi, kw, flush=True)
        got = eng.run_range(n, mode, lo, hi, **kw)
        want = orc.run(n, mode, lo, hi, **okw)
        keys = KEYS + (WALK_KEYS if mode == "canon" else ())
        if okw["detector"] == "streak":
            keys = ("cyc", "done", "maxfin", "turns", "undetermined", "lenhist")
        for k in keys:
            assert got[k] == want[k], (i, n, mode, lo, hi, kw, k, got[k], want[k])
        if okw.get("distinct"):
            assert got["distinct_hashset"] == want["distinct_keys"], (i, n, mode, lo, hi, kw)
