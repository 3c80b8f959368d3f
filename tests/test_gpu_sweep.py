"""Seeded sweeps over the request space against the oracle, one per entry point (war_run_range,
war_run_batch, the serving loop, war_run_range_noalloc): n (2..20, so 4-bit, 5-bit and held-card
piles), mode (full, half, canonical), rule, detector (Brent, the alternation and pair-run
checkpoint detectors, the streak test), kernel (persistent or plain), schedule (flow kernel or
staged pipeline, tile, waves per CU, first stage boundary, pile layout), histograms, cycle walks
(distinct, weights, shape, with and without the cycle-state table) and step caps, as each entry
point accepts them, drawn with numpy PCG64 from fixed seeds; each range small enough for the oracle.
Every output equals the oracle's (tolerance 0)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("cyc", "done", "maxfin", "turns", "undetermined", "hist", "lenhist")
WALK_KEYS = ("oncycle", "distinct_minkey", "dhist", "shape_nonstd", "shape_unchk")
STREAK_KEYS = ("cyc", "done", "maxfin", "turns", "undetermined", "lenhist")


def total(n, mode):
    return math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)


def draw(rng, target="single"):
    """One valid request for the target entry point: (n, mode, lo, hi, engine keywords, oracle
    keywords). single: war_run_range; batch: war_run_batch (persistent kernel, Brent or the
    alternation detector); serve: the serving loop (persistent kernel, no pair-run detector);
    noalloc: war_run_range_noalloc (no cycle walks, the flow schedule on the persistent kernel)."""
    n = int(rng.integers(2, 21))
    modes = ["full", "canon"] + (["half"] if n % 2 == 0 else [])
    mode = modes[int(rng.integers(len(modes)))]
    rule = int(rng.integers(2))
    t = total(n, mode)
    length = min(t, int(10 ** rng.uniform(1, 6 if n <= 17 else 5.5)))
    lo = int(rng.integers(0, t - length + 1))
    plain_ok = target in ("single", "noalloc")
    kernel = "plain" if plain_ok and rng.random() < 0.2 else "persistent"
    det = "brent"
    if n % 2 == 1 and rule == 0:
        u = rng.random()
        if u < 0.1 and plain_ok:
            det, kernel = "streak", "plain"
        elif u < 0.5:
            det = "alt"
    elif n % 4 == 2 and rule == 0 and kernel == "persistent" and target in ("single", "noalloc") and rng.random() < 0.3:
        det = "pair"
    kw = {"rule": rule, "kernel": kernel, "detector": det}
    okw = {"rule": rule, "detector": det}
    if det != "streak":
        kw["hbins"] = okw["hbins"] = int(rng.choice([0, 64, 256]))
    kw["lbins"] = okw["lbins"] = int(rng.choice([0, 128, 512]))
    walks = target != "noalloc" and mode == "canon" and det != "streak" and kernel == "persistent" and \
        (n <= 15 or n == 16) and rng.random() < 0.5
    if walks:
        if n <= 15 and rng.random() < 0.6:
            kw["distinct"] = okw["distinct"] = True
        if rng.random() < 0.6:
            kw["dbins"] = okw["dbins"] = 64
        if n <= 15 and (rng.random() < 0.6 or not (kw.get("distinct") or kw.get("dbins"))):
            kw["shape_blk"] = okw["shape_blk"] = n & -n
        if rule == 0 and n <= 14:
            kw["cycle_table"] = [None, True, False][int(rng.integers(3))]
    flow = True
    tun = {}
    if kernel == "persistent" and target in ("single", "noalloc") and rng.random() < 0.5:
        tun = {"ranks_per_refill": 64 * int(rng.integers(1, 129)),
               "waves_per_cu": int(rng.choice([1, 2, 4, 7, 8, 16, 28, 32])),
               "brent_cap": int(rng.choice([7, 15, 31, 63, 127]))}
        if det != "pair" and target == "single":
            tun["stage_kind"] = str(rng.choice(["flow", "flow", "refill", "lockstep"]))
            flow = tun["stage_kind"] == "flow"
    if n == 17 and kernel == "persistent" and not walks and det in ("brent", "alt") and rng.random() < 0.4:
        tun["pile_layout"] = int(rng.choice([5, 6]))
    if tun:
        kw["tuning"] = tun
    if not walks and det != "streak" and rng.random() < 0.15:
        if kernel == "plain":
            kw["step_cap"] = okw["step_cap"] = int(rng.integers(8, 200))
        elif flow:
            kw["step_cap"] = okw["step_cap"] = 32 * int(rng.integers(1, 6))
    return n, mode, lo, lo + length, kw, okw


def check(i, got, want, n, mode, lo, hi, kw, okw):
    keys = KEYS + (WALK_KEYS if mode == "canon" else ())
    if okw["detector"] == "streak":
        keys = STREAK_KEYS
    for k in keys:
        assert got[k] == want[k], (i, n, mode, lo, hi, kw, k, got[k], want[k])
    if okw.get("distinct"):
        assert got["distinct_hashset"] == want["distinct_keys"], (i, n, mode, lo, hi, kw)


def test_sweep_single(eng, orc):
    rng = np.random.Generator(np.random.PCG64(20261204))
    for i in range(1500):
        n, mode, lo, hi, kw, okw = draw(rng)
        print("draw", i, n, mode, lo, hi, kw, flush=True)
        got = eng.run_range(n, mode, lo, hi, **kw)
        check(i, got, orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)


def test_sweep_batches(eng, orc):
    """Batches of 1 to 12 requests of mixed n, mode, rule, detector and walks in shared launches."""
    rng = np.random.Generator(np.random.PCG64(20261205))
    for b in range(150):
        qs = [draw(rng, "batch") for _ in range(int(rng.integers(1, 13)))]
        print("batch", b, [(q[0], q[1], q[2], q[3], q[4]) for q in qs], flush=True)
        got, launches = eng.run_batch([(n, mode, lo, hi, kw) for n, mode, lo, hi, kw, _ in qs])
        assert launches >= 1
        for i, ((n, mode, lo, hi, kw, okw), g) in enumerate(zip(qs, got)):
            check((b, i), g, orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)


@pytest.mark.parametrize("step_positions", [1 << 16, 1 << 22])
def test_sweep_served(eng, orc, step_positions):
    """Requests queued into the serving loop at once, steps small and large."""
    rng = np.random.Generator(np.random.PCG64(20261206 + step_positions))
    qs = [draw(rng, "serve") for _ in range(150)]
    eng.serve_start(max_slots=16, step_positions=step_positions)
    try:
        ids = []
        for i, (n, mode, lo, hi, kw, _) in enumerate(qs):
            print("serve", i, n, mode, lo, hi, kw, flush=True)
            ids.append(eng.submit(n, mode, lo, hi, **kw))
        served = [eng.wait(r) for r in ids]
    finally:
        eng.serve_stop()
    for i, ((n, mode, lo, hi, kw, okw), g) in enumerate(zip(qs, served)):
        check(i, g, orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)


def test_sweep_noalloc(eng, orc):
    rng = np.random.Generator(np.random.PCG64(20261207))
    for i in range(400):
        n, mode, lo, hi, kw, okw = draw(rng, "noalloc")
        print("noalloc", i, n, mode, lo, hi, kw, flush=True)
        got = eng.run_range_noalloc(n, mode, lo, hi, **kw)
        check(i, got, orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)


def test_sweep_large_cross_kernel(eng):
    """Ranges of 10^7.5 to 10^10 positions (10^9.3 at n = 17, 10^8.7 for the 5-bit n = 18..20), longer than the oracle plays
    in a test: the persistent kernel at a random schedule, tile, waves per CU and layout against the
    plain kernel (one game per lane, no pools or tortoise schedule) with the same detector (Brent's
    counts for the pair-run detector, whose turns are its own); cycle walks on canonical n <= 14 (the
    state table on by default from 2^28 positions). Every output equal."""
    rng = np.random.Generator(np.random.PCG64(20261208))
    for i in range(40):
        n = int(rng.integers(13, 21))
        modes = ["full", "canon"] + (["half"] if n % 2 == 0 else [])
        mode = modes[int(rng.integers(len(modes)))]
        rule = int(rng.random() < 0.3)
        top = 10.0 if n <= 16 else (9.3 if n == 17 else 8.7)
        length = min(total(n, mode), int(10 ** rng.uniform(7.5, top)))
        lo = int(rng.integers(0, total(n, mode) - length + 1))
        det = "brent"
        if rule == 0 and n % 2 == 1 and rng.random() < 0.6:
            det = "alt"
        elif rule == 0 and n % 4 == 2 and rng.random() < 0.5:
            det = "pair"
        kw = {"rule": rule, "hbins": 256, "lbins": 512}
        if mode == "canon" and n <= 14 and rng.random() < 0.4:
            kw.update(distinct=True, dbins=64, shape_blk=n & -n)
        tun = {"ranks_per_refill": 64 * int(rng.integers(1, 129)), "waves_per_cu": int(rng.choice([2, 4, 8, 16, 28])),
               "brent_cap": int(rng.choice([15, 31, 63]))}
        if det != "pair":
            tun["stage_kind"] = str(rng.choice(["flow", "flow", "refill", "lockstep"]))
        if n == 17 and not kw.get("distinct"):
            tun["pile_layout"] = int(rng.choice([5, 6]))
        print("large", i, n, mode, lo, lo + length, det, kw, tun, flush=True)
        got = eng.run_range(n, mode, lo, lo + length, kernel="persistent", detector=det, tuning=tun, **kw)
        ref = eng.run_range(n, mode, lo, lo + length, kernel="plain", detector="brent" if det == "pair" else det, **kw)
        keys = ("cyc", "done", "maxfin", "undetermined", "hist", "lenhist") + (() if det == "pair" else ("turns",))
        if mode == "canon":
            keys += WALK_KEYS + ("distinct_hashset",)
        for k in keys:
            assert got[k] == ref[k], (i, n, mode, lo, lo + length, det, kw, tun, k, got[k], ref[k])
        print("  %.1f ms persistent, %.1f ms plain" % (got["device_ms"], ref["device_ms"]), flush=True)
