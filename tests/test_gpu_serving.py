"""The continuous-batching loop: requests submitted while others run share the persistent launches,
every served result equals the oracle's for the same range, rule and flags, and the same request run
alone (tolerance 0)."""
import math
import random

import pytest

from war_amd import WarError

pytestmark = pytest.mark.gpu

KEYS = ("cyc", "done", "maxfin", "turns", "hist", "lenhist", "oncycle", "distinct_minkey", "distinct_hashset",
        "dhist", "shape_nonstd", "shape_unchk")


def queries(seed, count):
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        n = rng.randint(5, 17)
        mode = rng.choice(["full", "canon", "full", "half"]) if n % 2 == 0 else rng.choice(["full", "canon"])
        if n >= 17 and mode == "half":
            mode = "full"
        import math
        total = math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)
        ln = min(total, int(10 ** rng.uniform(2, 7)))
        lo = rng.randrange(0, total - ln + 1)
        kw = {"hbins": 128, "lbins": 256}
        if mode == "canon" and n <= 13 and rng.random() < 0.4:
            kw.update(distinct=True, dbins=256, shape_blk=n & -n)
        out.append((n, mode, lo, lo + ln, rng.choice([0, 0, 1]), kw))
    return out


def test_alone_equals_in_the_mix(eng):
    qs = queries(7, 40)
    alone = [eng.run_range(n, mode, lo, hi, rule=rule, **kw) for n, mode, lo, hi, rule, kw in qs]
    eng.serve_start(max_slots=16, step_positions=1 << 22)
    try:
        ids = [eng.submit(n, mode, lo, hi, rule=rule, **kw) for n, mode, lo, hi, rule, kw in qs]
        mixed = [eng.wait(i) for i in ids]
    finally:
        eng.serve_stop()
    for q, a, m in zip(qs, alone, mixed):
        for k in KEYS:
            assert a[k] == m[k], (q, k, a[k], m[k])
        assert m["latency_ms"] > 0


def test_poll_and_late_arrivals(eng, orc):
    eng.serve_start(max_slots=4, step_positions=1 << 20)
    try:
        first = [eng.submit(12, "canon", i * 5_000_000, (i + 1) * 5_000_000, hbins=64) for i in range(6)]
        late = eng.submit(9, "full", hbins=64)
        res = [eng.wait(i) for i in first]
        r9 = eng.wait(late)
    finally:
        eng.serve_stop()
    want = orc.run(9, "full", hbins=64)
    assert (r9["cyc"], r9["maxfin"], r9["hist"]) == (want["cyc"], want["maxfin"], want["hist"])
    whole = orc.run(12, "canon", 0, 30_000_000, hbins=64)
    assert sum(r["cyc"] for r in res) == whole["cyc"]
    assert sum(r["turns"] for r in res) == whole["turns"]


def test_serving_refuses_pair_detector(eng):
    eng.serve_start(max_slots=4, step_positions=1 << 20)
    try:
        with pytest.raises(WarError, match="pair-run"):
            eng.submit(14, "canon", 0, 1000, detector="pair")
    finally:
        eng.serve_stop()


@pytest.mark.parametrize("stage_kind", ["flow", "lockstep"])
def test_mixed_detectors_on_both_pipelines(eng, stage_kind):
    """The mixed batches of the serving loop on the flow kernel (the default) and on the staged
    pipeline, with the alternation detector on the odd-n rule-0 queries: each result equals the same
    request alone (single-request flow batches)."""
    qs = []
    for n, mode, lo, hi, rule, kw in queries(11, 36):
        kw = dict(kw)
        if n % 2 == 1 and rule == 0 and len(qs) % 2 == 0:
            kw["detector"] = "alt"
        qs.append((n, mode, lo, hi, rule, kw))
    alone = [eng.run_range(n, mode, lo, hi, rule=rule, **kw) for n, mode, lo, hi, rule, kw in qs]
    eng.set_tuning(stage_kind=stage_kind)
    eng.serve_start(max_slots=12, step_positions=1 << 21)
    try:
        ids = [eng.submit(n, mode, lo, hi, rule=rule, **kw) for n, mode, lo, hi, rule, kw in qs]
        mixed = [eng.wait(i) for i in ids]
    finally:
        eng.serve_stop()
        eng.set_tuning()
    for q, a, m in zip(qs, alone, mixed):
        for k in KEYS:
            assert a[k] == m[k], (stage_kind, q, k, a[k], m[k])


def oracle_queries(seed=5):
    """Every n of 7..17 (and 18), each mode that applies, both rules, distinct + dbins + shape on
    canonical n <= 13, dbins on canonical n = 16, the alternation detector on odd n, the held-card
    (n = 17, default) and 5-bit (n = 17 forced, n = 18) layouts, a few step caps; lengths
    log-uniform in [10^2, 10^7] so the oracle checks them all in seconds."""
    rng = random.Random(seed)
    out = []
    for n in range(7, 19):
        modes = ["full", "canon"] + (["half"] if n % 2 == 0 and n < 17 else [])
        for mode in modes:
            for rule in (0, 1):
                total = math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)
                ln = min(total, int(10 ** rng.uniform(2, 7)))
                lo = rng.randrange(0, total - ln + 1)
                kw = {"hbins": 128, "lbins": 256, "rule": rule}
                if mode == "canon" and n <= 13 and rule == 0:
                    kw.update(distinct=True, dbins=256, shape_blk=n & -n)
                if mode == "canon" and n == 16:
                    kw.update(dbins=64)
                if n % 2 == 1 and rule == 0 and mode != "canon":
                    kw["detector"] = "alt"
                if n == 17 and mode == "full" and rule == 1:
                    kw["tuning"] = {"pile_layout": 5}
                if rng.random() < 0.15 and not kw.get("distinct"):
                    kw["step_cap"] = 32 * rng.randint(1, 4)
                out.append((n, mode, lo, lo + ln, kw))
    return out


ORC_KEYS = ("cyc", "done", "maxfin", "turns", "undetermined", "hist", "lenhist", "oncycle", "distinct_minkey",
            "dhist", "shape_nonstd", "shape_unchk")


def test_served_results_equal_oracle(eng, orc):
    qs = oracle_queries()
    eng.serve_start(max_slots=12, step_positions=1 << 22)
    try:
        ids = [eng.submit(n, mode, lo, hi, **kw) for n, mode, lo, hi, kw in qs]
        served = [eng.wait(i) for i in ids]
    finally:
        eng.serve_stop()
    for (n, mode, lo, hi, kw), g in zip(qs, served):
        okw = {k: v for k, v in kw.items() if k != "tuning"}
        want = orc.run(n, mode, lo, hi, **okw)
        for k in ORC_KEYS:
            assert g[k] == want[k], (n, mode, lo, hi, kw, k, g[k], want[k])
        if kw.get("distinct"):
            assert g["distinct_hashset"] == want["distinct_keys"], (n, mode, lo, hi)
