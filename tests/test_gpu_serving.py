"""The continuous-batching loop: requests submitted while others run share the persistent launches,
and every result equals the same request run alone (tolerance 0)."""
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
