"""The autotuner (war_autotune / war_autotune_batch): every candidate schedule must reproduce the
oracle's counts on the tuning slice exactly (the caller passes them), the chosen schedule becomes the
default of its workload, and a run with it equals the oracle. Also the default reference of the
search (the plain kernel's counts) and a tuned mixed-n batch."""
import math

import pytest

from war_amd import WarError

pytestmark = pytest.mark.gpu

KEYS = ("cyc", "done", "maxfin", "turns", "hist", "lenhist")
GRID = {"stage_kind": {"flow", "refill", "lockstep"}, "ranks_per_refill": {64, 128, 256, 512, 1024, 2048, 4096, 8192},
        "waves_per_cu": {1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 20, 24, 28, 32}}


@pytest.mark.parametrize("n,mode,lo,hi", [
    (11, "full", 0, math.factorial(11)),
    (12, "canon", 200_000_000, 200_000_000 + (1 << 26)),
])
def test_autotune_against_oracle(eng, orc, n, mode, lo, hi):
    print("case", n, mode, lo, hi, flush=True)
    want = orc.run(n, mode, lo, hi, hbins=128, lbins=256)
    t = eng.autotune_batch([(n, mode, lo, hi, {"hbins": 128, "lbins": 256})], budget_ms=60000, expect=[want])
    for k, vals in GRID.items():
        assert t[k] in vals, (k, t[k])
    assert eng.tuning_for(n, mode) == t          # kept as the default of (n, mode, rule 0, Brent)
    got = eng.run_range(n, mode, lo, hi, hbins=128, lbins=256)
    for k in KEYS:
        assert got[k] == want[k], (n, mode, k)
    # a wrong expectation is caught: the search fails instead of keeping a schedule
    bad = dict(want)
    bad["turns"] += 1
    with pytest.raises(WarError, match="exactness"):
        eng.autotune_batch([(n, mode, lo, hi, {"hbins": 128, "lbins": 256})], budget_ms=1000, expect=[bad])


def test_autotune_default_reference_is_plain(eng, orc):
    """war_autotune without expectations checks every candidate against the plain kernel; the
    n = 10 full range fits the slice whole, so the kept schedule must give the oracle's counts."""
    t = eng.autotune(10, "full", 0, budget_ms=20000)
    assert eng.tuning_for(10, "full") == t
    got = eng.run_range(10, "full", hbins=64)
    want = orc.run(10, "full", hbins=64)
    assert (got["cyc"], got["turns"], got["hist"]) == (want["cyc"], want["turns"], want["hist"])


def test_autotune_mixed_batch(eng, orc):
    """A mixed-n batch (n = 5..13, full and canonical, distinct on some) tuned as a whole against the
    oracle's per-request counts; the batch then runs with that schedule as one graph launch."""
    from test_gpu_batch import mixed_queries
    qs = mixed_queries(seed=20261203, count=48, top=5.5)
    qs = [(n, m, lo, hi, kw) for n, m, lo, hi, kw in qs if n <= 13]
    want = [orc.run(n, m, lo, hi, **kw) for n, m, lo, hi, kw in qs]
    t = eng.autotune_batch(qs, budget_ms=30000, expect=want)
    tuned = [(n, m, lo, hi, dict(kw, tuning=t)) for n, m, lo, hi, kw in qs]
    got, launches = eng.run_batch(tuned)
    assert launches == (1 if t["stage_kind"] == "flow" else launches)
    for q, g, w in zip(qs, got, want):
        for k in ("cyc", "done", "maxfin", "turns", "hist", "lenhist", "oncycle", "distinct_minkey", "dhist"):
            assert g[k] == w[k], (q, k)
