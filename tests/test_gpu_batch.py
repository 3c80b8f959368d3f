"""war_run_batch: requests of mixed n, mode and rule in shared persistent launches, every kernel
launch of the call a branch of one captured graph, launched once. Each result must equal the
oracle's for that request (tolerance 0).

The mixed-n workload: seeded numpy PCG64 (seed 20261203), 240 requests (the engine has 248 request
slots), n uniform over 5..15, mode full or canonical, with distinct + dhist for a third of the
canonical requests of n <= 13, length floor(10^U(3, 6.5)) clipped to the total, lo uniform over
[0, total - length]. The lengths stop at 10^6.5 rather than 10^9 so the CPU oracle checks all 240
requests in seconds; the launch packing is tested at full batch size against single requests."""
import math

import numpy as np
import pytest

from war_amd import WarError

pytestmark = pytest.mark.gpu

KEYS = ("cyc", "done", "maxfin", "turns", "hist", "lenhist", "oncycle", "distinct_minkey", "dhist")


def total(n, mode):
    return math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)


def mixed_queries(seed=20261203, count=240, top=6.5):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(rng.integers(5, 16))
        mode = "canon" if rng.random() < 0.5 else "full"
        t = total(n, mode)
        ln = min(t, int(10 ** rng.uniform(3, top)))
        lo = int(rng.integers(0, t - ln + 1))
        kw = {"hbins": 128, "lbins": 128}
        if mode == "canon" and n <= 13 and rng.random() < 1 / 3:
            kw.update(distinct=True, dbins=256)
        out.append((n, mode, lo, lo + ln, kw))
    return out


def test_mixed_n_one_graph_launch_equals_oracle(eng, orc):
    qs = mixed_queries()
    assert sum(hi - lo for _, _, lo, hi, _ in qs) < 1 << 28       # fits one batch with cycle walks
    got, launches = eng.run_batch(qs)
    # three kernel launches (without cycle walks: deal blocks, the rest; with walks), one graph
    assert launches == 1
    for (n, mode, lo, hi, kw), g in zip(qs, got):
        want = orc.run(n, mode, lo, hi, **kw)
        for k in KEYS:
            if k in want:
                assert g[k] == want[k], (n, mode, lo, hi, k, g[k], want[k])
        assert g["done"] == hi - lo


def test_groups_and_packing_equal_single_requests(eng):
    """Rules, detectors, layouts and cycle walks form separate groups (one launch each when they
    fit, deal blocks of 24 orders apart from the rest); a request larger than a batch is split
    over launches (here 5e9 deals of n = 14: two launches of the flow kernel's 2^32), and cycle
    walks bound a launch at 2^28 positions."""
    qs = [
        (14, "full", 10**9, 10**9 + 5 * 10**9, {"hbins": 256, "lbins": 256}),
        (13, "full", 7, 7 + 3_000_001, {"hbins": 64, "detector": "alt"}),
        (11, "canon", 12345, 12345 + 2_000_000, {"hbins": 64, "rule": 1}),
        (13, "canon", 10**8, 10**8 + 2 * 10**8, {"hbins": 128, "distinct": True, "dbins": 128}),
        (13, "canon", 5 * 10**8, 5 * 10**8 + 10**8, {"hbins": 128, "distinct": True, "dbins": 128}),
        (17, "full", 10**14, 10**14 + 1_000_000, {"hbins": 128}),
        (9, "full", 0, None, {"hbins": 64}),
    ]
    qs = [(n, m, lo, total(n, m) if hi is None else hi, kw) for n, m, lo, hi, kw in qs]
    got, launches = eng.run_batch(qs)
    # layout 4, rule 0, Brent, no walks (n14 full, n9): the deal blocks, 5e9 - 8 + 362880
    # positions, two launches of at most 2^32, and the 8 deals before n14's first block boundary,
    # one; the same with walks (the two distinct n13 ranges): 3e8 positions, two launches of at
    # most 2^28, one after the other on one branch; layout 4 rule 0 alternation: blocks 1, the ends
    # of the range 1; layout 4 rule 1 (canonical): 1; layout 6 (n = 17): 1. Nine kernel launches,
    # one graph launch.
    assert launches == 1
    for (n, mode, lo, hi, kw), g in zip(qs, got):
        alone = eng.run_range(n, mode, lo, hi, **kw)
        for k in KEYS:
            assert g[k] == alone[k], (n, mode, lo, hi, k)


def test_staged_batch_runs_launch_by_launch(eng, orc):
    """A staged schedule (the queue buffers of a layout are shared) runs each kernel launch as its own
    graph; the counts are the same."""
    qs = [(10, "full", 0, 3628800, {"hbins": 64, "tuning": {"stage_kind": "lockstep"}}),
          (11, "canon", 0, 10**6, {"hbins": 64, "tuning": {"stage_kind": "lockstep"}, "rule": 1})]
    got, launches = eng.run_batch(qs)
    assert launches == 2
    for (n, m, lo, hi, kw), g in zip(qs, got):
        w = orc.run(n, m, lo, hi, hbins=64, rule=kw.get("rule", 0))
        assert (g["cyc"], g["turns"], g["hist"]) == (w["cyc"], w["turns"], w["hist"])


def test_batch_refusals(eng):
    with pytest.raises(WarError, match="persistent kernel"):
        eng.run_batch([(9, "full", 0, 100, {"kernel": "plain"})])
    with pytest.raises(WarError, match="Brent or the alternation"):
        eng.run_batch([(10, "full", 0, 100, {"detector": "pair"})])
    with pytest.raises(WarError, match="free request slots"):
        eng.run_batch([(9, "full", 0, 100, {})] * 300)
    got, launches = eng.run_batch([])
    assert got == [] and launches == 0
    # an empty range is a result of zeros, with no launch
    got, launches = eng.run_batch([(9, "full", 50, 50, {"hbins": 8})])
    assert launches == 0 and got[0]["done"] == 0 and got[0]["cyc"] == 0
