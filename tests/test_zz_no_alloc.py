"""war_run_range_noalloc (the entry point that allocates nothing): through the C ABI it must give
exactly war_run_range's counts, on two calls in a row with the second smaller (the scratch array
keeps the first call's contents, which the second must clear), on both kernels, every pile layout
and detector it accepts; and refuse what it does not accept."""
import pytest

from war_amd import WarError

pytestmark = pytest.mark.gpu

KEYS = ("cyc", "done", "maxfin", "turns", "undetermined", "hist", "lenhist")

CASES = [
    # (n, mode, lo, hi, kw) pairs: a larger call, then a smaller one
    ((13, "full", 10**9, 10**9 + 50_000_000, {"hbins": 256, "lbins": 512, "detector": "alt"}),
     (13, "full", 10**9 + 7, 10**9 + 1_000_007, {"hbins": 64, "lbins": 32, "detector": "alt"})),
    ((14, "canon", 3 * 10**9, 3 * 10**9 + 40_000_000, {"hbins": 4096, "lbins": 4096}),
     (14, "canon", 3 * 10**9 + 5, 3 * 10**9 + 300_005, {"hbins": 128})),
    ((9, "full", 0, None, {"hbins": 64, "lbins": 64, "kernel": "plain"}),
     (9, "full", 100, 5000, {"lbins": 16, "kernel": "plain", "detector": "streak"})),
    ((17, "full", 10**14, 10**14 + 20_000_000, {"hbins": 128, "lbins": 128}),
     (17, "full", 10**14, 10**14 + 1_000_000, {"hbins": 128, "tuning": {"pile_layout": 5}})),
    ((18, "canon", 10**12, 10**12 + 2_000_000, {"hbins": 256, "kernel": "plain"}),
     (12, "half", 10**6, 2 * 10**6, {"hbins": 64, "lbins": 64})),
    ((11, "full", 0, None, {"hbins": 128, "rule": 1}),
     (10, "canon", 0, 1_000_000, {"hbins": 128, "rule": 1, "step_cap": 64})),
    ((14, "canon", 0, 20_000_000, {"hbins": 128, "detector": "pair"}),
     (9, "full", 0, None, {"hbins": 64, "step_cap": 32})),
]


@pytest.mark.parametrize("pair", range(len(CASES)))
def test_noalloc_equals_run_range(eng, pair):
    for n, mode, lo, hi, kw in CASES[pair]:
        print("case", pair, n, mode, lo, hi, kw, flush=True)
        want = eng.run_range(n, mode, lo, hi, **kw)
        got = eng.run_range_noalloc(n, mode, lo, hi, **kw)
        for k in KEYS:
            assert got[k] == want[k], (n, mode, lo, hi, kw, k, got[k], want[k])


def test_noalloc_refusals(eng):
    with pytest.raises(WarError, match="cycle walks"):
        eng.run_range_noalloc(11, "canon", 0, 1000, distinct=True)
    with pytest.raises(WarError, match="cycle walks"):
        eng.run_range_noalloc(11, "canon", 0, 1000, shape_blk=1)
    with pytest.raises(WarError, match="bins must be in"):
        eng.run_range_noalloc(11, "full", 0, 1000, hbins=5000)
    with pytest.raises(WarError, match="flow schedule"):
        eng.run_range_noalloc(11, "full", 0, 1000, tuning={"stage_kind": "lockstep"})
    # after refusals, a call still gives war_run_range's counts
    a = eng.run_range(10, "full", hbins=32)
    b = eng.run_range_noalloc(10, "full", hbins=32)
    assert (a["cyc"], a["turns"], a["hist"]) == (b["cyc"], b["turns"], b["hist"])
