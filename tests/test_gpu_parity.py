"""Parity of the HIP engine (through the C ABI) with the CPU oracle and the published goldens.
Integer counts: bit-exact, tolerance 0."""
import random

import pytest

from war_amd import WarError

pytestmark = pytest.mark.gpu

FIELDS = ("cyc", "done", "maxfin", "turns", "hist", "lenhist")


def check(eng, orc, n, mode, lo, hi, rule=0, kernel="persistent", fields=FIELDS, tuning=None, **kw):
    want = orc.run(n, mode, lo, hi, rule=rule, hbins=256, lbins=512, **kw)
    got = eng.run_range(n, mode, lo, hi, rule=rule, kernel=kernel, hbins=256, lbins=512, tuning=tuning, **kw)
    for k in fields:
        assert got[k] == want[k], (n, mode, lo, hi, rule, kernel, k, got[k], want[k])
    return got, want


@pytest.mark.parametrize("kernel", ["plain", "persistent"])
@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("mode", ["full", "canon"])
@pytest.mark.parametrize("n", list(range(1, 11)))
def test_whole_range_small_n(eng, orc, n, mode, rule, kernel):
    print("case", n, mode, rule, kernel, flush=True)
    check(eng, orc, n, mode, 0, None, rule=rule, kernel=kernel)


@pytest.mark.parametrize("kernel", ["plain", "persistent"])
@pytest.mark.parametrize("n,mode", [(11, "full"), (12, "full"), (13, "full"), (14, "full"), (15, "full"),
                                    (16, "full"), (11, "canon"), (12, "canon"), (13, "canon"), (14, "canon"),
                                    (15, "canon"), (16, "canon")])
def test_sampled_ranges(eng, orc, n, mode, kernel):
    rng = random.Random(1000 * n + len(mode))
    total = orc.total(n, mode)
    for _ in range(3):
        ln = rng.choice([1, 63, 64, 65, 1000, 1 << 16, 300_001])
        lo = rng.randrange(0, total - ln)
        print("case", n, mode, kernel, lo, ln, flush=True)
        check(eng, orc, n, mode, lo, lo + ln, kernel=kernel)
    # the last positions of the range
    check(eng, orc, n, mode, total - 5000, total, kernel=kernel)


@pytest.mark.parametrize("n", [4, 8, 10, 12])
def test_half_mode(eng, orc, n):
    total = orc.total(n, "full")
    lo, hi = (0, total) if n <= 10 else (total // 3, total // 3 + 2_000_000)
    for kernel in ("plain", "persistent"):
        check(eng, orc, n, "half", lo, hi, kernel=kernel)


@pytest.mark.parametrize("n", [7, 9, 11])
def test_distinct_shape_whole_range(eng, orc, goldens, n):
    got, want = check(eng, orc, n, "canon", 0, None, dbins=4096, distinct=True, shape_blk=n & -n,
                      fields=FIELDS + ("oncycle", "distinct_minkey", "dhist", "shape_nonstd", "shape_unchk"))
    assert got["distinct_hashset"] == want["distinct_keys"] == goldens["canon"][str(n)]["cycles"]


@pytest.mark.parametrize("n", [11, 13])
def test_distinct_partial_ranges(eng, orc, n):
    rng = random.Random(n)
    total = orc.total(n, "canon")
    for _ in range(3):
        lo = rng.randrange(0, total - 3_000_000)
        got, want = check(eng, orc, n, "canon", lo, lo + 3_000_000, dbins=4096, distinct=True,
                          fields=FIELDS + ("oncycle", "distinct_minkey", "dhist"))
        assert got["distinct_hashset"] == want["distinct_keys"]


@pytest.mark.parametrize("tuning", [
    {"brent_cap": 7, "ranks_per_refill": 64, "waves_per_cu": 4},
    {"brent_cap": 31, "ranks_per_refill": 4096, "waves_per_cu": 32},
    {"brent_cap": 127, "ranks_per_refill": 256, "waves_per_cu": 8},
    {"brent_cap": 1023, "ranks_per_refill": 1024, "waves_per_cu": 16},
    {"brent_cap": 1, "ranks_per_refill": 64, "waves_per_cu": 2},      # sparse tiles: grabs span many tiles
    {"brent_cap": 3, "ranks_per_refill": 8192, "waves_per_cu": 1},
    {"brent_cap": 15, "ranks_per_refill": 1024, "waves_per_cu": 32, "stage_kind": "lockstep"},
    {"brent_cap": 1, "ranks_per_refill": 256, "waves_per_cu": 8, "stage_kind": "lockstep"},
    {"stage_kind": "flow", "ranks_per_refill": 64, "waves_per_cu": 4},
    {"stage_kind": "flow", "ranks_per_refill": 8192, "waves_per_cu": 32},
    {"stage_kind": "flow", "ranks_per_refill": 128, "waves_per_cu": 1},
])
def test_schedules_are_exact(eng, orc, tuning):
    check(eng, orc, 13, "canon", 123_456_789, 123_456_789 + 2_000_000, tuning=tuning)
    check(eng, orc, 13, "full", 3_000_000_000, 3_000_000_000 + 1_000_000, tuning=tuning)
    check(eng, orc, 8, "full", 0, None, rule=1, tuning=tuning)          # short games: sparse later stages
    check(eng, orc, 12, "canon", 7_000_000, 7_000_000 + 500_000, tuning=tuning, dbins=64, distinct=True,
          shape_blk=4, fields=FIELDS + ("oncycle", "distinct_minkey", "dhist", "shape_nonstd"))


FLOW = {"stage_kind": "flow"}


@pytest.mark.parametrize("n,mode,rule", [(n, m, r) for n in (2, 3, 4, 5, 6, 7, 8, 9, 10) for m in ("full", "canon", "half")
                                         for r in (0, 1) if m != "half" or n % 2 == 0])
def test_flow_brent_whole_ranges(eng, orc, n, mode, rule):
    print("case flow brent", n, mode, rule, flush=True)
    check(eng, orc, n, mode, 0, None, rule=rule, tuning=FLOW)


def test_flow_brent_sampled(eng, orc):
    rng = random.Random(11)
    for n, mode, rule, lay in [(12, "half", 0, 0), (13, "full", 1, 0), (14, "canon", 0, 0), (14, "full", 0, 0),
                               (15, "canon", 1, 0), (16, "full", 0, 0), (17, "full", 0, 6), (17, "full", 0, 5),
                               (17, "canon", 1, 0), (18, "full", 0, 0), (20, "full", 1, 0)]:
        total = orc.total(n, mode)
        for ln in (1, 63, 150_001):
            lo = rng.randrange(0, total - ln)
            tuning = dict(FLOW, pile_layout=lay) if lay else FLOW
            print("case flow brent", n, mode, rule, lay, lo, ln, flush=True)
            check(eng, orc, n, mode, lo, lo + ln, rule=rule, tuning=tuning)


@pytest.mark.parametrize("n", [5, 6, 7, 9, 11, 12, 13, 14, 15, 16])
def test_flow_deal_blocks(eng, orc, n):
    """The flow kernel steps 4-bit deals through blocks of 24 orders that differ in their last four
    cards, one swap per deal: ranges on and off a block boundary, ending mid-block, of every lane-run
    length the tile sizes give, with both detectors and both rules."""
    rng = random.Random(24 * n)
    total = orc.total(n, "full")
    for ln in (1, 23, 24, 25, 47, 1535, 1537, 100_003):
        if ln > total:
            continue
        lo = min(rng.randrange(0, total - ln) // 24 * 24 + rng.choice([0, 1, 23]), total - ln)
        for tuning in (None, {"ranks_per_refill": 64}, {"ranks_per_refill": 8192}):
            for det in (["brent", "alt"] if n % 2 else ["brent"]):
                print("case deal blocks", n, lo, ln, tuning, det, flush=True)
                check(eng, orc, n, "full", lo, lo + ln, tuning=tuning, detector=det)
        print("case deal blocks rule 1", n, lo, ln, flush=True)
        check(eng, orc, n, "full", lo, lo + ln, rule=1)


def test_5bit_ranges(eng, orc):
    for n, lo, ln in [(17, 10**14, 200_000), (17, 0, 100_000), (18, 5 * 10**15, 50_000), (20, 10**18, 20_000)]:
        for kernel in ("plain", "persistent"):
            print("case 5bit", n, lo, ln, kernel, flush=True)
            check(eng, orc, n, "full", lo, lo + ln, kernel=kernel)
    check(eng, orc, 17, "canon", 10**13, 10**13 + 100_000)


def test_edges_and_errors(eng):
    r = eng.run_range(9, "full", 100, 100)
    assert r["done"] == 0 and r["cyc"] == 0
    r = eng.run_range(1, "full")
    assert (r["done"], r["cyc"], r["maxfin"]) == (1, 0, 0)
    r = eng.run_range(2, "canon")
    assert (r["done"], r["cyc"], r["maxfin"]) == (1, 0, 1)
    for bad in [dict(n=0), dict(n=21), dict(n=9, hi=10**9), dict(n=9, mode="half"), dict(n=9, distinct=True),
                dict(n=16, mode="canon", distinct=True), dict(n=9, hbins=5000),
                dict(n=9, tuning={"brent_cap": 10})]:
        kw = dict(bad)
        n = kw.pop("n")
        with pytest.raises(WarError):
            eng.run_range(n, **kw)


def test_n13_full_goldens(eng, goldens):
    g = goldens["deals"]["13"]
    r = eng.run_range(13, "full", hbins=256, lbins=512)
    assert r["done"] == g["deals"] and r["cyc"] == g["never_end"] and r["maxfin"] == g["longest"]
    assert {str(i): v for i, v in enumerate(r["hist"]) if v} == g["hist"]
    assert sum(r["lenhist"]) == g["deals"] - g["never_end"]
    print("n13 full persistent device_ms", r["device_ms"], flush=True)


def test_n14_canon_goldens(eng, goldens):
    g = goldens["long"]["canon14"]
    r = eng.run_range(14, "canon", hbins=256, lbins=512, dbins=4096, distinct=True, shape_blk=2)
    assert r["done"] == g["positions"] and r["cyc"] == g["cycling"] and r["maxfin"] == g["longest"]
    assert {str(i): v for i, v in enumerate(r["hist"]) if v} == g["hist"]
    assert {str(i): v for i, v in enumerate(r["dhist"]) if v} == g["weights"]
    assert (r["shape_nonstd"], r["shape_unchk"]) == (g["shape"]["nonstd"], g["shape"]["unchk"])
    assert r["distinct_minkey"] == r["distinct_hashset"] == g["distinct"]
    assert r["oncycle"] == goldens["canon"]["14"]["on_cycle"]
    print("n14 canon persistent device_ms", r["device_ms"], flush=True)


# ---- the streak test (gpu/war24.py:98-125): plain kernel, odd n, rule 0
STREAK_FIELDS = ("cyc", "done", "maxfin", "turns", "lenhist", "undetermined")


def check_streak(eng, orc, n, mode, lo, hi):
    want = orc.run(n, mode, lo, hi, lbins=512, detector="streak")
    got = eng.run_range(n, mode, lo, hi, kernel="plain", detector="streak", lbins=512)
    for k in STREAK_FIELDS:
        assert got[k] == want[k], (n, mode, lo, hi, k, got[k], want[k])
    return got


@pytest.mark.parametrize("n,mode", [(3, "full"), (5, "full"), (7, "full"), (9, "full"), (11, "full"),
                                    (5, "canon"), (7, "canon"), (9, "canon"), (11, "canon")])
def test_streak_whole_range(eng, orc, n, mode):
    print("case streak", n, mode, flush=True)
    got = check_streak(eng, orc, n, mode, 0, orc.total(n, mode))
    brent = eng.run_range(n, mode, kernel="persistent")
    assert (got["cyc"], got["maxfin"]) == (brent["cyc"], brent["maxfin"])


def test_streak_sampled_ranges(eng, orc):
    rng = random.Random(24)
    for n in (13, 15, 17, 19):
        total = orc.total(n, "full")
        for ln in (1, 65, 100_003):
            lo = rng.randrange(0, total - ln)
            print("case streak", n, lo, ln, flush=True)
            check_streak(eng, orc, n, "full", lo, lo + ln)
    check_streak(eng, orc, 13, "canon", 10**9, 10**9 + 200_000)


def test_streak_refusals(eng):
    for bad in [dict(n=14), dict(n=13, rule=1), dict(n=13, kernel="persistent"), dict(n=13, hbins=16),
                dict(n=13, mode="canon", dbins=16), dict(n=13, mode="canon", shape_blk=1)]:
        kw = dict(bad)
        n = kw.pop("n")
        kw.setdefault("kernel", "plain")
        with pytest.raises(WarError, match="streak"):
            eng.run_range(n, lo=0, hi=100, detector="streak", **kw)


# ---- pile layouts: n = 17 with card 16 held apart (layout 6) against the 5-bit layout and the oracle
@pytest.mark.parametrize("rule", [0, 1])
def test_held_layout_n17(eng, orc, rule):
    rng = random.Random(17 + rule)
    cases = [("full", 0, 70_000), ("full", orc.total(17, "full") - 50_000, 50_000), ("canon", 0, 40_000)]
    for _ in range(3):
        mode = rng.choice(["full", "canon"])
        ln = rng.choice([1, 64, 65, 100_000])
        cases.append((mode, rng.randrange(0, orc.total(17, mode) - ln), ln))
    for mode, lo, ln in cases:
        print("case held", rule, mode, lo, ln, flush=True)
        g6, want = check(eng, orc, 17, mode, lo, lo + ln, rule=rule, tuning={"pile_layout": 6})
        g5, _ = check(eng, orc, 17, mode, lo, lo + ln, rule=rule, tuning={"pile_layout": 5})
        assert g6["turns"] == g5["turns"] == want["turns"]
    assert eng.tuning_for(17, "full")["pile_layout"] == 0           # auto: 6 where it applies


def test_layout_5_for_small_n(eng, orc):
    """The 5-bit layout is exact for any n: the same counts as the 4-bit one."""
    for n, mode, lo, ln in [(13, "full", 10**9, 300_000), (14, "canon", 7 * 10**9, 200_000), (9, "full", 0, 362_880)]:
        print("case layout5", n, mode, lo, ln, flush=True)
        check(eng, orc, n, mode, lo, lo + ln, tuning={"pile_layout": 5})


def test_layout_refusals(eng):
    for n, kw in [(16, {}), (17, {"kernel": "plain"}), (17, {"mode": "canon", "shape_blk": 1}),
                  (18, {})]:
        kw = dict(kw)
        with pytest.raises(WarError, match="pile_layout 6"):
            eng.run_range(n, lo=0, hi=100, tuning={"pile_layout": 6}, **kw)
    with pytest.raises(WarError, match="pile_layout 4"):
        eng.run_range(17, lo=0, hi=100, tuning={"pile_layout": 4})
    with pytest.raises(WarError, match="4-bit"):
        eng.run_range(11, "canon", lo=0, hi=100, distinct=True, tuning={"pile_layout": 5})


# ---- the alternation detector (odd n, rule 0): every output equal to the oracle's, turns its own
ALT_FIELDS = ("cyc", "done", "maxfin", "turns", "hist", "lenhist", "undetermined")


def check_alt(eng, orc, n, mode, lo, hi, kernel="persistent", tuning=None, **kw):
    want = orc.run(n, mode, lo, hi, hbins=256, lbins=512, detector="alt", **kw)
    got = eng.run_range(n, mode, lo, hi, kernel=kernel, hbins=256, lbins=512, detector="alt", tuning=tuning, **kw)
    walk_fields = ("oncycle", "distinct_minkey", "dhist", "shape_nonstd", "shape_unchk") if kw else ()
    for k in ALT_FIELDS + walk_fields:
        assert got[k] == want[k], (n, mode, lo, hi, kernel, tuning, k, got[k], want[k])
    return got, want


@pytest.mark.parametrize("kernel", ["plain", "persistent"])
@pytest.mark.parametrize("n,mode", [(3, "full"), (5, "full"), (7, "full"), (9, "full"), (11, "full"),
                                    (5, "canon"), (7, "canon"), (9, "canon"), (11, "canon")])
def test_alt_whole_range(eng, orc, n, mode, kernel):
    print("case alt", n, mode, kernel, flush=True)
    got, _ = check_alt(eng, orc, n, mode, 0, orc.total(n, mode), kernel=kernel)
    brent = eng.run_range(n, mode, hbins=256, lbins=512)
    assert (got["cyc"], got["maxfin"], got["hist"], got["lenhist"]) == \
           (brent["cyc"], brent["maxfin"], brent["hist"], brent["lenhist"])


def test_alt_sampled_and_layouts(eng, orc):
    rng = random.Random(2)
    for n, mode in [(13, "full"), (15, "full"), (13, "canon"), (15, "canon"), (17, "full"), (17, "canon"),
                    (19, "full")]:
        total = orc.total(n, mode)
        for ln in (1, 65, 200_003):
            lo = rng.randrange(0, total - ln)
            for kernel in ("plain", "persistent"):
                print("case alt", n, mode, lo, ln, kernel, flush=True)
                check_alt(eng, orc, n, mode, lo, lo + ln, kernel=kernel)
    for lay in (5, 6):
        check_alt(eng, orc, 17, "full", 10**14, 10**14 + 300_000, tuning={"pile_layout": lay})


@pytest.mark.parametrize("stage_kind", ["flow", "refill", "lockstep"])
def test_alt_held_card_to_player2(eng, orc, stage_kind):
    """n = 17 deals whose card 16 goes to player 2 (deck position 13 at the range's start): layout 6
    names the players by who holds card 16, and the detector must still count the deal's player 1
    as the winner of the turn before the first. With the holder counted instead, `turns` here is
    22 862 471 over the whole range instead of 22 749 831 (every other output the same)."""
    lo = 302_876_792_870_854
    assert list(orc.unrank(lo, 17)).index(16) == 13
    for lay in (6, 5):
        check_alt(eng, orc, 17, "full", lo, lo + 196_845, tuning={"pile_layout": lay, "stage_kind": stage_kind})


@pytest.mark.parametrize("n", [9, 11])
def test_alt_walks(eng, orc, n):
    print("case alt walks", n, flush=True)
    check_alt(eng, orc, n, "canon", 0, orc.total(n, "canon"), dbins=4096, distinct=True, shape_blk=n & -n)


@pytest.mark.parametrize("tuning", [{"brent_cap": 7, "stage_kind": "lockstep"},
                                    {"brent_cap": 15, "ranks_per_refill": 64, "stage_kind": "lockstep"},
                                    {"brent_cap": 63, "waves_per_cu": 4, "stage_kind": "lockstep"},
                                    {"brent_cap": 255, "ranks_per_refill": 8192, "stage_kind": "lockstep"},
                                    {"stage_kind": "flow", "ranks_per_refill": 64},
                                    {"stage_kind": "flow", "ranks_per_refill": 8192, "waves_per_cu": 4},
                                    {"stage_kind": "flow", "ranks_per_refill": 128, "waves_per_cu": 32}])
def test_alt_schedules_are_exact(eng, orc, tuning):
    print("case alt schedule", tuning, flush=True)
    check_alt(eng, orc, 11, "full", 0, orc.total(11, "full"), tuning=tuning)
    check_alt(eng, orc, 13, "canon", 3 * 10**9, 3 * 10**9 + 400_000, tuning=tuning, dbins=64, distinct=True,
              shape_blk=1)


def test_alt_n13_goldens_and_refusals(eng, goldens):
    g = goldens["deals"]["13"]
    r = eng.run_range(13, "full", hbins=256, lbins=512, detector="alt")
    assert r["done"] == g["deals"] and r["cyc"] == g["never_end"] and r["maxfin"] == g["longest"]
    assert {str(i): v for i, v in enumerate(r["hist"]) if v} == g["hist"]
    print("n13 full alt device_ms", r["device_ms"], flush=True)
    for kw in ({"n": 14}, {"n": 13, "rule": 1}, {"n": 13, "tuning": {"brent_cap": 1}}):
        kw = dict(kw)
        n = kw.pop("n")
        with pytest.raises(WarError, match="alternation"):
            eng.run_range(n, lo=0, hi=100, detector="alt", **kw)


# ---- the pair-run detector (n = 2 mod 4, rule 0; flow kernel): every output equal to the oracle's
def check_pair(eng, orc, n, mode, lo, hi, tuning=None, **kw):
    want = orc.run(n, mode, lo, hi, hbins=256, lbins=512, detector="pair", **kw)
    got = eng.run_range(n, mode, lo, hi, hbins=256, lbins=512, detector="pair", tuning=tuning, **kw)
    walk_fields = ("oncycle", "distinct_minkey", "dhist", "shape_nonstd", "shape_unchk") if kw else ()
    for k in ALT_FIELDS + walk_fields:
        assert got[k] == want[k], (n, mode, lo, hi, tuning, k, got[k], want[k])
    return got, want


@pytest.mark.parametrize("n,mode", [(6, "full"), (6, "canon"), (6, "half"), (10, "full"), (10, "canon"),
                                    (10, "half")])
def test_pair_whole_range(eng, orc, n, mode):
    print("case pair", n, mode, flush=True)
    got, _ = check_pair(eng, orc, n, mode, 0, orc.total(n, mode))
    brent = eng.run_range(n, mode, hbins=256, lbins=512)
    assert (got["cyc"], got["maxfin"], got["hist"], got["lenhist"]) == \
           (brent["cyc"], brent["maxfin"], brent["hist"], brent["lenhist"])


def test_pair_sampled_and_tunings(eng, orc):
    rng = random.Random(3)
    for n, mode in [(14, "full"), (14, "canon"), (14, "half"), (18, "full"), (18, "canon")]:
        total = orc.total(n, mode)
        for ln in (1, 65, 200_003):
            lo = rng.randrange(0, total - ln)
            print("case pair", n, mode, lo, ln, flush=True)
            check_pair(eng, orc, n, mode, lo, lo + ln)
    for tuning in ({"ranks_per_refill": 64}, {"ranks_per_refill": 8192, "waves_per_cu": 4},
                   {"ranks_per_refill": 128, "waves_per_cu": 32}):
        print("case pair tuning", tuning, flush=True)
        check_pair(eng, orc, 10, "full", 0, orc.total(10, "full"), tuning=tuning)
        check_pair(eng, orc, 14, "canon", 4 * 10**10, 4 * 10**10 + 300_000, tuning=tuning)


def test_pair_walks(eng, orc):
    print("case pair walks", flush=True)
    check_pair(eng, orc, 10, "canon", 0, orc.total(10, "canon"), dbins=4096, distinct=True, shape_blk=2)
    check_pair(eng, orc, 14, "canon", 2 * 10**10, 2 * 10**10 + 1_000_000, dbins=64, distinct=True, shape_blk=2)


def test_pair_n14_canon_goldens_and_refusals(eng, goldens):
    r = eng.run_range(14, "canon", hbins=256, detector="pair")
    assert (r["cyc"], r["maxfin"], r["done"]) == (2872129173, 465, 80951270400)
    assert {i: v for i, v in enumerate(r["hist"]) if v} == {96: 43669233, 112: 2828459940}
    print("n14 canon pair device_ms", r["device_ms"], flush=True)
    for n, kw in ((13, {}), (12, {}), (14, {"rule": 1}), (14, {"kernel": "plain"}),
                  (14, {"tuning": {"stage_kind": "lockstep"}})):
        with pytest.raises(WarError, match="pair-run"):
            eng.run_range(n, lo=0, hi=100, detector="pair", **kw)
