"""The step cap of the oracle (orc_run_cap): a game that its detector has not resolved within the cap
counts as undetermined, with the cap's turns in `turns`.

What pins it: the same oracle without a cap, position by position. For Brent and the alternation
detector, a position resolved at round R (pinned by the published counts and the set-based model)
is undetermined under cap c iff R > c, charged c turns; for the streak test, whose test is the
reference's ("more than c turns played" at the top of its loop, gpu/war24.py:103), iff R > c + 1,
charged c + 1. So the capped totals are sums over single positions of the uncapped results: an
off-by-one in either form moves the counts below, at the caps chosen to cut through the
distribution of R."""
import pytest


def _per_position(orc, n, mode, lo, hi, detector, rule=0):
    out = []
    for p in range(lo, hi):
        r = orc.run(n, mode, p, p + 1, rule=rule, lbins=1, threads=1, detector=detector)
        out.append((r["turns"], r["cyc"] == 1, r["done"] == 1))
    return out


def _capped(per, cap, extra):
    """Totals under cap from uncapped (rounds, cycles, played) per position; extra = 1 for the
    streak test's 'more than cap' form."""
    und = cyc = turns = done = 0
    maxfin = 0
    for rounds, cy, played in per:
        if not played:
            continue
        done += 1
        if rounds > cap + extra:
            und += 1
            turns += cap + extra
            continue
        turns += rounds
        if cy:
            cyc += 1
        else:
            maxfin = max(maxfin, rounds)
    return {"undetermined": und, "cyc": cyc, "turns": turns, "done": done, "maxfin": maxfin}


@pytest.mark.parametrize("n,mode,lo,hi,det,rule", [
    (7, "full", 0, 5040, "brent", 0),
    (8, "canon", 0, 5040 * 7 // 2, "brent", 1),
    (9, "full", 100000, 103000, "alt", 0),
    (9, "full", 200000, 203000, "streak", 0),
    (10, "canon", 7777, 10777, "brent", 0),
])
def test_cap_equals_uncapped_cut(orc, n, mode, lo, hi, det, rule):
    per = _per_position(orc, n, mode, lo, hi, det, rule)
    rounds = sorted(r for r, _, p in per if p)
    extra = 1 if det == "streak" else 0
    # caps through the distribution of rounds: low, median, high, and one above every round
    for cap in (rounds[len(rounds) // 10], rounds[len(rounds) // 2], rounds[-2], rounds[-1] + 5):
        want = _capped(per, cap, extra)
        got = orc.run(n, mode, lo, hi, rule=rule, threads=2, detector=det, step_cap=cap)
        for k, v in want.items():
            assert got[k] == v, (n, mode, det, cap, k, got[k], v)
    # exactly at the largest round: nothing is cut (Brent: resolved at R <= cap is resolved)
    got = orc.run(n, mode, lo, hi, rule=rule, threads=2, detector=det, step_cap=rounds[-1] - extra)
    assert got["undetermined"] == 0


def test_streak_default_cap_is_the_reference():
    """cap 0 for the streak test is the reference's 2^22 (no n <= 9 game comes near it)."""
    import oracle
    a = oracle.run(9, "full", 0, 20000, detector="streak")
    b = oracle.run(9, "full", 0, 20000, detector="streak", step_cap=1 << 22)
    assert a == b and a["undetermined"] == 0
