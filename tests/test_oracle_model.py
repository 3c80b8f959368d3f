"""The oracle against mathematics it must satisfy, independent of its own code:

* lexicographic unranking = itertools.permutations order;
* per position, a set-based game model written from the rule (every state recorded, so the
  pre-period mu and cycle length lam come out exactly) plus the closed form of Brent's schedule:
  the tortoise jumps at rounds 2^k - 1 with power 2^k, so a cycle is seen at round R* + lam, where
  R* is the smallest R in {0, 1, 3, 7, ...} with R >= mu and power(R) >= lam (power(0) = 1);
* published single positions on cycles of 96 and 112 turns;
* additivity over split ranges and the histogram invariants.
"""
import itertools
import math
import random
from collections import deque

import pytest


def model(p1, p2, rule):
    a, b = deque(p1), deque(p2)
    seen = {}
    t = 0
    while a and b:
        key = (tuple(a), tuple(b))
        if key in seen:
            return True, seen[key], t - seen[key], None
        seen[key] = t
        x, y = a.popleft(), b.popleft()
        if x > y:
            a.extend((x, y) if rule == 0 else (y, x))
        else:
            b.extend((y, x) if rule == 0 else (x, y))
        t += 1
    return False, None, None, t


def brent_round(mu, lam):
    R, power = 0, 1
    while not (R >= mu and power >= lam):
        R, power = 2 * R + 1, 2 * power
    return R + lam


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 6])
def test_unrank_is_lexicographic(orc, m):
    for r, p in enumerate(itertools.permutations(range(m))):
        assert orc.unrank(r, m) == list(p)


def test_unrank_extremes(orc):
    for m in (10, 13, 16, 20):
        assert orc.unrank(0, m) == list(range(m))
        assert orc.unrank(math.factorial(m) - 1, m) == list(range(m))[::-1]
    # rank J*(m-2)! starts with the J-th ordered pair of distinct cards (cpu/war.cpp:40-46)
    m = 9
    pairs = [(i, j) for i in range(m) for j in range(m) if i != j]
    for J, (i, j) in enumerate(pairs):
        assert orc.unrank(J * math.factorial(m - 2), m)[:2] == [i, j]


def _deal(p):
    return list(p[0::2]), list(p[1::2])


def _canon(q, n, k):
    return [n - 1] + list(q[:k]), list(q[k:])


@pytest.mark.parametrize("n,rule", [(n, r) for n in range(1, 8) for r in (0, 1)])
def test_per_deal_against_model(orc, n, rule):
    for r, p in enumerate(itertools.permutations(range(n))):
        endless, mu, lam, length = model(*_deal(p), rule)
        got = orc.run(n, "full", r, r + 1, rule=rule, hbins=4096, lbins=4096, threads=1)
        assert got["done"] == 1
        assert got["cyc"] == int(endless)
        if endless:
            assert got["hist"][lam] == 1
            assert got["turns"] == brent_round(mu, lam)
        else:
            assert got["maxfin"] == length == got["turns"]
            assert got["lenhist"][length] == 1


@pytest.mark.parametrize("n,rule", [(n, r) for n in range(2, 8) for r in (0, 1)])
def test_per_canonical_position_against_model(orc, n, rule):
    pos = 0
    for q in itertools.permutations(range(n - 1)):
        for k in range(n - 1):
            endless, mu, lam, length = model(*_canon(q, n, k), rule)
            got = orc.run(n, "canon", pos, pos + 1, rule=rule, hbins=4096, threads=1)
            assert got["cyc"] == int(endless)
            if endless:
                assert got["hist"][lam] == 1 and got["turns"] == brent_round(mu, lam)
            else:
                assert got["maxfin"] == length == got["turns"]
            pos += 1


@pytest.mark.parametrize("p1,p2,lam", [
    ((13, 11, 10, 4), (8, 2, 12, 7, 5, 1, 14, 9, 6, 3), 112),        # figures/war_barcode.py:10
    ((1, 13, 11, 9, 4), (3, 5, 2, 14, 12, 10, 8, 7, 6), 96),         # figures/war_barcode.py:11
    ((7, 2, 14, 9, 10, 3), (13, 6, 5, 1, 12, 8, 11, 4), 112),        # cycles/spivey_categories.py:35
    ((3, 5, 1, 14, 9, 12, 7, 8, 6), (2, 13, 10, 11, 4), 96),         # cycles/spivey_categories.py:33
])
def test_published_positions(orc, p1, p2, lam):
    a = [c - 1 for c in p1]
    b = [c - 1 for c in p2]
    endless, rounds, got_lam = orc.play(14, a, b)
    assert endless and got_lam == lam
    e2, mu, lam2, _ = model(a, b, 0)
    assert e2 and lam2 == lam and rounds == brent_round(mu, lam)


def test_5bit_layout_matches_model(orc):
    """n >= 17 uses 5-bit piles in the oracle; check single random positions against the model."""
    rng = random.Random(17)
    for n in (17, 18, 20):
        for _ in range(30):
            p = list(range(n))
            rng.shuffle(p)
            k = rng.randrange(1, n)
            endless, rounds, lam = orc.play(n, p[:k], p[k:])
            e2, mu, lam2, length = model(p[:k], p[k:], 0)
            assert endless == e2
            if endless:
                assert lam == lam2 and rounds == brent_round(mu, lam2)
            else:
                assert rounds == length


@pytest.mark.parametrize("n,mode", [(10, "full"), (11, "canon"), (10, "half"), (17, "full")])
def test_additivity_and_invariants(orc, n, mode):
    total = orc.total(n, mode)
    rng = random.Random(n)
    lo = rng.randrange(0, total // 2)
    hi = min(total, lo + 200_000)
    cuts = sorted({lo, hi, *(rng.randrange(lo, hi) for _ in range(5))})
    whole = orc.run(n, mode, lo, hi, hbins=512, lbins=512)
    parts = [orc.run(n, mode, a, b, hbins=512, lbins=512) for a, b in zip(cuts, cuts[1:])]
    for k in ("cyc", "done", "turns"):
        assert whole[k] == sum(p[k] for p in parts)
    assert whole["maxfin"] == max(p["maxfin"] for p in parts)
    assert whole["hist"] == [sum(x) for x in zip(*(p["hist"] for p in parts))]
    assert whole["lenhist"] == [sum(x) for x in zip(*(p["lenhist"] for p in parts))]
    # invariants: sum(hist) = cyc; sum(lenhist) = ending games; top lenhist bin = maxfin
    assert sum(whole["hist"]) == whole["cyc"]
    assert sum(whole["lenhist"]) == whole["done"] - whole["cyc"]
    top = max(i for i, v in enumerate(whole["lenhist"]) if v)
    assert top == whole["maxfin"]
    if mode != "half":
        assert whole["done"] == hi - lo


def test_distinct_additive_on_partial_ranges(orc, goldens):
    n = 11
    total = orc.total(n, "canon")
    cuts = [0, 5_000_000, 12_345_678, 30_000_000, total]
    parts = [orc.run(n, "canon", a, b, distinct=True) for a, b in zip(cuts, cuts[1:])]
    assert sum(p["distinct_minkey"] for p in parts) == goldens["canon"]["11"]["cycles"]
    assert sum(p["oncycle"] for p in parts) == goldens["canon"]["11"]["on_cycle"]
    # cycles touched per part >= the cycles counted there (each counted cycle is touched)
    for p in parts:
        assert p["distinct_keys"] >= p["distinct_minkey"]


def test_errors(orc):
    with pytest.raises(ValueError):
        orc.run(9, "full", 0, 10**7)
    with pytest.raises(ValueError):
        orc.run(9, "half")
    with pytest.raises(ValueError):
        orc.run(9, "full", distinct=True)
    assert orc.run(9, "full", 5, 5)["done"] == 0
