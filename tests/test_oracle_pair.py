"""The oracle's pair-run checkpoint detector (n = 2 mod 4, rule 0): Brent's tortoise also jumps, at
ages that are multiples of 8 and at most 128, when the last 8 turns held a winner run of one turn
or of three or more, and after waiting 128 turns (then 256, ...). Pinned by:

* a model on plain Python deques of the same schedule (before the first turn player 1 counts as
  the last winner and the hands as having changed), for every deal and canonical position of
  n = 6, for seeded samples of n = 10 deals and canonical positions, and for n = 14 canonical
  positions that include cycles of both lengths (112, made of runs of two, and 96, which is not:
  there the pattern jumps stop at age 128 and a timed jump finds the cycle); outcome, game length,
  cycle length and detection turn;
* the set-based (mu, lam) model: a repeat is seen no earlier than mu + lam;
* Brent, over every deal and canonical position of n = 6 and 10 and sampled n = 14 ranges: the
  same endless count, longest game, cycle-length histogram, ending-length histogram, weights,
  distinct counts and shape counts (only the turn count differs, and it is smaller where games
  cycle).
"""
import itertools
import random
from collections import deque

import pytest

from test_oracle_model import _canon, _deal, model

K, L0, P = 8, 128, 128


def pair_model(p1, p2):
    a, b = deque(p1), deque(p2)
    if not a or not b:
        return False, 0, 0
    t = (tuple(a), tuple(b))
    R, age, L, viol, prevw, preveq = 0, 0, L0, False, 1, False
    while True:
        x, y = a.popleft(), b.popleft()
        if x > y:
            a.extend((x, y)); w = 1
        else:
            b.extend((y, x)); w = 0
        age += 1
        if not a or not b:
            return False, age, 0
        cur = (tuple(a), tuple(b))
        if cur == t:
            return True, age, age - R
        eq = w == prevw
        if eq == preveq:              # the winner changed twice, or repeated twice, in a row
            viol = True
        prevw, preveq = w, eq
        if age % K == 0:
            if age - R >= L:
                t, R, L = cur, age, 2 * L
            elif viol and age <= P:
                t, R = cur, age
            viol = False


def _check_one(orc, n, mode, idx, piles):
    endless, age, lam = pair_model(*piles)
    e2, mu, lam2, length = model(*piles, 0)
    assert endless == e2
    got = orc.run(n, mode, idx, idx + 1, hbins=4096, threads=1, detector="pair")
    assert (got["cyc"], got["turns"]) == (int(endless), age), (n, mode, idx)
    if endless:
        assert lam == lam2 and got["hist"][lam] == 1 and age >= mu + lam
    else:
        assert got["maxfin"] == length == age
    return endless, lam


def test_pair_against_models_n6(orc):
    n = 6
    for r, p in enumerate(itertools.permutations(range(n))):
        _check_one(orc, n, "full", r, _deal(p))
    m = n - 1
    for r, q in enumerate(itertools.permutations(range(m))):
        for k in range(m):
            _check_one(orc, n, "canon", r * m + k, _canon(q, n, k))


def test_pair_against_models_n10_sample(orc):
    n, rng = 10, random.Random(10)
    cycles = 0
    for _ in range(600):
        r = rng.randrange(orc.total(n, "full"))
        cycles += _check_one(orc, n, "full", r, _deal(orc.unrank(r, n)))[0]
    m = n - 1
    for _ in range(1200):
        pos = rng.randrange(orc.total(n, "canon"))
        r, k = divmod(pos, m)
        cycles += _check_one(orc, n, "canon", pos, _canon(orc.unrank(r, m), n, k))[0]
    assert cycles >= 60


def test_pair_against_models_n14_both_cycle_lengths(orc):
    n, m, rng = 14, 13, random.Random(14)
    lams = {96: 0, 112: 0}
    seen = 0
    while lams[96] < 3 or lams[112] < 30:
        pos = rng.randrange(orc.total(n, "canon"))
        r, k = divmod(pos, m)
        piles = _canon(orc.unrank(r, m), n, k)
        endless, _, lam = pair_model(*piles)
        seen += 1
        if endless or seen % 97 == 0:        # every cycling position and a thinned set of endings
            e, lam = _check_one(orc, n, "canon", pos, piles)
            if e:
                lams[lam] += 1
        assert seen < 200_000


@pytest.mark.parametrize("n,mode", [(6, "full"), (6, "canon"), (10, "full"), (10, "canon")])
def test_pair_equals_brent(orc, n, mode):
    kw = dict(hbins=256, lbins=128)
    if mode == "canon":
        kw.update(dbins=256, distinct=True, shape_blk=n & -n)
    a = orc.run(n, mode, detector="pair", **kw)
    b = orc.run(n, mode, **kw)
    for k in a:
        if k != "turns":
            assert a[k] == b[k], (n, mode, k)
    assert a["turns"] < b["turns"] or b["cyc"] == 0


def test_pair_equals_brent_n14_ranges(orc):
    rng = random.Random(1414)
    for mode in ("full", "canon"):
        total = orc.total(14, mode)
        for _ in range(2):
            lo = rng.randrange(total - 2_000_000)
            kw = dict(hbins=256, lbins=512)
            if mode == "canon":
                kw.update(dbins=64, distinct=True, shape_blk=2)
            a = orc.run(14, mode, lo, lo + 2_000_000, detector="pair", **kw)
            b = orc.run(14, mode, lo, lo + 2_000_000, **kw)
            for k in a:
                if k != "turns":
                    assert a[k] == b[k], (mode, lo, k)
            assert a["turns"] < b["turns"]


def test_pair_refusals(orc):
    for n, rule in ((13, 0), (12, 0), (14, 1)):
        with pytest.raises(ValueError, match="pair-run"):
            orc.run(n, "full", 0, 10, rule=rule, detector="pair")
