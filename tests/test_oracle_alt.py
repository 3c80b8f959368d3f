"""The oracle's alternation-checkpoint detector (odd n, rule 0): Brent's tortoise also jumps, at
ages that are multiples of 8, when one of the last 8 turns of odd age repeated the previous turn's
winner, and after waiting 128 turns (then 256, ...). Pinned by:

* a model on plain Python deques of the same schedule (before the first turn player 1 counts as
  the last winner), for every deal and canonical position of odd n <= 7 (outcome, game length,
  cycle length, detection turn);
* the set-based (mu, lam) model: a repeat is seen no earlier than mu + lam;
* Brent, over every deal of odd n <= 11 and every canonical position of odd n <= 9: the same
  endless count, longest game, cycle-length histogram, ending-length histogram, weights, distinct
  counts and shape counts (only the turn count differs, and it is smaller for n >= 9).
"""
import itertools
from collections import deque

import pytest

from test_oracle_model import _canon, _deal, model

K, L0 = 8, 128


def alt_model(p1, p2):
    a, b = deque(p1), deque(p2)
    t = (tuple(a), tuple(b))
    R, age, L, streak, prevw = 0, 0, L0, False, 1       # player 1 counts as the last winner at the start
    if not a or not b:
        return False, 0, 0
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
        if w == prevw and age % 2 == 1:
            streak = True
        prevw = w
        if age % K == 0:
            if age - R >= L:
                t, R, L = cur, age, 2 * L
            elif streak:
                t, R = cur, age
            streak = False


@pytest.mark.parametrize("n", [3, 5, 7])
def test_alt_against_models(orc, n):
    for r, p in enumerate(itertools.permutations(range(n))):
        endless, age, lam = alt_model(*_deal(p))
        e2, mu, lam2, length = model(*_deal(p), 0)
        assert endless == e2
        got = orc.run(n, "full", r, r + 1, hbins=4096, threads=1, detector="alt")
        assert (got["cyc"], got["turns"]) == (int(endless), age)
        if endless:
            assert lam == lam2 and got["hist"][lam] == 1 and age >= mu + lam
        else:
            assert got["maxfin"] == length == age
    m = n - 1
    for r, q in enumerate(itertools.permutations(range(m))):
        for k in range(m):
            endless, age, lam = alt_model(*_canon(q, n, k))
            got = orc.run(n, "canon", r * m + k, r * m + k + 1, hbins=4096, threads=1, detector="alt")
            assert (got["cyc"], got["turns"]) == (int(endless), age)


@pytest.mark.parametrize("n,mode", [(5, "full"), (7, "full"), (9, "full"), (11, "full"), (5, "canon"),
                                    (7, "canon"), (9, "canon")])
def test_alt_equals_brent(orc, n, mode):
    kw = dict(hbins=256, lbins=128)
    if mode == "canon":
        kw.update(dbins=256, distinct=True, shape_blk=n & -n)
    a = orc.run(n, mode, detector="alt", **kw)
    b = orc.run(n, mode, **kw)
    for k in a:
        if k != "turns":
            assert a[k] == b[k], (n, mode, k)
    # fewer turns where games are long enough to matter; the deals of n = 5 and 7 (games of a few
    # dozen turns) whose last repeated winner fell on a turn of even age wait for the timed jump
    if n >= 9:
        assert a["turns"] < b["turns"]


def test_alt_refusals(orc):
    for n, rule in ((8, 0), (9, 1)):
        with pytest.raises(ValueError, match="alternation"):
            orc.run(n, "full", 0, 10, rule=rule, detector="alt")
