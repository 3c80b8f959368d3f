"""The oracle's streak test (gpu/war24.py:98-125) against what pins it:

* a model on plain Python deques, written from the reference's description (anchor when the
  winner-to-be holds an even number ws of cards and the other ls >= 2; players alternating for
  lcm(n+1, ws, ls-1) turns back to the anchor = endless; a repeated winner drops the anchor), for
  every deal and canonical position of odd n <= 7: endless or not, game length, and the detection
  turn;
* the set-based model (exact pre-period mu and cycle length lam): a detected repeat can come no
  earlier than mu + lam, and a game that ends has the model's length;
* Brent: the same endless count, longest game and ending-length histogram over every deal of odd
  n <= 11 and every canonical position of odd n <= 9; the published a(n) (data/b_file.txt);
* its refusals: even n, rule 1, cycle-length outputs.
"""
import itertools
import math
from collections import deque

import pytest

from test_oracle_model import _canon, _deal, model


def streak_model(p1, p2, n):
    a, b = deque(p1), deque(p2)
    steps, anch, lam, since, prevw, anchor = 0, False, 0, 0, None, None
    while True:
        if not a or not b:
            return False, steps
        if not anch:
            ws = len(a) if a[0] > b[0] else len(b)
            ls = n - ws
            if ws % 2 == 0 and ls >= 2:
                anch, since, prevw = True, 0, None
                lam = math.lcm(n + 1, ws, ls - 1)
                anchor = (tuple(a), tuple(b))
        x, y = a.popleft(), b.popleft()
        if x > y:
            a.extend((x, y)); w = 1
        else:
            b.extend((y, x)); w = 0
        steps += 1
        if anch:
            if w == prevw:
                anch = False
            else:
                prevw = w
                since += 1
                if since == lam:
                    if (tuple(a), tuple(b)) == anchor:
                        return True, steps
                    anch = False


@pytest.mark.parametrize("n", [3, 5, 7])
def test_streak_against_models(orc, n):
    for r, p in enumerate(itertools.permutations(range(n))):
        endless, steps = streak_model(*_deal(p), n)
        e2, mu, lam, length = model(*_deal(p), 0)
        assert endless == e2
        got = orc.run(n, "full", r, r + 1, lbins=4096, threads=1, detector="streak")
        assert (got["cyc"], got["turns"], got["undetermined"]) == (int(endless), steps, 0)
        if endless:
            assert steps >= mu + lam
        else:
            assert got["maxfin"] == length == steps
    m = n - 1
    for r, q in enumerate(itertools.permutations(range(m))):
        for k in range(m):
            endless, steps = streak_model(*_canon(q, n, k), n)
            got = orc.run(n, "canon", r * m + k, r * m + k + 1, threads=1, detector="streak")
            assert (got["cyc"], got["turns"]) == (int(endless), steps)


@pytest.mark.parametrize("n,mode", [(5, "full"), (7, "full"), (9, "full"), (11, "full"), (5, "canon"),
                                    (7, "canon"), (9, "canon")])
def test_streak_equals_brent(orc, goldens, n, mode):
    s = orc.run(n, mode, lbins=128, detector="streak")
    b = orc.run(n, mode, lbins=128)
    assert (s["cyc"], s["done"], s["maxfin"], s["lenhist"], s["undetermined"]) == \
           (b["cyc"], b["done"], b["maxfin"], b["lenhist"], 0)
    if mode == "full":
        assert s["cyc"] == goldens["a"][str(n)]
    assert s["turns"] < b["turns"]         # the streak test sees most repeats earlier than Brent


def test_streak_refusals(orc):
    for kw in ({"n": 8}, {"n": 9, "rule": 1}, {"n": 9, "hbins": 16}, {"n": 9, "mode": "canon", "dbins": 16},
               {"n": 9, "mode": "canon", "shape_blk": 1}):
        kw = dict(kw)
        n = kw.pop("n")
        with pytest.raises(ValueError, match="streak"):
            orc.run(n, kw.pop("mode", "full"), 0, 10, detector="streak", **kw)
