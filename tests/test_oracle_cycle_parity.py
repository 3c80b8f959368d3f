"""The fact the flow kernel's repeat test rests on when it compares a game with its tortoise on every
other turn only: every War cycle is of even length, under either rule and for any n.

Over one cycle the state comes back, so player 1's card count does too; each turn moves one card
from the loser to the winner, so player 1 wins exactly as many of the cycle's turns as it loses, and
the cycle is twice that many turns long. A game therefore meets a tortoise taken on its cycle only an
even number of turns after the tortoise was taken.

* The lemma by brute force: every cycle met from the canonical positions of n <= 8, both rules,
  replayed by the set-based model (not the oracle): the length is even and player 1 wins exactly
  half of the cycle's turns.
* Every cycle length the reference publishes (deal and canonical histograms, distinct-cycle
  histograms, both rules where published) is even.
* The detection itself: Brent's loop with its repeat test only at an even number of turns after the
  tortoise (a model written here, independent of the kernel) returns the same round and cycle length
  as the oracle's Brent for every canonical position of n <= 8, both rules, and sampled deals of
  n = 9..12; testing at odd offsets instead misses or misreports cycles, so the check can fail.
"""
import itertools
import random
from collections import deque

import pytest

from test_oracle_model import _canon, model


def cycle_wins(p1, p2, rule, mu, lam):
    """Turns of its cycle that player 1 wins (set-based replay)."""
    a, b = deque(p1), deque(p2)
    wins = 0
    for t in range(mu + lam):
        x, y = a.popleft(), b.popleft()
        if x > y:
            a.extend((x, y) if rule == 0 else (y, x))
            wins += t >= mu
        else:
            b.extend((y, x) if rule == 0 else (x, y))
    return wins


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("n", [3, 4, 5, 6, 7, 8])
def test_every_cycle_is_even(n, rule):
    cycles = 0
    for q in itertools.permutations(range(n - 1)):
        for k in range(n - 1):
            p1, p2 = _canon(q, n, k)
            endless, mu, lam, _ = model(p1, p2, rule)
            if not endless:
                continue
            cycles += 1
            assert lam % 2 == 0, (p1, p2, rule, lam)
            assert 2 * cycle_wins(p1, p2, rule, mu, lam) == lam, (p1, p2, rule)
    if rule == 0:
        assert (cycles > 0) == (n in (5, 7))
    else:
        assert (cycles > 0) == (n in (3, 5, 6, 7))


def test_published_cycle_lengths_are_even(goldens):
    lengths = set()
    for d in goldens["deals"].values():
        lengths.update(int(x) for x in d.get("hist", {}))
    for d in goldens["distinct_minkey"].values():
        lengths.update(int(x) for x in d.get("distinct_hist", {}))
    lengths.update(int(x) for x in goldens["long"]["canon14"]["hist"])
    assert len(lengths) >= 10 and 112 in lengths and 96 in lengths
    assert all(x % 2 == 0 for x in lengths), sorted(lengths)


def brent_at_parity(p1, p2, rule, par):
    """Brent's loop (tortoise at rounds 2^k - 1) with the repeat test only when the number of turns
    since the tortoise was taken has parity `par`. Returns (endless, rounds, lam); rounds is None if
    no repeat is seen within 4096 turns."""
    a, b = deque(p1), deque(p2)
    t = (tuple(a), tuple(b))
    power, lam, r = 1, 0, 0
    while a and b and r < 4096:
        x, y = a.popleft(), b.popleft()
        if x > y:
            a.extend((x, y) if rule == 0 else (y, x))
        else:
            b.extend((y, x) if rule == 0 else (x, y))
        r += 1
        if not (a and b):
            return False, r, 0
        lam += 1
        if lam % 2 == par and (tuple(a), tuple(b)) == t:
            return True, r, lam
        if lam == power:
            t, power, lam = (tuple(a), tuple(b)), power * 2, 0
    return True, None, None


def _positions(n, rng):
    if n <= 8:
        for q in itertools.permutations(range(n - 1)):
            for k in range(n - 1):
                yield _canon(q, n, k)
    else:
        for _ in range(3000):
            d = list(range(n))
            rng.shuffle(d)
            yield tuple(d[0::2]), tuple(d[1::2])


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("n", [5, 6, 7, 8, 9, 10, 11, 12])
def test_even_offset_compares_detect_on_brents_round(orc, n, rule):
    rng = random.Random(1000 * n + rule)
    cycling = wrong_at_odd = 0
    for p1, p2 in _positions(n, rng):
        want = orc.play(n, list(p1), list(p2), rule)        # (endless, rounds, lam)
        got = brent_at_parity(p1, p2, rule, 0)
        assert got == (want[0], want[1], want[2] if want[0] else 0), (p1, p2, rule, got, want)
        if want[0]:
            cycling += 1
            wrong_at_odd += brent_at_parity(p1, p2, rule, 1) != got
    assert cycling > 0 or (rule == 0 and n in (6, 8, 12)) or (rule == 1 and n == 8)
    assert wrong_at_odd == cycling
