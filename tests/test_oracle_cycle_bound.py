"""The bound the flow kernel's Brent relies on to leave out, under rule 0, the windows of the tortoise
at ages 1 and 3 (which see only cycles of length <= 2 and <= 4): no rule-0 War cycle is 4 turns long
or shorter (rule 1 has cycles of 2 and 4 turns at n = 3 and 5, and keeps every window).

* The lemma: over one cycle each player plays every card it holds, so the cycle is at least as long
  as the larger pile of every state on it, hence at least ceil(n/2): checked by brute force on every
  cycle met from the canonical positions of n <= 7 and on the published positions of n = 14
  (a set-based model, not the oracle).
* For n <= 8 (ceil(n/2) <= 4) the cycle lengths themselves: every canonical position of n = 2..8
  under rule 0 (every cycle passes through a state with card n-1 on top of a pile, which is a
  canonical position or one with the players swapped), through the oracle, whose histogram is
  pinned by the published values; the shortest is 6.
* Brent's closed form (round R* + lam, R* the smallest 2^k - 1 >= mu with 2^k >= lam): leaving out
  the windows k = 1, 2 changes R* only when lam <= 4, so the detection round of every position of
  n <= 7 under rule 0 is the same under both schedules (set-based model), and under rule 1 it is not
  (which is why rule 1 keeps them)."""
import itertools

import pytest

from test_oracle_model import _canon, model


def cycle_states(p1, p2, rule, mu, lam):
    """The lam states of the cycle a position enters after mu turns (set-based replay)."""
    from collections import deque
    a, b = deque(p1), deque(p2)
    out = []
    for t in range(mu + lam):
        if t >= mu:
            out.append((tuple(a), tuple(b)))
        x, y = a.popleft(), b.popleft()
        if x > y:
            a.extend((x, y) if rule == 0 else (y, x))
        else:
            b.extend((y, x) if rule == 0 else (x, y))
    return out


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("n", [5, 6, 7])
def test_cycle_at_least_the_larger_pile(n, rule):
    cycles = 0
    for q in itertools.permutations(range(n - 1)):
        for k in range(n - 1):
            p1, p2 = _canon(q, n, k)
            endless, mu, lam, _ = model(p1, p2, rule)
            if not endless:
                continue
            cycles += 1
            states = cycle_states(p1, p2, rule, mu, lam)
            assert lam >= max(max(len(a), len(b)) for a, b in states) >= (n + 1) // 2
    assert cycles > 0 or n == 6 or rule == 1


def test_published_n14_cycles_bound():
    # positions on cycles of 96 and 112 turns (figures/war_barcode.py:10-11, cycles/spivey_categories.py:33-35)
    for p1, p2 in [((13, 11, 10, 4), (8, 2, 12, 7, 5, 1, 14, 9, 6, 3)),
                   ((7, 2, 14, 9, 10, 3), (13, 6, 5, 1, 12, 8, 11, 4))]:
        p1 = tuple(c - 1 for c in p1); p2 = tuple(c - 1 for c in p2)
        endless, mu, lam, _ = model(p1, p2, 0)
        assert endless
        states = cycle_states(p1, p2, 0, mu, lam)
        assert lam >= max(max(len(a), len(b)) for a, b in states) >= 7


@pytest.mark.parametrize("rule", [0, 1])
def test_shortest_cycles_up_to_n8(orc, rule):
    shortest = {}
    for n in range(2, 9):
        r = orc.run(n, "canon", hbins=64, rule=rule)
        lens = [i for i, v in enumerate(r["hist"]) if v]
        if lens:
            shortest[n] = min(lens)
    if rule == 0:
        assert shortest == {5: 6, 7: 8}          # none of 4 turns or fewer
    else:
        assert shortest == {3: 2, 5: 4, 6: 12, 7: 6}


def _brent_round_from(mu, lam, kmin):
    R, power, k = 0, 1, 0
    while not (R >= mu and power >= lam and k >= kmin):
        R, power, k = 2 * R + 1, 2 * power, k + 1
    return R + lam


@pytest.mark.parametrize("n", [5, 7])
def test_skipping_windows_one_and_two_keeps_every_round(n):
    for q in itertools.permutations(range(n - 1)):
        for k in range(n - 1):
            endless, mu, lam, _ = model(*_canon(q, n, k), 0)
            if endless:
                assert _brent_round_from(mu, lam, 0) == _brent_round_from(mu, lam, 3)


def test_rule1_needs_the_early_windows():
    """Under rule 1 a 5-card position on a 4-turn cycle is seen by the age-3 window; without it,
    the detection round would move."""
    moved = 0
    for q in itertools.permutations(range(4)):
        for k in range(4):
            endless, mu, lam, _ = model(*_canon(q, 5, k), 1)
            if endless and lam <= 4:
                moved += _brent_round_from(mu, lam, 0) != _brent_round_from(mu, lam, 3)
    assert moved > 0
