"""The engine's tools around the enumeration, through the C ABI on the GPU: cycle length and
pre-period of given positions (war_cycle_info_batch, cpu/cycle_length.c) against the oracle and
the published doubled-family lines, and cycle reachability (war_reach, cpu/reach_cycles.c) against
the published counts for n = 5..14 and the oracle."""
import random
import re

import pytest

import cycle_length as CL
import reach_cycles as RC
from war_amd import WarError

pytestmark = pytest.mark.gpu


def test_cycle_info_random_positions(eng, orc):
    rng = random.Random(255)
    pos = []
    for i in range(300):
        n = rng.choice([rng.randint(2, 20), rng.randint(21, 255)])
        cards = list(range(1, n + 1))
        rng.shuffle(cards)
        k = rng.randint(1, n - 1)
        pos.append((cards[:k], cards[k:]))
    for rule in (0, 1):
        print("case cycle_info rule", rule, flush=True)
        got = eng.cycle_info(pos, rule=rule, cap=3_000_000)
        for (p1, p2), g in zip(pos, got):
            assert g == orc.cycle_info(p1, p2, rule=rule, cap=3_000_000), (rule, p1, p2)


def test_cycle_info_cap_and_errors(eng):
    p1, p2 = CL.doubled(0, 1)
    assert eng.cycle_info([(p1, p2)], cap=50)[0] == {"status": "unsettled", "turns": 50, "lam": 0, "mu": 0}
    for bad in ([([], [1])], [([1, 2], [2])], [(list(range(1, 200)), list(range(200, 257)))]):
        with pytest.raises((WarError, OverflowError, TypeError)):
            eng.cycle_info(bad)


def test_doubled_family_published(eng, goldens, capsys):
    """Every doubled-family member of the published list whose cycle is at most 2e6 turns, through
    the cycle_length.py drop-in, line for line (data/checks.txt:795-1022)."""
    n_done = 0
    for c in goldens["cycle_length"]:
        lam = int(re.search(r"cycle length (\d+)", c["lines"][2]).group(1))
        if lam > 2_000_000:
            continue
        print("case doubled", c["args"], flush=True)
        capsys.readouterr()
        assert CL.main(c["args"]) == 0
        assert capsys.readouterr().out.splitlines() == c["lines"], c["source"]
        n_done += 1
    assert n_done >= 40


@pytest.mark.parametrize("n", range(5, 15))
def test_reach_published(eng, goldens, capsys, n):
    print("case reach", n, flush=True)
    capsys.readouterr()
    assert RC.main([str(n), "12"]) == 0
    assert capsys.readouterr().out.splitlines() == [goldens["reach"][str(n)]["line"]]


def test_reach_ranges_add_up(eng, orc):
    whole = eng.reach(11)
    assert whole["classes"] == orc.reach(11)["classes"] == 55592
    total = orc.total(11, "canon")
    cuts = [0, total // 7, total // 3, total - 12345, total]
    parts = [eng.reach(11, lo, hi) for lo, hi in zip(cuts, cuts[1:])]
    for k in ("classes", "reached", "not_reached", "undetermined"):
        assert sum(p[k] for p in parts) == whole[k]
    with pytest.raises(WarError):
        eng.reach(16)
