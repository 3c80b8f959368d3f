"""The oracle's single-position cycle tool (cpu/cycle_length.c) and cycle reachability
(cpu/reach_cycles.c) against what pins them:

* cycle_info: the set-based model (exact mu and lam) on random positions of up to 10 cards, both
  rules; the published doubled-family lines (data/checks.txt) for every member the CPU settles in
  a moment, with the piles built by the cycle_length.py drop-in; the reference's own
  cycle_length binary (compiled from its source into oracle/_ref) on random positions of up to
  120 cards, where it is built;
* reach: the published counts for n = 5..12 (data/checks.txt:259-266); the reference's own
  reach_cycles binary for n <= 11, where it is built.
"""
import os
import random
import re
import subprocess

import pytest

import cycle_length as CL
from test_oracle_model import model

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


@pytest.mark.parametrize("rule", [0, 1])
def test_cycle_info_against_model(orc, rule):
    rng = random.Random(5 + rule)
    for _ in range(400):
        n = rng.randint(2, 10)
        cards = list(range(n))
        rng.shuffle(cards)
        k = rng.randint(1, n - 1)
        p1, p2 = cards[:k], cards[k:]
        endless, mu, lam, length = model(p1, p2, rule)
        got = orc.cycle_info(p1, p2, rule=rule)
        if endless:
            assert (got["status"], got["lam"], got["mu"]) == ("cycles", lam, mu), (p1, p2)
        else:
            assert (got["status"], got["turns"]) == ("ends", length), (p1, p2)


def test_cycle_info_cap(orc):
    p1, p2 = CL.doubled(0, 1)
    assert orc.cycle_info(p1, p2, cap=50) == {"status": "unsettled", "turns": 50, "lam": 0, "mu": 0}


def test_doubled_family_published(orc, goldens):
    done = 0
    for c in goldens["cycle_length"]:
        lam = int(re.search(r"cycle length (\d+)", c["lines"][2]).group(1))
        if lam > 2_000_000:
            continue
        j, k = int(c["args"][1]), int(c["args"][2])
        p1, p2 = CL.doubled(j, k)
        got = orc.cycle_info(p1, p2)
        lines = ["doubled family j=%d k=%d" % (j, k)] + CL.lines_for(p1, p2, CL.DEFAULT_CAP, got)
        assert lines == c["lines"], c["source"]
        done += 1
    assert done >= 40


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "cycle_length")), reason="oracle/_ref not built")
def test_cycle_info_against_reference_binary(orc):
    rng = random.Random(42)
    for _ in range(60):
        n = rng.randint(4, 120)
        cards = list(range(1, n + 1))
        rng.shuffle(cards)
        k = rng.randint(1, n - 1)
        p1, p2 = cards[:k], cards[k:]
        out = subprocess.run([os.path.join(REF, "cycle_length"), ",".join(map(str, p1)), ",".join(map(str, p2)),
                              "20000000"], capture_output=True, text=True, check=True).stdout.splitlines()
        got = orc.cycle_info(p1, p2, cap=20000000)
        assert CL.lines_for(p1, p2, 20000000, got) == out[:len(CL.lines_for(p1, p2, 20000000, got))], (p1, p2)


@pytest.mark.parametrize("n", range(5, 13))
def test_reach_published(orc, goldens, n):
    r = orc.reach(n, threads=8)
    m = re.match(r"n=\d+: distinct cycles up to swap (\d+) \| reached by a deal (\d+) \| NOT reached (\d+) \| "
                 r"undetermined \(cap \d+\) (\d+)", goldens["reach"][str(n)]["line"])
    assert (r["classes"], r["reached"], r["not_reached"], r["undetermined"]) == tuple(map(int, m.groups()))
    if str(n) in goldens["distinct_minkey"]:
        assert r["classes"] == goldens["distinct_minkey"][str(n)]["distinct"]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "reach_cycles")), reason="oracle/_ref not built")
@pytest.mark.parametrize("n", [5, 7, 9, 11])
def test_reach_against_reference_binary(orc, n):
    out = subprocess.run([os.path.join(REF, "reach_cycles"), str(n), "4", "1000"], capture_output=True, text=True,
                         check=True).stdout.splitlines()
    r = orc.reach(n, threads=4, cap=1000)
    assert out[0] == ("n=%d: distinct cycles up to swap %d | reached by a deal %d | NOT reached %d | "
                      "undetermined (cap 1000) %d" % (n, r["classes"], r["reached"], r["not_reached"],
                                                       r["undetermined"]))
