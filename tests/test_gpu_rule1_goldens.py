"""Rule 1 (the losing card first) over whole ranges against the reference's published values
(tests/golden/goldens.json, from data/loser_first.tsv): all 13! deals, all 14! deals by halves
(the half with card 13 to player 1, doubled, as `war_gpu.py 14 --rule 1 --half` counts them) and
all canonical positions of n = 12; the persistent kernel at its default schedule and, for n = 13,
the staged pipeline too."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stage_kind", [None, "refill"])
def test_rule1_n13_deals(eng, goldens, stage_kind):
    g = goldens["rule1_deals"]["13"]
    tun = {"stage_kind": stage_kind} if stage_kind else None
    r = eng.run_range(13, "full", rule=1, tuning=tun)
    assert r["done"] == g["deals"]
    assert (r["cyc"], r["maxfin"]) == (g["never_end"], g["longest"])


def test_rule1_n14_deals_by_halves(eng, goldens):
    g = goldens["rule1_deals"]["14"]
    r = eng.run_range(14, "half", rule=1)
    assert (2 * r["cyc"], r["maxfin"]) == (g["never_end"], g["longest"])


def test_rule1_n12_canonical(eng, goldens):
    g = goldens["rule1_canon"]["12"]
    r = eng.run_range(12, "canon", rule=1)
    assert r["done"] == g["positions"]
    assert (r["cyc"], r["maxfin"]) == (g["cycling"], g["longest"])
