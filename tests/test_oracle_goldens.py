"""The CPU oracle against the reference's published results (tests/golden/goldens.json, transcribed
from the reference's data/ files). These pin the oracle; the GPU tests then compare the engine with it."""
import pytest


def hist_dict(h):
    return {str(i): v for i, v in enumerate(h) if v}


@pytest.mark.parametrize("n", list(range(1, 12)))
def test_full_deals_rule0(orc, goldens, n):
    g = goldens["deals"][str(n)]
    r = orc.run(n, "full", hbins=256)
    assert r["done"] == g["deals"]
    assert r["cyc"] == g["never_end"] == goldens["a"][str(n)]
    assert r["maxfin"] == g["longest"]
    assert hist_dict(r["hist"]) == g.get("hist", {})


def test_full_deals_n12(orc, goldens):
    r = orc.run(12, "full")
    assert (r["cyc"], r["maxfin"], r["done"]) == (0, 74, goldens["deals"]["12"]["deals"])


@pytest.mark.parametrize("n", list(range(1, 12)))
def test_full_deals_rule1(orc, goldens, n):
    g = goldens["rule1_deals"][str(n)]
    r = orc.run(n, "full", rule=1)
    assert (r["done"], r["cyc"], r["maxfin"]) == (g["deals"], g["never_end"], g["longest"])


@pytest.mark.parametrize("n", list(range(1, 12)))
def test_canon_rule0_weights_shape(orc, goldens, n):
    """war_gpu.py N --canon --distinct --shape (data/checks.txt:144-182)."""
    g = goldens["canon_gpu"][str(n)]
    blk = n & -n
    r = orc.run(n, "canon", dbins=4096 if n <= 15 else 0, distinct=n <= 15, shape_blk=blk)
    assert r["done"] == g["positions"]
    assert r["cyc"] == g["cycling"]
    assert r["maxfin"] == g["longest"]
    assert hist_dict(r["dhist"]) == g["weights"]
    assert sum(v // int(k) for k, v in hist_dict(r["dhist"]).items()) == g["distinct"]
    assert (r["shape_nonstd"], r["shape_unchk"]) == (g["shape"]["nonstd"], g["shape"]["unchk"])
    c = goldens["canon"][str(n)]
    assert (r["oncycle"], r["distinct_minkey"]) == (c["on_cycle"], c["cycles"])
    assert r["distinct_keys"] == c["cycles"]


@pytest.mark.parametrize("n", list(range(5, 12)))
def test_distinct_minkey(orc, goldens, n):
    """cpu/distinct_cycles.c output (data/checks.txt:202-221)."""
    g = goldens["distinct_minkey"][str(n)]
    r = orc.run(n, "canon", distinct=True, hbins=4096)
    assert (r["cyc"], r["oncycle"], r["distinct_minkey"]) == (g["cycling"], g["on_cycle"], g["distinct"])


@pytest.mark.parametrize("n", list(range(2, 12)))
def test_canon_rule1(orc, goldens, n):
    g = goldens["rule1_canon"][str(n)]
    r = orc.run(n, "canon", rule=1)
    assert (r["done"], r["cyc"], r["maxfin"]) == (g["positions"], g["cycling"], g["longest"])


def test_half_n10(orc, goldens):
    g = goldens["half"]["10"]
    r = orc.run(10, "half", hbins=256)
    assert r["done"] == g["played"]
    assert r["cyc"] == g["cycling_found"]
    assert 2 * r["cyc"] == goldens["a"]["10"]
    assert r["maxfin"] == g["longest"]


@pytest.mark.parametrize("n,blk,expect", [(10, 1, 287133), (9, 2, 153410)])
def test_shape_negative_control(orc, n, blk, expect):
    """A wrong run length rejects every cycle (data/checks.txt:197-198)."""
    r = orc.run(n, "canon", shape_blk=blk)
    assert r["cyc"] == expect and r["shape_nonstd"] == expect


def test_canonical_positions_totals(orc, goldens):
    for n in range(1, 12):
        g = goldens["canon"][str(n)]
        assert orc.total(n, "canon") == g["positions"]
