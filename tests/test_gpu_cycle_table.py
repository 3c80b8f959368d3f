"""The cycle walks' state table (WAR_FLAG_CYCLE_TABLE): with the table forced on, every output of a
canonical request with distinct / dbins / shape equals the oracle's and the same request with the
table off; whole ranges against the published distinct-cycle counts; the default (on for ranges of
at least 2^28 positions) against the published n = 13 canonical totals."""
import random

import pytest

pytestmark = pytest.mark.gpu

WALK_FIELDS = ("cyc", "done", "maxfin", "turns", "hist", "oncycle", "distinct_minkey", "dhist", "shape_nonstd",
               "shape_unchk")


def run_both(eng, n, lo, hi, **kw):
    on = eng.run_range(n, "canon", lo, hi, cycle_table=True, **kw)
    off = eng.run_range(n, "canon", lo, hi, cycle_table=False, **kw)
    for k in WALK_FIELDS + ("distinct_hashset",):
        assert on[k] == off[k], (n, lo, hi, k, on[k], off[k])
    return on


@pytest.mark.parametrize("n", [5, 7, 9, 10, 11])
def test_whole_range_table_on(eng, orc, goldens, n):
    kw = dict(hbins=256, dbins=4096, distinct=True, shape_blk=n & -n)
    got = run_both(eng, n, 0, None, **kw)
    want = orc.run(n, "canon", **kw)
    for k in WALK_FIELDS:
        assert got[k] == want[k], (n, k, got[k], want[k])
    assert got["distinct_hashset"] == want["distinct_keys"] == goldens["canon"][str(n)]["cycles"]


@pytest.mark.parametrize("n", [11, 13, 14])
def test_partial_ranges_table_on(eng, orc, n):
    """Ranges small enough for the oracle, the table forced on; several ranges of one n in a row
    each start from an empty table (a request's table is its own)."""
    rng = random.Random(100 + n)
    total = orc.total(n, "canon")
    for shape in (0, n & -n):
        lo = rng.randrange(0, total - 2_000_000)
        kw = dict(hbins=256, dbins=256, distinct=True, shape_blk=shape)
        got = run_both(eng, n, lo, lo + 2_000_000, **kw)
        want = orc.run(n, "canon", lo, lo + 2_000_000, **kw)
        for k in WALK_FIELDS:
            assert got[k] == want[k], (n, lo, shape, k, got[k], want[k])
        assert got["distinct_hashset"] == want["distinct_keys"]


def test_shape_only_and_dbins_only(eng, orc):
    """A table kept for the shape test alone (no distinct), and for dbins alone."""
    for kw in (dict(hbins=64, shape_blk=1), dict(hbins=64, dbins=64)):
        got = run_both(eng, 13, 10_000_000, 12_000_000, **kw)
        want = orc.run(13, "canon", 10_000_000, 12_000_000, **kw)
        for k in ("cyc", "done", "maxfin", "turns", "hist", "oncycle", "dhist", "shape_nonstd", "shape_unchk"):
            assert got[k] == want[k], (kw, k, got[k], want[k])


def test_n13_whole_range_default(eng, goldens):
    """All 5 748 019 200 canonical positions of n = 13 (the table on by default at this size)
    against the published totals, and against the same run with the table off."""
    kw = dict(hbins=256, dbins=4096, distinct=True, shape_blk=1)
    on = eng.run_range(13, "canon", **kw)
    off = eng.run_range(13, "canon", cycle_table=False, **kw)
    for k in WALK_FIELDS + ("distinct_hashset",):
        assert on[k] == off[k], (k, on[k], off[k])
    c = goldens["canon"]["13"]
    assert on["cyc"] == c["cycling"] and on["oncycle"] == c["on_cycle"]
    assert on["distinct_hashset"] == on["distinct_minkey"] == c["cycles"]
    assert on["shape_nonstd"] == 0 and on["shape_unchk"] == 0
