"""The oracle against the reference's own CPU programs, compiled from their sources into
oracle/_ref/ (make -C oracle ref). Skipped where they were not built (no reference tree at build time)."""
import os
import re
import subprocess

import pytest

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def _ref(prog, *args):
    path = os.path.join(REF, prog)
    if not os.path.exists(path):
        pytest.skip("reference program %s not built" % prog)
    return subprocess.run([path, *map(str, args)], capture_output=True, text=True, check=True).stdout


@pytest.mark.parametrize("n,mode", [(n, m) for n in range(2, 12) for m in (0, 1)])
def test_war_cpp(orc, n, mode):
    out = _ref("war", n, mode)
    m = re.search(r"deals=(\d+) cycling=(\d+) longest_finite=(\d+)", out)
    r = orc.run(n, "full", rule=mode)
    assert (r["done"], r["cyc"], r["maxfin"]) == tuple(map(int, m.groups()))


@pytest.mark.parametrize("n", list(range(5, 12)))
def test_distinct_cycles_c(orc, n):
    out = _ref("distinct_cycles", n, 8)
    m = re.search(r"never end (\d+); on a cycle (\d+); DISTINCT cycles up to swap (\d+)", out)
    r = orc.run(n, "canon", distinct=True)
    assert (r["cyc"], r["oncycle"], r["distinct_minkey"]) == tuple(map(int, m.groups()))
    hl = re.search(r"cycle lengths \(distinct cycles\):(.*)", out).group(1).split()
    # the lambda histogram of distinct cycles: count, per lam, the counted positions
    want = {int(a): int(b) for a, b in (t.split(":") for t in hl)}
    assert sum(want.values()) == r["distinct_minkey"]
