"""The drop-in war_gpu.py reproduces, line for line, the output the reference logged for the same
command lines (data/checks.txt, data/long_runs.txt; transcribed in tests/golden/goldens.json). Only
the elapsed-time field differs between machines and is masked."""
import re

import pytest

import war_gpu

pytestmark = pytest.mark.gpu

TIME = re.compile(r"  \d+\.\ds(  \(half\))?$")


def norm(line):
    return TIME.sub(lambda m: "  <t>s" + (m.group(1) or ""), line)


@pytest.mark.parametrize("case", range(7))
def test_cli_matches_reference_log(eng, goldens, capsys, case):
    c = goldens["cli"][case]
    print("case", c["args"], flush=True)
    capsys.readouterr()
    war_gpu.main(c["args"])
    got = [norm(ln) for ln in capsys.readouterr().out.splitlines() if ln.strip()]
    want = [norm(ln) for ln in c["lines"]]
    assert got == want, (c["source"], got, want)


@pytest.mark.parametrize("key", ["n17_100000000000000_100001000000000", "n17_355686642876416_355687428096000",
                                 "n17_0_8992587776"])
def test_n17_published_ranges(eng, goldens, key):
    """5-bit piles: the n = 17 rank ranges the reference logged (data/long_runs.txt:63-64,74-75)."""
    g = goldens["long"][key]
    r = eng.run_range(17, "full", g["lo"], g["hi"])
    assert r["done"] == g["hi"] - g["lo"]
    assert (r["cyc"], r["maxfin"]) == (g["cycling"], g["longest"])
