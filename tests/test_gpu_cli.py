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


LINE24 = re.compile(r"n=(\d+) det=(floyd|streak) range=\[(\d+),(\d+)\) ")
TIME24 = re.compile(r"time=[\d.]+s$")


@pytest.mark.parametrize("case", [0, 1, 2, 4, 5])
def test_war24_cli_matches_reference_log(eng, goldens, capsys, monkeypatch, tmp_path, case):
    """The war24.py drop-in prints the reference's result line for the same range and detector
    (data/long_runs.txt:62-64,74-75; the whole n = 17 run of :70 took the reference 6 days and
    is not repeated here). Only the elapsed time differs."""
    import war24
    monkeypatch.setattr(war24, "_HERE", str(tmp_path))
    c = goldens["cli24"][case]
    n, det, lo, hi = LINE24.match(c["line"]).groups()
    args = [n, "--det", det, "--start", lo, "--count", str(int(hi) - int(lo))]
    print("case war24", args, flush=True)
    capsys.readouterr()
    war24.main(args)
    got = capsys.readouterr().out.strip().splitlines()[-1]
    assert TIME24.sub("time=<t>s", got) == TIME24.sub("time=<t>s", c["line"]), (c["source"], got)


def test_resume_on_the_engine(eng, goldens, tmp_path, monkeypatch):
    """Resume and progress on the real engine (gpu/war_gpu.py:294-308, 329-341): the n = 13 canonical
    scan with distinct weights is interrupted after a few engine calls (the .progress file written
    after every call), resumed from its .progress file, and its totals equal the published ones and
    an uninterrupted run's."""
    monkeypatch.setattr(war_gpu, "SPAN_POSITIONS", 1 << 29)        # 11 calls of 12 * 2^29 / 12 positions
    clock = iter(range(0, 10 ** 9, 11))
    monkeypatch.setattr(war_gpu.time, "time", lambda: next(clock))  # progress written after every call
    out = str(tmp_path / "n13c")
    monkeypatch.setattr(war_gpu, "_engine", eng)
    calls = []
    real = eng.run_range

    def killed_after_3(*a, **kw):
        calls.append(a)
        if len(calls) > 3:
            raise KeyboardInterrupt("killed")
        return real(*a, **kw)
    monkeypatch.setattr(eng, "run_range", killed_after_3)
    with pytest.raises(KeyboardInterrupt):
        war_gpu.run(13, canon=True, out=out, quiet=True, hbins=128, dbins=4096)
    import json
    with open(out + ".progress") as f:
        mid = json.load(f)
    assert 0 < mid["next_start"] < mid["total"] and mid["cyc"] > 0
    monkeypatch.setattr(eng, "run_range", real)
    c, d, mx, el, h, sh, dd = war_gpu.run(13, canon=True, out=out, resume=True, quiet=True, hbins=128, dbins=4096)
    g = goldens["canon"]["13"]
    assert (c, d, mx) == (g["cycling"], g["positions"], goldens["canon_gpu"]["13"]["longest"])
    assert sum(v // k for k, v in dd.items()) == g["cycles"]
    assert sum(dd.values()) == g["on_cycle"]
    whole = war_gpu.run(13, canon=True, quiet=True, hbins=128, dbins=4096)
    assert (c, d, mx, h, dd) == (whole[0], whole[1], whole[2], whole[4], whole[6])
