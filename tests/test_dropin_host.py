"""Host logic of the war_gpu.py drop-in (no GPU): the .progress file has the reference's keys
(gpu/war_gpu.py:331-338), a run interrupted and resumed gives the same totals, a resume with other
parameters is refused (gpu/war_gpu.py:297,306), --half doubles as the reference does
(gpu/war_gpu.py:346-352), and the printed lines keep the reference's format (:384-398).
The engine is replaced by a stand-in that returns synthetic counts derived from the range only."""
import json
import math
import re

import pytest

import war_gpu


class FakeEngine:
    def __init__(self, fail_after=None):
        self.calls = 0
        self.fail_after = fail_after

    def run_range(self, n, mode, lo, hi, hbins=0, lbins=0, dbins=0, shape_blk=0, **kw):
        self.calls += 1
        if self.fail_after is not None and self.calls > self.fail_after:
            raise KeyboardInterrupt("killed")
        ln = hi - lo
        done = ln // 2 if mode == "half" else ln
        hist = [0] * hbins
        if hbins:
            hist[min(7, hbins - 1)] = done // 3
        dh = [0] * dbins
        if dbins:
            dh[3] = 3 * (ln // 100)
        return {"cyc": done // 3, "done": done, "maxfin": (lo % 97) + 5, "turns": 10 * done, "hist": hist,
                "lenhist": [0] * lbins, "dhist": dh, "shape_nonstd": ln // 50 if shape_blk else 0,
                "shape_unchk": 0}


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(war_gpu, "SPAN_POSITIONS", 1000)
    e = FakeEngine()
    monkeypatch.setattr(war_gpu, "_engine", e)
    return e


def test_progress_keys_and_resume(fake, tmp_path, monkeypatch):
    out = str(tmp_path / "n9")
    full = war_gpu.run(9, out=out, quiet=True, hbins=16)
    with open(out + ".progress") as f:
        st = json.load(f)
    assert set(st) == {"n", "half", "canon", "per", "next_start", "total", "cyc", "done", "maxfin", "hist",
                       "shape", "rule", "dhist", "elapsed_s", "slowest_launch_s", "updated"}
    assert st["next_start"] == st["total"] == math.factorial(9)
    # interrupted after 100 calls, then resumed: same totals
    out2 = str(tmp_path / "n9b")
    monkeypatch.setattr(war_gpu, "_engine", FakeEngine(fail_after=100))
    monkeypatch.setattr(war_gpu.time, "time", iter(range(0, 10**9, 11)).__next__)   # progress written every call
    with pytest.raises(KeyboardInterrupt):
        war_gpu.run(9, out=out2, quiet=True, hbins=16)
    with open(out2 + ".progress") as f:
        mid = json.load(f)
    assert 0 < mid["next_start"] < mid["total"]
    monkeypatch.setattr(war_gpu, "_engine", FakeEngine())
    again = war_gpu.run(9, out=out2, resume=True, quiet=True, hbins=16)
    assert again[:3] == full[:3] and again[4] == full[4]


def test_resume_refuses_other_parameters(fake, tmp_path):
    out = str(tmp_path / "x")
    war_gpu.run(8, out=out, quiet=True)
    with pytest.raises(AssertionError):
        war_gpu.run(8, out=out, resume=True, quiet=True, per=128)
    with pytest.raises(AssertionError):
        war_gpu.run(8, out=out, resume=True, quiet=True, rule=1)


def test_half_doubles(fake):
    c, d, mx, el, h, sh, dd = war_gpu.run(8, half=True, hbins=16)
    assert d == math.factorial(8)          # played half, reported doubled
    assert sum(h.values()) == c


def test_canon_units_and_shape(fake, monkeypatch):
    calls = []
    orig = fake.run_range

    def spy(n, mode, lo, hi, **kw):
        calls.append((mode, lo, hi, kw.get("shape_blk")))
        return orig(n, mode, lo, hi, **kw)
    monkeypatch.setattr(fake, "run_range", spy)
    c, d, mx, el, h, sh, dd = war_gpu.run(6, canon=True, shape=True, dbins=4096)
    assert d == math.factorial(5) * 5
    assert calls[0][0] == "canon" and calls[-1][2] == math.factorial(5) * 5 and calls[0][3] == 2
    assert sh[2] == 2 and dd is not None
    monkeypatch.setenv("WAR_BLK_OVERRIDE", "3")
    assert war_gpu.run(6, canon=True, shape=True)[5][2] == 3


def test_cli_lines(fake, capsys):
    war_gpu.main(["7", "--nmin", "6", "--hist", "16"])
    lines = capsys.readouterr().out.splitlines()
    assert re.match(r"n=6  a\(n\)=\d+  of 720  longest finite=\d+ rounds  \d+\.\ds$", lines[0])
    assert re.match(r"   loop lengths \(rounds: deals\): 7: \d+$", lines[1])
    assert lines[2].startswith("n=7  a(n)=")


def test_canon_n1_plays_nothing(fake):
    c, d, mx, el, h, sh, dd = war_gpu.run(1, canon=True, shape=True, dbins=4096)
    assert (c, d, mx) == (0, 0, 0) and fake.calls == 0
