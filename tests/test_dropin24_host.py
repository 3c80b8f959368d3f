"""Host logic of the war24.py drop-in (no GPU): the command line and its checks
(gpu/war24.py:159-173), the checkpoint / progress file names and keys (:177-181, :193-198), resume
from the checkpoint, the result line (:203-205) and the exit status on a deal-count mismatch
(:206). The engine is replaced by a stand-in that returns synthetic counts derived from the range."""
import json
import math
import re

import pytest

import war24


class FakeEngine:
    def __init__(self, fail_after=None, short=0):
        self.calls = []
        self.fail_after = fail_after
        self.short = short

    def run_range(self, n, mode, lo, hi, kernel="persistent", detector="brent", **kw):
        self.calls.append((n, mode, lo, hi, kernel, detector))
        if self.fail_after is not None and len(self.calls) > self.fail_after:
            raise KeyboardInterrupt("killed")
        ln = hi - lo
        return {"cyc": ln // 3, "done": ln - self.short, "maxfin": (lo % 89) + 3, "undetermined": 0}


@pytest.fixture
def fake(monkeypatch, tmp_path):
    e = FakeEngine()
    monkeypatch.setattr(war24, "_engine", e)
    monkeypatch.setattr(war24, "_HERE", str(tmp_path))
    return e


def _line_re():
    return re.compile(r"n=(\d+) det=(floyd|streak) range=\[(\d+),(\d+)\) deals_simulated=(\d+) \(expected (\d+)\) "
                      r"cycling=(\d+) longest_finite=(\d+) undetermined=(\d+) time=[\d.]+s$")


def test_line_files_and_detectors(fake, tmp_path, capsys):
    lo, cnt = 12345, 10 ** 6
    war24.main(["17", "--start", str(lo), "--count", str(cnt), "--det", "streak"])
    line = capsys.readouterr().out.strip().splitlines()[-1]
    m = _line_re().match(line)
    assert m, line
    assert int(m.group(5)) == int(m.group(6)) == cnt
    # the streak test runs in the plain kernel; floyd maps to the exact Brent pipeline
    assert {c[4:] for c in fake.calls} == {("plain", "streak")}
    prog = tmp_path / ("war24_n17_streak_%d_%d.progress" % (lo, lo + cnt))
    ckpt = tmp_path / ("war24_n17_streak_%d_%d.ckpt" % (lo, lo + cnt))
    st = json.loads(ckpt.read_text())
    assert set(st) == {"pos", "cyc", "max", "deals", "und"} and st["pos"] == lo + cnt and st["deals"] == cnt
    assert prog.read_text().splitlines()[-1] == line
    fake.calls.clear()
    war24.main(["13", "--det", "floyd", "--count", "5000"])
    assert {c[4:] for c in fake.calls} == {("persistent", "brent")}
    assert capsys.readouterr().out.startswith("n=13 det=floyd range=[0,5000) ")


def test_spans_cover_the_range_once(fake, monkeypatch):
    monkeypatch.setattr(war24, "SPAN_DEALS", 777)
    war24.main(["11", "--start", "100", "--count", "10000"])
    spans = [(c[2], c[3]) for c in fake.calls]
    assert spans[0][0] == 100 and spans[-1][1] == 10100
    assert all(a[1] == b[0] for a, b in zip(spans, spans[1:]))


def test_count_zero_means_to_the_end(fake, capsys):
    war24.main(["9", "--start", "362000", "--det", "floyd"])
    assert "range=[362000,%d)" % math.factorial(9) in capsys.readouterr().out


def test_argument_checks(fake):
    with pytest.raises(SystemExit, match="odd n"):
        war24.main(["14", "--det", "streak"])
    with pytest.raises(AssertionError):
        war24.main(["21"])
    with pytest.raises(AssertionError):
        war24.main(["9", "--threads", "1000"])
    war24.main(["14", "--det", "floyd", "--count", "10"])       # even n is fine with floyd


def test_resume_from_checkpoint(fake, tmp_path, monkeypatch, capsys):
    monkeypatch.setattr(war24, "SPAN_DEALS", 1000)
    monkeypatch.setattr(war24.time, "time", iter(range(0, 10 ** 9, 11)).__next__)   # checkpoint every span
    war24.main(["11", "--count", "50000", "--tag", "_a"])
    full = capsys.readouterr().out
    monkeypatch.setattr(war24, "_engine", FakeEngine(fail_after=20))
    with pytest.raises(KeyboardInterrupt):
        war24.main(["11", "--count", "50000", "--tag", "_b"])
    ck = json.loads((tmp_path / "war24_n11_streak_0_50000_b.ckpt").read_text())
    assert 0 < ck["pos"] < 50000
    e = FakeEngine()
    monkeypatch.setattr(war24, "_engine", e)
    war24.main(["11", "--count", "50000", "--tag", "_b", "--resume"])
    assert e.calls[0][2] == ck["pos"]                             # picks up where it stopped
    again = capsys.readouterr().out
    strip = lambda s: re.sub(r"time=[\d.]+s", "", s.strip())
    assert strip(again) == strip(full)


def test_mismatch_exits_1(monkeypatch, tmp_path, capsys):
    monkeypatch.setattr(war24, "_engine", FakeEngine(short=1))
    monkeypatch.setattr(war24, "_HERE", str(tmp_path))
    with pytest.raises(SystemExit) as ex:
        war24.main(["9", "--count", "100"])
    assert ex.value.code == 1
    assert capsys.readouterr().out.strip().endswith("ERROR: deal count mismatch")


class UndetEngine(FakeEngine):
    """Reports a few undetermined deals per call and records the step cap it was asked for."""

    def run_range(self, n, mode, lo, hi, kernel="persistent", detector="brent", **kw):
        r = super().run_range(n, mode, lo, hi, kernel, detector)
        self.caps = getattr(self, "caps", []) + [kw.get("step_cap")]
        r["undetermined"] = 7
        return r


def test_undetermined_is_counted_and_printed(monkeypatch, tmp_path, capsys):
    """Games unresolved within the reference's step cap (2^22, gpu/war24.py:57) are a count, summed
    over the calls into the checkpoint's `und` and printed as undetermined=N (:127, :203-205)."""
    e = UndetEngine()
    monkeypatch.setattr(war24, "_engine", e)
    monkeypatch.setattr(war24, "_HERE", str(tmp_path))
    monkeypatch.setattr(war24, "SPAN_DEALS", 1000)
    war24.main(["17", "--start", "0", "--count", "3500", "--det", "floyd"])
    line = capsys.readouterr().out.strip().splitlines()[-1]
    m = _line_re().match(line)
    assert m, line
    assert int(m.group(9)) == 7 * 4                      # four calls of at most 1000 deals
    assert set(e.caps) == {1 << 22}
    st = json.loads((tmp_path / "war24_n17_floyd_0_3500.ckpt").read_text())
    assert st["und"] == 28
