"""Host logic of the cycle_length.py and reach_cycles.py drop-ins (no GPU): the reference's argument
forms and error messages (cpu/cycle_length.c:90-111, cpu/reach_cycles.c:73), the doubled-family
builder, and the output lines (cpu/cycle_length.c:112-124, cpu/reach_cycles.c:103-105). The engine
is replaced by a stand-in."""
import pytest

import cycle_length as CL
import reach_cycles as RC


class FakeEngine:
    def __init__(self, info=None, reach=None):
        self.info = info
        self.reach_res = reach
        self.calls = []

    def cycle_info(self, positions, rule=0, cap=10**12):
        self.calls.append((positions, rule, cap))
        return [dict(self.info) for _ in positions]

    def reach(self, n, cap=200000000):
        self.calls.append((n, cap))
        return dict(self.reach_res)


def run(monkeypatch, capsys, mod, argv, **kw):
    monkeypatch.setattr(mod, "_engine", FakeEngine(**kw))
    rc = mod.main(argv)
    out = capsys.readouterr()
    return rc, out.out.splitlines(), out.err.strip(), mod._engine


def test_doubled_builder():
    p1, p2 = CL.doubled(0, 1)
    n = len(p1) + len(p2)
    assert (len(p1), len(p2)) == (4, 10) and sorted(p1 + p2) == list(range(1, n + 1))
    # letters A B B C | D E A B B C C D D E: A is strongest (highest), values fall in order of appearance
    assert p1 == [14, 12, 11, 8] and p2 == [5, 2, 13, 10, 9, 7, 6, 4, 3, 1]
    assert CL.doubled(4, 1) is not None and CL.doubled(4, 2) is None and CL.doubled(-1, 1) is None


def test_cycle_lines(monkeypatch, capsys):
    rc, out, err, e = run(monkeypatch, capsys, CL, ["doubled", "0", "1"],
                          info={"status": "cycles", "turns": 223, "lam": 96, "mu": 0})
    assert rc == 0 and out == ["doubled family j=0 k=1", "n=14: player 1 holds 4 cards, player 2 holds 10",
                               "the game cycles: cycle length 96, entered after 0 turns",
                               "n + 2^v = 16 divides the length: 96 = 16 x 6"]
    assert e.calls[0][2] == 10**12
    rc, out, err, e = run(monkeypatch, capsys, CL, ["1,3", "2,4,5", "777"],
                          info={"status": "cycles", "turns": 9, "lam": 7, "mu": 2})
    assert out[1:] == ["the game cycles: cycle length 7, entered after 2 turns",
                       "n + 2^v = 6 does NOT divide the length"] and e.calls[0][2] == 777
    rc, out, err, e = run(monkeypatch, capsys, CL, ["1,3", "2,4,5"], info={"status": "ends", "turns": 12})
    assert out == ["n=5: player 1 holds 2 cards, player 2 holds 3", "the game ends after 12 turns"]
    rc, out, err, e = run(monkeypatch, capsys, CL, ["1,3", "2,4,5", "50"], info={"status": "unsettled", "turns": 50})
    assert out[-1] == "not settled within 50 turns"


@pytest.mark.parametrize("argv,msg", [
    (["doubled", "0", "0"], "doubled j k: need j >= 0, k >= 1 and at most 255 cards"),
    (["doubled", "5", "3"], "doubled j k: need j >= 0, k >= 1 and at most 255 cards"),
    (["1,x", "2"], "each pile: comma-separated card values 1..255, top card first"),
    (["1,0", "2"], "each pile: comma-separated card values 1..255, top card first"),
    (["1,3", "2,3"], "the cards must be 1..4, each once"),
    (["1"], "usage: cycle_length P1 P2 [cap]   or   cycle_length doubled j k [cap]"),
])
def test_cycle_errors(monkeypatch, capsys, argv, msg):
    rc, out, err, e = run(monkeypatch, capsys, CL, argv, info={"status": "ends", "turns": 1})
    assert rc == 1 and err == msg and out == [] and e.calls == []


def test_reach_lines(monkeypatch, capsys):
    res = {"classes": 52, "reached": 52, "not_reached": 0, "undetermined": 0, "nolen": [0] * 4096}
    rc, out, err, e = run(monkeypatch, capsys, RC, ["7", "4"], reach=res)
    assert out == ["n=7: distinct cycles up to swap 52 | reached by a deal 52 | NOT reached 0 | "
                   "undetermined (cap 200000000) 0"]
    nl = [0] * 4096
    nl[28] = 3
    nl[44] = 1
    res = {"classes": 9, "reached": 5, "not_reached": 4, "undetermined": 0, "nolen": nl}
    rc, out, err, e = run(monkeypatch, capsys, RC, ["9", "4", "1000"], reach=res)
    assert out == ["n=9: distinct cycles up to swap 9 | reached by a deal 5 | NOT reached 4 | undetermined (cap 1000) 0",
                   "   loop lengths of cycles not reached: 28:3 44:1"]
    assert e.calls == [(9, 1000)]
