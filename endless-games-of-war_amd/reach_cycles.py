"""Is every War cycle entered by an ordinary deal? Drop-in for the reference's cpu/reach_cycles.c
command line and output; runs on the MI355X through libwar_engine.so (war_reach).

  python reach_cycles.py 13 12            n, threads (accepted; the GPU does the work)
  python reach_cycles.py 14 12 1000000    optional cap on positions visited per cycle

Every canonical position of n cards (player 1's top card is the highest) is played; each
distinct cycle, up to swapping the players, is taken once at its position of minimum key, and the
positions leading into it are searched backwards for one of deal size (ceil(n/2) and floor(n/2)
cards, either way round).

Reference behaviour followed (cpu/reach_cycles.c): arguments :73, output lines :103-105.
"""
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from war_amd import Engine  # noqa: E402

DEFAULT_CAP = 200000000
_engine = None


def engine(device=0):
    global _engine
    if _engine is None:
        _engine = Engine(device)
    return _engine


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    if len(argv) < 2:
        print("usage: reach_cycles n threads [cap]", file=sys.stderr)
        return 1
    n = int(argv[0])
    cap = int(argv[2]) if len(argv) > 2 else DEFAULT_CAP
    r = engine().reach(n, cap=cap)
    print("n=%d: distinct cycles up to swap %d | reached by a deal %d | NOT reached %d | undetermined (cap %d) %d"
          % (n, r["classes"], r["reached"], r["not_reached"], cap, r["undetermined"]), flush=True)
    if r["not_reached"]:
        print("   loop lengths of cycles not reached:" +
              "".join(" %d:%d" % (L, c) for L, c in enumerate(r["nolen"]) if c), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
