"""The cycle a War position ends in: cycle length, turns before it, and whether n + 2^v divides the
length (2^v = n & -n). Drop-in for the reference's cpu/cycle_length.c command line and output;
the game runs on the MI355X through libwar_engine.so (war_cycle_info_batch), one GPU lane per
position, byte cards in LDS rings, so positions of up to 255 cards are fine.

  python cycle_length.py 1,13,11,9,4 3,5,2,14,12,10,8,7,6      two piles, top card first
  python cycle_length.py doubled 0 1                           the doubled family, j >= 0, k >= 1
  python cycle_length.py doubled 2 3 100000000                 optional last argument: turn cap

The doubled family (as the reference's header describes it): T = D^j(ABBC), Z = D^j(DE),
Y = D^j(CDDE), where D replaces each letter x by x followed by the next letter; player 1 gets T^k
and player 2 gets Z T^k Y. Cards are numbered so that every letter beats every later letter: the
A cards get the highest values, and within a letter the values decrease in order of appearance,
player 1's pile first.

Reference behaviour followed (cpu/cycle_length.c): argument forms and messages :90-106, card
check :107-111, output lines :112,116-117,120,122-124.
"""
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from war_amd import Engine  # noqa: E402

DEFAULT_CAP = 10 ** 12
_engine = None


def engine(device=0):
    global _engine
    if _engine is None:
        _engine = Engine(device)
    return _engine


def parse_pile(arg):
    """Comma-separated card values 1..255 (at most 255 of them); None if malformed."""
    if not arg:
        return None
    out = []
    for part in arg.split(","):
        if not part.isdigit():
            return None
        v = int(part)
        if v < 1 or v > 255 or len(out) >= 255:
            return None
        out.append(v)
    return out


def _double(word, j):
    for _ in range(j):
        word = [y for x in word for y in (x, x + 1)]
    return word


def doubled(j, k):
    """Player 1 and player 2 piles of the doubled family member (j, k), or None if it has more
    than 255 cards (letters are 0 = A, 1 = B, ...)."""
    if j < 0 or k < 1:
        return None
    T, Z, Y = _double([0, 1, 1, 2], j), _double([3, 4], j), _double([2, 3, 3, 4], j)
    L1 = T * k
    L2 = Z + T * k + Y
    n = len(L1) + len(L2)
    if n > 255 or max(len(T), len(Z), len(Y)) > 256:
        return None
    counts = {}
    for x in L1 + L2:
        counts[x] = counts.get(x, 0) + 1
    nxt, v = {}, n
    for x in sorted(counts):                 # letter 0 is the strongest: the highest values
        nxt[x] = v
        v -= counts[x]
    p1, p2 = [], []
    for pile, src in ((p1, L1), (p2, L2)):
        for x in src:
            pile.append(nxt[x])
            nxt[x] -= 1
    return p1, p2


def lines_for(p1, p2, cap, info):
    n = len(p1) + len(p2)
    out = ["n=%d: player 1 holds %d cards, player 2 holds %d" % (n, len(p1), len(p2))]
    if info["status"] == "ends":
        out.append("the game ends after %d turns" % info["turns"])
        return out
    if info["status"] == "unsettled":
        out.append("not settled within %d turns" % cap)
        return out
    lam, mu = info["lam"], info["mu"]
    v = n & -n
    out.append("the game cycles: cycle length %d, entered after %d turns" % (lam, mu))
    if lam % (n + v) == 0:
        out.append("n + 2^v = %d divides the length: %d = %d x %d" % (n + v, lam, n + v, lam // (n + v)))
    else:
        out.append("n + 2^v = %d does NOT divide the length" % (n + v))
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    head = []
    if len(argv) >= 3 and argv[0] == "doubled":
        try:
            j, k = int(argv[1]), int(argv[2])
        except ValueError:
            j, k = -1, 0
        piles = doubled(j, k)
        if piles is None:
            print("doubled j k: need j >= 0, k >= 1 and at most 255 cards", file=sys.stderr)
            return 1
        p1, p2 = piles
        head.append("doubled family j=%d k=%d" % (j, k))
        rest = argv[3:]
    elif len(argv) >= 2:
        p1, p2 = parse_pile(argv[0]), parse_pile(argv[1])
        if not p1 or not p2:
            print("each pile: comma-separated card values 1..255, top card first", file=sys.stderr)
            return 1
        rest = argv[2:]
    else:
        print("usage: cycle_length P1 P2 [cap]   or   cycle_length doubled j k [cap]", file=sys.stderr)
        return 1
    cap = int(rest[0]) if rest else DEFAULT_CAP
    n = len(p1) + len(p2)
    if n > 255:
        print("at most 255 cards", file=sys.stderr)
        return 1
    if sorted(p1 + p2) != list(range(1, n + 1)):
        print("the cards must be 1..%d, each once" % n, file=sys.stderr)
        return 1
    for ln in head:
        print(ln, flush=True)
    info = engine().cycle_info([(p1, p2)], rule=0, cap=cap)[0]
    for ln in lines_for(p1, p2, cap, info):
        print(ln, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
