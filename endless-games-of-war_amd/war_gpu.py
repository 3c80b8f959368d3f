"""Count the deals of single-suit War that never end, on an MI355X (n <= 16; 17..20 with 5-bit piles).

Drop-in for the reference's gpu/war_gpu.py: the same command line, the same run() signature and
return tuple, the same output lines and the same .progress file. The counting runs in
libwar_engine.so (hand-written HIP for gfx950) through its C ABI; nothing is computed on the CPU.

Cards are 0..n-1 (0 lowest). Deck positions 0, 2, 4, ... go to player 1 and 1, 3, 5, ... to player
2, top card first. Each turn the higher top card wins, and the winner puts the winning card and
then the losing card under its own pile (--rule 1 reverses the order).

The command line is the reference's; for instance `war_gpu.py 15 --hist 128` prints a(15) with its
cycle-length histogram, and `--canon --out PREFIX` scans the canonical positions with a progress
file that `--resume` continues from.

Reference behaviour followed (gpu/war_gpu.py): run() :272-352 (progress keys :331-338, resume
asserts :297,306, --half doubling :346-352), CLI :355-398 (asserts :380-390, output :384-398).
Extra flags: --kernel plain|persistent (default persistent), --device D, --lbins B,
--det auto|brent|alt|pair (the cycle test; every printed number is the same either way).
"""
import argparse
import json
import math
import os
import sys
import time

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from war_amd import Engine  # noqa: E402

_engine = None
SPAN_POSITIONS = 1 << 30        # positions per engine call (between progress checks)


def engine(device=0):
    global _engine
    if _engine is None:
        _engine = Engine(device)
    return _engine


def detector_for(n, rule, kernel, det="auto"):
    """The cycle test a run uses: "auto" takes the alternation-checkpoint detector where it applies
    (odd n, rule 0, persistent kernel) and Brent elsewhere. "pair" (n = 2 mod 4) is there on request:
    it plays about 8 % fewer turns than Brent at n = 14 but is no faster on the flow kernel, whose
    scalar unit its per-turn run tracking loads. All are exact: every output of run() is the same;
    only the engine's own turn count differs."""
    if det != "auto":
        return det
    return "alt" if (n % 2 == 1 and rule == 0 and kernel == "persistent") else "brent"


def run(n, half=False, per=256, threads=256, batch_threads=1 << 18, out=None, resume=False, quiet=False, hbins=0,
        canon=False, shape=False, rule=0, dbins=0, kernel="persistent", lbins=0, device=0, det="auto"):
    """Same signature and return tuple as the reference's run() (gpu/war_gpu.py:272-352):
    (cycling, played, longest_finite, elapsed_s, hist{len: count}, shape (nonstd, unchecked, blk) or None,
    dhist{weight: count} or None). threads / batch_threads only shape the reference's launches and
    are accepted for compatibility. det: the cycle test (detector_for)."""
    eng = engine(device)
    detector = detector_for(n, rule, kernel, det)
    # units of `start`: deal ranks, or (canon) orders of the n-1 cards below the top one, n-1 splits each
    total = math.factorial(n - 1) if canon else math.factorial(n)
    unit = max(1, n - 1) if canon else 1               # positions per unit of `start`
    if not canon:
        dbins = 0
    blk = (n & -n) if (canon and shape) else 0          # the run length Theorem 3 predicts: lowest set bit of n
    if blk and os.environ.get("WAR_BLK_OVERRIDE"):      # the reference's check that the test can fail
        blk = int(os.environ["WAR_BLK_OVERRIDE"])
    cyc = done = maxfin = 0
    hist = [0] * max(hbins, 1)
    shp = [0, 0]
    dh = [0] * max(dbins, 1)
    start = 0
    prog = (out + ".progress") if out else None
    if resume and prog and os.path.exists(prog):
        with open(prog) as f:
            st = json.load(f)
        assert st["n"] == n and st["half"] == half and st["per"] == per and st.get("canon", False) == canon
        start = st["next_start"]
        cyc = st["cyc"]; done = st["done"]; maxfin = st["maxfin"]
        for k, v in st.get("hist", {}).items():
            hist[int(k)] = v
        if "shape" in st and st["shape"]:
            shp[0] = st["shape"][0]; shp[1] = st["shape"][1]
        for kk, vv in (st.get("dhist") or {}).items():
            dh[int(kk)] = vv
        assert st.get("rule", 0) == rule
        if not quiet:
            print("resuming at %d / %d" % (start, total), flush=True)
    mode = "canon" if canon else ("half" if half else "full")
    span = max(1, SPAN_POSITIONS // unit)
    tuning = {"plain_per": min(max(int(per), 1), 4096)} if kernel == "plain" else None
    t0 = time.time()
    last_write = 0.0
    slowest = 0.0
    while start < total:
        tl = time.time()
        stop = min(total, start + span)
        if canon and n < 2:      # one order of no cards and no split: nothing to play (gpu/war_gpu.py:209)
            start = stop
            continue
        r = eng.run_range(n, mode, start * unit, stop * unit, rule=rule, kernel=kernel, hbins=hbins, lbins=lbins,
                          dbins=dbins, shape_blk=blk, tuning=tuning, detector=detector)
        cyc += r["cyc"]; done += r["done"]; maxfin = max(maxfin, r["maxfin"])
        for i, v in enumerate(r["hist"]):
            hist[i] += v
        shp[0] += r["shape_nonstd"]; shp[1] += r["shape_unchk"]
        for i, v in enumerate(r["dhist"]):
            dh[i] += v
        slowest = max(slowest, time.time() - tl)
        start = stop
        if prog and (time.time() - last_write >= 10 or start >= total):
            last_write = time.time()
            st = {"n": n, "half": half, "canon": canon, "per": per, "next_start": min(start, total), "total": total,
                  "cyc": cyc, "done": done, "maxfin": maxfin,
                  "hist": {str(i): v for i, v in enumerate(hist) if v} if hbins else {},
                  "shape": [shp[0], shp[1]] if blk else None,
                  "rule": rule,
                  "dhist": {str(i): v for i, v in enumerate(dh) if v} if dbins else {},
                  "elapsed_s": round(time.time() - t0, 1), "slowest_launch_s": round(slowest, 3),
                  "updated": time.strftime("%Y-%m-%d %H:%M:%S")}
            tmp = prog + ".tmp"
            with open(tmp, "w") as f:
                json.dump(st, f)
            os.replace(tmp, prog)
            if not quiet:
                frac = min(start, total) / total
                el = time.time() - t0
                print("%.4f%%  cyc=%d  elapsed %.0fs  slowest launch %.3fs" % (100 * frac, cyc, el, slowest), flush=True)
    mult = 2 if half else 1
    c = cyc * mult
    d = done * mult
    h = {i: v * mult for i, v in enumerate(hist) if v} if hbins else {}
    sh = (shp[0], shp[1], blk) if blk else None
    dd = {i: v for i, v in enumerate(dh) if v} if dbins else None
    return c, d, maxfin, time.time() - t0, h, sh, dd


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("nmax", type=int)
    ap.add_argument("--nmin", type=int, default=None)
    ap.add_argument("--half", action="store_true", help="even n: use player-swap symmetry")
    ap.add_argument("--per", type=int, default=256,
                    help="deck orders per lane of the plain kernel (kept in the progress file)")
    ap.add_argument("--out", help="progress file prefix (enables resumable progress)")
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--hist", type=int, default=0, help="loop-length histogram bins (last bin = overflow)")
    ap.add_argument("--canon", action="store_true",
                    help="scan canonical positions (player 1's top card is n) instead of deals")
    ap.add_argument("--rule", type=int, default=0, choices=[0, 1],
                    help="put-back order: 0 = winner's card first (Spivey), 1 = loser's card first")
    ap.add_argument("--distinct", action="store_true",
                    help="canon only: count distinct cycles up to swapping the players")
    ap.add_argument("--shape", action="store_true",
                    help="canon only: check every loop alternates in runs of exactly 2^v (Spivey Thm 3 type)")
    ap.add_argument("--kernel", choices=["plain", "persistent"], default="persistent")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--lbins", type=int, default=0, help="histogram of ending-game lengths (not printed)")
    ap.add_argument("--det", choices=["auto", "brent", "alt", "pair"], default="auto",
                    help="cycle test (all exact): auto = alt for odd n, rule 0, persistent kernel; else brent; "
                         "pair: n = 2 mod 4, rule 0, persistent kernel")
    a = ap.parse_args(argv)
    lo = a.nmin if a.nmin is not None else a.nmax
    for n in range(lo, a.nmax + 1):
        half = a.half and n % 2 == 0 and not a.canon
        c, d, mx, el, h, sh, dd = run(n, half=half, per=a.per, out=a.out, resume=a.resume,
                                      quiet=a.out is None, hbins=a.hist, canon=a.canon, shape=a.shape,
                                      rule=a.rule, dbins=4096 if a.distinct else 0, kernel=a.kernel,
                                      lbins=a.lbins, device=a.device, det=a.det)
        expect = math.factorial(n - 1) * (n - 1) if a.canon else math.factorial(n)
        assert d == expect, (d, expect)
        if h:
            assert sum(h.values()) == c, ("histogram does not add up", sum(h.values()), c)
        what = "canonical positions cycling" if a.canon else "a(n)"
        print("n=%d  %s=%d  of %d  longest finite=%d rounds  %.1fs%s"
              % (n, what, c, d, mx, el, "  (half)" if half else ""), flush=True)
        if dd is not None:
            bad = {k: v for k, v in dd.items() if v % k}
            assert not bad, ("distinct-cycle weights not divisible", bad)
            assert 4095 not in dd, "a cycle had more than 4094 canonical states: raise dbins"
            print("   distinct cycles (up to swapping players): %d   [on-cycle canonical positions by weight: %s]"
                  % (sum(v // k for k, v in dd.items()), dict(sorted(dd.items()))), flush=True)
        if sh:
            print("   loop shape: %d cycling positions checked for runs of exactly %d; "
                  "not that shape: %d; too long to check: %d" % (c, sh[2], sh[0], sh[1]), flush=True)
        if h:
            print("   loop lengths (rounds: %s): " % ("positions" if a.canon else "deals") + ", ".join(
                ("%s%d: %d" % (">=" if k == a.hist - 1 else "", k, v)) for k, v in sorted(h.items())), flush=True)


if __name__ == "__main__":
    main()
