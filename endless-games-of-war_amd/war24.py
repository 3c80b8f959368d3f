"""Count the never-ending deals of single-suit War over a range of deal ranks, n <= 20, on an MI355X.

Drop-in for the reference's gpu/war24.py rank-range driver: the same command line, the same
checkpoint (.ckpt JSON) and progress (.progress text log) files next to this script, the same
result line and the same exit status. The counting runs in libwar_engine.so (hand-written HIP for
gfx950) through its C ABI; nothing is computed on the CPU.

  python war24.py 17 --start 100000000000000 --count 1000000000
  python war24.py 17 --det floyd --start 0 --count 8992587776 --resume --tag _idA

Deal ranks are lexicographic over the n! orders of cards 0..n-1 (the order war24.py enumerates);
deck positions 0, 2, 4, ... go to player 1, top card first. --det names the cycle detector:
floyd (any n) or streak (odd n only, as in the reference). Both are exact, so they give the same
cycling count and longest finite game; the engine runs the streak test as written for --det streak
(one deal per lane), and Brent on its persistent flow kernel for --det floyd (Floyd and Brent find
the same cycles and game lengths). A game still unresolved after the reference's step cap of 2^22
turns counts as undetermined and is printed as such (gpu/war24.py:57,96,103,127): the streak test
keeps the reference's test exactly; Brent stops a game unresolved after 2^22 turns of its own
(no valid game of n <= 20 comes near it). --chunk and --threads shaped the reference's launches;
they are accepted and checked as the reference checks them.

Reference behaviour followed (gpu/war24.py): main() :157-207 (arguments :159-167, asserts
:169-173, files :177-179, checkpoint keys :180-181, :193-195, progress line :197-198, result line
:203-205, exit on a count mismatch :206).
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

SPAN_DEALS = 1 << 32        # deals per engine call (between checkpoints)
STEP_CAP = 1 << 22          # the reference's CAP (gpu/war24.py:57)

_engine = None


def engine(device=0):
    global _engine
    if _engine is None:
        _engine = Engine(device)
    return _engine


def files(n, det, lo, hi, tag, where=None):
    """The reference's checkpoint and progress file names, next to this script (gpu/war24.py:177-179)."""
    where = _HERE if where is None else where
    t = "n%d_%s_%d_%d%s" % (n, det, lo, hi, tag)
    return os.path.join(where, "war24_%s.progress" % t), os.path.join(where, "war24_%s.ckpt" % t)


def run(n, det="streak", lo=0, hi=None, resume=False, tag="", device=0, where=None, span=None, log=True,
        step_cap=STEP_CAP):
    """Count deals [lo, hi) of n cards. Returns (cyc, maxfin, deals, undetermined, elapsed_s, line).
    step_cap: the turns after which an unresolved game counts as undetermined (the reference's 2^22;
    a multiple of 32 for --det floyd)."""
    hi = math.factorial(n) if hi is None else hi
    span = SPAN_DEALS if span is None else span
    prog, ckpt = files(n, det, lo, hi, tag, where)
    st = {"pos": lo, "cyc": 0, "max": 0, "deals": 0, "und": 0}
    if resume and os.path.exists(ckpt):
        with open(ckpt) as f:
            st = json.load(f)
    eng = engine(device)
    detector = "streak" if det == "streak" else "brent"
    kernel = "plain" if det == "streak" else "persistent"
    pos = st["pos"]
    pos0 = pos
    t0 = time.time()
    last = t0
    pf = open(prog, "a") if log else None
    while pos < hi:
        stop = min(hi, pos + span)
        r = eng.run_range(n, "full", pos, stop, kernel=kernel, detector=detector, step_cap=step_cap)
        st["cyc"] += r["cyc"]
        st["max"] = max(st["max"], r["maxfin"])
        st["deals"] += r["done"]
        st["und"] += r["undetermined"]
        pos = stop
        st["pos"] = pos
        now = time.time()
        if now - last > 10 or pos >= hi:
            with open(ckpt, "w") as f:
                json.dump(st, f)
            if pf:
                rate = (pos - pos0) / max(now - t0, 1e-9)
                pf.write("%s n=%d det=%s pos=%d/%d (%.4f%%) cyc=%d max=%d und=%d rate=%.3e/s\n"
                         % (time.strftime("%H:%M:%S"), n, det, pos, hi, 100 * (pos - lo) / (hi - lo), st["cyc"],
                            st["max"], st["und"], rate))
                pf.flush()
            last = now
    line = ("n=%d det=%s range=[%d,%d) deals_simulated=%d (expected %d) cycling=%d longest_finite=%d "
            "undetermined=%d time=%.1fs" % (n, det, lo, hi, st["deals"], hi - lo, st["cyc"], st["max"], st["und"],
                                           time.time() - t0))
    if pf:
        pf.write(line + "\n")
        pf.close()
    return st["cyc"], st["max"], st["deals"], st["und"], time.time() - t0, line


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("n", type=int)
    ap.add_argument("--det", choices=["floyd", "streak"], default="streak")
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--count", type=int, default=0, help="number of deals from --start (0 = to the end)")
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--threads", type=int, default=1 << 17)
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--tag", default="", help="suffix for the progress and checkpoint file names")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    n = a.n
    assert 2 <= n <= 20, "deal ranks are 64-bit: n <= 20"
    if a.det == "streak" and n % 2 == 0:
        sys.exit("streak detection is only valid for odd n (Spivey Thm 6)")
    assert a.threads % 256 == 0, "--threads must be a multiple of 256"
    total_all = math.factorial(n)
    lo = a.start
    hi = total_all if a.count == 0 else min(total_all, lo + a.count)
    cyc, mx, deals, und, el, line = run(n, a.det, lo, hi, resume=a.resume, tag=a.tag, device=a.device)
    print(line, flush=True)
    if deals != hi - lo:
        print("ERROR: deal count mismatch", flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
