"""ctypes binding of the CPU oracle (oracle/war_oracle.cpp -> oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the benchmark's CPU-baseline
leg, always as the checker or the CPU baseline, never as part of the product path.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

FULL, HALF, CANON = 0, 1, 2
_MODES = {"full": FULL, "half": HALF, "canon": CANON}

_U64P = ctypes.POINTER(ctypes.c_uint64)


class OrcRes(ctypes.Structure):
    _fields_ = [("cyc", ctypes.c_uint64), ("done", ctypes.c_uint64), ("maxfin", ctypes.c_uint64),
                ("turns", ctypes.c_uint64), ("oncycle", ctypes.c_uint64), ("distinct_minkey", ctypes.c_uint64),
                ("distinct_keys", ctypes.c_uint64), ("shape_nonstd", ctypes.c_uint64),
                ("shape_unchk", ctypes.c_uint64),
                ("hist", _U64P), ("hbins", ctypes.c_int),
                ("lenhist", _U64P), ("lbins", ctypes.c_int),
                ("dhist", _U64P), ("dbins", ctypes.c_int),
                ("undet", ctypes.c_uint64)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcRes)]
        L.orc_run.restype = ctypes.c_int
        L.orc_run_det.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcRes)]
        L.orc_run_det.restype = ctypes.c_int
        L.orc_run_cap.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                  ctypes.POINTER(OrcRes)]
        L.orc_run_cap.restype = ctypes.c_int
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_total.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_total.restype = ctypes.c_uint64
        L.orc_play.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                               ctypes.c_int, ctypes.c_int, _U64P, _U64P]
        L.orc_play.restype = ctypes.c_int
        L.orc_unrank.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_unrank.restype = None
        _lib = L
    return _lib


def total(n, mode="full"):
    return int(lib().orc_total(int(n), _MODES.get(mode, mode)))


def run(n, mode="full", lo=0, hi=None, rule=0, hbins=0, lbins=0, dbins=0, distinct=False, shape_blk=0, threads=0,
        detector="brent", step_cap=0):
    """Play every position of [lo, hi) on the CPU; returns a dict shaped like the engine's result.
    detector: "brent" (every output), "streak" (odd n, rule 0, no cycle-length outputs), "alt"
    (Brent with alternation checkpoints: odd n, rule 0, every output) or "pair" (Brent with pair-run
    checkpoints: n = 2 mod 4, rule 0, every output). step_cap: a game unresolved after step_cap turns
    counts as undetermined (0: none, or the reference's 2^22 for the streak test, whose test is "more
    than step_cap turns played" at the top of its loop)."""
    L = lib()
    m = _MODES.get(mode, mode)
    if hi is None:
        hi = total(n, m)
    H = (ctypes.c_uint64 * max(hbins, 1))()
    Ln = (ctypes.c_uint64 * max(lbins, 1))()
    D = (ctypes.c_uint64 * max(dbins, 1))()
    r = OrcRes()
    r.hist = ctypes.cast(H, _U64P); r.hbins = hbins
    r.lenhist = ctypes.cast(Ln, _U64P); r.lbins = lbins
    r.dhist = ctypes.cast(D, _U64P); r.dbins = dbins
    det = {"brent": 0, "streak": 1, "alt": 2, "pair": 3}[detector]
    rc = L.orc_run_cap(int(n), m, int(rule), int(lo), int(hi), int(threads), int(bool(distinct)), int(shape_blk),
                       det, int(step_cap), ctypes.byref(r))
    if rc:
        raise ValueError("oracle: %s" % L.orc_last_error().decode())
    return {"cyc": r.cyc, "done": r.done, "maxfin": r.maxfin, "turns": r.turns, "oncycle": r.oncycle,
            "distinct_minkey": r.distinct_minkey, "distinct_keys": r.distinct_keys,
            "shape_nonstd": r.shape_nonstd, "shape_unchk": r.shape_unchk, "undetermined": r.undet,
            "hist": [int(H[i]) for i in range(hbins)], "lenhist": [int(Ln[i]) for i in range(lbins)],
            "dhist": [int(D[i]) for i in range(dbins)]}


def play(n, p1, p2, rule=0):
    """One position (piles top card first). Returns (endless, rounds, lam)."""
    L = lib()
    a = (ctypes.c_int * max(1, len(p1)))(*p1)
    b = (ctypes.c_int * max(1, len(p2)))(*p2)
    rounds = ctypes.c_uint64(0)
    lam = ctypes.c_uint64(0)
    c = L.orc_play(int(n), a, len(p1), b, len(p2), int(rule), ctypes.byref(rounds), ctypes.byref(lam))
    return bool(c), rounds.value, lam.value


def unrank(r, m):
    p = (ctypes.c_int * max(1, m))()
    lib().orc_unrank(int(r), int(m), p)
    return [p[i] for i in range(m)]


def cycle_info(p1, p2, rule=0, cap=10**12):
    """One position of up to 255 byte cards (cpu/cycle_length.c). Returns
    {"status": "ends"|"cycles"|"unsettled", "turns", "lam", "mu"}."""
    L = lib()
    f = L.orc_cycle_info
    f.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                  ctypes.c_uint64, _U64P, _U64P, _U64P]
    f.restype = ctypes.c_int
    a = (ctypes.c_int * max(1, len(p1)))(*p1)
    b = (ctypes.c_int * max(1, len(p2)))(*p2)
    t, lam, mu = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
    rc = f(a, len(p1), b, len(p2), int(rule), int(cap), ctypes.byref(t), ctypes.byref(lam), ctypes.byref(mu))
    if rc < 0:
        raise ValueError("oracle: %s" % L.orc_last_error().decode())
    return {"status": {0: "ends", 1: "cycles", 2: "unsettled"}[rc], "turns": t.value,
            "lam": lam.value if rc == 1 else 0, "mu": mu.value if rc == 1 else 0}


def reach(n, threads=0, cap=200000000, lo=0, hi=None):
    """Distinct cycles met by the canonical positions [lo, hi) of n cards (each at its minimum-key
    position) and whether ordinary deals enter them (cpu/reach_cycles.c); max_depth: the deepest
    backward-search stack."""
    L = lib()
    f = L.orc_reach_range
    f.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64, _U64P, _U64P, _U64P,
                  _U64P, _U64P, _U64P]
    f.restype = ctypes.c_int
    if hi is None:
        hi = total(n, CANON)
    c, y, no, u, d = (ctypes.c_uint64(0) for _ in range(5))
    nl = (ctypes.c_uint64 * 4096)()
    rc = f(int(n), int(lo), int(hi), int(threads), int(cap), ctypes.byref(c), ctypes.byref(y), ctypes.byref(no),
           ctypes.byref(u), ctypes.cast(nl, _U64P), ctypes.byref(d))
    if rc:
        raise ValueError("oracle: %s" % L.orc_last_error().decode())
    return {"classes": c.value, "reached": y.value, "not_reached": no.value, "undetermined": u.value,
            "nolen": [int(nl[i]) for i in range(4096)], "max_depth": d.value}
