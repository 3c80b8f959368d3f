"""Python face of the MI355X War engine: range queries, autotuning and the serving loop.

All counting happens in libwar_engine.so on the GPU; this module only marshals requests and
results. Range units: deal ranks in [0, n!) for full/half mode, canonical positions
p = r*(n-1) + k in [0, (n-1)!(n-1)) for canon mode.
"""
import ctypes
import math
import threading

from . import _lib
from ._lib import (DETECTORS, KERNEL_PERSISTENT, KERNEL_PLAIN, KERNELS, MODE_CANON, MODE_FULL, MODE_HALF,
                   WarCycleInfo, WarDevinfo, WarPilePosition, WarReachResult, WarRequest, WarResult, WarTuning,
                   check)

MODES = {"full": MODE_FULL, "half": MODE_HALF, "canon": MODE_CANON}


def total_positions(n, mode):
    """Size of the range of (n, mode): n! deals, or (n-1)!(n-1) canonical positions."""
    mode = MODES.get(mode, mode)
    if n < 1:
        return 0
    if mode == MODE_CANON:
        return 0 if n < 2 else math.factorial(n - 1) * (n - 1)
    return math.factorial(n)


STAGE_KINDS = {"refill": 1, "lockstep": 2, "flow": 3}


def _tuning(t):
    if t is None:
        return WarTuning(0, 0, 0, 0, 0, 0)
    if isinstance(t, WarTuning):
        return t
    sk = t.get("stage_kind", 0)
    return WarTuning(int(t.get("waves_per_cu", 0)), int(t.get("ranks_per_refill", 0)),
                     int(t.get("brent_cap", 0)), int(t.get("plain_per", 0)), int(STAGE_KINDS.get(sk, sk)),
                     int(t.get("pile_layout", 0)))


def _tuning_dict(t):
    return {"waves_per_cu": t.waves_per_cu, "ranks_per_refill": t.ranks_per_refill, "brent_cap": t.brent_cap,
            "plain_per": t.plain_per, "stage_kind": {1: "refill", 2: "lockstep", 3: "flow"}.get(t.stage_kind, t.stage_kind),
            "pile_layout": t.pile_layout}


def make_request(n, mode="full", lo=0, hi=None, rule=0, kernel="persistent", hbins=0, lbins=0, dbins=0,
                 distinct=False, shape_blk=0, tuning=None, detector="brent", step_cap=0, undet_error=False,
                 cycle_table=None):
    """step_cap: turns after which an unresolved game counts as undetermined (0 = the default);
    undet_error: fail the request instead if any game is undetermined; cycle_table: True / False
    forces the cycle walks' state table on or off (None: the engine's default)."""
    mode = MODES.get(mode, mode)
    if hi is None:
        hi = total_positions(n, mode)
    return WarRequest(int(n), int(mode), int(rule), int(KERNELS.get(kernel, kernel)), int(lo), int(hi),
                      int(hbins), int(lbins), int(dbins), int(bool(distinct)), int(shape_blk),
                      int(DETECTORS.get(detector, detector)), _tuning(tuning), int(step_cap),
                      (_lib.FLAG_UNDET_ERROR if undet_error else 0) |
                      (0 if cycle_table is None else _lib.FLAG_CYCLE_TABLE if cycle_table else _lib.FLAG_NO_CYCLE_TABLE))


class _ResultBuf:
    """A WarResult with its caller-owned histogram arrays."""

    def __init__(self, req):
        self.req = req
        self.hist = (ctypes.c_uint64 * max(1, req.hbins))()
        self.lenhist = (ctypes.c_uint64 * max(1, req.lbins))()
        self.dhist = (ctypes.c_uint64 * max(1, req.dbins))()
        self.res = WarResult()
        self.res.hist = ctypes.cast(self.hist, ctypes.POINTER(ctypes.c_uint64))
        self.res.hbins = req.hbins
        self.res.lenhist = ctypes.cast(self.lenhist, ctypes.POINTER(ctypes.c_uint64))
        self.res.lbins = req.lbins
        self.res.dhist = ctypes.cast(self.dhist, ctypes.POINTER(ctypes.c_uint64))
        self.res.dbins = req.dbins

    def as_dict(self):
        r, q = self.res, self.req
        return {
            "n": q.n, "mode": q.mode, "rule": q.rule, "lo": q.lo, "hi": q.hi, "detector": q.detector,
            "cyc": r.cyc, "done": r.done, "maxfin": r.maxfin, "turns": r.turns,
            "oncycle": r.oncycle, "distinct_minkey": r.distinct_minkey, "distinct_hashset": r.distinct_hashset,
            "shape_nonstd": r.shape_nonstd, "shape_unchk": r.shape_unchk, "undetermined": r.undetermined,
            "hist": [int(self.hist[i]) for i in range(q.hbins)],
            "lenhist": [int(self.lenhist[i]) for i in range(q.lbins)],
            "dhist": [int(self.dhist[i]) for i in range(q.dbins)],
            "device_ms": r.device_ms, "latency_ms": r.latency_ms,
        }


class Engine:
    """One engine per process (the library holds a single device context)."""

    _lock = threading.Lock()

    def __init__(self, device=0, verbose=False):
        self.lib = _lib.load()
        with Engine._lock:
            check(self.lib.war_init(int(device), 1 if verbose else 0))
        self._pending = {}

    # ---- info / tuning
    def device_info(self):
        d = WarDevinfo()
        check(self.lib.war_device_info(ctypes.byref(d)))
        return {"cus": d.cus, "max_clock_mhz": d.max_clock_mhz, "wave_size": d.wave_size,
                "free_bytes": d.free_bytes, "total_bytes": d.total_bytes, "arch": d.arch.decode()}

    def get_tuning(self):
        t = WarTuning()
        check(self.lib.war_get_tuning(ctypes.byref(t)))
        return _tuning_dict(t)

    def tuning_for(self, n, mode="full", rule=0):
        """The tuning a request of (n, mode, rule) runs with when it sets none itself."""
        t = WarTuning()
        check(self.lib.war_tuning_for(int(n), int(MODES.get(mode, mode)), int(rule), ctypes.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, **kw):
        t = _tuning(kw)
        check(self.lib.war_set_tuning(ctypes.byref(t)))

    def autotune(self, n, mode="canon", rule=0, budget_ms=20000.0):
        t = WarTuning()
        check(self.lib.war_autotune(int(n), int(MODES.get(mode, mode)), int(rule), float(budget_ms), ctypes.byref(t)))
        return _tuning_dict(t)

    def autotune_batch(self, queries, budget_ms=20000.0, expect=None):
        """war_autotune_batch over queries [(n, mode, lo, hi, kw), ...] (run_range's keywords).
        expect: one result dict per query (keys as run_range returns them) that every candidate
        schedule must reproduce; None = the plain kernel's counts. Returns the chosen tuning."""
        reqs = [make_request(n, mode, lo, hi, **kw) for n, mode, lo, hi, kw in queries]
        arr = (WarRequest * len(reqs))(*reqs)
        exp = None
        keep = []
        if expect is not None:
            exp = (WarResult * len(reqs))()
            for i, (q, e) in enumerate(zip(reqs, expect)):
                r = exp[i]
                for k in ("cyc", "done", "maxfin", "turns", "oncycle", "shape_nonstd", "shape_unchk"):
                    setattr(r, k, int(e.get(k, 0)))
                r.distinct_minkey = int(e.get("distinct_minkey", 0))
                r.undetermined = int(e.get("undetermined", 0))
                for name, nb, key in (("hist", q.hbins, "hist"), ("lenhist", q.lbins, "lenhist"), ("dhist", q.dbins, "dhist")):
                    if nb:
                        a = (ctypes.c_uint64 * nb)(*[int(v) for v in e[key]])
                        keep.append(a)
                        setattr(r, name, ctypes.cast(a, ctypes.POINTER(ctypes.c_uint64)))
                        setattr(r, {"hist": "hbins", "lenhist": "lbins", "dhist": "dbins"}[name], nb)
        t = WarTuning()
        check(self.lib.war_autotune_batch(arr, len(reqs), float(budget_ms), exp, ctypes.byref(t)))
        return _tuning_dict(t)

    # ---- synchronous range query
    def run_range(self, n, mode="full", lo=0, hi=None, **kw):
        req = make_request(n, mode, lo, hi, **kw)
        buf = _ResultBuf(req)
        check(self.lib.war_run_range(ctypes.byref(req), ctypes.byref(buf.res)))
        return buf.as_dict()

    def run_range_noalloc(self, n, mode="full", lo=0, hi=None, **kw):
        """run_range through war_run_range_noalloc: the same counts with no device allocation (the
        engine's static scratch array); no cycle walks, histograms of at most 4096 bins."""
        req = make_request(n, mode, lo, hi, **kw)
        buf = _ResultBuf(req)
        check(self.lib.war_run_range_noalloc(ctypes.byref(req), ctypes.byref(buf.res)))
        return buf.as_dict()

    def run_batch(self, queries):
        """Several requests in shared persistent launches (war_run_batch). queries: [(n, mode, lo,
        hi, kw), ...] with run_range's keywords. Returns (results, launches): one result dict per
        query, each equal to run_range of that query, and the number of launches made."""
        reqs = [make_request(n, mode, lo, hi, **kw) for n, mode, lo, hi, kw in queries]
        count = len(reqs)
        arr = (WarRequest * max(1, count))(*reqs)
        bufs = [_ResultBuf(r) for r in reqs]
        res = (WarResult * max(1, count))(*[b.res for b in bufs])
        launches = ctypes.c_int32(0)
        check(self.lib.war_run_batch(arr, count, res, ctypes.byref(launches)))
        for i, b in enumerate(bufs):
            b.res = res[i]
        return [b.as_dict() for b in bufs], launches.value

    # ---- tools around the enumeration
    def reach(self, n, lo=0, hi=None, cap=200_000_000, nbins=4096):
        """Distinct cycles of canonical positions [lo, hi) and whether ordinary deals enter them."""
        if hi is None:
            hi = total_positions(n, MODE_CANON)
        nl = (ctypes.c_uint64 * max(1, nbins))()
        r = WarReachResult()
        r.nolen = ctypes.cast(nl, ctypes.POINTER(ctypes.c_uint64))
        r.nbins = nbins
        check(self.lib.war_reach(int(n), int(lo), int(hi), int(cap), ctypes.byref(r)))
        return {"classes": r.classes, "reached": r.reached, "not_reached": r.not_reached,
                "undetermined": r.undetermined, "nolen": [int(nl[i]) for i in range(nbins)],
                "device_ms": r.device_ms, "max_depth": r.max_depth}

    def cycle_info(self, positions, rule=0, cap=10**12):
        """positions: [(pile1, pile2), ...] of card values, top card first. Returns per position
        {"status": "ends"|"cycles"|"unsettled", "turns", "lam", "mu"}."""
        bufs = []
        arr = (WarPilePosition * max(1, len(positions)))()
        for i, (p1, p2) in enumerate(positions):
            b1 = (ctypes.c_uint8 * max(1, len(p1)))(*p1)
            b2 = (ctypes.c_uint8 * max(1, len(p2)))(*p2)
            bufs.append((b1, b2))
            arr[i].n1, arr[i].n2 = len(p1), len(p2)
            arr[i].p1 = ctypes.cast(b1, ctypes.POINTER(ctypes.c_uint8))
            arr[i].p2 = ctypes.cast(b2, ctypes.POINTER(ctypes.c_uint8))
        out = (WarCycleInfo * max(1, len(positions)))()
        check(self.lib.war_cycle_info_batch(arr, len(positions), int(rule), int(cap), out))
        names = {0: "ends", 1: "cycles", 2: "unsettled"}
        return [{"status": names[o.status], "turns": o.turns, "lam": o.lam, "mu": o.mu}
                for o in list(out)[:len(positions)]]

    # ---- serving loop
    def serve_start(self, max_slots=64, step_positions=0):
        check(self.lib.war_serve_start(int(max_slots), int(step_positions)))

    def serve_stop(self):
        check(self.lib.war_serve_stop())

    def submit(self, n, mode="full", lo=0, hi=None, **kw):
        req = make_request(n, mode, lo, hi, **kw)
        buf = _ResultBuf(req)
        rid = ctypes.c_uint64(0)
        check(self.lib.war_submit(ctypes.byref(req), ctypes.byref(buf.res), ctypes.byref(rid)))
        self._pending[rid.value] = buf
        return rid.value

    def poll(self, rid):
        d = ctypes.c_int32(0)
        check(self.lib.war_poll(ctypes.c_uint64(rid), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, rid):
        buf = self._pending.pop(rid)
        check(self.lib.war_wait(ctypes.c_uint64(rid)))
        return buf.as_dict()

    def close(self):
        check(self.lib.war_shutdown())
