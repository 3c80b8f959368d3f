"""The mixed-n batch (war_run_batch) and the serving loop at the request sizes their workloads name:
ranges log-uniform up to 10^9 positions for a batch, up to 10^10 for served queries. The CPU oracle
cannot play ranges that long in a test's time, so each workload here cuts whole ranges whose totals
the reference publishes into consecutive pieces of those lengths (seeded numpy PCG64), runs the
pieces as separate requests in one batch / one serving session, and requires the pieces to add up
to the published totals exactly (counts, cycle-length histograms, distinct cycles, weights) and their
longest games to reach the published longest. A piece counted wrongly, attributed to another request
or cut at the wrong position breaks a sum. Shorter pieces are checked one by one against the oracle
in test_gpu_batch.py and test_gpu_serving.py.

Published values (tests/golden/goldens.json, transcribed from the reference's data/ files):
a(13) = 4880546113 with its histogram and longest 57; n = 13 canonical positions: 3609048460 cycling,
319680 distinct cycles, weights {7, 14, 35}; n = 12 canonical: none cycling, longest 79; n = 12 deals
under rule 1: 23880636 endless, longest 118; n = 14 canonical: 2872129173 cycling, longest 465,
histogram {96, 112}; every n = 16 deal ends; the n = 17 rank ranges the reference logged."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def total(n, mode):
    return math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)


def pieces(rng, lo, hi, top):
    """Consecutive pieces covering [lo, hi), each floor(10^U(3, top)) long (the last one clipped)."""
    out = []
    while lo < hi:
        ln = min(hi - lo, int(10 ** rng.uniform(3, top)))
        out.append((lo, lo + ln))
        lo += ln
    return out


def hist_sum(results, key):
    acc = {}
    for r in results:
        for i, v in enumerate(r[key]):
            if v:
                acc[str(i)] = acc.get(str(i), 0) + v
    return acc


def check_sums(results, want, npos, hist=True):
    assert sum(r["done"] for r in results) == npos
    assert sum(r["cyc"] for r in results) == want["cyc"]
    assert max(r["maxfin"] for r in results) == want["longest"]
    if hist and "hist" in want:
        assert hist_sum(results, "hist") == want["hist"]


def test_batch_pieces_up_to_1e9_sum_to_published(eng, goldens):
    rng = np.random.Generator(np.random.PCG64(20261203))
    d13, c13, c12 = goldens["deals"]["13"], goldens["canon_gpu"]["13"], goldens["canon_gpu"]["12"]
    r12 = goldens["rule1_deals"]["12"]
    groups = {
        "n13_full": (13, "full", {"hbins": 128, "lbins": 128},
                     {"cyc": d13["never_end"], "longest": d13["longest"], "hist": d13["hist"]}),
        "n13_canon_distinct": (13, "canon", {"hbins": 128, "distinct": True, "dbins": 64},
                               {"cyc": c13["cycling"], "longest": c13["longest"]}),
        "n12_canon": (12, "canon", {"hbins": 64}, {"cyc": c12["cycling"], "longest": c12["longest"]}),
        "n12_full_rule1": (12, "full", {"hbins": 256, "rule": 1},
                           {"cyc": r12["never_end"], "longest": r12["longest"]}),
    }
    qs, where = [], {}
    for name, (n, mode, kw, _) in groups.items():
        for lo, hi in pieces(rng, 0, total(n, mode), 9):
            where.setdefault(name, []).append(len(qs))
            qs.append((n, mode, lo, hi, kw))
    # whole small ranges of every other n of 5..11, one request each
    for n in range(5, 12):
        where.setdefault("small", []).append(len(qs))
        qs.append((n, "full", 0, total(n, "full"), {"hbins": 64}))
    assert len(qs) <= 240, len(qs)
    print("batch of %d requests, longest %d positions" % (len(qs), max(hi - lo for _, _, lo, hi, _ in qs)), flush=True)
    got, launches = eng.run_batch(qs)
    assert launches >= 1
    for name, (n, mode, kw, want) in groups.items():
        rs = [got[i] for i in where[name]]
        check_sums(rs, want, total(n, mode))
    rs = [got[i] for i in where["n13_canon_distinct"]]
    assert sum(r["distinct_minkey"] for r in rs) == c13["distinct"]
    assert hist_sum(rs, "dhist") == c13["weights"]
    assert sum(r["oncycle"] for r in rs) == goldens["canon"]["13"]["on_cycle"]
    for i, n in zip(where["small"], range(5, 12)):
        g = goldens["deals"][str(n)]
        assert (got[i]["cyc"], got[i]["maxfin"]) == (g["never_end"], g["longest"]), n


def test_served_pieces_up_to_1e10_sum_to_published(eng, goldens):
    rng = np.random.Generator(np.random.PCG64(20261205))
    c14, d13 = goldens["long"]["canon14"], goldens["deals"]["13"]
    qs, where = [], {}
    for name, n, mode in (("n14_canon", 14, "canon"), ("n13_full", 13, "full")):
        for lo, hi in pieces(rng, 0, total(n, mode), 10):
            where.setdefault(name, []).append(len(qs))
            qs.append((n, mode, lo, hi, {"hbins": 256}))
    # three n = 16 deal ranges of 10^9..10^10 (every n = 16 deal ends) and the n = 17 ranges the
    # reference logged
    t16 = total(16, "full")
    for _ in range(3):
        ln = int(10 ** rng.uniform(9, 10))
        lo = int(rng.integers(0, t16 - ln))
        where.setdefault("n16", []).append(len(qs))
        qs.append((16, "full", lo, lo + ln, {"hbins": 64}))
    for key in ("n17_0_8992587776", "n17_100000000000000_100001000000000"):
        g = goldens["long"][key]
        where.setdefault("n17", []).append((len(qs), g))
        qs.append((17, "full", g["lo"], g["hi"], {"hbins": 128}))
    order = rng.permutation(len(qs))
    print("serving %d queries, %.3e positions" % (len(qs), sum(hi - lo for _, _, lo, hi, _ in qs)), flush=True)
    eng.serve_start(max_slots=64)
    try:
        ids = {}
        for i in order:
            n, mode, lo, hi, kw = qs[i]
            ids[int(i)] = eng.submit(n, mode, lo, hi, **kw)
        got = {i: eng.wait(rid) for i, rid in ids.items()}
    finally:
        eng.serve_stop()
    check_sums([got[i] for i in where["n14_canon"]],
               {"cyc": c14["cycling"], "longest": c14["longest"], "hist": c14["hist"]}, c14["positions"])
    check_sums([got[i] for i in where["n13_full"]],
               {"cyc": d13["never_end"], "longest": d13["longest"], "hist": d13["hist"]}, d13["deals"])
    for i in where["n16"]:
        assert got[i]["cyc"] == goldens["deals"]["16"]["never_end"] == 0 and got[i]["maxfin"] > 0
        assert got[i]["done"] == qs[i][3] - qs[i][2]
    for i, g in where["n17"]:
        assert (got[i]["cyc"], got[i]["maxfin"], got[i]["done"]) == (g["cycling"], g["longest"], g["hi"] - g["lo"])
    assert all(r["latency_ms"] > 0 for r in got.values())
