"""Benchmark of the MI355X War engine; prints ONE JSON line.

Headline workload (value): the n = 13 full enumeration, every one of the 13! = 6,227,020,800
deals, played to its end or to its cycle, as the drop-in war_gpu.py runs it by default: the flow
kernel with the alternation-checkpoint detector (exact; odd n, rule 0). Counts are checked
bit-exactly against the published a(13) = 4,880,546,113, longest 57 and cycle-length histogram
every step. One step = one whole enumeration. Beside it, on the same GPU in the same run:
  brent_flow     the same enumeration with Brent's detector on the flow kernel (its turn count is
                 the reference's Sum of rounds), and with Brent on the staged pipeline;
  plain_kernel   the one-deal-per-lane kernel on the same n = 13 enumeration;
  canon14        the n = 14 canonical scan (13! * 13 positions, Brent; flow kernel and staged
                 pipeline) against the published counts, and the plain kernel on sampled chunks;
  n17_range      the n = 17 rank range [0, 8992587776) the reference logged: held-card layout
                 (4-bit piles, card 16 apart) and 5-bit piles, Brent and the alternation detector;
  serving        a continuous-batching run of mixed queries with p50/p99 latency, beside the same
                 queries one by one through the plain and the persistent kernels;
  mixed_batch    240 requests of n = 5..15 (full, canonical, distinct) in shared persistent launches
                 (war_run_batch), beside the same requests one by one;
  roofline       integer VALU: useful lane-ops/s (turns x VALU instructions per turn of the flow
                 kernel's loop x 64 lanes) against the VALU's dual-issue peak, CUs x 128 lane-ops
                 per clock, with the fraction of the single-issue rate and of the turn's own
                 instruction-mix ceiling beside it;
  cpu_baseline   the CPU oracle (OpenMP) on a bounded sample of the same n = 13 workload.

  python bench.py [--gpus 1] [--steps 3] [--warmup 1]
                  [--only main|brent13|staged13|canon14|canon14walk|plain13|copy]
                  [--no-extras]
"""
import argparse
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "endless-games-of-war_amd"))

METRIC = "deals/sec (n=13 full, n=14 canon), bit-exact; batched vs plain kernel; p99 latency"
DET13 = "auto"     # set in main(): the cycle test war_gpu.py picks for n = 13 (war_gpu.detector_for)
PUBLISHED_N13_DEALS_PER_S = 6227020800 / 15.3      # RTX 5080, war_gpu.py 13 (reference data/checks.txt:23)
A13 = {"cyc": 4880546113, "maxfin": 57, "hist": {42: 1475465743, 56: 2394985450, 70: 1010094920}}
C14 = {"cyc": 2872129173, "maxfin": 465, "hist": {96: 43669233, 112: 2828459940}}
# VALU instructions per War turn in the flow kernel's loop: 16 for the turn (its last subtraction
# also yields the end test as its borrow) and one compare of player 1's pile with the tortoise's,
# counted in the gfx950 assembly of k_flow (k_stage's age-synchronous loop also has 17); the
# algorithmic unit is one turn of the detector run.
VALU_OPS_PER_TURN = 17
# VALU issue on gfx950, measured with an issue-rate probe at 7 waves per SIMD: the VOP2
# v_add/v_sub/v_and/v_or/v_xor/v_mov_b32 and the 32- and 16-bit shifts right issue two wave64
# instructions per SIMD per 4 clocks (dual issue); 64-bit shifts, adds and compares, VOP3 forms,
# selects, compares and VOP2 forms with a scalar operand issue one per 4 clocks. The peak is
# therefore 2 x 64 lanes per SIMD per 4 clocks = CUs x 128 lane-ops per clock (78.6 T/s at 256 CUs
# and 2.4 GHz, the rate the published 157.3 TFLOP/s FP32 vector figure implies). Of the turn's 17
# instructions 7 can dual-issue (2 ANDs, 2 shifts right, 2 ORs, the length add) and 10 cannot, so
# the turn itself needs at least 10 x 4 + 7 x 2 = 54 clocks per wave: its instruction mix caps VALU
# issue at 17 x 4 / 54 = 1.26 instructions per SIMD per 4 clocks, 0.630 of the dual-issue peak.
VALU_LANE_OPS_PER_CU_CLOCK = 128
TURN_MIX_CEILING = (VALU_OPS_PER_TURN * 4.0 / (10 * 4 + 7 * 2)) / 2.0
# HBM bytes per k_flow launch on the n = 13 workload (2 launches; no records are written: 0.20 MB
# read, 66.5 MB of tile-grab atomics and counters), from the rocprofv3 byte counters calibrated on a
# copy of known size.
K_FLOW_HBM_BYTES_PER_LAUNCH = 6.67e7


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hist_of(r):
    return {i: v for i, v in enumerate(r["hist"]) if v}


def check_n13(r):
    assert r["done"] == math.factorial(13), r["done"]
    assert r["cyc"] == A13["cyc"] and r["maxfin"] == A13["maxfin"] and hist_of(r) == A13["hist"], r
    assert r["undetermined"] == 0


def chunks_spread(total, count, size, seed):
    rng = random.Random(seed)
    step = total // count
    out = []
    for i in range(count):
        lo = i * step + rng.randrange(0, step - size)
        out.append((lo, lo + size))
    return out


def cpu_baseline(n, mode, chunks, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.lib()
    t0 = time.time()
    res = []
    for lo, hi in chunks:
        res.append(oracle.run(n, mode, lo, hi, hbins=256, threads=threads))
    el = time.time() - t0
    return res, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--only", choices=["main", "brent13", "staged13", "canon14", "canon14walk", "plain13", "copy"], default=None,
                    help="run one workload only (for profiling); copy: a 1 GiB device-to-device copy that "
                         "calibrates the HBM byte counters")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-chunks", type=int, default=64)
    ap.add_argument("--autotune", action="store_true", help="autotune n=13 full and n=14 canon first")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (the rank space would be "
              "split across GPUs with one all-reduce of the counters, not built here)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    if a.only == "copy":
        import torch
        x = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        assert int(y[-1]) == 1
        log("copy: 1 GiB read + 1 GiB written")
        return

    from war_amd import Engine
    from war_gpu import detector_for
    global DET13
    DET13 = detector_for(13, 0, "persistent")
    eng = Engine(0)
    dev = eng.device_info()
    log("device", dev)
    n13 = math.factorial(13)
    if a.autotune:
        for n, mode in ((13, "full"), (14, "canon")):
            log("autotune", n, mode, eng.autotune(n, mode, budget_ms=60000))
    tun = eng.tuning_for(13, "full")

    if a.only == "main":
        r = eng.run_range(13, "full", hbins=256, detector=DET13)
        check_n13(r)
        log("n13 flow (%s) device_ms %.1f" % (DET13, r["device_ms"]))
        return
    if a.only == "brent13":
        r = eng.run_range(13, "full", hbins=256)
        check_n13(r)
        log("n13 flow (brent) device_ms %.1f" % r["device_ms"])
        return
    if a.only == "staged13":
        r = eng.run_range(13, "full", hbins=256, tuning={"stage_kind": "lockstep"})
        check_n13(r)
        log("n13 staged (brent) device_ms %.1f" % r["device_ms"])
        return
    if a.only == "canon14":
        r = eng.run_range(14, "canon", hbins=256)
        assert r["cyc"] == C14["cyc"] and r["maxfin"] == C14["maxfin"] and hist_of(r) == C14["hist"]
        log("canon14 persistent device_ms %.1f" % r["device_ms"])
        return
    if a.only == "canon14walk":
        r = eng.run_range(14, "canon", hbins=256, dbins=4096, distinct=True, shape_blk=2)
        assert r["cyc"] == C14["cyc"] and r["maxfin"] == C14["maxfin"] and hist_of(r) == C14["hist"]
        assert (r["distinct_minkey"], r["shape_nonstd"], r["oncycle"]) == (26640, 43669233, 496800)
        log("canon14 with distinct + shape device_ms %.1f" % r["device_ms"])
        return
    if a.only == "plain13":
        r = eng.run_range(13, "full", kernel="plain", hbins=256)
        check_n13(r)
        log("n13 plain device_ms %.1f" % r["device_ms"])
        return

    # ---- headline: n = 13 full enumeration as the drop-in runs it (flow kernel, alternation detector)
    for _ in range(a.warmup):
        check_n13(eng.run_range(13, "full", hbins=256, detector=DET13))
    t0 = time.perf_counter()
    dev_ms = 0.0
    turns = 0
    for _ in range(a.steps):
        r = eng.run_range(13, "full", hbins=256, lbins=512, detector=DET13)
        check_n13(r)
        dev_ms += r["device_ms"]
        turns += r["turns"]
    wall = time.perf_counter() - t0
    value = n13 * a.steps / wall
    peak_lane_ops = dev["cus"] * VALU_LANE_OPS_PER_CU_CLOCK * dev["max_clock_mhz"] * 1e6
    achieved = turns * VALU_OPS_PER_TURN / (dev_ms / 1e3)
    out = {
        "metric": METRIC, "value": value, "unit": "deals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": value / PUBLISHED_N13_DEALS_PER_S, "dtype": "u64",
        "data": "synthetic: exhaustive enumeration of all 13! deals (no input data)",
        "config": {"workload": "n13_full: all 13! = 6227020800 deals, flow kernel (persistent), as war_gpu.py 13 "
                               "runs it", "n": 13, "mode": "full", "rule": 0, "positions": n13,
                   "kernel": "persistent", "detector": DET13, "tuning": tun,
                   "tiles": "ranks_per_refill rounded to whole blocks of 24 deals per lane run (1536 per "
                            "tile); the kernel's deal-block variant steps each block with one card swap per deal",
                   "parity": "a(13), longest 57, cycle-length histogram equal the published values every step"},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak_lane_ops / 1e12,
                     "unit": "Tlane-op/s", "frac": achieved / peak_lane_ops, "traffic": K_FLOW_HBM_BYTES_PER_LAUNCH,
                     "peak_note": "VALU dual-issue peak: 2 wave64 instructions per SIMD per 4 clocks (measured for "
                                  "v_add/v_and/v_or; every other form in the turn issues one per 4 clocks)",
                     "frac_single_issue": 2.0 * achieved / peak_lane_ops,
                     "turn_mix_ceiling": TURN_MIX_CEILING,
                     "frac_of_turn_mix_ceiling": achieved / peak_lane_ops / TURN_MIX_CEILING,
                     "kernel": "k_flow<4, 0, alt, deal blocks>: one launch per batch of at most 2^32 deals "
                               "(whole tiles of 1536: 4294966272 and 1932054528, 2 launches per step)",
                     "traffic_note": "HBM bytes per k_flow launch at this workload from rocprofv3 FETCH_SIZE and "
                                     "WRITE_SIZE (scaled by a 1 GiB copy: x2.0 and x1.0), profiles/"
                                     "k_flow_n13_full_pmc.md; not measured inside this run",
                     "turns_per_step": turns // a.steps, "valu_ops_per_turn": VALU_OPS_PER_TURN,
                     "device_ms_per_step": dev_ms / a.steps,
                     "note": "achieved = the detector's turns x VALU ops per turn / device time (HIP events on "
                             "the engine's stream around each batch)"},
    }
    log("n13 flow (%s): %.3e deals/s, %.1f ms/step" % (DET13, value, 1e3 * wall / a.steps))

    if not a.no_extras:
        # ---- Brent on the same workload: flow kernel, and the staged pipeline
        bf = {}
        for name, t in (("flow", None), ("staged", {"stage_kind": "lockstep"})):
            r = eng.run_range(13, "full", hbins=256, tuning=t)
            check_n13(r)
            bf[name] = r
        bt = bf["flow"]["turns"]
        out["brent_flow"] = {"workload": "n13_full", "value": n13 / (bf["flow"]["device_ms"] / 1e3), "unit": "deals/s",
                             "device_ms": bf["flow"]["device_ms"], "turns": bt,
                             "frac": bt * VALU_OPS_PER_TURN / (bf["flow"]["device_ms"] / 1e3) / peak_lane_ops,
                             "staged_device_ms": bf["staged"]["device_ms"],
                             "staged_frac": bt * VALU_OPS_PER_TURN / (bf["staged"]["device_ms"] / 1e3) / peak_lane_ops,
                             "note": "Brent's turns are the reference's Sum of rounds (the same on both paths); "
                                     "frac as for the roofline, on these turns"}
        assert bf["staged"]["turns"] == bt
        log("n13 brent: flow %.1f ms, staged %.1f ms" % (bf["flow"]["device_ms"], bf["staged"]["device_ms"]))

        # ---- plain kernel on the same workload
        t0 = time.perf_counter()
        r = eng.run_range(13, "full", kernel="plain", hbins=256)
        pw = time.perf_counter() - t0
        check_n13(r)
        out["plain_kernel"] = {"workload": "n13_full", "value": n13 / pw, "unit": "deals/s",
                               "device_ms": r["device_ms"], "persistent_over_plain": value / (n13 / pw)}
        log("n13 plain: %.3e deals/s" % (n13 / pw))

        # ---- n = 14 canonical scan, persistent, whole range: flow kernel, then the staged pipeline
        t0 = time.perf_counter()
        r = eng.run_range(14, "canon", hbins=256)
        cw = time.perf_counter() - t0
        assert r["cyc"] == C14["cyc"] and r["maxfin"] == C14["maxfin"] and hist_of(r) == C14["hist"]
        tot14 = math.factorial(13) * 13
        rs = eng.run_range(14, "canon", hbins=256, tuning={"stage_kind": "lockstep"})
        assert (rs["cyc"], rs["maxfin"], rs["turns"]) == (r["cyc"], r["maxfin"], r["turns"])
        canon = {"workload": "n14_canon: all 13!*13 = 80951270400 positions", "persistent_value": tot14 / cw,
                 "unit": "positions/s", "persistent_device_ms": r["device_ms"],
                 "frac": r["turns"] * VALU_OPS_PER_TURN / (r["device_ms"] / 1e3) / peak_lane_ops,
                 "staged_device_ms": rs["device_ms"], "turns": r["turns"],
                 "parity": "cycling 2872129173, longest 465, hist {96,112} equal the published values"}
        # both kernels on the same sampled chunks (seed 14, whole q-ranks): batched / plain ratio
        ch = [(lo - lo % 13, lo - lo % 13 + 13 * (1 << 22)) for lo, _ in chunks_spread(tot14, 8, 13 * (1 << 22), 14)]
        tp = tq = 0.0
        for lo, hi in ch:
            rp = eng.run_range(14, "canon", lo, hi, kernel="plain")
            rq = eng.run_range(14, "canon", lo, hi, kernel="persistent")
            assert (rp["cyc"], rp["maxfin"], rp["turns"]) == (rq["cyc"], rq["maxfin"], rq["turns"])
            tp += rp["device_ms"]; tq += rq["device_ms"]
        npos = sum(hi - lo for lo, hi in ch)
        canon.update({"sample": "8 chunks of 13*2^22 positions (whole q-ranks), seed 14",
                      "plain_sample_value": npos / (tp / 1e3), "persistent_sample_value": npos / (tq / 1e3),
                      "persistent_over_plain": tp / tq, "tuning": eng.tuning_for(14, "canon")})
        out["canon14"] = canon
        log("canon14: %.3e pos/s persistent (whole), ratio %.2f on samples" % (tot14 / cw, tp / tq))

        # ---- n = 17, the published rank range [0, 8992587776): held-card layout against 5-bit piles
        lo17, hi17 = 0, 8992587776
        n17 = {"workload": "n17_range: deal ranks [0, 8992587776) of 17!", "unit": "deals/s",
               "parity": "cycling 8016845372, longest 94 (the reference's logged range)",
               "reference_s": 13.2}
        for lay in (6, 5):
            r = eng.run_range(17, "full", lo17, hi17, tuning={"pile_layout": lay})
            assert (r["cyc"], r["maxfin"], r["done"]) == (8016845372, 94, hi17 - lo17), r
            n17["layout%d_value" % lay] = (hi17 - lo17) / (r["device_ms"] / 1e3)
            n17["layout%d_device_ms" % lay] = r["device_ms"]
        n17["held_over_5bit"] = n17["layout6_value"] / n17["layout5_value"]
        r = eng.run_range(17, "full", lo17, hi17, detector="alt")
        assert (r["cyc"], r["maxfin"], r["done"]) == (8016845372, 94, hi17 - lo17), r
        n17["layout6_alt_value"] = (hi17 - lo17) / (r["device_ms"] / 1e3)
        n17["layout6_alt_device_ms"] = r["device_ms"]
        out["n17_range"] = n17
        log("n17 range: held layout %.3e deals/s, 5-bit %.3e" % (n17["layout6_value"], n17["layout5_value"]))

        # ---- serving: continuous batching of mixed queries
        out["serving"] = serving(eng)

        # ---- mixed-n batch: many requests of n = 5..15 in shared persistent launches
        out["mixed_batch"] = mixed_batch(eng)
        log("mixed batch: %.3e positions/s in %d launches; one by one %.3e"
            % (out["mixed_batch"]["value"], out["mixed_batch"]["launches"],
               out["mixed_batch"]["one_by_one_persistent_positions_per_s"]))

        # ---- CPU baseline: the oracle on sampled n = 13 chunks, compared chunk by chunk
        threads = min(16, os.cpu_count() or 1)
        ch = chunks_spread(n13, a.cpu_chunks, 1 << 22, 13)
        cres, cel = cpu_baseline(13, "full", ch, threads)
        for (lo, hi), c in zip(ch, cres):
            g = eng.run_range(13, "full", lo, hi, hbins=256)
            assert (g["cyc"], g["maxfin"], g["turns"], g["hist"]) == (c["cyc"], c["maxfin"], c["turns"], c["hist"])
        out["cpu_baseline"] = {"value": sum(hi - lo for lo, hi in ch) / cel, "unit": "deals/s", "cores": threads,
                               "kind": "port",
                               "sample": "%d chunks of 2^22 n=13 deal ranks spread over 13! (seed 13), "
                                         "each equal to the GPU's counts" % len(ch)}
        log("cpu baseline: %.3e deals/s on %d threads" % (out["cpu_baseline"]["value"], threads))
    print(json.dumps(out), flush=True)


def serving(eng, nreq=64, seed=20261203, rate=100.0):
    """Mixed (n, range, mode) queries. (1) Poisson arrivals at `rate`/s through the continuous-batching
    loop: p50/p99 latency. (2) All submitted at once: aggregate positions/s. (3) The same queries one
    by one, synchronously, through the plain kernel and through the persistent kernel. Every count
    must be identical across (1), (2) and (3)."""
    rng = random.Random(seed)
    qs = []
    for _ in range(nreq):
        n = rng.randint(7, 17)
        mode = rng.choice(["full", "canon"])
        total = math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)
        ln = min(total, int(10 ** rng.uniform(3, 8.5)))
        lo = rng.randrange(0, total - ln + 1)
        distinct = mode == "canon" and n <= 13 and rng.random() < 0.3
        qs.append((n, mode, lo, lo + ln, distinct))
    keys = ("cyc", "done", "maxfin", "turns", "hist", "distinct_minkey")

    def run_served(gaps):
        eng.serve_start(max_slots=64, step_positions=1 << 26)
        ids = []
        t0 = time.perf_counter()
        for q, g in zip(qs, gaps):
            if g:
                time.sleep(g)
            n, mode, lo, hi, d = q
            ids.append(eng.submit(n, mode, lo, hi, hbins=128, distinct=d))
        res = [eng.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        eng.serve_stop()
        return res, wall

    arrivals, wall_a = run_served([rng.expovariate(rate) for _ in qs])
    burst, wall_b = run_served([0.0] * len(qs))
    lat = sorted(r["latency_ms"] for r in arrivals)
    npos = sum(hi - lo for _, _, lo, hi, _ in qs)
    alone = {}
    for kernel in ("plain", "persistent"):
        t = 0.0
        for q, r, b in zip(qs, arrivals, burst):
            n, mode, lo, hi, d = q
            t1 = time.perf_counter()
            p = eng.run_range(n, mode, lo, hi, kernel=kernel, hbins=128, distinct=d)
            t += time.perf_counter() - t1
            for k in keys:
                assert p[k] == r[k] == b[k], (q, kernel, k)
        alone[kernel] = npos / t
    return {"requests": nreq, "arrival_rate_per_s": rate, "p50_ms": lat[len(lat) // 2],
            "p99_ms": lat[min(len(lat) - 1, int(0.99 * len(lat)))], "arrivals_wall_s": wall_a,
            "saturated_positions_per_s": npos / wall_b,
            "one_by_one_plain_positions_per_s": alone["plain"],
            "one_by_one_persistent_positions_per_s": alone["persistent"],
            "queries": "n uniform 7..17, full or canon (distinct for 30%% of canon n<=13), length log-uniform "
                       "[1e3, 10^8.5] (%d positions in all); counts identical alone, in the mix and in the "
                       "burst" % npos}


def mixed_batch(eng, nreq=240, seed=20261203):
    """The mixed-n single-launch workload: seeded numpy PCG64, `nreq` requests (the engine has 248
    request slots), n uniform over 5..15, full or canonical, distinct + dhist for a third of the
    canonical requests of n <= 13, length floor(10^U(3, 9)) clipped to the total, lo uniform. Run
    as one war_run_batch call (requests of one layout, rule and detector share each launch; with
    cycle walks a launch holds 2^28 positions), then one by one through war_run_range: every count
    must be identical."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    qs = []
    for _ in range(nreq):
        n = int(rng.integers(5, 16))
        mode = "canon" if rng.random() < 0.5 else "full"
        total = math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)
        ln = min(total, int(10 ** rng.uniform(3, 9)))
        lo = int(rng.integers(0, total - ln + 1))
        kw = {"hbins": 128, "lbins": 128}
        if mode == "canon" and n <= 13 and rng.random() < 1 / 3:
            kw.update(distinct=True, dbins=256)
        qs.append((n, mode, lo, lo + ln, kw))
    npos = sum(hi - lo for _, _, lo, hi, _ in qs)
    eng.run_batch(qs[:8])                                   # graphs of the batch shapes captured
    t0 = time.perf_counter()
    got, launches = eng.run_batch(qs)
    wall = time.perf_counter() - t0
    t = 0.0
    for q, g in zip(qs, got):
        n, mode, lo, hi, kw = q
        t1 = time.perf_counter()
        p = eng.run_range(n, mode, lo, hi, **kw)
        t += time.perf_counter() - t1
        for k in ("cyc", "done", "maxfin", "turns", "hist", "lenhist", "oncycle", "distinct_minkey", "dhist"):
            assert p[k] == g[k], (q, k)
    return {"workload": "%d requests, n uniform 5..15, full or canon (distinct for 1/3 of canon n<=13), length "
                        "log-uniform [1e3, 1e9] clipped, numpy PCG64 seed %d: %d positions" % (nreq, seed, npos),
            "value": npos / wall, "unit": "positions/s", "launches": launches, "wall_s": wall,
            "one_by_one_persistent_positions_per_s": npos / t,
            "parity": "every count of every request identical in the batch and alone"}


if __name__ == "__main__":
    main()
