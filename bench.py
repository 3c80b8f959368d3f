"""Benchmark of the MI355X War engine; prints ONE JSON line.

Headline workload (value): the n = 13 full enumeration, every one of the 13! = 6,227,020,800
deals, played to its end or to its cycle, as the drop-in war_gpu.py runs it by default: the flow
kernel with the alternation-checkpoint detector (exact; odd n, rule 0). Counts are checked
bit-exactly against the published a(13) = 4,880,546,113, longest 57 and cycle-length histogram
every step. One step = one whole enumeration. Beside it, on the same GPU in the same run:
  brent_flow     the same enumeration with Brent's detector on the flow kernel (its turn count is
                 the reference's Sum of rounds), and with Brent on the staged pipeline;
  plain_kernel   the one-deal-per-lane kernel on the same n = 13 enumeration (Brent, its fastest turn
                 form; the flow kernel's turn measured 18 % slower in it);
  canon14        the n = 14 canonical scan (13! * 13 positions, Brent; flow kernel and staged
                 pipeline) against the published counts, and the plain kernel on sampled chunks;
  n17_range      the n = 17 rank range [0, 8992587776) the reference logged: held-card layout
                 (4-bit piles, card 16 apart) and 5-bit piles, Brent and the alternation detector;
  serving        512 mixed queries (n = 7..17, lengths log-uniform to 10^10) through the
                 continuous-batching loop: saturated throughput (2^26- and 2^28-position steps), then
                 Poisson arrivals at 0.25, 0.5 and 0.9 of the saturated rate with p50/p99 latency per
                 rate, beside the same queries one by one through the plain and the persistent kernels;
  mixed_batch    256 requests of n = 5..15 (full, canonical, distinct) autotuned, then run as one
                 war_run_batch call (one graph launch), beside the same requests one by one;
  roofline       integer VALU: useful lane-ops/s (turns x VALU instructions per turn of the flow
                 kernel's loop x 64 lanes) against the VALU's dual-issue peak, CUs x 128 lane-ops
                 per clock, with the fraction of the single-issue rate and of the turn's own
                 instruction-mix ceiling beside it;
  cpu_baseline   the CPU oracle (OpenMP, every CPU the process may use) on bounded samples of the
                 n = 13 deals, all n = 9 deals and sampled n = 14 canonical chunks; beside it the
                 reference's own cpu/war.cpp (built from its source into oracle/_ref/ where the build
                 had the reference tree) over every n = 12 deal, with the oracle on the same deals.

  python bench.py [--gpus 1] [--steps 3] [--warmup 1]
                  [--only main|brent13|staged13|canon14|canon14walk|plain13|plaincanon14|serving|serveburst|
                          mixed|mixedbatch|copy]
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
# also yields the end test as its borrow) and one compare of player 1's pile with the tortoise's on
# every other turn (every cycle is of even length, so a game meets its tortoise only at an even number
# of turns after the tortoise was taken): 16.5 on average, counted in the gfx950 assembly of k_flow.
# k_stage's age-synchronous loop compares on every turn: 17. The algorithmic unit is one turn of the
# detector run.
VALU_OPS_PER_TURN = 16.5
STAGED_VALU_OPS_PER_TURN = 17
# VALU issue on gfx950, measured with an issue-rate probe at 7 waves per SIMD: the VOP2
# v_add/v_sub/v_and/v_or/v_xor/v_mov_b32 and the 32- and 16-bit shifts right issue two wave64
# instructions per SIMD per 4 clocks (dual issue); 64-bit shifts, adds and compares, VOP3 forms,
# selects, compares and VOP2 forms with a scalar operand issue one per 4 clocks. The peak is
# therefore 2 x 64 lanes per SIMD per 4 clocks = CUs x 128 lane-ops per clock (78.6 T/s at 256 CUs
# and 2.4 GHz, the rate the published 157.3 TFLOP/s FP32 vector figure implies). Of the turn's 16.5
# instructions 7 can dual-issue (2 ANDs, 2 shifts right, 2 ORs, the length add) and 9.5 cannot, so
# the turn itself needs at least 9.5 x 4 + 7 x 2 = 52 clocks per wave: its instruction mix caps VALU
# issue at 16.5 x 4 / 52 = 1.27 instructions per SIMD per 4 clocks, 0.635 of the dual-issue peak.
VALU_LANE_OPS_PER_CU_CLOCK = 128
TURN_MIX_CEILING = (VALU_OPS_PER_TURN * 4.0 / (9.5 * 4 + 7 * 2)) / 2.0
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


def canon14_chunks(tot14):
    """8 chunks of 13 * 2^22 n = 14 canonical positions (whole q-ranks), seed 14."""
    return [(lo - lo % 13, lo - lo % 13 + 13 * (1 << 22)) for lo, _ in chunks_spread(tot14, 8, 13 * (1 << 22), 14)]


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
    ap.add_argument("--only", choices=["main", "brent13", "staged13", "canon14", "canon14walk", "plain13",
                                       "plaincanon14", "serving", "serveburst", "mixed", "mixedbatch", "copy"],
                    default=None,
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
    if a.only == "plaincanon14":
        tot14 = math.factorial(13) * 13
        for lo, hi in canon14_chunks(tot14):
            r = eng.run_range(14, "canon", lo, hi, kernel="plain")
        log("canon14 plain (last chunk) device_ms %.1f" % r["device_ms"])
        return
    if a.only == "serving":
        print(json.dumps(serving(eng)), flush=True)
        return
    if a.only == "serveburst":
        qs, _ = serving_queries()
        res, wall = serve_burst(eng, qs)
        log("serving burst: %d queries, %.3e positions/s" % (len(qs), sum(hi - lo for _, _, lo, hi, _ in qs) / wall))
        return
    if a.only == "mixed":
        print(json.dumps(mixed_batch(eng)), flush=True)
        return
    if a.only == "mixedbatch":
        qs = mixed_queries()
        eng.run_batch(qs)
        t0 = time.perf_counter()
        got, launches = eng.run_batch(qs)
        log("mixed batch: %.3e positions/s, %d graph launch(es)"
            % (sum(hi - lo for _, _, lo, hi, _ in qs) / (time.perf_counter() - t0), launches))
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
                     "kernel": "k_flow<4, 0, alt, deal blocks>: one launch per batch of fewer than 2^32 deals "
                               "(whole tiles of 1536: 4293918720 and 1933102080, 2 launches per step)",
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
                             "staged_frac": bt * STAGED_VALU_OPS_PER_TURN / (bf["staged"]["device_ms"] / 1e3) / peak_lane_ops,
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
                               "device_ms": r["device_ms"], "persistent_over_plain": value / (n13 / pw),
                               "brent_flow_over_plain": r["device_ms"] / bf["flow"]["device_ms"],
                               "note": "the plain kernel runs Brent with its own fastest turn (the flow kernel's "
                                       "turn_top measured 18 % slower in it); brent_flow_over_plain compares the two "
                                       "kernels on the same detector, device time against device time"}
        log("n13 plain: %.3e deals/s" % (n13 / pw))

        # ---- n = 14 canonical scan, persistent, whole range: flow kernel, then the staged pipeline
        t0 = time.perf_counter()
        r = eng.run_range(14, "canon", hbins=256)
        cw = time.perf_counter() - t0
        assert r["cyc"] == C14["cyc"] and r["maxfin"] == C14["maxfin"] and hist_of(r) == C14["hist"]
        tot14 = math.factorial(13) * 13
        rs = eng.run_range(14, "canon", hbins=256, tuning={"stage_kind": "lockstep"})
        assert (rs["cyc"], rs["maxfin"], rs["turns"]) == (r["cyc"], r["maxfin"], r["turns"])
        # the flow kernel deals each canonical position after its first turn (player 1's top card n - 1
        # always wins it) and counts the positions that end on it without playing them: of Brent's
        # turns it plays all but one per position
        played = r["turns"] - tot14
        canon = {"workload": "n14_canon: all 13!*13 = 80951270400 positions", "persistent_value": tot14 / cw,
                 "unit": "positions/s", "persistent_device_ms": r["device_ms"],
                 "frac": played * VALU_OPS_PER_TURN / (r["device_ms"] / 1e3) / peak_lane_ops,
                 "frac_brent_turns": r["turns"] * VALU_OPS_PER_TURN / (r["device_ms"] / 1e3) / peak_lane_ops,
                 "frac_note": "frac: the turns the flow kernel plays (Brent's turns minus each position's first "
                              "turn, which the deal includes); frac_brent_turns: on Brent's own count",
                 "staged_device_ms": rs["device_ms"], "turns": r["turns"], "turns_played_flow": played,
                 "parity": "cycling 2872129173, longest 465, hist {96,112} equal the published values"}
        # both kernels on the same sampled chunks (seed 14, whole q-ranks): batched / plain ratio
        ch = canon14_chunks(tot14)
        tp = tq = 0.0
        for lo, hi in ch:
            rp = eng.run_range(14, "canon", lo, hi, kernel="plain")
            rq = eng.run_range(14, "canon", lo, hi, kernel="persistent")
            assert (rp["cyc"], rp["maxfin"], rp["turns"]) == (rq["cyc"], rq["maxfin"], rq["turns"])
            tp += rp["device_ms"]; tq += rq["device_ms"]
        npos = sum(hi - lo for lo, hi in ch)
        canon.update({"sample": "8 chunks of 13*2^22 positions (whole q-ranks), seed 14",
                      "plain_sample_value": npos / (tp / 1e3), "persistent_sample_value": npos / (tq / 1e3),
                      "persistent_over_plain": tp / tq, "tuning": eng.tuning_for(14, "canon"),
                      "tiles": "whole scan: lane runs of six whole orders (78 positions, tiles of 4992), the "
                               "engine's choice for single canonical ranges of at least 2^31 positions; the "
                               "sampled chunks: two orders (26 positions, tiles of 1664)"})
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

        # ---- mixed-n batch: 256 requests of n = 5..15 as one graph launch, autotuned
        out["mixed_batch"] = mixed_batch(eng)
        log("mixed batch: %.3e positions/s in %d graph launch(es); one by one %.3e"
            % (out["mixed_batch"]["value"], out["mixed_batch"]["graph_launches"],
               out["mixed_batch"]["one_by_one_persistent_positions_per_s"]))

        # ---- CPU baseline: the oracle on sampled chunks of the same workloads, compared chunk by chunk
        out["cpu_baseline"] = cpu_legs(eng, a.cpu_chunks)
        out["cpu_baseline"]["c3_n14_canon"]["gpu_over_cpu"] = (out["canon14"]["persistent_value"] /
                                                                out["cpu_baseline"]["c3_n14_canon"]["value"])
    print(json.dumps(out), flush=True)


def served_equal(a, b, keys):
    for k in keys:
        if a[k] != b[k]:
            return k
    return None


def pct(lat, p):
    lat = sorted(lat)
    return lat[min(len(lat) - 1, max(0, int(math.ceil(p * len(lat))) - 1))]


def serving_queries(nreq=512, seed=20261205):
    """The serving workload's queries (numpy PCG64): n uniform over 7..17, full or canonical
    (canonical n <= 13 with distinct), length floor(10^U(3, 10)) clipped to the total, lo uniform.
    Returns them and the generator (its later draws time the arrivals)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    qs = []
    for _ in range(nreq):
        n = int(rng.integers(7, 18))
        mode = "canon" if rng.random() < 0.5 else "full"
        total = math.factorial(n - 1) * (n - 1) if mode == "canon" else math.factorial(n)
        ln = min(total, int(10 ** rng.uniform(3, 10)))
        lo = int(rng.integers(0, total - ln + 1))
        kw = {"hbins": 128}
        if mode == "canon" and n <= 13:
            kw.update(distinct=True)
        qs.append((n, mode, lo, lo + ln, kw))
    return qs, rng


def serve_burst(eng, qs, slots=64, step=1 << 26):
    """All queries submitted at once to the serving loop; returns (results, wall seconds)."""
    eng.serve_start(max_slots=slots, step_positions=step)
    t0 = time.perf_counter()
    try:
        ids = [eng.submit(n, mode, lo, hi, **kw) for n, mode, lo, hi, kw in qs]
        res = [eng.wait(i) for i in ids]
    finally:
        eng.serve_stop()
    return res, time.perf_counter() - t0


def serving(eng, nreq=512, seed=20261205, fractions=(0.25, 0.5, 0.9), slots=64):
    """The serving workload: `nreq` queries, n uniform over 7..17, mode full or canonical (canonical
    n <= 13 with distinct), length floor(10^U(3, 10)) clipped to the total, lo uniform (numpy PCG64).
    (1) The queries one by one, synchronously, through the persistent kernel (the parity reference of
    every later run). (2) All submitted at once (saturation): aggregate positions/s and the saturated
    request rate. (3) Poisson arrivals at fractions of that rate: p50 / p99 latency (submit to
    result) per rate. (4) The queries one by one through the plain kernel. Every result must equal
    its one-by-one persistent result (tolerance 0); the first difference stops the bench."""
    qs, rng = serving_queries(nreq, seed)
    npos = sum(hi - lo for _, _, lo, hi, _ in qs)
    keys = ("cyc", "done", "maxfin", "turns", "hist", "distinct_minkey", "undetermined")
    alone, alone_lat = [], []
    t0 = time.perf_counter()
    for n, mode, lo, hi, kw in qs:
        t1 = time.perf_counter()
        alone.append(eng.run_range(n, mode, lo, hi, **kw))
        alone_lat.append(1e3 * (time.perf_counter() - t1))
    t_alone = time.perf_counter() - t0
    log("serving: %d queries, %d positions; one by one (persistent) %.1f s" % (nreq, npos, t_alone))

    def run_served(gaps, step=1 << 26):
        eng.serve_start(max_slots=slots, step_positions=step)
        ids = []
        t0 = time.perf_counter()
        try:
            for q, g in zip(qs, gaps):
                if g:
                    time.sleep(g)
                n, mode, lo, hi, kw = q
                ids.append(eng.submit(n, mode, lo, hi, **kw))
            res = [eng.wait(i) for i in ids]
        finally:
            eng.serve_stop()
        wall = time.perf_counter() - t0
        for q, r, a in zip(qs, res, alone):
            bad = served_equal(r, a, keys)
            assert bad is None, ("served result differs from the same query alone", q, bad)
        return res, wall

    burst, wall_b = run_served([0.0] * nreq)
    sat_rate = nreq / wall_b
    lb = [r["latency_ms"] for r in burst]
    out = {"workload": "%d queries, n uniform 7..17, full or canonical (canonical n <= 13 with distinct), length "
                       "floor(10^U(3, 10)) clipped to the total, lo uniform, numpy PCG64 seed %d: %d positions"
                       % (nreq, seed, npos),
           "slots": slots, "step_positions": 1 << 26,
           "saturated": {"positions_per_s": npos / wall_b, "requests_per_s": sat_rate, "wall_s": wall_b,
                         "p50_ms": pct(lb, 0.5), "p99_ms": pct(lb, 0.99)},
           "rates": []}
    log("serving saturated: %.3e positions/s, %.1f req/s" % (npos / wall_b, sat_rate))
    # the step size trades the latency of a newly arrived query (it waits for the steps in flight)
    # against the per-step cost; the same burst with 2^28-position steps
    burst28, wall_28 = run_served([0.0] * nreq, step=1 << 28)
    l28 = [r["latency_ms"] for r in burst28]
    out["saturated_step_2e28"] = {"step_positions": 1 << 28, "positions_per_s": npos / wall_28, "wall_s": wall_28,
                                  "p50_ms": pct(l28, 0.5), "p99_ms": pct(l28, 0.99)}
    log("serving saturated, 2^28-position steps: %.3e positions/s" % (npos / wall_28))
    for f in fractions:
        rate = f * sat_rate
        gaps = list(rng.exponential(1.0 / rate, nreq))
        res, wall = run_served(gaps)
        lat = [r["latency_ms"] for r in res]
        out["rates"].append({"fraction_of_saturation": f, "arrival_rate_per_s": rate,
                             "offered_positions_per_s": rate * npos / nreq, "p50_ms": pct(lat, 0.5),
                             "p99_ms": pct(lat, 0.99), "max_ms": max(lat), "wall_s": wall,
                             "achieved_positions_per_s": npos / wall})
        log("serving at %.1f req/s: p50 %.2f ms, p99 %.2f ms" % (rate, pct(lat, 0.5), pct(lat, 0.99)))
    t0 = time.perf_counter()
    plain_lat = []
    for q, a in zip(qs, alone):
        n, mode, lo, hi, kw = q
        t1 = time.perf_counter()
        p = eng.run_range(n, mode, lo, hi, kernel="plain", **kw)
        plain_lat.append(1e3 * (time.perf_counter() - t1))
        bad = served_equal(p, a, keys)
        assert bad is None, ("plain kernel differs", q, bad)
    t_plain = time.perf_counter() - t0
    out["one_by_one_persistent"] = {"positions_per_s": npos / t_alone, "p50_ms": pct(alone_lat, 0.5),
                                    "p99_ms": pct(alone_lat, 0.99)}
    out["one_by_one_plain"] = {"positions_per_s": npos / t_plain, "p50_ms": pct(plain_lat, 0.5),
                               "p99_ms": pct(plain_lat, 0.99)}
    out["parity"] = ("every served result (saturated and at each rate) and every plain-kernel result equals the "
                     "same query run alone on the persistent kernel (cyc, done, maxfin, turns, hist, minkey)")
    log("serving one by one: plain %.3e, persistent %.3e positions/s" % (npos / t_plain, npos / t_alone))
    return out


def mixed_queries(nreq=256, seed=20261203):
    """The mixed-n batch's requests (numpy PCG64): n uniform over 5..15, full or canonical (canonical
    n <= 13 with distinct + dhist), length floor(10^U(3, 9)) clipped to the total, lo uniform."""
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
        if mode == "canon" and n <= 13:
            kw.update(distinct=True, dbins=256)
        qs.append((n, mode, lo, lo + ln, kw))
    return qs


def mixed_batch(eng, nreq=256, seed=20261203):
    """The mixed-n single-launch workload: seeded numpy PCG64, `nreq` requests, n uniform over 5..15,
    full or canonical (canonical n <= 13 with distinct + dhist), length floor(10^U(3, 9)) clipped to the
    total, lo uniform. The schedule is autotuned first on the same requests clipped to 2^24 positions
    (13 % of the batch's positions; every candidate must reproduce the plain kernel's counts), then the
    whole batch runs with it as
    one war_run_batch call (every kernel launch a branch of one hipGraph), then one by one through
    war_run_range: every count must be identical."""
    qs = mixed_queries(nreq, seed)
    npos = sum(hi - lo for _, _, lo, hi, _ in qs)
    sample = [(n, m, lo, min(hi, lo + (1 << 24)), kw) for n, m, lo, hi, kw in qs]
    t0 = time.perf_counter()
    tun = eng.autotune_batch(sample, budget_ms=30000)
    t_tune = time.perf_counter() - t0
    log("mixed batch: autotuned in %.1f s: %s" % (t_tune, tun))
    tuned = [(n, m, lo, hi, dict(kw, tuning=tun)) for n, m, lo, hi, kw in qs]
    res = {}
    for name, batch in (("default", qs), ("tuned", tuned)):
        eng.run_batch(batch)                                # graphs of the batch's shape captured
        t0 = time.perf_counter()
        got, launches = eng.run_batch(batch)
        res[name] = (got, launches, time.perf_counter() - t0)
    t = 0.0
    for q, g, g2 in zip(qs, res["tuned"][0], res["default"][0]):
        n, mode, lo, hi, kw = q
        t1 = time.perf_counter()
        p = eng.run_range(n, mode, lo, hi, **kw)
        t += time.perf_counter() - t1
        for k in ("cyc", "done", "maxfin", "turns", "hist", "lenhist", "oncycle", "distinct_minkey", "dhist"):
            assert p[k] == g[k] == g2[k], (q, k)
    got, launches, wall = res["tuned"]
    return {"workload": "%d requests, n uniform 5..15, full or canon (canonical n <= 13 with distinct + dhist), "
                        "length floor(10^U(3, 9)) clipped, numpy PCG64 seed %d: %d positions" % (nreq, seed, npos),
            "value": npos / wall, "unit": "positions/s", "graph_launches": launches, "wall_s": wall,
            "tuning": tun, "autotune_s": t_tune,
            "autotune_sample": "the same requests clipped to 2^24 positions each (13 % of the positions); "
                               "candidates checked against the plain kernel's counts; a schedule replaces the "
                               "best so far only if at least 2 % faster",
            "default_tuning_value": npos / res["default"][2], "default_graph_launches": res["default"][1],
            "one_by_one_persistent_positions_per_s": npos / t,
            "parity": "every count of every request identical in the tuned batch, the default batch and alone"}


def host_cpus():
    """The host's CPUs as this process may use them: nproc, the affinity set, the cgroup CPU quota
    and the model name."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()
            if q != "max":
                quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["cpu_model"] = model
    usable = info["affinity"]
    if quota:
        usable = min(usable, max(1, int(quota)))
    info["usable"] = usable
    return info


def cpu_legs(eng, n13_chunks=64):
    """The CPU baseline: the oracle (C++/OpenMP restatement of cpu/war.cpp, cpu/war_all.cpp) on every
    CPU this process may use (the affinity set within the cgroup quota), on bounded samples of the
    same workloads: n = 13 deals (64 chunks of 2^22 ranks spread over 13!, seed 13), all n = 9 deals,
    and n = 14 canonical positions (64 chunks of 2^20 q-ranks spread over 13!, seed 14); every chunk
    is compared with the GPU's counts."""
    hc = host_cpus()
    threads = hc["usable"]
    n13 = math.factorial(13)
    ch = chunks_spread(n13, n13_chunks, 1 << 22, 13)
    cres, cel = cpu_baseline(13, "full", ch, threads)
    for (lo, hi), c in zip(ch, cres):
        g = eng.run_range(13, "full", lo, hi, hbins=256)
        assert (g["cyc"], g["maxfin"], g["turns"], g["hist"]) == (c["cyc"], c["maxfin"], c["turns"], c["hist"])
    c2 = sum(hi - lo for lo, hi in ch) / cel
    reps = 20
    r9, el9 = cpu_baseline(9, "full", [(0, math.factorial(9))] * reps, threads)
    assert r9[0]["cyc"] == 218680 and r9[0]["maxfin"] == 28
    q14 = math.factorial(13)
    qch = chunks_spread(q14, 64, 1 << 20, 14)
    pch = [(13 * lo, 13 * hi) for lo, hi in qch]
    cres, el14 = cpu_baseline(14, "canon", pch, threads)
    for (lo, hi), c in zip(pch, cres):
        g = eng.run_range(14, "canon", lo, hi, hbins=256)
        assert (g["cyc"], g["maxfin"], g["turns"], g["hist"]) == (c["cyc"], c["maxfin"], c["turns"], c["hist"])
    c3 = sum(hi - lo for lo, hi in pch) / el14
    out = {"value": c2, "unit": "deals/s", "cores": threads, "kind": "port",
           "sample": "%d chunks of 2^22 n=13 deal ranks spread over 13! (seed 13), %d deals, each chunk equal to "
                     "the GPU's counts" % (len(ch), len(ch) << 22),
           "host": hc,
           "threads_note": "OpenMP threads = the CPUs this process may use (affinity set within the cgroup quota)",
           "c1_n9_full": {"value": reps * math.factorial(9) / el9, "unit": "deals/s",
                          "sample": "all 9! deals, %d times; a(9) = 218680, longest 28" % reps},
           "c3_n14_canon": {"value": c3, "unit": "positions/s",
                            "sample": "64 chunks of 2^20 q-ranks (13 * 2^20 positions each) spread over 13! (seed "
                                      "14), each chunk equal to the GPU's counts",
                            "gpu_over_cpu": None}}
    log("cpu baseline (%d threads of %s): n13 %.3e deals/s, n9 %.3e, n14 canon %.3e positions/s"
        % (threads, hc["cpu_model"], c2, out["c1_n9_full"]["value"], c3))
    ref = reference_program(eng, threads)
    if ref:
        out["reference_program"] = ref
    return out


def reference_program(eng, threads, n=12):
    """The reference's own CPU program (cpu/war.cpp, compiled from its source into oracle/_ref/war
    by oracle/Makefile; present where the build had the reference tree) over every deal of n = 12
    (its command line takes a whole n; 4.79e8 deals), timed on the same threads, beside the oracle
    on the same range; its line is compared with the GPU's counts. None if the program is absent."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "war")
    if not os.access(exe, os.X_OK):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.time()
    try:
        p = subprocess.run([exe, str(n), "0"], env=env, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                           timeout=300, check=True, text=True)
    except (OSError, subprocess.SubprocessError) as e:
        log("reference cpu/war.cpp not timed: %s" % e)
        return None
    el = time.time() - t0
    f = dict(kv.split("=") for kv in p.stdout.split() if "=" in kv)
    g = eng.run_range(n, "full")
    assert int(f["deals"]) == math.factorial(n) == g["done"]
    assert (int(f["cycling"]), int(f["longest_finite"])) == (g["cyc"], g["maxfin"]), (p.stdout, g)
    res, elp = cpu_baseline(n, "full", [(0, math.factorial(n))], threads)
    assert (res[0]["cyc"], res[0]["maxfin"]) == (g["cyc"], g["maxfin"])
    out = {"value": math.factorial(n) / el, "unit": "deals/s", "cores": threads, "kind": "reference",
           "sample": "oracle/_ref/war %d 0: every n = %d deal (%d), the reference's cpu/war.cpp built from its "
                     "source; its line (cycling %s, longest finite %s) equal to the GPU's" %
                     (n, n, math.factorial(n), f["cycling"], f["longest_finite"]),
           "port_value": math.factorial(n) / elp, "port_over_reference": el / elp}
    log("reference cpu/war.cpp (%d threads): n%d %.3e deals/s; the oracle on the same deals %.3e"
        % (threads, n, out["value"], out["port_value"]))
    return out


if __name__ == "__main__":
    main()
