"""Summarise a rocprofv3 results database (kernel trace or PMC counters) as Markdown.

    python -m war_amd.profsum trace  <results.db>      per-kernel totals and per-batch stage times
    python -m war_amd.profsum pmc    <results.db>      counters per kernel (summed over dispatches)
    python -m war_amd.profsum busy   <results.db>      device occupancy: share of the span with a kernel
                                                       running, idle gaps, persistent kernels at once
    python -m war_amd.profsum issue  <results.db> [turns [valu_per_turn]]
                                                       VALU issue read from the SQ counters (the
                                                       pass of SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU,
                                                       SQ_ACTIVE_INST_VALU2, SQ_THREAD_CYCLES_VALU,
                                                       SQ_INSTS_SALU, GRBM_GUI_ACTIVE): per SIMD
                                                       quad-cycle, lanes active, scalar issue; with
                                                       turns, the useful share (turns x VALU
                                                       instructions per turn / 64 / VALU
                                                       instructions; 16.5 per turn by default, the
                                                       flow kernel's loop)

A batch of the persistent pipeline starts with k_setup, so the trace view also lists, for the
first and last batches, the duration of each of its kernels (the flow kernel, or the stage
kernels) in launch order.
"""
import collections
import sqlite3
import sys


def _short(name):
    return name.split("(")[0].replace("void ", "")


def trace(db):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, start, end from kernels order by start"))
    agg = collections.OrderedDict()
    for n, s, e in rows:
        a = agg.setdefault(_short(n), [0, 0.0])
        a[0] += 1
        a[1] += (e - s) / 1e6
    tot = sum(v[1] for v in agg.values())
    out = ["| kernel | calls | total ms | avg ms | share |", "|---|---:|---:|---:|---:|"]
    for n, (cnt, ms) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append("| `%s` | %d | %.3f | %.4f | %.1f%% |" % (n, cnt, ms, ms / cnt, 100 * ms / tot))
    batches = []
    for n, s, e in rows:
        if "k_setup" in n:
            batches.append([])
        elif batches and ("k_stage" in n or "k_refill" in n or "k_flow" in n or "k_walk" in n or "k_plain" in n):
            batches[-1].append((e - s) / 1e6)
    if batches:
        out += ["", "Per batch (ms, the batch's kernels in launch order):", ""]
        show = list(enumerate(batches))
        if len(show) > 6:
            show = show[:3] + show[-3:]
        for i, b in show:
            out.append("- batch %d: %s (sum %.2f)" % (i, " ".join("%.2f" % v for v in b), sum(b)))
        out.append("- batches: %d, mean batch %.3f ms" % (len(batches), sum(sum(b) for b in batches) / len(batches)))
    return "\n".join(out)


def pmc(db):
    c = sqlite3.connect(db)
    d = collections.OrderedDict()
    dur = collections.defaultdict(float)
    seen = set()
    for kn, cn, v, disp, du in c.execute(
            "select kernel_name, counter_name, value, dispatch_id, duration from counters_collection"):
        k = _short(kn)
        d.setdefault(k, collections.OrderedDict())
        d[k][cn] = d[k].get(cn, 0.0) + v
        if (disp, kn) not in seen:
            seen.add((disp, kn))
            dur[k] += du / 1e6
    names = sorted({cn for v in d.values() for cn in v})
    out = ["| kernel | ms | " + " | ".join(names) + " |", "|---|---:|" + "---:|" * len(names)]
    for k, v in d.items():
        out.append("| `%s` | %.2f | " % (k, dur[k]) + " | ".join("%.4g" % v.get(n, 0) for n in names) + " |")
    return "\n".join(out)


def issue(db, turns=None, simds=1024, cus=256, xcds=8, valu_per_turn=16.5):
    """VALU issue of each kernel of the run, from the SQ counters summed over its dispatches.
    GRBM_GUI_ACTIVE is summed over the XCDs (divided by `xcds` for the kernel's clocks);
    SQ_ACTIVE_INST_VALU counts the VALU instructions issued and SQ_ACTIVE_INST_VALU2 the SIMD
    quad-cycles in which two issued, so the quad-cycles with an issue are their difference."""
    c = sqlite3.connect(db)
    d = collections.OrderedDict()
    for kn, cn, v in c.execute("select kernel_name, counter_name, value from counters_collection"):
        k = _short(kn)
        d.setdefault(k, collections.defaultdict(float))[cn] += v
    out = ["| kernel | clocks | VALU per SIMD quad-cycle | quad-cycles: two / one / none | lanes active per "
           "VALU instruction | busy x lanes | SALU per CU per clock | useful share |",
           "|---|---:|---:|---:|---:|---:|---:|---:|"]
    top = max((v.get("GRBM_GUI_ACTIVE", 0) for v in d.values()), default=0)
    for k, v in d.items():
        # the kernels that hold the run (not the runtime's copies and fills)
        if not v.get("SQ_INSTS_VALU") or v.get("GRBM_GUI_ACTIVE", 0) < 0.01 * top:
            continue
        clocks = v["GRBM_GUI_ACTIVE"] / xcds
        qc = simds * clocks / 4
        iv = v["SQ_INSTS_VALU"]
        two = v.get("SQ_ACTIVE_INST_VALU2", 0) / qc
        busy = (v.get("SQ_ACTIVE_INST_VALU", iv) - v.get("SQ_ACTIVE_INST_VALU2", 0)) / qc
        lanes = v.get("SQ_THREAD_CYCLES_VALU", 0) / 64 / iv
        useful = "%.3f" % (turns * valu_per_turn / 64 / iv) if turns else "—"
        out.append("| `%s` | %.4g | %.3f | %.3f / %.3f / %.3f | %.3f | %.3f | %.3f | %s |" % (
            k, clocks, iv / qc, two, busy - two, 1 - busy, lanes, busy * lanes, v.get("SQ_INSTS_SALU", 0) / cus / clocks,
            useful))
    return "\n".join(out)


def busy(db):
    """Device occupancy over a kernel trace: the share of the traced span in which some kernel runs,
    the idle gaps between kernels, and how many persistent kernels (flow, walk) run at once."""
    c = sqlite3.connect(db)
    rows = sorted(c.execute("select name, start, end from kernels"), key=lambda r: r[1])
    if not rows:
        return "no kernels"
    t0, t1 = rows[0][1], max(r[2] for r in rows)
    on, gaps = 0, []
    s0, e0 = rows[0][1], rows[0][2]
    for _, s, e in rows[1:]:
        if s > e0:
            on += e0 - s0
            gaps.append(s - e0)
            s0, e0 = s, e
        else:
            e0 = max(e0, e)
    on += e0 - s0
    ev = []
    for n, s, e in rows:
        if "k_flow" in n or "k_walk" in n:
            ev += [(s, 1), (e, -1)]
    ev.sort()
    share = collections.Counter()
    k, last = 0, ev[0][0] if ev else 0
    for t, d in ev:
        share[k] += t - last
        k += d
        last = t
    tot = sum(share.values()) or 1
    gs = sorted(gaps) or [0]
    out = ["| traced span | some kernel running | idle gaps | gap p50 / p90 / max |", "|---:|---:|---:|---:|",
           "| %.1f ms | %.1f ms (%.1f %%) | %d, %.1f ms | %.3f / %.3f / %.3f ms |"
           % ((t1 - t0) / 1e6, on / 1e6, 100.0 * on / (t1 - t0), len(gaps), sum(gaps) / 1e6,
              gs[len(gs) // 2] / 1e6, gs[int(len(gs) * 0.9)] / 1e6, gs[-1] / 1e6),
           "", "Persistent kernels (flow, walk) running at once, share of their time: " +
           ", ".join("%d: %.3f" % (kk, v / tot) for kk, v in sorted(share.items()))]
    return "\n".join(out)


if __name__ == "__main__":
    if len(sys.argv) < 3 or sys.argv[1] not in ("trace", "pmc", "issue", "busy"):
        sys.exit(__doc__)
    if sys.argv[1] == "issue":
        print(issue(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else None,
                    valu_per_turn=float(sys.argv[4]) if len(sys.argv) > 4 else 16.5))
    elif sys.argv[1] == "busy":
        print(busy(sys.argv[2]))
    else:
        print(trace(sys.argv[2]) if sys.argv[1] == "trace" else pmc(sys.argv[2]))
