"""Transcribe the reference's published results into tests/golden/goldens.json.

Reads only the reference's data files (text), never its code:
  data/b_file.txt, data/deals.tsv, data/canonical_positions.tsv, data/loser_first.tsv,
  data/checks.txt, data/long_runs.txt
Run it where the reference tree is present:  python tests/golden/make_goldens.py [REFDIR]
The output is committed; the tests only read the JSON.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
D = os.path.join(REF, "data")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "goldens.json")


def lines(name):
    with open(os.path.join(D, name)) as f:
        return f.read().splitlines()


def hist_pairs(s):
    """'20:98503 30:120177' -> {'20': 98503, '30': 120177}"""
    return {str(int(k)): int(v) for k, v in re.findall(r"(\d+)\s*:\s*(\d+)", s)}


def main():
    g = {"source": "reference data/ files (text), transcribed by tests/golden/make_goldens.py"}

    # a(n), n = 1..17 (data/b_file.txt)
    g["a"] = {}
    for ln in lines("b_file.txt"):
        if ln.strip():
            n, v = ln.split()
            g["a"][n] = int(v)

    # deals: n, deals, never_end, longest, lambda-hist (data/deals.tsv)
    g["deals"] = {}
    for ln in lines("deals.tsv"):
        if ln.startswith("#") or not ln.strip():
            continue
        f = ln.split("\t")
        row = {"deals": int(f[1]), "never_end": int(f[2])}
        if len(f) > 3 and f[3].strip():
            row["longest"] = int(f[3])
        if len(f) > 4 and f[4].strip():
            row["hist"] = hist_pairs(f[4])
        g["deals"][f[0]] = row

    # canonical positions (data/canonical_positions.tsv)
    g["canon"] = {}
    for ln in lines("canonical_positions.tsv"):
        if ln.startswith("#") or not ln.strip():
            continue
        f = ln.split("\t")
        row = {"positions": int(f[1])}
        if f[2].strip():
            row["cycling"] = int(f[2])
        row["on_cycle"] = int(f[3])
        row["cycles"] = int(f[4])
        g["canon"][f[0]] = row

    # rule 1 deals (data/loser_first.tsv)
    g["rule1_deals"] = {}
    for ln in lines("loser_first.tsv"):
        if ln.startswith("#") or not ln.strip():
            continue
        f = ln.split("\t")
        g["rule1_deals"][f[0]] = {"deals": int(f[1]), "never_end": int(f[2]), "longest": int(f[3])}

    # checks.txt: canonical scans with weights and shape; rule-1 canon; distinct (min-key); half
    chk = lines("checks.txt")
    section = ""
    g["canon_gpu"] = {}        # rule 0, war_gpu.py --canon --distinct --shape output
    g["rule1_canon"] = {}
    g["distinct_minkey"] = {}  # cpu/distinct_cycles.c output
    g["half"] = {}
    cur = None
    for i, ln in enumerate(chk):
        if ln.startswith("== "):
            section = ln
            cur = None
            continue
        m = re.match(r"n=(\d+)\s+canonical positions cycling=(\d+)\s+of (\d+)\s+longest finite=(\d+) rounds", ln)
        if m:
            n, cyc, tot, lng = m.groups()
            row = {"cycling": int(cyc), "positions": int(tot), "longest": int(lng)}
            if "Losing card first" in section:
                g["rule1_canon"][n] = row
                cur = None
            elif "Theorem-3 test, n = 1..13" in section:
                g["canon_gpu"][n] = row
                cur = row
            continue
        m = re.match(r"\s+distinct cycles \(up to swapping players\): (\d+)\s+\[on-cycle canonical positions by weight: \{(.*)\}\]", ln)
        if m and cur is not None:
            cur["distinct"] = int(m.group(1))
            cur["weights"] = hist_pairs(m.group(2).replace(" ", "")) if m.group(2).strip() else {}
            continue
        m = re.match(r"\s+loop shape: (\d+) cycling positions checked for runs of exactly (\d+); not that shape: (\d+); too long to check: (\d+)", ln)
        if m and cur is not None:
            cur["shape"] = {"blk": int(m.group(2)), "nonstd": int(m.group(3)), "unchk": int(m.group(4))}
            continue
        m = re.match(r"n=(\d+): canonical positions that never end (\d+); on a cycle (\d+); DISTINCT cycles up to swap (\d+)", ln)
        if m and "Distinct cycles, second method" in section:
            n = m.group(1)
            row = {"cycling": int(m.group(2)), "on_cycle": int(m.group(3)), "distinct": int(m.group(4))}
            nxt = chk[i + 1]
            hm = re.match(r"\s+cycle lengths \(distinct cycles\):(.*)", nxt)
            row["distinct_hist"] = hist_pairs(hm.group(1)) if hm and hm.group(1).strip() else {}
            g["distinct_minkey"][n] = row
            continue
        m = re.match(r"n=(\d+) half=True deals_simulated=(\d+) \(expected (\d+)\) cycling_found=(\d+) a\(n\)=(\d+) longest_finite=(\d+)", ln)
        if m:
            g["half"][m.group(1)] = {"played": int(m.group(2)), "cycling_found": int(m.group(4)),
                                     "a": int(m.group(5)), "longest": int(m.group(6))}

    # long_runs.txt: n = 14 canonical (with weights, shape, hist), n = 16 canonical, n = 17 ranges
    lr = lines("long_runs.txt")
    g["long"] = {}
    for i, ln in enumerate(lr):
        m = re.match(r"n=14\s+canonical positions cycling=(\d+)\s+of (\d+)\s+longest finite=(\d+) rounds", ln)
        if m:
            row = {"cycling": int(m.group(1)), "positions": int(m.group(2)), "longest": int(m.group(3))}
            for j in range(i + 1, i + 4):
                dm = re.match(r"\s+distinct cycles \(up to swapping players\): (\d+)\s+\[on-cycle canonical positions by weight: \{(.*)\}\]", lr[j])
                if dm:
                    row["distinct"] = int(dm.group(1))
                    row["weights"] = hist_pairs(dm.group(2).replace(" ", ""))
                sm = re.match(r"\s+loop shape: (\d+) cycling positions checked for runs of exactly (\d+); not that shape: (\d+); too long to check: (\d+)", lr[j])
                if sm:
                    row["shape"] = {"blk": int(sm.group(2)), "nonstd": int(sm.group(3)), "unchk": int(sm.group(4))}
                hm = re.match(r"\s+loop lengths \(rounds: positions\): (.*)", lr[j])
                if hm:
                    row["hist"] = hist_pairs(hm.group(1))
            g["long"]["canon14"] = row
        m = re.match(r"n=16\s+canonical positions cycling=(\d+)\s+of (\d+)\s+longest finite=(\d+) rounds\s+([\d.]+)s$", ln)
        if m and "canon16_rule0" not in g["long"]:
            g["long"]["canon16_rule0"] = {"cycling": int(m.group(1)), "positions": int(m.group(2)), "longest": int(m.group(3))}
        m = re.match(r"n=14\s+a\(n\)=(\d+)\s+of (\d+)\s+longest finite=(\d+) rounds\s+([\d.]+)s\s+\(half\)", ln)
        if m and "half14_rule0" not in g["long"]:
            g["long"]["half14_rule0"] = {"a": int(m.group(1)), "deals": int(m.group(2)), "longest": int(m.group(3))}
        m = re.match(r"n=17 det=(floyd|streak) range=\[(\d+),(\d+)\) deals_simulated=(\d+) \(expected (\d+)\) cycling=(\d+) longest_finite=(\d+)", ln)
        if m:
            lo, hi = int(m.group(2)), int(m.group(3))
            key = "n17_%d_%d" % (lo, hi)
            g["long"][key] = {"n": 17, "lo": lo, "hi": hi, "cycling": int(m.group(6)), "longest": int(m.group(7))}

    # war_gpu.py command lines with their logged output (timing fields differ by hardware)
    def block(lines_, start_pat, stop_pat):
        out, on = [], False
        for ln in lines_:
            if on and (ln.startswith("$") or ln.startswith("==") or not ln.strip() or re.match(stop_pat, ln)):
                break
            if on:
                out.append(ln)
            if ln.startswith("$") and start_pat in ln:
                on = True
        return out
    g["cli"] = []
    deals = block(chk, "war_gpu.py 13 --nmin 1 --hist 256", r"^$")
    g["cli"].append({"args": ["13", "--nmin", "1", "--hist", "256"], "lines": deals, "source": "data/checks.txt:5-24"})
    canon = block(chk, "war_gpu.py 4 --nmin 1 --canon --distinct --shape", r"^$")
    g["cli"].append({"args": ["4", "--nmin", "1", "--canon", "--distinct", "--shape"],
                     "lines": [ln for ln in canon if not re.match(r"n=(\d+)", ln) or int(re.match(r"n=(\d+)", ln).group(1)) <= 4][:12],
                     "source": "data/checks.txt:146-157"})
    g["cli"].append({"args": ["13", "--nmin", "5", "--canon", "--distinct", "--shape"], "lines": canon[12:],
                     "source": "data/checks.txt:158-182"})
    r1 = block(chk, "war_gpu.py 13 --nmin 2 --rule 1", r"^loser-first")
    g["cli"].append({"args": ["13", "--nmin", "2", "--rule", "1"], "lines": [ln for ln in r1 if ln.startswith("n=")],
                     "source": "data/checks.txt:370-381"})
    r1c = block(chk, "war_gpu.py 12 --nmin 2 --canon --rule 1", r"^n=\d+ rule=")
    g["cli"].append({"args": ["12", "--nmin", "2", "--canon", "--rule", "1"], "lines": r1c,
                     "source": "data/checks.txt:424-434"})
    c14 = block(lr, "war_gpu.py 14 --canon --distinct --shape --hist 256", r"^$")
    g["cli"].append({"args": ["14", "--canon", "--distinct", "--shape", "--hist", "256"], "lines": c14,
                     "source": "data/long_runs.txt:19-23"})
    h14 = block(lr, "war_gpu.py 14 --half --hist 256", r"^n=15")
    g["cli"].append({"args": ["14", "--half", "--hist", "256"], "lines": h14[:2], "source": "data/long_runs.txt:13-15"})

    # war24.py result lines (data/long_runs.txt): the command line is rebuilt from the line itself
    g["cli24"] = []
    for i, ln in enumerate(lr):
        if re.match(r"n=\d+ det=(floyd|streak) range=\[\d+,\d+\) deals_simulated=", ln):
            g["cli24"].append({"line": ln, "source": "data/long_runs.txt:%d" % (i + 1)})

    # cycle_length doubled j k (data/checks.txt, "Cycle lengths of every family member with n <= 255")
    g["cycle_length"] = []
    i0 = next(i for i, ln in enumerate(chk) if "./cycle_length doubled $j $k" in ln)
    cur = None
    for i in range(i0 + 1, len(chk)):
        ln = chk[i]
        if not ln.strip() or ln.startswith("=="):
            break
        m = re.match(r"doubled family j=(\d+) k=(\d+)$", ln)
        if m:
            cur = {"args": ["doubled", m.group(1), m.group(2)], "lines": [ln], "source": "data/checks.txt:%d" % (i + 1)}
            g["cycle_length"].append(cur)
        else:
            cur["lines"].append(ln)

    # reach_cycles n threads (data/checks.txt, "Cycles entered by an ordinary deal")
    g["reach"] = {}
    for i, ln in enumerate(chk):
        m = re.match(r"n=(\d+): distinct cycles up to swap (\d+) \| reached by a deal (\d+) \| NOT reached (\d+) \| "
                     r"undetermined \(cap (\d+)\) (\d+)$", ln)
        if m:
            g["reach"][m.group(1)] = {"line": ln, "source": "data/checks.txt:%d" % (i + 1)}

    with open(OUT, "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
