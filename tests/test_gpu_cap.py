"""The step cap on the GPU (war_request.step_cap): games unresolved within the cap are counted as
undetermined, not an error, exactly as the oracle counts them (tolerance 0), on the plain kernel
(Brent, the alternation detector and the streak test, whose cap keeps the reference's "more than
cap turns" form, gpu/war24.py:103) and on the persistent flow kernel (caps that are multiples of
its spans), alone, in batches and in the serving loop; the refusals; the opt-in error; and the
war24.py drop-in printing undetermined=N."""
import pytest

from war_amd import WarError

pytestmark = pytest.mark.gpu

F = ("cyc", "done", "maxfin", "turns", "undetermined", "hist", "lenhist")


def check(eng, orc, n, mode, lo, hi, cap, kernel, detector="brent", rule=0):
    print("case", n, mode, lo, hi, cap, kernel, detector, rule, flush=True)
    hb = 0 if detector == "streak" else 128
    want = orc.run(n, mode, lo, hi, rule=rule, hbins=hb, lbins=256, detector=detector, step_cap=cap)
    got = eng.run_range(n, mode, lo, hi, rule=rule, kernel=kernel, hbins=hb, lbins=256, detector=detector,
                        step_cap=cap)
    for k in F:
        if k == "hist" and not hb:
            continue
        assert got[k] == want[k], (n, mode, cap, kernel, detector, k, got[k], want[k])
    return got


@pytest.mark.parametrize("kernel", ["plain", "persistent"])
@pytest.mark.parametrize("n,mode,lo,hi,cap,det,rule", [
    (9, "full", 0, None, 64, "brent", 0),
    (9, "full", 0, None, 32, "alt", 0),
    (12, "canon", 10**7, 10**7 + 3_000_000, 32, "brent", 0),
    (11, "full", 123457, 123457 + 2_000_000, 96, "brent", 1),
    (13, "full", 10**9, 10**9 + 1_000_000, 64, "alt", 0),
    (14, "canon", 5 * 10**9, 5 * 10**9 + 1_000_000, 128, "brent", 0),
    (17, "full", 10**14, 10**14 + 300_000, 96, "brent", 0),
])
def test_cap_matches_oracle(eng, orc, kernel, n, mode, lo, hi, cap, det, rule):
    got = check(eng, orc, n, mode, lo, hi, cap, kernel, det, rule)
    assert got["undetermined"] > 0     # the caps cut into the game lengths


@pytest.mark.parametrize("cap", [20, 45, 1 << 22])
def test_streak_cap_plain(eng, orc, cap):
    check(eng, orc, 9, "full", 0, None, cap, "plain", "streak")
    check(eng, orc, 13, "full", 10**9, 10**9 + 500_000, cap, "plain", "streak")


def test_cap_refusals_and_error_flag(eng):
    with pytest.raises(WarError, match="multiple of 32"):
        eng.run_range(9, "full", kernel="persistent", step_cap=50)
    with pytest.raises(WarError, match="flow schedule"):
        eng.run_range(9, "full", kernel="persistent", step_cap=64, tuning={"stage_kind": "lockstep"})
    with pytest.raises(WarError, match="2\\^24"):
        eng.run_range(9, "full", step_cap=(1 << 24) + 32)
    with pytest.raises(WarError, match="not resolved within 64 turns"):
        eng.run_range(9, "full", step_cap=64, undet_error=True)
    # the default cap: every valid game resolves
    r = eng.run_range(9, "full")
    assert r["undetermined"] == 0


def test_cap_in_batch_and_serving(eng, orc):
    qs = [(9, "full", 0, 362880, {"hbins": 64, "step_cap": 64}),
          (9, "full", 0, 362880, {"hbins": 64}),
          (12, "canon", 10**6, 3 * 10**6, {"hbins": 64, "step_cap": 32}),
          (11, "full", 10**6, 2 * 10**6, {"hbins": 64, "detector": "alt", "step_cap": 32}),
          (13, "canon", 10**8, 10**8 + 10**6, {"hbins": 64, "distinct": True, "dbins": 64})]
    want = [orc.run(n, m, lo, hi, hbins=64, detector=kw.get("detector", "brent"), step_cap=kw.get("step_cap", 0),
                    distinct=kw.get("distinct", False), dbins=kw.get("dbins", 0)) for n, m, lo, hi, kw in qs]
    got, _ = eng.run_batch(qs)
    eng.serve_start(max_slots=8, step_positions=1 << 20)
    try:
        ids = [eng.submit(n, m, lo, hi, **kw) for n, m, lo, hi, kw in qs]
        served = [eng.wait(i) for i in ids]
    finally:
        eng.serve_stop()
    for q, w, g, s in zip(qs, want, got, served):
        for k in ("cyc", "done", "maxfin", "turns", "undetermined", "hist"):
            assert g[k] == w[k] == s[k], (q, k, g[k], s[k], w[k])


def test_war24_prints_undetermined(eng, orc, capsys, tmp_path, monkeypatch):
    import war24
    monkeypatch.setattr(war24, "_HERE", str(tmp_path))
    for det, cap in (("streak", 40), ("floyd", 64)):
        cyc, mx, deals, und, _, line = war24.run(9, det, 0, 362880, step_cap=cap, log=False)
        want = orc.run(9, "full", detector="streak" if det == "streak" else "brent", step_cap=cap)
        assert (cyc, mx, deals, und) == (want["cyc"], want["maxfin"], want["done"], want["undetermined"])
        assert und > 0 and ("undetermined=%d " % und) in line
