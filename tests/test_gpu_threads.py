"""Calls from several host threads at once: direct ranges, batches, no-allocation calls and the serving
loop's submit / wait, each from its own thread (the ctypes calls release the interpreter lock, so the
calls overlap inside the library), plus a thread that keeps making invalid calls. The engine mutex
serialises the GPU work and each entry point makes the engine's device current on its calling thread;
every result equals the oracle's, and a failed call's message stays with the thread that made it
(war_last_error is empty on a thread whose last call succeeded)."""
import threading

import numpy as np
import pytest

from test_gpu_sweep import check, draw

pytestmark = pytest.mark.gpu


def test_concurrent_entry_points(eng, orc):
    from war_amd import WarError
    rng = np.random.Generator(np.random.PCG64(20261209))
    singles = [[draw(rng) for _ in range(120)] for _ in range(3)]
    batches = [[draw(rng, "batch") for _ in range(int(rng.integers(1, 9)))] for _ in range(40)]
    noallocs = [draw(rng, "noalloc") for _ in range(120)]
    served = [draw(rng, "serve") for _ in range(120)]
    out = {}
    errors = []
    bad_msgs = []
    stop = threading.Event()

    def guard(fn):
        def run():
            try:
                fn()
            except BaseException as e:     # reported by the main thread
                errors.append(e)
        return run

    def single(t, qs):
        for i, (n, mode, lo, hi, kw, _) in enumerate(qs):
            out[("single", t, i)] = eng.run_range(n, mode, lo, hi, **kw)
            assert eng.lib.war_last_error() == b""

    def batch():
        for b, qs in enumerate(batches):
            got, _ = eng.run_batch([(n, mode, lo, hi, kw) for n, mode, lo, hi, kw, _ in qs])
            for i, g in enumerate(got):
                out[("batch", b, i)] = g

    def noalloc():
        for i, (n, mode, lo, hi, kw, _) in enumerate(noallocs):
            out[("noalloc", i)] = eng.run_range_noalloc(n, mode, lo, hi, **kw)

    def serve():
        ids = [eng.submit(n, mode, lo, hi, **kw) for n, mode, lo, hi, kw, _ in served]
        for i, r in enumerate(ids):
            out[("serve", i)] = eng.wait(r)

    def invalid():
        while not stop.is_set():
            with pytest.raises(WarError) as ei:
                eng.run_range(13, "half", 0, 10)            # half mode needs even n
            bad_msgs.append(str(ei.value))
            assert b"half mode needs even n" in eng.lib.war_last_error()

    eng.serve_start(max_slots=8, step_positions=1 << 18)
    try:
        threads = [threading.Thread(target=guard(lambda t=t: single(t, singles[t]))) for t in range(3)]
        threads += [threading.Thread(target=guard(f)) for f in (batch, noalloc, serve)]
        bad = threading.Thread(target=guard(invalid))
        print("threads start", len(threads) + 1, flush=True)
        for th in threads + [bad]:
            th.start()
        for th in threads:
            th.join(timeout=240)
        stop.set()
        bad.join(timeout=60)
        assert not any(th.is_alive() for th in threads + [bad]), "a thread did not finish"
    finally:
        eng.serve_stop()
    print("threads done", len(out), "results", len(bad_msgs), "invalid calls", flush=True)
    assert not errors, errors
    assert bad_msgs and all("half mode needs even n" in m for m in bad_msgs)
    for t in range(3):
        for i, (n, mode, lo, hi, kw, okw) in enumerate(singles[t]):
            check(("single", t, i), out[("single", t, i)], orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)
    for b, qs in enumerate(batches):
        for i, (n, mode, lo, hi, kw, okw) in enumerate(qs):
            check(("batch", b, i), out[("batch", b, i)], orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)
    for i, (n, mode, lo, hi, kw, okw) in enumerate(noallocs):
        check(("noalloc", i), out[("noalloc", i)], orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)
    for i, (n, mode, lo, hi, kw, okw) in enumerate(served):
        check(("serve", i), out[("serve", i)], orc.run(n, mode, lo, hi, **okw), n, mode, lo, hi, kw, okw)
