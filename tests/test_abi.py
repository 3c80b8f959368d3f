"""The C ABI library (no GPU needed): it loads, exports every entry point include/war_engine.h
declares, its structs have the layout the header gives (checked against gcc on the header), and it
reports errors instead of aborting."""
import ctypes
import os
import subprocess

import pytest

import war_amd
from war_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header():
    lib = war_amd.load()
    names = war_amd.header_functions()
    assert len(names) >= 14
    for nm in names:
        assert hasattr(lib, nm), nm
    assert lib.war_abi_version() == 3


def test_no_python_fallback(monkeypatch, tmp_path):
    with pytest.raises(OSError):
        _lib.load(str(tmp_path / "missing.so"))


STRUCTS = {
    "war_tuning": (_lib.WarTuning, ["waves_per_cu", "ranks_per_refill", "brent_cap", "plain_per", "stage_kind",
                                    "pile_layout"]),
    "war_request": (_lib.WarRequest, ["n", "mode", "rule", "kernel", "lo", "hi", "hbins", "lbins", "dbins",
                                      "distinct", "shape_blk", "detector", "tuning", "step_cap", "flags"]),
    "war_reach_result": (_lib.WarReachResult, ["classes", "reached", "not_reached", "undetermined", "nolen",
                                               "nbins", "device_ms", "max_depth"]),
    "war_pile_position": (_lib.WarPilePosition, ["n1", "n2", "p1", "p2"]),
    "war_cycle_info": (_lib.WarCycleInfo, ["status", "turns", "lam", "mu"]),
    "war_result": (_lib.WarResult, ["cyc", "done", "maxfin", "turns", "oncycle", "distinct_minkey",
                                    "distinct_hashset", "shape_nonstd", "shape_unchk", "undetermined", "hist",
                                    "hbins", "lenhist", "lbins", "dhist", "dbins", "device_ms", "latency_ms"]),
    "war_devinfo": (_lib.WarDevinfo, ["cus", "max_clock_mhz", "wave_size", "free_bytes", "total_bytes", "arch"]),
}


def test_struct_layout_matches_header(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "war_engine.h"', "int main(void) {"]
    for s, (_, fields) in STRUCTS.items():
        src.append('printf("%s %%zu\\n", sizeof(%s));' % (s, s))
        for f in fields:
            src.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (s, f, s, f))
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(c)])
    out = dict(line.split() for line in subprocess.check_output([str(exe)], text=True).splitlines())
    for s, (cls, fields) in STRUCTS.items():
        assert int(out[s]) == ctypes.sizeof(cls), s
        for f in fields:
            assert int(out["%s.%s" % (s, f)]) == getattr(cls, f).offset, (s, f)


def test_errors_without_engine():
    lib = war_amd.load()
    req = war_amd.make_request(9, "full")
    res = _lib.WarResult()
    rc = lib.war_run_range(ctypes.byref(req), ctypes.byref(res))
    if rc == 0:
        pytest.skip("an engine is live in this process")
    assert lib.war_last_error()
    launches = ctypes.c_int32(-1)
    reqs = (_lib.WarRequest * 1)(req)
    out = (_lib.WarResult * 1)()
    assert lib.war_run_batch(reqs, 1, out, ctypes.byref(launches)) != 0
    assert b"war_init" in lib.war_last_error()
    assert launches.value == -1


def test_make_request_totals():
    r = war_amd.make_request(14, "canon")
    assert r.hi == 80951270400 and r.mode == 2
    assert war_amd.total_positions(13, "full") == 6227020800
    assert war_amd.total_positions(1, "canon") == 0
