"""The C-ABI library loads and exports every entry point declared in include/paising.h
(no GPU calls)."""
import ctypes
import os
import re

import paising_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "paising.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pa_run_create" in names and "pa_server_submit" in names and len(names) > 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pa.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    assert set(declared_functions()) <= set(pa.SIGNATURES), set(declared_functions()) - set(pa.SIGNATURES)


def test_schedule_steps_host_only():
    lib = pa.lib()
    assert lib.pa_schedule_steps(0.0, 0.005, 1.0) == 199       # 0.995000...08 is the last beta
    assert lib.pa_schedule_steps(0.0, 0.002, 1.0) == 499
    assert lib.pa_schedule_steps(0.0, 0.005, 0.44) == 87
    assert lib.pa_schedule_steps(0.0, 0.001, 1.0) == 999


def test_footprint_host_only():
    p = pa.make_params(L=64, coding=pa.MSC, msbits=32, Rinit=20000)
    fp = pa.footprint(p)
    # two spin buffers of >= 625 groups x 16 KiB plus per-replica scalars
    assert 2 * 625 * 16384 < fp < 64 * 2**20
    bad = pa.make_params(L=63)
    assert pa.footprint(bad) == 0
