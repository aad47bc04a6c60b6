"""MHR build on the GPU against the oracle: energy histograms (integer, exact), the host
reweighting (bit-exact: same operation order and C library on both sides), and the
drop-in executables' PA_MHR_results files with the reference's BF[0] carry-over across -M
runs (PAisingSSC.cu:435,695-791)."""
import os
import subprocess

import pytest

import oracle
import paising_amd as pa
from tests.helpers import compare_rows, engine_params

pytestmark = pytest.mark.gpu

CASES = [(16, pa.SSC, 32, 700), (24, pa.SSC, 32, 300), (16, pa.MSC, 32, 900), (32, pa.MSC, 64, 800)]


@pytest.mark.parametrize("L,coding,w,R", CASES, ids=lambda c: str(c))
def test_histogram_and_mhr_match_oracle(L, coding, w, R):
    print(f"case L={L} coding={coding} w={w} R={R}", flush=True)
    N = L * L
    rc, runs, _, _, H, M = oracle.run_mhr(L=L, coding=coding, msbits=w, Rinit=R, eqsweeps=3,
                                          dB=0.02, Bfin=1.0, seed=41)
    assert rc == 0
    rows, _, hist = pa.run_all_hist(engine_params(L, coding, w, R, 3, 0.02, 1.0, 41, 0, mhr=1))
    compare_rows(runs[0], rows)
    assert hist == H[0]
    out, _ = pa.mhr_analysis(N, [r.B for r in rows], pa.unnormalised_bf(rows, N),
                             [r.R for r in rows], hist)
    assert [[x.hex() for x in r] for r in out] == [[x.hex() for x in r] for r in M[0]]


def test_histogram_refused_without_mhr():
    r = pa.Run(engine_params(16, pa.SSC, 32, 100, 1, 0.1, 0.3, 1, 0))
    r.step(-1)
    r.sync()
    import ctypes
    buf = (ctypes.c_int32 * 257)()
    assert pa.lib().pa_run_histogram(r.h, buf, 257) != 0
    r.close()


@pytest.mark.parametrize("exe,coding", [("PAisingSSC", oracle.SSC), ("PAisingMSC", oracle.MSC)])
def test_cli_mhr_files_with_two_runs(exe, coding, tmp_path):
    r = subprocess.run([os.path.join(pa.BIN_DIR, exe), "-H", "-R", "500", "-t", "3", "-d", "0.02",
                        "-s", "19", "-M", "2", "-L", "16", "-P", "13"], capture_output=True,
                       text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    d = tmp_path / f"data{'SSC' if coding == oracle.SSC else 'MSC'}_L16_R500_EqSw3_dB0.020000"
    rc, runs, _, _, H, M = oracle.run_mhr(L=16, coding=coding, Rinit=500, eqsweeps=3, dB=0.02,
                                          Bfin=1.0, seed=19, runs=2)
    assert rc == 0
    for k in range(2):
        assert (d / f"PA_results_run_{k}.dat").read_text() == oracle.format_rows(runs[k], 13)
        assert (d / f"PA_MHR_results_run_{k}.dat").read_text() == oracle.format_mhr_rows(M[k], 13)
