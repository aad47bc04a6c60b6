"""Drop-in executables PAisingSSC / PAisingMSC: options, exit codes, directory naming and the
results files (PAisingSSC.cu:381-806)."""
import os
import subprocess

import pytest

import oracle
import paising_amd as pa

BIN = pa.BIN_DIR


def run(exe, *args, cwd=None):
    return subprocess.run([os.path.join(BIN, exe), *map(str, args)], capture_output=True, text=True,
                          cwd=cwd, timeout=300)


@pytest.mark.parametrize("exe", ["PAisingSSC", "PAisingMSC"])
def test_usage_and_bad_options(exe, tmp_path):
    r = run(exe, "-?", cwd=tmp_path)
    assert r.returncode == 1 and "Usage: " + exe in r.stdout and "-R Rinit" in r.stdout
    r = run(exe, "-x", cwd=tmp_path)
    assert r.returncode == 1 and "illegal option -- x" in r.stdout
    r = run(exe, "stray", cwd=tmp_path)
    assert r.returncode == 1 and "Non-option argument stray" in r.stderr


@pytest.mark.parametrize("exe", ["PAisingSSC", "PAisingMSC"])
def test_nbmax_exceeded(exe, tmp_path):
    r = run(exe, "-d", "1e-5", "-s", "3", cwd=tmp_path)
    assert r.returncode == 1 and "exceeds nBmax=10000" in r.stderr
    assert not any(p.name.endswith(".dat") for p in tmp_path.rglob("*"))


@pytest.mark.gpu
@pytest.mark.parametrize("exe,coding,L,w", [("PAisingSSC", oracle.SSC, 16, 32), ("PAisingMSC", oracle.MSC, 16, 32),
                                            ("PAisingMSC", oracle.MSC, 32, 64), ("PAisingMSC", oracle.MSC, 16, 8),
                                            ("PAisingMSC", oracle.MSC, 32, 16)])
def test_results_files_match_oracle(exe, coding, L, w, tmp_path):
    args = ["-R", 700, "-t", 5, "-d", 0.05, "-f", 0.6, "-s", 100, "-M", 2, "-L", L, "-P", 13]
    if coding == oracle.MSC:
        args += ["-w", w]
    r = run(exe, *args, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "RNG initial seed: 100" in r.stdout and "Time per spin-flip:" in r.stdout
    d = tmp_path / f"data{'SSC' if coding == oracle.SSC else 'MSC'}_L{L}_R700_EqSw5_dB0.050000"
    assert d.is_dir()
    rc, runs, _, _ = oracle.run(L=L, coding=coding, msbits=w, Rinit=700, eqsweeps=5, dB=0.05, Bfin=0.6,
                                runs=2, seed=100)
    assert rc == 0
    for k in range(2):
        got = (d / f"PA_results_run_{k}.dat").read_text()
        assert got == oracle.format_rows(runs[k], precision=13)
    assert (d / "PA_results_run_0.dat").read_text() != (d / "PA_results_run_1.dat").read_text()


@pytest.mark.gpu
def test_output_directory_option(tmp_path):
    r = run("PAisingSSC", "-R", 200, "-t", 2, "-d", 0.2, "-f", 0.5, "-s", 7, "-L", 16, "-o", tmp_path / "out",
            cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "out" / "PA_results_run_0.dat").exists()
