"""Adaptive inverse-temperature step on the GPU (the reference's AdaptiveStep build,
PAisingSSC.cu:644-660, PAisingMSC.cu:795-811) against the oracle: the decided schedule and
every step record bit-exact; later -M runs replay run 0's schedule (PAisingSSC.cu:662)."""
import os
import subprocess

import pytest

import oracle
import paising_amd as pa
from tests.helpers import compare_rows, engine_params, oracle_rows

pytestmark = pytest.mark.gpu

CASES = [
    # (L, coding, msbits, Rinit, eqsweeps, Bfin, seed, min_overlap, max_overlap)
    (16, pa.SSC, 32, 600, 4, 1.0, 31, 0.0, 0.0),
    (24, pa.SSC, 32, 400, 3, 0.7, 5, 0.0, 0.0),
    (16, pa.MSC, 32, 1000, 4, 1.0, 7, 0.0, 0.0),
    (32, pa.MSC, 64, 900, 3, 0.8, 9, 0.0, 0.0),
    (16, pa.MSC, 32, 700, 3, 0.6, 13, 0.93, 0.95),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"L{c[0]}_{'SSC' if c[1] == 0 else 'MSC%d' % c[2]}_R{c[3]}")
def test_adaptive_schedule_and_rows_match_oracle(case):
    L, coding, w, R, eqs, Bfin, seed, lo, hi = case
    print(f"case {case}", flush=True)
    kw = dict(adaptive=1, min_overlap=lo, max_overlap=hi)
    orows = oracle_rows(L, coding, w, R, eqs, 0.005, Bfin, seed, 0, **kw)
    erows, _ = pa.run_all(engine_params(L, coding, w, R, eqs, 0.005, Bfin, seed, 0, **kw))
    assert [r.B.hex() for r in erows] == [r.B.hex() for r in orows]
    assert compare_rows(orows, erows) == len(orows)
    assert len(orows) > 20


def test_adaptive_truncated_and_stepwise():
    """max_steps truncates the adaptive schedule; stepping one beta step at a time gives the
    same records as one call."""
    args = (16, pa.SSC, 32, 500, 3, 0.005, 1.0, 3, 0)
    full, _ = pa.run_all(engine_params(*args, adaptive=1))
    part = engine_params(16, pa.SSC, 32, 500, 3, 0.005, 1.0, 3, 7, adaptive=1)
    got, _ = pa.run_all(part)
    assert len(got) == 8 and [r.spin_hash for r in got] == [r.spin_hash for r in full[:8]]
    r = pa.Run(engine_params(*args, adaptive=1))
    while True:
        r.step(1)
        done, total = r.steps_done()
        if done == total:
            break
    rows = r.results()
    r.close()
    assert [x.B for x in rows] == [x.B for x in full]
    assert [x.spin_hash for x in rows] == [x.spin_hash for x in full]


def test_explicit_schedule_replays_run0_like_the_reference():
    rc, runs, _, _ = oracle.run(L=16, coding=oracle.MSC, Rinit=800, eqsweeps=3, dB=0.005, Bfin=1.0,
                                seed=17, adaptive=1, runs=2)
    assert rc == 0
    run0, is_end = pa.run_all(engine_params(16, pa.MSC, 32, 800, 3, 0.005, 1.0, 17, 0, adaptive=1))
    compare_rows(runs[0], run0)
    p1 = engine_params(16, pa.MSC, 32, 800, 3, 0.005, 1.0, 17, 0, initial_sequence=is_end,
                       schedule=[r.B for r in run0])
    run1, _ = pa.run_all(p1)
    compare_rows(runs[1], run1)


def test_adaptive_runs_are_refused_by_batch_and_server():
    p = engine_params(16, pa.SSC, 32, 300, 2, 0.005, 1.0, 1, 0, adaptive=1)
    with pytest.raises(pa.PAError):
        pa.Batch([p])
    srv = pa.Server(hbm_budget_bytes=1 << 30)
    with pytest.raises(pa.PAError):
        srv.submit(p)
    srv.close()


@pytest.mark.parametrize("exe,coding", [("PAisingSSC", oracle.SSC), ("PAisingMSC", oracle.MSC)])
def test_cli_adaptive_with_two_runs(exe, coding, tmp_path):
    r = subprocess.run([os.path.join(pa.BIN_DIR, exe), "-a", "-R", "600", "-t", "3", "-f", "0.8",
                        "-s", "23", "-M", "2", "-L", "16"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    d = tmp_path / f"data{'SSC' if coding == oracle.SSC else 'MSC'}_L16_R600_EqSw3"
    rc, runs, _, _ = oracle.run(L=16, coding=coding, Rinit=600, eqsweeps=3, dB=0.005, Bfin=0.8,
                                runs=2, seed=23, adaptive=1)
    assert rc == 0
    for k in range(2):
        assert (d / f"PA_results_run_{k}.dat").read_text() == oracle.format_rows(runs[k])
