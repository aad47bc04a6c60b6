"""The benchmark workloads on the GPU against committed oracle traces
(tests/golden/make_config_traces.py):

  config 1  L=16 R=1000 EqSw=10, single-spin coding: the whole 199-step schedule;
  config 3  L=256 R=200000 EqSw=200 dB=0.002, 64 replicas per word: the first two steps bit
            for bit, then five full-size steps checked through properties the domain fixes;
  config 4  the 64-run batch mix, batched: every run's first three steps against the oracle,
            every full schedule against the same run on the plain path;
  config 5  the first 16 served requests: the same two checks for served runs.
"""
import math

import numpy as np
import pytest

import paising_amd as pa
from paising_amd import exact, workloads
from tests.helpers import compare_trace, load_golden, params_of, same_rows

pytestmark = pytest.mark.gpu


def test_config1_whole_schedule_matches_oracle():
    g = load_golden("config1")
    p = params_of(g["params"], trace=1)
    print("config 1: L=16 R=1000 EqSw=10 SSC, 199 steps", flush=True)
    rows, _ = pa.run_all(p)
    assert len(rows) == 200 == len(g["trace"])
    compare_trace(g["trace"], rows)


def test_config3_first_steps_match_oracle():
    g = load_golden("config3")
    p = params_of(g["params"], max_steps=2, trace=1)
    print("config 3: L=256 R=200000 MSC64 EqSw=200, 2 steps", flush=True)
    rows, _ = pa.run_all(p)
    compare_trace(g["trace"], rows)


def test_config3_full_size_properties():
    """Five steps at full size: the copy counts sum to the new population, every energy is
    a multiple of 4 within [-2N, 2N], magnetisations have the parity of N, the Philox counter
    advances by the reference's rule, and the free energy matches Kaufman's exact value."""
    L, R0, w, eqs, dB = 256, 200000, 64, 200, 0.002
    N = L * L
    p = pa.make_params(L=L, coding=pa.MSC, msbits=w, Rinit=R0, eqsweeps=eqs, dB=dB, Bfin=1.0,
                       seed=100, max_steps=5)
    print("config 3: L=256 R=200000 MSC64 EqSw=200, 5 steps with properties", flush=True)
    r = pa.Run(p)
    r.step(-1)
    r.sync()
    rows = r.results()
    R = rows[-1].R
    Rprev = rows[-1].Rprev
    roff = r.array(2)[: math.ceil(Rprev / w) * w]
    active_prev_groups = math.ceil(Rprev / w)
    IE = r.array(0)
    M = r.array(1)
    masks = r.array(4)
    r.close()
    # copy counts of the last resampling: their sum is the new population (CalcParSum)
    assert int(roff.astype(np.int64).sum()) == R
    # the current population: lanes of the transposed placement n -> (n mod rg, n div rg)
    rg = math.ceil(R / w)
    act = np.array([[(int(masks[g]) >> k) & 1 for k in range(w)] for g in range(rg)], bool)
    assert act.sum() == R
    e = IE.reshape(-1, w)[:rg][act]
    m = M.reshape(-1, w)[:rg][act]
    assert (e % 4 == 0).all() and (np.abs(e) <= 2 * N).all()
    assert (m % 2 == N % 2).all() and (np.abs(m) <= N).all()
    assert int(e.astype(np.int64).sum()) == rows[-1].sumE
    # Philox counter: init rg0 * 128, then per step w * groups (copy counts) + 128 * groups
    IS = math.ceil(R0 / w) * 128
    for row in rows[1:]:
        IS += math.ceil(row.Rprev / w) * w + math.ceil(row.R / w) * 128
        assert row.IS_after == IS
    assert active_prev_groups > 0
    # free energy per site against the exact finite-lattice value
    for row in rows[1:]:
        assert abs(row.BF - exact.beta_f_per_site(row.B, L)) < 1e-6, row.B


def test_config4_batch_matches_oracle_and_plain():
    g = load_golden("config4")
    plist = workloads.batch_mix()
    assert len(plist) == len(g["runs"]) == 64
    for p, gr in zip(plist, g["runs"]):
        assert p.L == gr["params"]["L"] and p.Rinit == gr["params"]["Rinit"] and p.seed == gr["params"]["seed"]
        p.trace = 1
    print("config 4: 64-run batch mix, full schedules", flush=True)
    b = pa.Batch(plist, use_graph=True)
    b.run()
    brows = [b.results(i) for i in range(len(plist))]
    b.close()
    for i, (gr, rows) in enumerate(zip(g["runs"], brows)):
        print(f"  run {i}: L={plist[i].L} coding={plist[i].coding} R={plist[i].Rinit}", flush=True)
        assert compare_trace(gr["trace"], rows) == g["steps"] + 1
    for i, p in enumerate(plist):
        same_rows(pa.run_all(p)[0], brows[i])


@pytest.mark.parametrize("policy", [pa.SERVE_SHORTEST])
def test_config5_served_match_oracle_and_plain(policy):
    g = load_golden("config5")
    reqs = [p for _, p in workloads.serve_requests(16)]
    for p, gr in zip(reqs, g["runs"]):
        assert p.L == gr["params"]["L"] and p.Rinit == gr["params"]["Rinit"] and p.seed == gr["params"]["seed"]
        p.trace = 1
    print("config 5: 16 served requests, full schedules", flush=True)
    srv = pa.Server(hbm_budget_bytes=200 * 10**9, max_runs=256, steps_per_tick=32, policy=policy)
    tickets = []
    for p in reqs:
        tickets.append(srv.submit(p))
        srv.tick()  # later requests arrive while earlier ones are on the device
    while not all(srv.done(t) for t in tickets):
        srv.tick()
    served = [srv.results(t) for t in tickets]
    srv.close()
    for i, (gr, rows) in enumerate(zip(g["runs"], served)):
        assert compare_trace(gr["trace"], rows) == g["steps"] + 1, i
    for p, rows in zip(reqs, served):
        same_rows(pa.run_all(p)[0], rows)
