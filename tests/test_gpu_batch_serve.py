"""Batched, graph-replayed and served runs are bit-exact to the oracle and to the same run on
the plain path."""
import random

import pytest

import paising_amd as pa
from tests.helpers import compare_rows, engine_params, oracle_rows, same_rows

pytestmark = pytest.mark.gpu

same = same_rows


def mix_cases(n, seed):
    """(L, coding, msbits, Rinit, EqSw, dB, Bfin, seed, steps) of a small mixed run list."""
    rng = random.Random(seed)
    out = []
    for k in range(n):
        L = rng.choice([16, 24, 32, 48, 64])
        coding = rng.choice([pa.SSC, pa.MSC])
        w = rng.choice([32, 64]) if coding == pa.MSC else 32
        out.append((L, coding, w, rng.randint(100, 1500), rng.choice([2, 5, 10]),
                    rng.choice([0.01, 0.02]), 1.0, 100 + k, rng.randint(3, 9)))
    return out


def mix(n, seed):
    return [engine_params(*c) for c in mix_cases(n, seed)]


def test_graph_replay_matches_plain():
    p = engine_params(32, pa.MSC, 32, 3000, 10, 0.005, 1.0, 100, 12)
    plain, _ = pa.run_all(p)
    r = pa.Run(p)
    r.set_graph(True)
    r.step(5)
    r.sync()          # host work between replays
    r.step(-1)
    got = r.results()
    r.close()
    same(plain, got)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_oracle_and_plain(use_graph):
    cases = mix_cases(12, 7)
    plist = [engine_params(*c) for c in cases]
    b = pa.Batch(plist, use_graph=use_graph)
    b.run()
    for i, (c, p) in enumerate(zip(cases, plist)):
        got = b.results(i)
        compare_rows(oracle_rows(*c), got)
        plain, _ = pa.run_all(p)
        same(plain, got)
    b.close()


@pytest.mark.parametrize("policy", [pa.SERVE_SHORTEST, pa.SERVE_SHARE])
def test_server_matches_oracle_and_plain_alone_and_in_mix(policy):
    cases = mix_cases(10, 11)
    plist = [engine_params(*c) for c in cases]
    srv = pa.Server(hbm_budget_bytes=64 * 2**20, max_runs=4, steps_per_tick=3, policy=policy)
    tickets = []
    for i, p in enumerate(plist):
        tickets.append(srv.submit(p))
        srv.tick()            # requests arrive while others are on the device
    while not all(srv.done(t) for t in tickets):
        srv.tick()
    for t, c, p in zip(tickets, cases, plist):
        got = srv.results(t)
        compare_rows(oracle_rows(*c), got)
        plain, _ = pa.run_all(p)
        same(plain, got)
        a, b = srv.latency(t)
        assert a >= b >= 0
    q, res, used = srv.stats()
    assert q == 0 and res == 0 and used == 0
    srv.close()


def test_batch_of_all_word_widths_matches_oracle():
    """Runs of 8, 16, 32 and 64 replicas per word and single-spin runs in one batch."""
    cases = [(32, pa.MSC, 8, 700, 4, 0.02, 1.0, 51, 5), (24, pa.MSC, 16, 900, 5, 0.02, 1.0, 52, 4),
             (16, pa.MSC, 32, 1200, 3, 0.02, 1.0, 53, 6), (32, pa.MSC, 64, 600, 4, 0.02, 1.0, 54, 3),
             (16, pa.SSC, 32, 400, 5, 0.02, 1.0, 55, 4), (48, pa.MSC, 8, 12000, 2, 0.05, 1.0, 56, 3)]
    plist = [engine_params(*c) for c in cases]
    b = pa.Batch(plist, use_graph=True)
    b.run()
    for i, c in enumerate(cases):
        compare_rows(oracle_rows(*c), b.results(i))
    b.close()


def test_server_ticks_allocate_no_device_memory_once_reserved():
    """With the pool reserved up front, admitting and retiring a stream of requests makes no
    further device allocation, and every result stays bit-exact to the plain path."""
    cases = mix_cases(14, 23)
    plist = [engine_params(*c) for c in cases]
    srv = pa.Server(hbm_budget_bytes=256 * 2**20, max_runs=6, steps_per_tick=2)
    srv.reserve(256 * 2**20)
    warm = srv.submit(plist[0])            # first tick creates the table pool and graph
    while not srv.done(warm):
        srv.tick()
    reserved, allocs = srv.memory()
    assert reserved >= 256 * 2**20
    tickets = []
    for p in plist:
        tickets.append(srv.submit(p))
        srv.tick()
    while not all(srv.done(t) for t in tickets):
        srv.tick()
    assert srv.memory() == (reserved, allocs)
    for t, p in zip(tickets, plist):
        same(pa.run_all(p)[0], srv.results(t))
    srv.close()


def test_server_request_that_cannot_be_created_fails_alone():
    """A request whose device state cannot be allocated (about 2 TB: more than the device
    holds, though within the stated budget) fails by itself; the requests around it run and
    stay bit-exact, and the queue moves on."""
    small = [engine_params(32, pa.MSC, 32, 800, 3, 0.02, 1.0, 61 + i, 4) for i in range(2)]
    huge = pa.make_params(L=64, coding=pa.MSC, msbits=64, Rinit=2_000_000_000, eqsweeps=1,
                          dB=0.01, Bfin=1.0, seed=3, max_steps=1)
    assert pa.footprint(huge) > 1.5e12
    srv = pa.Server(hbm_budget_bytes=4 * 10**12, max_runs=8, steps_per_tick=2)
    t0 = srv.submit(small[0])
    th = srv.submit(huge)
    t1 = srv.submit(small[1])
    for _ in range(10000):
        if srv.done(t0) and srv.done(t1):
            break
        srv.tick()
    with pytest.raises(pa.PAError):
        srv.done(th)
    for t, p in zip((t0, t1), small):
        same(pa.run_all(p)[0], srv.results(t))
    srv.close()


@pytest.mark.parametrize("policy", [pa.SERVE_SHORTEST, pa.SERVE_SHARE])
def test_server_shares_device_time(policy):
    """A cheap request submitted after an expensive one finishes first (processor sharing:
    each resident run gets about the same device time per tick; shortest remaining work: the
    cheap run is scheduled first), and both stay bit-exact."""
    big = pa.make_params(L=64, coding=pa.MSC, msbits=32, Rinit=20000, eqsweeps=50, dB=0.005,
                         Bfin=1.0, seed=71)
    small = pa.make_params(L=16, coding=pa.SSC, Rinit=500, eqsweeps=5, dB=0.005, Bfin=1.0, seed=72)
    srv = pa.Server(hbm_budget_bytes=1 << 30, max_runs=8, steps_per_tick=32, policy=policy)
    tb = srv.submit(big)
    ts = srv.submit(small)
    while not srv.done(ts):
        srv.tick()
    assert not srv.done(tb)
    while not srv.done(tb):
        srv.tick()
    same(pa.run_all(big)[0], srv.results(tb))
    same(pa.run_all(small)[0], srv.results(ts))
    srv.close()


def test_server_shortest_first_pauses_long_runs():
    """Under the shortest-remaining-work policy, once short runs fill the device a long run
    waits; it still finishes bit-exact."""
    long_run = pa.make_params(L=64, coding=pa.MSC, msbits=32, Rinit=40000, eqsweeps=20,
                              dB=0.05, Bfin=1.0, seed=81)
    shorts = [pa.make_params(L=32, coding=pa.SSC, Rinit=2000, eqsweeps=5, dB=0.05, Bfin=1.0,
                             seed=82 + i) for i in range(3)]
    srv = pa.Server(hbm_budget_bytes=1 << 30, max_runs=8, steps_per_tick=4,
                    policy=pa.SERVE_SHORTEST)
    tl = srv.submit(long_run)
    ts = [srv.submit(p) for p in shorts]
    while not all(srv.done(t) for t in ts):
        srv.tick()
    assert not srv.done(tl)
    while not srv.done(tl):
        srv.tick()
    same(pa.run_all(long_run)[0], srv.results(tl))
    for t, p in zip(ts, shorts):
        same(pa.run_all(p)[0], srv.results(t))
    with pytest.raises(pa.PAError):
        pa.Server(hbm_budget_bytes=1 << 20, policy=7)
    srv.close()


def test_server_respects_budget():
    p = engine_params(64, pa.MSC, 32, 20000, 1, 0.005, 1.0, 1, 2)
    need = pa.footprint(p)
    srv = pa.Server(hbm_budget_bytes=int(need * 1.5), max_runs=8, steps_per_tick=1)
    t1, t2 = srv.submit(p), srv.submit(p)
    srv.tick()
    q, res, used = srv.stats()
    assert res == 1 and q == 1 and used == need   # the second waits for memory
    while not (srv.done(t1) and srv.done(t2)):
        srv.tick()
    same(srv.results(t1), srv.results(t2))        # same parameters, same results
    with pytest.raises(pa.PAError):
        srv.submit(engine_params(64, pa.MSC, 32, 10**6, 1, 0.005, 1.0, 1, 2))
    srv.close()
