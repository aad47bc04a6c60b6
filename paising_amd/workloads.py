"""Synthetic workloads of the batched and served benchmarks, shared by bench.py and the tests.

batch_mix: 64 independent PA runs of uneven shape (L in {16,24,32,48,64,96}, R uniform in
[1000, 50000], EqSw in {10,50,100}, single-spin or 32-bit multi-spin coding, dB = 0.005,
beta 0 -> 1, seeds 100 + k), drawn from Python's Mersenne Twister seeded with 2026.

serve_requests: Poisson arrivals (mean 20 per second) of multi-spin-coded runs (32 replicas
per word) with L uniform in {16, 32, ..., 256}, R log-uniform in [1e3, 1e5], dB in
{0.001, 0.005}, final beta uniform in [0.44, 1.0], EqSw = 50, seeds 1000 + k.

Every run starts at beta 0 with subsequence counter 0, as a fresh reference run does.
"""
import random

from . import MSC, SSC, make_params

BATCH_DB = 0.005


def batch_mix(nruns=64, seed=2026, steps=0):
    """The batch mix; steps > 0 truncates every run to its first `steps` beta steps."""
    rng = random.Random(seed)
    plist = []
    for k in range(nruns):
        Lk = rng.choice([16, 24, 32, 48, 64, 96])
        Rk = rng.randint(1000, 50000)
        eq = rng.choice([10, 50, 100])
        coding = rng.choice([SSC, MSC])
        plist.append(make_params(L=Lk, coding=coding, msbits=32, Rinit=Rk, eqsweeps=eq,
                                 dB=BATCH_DB, Bfin=1.0, seed=100 + k, max_steps=steps))
    return plist


def small_mix(nruns=64, seed=7):
    """Latency-bound mix: small runs (L in {16,24,32}, R in [500,2000], EqSw=10, both
    codings), each too small to fill the device alone."""
    rng = random.Random(seed)
    return [make_params(L=rng.choice([16, 24, 32]), coding=rng.choice([SSC, MSC]), msbits=32,
                        Rinit=rng.randint(500, 2000), eqsweeps=10, dB=BATCH_DB, Bfin=1.0,
                        seed=500 + k) for k in range(nruns)]


def serve_requests(n, seed=5, rate=20.0, steps=0):
    """[(arrival time in s, params)] of the served load; steps > 0 truncates every run."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.expovariate(rate)
        L = 16 * rng.randint(1, 16)
        R = int(round(10 ** rng.uniform(3, 5)))
        out.append((t, make_params(L=L, coding=MSC, msbits=32, Rinit=R, eqsweeps=50,
                                   dB=rng.choice([0.001, 0.005]), Bfin=rng.uniform(0.44, 1.0),
                                   seed=1000 + k, max_steps=steps)))
    return out


def attempts(p, rows):
    """Spin-flip attempts of a finished run: EqSw * N * sum of the populations after each
    resampling (the reference's totPop, PAisingSSC.cu:632)."""
    return p.eqsweeps * p.L * p.L * sum(r.R for r in rows[1:])
