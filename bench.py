#!/usr/bin/env python3
"""Benchmark of the population-annealing engine on one MI355X.

Default (the headline line): the plain multi-spin-coded PA path at the reference sample
shape, L=64, R=20000 (32 replicas per uint32 word), EqSw=100, dB=0.005, seed 100.  A step is
one beta step of that run (copy counts, resampling, 100 checkerboard sweeps, energies,
averages) with the population resident in HBM.  The metric is spin-flip attempts per ns,
EqSw * N * sum_i R_i / t, the reciprocal of the reference's "Time per spin-flip"
(PAisingMSC.cu:957).

Extra fields: the sweep kernel's VALU roofline (attempts x VALU issue slots per attempt vs
the chip's issue rate), the resampling kernel's HBM roofline, and a CPU baseline (the
oracle, OpenMP, on this host) over a bounded sample.  --mode batch / serve measure the
batched (configs[3]-style mix) and served (Poisson arrivals) paths instead.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import paising_amd as pa  # noqa: E402

L, RINIT, EQSW, DB, SEED = 64, 20000, 100, 0.005, 100
N = L * L

# VALU issue slots per spin-flip attempt of the MSC32 sweep (lane-cycles), counted on the
# algorithm of PAisingMSC.cu:207-229 in its cheapest gfx950 form: per replica bit one LCG
# value (v_mul_lo_u32 + v_add_u32, jump-ahead form) and two threshold tests each a
# compare plus a bit insert; per 32-bit word a quarter of a Philox4x32-10 block (10 rounds
# x 2 v_mad_u64_u32 + 2 v_xor3) and the neighbour adder/acceptance logic (17 bitwise ops).
# MUL_SLOTS is the issue cost of one v_mul_lo_u32 / v_mad_u64_u32 in VALU cycles relative
# to v_add_u32 (measured with _work-style microbenchmarks on the box).
MUL_SLOTS = float(os.environ.get("PA_MUL_SLOTS", "1.0"))


def valu_slots_per_attempt(w=32):
    per_bit = MUL_SLOTS + 1 + 2 * 2
    philox_block = 10 * (2 * MUL_SLOTS + 2)
    per_word = philox_block / 4 + 17
    return per_bit + per_word / w


def peak_valu_slots_per_s(cu, clock_hz):
    return cu * 64 * clock_hz  # 4 SIMD x 16 lanes per CU per cycle


def device_clock_hz():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        v = ctypes.c_int(0)
        # hipDeviceAttributeClockRate = 5 in hip_runtime_api enum order (kHz)
        if hip.hipDeviceGetAttribute(ctypes.byref(v), 5, 0) == 0 and v.value > 0:
            return v.value * 1e3
    except OSError:
        pass
    return 2.4e9


def cpu_baseline(sample_steps=1, threads=None):
    """Oracle (host C++/OpenMP restatement of the reference) on a bounded sample of the same
    workload: the first `sample_steps` beta steps of config 2, timed on this host."""
    import oracle
    threads = threads or os.cpu_count() or 1
    rc, runs, el, tot = oracle.run(L=L, coding=oracle.MSC, msbits=32, Rinit=RINIT, eqsweeps=EQSW,
                                   dB=DB, Bfin=1.0, seed=SEED, max_steps=sample_steps,
                                   threads=threads)
    assert rc == 0
    att = EQSW * N * tot
    return {"value": att / (el * 1e9), "unit": "spin-flip attempts/ns", "cores": threads,
            "kind": "port",
            "sample": f"oracle, first {sample_steps} beta step(s) of L=64 R=20000 EqSw=100 MSC32, "
                      f"{el:.1f} s wall on {threads} OpenMP threads"}


def bench_plain(steps, warmup):
    total = pa.schedule_steps(0.0, DB, 1.0)
    if warmup + steps > total:
        raise SystemExit(f"--warmup + --steps must be <= {total}")
    p = pa.make_params(L=L, coding=pa.MSC, msbits=32, Rinit=RINIT, eqsweeps=EQSW, dB=DB, Bfin=1.0,
                       seed=SEED)
    run = pa.Run(p)
    run.step(warmup)
    run.sync()
    rows0 = run.results()
    t0 = time.perf_counter()
    run.step(steps)
    run.sync()
    t1 = time.perf_counter()
    rows = run.results()
    timed = rows[warmup + 1: warmup + 1 + steps]
    attempts = EQSW * N * sum(r.R for r in timed)
    wall = t1 - t0
    # sweep-kernel timing with HIP events on the run's stream (separate pass, same steps)
    run2 = pa.Run(p)
    run2.step(warmup)
    run2.sync()
    run2.set_timing(True)
    run2.step(steps)
    run2.sync()
    sweep_ms, _ = run2.timing()
    run.close()
    run2.close()
    return attempts, wall, sweep_ms, rows0, rows


def bench_config3(args):
    """MSC 64 replicas/uint64, L=256, R=200000, EqSw=200, dB=0.002, timed on
    the first --steps beta steps after --warmup (per-attempt work does not depend on beta)."""
    p = pa.make_params(L=256, coding=pa.MSC, msbits=64, Rinit=200000, eqsweeps=200, dB=0.002,
                       Bfin=1.0, seed=100, max_steps=args.warmup + args.steps)
    run = pa.Run(p)
    run.step(args.warmup)
    run.sync()
    t0 = time.perf_counter()
    run.step(args.steps)
    run.sync()
    wall = time.perf_counter() - t0
    rows = run.results()
    run.close()
    att = 200 * 65536 * sum(r.R for r in rows[args.warmup + 1:])
    return {"metric": "spin-flip attempts/ns, L=256 R=2e5 MSC64 (configs[2], truncated schedule)",
            "value": round(att / (wall * 1e9), 2), "unit": "spin-flip attempts/ns",
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 2),
            "R": [r.R for r in rows]}


def batch_mix(nruns=64, seed=2026, steps=0):
    """independent runs of uneven L, R and EqSw, both codings, dB=0.005, beta 0->1."""
    import random
    rng = random.Random(seed)
    plist = []
    for k in range(nruns):
        Lk = rng.choice([16, 24, 32, 48, 64, 96])
        Rk = rng.randint(1000, 50000)
        eq = rng.choice([10, 50, 100])
        coding = rng.choice([pa.SSC, pa.MSC])
        plist.append(pa.make_params(L=Lk, coding=coding, msbits=32, Rinit=Rk, eqsweeps=eq, dB=DB,
                                    Bfin=1.0, seed=100 + k, max_steps=steps))
    return plist


def bench_batch(args):
    """Batched path (one launch sequence per step for all runs, hipGraph replay) against the
    plain path launched one run at a time, on the same runs; every batched run is checked
    bit-exact against its plain counterpart."""
    plist = batch_mix(args.nruns, steps=args.batch_steps)
    t0 = time.perf_counter()
    b = pa.Batch(plist, use_graph=True)
    b.run()
    t_batch = time.perf_counter() - t0
    brows = [b.results(i) for i in range(len(plist))]
    b.close()
    t0 = time.perf_counter()
    prows = [pa.run_all(p)[0] for p in plist]
    t_plain = time.perf_counter() - t0
    exact = all(pa.format_rows(x, 17) == pa.format_rows(y, 17) and
                [r.spin_hash for r in x] == [r.spin_hash for r in y] for x, y in zip(brows, prows))
    steps = sum(len(r) - 1 for r in brows)
    return {
        "metric": "batched PA runs/s (configs[3] mix) vs plain one-at-a-time",
        "value": round(len(plist) / t_batch, 4), "unit": "runs/s", "n_gpus": 1,
        "higher_is_better": True, "plain_runs_per_s": round(len(plist) / t_plain, 4),
        "speedup_vs_plain": round(t_plain / t_batch, 3), "bit_exact_vs_plain": exact,
        "wall_s_batched": round(t_batch, 3), "wall_s_plain": round(t_plain, 3),
        "beta_steps_total": steps,
        "config": {"workload": f"{len(plist)} runs, L in {{16,24,32,48,64,96}}, R in [1000,50000], "
                               f"EqSw in {{10,50,100}}, SSC/MSC32 mixed, dB=0.005, beta 0->1"
                               + (f", truncated to {args.batch_steps} steps" if args.batch_steps else "")},
    }


def serve_requests(n, seed=5):
    """Poisson arrivals (20/s), L in {16..256 step 16}, R log-uniform in [1e3,1e5],
    dB in {0.001, 0.005}, Bfin uniform in [0.44, 1.0], EqSw=50, MSC32."""
    import random
    rng = random.Random(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.expovariate(20.0)
        L = 16 * rng.randint(1, 16)
        R = int(round(10 ** rng.uniform(3, 5)))
        out.append((t, pa.make_params(L=L, coding=pa.MSC, msbits=32, Rinit=R, eqsweeps=50,
                                      dB=rng.choice([0.001, 0.005]), Bfin=rng.uniform(0.44, 1.0),
                                      seed=1000 + k)))
    return out


def bench_serve(args):
    reqs = serve_requests(args.nreq)
    srv = pa.Server(hbm_budget_bytes=200 * 10**9, max_runs=256, steps_per_tick=args.tick_steps)
    t0 = time.perf_counter()
    pending = list(reqs)
    tickets = []
    while pending or not all(srv.done(t) for t, _ in tickets):
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            _, p = pending.pop(0)
            tickets.append((srv.submit(p), p))
        if tickets:
            srv.tick()
        else:
            time.sleep(min(0.001, max(0.0, pending[0][0] - now)))
    wall = time.perf_counter() - t0
    lat = sorted(srv.latency(t)[0] for t, _ in tickets)
    def pct(q):
        return lat[min(len(lat) - 1, int(math.ceil(q * len(lat))) - 1)]
    exact = True
    if args.check:
        for t, p in tickets[: args.check]:
            exact &= pa.format_rows(srv.results(t), 17) == pa.format_rows(pa.run_all(p)[0], 17)
    srv.close()
    return {
        "metric": "served PA runs: p50/p99 completion latency (configs[4] arrivals)",
        "value": round(pct(0.5), 4), "unit": "s (p50 latency)", "p99_s": round(pct(0.99), 4),
        "runs_per_s": round(len(lat) / wall, 4), "n_requests": len(lat), "n_gpus": 1,
        "higher_is_better": False, "bit_exact_vs_plain_checked": args.check if exact else "FAILED",
        "config": {"workload": "Poisson 20 req/s, L in {16..256 step 16}, R log-uniform [1e3,1e5], "
                               "dB in {0.001,0.005}, Bfin U[0.44,1], EqSw=50, MSC32, 200 GB budget"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="plain", choices=["plain", "batch", "serve", "copy", "config3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nruns", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=0, help="truncate batched runs (0 = full)")
    ap.add_argument("--nreq", type=int, default=16)
    ap.add_argument("--tick-steps", type=int, default=8)
    ap.add_argument("--check", type=int, default=4, help="served runs re-run on the plain path")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs each PA population on a single "
              "device (see DESIGN.md for how it would shard); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.mode == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.mode == "copy":
        gbps = pa.bandwidth_probe(1 << 30, 10)
        print(json.dumps({"metric": "streaming device copy (read + write)", "value": round(gbps, 1),
                          "unit": "GB/s", "bytes": 1 << 30, "iters": 10}), flush=True)
        return
    if args.mode == "config3":
        print(json.dumps(bench_config3(args)), flush=True)
        return
    name, cu, _ = pa.device_info()
    clock = device_clock_hz()
    attempts, wall, sweep_ms, rows0, rows = bench_plain(args.steps, args.warmup)
    value = attempts / (wall * 1e9)
    slots = valu_slots_per_attempt(32)
    peak = peak_valu_slots_per_s(cu, clock)
    achieved = attempts * slots / (sweep_ms * 1e-3)
    out = {
        "metric": "spin-flip attempts/ns at L=64 R=2e4 (plain MSC32 PA step)",
        "value": round(value, 4),
        "unit": "spin-flip attempts/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded Philox initial spins, reference schedule)",
        "config": {"workload": "plain PA, MSC 32 replicas/uint32, L=64, R=20000, EqSw=100, "
                               "dB=0.005, beta 0->1, seed 100",
                   "device": name},
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": round(peak / 1e12, 3), "unit": "Tlane-op/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "slots_per_attempt": round(slots, 3),
                     "sweep_attempts_per_ns": round(attempts / (sweep_ms * 1e6), 4)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
