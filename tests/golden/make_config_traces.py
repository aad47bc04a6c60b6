"""Write the oracle's golden traces of the benchmark workloads (per-step internals: R, q,
lnQ, FP64 sums as hex, integer sums, copy-count and spin hashes, Philox counter):

  config1  L=16 R=1000 EqSw=10 dB=0.005 beta 0->1, single-spin coding, seed 100, the whole
           199-step schedule;
  config3  L=256 R=200000 EqSw=200 dB=0.002, 64 replicas per word, seed 100, the first two
           beta steps (about 10 min of oracle time on 8 cores);
  config4  the 64-run batch mix (paising_amd.workloads.batch_mix), the first 3 beta steps
           of every run;
  config5  the first 16 served requests (paising_amd.workloads.serve_requests), the first
           3 beta steps of every run.

The GPU tests compare the engine's batched, served and plain runs with these without
rerunning the oracle at full size.  A run's first k steps do not depend on how many steps
follow, so the traces of truncated runs check the first k rows of the full runs too.
Usage: python tests/golden/make_config_traces.py [config1] [config3] [config4] [config5]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from paising_amd import workloads  # noqa: E402
from tests.golden.make_oracle_traces import trace_of  # noqa: E402

FIELDS = ("L", "coding", "msbits", "Rinit", "eqsweeps", "dB", "Binit", "Bfin", "seed",
          "initial_sequence", "uniform", "max_steps")


def params_dict(p):
    return {f: getattr(p, f) for f in FIELDS}


def oracle_trace(d, steps):
    rc, runs, el, _ = oracle.run(L=d["L"], coding=d["coding"], msbits=d["msbits"],
                                 Rinit=d["Rinit"], eqsweeps=d["eqsweeps"], dB=d["dB"],
                                 Binit=d["Binit"], Bfin=d["Bfin"], seed=d["seed"],
                                 initial_sequence=d["initial_sequence"],
                                 umap=oracle.U32_OPEN_CLOSED if d["uniform"] == 0
                                 else oracle.U32_MIDPOINT, max_steps=steps)
    assert rc == 0
    return trace_of(runs[0]), el


def single(L, coding, msbits, Rinit, eqsweeps, dB, seed, steps):
    d = dict(L=L, coding=coding, msbits=msbits, Rinit=Rinit, eqsweeps=eqsweeps, dB=dB,
             Binit=0.0, Bfin=1.0, seed=seed, initial_sequence=0, uniform=0, max_steps=steps)
    tr, el = oracle_trace(d, steps)
    return dict(params=d, trace=tr, elapsed_s=el)


def many(plist, steps):
    runs = []
    for i, p in enumerate(plist):
        d = params_dict(p)
        tr, el = oracle_trace(d, steps)
        runs.append(dict(params=d, trace=tr, elapsed_s=el))
        print(f"  run {i}: L={d['L']} coding={d['coding']} R={d['Rinit']} {el:.1f} s", flush=True)
    return dict(steps=steps, runs=runs)


def make(name):
    t0 = time.time()
    if name == "config1":
        out = single(16, oracle.SSC, 32, 1000, 10, 0.005, 100, 0)
    elif name == "config3":
        out = single(256, oracle.MSC, 64, 200000, 200, 0.002, 100, 2)
    elif name == "config4":
        out = many(workloads.batch_mix(), 3)
    elif name == "config5":
        out = many([p for _, p in workloads.serve_requests(16)], 3)
    else:
        raise SystemExit(f"unknown trace {name}")
    out["name"] = name
    with open(os.path.join(HERE, f"oracle_{name}.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(name, "done in %.0f s" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or ["config1", "config5", "config4", "config3"]:
        make(n)
