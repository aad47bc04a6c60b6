"""Write the oracle's full config-2 runs (L=64, R=20000, EqSw=100, dB=0.005, beta 0->1,
seed 100) as golden traces: the printed results text plus per-step internals (R, q, lnQ,
FP64 sums as hex, integer sums, copy-count and spin hashes, Philox counter).

The engine's GPU tests compare against these without rerunning the oracle at full size.
Usage: python tests/golden/make_oracle_traces.py [msc] [ssc101] [ssc100]   (about 10 min
for MSC and 35 min per SSC variant on 8 cores)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

CASES = {
    "msc": (oracle.MSC, oracle.U32_OPEN_CLOSED),
    "ssc101": (oracle.SSC, oracle.U32_OPEN_CLOSED),
    "ssc100": (oracle.SSC, oracle.U32_MIDPOINT),
}


def trace_of(rows):
    return [dict(i=i, R=r.R, Rprev=r.Rprev, q=r.q.hex(), lnQ=r.lnQ.hex(),
                 sums=[x.hex() for x in r.sums], sumE=r.sumE_int, sumM=r.sumM_int,
                 roff_hash=r.roff_hash, spin_hash=r.spin_hash, IS=r.IS_after)
            for i, r in enumerate(rows)]


def main(names):
    for name in names:
        coding, umap = CASES[name]
        rc, runs, el, tot = oracle.run(L=64, coding=coding, msbits=32, Rinit=20000, eqsweeps=100,
                                       dB=0.005, Bfin=1.0, seed=100, umap=umap)
        assert rc == 0
        d = dict(name=name, params=dict(L=64, coding=coding, msbits=32, Rinit=20000, eqsweeps=100,
                                        dB=0.005, Bfin=1.0, seed=100, umap=umap),
                 elapsed_s=el, total_pop=tot, text=oracle.format_rows(runs[0]),
                 trace=trace_of(runs[0]))
        with open(os.path.join(HERE, f"oracle_config2_{name}.json"), "w") as f:
            json.dump(d, f)
        print(name, "done in %.0f s" % el, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
