"""Full L=64, R=20000, EqSw=100, dB=0.005, beta 0->1, seed 100 runs on the GPU against the
committed golden traces (the oracle's per-step internals) and the reference's sample outputs:
every one of the 200 steps bit-exact (R, q, lnQ, FP64 sums, integer sums, copy-count and
spin hashes, Philox counter), and the printed file byte-identical."""
import json
import os

import pytest

import paising_amd as pa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
GOLD = os.path.join(HERE, "golden")
FIX = os.path.join(GOLD, "reference_fixtures")

CASES = [
    ("msc", pa.MSC, pa.UNIFORM_CUDA101, "msc_L64_R20000_EqSw100_dB0.005_s100.dat"),
    ("ssc101", pa.SSC, pa.UNIFORM_CUDA101, "ssc_cuda101_L64_R20000_EqSw100_dB0.005_s100.dat"),
    ("ssc100", pa.SSC, pa.UNIFORM_CUDA100, "ssc_cuda100_L64_R20000_EqSw100_dB0.005_s100.dat"),
]


@pytest.mark.parametrize("name,coding,uniform,fixture", CASES, ids=[c[0] for c in CASES])
def test_full_run_matches_golden_trace_and_reference_output(name, coding, uniform, fixture):
    path = os.path.join(GOLD, f"oracle_config2_{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"golden trace {name} not generated (tests/golden/make_oracle_traces.py)")
    gold = json.load(open(path))
    print(f"case {name}", flush=True)
    rows, _ = pa.run_all(pa.make_params(L=64, coding=coding, msbits=32, Rinit=20000, eqsweeps=100,
                                        dB=0.005, Bfin=1.0, seed=100, uniform=uniform, trace=1))
    tr = gold["trace"]
    assert len(rows) == len(tr) == 200
    for i, (r, g) in enumerate(zip(rows, tr)):
        assert r.R == g["R"], (i, "R")
        if i == 0:
            continue
        got = dict(Rprev=r.Rprev, q=r.q.hex(), lnQ=r.lnQ.hex(), sums=[x.hex() for x in r.sums],
                   sumE=r.sumE, sumM=r.sumM, roff_hash=r.roff_hash, spin_hash=r.spin_hash,
                   IS=r.IS_after)
        for k, v in got.items():
            assert v == g[k], (i, k, v, g[k])
    text = pa.format_rows(rows)
    assert text == gold["text"]
    assert text == open(os.path.join(FIX, fixture)).read()
