"""Engine vs oracle, bit-exact per step (R, q, lnQ, FP64 sums, integer sums, copy-count and
spin-configuration hashes, Philox counter, printed thermodynamics)."""
import pytest

import oracle
import paising_amd as pa
from tests.helpers import compare_rows, engine_params, oracle_rows

pytestmark = pytest.mark.gpu

CASES = [
    # (L, coding, msbits, Rinit, eqsweeps, dB, Bfin, seed, steps)
    (16, pa.SSC, 32, 1000, 10, 0.005, 1.0, 100, 8),      # L=16 R=1000 EqSw=10 shape, truncated
    (32, pa.SSC, 32, 500, 5, 0.01, 1.0, 7, 6),
    (24, pa.SSC, 32, 300, 3, 0.02, 1.0, 11, 5),           # N/2 not a multiple of 128
    (16, pa.MSC, 32, 1000, 10, 0.005, 1.0, 100, 8),
    (64, pa.MSC, 32, 2000, 10, 0.005, 1.0, 100, 5),
    (48, pa.MSC, 32, 777, 4, 0.01, 1.0, 5, 5),            # generic (unaligned) sweep path
    (24, pa.MSC, 32, 300, 3, 0.02, 1.0, 11, 5),
    (32, pa.MSC, 64, 1500, 6, 0.005, 1.0, 3, 5),
    (16, pa.MSC, 64, 200, 4, 0.05, 1.0, 9, 6),
    # populations whose parent chunks exceed the LDS budget: resampling by windows
    (16, pa.MSC, 32, 60000, 2, 0.05, 1.0, 13, 3),
    (16, pa.MSC, 64, 110000, 2, 0.05, 1.0, 17, 3),
    # 8 and 16 replicas per word (PAisingMSC.cu:34 MSbits = 8, 16)
    (16, pa.MSC, 8, 1000, 10, 0.005, 1.0, 100, 8),
    (64, pa.MSC, 8, 2000, 10, 0.005, 1.0, 100, 5),
    (16, pa.MSC, 16, 1000, 10, 0.005, 1.0, 100, 8),
    (64, pa.MSC, 16, 2000, 10, 0.005, 1.0, 100, 5),
    (24, pa.MSC, 8, 500, 4, 0.02, 1.0, 23, 5),
    (96, pa.MSC, 16, 900, 3, 0.01, 1.0, 29, 4),
    (16, pa.MSC, 8, 20000, 2, 0.05, 1.0, 13, 3),      # windowed resampling of 8-bit words
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"L{c[0]}_{'SSC' if c[1] == 0 else 'MSC%d' % c[2]}_R{c[3]}")
def test_engine_matches_oracle(case):
    L, coding, msbits, Rinit, eqs, dB, Bfin, seed, steps = case
    print(f"case L={L} coding={coding} w={msbits} R={Rinit} steps={steps}", flush=True)
    orows = oracle_rows(L, coding, msbits, Rinit, eqs, dB, Bfin, seed, steps)
    erows, _ = pa.run_all(engine_params(L, coding, msbits, Rinit, eqs, dB, Bfin, seed, steps))
    assert compare_rows(orows, erows) == steps + 1


@pytest.mark.parametrize("L,w,sw,nt", [(64, 16, 8, 256), (64, 8, 8, 128), (32, 16, 16, 1024),
                                       (128, 8, 0, 512)])
def test_narrow_word_block_shapes_match_oracle(L, w, sw, nt):
    args = (L, pa.MSC, w, 600, 3, 0.01, 1.0, 31, 3)
    print(f"case L={L} w={w} sw={sw} nt={nt}", flush=True)
    orows = oracle_rows(*args)
    erows, _ = pa.run_all(engine_params(*args, slice_bits=sw, threads=nt, balance=1))
    assert compare_rows(orows, erows) == 4


@pytest.mark.parametrize("uniform", [pa.UNIFORM_CUDA101, pa.UNIFORM_CUDA100])
def test_uniform_conventions(uniform):
    args = (32, pa.SSC, 32, 400, 4, 0.01, 1.0, 21, 4)
    orows = oracle_rows(*args, uniform=uniform)
    erows, _ = pa.run_all(engine_params(*args, uniform=uniform))
    compare_rows(orows, erows)


@pytest.mark.parametrize("L,w,R,eqs,steps", [(64, 32, 20000, 20, 3), (32, 32, 9000, 7, 4),
                                            (64, 64, 30000, 5, 3), (16, 32, 3000, 9, 3)])
def test_balanced_sweeps_match_unbalanced(L, w, R, eqs, steps):
    """Makespan-balanced MSC sweeps (units cut across workers) give the same bits."""
    ref, _ = pa.run_all(engine_params(L, pa.MSC, w, R, eqs, 0.005, 1.0, 5, steps, balance=-1))
    got, _ = pa.run_all(engine_params(L, pa.MSC, w, R, eqs, 0.005, 1.0, 5, steps, balance=1))
    for x, y in zip(ref, got):
        assert (x.R, x.sumE, x.sumM, x.spin_hash, x.roff_hash) == (y.R, y.sumE, y.sumM, y.spin_hash, y.roff_hash)
        assert x.lnQ.hex() == y.lnQ.hex()


def test_balanced_batch_of_mixed_sizes():
    plist = [engine_params(L, pa.MSC, 32, R, e, 0.01, 1.0, 40 + i, 3, balance=1)
             for i, (L, R, e) in enumerate([(64, 7000, 9), (32, 12000, 17), (96, 4000, 5), (64, 9000, 13)])]
    b = pa.Batch(plist, use_graph=True)
    b.run()
    for i, p in enumerate(plist):
        p.balance = -1
        plain, _ = pa.run_all(p)
        got = b.results(i)
        assert [r.spin_hash for r in plain] == [r.spin_hash for r in got]
        assert [r.lnQ.hex() for r in plain] == [r.lnQ.hex() for r in got]
    b.close()


@pytest.mark.parametrize("L,w,sw,nt", [(32, 32, 16, 128), (32, 32, 8, 256), (64, 32, 16, 512),
                                       (32, 64, 32, 256), (32, 64, 16, 512), (48, 32, 8, 128),
                                       (64, 32, 32, 128), (64, 32, 32, 256), (64, 32, 16, 1024),
                                       (128, 64, 8, 1024), (32, 32, 32, 1024), (128, 32, 32, 256),
                                       (256, 64, 16, 1024), (256, 32, 16, 512), (96, 32, 32, 256),
                                       (48, 32, 32, 256), (80, 32, 32, 512), (144, 32, 16, 1024),
                                       (10, 32, 32, 256), (16, 64, 64, 512), (112, 32, 0, 0),
                                       (176, 64, 0, 0), (24, 32, 0, 0), (40, 32, 0, 0),
                                       (192, 32, 32, 1024), (192, 32, 0, 0), (160, 32, 0, 0),
                                       (256, 32, 32, 0), (256, 64, 32, 0), (256, 64, 32, 1024),
                                       (384, 32, 8, 0), (384, 64, 8, 512)])
def test_slices_and_block_shapes_match_oracle(L, w, sw, nt):
    """Every replica-bit slice width and block shape the autotuner may pick is bit-exact,
    on lattices whose half sweeps split into whole Philox blocks per stream and on lattices
    whose do not (L = 10, 16, 24, 40, 48, 80, 112, 144, 176); sw = nt = 0 is the default plan.
    L = 256 with 32-bit slices keeps one sublattice in LDS and the other in registers.
    L = 384 (L/2 = 192 > 128: two sites of a stream in row 0) must not take the
    periodic-column path of L = 96 and 192."""
    args = (L, pa.MSC, w, 700, 4, 0.01, 1.0, 13, 4)
    print(f"case L={L} w={w} sw={sw} nt={nt}", flush=True)
    orows = oracle_rows(*args)
    erows, _ = pa.run_all(engine_params(*args, slice_bits=sw, threads=nt))
    assert compare_rows(orows, erows) == 5


@pytest.mark.parametrize("nt", [128, 256, 512])
def test_ssc_block_shapes_match_oracle(nt):
    args = (64, pa.SSC, 32, 300, 3, 0.01, 1.0, 17, 3)
    orows = oracle_rows(*args)
    erows, _ = pa.run_all(engine_params(*args, threads=nt))
    assert compare_rows(orows, erows) == 4


@pytest.mark.parametrize("L,nt,sb", [(96, 0, 0), (96, 256, 0), (96, 128, 0), (192, 0, 0), (192, 256, 0),
                                     (48, 0, 0), (16, 0, 0), (10, 0, 0), (6, 0, 0), (4, 0, 0),
                                     (130, 0, 0), (160, 0, 0), (64, 128, 0), (128, 0, 0), (128, 128, 0),
                                     (256, 0, 0), (80, 0, 0), (144, 256, 0), (112, 128, 0), (384, 0, 0),
                                     (384, 128, 0), (24, 64, 2), (16, 0, 2), (40, 0, 0), (208, 0, 0),
                                     # the site-per-thread kernels where the packed ones are the plan
                                     (48, 0, 1), (16, 0, 1), (80, 0, 1), (144, 256, 1), (112, 128, 1),
                                     (384, 128, 1),
                                     # the general packed kernel on lattices whose half sweeps
                                     # split into whole Philox blocks (sb = 2)
                                     (64, 0, 2), (96, 320, 2), (128, 64, 2), (256, 0, 2)])
def test_ssc_lattices_off_the_fast_path_match_oracle(L, nt, sb):
    """SSC lattices whose half sweeps do not split into whole Philox blocks per thread
    (stream walk, packed-byte walk), or need several blocks per thread, at every block shape;
    sb = 1 keeps the site-per-thread kernels, sb = 2 takes the general packed-byte kernel."""
    args = (L, pa.SSC, 32, 120, 3, 0.02, 1.0, 29, 3)
    print(f"case L={L} nt={nt} sb={sb}", flush=True)
    orows = oracle_rows(*args)
    erows, _ = pa.run_all(engine_params(*args, threads=nt, slice_bits=sb))
    assert compare_rows(orows, erows) == 4
