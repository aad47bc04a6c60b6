"""Single stages of a beta step (pa_stage_*) against the oracle's stage functions, on inputs
the full runs do not reach: random populations and energies, lattices off the fast paths,
inactive MSC lanes, beta = 0, both uniform mappings, every sweep variant family.  Integers,
spins and the bit patterns of FP64 results must be identical."""
import math

import numpy as np
import pytest

import oracle
import paising_amd as pa
from tests.helpers import UMAP

pytestmark = pytest.mark.gpu

P = oracle._ptr
lib = oracle.lib()


def params(L, coding, w=32, Rinit=1000, eqsweeps=3, seed=4242, uniform=0, **kw):
    return pa.make_params(L=L, coding=coding, msbits=w, Rinit=Rinit, eqsweeps=eqsweeps,
                          seed=seed, uniform=uniform, **kw)


DT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def ssc_pop(L, R, rng):
    return rng.choice(np.array([-1, 1], np.int8), size=(R, L * L))


def msc_pop(L, w, R, rng):
    rg = math.ceil(R / w)
    dt = DT[w]
    return rng.integers(0, 2**w, size=(rg, L * L), dtype=np.uint64).astype(dt)


@pytest.mark.parametrize("L,beta,uniform,nt", [(16, 0.44, 0, 0), (24, 1.0, 1, 0), (64, 0.0, 0, 0),
                                               (64, 0.3, 1, 256), (96, 0.7, 0, 0), (96, 0.44, 1, 128),
                                               (192, 0.44, 0, 0), (48, 0.5, 0, 0),
                                               (144, 0.44, 0, 0),
                                               # packed-byte sweeps (default plan, L % 32 == 0)
                                               (32, 0.44, 0, 0), (128, 0.44, 1, 0), (160, 0.7, 0, 0),
                                               (224, 0.3, 1, 0), (256, 0.44, 0, 0), (64, 1.0, 1, 0),
                                               # thresholds of 0 (exp(-8 beta) 2^32 < 1): wide compares
                                               (64, 3.0, 0, 0), (96, 6.0, 1, 0), (256, 6.0, 0, 0),
                                               # packed-byte sweeps of any L % 8 == 0: several
                                               # lattices per workgroup (one part per stream
                                               # quad; nt < 0 asks for them with -nt threads
                                               # where a small population would take the site
                                               # kernels) or one lattice over several parts
                                               (8, 0.44, 0, -256), (16, 0.7, 1, -64), (24, 0.44, 0, -1024),
                                               (32, 0.6, 1, -128), (40, 0.3, 1, 128), (48, 0.44, 0, 512),
                                               (56, 0.6, 0, 0), (80, 0.44, 1, 0), (112, 1.0, 0, 256),
                                               (144, 0.44, 0, 1024), (208, 0.5, 1, 0), (240, 0.44, 0, 0),
                                               (48, 6.0, 0, 0), (144, 6.0, 1, 0), (320, 0.44, 0, 0),
                                               (16, 6.0, 0, -256),
                                               # the two-sweep step (L = 24: three or two dwords
                                               # per quad; L = 32: four), both threshold forms
                                               (24, 0.44, 0, -256), (24, 6.0, 1, -64),
                                               (24, 0.7, 1, -512), (32, 6.0, 0, -512),
                                               (24, 6.0, 0, -256), (24, 1.0, 1, -128)])
def test_ssc_sweep_stage(L, beta, uniform, nt):
    rng = np.random.default_rng(L + abs(nt))
    R, IS, sweeps = 37, 9876543210, 3
    p = params(L, pa.SSC, eqsweeps=sweeps, uniform=uniform, threads=abs(nt),
               slice_bits=2 if nt < 0 else 0)
    pop = ssc_pop(L, R, rng)
    print(f"case SSC sweep L={L} beta={beta} uniform={uniform} nt={nt}", flush=True)
    got, IE, M = pa.stage_sweep(p, pop, R, IS, beta)
    ref = np.ascontiguousarray(pop.copy())
    lib.ora_ssc_sweep(L, R, P(ref), sweeps, p.seed, IS, beta, UMAP[uniform])
    assert (got == ref).all()
    rIE = np.zeros(R, np.int32)
    rM = np.zeros(R, np.int32)
    lib.ora_ssc_energy(L, R, P(ref), P(rIE), P(rM))
    assert (IE == rIE).all() and (M == rM).all()


@pytest.mark.parametrize("L,R,sb", [(64, 2100, 0), (48, 37, 0), (16, 37, 2), (80, 37, 0),
                                    (24, 37, 2), (32, 37, 2)])
def test_ssc_sweep_stage_subsequence_carry(L, R, sb):
    """Packed SSC sweeps evaluate the Philox blocks of four adjacent subsequences together
    and share work when the four have the same high 32 bits; a counter starting just below
    2^32 puts some stream quads across that boundary (their blocks are evaluated one by
    one) and the rest beyond it.  L=64 with 2100 replicas takes the aligned packed kernel,
    L=48 the one-part walk, L=16 its one-dword walk, L=80 two parts per quad, L=24 and 32
    the two-sweep step."""
    rng = np.random.default_rng(L + R)
    IS, sweeps, beta = (1 << 32) - 130, 2, 0.44
    p = params(L, pa.SSC, Rinit=R, eqsweeps=sweeps, slice_bits=sb)
    pop = ssc_pop(L, R, rng)
    print(f"case SSC sweep across 2^32 L={L} R={R} sb={sb}", flush=True)
    got, IE, M = pa.stage_sweep(p, pop, R, IS, beta)
    ref = np.ascontiguousarray(pop.copy())
    lib.ora_ssc_sweep(L, R, P(ref), sweeps, p.seed, IS, beta, UMAP[0])
    assert (got == ref).all()
    rIE = np.zeros(R, np.int32)
    rM = np.zeros(R, np.int32)
    lib.ora_ssc_energy(L, R, P(ref), P(rIE), P(rM))
    assert (IE == rIE).all() and (M == rM).all()


@pytest.mark.parametrize("L,beta,uniform,nt", [(16, 0.44, 0, 0), (24, 1.0, 1, 0), (48, 0.5, 0, 0),
                                               (80, 0.44, 1, 0), (144, 0.44, 0, 256),
                                               (112, 0.7, 0, 128), (64, 0.44, 0, 0)])
def test_ssc_sweep_stage_site_kernels(L, beta, uniform, nt):
    """slice_bits = 1 keeps SSC on the site-per-thread kernels (stream walks, block loops),
    which must stay bit-exact where the packed-byte kernels now take the default plan."""
    rng = np.random.default_rng(3 * L + nt)
    R, IS, sweeps = 29, 1234567, 3
    p = params(L, pa.SSC, eqsweeps=sweeps, uniform=uniform, threads=nt, slice_bits=1)
    pop = ssc_pop(L, R, rng)
    print(f"case SSC site-kernel sweep L={L} beta={beta} uniform={uniform} nt={nt}", flush=True)
    got, IE, M = pa.stage_sweep(p, pop, R, IS, beta)
    ref = np.ascontiguousarray(pop.copy())
    lib.ora_ssc_sweep(L, R, P(ref), sweeps, p.seed, IS, beta, UMAP[uniform])
    assert (got == ref).all()
    rIE = np.zeros(R, np.int32)
    rM = np.zeros(R, np.int32)
    lib.ora_ssc_energy(L, R, P(ref), P(rIE), P(rM))
    assert (IE == rIE).all() and (M == rM).all()


@pytest.mark.parametrize("L,w,sw,nt,bal,beta", [(64, 32, 0, 0, 0, 0.44), (64, 32, 8, 512, -1, 1.0),
                                                (32, 64, 16, 256, -1, 0.2), (16, 32, 0, 0, 0, 0.0),
                                                (80, 32, 0, 0, 0, 0.6), (128, 64, 8, 1024, -1, 0.44),
                                                (64, 32, 32, 512, 1, 0.44), (48, 64, 0, 0, 0, 0.9),
                                                (256, 64, 32, 0, 0, 0.44), (256, 32, 32, 0, 0, 0.0),
                                                (64, 8, 0, 0, 0, 0.44), (16, 16, 0, 0, 0, 0.3),
                                                (64, 16, 8, 256, -1, 0.6), (48, 8, 0, 0, 0, 1.0),
                                                (96, 16, 0, 1024, 0, 0.44), (256, 16, 0, 0, 0, 0.44),
                                                # one site per stream and half sweep (L <= 16):
                                                # a Philox block for two sweeps, odd sweep counts
                                                (16, 32, 8, 256, -1, 0.44), (16, 64, 16, 512, -1, 0.7),
                                                (10, 32, 0, 0, 0, 0.44), (16, 32, 32, 256, 1, 0.44),
                                                (12, 64, 32, 128, -1, 1.0),
                                                # columns repeating every 3rd, 5th, 7th site of
                                                # a stream (L = 192, 160, 224)
                                                (192, 32, 16, 0, 0, 0.44), (160, 32, 0, 0, 0, 0.44),
                                                (160, 64, 32, 256, 1, 0.3), (224, 32, 16, 512, -1, 0.7),
                                                (224, 16, 8, 1024, -1, 0.44),
                                                # L = 32 over four parts (one block word each)
                                                (32, 32, 0, 512, -1, 0.44), (32, 64, 32, 512, -1, 0.6),
                                                (32, 32, 16, 512, 0, 0.3), (32, 16, 8, 512, -1, 1.0)])
def test_msc_sweep_stage(L, w, sw, nt, bal, beta):
    rng = np.random.default_rng(L * w + sw)
    R, IS, sweeps = 5 * w - 7, 123456789, 3 if L <= 32 else 2
    rg = math.ceil(R / w)
    p = params(L, pa.MSC, w, eqsweeps=sweeps, slice_bits=sw, threads=nt, balance=bal)
    pop = msc_pop(L, w, R, rng)
    print(f"case MSC sweep L={L} w={w} sw={sw} nt={nt} bal={bal} beta={beta}", flush=True)
    got, IE, M = pa.stage_sweep(p, pop, R, IS, beta)
    ref = np.ascontiguousarray(pop.copy())
    lib.ora_msc_sweep(L, w, rg, P(ref), sweeps, p.seed, IS, beta)
    assert (got == ref).all()
    rIE = np.zeros(rg * w, np.int32)
    rM = np.zeros(rg * w, np.int32)
    lib.ora_msc_energy(L, w, rg, P(ref), P(rIE), P(rM))
    assert (IE == rIE).all() and (M == rM).all()


@pytest.mark.parametrize("coding,L,w", [(pa.SSC, 10, 32), (pa.SSC, 64, 32), (pa.MSC, 24, 32),
                                        (pa.MSC, 64, 64), (pa.MSC, 32, 8), (pa.MSC, 48, 16)])
def test_energy_stage(coding, L, w):
    rng = np.random.default_rng(L + w)
    R = 70
    p = params(L, coding, w)
    if coding == pa.SSC:
        pop = ssc_pop(L, R, rng)
        pop[0] = 1  # ground state
        IE, M = pa.stage_energy(p, pop, R)
        rIE, rM = np.zeros(R, np.int32), np.zeros(R, np.int32)
        lib.ora_ssc_energy(L, R, P(np.ascontiguousarray(pop)), P(rIE), P(rM))
        assert IE[0] == -2 * L * L and M[0] == L * L
    else:
        rg = math.ceil(R / w)
        pop = msc_pop(L, w, R, rng)
        IE, M = pa.stage_energy(p, pop, R)
        rIE, rM = np.zeros(rg * w, np.int32), np.zeros(rg * w, np.int32)
        lib.ora_msc_energy(L, w, rg, P(np.ascontiguousarray(pop)), P(rIE), P(rM))
    assert (IE == rIE).all() and (M == rM).all()


@pytest.mark.parametrize("coding,w,R,Rinit,uniform", [(pa.SSC, 32, 20000, 20000, 0),
                                                      (pa.SSC, 32, 3001, 2500, 1),
                                                      (pa.MSC, 32, 20000, 20000, 0),
                                                      (pa.MSC, 64, 5003, 5000, 1),
                                                      (pa.MSC, 32, 1500, 2000, 0),
                                                      (pa.MSC, 8, 7001, 7000, 0),
                                                      (pa.MSC, 16, 3000, 2900, 1)])
def test_copy_counts_stage(coding, w, R, Rinit, uniform):
    """QKer + lnQ + CalcTauKer + CalcParSum on random energies (and random inactive lanes)."""
    rng = np.random.default_rng(R + w)
    L, dB, IS = 32, 0.01, 777777
    p = params(L, coding, w, Rinit=Rinit, uniform=uniform)
    n = R if coding == pa.SSC else math.ceil(R / w) * w
    IE = (4 * rng.integers(-400, 200, n)).astype(np.int32)
    Emean = float(IE[:R].mean())
    active = None
    if coding == pa.MSC:
        active = (np.arange(n) < R).astype(np.uint8)
        active[rng.integers(0, R, 40)] = 0  # a few inactive lanes inside the population
    print(f"case copy counts coding={coding} w={w} R={R}", flush=True)
    q, lnQ, Roff, pre, Rn = pa.stage_copy_counts(p, IE, R, dB, Emean, IS, active)
    rRoff = np.zeros(n, np.uint32)
    if coding == pa.SSC:
        rq = lib.ora_ssc_q(R, P(IE), dB, Emean)
        rlnQ = lib.ora_lnq(dB, Emean, rq, R)
        tot = lib.ora_ssc_tau(R, P(IE), Rinit, dB, rlnQ, p.seed, IS, UMAP[uniform], P(rRoff))
    else:
        rg = n // w
        rq = lib.ora_msc_q(w, rg, P(active), P(IE), dB, Emean)
        rlnQ = lib.ora_lnq(dB, Emean, rq, R)
        tot = lib.ora_msc_tau(w, rg, P(active), P(IE), Rinit, R, dB, rlnQ, p.seed, IS,
                              UMAP[uniform], P(rRoff))
    assert q.hex() == rq.hex() and lnQ.hex() == rlnQ.hex()
    assert (Roff == rRoff).all() and Rn == tot
    excl = np.concatenate([[0], np.cumsum(rRoff.astype(np.int64))[:-1]])
    assert (pre.astype(np.int64) == excl).all()


@pytest.mark.parametrize("coding,L,w", [(pa.SSC, 16, 32), (pa.SSC, 24, 32), (pa.MSC, 16, 32),
                                        (pa.MSC, 64, 32), (pa.MSC, 32, 64), (pa.MSC, 24, 64),
                                        (pa.MSC, 16, 8), (pa.MSC, 24, 16), (pa.MSC, 64, 8)])
def test_resample_stage(coding, L, w):
    rng = np.random.default_rng(L * w + coding)
    R = 300
    p = params(L, coding, w)
    if coding == pa.SSC:
        pop = ssc_pop(L, R, rng)
        Roff = rng.integers(0, 3, R).astype(np.uint32)
        got, _, Rn = pa.stage_resample(p, pop, R, Roff)
        ref = np.zeros((int(Roff.sum()), L * L), np.int8)
        lib.ora_ssc_resample(L, R, P(np.ascontiguousarray(pop)), P(Roff), P(ref))
        assert Rn == int(Roff.sum()) and (got == ref).all()
    else:
        rg = math.ceil(R / w)
        n = rg * w
        pop = msc_pop(L, w, R, rng)
        active = (np.arange(n) < R).astype(np.uint8)
        Roff = rng.integers(0, 3, n).astype(np.uint32) * active
        Rn_ref = int(Roff.sum())
        got, an, Rn = pa.stage_resample(p, pop, R, Roff, active)
        rgn = math.ceil(Rn_ref / w)
        dt = DT[w]
        ref = np.zeros((rgn, L * L), dt)
        ran = np.zeros(rgn * w, np.uint8)
        IE = np.zeros(n, np.int32)
        IEn = np.zeros(rgn * w, np.int32)
        lib.ora_msc_resample(L, w, rg, P(np.ascontiguousarray(pop)), P(active), P(IE), P(Roff),
                             Rn_ref, P(ref), P(ran), P(IEn))
        assert Rn == Rn_ref and (got == ref).all() and (an == ran).all()


@pytest.mark.parametrize("L,w,R,pattern", [(16, 32, 60000, "uniform"), (16, 32, 60000, "shrink"),
                                           (24, 64, 100003, "uniform"), (16, 64, 100000, "shrink"),
                                           (16, 64, 100000, "few"), (20, 32, 60000, "uniform"),
                                           (16, 8, 20000, "uniform"), (16, 8, 20000, "few"),
                                           (16, 16, 30000, "shrink"), (24, 8, 15000, "uniform")])
def test_resample_stage_large_populations(L, w, R, pattern):
    """Populations whose parent chunks exceed the LDS budget (resampling by windows of new
    groups where N is a multiple of the 128-byte chunk, L=20 with 16-byte chunks otherwise):
    copy counts of 0-2, of 0-1 with the population shrinking fourfold (windows processed in
    halves), and of 0 or 50 (few parents, long runs of children); inactive lanes inside."""
    rng = np.random.default_rng(R + w + L)
    p = params(L, pa.MSC, w, Rinit=R)
    rg = math.ceil(R / w)
    n = rg * w
    pop = msc_pop(L, w, R, rng)
    active = (np.arange(n) < R).astype(np.uint8)
    active[rng.integers(0, R, 200)] = 0
    if pattern == "uniform":
        Roff = rng.integers(0, 3, n)
    elif pattern == "shrink":
        Roff = (rng.random(n) < 0.25).astype(np.int64)
    else:
        Roff = np.where(rng.random(n) < 0.02, 50, 0)
    Roff = (Roff * active).astype(np.uint32)
    Rn_ref = int(Roff.sum())
    print(f"case MSC resample L={L} w={w} R={R} {pattern} Rn={Rn_ref}", flush=True)
    got, an, Rn = pa.stage_resample(p, pop, R, Roff, active)
    rgn = math.ceil(Rn_ref / w)
    dt = DT[w]
    ref = np.zeros((rgn, L * L), dt)
    ran = np.zeros(rgn * w, np.uint8)
    IE = np.zeros(n, np.int32)
    IEn = np.zeros(rgn * w, np.int32)
    lib.ora_msc_resample(L, w, rg, P(np.ascontiguousarray(pop)), P(active), P(IE), P(Roff),
                         Rn_ref, P(ref), P(ran), P(IEn))
    assert Rn == Rn_ref and (got == ref).all() and (an == ran).all()


@pytest.mark.parametrize("coding,w,R", [(pa.SSC, 32, 5000), (pa.MSC, 32, 4001), (pa.MSC, 64, 3000),
                                        (pa.MSC, 8, 2003), (pa.MSC, 16, 3000)])
def test_averages_stage(coding, w, R):
    rng = np.random.default_rng(R)
    L = 64
    p = params(L, coding, w)
    n = R if coding == pa.SSC else math.ceil(R / w) * w
    IE = (4 * rng.integers(-2048, 0, n)).astype(np.int32)
    M = (2 * rng.integers(-2048, 2048, n)).astype(np.int32)
    active = None if coding == pa.SSC else (np.arange(n) < R).astype(np.uint8)
    sums, se, sm = pa.stage_averages(p, IE, M, R, active)
    out = (np.zeros(5))
    if coding == pa.SSC:
        lib.ora_ssc_averages(R, P(IE), P(M), P(out))
    else:
        lib.ora_msc_averages(w, n // w, P(active), P(IE), P(M), P(out))
    assert [x.hex() for x in sums] == [float(x).hex() for x in out]
    mask = np.ones(n, bool) if active is None else active.astype(bool)
    assert se == int(IE[mask].astype(np.int64).sum()) and sm == int(M[mask].astype(np.int64).sum())
