"""Oracle: the adaptive inverse-temperature step (the reference's AdaptiveStep build,
PAisingSSC.cu:308-334,644-660; PAisingMSC.cu:466-499,795-811).  No reference fixture covers
this build, so the overlap is pinned by closed forms and the schedule by its invariants."""
import ctypes
import math

import numpy as np
import pytest

import oracle

P = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(P)


def ssc_overlap(IE, Rinit, dB, Emean):
    IE = np.ascontiguousarray(IE, dtype=np.int32)
    return oracle.lib().ora_ssc_overlap(len(IE), _ptr(IE), Rinit, dB, Emean)


def msc_overlap(IE_lanes, active, w, Rinit, R, dB, Emean):
    rg = len(active) // w
    return oracle.lib().ora_msc_overlap(w, rg, _ptr(active), _ptr(IE_lanes), Rinit, R, dB, Emean)


@pytest.mark.parametrize("R,Rinit", [(1000, 1000), (1500, 1000), (700, 1000), (2049, 2000)])
def test_overlap_vanishing_step_is_min_one_ratio(R, Rinit):
    """dB -> 0: every replica keeps weight Rinit/R, so the overlap is min(1, Rinit/R)."""
    rng = np.random.default_rng(R)
    IE = rng.integers(-800, 801, size=R) * 4
    got = ssc_overlap(IE, Rinit, 1e-300, float(IE.mean()))
    assert got == pytest.approx(min(1.0, Rinit / R), rel=1e-13)


@pytest.mark.parametrize("dB", [0.001, 0.01, 0.1, 1.0])
def test_overlap_of_a_flat_population_does_not_depend_on_the_step(dB):
    IE = np.full(3000, -1234 * 4, dtype=np.int32)
    for Rinit in (2500, 3000, 4000):
        assert ssc_overlap(IE, Rinit, dB, float(IE[0])) == pytest.approx(min(1.0, Rinit / 3000),
                                                                        rel=1e-13)


@pytest.mark.parametrize("dB", [0.0005, 0.003, 0.02, 0.2])
def test_overlap_of_a_two_level_population_closed_form(dB):
    """Half the replicas at E0 < E1, Rinit = R: the reweighted histogram keeps all of the
    lower level and a fraction of the upper, overlap = 1/2 + 1/(1 + exp(dB (E1 - E0)))."""
    R, E0, E1 = 4096, -2000, -1800
    IE = np.array([E0, E1] * (R // 2), dtype=np.int32)
    expect = 0.5 + 1.0 / (1.0 + math.exp(dB * (E1 - E0)))
    assert ssc_overlap(IE, R, dB, float(IE.mean())) == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("w", [32, 64])
def test_msc_overlap_equals_ssc_overlap_on_the_same_population(w):
    """Lane/group indexing: the active lanes of the MSC layout hold the same energies as a
    replica list; only the summation order differs."""
    rng = np.random.default_rng(w)
    R, Rinit, dB = 3000, 2900, 0.013
    rg = (R + w - 1) // w
    IE = (rng.integers(-500, 501, size=R) * 4).astype(np.int32)
    lanes = np.zeros(rg * w, dtype=np.int32)
    active = np.zeros(rg * w, dtype=np.uint8)
    for n in range(R):  # lane layout of the resampled population: n -> (n mod rg, n div rg)
        k = (n % rg) * w + n // rg
        lanes[k] = IE[n]
        active[k] = 1
    a = msc_overlap(lanes, active, w, Rinit, R, dB, float(IE.mean()))
    b = ssc_overlap(IE, Rinit, dB, float(IE.mean()))
    assert a == pytest.approx(b, rel=1e-12)
    assert 0.0 < a < 1.0


def test_overlap_decreases_with_the_step():
    rng = np.random.default_rng(3)
    IE = (rng.normal(-3000, 60, size=5000).round() * 4).astype(np.int32)
    vals = [ssc_overlap(IE, 5000, dB, float(IE.mean())) for dB in (0.0, 0.001, 0.003, 0.01, 0.03)]
    assert all(x > y for x, y in zip(vals, vals[1:]))


@pytest.mark.parametrize("coding", [oracle.SSC, oracle.MSC])
def test_adaptive_schedule_invariants(coding):
    rc, runs, _, _ = oracle.run(L=16, coding=coding, Rinit=600, eqsweeps=4, dB=0.005, Bfin=1.0,
                                seed=31, adaptive=1, runs=2, threads=4)
    assert rc == 0
    B = [r.B for r in runs[0]]
    assert B[0] == 0.0 and B[1] == 0.005                 # the first step is dBinit (SSC:508)
    assert all(y > x for x, y in zip(B, B[1:]))
    assert abs(B[-1] - 1.0) < 1e-15                      # the last step lands on Bfin
    assert len(B) < 200                                  # steps grow beyond dBinit
    assert [r.B for r in runs[1]] == B                   # later runs replay run 0 (SSC:662)
    assert runs[1][1].IS_after > runs[0][-1].IS_after    # with fresh streams


def test_adaptive_tighter_overlap_window_takes_more_steps():
    kw = dict(L=16, coding=oracle.SSC, Rinit=600, eqsweeps=4, dB=0.005, Bfin=1.0, seed=31,
              adaptive=1, threads=4)
    n_default = len(oracle.run(**kw)[1][0])
    n_strict = len(oracle.run(min_overlap=0.93, max_overlap=0.95, **kw)[1][0])
    assert n_strict > n_default
