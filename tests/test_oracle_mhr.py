"""Oracle: energy histograms and multi-histogram reweighting (the reference's MHR build,
PAisingSSC.cu:294-304,459-463,496-501,608-611,695-791).  No reference fixture covers this
build: the analysis is pinned by an exact density of states (L=4, all 2^16 states), the
histograms by their counts."""
import ctypes
import math

import numpy as np
import pytest

import oracle

P = ctypes.c_void_p


def exact_dos(L):
    """Omega(E) of the periodic L x L Ising lattice by enumeration (bins E = 4k - 2N)."""
    N = L * L
    s = ((np.arange(2 ** N)[:, None] >> np.arange(N)) & 1).astype(np.int8) * 2 - 1
    g = s.reshape(-1, L, L)
    E = -(g * np.roll(g, 1, 1)).sum((1, 2)) - (g * np.roll(g, 1, 2)).sum((1, 2))
    return np.bincount((E + 2 * N) // 4, minlength=N + 1).astype(np.float64)


def mhr(N, B, BF, R, hist, niter=1):
    B = np.ascontiguousarray(B, dtype=np.float64)
    BF = np.ascontiguousarray(BF, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.int32)
    H = np.ascontiguousarray(hist, dtype=np.int32)
    out = np.zeros(len(B) * (1 + 3 * niter))
    oracle.lib().ora_mhr_analysis(N, len(B), B.ctypes.data_as(P), BF.ctypes.data_as(P),
                                  R.ctypes.data_as(P), H.ctypes.data_as(P), niter,
                                  out.ctypes.data_as(P))
    return out.reshape(len(B), -1), BF


def test_mhr_recovers_the_exact_thermodynamics_of_a_small_lattice():
    """Histograms of exact canonical samples (R_i = 1e8 each) with exact free energies:
    the reweighted <E>, C and BF equal the exact ones to the histogram rounding."""
    L, N = 4, 16
    om = exact_dos(L)
    Ei = 4 * np.arange(N + 1) - 2 * N
    B = np.linspace(0.0, 1.0, 11)
    lnZ = np.array([math.log((om * np.exp(-b * Ei)).sum()) for b in B])
    R = np.full(len(B), 10 ** 8)
    H = np.zeros(N + 1)
    for b, lz, r in zip(B, lnZ, R):
        H += r * om * np.exp(-b * Ei - lz)
    out, BFnew = mhr(N, B, -lnZ, R, np.rint(H))
    for i, b in enumerate(B):
        p = om * np.exp(-b * Ei - lnZ[i])
        e = (p * Ei).sum()
        c = b * b * (p * (Ei - e) ** 2).sum()
        assert out[i, 0] == b
        assert out[i, 1] == pytest.approx(e / N, rel=1e-6, abs=1e-7)
        assert out[i, 2] == pytest.approx(c / N, rel=1e-5, abs=1e-7)
        assert out[i, 3] == pytest.approx(-lnZ[i] / N, rel=1e-7)
    assert np.allclose(BFnew, out[:, 3] * N, rtol=0, atol=0)   # BF updated in place


def test_mhr_iterations_converge_from_rough_free_energies():
    L, N = 4, 16
    om = exact_dos(L)
    Ei = 4 * np.arange(N + 1) - 2 * N
    B = np.linspace(0.0, 0.8, 9)
    lnZ = np.array([math.log((om * np.exp(-b * Ei)).sum()) for b in B])
    R = np.full(len(B), 10 ** 8)
    H = sum(r * om * np.exp(-b * Ei - lz) for b, lz, r in zip(B, lnZ, R))
    rough = -lnZ + np.linspace(0, 0.3, len(B))
    rough[0] = -lnZ[0]
    out, _ = mhr(N, B, rough, R, np.rint(H), niter=100)
    # the fixed point is defined up to one additive constant: compare free-energy differences
    bf = out[:, 3 * 99 + 3] * N
    assert np.allclose(bf - bf[0], -(lnZ - lnZ[0]), rtol=1e-6, atol=1e-6)
    first = out[:, 3] * N
    assert not np.allclose(first - first[0], -(lnZ - lnZ[0]), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("coding,w", [(oracle.SSC, 32), (oracle.MSC, 32), (oracle.MSC, 64)])
def test_run_histograms_count_every_replica_of_every_step(coding, w):
    rc, runs, _, _, H, M = oracle.run_mhr(L=8, coding=coding, msbits=w, Rinit=700, eqsweeps=3,
                                          dB=0.05, Bfin=1.0, seed=11, runs=2, threads=4)
    assert rc == 0
    for r in range(2):
        assert sum(H[r]) == sum(x.R for x in runs[r])        # includes the initial population
        assert len(M[r]) == len(runs[r]) and all(m[0] == x.B for m, x in zip(M[r], runs[r]))
    # the reference keeps the reweighted BF[0] for the next run's row 0 (SSC:435, 758)
    assert runs[1][0].BF == M[0][0][3]
    assert runs[0][0].BF == -math.log(2.0)
