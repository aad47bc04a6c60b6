"""Physics cross-checks: Kaufman's exact finite-lattice free energy (pinned by brute force),
the reference's sample output against it, and the engine's full config-2 runs against both."""
import math
import os

import numpy as np
import pytest

from paising_amd import exact

FIX = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures")


def brute_ln_z(beta, L):
    N = L * L
    vals = []
    for bits in range(1 << N):
        s = np.array([1 if (bits >> i) & 1 else -1 for i in range(N)]).reshape(L, L)
        E = -(s * np.roll(s, 1, 0)).sum() - (s * np.roll(s, 1, 1)).sum()
        vals.append(-beta * E)
    v = np.array(vals, float)
    return v.max() + math.log(np.exp(v - v.max()).sum())


@pytest.mark.parametrize("L", [2, 4])
@pytest.mark.parametrize("beta", [0.05, 0.3, 0.4406868, 0.8])
def test_kaufman_matches_brute_force(L, beta):
    assert exact.ln_z(beta, L) == pytest.approx(brute_ln_z(beta, L), rel=1e-13, abs=1e-12)


def test_large_lattice_tends_to_onsager():
    """Disordered phase: the torus converges to Onsager's bulk value; ordered phase: the two
    degenerate magnetised states add ln 2 to ln Z, i.e. -ln2/N to beta F / N."""
    assert exact.beta_f_per_site(0.3, 256) == pytest.approx(exact.onsager_beta_f(0.3), abs=1e-12)
    L = 256
    assert exact.beta_f_per_site(0.6, L) + math.log(2) / L**2 == pytest.approx(
        exact.onsager_beta_f(0.6), abs=1e-9)


def fixture_rows(name):
    return [list(map(float, l.split())) for l in open(os.path.join(FIX, name))]


@pytest.mark.parametrize("name", ["msc_L64_R20000_EqSw100_dB0.005_s100.dat",
                                  "ssc_cuda101_L64_R20000_EqSw100_dB0.005_s100.dat"])
def test_reference_output_close_to_exact(name):
    """The reference's PA free energy is within 3e-5 per site of the exact value (L=64)."""
    for row in fixture_rows(name)[1:]:
        assert abs(row[6] - exact.beta_f_per_site(row[0], 64)) < 3e-5, row[0]


@pytest.mark.gpu
def test_engine_full_config2_msc_equals_reference_output_and_exact():
    """All 200 rows of the engine's config-2 MSC run equal the reference's sample output
    (fixture rung C8 on the GPU) and track Kaufman's free energy."""
    import paising_amd as pa
    rows, _ = pa.run_all(pa.make_params(L=64, coding=pa.MSC, msbits=32, Rinit=20000, eqsweeps=100,
                                        dB=0.005, Bfin=1.0, seed=100))
    text = pa.format_rows(rows)
    ref = open(os.path.join(FIX, "msc_L64_R20000_EqSw100_dB0.005_s100.dat")).read()
    assert text == ref
    worst = max(abs(r.BF - exact.beta_f_per_site(r.B, 64)) for r in rows[1:])
    assert worst < 3e-5
