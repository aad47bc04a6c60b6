"""Exact free energy of the 2-D ferromagnetic Ising model on an L x L torus (J = 1).

B. Kaufman, "Crystal statistics. II. Partition function evaluated by spinor analysis",
Phys. Rev. 76, 1232 (1949):  Z = 1/2 (2 sinh 2K)^(N/2) (Z1 + Z2 + Z3 + Z4) with
Z1,2 = prod_r 2 cosh / sinh (L gamma_{2r+1} / 2),  Z3,4 = prod_r 2 cosh / sinh (L gamma_{2r} / 2),
cosh gamma_l = cosh 2K coth 2K - cos(l pi / L), and gamma_0 = 2K + ln tanh K (signed).
Population annealing estimates beta F = -ln Z as BF = -N ln 2 - sum_i lnQ_i (the BF column).
"""
import math

import numpy as np


def ln_z(beta, L):
    """ln Z(beta) of the L x L periodic lattice, evaluated in log space."""
    K = float(beta)
    if K <= 0.0:
        return L * L * math.log(2.0)
    n = L
    ell = np.arange(2 * n)
    c = math.cosh(2 * K) / math.tanh(2 * K) - np.cos(ell * math.pi / n)
    gam = np.arccosh(np.maximum(c, 1.0))
    gam[0] = 2 * K + math.log(math.tanh(K))
    odd, even = gam[1::2], gam[0::2]

    def ln_cosh_prod(g):  # sum log(2 cosh x) = |x| + log1p(exp(-2|x|))
        x = np.abs(L * g / 2)
        return float(np.sum(x + np.log1p(np.exp(-2 * x))))

    def ln_sinh_prod(g):  # sum log|2 sinh x| = |x| + log1p(-exp(-2|x|))
        y = L * g / 2
        x = np.abs(y)
        return float(np.sum(x + np.log1p(-np.exp(-2 * x)))), int(np.prod(np.sign(y)))

    l1 = ln_cosh_prod(odd)
    l2, s2 = ln_sinh_prod(odd)
    l3 = ln_cosh_prod(even)
    l4, s4 = ln_sinh_prod(even)
    mx = max(l1, l2, l3, l4)
    tot = (math.exp(l1 - mx) + s2 * math.exp(l2 - mx) + math.exp(l3 - mx) + s4 * math.exp(l4 - mx))
    return math.log(0.5) + (L * L / 2) * math.log(2 * math.sinh(2 * K)) + mx + math.log(tot)


def beta_f_per_site(beta, L):
    """beta * F / N = -ln Z / N: the quantity of the results file's BF/N column."""
    return -ln_z(beta, L) / (L * L)


def onsager_beta_f(beta):
    """Infinite-lattice beta F per site (Onsager 1944), by numerical quadrature."""
    K = float(beta)
    k = 2 * math.sinh(2 * K) / math.cosh(2 * K) ** 2
    th = np.linspace(0, math.pi, 200001)
    integrand = np.log(0.5 * (1 + np.sqrt(np.maximum(1 - k * k * np.sin(th) ** 2, 0.0))))
    integral = np.trapz(integrand, th) / (2 * math.pi)
    return -(math.log(2 * math.cosh(2 * K)) + integral)
