"""CPU oracle for the spectral front end -- TEST INFRASTRUCTURE ONLY.

Restated (reference file:line):
  expn_coef_alpha_seq   numerical.h:304-316  (the running product sn *= s2)
  GreenFunctionImag()   lattice.h:618-624    f(e) = sum g_n Im(alpha_n) mu_n
  density_of_states_seq functions.h:27-30    DoS(e) = -f(e)/(R pi)
  GreenFunctionImag::fft lattice.h:625-640   FFTW REDFT01 of g*mu, restated
                         with scipy.fft.dct(type=3), which computes the same
                         unnormalised transform
  Kubo-Bastin           snippets.h:337-349, 406-415: Gamma_ij = sum over sites
                         and vectors of conj(aL_i) aR_j, and the direct
                         double sum zeta(e) = sum_ij Gamma_ij g_mn(e, i, j)
  sigma integration     snippets.h:397-402, 420-449: Chebyshev-node grid, N/24
                         nodes cut at each end, integral of Re zeta anchored
                         at e = 0 (3/4-1/4 split of the bracketing interval)
                         and continued outward by the trapezoid rule
"""
from __future__ import annotations

import cmath
import math

import numpy as np
import scipy.fft


def expn_coef_alpha_seq(z, n):
    z = complex(z)
    s1 = cmath.sqrt(1.0 - z * z)
    s2 = z - 1j * s1
    sn = -2j / s1
    g = [0j] * n
    g[0] = -1j / s1
    for i in range(1, n):
        sn *= s2
        g[i] = sn
    return g


def green_imag(mu, g, e):
    out = []
    for ee in np.atleast_1d(e):
        alpha = expn_coef_alpha_seq(complex(ee), len(mu))
        s = 0.0
        for gn, an, m in zip(g, alpha, mu):
            s += gn * an.imag * m
        out.append(s)
    return np.array(out)


def density_of_states(mu, g, e, R):
    return np.array([-f / R / math.pi for f in green_imag(mu, g, e)])


def dct_reconstruction(mu, g):
    N = len(mu)
    theta = 0.5 * math.pi / N + (math.pi / N) * np.arange(N)
    y = scipy.fft.dct(np.asarray(g) * np.asarray(mu), type=3)
    e = np.cos(theta)
    return e, y / np.sqrt(1.0 - e * e)


def kubo_gamma(aL, aR):
    """Gamma[i, j] = sum_{sites, vectors} conj(aL_i) * aR_j (snippets.h:337-342),
    aL, aR arrays (N, rows, R)."""
    N = aL.shape[0]
    A = aL.reshape(N, -1)
    B = aR.reshape(N, -1)
    return np.conj(A) @ B.T


def g_mn(e, m, n):
    """snippets.h:343-349 with a_d(0) = 1, a_d(k>0) = 2."""
    ad = lambda k: 1.0 if k == 0 else 2.0  # noqa: E731
    pre = ad(m) * ad(n)
    theta = math.acos(e)
    es = math.sqrt(1 - e * e)
    Tm, Tn = math.cos(m * theta), math.cos(n * theta)
    return pre * ((e - 1j * n * es) * cmath.exp(1j * n * theta) * Tm
                  + (e + 1j * m * es) * cmath.exp(-1j * m * theta) * Tn)


def kubo_zeta(gamma, e, volume=1.0):
    """zeta(e) = sum_ij Gamma_ij g_mn(e, i, j) / (1 - e^2)^2 / Volume
    (snippets.h:404, 406-415)."""
    N = gamma.shape[0]
    out = []
    for ee in np.atleast_1d(e):
        s = 0j
        for i in range(N):
            for j in range(N):
                s += gamma[i, j] * g_mn(ee, i, j)
        out.append(s / (1 - ee * ee) ** 2 / volume)
    return np.array(out)


def kubo_nodes(N):
    """epsilon_k = cos((k + 0.5)/N pi), k < N (snippets.h:397-402)."""
    return np.array([math.cos((k + 0.5) / N * math.pi) for k in range(N)])


def kubo_sigma_reference(epsilon, zeta):
    """snippets.h:420-449: drop N/24 nodes at each end, reverse to ascending,
    find E0_i = upper_bound(E, 0.0), seed the two entries around e = 0 with
    the 3/4-1/4 rule, then trapezoid outward.  Returns (E, Re-zeta integral).
    Only the real part is restated (sigma_xy = integral of Re zeta)."""
    N = len(epsilon)
    cut = N // 24
    E = list(epsilon[cut:cut + (N - 2 * cut)])[::-1]
    zr = [complex(z).real for z in zeta[cut:cut + (N - 2 * cut)]][::-1]
    E0 = next(i for i, x in enumerate(E) if x > 0.0)
    out = [0.0] * len(E)
    out[E0] = (zr[E0] * 3 / 4 + zr[E0 - 1] / 4) * (E[E0] - E[E0 - 1]) / 2
    out[E0 - 1] = -(zr[E0] / 4 + zr[E0 - 1] * 3 / 4) * (E[E0] - E[E0 - 1]) / 2
    for i in range(E0 + 1, len(E)):
        out[i] = out[i - 1] + (zr[i] + zr[i - 1]) / 2 * (E[i] - E[i - 1])
    for i in range(E0 - 2, -1, -1):
        out[i] = out[i + 1] - (zr[i] + zr[i + 1]) / 2 * (E[i + 1] - E[i])
    return np.array(E), np.array(out)
