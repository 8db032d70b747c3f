"""Kernel damping and spectral reconstruction from Chebyshev moments.

Host-side mirror of the reference's front end:
  kernels                 numerical.h:393-447 (Dirichlet, Fejer, Lorentz,
                          Jackson, Lanczos, Wang-Zunger, kernel_seq)
  expn_coef_alpha_seq     numerical.h:304-316: 1/(z - x) = sum alpha_n(z) T_n(x)
  green_imag              lattice.h:618-624: f(e) = sum g_n Im alpha_n(e) mu_n
  density_of_states       functions.h:27-30: DoS(e) = -f(e)/(R pi), e in
                          rescaled units (E = a e + b; divide by a for
                          states per energy)
  dct_reconstruction      lattice.h:625-640: REDFT01 of g*mu at the Chebyshev
                          nodes e_k = cos((k + 1/2) pi / N), times 1/sqrt(1-e^2)
For real e in (-1, 1), alpha_n(e) = -2i e^{-i n theta}/sin(theta) (n >= 1),
alpha_0 = -i/sin(theta), e = cos(theta); green_imag evaluates that closed
form (the reference multiplies out powers of e^{-i theta}, equal up to
rounding).
"""
from __future__ import annotations

import math

import numpy as np


# ------------------------------------------------------------------ kernels
def dirichlet(N):
    return np.ones(int(N))


def fejer(N):
    n = np.arange(int(N), dtype=np.float64)
    return 1.0 - n / N


def lorentz(N, lam):
    n = np.arange(int(N), dtype=np.float64)
    return np.sinh(lam * (1.0 - n / N)) / math.sinh(lam)


def jackson(N):
    N1 = int(N) + 1
    n = np.arange(int(N), dtype=np.float64)
    s1 = math.pi / N1
    s2 = n * s1
    return ((N1 - n) * np.cos(s2) + np.sin(s2) / math.tan(s1)) / N1


def lanczos(N, M):
    """As numerical.h:428-435: (sin(n pi/N) / (pi/N))^M -- the reference's
    formula, without the 1/n of the textbook sigma factor."""
    n = np.arange(int(N), dtype=np.float64)
    pi_N = math.pi / N
    return np.power(np.sin(n * pi_N) / pi_N, M)


def wang_zunger(N, alpha, beta):
    n = np.arange(int(N), dtype=np.float64)
    return np.exp(-np.power(n * (alpha / N), beta))


# ------------------------------------------------------- Chebyshev expansion
def expn_coef_alpha_seq(z, n):
    """alpha_0..alpha_{n-1} of 1/(z - x) = sum alpha_n(z) T_n(x)
    (numerical.h:304-316), for complex z."""
    z = complex(z)
    s1 = np.sqrt(1.0 - z * z + 0j)
    s2 = z - 1j * s1
    k = np.arange(int(n))
    g = -2j / s1 * np.power(s2, k)
    g[0] = -1j / s1
    return g


def cheb_T_table(N, e):
    """T_n(e) for n < N and |e| <= 1, shape (N, len(e)), by the recursion."""
    e = np.asarray(e, dtype=np.float64)
    T = np.empty((int(N), e.size))
    T[0] = 1.0
    if N > 1:
        T[1] = e
    for n in range(2, int(N)):
        T[n] = 2.0 * e * T[n - 1] - T[n - 2]
    return T


def green_imag(mu, g, e):
    """f(e) = sum_n g_n Im(alpha_n(e)) mu_n for rescaled e in (-1, 1)."""
    mu = np.asarray(mu, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    e = np.atleast_1d(np.asarray(e, dtype=np.float64))
    c = g * mu
    T = cheb_T_table(mu.size, e)
    s = c[0] * T[0] + 2.0 * (c[1:, None] * T[1:]).sum(axis=0)
    return -s / np.sqrt(1.0 - e * e)


def density_of_states(mu, g, e, R):
    """DoS(e) = -f(e)/(R pi) (functions.h:27-30), e rescaled."""
    return -green_imag(mu, g, e) / (R * math.pi)


def dct_reconstruction(mu, g):
    """(e_k, f_k): DCT-III (REDFT01) of g*mu at e_k = cos((k+1/2) pi/N),
    divided by sqrt(1 - e_k^2) (lattice.h:625-640)."""
    c = np.asarray(g, dtype=np.float64) * np.asarray(mu, dtype=np.float64)
    N = c.size
    k = np.arange(N)
    theta = (k + 0.5) * math.pi / N
    # REDFT01: y_k = c_0 + 2 sum_{n>=1} c_n cos(pi n (k + 1/2)/N)
    y = c[0] + 2.0 * np.cos(np.outer(theta, np.arange(1, N))) @ c[1:]
    e = np.cos(theta)
    return e, y / np.sqrt(1.0 - e * e)


def linspace(start, end, n, endpoint=False):
    """utils.h:15-21: endpoint excluded by default."""
    step = (end - start) / n if not endpoint else (end - start) / (n - 1)
    return start + step * np.arange(int(n))


# ------------------------------------------------------------ Kubo-Bastin
def _kubo_vectors(N, e):
    """u, v, p, q of zeta(e) = u^T G v + p^T G q (the separable form of the
    double sum of snippets.h:343-349, 406-415 with a_0 = 1, a_{n>0} = 2)."""
    e = np.atleast_1d(np.asarray(e, dtype=np.float64))
    n = np.arange(N, dtype=np.float64)[:, None]
    ad = np.where(n == 0, 1.0, 2.0)
    th = np.arccos(e)[None, :]
    s = np.sqrt(1.0 - e * e)[None, :]
    T = np.cos(n * th)
    u = ad * T
    v = ad * (e[None, :] - 1j * n * s) * np.exp(1j * n * th)
    p = ad * (e[None, :] + 1j * n * s) * np.exp(-1j * n * th)
    return u, v, p, T * ad


def kubo_zeta(gamma, e, volume=1.0):
    """zeta(e) = sum_ij Gamma_ij g_mn(e, i, j) / (1 - e^2)^2 / Volume at rescaled
    energies e (snippets.h:400-405; no 1/R and no e^2/h prefactor, as there)."""
    gamma = np.asarray(gamma)
    N = gamma.shape[0]
    e = np.atleast_1d(np.asarray(e, dtype=np.float64))
    u, v, p, q = _kubo_vectors(N, e)
    z = np.sum(u * (gamma @ v), axis=0) + np.sum(p * (gamma @ q), axis=0)
    return z / (1.0 - e * e) ** 2 / volume


def kubo_sigma(gamma, e_fermi, volume=1.0, trim=1.0 / 24, points_per_moment=8):
    """sigma(mu) = integral of Re zeta from the lower edge of the trimmed
    spectral interval to mu, for rescaled Fermi energies mu (T = 0).  The
    interval edges e = +-cos(trim*pi) are cut as the reference cuts N/24
    Chebyshev nodes at each end (snippets.h:421-426), where (1 - e^2)^-2
    diverges.  Trapezoid rule on a uniform grid in theta = arccos(e).
    Returns (sigma at e_fermi, the grid e, Re zeta on the grid)."""
    gamma = np.asarray(gamma)
    N = gamma.shape[0]
    M = max(64, points_per_moment * N)
    th = np.linspace(math.pi * (1 - trim), math.pi * trim, M)  # e ascending
    e = np.cos(th)
    zr = np.real(kubo_zeta(gamma, e, volume))
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (zr[1:] + zr[:-1]) * np.diff(e))])
    ef = np.atleast_1d(np.asarray(e_fermi, dtype=np.float64))
    return np.interp(ef, e, cum), e, zr
