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
  kubo_zeta / kubo_sigma   snippets.h:343-449 (the reference's unnormalised
                          zeta and its integral anchored at e = 0)
  kubo_sigma_physical     sigma_xy in units of e^2/h (not in the reference;
                          the conversion is derived below)
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


def kubo_nodes(N):
    """Chebyshev nodes e_k = cos((k + 1/2) pi / N), k < N (descending), where
    the reference evaluates the Kubo-Bastin integrand (snippets.h:397-402)."""
    k = np.arange(int(N), dtype=np.float64)
    return np.cos((k + 0.5) / N * math.pi)


def integrate_from_zero(E, z):
    """Cumulative integral of z over the ascending grid E, anchored at e = 0
    (snippets.h:436-449): the interval [E[i0-1], E[i0]] that brackets 0
    (i0 = first index with E > 0) is split at its midpoint, with the midpoint
    value taken as the mean of its ends; from there the trapezoid rule runs
    outward in both directions."""
    E = np.asarray(E, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    i0 = int(np.searchsorted(E, 0.0, side="right"))
    if i0 < 1 or i0 >= E.size:
        raise ValueError("the energy grid must bracket e = 0")
    out = np.zeros(E.size)
    h = E[i0] - E[i0 - 1]
    out[i0] = (z[i0] * 3 / 4 + z[i0 - 1] / 4) * h / 2
    out[i0 - 1] = -(z[i0] / 4 + z[i0 - 1] * 3 / 4) * h / 2
    acc = out[i0]
    for i in range(i0 + 1, E.size):
        acc = acc + (z[i] + z[i - 1]) / 2 * (E[i] - E[i - 1])
        out[i] = acc
    acc = out[i0 - 1]
    for i in range(i0 - 2, -1, -1):
        acc = acc - (z[i] + z[i + 1]) / 2 * (E[i + 1] - E[i])
        out[i] = acc
    return out


def kubo_sigma(gamma, e_fermi, volume=1.0, zeta=None):
    """sigma_xy at rescaled Fermi energies, by the reference's scheme
    (snippets.h:397-449): zeta at the N Chebyshev nodes, N/24 nodes cut at
    each end (where (1 - e^2)^-2 diverges), Re zeta integrated from e = 0
    outward (integrate_from_zero), then interpolated linearly to e_fermi.
    zeta: callable (gamma, energies, volume) -> complex zeta (default: the
    host evaluation kubo_zeta; kpm.kubo_zeta evaluates it on the device).
    Returns (sigma at e_fermi, grid E ascending, Re zeta on E, integral on E)."""
    gamma = np.asarray(gamma)
    N = gamma.shape[0]
    eps = kubo_nodes(N)
    z = (zeta or kubo_zeta)(gamma, eps, volume)
    cut = N // 24
    E = eps[cut:N - cut][::-1]
    zr = np.real(z)[cut:N - cut][::-1]
    integral = integrate_from_zero(E, zr)
    ef = np.atleast_1d(np.asarray(e_fermi, dtype=np.float64))
    return np.interp(ef, E, integral), E, zr, integral


# ------------------------------------------- sigma_xy in physical units
def sigma_xy_e2h_factor(n_sites, R, area, volume=1.0):
    """c in sigma_xy [e^2/h] = c * integral of Re zeta(e) de over rescaled e.

    The reference's zeta (snippets.h:343-405) is built from v = [r, H~]
    (no factor i, scaled H~), stochastic kets of norm 1 (elements
    e^{i phi}/sqrt(n)), no 1/R and no e^2/h.  With hbar = 1 the Bastin
    formula is

      sigma_xy = (i e^2 / Omega) int de f(e) Tr[v_x d(e-H) v_y G+'(e)
                                                - v_x G-'(e) v_y d(e-H)]

    with physical velocities v = i[H, r] = -i a [r, H~].  Expanding d and G+-
    in Chebyshev polynomials turns the reference's double sum into
    zeta(e) Volume = i pi x (that trace, in rescaled units and with v = [r, H~]);
    the a's of v^2 (-a^2), of d (1/a), of G' (1/a^2) and of de (a) cancel.
    The stochastic trace estimates sum_r <r|X|r> = (R/n) Tr X.  So

      sigma_xy / (e^2/h) = -(2 Volume n / (R Omega)) int_{-1}^{mu} Re zeta(e) de,

    Omega the sample area (2D; a volume in 3D, where sigma is per length) in
    the length unit of the lattice vectors.  The imaginary part of zeta
    integrates to zero.  Checked against the exact TKNN sum of a magnetic
    torus, where it gives the integer Landau-level plateaus."""
    return -2.0 * float(volume) * float(n_sites) / (float(R) * float(area))


def integrate_from_bottom(E, z):
    """Cumulative trapezoid integral of z over the ascending grid E, zero at
    E[0]."""
    E = np.asarray(E, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    out = np.zeros(E.size)
    if E.size > 1:
        out[1:] = np.cumsum((z[1:] + z[:-1]) / 2 * np.diff(E))
    return out


def kubo_sigma_physical(gamma, e_fermi, n_sites, R, area, volume=1.0, zeta=None):
    """sigma_xy in units of e^2/h at rescaled Fermi energies.

    zeta at the N Chebyshev nodes as the reference evaluates it, N/24 nodes
    cut at each end (kubo_sigma), Re zeta integrated from the lowest kept
    node upward (the band bottom, where the occupied states start, instead
    of the reference's anchor at e = 0), times sigma_xy_e2h_factor.  The
    cut nodes must hold no states: keep the spectrum inside
    |e| < cos(pi (N/24 + 1/2)/N) (about 0.99) by the padding of [E_min, E_max].
    Returns (sigma at e_fermi, grid E ascending, sigma on E)."""
    gamma = np.asarray(gamma)
    N = gamma.shape[0]
    eps = kubo_nodes(N)
    z = (zeta or kubo_zeta)(gamma, eps, volume)
    cut = N // 24
    E = eps[cut:N - cut][::-1]
    zr = np.real(z)[cut:N - cut][::-1]
    grid = sigma_xy_e2h_factor(n_sites, R, area, volume) * integrate_from_bottom(E, zr)
    ef = np.atleast_1d(np.asarray(e_fermi, dtype=np.float64))
    return np.interp(ef, E, grid), E, grid
