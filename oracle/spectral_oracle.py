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
  Kubo-Bastin, live     snippets.h:351-405 over numerical.h:137-284: the route
    route                the reference actually runs -- dct_iii_complex of the
                         weighted stacks (phi_3), fft_half_conj of
                         e^{i pi n/2N} a_n aL_n (phi_1) and of n times that
                         (phi_2), four trace_braket_complex sums per node, and
                         zeta_k = (e_k p_k + i s_k w_k)/(1 - e_k^2)^2/Volume;
                         restated step by step (kubo_zeta_fft), including the
                         split-radix recursion with its zero-padding add_sub
                         and the sqrt(2) C[N/2] term
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


def g_mn_matrix(e, N):
    """g_mn(e, m, n) of snippets.h:343-349 for all m, n < N at one energy,
    element by element (row m, column n)."""
    k = np.arange(N, dtype=np.float64)
    ad = np.where(k == 0, 1.0, 2.0)
    theta = math.acos(e)
    es = math.sqrt(1 - e * e)
    T = np.cos(k * theta)
    m, n = k[:, None], k[None, :]
    pre = ad[:, None] * ad[None, :]
    return pre * ((e - 1j * n * es) * np.exp(1j * n * theta) * T[:, None]
                  + (e + 1j * m * es) * np.exp(-1j * m * theta) * T[None, :])


def kubo_zeta(gamma, e, volume=1.0):
    """zeta(e) = sum_ij Gamma_ij g_mn(e, i, j) / (1 - e^2)^2 / Volume
    (snippets.h:404, 406-415): the direct double sum, one full N x N weight
    matrix per energy."""
    gamma = np.asarray(gamma)
    N = gamma.shape[0]
    out = []
    for ee in np.atleast_1d(e):
        s = np.sum(gamma * g_mn_matrix(float(ee), N))
        out.append(s / (1 - ee * ee) ** 2 / volume)
    return np.array(out)


# ------------------------------------------ the reference's live zeta route
def _slice_index(start, step, i, total):
    """numerical.h:151-156: start + i step, wrapped once when negative."""
    j = start + i * step
    return j + total if j < 0 else j


def _add_sub(x, i, j):
    """numerical.h:157-163: (x_i + x_j, x_i - x_j), with x read as zero past
    its end (the half-length input of fft_half_conj)."""
    n = x.shape[0]
    if i < n and j < n:
        return x[i] + x[j], x[i] - x[j]
    if j >= n:
        return x[i], x[i]
    return x[j], -x[j]


def _radix_fft_conj(x, start, step, n, total, select=0):
    """numerical.h:196-227: split-radix transform with e^{+2 pi i j k / n}
    over x[start + k step]; select > 0 keeps the first half of the output."""
    if n <= 1:
        return [x[_slice_index(start, step, 0, total)]]
    if n == 2:
        return list(_add_sub(x, _slice_index(start, step, 0, total), _slice_index(start, step, 1, total)))
    if n == 4:
        i0, i1, i2, i3 = (_slice_index(start, step, k, total) for k in range(4))
        a, b = _add_sub(x, i0, i2)
        c, d = _add_sub(x, i1, i3)
        if select > 0:
            return [a + c, b + 1j * d]
        return [a + c, b + 1j * d, a - c, b - 1j * d]
    U = _radix_fft_conj(x, start, 2 * step, n // 2, total)
    Z = _radix_fft_conj(x, start + step, 4 * step, n // 4, total)
    Zm = _radix_fft_conj(x, start - step, 4 * step, n // 4, total)
    X = [None] * (n // 2 if select > 0 else n)
    q = n // 4
    for i in range(q):
        w = cmath.exp(complex(0.0, -2.0 * math.pi * i / n))
        wz = Z[i] * w.conjugate()
        wzm = Zm[i] * w
        s_ = wz + wzm
        d_ = -1j * (wz - wzm)
        X[i] = U[i] + s_
        X[i + q] = U[i + q] - d_
        if select > 0:
            continue
        X[i + 2 * q] = U[i] - s_
        X[i + 3 * q] = U[i + q] + d_
    return X


def dct_iii_complex(C):
    """numerical.h:229-249: the DCT-III of a power-of-two sequence of blocks
    through one length-N conjugate FFT (pre-twiddle e^{-+ 2 pi i k/4N},
    sqrt(2) on the middle term, even/odd output interleave)."""
    C = np.asarray(C)
    N = C.shape[0]
    assert N > 0 and N & (N - 1) == 0, "dct_iii_complex: N must be power of 2"
    Z = np.empty_like(C)
    Z[0] = C[0]
    for i in range(1, N // 2):
        w = cmath.exp(complex(0.0, -2.0 * math.pi * i / (4 * N)))
        Z[i] = w.conjugate() * (C[i] - 1j * C[N - i])
        Z[N - i] = w * (C[i] + 1j * C[N - i])
    if N > 1:
        Z[N // 2] = math.sqrt(2) * C[N // 2]
    y = _radix_fft_conj(Z, 0, 1, N, N)
    x = [None] * N
    for i in range(N // 2):
        x[2 * i] = y[i]
        x[2 * i + 1] = y[N - i - 1]
    if N == 1:
        x[0] = y[0]
    return np.array(x)


def fft_half_conj(C):
    """numerical.h:276-283: X_k = sum_{j<N} C_j e^{+2 pi i j k / 2N}, k < N
    (length-2N conjugate transform of the zero-padded input, first half)."""
    C = np.asarray(C)
    N = C.shape[0]
    assert N > 0 and N & (N - 1) == 0, "fft_half_conj: N must be power of 2"
    return np.array(_radix_fft_conj(C, 0, 1, 2 * N, 2 * N, 1))


def kubo_zeta_fft(aL, aR, volume=1.0):
    """zeta at the N Chebyshev nodes by the reference's live route
    (snippets.h:351-405).  aL, aR: (N, K) complex stacks, already weighted by
    the kernel (aL_n = g_n T_n|r>, aR_m = g_m v_x T_m v_y|r>, flattened over
    sites and vectors).  Returns (epsilon_k, zeta_k), k < N."""
    aL = np.asarray(aL, dtype=np.complex128)
    aR = np.asarray(aR, dtype=np.complex128)
    N = aL.shape[0]
    ad = np.array([1.0 if i == 0 else 2.0 for i in range(N)])
    phi3L, phi3R = dct_iii_complex(aL), dct_iii_complex(aR)
    c = np.array([cmath.exp(1j * math.pi * i / (2 * N)) * ad[i] for i in range(N)])
    bL, bR = c[:, None] * aL, c[:, None] * aR
    phi1L, phi1R = fft_half_conj(bL), fft_half_conj(bR)
    nn = np.arange(N, dtype=np.float64)[:, None]
    phi2L, phi2R = fft_half_conj(nn * bL), fft_half_conj(nn * bR)

    def tr(A, B):  # trace_braket_complex, numerical.h:137-144
        return np.sum(np.conj(A) * B)

    eps = np.empty(N)
    zeta = np.empty(N, dtype=np.complex128)
    for k in range(N):
        p = tr(phi3L[k], phi1R[k]) + tr(phi1L[k], phi3R[k])
        w = tr(phi2L[k], phi3R[k]) - tr(phi3L[k], phi2R[k])
        e = math.cos((k + 0.5) / N * math.pi)
        eps[k] = e
        zeta[k] = (e * p + 1j * math.sqrt(1 - e * e) * w) / (1 - e * e) ** 2 / volume
    return eps, zeta


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
