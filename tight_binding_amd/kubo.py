"""Kubo-Bastin sigma_xy on the device: the flow of the reference's
conductivity_Kubo_Bastin_test (snippets.h:306-456).

  H~ = (H - b)/a with a, b from [E_min - 1e-8, E_max + 1e-8]  (snippets.h:311)
  v_x = X H~ - H~ X, v_y = Y H~ - H~ Y                          (snippets.h:316-317)
  R random-phase vectors, Lorentz kernel g_n (lambda = 6)       (snippets.h:318-320)
  Gamma_nm = sum conj(g_n T_n(H~)|r>) (g_m v_x T_m(H~) v_y|r>)  (snippets.h:326-342)
  zeta(e) = sum Gamma_nm g_mn(e) / (1 - e^2)^2 / Volume          (snippets.h:343-349, 400-405)
  sigma(mu) = integral of Re zeta up to mu                       (snippets.h:436-449)

Gamma is computed on the GPU (two vector stacks in HBM and a complex
contraction on matrix cores, fp64 or split-bf16 x3); zeta and sigma are
evaluated on the host from the N x N Gamma.
"""
from __future__ import annotations

import numpy as np

from . import kpm, spectral
from . import lattice as L


def velocity_operators(Hs, X, Y):
    """v_x, v_y = [X, H~], [Y, H~] (no factor i, as snippets.h:316-317)."""
    return L.commutator_diag(X, Hs), L.commutator_diag(Y, Hs)


def conductivity_kubo_bastin(ctx, H, X, Y, N, E_min, E_max, volume=1.0, R=4, lam=6.0, seed=1,
                             precision=kpm.GAMMA_FP64, fermi_energies=None, kets=None):
    """Returns dict(gamma, a, b, zeta_e, zeta, sigma, fermi_energies, stats).
    H: complex CSC Hamiltonian; X, Y: diagonal position operators."""
    lo, hi = E_min - 1e-8, E_max + 1e-8
    a, b = (hi - lo) / 2, (hi + lo) / 2
    Hs = L.rescale(H, a, b)
    vx, vy = velocity_operators(Hs, X, Y)
    g = spectral.lorentz(N, lam)
    MH, MX, MY = ctx.matrix(Hs), ctx.matrix(vx), ctx.matrix(vy)
    K = ctx.block(kets) if kets is not None else ctx.random_block(Hs.n, R, seed=seed)
    gamma, st = kpm.kubo_gamma(ctx, MH, MX, MY, K, N, precision=precision, g=g)
    out = dict(gamma=gamma, a=a, b=b, stats=st)
    if fermi_energies is not None:
        ef = (np.asarray(fermi_energies, dtype=np.float64) - b) / a
        sigma, e, zr = spectral.kubo_sigma(gamma, ef, volume)
        out.update(fermi_energies=np.asarray(fermi_energies), sigma=sigma, zeta_e=e, zeta=zr)
    for m in (MH, MX, MY, K):
        m.close()
    return out
