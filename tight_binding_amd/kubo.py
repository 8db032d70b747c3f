"""Kubo-Bastin sigma_xy on the device: the flow of the reference's
conductivity_Kubo_Bastin_test (snippets.h:306-456).

  H~ = (H - b)/a with a, b from [E_min - 1e-8, E_max + 1e-8]  (snippets.h:311)
  v_x = X H~ - H~ X, v_y = Y H~ - H~ Y                          (snippets.h:316-317)
  R random-phase vectors, Lorentz kernel g_n (lambda = 6)       (snippets.h:318-320)
  Gamma_nm = sum conj(g_n T_n(H~)|r>) (g_m v_x T_m(H~) v_y|r>)  (snippets.h:326-342)
  zeta(e) = sum Gamma_nm g_mn(e) / (1 - e^2)^2 / Volume          (snippets.h:343-349, 400-405)
  sigma(mu) = integral of Re zeta from e = 0 to mu               (snippets.h:420-449)

With the sample area given, solve() also reports sigma_xy in units of e^2/h
(spectral.kubo_sigma_physical): -(2 Volume n/(R area)) times the integral of
Re zeta from the band bottom, the reference's zeta converted to the Bastin
formula with physical velocities i[H, r].

Gamma is computed on the GPU (two vector stacks in HBM and a complex
contraction on matrix cores, fp64 or split-bf16 x3); zeta at the Chebyshev
nodes is evaluated on the GPU from the N x N Gamma, and the integral over
those N nodes on the host.
"""
from __future__ import annotations

import time

import numpy as np

from . import kpm, spectral
from . import lattice as L


def velocity_operators(Hs, X, Y):
    """v_x, v_y = [X, H~], [Y, H~] (no factor i, as snippets.h:316-317)."""
    return L.commutator_diag(X, Hs), L.commutator_diag(Y, Hs)


class KuboSolver:
    """One Hamiltonian resident on the device with its velocity operators;
    solve() runs one sigma_xy computation from fresh random vectors.
    H: complex CSC Hamiltonian; X, Y: diagonal position operators (CSC)."""

    def __init__(self, ctx: kpm.Context, H, X, Y, E_min, E_max, volume=1.0, lam=6.0):
        lo, hi = E_min - 1e-8, E_max + 1e-8
        self.a, self.b = (hi - lo) / 2, (hi + lo) / 2
        self.ctx, self.volume, self.lam = ctx, float(volume), float(lam)
        self.Hs = L.rescale(H, self.a, self.b)
        self.vx, self.vy = velocity_operators(self.Hs, X, Y)  # host CSC copies
        self.n = self.Hs.n
        self.MH, self.MX, self.MY = ctx.matrix(self.Hs), ctx.matrix(self.vx), ctx.matrix(self.vy)

    def close(self):
        for m in (self.MH, self.MX, self.MY):
            m.close()

    def tune(self, R=4, steps=16):
        """Autotune the context's launch shape for this Hamiltonian's
        recursion steps at R vectors (tbk_autotune); the tuning stays
        installed on the context.  Returns (tuning, ms per step)."""
        K = self.ctx.random_block(self.n, R, seed=1)
        try:
            return self.ctx.autotune(self.MH, K, steps=steps)
        finally:
            K.close()

    def zeta(self, gamma, e, volume):
        return kpm.kubo_zeta(self.ctx, gamma, e, volume)[0]

    def solve(self, N, R=4, seed=1, precision=kpm.GAMMA_FP64, fermi_energies=None, kets=None, area=None):
        """Returns dict(gamma, a, b, sigma, sigma_grid, zeta_e, zeta,
        fermi_energies, stats, timings); fermi_energies in the units of H
        (None skips sigma).  With area (the sample's area in lattice length
        units) also sigma_xy_e2h and sigma_xy_e2h_grid: sigma_xy in e^2/h,
        integrated from the band bottom."""
        t0 = time.perf_counter()
        g = spectral.lorentz(N, self.lam)
        K = self.ctx.block(kets) if kets is not None else self.ctx.random_block(self.n, R, seed=seed)
        try:
            gamma, st = kpm.kubo_gamma(self.ctx, self.MH, self.MX, self.MY, K, N, precision=precision, g=g)
        finally:
            K.close()
        t1 = time.perf_counter()
        out = dict(gamma=gamma, a=self.a, b=self.b, stats=st)
        if fermi_energies is not None:
            ef = (np.asarray(fermi_energies, dtype=np.float64) - self.b) / self.a
            z_nodes = self.zeta(gamma, spectral.kubo_nodes(N), self.volume)
            sigma, e, zr, integral = spectral.kubo_sigma(gamma, ef, self.volume, zeta=lambda *_: z_nodes)
            out.update(fermi_energies=np.asarray(fermi_energies), sigma=sigma, zeta_e=e, zeta=zr,
                       sigma_grid=integral)
            if area is not None:
                r_cols = kets.shape[1] if kets is not None else R
                sp, _, grid = spectral.kubo_sigma_physical(gamma, ef, self.n, r_cols, area, self.volume,
                                                           zeta=lambda *_: z_nodes)
                out.update(sigma_xy_e2h=sp, sigma_xy_e2h_grid=grid)
        t2 = time.perf_counter()
        out["timings"] = {"gamma_s": t1 - t0, "sigma_s": t2 - t1, "total_s": t2 - t0}
        return out


def conductivity_kubo_bastin(ctx, H, X, Y, N, E_min, E_max, volume=1.0, R=4, lam=6.0, seed=1,
                             precision=kpm.GAMMA_FP64, fermi_energies=None, kets=None, area=None):
    """One-shot form of KuboSolver.solve; returns the same dict."""
    s = KuboSolver(ctx, H, X, Y, E_min, E_max, volume, lam)
    try:
        return s.solve(N, R, seed, precision, fermi_energies, kets, area)
    finally:
        s.close()
