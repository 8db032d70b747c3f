"""Spectral front end: kernels, alpha_n, DoS and the DCT route, product
(tight_binding_amd.spectral) against the oracle and against closed forms."""
import math

import numpy as np
import pytest

import oracle
from oracle import spectral_oracle as SO
from tests.helpers import build_pair, cheb_T
from tight_binding_amd import spectral as S


@pytest.mark.parametrize("N", [16, 512, 2048])
def test_kernels_match_oracle(N):
    assert np.allclose(S.jackson(N), oracle.kernel_jackson(N), rtol=0, atol=1e-15)
    assert np.allclose(S.lorentz(N, 4.0), oracle.kernel_lorentz(N, 4.0), rtol=0, atol=1e-15)
    assert np.allclose(S.fejer(N), oracle.kernel_fejer(N), rtol=0, atol=0)


@pytest.mark.parametrize("z", [0.3, -0.77, 0.3 + 0.05j, -0.6 + 0.1j])
def test_alpha_expansion_reproduces_resolvent(z):
    """sum alpha_n(z) T_n(x) -> 1/(z - x) (the defining identity), for z on
    the spectral interval with a positive imaginary part -- the domain the
    reference evaluates it on (lattice.h:618-624)."""
    N = 4000
    a = S.expn_coef_alpha_seq(z, N)
    ref = np.array(SO.expn_coef_alpha_seq(z, 64))
    assert np.allclose(a[:64], ref, rtol=1e-12, atol=1e-13)
    if abs(np.imag(z)) > 0 or abs(z) > 1:
        for x in (-0.9, -0.2, 0.0, 0.6):
            s = np.sum(a * np.cos(np.arange(N) * math.acos(x)))
            assert abs(s - 1 / (z - x)) < 1e-6


def test_dos_product_matches_oracle_and_dct_route():
    p, _ = build_pair("square", (20, 20), [("A", 1, 1.0)], W=1.0, seed=3)
    from oracle import lattice_oracle as O

    Hs = O.rescale(p, 4.6, 0.0)
    R = 4
    kets = O.random_phase_kets(Hs.n, R, 8)
    N = 256
    mu = oracle.cheb_braket(Hs, N, kets, kets)
    g = S.lorentz(N, 4.0)
    e = S.linspace(-0.98, 0.98, 200)
    assert np.allclose(S.density_of_states(mu, g, e, R), SO.density_of_states(mu, g, e, R), rtol=1e-10, atol=1e-12)
    e1, f1 = S.dct_reconstruction(mu, g)
    e2, f2 = SO.dct_reconstruction(mu, g)
    assert np.allclose(e1, e2, rtol=0, atol=1e-15)
    assert np.allclose(f1, f2, rtol=1e-10, atol=1e-10 * np.max(np.abs(f2)))
    # DCT route at its own nodes equals the direct sum
    assert np.allclose(f1, S.green_imag(mu, g, e1) * -1.0, rtol=1e-9, atol=1e-9 * np.max(np.abs(f1)))


def test_dos_of_exact_moments_integrates_and_matches_eigenvalue_histogram():
    """Full-basis moments of a small lattice: the Jackson-damped KPM DoS is
    normalised and its integral over each window matches the eigenvalue
    count within the kernel's resolution."""
    p, _ = build_pair("honeycomb", (10, 10), [("A", 1, -2.7)], W=0.8, seed=5)
    D = p.to_dense()
    a = 1.02 * np.max(np.abs(np.linalg.eigvalsh(D)))
    evals = np.linalg.eigvalsh(D) / a
    n = D.shape[0]
    N = 512
    mu = cheb_T(N, evals).sum(axis=1)  # exact full-basis moments
    g = S.jackson(N)
    e = S.linspace(-0.999, 0.999, 8000, endpoint=True)
    dos = S.density_of_states(mu, g, e, n)  # per state
    de = e[1] - e[0]
    assert abs(np.sum(dos) * de - 1.0) < 2e-3
    assert dos.min() > -1e-8  # Jackson kernel keeps the DoS positive
    for lo, hi in [(-0.9, -0.4), (-0.4, 0.0), (0.1, 0.6)]:
        m = (e >= lo) & (e < hi)
        count = np.sum((evals >= lo) & (evals < hi)) / n
        assert abs(np.sum(dos[m]) * de - count) < 0.02


# ---------------------------------------------------------- Kubo-Bastin sigma
@pytest.mark.parametrize("N", [48, 64, 512])
def test_sigma_integration_oracle_closed_forms(N):
    """The oracle's zero-anchored trapezoid (snippets.h:420-449) integrates
    degree-1 integrands exactly: the Chebyshev nodes are symmetric about 0,
    so the split point of the bracketing interval is e = 0 itself."""
    eps = SO.kubo_nodes(N)
    E, I1 = SO.kubo_sigma_reference(eps, np.ones(N))
    assert np.allclose(I1, E, rtol=0, atol=1e-14)
    E, Ie = SO.kubo_sigma_reference(eps, eps * 3.0 - 0.5)
    assert np.allclose(Ie, 1.5 * E * E - 0.5 * E, rtol=0, atol=1e-14)
    cut = N // 24
    assert E.size == N - 2 * cut and np.all(np.diff(E) > 0)


def test_sigma_integration_product_matches_oracle():
    rng = np.random.default_rng(11)
    for N in (48, 100, 513):
        eps = S.kubo_nodes(N)
        assert np.array_equal(eps, SO.kubo_nodes(N))
        z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        E_ref, I_ref = SO.kubo_sigma_reference(eps, z)
        cut = N // 24
        E = eps[cut:N - cut][::-1]
        I = S.integrate_from_zero(E, np.real(z)[cut:N - cut][::-1])
        assert np.array_equal(E, E_ref) and np.array_equal(I, I_ref)


def test_host_zeta_separable_form_matches_direct_double_sum():
    rng = np.random.default_rng(5)
    N = 24
    G = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    e = np.array([-0.93, -0.4, 0.0, 0.17, 0.8])
    assert np.allclose(S.kubo_zeta(G, e, 2.5), SO.kubo_zeta(G, e, 2.5), rtol=1e-12, atol=1e-10)


def test_kubo_sigma_matches_oracle_pipeline():
    rng = np.random.default_rng(8)
    N = 48
    G = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    ef = np.linspace(-0.7, 0.7, 21)
    sig, E, zr, I = S.kubo_sigma(G, ef, 3.0)
    eps = SO.kubo_nodes(N)
    E_ref, I_ref = SO.kubo_sigma_reference(eps, SO.kubo_zeta(G, eps, 3.0))
    assert np.allclose(E, E_ref, rtol=0, atol=0)
    assert np.allclose(I, I_ref, rtol=1e-11, atol=1e-11 * np.abs(I_ref).max())
    assert np.allclose(sig, np.interp(ef, E_ref, I_ref), rtol=1e-11, atol=1e-11 * np.abs(I_ref).max())
