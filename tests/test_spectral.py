"""Spectral front end: kernels, alpha_n, DoS and the DCT route, product
(tight_binding_amd.spectral) against the oracle and against closed forms."""
import math

import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
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


# ---------------------------------- the reference's live zeta route (FFT)
@pytest.mark.parametrize("N", [1, 2, 4, 8, 64, 256])
def test_oracle_fft_steps_equal_textbook_transforms(N):
    """dct_iii_complex (numerical.h:229-249) is the unnormalised DCT-III,
    C_0 + 2 sum_n C_n cos(n pi (k + 1/2)/N), i.e. scipy's type 3; and
    fft_half_conj (numerical.h:276-283) is sum_j C_j e^{i pi j k/N}, k < N."""
    import scipy.fft

    rng = np.random.default_rng(N)
    C = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
    d = SO.dct_iii_complex(C)
    assert np.max(np.abs(d - scipy.fft.dct(C, type=3, axis=0))) <= 1e-13 * max(1.0, np.abs(d).max())
    f = SO.fft_half_conj(C)
    k = np.arange(N)
    want = np.exp(1j * math.pi * np.outer(k, k) / N) @ C
    assert np.max(np.abs(f - want)) <= 1e-12 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("N", [32, 256, 512])
def test_live_zeta_route_equals_direct_double_sum(N):
    """The route the reference runs (snippets.h:351-405: DCT-III, two
    half-conjugate FFTs, four trace sums per node) against its commented
    direct double sum (snippets.h:337-349, 406-415), on random weighted
    stacks; and the product's separable form against the same double sum.
    At N = 512 every eighth node is compared (the double sum costs N^2 per
    node)."""
    rng = np.random.default_rng(7 + N)
    K = 6
    aL = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    aR = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    eps, z_fft = SO.kubo_zeta_fft(aL, aR, 1.7)
    assert np.array_equal(eps, S.kubo_nodes(N))
    G = SO.kubo_gamma(aL[:, None, :], aR[:, None, :])
    sel = slice(None) if N <= 256 else slice(None, None, 8)
    z_dir = SO.kubo_zeta(G, eps[sel], 1.7)
    scale = np.abs(z_dir).max()
    assert np.max(np.abs(z_fft[sel] - z_dir)) <= 1e-10 * scale
    assert np.max(np.abs(S.kubo_zeta(G, eps[sel], 1.7) - z_dir)) <= 1e-12 * scale


# ------------------------------ sigma_xy in e^2/h against the exact TKNN sum
def _magnetic_torus(L, q):
    """L x L periodic square lattice, t = 1, Landau gauge A = (0, B x, 0) with
    B = 2 pi/q (L a multiple of q, so the phases close around the torus):
    H dense, and the bond displacements (dx, dy) of every pair (minimal
    image), sites numbered x L + y."""
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (L, L))
    lat.gauge_field = O.landau_gauge(2 * math.pi / q)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), True).to_dense()
    i = np.arange(L * L)
    x, y = i // L, i % L
    dx = (x[None, :] - x[:, None] + L // 2) % L - L // 2
    dy = (y[None, :] - y[:, None] + L // 2) % L - L // 2
    return H, dx, dy


def test_sigma_xy_physical_matches_tknn_on_magnetic_torus():
    """Full-basis pin of the zeta route and of its conversion to e^2/h.

    24 x 24 torus at flux 1/8 per plaquette.  Exact: the TKNN sum
    sigma = -(4 pi/Omega) sum_{a occ, b empty} Im[(v_x)_ab (v_y)_ba]/(E_a - E_b)^2
    with v = i[H, r] from the bond displacements, which gives the integer
    Landau-level plateaus 1 and 2 in the two lowest gaps.  KPM: Gamma_nm of
    the reference's definition (v = [r, H~], Lorentz lambda = 6) from the
    exact eigenbasis with kets = the identity (R = n), the product's zeta and
    kubo_sigma_physical.  The Lorentz kernel's broadening leaves an O(1/N)
    error: about 1% at N = 1024, shrinking with N; a wrong sign or factor
    would miss by 100% or more."""
    L, q = 24, 8
    H, dx, dy = _magnetic_torus(L, q)
    n = H.shape[0]
    off = ~np.eye(n, dtype=bool)
    vx, vy = 1j * dx * H * off, 1j * dy * H * off  # i[H, r] over the bonds
    E, U = np.linalg.eigh(H)
    area = float(L * L)
    Vx, Vy = U.conj().T @ vx @ U, U.conj().T @ vy @ U

    def tknn(mu):
        occ = E < mu
        sel = occ[:, None] & ~occ[None, :]
        D = E[:, None] - E[None, :]
        return -4 * math.pi / area * np.sum(np.imag(Vx * Vy.T)[sel] / D[sel] ** 2)

    gaps = sorted(np.argsort(np.diff(E))[::-1][:4])[:2]  # the two lowest of the four widest gaps
    mus = [0.5 * (E[g] + E[g + 1]) for g in gaps]
    exact = [tknn(mu) for mu in mus]
    assert abs(exact[0] - 1.0) < 1e-3 and abs(exact[1] - 2.0) < 1e-3, exact

    lo, hi = E[0] - 0.3, E[-1] + 0.3
    a, b = (hi - lo) / 2, (hi + lo) / 2
    et = (E - b) / a
    # the reference's velocity [r, H~] = i v / a in the eigenbasis
    X = (1j * Vx / a) * (1j * Vy / a).T
    errs = []
    for N in (512, 1024):
        g = S.lorentz(N, 6.0)
        GT = g[:, None] * np.cos(np.arange(N)[:, None] * np.arccos(et)[None, :])
        gamma = GT @ X @ GT.T  # g_n g_m Tr[T_n v_x T_m v_y]
        sig, _, _ = S.kubo_sigma_physical(gamma, (np.array(mus) - b) / a, n, n, area)
        errs.append(np.abs(sig - exact) / np.abs(exact))
    assert np.all(errs[1] < 0.015), errs
    assert np.all(errs[1] < errs[0]), errs
