"""Device reconstruction of the expansion from its moments
(tbk_cheb_reconstruct): the DoS and GreenFunctionImag's direct sum against
the oracle's alpha_n sum (lattice.h:617-623, functions.h:27-30 restated),
fft() against the oracle's DCT-III (FFTW REDFT01, lattice.h:625-640), and a
closed form: with mu_n = T_n(x0) and g = 1 the sum collapses to two Dirichlet
kernels."""
import math

import numpy as np
import pytest

from scipy.fft import dct

import oracle
from oracle import spectral_oracle as SO
from tests.test_gpu_batch import model_problem
from tight_binding_amd import _native as Nn
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu


def lattice_moments(ctx, N=512, R=8):
    H, a = model_problem("honeycomb", 24, 1.0, 3)
    M = ctx.matrix(H)
    K = ctx.random_block(H.n, R, seed=5)
    mu, _, _ = ctx.moments(M, K, N, a=a)
    M.close()
    K.close()
    return mu


def test_dos_and_direct_sum_match_oracle(gpu_ctx):
    mu = lattice_moments(gpu_ctx)
    g = oracle.kernel_jackson(mu.size)
    e = np.linspace(-0.999, 0.999, 401)
    got = kpm.density_of_states(gpu_ctx, mu, g, e, 8)
    want = SO.density_of_states(mu, g, e, 8)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # scale -1 is GreenFunctionImag::operator(); no context: the process-wide one
    f, ms = kpm.reconstruct(None, mu, g, e[::7], -1.0)
    fw = SO.green_imag(mu, g, e[::7])
    assert np.max(np.abs(f - fw)) <= 1e-12 * np.max(np.abs(fw)) and ms >= 0


def redft01_over_sin(x, M):
    """REDFT01 (scipy's DCT-III) of x zero-padded to M, over sin(theta_k),
    theta_k = (k + 1/2) pi / M formed directly."""
    y = dct(np.concatenate([x, np.zeros(M - x.size)]), type=3)
    return y / np.sin((np.arange(M) + 0.5) * np.pi / M)


@pytest.mark.parametrize("N", [1, 2, 7, 512, 2048])
def test_fft_nodes_match_redft01(gpu_ctx, N):
    """The device uses sin(theta_k); the reference's weight 1/sqrt(1 - E^2) at
    E = cos(theta_k) loses digits to cancellation at the edge nodes (1 - E^2
    ~ 6e-7 at theta_0 = pi/4096: ~1e-10 relative), so the oracle's form is
    held to 1e-9 and the exact weight to 1e-12."""
    rng = np.random.default_rng(N)
    mu = rng.standard_normal(N)
    g = oracle.kernel_lorentz(N, 4.0)
    E, r = kpm.green_fft(gpu_ctx, mu, g)
    Ew, rw = SO.dct_reconstruction(mu, g)
    assert np.max(np.abs(E - Ew)) <= 1e-15
    assert np.max(np.abs(r - rw)) <= 1e-9 * np.max(np.abs(rw))
    exact = redft01_over_sin(g * mu, N)
    assert np.max(np.abs(r - exact)) <= 1e-12 * np.max(np.abs(exact))
    # fft(4): four times as many nodes, moments zero-padded
    E4, r4 = kpm.green_fft(gpu_ctx, mu, g, n=4)
    assert np.max(np.abs(E4 - np.cos((np.arange(4 * N) + 0.5) * np.pi / (4 * N)))) <= 1e-15
    exact4 = redft01_over_sin(g * mu, 4 * N)
    assert np.max(np.abs(r4 - exact4)) <= 1e-12 * np.max(np.abs(exact4))


def test_closed_form_dirichlet(gpu_ctx):
    """mu_n = cos(n theta0), g = 1:
    sum_n c_n cos(n theta0) cos(n theta) = D(theta - theta0) + D(theta + theta0),
    D(phi) = sin((N - 1/2) phi) / (2 sin(phi / 2))."""
    N, x0 = 300, 0.3141
    th0 = math.acos(x0)
    mu = np.cos(np.arange(N) * th0)
    e = np.cos(np.linspace(0.01, math.pi - 0.01, 257))
    th = np.arccos(e)

    def D(p):
        return np.sin((N - 0.5) * p) / (2 * np.sin(p / 2))

    want = (D(th - th0) + D(th + th0)) / np.sin(th)
    got, _ = kpm.reconstruct(gpu_ctx, mu, None, e, 1.0)
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))


def test_edges_and_errors(gpu_ctx):
    mu = np.ones(16)
    v, _ = kpm.reconstruct(gpu_ctx, mu, None, [-1.0, 1.0, 0.0], 1.0)
    assert not np.isfinite(v[0]) and not np.isfinite(v[1]) and np.isfinite(v[2])
    out, _ = kpm.reconstruct(gpu_ctx, mu, None, np.zeros(0), 1.0)
    assert out.size == 0
    with pytest.raises(Nn.NativeError) as e:
        kpm.reconstruct(gpu_ctx, np.zeros(0), None, [0.0], 1.0)
    assert e.value.code == Nn.TBK_E_INVALID
