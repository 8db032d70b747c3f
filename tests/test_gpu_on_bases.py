"""Stored Chebyshev blocks (Cheb_first_kind_seq_on_bases, numerical.h:334-344)
on the device: tbk_cheb_seq_on_bases (every block resident) and the
host-buffer drop-in tbk_cheb_first_kind_seq_on_bases, against the oracle's
stored blocks (oracle_cheb_on_bases, pinned to exact diagonalisation in
test_oracle_pins.py).

Bar: fp64 blocks bit-identical to the oracle's at a = 1, b = 0 (M already
rescaled, as the reference calls it); with the rescale fused (a, b != 1, 0)
within 1e-13 of the oracle on the rescaled matrix (the fused form rounds
(M x - b x)/a differently); fp32 blocks T_k within 4 (k + 1)^2 2^-24 of the
oracle's relative to the block's largest entry (rounding errors of step j
reach T_k through U_{k-j-1}, of norm at most k - j)."""
import math

import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from tests.helpers import build_pair
from tight_binding_amd import _native as Nn
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu


def problem(cplx):
    if cplx:
        p, _ = build_pair("square", (20, 18), [("A", 1, 1.0)], builder="open", is_complex=True, W=1.5, seed=6,
                          gauge="landau", B=2 * math.pi / 40)
    else:
        p, _ = build_pair("honeycomb", (14, 12), [("A", 1, -2.7)], W=0.5, seed=6)
    return p


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("R", [3, 16])
def test_resident_blocks_bit_exact(gpu_ctx, cplx, R):
    Hs = O.rescale(problem(cplx), 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, R, 21)
    n = 11
    want = oracle.cheb_on_bases(Hs, n, kets)
    M = gpu_ctx.matrix(Hs)
    V = gpu_ctx.block(kets)
    blocks, ms = gpu_ctx.cheb_on_bases(M, V, n)
    assert ms >= 0.0
    for k, b in enumerate(blocks):
        assert np.array_equal(b.download(), want[k]), (cplx, R, k)
    # out[0] may be v itself: no copy, same later blocks
    again, _ = gpu_ctx.cheb_on_bases(M, V, n, out=[V] + blocks[1:])
    for k, b in enumerate(again):
        assert np.array_equal(b.download(), want[k]), ("in place", k)
    for x in blocks + [M, V]:
        x.close()


def test_blocks_reproduce_moments_and_product_identity(gpu_ctx):
    """Re<T_0|T_k> is the moment mu_k of tbk_moments; 2 Re<T_m|T_k> =
    mu_{m+k} + mu_{|m-k|} (T_m T_k = (T_{m+k} + T_{|m-k|})/2)."""
    p = problem(True)
    M = gpu_ctx.matrix(p)
    V = gpu_ctx.random_block(p.n, 8, seed=4)
    n, a, b = 24, 5.3, 0.2
    blocks, _ = gpu_ctx.cheb_on_bases(M, V, n, a=a, b=b)
    T = [x.download() for x in blocks]
    mu, _, _ = gpu_ctx.moments(M, V, 2 * n, a=a, b=b, plain=True)
    for k in range(n):
        assert abs(np.real(np.vdot(T[0], T[k])) - mu[k]) <= 1e-12 * mu[0]
    for m in range(n):
        for k in range(0, n, 5):
            assert abs(2 * np.real(np.vdot(T[m], T[k])) - mu[m + k] - mu[abs(m - k)]) <= 1e-11 * mu[0]


def test_fused_rescale_matches_oracle(gpu_ctx):
    p = problem(True)
    a, b = 5.3, 0.2
    Hs = O.rescale(p, a, b)
    kets = O.random_phase_kets(p.n, 5, 8)
    want = oracle.cheb_on_bases(Hs, 16, kets)
    blocks, _ = gpu_ctx.cheb_on_bases(gpu_ctx.matrix(p), gpu_ctx.block(kets), 16, a=a, b=b)
    for k, x in enumerate(blocks):
        assert np.max(np.abs(x.download() - want[k])) <= 1e-13 * max(1.0, np.max(np.abs(want[k]))), k


def test_fp32_blocks(gpu_ctx):
    Hs = O.rescale(problem(False), 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, 8, 3)
    n = 32
    want = oracle.cheb_on_bases(Hs, n, kets)
    M = gpu_ctx.matrix(Hs)
    blocks, _ = gpu_ctx.cheb_on_bases(M, gpu_ctx.block(kets, kpm.FP32), n)
    for k, x in enumerate(blocks):
        err = np.max(np.abs(x.download() - want[k])) / np.max(np.abs(want[k]))
        assert err <= 4 * (k + 1) ** 2 * 2.0 ** -24, (k, err)


@pytest.mark.parametrize("cplx", [False, True])
def test_host_dropin_bit_exact(cplx):
    Hs = O.rescale(problem(cplx), 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, 4, 13)
    n = 9
    got = kpm.cheb_first_kind_seq_on_bases(Hs, n, kets)
    want = oracle.cheb_on_bases(Hs, n, kets)
    assert got.shape == want.shape == (n, Hs.n, 4)
    assert np.array_equal(got, want)
    one = kpm.cheb_first_kind_seq_on_bases(Hs, 1, kets)
    assert one.shape == (1, Hs.n, 4) and np.array_equal(one[0], kets)
    col = kpm.cheb_first_kind_seq_on_bases(Hs, 3, kets[:, 2])  # a single vector
    assert np.array_equal(col[:, :, 0], want[:3, :, 2])


def test_errors(gpu_ctx):
    Hs = O.rescale(problem(False), 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, 4, 1)
    with pytest.raises(Nn.NativeError) as e:
        kpm.cheb_first_kind_seq_on_bases(Hs, 0, kets)
    assert e.value.code == Nn.TBK_E_INVALID
    M = gpu_ctx.matrix(Hs)
    V = gpu_ctx.block(kets)
    X = [gpu_ctx.zeros_block(Hs.n, 4) for _ in range(3)]
    cases = [
        [X[0], X[1], X[1]],                            # repeated block
        [X[0], V, X[2]],                               # v past out[0]
        [X[0], X[1], gpu_ctx.zeros_block(Hs.n, 5)],    # another R
        [X[0], X[1], gpu_ctx.zeros_block(Hs.n, 4, kpm.FP32)],  # another precision
    ]
    for out in cases:
        with pytest.raises(Nn.NativeError) as e:
            gpu_ctx.cheb_on_bases(M, V, 3, out=out)
        assert e.value.code == Nn.TBK_E_INVALID
    for a in (0.0, float("nan")):
        with pytest.raises(Nn.NativeError) as e:
            gpu_ctx.cheb_on_bases(M, V, 3, a=a, out=X)
        assert e.value.code == Nn.TBK_E_INVALID
    with pytest.raises(Nn.NativeError):
        gpu_ctx.cheb_on_bases(M, gpu_ctx.zeros_block(Hs.n + 1, 4), 2)


@pytest.mark.parametrize("cplx", [False, True])
def test_single_steps_are_the_stored_blocks(gpu_ctx, cplx):
    """tbk_cheb_step (the device side of Cheb_first_kind_seq_view::next,
    numerical.h:346-361): y = H~ x, then y = 2 H~ x - z, in place over z as
    the view rotates its blocks -- bit for bit the oracle's T_k."""
    Hs = O.rescale(problem(cplx), 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, 5, 17)
    want = oracle.cheb_on_bases(Hs, 8, kets)
    M = gpu_ctx.matrix(Hs)
    v0, v1 = gpu_ctx.block(kets), gpu_ctx.zeros_block(Hs.n, 5)
    gpu_ctx.cheb_step(M, v0, v1)
    for k in range(8):
        assert np.array_equal(v0.download(), want[k]), k
        gpu_ctx.cheb_step(M, v1, v0, z=v0)  # T_{k+2} over T_k
        v0, v1 = v1, v0
    # separate output, fused rescale
    a, b = 5.3, 0.2
    Hr = O.rescale(problem(cplx), a, b)
    w = oracle.cheb_on_bases(Hr, 3, kets)
    P = gpu_ctx.matrix(problem(cplx))
    x0, x1, x2 = gpu_ctx.block(kets), gpu_ctx.zeros_block(Hs.n, 5), gpu_ctx.zeros_block(Hs.n, 5)
    gpu_ctx.cheb_step(P, x0, x1, a=a, b=b)
    gpu_ctx.cheb_step(P, x1, x2, z=x0, a=a, b=b)
    for k, x in ((1, x1), (2, x2)):
        assert np.max(np.abs(x.download() - w[k])) <= 1e-13 * np.max(np.abs(w[k]))
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.cheb_step(M, v0, v0)
    assert e.value.code == Nn.TBK_E_INVALID
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.cheb_step(M, v0, gpu_ctx.zeros_block(Hs.n, 4))
    assert e.value.code == Nn.TBK_E_INVALID


@pytest.mark.parametrize("cplx", [False, True])
def test_expansion_is_the_ordered_sum_of_the_blocks(gpu_ctx, cplx):
    """tbk_cheb_expand: out = sum_k c_k T_k(H~) v with complex c_k, each term
    (c.re t.re - c.im t.im, c.re t.im + c.im t.re) added in order of k --
    bit for bit the same sum over the oracle's T_k; real coefficients scale
    each part."""
    Hs = O.rescale(problem(cplx), 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, 6, 29)
    n = 40
    T = oracle.cheb_on_bases(Hs, n, kets)
    rng = np.random.default_rng(3)
    M = gpu_ctx.matrix(Hs)
    V = gpu_ctx.block(kets)
    for c in (rng.standard_normal(n) + 1j * rng.standard_normal(n), rng.standard_normal(n) + 0j):
        wr = np.zeros(T[0].shape)
        wi = np.zeros(T[0].shape)
        for k in range(n):
            tr, ti = T[k].real, T[k].imag
            if c[k].imag == 0.0:
                pr, pi = c[k].real * tr, c[k].real * ti
            else:
                pr = c[k].real * tr - c[k].imag * ti
                pi = c[k].real * ti + c[k].imag * tr
            wr, wi = (pr, pi) if k == 0 else (wr + pr, wi + pi)
        out, ms = gpu_ctx.cheb_expand(M, V, c)
        assert ms >= 0.0
        assert np.array_equal(out.download(), wr + 1j * wi)
        out.close()
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.cheb_expand(M, V, [1.0, float("nan")])
    assert e.value.code == Nn.TBK_E_INVALID
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.cheb_expand(M, V, [1.0], out=V)
    assert e.value.code == Nn.TBK_E_INVALID
