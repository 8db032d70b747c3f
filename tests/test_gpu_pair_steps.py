"""Two recursion steps per launch (tbk_tuning.pair_steps: cheb_pair_kernel
and its prologue) against one step per launch and against the oracle's
direct recursion (numerical.h:363-381).

Bar: the fp64 and fp32 vectors of the last step bit-identical to the
one-step path (the rows are computed by the same per-row code in the same
order); fp64 moments within the 1e-10 bar of test_gpu_moments against the
oracle, and within 1e-13 of the one-step path (only the order in which the
per-tile partial sums are added differs)."""
import math

import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from tests.helpers import build_pair, same_bits
from tests.test_gpu_moments import assert_moments_close
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu


def honeycomb(L, seed, builder="periodic"):
    p, _ = build_pair("honeycomb", (L, L), [("A", 1, -2.7)], builder=builder, W=0.5, seed=seed)
    return p, 8.35 * 1.01


def landau_square(L, seed):
    p, _ = build_pair("square", (L, L), [("A", 1, 1.0)], builder="open", is_complex=True, W=1.0, seed=seed,
                      gauge="landau", B=2 * math.pi / 32)
    return p, 4.6


def run_both(ctx, p, a, kets, N, plain=False, tile_rows=0, precision=kpm.FP64):
    """(mu, stats, last) one step per launch, then two."""
    saved = ctx.get_tuning()
    M = ctx.matrix(p)
    K = ctx.block(kets, precision=precision)
    try:
        ctx.set_tuning(**dict(saved, pair_steps=0))
        one = ctx.moments(M, K, N, a=a, plain=plain, want_last=True)
        ctx.set_tuning(**dict(saved, pair_steps=1, pair_tile_rows=tile_rows))
        two = ctx.moments(M, K, N, a=a, plain=plain, want_last=True)
    finally:
        ctx.set_tuning(**saved)
        M.close()
        K.close()
    return one, two


@pytest.mark.parametrize("tile_rows", [16, 64, 256])
@pytest.mark.parametrize("plain", [False, True])
def test_pair_steps_match_one_step_and_oracle(gpu_ctx, plain, tile_rows):
    """131,072-site periodic honeycomb (the benchmark's lattice at 1/16 of
    its sites), R = 8."""
    p, a = honeycomb(256, 41)
    Hs = O.rescale(O.CSC(p.n, p.colptr, p.rowidx, p.values), a, 0.0)
    kets = O.random_phase_kets(p.n, 8, 13)
    N = 96 if not plain else 49
    print("honeycomb 256x256 R=8 N=%d plain=%s tile_rows=%d" % (N, plain, tile_rows), flush=True)
    (mu1, st1, last1), (mu2, st2, last2) = run_both(gpu_ctx, p, a, kets, N, plain=plain, tile_rows=tile_rows)
    assert st1["pair_launches"] == 0
    assert st2["pair_launches"] == (st2["steps"] - 2) // 2 and st2["pair_launches"] > 0
    assert st2["steps"] == st1["steps"]
    assert same_bits(last1, last2)
    assert np.max(np.abs(mu2 - mu1)) <= 1e-13 * np.max(np.abs(mu1))
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    assert_moments_close(mu2, ref)


def test_pair_steps_complex_open_lattice_odd_steps(gpu_ctx):
    """Complex H (Landau gauge, open square 192 x 192), an odd step count:
    the last step runs alone after the two-step launches."""
    p, a = landau_square(192, 5)
    Hs = O.rescale(O.CSC(p.n, p.colptr, p.rowidx, p.values), a, 0.0)
    kets = O.random_phase_kets(p.n, 16, 3)
    N = 70  # 35 steps: first, one, 16 pairs, last
    print("landau square 192x192 R=16 N=%d" % N, flush=True)
    (mu1, st1, last1), (mu2, st2, last2) = run_both(gpu_ctx, p, a, kets, N, tile_rows=32)
    assert st2["steps"] == 35 and st2["pair_launches"] == 16
    assert same_bits(last1, last2)
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    assert_moments_close(mu2, ref)
    assert np.max(np.abs(mu2 - mu1)) <= 1e-13 * np.max(np.abs(mu1))


def test_pair_steps_fp32_bit_identical_to_one_step(gpu_ctx):
    p, a = honeycomb(256, 43)
    kets = O.random_phase_kets(p.n, 16, 17)
    print("honeycomb 256x256 R=16 fp32", flush=True)
    (mu1, st1, last1), (mu2, st2, last2) = run_both(gpu_ctx, p, a, kets, 64, precision=kpm.FP32)
    assert st2["pair_launches"] > 0
    assert same_bits(last1, last2)
    assert np.max(np.abs(mu2 - mu1)) <= 1e-12 * np.max(np.abs(mu1))


def test_pair_steps_fall_back_where_the_band_is_too_large(gpu_ctx):
    """A 24 x 24 lattice has too few rows per slab: the launches stay one
    step each and the moments are those of the one-step path."""
    p, a = honeycomb(24, 7)
    kets = O.random_phase_kets(p.n, 8, 1)
    (mu1, _, last1), (mu2, st2, last2) = run_both(gpu_ctx, p, a, kets, 40)
    assert st2["pair_launches"] == 0
    assert same_bits(last1, last2) and same_bits(mu1, mu2)
