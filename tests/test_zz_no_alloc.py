"""tbk_moments_noalloc (the moment recursion without device allocation)
against tbk_moments on the same context: the same moment bits and the same
last Chebyshev block, on calls in a row with a smaller second input, in fp64
with and without the doubling identities, with the fused spectral shift, and
in fp32; sizes beyond the fixed scratch array and bad work buffers are
rejected before any launch."""
import ctypes as C
import math

import numpy as np
import pytest

from tests.helpers import build_pair
from tight_binding_amd import _native as N
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu


def dev_ptr(block):
    p = C.c_void_p()
    N.check_kpm(N.kpm().tbk_block_device_ptr(block._h, C.byref(p)))
    return p


def noalloc(ctx, M, kets, n, work0, work1, mu_blk, a, b, plain=False):
    flags = N.TBK_FLAG_PLAIN if plain else 0
    rc = N.kpm().tbk_moments_noalloc(ctx._h, M._h, float(a), float(b), None, kets._h, int(n), flags, dev_ptr(work0),
                                     dev_ptr(work1), dev_ptr(mu_blk), None)
    return rc


def read_mu(mu_blk, n):
    """n doubles stored in a (ceil(n/2) x 1) complex128 block."""
    c = np.ascontiguousarray(mu_blk.download()[:, 0])
    return c.view(np.float64)[:n]


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def run_case(ctx, M, n, R, N_, a, b, plain, seed, precision=kpm.FP64):
    K = ctx.random_block(n, R, seed=seed, precision=precision)
    mu_ref, st, last = ctx.moments(M, K, N_, a=a, b=b, plain=plain, want_last=True)
    W0, W1 = ctx.zeros_block(n, R, precision), ctx.zeros_block(n, R, precision)
    MU = ctx.zeros_block((N_ + 1) // 2, 1)
    print("noalloc case: n=%d R=%d N=%d plain=%d precision=%d" % (n, R, N_, plain, precision), flush=True)
    N.check_kpm(noalloc(ctx, M, K, N_, W0, W1, MU, a, b, plain))
    N.check_kpm(N.kpm().tbk_context_synchronize(ctx._h))
    mu = read_mu(MU, N_)
    assert np.array_equal(bits(mu), bits(mu_ref)), np.max(np.abs(mu - mu_ref))
    steps = st["steps"]
    T = (W1 if steps % 2 else W0).download()
    assert np.array_equal(bits(T), bits(last))
    for blk in (K, W0, W1, MU):
        blk.close()
    return steps


def test_noalloc_moments_equal_tbk_moments(gpu_ctx):
    # 4608-site honeycomb, 16 vectors, 300 moments with doubling (150 steps:
    # four full rings of partials and one of 22), fused shift b != 0
    p, _ = build_pair("honeycomb", (48, 48), [("A", 1, -2.7)], W=0.5, seed=3)
    M1 = gpu_ctx.matrix(p)
    assert run_case(gpu_ctx, M1, p.n, 16, 300, 8.4, 0.1, False, 5) == 150
    # then a smaller input: 400-site complex Landau square, 8 vectors, one
    # moment per step over 76 steps
    q, _ = build_pair("square", (20, 20), [("A", 1, 1.0)], builder="open", is_complex=True, W=1.0, seed=4,
                      gauge="landau", B=2 * math.pi / 20)
    M2 = gpu_ctx.matrix(q)
    assert run_case(gpu_ctx, M2, q.n, 8, 77, 4.8, 0.0, True, 6) == 76
    # fp32 recursion on the same matrix (its fp32 values formed by tbk_moments)
    assert run_case(gpu_ctx, M2, q.n, 8, 64, 4.8, 0.0, False, 7, precision=kpm.FP32) == 32
    M1.close()
    M2.close()


def test_noalloc_rejects_before_launch(gpu_ctx):
    p, _ = build_pair("square", (8, 8), [("A", 1, 1.0)], W=0.5, seed=2)
    M = gpu_ctx.matrix(p)
    K = gpu_ctx.random_block(p.n, 4, seed=1)
    W0, W1 = gpu_ctx.zeros_block(p.n, 4), gpu_ctx.zeros_block(p.n, 4)
    MU = gpu_ctx.zeros_block(4, 1)
    # more steps than the scratch array holds sums for
    assert noalloc(gpu_ctx, M, K, 70000, W0, W1, MU, 4.6, 0.0) == N.TBK_E_UNSUPPORTED
    assert noalloc(gpu_ctx, M, K, 40000, W0, W1, MU, 4.6, 0.0, plain=True) == N.TBK_E_UNSUPPORTED
    assert "32768" in N.kpm().tbk_last_error().decode()
    # aliased work buffers
    assert noalloc(gpu_ctx, M, K, 8, W0, W0, MU, 4.6, 0.0) == N.TBK_E_INVALID
    assert noalloc(gpu_ctx, M, K, 8, K, W1, MU, 4.6, 0.0) == N.TBK_E_INVALID
    # fp32 kets on a matrix whose fp32 values were never formed
    K32 = gpu_ctx.random_block(p.n, 4, seed=1, precision=kpm.FP32)
    assert noalloc(gpu_ctx, M, K32, 8, W0, W1, MU, 4.6, 0.0) == N.TBK_E_INVALID
    assert "fp32" in N.kpm().tbk_last_error().decode()
    for o in (K, K32, W0, W1, MU, M):
        o.close()
