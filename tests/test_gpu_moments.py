"""GPU parity of the HIP moment engine against the CPU oracle (the
restatement of numerical.h:363-381) and against exact diagonalisation.

Bar: fp64 moments within 1e-10 relative (normwise |dmu| <= 1e-10 max|mu|,
and elementwise where |mu_n| > 1e-6 mu_0); fp64 recursion vectors T_n
bit-identical to the oracle's; fp32 recursion within the stated bound."""
import math

import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from tests.helpers import build_pair, cheb_T, same_bits
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu

TOL = 1e-10


def assert_moments_close(mu, ref, tol=TOL):
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(mu - ref)) <= tol * scale, np.max(np.abs(mu - ref)) / scale
    big = np.abs(ref) > 1e-6 * abs(ref[0])
    if big.any():
        rel = np.max(np.abs(mu[big] - ref[big]) / np.abs(ref[big]))
        assert rel <= tol * 1e3, rel  # elementwise: small moments carry the normwise error


def c1_problem(L=100, R=16, seed=5):
    """100x100 periodic square, t = 1, V ~ U[-0.5, 0.5], a = 4.6."""
    p, _ = build_pair("square", (L, L), [("A", 1, 1.0)], W=1.0, seed=seed)
    Hs = O.rescale(p, 4.6, 0.0)
    kets = O.random_phase_kets(Hs.n, R, seed + 1)
    return p, Hs, kets


def test_dropin_c1_doubling_and_general_match_oracle():
    _, Hs, kets = c1_problem()
    N = 512
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    mu_d = kpm.cheb_first_kind_seq_braket(Hs, N, kets, kets)  # same object -> doubling
    mu_g = kpm.cheb_first_kind_seq_braket(Hs, N, kets.copy(), kets)  # distinct bras -> one moment/step
    assert abs(ref[0] - 16.0) < 1e-9  # sequential sum of 160k terms
    assert_moments_close(mu_g, ref)
    assert_moments_close(mu_d, ref)


@pytest.mark.parametrize("cplx", [False, True])
def test_dropin_non_hermitian_same_kets_takes_direct_path(cplx):
    """bras_T is kets but M is not Hermitian: the doubling identities do not
    hold, so the drop-in must form every moment directly, as
    numerical.h:363-381 does."""
    if cplx:
        p, _ = build_pair("square", (16, 16), [("A", 1, 1.0)], builder="open", is_complex=True, W=1.0, seed=4,
                          gauge="landau", B=2 * math.pi / 16)
    else:
        p, _ = build_pair("square", (16, 16), [("A", 1, 1.0)], W=1.0, seed=4)
    Hs = O.rescale(p, 4.8, 0.0)
    cols = np.repeat(np.arange(Hs.n), np.diff(Hs.colptr))
    v = Hs.values.copy()
    v[Hs.rowidx < cols] *= 0.9  # upper triangle scaled: M != M^H; moments stay bounded (max |mu| = R)
    M = O.CSC(Hs.n, Hs.colptr, Hs.rowidx, v)
    kets = O.random_phase_kets(M.n, 8, 21)
    N = 96
    ref = oracle.cheb_braket(M, N, kets, kets)
    mu = kpm.cheb_first_kind_seq_braket(M, N, kets, kets)
    assert_moments_close(mu, ref)
    # the doubling identities would give a different answer on this M
    herm = kpm.cheb_first_kind_seq_braket(Hs, N, kets, kets)
    assert np.max(np.abs(herm - mu)) > 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("R", [8, 64])
def test_recursion_vectors_bit_exact(gpu_ctx, cplx, R):
    if cplx:
        p, _ = build_pair("square", (24, 24), [("A", 1, 1.0)], builder="open", is_complex=True, W=2.0, seed=3,
                          gauge="landau", B=2 * math.pi / 64)
    else:
        p, _ = build_pair("honeycomb", (20, 20), [("A", 1, -2.7)], W=0.5, seed=3)
    Hs = O.rescale(p, 8.6, 0.0)
    kets = O.random_phase_kets(Hs.n, R, 11)
    N = 33
    M = gpu_ctx.matrix(Hs)
    K = gpu_ctx.block(kets)
    # plain: last block is T_{N-1}
    mu, st, last = gpu_ctx.moments(M, K, N, plain=True, want_last=True)
    ref, T = oracle.cheb_braket(Hs, N, kets, kets, last_index=N - 1)
    assert st["doubling"] == 0 and st["steps"] == N - 1
    assert np.array_equal(last, T), "T_{N-1} differs from the oracle"
    assert_moments_close(mu, ref)
    # doubling: last block is T_{ceil(N/2)}
    mu2, st2, last2 = gpu_ctx.moments(M, K, N, want_last=True)
    steps = (N + 1) // 2
    _, T2 = oracle.cheb_braket(Hs, N, kets, kets, last_index=steps)
    assert st2["doubling"] == 1 and st2["steps"] == steps
    assert np.array_equal(last2, T2)
    assert_moments_close(mu2, ref)


def test_distinct_bras_complex_landau(gpu_ctx):
    p, _ = build_pair("square", (32, 32), [("A", 1, 1.0)], builder="open", is_complex=True, W=2.0, seed=8,
                      gauge="landau", B=2 * math.pi / 64)
    Hs = O.rescale(p, 5.2, 0.0)
    kets = O.random_phase_kets(Hs.n, 32, 1)
    bras = O.random_phase_kets(Hs.n, 32, 2)
    N = 128
    ref = oracle.cheb_braket(Hs, N, bras, kets)
    mu, st, _ = gpu_ctx.moments(gpu_ctx.matrix(Hs), gpu_ctx.block(kets), N, bras=gpu_ctx.block(bras))
    assert st["doubling"] == 0
    assert np.max(np.abs(mu - ref)) <= TOL * np.max(np.abs(ref))


@pytest.mark.parametrize("cplx", [False, True])
def test_fused_rescale_matches_prescaled_oracle(gpu_ctx, cplx):
    p, _ = build_pair("honeycomb", (16, 16), [("A", 1, -2.7)], is_complex=cplx, W=1.0, seed=4)
    a, b = 9.1, 0.35
    Hs = O.rescale(p, a, b)
    kets = O.random_phase_kets(p.n, 16, 3)
    N = 256
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    M = gpu_ctx.matrix(p)  # raw H; (H - b)/a fused in the step kernel
    for plain in (False, True):
        mu, _, _ = gpu_ctx.moments(M, gpu_ctx.block(kets), N, a=a, b=b, plain=plain)
        assert_moments_close(mu, ref)


def test_no_explicit_diagonal_with_shift(gpu_ctx):
    """bare builder: no diagonal entries; the shift term loads x_i itself."""
    p, _ = build_pair("square", (20, 20), [("A", 1, 1.0)], builder="bare")
    kets = O.random_phase_kets(p.n, 4, 7)
    a, b = 4.5, -0.2
    ref = oracle.cheb_braket(O.rescale(p, a, b), 64, kets, kets)
    mu, _, _ = gpu_ctx.moments(gpu_ctx.matrix(p), gpu_ctx.block(kets), 64, a=a, b=b)
    assert_moments_close(mu, ref)


def test_fp32_recursion_error_vs_fp64(gpu_ctx):
    _, Hs, kets = c1_problem()
    N = 512
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    M = gpu_ctx.matrix(Hs)
    for plain in (False, True):
        mu, st, _ = gpu_ctx.moments(M, gpu_ctx.block(kets, precision=kpm.FP32), N, plain=plain)
        err = np.max(np.abs(mu - ref)) / np.max(np.abs(ref))
        print("fp32 recursion, plain=%s: normwise error vs fp64 oracle %.3e" % (plain, err))
        assert err < 2e-5


@pytest.mark.parametrize("R", [1, 2, 3, 5, 16, 100, 128, 256])
def test_block_widths(gpu_ctx, R):
    p, _ = build_pair("kagome", (7, 6), [("A", 1, 1.0), ("B", 1, 1.0), ("C", 1, 1.0)], W=1.0)
    Hs = O.rescale(p, 6.5, 0.0)
    kets = O.random_phase_kets(Hs.n, R, R)
    N = 40
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    M, K = gpu_ctx.matrix(Hs), gpu_ctx.block(kets)
    mu, st, last = gpu_ctx.moments(M, K, N, plain=True, want_last=True)
    _, T = oracle.cheb_braket(Hs, N, kets, kets, last_index=N - 1)
    assert np.array_equal(last, T)
    assert_moments_close(mu, ref)
    mu2, _, _ = gpu_ctx.moments(M, K, N)
    assert_moments_close(mu2, ref)


@pytest.mark.parametrize("model,nn,R,prec", [
    ("honeycomb", [("A", 1, -2.7)], 64, "fp64"),
    ("square", [("A", 1, 1.0)], 16, "fp64"),
    ("fcc", [("A", 1, 1.0)], 32, "fp64"),
    ("cubic", [("A", 1, 1.0), ("A", 2, 0.3), ("A", 3, 0.1)], 8, "fp64"),  # 27 entries/row: row kernel
    ("honeycomb", [("A", 1, -2.7)], 64, "fp32"),
    ("square", [("A", 1, 1.0)], 128, "fp64"),  # two values per lane: row kernel, and tiled at a 128-column slab
])
def test_tiled_and_row_kernels_and_slabs_agree(gpu_ctx, model, nn, R, prec):
    """Every kernel family and slab width produces the oracle's T_n bit for
    bit (fp64) and the same moments; kernel_kind / slab_cols report the
    choice."""
    span = (10, 10, 10) if model in ("fcc", "cubic") else (40, 36)
    p, _ = build_pair(model, span, nn, W=1.0, seed=2)
    gersh = np.bincount(p.rowidx, weights=np.abs(p.values), minlength=p.n).max()
    Hs = O.rescale(p, 1.01 * gersh, 0.0)
    kets = O.random_phase_kets(Hs.n, R, 5)
    N = 24
    ref, T = oracle.cheb_braket(Hs, N, kets, kets, last_index=N - 1)
    precision = kpm.FP64 if prec == "fp64" else kpm.FP32
    M, K = gpu_ctx.matrix(Hs), gpu_ctx.block(kets, precision=precision)
    seen = set()
    try:
        for slab, chunk_rows, period, nt in [(0, 0, 0, 0), (8, 64, 3, -1), (16, 0, -1, 0), (32, 256, 0, -1),
                                             (64, 32, 0, 0), (64, 64, 7, 1), (128, 0, 0, 0)]:
            for row in (False, True):
                gpu_ctx.set_tuning(slab_cols=slab, chunk_rows=chunk_rows, chunk_period=period, stream_nt=nt,
                                   use_tiled=0 if row else 1)
                mu, st, last = gpu_ctx.moments(M, K, N, plain=True, want_last=True, row_kernel=row)
                seen.add((st["kernel_kind"], st["slab_cols"]))
                if prec == "fp64":
                    assert np.array_equal(last, T), (slab, row, st)
                    assert_moments_close(mu, ref)
                else:
                    assert np.max(np.abs(mu - ref)) < 1e-5 * np.max(np.abs(ref))
                mu2, _, _ = gpu_ctx.moments(M, K, N, row_kernel=row)
                assert np.max(np.abs(mu2 - mu)) <= 1e-12 * np.max(np.abs(mu)) * (1 if prec == "fp64" else 1e6)
    finally:
        gpu_ctx.set_tuning()
    if model == "cubic":
        assert all(k == 0 for k, _ in seen)
    else:
        assert any(k > 0 for k, _ in seen) and any(k == 0 for k, _ in seen)


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_smallest_moment_counts(gpu_ctx, N):
    _, Hs, kets = c1_problem(L=10, R=4)
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    for plain in (False, True):
        mu, st, _ = gpu_ctx.moments(gpu_ctx.matrix(Hs), gpu_ctx.block(kets), N, plain=plain)
        assert mu.shape == (N,)
        assert_moments_close(mu, ref)


def test_tiny_matrices_and_empty_rows(gpu_ctx):
    # 1 x 1
    one = O.CSC(1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([0.3]))
    k = np.array([[1.0 + 0j]])
    mu, _, _ = gpu_ctx.moments(gpu_ctx.matrix(one), gpu_ctx.block(k), 6)
    assert np.allclose(mu, np.cos(np.arange(6) * math.acos(0.3)), atol=1e-15)
    # vacancies: trailing empty rows and columns (lattice.h:300)
    p, _ = build_pair("honeycomb", (6, 6), [("A", 1, 1.0)], builder="open", W=1.0,
                      removed=("A", [(0, 0), (2, 3)]))
    Hs = O.rescale(p, 3.8, 0.0)
    kets = O.random_phase_kets(Hs.n, 8, 2)
    ref = oracle.cheb_braket(Hs, 50, kets, kets)
    mu, _, _ = gpu_ctx.moments(gpu_ctx.matrix(Hs), gpu_ctx.block(kets), 50)
    assert_moments_close(mu, ref)


def test_full_basis_moments_equal_exact_diagonalisation(gpu_ctx):
    p, _ = build_pair("lieb", (5, 5), [("A", 1, 1.0), ("B", 1, 1.0), ("C", 1, 1.0)], W=1.0, seed=9)
    D = p.to_dense()
    a = 1.01 * np.max(np.abs(np.linalg.eigvalsh(D)))
    Hs = O.rescale(p, a, 0.0)
    n = Hs.n
    N = 128
    e = np.linalg.eigvalsh(Hs.to_dense())
    exact = cheb_T(N, e).sum(axis=1)
    kets = np.eye(n, dtype=np.complex128)
    mu, _, _ = gpu_ctx.moments(gpu_ctx.matrix(Hs), gpu_ctx.block(kets), N)
    assert np.max(np.abs(mu - exact)) <= 1e-11 * n


def test_closed_form_bloch_moments_periodic_honeycomb(gpu_ctx):
    """Full-basis moments of a clean periodic honeycomb are sums of T_n over
    the Bloch band energies -- no diagonalisation involved."""
    L0, L1, t = 8, 6, -2.7
    p, _ = build_pair("honeycomb", (L0, L1), [("A", 1, t)])
    a = 8.2
    Hs = O.rescale(p, a, 0.0)
    from tests.helpers import honeycomb_bloch_eigs

    exact = cheb_T(96, honeycomb_bloch_eigs(L0, L1, t) / a).sum(axis=1)
    kets = np.eye(Hs.n, dtype=np.complex128)
    mu, _, _ = gpu_ctx.moments(gpu_ctx.matrix(Hs), gpu_ctx.block(kets), 96)
    assert np.max(np.abs(mu - exact)) <= 1e-11 * Hs.n


def test_deterministic_repeat(gpu_ctx):
    _, Hs, kets = c1_problem(L=60, R=16)
    M, K = gpu_ctx.matrix(Hs), gpu_ctx.block(kets)
    a, _, _ = gpu_ctx.moments(M, K, 200)
    b, _, _ = gpu_ctx.moments(M, K, 200)
    assert same_bits(a, b)


def test_device_random_phase_block(gpu_ctx):
    blk = gpu_ctx.random_block(1000, 5, seed=42)
    k = blk.download()
    assert k.shape == (1000, 5)
    assert np.allclose(np.abs(k), 1 / math.sqrt(1000), rtol=1e-14)
    host = O.random_phase_kets(1000, 5, 42)
    assert np.max(np.abs(k - host)) < 1e-15  # same generator, libm vs device sin/cos


def test_errors_are_reported(gpu_ctx):
    from tight_binding_amd import _native as N

    _, Hs, kets = c1_problem(L=10, R=4)
    M, K = gpu_ctx.matrix(Hs), gpu_ctx.block(kets)
    with pytest.raises(N.NativeError) as e:
        gpu_ctx.moments(M, K, 1)
    assert e.value.code == N.TBK_E_INVALID
    with pytest.raises(N.NativeError):
        gpu_ctx.moments(M, gpu_ctx.block(kets[:50]), 8)
    with pytest.raises(N.NativeError) as e:
        gpu_ctx.block(np.zeros((10, 300), complex))
    assert e.value.code == N.TBK_E_UNSUPPORTED


def test_large_honeycomb_size_independent_properties(gpu_ctx):
    """512x512-cell honeycomb (524,288 sites), R = 64: mu_0 = R, |mu_n| <= mu_0,
    doubling and plain agree, and both leave the same T_k bit for bit."""
    p, _ = build_pair("honeycomb", (512, 512), [("A", 1, -2.7)], W=0.5, seed=21)
    M = gpu_ctx.matrix(p)
    K = gpu_ctx.random_block(p.n, 64, seed=77)
    a = 8.35 * 1.01
    N = 64
    mu_d, st_d, last_d = gpu_ctx.moments(M, K, N, a=a, want_last=True)
    mu_p, st_p, last_p = gpu_ctx.moments(M, K, st_d["steps"] + 1, a=a, plain=True, want_last=True)
    assert np.array_equal(last_d, last_p)
    mu_p2, _, _ = gpu_ctx.moments(M, K, N, a=a, plain=True)
    assert abs(mu_d[0] - 64.0) < 1e-10
    assert np.all(np.abs(mu_d) <= mu_d[0] * (1 + 1e-12))
    assert_moments_close(mu_d, mu_p2)


def test_autotune_installs_a_valid_tuning(gpu_ctx):
    p, _ = build_pair("honeycomb", (96, 96), [("A", 1, -2.7)], W=0.5, seed=1)
    M = gpu_ctx.matrix(p)
    K = gpu_ctx.random_block(p.n, 64, seed=3)
    a = 8.5
    ref, _, _ = gpu_ctx.moments(M, K, 64, a=a)
    try:
        best, ms = gpu_ctx.autotune(M, K, a=a, steps=4)
        assert ms > 0 and gpu_ctx.get_tuning() == best
        mu, st, _ = gpu_ctx.moments(M, K, 64, a=a)
        assert_moments_close(mu, ref, tol=1e-12)
    finally:
        gpu_ctx.set_tuning()


def test_bandwidth_probe(gpu_ctx):
    """The attainable-bandwidth references run and give physical figures:
    positive, below the 8 TB/s HBM peak (the buffers exceed the caches), and
    reading alone at least as fast as the copy."""
    from tight_binding_amd import _native as Nn

    b = gpu_ctx.bandwidth_probe(1 << 30, repeats=2)
    assert b["bytes"] == float(1 << 30)
    for k in ("copy_GBps", "stream_2r1w_GBps", "read_GBps"):
        assert 100.0 < b[k] < 8000.0, (k, b)
    assert b["read_GBps"] >= 0.9 * b["copy_GBps"]
    with pytest.raises(Nn.NativeError):
        gpu_ctx.bandwidth_probe(1024)


@pytest.mark.parametrize("prec,R,slab,cplx", [(kpm.FP64, 64, 16, False), (kpm.FP64, 64, 32, True),
                                              (kpm.FP32, 32, 8, False), (kpm.FP64, 128, 64, False)])
def test_row_slabs_give_the_same_vectors(gpu_ctx, prec, R, slab, cplx):
    """row_slab_cols sweeps the rows a few vector columns at a time: every
    T_n element is formed by the same per-row sum, so the last block is the
    same bits and the moments agree to rounding (partials grouped by sweep)."""
    if cplx:
        p, _ = build_pair("square", (40, 40), [("A", 1, 1.0)], builder="open", is_complex=True, W=1.0, seed=6,
                          gauge="landau", B=2 * math.pi / 20)
    else:
        p, _ = build_pair("honeycomb", (40, 40), [("A", 1, -2.7)], W=0.5, seed=6)
    M = gpu_ctx.matrix(p)
    K = gpu_ctx.random_block(p.n, R, seed=12, precision=prec)
    saved = gpu_ctx.get_tuning()
    try:
        for plain in (False, True):
            mu0, st0, last0 = gpu_ctx.moments(M, K, 65, a=8.6, b=0.1, plain=plain, want_last=True)
            gpu_ctx.set_tuning(**dict(saved, row_slab_cols=slab))
            mu1, st1, last1 = gpu_ctx.moments(M, K, 65, a=8.6, b=0.1, plain=plain, want_last=True)
            gpu_ctx.set_tuning(**saved)
            assert st1["slab_cols"] == slab and st0["slab_cols"] == st0["padded_R"]
            assert np.array_equal(np.ascontiguousarray(last0).view(np.uint64), np.ascontiguousarray(last1).view(np.uint64))
            assert np.max(np.abs(mu1 - mu0)) <= 1e-12 * np.max(np.abs(mu0))
    finally:
        gpu_ctx.set_tuning(**saved)
    M.close()
    K.close()


def test_autotune_cache_and_plain_mode(gpu_ctx, tmp_path):
    """A second autotune of the same problem shape is served from the cache
    (same tuning, no trials); refresh runs the trials again; the
    one-moment-per-step mode has its own entry; a saved cache read back into
    an empty one gives the same entries."""
    import time

    p, _ = build_pair("honeycomb", (64, 64), [("A", 1, -2.7)], W=0.5, seed=1)
    M = gpu_ctx.matrix(p)
    K = gpu_ctx.random_block(p.n, 16, seed=3)
    saved = gpu_ctx.get_tuning()
    try:
        kpm.tuning_cache_clear()
        t0 = time.perf_counter()
        best, ms = gpu_ctx.autotune(M, K, a=8.5, steps=4)
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        best2, ms2 = gpu_ctx.autotune(M, K, a=8.5, steps=4)
        second = time.perf_counter() - t0
        assert best2 == best and ms2 == ms and second < first
        best_p, ms_p = gpu_ctx.autotune(M, K, a=8.5, steps=4, plain=True)
        assert ms_p > 0
        _, ms3 = gpu_ctx.autotune(M, K, a=8.5, steps=4, refresh=True)
        assert ms3 > 0
        path = tmp_path / "cache.txt"
        kpm.tuning_cache_save(str(path))
        kpm.tuning_cache_clear()
        assert kpm.tuning_cache_load(str(path)) == 2
        best4, _ = gpu_ctx.autotune(M, K, a=8.5, steps=4, plain=True)
        assert best4 == best_p
        # the installed tuning gives the moments of the default shape
        ref, _, _ = gpu_ctx.moments(M, K, 64, a=8.5, plain=True)
        gpu_ctx.set_tuning(**saved)
        mu, _, _ = gpu_ctx.moments(M, K, 64, a=8.5, plain=True)
        assert_moments_close(mu, ref, tol=1e-12)
    finally:
        gpu_ctx.set_tuning(**saved)
        kpm.tuning_cache_clear()
    M.close()
    K.close()
