"""Kubo-Bastin Gamma_nm on the GPU against the oracle (stored-block
recursion numerical.h:334-344 + the double-sum definition snippets.h:337-342)
and against exact diagonalisation; split-bf16 x3 error against fp64."""
import math

import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from oracle import spectral_oracle as SO
from tests.helpers import build_pair, cheb_T
from tight_binding_amd import kpm
from tight_binding_amd import lattice as L
from tight_binding_amd import spectral as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,K", [(37, 1000), (64, 4096), (130, 777), (1, 5), (256, 3000), (300, 77), (512, 2049),
                                 (383, 20001)])
def test_stack_contraction_matches_numpy(gpu_ctx, N, K):
    rng = np.random.default_rng(N * 7 + K)
    A = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    B = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    ref = np.conj(A) @ B.T
    g64, _ = kpm.stack_contract(gpu_ctx, A, B, kpm.GAMMA_FP64)
    err64 = np.max(np.abs(g64 - ref)) / np.max(np.abs(ref))
    assert err64 < 1e-13, err64
    gbf, _ = kpm.stack_contract(gpu_ctx, A, B, kpm.GAMMA_BF16X3)
    errbf = np.max(np.abs(gbf - ref)) / np.max(np.abs(ref))
    print("contraction N=%d K=%d: fp64 %.2e, bf16x3 %.2e" % (N, K, err64, errbf))
    assert errbf < 1e-4, errbf


def landau_problem(L0, B, W, seed, a):
    p, _ = build_pair("square", (L0, L0), [("A", 1, 1.0)], builder="open", is_complex=True, W=W, seed=seed,
                      gauge="landau", B=B)
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (L0, L0))
    X, Y = lat.op_position(0), lat.op_position(1)
    Hs = L.rescale(p, a, 0.0)
    vx, vy = L.commutator_diag(X, Hs), L.commutator_diag(Y, Hs)
    # the velocity operators agree bit for bit with the oracle's restatement
    ovx = O.commutator_with_diagonal(X, O.rescale(O.CSC(p.n, p.colptr, p.rowidx, p.values), a, 0.0))
    assert np.array_equal(vx.values, ovx.values) and np.array_equal(vx.rowidx, ovx.rowidx)
    return Hs, vx, vy


def oracle_gamma(Hs, vx, vy, kets, N, g):
    TL = oracle.cheb_on_bases(Hs, N, kets)
    w = oracle.cheb_on_bases(vy, 2, kets)[1]
    TR = oracle.cheb_on_bases(Hs, N, w)
    aR = np.stack([oracle.cheb_on_bases(vx, 2, TR[m])[1] for m in range(N)])
    aL = TL * g[:, None, None]
    aR = aR * g[:, None, None]
    return SO.kubo_gamma(aL, aR)


def test_kubo_gamma_matches_oracle(gpu_ctx):
    Hs, vx, vy = landau_problem(10, 2 * math.pi / 16, 1.0, 4, 5.2)
    kets = O.random_phase_kets(Hs.n, 4, 3)
    N = 32
    g = S.lorentz(N, 6.0)
    ref = oracle_gamma(Hs, vx, vy, kets, N, g)
    got, st = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                             gpu_ctx.block(kets), N, g=g)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    # fused rescaling of the raw H gives the same Gamma to rounding
    zeta_e = np.linspace(-0.8, 0.8, 9)
    assert np.allclose(S.kubo_zeta(got, zeta_e), SO.kubo_zeta(ref, zeta_e), rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_kubo_gamma_full_basis_equals_exact_diagonalisation(gpu_ctx):
    Hs, vx, vy = landau_problem(6, 2 * math.pi / 6, 0.5, 1, 5.0)
    n = Hs.n
    N = 24
    e, V = np.linalg.eigh(Hs.to_dense())
    VX = V.conj().T @ vx.to_dense() @ V
    VY = V.conj().T @ vy.to_dense() @ V
    T = cheb_T(N, e)  # (N, n)
    ref = np.einsum("na,ab,mb,ba->nm", T, VX, T, VY)
    got, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                            gpu_ctx.block(np.eye(n, dtype=np.complex128)), N)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_kubo_bf16x3_error_against_fp64(gpu_ctx):
    Hs, vx, vy = landau_problem(40, 2 * math.pi / 16, 1.0, 7, 5.2)
    N = 64
    g = S.lorentz(N, 6.0)
    M = [gpu_ctx.matrix(m) for m in (Hs, vx, vy)]
    K = gpu_ctx.random_block(Hs.n, 8, seed=5)
    g64, st64 = kpm.kubo_gamma(gpu_ctx, *M, K, N, g=g)
    gbf, stbf = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=kpm.GAMMA_BF16X3, g=g)
    gerr = np.max(np.abs(gbf - g64)) / np.max(np.abs(g64))
    ef = np.linspace(-0.6, 0.6, 41)
    s64 = S.kubo_sigma(g64, ef)[0]
    sbf = S.kubo_sigma(gbf, ef)[0]
    serr = np.max(np.abs(sbf - s64)) / np.max(np.abs(s64))
    print("bf16x3 vs fp64: Gamma %.2e, sigma %.2e" % (gerr, serr))
    assert gerr < 1e-4 and serr < 1e-4


@pytest.mark.parametrize("N", [1, 7, 48])
def test_device_zeta_matches_direct_double_sum(gpu_ctx, N):
    rng = np.random.default_rng(N)
    G = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    e = np.concatenate([SO.kubo_nodes(max(N, 2)), rng.uniform(-0.95, 0.95, 9)])
    z, _ = kpm.kubo_zeta(gpu_ctx, G, e, 1.7)
    ref = SO.kubo_zeta(G, e, 1.7)
    assert np.max(np.abs(z - ref)) <= 1e-11 * np.max(np.abs(ref)), np.max(np.abs(z - ref))


@pytest.mark.parametrize("N", [600, 1000, 2048])
def test_device_zeta_large_orders_match_host_form(gpu_ctx, N):
    """N = 600 and 1000 run two energies per workgroup, N = 2048 one."""
    rng = np.random.default_rng(N)
    G = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    e = np.concatenate([rng.uniform(-0.9, 0.9, 13), [0.0]])
    z, _ = kpm.kubo_zeta(gpu_ctx, G, e)
    ref = S.kubo_zeta(G, e)
    assert np.max(np.abs(z - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_device_zeta_errors(gpu_ctx):
    from tight_binding_amd import _native as Nn

    G = np.eye(3, dtype=np.complex128)
    with pytest.raises(Nn.NativeError) as e:
        kpm.kubo_zeta(gpu_ctx, G, [0.2, 1.0])
    assert e.value.code == Nn.TBK_E_INVALID
    with pytest.raises(Nn.NativeError) as e:
        kpm.kubo_zeta(gpu_ctx, np.zeros((3000, 3000), np.complex128), [0.1])
    assert e.value.code == Nn.TBK_E_UNSUPPORTED


def test_sigma_full_basis_equals_exact_diagonalisation(gpu_ctx):
    """sigma_xy from the device path (full-basis Gamma, device zeta, zero-anchored
    integral) equals sigma_xy built from dense exact diagonalisation with the
    same Lorentz damping and the oracle's direct double sum and integration."""
    Hs, vx, vy = landau_problem(6, 2 * math.pi / 5, 0.5, 2, 5.0)
    n = Hs.n
    N = 48
    g = S.lorentz(N, 6.0)
    e_all, V = np.linalg.eigh(Hs.to_dense())
    VX = V.conj().T @ vx.to_dense() @ V
    VY = V.conj().T @ vy.to_dense() @ V
    T = cheb_T(N, e_all)
    G_ed = np.einsum("na,ab,mb,ba->nm", T, VX, T, VY) * np.outer(g, g)
    eps = SO.kubo_nodes(N)
    E_ref, I_ref = SO.kubo_sigma_reference(eps, SO.kubo_zeta(G_ed, eps, float(n)))
    ef = np.linspace(-0.5, 0.5, 11)
    sig_ref = np.interp(ef, E_ref, I_ref)

    got, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                            gpu_ctx.block(np.eye(n, dtype=np.complex128)), N, g=g)
    sig, E, _, I = S.kubo_sigma(got, ef, float(n), zeta=lambda G, e, v: kpm.kubo_zeta(gpu_ctx, G, e, v)[0])
    scale = np.max(np.abs(I_ref))
    assert scale > 0
    assert np.max(np.abs(I - I_ref)) <= 1e-9 * scale
    assert np.max(np.abs(sig - sig_ref)) <= 1e-9 * scale


def test_tuned_solver_gives_the_same_gamma(gpu_ctx):
    """KuboSolver.tune changes only the launch shape of the stack recursion:
    every stored T_n is formed by the same per-row sums, so Gamma is the same
    bits before and after."""
    from tight_binding_amd import kubo

    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (24, 24))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * np.pi / 24)
    lat.set_disorder(1.0, 4)
    H = lat.get_Hamiltonian_sparse(True)
    saved = gpu_ctx.get_tuning()
    s = kubo.KuboSolver(gpu_ctx, H, lat.op_position(0), lat.op_position(1), -5.1, 5.1, volume=float(H.n))
    try:
        g0 = s.solve(64, R=8, seed=3)["gamma"]
        tuning, ms = s.tune(R=8, steps=4)
        assert ms > 0 and tuning["row_lane_bytes"] in (0, 8, 16, 32)
        g1 = s.solve(64, R=8, seed=3)["gamma"]
        assert np.array_equal(g0, g1)
    finally:
        s.close()
        gpu_ctx.set_tuning(**saved)


@pytest.mark.parametrize("N", [1, 2, 3, 40])
def test_fused_velocity_stacks_are_the_same_bits(gpu_ctx, N):
    """When v_x has H~'s pattern (v_x = [X, H~]) the right stack forms
    v_x T_{m-1} inside the step that makes T_m; when v_x is uploaded with a
    different pattern (an explicit zero dropped), the separate launches run.
    Both give the same Gamma bits, and both match the oracle's stacks."""
    import scipy.sparse as sp

    from tight_binding_amd import sparse as SP

    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (14, 11))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * np.pi / 14)
    lat.set_disorder(1.5, 6)
    H = lat.get_Hamiltonian_sparse(True)
    Hs = L.rescale(H, 5.0, 0.1)
    vx = L.commutator_diag(lat.op_position(0), Hs)
    vy = L.commutator_diag(lat.op_position(1), Hs)
    # the same operator without its stored zeros: another pattern
    csc = sp.csc_matrix((vx.values.copy(), vx.rowidx.copy(), vx.colptr.copy()), shape=(vx.n, vx.n))
    csc.eliminate_zeros()
    vx2 = SP.CSC(vx.n, csc.indptr.astype(np.int32), csc.indices.astype(np.int32), csc.data.astype(np.complex128))
    assert vx2.values.size < vx.values.size
    K = gpu_ctx.random_block(Hs.n, 4, seed=21)
    g1, st1 = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy), K, N)
    g2, st2 = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx2), gpu_ctx.matrix(vy), K, N)
    assert np.array_equal(g1, g2)
    kets = K.download()
    TL = oracle.cheb_on_bases(Hs, N, kets)
    w = oracle.cheb_on_bases(vy, 2, kets)[1]
    TR = oracle.cheb_on_bases(Hs, N, w)
    aR = np.stack([oracle.cheb_on_bases(vx, 2, TR[m])[1] for m in range(N)])
    ref = SO.kubo_gamma(TL, aR)
    assert np.max(np.abs(g1 - ref)) <= 1e-12 * np.max(np.abs(ref))
    K.close()


@pytest.mark.parametrize("R,lane_bytes", [(2, 16), (2, 0), (4, 32), (8, 16), (8, 0), (32, 16), (32, 32)])
def test_presplit_stacks_give_the_bf16x3_gamma_of_fp64_stacks(gpu_ctx, R, lane_bytes):
    """Split-BF16x3 solves with N >= 256 on complex operators write their
    stacks pre-split (the recursion stores the hi/lo bf16 words the
    contraction would form).  Gamma must be the same bits as the BF16x3
    contraction of the fp64 stacks -- here the oracle's, which the device's
    equal bit for bit at a = 1, b = 0 -- with v_x fused (both chains then run
    as one block of 2 Rp columns) or not, at one (16-byte lanes) or two
    complex values per lane (32-byte lanes, and the paired block's default)."""
    import scipy.sparse as sp

    from tight_binding_amd import sparse as SP

    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (12, 10))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * np.pi / 12)
    lat.set_disorder(1.0, 9)
    Hs = L.rescale(lat.get_Hamiltonian_sparse(True), 5.0, 0.0)
    vx = L.commutator_diag(lat.op_position(0), Hs)
    vy = L.commutator_diag(lat.op_position(1), Hs)
    N = 256
    saved = gpu_ctx.get_tuning()
    gpu_ctx.set_tuning(row_lane_bytes=lane_bytes)
    K = gpu_ctx.random_block(Hs.n, R, seed=17)
    kets = K.download()
    g_ps, st = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy), K, N,
                              precision=kpm.GAMMA_BF16X3)
    Rp = 1 << (R - 1).bit_length()
    assert st["stack_bytes"] == 2 * N * Hs.n * Rp * 8  # half-size entries
    TL = oracle.cheb_on_bases(Hs, N, kets)
    w = oracle.cheb_on_bases(vy, 2, kets)[1]
    TR = oracle.cheb_on_bases(Hs, N, w)
    aR = np.stack([oracle.cheb_on_bases(vx, 2, TR[m])[1] for m in range(N)])
    ref, _ = kpm.stack_contract(gpu_ctx, TL.reshape(N, -1), aR.reshape(N, -1), kpm.GAMMA_BF16X3)
    assert np.array_equal(g_ps, ref)
    g64, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy), K, N)
    assert np.max(np.abs(g_ps - g64)) <= 1e-4 * np.max(np.abs(g64))
    # v_x without its stored zeros: the unfused pre-split path, the same bits
    csc = sp.csc_matrix((vx.values.copy(), vx.rowidx.copy(), vx.colptr.copy()), shape=(vx.n, vx.n))
    csc.eliminate_zeros()
    vx2 = SP.CSC(vx.n, csc.indptr.astype(np.int32), csc.indices.astype(np.int32), csc.data.astype(np.complex128))
    g_ps2, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx2), gpu_ctx.matrix(vy), K, N,
                              precision=kpm.GAMMA_BF16X3)
    assert np.array_equal(g_ps2, ref)
    gpu_ctx.set_tuning(**saved)
    K.close()


def test_sigma_xy_e2h_device_path_matches_tknn_on_magnetic_torus(gpu_ctx):
    """sigma_xy in e^2/h through the device: Gamma from tbk_kubo_gamma with the
    full basis (kets = identity, R = n = 256), zeta on the device, and
    spectral.kubo_sigma_physical, against the exact TKNN sum of the same
    16 x 16 torus at flux 1/8 (integer plateaus 1 and 2 in the two lowest
    gaps).  The velocities are the reference's [r, H~] in bond form,
    (x_i - x_j) H~_ij with x_i - x_j the minimal-image bond displacement, so
    that they are defined on the torus.  Device Gamma against the eigenbasis
    Gamma to 1e-10; sigma within 3% at N = 512 (Lorentz broadening)."""
    from tight_binding_amd.sparse import CSC

    L0, q, N = 16, 8, 512
    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (L0, L0))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * math.pi / q)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), True)
    n = H.n
    Hd = H.to_dense()
    E, U = np.linalg.eigh(Hd)
    lo, hi = E[0] - 0.3, E[-1] + 0.3
    a, b = (hi - lo) / 2, (hi + lo) / 2
    Hs = L.rescale(H, a, b)
    cols = np.repeat(np.arange(n), np.diff(Hs.colptr))
    rows = Hs.rowidx
    xr, yr, xc, yc = rows // L0, rows % L0, cols // L0, cols % L0
    dx = (xc - xr + L0 // 2) % L0 - L0 // 2  # bond displacement row -> column
    dy = (yc - yr + L0 // 2) % L0 - L0 // 2
    vx = CSC(n, Hs.colptr, Hs.rowidx, (-dx) * Hs.values)  # (x_i - x_j) H~_ij
    vy = CSC(n, Hs.colptr, Hs.rowidx, (-dy) * Hs.values)
    g = S.lorentz(N, 6.0)
    got, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                            gpu_ctx.block(np.eye(n, dtype=np.complex128)), N, g=g)
    Vx, Vy = U.conj().T @ vx.to_dense() @ U, U.conj().T @ vy.to_dense() @ U
    et = (E - b) / a
    GT = g[:, None] * np.cos(np.arange(N)[:, None] * np.arccos(np.clip(et, -1, 1))[None, :])
    ref = GT @ (Vx * Vy.T) @ GT.T
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))

    area = float(L0 * L0)
    Px, Py = -1j * a * Vx, -1j * a * Vy  # physical v = i[H, r] = -i a [r, H~]

    def tknn(mu):
        occ = E < mu
        sel = occ[:, None] & ~occ[None, :]
        D = E[:, None] - E[None, :]
        return -4 * math.pi / area * np.sum(np.imag(Px * Py.T)[sel] / D[sel] ** 2)

    gaps = sorted(np.argsort(np.diff(E))[::-1][:4])[:2]
    mus = np.array([0.5 * (E[k] + E[k + 1]) for k in gaps])
    exact = np.array([tknn(mu) for mu in mus])
    assert np.allclose(exact, [1.0, 2.0], atol=0.02), exact
    sig, _, _ = S.kubo_sigma_physical(got, (mus - b) / a, n, n, area,
                                      zeta=lambda G, e, v: kpm.kubo_zeta(gpu_ctx, G, e, v)[0])
    print("torus 16x16 flux 1/8: TKNN", exact, "KPM N=512", sig)
    assert np.all(np.abs(sig - exact) <= 0.03 * np.abs(exact)), (sig, exact)


@pytest.mark.parametrize("prec,tol", [(kpm.GAMMA_FP64, 1e-12), (kpm.GAMMA_BF16X3, 1e-5)])
def test_kubo_gamma_in_vector_groups_under_a_stack_cap(gpu_ctx, prec, tol):
    """tbk_kubo_gamma with the stacks capped (kubo_stack_mib) below what the
    whole 12-vector block needs: the vectors run in groups (the last one
    partial) and Gamma equals the whole block's to rounding."""
    Hs, vx, vy = landau_problem(48, 2 * math.pi / 16, 1.0, 5, 5.2)
    kets = O.random_phase_kets(Hs.n, 12, 17)
    N = 96
    g = S.lorentz(N, 6.0)
    M = [gpu_ctx.matrix(x) for x in (Hs, vx, vy)]
    K = gpu_ctx.block(kets)
    whole, st = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=prec, g=g)
    assert st["vector_groups"] == 1
    saved = gpu_ctx.get_tuning()
    try:
        # room for about a third of the whole block's stacks
        gpu_ctx.set_tuning(**dict(saved, kubo_stack_mib=max(1, int(st["stack_bytes"] / 3 / 2**20))))
        gpu_ctx.release_work()
        grouped, stg = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=prec, g=g)
    finally:
        gpu_ctx.set_tuning(**saved)
        gpu_ctx.release_work()
    assert stg["vector_groups"] >= 3
    assert np.max(np.abs(grouped - whole)) <= tol * np.max(np.abs(whole))
    if prec == kpm.GAMMA_FP64:
        ref = oracle_gamma(Hs, vx, vy, kets, N, g)
        assert np.max(np.abs(grouped - ref)) <= 1e-10 * np.max(np.abs(ref))
    for x in M + [K]:
        x.close()


@pytest.mark.parametrize("prec,tol", [(kpm.GAMMA_FP64, 1e-12), (kpm.GAMMA_BF16X3, 1e-5)])
def test_kubo_gamma_in_row_ranges_when_one_vector_exceeds_the_cap(gpu_ctx, prec, tol):
    """tbk_kubo_gamma with a 1 MiB stack cap, below the 7 MB the stacks of one
    random vector need: every group is one vector and its stacks hold a range
    of rows of every entry at a time (chains rerun per range); Gamma equals
    the whole block's to rounding and the oracle's at 1e-10 (fp64)."""
    Hs, vx, vy = landau_problem(48, 2 * math.pi / 16, 1.0, 5, 5.2)
    kets = O.random_phase_kets(Hs.n, 3, 23)
    N = 96
    g = S.lorentz(N, 6.0)
    M = [gpu_ctx.matrix(x) for x in (Hs, vx, vy)]
    K = gpu_ctx.block(kets)
    whole, st = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=prec, g=g)
    assert st["row_ranges"] == 1
    saved = gpu_ctx.get_tuning()
    try:
        gpu_ctx.set_tuning(**dict(saved, kubo_stack_mib=1))
        gpu_ctx.release_work()
        ranged, str_ = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=prec, g=g)
    finally:
        gpu_ctx.set_tuning(**saved)
        gpu_ctx.release_work()
    assert str_["vector_groups"] == 3 and str_["row_ranges"] >= 8, str_
    assert np.max(np.abs(ranged - whole)) <= tol * np.max(np.abs(whole))
    if prec == kpm.GAMMA_FP64:
        ref = oracle_gamma(Hs, vx, vy, kets, N, g)
        assert np.max(np.abs(ranged - ref)) <= 1e-10 * np.max(np.abs(ref))
    for x in M + [K]:
        x.close()


def _f32_rows(M):
    """CSR of M padded to its longest row: column indices, fp32 real and
    imaginary parts (the device's rounding of the fp64 values) and a mask."""
    rp, ci, val = M.to_csr()
    n = M.n
    cnt = np.diff(rp)
    m = int(cnt.max())
    cols = np.zeros((n, m), dtype=np.int64)
    hr = np.zeros((n, m), dtype=np.float32)
    hi = np.zeros((n, m), dtype=np.float32)
    mask = np.zeros((n, m), dtype=bool)
    for j in range(m):
        rows = np.nonzero(cnt > j)[0]
        p = rp[rows] + j
        cols[rows, j] = ci[p]
        hr[rows, j] = np.real(val[p]).astype(np.float32)
        hi[rows, j] = np.imag(val[p]).astype(np.float32)
        mask[rows, j] = True
    return cols, hr, hi, mask


def _spmm_f32(rows, xr, xi):
    """sum over a row's entries in ascending column order of h * x, each
    complex product formed as the step kernel forms it (products rounded,
    then added), all in fp32."""
    cols, hr, hi, mask = rows
    ar = np.zeros_like(xr)
    ai = np.zeros_like(xi)
    for j in range(cols.shape[1]):
        gr, gi = xr[cols[:, j]], xi[cols[:, j]]
        h_r, h_i = hr[:, j, None], hi[:, j, None]
        re = h_r * gr - h_i * gi
        im = h_r * gi + h_i * gr
        mk = mask[:, j, None]
        ar = np.where(mk, ar + re, ar)
        ai = np.where(mk, ai + im, ai)
    return ar, ai


def _fp32_paired_stacks(Hs, vx, vy, kets, N):
    """The stacks of the fp32 paired recursion restated on the CPU in numpy
    fp32: X_0 = [kets | v_y kets] rounded to complex64 (v_y kets in fp64, the
    oracle's, equal to the device's), X_1 = H X_0, X_{k+1} = 2 H X_k - X_{k-1};
    L_0 = kets, L_n = left half of X_n, R_m = v_x (right half of X_m), as
    complex128 (exact) for the contraction."""
    n, R = kets.shape
    w = oracle.cheb_on_bases(vy, 2, kets)[1]
    X0 = np.concatenate([kets, w], axis=1).astype(np.complex64)
    rows_h, rows_x = _f32_rows(Hs), _f32_rows(vx)
    xr, xi = np.ascontiguousarray(X0.real), np.ascontiguousarray(X0.imag)
    zr = zi = None
    Ls = [kets.astype(np.complex128)]
    Rs = []
    two = np.float32(2.0)
    for k in range(N):
        ar, ai = _spmm_f32(rows_h, xr, xi)
        br, bi = _spmm_f32(rows_x, xr[:, R:], xi[:, R:])
        Rs.append((br + 1j * bi.astype(np.float64)).astype(np.complex128))
        if k == 0:
            yr, yi = ar, ai
        else:
            yr, yi = two * ar - zr, two * ai - zi
        if k + 1 < N:
            Ls.append((yr[:, :R].astype(np.float64) + 1j * yi[:, :R].astype(np.float64)))
        zr, zi, xr, xi = xr, xi, yr, yi
    return np.stack(Ls), np.stack(Rs)


@pytest.mark.parametrize("R,lane_bytes", [(2, 0), (4, 16), (8, 32), (16, 8), (32, 0)])
def test_fp32_recursion_stacks_match_an_fp32_restatement(gpu_ctx, R, lane_bytes):
    """TBK_GAMMA_BF16X3_FP32 builds the paired stacks with an fp32 recursion
    (complex64 vectors, fp32 H and v_x values) and splits the fp32 values
    into bf16 words.  Gamma must be the same bits as the split-bf16x3
    contraction of the same recursion restated in numpy fp32 (an fp32 value
    splits to the same words through the fp64 split), and within 1e-4 of the
    fp64 Gamma."""
    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (12, 10))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * np.pi / 12)
    lat.set_disorder(1.0, 9)
    Hs = L.rescale(lat.get_Hamiltonian_sparse(True), 5.0, 0.0)
    vx = L.commutator_diag(lat.op_position(0), Hs)
    vy = L.commutator_diag(lat.op_position(1), Hs)
    N = 256
    saved = gpu_ctx.get_tuning()
    gpu_ctx.set_tuning(row_lane_bytes=lane_bytes)
    try:
        K = gpu_ctx.random_block(Hs.n, R, seed=23)
        kets = K.download()
        M = [gpu_ctx.matrix(m) for m in (Hs, vx, vy)]
        g32, st = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=kpm.GAMMA_BF16X3_FP32)
        assert st["recursion_fp32"] == 1
        Ls, Rs = _fp32_paired_stacks(Hs, vx, vy, kets, N)
        ref, _ = kpm.stack_contract(gpu_ctx, Ls.reshape(N, -1), Rs.reshape(N, -1), kpm.GAMMA_BF16X3)
        assert np.array_equal(g32, ref)
        g64, _ = kpm.kubo_gamma(gpu_ctx, *M, K, N)
        err = np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))
        print("fp32 recursion + bf16x3 vs fp64: Gamma %.2e" % err)
        assert err <= 1e-4
        K.close()
    finally:
        gpu_ctx.set_tuning(**saved)


def test_fp32_recursion_sigma_error_against_fp64(gpu_ctx):
    """sigma_xy from the fp32-recursion stacks within 1e-4 of fp64 on a
    Landau lattice at N = 256; problems outside the paired path (one vector
    per chain) take the fp64 recursion and say so."""
    Hs, vx, vy = landau_problem(40, 2 * math.pi / 16, 1.0, 7, 5.2)
    N = 256
    g = S.lorentz(N, 6.0)
    M = [gpu_ctx.matrix(m) for m in (Hs, vx, vy)]
    K = gpu_ctx.random_block(Hs.n, 8, seed=5)
    g64, _ = kpm.kubo_gamma(gpu_ctx, *M, K, N, g=g)
    g32, st = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=kpm.GAMMA_BF16X3_FP32, g=g)
    assert st["recursion_fp32"] == 1
    ef = np.linspace(-0.6, 0.6, 41)
    s64 = S.kubo_sigma(g64, ef)[0]
    s32 = S.kubo_sigma(g32, ef)[0]
    gerr = np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))
    serr = np.max(np.abs(s32 - s64)) / np.max(np.abs(s64))
    print("fp32 recursion + bf16x3 vs fp64: Gamma %.2e, sigma %.2e" % (gerr, serr))
    assert gerr < 1e-4 and serr < 1e-4
    K.close()
    K1 = gpu_ctx.random_block(Hs.n, 1, seed=5)
    g1, st1 = kpm.kubo_gamma(gpu_ctx, *M, K1, N, precision=kpm.GAMMA_BF16X3_FP32, g=g)
    gb, _ = kpm.kubo_gamma(gpu_ctx, *M, K1, N, precision=kpm.GAMMA_BF16X3, g=g)
    assert st1["recursion_fp32"] == 0 and np.array_equal(g1, gb)
    K1.close()


@pytest.mark.parametrize("N,L0,R", [(300, 9, 3), (512, 12, 2), (256, 7, 5), (256, 2, 1), (256, 4, 3)])
def test_wide_contraction_tiles_give_the_same_bits(gpu_ctx, N, L0, R):
    """The pre-split contraction's 256 x 128 tiles (64 x 64 per wave, the
    default) and its 128 x 128 tiles form the same sums: Gamma bit-identical
    between the two and to the BF16x3 contraction of the oracle's fp64
    stacks, at N not a multiple of the tile (rows past N clamped and not
    stored) and with K = sites x R not a multiple of the 16-column stage (a
    short last stage through registers); K = 6 (one short stage) and K = 80
    (five full stages: one pass of the three-buffer main loop, then its
    last steps) cover the pipeline's ends."""
    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (L0, L0 + 1))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * np.pi / 9)
    lat.set_disorder(1.0, 4)
    Hs = L.rescale(lat.get_Hamiltonian_sparse(True), 5.0, 0.0)
    vx = L.commutator_diag(lat.op_position(0), Hs)
    vy = L.commutator_diag(lat.op_position(1), Hs)
    saved = gpu_ctx.get_tuning()
    try:
        K = gpu_ctx.random_block(Hs.n, R, seed=41)
        kets = K.download()
        M = [gpu_ctx.matrix(m) for m in (Hs, vx, vy)]
        gpu_ctx.set_tuning(**dict(saved, gamma_tile_rows=0))
        g_wide, _ = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=kpm.GAMMA_BF16X3)
        gpu_ctx.set_tuning(**dict(saved, gamma_tile_rows=128))
        g_sq, _ = kpm.kubo_gamma(gpu_ctx, *M, K, N, precision=kpm.GAMMA_BF16X3)
        assert np.array_equal(g_wide, g_sq)
        TL = oracle.cheb_on_bases(Hs, N, kets)
        w = oracle.cheb_on_bases(vy, 2, kets)[1]
        TR = oracle.cheb_on_bases(Hs, N, w)
        aR = np.stack([oracle.cheb_on_bases(vx, 2, TR[m])[1] for m in range(N)])
        Rp = 1 << (R - 1).bit_length()
        pad = lambda T: np.concatenate([T, np.zeros(T.shape[:2] + (Rp - R,))], axis=2) if Rp > R else T
        ref, _ = kpm.stack_contract(gpu_ctx, pad(TL).reshape(N, -1), pad(aR).reshape(N, -1), kpm.GAMMA_BF16X3)
        assert np.array_equal(g_wide, ref)
        K.close()
    finally:
        gpu_ctx.set_tuning(**saved)
