"""Hamiltonian assembly on the device (tbk_matrix_assemble) against the host
builders (get_Hamiltonian_sparse, get_Hamiltonian_periodic_extend_bare_sparse,
get_Hamiltonian_bare_sparse, restated from lattice.h:216-305, 347-377 and
pinned against the oracle in test_assembly.py): the same row pointers,
column indices and value bits, for every builder, 1-3 dimensional models,
periodic spans of 1 and 2 (self-loops and duplicates summed in insertion
order), second shells, both gauges (the host's phase table) and site-keyed
disorder; the engine's matrix statistics and moments agree with those of the
host-built matrix."""
import math
import time

import numpy as np
import pytest

from oracle import lattice_oracle as O
from tight_binding_amd import _native as Nn
from tight_binding_amd import kpm
from tight_binding_amd import lattice as L

pytestmark = pytest.mark.gpu


def lattice(model, span, shells=(1,), t=1.0, W=0.0, seed=1, gauge=0, B=0.0, all_atoms=False):
    c = L.UnitCell.model(model)
    atoms = [a for a, _ in c.atoms()] if all_atoms and hasattr(c, "atoms") else ["A"]
    for k in shells:
        for a in atoms:
            c.add_hopping_nn(a, k, t * (1.0 if k == 1 else 0.37))
    lat = L.PartialLattice(c, span)
    if gauge:
        lat.set_gauge(gauge, B)
    if W:
        lat.set_disorder(W, seed)
    return lat


def oracle_csr(model, span, shells, t, W, seed, gauge, B, builder, extend, is_complex):
    """The same Hamiltonian from the oracle's restatement of the builders, as
    CSR arrays (lattice.h:216-305)."""
    c = O.model(model)
    for k in shells:
        c.add_hopping_nn("A", k, t * (1.0 if k == 1 else 0.37))
    lat = O.PartialLattice(c, span)
    if gauge:
        lat.gauge_field = (O.landau_gauge if gauge == L.GAUGE_LANDAU else O.symmetric_gauge)(B)
    if W:
        lat.set_scalar_field_indexed(O.uniform_disorder(W, seed))
    if builder == L.BUILD_OPEN:
        H = lat.get_Hamiltonian_sparse(is_complex)
    elif builder == L.BUILD_PERIODIC:
        H = lat.get_Hamiltonian_periodic_extend_bare_sparse(extend, is_complex)
    else:
        H = lat.get_Hamiltonian_bare_sparse(is_complex)
    return H.to_csr()


def host_csr(lat, builder, extend, is_complex):
    if builder == L.BUILD_OPEN:
        H = lat.get_Hamiltonian_sparse(is_complex)
    elif builder == L.BUILD_PERIODIC:
        H = lat.get_Hamiltonian_periodic_extend_bare_sparse(extend, is_complex)
    else:
        H = lat.get_Hamiltonian_bare_sparse(is_complex)
    return H, H.to_csr()


def bits(v):
    v = np.ascontiguousarray(v)
    return v.view(np.uint64) if v.dtype == np.float64 else v.view(np.float64).view(np.uint64)


CASES = [
    # (name, model, span, shells, W, gauge, B, builder, extend, complex)
    ("square periodic 7x7 W", "square", (7, 7), (1,), 1.0, 0, 0.0, L.BUILD_PERIODIC, (1, 1), False),
    ("square periodic 2x2 (duplicates)", "square", (2, 2), (1,), 0.7, 0, 0.0, L.BUILD_PERIODIC, (1, 1), True),
    ("square periodic 1x3 (self-loops)", "square", (1, 3), (1,), 0.5, 0, 0.0, L.BUILD_PERIODIC, (1, 1), False),
    ("square periodic 3x3 NNN", "square", (3, 3), (1, 2), 0.3, 0, 0.0, L.BUILD_PERIODIC, (1, 1), True),
    ("honeycomb open Landau", "honeycomb", (9, 7), (1,), 0.5, L.GAUGE_LANDAU, 2 * math.pi / 9, L.BUILD_OPEN, None, True),
    ("honeycomb periodic symmetric", "honeycomb", (8, 8), (1,), 0.0, L.GAUGE_SYMMETRIC, 0.05, L.BUILD_PERIODIC, (1, 1), True),
    ("kagome periodic x only", "kagome", (6, 5), (1,), 2.0, 0, 0.0, L.BUILD_PERIODIC, (1, 0), False),
    ("lieb bare", "lieb", (6, 6), (1,), 1.0, L.GAUGE_LANDAU, 0.3, L.BUILD_BARE, None, True),
    ("cubic periodic 5x4x3", "cubic", (5, 4, 3), (1,), 1.5, 0, 0.0, L.BUILD_PERIODIC, (1, 1, 1), False),
    ("fcc open 4x4x4 symmetric", "fcc", (4, 4, 4), (1,), 0.2, L.GAUGE_SYMMETRIC, 0.1, L.BUILD_OPEN, None, True),
    ("chain periodic 5", "chain", (5,), (1, 2), 1.0, 0, 0.0, L.BUILD_PERIODIC, (1,), False),
    ("square open Landau real (phase dropped)", "square", (6, 5), (1,), 0.4, L.GAUGE_LANDAU, 0.2, L.BUILD_OPEN, None, False),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_assembly_is_bit_identical(gpu_ctx, case):
    name, model, span, shells, W, gauge, B, builder, ext, cplx = CASES[case]
    lat = lattice(model, span, shells, W=W, seed=7 + case, gauge=gauge, B=B)
    H, (rp, ci, val) = host_csr(lat, builder, ext, cplx)
    M, ms = gpu_ctx.assemble(lat, builder, ext, cplx)
    drp, dci, dval = M.download()
    assert np.array_equal(drp, rp), name
    assert np.array_equal(dci, ci), name
    assert np.array_equal(bits(dval), bits(val)), name
    # and directly against the oracle, independently of the host builder
    orp, oci, oval = oracle_csr(model, span, shells, 1.0, W, 7 + case, gauge, B, builder, ext, cplx)
    assert np.array_equal(drp, orp), name
    assert np.array_equal(dci, oci), name
    assert np.array_equal(bits(dval), bits(np.ascontiguousarray(oval))), name
    # the engine's statistics: as for the uploaded host matrix
    Mh = gpu_ctx.matrix(H)
    assert M.info() == Mh.info(), name
    bd, bh = gpu_ctx.spectral_bounds(M, max_iters=8), gpu_ctx.spectral_bounds(Mh, max_iters=8)
    for k in ("gersh_min", "gersh_max"):
        assert abs(bd[k] - bh[k]) <= 4e-16 * max(1.0, abs(bh[k])), (name, k)
    M.close()
    Mh.close()


def test_moments_from_device_assembly_equal_host_matrix(gpu_ctx):
    lat = lattice("honeycomb", (40, 40), W=0.8, seed=3, gauge=L.GAUGE_LANDAU, B=2 * math.pi / 40)
    H = lat.get_Hamiltonian_sparse(True)
    M, _ = gpu_ctx.assemble(lat, L.BUILD_OPEN, None, True)
    Mh = gpu_ctx.matrix(H)
    K = gpu_ctx.random_block(H.n, 16, seed=9)
    mu_d, _, _ = gpu_ctx.moments(M, K, 256, a=3.6)
    mu_h, _, _ = gpu_ctx.moments(Mh, K, 256, a=3.6)
    assert np.array_equal(mu_d, mu_h)
    for m in (M, Mh, K):
        m.close()


def test_c2_sized_assembly(gpu_ctx):
    """The 2,097,152-site honeycomb of the C2 workload: bit-identical to the
    host builder, and the time of each path."""
    from bench import build_c2

    t0 = time.perf_counter()
    H, a = build_c2(1024)
    host_s = time.perf_counter() - t0
    lat = lattice("honeycomb", (1024, 1024), t=-2.7, W=0.5, seed=2026)
    t0 = time.perf_counter()
    M, ms = gpu_ctx.assemble(lat, L.BUILD_PERIODIC, (1, 1), False)
    dev_s = time.perf_counter() - t0
    rp, ci, val = H.to_csr()
    drp, dci, dval = M.download()
    assert np.array_equal(drp, rp) and np.array_equal(dci, ci) and np.array_equal(bits(dval), bits(val))
    print("C2 assembly: host builder %.3f s, device %.1f ms (call %.3f s)" % (host_s, ms, dev_s))
    M.close()


def test_assembly_limits(gpu_ctx):
    lat = lattice("square", (4, 4))
    lat.remove_atoms("A", np.array([[1, 1]], dtype=np.int32))
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.assemble(lat, L.BUILD_OPEN, None, False)
    assert e.value.code == Nn.TBK_E_UNSUPPORTED
    # the periodic builder ignores removed sites (lattice.h:216-258): allowed
    M, _ = gpu_ctx.assemble(lat, L.BUILD_PERIODIC, (1, 1), False)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1))
    assert np.array_equal(M.download()[1], H.to_csr()[1])
    M.close()
    # more than 32 pushes touching one atom
    big = lattice("square", (12, 12), shells=tuple(range(1, 10)))
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.assemble(big, L.BUILD_PERIODIC, (1, 1), False)
    assert e.value.code == Nn.TBK_E_UNSUPPORTED


def test_server_assembles_lattice_requests(gpu_ctx):
    """DoS requests that carry a lattice (assembled on the device at
    admission) give the same bits as requests that carry its host-built
    matrix; with a = 0 the bounds come from the spectral range."""
    specs = [
        (lattice("honeycomb", (30, 30), W=0.5, seed=4), L.BUILD_PERIODIC, (1, 1), False, kpm.FP64, 3.2),
        (lattice("square", (25, 20), W=1.0, seed=5, gauge=L.GAUGE_LANDAU, B=0.2), L.BUILD_OPEN, None, True,
         kpm.FP32, 4.8),
    ]
    with kpm.Server(0, lazy_arena=True, chunk_steps=8) as srv:
        tl = [srv.submit_dos(lat, 96, a, 0.0, 8, 40 + i, prec, builder=b, extend=e, is_complex=c)
              for i, (lat, b, e, c, prec, a) in enumerate(specs)]
        tc = [srv.submit_dos(host_csr(lat, b, e, c)[0], 96, a, 0.0, 8, 40 + i, prec)
              for i, (lat, b, e, c, prec, a) in enumerate(specs)]
        ta = srv.submit_dos(specs[0][0], 64, 0.0, 0.0, 8, 50, kpm.FP64, builder=L.BUILD_PERIODIC, extend=(1, 1))
        for x, y in zip(tl, tc):
            assert srv.wait(x, 60.0)["state"] == Nn.TBK_REQ_DONE and srv.wait(y, 60.0)["state"] == Nn.TBK_REQ_DONE
            assert np.array_equal(srv.take(x)[0], srv.take(y)[0])
        assert srv.wait(ta, 60.0)["state"] == Nn.TBK_REQ_DONE
        mu, info = srv.take(ta)
    H = host_csr(specs[0][0], L.BUILD_PERIODIC, (1, 1), False)[0]
    ev = np.linalg.eigvalsh(H.to_dense())
    a, b = info["a"], info["b"]
    assert (ev[-1] - b) / a <= 1 - 0.004 and (ev[0] - b) / a >= -1 + 0.004
    assert np.max(np.abs(mu)) <= mu[0] * (1 + 1e-12)


@pytest.mark.parametrize("model,span,z", [("square", (20, 18), 4), ("cubic", (7, 6, 5), 6)])
def test_server_assembles_sigma_xy_lattice_requests(gpu_ctx, model, span, z):
    """sigma_xy requests that carry a lattice (H, H~ = H/a and v_x, v_y all
    formed on the device at admission) give the Gamma of the same request
    carrying the host-built H~, v_x, v_y, bit for bit, and the Gamma of the
    oracle's lattice, rescale and commutators (lattice.h:314-345) with the
    same kets within 1e-10."""
    import oracle
    from oracle import spectral_oracle as SO
    from tight_binding_amd import kubo as Kb
    from tight_binding_amd import spectral as S

    B, W, seed, N, R, kseed = 2 * math.pi / 16, 1.0, 61, 48, 4, 91
    a = 1.02 * (z + W / 2)
    lat = lattice(model, span, W=W, seed=seed, gauge=L.GAUGE_LANDAU, B=B)
    Hs = L.rescale(lat.get_Hamiltonian_sparse(True), a, 0.0)
    vx, vy = Kb.velocity_operators(Hs, lat.op_position(0), lat.op_position(1))
    g = S.lorentz(N, 4.0)
    with kpm.Server(0, lazy_arena=True) as srv:
        tl = srv.submit_kubo_lattice(lat, N, a, 0.0, R=R, seed=kseed, g=g)
        th = srv.submit_kubo(Hs, vx, vy, N, R=R, seed=kseed, g=g)
        for t in (tl, th):
            assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        (zl, Gl), info = srv.take(tl)
        (zh, Gh), _ = srv.take(th)
    assert info["vector_groups"] == 1 and info["row_ranges"] == 1
    assert np.array_equal(Gl, Gh) and np.array_equal(zl, zh)
    # the oracle's Hamiltonian, rescale and velocities, the server's kets
    ocell = O.model(model)
    ocell.add_hopping_nn("A", 1, 1.0)
    olat = O.PartialLattice(ocell, span)
    olat.gauge_field = O.landau_gauge(B)
    olat.set_scalar_field_indexed(O.uniform_disorder(W, seed))
    oHs = O.rescale(olat.get_Hamiltonian_sparse(True), a, 0.0)
    ovx = O.commutator_with_diagonal(olat.op_position(0), oHs)
    ovy = O.commutator_with_diagonal(olat.op_position(1), oHs)
    kets = gpu_ctx.random_block(oHs.n, R, seed=kseed).download()
    TL = oracle.cheb_on_bases(oHs, N, kets)
    w = oracle.cheb_on_bases(ovy, 2, kets)[1]
    TR = oracle.cheb_on_bases(oHs, N, w)
    aR = np.stack([oracle.cheb_on_bases(ovx, 2, TR[m])[1] for m in range(N)])
    ref = SO.kubo_gamma(TL * g[:, None, None], aR * g[:, None, None])
    assert np.max(np.abs(Gl - ref)) <= 1e-10 * np.max(np.abs(ref))
