"""sigma_xy operators formed on the device (tbk_matrix_rescale,
tbk_matrix_velocity) against the oracle's rescale (lattice.h:550-559) and
commutator_with_diagonal of op_position (lattice.h:314-345): the same row
pointers, columns and value bits; a lattice sigma_xy request, whose three
operators the server forms on the device, gives the Gamma bits of the
direct path on host-formed operators."""
import math
import time

import numpy as np
import pytest

from oracle import lattice_oracle as O
from tests.helpers import lattice_pair
from tight_binding_amd import _native as Nn
from tight_binding_amd import kpm
from tight_binding_amd import lattice as L
from tight_binding_amd import spectral as S

pytestmark = pytest.mark.gpu

BUILDERS = {"open": L.BUILD_OPEN, "periodic": L.BUILD_PERIODIC}

CASES = [
    # (model, span, shells, W, gauge, B, builder, complex, a, b)
    ("square", (9, 8), [("A", 1, 1.0)], 1.0, "landau", 2 * math.pi / 8, "open", True, 5.2, 0.3),
    ("honeycomb", (7, 6), [("A", 1, -2.7)], 0.0, None, 0.0, "open", False, 8.4, 0.0),  # no diagonal entries
    ("honeycomb", (6, 6), [("A", 1, 1.0)], 0.5, "symmetric", 0.07, "periodic", True, 3.6, -0.2),
    ("kagome", (5, 4), [("A", 1, 1.0), ("B", 1, 1.0)], 2.0, None, 0.0, "open", False, 6.0, 0.1),
    ("cubic", (4, 3, 3), [("A", 1, 1.0)], 1.5, None, 0.0, "periodic", False, 7.6, 0.0),
    ("square", (2, 2), [("A", 1, 0.8)], 0.3, None, 0.0, "periodic", False, 4.0, 0.5),  # summed duplicates
    ("lieb", (5, 5), [("A", 1, 1.0)], 0.0, "landau", 0.3, "open", True, -4.1, 0.2),  # a < 0: signed zeros
]


def bits(v):
    v = np.ascontiguousarray(v)
    return v.view(np.uint64) if v.dtype == np.float64 else v.view(np.float64).view(np.uint64)


def oracle_csr(M):
    rp, ci, val = M.to_csr()
    return rp, ci, val


def same_csr(dev, want):
    return (np.array_equal(dev[0], want[0]) and np.array_equal(dev[1], want[1])
            and np.array_equal(bits(dev[2]), bits(want[2])))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_rescale_and_velocities_bit_identical(gpu_ctx, case):
    model, span, nn, W, gauge, B, builder, cplx, a, b = CASES[case]
    lat, ol = lattice_pair(model, span, nn, W=W, seed=11 + case, gauge=gauge, B=B)
    M, _ = gpu_ctx.assemble(lat, BUILDERS[builder], (1,) * len(span) if builder == "periodic" else None, cplx)
    oH = (ol.get_Hamiltonian_periodic_extend_bare_sparse((1,) * len(span), cplx) if builder == "periodic"
          else ol.get_Hamiltonian_sparse(cplx))
    oHs = O.rescale(oH, a, b)
    Hs = gpu_ctx.rescale(M, a, b)
    assert same_csr(Hs.download(), oracle_csr(oHs)), (case, "rescale")
    inf = Hs.info()
    assert inf["nnz"] == oHs.colptr[-1]
    assert inf["max_row_nnz"] == int(np.max(np.diff(oracle_csr(oHs)[0])))
    for axis in range(len(span)):
        v = gpu_ctx.velocity(Hs, lat, axis)
        want = O.commutator_with_diagonal(ol.op_position(axis), oHs)
        assert same_csr(v.download(), oracle_csr(want)), (case, "velocity", axis)
        assert v.info() == Hs.info()
        v.close()
    # an uploaded matrix rescales to the same bits as the assembled one
    Mu = gpu_ctx.matrix(oH)
    Hu = gpu_ctx.rescale(Mu, a, b)
    assert same_csr(Hu.download(), oracle_csr(oHs))
    for m in (M, Hs, Mu, Hu):
        m.close()


def test_rescaled_matrix_moments_equal_fused_rescale(gpu_ctx):
    """Moments of the materialised H~ at a = 1, b = 0 against the step's fused
    rescale of H, and against the uploaded oracle H~ (bits)."""
    lat, ol = lattice_pair("honeycomb", (20, 20), [("A", 1, 1.0)], W=0.8, seed=2, gauge="landau", B=0.1)
    M, _ = gpu_ctx.assemble(lat, L.BUILD_OPEN, None, True)
    a, b = 3.7, 0.05
    Hs = gpu_ctx.rescale(M, a, b)
    K = gpu_ctx.random_block(M.n, 8, seed=3)
    mu_m, _, _ = gpu_ctx.moments(Hs, K, 128)
    mu_f, _, _ = gpu_ctx.moments(M, K, 128, a=a, b=b)
    mu_o, _, _ = gpu_ctx.moments(gpu_ctx.matrix(O.rescale(ol.get_Hamiltonian_sparse(True), a, b)), K, 128)
    assert np.array_equal(mu_m, mu_o)
    assert np.max(np.abs(mu_m - mu_f)) <= 1e-11 * mu_m[0]


def test_operator_errors(gpu_ctx):
    lat, _ = lattice_pair("square", (4, 4), [("A", 1, 1.0)])
    M, _ = gpu_ctx.assemble(lat, L.BUILD_OPEN, None, False)
    for a in (0.0, float("inf"), float("nan")):
        with pytest.raises(Nn.NativeError) as e:
            gpu_ctx.rescale(M, a, 0.0)
        assert e.value.code == Nn.TBK_E_INVALID
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.velocity(M, lat, 2)  # a 2-d lattice has axes 0 and 1
    assert e.value.code == Nn.TBK_E_INVALID
    other, _ = lattice_pair("square", (5, 4), [("A", 1, 1.0)])
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.velocity(M, other, 0)  # not this matrix's sites
    assert e.value.code == Nn.TBK_E_INVALID
    holes, _ = lattice_pair("square", (4, 4), [("A", 1, 1.0)], removed=("A", np.array([[1, 1]], dtype=np.int32)))
    with pytest.raises(Nn.NativeError) as e:
        gpu_ctx.velocity(M, holes, 0)
    assert e.value.code == Nn.TBK_E_UNSUPPORTED


@pytest.mark.parametrize("prec", [kpm.GAMMA_FP64, kpm.GAMMA_BF16X3])
def test_lattice_kubo_request_matches_host_operators(gpu_ctx, prec):
    """The server assembles H, forms H~ and v_x, v_y on the device; Gamma is
    bit for bit the direct path's on the host-formed H~, v_x, v_y (the
    KuboSolver flow: rescale, commutator_diag of op_position)."""
    lat, _ = lattice_pair("square", (10, 10), [("A", 1, 1.0)], W=1.0, seed=3, gauge="landau", B=2 * math.pi / 8)
    a, b = 5.2, 0.1
    H = lat.get_Hamiltonian_sparse(True)
    Hs = L.rescale(H, a, b)
    vx = L.commutator_diag(lat.op_position(0), Hs)
    vy = L.commutator_diag(lat.op_position(1), Hs)
    g = S.lorentz(32, 6.0)
    N = g.size
    ref, _ = kpm.kubo_gamma(gpu_ctx, gpu_ctx.matrix(Hs), gpu_ctx.matrix(vx), gpu_ctx.matrix(vy),
                            gpu_ctx.random_block(Hs.n, 4, seed=9), N, precision=prec, g=g)
    with kpm.Server(0, lazy_arena=True) as srv:
        t = srv.submit_kubo_lattice(lat, N, a, b, R=4, seed=9, precision=prec, g=g)
        t2 = srv.submit_kubo(Hs, vx, vy, N, R=4, seed=9, precision=prec, g=g)
        assert srv.wait(t, 60.0)["state"] == Nn.TBK_REQ_DONE
        assert srv.wait(t2, 60.0)["state"] == Nn.TBK_REQ_DONE
        (z, G), info = srv.take(t)
        (z2, G2), _ = srv.take(t2)
    assert np.array_equal(G, ref) and np.array_equal(G2, ref) and np.array_equal(z, z2)
    assert info["a"] == a and info["b"] == b


def test_lattice_kubo_request_errors():
    lat, _ = lattice_pair("square", (6, 6), [("A", 1, 1.0)])
    with kpm.Server(0, lazy_arena=True) as srv:
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_kubo_lattice(lat, 16, 0.0)  # sigma_xy needs a, b
        assert e.value.code == Nn.TBK_E_INVALID
        chain, _ = lattice_pair("chain", (12,), [("A", 1, 1.0)])
        with pytest.raises(Nn.NativeError) as e:
            srv.submit_kubo_lattice(chain, 16, 2.5)  # one axis: no v_y
        assert e.value.code == Nn.TBK_E_INVALID


def test_c3_sized_operators(gpu_ctx):
    """The sigma_xy workload's operators (square 512 x 512, open, Landau
    flux 1/64, W = 2): H~, v_x, v_y formed on the device equal the host's,
    and the time of each path."""
    c = L.UnitCell.model("square")
    c.add_hopping_nn("A", 1, 1.0)
    lat = L.PartialLattice(c, (512, 512))
    lat.set_gauge(L.GAUGE_LANDAU, 2 * math.pi / 64)
    lat.set_disorder(2.0, 5)
    a, b = 5.1, 0.0
    t0 = time.perf_counter()
    Hs = L.rescale(lat.get_Hamiltonian_sparse(True), a, b)
    vx = L.commutator_diag(lat.op_position(0), Hs)
    vy = L.commutator_diag(lat.op_position(1), Hs)
    host_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    M, _ = gpu_ctx.assemble(lat, L.BUILD_OPEN, None, True)
    D = gpu_ctx.rescale(M, a, b)
    DX, DY = gpu_ctx.velocity(D, lat, 0), gpu_ctx.velocity(D, lat, 1)
    dev_s = time.perf_counter() - t0
    for dev, host in ((D, Hs), (DX, vx), (DY, vy)):
        assert same_csr(dev.download(), host.to_csr())
    print("C3 operators: host %.3f s, device %.1f ms (H %d rows, %d entries)" % (host_s, 1e3 * dev_s, D.n, D.nnz))
    for m in (M, D, DX, DY):
        m.close()
