"""Pins of the CPU oracle itself, against things the oracle does not compute
from its own formulas: libstdc++'s real priority_queue, scipy's sparse
construction, brute-force neighbour shells, closed-form lattice spectra,
exact diagonalisation, Chebyshev identities and kernel identities.

The reference ships no tests or golden vectors (no tests/ directory, no
fixtures), so these checks are what fixes the oracle's behaviour."""
import itertools
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle import lattice_oracle as O
from tests.helpers import cheb_T, honeycomb_bloch_eigs, square_bloch_eigs

# ------------------------------------------------------------------ heap


HEAP_CPP = r"""
#include <cstdio>
#include <queue>
#include <tuple>
#include <vector>
using Item = std::tuple<double, int>;
struct Cmp { bool operator()(const Item& a, const Item& b) const { return std::get<0>(a) > std::get<0>(b); } };
int main() {
    std::priority_queue<Item, std::vector<Item>, Cmp> q;
    int op; double d; int id;
    while (std::scanf("%d", &op) == 1) {
        if (op == 0) { std::scanf("%lf %d", &d, &id); q.emplace(d, id); }
        else { std::printf("%d\n", std::get<1>(q.top())); q.pop(); }
    }
}
"""


def test_heap_restatement_matches_libstdcxx_priority_queue():
    """Equal-distance ties must pop in libstdc++'s order; compare the Python
    restatement with a real std::priority_queue on a tie-heavy sequence."""
    rng = np.random.default_rng(3)
    ops = []
    live = 0
    for i in range(3000):
        if live == 0 or rng.random() < 0.6:
            ops.append((0, float(rng.integers(0, 6)) * 0.5, i))
            live += 1
        else:
            ops.append((1,))
            live -= 1
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "h.cpp")
        exe = os.path.join(td, "h")
        open(src, "w").write(HEAP_CPP)
        subprocess.check_call(["g++", "-O1", "-o", exe, src])
        inp = "\n".join("0 %r %d" % (o[1], o[2]) if o[0] == 0 else "1" for o in ops) + "\n"
        out = subprocess.run([exe], input=inp.encode(), capture_output=True, check=True).stdout.decode().split()
    q = O.StdPriorityQueue(lambda n, m: n[0] > m[0])
    got = []
    for o in ops:
        if o[0] == 0:
            q.push((o[1], o[2]))
        else:
            got.append(q.top()[1])
            q.pop()
    assert got == [int(x) for x in out]


# ----------------------------------------------------------- NN shells


def brute_shells(cell, frm, kmax, reach=3):
    """k-th distinct distance shells (1e-7 ties) by enumerating every atom of
    every cell in [-reach, reach]^dim, independent of the heap search."""
    items = []
    for dr in itertools.product(range(-reach, reach + 1), repeat=cell.dim):
        for b in range(cell.n_atoms):
            hp = (frm, b, tuple(dr))
            items.append((cell.distance_of_hopping(hp), hp))
    items.sort(key=lambda x: x[0])
    shells, cur, d0 = [], [], None
    for d, hp in items:
        if d0 is None or abs(d - d0) >= 1e-7:
            if cur:
                shells.append(cur)
            cur, d0 = [], d
        cur.append(hp)
    shells.append(cur)
    return [set(s) for s in shells[: kmax + 1]]


@pytest.mark.parametrize("model,atom,k", [
    ("square", "A", 1), ("square", "A", 2), ("square", "A", 3), ("honeycomb", "A", 1), ("honeycomb", "B", 1),
    ("honeycomb", "A", 2), ("lieb", "A", 1), ("lieb", "B", 1), ("kagome", "A", 1), ("kagome", "C", 1),
    ("fcc", "A", 1), ("cubic", "A", 1), ("cubic", "A", 2), ("chain", "A", 2),
])
def test_nn_shells_equal_brute_force(model, atom, k):
    cell = O.model(model)
    a = cell.atom_no[atom]
    gen = O.UnitCellNN(a)
    gen.generate_hoppings_to(cell, k)
    want = brute_shells(cell, a, k)
    for j in range(k + 1):
        assert set(gen.NN_solved[j]) == want[j], (model, atom, j)


def test_closed_form_coordination_numbers():
    expect = {("square", 1): 4, ("honeycomb", 1): 3, ("fcc", 1): 12, ("cubic", 1): 6, ("cubic", 2): 12,
              ("lieb", 1): 4, ("kagome", 1): 4}
    for (model, k), z in expect.items():
        cell = O.model(model)
        gen = O.UnitCellNN(0)
        gen.generate_hoppings_to(cell, k)
        assert len(gen.NN_solved[k]) == z
    # canonical storage keeps one of each +-dR pair when from == to
    cell = O.model("fcc")
    cell.add_hopping_nn("A", 1, 1.0)
    assert len(cell.hoppings) == 6
    assert all(hp[0] >= hp[1] for hp, _ in cell.hoppings)


def test_canonical_orientation_and_conjugation():
    cell = O.model("honeycomb")
    cell.add_hopping("A", "B", (0, 0), 0.3 + 0.4j)  # from < to -> stored as h.c.
    assert cell.hoppings == [((1, 0, (0, 0)), 0.3 - 0.4j)]
    assert cell.add_hopping_direct((1, 0, (0, 0)), 1.0) is False  # duplicate
    assert cell.add_hopping_direct((0, 1, (0, 0)), 1.0) is False  # its h.c.
    cell.add_hopping("A", "A", (0, 0), 1j)  # complex on-site rejected
    assert len(cell.hoppings) == 1


# --------------------------------------------------------- setFromTriplets


def test_set_from_triplets_matches_scipy_and_insertion_order_sum():
    rng = np.random.default_rng(5)
    n, m = 40, 900
    rows = rng.integers(0, n, m)
    cols = rng.integers(0, n, m)
    vals = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    csc = O.set_from_triplets(n, rows, cols, vals)
    ref = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsc()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(csc.colptr, ref.indptr)
    assert np.array_equal(csc.rowidx, ref.indices)
    assert np.allclose(csc.values, ref.data, rtol=1e-14, atol=1e-14)
    # exact insertion-order summation, brute force
    acc = {}
    for r, c, v in zip(rows, cols, vals):
        acc[(r, c)] = acc[(r, c)] + v if (r, c) in acc else v
    D = csc.to_dense()
    for (r, c), v in acc.items():
        assert D[r, c] == v


# --------------------------------------------------- assembly closed forms


def ring(L):
    A = np.zeros((L, L))
    for i in range(L):
        A[i, (i + 1) % L] += 1
        A[i, (i - 1) % L] += 1
    return A


@pytest.mark.parametrize("L0,L1", [(3, 3), (4, 5), (7, 3)])
def test_periodic_square_is_kron_of_rings(L0, L1):
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (L0, L1))
    W = 0.8
    lat.set_scalar_field_indexed(O.uniform_disorder(W, 11))
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1)).to_dense()
    V = O.uniform_disorder(W, 11)(np.arange(L0 * L1))
    ref = np.kron(ring(L0), np.eye(L1)) + np.kron(np.eye(L0), ring(L1)) + np.diag(V)
    assert np.array_equal(H, ref)


def test_periodic_spectra_equal_bloch_closed_forms():
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    H = O.PartialLattice(cell, (6, 8)).get_Hamiltonian_periodic_extend_bare_sparse((1, 1)).to_dense()
    assert np.allclose(np.sort(np.linalg.eigvalsh(H)), np.sort(square_bloch_eigs(6, 8)), atol=1e-12)
    cell = O.model("honeycomb")
    cell.add_hopping_nn("A", 1, -2.7)
    H = O.PartialLattice(cell, (5, 7)).get_Hamiltonian_periodic_extend_bare_sparse((1, 1)).to_dense()
    assert np.allclose(np.sort(np.linalg.eigvalsh(H)), np.sort(honeycomb_bloch_eigs(5, 7, -2.7)), atol=1e-11)


def test_peierls_phase_two_site_chain_by_hand():
    """Two sites of an open chain, t = -1.3, constant A = (0.7, 0, 0).  The
    stored hopping is A->A over dR = +1 or -1; either way lattice.h:284 gives
    H[0,1] = t exp(+i A.d) with d = (1, 0, 0), i.e. -1.3 (cos 0.7 + i sin 0.7),
    and H[1,0] its conjugate."""
    cell = O.model("chain")
    cell.add_hopping_nn("A", 1, -1.3)
    lat = O.PartialLattice(cell, (2,))
    lat.gauge_field = lambda pos: np.tile(np.array([0.7, 0.0, 0.0]), (pos.shape[0], 1))
    H = lat.get_Hamiltonian_sparse(True).to_dense()
    h01 = complex(-1.3 * 0.7648421872844885, -1.3 * 0.644217687237691)
    assert abs(H[0, 1] - h01) < 1e-15
    assert abs(H[1, 0] - h01.conjugate()) < 1e-15
    assert H[0, 0] == 0 and H[1, 1] == 0


@pytest.mark.parametrize("gauge", ["landau", "symmetric"])
def test_peierls_flux_per_plaquette(gauge):
    """t*exp(+i A(r_from).d) (lattice.h:284: `1.0i * A.dot(d)` with A and d
    real): the product H[a,b] H[b,c] H[c,d] H[d,a] around a counter-clockwise
    unit plaquette has phase +B (gauge invariant); a reversed sign gives -B."""
    B = 2 * math.pi / 7
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    L = 6
    lat = O.PartialLattice(cell, (L, L))
    lat.gauge_field = (O.landau_gauge if gauge == "landau" else O.symmetric_gauge)(B)
    H = lat.get_Hamiltonian_sparse(True).to_dense()
    idx = lambda x, y: x * L + y  # noqa: E731  (last index fastest)
    for x in range(L - 1):
        for y in range(L - 1):
            loop = (H[idx(x, y), idx(x + 1, y)] * H[idx(x + 1, y), idx(x + 1, y + 1)]
                    * H[idx(x + 1, y + 1), idx(x, y + 1)] * H[idx(x, y + 1), idx(x, y)])
            assert abs(np.angle(loop) - B) < 1e-12
    assert np.allclose(H, H.conj().T)


def test_rescale_union_pattern():
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    H = O.PartialLattice(cell, (4, 4)).get_Hamiltonian_bare_sparse()  # no diagonal entries
    assert not np.any(np.diag(H.to_dense()))
    Hs = O.rescale(H, 2.5, 0.5)
    D = Hs.to_dense()
    assert Hs.nnz == H.nnz + 16
    assert np.array_equal(np.diag(D), np.full(16, (0.0 - 0.5) / 2.5))
    assert np.allclose(D, (H.to_dense() - 0.5 * np.eye(16)) / 2.5, rtol=0, atol=0)


# ------------------------------------------------ Chebyshev recursion oracle


def _ed_moments(Hs_dense, N, bras, kets):
    e, V = np.linalg.eigh(Hs_dense)
    T = cheb_T(N, e)
    Bv = V.conj().T @ bras
    Kv = V.conj().T @ kets
    # Re tr(B^H V diag(T_n) V^H K)
    return np.array([np.real(np.sum(np.conj(Bv) * (T[n][:, None] * Kv))) for n in range(N)])


@pytest.mark.parametrize("case", ["square_real", "honeycomb_real", "landau_complex"])
def test_oracle_recursion_full_basis_equals_exact_diagonalisation(case):
    if case == "square_real":
        cell, span, t, cplx = O.model("square"), (6, 5), 1.0, False
    elif case == "honeycomb_real":
        cell, span, t, cplx = O.model("honeycomb"), (4, 4), -2.7, False
    else:
        cell, span, t, cplx = O.model("square"), (5, 5), 1.0, True
    cell.add_hopping_nn("A", 1, t)
    lat = O.PartialLattice(cell, span)
    lat.set_scalar_field_indexed(O.uniform_disorder(1.0, 2))
    if cplx:
        lat.gauge_field = O.landau_gauge(2 * math.pi / 5)
        H = lat.get_Hamiltonian_sparse(True)
    else:
        H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1))
    a = 1.05 * (np.max(np.abs(np.linalg.eigvalsh(H.to_dense()))) + 0.1)  # spectrum of H~ inside (-1, 1)
    Hs = O.rescale(H, a, 0.1)
    n = Hs.n
    N = 40
    kets = np.eye(n, dtype=np.complex128)
    mu = oracle.cheb_braket(Hs, N, kets, kets)
    ref = _ed_moments(Hs.to_dense(), N, kets, kets)
    assert np.max(np.abs(mu - ref)) <= 1e-11 * n
    assert abs(mu[0] - n) < 1e-12  # mu_0 = tr 1
    # distinct bras: random blocks
    rng = np.random.default_rng(0)
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    K = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    mu2 = oracle.cheb_braket(Hs, N, B, K)
    ref2 = _ed_moments(Hs.to_dense(), N, B, K)
    assert np.max(np.abs(mu2 - ref2)) <= 1e-11 * np.max(np.abs(ref2))


def test_oracle_chebyshev_product_identity_and_stored_blocks():
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (8, 8))
    lat.set_scalar_field_indexed(O.uniform_disorder(0.5, 4))
    Hs = O.rescale(lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1)), 4.6, 0.0)
    kets = O.random_phase_kets(Hs.n, 4, 9)
    N = 30
    mu = oracle.cheb_braket(Hs, N, kets, kets)
    T = oracle.cheb_on_bases(Hs, N, kets)
    # T_n stored blocks reproduce the moments, and T_m T_k = (T_{m+k} + T_{|m-k|})/2
    for k in range(N):
        assert abs(np.real(np.sum(np.conj(kets) * T[k])) - mu[k]) < 1e-12
    for m in range(0, 14):
        for k in range(0, 14):
            lhs = 2 * np.real(np.sum(np.conj(T[m]) * T[k]))
            assert abs(lhs - mu[m + k] - mu[abs(m - k)]) < 1e-11
    assert abs(mu[0] - 4.0) < 1e-12
    assert np.all(np.abs(mu) <= mu[0] + 1e-12)


def test_oracle_stored_blocks_match_eigendecomposition():
    """T_k(H) V = U cos(k arccos Lambda) U^H V for Hermitian H with spectrum
    in [-1, 1] (complex Landau-gauge hoppings, so a transposed or unconjugated
    operand fails); n = 1 is V alone."""
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (6, 5))
    lat.set_scalar_field_indexed(O.uniform_disorder(1.0, 2))
    lat.gauge_field = O.landau_gauge(2 * math.pi / 7)
    H = lat.get_Hamiltonian_sparse(True)
    Hs = O.rescale(H, 5.1, 0.1)
    D = Hs.to_dense()
    assert np.abs(D - D.conj().T).max() == 0 and np.abs(D.imag).max() > 0.01
    lam, U = np.linalg.eigh(D)
    assert np.max(np.abs(lam)) < 1
    rng = np.random.default_rng(3)
    V = rng.standard_normal((Hs.n, 3)) + 1j * rng.standard_normal((Hs.n, 3))
    n = 40
    T = oracle.cheb_on_bases(Hs, n, V)
    W = U.conj().T @ V
    for k in range(n):
        want = U @ (np.cos(k * np.arccos(lam))[:, None] * W)
        assert np.max(np.abs(T[k] - want)) <= 1e-12 * np.max(np.abs(V)) * (k + 1), k
    assert np.array_equal(oracle.cheb_on_bases(Hs, 1, V)[0], V)


def test_trace_braket_matches_numpy():
    rng = np.random.default_rng(1)
    B = rng.standard_normal((50, 7)) + 1j * rng.standard_normal((50, 7))
    K = rng.standard_normal((50, 7)) + 1j * rng.standard_normal((50, 7))
    assert abs(oracle.trace_braket(B, K) - np.real(np.vdot(B, K))) < 1e-12


# ---------------------------------------------------------------- kernels


@pytest.mark.parametrize("N", [8, 64, 513])
def test_jackson_kernel_identities(N):
    g = oracle.kernel_jackson(N)
    assert abs(g[0] - 1.0) < 1e-14
    assert abs(g[1] - math.cos(math.pi / (N + 1))) < 1e-14
    phi = np.linspace(0, np.pi, 2001)
    K = g[0] + 2 * np.sum(g[1:, None] * np.cos(np.arange(1, N)[:, None] * phi[None, :]), axis=0)
    assert K.min() > -1e-9 * K.max()  # the Jackson kernel is positive


@pytest.mark.parametrize("N,lam", [(64, 4.0), (512, 6.0)])
def test_lorentz_kernel_identities(N, lam):
    g = oracle.kernel_lorentz(N, lam)
    assert abs(g[0] - 1.0) < 1e-14
    c = 2 * math.cosh(lam / N)
    assert np.max(np.abs(g[:-2] + g[2:] - c * g[1:-1])) < 1e-12
    # continued one step, the recursion reaches sinh(0) = 0 at n = N
    assert abs(c * g[-1] - g[-2]) < 1e-12


def test_fejer_kernel():
    g = oracle.kernel_fejer(10)
    assert g[0] == 1.0 and np.allclose(np.diff(g), -0.1)


@pytest.mark.parametrize("cplx,threads", [(False, 1), (False, 3), (True, 4)])
def test_all_cores_timing_variant_equals_the_oracle_loop(cplx, threads):
    """The OpenMP timing variant of the CPU baseline runs the oracle's
    recursion per block column: the same T_n, the moments differ only in the
    summation order of the trace."""
    cell = O.model("honeycomb")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (9, 7))
    lat.set_scalar_field_indexed(O.uniform_disorder(0.8, 5))
    if cplx:
        lat.gauge_field = O.symmetric_gauge(0.05)
    Hs = O.rescale(lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), cplx), 3.5, 0.1)
    kets = O.random_phase_kets(Hs.n, 5, 12)
    steps = 30
    ref = oracle.cheb_braket(Hs, steps + 2, kets, kets)
    got = oracle.cheb_steps_threads(Hs, steps, kets, threads)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
