"""The C++ drop-in (tight_binding_amd/host: numerics_cuda.cuh, green.h) used
the way the reference's callers use it (tests/cpp/dropin_dos.cpp):
compiles against the package headers and libtbkpm.so on the CPU; on the GPU
its GreenFunctionImag moments and density_of_states_seq values match the
oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from oracle import spectral_oracle as SO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(tmp_path, name="dropin_dos"):
    exe = os.path.join(str(tmp_path), name)
    subprocess.check_call([
        "g++", "-std=c++20", "-O2", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
        "-I" + os.path.join(ROOT, "tight_binding_amd", "host"), "-o", exe,
        os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-L" + os.path.join(ROOT, "tight_binding_amd"),
        "-ltbkpm", "-Wl,-rpath," + os.path.join(ROOT, "tight_binding_amd"),
    ])
    return exe


@pytest.mark.parametrize("name", ["dropin_dos", "kubo_xy_entry", "kubo_bastin", "green_direct"])
def test_cpp_dropin_compiles_and_links(tmp_path, name):
    assert os.path.exists(build(tmp_path, name))


def cpp_kets(n, m, seed):
    """The C++ generator of bases_of_uniform_ampl_with_random_phases."""
    u = O.counter_uniform(seed, np.arange(n, dtype=np.uint64)[:, None] * np.uint64(m) + np.arange(m, dtype=np.uint64)[None, :])
    inv = 1.0 / math.sqrt(n)
    k = np.empty((n, m), dtype=np.complex128)
    for i in range(n):
        for r in range(m):
            th = 2.0 * math.pi * float(u[i, r])
            k[i, r] = complex(math.cos(th) * inv, math.sin(th) * inv)
    return k


@pytest.mark.gpu
def test_cpp_dropin_matches_oracle(tmp_path):
    L, N = 16, 64
    out = subprocess.run([build(tmp_path), str(L), str(N)], capture_output=True, text=True, check=True, timeout=120)
    lines = out.stdout.split("\n")
    mu = np.array([float(x.split()[1]) for x in lines if x.startswith("mu ")])
    dos = [(float(x.split()[1]), float(x.split()[2])) for x in lines if x.startswith("dos ")]
    assert any(x == "empty 0" for x in lines)
    assert any(x == "shape_error thrown" for x in lines)
    assert "n must be greater than or equal to 2" in out.stderr

    cell = O.model("honeycomb")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (L, L))
    lat.set_scalar_field_indexed(O.uniform_disorder(0.2, 99))
    lat.gauge_field = O.symmetric_gauge(0.01)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), True)
    Emin, Emax = -3.5 - 1e-8, 3.5 + 1e-8
    Hs = O.rescale(H, (Emax - Emin) / 2, (Emax + Emin) / 2)
    n = Hs.n
    k1 = cpp_kets(n, 4, 0x5EED)
    ref = oracle.cheb_braket(Hs, N, k1, k1)
    assert mu.shape == (N,)
    assert np.max(np.abs(mu - ref)) <= 1e-10 * np.max(np.abs(ref))
    # fft(): REDFT01 of g_n mu_n (scipy's DCT-III is the same transform) over
    # sqrt(1 - E_k^2) at E_k = cos((k + 1/2) pi / N), Lorentz(N, 4)
    from scipy.fft import dct

    fl = np.array([[float(v) for v in x.split()[1:]] for x in lines if x.startswith("fft ")])
    gL = oracle.kernel_lorentz(N, 4.0)
    Ek = np.cos((np.arange(N) + 0.5) * np.pi / N)
    want = dct(gL * ref, type=3) / np.sin((np.arange(N) + 0.5) * np.pi / N)
    assert np.max(np.abs(fl[:, 0] - Ek)) <= 1e-15  # libm and numpy cosines may differ in the last bit
    assert np.max(np.abs(fl[:, 1] - want)) <= 1e-10 * np.max(np.abs(want))
    # Cheb_first_kind_seq_on_bases(scaled_HE, 6, kets): column 1 of each block
    tn = np.array([[float(v) for v in x.split()[1:]] for x in lines if x.startswith("tn ")])
    assert tn.shape == (6 * n, 4)
    got = (tn[:, 2] + 1j * tn[:, 3]).reshape(6, n)
    want = oracle.cheb_on_bases(Hs, 6, k1)[:, :, 1]
    assert np.array_equal(got[0], k1[:, 1])
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    # density_of_states_seq draws the next seed and uses Lorentz(N, 4)
    k2 = cpp_kets(n, 4, 0x5EED + 1)
    mu2 = oracle.cheb_braket(Hs, N, k2, k2)
    g = oracle.kernel_lorentz(N, 4.0)
    want = SO.density_of_states(mu2, g, [e for e, _ in dos], 4)
    got = np.array([d for _, d in dos])
    assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))
    # GreenFunction(N, H, H): E_min, E_max = the extreme eigenvalues of H
    # (dense eigvalsh), a = (E_max - E_min) / 1.99, b = (E_max + E_min) / 2
    rng = [float(v) for x in lines if x.startswith("range ") for v in x.split()[1:]]
    ev = np.linalg.eigvalsh(H.to_dense())
    assert abs(rng[0] - ev[0]) <= 1e-9 and abs(rng[1] - ev[-1]) <= 1e-9
    assert abs(rng[2] - (rng[1] - rng[0]) / 1.99) <= 1e-15 * rng[2] and abs(rng[3] - (rng[1] + rng[0]) / 2) <= 1e-15


def kubo_xy_lattice(L):
    """test_conductivity_Kubo_xy's lattice (snippets.h:745-766)."""
    cell = O.model("honeycomb")
    cell.add_hopping_nn("A", 1, 1.0)
    t = 0.0
    for atom, sgn in (("A", 1.0), ("B", -1.0)):
        for dr in ((0, 1), (1, -1), (-1, 0)):
            cell.add_hopping(atom, atom, dr, complex(-0.0 if sgn < 0 else 0.0, sgn) * t)
    lat = O.PartialLattice(cell, (L, L))
    lat.gauge_field = O.symmetric_gauge(300.0 / 205102.0 * 2 * math.pi)
    return lat


@pytest.mark.gpu
def test_named_entry_kubo_xy_at_full_size(tmp_path):
    """The reference's entry test_conductivity_Kubo_xy at its own sizes
    (graphene 128 x 128, N = 1024, 2000 energies): the DoS printed by the C++
    program against the oracle's moments and reconstruction on the same
    lattice and kets, 1e-9 relative; not finite at e = -1, where
    1/sqrt(1 - e^2) is infinite; the open Hamiltonian's size."""
    L, N, nE = 128, 1024, 2000
    out = subprocess.run([build(tmp_path, "kubo_xy_entry"), str(L), str(N), str(nE)], capture_output=True, text=True,
                         check=True, timeout=300)
    lines = out.stdout.split("\n")
    dos = np.array([[float(v) for v in x.split()[1:]] for x in lines if x.startswith("dos ")])
    assert dos.shape == (nE, 2)
    E = -1.0 + np.arange(nE) * (2.0 / nE)
    assert np.array_equal(dos[:, 0], E)
    lat = kubo_xy_lattice(L)
    H = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), True)
    Emin, Emax = -6 - 1e-8, 6 + 1e-8
    Hs = O.rescale(H, (Emax - Emin) / 2, (Emax + Emin) / 2)
    k = cpp_kets(Hs.n, 4, 0x5EED)
    mu = oracle.cheb_braket(Hs, N, k, k)
    # the oracle's complex division raises at e = -1; the C++ result is not finite there
    want = SO.density_of_states(mu, oracle.kernel_lorentz(N, 4.0), E[1:], 4)
    assert not np.isfinite(dos[0, 1]) and np.isfinite(dos[1:, 1]).all()
    assert np.max(np.abs(dos[1:, 1] - want)) <= 1e-9 * np.max(np.abs(want))
    Ho = lat.get_Hamiltonian_sparse(True)
    opn = [x.split()[1:] for x in lines if x.startswith("open ")][0]
    assert (int(opn[0]), int(opn[1])) == (Ho.n, int(Ho.colptr[-1]))


@pytest.mark.gpu
def test_cpp_conductivity_kubo_bastin_matches_oracle(tmp_path):
    """conductivity_Kubo_Bastin_test (host/kubo.h; snippets.h:306-456) on a
    12 x 12 open square lattice with a Landau gauge (B = 2 pi / 8), site
    disorder and Volume = 2, against the oracle: Gamma from the stored
    blocks, zeta by the reference's live DCT-III / FFT route, the kept nodes
    and the integral of Re zeta by the reference's rule (kubo_sigma_reference)
    and of Im zeta by the same loops; trace_braket_complex of two seeded
    blocks; density_of_states_function = -f(E/a)/(4 pi) with Lorentz(N, 4)."""
    L, N, vol = 12, 64, 2.0
    out = subprocess.run([build(tmp_path, "kubo_bastin"), str(L), str(N)], capture_output=True, text=True,
                         check=True, timeout=120)
    lines = out.stdout.split("\n")
    kub = np.array([[float(v) for v in x.split()[1:]] for x in lines if x.startswith("kubo ")])
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (L, L))
    lat.set_scalar_field_indexed(O.uniform_disorder(1.0, 31))
    lat.gauge_field = O.landau_gauge(2 * math.pi / 8)
    H = lat.get_Hamiltonian_sparse(True)
    Emin, Emax = -5.0 - 1e-8, 5.0 + 1e-8
    Hs = O.rescale(H, (Emax - Emin) / 2, (Emax + Emin) / 2)
    vx = O.commutator_with_diagonal(lat.op_position(0), Hs)
    vy = O.commutator_with_diagonal(lat.op_position(1), Hs)
    n = Hs.n
    kets = cpp_kets(n, 4, 0x5EED)
    g = oracle.kernel_lorentz(N, 6.0)
    TL = oracle.cheb_on_bases(Hs, N, kets) * g[:, None, None]
    w = oracle.cheb_on_bases(vy, 2, kets)[1]
    TR = oracle.cheb_on_bases(Hs, N, w)
    aR = np.stack([oracle.cheb_on_bases(vx, 2, TR[m])[1] for m in range(N)]) * g[:, None, None]
    eps, zeta = SO.kubo_zeta_fft(TL.reshape(N, -1), aR.reshape(N, -1), vol)
    E, re_int = SO.kubo_sigma_reference(eps, zeta)
    cut = N // 24
    zr = np.real(zeta[cut:N - cut])[::-1]
    zi = np.imag(zeta[cut:N - cut])[::-1]
    assert kub.shape == (N - 2 * cut, 5)
    assert np.max(np.abs(kub[:, 0] - E)) <= 1e-15
    scale = np.max(np.abs(zeta[cut:N - cut]))
    assert np.max(np.abs(kub[:, 1] - zr)) <= 1e-9 * scale
    assert np.max(np.abs(kub[:, 2] - zi)) <= 1e-9 * scale
    assert np.max(np.abs(kub[:, 3] - re_int)) <= 1e-9 * np.max(np.abs(re_int))
    # Im zeta: the same outward loops, without the seeded pair (snippets.h:440-449)
    e0 = int(np.searchsorted(E, 0.0, side="right"))
    im_int = np.zeros_like(E)
    for i in range(e0 + 1, len(E)):
        im_int[i] = im_int[i - 1] + (zi[i] + zi[i - 1]) / 2 * (E[i] - E[i - 1])
    for i in range(e0 - 2, -1, -1):
        im_int[i] = im_int[i + 1] - (zi[i] + zi[i + 1]) / 2 * (E[i + 1] - E[i])
    assert np.max(np.abs(kub[:, 4] - im_int)) <= 1e-9 * max(np.max(np.abs(im_int)), scale * 1e-3)
    # trace_braket_complex(a, b) = sum conj(a) b over seeded blocks 77 and 78
    tr = [complex(float(x.split()[1]), float(x.split()[2])) for x in lines if x.startswith("trace ")][0]
    a, b = cpp_kets(n, 3, 77), cpp_kets(n, 3, 78)
    assert abs(tr - np.sum(np.conj(a) * b)) <= 1e-14
    # density_of_states_function on the periodic H, GreenFunction(N, Hp, -5, 5): seed 79
    dosf = np.array([[float(v) for v in x.split()[1:]] for x in lines if x.startswith("dosf ")])
    Hp = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), True)
    Hps = O.rescale(Hp, 5.0, 0.0)
    k3 = cpp_kets(n, 4, 79)
    mu = oracle.cheb_braket(Hps, N, k3, k3)
    gl = oracle.kernel_lorentz(N, 4.0)
    want = []
    for e in dosf[:, 0]:
        al = SO.expn_coef_alpha_seq(complex(e / 5.0, 0.0), N)
        want.append(-1.0 / 4 / math.pi * float(np.sum(gl * np.imag(al) * mu)))
    want = np.array(want)
    assert np.max(np.abs(dosf[:, 1] - want)) <= 1e-9 * np.max(np.abs(want))


def _alpha(z, N):
    """1/(z - x) = sum alpha_n(z) T_n(x) (numerical.h:304-316)."""
    s1 = np.sqrt(1.0 - z * z + 0j)
    s2 = z - 1j * s1
    a = np.empty(N, dtype=np.complex128)
    a[0] = -1j / s1
    sn = -2j / s1
    for i in range(1, N):
        sn = sn * s2
        a[i] = sn
    return a


def _entries(lines, tag, n):
    G = np.zeros((n, n), dtype=np.complex128)
    for x in lines:
        if x.startswith(tag + " "):
            _, i, j, re, im = x.split()
            G[int(i), int(j)] = complex(float(re), float(im))
    return G


@pytest.mark.gpu
def test_cpp_small_system_green_function_and_view(tmp_path):
    """GreenFunction(N, H) and direct_eval (lattice.h:570-592) and
    Cheb_first_kind_seq_view (numerical.h:346-361) through the C++ layer.
    direct_eval = sum_n Im(alpha_n(z~)) T_n(H~) / a: against the oracle's T_n
    (identity kets) summed in order, and -- independently of any Chebyshev
    code -- against the dense resolvent, whose anti-Hermitian part
    (G - G^H) / 2i with G = (z~ - H~)^-1 is that sum once the series has
    converged (Im z~ ~ 0.07 at N = 600: |s2|^N ~ 1e-19).  The view's blocks
    are the oracle's T_i bit for bit."""
    L, N = 6, 600
    out = subprocess.run([build(tmp_path, "green_direct"), str(L), str(N)], capture_output=True, text=True,
                         check=True, timeout=120)
    lines = out.stdout.split("\n")
    assert "Warning: Too big to implement direct eval." in out.stderr
    assert "big_nnz 0" in lines
    n = L * L
    cell = O.model("square")
    cell.add_hopping_nn("A", 1, 1.0)
    lat = O.PartialLattice(cell, (L, L))
    lat.set_scalar_field_indexed(O.uniform_disorder(1.0, 5))
    Hr = lat.get_Hamiltonian_periodic_extend_bare_sparse((1, 1), False)
    lat.gauge_field = O.landau_gauge(2 * math.pi / 6)
    Hc = lat.get_Hamiltonian_sparse(True)
    eye = np.eye(n, dtype=np.complex128)
    for tag, H, z in (("c", Hc, complex(0.3, 0.3)), ("r", Hr, complex(-1.1, 0.4))):
        a, b = [float(v) for x in lines if x.startswith("scale_%s " % tag) for v in x.split()[1:]]
        ev = np.linalg.eigvalsh(H.to_dense())
        assert abs(a - (ev[-1] - ev[0]) / 1.99) <= 1e-9 and abs(b - (ev[-1] + ev[0]) / 2) <= 1e-9
        Hs = O.rescale(H, a, b)
        got = _entries(lines, "g" + tag, n)
        zt = complex((z.real - b) / a, z.imag / a)
        al = _alpha(zt, N)
        T = oracle.cheb_on_bases(Hs, N, eye)
        want = np.zeros((n, n), dtype=np.complex128)
        for k in range(N):
            want = want + al[k].imag * T[k]
        want = want / a
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
        Gz = np.linalg.inv(zt * np.eye(n) - Hs.to_dense())
        anti = (Gz - Gz.conj().T) / 2j / a
        assert np.max(np.abs(got - anti)) <= 1e-10 * np.max(np.abs(anti))
        if tag == "r":
            assert np.all(got.imag == 0)
    # the view: T_0 .. T_5 of the seeded block on the complex H~
    a, b = [float(v) for x in lines if x.startswith("scale_c ") for v in x.split()[1:]]
    Hs = O.rescale(Hc, a, b)
    v = cpp_kets(n, 3, 0x77)
    rows = np.array([[float(t) for t in x.split()[1:]] for x in lines if x.startswith("view ")])
    assert rows.shape == (6 * n, 4)
    assert np.array_equal(rows[:, 0].astype(int), np.repeat(np.arange(6), n))
    got = (rows[:, 2] + 1j * rows[:, 3]).reshape(6, n)
    want = oracle.cheb_on_bases(Hs, 6, v)[:, :, 1]
    assert np.array_equal(got, want)
