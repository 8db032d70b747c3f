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


@pytest.mark.parametrize("name", ["dropin_dos", "kubo_xy_entry"])
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
