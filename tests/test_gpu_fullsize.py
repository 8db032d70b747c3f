"""Parity at the moment counts and sizes of the benchmark workloads.

- N = 2048 moments (the honeycomb workload's count), doubling and one moment
  per step, against the oracle's direct recursion (numerical.h:363-381) on a
  honeycomb the oracle runs in seconds: 1e-10 normwise and the elementwise
  bar of test_gpu_moments.
- The full 2,097,152-site honeycomb at R = 64, N = 2048: mu_0 = R,
  |mu_n| <= mu_0, and the doubling identities against the direct recursion
  (size-independent properties; the oracle would need hours).
- The full 262,144-site Landau square at N = 512, R = 32: sigma_xy from the
  split-bf16 x3 contraction within 1e-4 of fp64 on the same random vectors,
  in the reference's units and in e^2/h.
"""
import numpy as np
import pytest

import oracle
from oracle import lattice_oracle as O
from tests.helpers import build_pair
from tests.test_gpu_moments import assert_moments_close
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu


def test_n2048_honeycomb_doubling_and_plain_match_oracle(gpu_ctx):
    p, _ = build_pair("honeycomb", (40, 40), [("A", 1, -2.7)], W=0.5, seed=31)
    a = 8.35 * 1.01
    Hs = O.rescale(O.CSC(p.n, p.colptr, p.rowidx, p.values), a, 0.0)
    kets = O.random_phase_kets(p.n, 64, 9)
    N = 2048
    ref = oracle.cheb_braket(Hs, N, kets, kets)
    M, K = gpu_ctx.matrix(p), gpu_ctx.block(kets)
    mu_d, st_d, _ = gpu_ctx.moments(M, K, N, a=a)  # raw H, rescale fused in the step
    mu_p, st_p, _ = gpu_ctx.moments(M, K, N, a=a, plain=True)
    assert st_d["doubling"] == 1 and st_d["steps"] == 1024
    assert st_p["doubling"] == 0 and st_p["steps"] == 2047
    assert abs(ref[0] - 64.0) < 1e-9
    assert_moments_close(mu_d, ref)
    assert_moments_close(mu_p, ref)
    print("N=2048 honeycomb 40x40 R=64: doubling %.2e, plain %.2e (normwise vs oracle)"
          % (np.max(np.abs(mu_d - ref)) / 64.0, np.max(np.abs(mu_p - ref)) / 64.0))
    M.close()
    K.close()


def test_c2_full_size_n2048_properties(gpu_ctx):
    """2,097,152 sites, assembled on the device, R = 64, N = 2048."""
    from tight_binding_amd import lattice as L

    cell = L.UnitCell.model("honeycomb")
    cell.add_hopping_nn("A", 1, -2.7)
    lat = L.PartialLattice(cell, (1024, 1024))
    lat.set_disorder(0.5, 2026)
    M, _ = gpu_ctx.assemble(lat, L.BUILD_PERIODIC, (1, 1), False)
    a = (2 * (3 * 2.7 + 0.25)) / (2 - 0.01)
    K = gpu_ctx.random_block(M.n, 64, seed=2027)
    N = 2048
    mu_d, st_d, _ = gpu_ctx.moments(M, K, N, a=a)
    mu_p, st_p, _ = gpu_ctx.moments(M, K, N, a=a, plain=True)
    assert st_d["steps"] == 1024 and st_p["steps"] == 2047
    assert abs(mu_d[0] - 64.0) < 1e-9 and abs(mu_p[0] - 64.0) < 1e-9
    assert np.all(np.abs(mu_d) <= mu_d[0] * (1 + 1e-12))
    assert np.all(np.abs(mu_p) <= mu_p[0] * (1 + 1e-12))
    err = np.max(np.abs(mu_d - mu_p)) / 64.0
    print("C2 full size N=2048: doubling vs direct %.2e normwise" % err)
    assert_moments_close(mu_d, mu_p)
    M.close()
    K.close()
    gpu_ctx.release_work()


def test_c3_full_size_bf16x3_sigma_within_1e4_of_fp64(gpu_ctx):
    """Square 512 x 512 open, Landau flux 1/64, W = 2, N = 512, R = 32, the
    same random vectors for both precisions."""
    from bench import build_c3
    from tight_binding_amd import kubo

    H, X, Y, emax = build_c3(512)
    solver = kubo.KuboSolver(gpu_ctx, H, X, Y, -1.02 * emax, 1.02 * emax, volume=float(H.n), lam=6.0)
    ef = np.linspace(-4.0, 4.0, 201)
    try:
        r64 = solver.solve(512, R=32, seed=11, precision=kpm.GAMMA_FP64, fermi_energies=ef, area=float(H.n))
        rbf = solver.solve(512, R=32, seed=11, precision=kpm.GAMMA_BF16X3, fermi_energies=ef, area=float(H.n))
    finally:
        solver.close()
        gpu_ctx.release_work()  # the two 68.7 GB stacks
    s64, sbf = r64["sigma"], rbf["sigma"]
    err = np.max(np.abs(sbf - s64)) / np.max(np.abs(s64))
    gerr = np.max(np.abs(rbf["gamma"] - r64["gamma"])) / np.max(np.abs(r64["gamma"]))
    perr = (np.max(np.abs(rbf["sigma_xy_e2h"] - r64["sigma_xy_e2h"]))
            / np.max(np.abs(r64["sigma_xy_e2h"])))
    print("C3 full size: sigma bf16x3 vs fp64 %.2e, e^2/h %.2e, Gamma %.2e" % (err, perr, gerr))
    assert np.max(np.abs(s64)) > 0
    assert err <= 1e-4 and perr <= 1e-4
