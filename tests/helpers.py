"""Shared builders for the tests: the same lattice through the product
(libtbhost) and through the oracle, and closed-form spectra."""
from __future__ import annotations

import math

import numpy as np

from oracle import lattice_oracle as O
from tight_binding_amd import lattice as L

GAUGES = {None: (0, None), "landau": (1, O.landau_gauge), "symmetric": (2, O.symmetric_gauge)}


def lattice_pair(model, span, nn, W=0.0, seed=1, gauge=None, B=0.0, removed=None, explicit=None):
    """(product PartialLattice, oracle PartialLattice) of the same lattice.
    nn: [(atom, k, t)] shells; explicit: [(from, to, dR, t)] hoppings."""
    c = L.UnitCell.model(model)
    oc = O.model(model)
    for a, k, t in nn:
        c.add_hopping_nn(a, k, t)
        oc.add_hopping_nn(a, k, t)
    for f, to, dr, t in explicit or []:
        c.add_hopping(f, to, dr, t)
        oc.add_hopping(f, to, dr, t)
    lat = L.PartialLattice(c, span)
    ol = O.PartialLattice(oc, span)
    if W:
        lat.set_disorder(W, seed)
        ol.set_scalar_field_indexed(O.uniform_disorder(W, seed))
    if gauge:
        kind, fn = GAUGES[gauge]
        lat.set_gauge(kind, B)
        ol.gauge_field = fn(B)
    if removed:
        atom, cells = removed
        lat.remove_atoms(atom, cells)
        ol.remove_atoms(atom, cells)
    return lat, ol


def build_pair(model, span, nn, builder="periodic", extend=None, is_complex=False, W=0.0, seed=1,
               gauge=None, B=0.0, removed=None, explicit=None):
    """Return (product CSC, oracle CSC) of the same Hamiltonian (lattice_pair's
    arguments, then the builder)."""
    lat, ol = lattice_pair(model, span, nn, W, seed, gauge, B, removed, explicit)
    if builder == "periodic":
        ext = extend or [1] * len(span)
        return (lat.get_Hamiltonian_periodic_extend_bare_sparse(ext, is_complex),
                ol.get_Hamiltonian_periodic_extend_bare_sparse(ext, is_complex))
    if builder == "open":
        return lat.get_Hamiltonian_sparse(is_complex), ol.get_Hamiltonian_sparse(is_complex)
    return lat.get_Hamiltonian_bare_sparse(is_complex), ol.get_Hamiltonian_bare_sparse(is_complex)


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    return bool(np.array_equal(a.view(np.uint8), b.view(np.uint8)))


def csc_identical(p, o):
    return (p.n == o.n and np.array_equal(p.colptr, o.colptr) and np.array_equal(p.rowidx, o.rowidx)
            and same_bits(p.values, o.values))


def square_bloch_eigs(L0, L1, t=1.0):
    """Periodic square lattice, nearest-neighbour t: 2t(cos kx + cos ky)."""
    kx = 2 * np.pi * np.arange(L0) / L0
    ky = 2 * np.pi * np.arange(L1) / L1
    return (2 * t * (np.cos(kx)[:, None] + np.cos(ky)[None, :])).ravel()


def honeycomb_bloch_eigs(L0, L1, t=1.0):
    """Periodic honeycomb, A-B hopping t: +-|t| |1 + e^{i k.a1} + e^{i k.a2}|,
    with k.a_j = 2 pi m_j / L_j."""
    p = 2 * np.pi * np.arange(L0) / L0
    q = 2 * np.pi * np.arange(L1) / L1
    f = np.abs(1 + np.exp(1j * p)[:, None] + np.exp(1j * q)[None, :]).ravel()
    return np.concatenate([abs(t) * f, -abs(t) * f])


def cheb_T(n_max, x):
    """T_0..T_{n_max-1} at x by the trigonometric closed form (|x| <= 1)."""
    th = np.arccos(np.clip(x, -1.0, 1.0))
    return np.cos(np.arange(n_max)[:, None] * th[None, :])


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
