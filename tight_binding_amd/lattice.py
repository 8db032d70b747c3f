"""Python mirror of the reference's model layer over libtbhost.so.

UnitCell / PartialLattice keep the reference's names and argument meaning
(unit_cell.h:66-156, lattice.h:22-103); the work is done by the C++ classes
of tight_binding_amd/host through include/tbhost.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .sparse import CSC

GAUGE_NONE, GAUGE_LANDAU, GAUGE_SYMMETRIC = 0, 1, 2
BUILD_OPEN, BUILD_PERIODIC, BUILD_BARE = 0, 1, 2


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _csc_from_handle(h) -> CSC:
    lib = N.host()
    try:
        n = C.c_int32()
        nnz = C.c_int64()
        cplx = C.c_int32()
        N.check_host(lib.tbh_csc_info(h, C.byref(n), C.byref(nnz), C.byref(cplx)))
        colptr = np.zeros(n.value + 1, dtype=np.int32)
        rowidx = np.zeros(nnz.value, dtype=np.int32)
        vals = np.zeros(nnz.value * (2 if cplx.value else 1), dtype=np.float64)
        N.check_host(lib.tbh_csc_copy(h, _ip(colptr), _ip(rowidx), _dp(vals)))
        if cplx.value:
            vals = vals.view(np.complex128)
        return CSC(n.value, colptr, rowidx, vals)
    finally:
        lib.tbh_csc_destroy(h)


def _csc_handle(m):
    lib = N.host()
    cp = np.ascontiguousarray(m.colptr, dtype=np.int32)
    ri = np.ascontiguousarray(m.rowidx, dtype=np.int32)
    cplx = np.iscomplexobj(m.values)
    v = np.ascontiguousarray(m.values, dtype=np.complex128 if cplx else np.float64)
    vv = v.view(np.float64) if cplx else v
    h = C.c_void_p()
    N.check_host(lib.tbh_csc_create(int(m.n), int(ri.size), _ip(cp), _ip(ri), _dp(vv), int(cplx), C.byref(h)))
    return h


def rescale(m, a, b) -> CSC:
    """H~ = (H - b I)/a over the union pattern of H and I (lattice.h:550-559)."""
    lib = N.host()
    h = _csc_handle(m)
    try:
        h2 = C.c_void_p()
        N.check_host(lib.tbh_csc_rescale(h, float(a), float(b), C.byref(h2)))
    finally:
        lib.tbh_csc_destroy(h)
    return _csc_from_handle(h2)


def commutator_diag(X, M) -> CSC:
    """X M - M X for a diagonal real X: x_i M_ij - M_ij x_j on M's pattern --
    the velocity operator v = [x, H~] of snippets.h:316-317."""
    lib = N.host()
    hx, hm = _csc_handle(X), _csc_handle(M)
    try:
        h2 = C.c_void_p()
        N.check_host(lib.tbh_csc_commutator_diag(hx, hm, C.byref(h2)))
    finally:
        lib.tbh_csc_destroy(hx)
        lib.tbh_csc_destroy(hm)
    return _csc_from_handle(h2)


class UnitCell:
    def __init__(self, dim, lattice_vectors, atoms, _handle=None):
        lib = N.host()
        if _handle is not None:
            self._h = _handle
        else:
            lv = np.ascontiguousarray(np.asarray(lattice_vectors, dtype=np.float64).reshape(dim, 3))
            names = [a[0].encode() for a in atoms]
            pos = np.ascontiguousarray(np.asarray([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3))
            arr = (C.c_char_p * len(names))(*names)
            h = C.c_void_p()
            N.check_host(lib.tbh_cell_create(dim, _dp(lv), len(names), arr, _dp(pos), C.byref(h)))
            self._h = h
        d = C.c_int32()
        N.check_host(lib.tbh_cell_dim(self._h, C.byref(d)))
        self.dim = d.value

    @classmethod
    def model(cls, name: str) -> "UnitCell":
        h = C.c_void_p()
        N.check_host(N.host().tbh_cell_model(name.encode(), C.byref(h)))
        return cls(0, None, None, _handle=h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._host is not None:
            N._host.tbh_cell_destroy(h)
            self._h = None

    def add_hopping(self, frm: str, to: str, lattice_vector, t):
        t = complex(t)
        lv = np.ascontiguousarray(np.asarray(lattice_vector, dtype=np.int32).reshape(self.dim))
        N.check_host(N.host().tbh_cell_add_hopping(self._h, frm.encode(), to.encode(), _ip(lv), t.real, t.imag))
        return self

    def add_hopping_nn(self, frm: str, k: int, t):
        t = complex(t)
        N.check_host(N.host().tbh_cell_add_hopping_nn(self._h, frm.encode(), int(k), t.real, t.imag))
        return self

    def hoppings(self):
        """[((from, to, lattice_vector), t)] in table order."""
        lib = N.host()
        n = C.c_int32()
        N.check_host(lib.tbh_cell_hopping_count(self._h, C.byref(n)))
        k = n.value
        f = np.zeros(max(k, 1), dtype=np.int32)
        to = np.zeros(max(k, 1), dtype=np.int32)
        lv = np.zeros(max(k, 1) * self.dim, dtype=np.int32)
        t = np.zeros(2 * max(k, 1), dtype=np.float64)
        N.check_host(lib.tbh_cell_hoppings(self._h, _ip(f), _ip(to), _ip(lv), _dp(t)))
        return [
            ((int(f[i]), int(to[i]), tuple(int(x) for x in lv[i * self.dim:(i + 1) * self.dim])),
             complex(t[2 * i], t[2 * i + 1]))
            for i in range(k)
        ]


class PartialLattice:
    def __init__(self, cell: UnitCell, span):
        self.dim = cell.dim
        sp = np.ascontiguousarray(np.asarray(span, dtype=np.int32).reshape(self.dim))
        h = C.c_void_p()
        N.check_host(N.host().tbh_lattice_create(cell._h, _ip(sp), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._host is not None:
            N._host.tbh_lattice_destroy(h)
            self._h = None

    def get_total_number(self) -> int:
        n = C.c_int32()
        N.check_host(N.host().tbh_lattice_total_number(self._h, C.byref(n)))
        return n.value

    def set_gauge(self, kind: int, B: float):
        N.check_host(N.host().tbh_lattice_set_gauge(self._h, int(kind), float(B)))
        return self

    def set_disorder(self, W: float, seed: int):
        N.check_host(N.host().tbh_lattice_set_disorder(self._h, float(W), int(seed)))
        return self

    def remove_atoms(self, atom: str, cells):
        cells = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, self.dim))
        N.check_host(N.host().tbh_lattice_remove_atoms(self._h, atom.encode(), _ip(cells), cells.shape[0]))
        return self

    def _build(self, builder, extend, is_complex) -> CSC:
        h = C.c_void_p()
        ext = np.ascontiguousarray(np.asarray(extend if extend is not None else [0] * self.dim, dtype=np.int32))
        N.check_host(N.host().tbh_lattice_hamiltonian(self._h, builder, _ip(ext), int(bool(is_complex)), C.byref(h)))
        return _csc_from_handle(h)

    def get_Hamiltonian_sparse(self, is_complex=False) -> CSC:
        return self._build(BUILD_OPEN, None, is_complex)

    def get_Hamiltonian_periodic_extend_bare_sparse(self, extend, is_complex=False) -> CSC:
        return self._build(BUILD_PERIODIC, extend, is_complex)

    def get_Hamiltonian_bare_sparse(self, is_complex=False) -> CSC:
        return self._build(BUILD_BARE, None, is_complex)

    def op_position(self, axis: int) -> CSC:
        h = C.c_void_p()
        N.check_host(N.host().tbh_lattice_op_position(self._h, int(axis), C.byref(h)))
        return _csc_from_handle(h)

    def op_velocity(self, axis: int, is_complex=False) -> CSC:
        h = C.c_void_p()
        N.check_host(N.host().tbh_lattice_op_velocity(self._h, int(axis), int(bool(is_complex)), C.byref(h)))
        return _csc_from_handle(h)

    def hamiltonian_scaled(self, a, b, builder=BUILD_PERIODIC, extend=None, is_complex=False) -> CSC:
        """H~ = (H - bI)/a of the chosen builder, rescaled by the host library
        (lattice.h:550-559)."""
        h = C.c_void_p()
        lib = N.host()
        ext = np.ascontiguousarray(np.asarray(extend if extend is not None else [1] * self.dim, dtype=np.int32))
        N.check_host(lib.tbh_lattice_hamiltonian(self._h, builder, _ip(ext), int(bool(is_complex)), C.byref(h)))
        try:
            h2 = C.c_void_p()
            N.check_host(lib.tbh_csc_rescale(h, float(a), float(b), C.byref(h2)))
        finally:
            lib.tbh_csc_destroy(h)
        return _csc_from_handle(h2)
