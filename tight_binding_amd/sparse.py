"""Compressed sparse matrices as numpy arrays (CSC as the reference stores
Eigen::SparseMatrix, plus the CSR view the device engine consumes)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class CSC:
    """Column-compressed n x n matrix: colptr[n+1], rowidx[nnz], values[nnz]
    (float64 or complex128)."""

    n: int
    colptr: np.ndarray
    rowidx: np.ndarray
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rowidx.size)

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.values)

    def to_dense(self) -> np.ndarray:
        D = np.zeros((self.n, self.n), dtype=self.values.dtype)
        cols = np.repeat(np.arange(self.n), np.diff(self.colptr))
        D[self.rowidx, cols] = self.values
        return D

    def to_csr(self):
        """(rowptr, colidx, values): the same matrix row-compressed, columns
        ascending within each row; an exact reordering."""
        cols = np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.colptr))
        order = np.lexsort((cols, self.rowidx))
        rp = np.zeros(self.n + 1, dtype=np.int32)
        np.cumsum(np.bincount(self.rowidx, minlength=self.n), out=rp[1:])
        return rp, cols[order].astype(np.int32), self.values[order]

    def as_tbk_sparse(self):
        return tbk_sparse_of(self)


def tbk_sparse_of(m):
    """tbk_sparse descriptor over any column-compressed matrix object with
    n, colptr, rowidx and values; the returned tuple keeps the arrays alive
    while the descriptor is used."""
    cp = np.ascontiguousarray(m.colptr, dtype=np.int32)
    ri = np.ascontiguousarray(m.rowidx, dtype=np.int32)
    cplx = np.iscomplexobj(m.values)
    if cplx:
        v = np.ascontiguousarray(m.values, dtype=np.complex128).view(np.float64)
    else:
        v = np.ascontiguousarray(m.values, dtype=np.float64)
    s = N.tbk_sparse(
        int(m.n),
        int(m.n),
        int(ri.size),
        cp.ctypes.data_as(C.POINTER(C.c_int32)),
        ri.ctypes.data_as(C.POINTER(C.c_int32)),
        v.ctypes.data_as(C.POINTER(C.c_double)),
        1 if cplx else 0,
        N.TBK_CSC,
    )
    return s, (cp, ri, v)
