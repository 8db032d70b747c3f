"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

Restatements of the reference's algorithm used to check the HIP engine:
  cheb_oracle.c     -> liboracle.so: the CPU moment recursion
                       (numerical.h:363-381), stored-block recursion
                       (numerical.h:334-344) and the damping kernels
  lattice_oracle.py -> the model and Hamiltonian-assembly path
                       (unit_cell.h, lattice.h)
  spectral_oracle.py -> alpha_n, DoS reconstruction, Kubo-Bastin contraction

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, and only as the checker / the timed CPU reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        I32, DP, IP = C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.oracle_cheb_braket.argtypes = [C.c_int, I32, IP, IP, DP, I32, DP, DP, I32, DP, DP, I32]
        L.oracle_cheb_braket.restype = C.c_int
        L.oracle_cheb_steps.argtypes = [C.c_int, I32, IP, IP, DP, I32, DP, I32, DP]
        L.oracle_cheb_steps.restype = C.c_int
        L.oracle_cheb_steps_threads.argtypes = [C.c_int, I32, IP, IP, DP, I32, DP, I32, DP, I32]
        L.oracle_cheb_steps_threads.restype = C.c_int
        L.oracle_cheb_on_bases.argtypes = [C.c_int, I32, IP, IP, DP, I32, DP, I32, DP]
        L.oracle_cheb_on_bases.restype = C.c_int
        L.oracle_trace_braket.argtypes = [DP, DP, I32, I32]
        L.oracle_trace_braket.restype = C.c_double
        L.oracle_kernel_jackson.argtypes = [I32, DP]
        L.oracle_kernel_lorentz.argtypes = [I32, C.c_double, DP]
        L.oracle_kernel_fejer.argtypes = [I32, DP]
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _csc_arrays(colptr, rowidx, values):
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    ri = np.ascontiguousarray(rowidx, dtype=np.int32)
    cplx = np.iscomplexobj(values)
    v = np.ascontiguousarray(values, dtype=np.complex128 if cplx else np.float64)
    vv = v.view(np.float64) if cplx else v
    return cp, ri, vv, cplx


def _colmajor(a):
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim == 1:
        a = a[:, None]
    return np.asfortranarray(a)


def cheb_braket(M, n, bras_T, kets, last_index=None):
    """mu[i] = Re tr(bras^H T_i(M) kets) by the reference's CPU loop.  M is a
    CSC-like object with n, colptr, rowidx, values.  Returns mu, or
    (mu, T_last) when last_index is given."""
    cp, ri, vv, cplx = _csc_arrays(M.colptr, M.rowidx, M.values)
    k = _colmajor(kets)
    b = _colmajor(bras_T)
    rows, R = k.shape
    mu = np.zeros(int(n))
    last = np.zeros((rows, R), dtype=np.complex128, order="F") if last_index is not None else None
    rc = lib().oracle_cheb_braket(
        int(cplx), rows, _ip(cp), _ip(ri), _dp(vv), int(n),
        _dp(b.ravel(order="F").view(np.float64)), _dp(k.ravel(order="F").view(np.float64)), R, _dp(mu),
        _dp(last.ravel(order="F").view(np.float64)) if last is not None else None,
        int(last_index) if last_index is not None else 0,
    )
    if rc != 0:
        raise ValueError("oracle_cheb_braket rejected its arguments")
    return (mu, last) if last is not None else mu


def cheb_steps_threads(M, steps, kets, nthreads=0):
    """mu[0 .. steps+1] of bras = kets by the same recursion with each step
    split over the block columns across OpenMP threads (the all-cores timing
    variant; nthreads <= 0: the OpenMP default)."""
    cp, ri, vv, cplx = _csc_arrays(M.colptr, M.rowidx, M.values)
    k = _colmajor(kets)
    rows, R = k.shape
    mu = np.zeros(int(steps) + 2)
    rc = lib().oracle_cheb_steps_threads(int(cplx), rows, _ip(cp), _ip(ri), _dp(vv), int(steps),
                                         _dp(k.ravel(order="F").view(np.float64)), R, _dp(mu), int(nthreads))
    if rc != 0:
        raise ValueError("oracle_cheb_steps_threads rejected its arguments")
    return mu


def cheb_on_bases(M, n, V):
    """[T_0(M)V, ..., T_{n-1}(M)V] as an array (n, rows, R)."""
    cp, ri, vv, cplx = _csc_arrays(M.colptr, M.rowidx, M.values)
    Vc = _colmajor(V)
    rows, R = Vc.shape
    out = np.zeros(int(n) * rows * R * 2)
    rc = lib().oracle_cheb_on_bases(int(cplx), rows, _ip(cp), _ip(ri), _dp(vv), int(n),
                                    _dp(Vc.ravel(order="F").view(np.float64)), R, _dp(out))
    if rc != 0:
        raise ValueError("oracle_cheb_on_bases rejected its arguments")
    return out.view(np.complex128).reshape(int(n), R, rows).transpose(0, 2, 1)


def trace_braket(bras_T, kets):
    b, k = _colmajor(bras_T), _colmajor(kets)
    return lib().oracle_trace_braket(_dp(b.ravel(order="F").view(np.float64)),
                                      _dp(k.ravel(order="F").view(np.float64)), b.shape[0], b.shape[1])


def kernel_jackson(N):
    g = np.zeros(int(N))
    lib().oracle_kernel_jackson(int(N), _dp(g))
    return g


def kernel_lorentz(N, lam):
    g = np.zeros(int(N))
    lib().oracle_kernel_lorentz(int(N), float(lam), _dp(g))
    return g


def kernel_fejer(N):
    g = np.zeros(int(N))
    lib().oracle_kernel_fejer(int(N), _dp(g))
    return g
