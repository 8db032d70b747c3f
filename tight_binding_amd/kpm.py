"""Python interface of the HIP moment engine (libtbkpm.so).

Context / Matrix / Block are device-resident objects; Context.moments runs
the fused Chebyshev recursion.  cheb_first_kind_seq_braket is the drop-in of
cuda::Cheb_first_kind_seq_braket (numerics_cuda.cuh:10-13).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .sparse import CSC, tbk_sparse_of

FP64 = N.TBK_FP64
FP32 = N.TBK_FP32


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def device_count() -> int:
    n = C.c_int32()
    N.check_kpm(N.kpm().tbk_device_count(C.byref(n)))
    return n.value


def _as_colmajor_c128(a: np.ndarray) -> np.ndarray:
    """(rows, R) complex array -> column-major complex128 buffer as float64 view."""
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim == 1:
        a = a[:, None]
    return np.asfortranarray(a).ravel(order="F").view(np.float64)


class Matrix:
    def __init__(self, ctx: "Context", m: CSC = None, handle=None, is_complex: bool = False):
        """Upload a CSC matrix, or wrap a resident matrix handle (assembly on
        the device)."""
        self.ctx = ctx
        if handle is None:
            s, keep = tbk_sparse_of(m)
            handle = C.c_void_p()
            N.check_kpm(N.kpm().tbk_matrix_create(ctx._h, C.byref(s), C.byref(handle)))
            del keep
            is_complex = bool(np.iscomplexobj(m.values))
        self._h = handle
        self.is_complex = bool(is_complex)
        inf = self.info()
        self.n, self.nnz = inf["rows"], inf["nnz"]

    def info(self) -> dict:
        rows, nnz, mx, bw = C.c_int32(), C.c_int64(), C.c_int32(), C.c_int32()
        N.check_kpm(N.kpm().tbk_matrix_info(self._h, C.byref(rows), C.byref(nnz), C.byref(mx), C.byref(bw)))
        return {"rows": rows.value, "nnz": nnz.value, "max_row_nnz": mx.value, "bandwidth95": bw.value}

    def download(self):
        """(row pointers, column indices, values) of the resident CSR matrix."""
        rp = np.empty(self.n + 1, dtype=np.int32)
        ci = np.empty(max(self.nnz, 1), dtype=np.int32)
        v = np.empty(max(self.nnz, 1) * (2 if self.is_complex else 1), dtype=np.float64)
        N.check_kpm(N.kpm().tbk_matrix_download(self._h, _ip(rp), _ip(ci), _dp(v)))
        vals = v[: self.nnz * (2 if self.is_complex else 1)]
        return rp, ci[: self.nnz], (vals.view(np.complex128) if self.is_complex else vals)

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.kpm().tbk_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Block:
    def __init__(self, ctx: "Context", handle, rows, R, precision):
        self.ctx, self._h, self.rows, self.R, self.precision = ctx, handle, rows, R, precision

    def download(self) -> np.ndarray:
        out = np.zeros(2 * self.rows * self.R, dtype=np.float64)
        N.check_kpm(N.kpm().tbk_block_download(self._h, _dp(out)))
        return out.view(np.complex128).reshape(self.R, self.rows).T

    def copy_to(self, dst: "Block") -> float:
        """dst = self on the device; returns the copy kernel's time in ms."""
        ms = C.c_float()
        N.check_kpm(N.kpm().tbk_block_copy(self._h, dst._h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.kpm().tbk_block_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One device, one HIP stream.  Not safe for concurrent use."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_context_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.kpm().tbk_context_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def memory(self) -> int:
        b = C.c_int64()
        N.check_kpm(N.kpm().tbk_context_memory(self._h, C.byref(b)))
        return b.value

    def release_work(self):
        """Free the work buffers kept between calls (recursion blocks, Kubo
        stacks, Gamma buffers)."""
        N.check_kpm(N.kpm().tbk_context_release_work(self._h))

    def synchronize(self):
        N.check_kpm(N.kpm().tbk_context_synchronize(self._h))

    def set_tuning(self, slab_cols=0, blocks_per_cu=0, xcd_split=0, l2_budget_bytes=0, stream_nt=0,
                   use_tiled=0, chunk_rows=0, chunk_period=0, row_lane_bytes=0, row_slab_cols=0, kubo_stack_mib=0,
                   gamma_tile_rows=0):
        """Launch-shape choices (0 = default)."""
        t = N.tbk_tuning(int(slab_cols), int(blocks_per_cu), int(xcd_split), int(l2_budget_bytes),
                         int(stream_nt), int(use_tiled), int(chunk_rows), int(chunk_period), int(row_lane_bytes),
                         int(row_slab_cols), int(kubo_stack_mib), int(gamma_tile_rows))
        N.check_kpm(N.kpm().tbk_context_set_tuning(self._h, C.byref(t)))

    def autotune(self, M: "Matrix", kets: "Block", a=1.0, b=0.0, steps=16, plain=False, refresh=False):
        """Search launch shapes for (M, kets) -- the one-moment-per-step
        recursion with plain=True -- or take them from the process's tuning
        cache (unless refresh); installs and returns the best tuning and its
        time per step in ms."""
        t = N.tbk_tuning()
        ms = C.c_double()
        flags = (N.TBK_FLAG_PLAIN if plain else 0) | (N.TBK_AUTOTUNE_REFRESH if refresh else 0)
        N.check_kpm(N.kpm().tbk_autotune_flags(self._h, M._h, float(a), float(b), kets._h, int(steps), flags,
                                               C.byref(t), C.byref(ms)))
        return {k: getattr(t, k) for k, _ in t._fields_}, ms.value

    def get_tuning(self) -> dict:
        t = N.tbk_tuning()
        N.check_kpm(N.kpm().tbk_context_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def matrix(self, m: CSC) -> Matrix:
        return Matrix(self, m)

    def assemble(self, lattice, builder: int = 1, extend=None, is_complex: bool = False) -> Matrix:
        """The lattice's Hamiltonian assembled on the device (tbk_matrix_assemble):
        the matrix of lattice.get_Hamiltonian_sparse (builder 0),
        get_Hamiltonian_periodic_extend_bare_sparse(extend) (1) or
        get_Hamiltonian_bare_sparse (2), bit for bit, never formed on the host.
        Returns (Matrix, wall ms)."""
        d, keep = lattice_desc_of(lattice, builder, extend, is_complex)
        h = C.c_void_p()
        ms = C.c_double()
        N.check_kpm(N.kpm().tbk_matrix_assemble(self._h, C.byref(d), C.byref(h), C.byref(ms)))
        del keep
        return Matrix(self, handle=h, is_complex=is_complex), ms.value

    def rescale(self, M: Matrix, a: float, b: float) -> Matrix:
        """H~ = (M - b)/a over the union pattern of M and I as a new resident
        matrix (tbk_matrix_rescale): the host rescale's entries and bits."""
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_matrix_rescale(self._h, M._h, float(a), float(b), C.byref(h)))
        return Matrix(self, handle=h, is_complex=M.is_complex)

    def velocity(self, M: Matrix, lattice, axis: int) -> Matrix:
        """v = X M - M X (no factor i) for the lattice's position operator
        along axis, on M's pattern (tbk_matrix_velocity): the bits of
        lattice.commutator_diag(lattice.op_position(axis), M)."""
        info = N.tbh_lattice_info()
        N.check_host(N.host().tbh_lattice_info_get(lattice._h, C.byref(info)))
        if info.removed_sites:
            raise N.NativeError(N.TBK_E_UNSUPPORTED, "removed sites: positions use compressed indices")
        d, keep = lattice_desc_of(lattice, 2, None, M.is_complex)
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_matrix_velocity(self._h, M._h, C.byref(d), int(axis), C.byref(h)))
        del keep
        return Matrix(self, handle=h, is_complex=M.is_complex)

    def bandwidth_probe(self, nbytes: int, repeats: int = 3) -> dict:
        """Attainable HBM references on buffers of nbytes (tbk_bandwidth_probe):
        non-temporal copy, two-reads-one-write stream, read-only, in GB/s."""
        b = N.tbk_bandwidth()
        N.check_kpm(N.kpm().tbk_bandwidth_probe(self._h, int(nbytes), int(repeats), C.byref(b)))
        return b.as_dict()

    def spectral_bounds(self, M: Matrix, max_iters: int = 1000, tol: float = 1e-8, seed: int = 1) -> dict:
        """Extreme eigenvalues of the Hermitian matrix M by FP64 Lanczos on the
        device (tbk_spectral_bounds): ritz_min/max, their residual bounds,
        the Gershgorin interval, lower/upper, iters, converged."""
        b = N.tbk_bounds()
        N.check_kpm(N.kpm().tbk_spectral_bounds(self._h, M._h, int(max_iters), float(tol), int(seed), C.byref(b)))
        return b.as_dict()

    def block(self, kets: np.ndarray, precision: int = FP64) -> Block:
        kets = np.asarray(kets)
        if kets.ndim == 1:
            kets = kets[:, None]
        rows, R = kets.shape
        buf = _as_colmajor_c128(kets)
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_block_create(self._h, _dp(buf), rows, R, int(precision), C.byref(h)))
        return Block(self, h, rows, R, precision)

    def zeros_block(self, rows: int, R: int, precision: int = FP64) -> Block:
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_block_create(self._h, None, int(rows), int(R), int(precision), C.byref(h)))
        return Block(self, h, rows, R, precision)

    def random_block(self, rows: int, R: int, seed: int, precision: int = FP64) -> Block:
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_block_random_phase(self._h, rows, R, int(seed), int(precision), C.byref(h)))
        return Block(self, h, rows, R, precision)

    def cheb_on_bases(self, M: Matrix, v: Block, n: int, a: float = 1.0, b: float = 0.0,
                      out: list | None = None):
        """Stored blocks [T_0(H~) v, ..., T_{n-1}(H~) v], H~ = (M - b)/a, all
        resident on the device (tbk_cheb_seq_on_bases; Cheb_first_kind_seq_on_bases,
        numerical.h:334-344).  out: n blocks of v's shape to fill (made here
        when None; out[0] may be v).  Returns (blocks, kernel ms)."""
        n = int(n)
        if out is None:
            out = [self.zeros_block(v.rows, v.R, v.precision) for _ in range(n)]
        if len(out) != n:
            raise ValueError("out must hold n blocks")
        arr = (C.c_void_p * max(n, 1))(*[o._h for o in out])
        ms = C.c_float()
        N.check_kpm(N.kpm().tbk_cheb_seq_on_bases(self._h, M._h, float(a), float(b), v._h, n,
                                                  C.cast(arr, N.PP), C.byref(ms)))
        return out, ms.value

    def cheb_step(self, M: Matrix, x: Block, y: Block, z: Block | None = None, a: float = 1.0, b: float = 0.0):
        """One step on resident blocks (tbk_cheb_step): y = H~ x when z is None,
        else y = 2 H~ x - z (y may be z), H~ = (M - b)/a."""
        N.check_kpm(N.kpm().tbk_cheb_step(self._h, M._h, float(a), float(b), x._h,
                                          z._h if z is not None else None, y._h))
        return y

    def cheb_expand(self, M: Matrix, v: Block, coef, a: float = 1.0, b: float = 0.0, out: Block | None = None):
        """out = sum_k coef[k] T_k(H~) v over the given (complex) coefficients
        (tbk_cheb_expand).  Returns (out, kernel ms)."""
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
        n = c.size // 2
        if out is None:
            out = self.zeros_block(v.rows, v.R, v.precision)
        ms = C.c_float()
        N.check_kpm(N.kpm().tbk_cheb_expand(self._h, M._h, float(a), float(b), v._h, int(n), _dp(c), out._h,
                                            C.byref(ms)))
        return out, ms.value

    def moments(self, M: Matrix, kets: Block, n: int, bras: Block | None = None, a: float = 1.0,
                b: float = 0.0, plain: bool = False, time_launches: bool = False, want_last: bool = False,
                row_kernel: bool = False):
        """mu[i] = Re <bras| T_i((M - b)/a) |kets>, i < n.  Returns (mu, stats, last)."""
        mu = np.zeros(int(n), dtype=np.float64)
        last = np.zeros(2 * kets.rows * kets.R, dtype=np.float64) if want_last else None
        st = N.tbk_stats()
        flags = ((N.TBK_FLAG_PLAIN if plain else 0) | (N.TBK_FLAG_TIME_LAUNCHES if time_launches else 0)
                 | (N.TBK_FLAG_ROW_KERNEL if row_kernel else 0))
        N.check_kpm(
            N.kpm().tbk_moments(
                self._h, M._h, float(a), float(b), bras._h if bras is not None else None, kets._h, int(n), flags,
                _dp(mu), _dp(last) if last is not None else None, C.byref(st),
            )
        )
        if last is not None:
            last = last.view(np.complex128).reshape(kets.R, kets.rows).T
        return mu, st.as_dict(), last


GAMMA_FP64 = N.TBK_GAMMA_FP64
GAMMA_BF16X3 = N.TBK_GAMMA_BF16X3
# the split-bf16x3 contraction over stacks from an fp32 recursion (paired
# stacks: complex H, v_x with H's pattern, n >= 256); stats["recursion_fp32"]
GAMMA_BF16X3_FP32 = N.TBK_GAMMA_BF16X3_FP32


def tuning_cache_save(path: str):
    """Write the process's autotuner cache (tbk_tuning_cache_save)."""
    N.check_kpm(N.kpm().tbk_tuning_cache_save(str(path).encode()))


def tuning_cache_load(path: str) -> int:
    """Add a saved autotuner cache to the process's; returns its entry count."""
    k = C.c_int32()
    N.check_kpm(N.kpm().tbk_tuning_cache_load(str(path).encode(), C.byref(k)))
    return k.value


def tuning_cache_clear():
    N.check_kpm(N.kpm().tbk_tuning_cache_clear())


def kubo_gamma(ctx: Context, H: Matrix, vx: Matrix, vy: Matrix, kets: Block, n: int, a: float = 1.0,
               b: float = 0.0, precision: int = GAMMA_FP64, g=None):
    """Gamma[n, m] = sum conj(g_n T_n(H~)|r>) * (g_m v_x T_m(H~) v_y|r>) over sites
    and vectors (tbkpm.h tbk_kubo_gamma).  Returns (Gamma (n, n) complex, stats)."""
    out = np.zeros(2 * n * n, dtype=np.float64)
    gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
    st = N.tbk_kubo_stats()
    N.check_kpm(N.kpm().tbk_kubo_gamma(ctx._h, H._h, float(a), float(b), vx._h, vy._h, kets._h, int(n),
                                       int(precision), _dp(gg) if gg is not None else None, _dp(out),
                                       C.byref(st)))
    return out.view(np.complex128).reshape(n, n), st.as_dict()


def stack_contract(ctx: Context, A: np.ndarray, B: np.ndarray, precision: int = GAMMA_FP64):
    """Gamma = conj(A) @ B.T for A, B of shape (N, K) complex, on matrix cores.
    Returns (Gamma, milliseconds)."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    assert A.shape == B.shape and A.ndim == 2
    n, k = A.shape
    out = np.zeros(2 * n * n, dtype=np.float64)
    ms = C.c_double()
    N.check_kpm(N.kpm().tbk_stack_contract(ctx._h, _dp(A.view(np.float64)), _dp(B.view(np.float64)), n, k,
                                           int(precision), _dp(out), C.byref(ms)))
    return out.view(np.complex128).reshape(n, n), ms.value


def lattice_desc_of(lattice, builder: int = 1, extend=None, is_complex: bool = False):
    """tbk_lattice_desc of a lattice.PartialLattice (its hopping table, the
    host's phase table when a gauge field is set, the site-keyed disorder);
    the returned tuple keeps the arrays alive while the descriptor is used."""
    info = N.tbh_lattice_info()
    N.check_host(N.host().tbh_lattice_info_get(lattice._h, C.byref(info)))
    if info.position_field:
        raise N.NativeError(N.TBK_E_UNSUPPORTED, "a position-dependent scalar field cannot be assembled on the device")
    if builder == 0 and info.removed_sites:
        raise N.NativeError(N.TBK_E_UNSUPPORTED, "removed sites cannot be assembled on the device")
    nh, dim = info.n_hops, info.dim
    fr = np.zeros(max(nh, 1), dtype=np.int32)
    to = np.zeros(max(nh, 1), dtype=np.int32)
    lv = np.zeros(max(nh * dim, 1), dtype=np.int32)
    t = np.zeros(max(2 * nh, 2), dtype=np.float64)
    if nh:
        N.check_host(N.host().tbh_lattice_hoppings(lattice._h, _ip(fr), _ip(to), _ip(lv), _dp(t)))
    phase = None
    if info.gauge != 0 and builder != 2:
        phase = np.zeros(max(2 * info.cell_number * nh, 2), dtype=np.float64)
        N.check_host(N.host().tbh_lattice_phase_table(lattice._h, _dp(phase)))
    d = N.tbk_lattice_desc()
    d.dim = dim
    for k in range(3):
        d.span[k] = info.span[k]
        d.extend[k] = int(extend[k]) if (extend is not None and k < dim) else 0
    d.n_atoms, d.n_hops, d.builder, d.is_complex = info.n_atoms, nh, int(builder), int(bool(is_complex))
    d.hop_from, d.hop_to, d.hop_lv, d.hop_t = _ip(fr), _ip(to), _ip(lv), _dp(t)
    d.phase = _dp(phase) if phase is not None else None
    d.has_disorder, d.disorder_W, d.disorder_seed = info.has_disorder, info.disorder_W, info.disorder_seed
    bases = np.zeros(9, dtype=np.float64)
    apos = np.zeros(3 * max(info.n_atoms, 1), dtype=np.float64)
    N.check_host(N.host().tbh_lattice_geometry(lattice._h, _dp(bases), _dp(apos)))
    for k in range(9):
        d.bases[k] = bases[k]
    d.atom_pos = _dp(apos)
    return d, (fr, to, lv, t, phase, apos)


def reconstruct(ctx, mu, g, e, scale: float = 1.0):
    """scale * [g_0 mu_0 + 2 sum_n g_n mu_n cos(n theta)] / sin(theta) at the
    rescaled energies e = cos(theta), on the device (tbk_cheb_reconstruct);
    e an int: at the e nodes theta_k = (k + 1/2) pi / e.  ctx None: the
    process-wide context.  Returns (values, kernel ms)."""
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
    if gg is not None:
        assert gg.size >= mu.size
    if isinstance(e, int):  # node mode: theta_k = (k + 1/2) pi / e
        n_e, ep = e, None
    else:
        e = np.ascontiguousarray(np.atleast_1d(e), dtype=np.float64)
        n_e, ep = e.size, _dp(e)
    out = np.empty(n_e, dtype=np.float64)
    ms = C.c_double()
    N.check_kpm(N.kpm().tbk_cheb_reconstruct(None if ctx is None else ctx._h, _dp(mu),
                                             None if gg is None else _dp(gg), int(mu.size), ep, int(n_e),
                                             float(scale), _dp(out), C.byref(ms)))
    return out, ms.value


def density_of_states(ctx, mu, g, e, R):
    """DoS(e) = -f(e)/(R pi) of density_of_states_seq (functions.h:27-30) on
    the device; e in rescaled units."""
    return reconstruct(ctx, mu, g, e, 1.0 / (R * np.pi))[0]


def green_fft(ctx, mu, g, n: int = 1):
    """GreenFunctionImag::fft(n) (lattice.h:625-640): the nodes
    E_k = cos((k + 1/2) pi / (n N)) and REDFT01 of g mu zero-padded to n N,
    over sin(theta_k)."""
    m = n * np.asarray(mu).size
    E = np.cos((np.arange(m) + 0.5) * np.pi / m)
    return E, reconstruct(ctx, mu, g, int(m), 1.0)[0]


def spectral_range(m: CSC):
    """(E_min, E_max) of a Hermitian CSC matrix, the one-shot form of the
    reference's matrix_spectral_range_sparse (numerical.h:23-46) without its
    5000-row limit (tbk_spectral_range, process-wide context on device 0)."""
    s, keep = tbk_sparse_of(m)
    lo, hi = C.c_double(), C.c_double()
    N.check_kpm(N.kpm().tbk_spectral_range(C.byref(s), C.byref(lo), C.byref(hi)))
    del keep
    return lo.value, hi.value


def rescale_from_range(e_min: float, e_max: float, epsilon: float = 0.01):
    """a, b of GreenFunction(N, HE, ref_HE, epsilon) (lattice.h:560-569):
    a = (E_max - E_min) / (2 - epsilon), b = (E_max + E_min) / 2."""
    return (e_max - e_min) / (2.0 - epsilon), (e_max + e_min) / 2.0


def kubo_zeta(ctx: Context, gamma: np.ndarray, e, volume: float = 1.0):
    """Kubo-Bastin integrand zeta(e_j) = sum Gamma_nm g_mn(e_j) / (1 - e_j^2)^2 /
    volume at rescaled energies |e_j| < 1, on the device (tbk_kubo_zeta).
    Returns (zeta complex array, kernel milliseconds)."""
    G = np.ascontiguousarray(gamma, dtype=np.complex128)
    n = G.shape[0]
    assert G.shape == (n, n)
    e = np.ascontiguousarray(np.atleast_1d(e), dtype=np.float64)
    out = np.zeros(2 * e.size, dtype=np.float64)
    ms = C.c_double()
    N.check_kpm(N.kpm().tbk_kubo_zeta(ctx._h, _dp(G.view(np.float64)), n, _dp(e), int(e.size), float(volume),
                                      _dp(out), C.byref(ms)))
    return out.view(np.complex128), ms.value


class Batch:
    """Independent moment problems run as one block-diagonal recursion
    (tbk_batch_create): mats[p] with rescaling a[p], b[p], n_moments[p]
    moments, R device-generated random-phase vectors from seeds[p].  All
    problems share R, the precision and real-or-complex values."""

    def __init__(self, ctx: Context, mats, a, b, n_moments, R: int, precision: int = FP64, seeds=None):
        n = len(mats)
        arr = (N.tbk_sparse * n)()
        keep = []
        for i, m in enumerate(mats):
            s_, k_ = tbk_sparse_of(m)
            arr[i] = s_
            keep.append(k_)
        self.n_moments = np.ascontiguousarray(n_moments, dtype=np.int32)
        assert self.n_moments.size == n
        aa = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
        bb = np.ascontiguousarray(np.broadcast_to(np.asarray(b, dtype=np.float64), (n,)))
        sd = np.ascontiguousarray(seeds if seeds is not None else np.arange(n), dtype=np.uint64)
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_batch_create(ctx._h, n, arr, _dp(aa), _dp(bb),
                                             self.n_moments.ctypes.data_as(C.POINTER(C.c_int32)), int(R),
                                             int(precision), sd.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(h)))
        del keep
        self.ctx, self._h, self.R, self.precision = ctx, h, int(R), int(precision)
        self.offsets = np.concatenate([[0], np.cumsum(self.n_moments)])

    def run(self, graphs: bool = False, persistent: bool = False, synced: bool = False):
        """Returns (list of moment arrays in problem order, stats).  graphs:
        capture each chunk of 64 steps in a hipGraph; persistent: one
        persistent launch per chunk (problems on XCD queues); synced: one
        persistent launch per chunk over the whole GPU, grid barriers between
        steps (the per-step launches' work and moments)."""
        mu = np.zeros(int(self.offsets[-1]), dtype=np.float64)
        st = N.tbk_batch_stats()
        flags = ((N.TBK_BATCH_GRAPHS if graphs else 0) | (N.TBK_BATCH_PERSISTENT if persistent else 0)
                 | (N.TBK_BATCH_SYNCED if synced else 0))
        N.check_kpm(N.kpm().tbk_batch_run(self._h, flags, _dp(mu), C.byref(st)))
        return [mu[self.offsets[i]:self.offsets[i + 1]] for i in range(self.n_moments.size)], st.as_dict()

    def set_shape(self, cpl=1, blocks_per_cu=4):
        N.check_kpm(N.kpm().tbk_batch_set_shape(self._h, int(cpl), int(blocks_per_cu)))

    def autotune(self, steps=8) -> float:
        ms = C.c_double()
        N.check_kpm(N.kpm().tbk_batch_autotune(self._h, int(steps), C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.kpm().tbk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server on one device (tbk_server_*).
    Requests are copied at submission; results are taken by ticket."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_slots: int = 0, max_kubo: int = 0,
                 chunk_steps: int = 0, cpl: int = 0, max_moments: int = 0, max_kubo_moments: int = 0,
                 max_wait_bypass_ms: float = 0.0, lazy_arena: bool = False):
        cfg = N.tbk_server_config(int(hbm_budget_bytes), int(max_slots), int(max_kubo), int(chunk_steps), int(cpl),
                                  int(max_moments), int(max_kubo_moments), float(max_wait_bypass_ms),
                                  int(bool(lazy_arena)), 0)
        h = C.c_void_p()
        N.check_kpm(N.kpm().tbk_server_create(int(device), C.byref(cfg), C.byref(h)))
        self._h = h
        self._kind = {}

    def _submit(self, req, keep):
        t = C.c_int64()
        N.check_kpm(N.kpm().tbk_server_submit(self._h, C.byref(req), C.byref(t)))
        del keep
        return t.value

    def submit_dos(self, H, n: int, a: float, b: float = 0.0, R: int = 8, seed: int = 0,
                   precision: int = FP64, builder: int = 1, extend=None, is_complex: bool = False) -> int:
        """Moments mu_0..mu_{n-1} of (H - b)/a from R device-generated random-phase vectors.
        H: a CSC matrix, or a lattice.PartialLattice that the server assembles
        on the device at admission (builder, extend, is_complex as for
        Context.assemble).  a = 0: the server estimates the bounds at
        admission (Lanczos range, contracted by 0.01 as
        GreenFunction(N, HE, ref_HE) does); the a, b it used are in
        info()["a"], info()["b"]."""
        if hasattr(H, "get_total_number"):  # a lattice
            d, keep = lattice_desc_of(H, builder, extend, is_complex)
            req = N.tbk_request(N.TBK_REQ_DOS, int(precision), int(n), int(R), int(seed), float(a), float(b),
                                None, None, None, None, C.pointer(d))
            t = self._submit(req, (d, keep))
            self._kind[t] = ("dos", int(n))
            return t
        sH, kH = tbk_sparse_of(H)
        req = N.tbk_request(N.TBK_REQ_DOS, int(precision), int(n), int(R), int(seed), float(a), float(b),
                            C.pointer(sH), None, None, None)
        t = self._submit(req, (sH, kH))
        self._kind[t] = ("dos", int(n))
        return t

    def submit_kubo(self, H: CSC, vx: CSC, vy: CSC, N_: int, a: float = 1.0, b: float = 0.0, R: int = 8,
                    seed: int = 0, precision: int = N.TBK_GAMMA_FP64, g=None) -> int:
        """Kubo-Bastin Gamma (N_ x N_, weighted by g) and zeta at the N_ Chebyshev nodes."""
        sH, kH = tbk_sparse_of(H)
        sx, kx = tbk_sparse_of(vx)
        sy, ky = tbk_sparse_of(vy)
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        req = N.tbk_request(N.TBK_REQ_KUBO, int(precision), int(N_), int(R), int(seed), float(a), float(b),
                            C.pointer(sH), C.pointer(sx), C.pointer(sy), _dp(gg) if gg is not None else None)
        t = self._submit(req, (sH, kH, sx, kx, sy, ky, gg))
        self._kind[t] = ("kubo", int(N_))
        return t

    def submit_kubo_lattice(self, lattice, N_: int, a: float, b: float = 0.0, R: int = 8, seed: int = 0,
                            precision: int = N.TBK_GAMMA_FP64, g=None, builder: int = 0, extend=None,
                            is_complex: bool = True) -> int:
        """sigma_xy of a lattice.PartialLattice formed entirely on the device at
        admission: H assembled (builder, extend, is_complex as for
        Context.assemble), H~ = (H - b)/a, v_x, v_y = [X, H~], [Y, H~]; the
        result as submit_kubo's for those three matrices with a = 1, b = 0."""
        d, keep = lattice_desc_of(lattice, builder, extend, is_complex)
        gg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        req = N.tbk_request(N.TBK_REQ_KUBO, int(precision), int(N_), int(R), int(seed), float(a), float(b),
                            None, None, None, _dp(gg) if gg is not None else None, C.pointer(d))
        t = self._submit(req, (d, keep, gg))
        self._kind[t] = ("kubo", int(N_))
        return t

    def info(self, ticket: int) -> dict:
        inf = N.tbk_request_info()
        N.check_kpm(N.kpm().tbk_server_info(self._h, int(ticket), C.byref(inf)))
        return inf.as_dict()

    def wait(self, ticket: int, timeout: float = -1.0) -> dict:
        inf = N.tbk_request_info()
        N.check_kpm(N.kpm().tbk_server_wait(self._h, int(ticket), float(timeout), C.byref(inf)))
        return inf.as_dict()

    def take(self, ticket: int):
        """(result, info): DoS -> moments (n,); sigma_xy -> (zeta (N,), Gamma (N, N)) complex."""
        kind, n = self._kind.pop(int(ticket))
        size = n if kind == "dos" else 2 * n + 2 * n * n
        out = np.zeros(size, dtype=np.float64)
        inf = N.tbk_request_info()
        N.check_kpm(N.kpm().tbk_server_take(self._h, int(ticket), _dp(out), int(size), C.byref(inf)))
        if kind == "dos":
            return out, inf.as_dict()
        z = out[:2 * n].view(np.complex128)
        G = out[2 * n:].view(np.complex128).reshape(n, n)
        return (z, G), inf.as_dict()

    def cancel(self, ticket: int) -> None:
        """Cancel a queued or running request (tbk_server_cancel): it ends
        failed with TBK_E_CANCELLED (take raises); a running sigma_xy
        request's enqueued work still runs."""
        N.check_kpm(N.kpm().tbk_server_cancel(self._h, int(ticket)))

    def stats(self) -> dict:
        st = N.tbk_server_stats()
        N.check_kpm(N.kpm().tbk_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.kpm().tbk_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cheb_first_kind_seq_on_bases(M: CSC, n: int, v: np.ndarray) -> np.ndarray:
    """Drop-in of Chebyshev::Cheb_first_kind_seq_on_bases for a sparse M (already
    rescaled, CSC): [T_0(M) v, ..., T_{n-1}(M) v] as an array (n, rows, R)
    complex, formed in FP64 on the device."""
    s, keep = tbk_sparse_of(M)
    vb = _as_colmajor_c128(v)
    rows, R = (v.shape[0], 1) if np.ndim(v) == 1 else v.shape
    out = np.zeros(2 * max(int(n), 0) * rows * R, dtype=np.float64)
    N.check_kpm(N.kpm().tbk_cheb_first_kind_seq_on_bases(C.byref(s), int(n), _dp(vb), rows, R, _dp(out)))
    del keep
    return out.view(np.complex128).reshape(int(n), R, rows).transpose(0, 2, 1)


def cheb_first_kind_seq_braket(M: CSC, n: int, bras_T: np.ndarray, kets: np.ndarray) -> np.ndarray:
    """Drop-in of cuda::Cheb_first_kind_seq_braket: M already rescaled (CSC),
    bras_T and kets (rows, R) complex; passing the same array object for both
    selects the two-moments-per-step path."""
    s, keep = tbk_sparse_of(M)
    same = bras_T is kets
    kb = _as_colmajor_c128(kets)
    bb = kb if same else _as_colmajor_c128(bras_T)
    rows, R = (kets.shape[0], 1) if np.ndim(kets) == 1 else kets.shape
    mu = np.zeros(int(n), dtype=np.float64)
    N.check_kpm(N.kpm().tbk_cheb_first_kind_seq_braket(C.byref(s), int(n), _dp(bb), _dp(kb), rows, R, _dp(mu)))
    del keep
    return mu
