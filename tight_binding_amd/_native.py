"""ctypes bindings of the package's native libraries.

libtbkpm.so (HIP engine, include/tbkpm.h) and libtbhost.so (host model
layer, include/tbhost.h) are loaded from this directory.  There is no
fallback: if a library is missing the import raises, so a caller can never
silently run something other than the native code.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

TBK_OK = 0
TBK_E_INVALID = -1
TBK_E_DEVICE = -2
TBK_E_OOM = -3
TBK_E_UNSUPPORTED = -4
TBK_E_CANCELLED = -5
TBK_FP64 = 0
TBK_FP32 = 1
TBK_CSC = 0
TBK_CSR = 1
TBK_FLAG_PLAIN = 1
TBK_AUTOTUNE_REFRESH = 8
TBK_FLAG_TIME_LAUNCHES = 2
TBK_FLAG_ROW_KERNEL = 4


class NativeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class tbk_sparse(C.Structure):
    _fields_ = [
        ("rows", C.c_int32),
        ("cols", C.c_int32),
        ("nnz", C.c_int64),
        ("outer", C.POINTER(C.c_int32)),
        ("inner", C.POINTER(C.c_int32)),
        ("values", C.POINTER(C.c_double)),
        ("is_complex", C.c_int32),
        ("layout", C.c_int32),
    ]


class tbk_bandwidth(C.Structure):
    _fields_ = [("copy_GBps", C.c_double), ("stream_2r1w_GBps", C.c_double), ("read_GBps", C.c_double),
                ("bytes", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class tbk_lattice_desc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("span", C.c_int32 * 3),
        ("extend", C.c_int32 * 3),
        ("n_atoms", C.c_int32),
        ("n_hops", C.c_int32),
        ("builder", C.c_int32),
        ("is_complex", C.c_int32),
        ("hop_from", C.POINTER(C.c_int32)),
        ("hop_to", C.POINTER(C.c_int32)),
        ("hop_lv", C.POINTER(C.c_int32)),
        ("hop_t", C.POINTER(C.c_double)),
        ("phase", C.POINTER(C.c_double)),
        ("has_disorder", C.c_int32),
        ("disorder_W", C.c_double),
        ("disorder_seed", C.c_uint64),
        ("bases", C.c_double * 9),
        ("atom_pos", C.POINTER(C.c_double)),
    ]


class tbh_lattice_info(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n_atoms", C.c_int32),
        ("n_hops", C.c_int32),
        ("span", C.c_int32 * 3),
        ("cell_number", C.c_int32),
        ("total_number", C.c_int32),
        ("gauge", C.c_int32),
        ("B", C.c_double),
        ("has_disorder", C.c_int32),
        ("disorder_W", C.c_double),
        ("disorder_seed", C.c_uint64),
        ("removed_sites", C.c_int32),
        ("position_field", C.c_int32),
    ]


class tbk_stats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("steps", C.c_int64),
        ("moments", C.c_int64),
        ("bytes_per_step", C.c_double),
        ("bytes_total", C.c_double),
        ("doubling", C.c_int32),
        ("blocks", C.c_int32),
        ("rows_per_tile", C.c_int32),
        ("padded_R", C.c_int32),
        ("kernel_kind", C.c_int32),
        ("slab_cols", C.c_int32),
        ("chunk_rows", C.c_int32),
        ("chunk_period", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class tbk_tuning(C.Structure):
    _fields_ = [
        ("slab_cols", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("xcd_split", C.c_int32),
        ("l2_budget_bytes", C.c_int32),
        ("stream_nt", C.c_int32),
        ("use_tiled", C.c_int32),
        ("chunk_rows", C.c_int32),
        ("chunk_period", C.c_int32),
        ("row_lane_bytes", C.c_int32),
        ("row_slab_cols", C.c_int32),
        ("kubo_stack_mib", C.c_int32),
        ("gamma_tile_rows", C.c_int32),
    ]


class tbk_kubo_stats(C.Structure):
    _fields_ = [
        ("recursion_ms", C.c_double),
        ("contraction_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("flops", C.c_double),
        ("stack_bytes", C.c_double),
        ("nchunks", C.c_int32),
        ("vector_groups", C.c_int32),
        ("recursion_bytes", C.c_double),
        ("row_ranges", C.c_int32),
        ("recursion_fp32", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class tbk_bounds(C.Structure):
    _fields_ = [
        ("ritz_min", C.c_double),
        ("ritz_max", C.c_double),
        ("resid_min", C.c_double),
        ("resid_max", C.c_double),
        ("gersh_min", C.c_double),
        ("gersh_max", C.c_double),
        ("lower", C.c_double),
        ("upper", C.c_double),
        ("iters", C.c_int32),
        ("converged", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class tbk_batch_stats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("launches", C.c_int64),
        ("graphs", C.c_int64),
        ("bytes_total", C.c_double),
        ("rows_total", C.c_int64),
        ("workgroups", C.c_int64),
        ("cpl", C.c_int32),
        ("blocks_per_cu", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TBK_BATCH_GRAPHS = 1
TBK_BATCH_PERSISTENT = 2
TBK_BATCH_SYNCED = 4

TBK_REQ_DOS, TBK_REQ_KUBO = 0, 1
TBK_REQ_QUEUED, TBK_REQ_RUNNING, TBK_REQ_DONE, TBK_REQ_FAILED = 0, 1, 2, 3


class tbk_server_config(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_int64),
        ("max_slots", C.c_int32),
        ("max_kubo", C.c_int32),
        ("chunk_steps", C.c_int32),
        ("cpl", C.c_int32),
        ("max_moments", C.c_int32),
        ("max_kubo_moments", C.c_int32),
        ("max_wait_bypass_ms", C.c_double),
        ("lazy_arena", C.c_int32),
        ("reserved", C.c_int32),
    ]


class tbk_request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("precision", C.c_int32),
        ("n_moments", C.c_int32),
        ("R", C.c_int32),
        ("seed", C.c_uint64),
        ("a", C.c_double),
        ("b", C.c_double),
        ("H", C.POINTER(tbk_sparse)),
        ("vx", C.POINTER(tbk_sparse)),
        ("vy", C.POINTER(tbk_sparse)),
        ("g", C.POINTER(C.c_double)),
        ("lattice", C.POINTER(tbk_lattice_desc)),
    ]


class tbk_request_info(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("error_code", C.c_int32),
        ("queue_ms", C.c_double),
        ("run_ms", C.c_double),
        ("latency_ms", C.c_double),
        ("footprint_bytes", C.c_int64),
        ("result_doubles", C.c_int64),
        ("a", C.c_double),
        ("b", C.c_double),
        ("vector_groups", C.c_int32),
        ("row_ranges", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class tbk_server_stats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "failed", "queued", "running_dos",
                                          "running_kubo", "kubo_done", "chunks", "used_bytes", "budget_bytes",
                                          "arena_bytes", "arena_slabs", "cancelled")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

TBK_GAMMA_FP64 = 0
TBK_GAMMA_BF16X3 = 1
TBK_GAMMA_BF16X3_FP32 = 2


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise ImportError(
            "native library %s is not built; run __graft_entry__.build() or "
            "`make -C tight_binding_amd`" % path
        )
    return C.CDLL(path)


_kpm = None
_host = None

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I32 = C.c_int32
I64 = C.c_int64
U64 = C.c_uint64
D = C.c_double
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int32)

KPM_SIGNATURES = {
    "tbk_version": ([], C.c_char_p),
    "tbk_last_error": ([], C.c_char_p),
    "tbk_device_count": ([IP], I32),
    "tbk_cheb_first_kind_seq_braket": ([C.POINTER(tbk_sparse), I32, DP, DP, I32, I32, DP], I32),
    "tbk_cheb_first_kind_seq_on_bases": ([C.POINTER(tbk_sparse), I32, DP, I32, I32, DP], I32),
    "tbk_cheb_seq_on_bases": ([P, P, D, D, P, I32, PP, C.POINTER(C.c_float)], I32),
    "tbk_cheb_step": ([P, P, D, D, P, P, P], I32),
    "tbk_cheb_expand": ([P, P, D, D, P, I32, DP, P, C.POINTER(C.c_float)], I32),
    "tbk_context_create": ([I32, PP], I32),
    "tbk_context_destroy": ([P], I32),
    "tbk_context_synchronize": ([P], I32),
    "tbk_context_memory": ([P, C.POINTER(I64)], I32),
    "tbk_context_release_work": ([P], I32),
    "tbk_context_set_tuning": ([P, C.POINTER(tbk_tuning)], I32),
    "tbk_context_get_tuning": ([P, C.POINTER(tbk_tuning)], I32),
    "tbk_matrix_create": ([P, C.POINTER(tbk_sparse), PP], I32),
    "tbk_matrix_destroy": ([P], I32),
    "tbk_matrix_info": ([P, IP, C.POINTER(I64), IP, IP], I32),
    "tbk_block_create": ([P, DP, I32, I32, I32, PP], I32),
    "tbk_block_random_phase": ([P, I32, I32, U64, I32, PP], I32),
    "tbk_block_download": ([P, DP], I32),
    "tbk_block_copy": ([P, P, C.POINTER(C.c_float)], I32),
    "tbk_block_destroy": ([P], I32),
    "tbk_moments": ([P, P, D, D, P, P, I32, I32, DP, DP, C.POINTER(tbk_stats)], I32),
    "tbk_moments_noalloc": ([P, P, D, D, P, P, I32, I32, P, P, P, P], I32),
    "tbk_block_device_ptr": ([P, C.POINTER(C.c_void_p)], I32),
    "tbk_autotune": ([P, P, D, D, P, I32, C.POINTER(tbk_tuning), DP], I32),
    "tbk_autotune_flags": ([P, P, D, D, P, I32, I32, C.POINTER(tbk_tuning), DP], I32),
    "tbk_tuning_cache_save": ([C.c_char_p], I32),
    "tbk_tuning_cache_load": ([C.c_char_p, C.POINTER(C.c_int32)], I32),
    "tbk_tuning_cache_clear": ([], I32),
    "tbk_kubo_gamma": ([P, P, D, D, P, P, P, I32, I32, DP, DP, C.POINTER(tbk_kubo_stats)], I32),
    "tbk_stack_contract": ([P, DP, DP, I32, I64, I32, DP, DP], I32),
    "tbk_kubo_zeta": ([P, DP, I32, DP, I32, D, DP, DP], I32),
    "tbk_bandwidth_probe": ([P, I64, I32, C.POINTER(tbk_bandwidth)], I32),
    "tbk_matrix_assemble": ([P, C.POINTER(tbk_lattice_desc), PP, DP], I32),
    "tbk_matrix_rescale": ([P, P, D, D, PP], I32),
    "tbk_matrix_velocity": ([P, P, C.POINTER(tbk_lattice_desc), I32, PP], I32),
    "tbk_matrix_download": ([P, IP, IP, DP], I32),
    "tbk_cheb_reconstruct": ([P, DP, DP, I32, DP, I32, D, DP, DP], I32),
    "tbk_spectral_bounds": ([P, P, I32, D, U64, C.POINTER(tbk_bounds)], I32),
    "tbk_spectral_range": ([C.POINTER(tbk_sparse), DP, DP], I32),
    "tbk_batch_create": ([P, I32, C.POINTER(tbk_sparse), DP, DP, IP, I32, I32, C.POINTER(U64), PP], I32),
    "tbk_batch_run": ([P, I32, DP, C.POINTER(tbk_batch_stats)], I32),
    "tbk_batch_set_shape": ([P, I32, I32], I32),
    "tbk_batch_autotune": ([P, I32, DP], I32),
    "tbk_batch_destroy": ([P], I32),
    "tbk_server_create": ([I32, C.POINTER(tbk_server_config), PP], I32),
    "tbk_server_submit": ([P, C.POINTER(tbk_request), C.POINTER(I64)], I32),
    "tbk_server_info": ([P, I64, C.POINTER(tbk_request_info)], I32),
    "tbk_server_wait": ([P, I64, D, C.POINTER(tbk_request_info)], I32),
    "tbk_server_take": ([P, I64, DP, I64, C.POINTER(tbk_request_info)], I32),
    "tbk_server_get_stats": ([P, C.POINTER(tbk_server_stats)], I32),
    "tbk_server_destroy": ([P], I32),
    "tbk_server_cancel": ([P, I64], I32),
}

HOST_SIGNATURES = {
    "tbh_last_error": ([], C.c_char_p),
    "tbh_cell_create": ([I32, DP, I32, C.POINTER(C.c_char_p), DP, PP], I32),
    "tbh_cell_model": ([C.c_char_p, PP], I32),
    "tbh_cell_destroy": ([P], I32),
    "tbh_cell_dim": ([P, IP], I32),
    "tbh_cell_add_hopping": ([P, C.c_char_p, C.c_char_p, IP, D, D], I32),
    "tbh_cell_add_hopping_nn": ([P, C.c_char_p, U64, D, D], I32),
    "tbh_cell_hopping_count": ([P, IP], I32),
    "tbh_cell_hoppings": ([P, IP, IP, IP, DP], I32),
    "tbh_lattice_create": ([P, IP, PP], I32),
    "tbh_lattice_destroy": ([P], I32),
    "tbh_lattice_total_number": ([P, IP], I32),
    "tbh_lattice_set_gauge": ([P, I32, D], I32),
    "tbh_lattice_set_disorder": ([P, D, U64], I32),
    "tbh_lattice_remove_atoms": ([P, C.c_char_p, IP, I32], I32),
    "tbh_lattice_info_get": ([P, C.POINTER(tbh_lattice_info)], I32),
    "tbh_lattice_hoppings": ([P, IP, IP, IP, DP], I32),
    "tbh_lattice_phase_table": ([P, DP], I32),
    "tbh_lattice_geometry": ([P, DP, DP], I32),
    "tbh_lattice_hamiltonian": ([P, I32, IP, I32, PP], I32),
    "tbh_lattice_op_position": ([P, I32, PP], I32),
    "tbh_lattice_op_velocity": ([P, I32, I32, PP], I32),
    "tbh_csc_create": ([I32, I64, IP, IP, DP, I32, PP], I32),
    "tbh_csc_rescale": ([P, D, D, PP], I32),
    "tbh_csc_commutator_diag": ([P, P, PP], I32),
    "tbh_csc_info": ([P, IP, C.POINTER(I64), IP], I32),
    "tbh_csc_copy": ([P, IP, IP, DP], I32),
    "tbh_csc_copy_csr": ([P, IP, IP, DP], I32),
    "tbh_csc_destroy": ([P], I32),
}


def _bind(lib, sigs):
    for name, (args, res) in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def kpm():
    global _kpm
    if _kpm is None:
        _kpm = _bind(_load("libtbkpm.so"), KPM_SIGNATURES)
    return _kpm


def host():
    global _host
    if _host is None:
        _host = _bind(_load("libtbhost.so"), HOST_SIGNATURES)
    return _host


def check_kpm(rc):
    if rc != TBK_OK:
        raise NativeError(rc, kpm().tbk_last_error().decode())
    return rc


def check_host(rc):
    if rc != TBK_OK:
        raise NativeError(rc, host().tbh_last_error().decode())
    return rc


def loaded_paths():
    """Paths of the native libraries this process has loaded."""
    out = []
    for lib in (_kpm, _host):
        if lib is not None:
            out.append(lib._name)
    return out
