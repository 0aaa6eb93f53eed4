"""Loads libboojum_amd.so and declares its C-ABI (include/boojum_amd.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# BOOJUM_AMD_LIBRARY overrides the in-tree build (e.g. to compare two builds of the library)
library_path = os.environ.get("BOOJUM_AMD_LIBRARY") or os.path.join(os.path.dirname(_HERE), "libboojum_amd.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_u32, _u64, _i32, _sz = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t

class Job(ctypes.Structure):
    """ba_job (include/boojum_amd.h)."""
    _fields_ = [("kind", _u32), ("log_n", _u32), ("num_cols", _u32), ("bitrev_inputs", _u32),
                ("inverse", _u32), ("log_extension_degree", _u32), ("coset_index", _u32),
                ("log_blowup", _u32), ("inputs", _vp), ("outputs", _vp), ("monomials", _vp),
                ("stride_in", _u64), ("stride_out", _u64), ("out_coset_stride", _u64)]


class TunedPlan(ctypes.Structure):
    """ba_tuned_plan (include/boojum_amd.h)."""
    _fields_ = [("num_passes", _u32), ("bits", _u32 * 8), ("reps", _u32 * 8), ("contig_waves", _u32),
                ("tile_log", _u32)]


class ServeRequest(ctypes.Structure):
    """ba_serve_request (include/boojum_amd.h)."""
    _fields_ = [("log_n", _u32), ("num_cols", _u32), ("log_blowup", _u32), ("input_kind", _u32),
                ("evals", _vp), ("in_stride", _u64), ("seed", _u64)]


class ServeResult(ctypes.Structure):
    """ba_serve_result (include/boojum_amd.h)."""
    _fields_ = [("status", _i32), ("reserved", _u32), ("outputs", _vp), ("out_col_stride", _u64),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("batch", _u64)]


class ServeStats(ctypes.Structure):
    """ba_serve_stats (include/boojum_amd.h)."""
    _fields_ = [("completed", _u64), ("failed", _u64), ("batches", _u64), ("output_elems", _u64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double),
                ("max_ms", ctypes.c_double), ("elapsed_s", ctypes.c_double),
                ("mean_batch_jobs", ctypes.c_double), ("budget_bytes", _u64), ("peak_bytes", _u64),
                ("in_use_bytes", _u64)]


# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "ba_get_device_count": [ctypes.POINTER(_i32)],
    "ba_set_device": [_i32],
    "ba_get_device": [ctypes.POINTER(_i32)],
    "ba_device_synchronize": [],
    "ba_mem_get_info": [ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "ba_device_get_attribute": [ctypes.POINTER(_i32), _i32, _i32],
    "ba_malloc": [ctypes.POINTER(_vp), _sz],
    "ba_free": [_vp],
    "ba_malloc_host": [ctypes.POINTER(_vp), _sz],
    "ba_free_host": [_vp],
    "ba_memcpy": [_vp, _vp, _sz, _i32],
    "ba_memcpy_async": [_vp, _vp, _sz, _i32, _vp],
    "ba_memset_async": [_vp, _i32, _sz, _vp],
    "ba_stream_create": [ctypes.POINTER(_vp)],
    "ba_stream_destroy": [_vp],
    "ba_stream_synchronize": [_vp],
    "ba_stream_wait_event": [_vp, _vp],
    "ba_event_create": [ctypes.POINTER(_vp)],
    "ba_event_destroy": [_vp],
    "ba_event_record": [_vp, _vp],
    "ba_event_synchronize": [_vp],
    "ba_event_query": [_vp],
    "ba_event_elapsed_ms": [ctypes.POINTER(ctypes.c_float), _vp, _vp],
    "ba_get_last_error": [],
    "ba_error_name": [_i32],
    "ba_stream_begin_capture": [_vp],
    "ba_stream_end_capture": [_vp, ctypes.POINTER(_vp)],
    "ba_graph_launch": [_vp, _vp],
    "ba_graph_destroy": [_vp],
    "ba_context_create": [_u32, _u32, ctypes.POINTER(_vp)],
    "ba_context_destroy": [_vp],
    "ba_batch_ntt_internal": [_vp, _vp, _u32, _u32, _u32, _u32, _i32, _i32, _u32, _u32, _vp],
    "ba_batch_ntt_out_of_place": [_vp, _sz, _vp, _sz, _u32, _u32, _u32, _u32, _u32, _u32, _i32,
                                  _i32, _u32, _u32, _vp],
    "ba_batch_ntt_in_place": [_vp, _sz, _u32, _u32, _u32, _u32, _i32, _i32, _u32, _u32, _vp],
    "ba_batch_ntt_ctx": [_vp, _vp, _vp, _u32, _u32, _u64, _u64, _i32, _i32, _u32, _u32, _vp],
    "ba_bit_reverse": [_vp, _u64, _vp, _u64, _u32, _u32, _vp],
    "ba_bit_reverse_in_place": [_vp, _u64, _u32, _u32, _vp],
    "ba_lde_batch": [_vp, _vp, _u64, _vp, _vp, _u64, _u64, _u32, _u32, _u32, _vp],
    "ba_lde_batch_ex": [_vp, _vp, _u64, _vp, _vp, _u64, _u64, _u32, _u32, _u32, _u32, _vp],
    "ba_batch_plan_create": [_vp, ctypes.POINTER(Job), _u32, ctypes.POINTER(_vp)],
    "ba_batch_plan_launch": [_vp, _vp],
    "ba_batch_plan_stats": [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u64)],
    "ba_batch_plan_destroy": [_vp],
    "ba_autotune": [_vp, _u32, _u32, _u64, _vp],
    "ba_get_powers_of_w": [_u32, _u32, _i32, _i32, _vp, _u32, _vp],
    "ba_get_powers_of_g": [_u32, _u32, _i32, _i32, _vp, _u32, _vp],
    "ba_get_powers_by_val_bf": [_u64, _u32, _i32, _vp, _u32, _vp],
    "ba_get_powers_by_ref_bf": [_vp, _u32, _i32, _vp, _u32, _vp],
    "ba_get_powers_by_val_ef": [_u64, _u64, _u32, _i32, _vp, _u32, _vp],
    "ba_get_powers_by_ref_ef": [_vp, _u32, _i32, _vp, _u32, _vp],
    "ba_omega_shift": [_vp, _u32, _u32, _i32, _u32, _vp, _u32, _vp],
    "ba_omega_shift_in_place": [_vp, _u32, _u32, _i32, _u32, _u32, _vp],
    "ba_barycentric_precompute_lagrange_coeffs": [_i32, _vp, _vp, _u64, _vp, _u32, _i32, _vp],
    "ba_barycentric_batch_eval_temp_elems": [_i32, _u32, _u32, ctypes.POINTER(_u64)],
    "ba_barycentric_batch_eval": [_i32, _vp, _u32, _u32, _vp, _vp, _u64, _vp, _vp],
    "ba_barycentric_batch_eval_noalloc": [_i32, _vp, _u32, _u32, _vp, _vp, _vp],
    "ba_diag_butterfly_rate": [_vp, _vp, _u32, _u32, _vp],
    "ba_poseidon2_gl_set_round_constants": [_vp, _vp],
    "ba_poseidon2_gl_compute_leaf_hashes": [_vp, _sz, _vp, _sz, _u32, _i32, _i32, _vp],
    "ba_poseidon2_gl_compute_node_hashes": [_vp, _sz, _vp, _sz, _vp],
    "ba_poseidon2_gl_build_merkle_tree_leaves": [_vp, _sz, _vp, _sz, _u32, _i32, _i32, _vp],
    "ba_poseidon2_gl_build_merkle_tree_nodes": [_vp, _sz, _vp, _sz, _u32, _vp],
    "ba_poseidon2_gl_build_merkle_tree": [_vp, _sz, _vp, _sz, _u32, _u32, _vp],
    "ba_server_create": [_vp, _sz, _u32, ctypes.POINTER(_vp)],
    "ba_server_set_max_bypass": [_vp, _u32],
    "ba_server_submit": [_vp, ctypes.POINTER(ServeRequest), ctypes.POINTER(_u64)],
    "ba_server_wait": [_vp, _u64, ctypes.POINTER(ServeResult)],
    "ba_server_poll": [_vp, _u64, ctypes.POINTER(_i32), ctypes.POINTER(ServeResult)],
    "ba_server_release": [_vp, _u64],
    "ba_server_stats": [_vp, ctypes.POINTER(ServeStats)],
    "ba_server_destroy": [_vp],
    "ba_tuning_get": [_vp, _u32, _i32, ctypes.POINTER(TunedPlan)],
    "ba_tuning_set": [_vp, _u32, _i32, ctypes.POINTER(TunedPlan)],
    "ba_tuning_candidates": [_u32, ctypes.POINTER(TunedPlan), _u32, ctypes.POINTER(_u32)],
    "ba_profile_begin": [ctypes.POINTER(_vp), _u32],
    "ba_profile_end": [ctypes.POINTER(_u32)],
    "ba_profile_launch_name": [_u32],
    "ba_fill_uniform": [_vp, _u64, _u64, _u32, _u64, _vp],
    "ba_field_op": [_i32, _vp, _vp, _vp, _sz, _vp],
}
_RESTYPES = {"ba_error_name": ctypes.c_char_p, "ba_profile_launch_name": ctypes.c_char_p}

_lib = None


class BoojumAmdError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        name = "?"
        try:
            name = lib().ba_error_name(code).decode()
        except Exception:  # pragma: no cover - only when the library itself is broken
            pass
        super().__init__(f"{what}: status {code} ({name})")


def lib():
    """The loaded library; raises if it was not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise RuntimeError(f"libboojum_amd.so not built at {library_path}; run __graft_entry__.build()")
        L = ctypes.CDLL(library_path)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc, what="boojum_amd call"):
    if rc != 0:
        raise BoojumAmdError(rc, what)
