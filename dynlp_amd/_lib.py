"""ctypes binding of libdynlp_hip.so (the C ABI declared in include/dynlp.h).

The library is loaded from this package's folder.  There is no fallback: if it is missing
or fails to load, every call raises, naming the path and the build command.
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdynlp_hip.so")

F32, F64 = 0, 1
_lib = None

P = ct.c_void_p
I32 = ct.c_int32
I64 = ct.c_int64
DBL = ct.c_double
FLT = ct.c_float
SZ = ct.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "dlp_last_error": (ct.c_char_p, []),
    "dlp_abi_version": (ct.c_int, []),
    "dlp_spmm_axpy": (ct.c_int, [I32, I32, P, P, P, P, I64, P, DBL, P, I64, I32, DBL, P]),
    "dlp_row_operator": (ct.c_int, [I32, P, I32, P, P, P, P, I32, P, P, I64, I32, P, DBL, P, I64,
                                    P, I64, I32, I32, DBL, P]),
    "dlp_rho_cols": (ct.c_int, [I32, P, I32, P, I64, I32, P, P, P]),
    "dlp_violating_rows": (ct.c_int, [I32, P, I32, P, I64, I32, P, P, P, P, P]),
    "dlp_frontier_jacobi": (ct.c_int, [I32, P, I32, I32, P, P, P, P, P, P, DBL, P, P, P, I32, DBL, P]),
    "dlp_push_scatter": (ct.c_int, [I32, P, I32, P, P, P, P, P, P, P, I32, DBL, P]),
    "dlp_mark_neighbors": (ct.c_int, [I32, P, P, P, P, P, P]),
    "dlp_flatnonzero_u8": (ct.c_int, [I32, P, P, P, P, SZ, P]),
    "dlp_scan_workspace_bytes": (SZ, [I32]),
    "dlp_cc_edges": (ct.c_int, [I32, I64, P, P, P, P, P, SZ, P]),
    "dlp_cc_csr": (ct.c_int, [I32, P, P, P, P, P, P, P, SZ, P]),
    "dlp_cc_workspace_bytes": (SZ, [I32]),
    "dlp_lp_create": (ct.c_int, [P, P]),
    "dlp_lp_destroy": (ct.c_int, [P]),
    "dlp_lp_set_store": (ct.c_int, [P, P, P, P, P]),
    "dlp_lp_set_eps": (ct.c_int, [P, P, DBL, P]),
    "dlp_lp_begin_batch": (ct.c_int, [P, P]),
    "dlp_lp_apply": (ct.c_int, [P, I32, P, P, I32, P, P, P, P]),
    "dlp_lp_apply_ops": (ct.c_int, [P, I32, P, P, P, P]),
    "dlp_lp_regrow_plan": (ct.c_int, [P, I32, DBL, P, P, P]),
    "dlp_lp_regrow_commit": (ct.c_int, [P, P, P, P, I32, P, P, P, P, P]),
    "dlp_lp_ground": (ct.c_int, [P, P]),
    "dlp_lp_ground_incremental": (ct.c_int, [P, I32, P, P, I32, P]),
    "dlp_lp_ground_flags": (ct.c_int, [P, P, P]),
    "dlp_lp_refresh": (ct.c_int, [P, I32, I32, P]),
    "dlp_lp_push": (ct.c_int, [P, I32, I32, I32, I64, P, P, P]),
    "dlp_lp_push64_round": (ct.c_int, [P, I32, P, P]),
    "dlp_lp_push64_next": (ct.c_int, [P, P, P, P]),
    "dlp_lp_violators": (ct.c_int, [P, P, P]),
    "dlp_lp_status": (ct.c_int, [P, P, P, P]),
    "dlp_lp_pcg": (ct.c_int, [P, I32, I32, I32, I32, P, P, P]),
    "dlp_lp_pcg_verify": (ct.c_int, [P, I32, I32, I32, I32, P, P, P]),
    "dlp_lp_pcg_stats": (ct.c_int, [P, P]),
    "dlp_lp_reserve": (ct.c_int, [P, I32, I32, P]),
    "dlp_push_plan_create": (ct.c_int, [P, I32, P]),
    "dlp_push_plan_destroy": (ct.c_int, [P]),
    "dlp_lp_push_multi": (ct.c_int, [P, P, I32, I32, I32, P, P, P, P]),
    "dlp_cc_csr_static": (ct.c_int, [I32, P, P, P, P, P, P, P]),
    "dlp_cc_static_max_n": (ct.c_int, []),
    "dlp_push_rounds": (ct.c_int, [I32, I32, I64, I32, P, P, P, P, P, P, P, P, P, P, P, P, P, P, I32, P]),
    "dlp_cg_residual": (ct.c_int, [I32, I32, I64, I32, P, P, P, P, I32, P, DBL, P, P, I64, P, P, P, P, P]),
    "dlp_cg_direction": (ct.c_int, [I32, I32, I64, I32, P, P, P, P, I32, P]),
    "dlp_cg_iteration": (ct.c_int, [I32, I32, I64, I32, P, P, P, P, I32, P, DBL, P, P, P, P, P, P, P, P]),
    "dlp_park_rows": (ct.c_int, [I32, P, I32, I64, I32, P, P, P]),
    "dlp_rows_copy": (ct.c_int, [I32, P, I32, P, I64, I32, P, I64, I32, I32, P]),
    "dlp_supernode_average": (ct.c_int, [I32, P, P, I32, P, P, P, P, I32, P, I64, P, I32, I32, P, DBL, P, I64, I32,
                                         P, SZ, P]),
    "dlp_supernode_workspace_bytes": (SZ, [I32, I32, I32]),
    "dlp_frontier_commit": (ct.c_int, [I32, P, I32, I64, P, P, P, P, P, P, P, P, I32, P]),
    "dlp_dense_operator_f64": (ct.c_int, [I32, P, P, P, P, I32, P, P]),
    "dlp_spd_inverse_f64": (ct.c_int, [I32, P, P, P]),
    "dlp_dense_apply_f64": (ct.c_int, [I32, I32, P, P, I64, P, I64, I32, P]),
    "dlp_scatter": (ct.c_int, [I32, I32, I64, I64, P, I64, P, I32, P, I64, P, P]),
    "dlp_fh_components": (ct.c_int, [I32, I64, P, P, P, DBL, P, P, P, P, SZ, P]),
    "dlp_fh_workspace_bytes": (SZ, [I32]),
    "dlp_fh_scale": (ct.c_int, [I32, I64, P, P, I32, DBL, P, P, P]),
    "dlp_fh_dissimilarity": (ct.c_int, [I64, P, I32, P, P, P, P, DBL, P, P, P, SZ, P]),
    "dlp_fh_dissimilarity_workspace_bytes": (SZ, []),
    "dlp_lp_device_bytes": (ct.c_int, [P, P]),
    "dlp_lp_handle_bytes": (ct.c_int64, [I32, I32, I64, I32, I32, I32, I32, I32]),
    "dlp_lp_bpcg": (ct.c_int, [P, I32, I32, I32, I32, P, P, P]),
    "dlp_gram": (ct.c_int, [I32, I32, P, I64, P, I64, P, P, I32, P]),
    "dlp_chol_solve": (ct.c_int, [I32, P, P, P, DBL, P]),
    "dlp_block_update": (ct.c_int, [I32, I32, I32, P, P, P, P, P, I64, P, P, P, P, P]),
    "dlp_lp_time_operator": (ct.c_int, [P, I32, P, P]),
    "dlp_lp_scores": (ct.c_int, [P, P, I64, P]),
    "dlp_lp_set_pcg_directions": (ct.c_int, [P, I32]),
    "dlp_lp_set_coarse": (ct.c_int, [P, I32, I32]),
    "dlp_lp_coarse_info": (ct.c_int, [P, P, P, P]),
    "dlp_tune_set": (ct.c_int, [I32, I32]),
    "dlp_row_operator_dot": (ct.c_int, [I32, I32, P, P, P, P, P, P, I64, P, FLT, P, I64, P, P]),
    "dlp_coldot": (ct.c_int, [I32, I32, P, I64, P, I64, P, P, I32, P]),
    "dlp_colop": (ct.c_int, [I32, I32, I32, DBL, P, P, I64, P, P, I64, I32, P]),
    "dlp_amg_match": (ct.c_int, [I32, P, P, P, I32, I32, P, P, P, P, P, SZ, P, P]),
    "dlp_amg_compose": (ct.c_int, [I32, P, P, P, P]),
    "dlp_amg_coarsen_fill": (ct.c_int, [I32, I32, P, P, P, I32, P, P, P, P, P, P, SZ, P]),
    "dlp_amg_coarsen_merge": (ct.c_int, [I32, P, P, P, I32, P, P, SZ, P, P]),
    "dlp_amg_coarsen_compact": (ct.c_int, [I32, P, P, P, I32, P, P, P, P]),
    "dlp_amg_restrict": (ct.c_int, [I32, I32, P, P, I64, P, I64, I32, P]),
    "dlp_amg_prolong_add": (ct.c_int, [I32, I32, P, P, I64, P, I64, I32, P]),
    "dlp_amg_level": (ct.c_int, [I32, I32, P, P, P, P, P, P, I64, DBL, I32, P, P]),
    "dlp_check_edge_batch": (ct.c_int, [I64, I64, P, P, I64, P, P, P, I32, P]),
}


class LPDesc(ct.Structure):
    """dlp_lp_desc of include/dynlp.h."""
    _fields_ = [("n", I32), ("C1", I32), ("K2", I32), ("pad_", I32), ("eta", DBL), ("prior", DBL),
                ("indptr", P), ("rowlen", P), ("col", P), ("w", P), ("labeled", P), ("y", P),
                ("solved", P), ("labcnt", P), ("Z32", P), ("R32", P), ("ld", I64), ("diag64", P),
                ("diag32", P), ("Z64", P), ("R64", P)]


class DLPError(RuntimeError):
    pass


def load():
    """The loaded library (raises DLPError if it is not built or does not load)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DLPError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}/csrc` "
                       "(or __graft_entry__.build()); there is no fallback path")
    try:
        lib = ct.CDLL(LIB_PATH)
    except OSError as e:
        raise DLPError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


SIGS = _SIGS


def exported_symbols():
    return list(_SIGS)


def call(name, *args):
    """Call a status-returning entry point; raise DLPError with the library's message."""
    fn = getattr(load(), name)
    rc = fn(*args)
    if rc != 0:
        msg = load().dlp_last_error().decode(errors="replace")
        raise DLPError(f"{name} failed ({rc}): {msg}")
    return rc
