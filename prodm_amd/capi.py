"""ctypes binding of libprodm_mdr.so (include/prodm_mdr.h).

Device memory and streams come from PyTorch (torch tensors on cuda:0 and
torch.cuda.current_stream()); all compute runs in the HIP library. Loading
fails loudly if the library has not been built: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PRODM_MDR_LIB selects another build of the same library (same ABI)
LIB_PATH = os.environ.get("PRODM_MDR_LIB") or os.path.join(HERE, "libprodm_mdr.so")

F32, F64 = 0, 1
ORTHOGONAL, HIERARCHICAL = 0, 1
NEGABINARY, GROUPED, PERBIT = 0, 1, 2
QOI_VELOCITY_MAGNITUDE = 0

_lib = None

vp = C.c_void_p
u32, u64, i32, sz = C.c_uint32, C.c_uint64, C.c_int32, C.c_size_t

_SIGS = {
    "prodm_version": (C.c_int, []),
    "prodm_last_error": (C.c_char_p, []),
    "prodm_level_geometry": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
    "prodm_mgard_workspace_bytes": (sz, [vp, C.c_int, C.c_int, C.c_int]),
    "prodm_decompose_f32": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_decompose_f64": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_recompose_f32": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_recompose_f64": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_recompose_adjoint_workspace_bytes": (sz, [vp, C.c_int, C.c_int]),
    "prodm_recompose_adjoint_f32": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_recompose_adjoint_f64": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_decompose_hierarchical_f32": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_decompose_hierarchical_f64": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_recompose_hierarchical_f32": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_recompose_hierarchical_f64": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp, vp, sz, vp]),
    "prodm_mdr_plan_set_decomposer": (C.c_int, [vp, C.c_int]),
    "prodm_interleave_f32": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp]),
    "prodm_interleave_f64": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp]),
    "prodm_reposition_f32": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp]),
    "prodm_reposition_f64": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, vp, vp]),
    "prodm_level_maxabs_f32": (C.c_int, [vp, u64, vp, vp]),
    "prodm_level_maxabs_f64": (C.c_int, [vp, u64, vp, vp]),
    "prodm_negabinary_encode_f32": (C.c_int, [vp, u32, i32, C.c_int, vp, u64, vp]),
    "prodm_negabinary_encode_f64": (C.c_int, [vp, u32, i32, C.c_int, vp, u64, vp]),
    "prodm_negabinary_decode_f32": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, vp]),
    "prodm_negabinary_decode_f64": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, vp]),
    "prodm_grouped_workspace_bytes": (sz, [u32, C.c_int, C.c_int]),
    "prodm_grouped_encode_f32": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, vp, u64, vp, vp, sz, vp]),
    "prodm_grouped_encode_f64": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, vp, u64, vp, vp, sz, vp]),
    "prodm_grouped_encode_noalloc_f32": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, vp, u64, vp, vp]),
    "prodm_grouped_state_create": (C.c_int, [vp, C.c_int, C.c_int]),
    "prodm_grouped_state_destroy": (C.c_int, [vp]),
    "prodm_grouped_decode_f32": (C.c_int, [vp, C.c_int, vp, u32, i32, C.c_int, C.c_int, vp, vp]),
    "prodm_grouped_decode_f64": (C.c_int, [vp, C.c_int, vp, u32, i32, C.c_int, C.c_int, vp, vp]),
    "prodm_perbit_stream_words": (u64, [u32]),
    "prodm_perbit_workspace_bytes": (sz, [u32, C.c_int, C.c_int]),
    "prodm_perbit_encode_f32": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, u64, vp, vp, sz, vp]),
    "prodm_perbit_encode_f64": (C.c_int, [vp, u32, i32, C.c_int, C.c_int, vp, u64, vp, vp, sz, vp]),
    "prodm_perbit_state_create": (C.c_int, [vp]),
    "prodm_perbit_state_destroy": (C.c_int, [vp]),
    "prodm_perbit_decode_f32": (C.c_int, [vp, C.c_int, vp, vp, u32, i32, C.c_int, C.c_int, vp]),
    "prodm_perbit_decode_f64": (C.c_int, [vp, C.c_int, vp, vp, u32, i32, C.c_int, C.c_int, vp]),
    "prodm_collect_max_f32": (None, [C.c_float, C.c_int, vp]),
    "prodm_collect_max_f64": (None, [C.c_double, C.c_int, vp]),
    "prodm_collect_squared_workspace_bytes": (sz, [u64, C.c_int, C.c_int]),
    "prodm_collect_squared_f32": (C.c_int, [vp, u64, C.c_int, C.c_float, vp, vp, sz, vp]),
    "prodm_collect_squared_f64": (C.c_int, [vp, u64, C.c_int, C.c_double, vp, vp, sz, vp]),
    "prodm_mdr_plan_create": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "prodm_mdr_plan_destroy": (C.c_int, [vp]),
    "prodm_mdr_plan_device_bytes": (u64, [vp]),
    "prodm_mdr_refactor": (C.c_int, [vp, vp, vp]),
    "prodm_mdr_level_info": (C.c_int, [vp, C.c_int, vp, vp, vp]),
    "prodm_mdr_plane": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
    "prodm_mdr_level_rb": (C.c_int, [vp, C.c_int, vp, vp]),
    "prodm_mdr_level_coefficients": (C.c_int, [vp, C.c_int, vp]),
    "prodm_mdr_retrieve": (C.c_int, [vp, vp, vp, vp]),
    "prodm_mdr_profile": (C.c_int, [vp, C.c_int]),
    "prodm_mdr_stage_times": (C.c_int, [vp, vp, vp, C.c_int]),
    "prodm_mdr_refactor_graph": (C.c_int, [vp, vp, vp]),
    "prodm_mdr_retrieve_graph": (C.c_int, [vp, vp, vp, vp]),
    "prodm_mdr_plane_stream": (C.c_int, [vp, C.c_int, C.c_int, vp, u64, vp, vp]),
    "prodm_mdr_load_plane_stream": (C.c_int, [vp, C.c_int, C.c_int, vp, u64, vp]),
    "prodm_mdr_set_level_max": (C.c_int, [vp, C.c_int, C.c_double, vp]),
    "prodm_interpret_retrieve_size": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                                vp, vp, vp, C.c_double, vp, vp]),
    "prodm_max_estimator_ob_constant": (C.c_double, [C.c_int, C.c_int]),
    "prodm_recompose_amplification": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
    "prodm_qoi_interpret_retrieve_size": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, u64,
                                                    C.c_double, vp, vp]),
    "prodm_level_errors_workspace_bytes": (sz, [u64, C.c_int]),
    "prodm_negabinary_level_errors_f32": (C.c_int, [vp, u64, i32, C.c_int, vp, vp, sz, vp]),
    "prodm_grouped_level_errors_f32": (C.c_int, [vp, u64, i32, C.c_int, vp, vp, sz, vp]),
    "prodm_mdr_level_errors": (C.c_int, [vp, C.c_int, vp, vp]),
    "prodm_mdr_batch_create": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "prodm_mdr_batch_destroy": (C.c_int, [vp]),
    "prodm_mdr_batch_fields": (C.c_int, [vp]),
    "prodm_mdr_batch_device_bytes": (u64, [vp]),
    "prodm_mdr_batch_plan": (vp, [vp, C.c_int]),
    "prodm_mdr_batch_set_input": (C.c_int, [vp, C.c_int, vp]),
    "prodm_mdr_batch_set_decomposer": (C.c_int, [vp, C.c_int]),
    "prodm_mdr_batched_refactor": (C.c_int, [vp, vp]),
    "prodm_mdr_batched_refactor_graph": (C.c_int, [vp, vp]),
    "prodm_mdr_batch_autotune": (C.c_int, [vp, C.c_char_p, C.c_int, vp]),
    "prodm_mdr_batch_describe": (C.c_int, [vp, vp, u64, vp]),
}


class FieldDesc(C.Structure):
    """prodm_field_desc: one field of a batched refactor."""
    _fields_ = [("d_in", C.c_void_p), ("dims", C.c_uint32 * 3), ("ndims", C.c_int), ("target_level", C.c_int)]


STAGES = ["decompose_load_top", "decompose_load_rest", "decompose_solve_top", "decompose_solve_rest",
          "decompose_copy", "encode_top", "encode_rest", "decode_top", "decode_rest", "recompose_load_top",
          "recompose_load_rest", "recompose_solve_top", "recompose_solve_rest", "recompose_interp_top",
          "recompose_interp_rest"]


class ProdmError(RuntimeError):
    pass


def lib():
    """The loaded HIP library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ProdmError(f"{LIB_PATH} not built: run `python -m prodm_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc, what=""):
    if rc != 0:
        msg = lib().prodm_last_error().decode(errors="replace")
        raise ProdmError(f"{what} failed ({rc}): {msg}")
    return rc
