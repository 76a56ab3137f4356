"""ctypes binding of libvnec_mi355x.so (the C-ABI declared in
include/OpenSparseBLAS.h and include/vnec_ext.h).

Loading fails loudly when the library is missing: there is no CPU
fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvnec_mi355x.so")

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
u64 = C.c_uint64
Info = C.c_int


class Tuning(C.Structure):
    _fields_ = [("variant", C.c_int), ("ipt", C.c_int), ("lanes", C.c_int),
                ("encoding", C.c_int), ("nrhs", C.c_int), ("grid", C.c_int)]


class Coord(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int)]


class VNEC_L(C.Structure):
    _fields_ = [("num_merge_items", C.c_int), ("items_per_thread", C.c_int),
                ("diagonal_start", C.POINTER(C.c_int)), ("diagonal_end", C.POINTER(C.c_int)),
                ("thread_coord_start", C.POINTER(Coord)), ("thread_coord_end", C.POINTER(Coord)),
                ("nz_indices", C.POINTER(C.c_int))]


class VNEC_D_FP64(C.Structure):
    _fields_ = [("thread_coord_start", C.POINTER(Coord)), ("thread_coord_end", C.POINTER(Coord)),
                ("ecr_indices", C.POINTER(C.c_int)), ("NEC_NUM", C.POINTER(C.c_int)),
                ("use_x_indices", C.POINTER(C.POINTER(C.c_int))),
                ("ecr_xx_val", C.POINTER(C.POINTER(C.c_double))),
                ("v_row_ptr", C.POINTER(C.c_int)), ("col_start", C.POINTER(C.c_int)),
                ("val_align", C.POINTER(C.c_double))]


class CsrF64(C.Structure):
    _fields_ = [("nnz", C.POINTER(C.c_double)), ("col", C.POINTER(C.c_int)),
                ("row_begin", C.POINTER(C.c_int)), ("row_end", C.POINTER(C.c_int)),
                ("task_start", C.POINTER(C.c_int)), ("task_end", C.POINTER(C.c_int))]


class VNEC_S_FP64(C.Structure):
    _fields_ = [("NEC_NUM", C.POINTER(C.c_int)), ("use_x_indices", C.POINTER(C.POINTER(C.c_int))),
                ("ecr_xx_val", C.POINTER(C.POINTER(C.c_double))),
                ("spvv8_len", C.POINTER(C.c_int)), ("tasks", C.POINTER(C.POINTER(C.c_int))),
                ("reorder_mat", CsrF64)]


class VNEC_D_FP32(C.Structure):
    _fields_ = [("thread_coord_start", C.POINTER(Coord)), ("thread_coord_end", C.POINTER(Coord)),
                ("ecr_indices", C.POINTER(C.c_int)), ("NEC_NUM", C.POINTER(C.c_int)),
                ("use_x_indices", C.POINTER(C.POINTER(C.c_int))),
                ("ecr_xx_val", C.POINTER(C.POINTER(C.c_float))),
                ("v_row_ptr", C.POINTER(C.c_int)), ("col_start", C.POINTER(C.c_int)),
                ("val_align", C.POINTER(C.c_float))]


class CsrF32(C.Structure):
    _fields_ = [("nnz", C.POINTER(C.c_float)), ("col", C.POINTER(C.c_int)),
                ("row_begin", C.POINTER(C.c_int)), ("row_end", C.POINTER(C.c_int)),
                ("task_start", C.POINTER(C.c_int)), ("task_end", C.POINTER(C.c_int))]


class VNEC_S_FP32(C.Structure):
    _fields_ = [("NEC_NUM", C.POINTER(C.c_int)), ("use_x_indices", C.POINTER(C.POINTER(C.c_int))),
                ("ecr_xx_val", C.POINTER(C.POINTER(C.c_float))),
                ("spvv16_len", C.POINTER(C.c_int)), ("tasks", C.POINTER(C.POINTER(C.c_int))),
                ("reorder_mat", CsrF32)]


class SpB_Object_Format_ext(C.Union):
    _fields_ = [("vector_format", C.c_int), ("matrix_format", C.c_int)]


# name -> (restype, argtypes); every symbol include/*.h declares
SIGNATURES = {
    # OpenSparseBLAS.h
    "SpB_init": (None, []),
    "SpB_finalize": (None, []),
    "SpB_Vector_new": (Info, [C.POINTER(vp), C.c_int, u64]),
    "SpB_Vector_free": (Info, [C.POINTER(vp)]),
    "SpB_Vector_nvals": (Info, [C.POINTER(u64), vp]),
    "SpB_Vector_setElement_FP64": (Info, [vp, C.c_double, u64]),
    "SpB_Vector_extractElement_FP64": (Info, [C.POINTER(C.c_double), vp, u64]),
    "SpB_Vector_extractTuples_FP32": (Info, [vp, vp, C.POINTER(u64), vp]),
    "SpB_Vector_extractTuples_FP64": (Info, [vp, vp, C.POINTER(u64), vp]),
    "SpB_Vector_setElement_FP32": (Info, [vp, C.c_float, u64]),
    "SpB_Vector_extractElement_FP32": (Info, [C.POINTER(C.c_float), vp, u64]),
    "SpB_Matrix_nrows": (Info, [C.POINTER(u64), vp]),
    "SpB_Matrix_ncols": (Info, [C.POINTER(u64), vp]),
    "SpB_Matrix_new": (Info, [C.POINTER(vp), C.c_int, u64, u64]),
    "SpB_Matrix_free": (Info, [C.POINTER(vp)]),
    "SpB_Matrix_build_FP64": (Info, [vp, vp, vp, vp, u64, C.c_int]),
    "SpB_Matrix_build_FP32": (Info, [vp, vp, vp, vp, u64, C.c_int]),
    "SpB_mxv": (Info, [vp, vp, vp, vp]),
    "IRD_VNEC": (C.c_int, [vp, C.c_int]),
    "Build_VNEC_L": (C.POINTER(VNEC_L), [vp]),
    "Build_VNEC_D_FP64": (C.POINTER(VNEC_D_FP64), [vp]),
    "Build_VNEC_S_FP64": (C.POINTER(VNEC_S_FP64), [vp]),
    "SpMV_VNEC_L_FP64": (None, [vp, vp, vp, vp]),
    "SpMV_VNEC_D_FP64": (None, [vp, vp, vp, vp]),
    "SpMV_VNEC_S_FP64": (None, [vp, vp, vp, vp]),
    "Build_VNEC_D_FP32": (C.POINTER(VNEC_D_FP32), [vp]),
    "Build_VNEC_S_FP32": (C.POINTER(VNEC_S_FP32), [vp]),
    "SpMV_VNEC_L_FP32": (None, [vp, vp, vp, vp]),
    "SpMV_VNEC_D_FP32": (None, [vp, vp, vp, vp]),
    "SpMV_VNEC_S_FP32": (None, [vp, vp, vp, vp]),
    "SpB_Global_Option_set_ext": (Info, [C.c_int, SpB_Object_Format_ext]),
    # vnec_ext.h
    "vnec_last_error": (C.c_char_p, []),
    "vnec_device_count_ext": (C.c_int, []),
    "SpB_Matrix_import_csr_ext": (Info, [vp, vp, vp, vp]),
    "SpB_Matrix_import_csr_fp32_ext": (Info, [vp, vp, vp, vp]),
    "SpB_Matrix_nvals_ext": (Info, [C.POINTER(u64), vp]),
    "vnec_batch_import_ext": (Info, [vp, i64, vp, vp, vp, vp, vp]),
    "SpB_Vector_import_ext": (Info, [vp, vp]),
    "SpB_Vector_export_ext": (Info, [vp, vp]),
    "SpB_Vector_import_fp32_ext": (Info, [vp, vp]),
    "SpB_Vector_export_fp32_ext": (Info, [vp, vp]),
    "SpB_Vector_device_resident_ext": (Info, [vp, C.c_bool]),
    "SpB_Vector_sync_host_ext": (Info, [vp]),
    "SpB_Vector_device_ptr_ext": (vp, [vp]),
    "vnec_plan_build_ext": (Info, [C.POINTER(vp), vp, C.POINTER(Tuning)]),
    "vnec_plan_free_ext": (Info, [C.POINTER(vp)]),
    "vnec_plan_info_ext": (Info, [vp, C.POINTER(Tuning), C.POINTER(i64), C.POINTER(i64)]),
    "vnec_spmv_ext": (Info, [vp, vp, vp]),
    "vnec_spmv_dev_ext": (Info, [vp, vp, vp]),
    "vnec_spmv_graph_ext": (Info, [vp, vp, vp]),
    "vnec_spmm_colmajor_dev_ext": (Info, [vp, vp, vp]),
    "vnec_time_dev_ext": (Info, [vp, vp, vp, C.c_int, C.c_int, C.c_size_t,
                                 C.POINTER(C.c_double)]),
    "vnec_time_kernels_ext": (Info, [vp, vp, vp, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "vnec_copy_bandwidth_ext": (Info, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "vnec_read_bandwidth_ext": (Info, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "vnec_plan_bytes_ext": (Info, [vp, C.POINTER(i64), C.POINTER(i64)]),
    "vnec_dev_alloc_ext": (vp, [C.c_size_t]),
    "vnec_dev_free_ext": (Info, [vp]),
    "vnec_memcpy_h2d_ext": (Info, [vp, vp, C.c_size_t]),
    "vnec_memcpy_d2h_ext": (Info, [vp, vp, C.c_size_t]),
    "vnec_sync_ext": (Info, []),
    "vnec_partition_ext": (Info, [vp, i64, C.c_int, vp]),
    "vnec_ecr_ext": (Info, [vp, i64, vp, C.c_int, vp, vp, vp, vp]),
    "vnec_d_layout_ext": (Info, [vp, i64, C.c_int, vp, vp, vp, C.POINTER(i64)]),
    "vnec_s_order_ext": (Info, [vp, i64, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]),
    "vnec_ird_ext": (Info, [vp, i64, C.c_int, C.c_double, C.POINTER(C.c_float),
                            C.POINTER(C.c_int)]),
    "vnec_read_graph_ext": (Info, [C.c_char_p, C.POINTER(i64), C.POINTER(vp), C.POINTER(vp)]),
    "vnec_host_free_ext": (None, [vp]),
    "SpB_Matrix_read_ext": (Info, [C.POINTER(vp), C.c_char_p, C.c_int, C.c_int]),
    "vnec_candidates_ext": (C.c_int, [vp, C.c_int]),
    "vnec_candidate_mask_ext": (Info, [vp, C.POINTER(C.c_uint32)]),
    "vnec_autotune_ext": (Info, [vp, C.c_int, vp, C.POINTER(C.c_int)]),
    "vnec_feature_ints_ext": (Info, [vp, vp]),
    "vnec_features_from_ints_ext": (None, [vp, vp]),
    "vnec_candidate_mask_from_ints_ext": (C.c_uint32, [vp]),
    "vnec_mlp_forward_ext": (Info, [vp, vp, i64, vp]),
    "vnec_mlp_grad_ext": (Info, [vp, vp, vp, vp, i64, vp, C.POINTER(C.c_float)]),
    "vnec_mlp_train_ext": (Info, [vp, vp, vp, vp, i64, C.c_int, C.c_int, C.c_float, C.c_float,
                                  u64, vp]),
    "vnec_mlp_select_ext": (Info, [vp, vp, vp, vp, vp, i64, vp]),
    "vnec_selector_install_ext": (Info, [vp, vp, vp]),
    "vnec_selector_info_ext": (Info, [C.POINTER(C.c_int), C.POINTER(i64)]),
    "vnec_select_ext": (Info, [vp, C.c_int, C.POINTER(C.c_int)]),
    "VNEC_D_free_fp64_ext": (Info, [vp]),
    "VNEC_S_free_fp64_ext": (Info, [vp]),
    "VNEC_D_free_fp32_ext": (Info, [vp]),
    "vnec_pattern_ext": (Info, [vp, vp, C.POINTER(C.c_int), C.POINTER(i64)]),
    "vnec_gather_ceiling_ext": (Info, [vp, vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "vnec_pattern_match_ext": (Info, [vp, vp, C.c_int, vp, vp]),
    "VNEC_S_free_fp32_ext": (Info, [vp]),
    "VNEC_L_free_ext": (Info, [vp]),
    "VNEC_plan_of_ext": (vp, [vp]),
    "vnec_csr_spmv_noalloc_ext": (Info, [i64, i64, i64, vp, vp, vp, vp, vp, vp]),
    # vnec_hip.h (thin C-ABI to the kernels)
    "vnec_hip_partition": (C.c_int, [vp, i64, i64, i64, C.c_int, vp, vp]),
    "vnec_hip_scan_excl_i32": (C.c_int, [vp, vp, i64, C.POINTER(i64), vp]),
    "vnec_hip_ecr_tiles_pass1": (C.c_int, [vp, vp, i64, vp, vp, vp, vp, vp, vp]),
    "vnec_hip_ecr_tiles_pass2": (C.c_int, [vp, i64, vp, vp, vp, vp, vp, vp]),
    "vnec_hip_max_tile_nnz": (C.c_int, [vp, i64, vp, vp]),
    "vnec_hip_ecr_dense": (C.c_int, [vp, i64, i64, vp, i64, vp, vp, vp, vp, vp, vp]),
    "vnec_hip_d_count": (C.c_int, [vp, vp, i64, vp, i64, C.c_int, vp, vp, vp]),
    "vnec_hip_d_fill": (C.c_int, [vp, vp, vp, i64, vp, i64, C.c_int, vp, vp, vp, C.c_int, vp]),
    "vnec_hip_split_hilo": (C.c_int, [vp, i64, vp, vp, vp]),
    "vnec_hip_rhs_layout": (C.c_int, [vp, i64, C.c_int, C.c_int, vp, vp]),
    "vnec_hip_csort": (C.c_int, [vp, i64, vp, vp, i64, i64, C.c_int, vp, vp, vp, vp, vp, vp, vp]),
    "vnec_hip_tile_colspan": (C.c_int, [vp, i64, vp, vp, vp, vp]),
    "vnec_hip_spmv_merge": (C.c_int, [vp, vp]),
    "vnec_hip_carry_fixup": (C.c_int, [vp, vp, i64, vp, i64, C.c_int, vp]),
    "vnec_hip_carry_fixup_f32": (C.c_int, [vp, vp, i64, vp, i64, vp]),
    "vnec_hip_merge_supported": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "vnec_hip_stream_grid": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "vnec_hip_stream_supported": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "vnec_hip_spmv_stream": (C.c_int, [vp, C.c_int, vp, vp, vp]),
    "vnec_hip_wg_fixup": (C.c_int, [vp, vp, C.c_int, vp, i64, vp]),
    "vnec_hip_strided_supported": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "vnec_hip_strided_grid": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "vnec_hip_spmv_strided": (C.c_int, [vp, C.c_int, vp]),
    "vnec_hip_stream_copy": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "vnec_hip_stream_read": (C.c_int, [vp, C.c_size_t, vp, vp]),
    "vnec_hip_flush_read": (C.c_int, [vp, C.c_size_t, vp, vp]),
    "vnec_hip_slice_len": (C.c_int, [vp, i64, i64, vp, vp, vp]),
    "vnec_hip_slice_tile_rows": (C.c_int, [vp, i64, vp, i64, vp, vp]),
    "vnec_hip_slice_fill": (C.c_int, [vp, vp, vp, vp, i64, i64, vp, vp, vp, vp, vp, vp, vp, vp]),
    "vnec_hip_spmv_slice": (C.c_int, [vp, vp]),
    "vnec_hip_scribble": (C.c_int, [vp, C.c_size_t, vp]),
    "vnec_hip_zero_f64": (C.c_int, [vp, i64, vp]),
    "vnec_hip_gather_probe": (C.c_int, [vp, vp, i64, vp, C.c_int, vp, C.c_int, vp]),
    "vnec_hip_pattern_match": (C.c_int, [vp, vp, i64, vp, C.c_int, vp, vp, vp]),
    "vnec_hip_pattern_split": (C.c_int, [vp, vp, i64, vp, vp]),
    "vnec_hip_pattern_fill": (C.c_int, [vp, vp, vp, i64, C.c_int, vp, vp]),
    "vnec_hip_spmv_pattern": (C.c_int, [vp, vp]),
    "vnec_hip_scatter_rows": (C.c_int, [vp, vp, i64, vp, vp]),
    "vnec_hip_scatter_rows_f32": (C.c_int, [vp, vp, i64, vp, vp]),
    "vnec_hip_zero_f32": (C.c_int, [vp, i64, vp]),
    "vnec_hip_widen_f32": (C.c_int, [vp, i64, vp, vp]),
    "vnec_hip_hot_supported": (C.c_int, [C.c_int, C.c_int]),
    "vnec_hip_hot_grid": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "vnec_hip_spmv_hot": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp]),
    "vnec_hip_col_degree": (C.c_int, [vp, i64, vp, vp]),
    "vnec_hip_hot_order": (C.c_int, [vp, i64, vp, vp]),
    "vnec_hip_hot_encode": (C.c_int, [vp, i64, vp, i64, i64, vp, vp, vp]),
    "vnec_hip_pack_order": (C.c_int, [vp, i64, i64, vp, vp, vp, vp]),
    "vnec_hip_relabel_cols": (C.c_int, [vp, i64, vp, i64, vp, vp, vp]),
    "vnec_hip_gather_hot": (C.c_int, [vp, vp, i64, vp, vp]),
    "vnec_hip_sort_pairs_u64": (C.c_int, [vp, vp, i64, C.c_int, vp]),
    "vnec_hip_s_order": (C.c_int, [vp, i64, vp, i64, C.c_int, C.c_int, vp, vp, vp]),
    "vnec_hip_s_reorder": (C.c_int, [vp, vp, vp, vp, i64, i64, vp, vp, i64, C.c_int, vp, vp, vp,
                                     vp, vp]),
    "vnec_hip_d_interleave": (C.c_int, [vp, i64, C.c_int, vp, vp]),
    "vnec_hip_d_interleave_f32": (C.c_int, [vp, i64, C.c_int, vp, vp]),
    "vnec_hip_spmv_d": (C.c_int, [vp, vp]),
    "vnec_hip_d_supported": (C.c_int, [C.c_int, C.c_int]),
    "vnec_hip_feature_rows": (C.c_int, [vp, i64, vp, vp, vp]),
    "vnec_hip_feature_cols": (C.c_int, [vp, vp, i64, i64, vp, vp]),
    "vnec_hip_mlp_forward": (C.c_int, [vp, vp, i64, vp, vp]),
    "vnec_hip_mlp_grad": (C.c_int, [vp, vp, vp, vp, vp, i64, C.c_int, C.c_float, vp, vp, vp]),
    "vnec_hip_mlp_reduce_adam": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, C.c_float, C.c_float,
                                           C.c_float, C.c_float, vp]),
    "vnec_hip_mlp_select": (C.c_int, [vp, vp, vp, vp, vp, i64, vp, vp]),
}

# Constants of include/vnec_ext.h used by the selector bindings.
NCAND = 23
NFEAT = 32
NFEAT_INTS = 32
MLP_HIDDEN = 64
MLP_NPARAMS = NFEAT * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN \
    + MLP_HIDDEN * NCAND + NCAND
SELECT_LEARNED, SELECT_AUTOTUNE, SELECT_RULE, SELECT_REFERENCE = -1, -2, -3, -4

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    e = load().vnec_last_error()
    return e.decode() if e else ""


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")
