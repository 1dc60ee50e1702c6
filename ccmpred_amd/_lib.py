"""ctypes binding of libccmpred_amd.so (the HIP library).

The library must have been built (``make -C ccmpred_amd/csrc`` or
``__graft_entry__.build()``); importing without it raises, there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# CCMPRED_AMD_LIB selects another build of the library (same ABI)
LIB_PATH = os.environ.get("CCMPRED_AMD_LIB") or os.path.join(_HERE, "lib", "libccmpred_amd.so")
BIN_PATH = os.path.join(_HERE, "bin", "ccmpred")

_lib = None


class UserData(C.Structure):
    """ccmpred.h:39-92, field for field."""
    _fields_ = [
        ("msa", C.POINTER(C.c_ubyte)),
        ("weights", C.POINTER(C.c_float)),
        ("ncol", C.c_int),
        ("nrow", C.c_int),
        ("nsingle", C.c_int),
        ("nvar", C.c_int),
        ("extra", C.c_void_p),
        ("lambda_single", C.c_float),
        ("lambda_pair", C.c_float),
        ("reweighting_threshold", C.c_float),
        ("meta_steps", C.c_void_p),
    ]


class HipConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("bf16_passes", C.c_int),
                ("bf16_fwd_passes", C.c_int), ("verbose", C.c_int), ("chunk_rows", C.c_int),
                ("gemm_variant", C.c_int), ("sparse_fwd", C.c_int)]


class CGParam(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("epsilon", C.c_double), ("k", C.c_int),
                ("max_linesearch", C.c_int), ("alpha_mul", C.c_double), ("ftol", C.c_double),
                ("wolfe", C.c_double), ("line_search", C.c_int)]


EVALUATE_T = C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int)
PROGRESS_T = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                         C.c_float, C.c_float, C.c_int, C.c_int, C.c_int)

# every symbol the public headers declare (include/ccmpred_amd.h, include/ccmpred_amd_kernels.h)
EXPORTED = [
    "ccm_nsingle_padded", "ccm_nvar_padded", "ccm_hip_default_config", "init_hip", "init_hip_ex",
    "init_hip_bf16", "evaluate_hip", "destroy_hip", "ccm_hip_evaluate", "ccm_hip_stream",
    "ccm_hip_set_timing", "ccm_hip_get_timing", "ccm_hip_step_names", "ccm_hip_objective", "ccm_hip_forward",
    "ccm_hip_backward", "ccm_hip_device_bytes", "ccm_hip_objective_resolution", "ccm_reg_only_blocks", "ccm_launch_reg_only",
    "ccm_conjugrad_init", "ccm_conjugrad_gpu", "ccm_conjugrad_gpu_ex", "ccm_read_msa",
    "ccm_init_bias", "ccm_sum_submatrices", "ccm_apc", "ccm_normalize", "ccm_write_matrix",
    "ccm_write_raw", "ccm_read_raw", "ccm_write_raw_msgpack",
    "ccm_reweight_words", "ccm_launch_identity_counts", "ccm_launch_weights_from_counts",
    "ccm_launch_fill_f32", "ccm_launch_gather_fwd_f32", "ccm_launch_gather_grad_f32",
    "ccm_launch_neg_weighted_onehot", "ccm_seg_meta_bytes", "ccm_launch_gather_grad_seg_f32", "ccm_softmax_rows_blocks", "ccm_softmax_rows_nblk",
    "ccm_softmax_cols_blocks", "ccm_softmax_cols_nblk", "ccm_softmax_cols_granule", "ccm_launch_softmax_rows",
    "ccm_launch_softmax_cols", "ccm_launch_convert_wt", "ccm_launch_gemm_fwd",
    "ccm_launch_gemm_grad", "ccm_fwd_softmax_blocks", "ccm_fwd_softmax_default_variant",
    "ccm_launch_gemm_fwd_softmax", "ccm_fwd_softmax_fits", "ccm_fwd_softmax_tile_count",
    "ccm_fwd_softmax_tile_rows", "ccm_launch_gemm_fwd_softmax_grouped", "ccm_gemm_variants", "ccm_gemm_default_variant", "ccm_pair_epilogue_blocks", "ccm_singles_blocks",
    "ccm_launch_pair_epilogue", "ccm_launch_hsym", "ccm_launch_singles_finalize",
    "ccm_launch_sum3", "ccm_dots_partial_len", "ccm_launch_dots", "ccm_launch_axpy",
    "ccm_launch_direction", "ccm_launch_negate", "ccm_contact_map_work_len", "ccm_launch_contact_map",
    "ccm_contact_map_noalloc_max_l", "ccm_launch_contact_map_noalloc",
    "ccm_sparse_ws_len", "ccm_sparse_steps", "ccm_sparse_index_ld", "ccm_sparse_fwd_fits",
    "ccm_launch_convert_ws", "ccm_launch_sparse_index", "ccm_launch_sparse_fwd_softmax",
    "ccm_convert_ws_blocks", "ccm_sparse_fwd_tile_rows", "ccm_sparse_fwd_tile_count",
    "ccm_launch_convert_ws_grouped", "ccm_launch_sparse_fwd_softmax_grouped",
    "ccm_batch_create", "ccm_batch_destroy", "ccm_batch_size", "ccm_batch_x", "ccm_batch_g",
    "ccm_batch_write_x", "ccm_batch_read", "ccm_batch_contact_map", "ccm_batch_stream", "ccm_batch_device_bytes",
    "ccm_batch_set_variants", "ccm_batch_get_variants", "ccm_batch_autotune", "ccm_batch_evaluate",
    "ccm_batch_solve", "ccm_batch_result",
    "ccm_convert_wt_blocks", "ccm_gemm_tile_count", "ccm_gemm_variant_ok",
    "ccm_gemm_default_variant_for", "ccm_gemm_variant_direct", "ccm_gemm_tile_shape", "ccm_dots_blocks", "ccm_cg_vec_blocks",
    "ccm_launch_convert_wt_grouped", "ccm_launch_gemm_fwd_grouped", "ccm_launch_gemm_grad_grouped", "ccm_launch_gemm_grad_banded_desc",
    "ccm_launch_softmax_cols_grouped", "ccm_launch_pair_epilogue_grouped",
    "ccm_launch_singles_finalize_grouped", "ccm_launch_sum3_grouped", "ccm_launch_dots_grouped",
    "ccm_launch_cg_vec_grouped", "ccm_launch_cg_ctrl", "ccm_launch_pad_scan",
    "ccm_launch_pad_scan_x", "ccm_launch_dots_padded", "ccm_launch_axpy_padded",
    "ccm_launch_direction_padded", "ccm_launch_negate_padded",
    "ccm_assemble_blocks", "ccm_launch_assemble", "ccm_launch_assemble_grouped",
    "ccm_convert_ws_reg_parts", "ccm_launch_convert_ws_reg", "ccm_assemble_fs_blocks",
    "ccm_launch_assemble_fs", "ccm_launch_sparse_fwd_softmax_pad",
]


def lib():
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime under the same soname as the system
        # one; load it first so that this library binds to the same runtime
        # instead of pulling in a second copy that torch then fails to use.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C ccmpred_amd/csrc` "
                "(there is no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, ip = C.c_void_p, C.c_int
    L.ccm_nsingle_padded.argtypes = [ip]
    L.ccm_nsingle_padded.restype = C.c_long
    L.ccm_nvar_padded.argtypes = [ip]
    L.ccm_nvar_padded.restype = C.c_long
    L.ccm_hip_default_config.argtypes = [C.POINTER(HipConfig)]
    L.init_hip.argtypes = [vp]
    L.init_hip.restype = ip
    L.init_hip_ex.argtypes = [vp, C.POINTER(HipConfig)]
    L.init_hip_ex.restype = ip
    L.init_hip_bf16.argtypes = [vp]
    L.init_hip_bf16.restype = ip
    L.evaluate_hip.argtypes = [vp, vp, vp, ip]
    L.evaluate_hip.restype = C.c_float
    L.destroy_hip.argtypes = [vp]
    L.destroy_hip.restype = ip
    L.ccm_hip_evaluate.argtypes = [vp, vp, vp, ip, C.POINTER(C.c_double)]
    L.ccm_hip_evaluate.restype = ip
    L.ccm_hip_objective_resolution.argtypes = [vp]
    L.ccm_hip_objective_resolution.restype = C.c_double
    L.ccm_hip_stream.argtypes = [vp]
    L.ccm_hip_stream.restype = vp
    L.ccm_hip_objective.argtypes = [vp, vp, ip, C.POINTER(C.c_double)]
    L.ccm_hip_objective.restype = ip
    L.ccm_hip_forward.argtypes = [vp, vp, ip, C.POINTER(C.c_double)]
    L.ccm_hip_forward.restype = ip
    L.ccm_hip_backward.argtypes = [vp, vp, vp, ip]
    L.ccm_hip_backward.restype = ip
    L.ccm_hip_device_bytes.argtypes = [vp]
    L.ccm_hip_device_bytes.restype = C.c_size_t
    L.ccm_hip_set_timing.argtypes = [vp, ip]
    L.ccm_hip_set_timing.restype = ip
    L.ccm_hip_get_timing.argtypes = [vp, C.POINTER(C.c_double), ip]
    L.ccm_hip_get_timing.restype = ip
    L.ccm_hip_step_names.argtypes = [vp, C.POINTER(C.c_char_p), ip]
    L.ccm_hip_step_names.restype = ip
    L.ccm_conjugrad_init.argtypes = [C.POINTER(CGParam)]
    L.ccm_conjugrad_gpu_ex.argtypes = [ip, vp, C.POINTER(C.c_double), EVALUATE_T, vp, vp,
                                       C.POINTER(CGParam), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_double)]
    L.ccm_conjugrad_gpu_ex.restype = ip
    L.ccm_init_bias.argtypes = [vp, vp, ip, ip]
    L.ccm_sum_submatrices.argtypes = [vp, vp, ip]
    L.ccm_apc.argtypes = [vp, ip]
    L.ccm_normalize.argtypes = [vp, ip]
    L.ccm_read_raw.argtypes = [C.c_char_p, ip, vp]
    L.ccm_read_raw.restype = ip
    # batched form
    L.ccm_batch_create.argtypes = [vp, ip, C.POINTER(HipConfig), C.POINTER(C.c_void_p)]
    L.ccm_batch_create.restype = ip
    L.ccm_batch_destroy.argtypes = [vp]
    L.ccm_batch_destroy.restype = None
    L.ccm_batch_size.argtypes = [vp]
    L.ccm_batch_size.restype = ip
    L.ccm_batch_x.argtypes = [vp, ip]
    L.ccm_batch_x.restype = vp
    L.ccm_batch_g.argtypes = [vp, ip]
    L.ccm_batch_g.restype = vp
    L.ccm_batch_write_x.argtypes = [vp, ip, vp]
    L.ccm_batch_write_x.restype = ip
    L.ccm_batch_read.argtypes = [vp, ip, vp, vp]
    L.ccm_batch_read.restype = ip
    L.ccm_batch_contact_map.argtypes = [vp, ip, ip, ip, vp]
    L.ccm_batch_contact_map.restype = ip
    L.ccm_contact_map_work_len.argtypes = [ip]
    L.ccm_contact_map_work_len.restype = ip
    L.ccm_launch_contact_map.argtypes = [vp, ip, ip, ip, vp, vp, vp]
    L.ccm_launch_contact_map.restype = ip
    L.ccm_contact_map_noalloc_max_l.argtypes = []
    L.ccm_contact_map_noalloc_max_l.restype = ip
    L.ccm_launch_contact_map_noalloc.argtypes = [vp, ip, ip, ip, vp, vp]
    L.ccm_launch_contact_map_noalloc.restype = ip
    L.ccm_batch_stream.argtypes = [vp]
    L.ccm_batch_stream.restype = vp
    L.ccm_batch_device_bytes.argtypes = [vp]
    L.ccm_batch_device_bytes.restype = C.c_size_t
    L.ccm_batch_set_variants.argtypes = [vp, vp, vp]
    L.ccm_batch_set_variants.restype = ip
    L.ccm_batch_get_variants.argtypes = [vp, vp, vp]
    L.ccm_batch_get_variants.restype = ip
    L.ccm_batch_autotune.argtypes = [vp, ip, vp]
    L.ccm_batch_autotune.restype = ip
    L.ccm_batch_evaluate.argtypes = [vp, vp]
    L.ccm_batch_evaluate.restype = ip
    L.ccm_batch_solve.argtypes = [vp, C.POINTER(CGParam), ip, C.POINTER(C.c_int)]
    L.ccm_batch_solve.restype = ip
    L.ccm_batch_result.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), vp]
    L.ccm_batch_result.restype = ip
    L.ccm_gemm_tile_count.argtypes = [ip, ip, ip]
    L.ccm_gemm_tile_count.restype = C.c_uint32
    L.ccm_gemm_variant_ok.argtypes = [ip, ip, ip, ip]
    L.ccm_gemm_variant_ok.restype = ip
    L.ccm_gemm_default_variant_for.argtypes = [ip, ip, ip]
    L.ccm_gemm_default_variant_for.restype = ip


def evaluate_fnptr():
    """The library's evaluate_hip as a C function pointer usable as a CG callback."""
    return C.cast(lib().evaluate_hip, EVALUATE_T)
