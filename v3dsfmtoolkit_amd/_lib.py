"""ctypes binding of libv3dmatch.so (the C ABI in include/v3dmatch.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C v3dsfmtoolkit_amd/csrc``).  Loading fails loudly when it is
missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libv3dmatch.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")
HEADER_PATH = os.path.join(INCLUDE_DIR, "v3dmatch.h")
HEADER_PATHS = [os.path.join(INCLUDE_DIR, h) for h in ("v3dmatch.h", "v3dm_io.h", "v3dm_voctree.h")]

V3DM_OK = 0
V3DM_ERR_ARG = -1
V3DM_ERR_STATE = -2
V3DM_ERR_DEVICE = -3
V3DM_ERR_NOMEM = -4
V3DM_ERR_UNSUPPORTED = -5
V3DM_ERR_CAPACITY = -6

KERNEL_AUTO = 0
KERNEL_PLAIN = 1
KERNEL_I8 = 2
KERNEL_BF16 = 3

_c_int_p = ctypes.POINTER(ctypes.c_int)
_i32_p = ctypes.POINTER(ctypes.c_int32)
_u8_p = ctypes.POINTER(ctypes.c_uint8)
_f32_p = ctypes.POINTER(ctypes.c_float)
_f64_p = ctypes.POINTER(ctypes.c_double)
_i64_p = ctypes.POINTER(ctypes.c_int64)
_ctx_p = ctypes.c_void_p

_SIGNATURES = {
    "v3dm_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_destroy": (None, [_ctx_p]),
    "v3dm_last_error": (ctypes.c_char_p, [_ctx_p]),
    "v3dm_set_rounding_warning": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_set_reduction_path": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_allocate": (ctypes.c_int, [_ctx_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "v3dm_deallocate": (ctypes.c_int, [_ctx_p]),
    "v3dm_set_left_f32": (ctypes.c_int, [_ctx_p, ctypes.c_int, _f32_p]),
    "v3dm_set_right_f32": (ctypes.c_int, [_ctx_p, ctypes.c_int, _f32_p]),
    "v3dm_set_left_u8": (ctypes.c_int, [_ctx_p, ctypes.c_int, _u8_p]),
    "v3dm_set_right_u8": (ctypes.c_int, [_ctx_p, ctypes.c_int, _u8_p]),
    "v3dm_get_counts": (ctypes.c_int, [_ctx_p, _c_int_p, _c_int_p]),
    "v3dm_get_device_buffers": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                               ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                               ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_find_putative_matches": (ctypes.c_int, [_ctx_p, ctypes.c_float, ctypes.c_float, _i32_p, ctypes.c_int, _c_int_p]),
    "v3dm_run_guided_matching": (ctypes.c_int, [_ctx_p, _f64_p, ctypes.c_float, _f32_p, _f32_p, ctypes.c_float,
                                                ctypes.c_float, _i32_p, ctypes.c_int, _c_int_p]),
    "v3dm_run_guided_homography_matching": (ctypes.c_int, [_ctx_p, _f64_p, ctypes.c_float, _f32_p, _f32_p,
                                                           ctypes.c_float, ctypes.c_float, _i32_p, ctypes.c_int,
                                                           _c_int_p]),
    "v3dm_set_kernel": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_get_last_kernel": (ctypes.c_int, [_ctx_p, _c_int_p]),
    "v3dm_get_stream": (ctypes.c_int, [_ctx_p, ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_find_async": (ctypes.c_int, [_ctx_p, ctypes.c_float, ctypes.c_float]),
    "v3dm_fetch_matches": (ctypes.c_int, [_ctx_p, _i32_p, ctypes.c_int, _c_int_p]),
    "v3dm_synchronize": (ctypes.c_int, [_ctx_p]),
    "v3dm_profile_enable": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_profile_read": (ctypes.c_int, [_ctx_p, _f64_p, _f64_p, _c_int_p]),
    "v3dm_batch_set_images_u8": (ctypes.c_int, [_ctx_p, ctypes.c_int, _i32_p, _u8_p]),
    "v3dm_batch_set_images_f32": (ctypes.c_int, [_ctx_p, ctypes.c_int, _i32_p, _f32_p]),
    "v3dm_batch_set_pairs": (ctypes.c_int, [_ctx_p, ctypes.c_int, _i32_p]),
    "v3dm_batch_set_positions": (ctypes.c_int, [_ctx_p, _f32_p]),
    "v3dm_batch_set_guides": (ctypes.c_int, [_ctx_p, _i32_p, _f64_p, _f32_p]),
    "v3dm_batch_match": (ctypes.c_int, [_ctx_p, ctypes.c_float, ctypes.c_float, _i32_p, _i64_p]),
    "v3dm_batch_match_async": (ctypes.c_int, [_ctx_p, ctypes.c_float, ctypes.c_float]),
    "v3dm_batch_fetch": (ctypes.c_int, [_ctx_p, _i32_p, _i64_p]),
    "v3dm_batch_set_variant": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_batch_get_variant": (ctypes.c_int, [_ctx_p, _c_int_p]),
    "v3dm_batch_autotune": (ctypes.c_int, [_ctx_p, ctypes.c_float, ctypes.c_float, ctypes.c_int, _f64_p, _c_int_p]),
    "v3dm_batch_set_graph": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_batch_info": (ctypes.c_int, [_ctx_p, _c_int_p, _i64_p, _c_int_p, _i64_p]),
    "v3dm_scan_variant_count": (ctypes.c_int, []),
    "v3dm_scan_variant_count_f32": (ctypes.c_int, []),
    "v3dm_set_insert_form": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "v3dm_find_u8_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "v3dm_get_last_insert_form": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "v3dm_scan_variant_info": (ctypes.c_int, [ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]),
    "v3dm_synth_pair": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, _u8_p, _u8_p, _f32_p, _f32_p, _i32_p]),
    "v3dm_synth_seed": (ctypes.c_uint64, [ctypes.c_int, ctypes.c_int]),
    "v3dm_device_count": (ctypes.c_int, [_c_int_p]),
    "v3dm_mfma_layout_probe": (ctypes.c_int, [ctypes.c_int, _i32_p]),
    "v3dm_scan_variant_form": (ctypes.c_int, [ctypes.c_int, _i32_p]),
    "v3dm_mfma_ceiling_probe": (ctypes.c_int, [ctypes.c_int, _u8_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double)]),
    "v3dm_copy_probe": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int]),
    "v3dm_set_scan_variant": (ctypes.c_int, [_ctx_p, ctypes.c_int]),
    "v3dm_get_last_scan_variant": (ctypes.c_int, [_ctx_p, _c_int_p]),
    "v3dm_get_last_scan_inline": (ctypes.c_int, [_ctx_p, _c_int_p]),
    "v3dm_fetch_left_top2": (ctypes.c_int, [_ctx_p, ctypes.c_int, _f32_p, _f32_p, _i32_p]),
    "v3dm_mlp_param_count": (ctypes.c_int, []),
    "v3dm_mlp_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_mlp_destroy": (None, [ctypes.c_void_p]),
    "v3dm_mlp_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "v3dm_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _f32_p]),
    "v3dm_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _f32_p]),
    "v3dm_mlp_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "v3dm_mlp_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "v3dm_mlp_loss_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _u8_p, _u8_p, ctypes.c_float, _f32_p,
                                          _f32_p]),
    "v3dm_mlp_set_training_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _u8_p, _u8_p]),
    "v3dm_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_uint64, _f32_p]),
    "v3dm_mlp_embed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _u8_p, _f32_p]),
    "v3dm_learned_find": (ctypes.c_int, [_ctx_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, _i32_p, ctypes.c_int, _c_int_p, _i64_p]),
    "v3dm_learned_candidates": (ctypes.c_int, [_ctx_p, ctypes.c_int, _i32_p, _c_int_p]),
    # include/v3dm_io.h (host only)
    "v3dm_features_info": (ctypes.c_int, [ctypes.c_char_p, _c_int_p, _c_int_p]),
    "v3dm_features_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _i32_p, _f32_p, _f32_p, _f32_p, _f32_p,
                                          _f32_p, _c_int_p]),
    "v3dm_features_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _i32_p, _f32_p, _f32_p, _f32_p, _f32_p,
                                           _f32_p, ctypes.c_int]),
    "v3dm_truncated_feature_count": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int]),
    "v3dm_sequence_pairs": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i32_p, ctypes.c_int, _c_int_p]),
    "v3dm_descriptors_to_u8": (ctypes.c_int, [ctypes.c_int64, _f32_p, _u8_p]),
    "v3dm_pairwise_match_size": (ctypes.c_size_t, [ctypes.c_int]),
    "v3dm_encode_pairwise_match": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _f64_p, _f64_p, ctypes.c_int,
                                                  ctypes.c_void_p, _u8_p, ctypes.c_size_t,
                                                  ctypes.POINTER(ctypes.c_size_t)]),
    "v3dm_decode_pairwise_match": (ctypes.c_int, [_u8_p, ctypes.c_size_t, _c_int_p, _c_int_p, _f64_p, _f64_p,
                                                  ctypes.c_void_p, ctypes.c_int, _c_int_p]),
    "v3dm_match_db_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_match_db_put": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64_p, _f64_p,
                                         ctypes.c_int, ctypes.c_void_p]),
    "v3dm_match_db_close": (ctypes.c_int, [ctypes.c_void_p]),
    "v3dm_fundamental_from_pose": (ctypes.c_int, [_f64_p, _f64_p, _f64_p, _f64_p, _f64_p]),
    "v3dm_normalize_correspondences": (ctypes.c_int, [_f64_p, _f64_p, ctypes.c_int, ctypes.c_void_p]),
    "v3dm_effective_inlier_count": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                   _c_int_p]),
    "v3dm_filter_in_front": (ctypes.c_int, [_f64_p, _f64_p, _f64_p, _f64_p, ctypes.c_int, ctypes.c_void_p,
                                            _c_int_p]),
    "v3dm_calibration_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _f64_p, _f64_p, _i32_p, _c_int_p]),
    "v3dm_sampson_inliers": (ctypes.c_int, [_f64_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, _i32_p,
                                            _c_int_p]),
    # include/v3dm_voctree.h
    "v3dm_voctree_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32_p,
                                           ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_voctree_load": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_voctree_destroy": (None, [ctypes.c_void_p]),
    "v3dm_voctree_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "v3dm_voctree_info": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "v3dm_voctree_quantize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _f32_p, _i32_p]),
    "v3dm_voctree_set_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "v3dm_voctree_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _f32_p]),
    "v3dm_voctree_run": (ctypes.c_int, [ctypes.c_void_p]),
    "v3dm_voctree_fetch": (ctypes.c_int, [ctypes.c_void_p, _i32_p]),
    "v3dm_voctree_run_times": (ctypes.c_int, [ctypes.c_void_p, _f64_p, _f64_p]),
    "v3dm_leaf_documents_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "v3dm_leaf_documents_destroy": (None, [ctypes.c_void_p]),
    "v3dm_leaf_documents_insert": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _i32_p, _c_int_p]),
    "v3dm_leaf_documents_count": (ctypes.c_int, [ctypes.c_void_p, _c_int_p]),
    "v3dm_leaf_documents_histogram": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _i32_p, _i32_p, ctypes.c_int]),
    "v3dm_leaf_documents_sizes": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _i64_p]),
    "v3dm_leaf_documents_export": (ctypes.c_int, [ctypes.c_void_p, _i32_p, _i32_p, _i32_p, _i32_p, _f32_p]),
    "v3dm_leaf_documents_import": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _i32_p, _i32_p, _i32_p,
                                                  _i32_p, _f32_p]),
    "v3dm_leaf_documents_binomial_scores": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _i32_p,
                                                           ctypes.c_int, _f32_p, _i32_p, ctypes.c_int, _c_int_p]),
    "v3dm_leaf_documents_l1_scores": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _i32_p, ctypes.c_int, _f32_p,
                                                     _i32_p, ctypes.c_int, _c_int_p]),
    "v3dm_voctree_candidate_pairs": (ctypes.c_int, [ctypes.c_int, _i32_p, _i32_p, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_double, _i32_p, _f32_p, ctypes.c_int, _c_int_p]),
}

_lib = None


def header_symbols(paths=None) -> list[str]:
    """Names of every function declared in include/v3dmatch.h and include/v3dm_io.h."""
    names = set()
    for path in paths or HEADER_PATHS:
        text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
        names.update(re.findall(r"\b(v3dm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def lib() -> ctypes.CDLL:
    """Load libv3dmatch.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C v3dsfmtoolkit_amd/csrc`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def ptr(a: np.ndarray, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class V3DMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"v3dmatch error {code}: {msg}")
        self.code = code


def check(code: int, ctx=None) -> None:
    if code != V3DM_OK:
        raise V3DMError(code, lib().v3dm_last_error(ctx).decode(errors="replace"))


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().v3dm_device_count(ctypes.byref(n)))
    return n.value


def synth_seed(config: int, pair_index: int) -> int:
    return int(lib().v3dm_synth_seed(config, pair_index))


def synth_pair(seed: int, n_left: int, n_right: int, rho: float = 0.4, sigma: float = 0.05, *,
               u8: bool = True, f32: bool = False, threads: int = 0):
    """Synthetic SIFT-like pair.  Returns a dict with 'L', 'R' (uint8),
    'Lf', 'Rf' (float32, unquantised) as requested, and 'src' (right -> left
    source row or -1)."""
    out = {}
    L = np.empty((n_left, 128), np.uint8) if u8 else None
    R = np.empty((n_right, 128), np.uint8) if u8 else None
    Lf = np.empty((n_left, 128), np.float32) if f32 else None
    Rf = np.empty((n_right, 128), np.float32) if f32 else None
    src = np.empty(n_right, np.int32)
    check(lib().v3dm_synth_pair(seed, n_left, n_right, rho, sigma, threads, ptr(L, ctypes.c_uint8),
                                ptr(R, ctypes.c_uint8), ptr(Lf, ctypes.c_float), ptr(Rf, ctypes.c_float),
                                ptr(src, ctypes.c_int32)))
    if u8:
        out["L"], out["R"] = L, R
    if f32:
        out["Lf"], out["Rf"] = Lf, Rf
    out["src"] = src
    return out


def mfma_layout_probe(device: int = 0) -> np.ndarray:
    out = np.zeros((64, 16), np.int32)
    check(lib().v3dm_mfma_layout_probe(device, ptr(out, ctypes.c_int32)))
    return out


def mfma_ceiling_probe(rows: np.ndarray, from_lds: bool = True, iters: int = 20000, device: int = 0) -> float:
    """TOP/s of the int8 scan's step loop alone on the given 128 stored rows
    (uint8 view of the scan's u ^ 0x80 bytes); see v3dm_mfma_ceiling_probe."""
    r = np.ascontiguousarray(rows, dtype=np.uint8)
    if r.shape != (128, 128):
        raise ValueError("mfma_ceiling_probe: rows must be 128 x 128 bytes")
    out = ctypes.c_double(0.0)
    check(lib().v3dm_mfma_ceiling_probe(device, ptr(r, ctypes.c_uint8), 1 if from_lds else 0, int(iters),
                                         ctypes.byref(out)))
    return out.value


def copy_probe(nbytes: int, reps: int = 5, device: int = 0) -> None:
    """Plain device copy of nbytes (read once, written once) per launch."""
    check(lib().v3dm_copy_probe(device, nbytes, reps))


def scan_variants_f32() -> int:
    """Number of scan variants (the first ones) that also run float descriptors."""
    return int(lib().v3dm_scan_variant_count_f32())


def scan_variant_whole(variant: int) -> bool:
    """True for the whole-tile scan form (k_i8_scan_whole), False for the ring form."""
    w = ctypes.c_int32(0)
    check(lib().v3dm_scan_variant_form(variant, ctypes.byref(w)))
    return bool(w.value)


def scan_variants() -> list[tuple[int, int, int]]:
    """(waves per workgroup, tile rows, tile columns) of each scan variant."""
    out = []
    for v in range(lib().v3dm_scan_variant_count()):
        w, r, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().v3dm_scan_variant_info(v, ctypes.byref(w), ctypes.byref(r), ctypes.byref(c)))
        out.append((w.value, r.value, c.value))
    return out
