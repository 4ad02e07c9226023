"""ctypes wrapper of the CPU oracle (oracle/v3dm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
CPU-baseline leg of the benchmark, always as the checker and never as the
thing measured for the GPU path or shipped.  See v3dm_oracle.c for the
reference lines each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32 = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_f32 = ctypes.POINTER(ctypes.c_float)
_f64 = ctypes.POINTER(ctypes.c_double)
_int = ctypes.POINTER(ctypes.c_int)

_SIG = {
    "orc_scores_u8": (None, [ctypes.c_int, _u8, ctypes.c_int, _u8, _i32]),
    "orc_scores_u8_as_f32": (None, [ctypes.c_int, _u8, ctypes.c_int, _u8, _f32]),
    "orc_scores_f32": (None, [ctypes.c_int, _f32, ctypes.c_int, _f32, _f32]),
    "orc_scores_f64": (None, [ctypes.c_int, _f32, ctypes.c_int, _f32, _f64]),
    "orc_scores_f32_simple": (None, [ctypes.c_int, _f32, ctypes.c_int, _f32, _f32]),
    "orc_scores_f64_simple": (None, [ctypes.c_int, _f32, ctypes.c_int, _f32, _f64]),
    "orc_epipolar_masked": (ctypes.c_int, [_f32, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float]),
    "orc_homography_masked": (ctypes.c_int, [_f32, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_float]),
    "orc_apply_masks_f32": (None, [ctypes.c_int, ctypes.c_int, _f32, ctypes.c_int, _int, _f32, _f32, _f32, _f32]),
    "orc_column_pass": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32, _i32, _f32, _f32]),
    "orc_row_pass": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32, _i32, _f32, _f32]),
    "orc_mutual": (ctypes.c_int, [ctypes.c_int, _i32, _i32, _i32]),
    "orc_mutual_table": (None, [ctypes.c_int, _i32, _i32, _i32]),
    "orc_select_f32": (None, [ctypes.c_int, ctypes.c_int, _f32, ctypes.c_float, ctypes.c_float, _i32, _i32, _i32]),
    "orc_find_u8": (ctypes.c_int, [ctypes.c_int, _u8, ctypes.c_int, _u8, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_int, _int, _f32, _f32, _f32, _f32, _i32, _i32, _i32]),
    "orc_find_f32": (ctypes.c_int, [ctypes.c_int, _f32, ctypes.c_int, _f32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_int, _int, _f32, _f32, _f32, _f32, _i32, _i32, _i32]),
    "orc_find_f64": (ctypes.c_int, [ctypes.c_int, _f32, ctypes.c_int, _f32, ctypes.c_double, ctypes.c_double,
                                    _i32, _i32, _i32, _f64, _f64]),
    "orc_select_f64": (None, [ctypes.c_int, ctypes.c_int, _f64, ctypes.c_double, ctypes.c_double, _i32, _i32, _i32,
                              _f64, _f64]),
    "orc_find_u8_blocked": (ctypes.c_int, [ctypes.c_int, _u8, ctypes.c_int, _u8, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_int, _i32, _i32, _i32]),
    "orc_max_threads": (ctypes.c_int, []),
    "orc_voctree_word": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _f32, _f32]),
    "orc_voctree_quantize": (None, [ctypes.c_int, ctypes.c_int, _f32, ctypes.c_int64, _f32, _i32]),
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _masks(masks):
    """masks: list of (kind, 3x3 matrix, dist, lpos (n,2), rpos (m,2)).  The
    oracle takes one position set for all masks, so every mask must carry the
    same positions (as the callers do)."""
    if not masks:
        return 0, None, None, None, None, None
    kinds = np.array([m[0] for m in masks], np.int32)
    mats = np.ascontiguousarray(np.array([np.asarray(m[1], np.float64).reshape(9).astype(np.float32)
                                          for m in masks], np.float32))
    dist = np.array([m[2] for m in masks], np.float32)
    lpos = np.ascontiguousarray(np.asarray(masks[0][3], np.float32).reshape(-1, 2))
    rpos = np.ascontiguousarray(np.asarray(masks[0][4], np.float32).reshape(-1, 2))
    for m in masks[1:]:
        assert np.array_equal(np.asarray(m[3], np.float32).reshape(-1, 2), lpos)
        assert np.array_equal(np.asarray(m[4], np.float32).reshape(-1, 2), rpos)
    return len(masks), kinds, mats, dist, lpos, rpos


def find_u8(L: np.ndarray, R: np.ndarray, min_score: float = 0.9, min_ratio: float = 0.8, masks=None):
    """Reference rule on u8 descriptors (floats u/512).  Returns
    (colMax, rowMax, matchOfLeft) int32 arrays."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    nL, nR = L.shape[0], R.shape[0]
    col = np.empty(nL, np.int32)
    row = np.empty(nR, np.int32)
    tab = np.empty(nL, np.int32)
    nm, kinds, mats, dist, lpos, rpos = _masks(masks)
    rc = lib().orc_find_u8(nL, _p(L, _u8), nR, _p(R, _u8), min_score, min_ratio, nm, _p(kinds, _int),
                           _p(mats, _f32), _p(dist, _f32), _p(lpos, _f32), _p(rpos, _f32), _p(col, _i32),
                           _p(row, _i32), _p(tab, _i32))
    assert rc == 0
    return col, row, tab


def find_f32(L: np.ndarray, R: np.ndarray, min_score: float = 0.9, min_ratio: float = 0.8, masks=None):
    L = np.ascontiguousarray(L, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    nL, nR = L.shape[0], R.shape[0]
    col = np.empty(nL, np.int32)
    row = np.empty(nR, np.int32)
    tab = np.empty(nL, np.int32)
    nm, kinds, mats, dist, lpos, rpos = _masks(masks)
    rc = lib().orc_find_f32(nL, _p(L, _f32), nR, _p(R, _f32), min_score, min_ratio, nm, _p(kinds, _int),
                            _p(mats, _f32), _p(dist, _f32), _p(lpos, _f32), _p(rpos, _f32), _p(col, _i32),
                            _p(row, _i32), _p(tab, _i32))
    assert rc == 0
    return col, row, tab


def find_f64(L: np.ndarray, R: np.ndarray, min_score: float = 0.9, min_ratio: float = 0.8):
    L = np.ascontiguousarray(L, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    nL, nR = L.shape[0], R.shape[0]
    col = np.empty(nL, np.int32)
    row = np.empty(nR, np.int32)
    tab = np.empty(nL, np.int32)
    best = np.empty(nL, np.float64)
    nxt = np.empty(nL, np.float64)
    rc = lib().orc_find_f64(nL, _p(L, _f32), nR, _p(R, _f32), min_score, min_ratio, _p(col, _i32), _p(row, _i32),
                            _p(tab, _i32), _p(best, _f64), _p(nxt, _f64))
    assert rc == 0
    return col, row, tab, best, nxt


def find_u8_blocked(L: np.ndarray, R: np.ndarray, min_score: float = 0.9, min_ratio: float = 0.8,
                    threads: int = 0):
    """Blocked multithreaded rule (the CPU baseline).  Returns
    (colMax, rowMax, matchOfLeft, threads_used)."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    nL, nR = L.shape[0], R.shape[0]
    col = np.empty(nL, np.int32)
    row = np.empty(nR, np.int32)
    tab = np.empty(nL, np.int32)
    used = lib().orc_find_u8_blocked(nL, _p(L, _u8), nR, _p(R, _u8), min_score, min_ratio, threads,
                                     _p(col, _i32), _p(row, _i32), _p(tab, _i32))
    assert used > 0
    return col, row, tab, used


def scores_u8(L: np.ndarray, R: np.ndarray) -> np.ndarray:
    """Exact integer scores, reference layout S[r, l]."""
    L = np.ascontiguousarray(L, np.uint8)
    R = np.ascontiguousarray(R, np.uint8)
    S = np.empty((R.shape[0], L.shape[0]), np.int32)
    lib().orc_scores_u8(L.shape[0], _p(L, _u8), R.shape[0], _p(R, _u8), _p(S, _i32))
    return S


def scores_f32(L: np.ndarray, R: np.ndarray) -> np.ndarray:
    L = np.ascontiguousarray(L, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    S = np.empty((R.shape[0], L.shape[0]), np.float32)
    lib().orc_scores_f32(L.shape[0], _p(L, _f32), R.shape[0], _p(R, _f32), _p(S, _f32))
    return S


def scores_f32_simple(L: np.ndarray, R: np.ndarray) -> np.ndarray:
    """scores_f32 one dot product at a time (the plain statement of the order)."""
    L = np.ascontiguousarray(L, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    S = np.empty((R.shape[0], L.shape[0]), np.float32)
    lib().orc_scores_f32_simple(L.shape[0], _p(L, _f32), R.shape[0], _p(R, _f32), _p(S, _f32))
    return S


def scores_f64(L: np.ndarray, R: np.ndarray, simple: bool = False) -> np.ndarray:
    """FP64 scores S[r, l] of float descriptors (blocked, or one at a time)."""
    L = np.ascontiguousarray(L, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    S = np.empty((R.shape[0], L.shape[0]), np.float64)
    f = lib().orc_scores_f64_simple if simple else lib().orc_scores_f64
    f(L.shape[0], _p(L, _f32), R.shape[0], _p(R, _f32), _p(S, _f64))
    return S


def select_f64(S: np.ndarray, min_score: float, min_ratio: float):
    """The FP64 rule (find_f64's passes) on an FP64 score matrix S[r, l];
    returns (colMax, rowMax, matchOfLeft)."""
    S = np.ascontiguousarray(S, np.float64)
    nR, nL = S.shape
    col = np.empty(nL, np.int32)
    row = np.empty(nR, np.int32)
    tab = np.empty(nL, np.int32)
    lib().orc_select_f64(nL, nR, _p(S, _f64), min_score, min_ratio, _p(col, _i32), _p(row, _i32), _p(tab, _i32),
                         None, None)
    return col, row, tab


def select_f32(S: np.ndarray, min_score: float, min_ratio: float):
    """Column + row passes + mutual on an explicit score matrix S[r, l]."""
    S = np.ascontiguousarray(S, np.float32)
    nR, nL = S.shape
    col = np.empty(nL, np.int32)
    row = np.empty(nR, np.int32)
    tab = np.empty(nL, np.int32)
    lib().orc_select_f32(nL, nR, _p(S, _f32), min_score, min_ratio, _p(col, _i32), _p(row, _i32), _p(tab, _i32))
    return col, row, tab


def column_pass(S: np.ndarray, min_score: float, min_ratio: float):
    S = np.ascontiguousarray(S, np.float32)
    nR, nL = S.shape
    arg = np.empty(nL, np.int32)
    best = np.empty(nL, np.float32)
    nxt = np.empty(nL, np.float32)
    lib().orc_column_pass(nR, nL, min_score, min_ratio, _p(S, _f32), _p(arg, _i32), _p(best, _f32), _p(nxt, _f32))
    return arg, best, nxt


def row_pass(S: np.ndarray, min_score: float, min_ratio: float):
    S = np.ascontiguousarray(S, np.float32)
    nR, nL = S.shape
    arg = np.empty(nR, np.int32)
    best = np.empty(nR, np.float32)
    nxt = np.empty(nR, np.float32)
    lib().orc_row_pass(nR, nL, min_score, min_ratio, _p(S, _f32), _p(arg, _i32), _p(best, _f32), _p(nxt, _f32))
    return arg, best, nxt


def epipolar_masked(F, sqr_thr, p, q) -> bool:
    Fm = np.ascontiguousarray(np.asarray(F, np.float64).reshape(9).astype(np.float32))
    return bool(lib().orc_epipolar_masked(_p(Fm, _f32), sqr_thr, p[0], p[1], q[0], q[1]))


def homography_masked(H, sqr_thr, p, q) -> bool:
    Hm = np.ascontiguousarray(np.asarray(H, np.float64).reshape(9).astype(np.float32))
    return bool(lib().orc_homography_masked(_p(Hm, _f32), sqr_thr, p[0], p[1], q[0], q[1]))


def table_to_pairs(tab: np.ndarray) -> list[tuple[int, int]]:
    return [(int(k), int(l)) for k, l in enumerate(tab) if l >= 0]


def max_threads() -> int:
    return int(lib().orc_max_threads())
