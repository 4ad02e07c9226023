"""The host-side arithmetic around the drivers' robust orientation step,
through the library (include/v3dm_io.h).

After the putative matches, the reference drivers estimate each pair's
relative pose (RANSAC, outside this library), form the fundamental matrix
F = K2^-T [T]x R K1^-1 for guided matching, check the effective inlier count,
and store the correspondences with positions normalised by K^-1
(SfM/match_sequence/src/match_sequence.cpp:229-346,
SfM/match_using_voctree/src/match_using_voctree.cpp:325-491).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import V3DM_ERR_CAPACITY, V3DM_OK, check, lib, ptr
from .storage import CORRESPONDENCE


def _m(a, n) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, np.float64).reshape(n))


def _corrs(c) -> np.ndarray:
    c = np.ascontiguousarray(c)
    if c.dtype != CORRESPONDENCE:
        raise TypeError("correspondences must have the CORRESPONDENCE dtype")
    return c


def fundamental_from_pose(K1, K2, R, T) -> np.ndarray:
    """F = invertedMatrix(K2)^T (crossProductMatrix(T) R) invertedMatrix(K1) with the
    reference's double operations (match_sequence.cpp:279-283); (3, 3) float64."""
    F = np.zeros(9)
    check(lib().v3dm_fundamental_from_pose(ptr(_m(K1, 9), ctypes.c_double), ptr(_m(K2, 9), ctypes.c_double),
                                           ptr(_m(R, 9), ctypes.c_double), ptr(_m(T, 3), ctypes.c_double),
                                           ptr(F, ctypes.c_double)))
    return F.reshape(3, 3)


def normalize_correspondences(K1, K2, corrs) -> np.ndarray:
    """Positions mapped by K1^-1 (left) and K2^-1 (right), as the drivers store them
    (match_sequence.cpp:327-338); returns a new array."""
    c = _corrs(corrs).copy()
    check(lib().v3dm_normalize_correspondences(ptr(_m(K1, 9), ctypes.c_double), ptr(_m(K2, 9), ctypes.c_double),
                                               len(c), c.ctypes.data_as(ctypes.c_void_p)))
    return c


def effective_inlier_count(corrs, width: int, height: int) -> int:
    """getEffectiveInlierCount (match_sequence.cpp:34-50): the inlier count scaled by the
    mean image coverage of the left and right positions."""
    c = _corrs(corrs)
    n = ctypes.c_int()
    check(lib().v3dm_effective_inlier_count(len(c), c.ctypes.data_as(ctypes.c_void_p), int(width), int(height),
                                            ctypes.byref(n)))
    return n.value


def filter_in_front(K1, K2, R, T, corrs) -> np.ndarray:
    """The correspondences whose linear triangulation lies in front of view 1
    (match_using_voctree.cpp:397-418); returns a new array."""
    c = _corrs(corrs).copy()
    kept = ctypes.c_int()
    check(lib().v3dm_filter_in_front(ptr(_m(K1, 9), ctypes.c_double), ptr(_m(K2, 9), ctypes.c_double),
                                     ptr(_m(R, 9), ctypes.c_double), ptr(_m(T, 3), ctypes.c_double), len(c),
                                     c.ctypes.data_as(ctypes.c_void_p), ctypes.byref(kept)))
    return c[:kept.value].copy()


def sampson_inliers(F, threshold: float, corrs) -> np.ndarray:
    """Indices of the correspondences within `threshold` (squared Sampson distance
    <= threshold^2, double): the drivers' known-pose inlier rule."""
    c = _corrs(corrs)
    out = np.zeros(max(len(c), 1), np.int32)
    n = ctypes.c_int()
    check(lib().v3dm_sampson_inliers(ptr(_m(F, 9), ctypes.c_double), float(threshold), len(c),
                                     c.ctypes.data_as(ctypes.c_void_p), ptr(out, ctypes.c_int32), ctypes.byref(n)))
    return out[:n.value].copy()


def read_calibration(path) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """CalibrationDatabase (SfM/reconstruction/src/reconstruction_common.cpp:17-38):
    K (n, 3, 3), distortion (n, 4), image size (n, 2) per view."""
    n = ctypes.c_int()
    rc = lib().v3dm_calibration_read(str(path).encode(), 0, None, None, None, ctypes.byref(n))
    if rc not in (V3DM_OK, V3DM_ERR_CAPACITY):
        check(rc)
    K = np.zeros((max(n.value, 1), 9))
    dist = np.zeros((max(n.value, 1), 4))
    dims = np.zeros((max(n.value, 1), 2), np.int32)
    check(lib().v3dm_calibration_read(str(path).encode(), n.value, ptr(K, ctypes.c_double),
                                      ptr(dist, ctypes.c_double), ptr(dims, ctypes.c_int32), ctypes.byref(n)))
    return K[:n.value].reshape(-1, 3, 3), dist[:n.value], dims[:n.value]


def write_calibration(path, K, dims, distortion=None) -> None:
    """The text form CalibrationDatabase reads: the view count, then per view
    K[0][0] K[0][1] K[0][2] K[1][1] K[1][2], four distortion terms, width, height."""
    K = np.asarray(K, np.float64).reshape(-1, 3, 3)
    dist = np.zeros((len(K), 4)) if distortion is None else np.asarray(distortion, np.float64).reshape(-1, 4)
    with open(path, "w") as f:
        f.write(f"{len(K)}\n")
        for k, d, (w, h) in zip(K, dist, np.asarray(dims).reshape(-1, 2)):
            vals = [k[0, 0], k[0, 1], k[0, 2], k[1, 1], k[1, 2], *d]
            f.write(" ".join(repr(float(x)) for x in vals) + f" {int(w)} {int(h)}\n")


__all__ = ["fundamental_from_pose", "normalize_correspondences", "effective_inlier_count", "filter_in_front",
           "sampson_inliers", "read_calibration", "write_calibration"]
