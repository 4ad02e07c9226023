"""The matching drivers' bookkeeping and their batched matching step.

The reference drivers (SfM/match_sequence/src/match_sequence.cpp:134-227,
SfM/match_using_voctree/src/match_using_voctree.cpp:251-323) load each view's
features, keep min(MATCHING_MAX_FEATURES, size & 0xffe0) of them, and call
setRightFeatures + findPutativeMatches once per view pair, turning each
(left, right) index pair into a PointCorrespondence of feature ids and pixel
positions.  ``match_views`` does the same for a whole pair list as batched
jobs on the GPU; the native counterpart application is
``v3dsfmtoolkit_amd/bin/v3dm_match_sequence``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import V3DM_ERR_CAPACITY, V3DM_OK, V3DMError, check, lib, ptr
from .features import Features
from .storage import CORRESPONDENCE

SCORE_THRESHOLD = 0.9  # match_sequence.cpp:79, match_using_voctree.cpp:102


def truncated_feature_count(size: int, n_max: int) -> int:
    """min(nMax, int(size & 0xffe0)) (match_sequence.cpp:152,184); wraps at 65536 as the reference does."""
    return int(lib().v3dm_truncated_feature_count(int(size), int(n_max)))


def sequence_pairs(n_views: int, radius: int) -> np.ndarray:
    """(i, j) view pairs of match_sequence in its loop order: j = i+1 .. i+2*radius-1, j < n_views."""
    n = ctypes.c_int()
    rc = lib().v3dm_sequence_pairs(n_views, radius, None, 0, ctypes.byref(n))
    if rc not in (V3DM_OK, V3DM_ERR_CAPACITY):
        check(rc)
    out = np.zeros((n.value, 2), np.int32)
    check(lib().v3dm_sequence_pairs(n_views, radius, ptr(out, ctypes.c_int32), n.value, ctypes.byref(n)))
    return out


def descriptors_to_u8(desc: np.ndarray):
    """uint8 u when every descriptor value is exactly u/512 (0 <= u <= 255), else None."""
    d = np.ascontiguousarray(desc, np.float32).reshape(-1, 128)
    out = np.zeros(d.shape, np.uint8)
    rc = lib().v3dm_descriptors_to_u8(len(d), ptr(d, ctypes.c_float), ptr(out, ctypes.c_uint8))
    return out if rc == V3DM_OK else None


def kept_count(n_features: int, n_max: int) -> int:
    """Features a driver passes to the matcher: the truncation, then setLeft/RightFeatures'
    rounding to a multiple of 32 (which matters only for an nMax that is not one)."""
    return truncated_feature_count(n_features, n_max) & ~31


def correspondences_from_table(table: np.ndarray, v1: int, v2: int, f1: Features, f2: Features) -> np.ndarray:
    """The PointCorrespondences the drivers build from a match list (match_sequence.cpp:205-222),
    in ascending left order."""
    left = np.nonzero(table >= 0)[0]
    right = table[left]
    c = np.zeros(len(left), CORRESPONDENCE)
    c["left_x"], c["left_y"] = f1.positions[left, 0], f1.positions[left, 1]
    c["left_view"], c["left_id"] = v1, f1.ids[left]
    c["right_x"], c["right_y"] = f2.positions[right, 0], f2.positions[right, 1]
    c["right_view"], c["right_id"] = v2, f2.ids[right]
    return c


def match_views(matcher, views: list[Features], pairs, n_max: int = 7168, min_score: float = SCORE_THRESHOLD,
                min_ratio: float = 0.8, pairs_per_job: int = 4096) -> list[np.ndarray]:
    """Putative correspondences of every (left view, right view) pair, as batched jobs on
    ``matcher`` (a SIFTMatching).  Uint8-representable descriptors use the exact int8 pool,
    others the BF16 + FP32-rescoring pool; either way each pair's list equals
    findPutativeMatches on that pair."""
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    kept = [kept_count(len(f), n_max) for f in views]
    floats = [np.ascontiguousarray(f.descriptors[:k], np.float32) for f, k in zip(views, kept)]
    u8 = [descriptors_to_u8(d) for d in floats]
    images = u8 if all(x is not None for x in u8) else floats
    matcher.setImages(images)
    out = []
    for p0 in range(0, len(pairs), pairs_per_job):
        chunk = pairs[p0:p0 + pairs_per_job]
        matcher.setPairs([tuple(map(int, p)) for p in chunk])
        tables = matcher.batchMatch(min_score, min_ratio)
        for (a, b), t in zip(chunk, tables):
            out.append(correspondences_from_table(np.asarray(t), int(a), int(b), views[a], views[b]))
    return out


__all__ = ["SCORE_THRESHOLD", "truncated_feature_count", "sequence_pairs", "descriptors_to_u8", "kept_count",
           "correspondences_from_table", "match_views", "V3DMError"]
