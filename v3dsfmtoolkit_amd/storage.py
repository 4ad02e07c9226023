"""Pairwise-match records and the match database, through the library
(include/v3dm_io.h).

PairwiseMatch (SfM/reconstruction/include/reconstruction_common.h:101-119) is
what the reference drivers store per matched view pair: the view pair, a
rotation, a translation and the point correspondences.  Its blob is the
reference's BlobOArchive byte stream, stored with the ViewPair in the SQLite
tables matches_data / matches_list (V3D/src/Base/v3d_storage.cpp:70-216).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib, ptr

# PointCorrespondence (V3D/include/Geometry/v3d_mviewutilities.h:45-68) as it
# serialises: left (x, y, view, id) then right, 32 bytes.
CORRESPONDENCE = np.dtype([("left_x", "<f4"), ("left_y", "<f4"), ("left_view", "<i4"), ("left_id", "<i4"),
                           ("right_x", "<f4"), ("right_y", "<f4"), ("right_view", "<i4"), ("right_id", "<i4")])
assert CORRESPONDENCE.itemsize == 32


@dataclass
class PairwiseMatch:
    views: tuple[int, int]
    rotation: np.ndarray = field(default_factory=lambda: np.eye(3))        # (3, 3) float64, row-major
    translation: np.ndarray = field(default_factory=lambda: np.zeros(3))   # (3,) float64
    corrs: np.ndarray = field(default_factory=lambda: np.zeros(0, CORRESPONDENCE))


def correspondences(n: int) -> np.ndarray:
    return np.zeros(n, CORRESPONDENCE)


def _corr_array(corrs) -> np.ndarray:
    c = np.ascontiguousarray(corrs)
    if c.dtype != CORRESPONDENCE:
        raise TypeError("correspondences must be a CORRESPONDENCE array")
    return c


def pairwise_match_size(n_corrs: int) -> int:
    return int(lib().v3dm_pairwise_match_size(n_corrs))


def encode_pairwise_match(m: PairwiseMatch) -> bytes:
    """BlobOArchive << PairwiseMatch."""
    c = _corr_array(m.corrs)
    R = np.ascontiguousarray(m.rotation, np.float64).reshape(9)
    T = np.ascontiguousarray(m.translation, np.float64).reshape(3)
    size = ctypes.c_size_t()
    buf = np.zeros(pairwise_match_size(len(c)), np.uint8)
    check(lib().v3dm_encode_pairwise_match(int(m.views[0]), int(m.views[1]), ptr(R, ctypes.c_double),
                                           ptr(T, ctypes.c_double), len(c), c.ctypes.data if len(c) else None,
                                           ptr(buf, ctypes.c_uint8), buf.nbytes, ctypes.byref(size)))
    return buf[:size.value].tobytes()


def decode_pairwise_match(blob: bytes) -> PairwiseMatch:
    """BlobIArchive >> PairwiseMatch."""
    b = np.frombuffer(blob, np.uint8).copy()
    v0, v1, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    R, T = np.zeros(9), np.zeros(3)
    rc = lib().v3dm_decode_pairwise_match(ptr(b, ctypes.c_uint8), b.nbytes, ctypes.byref(v0), ctypes.byref(v1),
                                          ptr(R, ctypes.c_double), ptr(T, ctypes.c_double), None, 0, ctypes.byref(n))
    c = correspondences(n.value)
    if n.value or rc != 0:
        check(lib().v3dm_decode_pairwise_match(ptr(b, ctypes.c_uint8), b.nbytes, ctypes.byref(v0), ctypes.byref(v1),
                                               ptr(R, ctypes.c_double), ptr(T, ctypes.c_double),
                                               c.ctypes.data if n.value else None, n.value, ctypes.byref(n)))
    return PairwiseMatch((v0.value, v1.value), R.reshape(3, 3), T, c)


class MatchDatabase:
    """SQLite3_Database with the tables matches_data / matches_list, as the
    reference drivers create them (match_sequence.cpp:125-127); rows are
    committed by close()."""

    def __init__(self, path, truncate: bool = True):
        h = ctypes.c_void_p()
        check(lib().v3dm_match_db_open(os.fsencode(os.fspath(path)), int(truncate), ctypes.byref(h)))
        self._h = h

    def put(self, oid: int, m: PairwiseMatch) -> None:
        """updateObject("matches_data", oid, m) and updateObject("matches_list", oid, m.views)."""
        c = _corr_array(m.corrs)
        R = np.ascontiguousarray(m.rotation, np.float64).reshape(9)
        T = np.ascontiguousarray(m.translation, np.float64).reshape(3)
        check(lib().v3dm_match_db_put(self._h, int(oid), int(m.views[0]), int(m.views[1]), ptr(R, ctypes.c_double),
                                      ptr(T, ctypes.c_double), len(c), c.ctypes.data if len(c) else None))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, None
            check(lib().v3dm_match_db_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ = ["CORRESPONDENCE", "PairwiseMatch", "correspondences", "pairwise_match_size", "encode_pairwise_match",
           "decode_pairwise_match", "MatchDatabase"]
