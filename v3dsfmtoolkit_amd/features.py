"""SIFT feature files (``<image>.features``) through the library's reader and
writer (include/v3dm_io.h).

A file holds a ``SerializableVector<SIFT_Feature>`` in the reference's binary
("V3DBIN") or text ("V3DTXT") archive (V3D/include/Base/v3d_serialization.h:884-947,
V3D/include/Base/v3d_feature.h:12-49): per feature an int id, position (x, y),
direction (x, y), scale, magnitude and a 128-float descriptor.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, ptr

BINARY = 1  # V3DM_FEATURES_BINARY
TEXT = 0    # V3DM_FEATURES_TEXT


@dataclass
class Features:
    """One image's features, field for field as SIFT_Feature holds them."""
    ids: np.ndarray          # (n,) int32
    positions: np.ndarray    # (n, 2) float32
    directions: np.ndarray   # (n, 2) float32
    scales: np.ndarray       # (n,) float32
    magnitudes: np.ndarray   # (n,) float32
    descriptors: np.ndarray  # (n, 128) float32

    def __len__(self) -> int:
        return len(self.ids)


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def feature_file_info(path) -> tuple[int, int]:
    """(record count, format) of a feature file; format is BINARY or TEXT."""
    n, fmt = ctypes.c_int(), ctypes.c_int()
    check(lib().v3dm_features_info(_path(path), ctypes.byref(n), ctypes.byref(fmt)))
    return n.value, fmt.value


def read_features(path) -> Features:
    """serializeDataFromFile(path, SerializableVector<SIFT_Feature>&)."""
    n, _ = feature_file_info(path)
    f = Features(np.zeros(n, np.int32), np.zeros((n, 2), np.float32), np.zeros((n, 2), np.float32),
                 np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros((n, 128), np.float32))
    got = ctypes.c_int()
    check(lib().v3dm_features_read(_path(path), n, ptr(f.ids, ctypes.c_int32), ptr(f.positions, ctypes.c_float),
                                         ptr(f.directions, ctypes.c_float), ptr(f.scales, ctypes.c_float),
                                         ptr(f.magnitudes, ctypes.c_float), ptr(f.descriptors, ctypes.c_float),
                                         ctypes.byref(got)))
    assert got.value == n
    return f


def write_features(path, descriptors, ids=None, positions=None, directions=None, scales=None, magnitudes=None,
                   binary: bool = True) -> None:
    """serializeDataToFile(path, SerializableVector<SIFT_Feature>, binary).

    Missing fields are written as zeros (ids as 0..n-1)."""
    d = np.ascontiguousarray(descriptors, np.float32).reshape(-1, 128)
    n = len(d)

    def arr(a, shape, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype)
        if a.size != int(np.prod(shape)):
            raise ValueError(f"expected {shape} values, got {a.shape}")
        return a

    ids = arr(ids, (n,), np.int32)
    positions = arr(positions, (n, 2), np.float32)
    directions = arr(directions, (n, 2), np.float32)
    scales = arr(scales, (n,), np.float32)
    magnitudes = arr(magnitudes, (n,), np.float32)
    check(lib().v3dm_features_write(_path(path), n, ptr(ids, ctypes.c_int32), ptr(positions, ctypes.c_float),
                                          ptr(directions, ctypes.c_float), ptr(scales, ctypes.c_float),
                                          ptr(magnitudes, ctypes.c_float), ptr(d, ctypes.c_float),
                                          BINARY if binary else TEXT))


__all__ = ["Features", "feature_file_info", "read_features", "write_features", "BINARY", "TEXT"]
