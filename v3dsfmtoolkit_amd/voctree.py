"""Vocabulary tree and inverted file (include/v3dm_voctree.h).

``VocabularyTree`` mirrors V3D::VocabularyTree<Depth, BranchFactor, 128>
(V3D/include/Base/v3d_voctree.h:19-103) with the quantisation on the GPU:
``getVisualWords`` returns, for many keys at once, the words the CPU tree's
getVisualWord gives (bit-identical).  ``LeafDocuments`` mirrors
V3D::LeafDocuments (:105-309) on the host.  ``candidate_pairs`` is the first
stage of SfM/match_using_voctree (match_using_voctree.cpp:160-207).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import V3DM_ERR_CAPACITY, V3DM_OK, V3DMError, check, lib, ptr


def tree_nodes(depth: int, branch_factor: int) -> int:
    return sum(branch_factor ** (d + 1) for d in range(depth))


def synthetic_tree(seed: int, depth: int, branch_factor: int, spread: float = 0.35) -> np.ndarray:
    """Synthetic node descriptors for tests and benchmarks (no trained tree
    ships with the reference): unit-norm, non-negative, each child a
    perturbation of its parent, roughly what hierarchical k-means on SIFT
    descriptors produces.  Returns (tree_nodes, 128) float32, level by level."""
    rng = np.random.default_rng(seed)
    tree = np.zeros((tree_nodes(depth, branch_factor), 128), np.float32)
    parents = np.abs(rng.normal(size=(1, 128)))
    start = 0
    for _ in range(depth):
        c = np.abs(parents[:, None, :] + spread * rng.normal(size=(len(parents), branch_factor, 128)))
        c = c.reshape(-1, 128)
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        tree[start:start + len(c)] = c
        start += len(c)
        parents = c
    return tree


class VocabularyTree:
    """A vocabulary tree resident on HIP device ``device``."""

    def __init__(self, depth: int = 3, branch_factor: int = 50, nodes=None, path=None, device: int = 0):
        h = ctypes.c_void_p()
        if path is not None:
            rc = lib().v3dm_voctree_load(device, depth, branch_factor, os.fsencode(os.fspath(path)), ctypes.byref(h))
        else:
            n = np.ascontiguousarray(nodes, np.float32)
            if n.size != tree_nodes(depth, branch_factor) * 128:
                raise ValueError(f"expected {tree_nodes(depth, branch_factor)} x 128 node values, got {n.shape}")
            rc = lib().v3dm_voctree_create(device, depth, branch_factor, ptr(n, ctypes.c_float), ctypes.byref(h))
        if rc != V3DM_OK:
            raise V3DMError(rc, lib().v3dm_voctree_last_error(None).decode(errors="replace"))
        self._h = h
        self.depth, self.branch_factor = depth, branch_factor
        self._n = 0

    def _check(self, rc):
        if rc != V3DM_OK:
            raise V3DMError(rc, lib().v3dm_voctree_last_error(self._h).decode(errors="replace"))

    AUTO, EXACT_VALU = 0, 1  # scoring modes (same words): MFMA filter + exact rescoring, or exact sums only

    def setMode(self, mode: int) -> None:
        self._check(lib().v3dm_voctree_set_mode(self._h, mode))

    def visualWordCount(self) -> int:
        leaves = ctypes.c_int()
        self._check(lib().v3dm_voctree_info(self._h, None, None, None, ctypes.byref(leaves)))
        return leaves.value

    def getVisualWords(self, keys) -> np.ndarray:
        """getVisualWord for every row of ``keys`` (n x 128 float32)."""
        self.upload(keys)
        self.run()
        return self.fetch()

    def upload(self, keys) -> None:
        k = np.ascontiguousarray(keys, np.float32).reshape(-1, 128)
        self._check(lib().v3dm_voctree_upload(self._h, len(k), ptr(k, ctypes.c_float)))
        self._n = len(k)

    def run(self) -> None:
        self._check(lib().v3dm_voctree_run(self._h))

    def fetch(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), np.int32)
        self._check(lib().v3dm_voctree_fetch(self._h, ptr(out, ctypes.c_int32)))
        return out[:self._n]

    def runTimes(self) -> tuple[float, float]:
        """(scoring kernels ms, whole run ms) of the last run."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(lib().v3dm_voctree_run_times(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self) -> None:
        if self._h:
            lib().v3dm_voctree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _scores(call, cap):
    s = np.zeros(max(cap, 1), np.float32)
    d = np.zeros(max(cap, 1), np.int32)
    n = ctypes.c_int()
    check(call(ptr(s, ctypes.c_float), ptr(d, ctypes.c_int32), cap, ctypes.byref(n)))
    return list(zip(s[:n.value].tolist(), d[:n.value].tolist()))


class LeafDocuments:
    """V3D::LeafDocuments: an inverted file of visual words over documents (views)."""

    def __init__(self, n_leaves: int):
        h = ctypes.c_void_p()
        check(lib().v3dm_leaf_documents_create(n_leaves, ctypes.byref(h)))
        self._h = h

    def __len__(self) -> int:
        n = ctypes.c_int()
        check(lib().v3dm_leaf_documents_count(self._h, ctypes.byref(n)))
        return n.value

    def insertDocument(self, words) -> int:
        w = np.ascontiguousarray(words, np.int32)
        doc = ctypes.c_int()
        check(lib().v3dm_leaf_documents_insert(self._h, len(w), ptr(w, ctypes.c_int32), ctypes.byref(doc)))
        return doc.value

    def computeBinomialScores(self, p_match: float, words, n_best: int = -1) -> list[tuple[float, int]]:
        w = np.ascontiguousarray(words, np.int32)
        return _scores(lambda s, d, cap, n: lib().v3dm_leaf_documents_binomial_scores(
            self._h, p_match, len(w), ptr(w, ctypes.c_int32), n_best, s, d, cap, n), len(self))

    def compute_L1_Scores(self, words, n_best: int = -1) -> list[tuple[float, int]]:
        w = np.ascontiguousarray(words, np.int32)
        return _scores(lambda s, d, cap, n: lib().v3dm_leaf_documents_l1_scores(
            self._h, len(w), ptr(w, ctypes.c_int32), n_best, s, d, cap, n), len(self))

    def close(self) -> None:
        if self._h:
            lib().v3dm_leaf_documents_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def candidate_pairs(counts, words, n_leaves: int, n_best: int = 20, p_match: float = 0.1):
    """[(earlier view, view, score)] as match_using_voctree's first stage picks them.

    ``p_match`` is used as given; the reference driver passes the float
    VISUAL_WORD_MATCH_PROPABILITY, i.e. float(np.float32(0.1)) for the default."""
    c = np.ascontiguousarray(counts, np.int32)
    w = np.ascontiguousarray(words, np.int32)
    n = ctypes.c_int()
    rc = lib().v3dm_voctree_candidate_pairs(len(c), ptr(c, ctypes.c_int32), ptr(w, ctypes.c_int32), n_leaves, n_best,
                                            p_match, None, None, 0, ctypes.byref(n))
    if rc not in (V3DM_OK, V3DM_ERR_CAPACITY):
        check(rc)
    pairs = np.zeros((max(n.value, 1), 2), np.int32)
    scores = np.zeros(max(n.value, 1), np.float32)
    check(lib().v3dm_voctree_candidate_pairs(len(c), ptr(c, ctypes.c_int32), ptr(w, ctypes.c_int32), n_leaves, n_best,
                                             p_match, ptr(pairs, ctypes.c_int32), ptr(scores, ctypes.c_float), n.value,
                                             ctypes.byref(n)))
    return [(int(a), int(b), float(s)) for (a, b), s in zip(pairs[:n.value], scores[:n.value])]


__all__ = ["VocabularyTree", "LeafDocuments", "candidate_pairs", "tree_nodes", "synthetic_tree"]
