"""CPU oracle for the vocabulary-tree stage of the matching drivers.

TEST INFRASTRUCTURE ONLY: imported by the test suite and the benchmark's
CPU-baseline leg as the checker, never as the thing measured or shipped.

* quantise / word: V3D::VocabularyTree::findBestLeafIndex
  (V3D/include/Base/v3d_voctree.h:63-99), restated in C (v3dm_oracle_voctree.c).
* LeafDocuments (v3d_voctree.h:105-309): insertDocument (:143-164),
  getDocumentHistogram (:166-178), compute_L1_Scores (:198-240) and
  computeBinomialScores (:242-289), restated below in plain Python with the
  reference's arithmetic: doubles and the C library's log (Python's math.log)
  for the binomial score, float32 in the reference's order for L1.
* candidate_pairs: the first stage of match_using_voctree
  (SfM/match_using_voctree/src/match_using_voctree.cpp:160-207).
"""
from __future__ import annotations

import math

import numpy as np

from .oracle import _f32, _i32, _p, lib

F32 = np.float32


def tree_size(depth: int, bf: int) -> int:
    return sum(bf ** (d + 1) for d in range(depth))


def word(depth: int, bf: int, tree: np.ndarray, key: np.ndarray) -> int:
    t = np.ascontiguousarray(tree, np.float32)
    k = np.ascontiguousarray(key, np.float32)
    return int(lib().orc_voctree_word(depth, bf, _p(t, _f32), _p(k, _f32)))


def quantise(depth: int, bf: int, tree: np.ndarray, keys: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(tree, np.float32)
    k = np.ascontiguousarray(keys, np.float32).reshape(-1, 128)
    out = np.zeros(len(k), np.int32)
    lib().orc_voctree_quantize(depth, bf, _p(t, _f32), len(k), _p(k, _f32), _p(out, _i32))
    return out


class LeafDocuments:
    """V3D::LeafDocuments (v3d_voctree.h:105-309)."""

    def __init__(self, n_leaves: int):
        self.n_leaves = n_leaves
        self.inverted = [[] for _ in range(n_leaves)]  # (document, count) in insertion order
        self.n_features = []
        self.rcp_n_features = []

    def insert(self, words) -> int:
        doc = len(self.n_features)
        hist = {}
        for w in words:
            hist[int(w)] = hist.get(int(w), 0) + 1
        for leaf in sorted(hist):
            self.inverted[leaf].append((doc, hist[leaf]))
        self.n_features.append(len(words))
        self.rcp_n_features.append(F32(1.0) / F32(len(words)) if len(words) else F32(np.inf))
        return doc

    def document_histogram(self, words) -> list[int]:
        h = [0] * len(self.n_features)
        for w in words:
            for doc, _ in self.inverted[int(w)]:
                h[doc] += 1
        return h

    def binomial_scores(self, p_match: float, words, n_best: int = -1) -> list[tuple[float, int]]:
        h = self.document_histogram(words)
        normalizer = 1.0 / len(self.inverted)
        nq = len(words)
        scores = []
        for i, k in enumerate(h):
            k = float(k)
            n = float(min(self.n_features[i], nq))
            p_random = max(self.n_features[i], nq) * normalizer
            log_p0 = math.log(p_random)
            log_1_p0 = math.log(1.0 - p_random)
            p_matching = min(1.0, p_random + p_match)
            log_p1 = math.log(p_matching)
            log_1_p1 = math.log(1.0 - p_matching)
            z0 = k * log_p0 + (n - k) * log_1_p0
            z1 = k * log_p1 + (n - k) * log_1_p1
            scores.append((float(F32(z0 - z1)), i))
        return _best(scores, n_best)

    def l1_scores(self, words, n_best: int = -1) -> list[tuple[float, int]]:
        hist = {}
        for w in words:
            hist[int(w)] = hist.get(int(w), 0) + 1
        rcp = F32(1.0) / F32(len(words))
        dist = [F32(2.0)] * len(self.n_features)
        for leaf in sorted(hist):
            qi = F32(hist[leaf]) * rcp
            for doc, count in self.inverted[leaf]:
                di = F32(count) * self.rcp_n_features[doc]
                dist[doc] = F32(dist[doc] + F32(abs(F32(di - qi))))
                dist[doc] = F32(dist[doc] - di)
                dist[doc] = F32(dist[doc] - qi)
        return _best([(float(d), i) for i, d in enumerate(dist)], n_best)


def _best(scores, n_best):
    scores = sorted(scores)
    return scores if n_best < 0 or len(scores) <= n_best else scores[:n_best]


def candidate_pairs(counts, words, n_leaves: int, n_best: int, p_match: float):
    """[(earlier view, view, score)] in the driver's matching order."""
    docs = LeafDocuments(n_leaves)
    chosen = {}
    off = 0
    for i, c in enumerate(counts):
        w = [int(x) for x in words[off:off + c]]
        for score, doc in docs.binomial_scores(p_match, w, n_best):
            chosen.setdefault(doc, {})[i] = score
        docs.insert(w)
        off += c
    return [(a, b, chosen[a][b]) for a in sorted(chosen) for b in sorted(chosen[a])]
