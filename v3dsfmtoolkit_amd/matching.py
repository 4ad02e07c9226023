"""Python mirror of V3D_CUDA::SIFT_Matching over the C ABI.

Same method names, argument meaning and defaults as the reference class
(V3D/include/CUDA/v3d_cudamatching.h:15-67): allocate / deallocate /
setLeftFeatures / setRightFeatures / findPutativeMatches /
runGuidedMatching / runGuidedHomographyMatching / enableRoundingWarning /
enableReductionPath.  Match lists are lists of (left, right) index tuples
in ascending left order.  Device errors raise V3DMError (the reference
exits the process instead, v3d_cudamatching.cuh:9-24).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr


class SIFTMatching:
    """One matcher instance bound to one GPU (one HIP stream)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.lib()
        h = ctypes.c_void_p()
        check(self._lib.v3dm_create(device, ctypes.byref(h)))
        self._ctx = h
        self._cap = 0

    # -- reference API ------------------------------------------------------
    def enableRoundingWarning(self, flag: bool = True) -> None:
        check(self._lib.v3dm_set_rounding_warning(self._ctx, int(flag)), self._ctx)

    def enableReductionPath(self, flag: bool = True) -> None:
        check(self._lib.v3dm_set_reduction_path(self._ctx, int(flag)), self._ctx)

    def allocate(self, nMaxLeftFeatures: int, nMaxRightFeatures: int | bool | None = None,
                 useGuidedMatching: bool = False) -> None:
        # allocate(nMax, useGuided) and allocate(nMaxL, nMaxR, useGuided)
        if nMaxRightFeatures is None or isinstance(nMaxRightFeatures, bool):
            if isinstance(nMaxRightFeatures, bool):
                useGuidedMatching = nMaxRightFeatures
            nMaxRightFeatures = nMaxLeftFeatures
        check(self._lib.v3dm_allocate(self._ctx, int(nMaxLeftFeatures), int(nMaxRightFeatures),
                                      int(useGuidedMatching)), self._ctx)
        self._cap = max(int(nMaxLeftFeatures), 1)

    def deallocate(self) -> None:
        check(self._lib.v3dm_deallocate(self._ctx), self._ctx)

    def setLeftFeatures(self, nFeatures: int, descriptors: np.ndarray) -> None:
        self._set(True, nFeatures, descriptors)

    def setRightFeatures(self, nFeatures: int, descriptors: np.ndarray) -> None:
        self._set(False, nFeatures, descriptors)

    def findPutativeMatches(self, minScore: float = 0.9, minRatio: float = 0.8) -> list[tuple[int, int]]:
        pairs, count = self._buffer()
        check(self._lib.v3dm_find_putative_matches(self._ctx, minScore, minRatio, ptr(pairs, ctypes.c_int32),
                                                   self._cap, ctypes.byref(count)), self._ctx)
        return _as_list(pairs, count.value)

    def runGuidedMatching(self, F, distanceThreshold: float, leftPositions, rightPositions,
                          minScore: float = 0.9, minRatio: float = 0.8) -> list[tuple[int, int]]:
        return self._guided(self._lib.v3dm_run_guided_matching, F, distanceThreshold, leftPositions,
                            rightPositions, minScore, minRatio)

    def runGuidedHomographyMatching(self, H, distanceThreshold: float, leftPositions, rightPositions,
                                    minScore: float = 0.9, minRatio: float = 0.8) -> list[tuple[int, int]]:
        return self._guided(self._lib.v3dm_run_guided_homography_matching, H, distanceThreshold, leftPositions,
                            rightPositions, minScore, minRatio)

    # -- additions ------------------------------------------------------------
    def setKernel(self, kernel: int) -> None:
        check(self._lib.v3dm_set_kernel(self._ctx, int(kernel)), self._ctx)

    def setScanVariant(self, variant: int) -> None:
        """Single-pair int8 scan variant; -1 (default) measures the fastest per shape bucket."""
        check(self._lib.v3dm_set_scan_variant(self._ctx, int(variant)), self._ctx)

    def lastScanVariant(self) -> int:
        k = ctypes.c_int(-1)
        check(self._lib.v3dm_get_last_scan_variant(self._ctx, ctypes.byref(k)), self._ctx)
        return k.value

    def lastScanInline(self) -> int:
        """1 if the last single-pair int8 job's scan applied its survivors itself, 0 if
        the insertion kernel did, -1 before any job."""
        k = ctypes.c_int(-1)
        check(self._lib.v3dm_get_last_scan_inline(self._ctx, ctypes.byref(k)), self._ctx)
        return k.value

    def setInsertForm(self, form: int) -> None:
        """Insertion pass of large int8 jobs: -1 automatic, 0 one wave per region, 1 one wave per tile."""
        check(self._lib.v3dm_set_insert_form(self._ctx, int(form)), self._ctx)

    def lastInsertForm(self) -> int:
        """Form of the last int8 job's insertion pass (0 / 1), -1 when the scan applied its survivors."""
        k = ctypes.c_int(-1)
        check(self._lib.v3dm_get_last_insert_form(self._ctx, ctypes.byref(k)), self._ctx)
        return k.value

    def lastKernel(self) -> int:
        k = ctypes.c_int(0)
        check(self._lib.v3dm_get_last_kernel(self._ctx, ctypes.byref(k)), self._ctx)
        return k.value

    def counts(self) -> tuple[int, int]:
        a, b = ctypes.c_int(0), ctypes.c_int(0)
        check(self._lib.v3dm_get_counts(self._ctx, ctypes.byref(a), ctypes.byref(b)), self._ctx)
        return a.value, b.value

    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(self._lib.v3dm_get_stream(self._ctx, ctypes.byref(s)), self._ctx)
        return s.value or 0

    def findAsync(self, minScore: float = 0.9, minRatio: float = 0.8) -> None:
        check(self._lib.v3dm_find_async(self._ctx, minScore, minRatio), self._ctx)

    def fetchMatches(self) -> list[tuple[int, int]]:
        pairs, count = self._buffer()
        check(self._lib.v3dm_fetch_matches(self._ctx, ptr(pairs, ctypes.c_int32), self._cap, ctypes.byref(count)),
              self._ctx)
        return _as_list(pairs, count.value)

    def matchTable(self, pairs: list[tuple[int, int]] | None = None) -> np.ndarray:
        """Per-left table (right index or -1) of the given or last list."""
        n_left, _ = self.counts()
        out = np.full(n_left, -1, np.int32)
        for k, l in (pairs if pairs is not None else self.fetchMatches()):
            out[k] = l
        return out

    def profileEnable(self, enable: bool = True) -> None:
        check(self._lib.v3dm_profile_enable(self._ctx, int(enable)), self._ctx)

    def profileRead(self) -> tuple[float, float, int]:
        """(scan ms, select ms, jobs) summed over the jobs since profileEnable."""
        a, b, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int(0)
        check(self._lib.v3dm_profile_read(self._ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)), self._ctx)
        return a.value, b.value, n.value

    def leftTop2(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(best, second, argBest) per left feature of the last top-2 job,
        over the candidates above the survivor threshold (-1e30 / -1 where none)."""
        n_left = self.counts()[0]
        best = np.empty(n_left, np.float32)
        second = np.empty(n_left, np.float32)
        arg = np.empty(n_left, np.int32)
        check(self._lib.v3dm_fetch_left_top2(self._ctx, n_left, ptr(best, ctypes.c_float), ptr(second, ctypes.c_float),
                                             ptr(arg, ctypes.c_int32)), self._ctx)
        return best, second, arg

    # -- learned proposal path ----------------------------------------------------
    def learnedFind(self, net: "EmbeddingNet", margin: float = 0.05, minScore: float = 0.9,
                    minRatio: float = 0.8, topK: int = 16, asArray: bool = False):
        """Matches the current pair through the learned proposal path (both
        sides uploaded as uint8).  Returns (pairs, stats); pairs as an (n, 2)
        int32 array when asArray."""
        pairs, count = self._buffer()
        st = np.zeros(6, np.int64)
        check(self._lib.v3dm_learned_find(self._ctx, net._h, topK, margin, minScore, minRatio,
                                          ptr(pairs, ctypes.c_int32),
                                          self._cap, ctypes.byref(count), ptr(st, ctypes.c_int64)), self._ctx)
        keys = ["n_left", "n_right", "candidates", "flagged_left", "flagged_right", "fallback_pairs"]
        stats = {k: int(x) for k, x in zip(keys, st)}
        if asArray:
            return pairs[:count.value].copy(), stats
        return _as_list(pairs, count.value), stats

    def learnedCandidates(self) -> np.ndarray:
        """(n_left, K) candidate right indices of the last learned job."""
        n = self.counts()[0]
        out = np.empty((n, 32), np.int32)
        k = ctypes.c_int(0)
        check(self._lib.v3dm_learned_candidates(self._ctx, n, ptr(out, ctypes.c_int32), ctypes.byref(k)), self._ctx)
        return out.reshape(-1)[:n * k.value].reshape(n, k.value).copy()

    # -- many pairs per job (v3dm_batch_*) ---------------------------------------
    def setImages(self, images) -> list[int]:
        """Uploads descriptor sets (list of (n_i, 128) arrays, all uint8 or all
        float32); returns the counts after rounding to multiples of 32."""
        counts = np.array([len(x) for x in images], np.int32)
        is_float = len(images) > 0 and np.asarray(images[0]).dtype != np.uint8
        dt = np.float32 if is_float else np.uint8
        desc = np.ascontiguousarray(np.concatenate([np.asarray(x, dt).reshape(-1, 128) for x in images])
                                    if len(images) else np.zeros((0, 128), dt))
        if is_float:
            check(self._lib.v3dm_batch_set_images_f32(self._ctx, len(images), ptr(counts, ctypes.c_int32),
                                                      ptr(desc, ctypes.c_float)), self._ctx)
        else:
            check(self._lib.v3dm_batch_set_images_u8(self._ctx, len(images), ptr(counts, ctypes.c_int32),
                                                     ptr(desc, ctypes.c_uint8)), self._ctx)
        self._img_counts = [int(c) - int(c) % 32 for c in counts]
        return self._img_counts

    def setPairs(self, pairs) -> None:
        pairs = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        check(self._lib.v3dm_batch_set_pairs(self._ctx, len(pairs), ptr(pairs, ctypes.c_int32)), self._ctx)
        self._pairs = pairs

    def setPositions(self, positions) -> None:
        """Feature positions of the uploaded images (list of (n_i, 2) float arrays, the
        counts of setImages): what runGuidedMatching takes as left/right positions."""
        pos = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).reshape(-1, 2) for x in positions])
                                   if len(positions) else np.zeros((0, 2), np.float32))
        check(self._lib.v3dm_batch_set_positions(self._ctx, ptr(pos, ctypes.c_float)), self._ctx)

    def setGuides(self, kinds=None, matrices=None, thresholds=None) -> None:
        """Per-pair guided masks of the current pair list: kind 0 none, 1 epipolar (F),
        2 homography (H); matrices (nPairs, 3, 3) float64, thresholds (nPairs,) float32.
        kinds=None clears them."""
        if kinds is None:
            check(self._lib.v3dm_batch_set_guides(self._ctx, None, None, None), self._ctx)
            return
        k = np.ascontiguousarray(np.asarray(kinds, np.int32).reshape(-1))
        m = np.ascontiguousarray(np.asarray(matrices, np.float64).reshape(-1, 9))
        t = np.ascontiguousarray(np.asarray(thresholds, np.float32).reshape(-1))
        if not (len(k) == len(m) == len(t) == len(self._pairs)):
            raise ValueError("setGuides: one kind, matrix and threshold per pair")
        check(self._lib.v3dm_batch_set_guides(self._ctx, ptr(k, ctypes.c_int32), ptr(m, ctypes.c_double),
                                              ptr(t, ctypes.c_float)), self._ctx)

    def batchInfo(self) -> dict:
        n, nl, nt, ps = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int64()
        check(self._lib.v3dm_batch_info(self._ctx, ctypes.byref(n), ctypes.byref(nl), ctypes.byref(nt),
                                        ctypes.byref(ps)), self._ctx)
        return {"pairs": n.value, "left_total": nl.value, "tiles": nt.value, "pairs_scored": ps.value}

    def batchMatchAsync(self, minScore: float = 0.9, minRatio: float = 0.8) -> None:
        check(self._lib.v3dm_batch_match_async(self._ctx, minScore, minRatio), self._ctx)

    def batchFetch(self) -> list[np.ndarray]:
        info = self.batchInfo()
        tab = np.empty(max(info["left_total"], 1), np.int32)
        off = np.empty(info["pairs"] + 1, np.int64)
        check(self._lib.v3dm_batch_fetch(self._ctx, ptr(tab, ctypes.c_int32), ptr(off, ctypes.c_int64)), self._ctx)
        return [tab[off[p]:off[p + 1]].copy() for p in range(info["pairs"])]

    def batchMatch(self, minScore: float = 0.9, minRatio: float = 0.8) -> list[np.ndarray]:
        """Per pair: table (right index or -1 per left feature)."""
        self.batchMatchAsync(minScore, minRatio)
        return self.batchFetch()

    def batchSetVariant(self, variant: int) -> None:
        check(self._lib.v3dm_batch_set_variant(self._ctx, int(variant)), self._ctx)

    def batchVariant(self) -> int:
        v = ctypes.c_int()
        check(self._lib.v3dm_batch_get_variant(self._ctx, ctypes.byref(v)), self._ctx)
        return v.value

    def batchAutotune(self, minScore: float = 0.9, minRatio: float = 0.8, reps: int = 3):
        n = self._lib.v3dm_scan_variant_count()
        ms = np.zeros(n, np.float64)
        best = ctypes.c_int()
        check(self._lib.v3dm_batch_autotune(self._ctx, minScore, minRatio, reps, ptr(ms, ctypes.c_double),
                                            ctypes.byref(best)), self._ctx)
        return best.value, ms

    def batchSetGraph(self, enable: bool = True) -> None:
        check(self._lib.v3dm_batch_set_graph(self._ctx, int(enable)), self._ctx)

    def synchronize(self) -> None:
        check(self._lib.v3dm_synchronize(self._ctx), self._ctx)

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self._lib.v3dm_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- helpers ---------------------------------------------------------------
    def _buffer(self):
        return np.empty((max(self._cap, 1), 2), np.int32), ctypes.c_int(0)

    def _set(self, left: bool, n: int, desc: np.ndarray) -> None:
        desc = np.ascontiguousarray(desc)
        if desc.ndim != 2 or desc.shape[1] != 128:
            raise ValueError("descriptors must be an (n, 128) array")
        if int(n) > desc.shape[0]:
            raise ValueError("nFeatures exceeds the number of descriptor rows")
        if desc.dtype == np.uint8:
            fn = self._lib.v3dm_set_left_u8 if left else self._lib.v3dm_set_right_u8
            check(fn(self._ctx, int(n), ptr(desc, ctypes.c_uint8)), self._ctx)
        else:
            desc = np.ascontiguousarray(desc, dtype=np.float32)
            fn = self._lib.v3dm_set_left_f32 if left else self._lib.v3dm_set_right_f32
            check(fn(self._ctx, int(n), ptr(desc, ctypes.c_float)), self._ctx)

    def _guided(self, fn, M, thr, lpos, rpos, minScore, minRatio):
        M = np.ascontiguousarray(np.asarray(M, np.float64).reshape(9))
        lpos = np.ascontiguousarray(np.asarray(lpos, np.float32).reshape(-1, 2))
        rpos = np.ascontiguousarray(np.asarray(rpos, np.float32).reshape(-1, 2))
        pairs, count = self._buffer()
        check(fn(self._ctx, ptr(M, ctypes.c_double), float(thr), ptr(lpos, ctypes.c_float),
                 ptr(rpos, ctypes.c_float), minScore, minRatio, ptr(pairs, ctypes.c_int32), self._cap,
                 ctypes.byref(count)), self._ctx)
        return _as_list(pairs, count.value)


def _as_list(pairs: np.ndarray, n: int) -> list[tuple[int, int]]:
    return list(map(tuple, pairs[:n].tolist()))


class EmbeddingNet:
    """The learned proposal network (v3dm_mlp_*): 128 -> 256 (GELU) -> 64,
    L2-normalised, BF16 weights; trained and run by the library's HIP kernels."""

    PARAMS = None

    def __init__(self, device: int = 0, seed: int = 1):
        self._lib = _lib.lib()
        h = ctypes.c_void_p()
        rc = self._lib.v3dm_mlp_create(device, seed, ctypes.byref(h))
        if rc != _lib.V3DM_OK:
            raise _lib.V3DMError(rc, self._lib.v3dm_mlp_last_error(None).decode(errors="replace"))
        self._h = h
        self.n_params = self._lib.v3dm_mlp_param_count()

    def _check(self, rc):
        if rc != _lib.V3DM_OK:
            raise _lib.V3DMError(rc, self._lib.v3dm_mlp_last_error(self._h).decode(errors="replace"))

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        self._check(self._lib.v3dm_mlp_get_params(self._h, ptr(out, ctypes.c_float)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._check(self._lib.v3dm_mlp_set_params(self._h, ptr(p, ctypes.c_float)))

    def save(self, path: str) -> None:
        """BF16 weights + FP32 biases with a checked header (v3dm_mlp_save)."""
        self._check(self._lib.v3dm_mlp_save(self._h, str(path).encode()))

    def load(self, path: str) -> None:
        """Parameters from a file written by save(); embeddings are then bit
        for bit those of the saved network."""
        self._check(self._lib.v3dm_mlp_load(self._h, str(path).encode()))

    def loss_grad(self, left: np.ndarray, right: np.ndarray, temperature: float) -> tuple[float, np.ndarray]:
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        assert left.shape == right.shape and left.shape[1] == 128
        loss = np.zeros(1, np.float32)
        grad = np.empty(self.n_params, np.float32)
        self._check(self._lib.v3dm_mlp_loss_grad(self._h, left.shape[0], ptr(left, ctypes.c_uint8),
                                                 ptr(right, ctypes.c_uint8), temperature, ptr(loss, ctypes.c_float),
                                                 ptr(grad, ctypes.c_float)))
        return float(loss[0]), grad

    def set_training_pairs(self, left: np.ndarray, right: np.ndarray) -> None:
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        assert left.shape == right.shape and left.shape[1] == 128
        self._check(self._lib.v3dm_mlp_set_training_pairs(self._h, left.shape[0], ptr(left, ctypes.c_uint8),
                                                          ptr(right, ctypes.c_uint8)))

    def train(self, steps: int, batch: int = 1024, lr: float = 1e-3, temperature: float = 0.05,
              seed: int = 7) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(self._lib.v3dm_mlp_train(self._h, steps, batch, lr, temperature, seed,
                                             ptr(losses, ctypes.c_float)))
        return losses[:steps]

    def embed(self, desc: np.ndarray) -> np.ndarray:
        desc = np.ascontiguousarray(desc, np.uint8)
        out = np.empty((desc.shape[0], 64), np.float32)
        self._check(self._lib.v3dm_mlp_embed(self._h, desc.shape[0], ptr(desc, ctypes.c_uint8),
                                             ptr(out, ctypes.c_float)))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.v3dm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
