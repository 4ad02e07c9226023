"""CPU restatement of the drivers' geometric stage around robust orientation
(SfM/match_sequence/src/match_sequence.cpp:229-346,
SfM/match_using_voctree/src/match_using_voctree.cpp:325-491).  Test
infrastructure only.

Each piece is stated from what the mathematics fixes, not from the
reference's operation order:

* coverage / effective inlier count (match_sequence.cpp:34-50,
  SfM/reconstruction/include/reconstruction_common.h:125-157): the reference
  computes each raster pixel's squared distance to the nearest point with an
  exact distance transform (reconstruction_common2.cpp:17-162); here a
  brute-force nearest-point search gives the same integers.  The float32
  steps around it (raster size, radius, clamping, the final ratio and
  rounding) follow the reference's types.
* known-pose inliers: the squared Sampson distance in float64 against
  threshold^2 (the rule of the drivers' --poses orientation).
* visibility (match_using_voctree.cpp:397-418): the depth of the
  least-squares triangulation, by numpy's lstsq.
* normalised positions (match_sequence.cpp:327-338): K^-1 (x, y, 1) by numpy's
  solve, rounded to float32.
* the stage itself: putative list, orientation, the inlier checks, guided
  matching (the oracle rule with an epipolar mask), the visibility test, the
  effective-inlier check, normalisation.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

PI_4 = np.float32(0.785398163397448)


def coverage(px, py, width: int, height: int) -> np.float32:
    f32 = np.float32
    dsf = f32(width) / f32(256)
    W, H = 256, int(f32(height) / dsf)
    occupied = set()
    for x, y in zip(px, py):
        X = max(f32(0), min(f32(W - 1), f32(x) / dsf))
        Y = max(f32(0), min(f32(H - 1), f32(y) / dsf))
        occupied.add((int(X), int(Y)))
    radius2 = f32(f32(W) * f32(H)) / f32(len(px) * 4) if len(px) else f32(np.inf)
    xs, ys = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    if occupied:
        d = np.full((W, H), np.iinfo(np.int64).max)
        for ox, oy in sorted(occupied):
            d = np.minimum(d, (xs - ox) ** 2 + (ys - oy) ** 2)
        dist = d.astype(np.float32)
    else:
        dist = np.full((W, H), f32(1e20))
    s = int((dist < radius2).sum())
    return f32(s) / (f32(W * H) * PI_4)


def effective_inlier_count(p, q, width: int, height: int) -> int:
    """p, q: (n, 2) left / right positions."""
    p = np.asarray(p, np.float32).reshape(-1, 2)
    q = np.asarray(q, np.float32).reshape(-1, 2)
    c1 = coverage(p[:, 0], p[:, 1], width, height)
    c2 = coverage(q[:, 0], q[:, 1], width, height)
    cov = np.float32(0.5) * c1 + np.float32(0.5) * c2
    return int(cov * np.float32(len(p)) + np.float32(0.5))


def sampson_sq(F, p, q) -> np.ndarray:
    F = np.asarray(F, np.float64).reshape(3, 3)
    ph = np.c_[np.asarray(p, np.float64).reshape(-1, 2), np.ones(len(p))]
    qh = np.c_[np.asarray(q, np.float64).reshape(-1, 2), np.ones(len(q))]
    Fp, Ftq = ph @ F.T, qh @ F
    num = np.einsum("ij,ij->i", qh, Fp) ** 2
    den = Fp[:, 0] ** 2 + Fp[:, 1] ** 2 + Ftq[:, 0] ** 2 + Ftq[:, 1] ** 2
    return num / den


def sampson_inliers(F, threshold: float, p, q) -> np.ndarray:
    return np.nonzero(sampson_sq(F, p, q) <= float(threshold) ** 2)[0]


def depth_in_view1(K1, K2, R, T, p, q) -> np.ndarray:
    P0 = np.asarray(K1, np.float64) @ np.c_[np.eye(3), np.zeros(3)]
    P1 = np.asarray(K2, np.float64) @ np.c_[np.asarray(R, np.float64).reshape(3, 3), np.asarray(T, np.float64)]
    out = np.zeros(len(p))
    for k in range(len(p)):
        rows, rhs = [], []
        for P, (x, y) in ((P0, p[k]), (P1, q[k])):
            x, y = float(x), float(y)
            rows += [x * P[2, :3] - P[0, :3], y * P[2, :3] - P[1, :3]]
            rhs += [P[0, 3] - x * P[2, 3], P[1, 3] - y * P[2, 3]]
        out[k] = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0][2]
    return out


def normalize(K, p) -> np.ndarray:
    p = np.asarray(p, np.float64).reshape(-1, 2)
    w = np.linalg.solve(np.asarray(K, np.float64), np.c_[p, np.ones(len(p))].T).T
    return w[:, :2].astype(np.float32)


def geometric_stage(L, R, lpos, rpos, K1, K2, dims1, dims2, pose, F, *, u8: bool, ratio: float, min_raw: int,
                    eg_threshold: float = 2.0, min_inliers: int = -1, min_inlier_ratio: float = 0.5,
                    effective: bool = True, guided: bool = True, visibility: bool = False):
    """The drivers' steps for one pair with a known pose (R, T) whose inliers are the
    Sampson rule's; F is the pair's fundamental matrix.  Returns None (not stored) or
    (left indices, right indices, normalised left (n, 2), normalised right (n, 2)).
    L, R: the kept descriptors (uint8 u or float32); lpos, rpos their pixel positions."""
    find = O.find_u8 if u8 else O.find_f32
    tab = find(L, R, 0.9, ratio)[2]
    il = np.nonzero(tab >= 0)[0]
    ir = tab[il]
    if len(il) < min_raw:
        return None
    if min_inliers < 0:
        min_inliers = min_raw
    max_width = max(max(dims1), max(dims2))
    thr = float(np.float32(max_width) * np.float32(eg_threshold / 1024))
    inl = sampson_inliers(F, thr, lpos[il], rpos[ir])
    if len(inl) < min_inliers or len(inl) < min_inlier_ratio * len(il):
        return None
    il, ir = il[inl], ir[inl]
    w, h = dims1
    if guided:
        mask = (1, np.asarray(F, np.float64), float(np.float32(thr)), lpos, rpos)
        tab = find(L, R, 0.9, ratio, masks=[mask])[2]
        il = np.nonzero(tab >= 0)[0]
        ir = tab[il]
        if visibility:
            keep = depth_in_view1(K1, K2, pose[0], pose[1], lpos[il], rpos[ir]) > 0
            il, ir = il[keep], ir[keep]
    n_eff = effective_inlier_count(lpos[il], rpos[ir], w, h) if effective else len(il)
    if n_eff < min_inliers:
        return None
    return il, ir, normalize(K1, lpos[il]), normalize(K2, rpos[ir])
