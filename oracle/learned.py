"""CPU restatement of the learned proposal path's decisions, given the
candidate table the retrieval produced.  Test infrastructure only.

Scored pairs: every (k, cand[k][i]).  Over those, each left and right gets
its exact top-2 (score, then smaller index, as the reference's first-index
rule, v3d_cudamatching.cu:17-33).  A left is flagged when
best - second < margin; a right when it has one scored pair or
best - second < margin.  A flagged left adds every (k, j) whose exact score
exceeds the survivor threshold of the rule (the pairs an exhaustive int8 scan
would keep); a flagged right every (k, j) likewise.  The reference rule
(ratio test, v3d_cudamatching.cu:117-194; mutual check,
v3d_cudamatching.cpp:121-126) then runs on the exact top-2 over all scored
pairs, in FP32 as the GPU finalize kernel evaluates it.
"""
import math

import numpy as np

from . import oracle as O

SCALE = np.float32(1.0 / 262144.0)


def survivor_threshold(min_score: float, min_ratio: float) -> int:
    """Integer survivor threshold (scores in units of 2^-18); -1 keeps all, 2^28 none."""
    ms = float(np.float32(min_score))
    mr = float(np.float32(min_ratio))
    if mr < 0:
        t = ms
    elif mr > 0:
        t = min(ms, 1.0 - (1.0 - ms) / (mr * (1.0 - 1e-5)))
    elif mr == 0:
        t = max(ms, 1.0)  # accepted only with best > 1 (the ratio is then negative)
    else:
        return 1 << 28  # NaN: nothing is accepted
    if t != t:
        return 1 << 28
    x = math.floor(t * 262144.0)
    return int(min(max(x, -1), 1 << 28))


def rule_accept(best, nxt, min_score, min_ratio) -> bool:
    best, nxt = np.float32(best), np.float32(nxt)
    ms, mr = np.float32(min_score), np.float32(min_ratio)
    if mr < 0:
        return bool(best > ms)
    nd = np.maximum(np.float32(0.001), np.float32(2.0) - np.float32(2.0) * nxt)
    bd = np.float32(2.0) - np.float32(2.0) * best
    return bool(best > ms and bd / nd < mr)


def _top2(vals: np.ndarray, mask: np.ndarray):
    """(best index, best value, second value) over masked entries; None when empty."""
    idx = np.flatnonzero(mask)
    if idx.size == 0:
        return None
    v = vals[idx]
    o = np.lexsort((idx, -v))  # score descending, then index ascending
    b = idx[o[0]]
    second = vals[idx[o[1]]] if idx.size > 1 else None
    return int(b), int(vals[b]), (None if second is None else int(second))


def _f(v):
    return np.float32(-1e30) if v is None else np.float32(v) * SCALE


def learned_match(L, R, cand, margin, min_score, min_ratio):
    """Returns (match table per left, flagged lefts, flagged rights)."""
    S = O.scores_u8(L, R).T.astype(np.int64)  # [left, right]
    nL, nR = S.shape
    scored = np.zeros((nL, nR), bool)
    for k in range(nL):
        c = cand[k]
        scored[k, c[c >= 0]] = True
    m32 = np.float32(margin)
    flagL, flagR = [], []
    for k in range(nL):
        t = _top2(S[k], scored[k])
        if t is not None and t[2] is not None and _f(t[1]) - _f(t[2]) < m32:
            flagL.append(k)
    for j in range(nR):
        t = _top2(S[:, j], scored[:, j])
        if t is None:
            continue
        if t[2] is None or _f(t[1]) - _f(t[2]) < m32:
            flagR.append(j)
    thr = survivor_threshold(min_score, min_ratio)
    if flagL:
        scored[flagL, :] |= S[flagL, :] > thr
    if flagR:
        scored[:, flagR] |= S[:, flagR] > thr
    table = np.full(nL, -1, np.int32)
    colTop = [_top2(S[:, j], scored[:, j]) for j in range(nR)]
    for k in range(nL):
        t = _top2(S[k], scored[k])
        if t is None or not rule_accept(_f(t[1]), _f(t[2]), min_score, min_ratio):
            continue
        j = t[0]
        u = colTop[j]
        if u is None or u[0] != k:
            continue
        if rule_accept(_f(u[1]), _f(u[2]), min_score, min_ratio):
            table[k] = j
    return table, np.array(flagL, np.int64), np.array(flagR, np.int64)
