"""CPU checks of the arithmetic the BF16 candidate path rests on.

The GPU path (V3DM_KERNEL_BF16) rounds float descriptors to BF16 (round to
nearest even), scores every pair on BF16 MFMA with FP32 accumulation, keeps
the pairs whose BF16 score exceeds lim = thr - (2^-7 + 2^-14)|x|max |y|max,
and rescores those with the plain kernel's FP32 arithmetic.  These tests pin
(1) the rounding against torch's float -> bfloat16 conversion, (2) the
per-pair error bound |S_bf16 - x.y| <= 2^-7 |x||y| (the 2^-14 slack covers
FP32 accumulation) on generator data, including that it is not vacuous, and
(3) that every pair whose FP32 score passes the survivor threshold is a
candidate.  They also state how often the spec's smaller rescoring set (the
BF16 top-4 of each query) would have contained the exact best and second.
"""
import numpy as np
import pytest
import torch

import v3dsfmtoolkit_amd as v
from oracle import oracle as O


def bf16_rne(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 -> FP32, round to nearest even (the prep kernel's rule, finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def survivor_threshold_f(min_score: float, min_ratio: float) -> float:
    if min_ratio < 0:
        t = float(min_score)
    elif min_ratio > 0:
        t = min(float(np.float32(min_score)),
                1.0 - (1.0 - float(np.float32(min_score))) / (float(np.float32(min_ratio)) * (1.0 - 1e-5)))
    else:
        return -np.inf
    f = np.float32(t)
    if float(f) > t:
        f = np.nextafter(f, np.float32(-np.inf))
    return float(f)


def floats(seed, nL, nR):
    d = v.synth_pair(seed, nL, nR, 0.4, 0.05, u8=False, f32=True)
    return d["Lf"], d["Rf"]


def test_rounding_matches_torch_bfloat16():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        rng.random(20000).astype(np.float32) * 0.2,
                        # exact ties: 1 + (2k+1) 2^-8 sits halfway between two BF16 values
                        (1.0 + (2 * np.arange(64) + 1) * 2.0 ** -8).astype(np.float32),
                        np.array([0.0, -0.0, 1e-30, -3.3e-39, 65504.7], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(bf16_rne(x), want)


@pytest.mark.parametrize("seed", [11, 12])
def test_bf16_error_bound_holds_and_is_tight(seed):
    L, R = floats(v.synth_seed(2, seed), 512, 640)
    S64 = L.astype(np.float64) @ R.astype(np.float64).T
    Sbf = bf16_rne(L).astype(np.float64) @ bf16_rne(R).astype(np.float64).T  # exact products, ~exact sum
    norms = np.linalg.norm(L.astype(np.float64), axis=1)[:, None] * np.linalg.norm(R.astype(np.float64), axis=1)[None]
    rel = np.abs(Sbf - S64) / norms
    assert rel.max() <= 2.0 ** -7
    # the bound is within a small factor of what the data reaches
    assert rel.max() > 2.0 ** -11


@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)])
def test_every_deciding_pair_is_a_candidate(rule):
    L, R = floats(v.synth_seed(2, 21), 768, 1024)
    thr = survivor_threshold_f(*rule)
    n2L = float(np.max(np.sum(L.astype(np.float32) ** 2, axis=1)))
    n2R = float(np.max(np.sum(R.astype(np.float32) ** 2, axis=1)))
    lim = thr - (2.0 ** -7 + 2.0 ** -14) * np.sqrt(n2L * n2R) * (1 + 1e-4)
    S32 = O.scores_f32(L, R)  # the plain kernel's FP32 scores, [right, left]
    Sbf = (bf16_rne(R) @ bf16_rne(L).T).astype(np.float32)  # FP32 accumulation
    deciding = S32 > thr
    assert deciding.any()
    assert np.all(Sbf[deciding] > lim)
    # candidates are few: the scan's survivor regions rarely overflow
    assert (Sbf > lim).sum() < 0.01 * Sbf.size


def test_bf16_top4_would_contain_exact_top2():
    """The spec's rescoring set (BF16 top-4 per query) versus the exact top-2:
    stated, not required -- the GPU path rescores the bounded set instead."""
    L, R = floats(v.synth_seed(2, 31), 1024, 1024)
    S32 = O.scores_f32(L, R).T.astype(np.float64)  # [left, right]
    Sbf = (bf16_rne(L) @ bf16_rne(R).T).astype(np.float32)
    top2 = np.argsort(-S32, axis=1, kind="stable")[:, :2]
    top4bf = np.argsort(-Sbf, axis=1, kind="stable")[:, :4]
    hit = np.array([set(top2[i]).issubset(set(top4bf[i])) for i in range(len(L))])
    print(f"exact top-2 inside the BF16 top-4 for {hit.mean() * 100:.2f}% of {len(L)} queries")
    assert hit.mean() > 0.95
