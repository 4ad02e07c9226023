"""GPU parity of the float-descriptor path (V3DM_KERNEL_BF16): BF16 MFMA
candidates + FP32 rescoring must give exactly the plain FP32 kernel's and the
oracle's FP32 match lists (v3d_cudamatching.cpp:97-127 on unquantised
floats), and its rescored scores are compared with an FP64 oracle."""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def floats(seed, nL, nR, rho=0.4, sigma=0.05):
    d = v.synth_pair(seed, nL, nR, rho, sigma, u8=False, f32=True)
    return d["Lf"], d["Rf"]


@pytest.fixture(scope="module")
def m(gpu_available):
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    mm.allocate(2048, 2048, True)
    yield mm
    mm.close()


def run(m, L, R, rule, kernel):
    m.setLeftFeatures(len(L), L)
    m.setRightFeatures(len(R), R)
    m.setKernel(kernel)
    out = m.findPutativeMatches(*rule)
    assert m.lastKernel() == kernel
    return out


SHAPES = [(256, 320), (992, 1472), (2048, 2048), (96, 2048)]
RULES = [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0), (0.7, 0.95)]


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("shape", SHAPES)
def test_bf16_equals_plain_and_oracle(m, shape, rule):
    nL, nR = shape
    say("bf16 vs plain", shape, rule)
    L, R = floats(v.synth_seed(3, nL + nR), nL, nR)
    got = run(m, L, R, rule, v.KERNEL_BF16)
    plain = run(m, L, R, rule, v.KERNEL_PLAIN)
    want = O.table_to_pairs(O.find_f32(L, R, *rule)[2])
    assert got == plain
    assert got == want


def test_auto_picks_bf16_for_float_descriptors(m):
    say("auto float")
    L, R = floats(v.synth_seed(3, 7), 512, 512)
    m.setLeftFeatures(512, L)
    m.setRightFeatures(512, R)
    m.setKernel(v.KERNEL_AUTO)
    got = m.findPutativeMatches(0.8, 0.8)
    assert m.lastKernel() == v.KERNEL_BF16
    assert got == O.table_to_pairs(O.find_f32(L, R, 0.8, 0.8)[2])


def test_no_filter_and_near_duplicates(m):
    """minRatio 0 keeps every pair as a candidate, and near-identical sets
    put hundreds of candidates in one region: both take the overflow path
    that rescores whole blocks."""
    say("bf16 overflow paths")
    L, R = floats(v.synth_seed(3, 9), 256, 320)
    got = run(m, L, R, (0.8, 0.0), v.KERNEL_BF16)
    assert got == O.table_to_pairs(O.find_f32(L, R, 0.8, 0.0)[2])
    rng = np.random.default_rng(4)
    Ld = np.repeat(L[:8], 32, axis=0) + rng.normal(0, 1e-4, (256, 128)).astype(np.float32)
    Rd = np.repeat(L[:8], 40, axis=0) + rng.normal(0, 1e-4, (320, 128)).astype(np.float32)
    for rule in [(0.8, 0.8), (0.5, -1.0)]:
        got = run(m, Ld, Rd, rule, v.KERNEL_BF16)
        assert got == O.table_to_pairs(O.find_f32(Ld, Rd, *rule)[2])


def test_switching_between_int8_and_bf16_jobs(m):
    """The top-2 arrays change key format between the int8 and BF16 paths."""
    say("kernel switching")
    d = v.synth_pair(v.synth_seed(3, 11), 1024, 1024, 0.4, 0.05, f32=True)
    want8 = O.table_to_pairs(O.find_u8(d["L"], d["R"], 0.8, 0.8)[2])
    wantf = O.table_to_pairs(O.find_f32(d["Lf"], d["Rf"], 0.8, 0.8)[2])
    for _ in range(3):
        assert run(m, d["L"], d["R"], (0.8, 0.8), v.KERNEL_I8) == want8
        assert run(m, d["Lf"], d["Rf"], (0.8, 0.8), v.KERNEL_BF16) == wantf


def test_guided_epipolar_with_bf16(m):
    say("bf16 guided")
    nL, nR = 1024, 992
    L, R = floats(v.synth_seed(3, 13), nL, nR)
    rng = np.random.default_rng(8)
    lp = (rng.random((nL, 2)) * 1000).astype(np.float32)
    rp = (rng.random((nR, 2)) * 1000).astype(np.float32)
    F = np.array([[0, -1e-3, 0.2], [1e-3, 0, -0.3], [-0.2, 0.3, 1.0]], np.float64)
    m.setLeftFeatures(nL, L)
    m.setRightFeatures(nR, R)
    for kernel in (v.KERNEL_BF16, v.KERNEL_PLAIN):
        m.setKernel(kernel)
        m.findPutativeMatches(0.8, 0.8)
        g = m.runGuidedMatching(F, 40.0, lp, rp, 0.8, 0.8)
        if kernel == v.KERNEL_BF16:
            got = g
    want = O.table_to_pairs(O.find_f32(L, R, 0.8, 0.8, masks=[(1, F, 40.0, lp, rp)])[2])
    unguided = O.table_to_pairs(O.find_f32(L, R, 0.8, 0.8)[2])
    say(f"guided {len(want)} of {len(unguided)} unguided pairs")
    assert got == g == want


def test_bf16_scores_and_lists_vs_fp64(m):
    """Rescored FP32 scores against the FP64 oracle, and the FP32 rule's list
    against the FP64 rule's list (the reference's own FP32 arithmetic is the
    only source of disagreement)."""
    say("bf16 vs fp64")
    nL = nR = 2048
    L, R = floats(v.synth_seed(3, 15), nL, nR)
    got = run(m, L, R, (0.8, 0.8), v.KERNEL_BF16)
    best, second, arg = m.leftTop2()
    have = arg >= 0
    assert have.sum() > 0.3 * nL
    S64 = np.einsum("ij,ij->i", L[have].astype(np.float64), R[arg[have]].astype(np.float64))
    err = np.abs(best[have].astype(np.float64) - S64)
    say(f"rescored best-score |FP32 - FP64|: max {err.max():.3e}, mean {err.mean():.3e} over {have.sum()} queries")
    assert err.max() < 1e-5
    _, _, tab64, best64, _ = O.find_f64(L, R, 0.8, 0.8)
    l64 = set(O.table_to_pairs(tab64))
    diff = set(got) ^ l64
    say(f"match lists: FP32 rule {len(got)} pairs, FP64 rule {len(l64)} pairs, {len(diff)} differ")
    assert len(diff) <= max(2, len(l64) // 1000)
