"""The benchmark workloads at their own sizes on the GPU, each checked
against the oracle (all through the C ABI).

* 256 image pairs, feature counts uniform in [500, 12000], 40 % noisy copies:
  one batched job (autotuned tile variant, hipGraph replay and eager), every
  pair's table equal to the oracle's blocked rule; the same pairs as float
  descriptors (BF16 candidates + FP32 rescoring), every pair's table equal to
  the FP32 rule, and the FP32 rule's score error and list disagreement against
  the FP64 rule stated.
* 64 pairs of 16384 x 16384 through the learned path, network trained on 2 M
  exact-labelled pairs: recall and precision against the exact int8 lists
  (which equal the oracle's on two of the pairs).
* One 65536 x 65536 pair (30 % copies, sigma 0.08): int8 list equal to the
  oracle's; learned path with top-32 candidates: recall and precision, and
  the top-32 path's list equal to the learned-decision oracle on a pair the
  Python oracle covers.
* Batches above 128 survivor regions per CU with both insertion forms, filter
  on and with every score a survivor (every region overflows).

The per-pair rule these batched jobs replace is the reference's one
findPutativeMatches per pair (SfM/match_sequence/src/match_sequence.cpp:134-227,
V3D/src/CUDA/v3d_cudamatching.cpp:97-127).
"""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import learned as OL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RULE = (0.8, 0.8)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


# ---- 256 uneven pairs --------------------------------------------------------

@pytest.fixture(scope="module")
def c3_images(gpu_available):
    rng = np.random.default_rng(2024)
    counts = rng.integers(500, 12001, size=(256, 2))
    u8, f32 = [], []
    for i in range(256):
        d = v.synth_pair(v.synth_seed(3, i), int(counts[i, 0]), int(counts[i, 1]), 0.4, 0.05, f32=True)
        u8 += [d["L"], d["R"]]
        f32 += [d["Lf"], d["Rf"]]
    return u8, f32


def _run_batch(images, graph):
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    c32 = m.setImages(images)
    m.setPairs([(2 * i, 2 * i + 1) for i in range(len(images) // 2)])
    best, _ = m.batchAutotune(*RULE, reps=1)
    m.batchSetVariant(best)
    m.batchSetGraph(False)
    eager = m.batchMatch(*RULE)
    m.batchSetGraph(graph)
    tabs = m.batchMatch(*RULE)
    tabs2 = m.batchMatch(*RULE)  # a replay
    m.close()
    for p, (a, b, c) in enumerate(zip(eager, tabs, tabs2)):
        np.testing.assert_array_equal(a, b, err_msg=f"pair {p}: graph job differs from the eager job")
        np.testing.assert_array_equal(a, c, err_msg=f"pair {p}: graph replay differs")
    return c32, tabs, best


def test_c3_u8_batch_every_pair_equals_oracle(c3_images):
    u8, _ = c3_images
    c32, tabs, best = _run_batch(u8, graph=True)
    say(f"256-pair u8 batch, variant {best}, {sum(int((t >= 0).sum()) for t in tabs)} matches")
    for p in range(256):
        L, R = u8[2 * p][:c32[2 * p]], u8[2 * p + 1][:c32[2 * p + 1]]
        want = O.find_u8_blocked(L, R, *RULE)[2]
        np.testing.assert_array_equal(tabs[p], want, err_msg=f"pair {p} ({len(L)} x {len(R)})")


@pytest.fixture(scope="module")
def c3_float(c3_images):
    _, f32 = c3_images
    c32, tabs, best = _run_batch(f32, graph=True)
    say(f"256-pair float batch, variant {best}")
    return f32, c32, tabs


def test_c3_float_batch_every_pair_equals_fp32_rule(c3_float):
    f32, c32, tabs = c3_float
    for p in range(256):
        L, R = f32[2 * p][:c32[2 * p]], f32[2 * p + 1][:c32[2 * p + 1]]
        want = O.select_f32(O.scores_f32(L, R), *RULE)[2]
        np.testing.assert_array_equal(tabs[p], want, err_msg=f"pair {p} ({len(L)} x {len(R)})")


def test_c3_float_rule_against_fp64(c3_float):
    """What FP32 arithmetic costs on these data: the largest |FP32 - FP64|
    score difference over all 1e10 scored pairs and the number of list
    entries where the FP32 rule (= the GPU's tables) and the FP64 rule
    differ.  Bounds: sequential fmaf over 128 terms of unit vectors is within
    128 * 2^-24 |x||y| (< 1e-5) of the exact dot; the lists may differ only
    where a decision sits within that of a threshold or a tie."""
    f32, c32, tabs = c3_float
    worst, differ, matches = 0.0, 0, 0
    for p in range(256):
        L, R = f32[2 * p][:c32[2 * p]], f32[2 * p + 1][:c32[2 * p + 1]]
        S32 = O.scores_f32(L, R)
        S64 = O.scores_f64(L, R)
        worst = max(worst, float(np.abs(S32.astype(np.float64) - S64).max()))
        want64 = O.select_f64(S64, *RULE)[2]
        differ += int(np.count_nonzero(tabs[p] != want64))
        matches += int((tabs[p] >= 0).sum())
    say(f"max |FP32 - FP64| score {worst:.3g}; {differ} of {matches} list entries differ from the FP64 rule")
    assert worst < 128 * 2.0 ** -24
    assert differ <= max(1, matches // 1000)


# ---- batches above 128 regions per CU: both insertion forms -----------------

@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.0, -1.0)])
def test_batch_insert_forms_above_region_switch(variant, rule):
    """260 pairs of 2048 x 2048 with tile variant 2 (8 waves, 512 x 512:
    33280 regions) or 3 (256 x 256: 66560 regions), run with each insertion
    form forced and with the automatic choice.  Rule (0, ratio off) keeps
    every positive score as a survivor, so every region overflows and is
    recomputed from the descriptors."""
    n_pairs = 260
    imgs = []
    for i in range(n_pairs):
        d = v.synth_pair(v.synth_seed(9, i), 2048, 2048, 0.4, 0.05)
        imgs += [d["L"], d["R"]]
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    m.setImages(imgs)
    m.setPairs([(2 * i, 2 * i + 1) for i in range(n_pairs)])
    m.batchSetVariant(variant)
    assert m.batchInfo()["tiles"] * v.scan_variants()[variant][0] > 32768
    got = {}
    for form in (0, 1, -1):
        m.setInsertForm(form)
        got[form] = m.batchMatch(*rule)
        ran = m.lastInsertForm()
        assert ran == (1 if form == -1 else form), (form, ran)
    m.close()
    say(f"variant {variant} rule {rule}: forms agree on {n_pairs} pairs")
    for p in range(n_pairs):
        want = O.find_u8_blocked(imgs[2 * p], imgs[2 * p + 1], *rule)[2]
        for form in (0, 1, -1):
            np.testing.assert_array_equal(got[form][p], want, err_msg=f"form {form} pair {p}")


# ---- learned path: 64 x 16384^2 and 65536^2 with top-32 ----------------------

def _train(pairs, rho, sigma, steps):
    from v3dsfmtoolkit_amd import learned as VL
    Lt, Rt = VL.labeled_pairs(pairs, n=16384, rho=rho, sigma=sigma)
    net = v.EmbeddingNet(0, seed=5)
    losses = VL.train(net, Lt, Rt, steps=steps, batch=1024, lr=2e-3)
    say(f"trained on {len(Lt)} exact-labelled pairs: loss {losses[0]:.3f} -> {np.mean(losses[-20:]):.3f}")
    return net


@pytest.fixture(scope="module")
def net_c4(gpu_available):
    net = _train(2_000_000, 0.4, 0.05, 2000)
    yield net
    net.close()


def test_c4_learned_64_pairs_16384(net_c4):
    n = 16384
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    m.allocate(n, n)
    hit = n_exact = n_got = flagged = 0
    worst = 1.0
    try:
        for i in range(64):
            d = v.synth_pair(v.synth_seed(4, i), n, n, 0.4, 0.05)
            m.setLeftFeatures(n, d["L"])
            m.setRightFeatures(n, d["R"])
            m.setKernel(v.KERNEL_I8)
            exact = m.findPutativeMatches(*RULE)
            if i < 2:
                assert exact == O.table_to_pairs(O.find_u8_blocked(d["L"], d["R"], *RULE)[2]), f"pair {i}"
            got, st = m.learnedFind(net_c4, 0.05, *RULE, topK=16)
            g, e = set(got), set(exact)
            hit += len(g & e)
            n_exact += len(e)
            n_got += len(g)
            flagged += st["flagged_left"]
            worst = min(worst, len(g & e) / max(len(e), 1))
    finally:
        m.close()
    rec, prec = hit / n_exact, hit / max(n_got, 1)
    say(f"64 x {n}^2: recall {rec:.5f} precision {prec:.5f} worst pair recall {worst:.5f} "
        f"fallback {flagged / (64 * n):.3f} of lefts ({n_exact} exact matches)")
    assert rec >= 0.995 and prec >= 0.995 and worst >= 0.99


@pytest.fixture(scope="module")
def c5_pair(gpu_available):
    n = 65536
    return v.synth_pair(v.synth_seed(5, 0), n, n, 0.3, 0.08)


@pytest.fixture(scope="module")
def net_c5(gpu_available):
    net = _train(1_000_000, 0.3, 0.08, 1500)
    yield net
    net.close()


def test_c5_exact_65536_equals_oracle(c5_pair):
    n = 65536
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    try:
        m.allocate(n, n)
        m.setLeftFeatures(n, c5_pair["L"])
        m.setRightFeatures(n, c5_pair["R"])
        m.setKernel(v.KERNEL_I8)
        got = m.findPutativeMatches(*RULE)
        form = m.lastInsertForm()
    finally:
        m.close()
    want = O.table_to_pairs(O.find_u8_blocked(c5_pair["L"], c5_pair["R"], *RULE)[2])
    say(f"65536^2 int8: {len(got)} matches, insertion form {form}")
    assert got == want


def test_c5_learned_top32_recall(c5_pair, net_c5):
    n = 65536
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    try:
        m.allocate(n, n)
        m.setLeftFeatures(n, c5_pair["L"])
        m.setRightFeatures(n, c5_pair["R"])
        m.setKernel(v.KERNEL_I8)
        exact = set(m.findPutativeMatches(*RULE))
        got, st = m.learnedFind(net_c5, 0.05, *RULE, topK=32)
        cand = m.learnedCandidates()
    finally:
        m.close()
    assert cand.shape == (n, 32)
    g = set(got)
    rec, prec = len(g & exact) / len(exact), len(g & exact) / max(len(g), 1)
    say(f"65536^2 top-32: recall {rec:.5f} precision {prec:.5f} fallback {st['flagged_left'] / n:.3f} of lefts")
    assert rec >= 0.995 and prec >= 0.995


def test_top32_list_equals_learned_oracle(net_c5):
    """The top-32 retrieval and selection kernels' candidates and the list they
    lead to, against the learned-decision oracle on the GPU's own candidate
    table (4096 x 4096, C5 generator settings)."""
    n = 4096
    d = v.synth_pair(v.synth_seed(5, 1), n, n, 0.3, 0.08)
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    try:
        m.allocate(n, n)
        m.setLeftFeatures(n, d["L"])
        m.setRightFeatures(n, d["R"])
        got, st = m.learnedFind(net_c5, 0.05, *RULE, topK=32)
        cand = m.learnedCandidates()
    finally:
        m.close()
    assert cand.shape == (n, 32) and np.all((cand >= 0) & (cand < n))
    srt = np.sort(cand, axis=1)
    assert np.all(srt[:, 1:] != srt[:, :-1])  # 32 distinct rights per query
    tab, fl, fr = OL.learned_match(d["L"], d["R"], cand, 0.05, *RULE)
    assert st["flagged_left"] == len(fl) and st["flagged_right"] == len(fr)
    assert got == O.table_to_pairs(tab)
    # the candidates are the embedding top-32 (ranked on the CPU from the GPU's BF16 embeddings)
    from oracle import mlp as OM
    p = net_c5.params()
    EL = OM.bf16_round(OM.embed(p, d["L"])).astype(np.float64)
    ER = OM.bf16_round(OM.embed(p, d["R"])).astype(np.float64)
    want = OM.topk(EL, ER, 32)
    same = np.mean([len(set(a) & set(b)) / 32 for a, b in zip(cand, want)])
    say(f"top-32 candidate overlap with the CPU ranking {same:.5f}; flagged {len(fl)} / {len(fr)}")
    assert same > 0.99
