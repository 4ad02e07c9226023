"""GPU parity of the batched matcher (v3dm_batch_*): every pair's table must
equal the oracle's rule on that pair, for every scan tile variant, in the
persistent-queue mode, and on repeated hipGraph replays."""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def make_images(seed, counts):
    """Images whose consecutive pairs share planted copies (pair i = (i, i+1))."""
    rng = np.random.default_rng(seed)
    imgs = []
    for i, n in enumerate(counts):
        fresh = v.synth_pair(v.synth_seed(2, 1000 + seed * 64 + i), n, 0)["L"]
        if i > 0 and n > 0 and len(imgs[-1]) > 0:
            # a third of the rows: noisy copies of rows of the previous image
            k = min(n, len(imgs[-1])) // 3
            idx = rng.choice(n, k, replace=False)
            src = rng.choice(len(imgs[-1]), k, replace=False)
            fresh[idx] = np.clip(imgs[-1][src].astype(np.int32) + rng.integers(-6, 7, (k, 128)), 0,
                                 255).astype(np.uint8)
        imgs.append(fresh)
    return imgs


def expected_tables(imgs, counts32, pairs, ms, mr):
    out = []
    for a, b in pairs:
        L = imgs[a][:counts32[a]]
        R = imgs[b][:counts32[b]]
        out.append(O.find_u8(L, R, ms, mr)[2])
    return out


@pytest.fixture(scope="module")
def m(gpu_available):
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    yield mm
    mm.close()


@pytest.fixture(scope="module")
def batch(m):
    counts = [517, 1200, 3000, 64, 2400, 999, 31, 1800, 2600, 700]
    imgs = make_images(7, counts)
    c32 = m.setImages(imgs)
    rng = np.random.default_rng(3)
    pairs = [(i, i + 1) for i in range(len(counts) - 1)]
    pairs += [(int(a), int(b)) for a, b in rng.integers(0, len(counts), (14, 2))]
    pairs += [(2, 2), (6, 4), (4, 6)]  # self pair, empty image both sides
    m.setPairs(pairs)
    return imgs, c32, pairs


@pytest.mark.parametrize("variant", list(range(len(v.scan_variants()))))
@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)])
def test_batch_variants_match_oracle(m, batch, variant, rule):
    imgs, c32, pairs = batch
    say("batch variant", variant, rule)
    m.batchSetGraph(False)
    m.batchSetVariant(variant)
    got = m.batchMatch(*rule)
    assert m.batchVariant() == variant
    want = expected_tables(imgs, c32, pairs, *rule)
    assert len(got) == len(want)
    for p, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(g, w, err_msg=f"pair {p} {pairs[p]}")


def test_batch_persistent_queue(m, gpu_available):
    """More tiles than resident workgroups: the persistent tile queue runs."""
    say("persistent queue")
    counts = [4096, 4096, 4096, 4096, 4096]
    imgs = make_images(11, counts)
    c32 = m.setImages(imgs)
    pairs = [(0, 1), (1, 2), (2, 3), (3, 4), (0, 4), (4, 0), (1, 3), (2, 0)]
    m.setPairs(pairs)
    m.batchSetGraph(False)
    m.batchSetVariant(3)  # 256 x 256 tiles: 8 pairs x 16 x 16 = 2048 tiles
    info = m.batchInfo()
    assert info["tiles"] == 8 * 16 * 16
    want = expected_tables(imgs, c32, pairs, 0.8, 0.8)
    for rep in range(3):
        got = m.batchMatch(0.8, 0.8)
        for p in range(len(pairs)):
            np.testing.assert_array_equal(got[p], want[p], err_msg=f"rep {rep} pair {p}")


def test_batch_pool_growth(gpu_available):
    """Image pools that grow and shrink between jobs: every pool buffer
    (descriptors and their row sums) must follow the new size."""
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    try:
        for seed, counts in [(21, [300, 640]), (22, [2048, 4096, 3072]), (23, [96, 512])]:
            say("pool growth", counts)
            imgs = make_images(seed, counts)
            c32 = mm.setImages(imgs)
            pairs = [(i, i + 1) for i in range(len(counts) - 1)] + [(len(counts) - 1, 0)]
            mm.setPairs(pairs)
            for variant in (0, 3):
                mm.batchSetVariant(variant)
                got = mm.batchMatch(0.8, 0.8)
                want = expected_tables(imgs, c32, pairs, 0.8, 0.8)
                for p in range(len(pairs)):
                    np.testing.assert_array_equal(got[p], want[p], err_msg=f"{counts} variant {variant} pair {p}")
    finally:
        mm.close()


def test_batch_graph_replays(m, batch):
    """Graph replays interleaved with eager jobs and other thresholds."""
    imgs, c32, pairs = batch
    m.setImages(imgs)
    m.setPairs(pairs)
    say("graph replays")
    want = expected_tables(imgs, c32, pairs, 0.8, 0.8)
    want9 = expected_tables(imgs, c32, pairs, 0.9, 0.8)
    m.batchSetVariant(0)
    for rep in range(4):
        m.batchSetGraph(True)
        got = m.batchMatch(0.8, 0.8)
        for p in range(len(pairs)):
            np.testing.assert_array_equal(got[p], want[p], err_msg=f"graph rep {rep} pair {p}")
        m.batchSetGraph(False)
        got9 = m.batchMatch(0.9, 0.8)
        for p in range(len(pairs)):
            np.testing.assert_array_equal(got9[p], want9[p], err_msg=f"eager rep {rep} pair {p}")
    m.batchSetGraph(True)
    for _ in range(5):
        m.batchMatchAsync(0.8, 0.8)
    got = m.batchFetch()
    for p in range(len(pairs)):
        np.testing.assert_array_equal(got[p], want[p])
    m.batchSetGraph(False)


def test_batch_autotune_and_auto_variant(m, batch, tmp_path, monkeypatch):
    imgs, c32, pairs = batch
    say("autotune")
    m.setImages(imgs)
    m.setPairs(pairs)
    best, ms = m.batchAutotune(0.8, 0.8, reps=2)
    assert 0 <= best < len(v.scan_variants())
    assert np.all(ms > 0) and ms[best] == ms.min()
    want = expected_tables(imgs, c32, pairs, 0.8, 0.8)
    m.batchSetVariant(-1)  # automatic: keeps the tuned variant for this pair set
    got = m.batchMatch(0.8, 0.8)
    for p in range(len(pairs)):
        np.testing.assert_array_equal(got[p], want[p])


def test_batch_equals_single_pair_api(m, batch):
    imgs, c32, pairs = batch
    say("batch vs single")
    m.setImages(imgs)
    m.setPairs(pairs[:5])
    m.batchSetVariant(0)
    got = m.batchMatch(0.8, 0.8)
    s = v.SIFTMatching(0)
    s.enableRoundingWarning(False)
    s.allocate(4096, 4096)
    for p, (a, b) in enumerate(pairs[:5]):
        s.setLeftFeatures(len(imgs[a]), imgs[a])
        s.setRightFeatures(len(imgs[b]), imgs[b])
        lst = s.findPutativeMatches(0.8, 0.8)
        assert lst == O.table_to_pairs(got[p])
    s.close()
