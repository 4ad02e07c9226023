"""GPU parity of the batched matcher on float images (v3dm_batch_set_images_f32:
BF16 candidate scan + FP32 rescoring): every pair's table must equal the
oracle's FP32 rule on that pair (v3d_cudamatching.cpp:97-127), for every tile
variant, in the persistent-queue mode, and on graph replays."""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def float_images(seed, counts):
    """Unit float images; a third of each image's rows are noisy copies of the previous image's."""
    rng = np.random.default_rng(seed)
    imgs = []
    for i, n in enumerate(counts):
        x = v.synth_pair(v.synth_seed(2, 2000 + seed * 64 + i), n, 0, u8=False, f32=True)["Lf"]
        if i > 0 and n > 0 and len(imgs[-1]) > 0:
            k = min(n, len(imgs[-1])) // 3
            idx = rng.choice(n, k, replace=False)
            src = rng.choice(len(imgs[-1]), k, replace=False)
            y = imgs[-1][src] + rng.normal(0, 0.02, (k, 128)).astype(np.float32)
            x[idx] = y / np.linalg.norm(y, axis=1, keepdims=True)
        imgs.append(np.ascontiguousarray(x, np.float32))
    return imgs


def expected(imgs, c32, pairs, ms, mr):
    return [O.find_f32(imgs[a][:c32[a]], imgs[b][:c32[b]], ms, mr)[2] for a, b in pairs]


@pytest.fixture(scope="module")
def m(gpu_available):
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    yield mm
    mm.close()


@pytest.fixture(scope="module")
def batch(m):
    counts = [517, 1200, 3000, 64, 2400, 999, 31, 1800]
    imgs = float_images(5, counts)
    c32 = m.setImages(imgs)
    rng = np.random.default_rng(4)
    pairs = [(i, i + 1) for i in range(len(counts) - 1)]
    pairs += [(int(a), int(b)) for a, b in rng.integers(0, len(counts), (8, 2))]
    pairs += [(2, 2), (6, 4)]
    m.setPairs(pairs)
    return imgs, c32, pairs


@pytest.mark.parametrize("variant", list(range(v.scan_variants_f32())))
@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)])
def test_float_batch_variants_match_oracle(m, batch, variant, rule):
    imgs, c32, pairs = batch
    say("float batch variant", variant, rule)
    m.setImages(imgs)
    m.setPairs(pairs)
    m.batchSetGraph(False)
    m.batchSetVariant(variant)
    got = m.batchMatch(*rule)
    want = expected(imgs, c32, pairs, *rule)
    for p, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(g, w, err_msg=f"pair {p} {pairs[p]}")


def test_float_batch_persistent_queue_and_graph(m, gpu_available):
    say("float persistent queue + graph")
    imgs = float_images(9, [4096] * 5)
    c32 = m.setImages(imgs)
    pairs = [(0, 1), (1, 2), (2, 3), (3, 4), (0, 4), (4, 0), (1, 3), (2, 0)]
    m.setPairs(pairs)
    m.batchSetVariant(3)  # 256 x 256 tiles: 2048 tiles, more than the resident workgroups
    want = expected(imgs, c32, pairs, 0.8, 0.8)
    for graph in (False, True, True, False):
        m.batchSetGraph(graph)
        got = m.batchMatch(0.8, 0.8)
        for p in range(len(pairs)):
            np.testing.assert_array_equal(got[p], want[p], err_msg=f"graph {graph} pair {p}")
    m.batchSetGraph(False)


def test_switching_pool_type(m, gpu_available):
    """A uint8 pool after a float pool (and back) on the same context."""
    say("pool type switch")
    d = v.synth_pair(v.synth_seed(2, 77), 1024, 1024, 0.4, 0.05, f32=True)
    for kind in ("f32", "u8", "f32"):
        imgs = [d["Lf"], d["Rf"]] if kind == "f32" else [d["L"], d["R"]]
        m.setImages(imgs)
        m.setPairs([(0, 1), (1, 0)])
        m.batchSetVariant(0)
        got = m.batchMatch(0.8, 0.8)
        fn = O.find_f32 if kind == "f32" else O.find_u8
        np.testing.assert_array_equal(got[0], fn(imgs[0], imgs[1], 0.8, 0.8)[2])
        np.testing.assert_array_equal(got[1], fn(imgs[1], imgs[0], 0.8, 0.8)[2])
