"""GPU parity of vocabulary-tree quantisation (v3dm_voctree_*): the words must
equal the CPU oracle's restatement of VocabularyTree::findBestLeafIndex
(V3D/include/Base/v3d_voctree.h:63-99) bit for bit -- on the driver's tree
shape (depth 3, branch factor 50, match_using_voctree.cpp:145), on other
shapes, and on the crafted cases that separate the sequential unfused float
sum from other summations -- in both scoring modes: the FP32-MFMA filter with
exact rescoring of the children within its error bound, and exact VALU sums
for every child."""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import voctree as OV
from v3dsfmtoolkit_amd import voctree as VT

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def kmeans_like_tree(seed, depth, bf):
    return VT.synthetic_tree(seed, depth, bf)


def keys_like_sift(seed, n):
    d = v.synth_pair(v.synth_seed(6, seed), n, 0, u8=False, f32=True)["Lf"]
    return np.ascontiguousarray(d, np.float32)


MODES = [VT.VocabularyTree.AUTO, VT.VocabularyTree.EXACT_VALU]
MODE_IDS = ["mfma_filter", "exact_valu"]


@pytest.fixture(scope="module")
def big_tree(gpu_available):
    say("building the 3 x 50 tree")
    tree = kmeans_like_tree(1, 3, 50)
    t = VT.VocabularyTree(3, 50, tree)
    yield tree, t
    t.close()


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_driver_tree_shape_matches_oracle(big_tree, mode):
    tree, t = big_tree
    t.setMode(mode)
    assert t.visualWordCount() == 125000
    keys = keys_like_sift(1, 20000)
    say("quantise 20000 unit keys")
    got = t.getVisualWords(keys)
    want = OV.quantise(3, 50, tree, keys)
    np.testing.assert_array_equal(got, want)
    assert len(np.unique(want)) > 2000  # the keys spread over the tree
    q = np.clip(np.rint(keys * 512), 0, 255).astype(np.float32) / 512  # u/512 keys
    np.testing.assert_array_equal(t.getVisualWords(q), OV.quantise(3, 50, tree, q))
    scoring, run = t.runTimes()
    assert 0 < scoring <= run


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_repeated_runs_and_smaller_jobs(big_tree, mode):
    tree, t = big_tree
    t.setMode(mode)
    for n in (5000, 1, 65, 0, 777):
        say("job of", n, "keys")
        keys = keys_like_sift(100 + n, max(n, 1))[:n]
        t.upload(keys)
        t.run()
        t.run()
        np.testing.assert_array_equal(t.fetch(), OV.quantise(3, 50, tree, keys))


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("depth,bf", [(1, 7), (2, 300), (5, 10), (4, 13), (2, 33), (2, 64), (3, 32)])
def test_other_tree_shapes(gpu_available, depth, bf, mode):
    say("tree shape", depth, bf, mode)
    tree = kmeans_like_tree(depth * 1000 + bf, depth, bf)
    t = VT.VocabularyTree(depth, bf, tree)
    t.setMode(mode)
    try:
        keys = keys_like_sift(depth * 7 + bf, 3000)
        np.testing.assert_array_equal(t.getVisualWords(keys), OV.quantise(depth, bf, tree, keys))
    finally:
        t.close()


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_crafted_rounding_and_tie_cases(gpu_available, mode):
    """Each key's answer differs between the reference's unfused sequential float
    sum and (a) the exact sum, (b) a fused multiply-add order; plus exact ties
    (first child wins) and a non-finite key (-1)."""
    depth, bf = 2, 4
    tree = np.zeros((VT.tree_nodes(depth, bf), 128), np.float32)
    keys = []
    # root child 0 vs 1: cancellation -- exact score of child 0 is 1, float sum 0; child 1 scores 0.5
    tree[0, :3] = 1.0
    tree[1, 3] = 0.5
    k = np.zeros(128, np.float32)
    k[:4] = [1e8, 1.0, -1e8, 1.0]
    keys.append(k)
    # root child 2 vs 3 for a key orthogonal to children 0/1: fused vs unfused products
    x = np.float32(1 + 2.0 ** -12)
    tree[2, 10], tree[2, 11] = -1.0, x
    tree[3, 12] = np.float32(2.0 ** -11 + 2.0 ** -25)
    k = np.zeros(128, np.float32)
    k[10], k[11], k[12] = 1.0, x, 1.0
    keys.append(k)
    # second level below child 1 (nodes 8..11): two identical children 9 and 10 -> 9 (index 1)
    tree[9, 20] = tree[10, 20] = 0.25
    k = np.zeros(128, np.float32)
    k[3], k[20] = 1.0, 1.0
    keys.append(k)
    k = keys[0].copy()
    k[50] = np.nan
    keys.append(k)
    keys = np.asarray(keys, np.float32)
    want = OV.quantise(depth, bf, tree, keys)
    assert want[0] // bf == 1 and want[1] // bf == 3 and want[2] == 1 * bf + 1 and want[3] == -1
    t = VT.VocabularyTree(depth, bf, tree)
    t.setMode(mode)
    try:
        np.testing.assert_array_equal(t.getVisualWords(keys), want)
        many = np.repeat(keys, 97, axis=0)  # the same cases spread over many waves and groups
        np.testing.assert_array_equal(t.getVisualWords(many), np.repeat(want, 97))
    finally:
        t.close()


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_extreme_magnitudes(gpu_available, mode):
    """Keys and nodes far from unit scale, where the filter's bound leans on
    its absolute terms: keys whose squares underflow (the norm bound's 2^-59
    term), subnormal components and a subnormal child (inputs the matrix core
    may flush), the largest finite node values, and products that overflow
    (non-finite scores: every child is rescored)."""
    depth, bf = 2, 16
    rng = np.random.default_rng(91)
    tree = rng.normal(0, 0.1, (VT.tree_nodes(depth, bf), 128)).astype(np.float32)
    tree[5] *= np.float32(1e-39 / 0.1)        # a subnormal child
    tree[bf + 7] = np.float32(3.4e38)          # near the largest finite float
    base = np.abs(rng.normal(0, 0.1, (64, 128))).astype(np.float32)
    keys = []
    for scale in (1e-30, 1e-22, 1e-19, 1e-3, 1.0, 1e18, 1e19, 3.0e38):
        keys.append((base * np.float32(scale)).astype(np.float32))
    sub = base.copy()
    sub[:, ::3] = np.float32(1e-40)            # subnormal components
    keys.append(sub)
    keys = np.concatenate(keys).astype(np.float32)
    want = OV.quantise(depth, bf, tree, keys)
    t = VT.VocabularyTree(depth, bf, tree)
    t.setMode(mode)
    try:
        say("extreme magnitudes", mode, len(keys))
        np.testing.assert_array_equal(t.getVisualWords(keys), want)
    finally:
        t.close()


def test_tree_file_round_trip(gpu_available, tmp_path):
    depth, bf = 2, 9
    tree = kmeans_like_tree(5, depth, bf)
    p = tmp_path / "tree.bin"
    tree.tofile(p)
    keys = keys_like_sift(9, 500)
    t = VT.VocabularyTree(depth, bf, path=p)
    try:
        np.testing.assert_array_equal(t.getVisualWords(keys), OV.quantise(depth, bf, tree, keys))
    finally:
        t.close()
    p.write_bytes(tree.tobytes()[:-4])
    with pytest.raises(v.V3DMError, match="needs"):
        VT.VocabularyTree(depth, bf, path=p)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_near_ties_in_the_driver_shape(gpu_available, mode):
    """The 3 x 50 tree with, under every node, children 2k and 2k+1 made equal or
    one ulp apart in one component, and keys drawn close to those children:
    most keys face a near-tie that only the exact sequential sums decide
    (the filter must send them to its exact path)."""
    depth, bf = 3, 50
    tree = kmeans_like_tree(77, depth, bf).copy()
    rng = np.random.default_rng(78)
    for start in (0, 50, 2550):
        level = tree[start:start + (bf if start == 0 else (2500 if start == 50 else 125000))]
        level[1::2] = level[0::2]
        flip = rng.random(len(level) // 2) < 0.5
        idx = np.nonzero(flip)[0] * 2 + 1
        level[idx, 7] = np.nextafter(level[idx, 7], np.float32(2))
    keys = np.abs(tree[2550 + rng.integers(0, 125000, 6000)] + rng.normal(0, 0.003, (6000, 128))).astype(np.float32)
    keys /= np.linalg.norm(keys, axis=1, keepdims=True)
    want = OV.quantise(depth, bf, tree, keys)
    t = VT.VocabularyTree(depth, bf, tree)
    t.setMode(mode)
    try:
        np.testing.assert_array_equal(t.getVisualWords(keys), want)
    finally:
        t.close()


def test_cpp_dropin_traversal_matches_oracle(big_tree, tmp_path):
    """V3D_CUDA::VocabularyTreeTraversal<50, 3> from include/CUDA/v3d_cudamatching.h
    (tests/cpp/voctree_check.cpp): leaf node ids = the CPU tree's words + the
    first leaf's node id (v3d_cudamatching.cu:514-521), in batches, with the
    reference's rounding warning."""
    import os
    import subprocess
    tree, _ = big_tree
    keys = keys_like_sift(55, 10000)
    tree.tofile(tmp_path / "tree.bin")
    keys.tofile(tmp_path / "keys.bin")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "voctree_check")
    say("voctree_check")
    r = subprocess.run([exe, str(tmp_path / "tree.bin"), str(tmp_path / "keys.bin"), str(len(keys)),
                        str(tmp_path / "leaves.bin")], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    assert "nMaxFeatures should be a multiple of 32" in r.stderr
    leaves = np.fromfile(tmp_path / "leaves.bin", np.int32)
    np.testing.assert_array_equal(leaves, OV.quantise(3, 50, tree, keys) + 2550)


def test_cpp_dropin_voctree_classes(big_tree, tmp_path):
    """V3D::VocabularyTree<3, 50, 128> and V3D::LeafDocuments from
    include/Base/v3d_voctree.h (tests/cpp/voctree_dropin_check.cpp, built with
    fused multiply-adds available and contraction allowed): getVisualWord on
    the host and getVisualWords on the GPU both give the oracle's words; the
    driver's per-view binomial (and L1) scores against the earlier views equal
    the inverted-file oracle's; a word outside the tree throws."""
    import os
    import subprocess
    tree, _ = big_tree
    keys = keys_like_sift(56, 3000)
    tree.tofile(tmp_path / "tree.bin")
    keys.tofile(tmp_path / "keys.bin")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "voctree_dropin_check")
    per_view, n_best, p_match = 500, 3, np.float32(0.1)
    say("voctree_dropin_check")
    r = subprocess.run([exe, str(tmp_path / "tree.bin"), str(tmp_path / "keys.bin"), str(len(keys)), str(per_view),
                        "0.1", str(n_best), str(tmp_path / "host.bin"), str(tmp_path / "gpu.bin"),
                        str(tmp_path / "cand.txt")], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, (r.returncode, r.stderr)
    want = OV.quantise(3, 50, tree, keys)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "host.bin", np.int32), want)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "gpu.bin", np.int32), want)
    docs = OV.LeafDocuments(125000)
    expect = []
    for view in range(len(keys) // per_view):
        w = [int(x) for x in want[view * per_view:(view + 1) * per_view]]
        expect += [f"{view} {d} {np.float32(sc)!r}" for sc, d in docs.binomial_scores(float(p_match), w, n_best)]
        expect += [f"L1 {view} {d} {np.float32(sc)!r}" for sc, d in docs.l1_scores(w, n_best)]
        docs.insert(w)
    got = []
    for line in (tmp_path / "cand.txt").read_text().splitlines():
        f = line.split()
        got.append(" ".join(f[:-1] + [repr(np.float32(float(f[-1])))]))
    assert got == expect
    # view v has v earlier documents to score
    assert sum(1 for line in got if not line.startswith("L1")) == sum(min(v, n_best) for v in range(len(keys) // per_view))
