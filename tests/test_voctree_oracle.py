"""The vocabulary-tree oracle pinned by what the mathematics fixes, and the
library's host-side inverted file (LeafDocuments, candidate pairs) checked
against it.  No GPU.

Pins of the tree descent (V3D/include/Base/v3d_voctree.h:63-99):
  * a constructed tree whose every path is known in closed form (exact sums),
  * first-index ties between identical siblings,
  * the sequential float order: a cancellation case where the exact sum and
    the left-to-right float sum pick different children,
  * products rounded before the add: a case where a fused multiply-add would
    pick the other child,
  * random trees against an independent greedy descent built on
    numpy's sequential float32 accumulate.
Pins of LeafDocuments (:143-289): the shared-word histogram by brute force,
the binomial score against scipy's binomial log-pmf, the L1 score against the
direct L1 distance of the normalised histograms."""
import math

import numpy as np
import pytest
from scipy.stats import binom

from oracle import voctree as OV
from v3dsfmtoolkit_amd import V3DMError
from v3dsfmtoolkit_amd import voctree as VT


def starts(depth, bf):
    out, s, ln = [], 0, 1
    for _ in range(depth):
        out.append(s)
        ln *= bf
        s += ln
    return out + [s]


def node_index(depth, bf, path):
    """Node id of the child path[d] at each level (leafId recursion of the reference)."""
    nxt, leaf = 0, 0
    for k in path:
        leaf = nxt + k
        nxt = (leaf + 1) * bf
    return leaf


def constructed_tree(depth=3, bf=4):
    """Level d node of path (a_0..a_d) = sum_j 2^-j e_{bf*j + a_j}: the score of
    key = sum_j (1 - 0.1 j) e_{bf*j + b_j} against a child is maximal exactly
    for the child on the key's path, and every sum is exact."""
    tree = np.zeros((starts(depth, bf)[-1], 128), np.float32)
    for d in range(depth):
        for path in np.ndindex(*([bf] * (d + 1))):
            v = np.zeros(128, np.float32)
            for j, a in enumerate(path):
                v[bf * j + a] = 2.0 ** -j
            tree[node_index(depth, bf, path)] = v
    return tree


def greedy_accumulate(depth, bf, tree, key):
    """Independent greedy descent: scores by numpy's sequential float32 accumulate."""
    nxt, leaf = 0, 0
    for _ in range(depth):
        nodes = tree[nxt:nxt + bf].astype(np.float32)
        scores = [np.add.accumulate((key.astype(np.float32) * n).astype(np.float32), dtype=np.float32)[-1]
                  for n in nodes]
        best, m = -1, np.float32(-1e30)
        for k, s in enumerate(scores):
            if s > m:
                best, m = k, s
        leaf = nxt + best
        nxt = (leaf + 1) * bf
    return leaf - starts(depth, bf)[depth - 1]


def test_constructed_tree_paths():
    depth, bf = 3, 4
    tree = constructed_tree(depth, bf)
    for path in np.ndindex(bf, bf, bf):
        key = np.zeros(128, np.float32)
        for j, a in enumerate(path):
            key[bf * j + a] = 1.0 - 0.1 * j
        want = path[0] * bf * bf + path[1] * bf + path[2]
        assert OV.word(depth, bf, tree, key) == want, path


def test_first_index_wins_ties():
    depth, bf = 2, 3
    tree = constructed_tree(depth, bf)
    tree[2] = tree[1]          # root children 1 and 2 identical
    key = np.zeros(128, np.float32)
    key[1] = 1.0
    key[3 + 2] = 0.5           # second-level child 2
    assert OV.word(depth, bf, tree, key) == 1 * bf + 2


def test_sequential_float_order_decides():
    """Child 0's exact score is 1 but its left-to-right float sum is 0
    ((1e8 + 1) - 1e8); child 1 scores 0.5 exactly.  The reference keeps child 1."""
    depth, bf = 1, 2
    tree = np.zeros((2, 128), np.float32)
    key = np.zeros(128, np.float32)
    key[:3] = [1e8, 1.0, -1e8]
    tree[0, :3] = [1.0, 1.0, 1.0]
    key[3] = 1.0
    tree[1, 3] = 0.5
    assert float(np.dot(key.astype(np.float64), tree[0].astype(np.float64))) == 1.0
    assert OV.word(depth, bf, tree, key) == 1
    assert greedy_accumulate(depth, bf, tree, key) == 1


def test_products_are_rounded_before_the_add():
    """Child 0: -1*1 then x*y with x = y = 1 + 2^-12: x*y = 1 + 2^-11 + 2^-24 rounds to
    1 + 2^-11, so the unfused score is 2^-11; a fused multiply-add would give
    2^-11 + 2^-24.  Child 1 scores 2^-11 + 2^-25 exactly: the unfused rule keeps it."""
    depth, bf = 1, 2
    x = np.float32(1 + 2.0 ** -12)
    tree = np.zeros((2, 128), np.float32)
    key = np.zeros(128, np.float32)
    key[0], tree[0, 0] = 1.0, -1.0
    key[1], tree[0, 1] = x, x
    key[2], tree[1, 2] = 1.0, np.float32(2.0 ** -11 + 2.0 ** -25)
    unfused = np.float32(np.float32(-1.0) + np.float32(x * x))
    fused = np.float32(-1.0 + float(x) * float(x))
    assert unfused == np.float32(2.0 ** -11) and fused == np.float32(2.0 ** -11 + 2.0 ** -24)
    assert unfused < tree[1, 2] < fused
    assert OV.word(depth, bf, tree, key) == 1


@pytest.mark.parametrize("depth,bf", [(1, 7), (2, 7), (3, 5)])
def test_random_trees_against_accumulate_descent(depth, bf):
    rng = np.random.default_rng(depth * 10 + bf)
    tree = rng.normal(size=(starts(depth, bf)[-1], 128)).astype(np.float32)
    keys = rng.normal(size=(60, 128)).astype(np.float32)
    got = OV.quantise(depth, bf, tree, keys)
    for k, g in zip(keys, got):
        assert g == greedy_accumulate(depth, bf, tree, k)
    assert got.min() >= 0 and got.max() < bf ** depth


def test_non_finite_key_gets_minus_one():
    tree = constructed_tree(2, 3)
    key = np.zeros(128, np.float32)
    key[5] = np.nan
    assert OV.word(2, 3, tree, key) == -1


# --- LeafDocuments ---------------------------------------------------------------------

def random_docs(seed, n_docs, n_leaves, sizes=(0, 300)):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, n_leaves, rng.integers(*sizes)).astype(np.int32) for _ in range(n_docs)]


def test_document_histogram_is_shared_word_count():
    n_leaves = 50
    docs = random_docs(1, 8, n_leaves, (1, 60))
    ld = OV.LeafDocuments(n_leaves)
    for d in docs:
        ld.insert(d)
    q = np.random.default_rng(2).integers(0, n_leaves, 80)
    want = [sum(1 for w in q if w in set(d.tolist())) for d in docs]
    assert ld.document_histogram(q) == want


def test_binomial_score_is_a_binomial_log_likelihood_ratio():
    n_leaves = 4000
    docs = random_docs(3, 12, n_leaves, (200, 900))
    ld = OV.LeafDocuments(n_leaves)
    for d in docs:
        ld.insert(d)
    q = np.concatenate([docs[4][:150], np.random.default_rng(4).integers(0, n_leaves, 300)])
    k = ld.document_histogram(q)
    checked = 0
    for score, doc in ld.binomial_scores(0.1, q):
        n = min(len(docs[doc]), len(q))
        if k[doc] > n:
            continue
        p0 = max(len(docs[doc]), len(q)) / n_leaves
        p1 = min(1.0, p0 + 0.1)
        want = binom.logpmf(k[doc], n, p0) - binom.logpmf(k[doc], n, p1)
        assert math.isclose(score, want, rel_tol=1e-6, abs_tol=1e-4), (doc, score, want)
        checked += 1
    assert checked >= 10
    best = ld.binomial_scores(0.1, q, 1)
    assert best[0][1] == 4  # the document sharing 150 words scores best


def test_l1_score_is_the_l1_distance_of_normalised_histograms():
    n_leaves = 300
    docs = random_docs(5, 10, n_leaves, (1, 200))
    ld = OV.LeafDocuments(n_leaves)
    for d in docs:
        ld.insert(d)
    q = np.random.default_rng(6).integers(0, n_leaves, 150)
    hq = np.bincount(q, minlength=n_leaves) / len(q)
    for score, doc in ld.l1_scores(q):
        hd = np.bincount(docs[doc], minlength=n_leaves) / len(docs[doc])
        assert math.isclose(score, np.abs(hd - hq).sum(), abs_tol=2e-5)


# --- the library's host inverted file against the oracle ----------------------------------

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_library_leaf_documents_equal_oracle(seed):
    n_leaves = 500
    docs = random_docs(10 + seed, 15, n_leaves, (1, 250))
    docs.append(docs[3].copy())  # an identical document: equal scores, order by document id
    lib_docs, orc_docs = VT.LeafDocuments(n_leaves), OV.LeafDocuments(n_leaves)
    for d in docs:
        assert lib_docs.insertDocument(d) == orc_docs.insert(d)
    assert len(lib_docs) == len(docs)
    q = np.concatenate([docs[3][:40], np.random.default_rng(seed).integers(0, n_leaves, 90)]).astype(np.int32)
    for n_best in (-1, 5, len(docs), 100):
        assert lib_docs.computeBinomialScores(0.1, q, n_best) == orc_docs.binomial_scores(0.1, q, n_best)
        assert lib_docs.compute_L1_Scores(q, n_best) == orc_docs.l1_scores(q, n_best)
    ranks = [d for _, d in lib_docs.computeBinomialScores(0.1, q)]
    assert ranks.index(3) < ranks.index(len(docs) - 1)


def test_leaf_documents_reject_bad_words():
    d = VT.LeafDocuments(10)
    with pytest.raises(V3DMError, match="outside"):
        d.insertDocument([1, 2, 10])
    with pytest.raises(V3DMError, match="outside"):
        d.computeBinomialScores(0.1, [-1])


def test_candidate_pairs_equal_oracle_and_closed_form():
    n_leaves = 800
    rng = np.random.default_rng(7)
    base = [rng.integers(0, n_leaves, 300) for _ in range(4)]
    views = []
    for i in range(12):  # views sharing words with a few earlier ones
        w = rng.integers(0, n_leaves, int(rng.integers(50, 400)))
        if i >= 3:
            w = np.concatenate([w, base[i % 4][:120]])
        views.append(w.astype(np.int32))
    counts = [len(w) for w in views]
    words = np.concatenate(views)
    p = float(np.float32(0.1))
    for n_best in (1, 3, 20):
        got = VT.candidate_pairs(counts, words, n_leaves, n_best, p)
        want = OV.candidate_pairs(counts, words, n_leaves, n_best, p)
        assert got == want
        assert all(a < b for a, b, _ in got)
    # every earlier view is a candidate when n_best covers all of them
    allp = VT.candidate_pairs(counts, words, n_leaves, 100, p)
    assert [(a, b) for a, b, _ in allp] == sorted((a, b) for b in range(12) for a in range(b))
