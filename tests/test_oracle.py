"""Pins of the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for this path, so
the oracle is pinned by what fixes it independently:
  * closed forms: exact integer scores against numpy int64 matmul, and the
    FP32 scores of u/512 inputs equal to int * 2^-18;
  * a literal Python transcription of the reference kernels' loops
    (v3d_cudamatching.cu:117-146 and :148-194, including the 32x32 tile loop
    of the row kernel) on small tie-heavy score matrices;
  * hand-computed special cases of the ratio rule;
  * invariants: left/right swap, permutation equivariance, L = R identity;
  * the blocked multithreaded variant against the literal one.
"""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32


# ---- literal transcription of the reference kernels -------------------------

def _decide(maxVal, nextVal, threshold, minRatio):
    threshold = F32(threshold)
    if minRatio < 0:  # computeColumnMaxima / computeRowMaxima: .cu:32, :68
        return bool(maxVal > threshold)
    nextDist = F32(max(0.001, F32(2) - F32(2) * nextVal))  # .cu:138 double max, stored as float
    bestDist = F32(2) - F32(2) * maxVal
    ratio = F32(bestDist / nextDist)
    return bool(maxVal > threshold and ratio < F32(minRatio))


def ref_column_kernel(S, threshold, minRatio):
    """_cuda_computeColumnMaximaWithRatioTest, one thread per column."""
    nRows, nCols = S.shape
    arg_out, best_out, next_out = [], [], []
    for col in range(nCols):
        maxVal, nextVal, argMax = F32(-1e30), F32(-1e30), -1
        for row in range(nRows):
            val = S[row, col]
            argMax = row if val > maxVal else argMax
            nextVal = maxVal if val > maxVal else max(val, nextVal)
            maxVal = max(maxVal, val)
        arg_out.append(argMax if _decide(maxVal, nextVal, threshold, minRatio) else -1)
        best_out.append(maxVal)
        next_out.append(nextVal)
    return np.array(arg_out, np.int32), np.array(best_out, F32), np.array(next_out, F32)


def ref_row_kernel(S, threshold, minRatio, NT=32):
    """_cuda_computeRowMaximaWithRatioTest: blocks of NT rows; each block walks
    the columns in NT x NT tiles staged through shared memory."""
    nRows, nCols = S.shape
    assert nRows % NT == 0 and nCols % NT == 0
    arg_out = np.empty(nRows, np.int32)
    best_out = np.empty(nRows, F32)
    next_out = np.empty(nRows, F32)
    for row0 in range(0, nRows, NT):
        maxVal = [F32(-1e30)] * NT
        nextVal = [F32(-1e30)] * NT
        argMax = [-1] * NT
        for col0 in range(0, nCols, NT):
            tile = S[row0:row0 + NT, col0:col0 + NT]  # scores[k][threadIdx.x] = Scores[startPos0 + k*nCols + ...]
            for tid in range(NT):
                for k in range(NT):
                    val = tile[tid, k]
                    argMax[tid] = col0 + k if val > maxVal[tid] else argMax[tid]
                    nextVal[tid] = maxVal[tid] if val > maxVal[tid] else max(val, nextVal[tid])
                    maxVal[tid] = max(maxVal[tid], val)
        for tid in range(NT):
            ok = _decide(maxVal[tid], nextVal[tid], threshold, minRatio)
            arg_out[row0 + tid] = argMax[tid] if ok else -1
            best_out[row0 + tid] = maxVal[tid]
            next_out[row0 + tid] = nextVal[tid]
    return arg_out, best_out, next_out


def ref_mutual(colMax, rowMax):
    """v3d_cudamatching.cpp:121-126"""
    out = []
    for k in range(len(colMax)):
        l = colMax[k]
        if l >= 0 and rowMax[l] == k:
            out.append((k, int(l)))
    return out


RULES = [(0.9, 0.8), (0.8, 0.8), (0.8, -1.0), (0.5, 0.95), (0.0, 0.0), (0.95, 0.6), (1.1, 0.8)]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("rule", RULES)
def test_passes_match_literal_kernels_on_tie_heavy_scores(seed, rule):
    rng = np.random.default_rng(seed)
    levels = np.array([0.3, 0.79, 0.8, 0.85, 0.9, 0.95, 0.999, 1.0, 1.2], F32)
    S = levels[rng.integers(0, len(levels), size=(64, 96))]
    ms, mr = rule
    a_ref, b_ref, n_ref = ref_column_kernel(S, ms, mr)
    a, b, n = O.column_pass(S, ms, mr)
    np.testing.assert_array_equal(a, a_ref)
    np.testing.assert_array_equal(b, b_ref)
    np.testing.assert_array_equal(n, n_ref)
    a_ref, b_ref, n_ref = ref_row_kernel(S, ms, mr)
    a, b, n = O.row_pass(S, ms, mr)
    np.testing.assert_array_equal(a, a_ref)
    np.testing.assert_array_equal(b, b_ref)
    np.testing.assert_array_equal(n, n_ref)
    col, row, tab = O.select_f32(S, ms, mr)
    assert O.table_to_pairs(tab) == ref_mutual(ref_column_kernel(S, ms, mr)[0], ref_row_kernel(S, ms, mr)[0])


@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)])
def test_passes_match_literal_kernels_on_synthetic_scores(rule):
    import v3dsfmtoolkit_amd as v

    d = v.synth_pair(v.synth_seed(0, 7), 96, 64, 0.4, 0.05)
    S = O.scores_u8(d["L"], d["R"]).astype(np.float64) / 262144.0
    S = S.astype(F32)
    ms, mr = rule
    col, row, tab = O.select_f32(S, ms, mr)
    np.testing.assert_array_equal(col, ref_column_kernel(S, ms, mr)[0])
    np.testing.assert_array_equal(row, ref_row_kernel(S, ms, mr)[0])
    # same rule straight from the descriptors
    col2, row2, tab2 = O.find_u8(d["L"], d["R"], ms, mr)
    np.testing.assert_array_equal(col, col2)
    np.testing.assert_array_equal(row, row2)
    np.testing.assert_array_equal(tab, tab2)


# ---- closed forms ---------------------------------------------------------------

def test_u8_scores_closed_form():
    rng = np.random.default_rng(5)
    L = rng.integers(0, 256, size=(150, 128), dtype=np.uint8)
    R = rng.integers(0, 256, size=(70, 128), dtype=np.uint8)
    S = O.scores_u8(L, R)
    np.testing.assert_array_equal(S, (R.astype(np.int64) @ L.astype(np.int64).T).astype(np.int32))
    # the reference's FP32 SGEMM on u/512 inputs is exact: int * 2^-18
    Sf = O.scores_f32(L.astype(F32) / F32(512), R.astype(F32) / F32(512))
    np.testing.assert_array_equal(Sf, (S.astype(np.float64) / 262144.0).astype(F32))
    assert np.all(S.astype(np.float64) / 262144.0 == Sf.astype(np.float64))


def test_f32_scores_are_sequential_fma():
    rng = np.random.default_rng(6)
    L = rng.random((9, 128), dtype=F32)
    R = rng.random((5, 128), dtype=F32)
    S = O.scores_f32(L, R)
    for r in range(5):
        for l in range(9):
            acc = 0.0
            for k in range(128):  # fmaf: exact product + sum, one rounding
                acc = float(F32(np.float64(L[l, k]) * np.float64(R[r, k]) + np.float64(acc)))
            assert S[r, l] == F32(acc)


@pytest.mark.parametrize("shape", [(37, 29), (64, 8), (5, 3), (130, 61)])
def test_blocked_float_scores_equal_one_at_a_time(shape):
    """The blocked FP32 / FP64 score functions (4 lefts x 8 rights side by
    side) must give every dot product bit for bit as the one-at-a-time loop,
    also for signs, zeros, subnormals and large magnitudes."""
    nL, nR = shape
    rng = np.random.default_rng(nL * 1000 + nR)
    L = rng.standard_normal((nL, 128)).astype(F32)
    R = rng.standard_normal((nR, 128)).astype(F32)
    L[::3, ::7] = F32(0.0)
    L[1::5, 3::11] = F32(-0.0)
    R[::4, 5::9] = F32(1e-40)  # subnormal
    R[2::7, ::13] = F32(3e18)
    np.testing.assert_array_equal(O.scores_f32(L, R).view(np.uint32), O.scores_f32_simple(L, R).view(np.uint32))
    np.testing.assert_array_equal(O.scores_f64(L, R).view(np.uint64), O.scores_f64(L, R, simple=True).view(np.uint64))


# ---- hand-computed special cases ------------------------------------------------

def _col(S, ms=0.8, mr=0.8):
    return O.column_pass(np.array(S, F32), ms, mr)


def test_duplicate_best_gives_ratio_one_and_rejects():
    arg, best, nxt = _col([[0.95], [0.3], [0.95]])
    assert best[0] == F32(0.95) and nxt[0] == F32(0.95)
    assert arg[0] == -1
    arg, _, _ = _col([[0.95], [0.3], [0.95]], 0.8, -1.0)  # ratio test off: first index wins
    assert arg[0] == 0


def test_exact_self_score_passes():
    arg, best, nxt = _col([[0.5], [1.0], [0.99]])  # 0/(2-1.98) = 0
    assert arg[0] == 1


def test_single_candidate_passes_with_init_second():
    arg, best, nxt = _col([[0.81]])
    assert nxt[0] == F32(-1e30) and arg[0] == 0
    arg, _, _ = _col([[0.8]])  # strict > on the score
    assert arg[0] == -1


def test_best_above_one_gives_negative_ratio_and_passes():
    arg, _, _ = _col([[1.2], [1.19]])  # (2-2.4)/(2-2.38) is positive? both negative -> ratio > 0
    # bestDist = -0.4, nextDist = max(0.001, -0.38) = 0.001 -> ratio = -400 < 0.8
    assert arg[0] == 0


def test_ratio_boundary_and_clamp():
    # best 0.9 -> bestDist 0.2 (float); second s with nextDist 0.25 -> ratio 0.8: not < 0.8
    s = F32(1) - F32(0.125)  # 2 - 2 s = 0.25 exactly
    b = F32(0.9)
    ratio = F32((F32(2) - F32(2) * b) / F32(0.25))
    arg, _, _ = _col([[b], [s]], 0.8, float(ratio))
    assert arg[0] == -1  # ratio == minRatio is rejected
    arg, _, _ = _col([[b], [s]], 0.8, float(np.nextafter(ratio, F32(1))))
    assert arg[0] == 0
    # second above 0.9995: nextDist clamps to 0.001
    arg, best, nxt = _col([[1.0], [0.9999]], 0.8, 0.8)
    assert arg[0] == 0  # bestDist = 0 -> ratio 0


def test_all_below_threshold_reject():
    arg, _, _ = _col([[0.5], [0.1]])
    assert arg[0] == -1


# ---- the GPU paths' survivor filter ----------------------------------------------

def _row_with_score(x, z, lo, hi):
    """A uint8 row y whose exact score with x, sum(x y) / 2^18, lies in (lo, hi]:
    y = round(512 w) for w = a x / 512 + b z, bisecting on the blend (scores fall as
    the weight moves from x to z; a scale above 1 reaches scores above 1)."""
    xf = x.astype(np.float64) / 512
    zf = z.astype(np.float64) / 512
    def score(t):
        w = np.clip((1 - t) * xf + t * zf, 0, None) * (1.02 if lo >= 0.999 else 1.0) * (1 - t / 50)
        y = np.clip(np.rint(w * 512), 0, 255).astype(np.uint8)
        return int(x.astype(np.int64) @ y.astype(np.int64)) / 262144.0, y
    a, b = 0.0, 1.0
    for _ in range(60):
        t = 0.5 * (a + b)
        sc, y = score(t)
        if lo < sc <= hi:
            return y
        if sc > hi:
            a = t
        else:
            b = t
    raise AssertionError(f"no row with a score in ({lo}, {hi}]")


def _filter_case():
    """SIFT-like rows, noisy copies, exact duplicates (ties), brighter copies (scores
    near 1.15), and rows built so that scores sit right at the filters' boundaries: best
    scores just above and just below 1 (the minRatio 0 rule), and a best just above 0.8
    with a second best just above 0.75 (the 0.8 / 0.8 rule's T)."""
    rng = np.random.default_rng(7)
    d = _pair(77, 300, 320, 0.5, 0.03)
    L, R = d["L"].copy(), d["R"].copy()
    R[:30] = L[:30] // 2 + rng.integers(0, 8, (30, 128), dtype=np.uint8)
    R[30:40] = np.clip(np.rint(L[:10] * 1.15), 0, 255).astype(np.uint8)  # scores near 1.15 with their lefts
    L[200:210] = L[190:200]
    z = d["R"][300:320]
    L[260:268] = _pair(78, 8, 0)["L"]  # fresh rows: no copy of them elsewhere in R
    for k in range(6):  # bests in (1, 1.002] and (0.998, 1]
        R[240 + k] = _row_with_score(L[240 + k], z[k], 1.0, 1.002)
        R[250 + k] = _row_with_score(L[250 + k], z[k + 6], 0.998, 1.0)
    for k in range(8):  # best in (0.8, 0.8015], second in (0.7501, 0.758] or just below 0.75
        R[260 + k] = _row_with_score(L[260 + k], z[k], 0.8, 0.8015)
        R[270 + k] = _row_with_score(L[260 + k], z[k + 8], 0.7501, 0.758) if k < 5 else \
            _row_with_score(L[260 + k], z[k + 8], 0.745, 0.75)
    return L, R


@pytest.mark.parametrize("rule", [(0.9, 0.8), (0.8, 0.8), (0.5, 0.95), (1.1, 0.8), (0.8, -1.0), (0.0, 0.0),
                                  (0.8, 0.0), (1.5, 0.0), (0.9, 0.05), (0.8, float("nan")),
                                  (float("nan"), 0.8)])
def test_survivor_filter_keeps_every_decision(rule):
    """The scans drop every score at or below the rule's survivor threshold (the
    proof at the top of k_i8.hip).  Dropping them -- the oracle's passes on the
    score matrix with those entries replaced by the passes' -1e30 initial value --
    must leave every decision as it is, and a threshold one step stronger than the
    proof allows must not (on the rows built at the boundaries)."""
    from oracle.learned import survivor_threshold
    L, R = _filter_case()
    S = O.scores_u8(L, R).astype(np.int64)  # exact integer scores [right, left]
    thr = survivor_threshold(*rule)

    def decide(t):
        kept = np.where(S > t, S / 262144.0, -1e30).astype(np.float32)
        return O.select_f32(kept, *rule)[2]

    full = O.select_f32((S / 262144.0).astype(np.float32), *rule)[2]
    np.testing.assert_array_equal(decide(thr), full)
    assert np.array_equal(full, O.find_u8(L, R, *rule)[2])
    if rule in ((0.0, 0.0), (0.8, 0.0)):
        # only rows whose best is above 1 pass; the boundary rows sit on both sides
        assert np.all(S.max(axis=0)[full >= 0] > 262144) and (full[240:246] >= 0).all()
        assert not np.array_equal(decide(int(262144 * 1.002)), full)
    if rule == (0.8, 0.8):
        assert thr < int(0.75 * 262144) + 1 and not np.array_equal(decide(int(0.758 * 262144)), full)
    if np.isnan(rule[0]) or np.isnan(rule[1]):
        assert (full < 0).all() and thr == 1 << 28


# ---- invariants --------------------------------------------------------------------

def _pair(seed, nL, nR, rho=0.4, sigma=0.05):
    import v3dsfmtoolkit_amd as v
    return v.synth_pair(v.synth_seed(0, seed), nL, nR, rho, sigma)


@pytest.mark.parametrize("rule", [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)])
def test_swap_left_right_transposes_list(rule):
    d = _pair(11, 200, 160)
    _, _, tab = O.find_u8(d["L"], d["R"], *rule)
    _, _, tab_sw = O.find_u8(d["R"], d["L"], *rule)
    fwd = set(O.table_to_pairs(tab))
    bwd = {(l, k) for k, l in O.table_to_pairs(tab_sw)}
    assert fwd == bwd and len(fwd) > 10


def test_permutation_equivariance():
    d = _pair(12, 180, 220)
    perm = np.random.default_rng(1).permutation(220)
    _, _, tab = O.find_u8(d["L"], d["R"])
    _, _, tab_p = O.find_u8(d["L"], d["R"][perm])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    expect = {(k, int(inv[l])) for k, l in O.table_to_pairs(tab)}
    assert set(O.table_to_pairs(tab_p)) == expect


def test_identity_matches_diagonal():
    d = _pair(13, 256, 32)
    L = d["L"]
    _, _, tab = O.find_u8(L, L, 0.8, 0.8)
    assert np.array_equal(tab, np.arange(256))


# ---- blocked CPU baseline == literal oracle ---------------------------------------

@pytest.mark.parametrize("shape", [(64, 64), (300, 257), (1, 50), (50, 1), (2048, 2048)])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_blocked_equals_literal(shape, threads):
    nL, nR = shape
    d = _pair(20 + nL, nL, nR)
    for rule in [(0.8, 0.8), (0.8, -1.0), (0.5, 0.95)]:
        c1, r1, t1 = O.find_u8(d["L"], d["R"], *rule)
        c2, r2, t2, used = O.find_u8_blocked(d["L"], d["R"], *rule, threads=threads)
        np.testing.assert_array_equal(c1, c2)
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(t1, t2)


def test_blocked_with_exact_duplicates():
    d = _pair(30, 128, 128)
    L = d["L"].copy()
    L[50] = L[10]  # exact duplicate rows -> tied scores, first index must win
    L[90] = L[10]
    for rule in [(0.8, -1.0), (0.8, 0.8)]:
        c1, r1, t1 = O.find_u8(L, d["R"], *rule)
        c2, r2, t2, _ = O.find_u8_blocked(L, d["R"], *rule, threads=4)
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(t1, t2)


# ---- guided masks: closed forms ----------------------------------------------------

def test_epipolar_sampson_closed_form():
    # F = [t]_x for t = (1, 0, 0): epipolar lines are horizontal; the Sampson
    # distance of (p, q) is (p_y - q_y)^2 / 2.
    F = np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)
    p = (F32(10.0), F32(4.0))
    for qy, thr in [(4.0, 0.1), (5.0, 0.70), (5.0, 0.71), (7.0, 2.0), (7.0, 2.2)]:
        sq = F32(thr) * F32(thr)
        d2 = (4.0 - qy) ** 2 / 2.0
        assert O.epipolar_masked(F, sq, p, (F32(3.0), F32(qy))) == (d2 > float(sq))


def test_homography_transfer_closed_form():
    H = np.array([[1, 0, 5], [0, 1, -3], [0, 0, 1]], np.float64)
    p = (F32(2.0), F32(2.0))  # maps to (7, -1)
    for q, thr in [((7.0, -1.0), 0.5), ((8.0, -1.0), 0.99), ((8.0, -1.0), 1.01), ((7.0, 1.0), 1.5)]:
        sq = F32(thr) * F32(thr)
        d2 = (7.0 - q[0]) ** 2 + (-1.0 - q[1]) ** 2
        assert O.homography_masked(H, sq, p, (F32(q[0]), F32(q[1]))) == (d2 > float(sq))


def test_guided_find_equals_masked_selection():
    d = _pair(40, 96, 128)
    rng = np.random.default_rng(3)
    lpos = (rng.random((96, 2)) * 100).astype(F32)
    rpos = (rng.random((128, 2)) * 100).astype(F32)
    F = np.array([[0, -0.001, 0.05], [0.001, 0, -0.2], [-0.04, 0.21, 0.003]], np.float64)
    masks = [(1, F, 3.0, lpos, rpos)]
    col, row, tab = O.find_u8(d["L"], d["R"], 0.8, 0.8, masks=masks)
    S = (O.scores_u8(d["L"], d["R"]).astype(np.float64) / 262144.0).astype(F32)
    sq = F32(3.0) * F32(3.0)
    for r in range(128):
        for l in range(96):
            if O.epipolar_masked(F, sq, lpos[l], rpos[r]):
                S[r, l] = F32(-1e20)
    c2, r2, t2 = O.select_f32(S, 0.8, 0.8)
    np.testing.assert_array_equal(tab, t2)
    np.testing.assert_array_equal(col, ref_column_kernel(S, 0.8, 0.8)[0])


# ---- generator ----------------------------------------------------------------------

def test_generator_is_deterministic_and_pinned():
    import v3dsfmtoolkit_amd as v

    d1 = v.synth_pair(v.synth_seed(0, 0), 256, 256, 0.4, 0.05, f32=True, threads=1)
    d2 = v.synth_pair(v.synth_seed(0, 0), 256, 256, 0.4, 0.05, f32=True, threads=8)
    for k in ("L", "R", "Lf", "Rf", "src"):
        np.testing.assert_array_equal(d1[k], d2[k])
    h = hashlib.sha256(d1["L"].tobytes() + d1["R"].tobytes() + d1["src"].tobytes()).hexdigest()
    assert h == GENERATOR_SHA256_256


GENERATOR_SHA256_256 = "299d2a7a2eb6d518c6d8a782aa93d1a3cc6b76889b871acae3a186d883d6f288"


def test_generator_properties():
    import v3dsfmtoolkit_amd as v

    d = v.synth_pair(v.synth_seed(1, 0), 500, 1000, 0.4, 0.05, f32=True)
    Lf, Rf, L, R, src = d["Lf"], d["Rf"], d["L"], d["R"], d["src"]
    np.testing.assert_allclose(np.linalg.norm(Lf, axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(Rf, axis=1), 1.0, atol=1e-6)
    assert Lf.min() >= 0
    np.testing.assert_array_equal(L, np.minimum(255, np.floor(512.0 * Lf.astype(np.float64) + 0.5)).astype(np.uint8))
    assert (src >= 0).sum() == 400  # floor(0.4 * 1000) distinct sources
    assert len(set(src[src >= 0].tolist())) == 400
    # copies are close to their sources, independent draws are not
    cos = np.einsum("ij,ij->i", Rf[src >= 0], Lf[src[src >= 0]])
    assert cos.mean() > 0.8
    cap = v.synth_pair(5, 10, 100, 0.4, 0.05)["src"]
    assert (cap >= 0).sum() == 10  # capped at nL
