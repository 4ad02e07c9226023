"""GPU parity: every kernel's match list against the CPU oracle, bit for bit.

All calls go through the C ABI (libv3dmatch.so) via the SIFTMatching mirror.
u8 descriptors (and floats u/512) must give identical lists on the int8-MFMA
kernel, the plain FP32 kernel and the oracle; general float descriptors give
identical lists on the plain kernel and the oracle's sequential-FMA scores.
"""
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RULES = [(0.9, 0.8), (0.8, 0.8), (0.8, -1.0), (0.5, 0.95), (0.95, 0.6), (0.0, 0.0), (1.1, 0.8)]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def m(gpu_available):
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    mm.allocate(16384, 16384, True)
    yield mm
    mm.close()


def run(m, L, R, ms, mr, kernel=v.KERNEL_AUTO):
    m.setKernel(kernel)
    m.setLeftFeatures(len(L), L)
    m.setRightFeatures(len(R), R)
    return m.findPutativeMatches(ms, mr)


def oracle_pairs(L, R, ms, mr, masks=None):
    return O.table_to_pairs(O.find_u8(L, R, ms, mr, masks=masks)[2])


def pair(cfg, idx, nL, nR, rho=0.4, sigma=0.05, f32=False):
    return v.synth_pair(v.synth_seed(cfg, idx), nL, nR, rho, sigma, f32=f32)


def test_mfma_accumulator_layout(gpu_available):
    say("mfma layout probe")
    out = v.mfma_layout_probe(0)
    lane = np.arange(64)[:, None]
    reg = np.arange(16)[None, :]
    i = 8 * (reg // 4) + 4 * (lane // 32) + reg % 4
    j = lane % 32
    np.testing.assert_array_equal(out, i + 32 * j)


@pytest.mark.parametrize("kernel", [v.KERNEL_I8, v.KERNEL_PLAIN])
@pytest.mark.parametrize("rule", RULES)
def test_c1_2048_matches_oracle(m, kernel, rule):
    say("c1 2048x2048", kernel, rule)
    d = pair(0, 0, 2048, 2048)
    got = run(m, d["L"], d["R"], *rule, kernel=kernel)
    assert m.lastKernel() == kernel
    assert got == oracle_pairs(d["L"], d["R"], *rule)


@pytest.mark.parametrize("rule", [(0.9, 0.05), (0.0, 0.0), (0.8, float("nan")), (float("nan"), 0.8)])
def test_rules_without_a_useful_filter(m, rule):
    """A tiny ratio (no survivor filter: the automatic choice takes the plain path;
    the int8 scan, forced, recomputes every overflowed region), minRatio 0 (only bests
    above 1 pass) and NaN thresholds (nothing passes) -- int8, plain and automatic
    paths all equal the oracle.  The pair mixes copies scaled up by 1.1 so that some
    scores exceed 1."""
    say("filterless rules", rule)
    d = pair(0, 3, 1024, 1088)
    L, R = d["L"], d["R"].copy()
    R[:64] = np.clip(np.rint(L[:64] * 1.1), 0, 255).astype(np.uint8)
    want = oracle_pairs(L, R, *rule)
    for kernel in (v.KERNEL_I8, v.KERNEL_PLAIN, v.KERNEL_AUTO):
        got = run(m, L, R, *rule, kernel=kernel)
        assert got == want, (rule, kernel)
        if kernel == v.KERNEL_AUTO:
            assert m.lastKernel() == (v.KERNEL_PLAIN if rule == (0.9, 0.05) else v.KERNEL_I8)
    if rule == (0.0, 0.0):
        assert len(want) > 0


def test_device_buffers_follow_the_path(m):
    """v3dm_get_device_buffers (the drop-in header's _d_* members): the score
    matrix exists after the plain path only, FP32 descriptors once the plain path
    converted them, positions after a guided call."""
    import ctypes
    say("device buffers")
    d = pair(0, 4, 256, 320)
    ptrs = [ctypes.c_void_p() for _ in range(5)]

    def query():
        assert m._lib.v3dm_get_device_buffers(m._ctx, *[ctypes.byref(p) for p in ptrs]) == 0
        return [bool(p.value) for p in ptrs]

    run(m, d["L"], d["R"], 0.8, 0.8, kernel=v.KERNEL_I8)
    assert query()[2] is False  # the int8 scan stores no scores
    run(m, d["L"], d["R"], 0.8, 0.8, kernel=v.KERNEL_PLAIN)
    left, right, scores, lpos, rpos = query()
    assert left and right and scores and not lpos and not rpos
    lp = np.zeros((256, 2), np.float32)
    rp = np.zeros((320, 2), np.float32)
    m.runGuidedMatching(np.eye(3), 5.0, lp, rp, 0.8, 0.8)
    assert query()[3:] == [True, True]


SHAPES = [(32, 32), (64, 96), (96, 1000), (1000, 96), (2080, 704), (256, 256), (288, 320), (4096, 128),
          (128, 4096), (320, 2336)]


@pytest.mark.parametrize("kernel", [v.KERNEL_I8, v.KERNEL_PLAIN])
@pytest.mark.parametrize("shape", SHAPES)
def test_shapes_match_oracle(m, kernel, shape):
    nL, nR = shape
    say("shape", shape, "kernel", kernel)
    d = pair(0, 100 + nL % 97, nL, nR)
    # the engine keeps the reference's rounding of counts down to multiples of 32
    L32, R32 = d["L"][:nL - nL % 32], d["R"][:nR - nR % 32]
    for rule in [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)]:
        assert run(m, d["L"], d["R"], *rule, kernel=kernel) == oracle_pairs(L32, R32, *rule)


@pytest.mark.parametrize("kernel", [v.KERNEL_I8, v.KERNEL_PLAIN])
def test_rounding_and_clamp(m, kernel):
    say("rounding/clamp", kernel)
    d = pair(0, 3, 1000, 777)
    m.setKernel(kernel)
    m.setLeftFeatures(1000, d["L"])
    m.setRightFeatures(777, d["R"])
    assert m.counts() == (992, 768)  # v3d_cudamatching.cpp:68-77: rounded down to multiples of 32
    assert m.findPutativeMatches(0.8, 0.8) == oracle_pairs(d["L"][:992], d["R"][:768], 0.8, 0.8)
    small = v.SIFTMatching(0)
    small.enableRoundingWarning(False)
    small.allocate(500, 300)  # rounded to 480 x 288
    small.setKernel(kernel)
    small.setLeftFeatures(1000, d["L"])
    small.setRightFeatures(777, d["R"])
    assert small.counts() == (480, 288)
    assert small.findPutativeMatches(0.8, 0.8) == oracle_pairs(d["L"][:480], d["R"][:288], 0.8, 0.8)
    small.close()


@pytest.mark.parametrize("kernel", [v.KERNEL_I8, v.KERNEL_PLAIN])
def test_empty_sides(m, kernel):
    say("empty", kernel)
    d = pair(0, 4, 64, 64)
    assert run(m, d["L"][:0], d["R"], 0.8, 0.8, kernel) == []
    assert run(m, d["L"], d["R"][:0], 0.8, 0.8, kernel) == []
    assert run(m, d["L"][:0], d["R"][:0], 0.8, 0.8, kernel) == []


@pytest.mark.parametrize("kernel", [v.KERNEL_I8, v.KERNEL_PLAIN])
def test_exact_duplicates_first_index_wins(m, kernel):
    say("duplicates", kernel)
    d = pair(0, 5, 512, 640)
    L, R = d["L"].copy(), d["R"].copy()
    L[300] = L[12]
    L[301] = L[12]
    R[100] = R[7]
    R[500:520] = R[200:220]
    src = d["src"]
    k = int(np.flatnonzero(src >= 0)[0])
    R[k + 1 if k + 1 < 640 else k - 1] = R[k]  # duplicated true match: ratio 1, rejected
    for rule in [(0.8, 0.8), (0.8, -1.0), (0.0, 0.0), (0.5, 0.95)]:
        assert run(m, L, R, *rule, kernel=kernel) == oracle_pairs(L, R, *rule)


def test_identical_sets_worst_case_survivors(m):
    # every score of the L = R diagonal survives; many near-duplicates
    say("identity")
    d = pair(0, 6, 1024, 64)
    L = d["L"]
    for rule in [(0.8, 0.8), (0.3, 0.99)]:
        assert run(m, L, L, *rule, kernel=v.KERNEL_I8) == oracle_pairs(L, L, *rule)


@pytest.mark.parametrize("n,inline", [(4096, 1), (16384, 0)])
def test_survivors_applied_in_the_scan_or_by_the_insertion_pass(gpu_available, n, inline):
    """A grid that fits the device at once applies its survivors inside the scan
    (regions, and exact recomputation of a block whose region overflows); a
    larger grid leaves them to the insertion kernel.  Both give the oracle's
    lists, with the filter on and with minRatio 0 (every score survives, every
    region overflows)."""
    import os
    d = pair(3, 17, n, n)
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    try:
        mm.allocate(n, n)
        mm.setScanVariant(0)
        for rule in [(0.8, 0.8), (0.0, 0.0)]:
            say("survivor application", n, rule)
            got = run(mm, d["L"], d["R"], *rule, kernel=v.KERNEL_I8)
            assert mm.lastScanInline() == inline
            tab = O.find_u8_blocked(d["L"], d["R"], *rule, threads=len(os.sched_getaffinity(0)))[2]
            assert got == O.table_to_pairs(tab)
    finally:
        mm.close()


@pytest.mark.parametrize("variant", list(range(len(v.scan_variants()))))
def test_single_pair_every_scan_variant(m, variant):
    """Every int8 scan variant, forced, on single pairs with ragged shapes
    (tiles cut at the padded edge of either side; for the whole-tile form,
    fewer than its 16 stages of rows on a side), with the filter on, with
    minRatio 0 (every score survives, regions overflow and are recomputed) and
    with the ratio test off: all equal the oracle."""
    for nL, nR in [(2080, 704), (320, 2336), (544, 512)]:
        d = pair(0, 300 + nL % 89, nL, nR)
        m.setScanVariant(variant)
        try:
            for rule in [(0.8, 0.8), (0.0, 0.0), (0.8, -1.0)]:
                say("scan variant", variant, (nL, nR), rule)
                got = run(m, d["L"], d["R"], *rule, kernel=v.KERNEL_I8)
                assert m.lastScanVariant() == variant
                assert got == oracle_pairs(d["L"], d["R"], *rule)
        finally:
            m.setScanVariant(-1)


@pytest.mark.parametrize("n,inline", [(8192, 1), (12288, 0)])
def test_whole_tile_form_resident_and_persistent(gpu_available, n, inline):
    """The whole-tile scan (one 512 x 512 tile per workgroup, all of it staged
    in LDS) applies its survivors itself when its grid fits the device at once
    (8192^2: 256 tiles) and leaves them to the insertion kernel when resident
    workgroups pull tiles from the queue (12288^2: 576 tiles); both equal the
    oracle with the filter on and with minRatio 0."""
    import os
    variant = len(v.scan_variants()) - 1
    d = pair(3, 41, n, n)
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    try:
        mm.allocate(n, n)
        mm.setScanVariant(variant)
        for rule in [(0.8, 0.8), (0.0, 0.0)]:
            say("whole-tile scan", n, rule)
            got = run(mm, d["L"], d["R"], *rule, kernel=v.KERNEL_I8)
            assert mm.lastScanVariant() == variant
            assert mm.lastScanInline() == inline
            tab = O.find_u8_blocked(d["L"], d["R"], *rule, threads=len(os.sched_getaffinity(0)))[2]
            assert got == O.table_to_pairs(tab)
    finally:
        mm.close()


@pytest.mark.parametrize("shape,per_tile", [((4, 256, 256), True), ((4, 256, 512), False)])
def test_insertion_pass_per_tile_and_per_region(gpu_available, shape, per_tile):
    """The insertion kernel takes one wave per tile above 32768 survivor
    regions and one wave per region below: a 32768 x 32768 pair has 65536
    regions with 256 x 256 tiles and 32768 with 256 x 512 tiles.  Both give the
    oracle's lists, with the filter on and with minRatio 0 (every region
    overflows and is recomputed from the descriptors)."""
    import os
    variant = v.scan_variants().index(shape)
    waves, rows, cols = shape
    n = 32768
    assert ((n // rows) * (n // cols) * waves > 32768) == per_tile
    d = pair(3, 29, n, n)
    mm = v.SIFTMatching(0)
    mm.enableRoundingWarning(False)
    try:
        mm.allocate(n, n)
        mm.setScanVariant(variant)
        for rule in [(0.8, 0.8), (0.0, 0.0)]:
            say("insertion pass", shape, rule)
            got = run(mm, d["L"], d["R"], *rule, kernel=v.KERNEL_I8)
            assert mm.lastScanInline() == 0
            tab = O.find_u8_blocked(d["L"], d["R"], *rule, threads=len(os.sched_getaffinity(0)))[2]
            assert got == O.table_to_pairs(tab)
    finally:
        mm.close()


def test_float_inputs_dispatch(m):
    say("float dispatch")
    d = pair(0, 8, 700, 900, f32=True)
    Lq = d["L"].astype(np.float32) / np.float32(512)
    Rq = d["R"].astype(np.float32) / np.float32(512)
    got = run(m, Lq, Rq, 0.8, 0.8, v.KERNEL_AUTO)
    assert m.lastKernel() == v.KERNEL_I8
    assert got == oracle_pairs(d["L"][:672], d["R"][:896], 0.8, 0.8)
    # unquantised floats: BF16 candidates + FP32 rescoring, bit-exact with the
    # oracle's sequential-FMA scores; the plain kernel stays selectable
    Lf, Rf = d["Lf"][:672], d["Rf"][:896]
    for rule in [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)]:
        want = O.table_to_pairs(O.find_f32(Lf, Rf, *rule)[2])
        got = run(m, Lf, Rf, *rule, v.KERNEL_AUTO)
        assert m.lastKernel() == v.KERNEL_BF16
        assert got == want
        assert run(m, Lf, Rf, *rule, v.KERNEL_PLAIN) == want
    m.setKernel(v.KERNEL_I8)
    with pytest.raises(v.V3DMError):
        m.findPutativeMatches(0.8, 0.8)
    m.setKernel(v.KERNEL_AUTO)


def _positions(d, nL, nR, seed):
    """Left positions uniform in [0, 1000)^2; a right copy sits near the
    horizontal epipolar line of its source (pure x-translation)."""
    rng = np.random.default_rng(seed)
    lpos = (rng.random((nL, 2)) * 1000).astype(np.float32)
    rpos = (rng.random((nR, 2)) * 1000).astype(np.float32)
    src = d["src"][:nR]
    cp = src >= 0
    rpos[cp, 1] = lpos[src[cp], 1] + rng.normal(0, 2.0, cp.sum()).astype(np.float32)
    rpos[cp, 0] = lpos[src[cp], 0] + 30 + rng.normal(0, 2.0, cp.sum()).astype(np.float32)
    return lpos, rpos


@pytest.mark.parametrize("kernel", [v.KERNEL_I8, v.KERNEL_PLAIN])
def test_guided_epipolar_and_homography(m, kernel):
    say("guided", kernel)
    nL, nR = 1024, 1280
    d = pair(0, 9, nL, nR)
    L, R = d["L"], d["R"]
    lpos, rpos = _positions(d, nL, nR, 1)
    F = np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)  # y-disparity Sampson test
    H = np.array([[1, 0, 30], [0, 1, 0], [0, 0, 1]], np.float64)  # shift by +30 in x
    m.setKernel(kernel)
    m.setLeftFeatures(nL, L)
    m.setRightFeatures(nR, R)
    for rule in [(0.8, 0.8), (0.9, 0.8), (0.8, -1.0)]:
        base = m.findPutativeMatches(*rule)
        assert base == oracle_pairs(L, R, *rule)
        g1 = m.runGuidedMatching(F, 3.0, lpos, rpos, *rule)
        assert g1 == oracle_pairs(L, R, *rule, masks=[(1, F, 3.0, lpos, rpos)])
        # a second guided call accumulates with the first (the score buffer is masked in place)
        g2 = m.runGuidedHomographyMatching(H, 5.0, lpos, rpos, *rule)
        assert g2 == oracle_pairs(L, R, *rule, masks=[(1, F, 3.0, lpos, rpos), (2, H, 5.0, lpos, rpos)])
        # findPutativeMatches recomputes the scores: masks are gone
        assert m.findPutativeMatches(*rule) == base


def test_guided_requires_allocation_flag(gpu_available):
    say("guided flag")
    mm = v.SIFTMatching(0)
    mm.allocate(256, 256)
    d = pair(0, 10, 64, 64)
    mm.setLeftFeatures(64, d["L"])
    mm.setRightFeatures(64, d["R"])
    mm.findPutativeMatches()
    with pytest.raises(v.V3DMError):
        mm.runGuidedMatching(np.eye(3), 1.0, np.zeros((64, 2)), np.zeros((64, 2)))
    mm.close()


def test_async_repeat_is_stable(m):
    say("async repeat")
    d = pair(0, 11, 3000, 2500)
    m.setKernel(v.KERNEL_I8)
    m.setLeftFeatures(3000, d["L"])
    m.setRightFeatures(2500, d["R"])
    want = oracle_pairs(d["L"][:2976], d["R"][:2496], 0.8, 0.8)
    for _ in range(50):
        m.findAsync(0.8, 0.8)
    assert m.fetchMatches() == want
    assert m.findPutativeMatches(0.8, 0.8) == want


def test_c2_8192_i8_equals_plain_and_oracle(m):
    say("c2 8192x8192")
    d = pair(1, 0, 8192, 8192)
    L, R = d["L"], d["R"]
    i8 = run(m, L, R, 0.8, 0.8, v.KERNEL_I8)
    plain = run(m, L, R, 0.8, 0.8, v.KERNEL_PLAIN)
    assert i8 == plain
    c, r, tab, _ = O.find_u8_blocked(L, R, 0.8, 0.8)
    assert i8 == O.table_to_pairs(tab)
    # size-independent property: every match of a planted copy points at its source
    src = d["src"]
    correct = sum(1 for k, l in i8 if src[l] == k)
    assert correct >= 0.95 * len(i8) and len(i8) > 0.3 * 8192


def _lcg_positions(seed, nL, nR, src):
    """Positions exactly as tests/cpp/wrapper_check.cpp makes them."""
    M = (1 << 64) - 1
    s = (seed * 2654435761 + 1) & M
    f32 = np.float32

    def nxt():
        nonlocal s
        s = (s * 6364136223846793005 + 1442695040888963407) & M
        return f32(f32((s >> 33) % 100000) / f32(100.0))

    lp = np.zeros((nL, 2), np.float32)
    rp = np.zeros((nR, 2), np.float32)
    for k in range(nL):
        lp[k, 0] = nxt()
        lp[k, 1] = nxt()
    for k in range(nR):
        rp[k, 0] = nxt()
        rp[k, 1] = nxt()
        if src[k] >= 0:
            rp[k, 0] = lp[src[k], 0] + f32(30.0)
            rp[k, 1] = lp[src[k], 1] + f32(k % 5 - 2)
    return lp, rp


def test_cpp_dropin_header_matches_oracle(gpu_available):
    """The C++ SIFT_Matching drop-in (float descriptors, guided F then H)."""
    import os
    import subprocess

    say("c++ drop-in header")
    exe = os.path.join(os.path.dirname(__file__), "cpp", "wrapper_check")
    nL, nR, seed = 992, 1216, 4242
    out = subprocess.run([exe, str(nL), str(nR), str(seed)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lists, cur = {}, None
    for line in out.stdout.splitlines():
        a, b = line.split()
        if a in ("find", "guided", "homography"):
            cur = a
            lists[cur] = []
        else:
            lists[cur].append((int(a), int(b)))
    d = v.synth_pair(seed, nL, nR, 0.4, 0.05)
    lp, rp = _lcg_positions(seed, nL, nR, d["src"])
    F = np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)
    H = np.array([[1, 0, 30], [0, 1, 0], [0, 0, 1]], np.float64)
    assert lists["find"] == oracle_pairs(d["L"], d["R"], 0.8, 0.8)
    assert lists["guided"] == oracle_pairs(d["L"], d["R"], 0.8, 0.8, masks=[(1, F, 1.2, lp, rp)])
    assert lists["homography"] == oracle_pairs(d["L"], d["R"], 0.8, 0.8,
                                               masks=[(1, F, 1.2, lp, rp), (2, H, 2.0, lp, rp)])
    assert len(lists["guided"]) < len(lists["find"])
