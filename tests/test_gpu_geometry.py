"""The geometric stage after the putative matches, on the GPU: batched guided
matching (one mask per pair, v3dm_batch_set_guides) against the per-pair
runGuidedMatching entry point and the oracle, and the driver applications
with robust orientation supplied by known poses or by a plugin, end to end
against the oracle's restatement of the reference drivers' stage
(SfM/match_sequence/src/match_sequence.cpp:229-346,
SfM/match_using_voctree/src/match_using_voctree.cpp:325-491) on synthetic
scenes of known pose."""
import os
import subprocess
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import geometry as OG
from oracle import oracle as O
from v3dsfmtoolkit_amd import drivers as D
from v3dsfmtoolkit_amd import features as Fe
from v3dsfmtoolkit_amd import geometry as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_sequence")
VOCTREE_APP = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_using_voctree")
PLUGIN = os.path.join(ROOT, "tests", "cpp", "orientation_plugin.so")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rot(rng, scale):
    w = rng.normal(0, scale, 3)
    th = np.linalg.norm(w)
    k = w / th
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx


def unit_noisy_copy(rng, x, sigma):
    w = np.maximum(x + rng.normal(0, sigma, x.shape), 0)
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w = np.minimum(w, 0.2)
    return w / np.linalg.norm(w, axis=1, keepdims=True)


def make_scene(seed, n_views, n_points=2400, n_distract=400, quantise=True, width=1024, height=768):
    """Views of one rigid scene: a quarter of the points in triplets that share
    a descriptor (repeated structure: the ratio test rejects them, guided
    matching recovers them), each view seeing noisy copies of the points' descriptors at
    their projections plus distractors.  Intrinsics with focal length 1024 and
    quarter-pixel principal points keep K^-1 exact."""
    rng = np.random.default_rng(seed)
    n_rep = n_points // 4  # points in triplets
    base = v.synth_pair(v.synth_seed(9, seed), n_points - n_rep + n_rep // 3, 0, u8=False, f32=True)["Lf"]
    base = base.astype(np.float64)
    base = np.r_[base[:n_points - n_rep], np.repeat(base[n_points - n_rep:], 3, axis=0)]
    X = rng.uniform([-4, -3, 8], [4, 3, 14], (len(base), 3))
    views = []
    for i in range(n_views):
        K = np.array([[1024.0, 0, 512 + 0.25 * (i % 5)], [0, 1024.0, 384 - 0.75 * (i % 3)], [0, 0, 1]])
        R = rot(rng, 0.03)
        c = np.array([0.35 * i, 0.05 * np.sin(i), 0.02 * i])
        T = -R @ c
        Xc = (R @ X.T).T + T
        x = (K @ Xc.T).T
        uv = x[:, :2] / x[:, 2:3] + rng.normal(0, 0.3, (len(X), 2))
        vis = np.nonzero((uv[:, 0] >= 0) & (uv[:, 0] < width) & (uv[:, 1] >= 0) & (uv[:, 1] < height))[0]
        desc = unit_noisy_copy(rng, base[vis], 0.03)
        dd = v.synth_pair(v.synth_seed(10, seed * 64 + i), n_distract, 0, u8=False, f32=True)["Lf"].astype(np.float64)
        pos = np.r_[uv[vis], rng.uniform([0, 0], [width, height], (n_distract, 2))]
        desc = np.r_[desc, dd]
        order = rng.permutation(len(desc))
        desc, pos = desc[order], pos[order]
        if quantise:
            desc = np.clip(np.rint(desc * 512), 0, 255) / 512
        views.append({"K": K, "R": R, "T": T, "desc": desc.astype(np.float32), "pos": pos.astype(np.float32),
                      "ids": rng.permutation(10 ** 6)[:len(desc)].astype(np.int32), "dims": (width, height)})
    return views


def relative_pose(va, vb):
    R = vb["R"] @ va["R"].T
    return R, vb["T"] - R @ va["T"]


def write_scene(tmp, views, conf, poses):
    names = []
    for i, vw in enumerate(views):
        name = f"view{i:02d}"
        Fe.write_features(tmp / f"{name}.features", vw["desc"], ids=vw["ids"], positions=vw["pos"])
        names.append(name)
    (tmp / "list.txt").write_text("\n".join(names) + "\n")
    (tmp / "run.conf").write_text("".join(f"{k} {val}\n" for k, val in conf.items()))
    G.write_calibration(tmp / "calibration_db.txt", [vw["K"] for vw in views], [vw["dims"] for vw in views])
    with open(tmp / "poses.txt", "w") as f:
        f.write("# view1 view2 R T\n")
        for (a, b), (R, T) in poses.items():
            f.write(f"{a} {b} " + " ".join(repr(float(x)) for x in [*np.ravel(R), *np.ravel(T)]) + "\n")


def expected(tmp, views, pairs, poses, conf, *, u8, visibility=False, effective_default=True, t_scale=1.0,
             reverse_inliers=False):
    """The oracle's rows: (views, R, T, correspondences) per stored pair, in pair order,
    plus the guided list sizes per pair that reached guided matching."""
    n_max = conf.get("MATCHING_MAX_FEATURES", 7168)
    ratio = conf["MATCHING_SCORE_RATIO"]
    min_raw = conf["MINIMUM_RAW_CORRESPONDENCES"]
    guided = bool(conf.get("USE_GUIDED_MATCHING", 1))
    effective = bool(conf.get("USE_EFFECTIVE_INLIERS", int(effective_default)))
    fs = [Fe.read_features(tmp / f"view{i:02d}.features") for i in range(len(views))]
    rows = []
    for a, b in pairs:
        if (a, b) not in poses:
            continue
        ka, kb = D.kept_count(len(fs[a]), n_max), D.kept_count(len(fs[b]), n_max)
        L, Rd = fs[a].descriptors[:ka], fs[b].descriptors[:kb]
        if u8:
            L, Rd = D.descriptors_to_u8(L), D.descriptors_to_u8(Rd)
        lpos, rpos = fs[a].positions[:ka], fs[b].positions[:kb]
        Rp, Tp = poses[(a, b)]
        K1, K2 = views[a]["K"], views[b]["K"]
        F = G.fundamental_from_pose(K1, K2, Rp, Tp)
        got = OG.geometric_stage(L, Rd, lpos, rpos, K1, K2, views[a]["dims"], views[b]["dims"], (Rp, Tp), F, u8=u8,
                                 ratio=ratio, min_raw=min_raw, guided=guided, effective=effective,
                                 visibility=visibility)
        if got is None:
            continue
        il, ir, pl, pr = got
        if reverse_inliers and not guided:
            il, ir, pl, pr = il[::-1], ir[::-1], pl[::-1], pr[::-1]
        corr = [(pl[k, 0], pl[k, 1], a, fs[a].ids[il[k]], pr[k, 0], pr[k, 1], b, fs[b].ids[ir[k]])
                for k in range(len(il))]
        scale = t_scale if guided else 1.0
        rows.append(((a, b), np.ravel(Rp), np.ravel(Tp) * scale, corr))
    return rows


def read_rows(path):
    import sqlite3
    import struct
    con = sqlite3.connect(str(path))
    data = con.execute("select OID, DATA from matches_data order by OID").fetchall()
    lst = dict(con.execute("select OID, DATA from matches_list").fetchall())
    con.close()
    out = []
    for oid, blob in data:
        v0, v1 = struct.unpack_from("<ii", blob)
        R = struct.unpack_from("<9d", blob, 16)
        T = struct.unpack_from("<3d", blob, 92)
        n = struct.unpack_from("<I", blob, 116)[0]
        corr = [struct.unpack_from("<ffiiffii", blob, 120 + 32 * k) for k in range(n)]
        assert lst[oid] == struct.pack("<ii", v0, v1)
        out.append((oid, (v0, v1), R, T, corr))
    return out


def check_rows(path, want):
    got = read_rows(path)
    assert [g[1] for g in got] == [w[0] for w in want]
    for oid, (g, (views, R, T, corr)) in enumerate(zip(got, want)):
        assert g[0] == oid
        np.testing.assert_array_equal(np.array(g[2]), R, err_msg=f"pair {views}")
        np.testing.assert_array_equal(np.array(g[3]), T, err_msg=f"pair {views}")
        np.testing.assert_array_equal(np.array(g[4], np.float64).reshape(-1, 8), np.array(corr, np.float64).reshape(-1, 8),
                                      err_msg=f"pair {views}")


def run(app, tmp, *args, env=None):
    say(os.path.basename(app), *args)
    r = subprocess.run([app, *args], capture_output=True, text=True, cwd=tmp, timeout=150,
                       env={**os.environ, **(env or {})})
    assert r.returncode == 0, r.stderr
    return r.stdout


CONF = {"MATCHING_MAX_FEATURES": 7168, "MATCHING_SCORE_RATIO": 0.8, "MINIMUM_RAW_CORRESPONDENCES": 20,
        "EG_INLIER_THRESHOLD": 2.0, "MINIMUM_INLIER_CORRESPONDENCES": 20, "MINIMUM_INLIER_MATCHES_RATIO": 0.5,
        "USE_EFFECTIVE_INLIERS": 1, "USE_GUIDED_MATCHING": 1}


def scene_poses(views, pairs, drop=(), wrong=()):
    rng = np.random.default_rng(77)
    poses = {}
    for a, b in pairs:
        if (a, b) in drop:
            continue
        R, T = relative_pose(views[a], views[b])
        if (a, b) in wrong:
            R, T = rot(rng, 0.3) @ R, T + rng.normal(0, 0.5, 3)
        poses[(a, b)] = (R, T)
    return poses


@pytest.mark.parametrize("quantise", [True, False], ids=["uint8_pool", "float_pool"])
def test_match_sequence_guided_with_known_poses(tmp_path, gpu_available, quantise):
    views = make_scene(3 if quantise else 4, 5 if quantise else 3, n_points=2400 if quantise else 1200,
                       quantise=quantise)
    pairs = [tuple(p) for p in D.sequence_pairs(len(views), 2).tolist()]
    poses = scene_poses(views, pairs, drop={pairs[1]}, wrong={pairs[-1]})
    write_scene(tmp_path, views, CONF, poses)
    out = run(APP, tmp_path, "--poses", "poses.txt", "list.txt", "run.conf", "2")
    assert f"{'uint8' if quantise else 'float'} pool)" in out
    want = expected(tmp_path, views, pairs, poses, CONF, u8=quantise)
    assert len(want) >= len(pairs) - 2  # every pair with its true pose is stored
    check_rows(tmp_path / "pairwise_matches.db", want)
    raw = [int(l.split("=")[1]) for l in out.splitlines() if l.startswith("matches.size() =")]
    guided = [int(l.split(":")[1]) for l in out.splitlines() if l.startswith("matches.size() after guided")]
    assert len(raw) == len(pairs) and len(guided) == len(want)
    # guided matching recovers the repeated structure the ratio test rejected
    stored = [p for p in pairs if p in poses and p != pairs[-1]]
    assert all(g > r for g, r in zip(guided, [raw[pairs.index(p)] for p in stored]))
    assert "no relative pose was found" in out and "too few inliers" in out


def test_match_sequence_plugin_hook_order_and_refinement(tmp_path, gpu_available):
    views = make_scene(5, 4)
    pairs = [tuple(p) for p in D.sequence_pairs(len(views), 2).tolist()]
    poses = scene_poses(views, pairs)
    for guided in (1, 0):
        conf = dict(CONF, USE_GUIDED_MATCHING=guided)
        write_scene(tmp_path, views, conf, poses)
        run(APP, tmp_path, "--orientation", PLUGIN, "--db", "plugin.db", "--quiet", "list.txt", "run.conf", "2",
            env={"V3DM_TEST_POSES": str(tmp_path / "poses.txt")})
        # the plugin doubles T at each refinement (before and after guided matching) and hands
        # its inliers back in descending order, which the unguided rows keep
        want = expected(tmp_path, views, pairs, poses, conf, u8=True, t_scale=4.0, reverse_inliers=True)
        assert len(want) == len(pairs)
        check_rows(tmp_path / "plugin.db", want)


def test_match_using_voctree_guided_with_visibility(tmp_path, gpu_available):
    from oracle import voctree as OV
    from v3dsfmtoolkit_amd import voctree as VT
    depth, bf = 3, 16
    tree = VT.synthetic_tree(8, depth, bf)
    tree.tofile(tmp_path / "tree.bin")
    views = make_scene(6, 5)
    conf = dict(CONF, POTENTIAL_VIEWS_COUNT=2, VISUAL_WORD_MATCH_PROPABILITY=0.1)
    del conf["USE_EFFECTIVE_INLIERS"]  # this driver's default (false)
    fs_desc = np.concatenate([vw["desc"] for vw in views])
    words = OV.quantise(depth, bf, tree, fs_desc)
    cands = OV.candidate_pairs([len(vw["desc"]) for vw in views], words, bf ** depth, 2, float(np.float32(0.1)))
    pairs = [(a, b) for a, b, _ in cands]
    poses = scene_poses(views, pairs)
    write_scene(tmp_path, views, conf, poses)
    out = run(VOCTREE_APP, tmp_path, "--tree-shape", str(depth), str(bf), "--poses", "poses.txt", "tree.bin",
              "list.txt", "run.conf")
    want = expected(tmp_path, views, pairs, poses, conf, u8=True, visibility=True, effective_default=False)
    assert len(want) == len(pairs)
    check_rows(tmp_path / "pairwise_matches.db", want)
    assert "after visibility test." in out and "inlierThreshold" not in out


def guided_cases(seed, n_images, counts, quantise):
    """Random images with random positions, pairs with random epipolar and
    homography masks (some without)."""
    rng = np.random.default_rng(seed)
    imgs, pos = [], []
    for i, n in enumerate(counts):
        d = v.synth_pair(v.synth_seed(11, seed * 64 + i), max(n, 1), 0, u8=False, f32=True)["Lf"][:n]
        if i % 2 and n:
            src = imgs[i - 1]
            k = min(n, len(src)) // 2
            d[:k] = unit_noisy_copy(rng, src[:k].astype(np.float64), 0.03)
        if quantise:
            d = np.clip(np.rint(d * 512), 0, 255) / 512
        imgs.append(d.astype(np.float32))
        pos.append(rng.uniform(0, 1000, (n, 2)).astype(np.float32))
    return imgs, pos


@pytest.mark.parametrize("quantise", [True, False], ids=["uint8_pool", "float_pool"])
def test_batch_guides_equal_per_pair_guided_and_oracle(gpu_available, quantise):
    counts = [512, 1024, 700, 1600, 96, 0, 1312, 800]
    imgs, pos = guided_cases(21, len(counts), counts, quantise)
    rng = np.random.default_rng(22)
    pairs = [(0, 1), (2, 3), (3, 2), (1, 6), (4, 7), (5, 0), (6, 7), (7, 1), (1, 1)]
    kinds, mats, thrs = [], [], []
    for k, (a, b) in enumerate(pairs):
        kind = [1, 2, 0][k % 3]
        K = np.array([[900.0, 0, 500], [0, 900, 400], [0, 0, 1]])
        R = rot(rng, 0.1)
        M = G.fundamental_from_pose(K, K, R, rng.normal(0, 1, 3)) if kind == 1 else K @ R @ np.linalg.inv(K)
        kinds.append(kind)
        mats.append(M)
        thrs.append([8.0, 40.0, 1.0][k % 3])
    m = v.SIFTMatching(0)
    single = v.SIFTMatching(0)
    try:
        m.enableRoundingWarning(False)
        data = [(x * 512).round().astype(np.uint8) for x in imgs] if quantise else imgs
        m.setImages(data)
        m.setPositions(pos)
        m.setPairs(pairs)
        m.setGuides(kinds, mats, thrs)
        single.allocate(2048, 2048, True)
        single.setKernel(v.KERNEL_I8 if quantise else v.KERNEL_BF16)
        for variant in range(v.scan_variants_f32() if not quantise else len(v.scan_variants())):
            say("variant", variant)
            m.batchSetVariant(variant)
            tabs = m.batchMatch(0.9, 0.8)
            for k, (a, b) in enumerate(pairs):
                # the batch keeps each image's first multiple-of-32 rows (700 -> 672)
                ka, kb = len(data[a]) & ~31, len(data[b]) & ~31
                mask = [(kinds[k], mats[k], thrs[k], pos[a][:ka], pos[b][:kb])] if kinds[k] else None
                find = O.find_u8 if quantise else O.find_f32
                want = find(data[a][:ka], data[b][:kb], 0.9, 0.8, masks=mask)[2]
                np.testing.assert_array_equal(np.asarray(tabs[k]), want, err_msg=f"variant {variant} pair {k}")
                if variant == 0 and ka and kb:
                    single.setLeftFeatures(ka, data[a][:ka])
                    single.setRightFeatures(kb, data[b][:kb])
                    single.findPutativeMatches(0.9, 0.8)
                    if kinds[k] == 1:
                        got = single.runGuidedMatching(mats[k], thrs[k], pos[a][:ka], pos[b][:kb], 0.9, 0.8)
                    elif kinds[k] == 2:
                        got = single.runGuidedHomographyMatching(mats[k], thrs[k], pos[a][:ka], pos[b][:kb], 0.9,
                                                                 0.8)
                    else:
                        got = single.findPutativeMatches(0.9, 0.8)
                    assert got == O.table_to_pairs(want)
        # the masks apply to later jobs until cleared; new pairs clear them
        m.setGuides(None)
        tabs = m.batchMatch(0.9, 0.8)
        find = O.find_u8 if quantise else O.find_f32
        for k, (a, b) in enumerate(pairs):
            ka, kb = len(data[a]) & ~31, len(data[b]) & ~31
            np.testing.assert_array_equal(np.asarray(tabs[k]), find(data[a][:ka], data[b][:kb], 0.9, 0.8)[2])
    finally:
        m.close()
        single.close()


def test_batch_guides_persistent_grid(gpu_available):
    """More tiles than the device holds at once (persistent scan, separate insertion
    pass), every pair masked."""
    counts = [2048] * 12
    imgs, pos = guided_cases(31, len(counts), counts, True)
    data = [(x * 512).round().astype(np.uint8) for x in imgs]
    pairs = [(i, j) for i in range(12) for j in range(12) if i != j][:48]
    rng = np.random.default_rng(32)
    K = np.array([[900.0, 0, 500], [0, 900, 400], [0, 0, 1]])
    mats = [G.fundamental_from_pose(K, K, rot(rng, 0.1), rng.normal(0, 1, 3)) for _ in pairs]
    thrs = [30.0] * len(pairs)
    m = v.SIFTMatching(0)
    try:
        m.setImages(data)
        m.setPositions(pos)
        m.setPairs(pairs)
        m.setGuides([1] * len(pairs), mats, thrs)
        m.batchSetVariant(0)
        tabs = m.batchMatch(0.9, 0.8)
        assert m.batchInfo()["tiles"] > 1024
        for k, (a, b) in enumerate(pairs):
            want = O.find_u8(data[a], data[b], 0.9, 0.8, masks=[(1, mats[k], thrs[k], pos[a], pos[b])])[2]
            np.testing.assert_array_equal(np.asarray(tabs[k]), want, err_msg=f"pair {k}")
    finally:
        m.close()
