"""The counterpart driver end to end on the GPU: feature files on disk ->
v3dm_match_sequence (batched jobs) -> SQLite rows, checked against the oracle
run pair by pair the way the reference's match_sequence loop does it
(SfM/match_sequence/src/match_sequence.cpp:134-229: truncation, minScore 0.9,
MATCHING_SCORE_RATIO, MINIMUM_RAW_CORRESPONDENCES, ids and pixel positions).
The database is decoded with sqlite3 + struct, independently of the library."""
import os
import sqlite3
import struct
import subprocess
import sys

import numpy as np
import pytest

import v3dsfmtoolkit_amd as v
from oracle import oracle as O
from v3dsfmtoolkit_amd import drivers as D
from v3dsfmtoolkit_amd import features as F

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_sequence")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def unit_noisy_copy(rng, x, sigma):
    """Noisy copy of unit SIFT-like rows: add N(0, sigma), clip at 0, normalise, cap 0.2, normalise."""
    w = np.maximum(x + rng.normal(0, sigma, x.shape), 0)
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w = np.minimum(w, 0.2)
    return w / np.linalg.norm(w, axis=1, keepdims=True)


def make_views(seed, counts, quantise, sigma=0.02, share=0.5):
    """Views where about half of each view's rows are noisy copies of the previous view's."""
    rng = np.random.default_rng(seed)
    views = []
    for i, n in enumerate(counts):
        x = v.synth_pair(v.synth_seed(5, seed * 64 + i), max(n, 1), 0, u8=False, f32=True)["Lf"][:n].astype(np.float64)
        if i and n and len(views[-1]):
            prev = views[-1]
            k = min(n, len(prev)) // 2
            dst, src = rng.choice(n, k, replace=False), rng.choice(len(prev), k, replace=False)
            x[dst] = unit_noisy_copy(rng, prev[src].astype(np.float64), sigma)
        if quantise:
            x = np.clip(np.rint(x * 512), 0, 255) / 512
        views.append(x.astype(np.float32))
    return views


def write_run(tmp, views, binary, seed, conf):
    rng = np.random.default_rng(seed)
    names = []
    for i, d in enumerate(views):
        n = len(d)
        name = f"img{i:03d}"
        F.write_features(tmp / f"{name}.features", d, ids=rng.permutation(10**6)[:n].astype(np.int32),
                         positions=rng.uniform(0, 1600, (n, 2)).astype(np.float32),
                         scales=rng.uniform(1, 20, n).astype(np.float32), binary=binary[i])
        names.append(name)
    (tmp / "list.txt").write_text("\n".join(names) + "\n")
    (tmp / "run.conf").write_text("".join(f"{k} {val}\n" for k, val in conf.items()))
    return names


def read_db(path):
    con = sqlite3.connect(str(path))
    rows = con.execute("select OID, DATA from matches_data order by OID").fetchall()
    lst = dict(con.execute("select OID, DATA from matches_list").fetchall())
    con.close()
    out = []
    for oid, blob in rows:
        v0, v1, r, c = struct.unpack_from("<iiii", blob)
        assert (r, c) == (3, 3)
        R = struct.unpack_from("<9d", blob, 16)
        assert struct.unpack_from("<i", blob, 88)[0] == 3
        T = struct.unpack_from("<3d", blob, 92)
        n = struct.unpack_from("<I", blob, 116)[0]
        assert len(blob) == 120 + 32 * n
        corr = [struct.unpack_from("<ffiiffii", blob, 120 + 32 * k) for k in range(n)]
        assert lst[oid] == struct.pack("<ii", v0, v1)
        out.append((oid, (v0, v1), R, T, corr))
    return out


def expected_rows(tmp, names, pairs, n_max, ratio, min_raw, use_u8):
    fs = [F.read_features(tmp / f"{nm}.features") for nm in names]
    rows, sizes = [], []
    for a, b in pairs:
        ka, kb = D.kept_count(len(fs[a]), n_max), D.kept_count(len(fs[b]), n_max)
        L, R = fs[a].descriptors[:ka], fs[b].descriptors[:kb]
        if use_u8:
            tab = O.find_u8(D.descriptors_to_u8(L), D.descriptors_to_u8(R), 0.9, ratio)[2]
        else:
            tab = O.find_f32(L, R, 0.9, ratio)[2]
        corr = [(fs[a].positions[i, 0], fs[a].positions[i, 1], a, fs[a].ids[i],
                 fs[b].positions[j, 0], fs[b].positions[j, 1], b, fs[b].ids[j]) for i, j in O.table_to_pairs(tab)]
        sizes.append(len(corr))
        if len(corr) >= min_raw:
            rows.append(((a, b), corr))
    return rows, sizes


def run_app(tmp, *args):
    say("v3dm_match_sequence", *args)
    r = subprocess.run([APP, "--db", str(tmp / "out.db"), *args], capture_output=True, text=True, cwd=tmp,
                       timeout=100)
    assert r.returncode == 0, r.stderr
    return r.stdout


def check_rows(tmp, want):
    got = read_db(tmp / "out.db")
    assert len(got) == len(want)
    for oid, (row, (views, corr)) in enumerate(zip(got, want)):
        assert row[0] == oid and row[1] == views
        assert row[2] == (1, 0, 0, 0, 1, 0, 0, 0, 1) and row[3] == (0, 0, 0)
        g = np.array(row[4], dtype=np.float64).reshape(-1, 8)
        w = np.array(corr, dtype=np.float64).reshape(-1, 8)
        np.testing.assert_array_equal(g, w, err_msg=f"pair {views}")


@pytest.mark.parametrize("quantise", [True, False], ids=["uint8_pool", "float_pool"])
def test_match_sequence_matches_per_pair_oracle(tmp_path, gpu_available, quantise):
    counts = [700, 1300, 2100, 64, 1500, 999, 40, 1200]
    views = make_views(11 if quantise else 12, counts, quantise)
    conf = {"MATCHING_MAX_FEATURES": 1024, "MATCHING_SCORE_RATIO": 0.8, "MINIMUM_RAW_CORRESPONDENCES": 20}
    binary = [True] * len(views)
    if not quantise:
        binary[2] = False  # one text archive (read back at six digits, as the reference's reader does)
    names = write_run(tmp_path, views, binary, 3, conf)
    out = run_app(tmp_path, "list.txt", "run.conf", "2")
    pairs = D.sequence_pairs(len(views), 2).tolist()
    want, sizes = expected_rows(tmp_path, names, pairs, 1024, 0.8, 20, quantise)
    assert f"{'uint8' if quantise else 'float'} pool)" in out
    assert [int(l.split("=")[1]) for l in out.splitlines() if l.startswith("matches.size()")] == sizes
    assert sum(1 for s in sizes if s >= 20) >= 3  # the data do produce stored pairs
    check_rows(tmp_path, want)


def test_pair_list_mode_and_job_chunks(tmp_path, gpu_available):
    views = make_views(21, [900, 1100, 800, 1000, 1200], True)
    conf = {"MATCHING_MAX_FEATURES": 7168, "MATCHING_SCORE_RATIO": 0.7, "MINIMUM_RAW_CORRESPONDENCES": 1}
    names = write_run(tmp_path, views, [True] * 5, 4, conf)
    pairs = [(0, 1), (3, 4), (1, 2), (0, 4), (2, 3), (1, 0)]
    (tmp_path / "putative_EG.dot").write_text("graph {\n" + "".join(f"{a} -- {b};\n" for a, b in pairs) + "}\n")
    run_app(tmp_path, "--pairs-per-job", "4", "--quiet", "--pairs", "putative_EG.dot", "list.txt", "run.conf")
    want, _ = expected_rows(tmp_path, names, pairs, 7168, 0.7, 1, True)
    check_rows(tmp_path, want)


def test_python_match_views_equals_oracle(tmp_path, gpu_available):
    views = make_views(31, [640, 2000, 1024, 33], False)
    fs = []
    for i, d in enumerate(views):
        F.write_features(tmp_path / f"v{i}.features", d, ids=np.arange(len(d), dtype=np.int32) + 1000 * i)
        fs.append(F.read_features(tmp_path / f"v{i}.features"))
    pairs = [(0, 1), (1, 2), (2, 3), (0, 2)]
    m = v.SIFTMatching(0)
    try:
        got = D.match_views(m, fs, pairs, n_max=1536, min_ratio=0.8)
    finally:
        m.close()
    for (a, b), c in zip(pairs, got):
        ka, kb = D.kept_count(len(fs[a]), 1536), D.kept_count(len(fs[b]), 1536)
        tab = O.find_f32(fs[a].descriptors[:ka], fs[b].descriptors[:kb], 0.9, 0.8)[2]
        want = O.table_to_pairs(tab)
        assert [(int(x) - 1000 * a, int(y) - 1000 * b) for x, y in zip(c["left_id"], c["right_id"])] == want
        assert (c["left_view"] == a).all() and (c["right_view"] == b).all()


VOCTREE_APP = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_using_voctree")


def test_match_using_voctree_matches_oracles(tmp_path, gpu_available):
    """Stage 1 (quantisation + binomial candidate choice) against the voctree
    oracles, the candidate graph file, and stage 2 against the per-pair oracle."""
    from oracle import voctree as OV
    from v3dsfmtoolkit_amd import voctree as VT
    depth, bf = 2, 40
    tree = VT.synthetic_tree(8, depth, bf)
    tree.tofile(tmp_path / "tree.bin")
    views = make_views(41, [900, 1100, 700, 1300, 800, 1000, 1200, 600], True)
    conf = {"MATCHING_MAX_FEATURES": 1024, "MATCHING_SCORE_RATIO": 0.8, "MINIMUM_RAW_CORRESPONDENCES": 10,
            "POTENTIAL_VIEWS_COUNT": 2, "VISUAL_WORD_MATCH_PROPABILITY": 0.1}
    names = write_run(tmp_path, views, [True] * len(views), 5, conf)
    say("v3dm_match_using_voctree")
    r = subprocess.run([VOCTREE_APP, "--db", str(tmp_path / "out.db"), "--tree-shape", str(depth), str(bf), "tree.bin",
                        "list.txt", "run.conf"], capture_output=True, text=True, cwd=tmp_path, timeout=100)
    assert r.returncode == 0, r.stderr
    fs = [F.read_features(tmp_path / f"{nm}.features") for nm in names]
    words = OV.quantise(depth, bf, tree, np.concatenate([f.descriptors for f in fs]))
    cands = OV.candidate_pairs([len(f) for f in fs], words, bf ** depth, 2, float(np.float32(0.1)))
    pairs = [(a, b) for a, b, _ in cands]
    assert len(pairs) == 2 * (len(views) - 1) - 1  # view 1 has one earlier view, the others two
    dot = (tmp_path / "putative_EG.dot").read_text()
    assert dot == "graph {\n" + "".join(f"{a} -- {b};\n" for a, b in pairs) + "}\n"
    printed, cur = {}, None
    for line in r.stdout.splitlines():
        if line.startswith("Potential matches for "):
            cur = line[len("Potential matches for "):-1]
            printed[cur] = []
        elif line.startswith(" ") and " with score " in line:
            printed[cur].append(line.split()[0])
    for i, nm in enumerate(names):
        mine = sorted((s, a) for a, b, s in cands if b == i)
        assert printed[nm] == [names[a] for _, a in mine], nm
    want, _ = expected_rows(tmp_path, names, pairs, 1024, 0.8, 10, True)
    assert len(want) >= 3
    check_rows(tmp_path, want)
