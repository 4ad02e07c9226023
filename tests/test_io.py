"""Feature files, pairwise-match records, the match database and the drivers'
bookkeeping (include/v3dm_io.h) -- all host code, no GPU.

The expected bytes are built here with struct.pack from the reference's
serialisers (V3D/include/Base/v3d_serialization.h:335-421 BinaryOStreamArchive,
:173-256 TextOStreamArchive, :591-689 BlobOArchive; records
V3D/include/Base/v3d_feature.h:41-49, SfM/reconstruction/include/reconstruction_common.h:65-119,
V3D/include/Math/v3d_linearbase.h:45-52,211-220), and the database is read back with
Python's own sqlite3 module, so neither side checks the library against itself."""
import os
import sqlite3
import struct
import subprocess

import numpy as np
import pytest

from v3dsfmtoolkit_amd import V3DMError, _lib
from v3dsfmtoolkit_amd import drivers as D
from v3dsfmtoolkit_amd import features as F
from v3dsfmtoolkit_amd import storage as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_sequence")


def feats(n, seed=0):
    rng = np.random.default_rng(seed)
    return dict(ids=rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32),
                positions=rng.uniform(0, 2000, (n, 2)).astype(np.float32),
                directions=rng.normal(size=(n, 2)).astype(np.float32),
                scales=rng.uniform(1, 30, n).astype(np.float32),
                magnitudes=rng.uniform(0, 5, n).astype(np.float32),
                descriptors=rng.uniform(0, 0.3, (n, 128)).astype(np.float32))


def v3dbin(f):
    """The reference's binary archive of a SerializableVector<SIFT_Feature>, by hand."""
    out = [b"V3DBIN", struct.pack("<I", len(f["ids"]))]
    for k in range(len(f["ids"])):
        out.append(struct.pack("<i", int(f["ids"][k])))
        out.append(struct.pack("<6f", *f["positions"][k], *f["directions"][k], f["scales"][k], f["magnitudes"][k]))
        out.append(struct.pack("<128f", *f["descriptors"][k]))
    return b"".join(out)


def g(x):
    """std::ostream's default float output (precision 6, %g)."""
    return "%g" % float(x)


def v3dtxt(f):
    """The reference's text archive: values then a blank; '{ ' opens a scope, '} ' + newline +
    the enclosing indentation closes it."""
    n = len(f["ids"])
    s = "V3DTXT\n%d { " % n
    for k in range(n):
        vals = [f["positions"][k][0], f["positions"][k][1], f["directions"][k][0], f["directions"][k][1],
                f["scales"][k], f["magnitudes"][k], *f["descriptors"][k]]
        s += "SIFT_Feature { %d " % int(f["ids"][k]) + "".join(g(v) + " " for v in vals) + "} \n  "
    return (s + "} \n").encode()


def assert_features_equal(got, f):
    for key in ("ids", "positions", "directions", "scales", "magnitudes", "descriptors"):
        np.testing.assert_array_equal(getattr(got, key), f[key], err_msg=key)


# --- feature files ---------------------------------------------------------------

def test_reads_hand_built_binary_archive(tmp_path):
    f = feats(5, 1)
    p = tmp_path / "a.features"
    p.write_bytes(v3dbin(f))
    assert os.path.getsize(p) == 6 + 4 + 5 * 540
    assert F.feature_file_info(p) == (5, F.BINARY)
    assert_features_equal(F.read_features(p), f)


def test_binary_writer_matches_hand_built_bytes(tmp_path):
    f = feats(7, 2)
    p = tmp_path / "b.features"
    F.write_features(p, **f)
    assert p.read_bytes() == v3dbin(f)
    assert_features_equal(F.read_features(p), f)


def test_writer_defaults_missing_fields(tmp_path):
    d = feats(3, 3)["descriptors"]
    p = tmp_path / "c.features"
    F.write_features(p, d)
    got = F.read_features(p)
    np.testing.assert_array_equal(got.ids, [0, 1, 2])
    assert not got.positions.any() and not got.scales.any()
    np.testing.assert_array_equal(got.descriptors, d)


def test_text_writer_matches_reference_layout(tmp_path):
    f = feats(3, 4)
    f["descriptors"][0, :4] = [0.5, 1e-5, 123456789.0, -0.0]
    p = tmp_path / "t.features"
    F.write_features(p, binary=False, **f)
    assert p.read_bytes() == v3dtxt(f)
    assert F.feature_file_info(p) == (3, F.TEXT)
    got = F.read_features(p)
    for key in ("positions", "directions", "scales", "magnitudes", "descriptors"):
        want = np.vectorize(lambda v: np.float32(float(g(v))))(f[key]).astype(np.float32)
        np.testing.assert_array_equal(getattr(got, key), want, err_msg=key)
    np.testing.assert_array_equal(got.ids, f["ids"])


def test_text_reader_takes_any_whitespace(tmp_path):
    vals = " ".join(["1.5"] * 6) + "\n" + "\t".join(["0.25"] * 128)
    p = tmp_path / "w.features"
    p.write_bytes(f"V3DTXT\n1\n{{\n SIFT_Feature\t{{ -7 {vals} }}\n}}".encode())
    got = F.read_features(p)
    assert got.ids.tolist() == [-7]
    assert (got.positions == 1.5).all() and (got.descriptors == 0.25).all()


def test_empty_files(tmp_path):
    for binary in (True, False):
        p = tmp_path / f"e{binary}.features"
        F.write_features(p, np.zeros((0, 128), np.float32), binary=binary)
        assert p.read_bytes() == (b"V3DBIN\0\0\0\0" if binary else b"V3DTXT\n0 { } \n")
        assert len(F.read_features(p)) == 0


@pytest.mark.parametrize("content,message", [
    (b"V3DXXX\0\0\0\0", "unknown archive magic"),
    (b"V3DB", "unknown archive magic"),
    (b"V3DBIN\x01\0", "truncated before the record count"),
    (b"V3DTXT\n1 { SIFT_Featur { 1", "tag mismatch"),
    (b"V3DTXT\n1 { SIFT_Feature { 1 2 3", "ends before"),
    (b"V3DTXT\n1 { SIFT_Feature { x", "bad feature id"),
])
def test_malformed_files_are_errors(tmp_path, content, message):
    p = tmp_path / "bad.features"
    p.write_bytes(content)
    with pytest.raises(V3DMError, match=message):
        F.read_features(p)


def test_short_binary_file_is_an_error(tmp_path):
    p = tmp_path / "s.features"
    p.write_bytes(v3dbin(feats(2, 5))[:-1])
    with pytest.raises(V3DMError, match="2 records need"):
        F.read_features(p)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(V3DMError, match="cannot open"):
        F.read_features(tmp_path / "nope.features")


def test_capacity_is_reported(tmp_path):
    import ctypes
    p = tmp_path / "k.features"
    F.write_features(p, **feats(4, 6))
    n = ctypes.c_int()
    rc = _lib.lib().v3dm_features_read(os.fsencode(str(p)), 3, None, None, None, None, None, None, ctypes.byref(n))
    assert rc == _lib.V3DM_ERR_CAPACITY and n.value == 4


# --- drivers' bookkeeping ------------------------------------------------------------

def test_truncated_feature_count_follows_the_drivers():
    for size in [0, 5, 31, 32, 33, 7168, 7200, 9000, 65535, 65536, 65567, 70000, 200000]:
        for n_max in [7168, 1000, 100000]:
            assert D.truncated_feature_count(size, n_max) == min(n_max, size & 0xffe0), (size, n_max)
    assert D.truncated_feature_count(65536 + 100, 7168) == 96  # the reference's wrap
    assert D.kept_count(5000, 1000) == 992                      # nMax 1000 -> setLeftFeatures rounds to 992


def test_sequence_pairs_follow_the_loop():
    for n in [0, 1, 2, 5, 9]:
        for r in [0, 1, 2, 3, 7]:
            want = [(i, j) for i in range(n) for j in range(i + 1, i + 2 * r) if j < n]
            assert D.sequence_pairs(n, r).tolist() == [list(p) for p in want], (n, r)


def test_descriptors_to_u8():
    rng = np.random.default_rng(7)
    u = rng.integers(0, 256, (10, 128)).astype(np.uint8)
    np.testing.assert_array_equal(D.descriptors_to_u8(u.astype(np.float32) / 512), u)
    for bad in [0.5 / 512, -1 / 512, 256 / 512, np.nan, np.inf]:
        d = u.astype(np.float32) / 512
        d[3, 17] = bad
        assert D.descriptors_to_u8(d) is None, bad


# --- pairwise-match records and the database ------------------------------------------

def corrs(n, seed=0):
    rng = np.random.default_rng(seed)
    c = S.correspondences(n)
    for side in ("left", "right"):
        c[side + "_x"] = rng.uniform(0, 1000, n)
        c[side + "_y"] = rng.uniform(0, 1000, n)
        c[side + "_view"] = rng.integers(0, 50, n)
        c[side + "_id"] = rng.integers(0, 10**6, n)
    return c


def pairwise_blob(v0, v1, R, T, c):
    b = struct.pack("<ii", v0, v1) + struct.pack("<ii", 3, 3) + struct.pack("<9d", *np.ravel(R))
    b += struct.pack("<i", 3) + struct.pack("<3d", *T) + struct.pack("<I", len(c))
    for r in c:
        b += struct.pack("<ffii", r["left_x"], r["left_y"], r["left_view"], r["left_id"])
        b += struct.pack("<ffii", r["right_x"], r["right_y"], r["right_view"], r["right_id"])
    return b


def test_pairwise_match_blob_layout_and_round_trip():
    R = np.arange(9, dtype=np.float64).reshape(3, 3) / 7
    T = np.array([0.5, -1.25, 3.0])
    for n in (0, 1, 13):
        c = corrs(n, n)
        m = S.PairwiseMatch((4, 9), R, T, c)
        blob = S.encode_pairwise_match(m)
        assert blob == pairwise_blob(4, 9, R, T, c)
        assert len(blob) == S.pairwise_match_size(n) == 120 + 32 * n
        back = S.decode_pairwise_match(blob)
        assert back.views == (4, 9)
        np.testing.assert_array_equal(back.rotation, R)
        np.testing.assert_array_equal(back.translation, T)
        np.testing.assert_array_equal(back.corrs, c)


def test_pairwise_match_decode_rejects_bad_blobs():
    blob = bytearray(S.encode_pairwise_match(S.PairwiseMatch((0, 1), corrs=corrs(2))))
    with pytest.raises(V3DMError, match="does not hold"):
        S.decode_pairwise_match(bytes(blob[:-1]))
    bad = bytearray(blob)
    bad[8:12] = struct.pack("<i", 4)
    with pytest.raises(V3DMError, match="not 3 x 3"):
        S.decode_pairwise_match(bytes(bad))


def test_match_database_tables_rows_and_truncation(tmp_path):
    p = tmp_path / "m.db"
    ms = [S.PairwiseMatch((i, i + 1), np.eye(3) * (i + 1), np.array([i, 0, 1.0]), corrs(5 + i, i)) for i in range(3)]
    with S.MatchDatabase(p) as db:
        for oid, m in enumerate(ms):
            db.put(oid, m)
    con = sqlite3.connect(str(p))
    schema = dict(con.execute("select name, sql from sqlite_master where type='table'").fetchall())
    assert schema == {"matches_data": "CREATE TABLE matches_data (DATA BLOB)",
                      "matches_list": "CREATE TABLE matches_list (DATA BLOB)"}
    data = con.execute("select OID, DATA from matches_data order by OID").fetchall()
    lst = con.execute("select OID, DATA from matches_list order by OID").fetchall()
    con.close()
    assert [o for o, _ in data] == [0, 1, 2] and [o for o, _ in lst] == [0, 1, 2]
    for (o, blob), m in zip(data, ms):
        assert blob == pairwise_blob(*m.views, m.rotation, m.translation, m.corrs)
    assert [b for _, b in lst] == [struct.pack("<ii", *m.views) for m in ms]
    # "insert or replace" without truncation keeps the other rows
    with S.MatchDatabase(p, truncate=False) as db:
        db.put(1, ms[0])
    con = sqlite3.connect(str(p))
    assert con.execute("select count(*) from matches_data").fetchone()[0] == 3
    assert con.execute("select DATA from matches_list where OID = 1").fetchone()[0] == struct.pack("<ii", 0, 1)
    con.close()
    with S.MatchDatabase(p, truncate=True):
        pass
    con = sqlite3.connect(str(p))
    assert con.execute("select count(*) from matches_data").fetchone()[0] == 0
    con.close()


# --- the counterpart application (no GPU calls here) -------------------------------------

def test_app_usage_and_input_errors(tmp_path):
    assert os.path.exists(APP), "build the library and apps first"
    r = subprocess.run([APP], capture_output=True, text=True)
    assert r.returncode == 2 and "Usage" in r.stderr
    (tmp_path / "list.txt").write_text("view0\nview1\n")
    (tmp_path / "m.conf").write_text("MATCHING_MAX_FEATURES 1024\n")
    r = subprocess.run([APP, "list.txt", "m.conf", "1"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "view0.features" in r.stderr and "cannot open" in r.stderr
    (tmp_path / "pairs.dot").write_text("graph {\n0 -- 5;\n}\n")
    r = subprocess.run([APP, "--pairs", "pairs.dot", "list.txt", "m.conf"], capture_output=True, text=True,
                       cwd=tmp_path)
    assert r.returncode == 1 and "outside the image list" in r.stderr
