/*
 * v3dm_io.h -- host-side C ABI of the matching stage's data formats and of
 * its drivers' bookkeeping.  Nothing here touches a GPU.
 *
 *  - SIFT feature files (<image>.features): a SerializableVector<SIFT_Feature>
 *    (V3D/include/Base/v3d_serialization.h:972-988, V3D/include/Base/v3d_feature.h:12-49)
 *    written by serializeDataToFile (v3d_serialization.h:884-903) and read by
 *    serializeDataFromFile (:905-947), in the binary archive (magic "V3DBIN",
 *    BinaryOStreamArchive :335-421: little-endian 32-bit integers, raw IEEE
 *    floats, no tags) or the text archive (magic "V3DTXT", TextOStreamArchive
 *    :173-256, TextIStreamArchive :258-324).
 *  - The pairwise-match records the drivers store: PairwiseMatch and ViewPair
 *    (SfM/reconstruction/include/reconstruction_common.h:65-119) serialised by
 *    BlobOArchive (v3d_serialization.h:591-689) into the SQLite tables
 *    matches_data / matches_list (V3D/src/Base/v3d_storage.cpp:70-216,
 *    V3D/include/Base/v3d_storage.h:259-278).
 *  - The drivers' pair enumeration and feature-count truncation
 *    (SfM/match_sequence/src/match_sequence.cpp:134-190).
 *
 * Status codes are those of v3dmatch.h; the message of the last failure of
 * any function here is v3dm_last_error(NULL) (per thread).
 */
#ifndef V3DM_IO_H
#define V3DM_IO_H

#include "v3dmatch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One image measurement of a correspondence, in the byte order PointMeasurement
 * serialises to (V3D/include/Geometry/v3d_mviewutilities.h:14-41: pos[0], pos[1],
 * view, id -- 16 bytes). */
typedef struct v3dm_point_measurement
{
   float x, y;
   int32_t view, id;
} v3dm_point_measurement;

/* PointCorrespondence (v3d_mviewutilities.h:45-68): left, then right; 32 bytes. */
typedef struct v3dm_correspondence
{
   v3dm_point_measurement left, right;
} v3dm_correspondence;

#define V3DM_FEATURES_TEXT 0   /* "V3DTXT" archive */
#define V3DM_FEATURES_BINARY 1 /* "V3DBIN" archive (serializeDataToFile's default) */

/* --- SIFT feature files ----------------------------------------------------- */

/* Header of a feature file: the record count and the archive format
 * (V3DM_FEATURES_BINARY / _TEXT).  serializeDataFromFile's magic check
 * (v3d_serialization.h:913-925): anything else is an error. */
int v3dm_features_info(const char *path, int *count, int *format);

/* serializeDataFromFile(path, SerializableVector<SIFT_Feature>&)
 * (v3d_serialization.h:905-947; record layout v3d_feature.h:41-49).  Fills, for
 * record k < *count: ids[k], positions[2k..2k+1], directions[2k..2k+1],
 * scales[k], magnitudes[k], descriptors[128k..128k+127].  Any output except
 * `count` may be NULL (that field is skipped).  cap is the number of records
 * the arrays hold; a larger file returns V3DM_ERR_CAPACITY with *count set.
 * A short or malformed file (which the reference reads as garbage) is an
 * error here. */
int v3dm_features_read(const char *path, int cap, int32_t *ids, float *positions, float *directions,
                       float *scales, float *magnitudes, float *descriptors, int *count);

/* serializeDataToFile(path, SerializableVector<SIFT_Feature>, format == BINARY)
 * (v3d_serialization.h:884-903).  The text form reproduces TextOStreamArchive's
 * output (six significant digits, as std::ostream prints floats).  NULL
 * positions/directions/scales/magnitudes are written as zeros, NULL ids as
 * 0..n-1; descriptors are required. */
int v3dm_features_write(const char *path, int n, const int32_t *ids, const float *positions,
                        const float *directions, const float *scales, const float *magnitudes,
                        const float *descriptors, int format);

/* --- drivers' bookkeeping ---------------------------------------------------- */

/* The drivers' feature-count truncation, min(nMax, int(size & 0xffe0))
 * (match_sequence.cpp:152,184; match_using_voctree.cpp:272,304): a multiple of
 * 32 that wraps for sizes of 65536 and more, as the reference's does.  Returns
 * the count (size < 0 is treated as 0). */
int v3dm_truncated_feature_count(int64_t size, int nMax);

/* The view pairs match_sequence matches, in its loop order
 * (match_sequence.cpp:134-190): for i < nViews, j = i+1 .. i+2*radius-1, j < nViews.
 * Writes (i, j) int32 pairs; cap in pairs; V3DM_ERR_CAPACITY with *count set
 * when cap is too small. */
int v3dm_sequence_pairs(int nViews, int radius, int32_t *pairs, int cap, int *count);

/* Whether n x 128 float descriptors are exactly u/512 for integers u in
 * [0, 255] (the form on which the int8 path is exact); if so writes u to out
 * (n x 128 bytes) and returns V3DM_OK, else returns V3DM_ERR_UNSUPPORTED with
 * out unspecified.  No counterpart in the reference (its matcher only takes
 * floats); the drivers use it to choose the uint8 image pool. */
int v3dm_descriptors_to_u8(int64_t n, const float *descriptors, uint8_t *out);

/* --- pairwise-match records --------------------------------------------------- */

/* Bytes of a PairwiseMatch blob with nCorrs correspondences:
 * views 8 + rotation (2 int32 + 9 doubles) 80 + translation (int32 + 3 doubles) 28
 * + count 4 + 32 per correspondence (reconstruction_common.h:101-119). */
size_t v3dm_pairwise_match_size(int nCorrs);

/* BlobOArchive << PairwiseMatch (reconstruction_common.h:101-119): ViewPair
 * (view0, view1), rotation row-major (3 x 3 doubles after the dimensions 3, 3),
 * translation (3 doubles after the size 3), then the correspondences
 * (uint32 count, 32 bytes each).  cap bytes at blob; V3DM_ERR_CAPACITY with
 * *size set when too small. */
int v3dm_encode_pairwise_match(int view0, int view1, const double rotation[9], const double translation[3],
                               int nCorrs, const v3dm_correspondence *corrs, uint8_t *blob, size_t cap,
                               size_t *size);

/* BlobIArchive >> PairwiseMatch: the inverse of v3dm_encode_pairwise_match.
 * Checks the matrix/vector dimensions and the blob length.  corrs may be NULL
 * with cap 0 to query *nCorrs (V3DM_ERR_CAPACITY when corrs cannot hold them). */
int v3dm_decode_pairwise_match(const uint8_t *blob, size_t size, int *view0, int *view1, double rotation[9],
                               double translation[3], v3dm_correspondence *corrs, int cap, int *nCorrs);

/* --- match database (SQLite, the reference drivers' output sink) --------------- */

typedef struct v3dm_match_db v3dm_match_db;

/* SQLite3_Database(path) + createTable("matches_data", truncate) +
 * createTable("matches_list", truncate) (v3d_storage.cpp:70-121,
 * match_sequence.cpp:125-127): tables "(DATA BLOB)" keyed by their OID,
 * "pragma synchronous = off" as SQLite3_Database's default mode sets.
 * The system SQLite library (libsqlite3.so.0) is loaded at run time; its
 * absence is reported as V3DM_ERR_UNSUPPORTED.  Rows are written inside one
 * transaction that v3dm_match_db_close commits. */
int v3dm_match_db_open(const char *path, int truncate, v3dm_match_db **out);

/* matchesDB.updateObject("matches_data", oid, PairwiseMatch) and
 * updateObject("matches_list", oid, ViewPair(view0, view1))
 * (match_sequence.cpp:345-346; "insert or replace", v3d_storage.cpp:190,203-216). */
int v3dm_match_db_put(v3dm_match_db *db, int oid, int view0, int view1, const double rotation[9],
                      const double translation[3], int nCorrs, const v3dm_correspondence *corrs);

/* Commits and closes (NULL is a no-op). */
int v3dm_match_db_close(v3dm_match_db *db);

#ifdef __cplusplus
}
#endif

#endif
