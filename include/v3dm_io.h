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

/* --- geometric verification around robust orientation --------------------------- */
/* After the putative matches the drivers estimate the relative pose of each
 * pair (computeRobustOrientationMLE: RANSAC on the essential matrix, then
 * refinePoseFromInliers), form F = K2^-T [T]x R K1^-1, run guided matching
 * with it, refine again and store the post-guided correspondences with
 * normalised positions (match_sequence.cpp:229-346,
 * match_using_voctree.cpp:325-491).  Robust orientation and pose refinement
 * are outside this library: the integrator supplies them through this hook
 * (the driver applications load it from a plugin library, --orientation, or
 * use known poses, --poses).  Matrices are row-major doubles; a pose maps
 * view-1 camera coordinates X to view-2 camera coordinates R X + T. */
typedef struct v3dm_orientation_hook
{
   void *user;
   /* computeRobustOrientationMLE(rawCorrs, K1, K2, inlierThreshold, 1000, result,
    * true, mode) (match_sequence.cpp:229-237): corrs hold pixel positions; writes
    * R, T and the indices into corrs of the inliers (at most n; the driver keeps
    * their order) and returns 0, or returns nonzero when no pose was found (the
    * pair is skipped). */
   int (*robust_orientation)(void *user, int view1, int view2, const double K1[9], const double K2[9],
                             double inlierThreshold, int n, const v3dm_correspondence *corrs, double R[9],
                             double T[3], int32_t *inliers, int *nInliers);
   /* refinePoseFromInliers(corrs, K1, K2, R, T) (match_sequence.cpp:53-67), called
    * before and after guided matching (:275, :321): refines R, T in place from pixel
    * correspondences; nonzero skips the pair.  NULL keeps the pose. */
   int (*refine_pose)(void *user, int view1, int view2, const double K1[9], const double K2[9], int n,
                      const v3dm_correspondence *corrs, double R[9], double T[3]);
} v3dm_orientation_hook;

/* The function a plugin library for the drivers' --orientation option exports
 * under the name V3DM_ORIENTATION_PLUGIN_SYMBOL: fills the hook, returns 0. */
typedef int (*v3dm_orientation_plugin_fn)(v3dm_orientation_hook *hook);
#define V3DM_ORIENTATION_PLUGIN_SYMBOL "v3dm_orientation_plugin"

/* F = invertedMatrix(K2)^T * (crossProductMatrix(T) * R) * invertedMatrix(K1)
 * (match_sequence.cpp:279-283) with the reference's double operations
 * (v3d_linear.h:513-530 adjugate inverse, v3d_linearbase.h:1168-1185 products). */
int v3dm_fundamental_from_pose(const double K1[9], const double K2[9], const double R[9], const double T[3],
                               double F[9]);

/* Pixel to normalised positions in place: pos = (K^-1 (x, y, 1))[0..1], K1 for
 * the left measurement, K2 for the right (match_sequence.cpp:327-338,
 * multiply_A_v_affine). */
int v3dm_normalize_correspondences(const double K1[9], const double K2[9], int n, v3dm_correspondence *corrs);

/* getEffectiveInlierCount(corrs, width, height) (match_sequence.cpp:34-50):
 * int(0.5 (coverage(left) + coverage(right)) n + 0.5), coverage as
 * estimateCoverage (SfM/reconstruction/include/reconstruction_common.h:125-157)
 * computes it on a 256-wide raster with the squared distance transform of
 * reconstruction_common2.cpp:17-162; width, height: the left view's image. */
int v3dm_effective_inlier_count(int n, const v3dm_correspondence *corrs, int width, int height, int *count);

/* The visibility test of match_using_voctree.cpp:397-418: keeps (in order, in
 * place) the correspondences whose linear triangulation with P0 = K1 [I | 0]
 * and P1 = K2 [R | T] (v3d_mviewutilities.h:209-258, a least-squares solve)
 * has positive depth in view 1; *kept = their number. */
int v3dm_filter_in_front(const double K1[9], const double K2[9], const double R[9], const double T[3], int n,
                         v3dm_correspondence *corrs, int *kept);

/* CalibrationDatabase(path) (SfM/reconstruction/src/reconstruction_common.cpp:17-38):
 * the view count, then per view K[0][0] K[0][1] K[0][2] K[1][1] K[1][2], four
 * distortion terms, width and height.  Fills K (9 per view, the rest of K
 * the identity's), distortion (4 per view) and dims (width, height per view);
 * each may be NULL.  V3DM_ERR_CAPACITY with *nViews set when cap is smaller. */
int v3dm_calibration_read(const char *path, int cap, double *K, double *distortion, int32_t *dims, int *nViews);

/* The inlier rule of the drivers' known-pose orientation (--poses): indices of
 * the correspondences whose squared Sampson distance under F, in double,
 * is at most threshold^2 -- the epipolar test guided matching applies
 * (v3d_cudamatching.cu:266-290).  No counterpart in the reference, whose
 * inliers come from its RANSAC. */
int v3dm_sampson_inliers(const double F[9], double threshold, int n, const v3dm_correspondence *corrs,
                         int32_t *inliers, int *nInliers);

#ifdef __cplusplus
}
#endif

#endif
