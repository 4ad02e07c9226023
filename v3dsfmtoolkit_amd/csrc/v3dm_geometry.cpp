// v3dm_geometry.cpp -- the host-side arithmetic around the drivers' robust
// orientation step (declared in include/v3dm_io.h): the fundamental matrix
// handed to guided matching, the effective-inlier count, the cheirality test
// of the voctree driver, the normalisation of stored positions and the
// calibration database.  Robust orientation itself (RANSAC, pose refinement)
// is the integrator's, through v3dm_orientation_hook.
//
// Every double and float operation follows the reference's expression order
// (V3D/include/Math/v3d_linear.h, v3d_linearbase.h, the drivers): built with
// -ffp-contract=off, so no multiply-add is fused.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/v3dm_io.h"

namespace v3dm_host {
extern thread_local std::string g_createError;
}
using v3dm_host::g_createError;

namespace {

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...)
{
   char buf[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof buf, fmt, ap);
   va_end(ap);
   g_createError = buf;
   return code;
}

int ok()
{
   g_createError.clear();
   return V3DM_OK;
}

// multiply_A_B (v3d_linearbase.h:1168-1185): accum = 0, then += a[r][k] * b[k][c]
// for k ascending.  A is rows x inner, B inner x cols, all row-major.
void mulAB(const double *A, const double *B, double *C, int rows, int inner, int cols)
{
   for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c)
      {
         double accum = 0;
         for (int k = 0; k < inner; ++k) accum += A[r * inner + k] * B[k * cols + c];
         C[r * cols + c] = accum;
      }
}

// invertedMatrix for 3 x 3 (v3d_linear.h:513-530): the adjugate scaled by
// 1.0 / det, det summed in the reference's term order.
void inv3(const double *A, double *R)
{
   const double a = A[0], b = A[1], c = A[2];
   const double d = A[3], e = A[4], f = A[5];
   const double g = A[6], h = A[7], i = A[8];
   const double det = a * e * i + b * f * g + c * d * h - c * e * g - b * d * i - a * f * h;
   R[0] = e * i - f * h;
   R[1] = c * h - b * i;
   R[2] = b * f - c * e;
   R[3] = f * g - d * i;
   R[4] = a * i - c * g;
   R[5] = c * d - a * f;
   R[6] = d * h - e * g;
   R[7] = b * g - a * h;
   R[8] = a * e - b * d;
   const double s = 1.0 / det;
   for (int k = 0; k < 9; ++k) R[k] *= s; // scaleMatrixIP (v3d_linearbase.h:944-951)
}

// multiply_A_v_affine (v3d_linearbase.h:986-1000) with a 3 x 3 matrix and a
// float 2-vector: dst[r] = m[r][2], then += m[r][0] in[0], += m[r][1] in[1].
void affine(const double *m, float x, float y, double *dst)
{
   for (int r = 0; r < 3; ++r) dst[r] = m[3 * r + 2];
   for (int r = 0; r < 3; ++r)
   {
      dst[r] += m[3 * r + 0] * (double)x;
      dst[r] += m[3 * r + 1] * (double)y;
   }
}

// One pass of the squared-distance transform of a sampled function along a
// line (Felzenszwalb & Huttenlocher, "Distance Transforms of Sampled
// Functions", 2004): the lower envelope of the parabolas (q - p)^2 + f(p).
// The arithmetic types are those of the reference's pass
// (SfM/reconstruction/src/reconstruction_common2.cpp:17-49): float samples,
// integer squares and denominators, float boundaries, +-1e20 as infinity.
void envelopePass(const float *f, int n, float *d, std::vector<int> &v, std::vector<float> &z)
{
   const float inf = (float)1e20;
   v.assign(n, 0);
   z.assign(n + 1, 0.0f);
   int k = 0;
   v[0] = 0;
   z[0] = -inf;
   z[1] = +inf;
   for (int q = 1; q <= n - 1; ++q)
   {
      float s = ((f[q] + q * q) - (f[v[k]] + v[k] * v[k])) / (2 * q - 2 * v[k]);
      while (s <= z[k])
      {
         --k;
         s = ((f[q] + q * q) - (f[v[k]] + v[k] * v[k])) / (2 * q - 2 * v[k]);
      }
      ++k;
      v[k] = q;
      z[k] = s;
      z[k + 1] = +inf;
   }
   k = 0;
   for (int q = 0; q <= n - 1; ++q)
   {
      while (z[k + 1] < q) ++k;
      d[q] = (q - v[k]) * (q - v[k]) + f[v[k]];
   }
}

// estimateCoverage (SfM/reconstruction/include/reconstruction_common.h:125-157,
// targetWidth 256): the share of a width-256 raster within the radius that n
// uniform points would need to cover it, of the nearest point.
float coverage(const std::vector<float> &px, const std::vector<float> &py, int width, int height)
{
   const int targetWidth = 256;
   const float downScaleFactor = float(width) / targetWidth;
   const int W = targetWidth;
   const int H = int(height / downScaleFactor);
   if (H <= 0) return 0.0f;
   // pixel (x, y) at img[x * H + y]: the column pass reads contiguous memory
   std::vector<float> img((size_t)W * H, float(W + H));
   const float radius2 = float(W) * H / (float)(px.size() * 4);
   for (size_t i = 0; i < px.size(); ++i)
   {
      const float X = std::max(0.0f, std::min(float(W - 1), px[i] / downScaleFactor));
      const float Y = std::max(0.0f, std::min(float(H - 1), py[i] / downScaleFactor));
      img[(size_t)int(X) * H + int(Y)] = 0.0f;
   }
   // computeEuclideanSqrDistanceTransform (reconstruction_common2.cpp:148-162):
   // non-zero samples become infinity, then a pass along columns and one along rows
   for (float &val : img)
      if (val != 0) val = (float)1e20;
   std::vector<float> f(std::max(W, H)), d(std::max(W, H));
   std::vector<int> v;
   std::vector<float> z;
   for (int x = 0; x < W; ++x)
   {
      for (int y = 0; y < H; ++y) f[y] = img[(size_t)x * H + y];
      envelopePass(f.data(), H, d.data(), v, z);
      for (int y = 0; y < H; ++y) img[(size_t)x * H + y] = d[y];
   }
   for (int y = 0; y < H; ++y)
   {
      for (int x = 0; x < W; ++x) f[x] = img[(size_t)x * H + y];
      envelopePass(f.data(), W, d.data(), v, z);
      for (int x = 0; x < W; ++x) img[(size_t)x * H + y] = d[x];
   }
   unsigned int sum = 0;
   for (float val : img)
      if (val < radius2) ++sum;
   const float pi_4 = 0.785398163397448;
   return float(sum) / float(W * H * pi_4);
}

// Least-squares solution of the 4 x 3 system A x = b by Householder QR
// (Golub & Van Loan, Matrix Computations, 5.3.3): reflect A's columns to upper
// triangular form, apply the same reflections to b, back-substitute.
// Returns false when A is rank deficient.
bool lstsq43(double A[4][3], double b[4], double x[3])
{
   for (int k = 0; k < 3; ++k)
   {
      double nrm = 0;
      for (int i = k; i < 4; ++i) nrm = std::hypot(nrm, A[i][k]);
      if (nrm == 0.0) return false;
      if (A[k][k] < 0) nrm = -nrm;
      for (int i = k; i < 4; ++i) A[i][k] /= nrm;
      A[k][k] += 1.0;
      for (int j = k + 1; j < 3; ++j)
      {
         double s = 0;
         for (int i = k; i < 4; ++i) s += A[i][k] * A[i][j];
         s = -s / A[k][k];
         for (int i = k; i < 4; ++i) A[i][j] += s * A[i][k];
      }
      double s = 0;
      for (int i = k; i < 4; ++i) s += A[i][k] * b[i];
      s = -s / A[k][k];
      for (int i = k; i < 4; ++i) b[i] += s * A[i][k];
      A[k][k] = -nrm; // the diagonal of R
   }
   for (int k = 2; k >= 0; --k)
   {
      double s = b[k];
      for (int j = k + 1; j < 3; ++j) s -= A[k][j] * x[j];
      x[k] = s / A[k][k];
   }
   return true;
}

// Rows of the linear triangulation system of one image measurement
// (V3D/include/Geometry/v3d_mviewutilities.h:209-258): x P[2] - P[0], y P[2] - P[1].
void triangulationRows(const double P[12], double x, double y, double A[2][3], double b[2])
{
   for (int c = 0; c < 3; ++c)
   {
      A[0][c] = x * P[8 + c] - P[c];
      A[1][c] = y * P[8 + c] - P[4 + c];
   }
   b[0] = P[3] - x * P[11];
   b[1] = P[7] - y * P[11];
}

} // namespace

extern "C" {

int v3dm_fundamental_from_pose(const double K1[9], const double K2[9], const double R[9], const double T[3],
                               double F[9])
{
   if (!K1 || !K2 || !R || !T || !F) return fail(V3DM_ERR_ARG, "fundamental_from_pose: NULL argument");
   double invK1[9], invK2[9], invK2t[9], Tx[9], E[9], tmp[9];
   inv3(K1, invK1);
   inv3(K2, invK2);
   for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) invK2t[3 * r + c] = invK2[3 * c + r];
   // crossProductMatrix (v3d_linearbase.h:856-865)
   Tx[0] = 0;
   Tx[1] = -T[2];
   Tx[2] = T[1];
   Tx[3] = T[2];
   Tx[4] = 0;
   Tx[5] = -T[0];
   Tx[6] = -T[1];
   Tx[7] = T[0];
   Tx[8] = 0;
   mulAB(Tx, R, E, 3, 3, 3);        // E = [T]x R
   mulAB(invK2t, E, tmp, 3, 3, 3);  // (K2^-T E) K1^-1, left to right as C++ evaluates it
   mulAB(tmp, invK1, F, 3, 3, 3);
   return ok();
}

int v3dm_normalize_correspondences(const double K1[9], const double K2[9], int n, v3dm_correspondence *corrs)
{
   if (!K1 || !K2 || n < 0 || (n > 0 && !corrs)) return fail(V3DM_ERR_ARG, "normalize_correspondences: bad arguments");
   double invK1[9], invK2[9];
   inv3(K1, invK1);
   inv3(K2, invK2);
   for (int i = 0; i < n; ++i)
   {
      double p[3], q[3];
      affine(invK1, corrs[i].left.x, corrs[i].left.y, p);
      affine(invK2, corrs[i].right.x, corrs[i].right.y, q);
      corrs[i].left.x = (float)p[0];
      corrs[i].left.y = (float)p[1];
      corrs[i].right.x = (float)q[0];
      corrs[i].right.y = (float)q[1];
   }
   return ok();
}

int v3dm_effective_inlier_count(int n, const v3dm_correspondence *corrs, int width, int height, int *count)
{
   if (n < 0 || (n > 0 && !corrs) || !count) return fail(V3DM_ERR_ARG, "effective_inlier_count: bad arguments");
   if (width <= 0 || height <= 0)
      return fail(V3DM_ERR_ARG, "effective_inlier_count: image size %d x %d", width, height);
   std::vector<float> lx(n), ly(n), rx(n), ry(n);
   for (int i = 0; i < n; ++i)
   {
      lx[i] = corrs[i].left.x;
      ly[i] = corrs[i].left.y;
      rx[i] = corrs[i].right.x;
      ry[i] = corrs[i].right.y;
   }
   const float coverage1 = coverage(lx, ly, width, height);
   const float coverage2 = coverage(rx, ry, width, height);
   const float cov = 0.5f * coverage1 + 0.5f * coverage2;
   *count = int(cov * float(n) + 0.5f);
   return ok();
}

int v3dm_filter_in_front(const double K1[9], const double K2[9], const double R[9], const double T[3], int n,
                         v3dm_correspondence *corrs, int *kept)
{
   if (!K1 || !K2 || !R || !T || n < 0 || (n > 0 && !corrs) || !kept)
      return fail(V3DM_ERR_ARG, "filter_in_front: bad arguments");
   // P = K [R | T] (CameraMatrix::getProjection, v3d_cameramatrix.h:167-171); view 1 canonical
   const double RT0[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
   const double RT1[12] = {R[0], R[1], R[2], T[0], R[3], R[4], R[5], T[1], R[6], R[7], R[8], T[2]};
   double P0[12], P1[12];
   mulAB(K1, RT0, P0, 3, 3, 4);
   mulAB(K2, RT1, P1, 3, 3, 4);
   int m = 0;
   for (int i = 0; i < n; ++i)
   {
      double A[4][3], b[4], X[3];
      triangulationRows(P0, corrs[i].left.x, corrs[i].left.y, A, b);
      triangulationRows(P1, corrs[i].right.x, corrs[i].right.y, A + 2, b + 2);
      if (lstsq43(A, b, X) && X[2] > 0.0) corrs[m++] = corrs[i];
   }
   *kept = m;
   return ok();
}

int v3dm_calibration_read(const char *path, int cap, double *K, double *distortion, int32_t *dims, int *nViews)
{
   if (!path || !nViews || cap < 0) return fail(V3DM_ERR_ARG, "calibration_read: bad arguments");
   std::ifstream is(path);
   if (!is) return fail(V3DM_ERR_ARG, "Could not open calibration database file %s.", path);
   int n = 0;
   if (!(is >> n) || n < 0) return fail(V3DM_ERR_ARG, "%s: no view count", path);
   *nViews = n;
   if (n > cap) return fail(V3DM_ERR_CAPACITY, "%s: %d views, capacity %d", path, n, cap);
   for (int i = 0; i < n; ++i)
   {
      // K[0][0] K[0][1] K[0][2] K[1][1] K[1][2], four distortion terms, width height
      // (SfM/reconstruction/src/reconstruction_common.cpp:17-38); the rest of K is the identity's
      double k00, k01, k02, k11, k12, dist[4];
      int w, h;
      if (!(is >> k00 >> k01 >> k02 >> k11 >> k12 >> dist[0] >> dist[1] >> dist[2] >> dist[3] >> w >> h))
         return fail(V3DM_ERR_ARG, "%s: view %d is incomplete", path, i);
      if (K)
      {
         double *k = K + 9 * (size_t)i;
         const double v[9] = {k00, k01, k02, 0, k11, k12, 0, 0, 1};
         std::copy(v, v + 9, k);
      }
      if (distortion) std::copy(dist, dist + 4, distortion + 4 * (size_t)i);
      if (dims)
      {
         dims[2 * (size_t)i] = w;
         dims[2 * (size_t)i + 1] = h;
      }
   }
   return ok();
}

int v3dm_sampson_inliers(const double F[9], double threshold, int n, const v3dm_correspondence *corrs,
                         int32_t *inliers, int *nInliers)
{
   if (!F || n < 0 || (n > 0 && (!corrs || !inliers)) || !nInliers)
      return fail(V3DM_ERR_ARG, "sampson_inliers: bad arguments");
   const double t2 = threshold * threshold;
   int m = 0;
   for (int i = 0; i < n; ++i)
   {
      const double px = corrs[i].left.x, py = corrs[i].left.y;
      const double qx = corrs[i].right.x, qy = corrs[i].right.y;
      const double Fp0 = F[0] * px + F[1] * py + F[2];
      const double Fp1 = F[3] * px + F[4] * py + F[5];
      const double Fp2 = F[6] * px + F[7] * py + F[8];
      const double Ftq0 = F[0] * qx + F[3] * qy + F[6];
      const double Ftq1 = F[1] * qx + F[4] * qy + F[7];
      const double num = qx * Fp0 + qy * Fp1 + Fp2;
      const double den = Fp0 * Fp0 + Fp1 * Fp1 + Ftq0 * Ftq0 + Ftq1 * Ftq1;
      if (num * num <= t2 * den) inliers[m++] = i;
   }
   *nInliers = m;
   return ok();
}

} // extern "C"
