/*
 * v3dm_oracle.c -- CPU oracle for SIFT putative matching.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product library never links or calls it.
 *
 * It restates, in plain C, the decision rule of the reference matcher
 * V3D_CUDA::SIFT_Matching (dddExperiments/V3DSfMToolkit):
 *
 *   scores      S[r][l] = <L_l, R_r>         V3D/src/CUDA/v3d_cudamatching.cpp:104-105
 *                                            (cublasSgemm('T','N'), read row-major
 *                                             with nRows = nR, nCols = nL,
 *                                             V3D/src/CUDA/v3d_cudamatching.cu:15)
 *   column pass per left l over rows r      v3d_cudamatching.cu:17-33  (no ratio)
 *                                           v3d_cudamatching.cu:117-146 (ratio test)
 *   row pass per right r over columns l     v3d_cudamatching.cu:35-69  (no ratio)
 *                                           v3d_cudamatching.cu:148-194 (ratio test)
 *   mutual check, ascending left index      v3d_cudamatching.cpp:121-126
 *   epipolar mask (-1e20f)                  v3d_cudamatching.cu:259-293, 407-418
 *   homography mask (-1e20f)                v3d_cudamatching.cu:295-325, 420-431
 *
 * Score arithmetic:
 *   u8 descriptors: the reference receives floats u/512; every product and
 *   partial sum is an integer multiple of 2^-18 below 2^6, so the FP32 SGEMM
 *   result equals (float)(sum u_a*u_b) * 2^-18 exactly for any summation
 *   order.  The oracle therefore computes exact int32 dot products.
 *   f32 descriptors: cuBLAS' summation order is unspecified; the oracle pins
 *   its own order: acc = fmaf(a[k], b[k], acc) for k = 0..127 (the same order
 *   the product's plain FP32 kernel uses).
 *   f64: plain double dot products, the accuracy reference for float inputs.
 *
 * Guided masks use an explicit, uncontracted left-to-right evaluation of the
 * reference expressions (this file is built with -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_D 128
#define ORC_INIT (-1e30f)   /* v3d_cudamatching.cu:22-23,122-124 (double literal -> float) */
#define ORC_MASKED (-1e20f) /* v3d_cudamatching.cu:290,322 */
#define ORC_SCALE (1.0f / 262144.0f) /* 2^-18 = (1/512)^2 */

/* ------------------------------------------------------------------------ */
/* scores                                                                    */
/* ------------------------------------------------------------------------ */

static int32_t dot_u8(const uint8_t *a, const uint8_t *b)
{
   int32_t s = 0;
   for (int k = 0; k < ORC_D; ++k) s += (int32_t)a[k] * (int32_t)b[k];
   return s;
}

static float dot_f32_seq(const float *a, const float *b)
{
   float acc = 0.0f;
   for (int k = 0; k < ORC_D; ++k) acc = fmaf(a[k], b[k], acc);
   return acc;
}

static double dot_f64(const float *a, const float *b)
{
   double acc = 0.0;
   for (int k = 0; k < ORC_D; ++k) acc += (double)a[k] * (double)b[k];
   return acc;
}

/* Exact integer score matrix, reference layout S[r*nL + l]. */
void orc_scores_u8(int nL, const uint8_t *L, int nR, const uint8_t *R, int32_t *S)
{
#pragma omp parallel for schedule(static)
   for (int r = 0; r < nR; ++r)
      for (int l = 0; l < nL; ++l)
         S[(size_t)r * nL + l] = dot_u8(L + (size_t)l * ORC_D, R + (size_t)r * ORC_D);
}

/* FP32 score matrix of u8 inputs as the reference sees them (u/512). */
void orc_scores_u8_as_f32(int nL, const uint8_t *L, int nR, const uint8_t *R, float *S)
{
#pragma omp parallel for schedule(static)
   for (int r = 0; r < nR; ++r)
      for (int l = 0; l < nL; ++l)
         S[(size_t)r * nL + l] = (float)dot_u8(L + (size_t)l * ORC_D, R + (size_t)r * ORC_D) * ORC_SCALE;
}

/* One dot product at a time (dot_f32_seq / dot_f64): the plain statement of
 * the score order, which the blocked functions below must equal bit for bit
 * (tests/test_oracle.py compares them). */
void orc_scores_f32_simple(int nL, const float *L, int nR, const float *R, float *S)
{
#pragma omp parallel for schedule(static)
   for (int r = 0; r < nR; ++r)
      for (int l = 0; l < nL; ++l)
         S[(size_t)r * nL + l] = dot_f32_seq(L + (size_t)l * ORC_D, R + (size_t)r * ORC_D);
}

void orc_scores_f64_simple(int nL, const float *L, int nR, const float *R, double *S)
{
#pragma omp parallel for schedule(static)
   for (int r = 0; r < nR; ++r)
      for (int l = 0; l < nL; ++l)
         S[(size_t)r * nL + l] = dot_f64(L + (size_t)l * ORC_D, R + (size_t)r * ORC_D);
}

/* The same scores in blocks of 4 lefts x 8 rights: every one of the 32 dot
 * products keeps its own sequential order (acc = fmaf(a[k], b[k], acc),
 * k = 0..127, from +0), the block only runs the independent chains side by
 * side so they vectorise across the 8 rights. */
#define ORC_BR 8
#define ORC_BL 4
void orc_scores_f32(int nL, const float *L, int nR, const float *R, float *S)
{
#pragma omp parallel
   {
      float Rt[ORC_D][ORC_BR];
#pragma omp for schedule(dynamic, 4)
      for (int r0 = 0; r0 < nR; r0 += ORC_BR)
      {
         const int nr = nR - r0 < ORC_BR ? nR - r0 : ORC_BR;
         for (int k = 0; k < ORC_D; ++k)
            for (int j = 0; j < ORC_BR; ++j) Rt[k][j] = j < nr ? R[(size_t)(r0 + j) * ORC_D + k] : 0.0f;
         int l = 0;
         for (; l + ORC_BL <= nL; l += ORC_BL)
         {
            float acc[ORC_BL][ORC_BR];
            for (int i = 0; i < ORC_BL; ++i)
               for (int j = 0; j < ORC_BR; ++j) acc[i][j] = 0.0f;
            for (int k = 0; k < ORC_D; ++k)
               for (int i = 0; i < ORC_BL; ++i)
               {
                  const float a = L[(size_t)(l + i) * ORC_D + k];
                  for (int j = 0; j < ORC_BR; ++j) acc[i][j] = fmaf(a, Rt[k][j], acc[i][j]);
               }
            for (int j = 0; j < nr; ++j)
               for (int i = 0; i < ORC_BL; ++i) S[(size_t)(r0 + j) * nL + l + i] = acc[i][j];
         }
         for (; l < nL; ++l)
            for (int j = 0; j < nr; ++j)
               S[(size_t)(r0 + j) * nL + l] = dot_f32_seq(L + (size_t)l * ORC_D, R + (size_t)(r0 + j) * ORC_D);
      }
   }
}

/* FP64 scores, blocked the same way (acc += (double)a[k] * (double)b[k]). */
void orc_scores_f64(int nL, const float *L, int nR, const float *R, double *S)
{
#pragma omp parallel
   {
      double Rt[ORC_D][ORC_BR];
#pragma omp for schedule(dynamic, 4)
      for (int r0 = 0; r0 < nR; r0 += ORC_BR)
      {
         const int nr = nR - r0 < ORC_BR ? nR - r0 : ORC_BR;
         for (int k = 0; k < ORC_D; ++k)
            for (int j = 0; j < ORC_BR; ++j) Rt[k][j] = j < nr ? (double)R[(size_t)(r0 + j) * ORC_D + k] : 0.0;
         int l = 0;
         for (; l + ORC_BL <= nL; l += ORC_BL)
         {
            double acc[ORC_BL][ORC_BR];
            for (int i = 0; i < ORC_BL; ++i)
               for (int j = 0; j < ORC_BR; ++j) acc[i][j] = 0.0;
            for (int k = 0; k < ORC_D; ++k)
               for (int i = 0; i < ORC_BL; ++i)
               {
                  const double a = (double)L[(size_t)(l + i) * ORC_D + k];
                  for (int j = 0; j < ORC_BR; ++j) acc[i][j] += a * Rt[k][j];
               }
            for (int j = 0; j < nr; ++j)
               for (int i = 0; i < ORC_BL; ++i) S[(size_t)(r0 + j) * nL + l + i] = acc[i][j];
         }
         for (; l < nL; ++l)
            for (int j = 0; j < nr; ++j)
               S[(size_t)(r0 + j) * nL + l] = dot_f64(L + (size_t)l * ORC_D, R + (size_t)(r0 + j) * ORC_D);
      }
   }
}

/* ------------------------------------------------------------------------ */
/* guided masks                                                              */
/* ------------------------------------------------------------------------ */

/* Epipolar Sampson test, v3d_cudamatching.cu:266-290.  F is row-major 3x3
 * (F1 = row 0, ...), p = left position, q = right position.  Returns 1 when
 * the score is replaced by -1e20f. */
int orc_epipolar_masked(const float F[9], float sqrThr, float px, float py, float qx, float qy)
{
   float Fpx = F[0] * px + F[1] * py + F[2];
   float Fpy = F[3] * px + F[4] * py + F[5];
   float Fpz = F[6] * px + F[7] * py + F[8];
   float Ftqx = F[0] * qx + F[3] * qy + F[6];
   float Ftqy = F[1] * qx + F[4] * qy + F[7];
   float num = qx * Fpx + qy * Fpy + Fpz;
   num *= num;
   float denom = Fpx * Fpx + Fpy * Fpy + Ftqx * Ftqx + Ftqy * Ftqy;
   float sampson = num / denom;
   return sampson > sqrThr;
}

/* Homography transfer test, v3d_cudamatching.cu:306-322. */
int orc_homography_masked(const float H[9], float sqrThr, float px, float py, float qx, float qy)
{
   float Hpx = H[0] * px + H[1] * py + H[2];
   float Hpy = H[3] * px + H[4] * py + H[5];
   float Hpz = H[6] * px + H[7] * py + H[8];
   Hpx /= Hpz;
   Hpy /= Hpz;
   float sqrDist = (Hpx - qx) * (Hpx - qx) + (Hpy - qy) * (Hpy - qy);
   return sqrDist > sqrThr;
}

/* Apply a list of masks in place (reference: each runGuided* call masks the
 * persistent score buffer, so consecutive calls accumulate).
 * kinds[m]: 1 = epipolar, 2 = homography; mats: 9 floats per mask (the
 * double matrix converted to float, v3d_cudamatching.cpp:161-164);
 * dist: the distance threshold per mask (squared in float, .cu:414).
 * lpos/rpos: interleaved (x,y) per feature. */
void orc_apply_masks_f32(int nL, int nR, float *S, int nMasks, const int *kinds, const float *mats,
                         const float *dist, const float *lpos, const float *rpos)
{
   for (int m = 0; m < nMasks; ++m)
   {
      const float *M = mats + 9 * m;
      float sqrThr = dist[m] * dist[m];
#pragma omp parallel for schedule(static)
      for (int r = 0; r < nR; ++r)
         for (int l = 0; l < nL; ++l)
         {
            int masked = kinds[m] == 1
                            ? orc_epipolar_masked(M, sqrThr, lpos[2 * l], lpos[2 * l + 1], rpos[2 * r], rpos[2 * r + 1])
                            : orc_homography_masked(M, sqrThr, lpos[2 * l], lpos[2 * l + 1], rpos[2 * r], rpos[2 * r + 1]);
            if (masked) S[(size_t)r * nL + l] = ORC_MASKED;
         }
   }
}

/* ------------------------------------------------------------------------ */
/* selection (literal transcriptions of the reference kernels)               */
/* ------------------------------------------------------------------------ */

/* Decision of one scan state, v3d_cudamatching.cu:32, 138-145, 187-193.
 * max(0.001, x) with a double literal equals fmaxf(0.001f, x) for float x. */
static int orc_accept(float maxVal, float nextVal, float threshold, float minRatio)
{
   if (minRatio < 0) return maxVal > threshold;
   float nextDist = fmaxf(0.001f, 2.0f - 2.0f * nextVal);
   float bestDist = 2.0f - 2.0f * maxVal;
   float ratio = bestDist / nextDist;
   return maxVal > threshold && ratio < minRatio;
}

/* Column pass: for every left l (a column of S), scan rows r = 0..nR-1.
 * v3d_cudamatching.cu:117-146 (and :17-33 when minRatio < 0: same argMax and
 * maxVal, nextVal unused).  Optional outputs best/next may be NULL. */
void orc_column_pass(int nR, int nL, float threshold, float minRatio, const float *S, int32_t *argMaxOut,
                     float *bestOut, float *nextOut)
{
   /* 64 adjacent columns at a time (each still scans its rows in increasing
    * order), so every row contributes one contiguous 256-byte read */
#pragma omp parallel for schedule(dynamic, 1)
   for (int col0 = 0; col0 < nL; col0 += 64)
   {
      const int nc = nL - col0 < 64 ? nL - col0 : 64;
      float maxVal[64], nextVal[64];
      int argMax[64];
      for (int c = 0; c < 64; ++c)
      {
         maxVal[c] = ORC_INIT;
         nextVal[c] = ORC_INIT;
         argMax[c] = -1;
      }
      for (int row = 0; row < nR; ++row)
      {
         const float *Srow = S + (size_t)row * nL + col0;
         for (int c = 0; c < nc; ++c)
         {
            float val = Srow[c];
            argMax[c] = (val > maxVal[c]) ? row : argMax[c];
            nextVal[c] = (val > maxVal[c]) ? maxVal[c] : fmaxf(val, nextVal[c]);
            maxVal[c] = fmaxf(maxVal[c], val);
         }
      }
      for (int c = 0; c < nc; ++c)
      {
         argMaxOut[col0 + c] = orc_accept(maxVal[c], nextVal[c], threshold, minRatio) ? argMax[c] : -1;
         if (bestOut) bestOut[col0 + c] = maxVal[c];
         if (nextOut) nextOut[col0 + c] = nextVal[c];
      }
   }
}

/* Row pass: for every right r (a row of S), scan columns l = 0..nL-1 in
 * increasing order (the tile loop of v3d_cudamatching.cu:161-181 visits
 * col0+k in increasing order).  v3d_cudamatching.cu:148-194 / :35-69. */
void orc_row_pass(int nR, int nL, float threshold, float minRatio, const float *S, int32_t *argMaxOut,
                  float *bestOut, float *nextOut)
{
#pragma omp parallel for schedule(static)
   for (int row = 0; row < nR; ++row)
   {
      float maxVal = ORC_INIT, nextVal = ORC_INIT;
      int argMax = -1;
      const float *Srow = S + (size_t)row * nL;
      for (int col = 0; col < nL; ++col)
      {
         float val = Srow[col];
         argMax = (val > maxVal) ? col : argMax;
         nextVal = (val > maxVal) ? maxVal : fmaxf(val, nextVal);
         maxVal = fmaxf(maxVal, val);
      }
      argMaxOut[row] = orc_accept(maxVal, nextVal, threshold, minRatio) ? argMax : -1;
      if (bestOut) bestOut[row] = maxVal;
      if (nextOut) nextOut[row] = nextVal;
   }
}

/* Mutual check, v3d_cudamatching.cpp:121-126.  pairs: 2 ints per match
 * (left k, right l), ascending k.  Returns the match count. */
int orc_mutual(int nL, const int32_t *colMax, const int32_t *rowMax, int32_t *pairs)
{
   int n = 0;
   for (int k = 0; k < nL; ++k)
   {
      int l = colMax[k];
      if (l >= 0 && rowMax[l] == k)
      {
         pairs[2 * n] = k;
         pairs[2 * n + 1] = l;
         ++n;
      }
   }
   return n;
}

/* Same, but as a per-left table: out[k] = l for a mutual match, else -1. */
void orc_mutual_table(int nL, const int32_t *colMax, const int32_t *rowMax, int32_t *out)
{
   for (int k = 0; k < nL; ++k)
   {
      int l = colMax[k];
      out[k] = (l >= 0 && rowMax[l] == k) ? l : -1;
   }
}

/* ------------------------------------------------------------------------ */
/* full rule on an explicit score matrix (the reference design)             */
/* ------------------------------------------------------------------------ */

/* findPutativeMatches on a given FP32 score matrix: column + row passes and
 * the mutual check; fills colMax/rowMax and the per-left match table. */
void orc_select_f32(int nL, int nR, const float *S, float minScore, float minRatio, int32_t *colMax,
                    int32_t *rowMax, int32_t *matchOfLeft)
{
   orc_column_pass(nR, nL, minScore, minRatio, S, colMax, NULL, NULL);
   orc_row_pass(nR, nL, minScore, minRatio, S, rowMax, NULL, NULL);
   orc_mutual_table(nL, colMax, rowMax, matchOfLeft);
}

/* u8 descriptors (reference float input u/512), optional guided masks. */
int orc_find_u8(int nL, const uint8_t *L, int nR, const uint8_t *R, float minScore, float minRatio, int nMasks,
                const int *kinds, const float *mats, const float *dist, const float *lpos, const float *rpos,
                int32_t *colMax, int32_t *rowMax, int32_t *matchOfLeft)
{
   float *S = (float *)malloc(sizeof(float) * (size_t)nL * (size_t)nR + 4);
   if (!S) return -1;
   orc_scores_u8_as_f32(nL, L, nR, R, S);
   if (nMasks > 0) orc_apply_masks_f32(nL, nR, S, nMasks, kinds, mats, dist, lpos, rpos);
   orc_select_f32(nL, nR, S, minScore, minRatio, colMax, rowMax, matchOfLeft);
   free(S);
   return 0;
}

/* f32 descriptors with the pinned sequential-fmaf score order. */
int orc_find_f32(int nL, const float *L, int nR, const float *R, float minScore, float minRatio, int nMasks,
                 const int *kinds, const float *mats, const float *dist, const float *lpos, const float *rpos,
                 int32_t *colMax, int32_t *rowMax, int32_t *matchOfLeft)
{
   float *S = (float *)malloc(sizeof(float) * (size_t)nL * (size_t)nR + 4);
   if (!S) return -1;
   orc_scores_f32(nL, L, nR, R, S);
   if (nMasks > 0) orc_apply_masks_f32(nL, nR, S, nMasks, kinds, mats, dist, lpos, rpos);
   orc_select_f32(nL, nR, S, minScore, minRatio, colMax, rowMax, matchOfLeft);
   free(S);
   return 0;
}

/* FP64 rule: double scores, double ratio, same scan semantics.  This is the
 * accuracy reference for float descriptors (no reference counterpart: it
 * states what exact arithmetic would decide). */
static int orc_accept_f64(double maxVal, double nextVal, double threshold, double minRatio)
{
   if (minRatio < 0) return maxVal > threshold;
   double nextDist = fmax(0.001, 2.0 - 2.0 * nextVal);
   double bestDist = 2.0 - 2.0 * maxVal;
   return maxVal > threshold && bestDist / nextDist < minRatio;
}

/* Column + row passes and the mutual check on an FP64 score matrix S[r*nL + l]. */
void orc_select_f64(int nL, int nR, const double *S, double minScore, double minRatio, int32_t *colMax,
                    int32_t *rowMax, int32_t *matchOfLeft, double *bestL, double *nextL)
{
#pragma omp parallel for schedule(dynamic, 1)
   for (int col0 = 0; col0 < nL; col0 += 64)
   {
      const int nc = nL - col0 < 64 ? nL - col0 : 64;
      double maxVal[64], nextVal[64];
      int argMax[64];
      for (int c = 0; c < 64; ++c)
      {
         maxVal[c] = -1e30;
         nextVal[c] = -1e30;
         argMax[c] = -1;
      }
      for (int row = 0; row < nR; ++row)
      {
         const double *Srow = S + (size_t)row * nL + col0;
         for (int c = 0; c < nc; ++c)
         {
            double val = Srow[c];
            argMax[c] = (val > maxVal[c]) ? row : argMax[c];
            nextVal[c] = (val > maxVal[c]) ? maxVal[c] : fmax(val, nextVal[c]);
            maxVal[c] = fmax(maxVal[c], val);
         }
      }
      for (int c = 0; c < nc; ++c)
      {
         colMax[col0 + c] = orc_accept_f64(maxVal[c], nextVal[c], minScore, minRatio) ? argMax[c] : -1;
         if (bestL) bestL[col0 + c] = maxVal[c];
         if (nextL) nextL[col0 + c] = nextVal[c];
      }
   }
#pragma omp parallel for schedule(static)
   for (int row = 0; row < nR; ++row)
   {
      double maxVal = -1e30, nextVal = -1e30;
      int argMax = -1;
      for (int col = 0; col < nL; ++col)
      {
         double val = S[(size_t)row * nL + col];
         argMax = (val > maxVal) ? col : argMax;
         nextVal = (val > maxVal) ? maxVal : fmax(val, nextVal);
         maxVal = fmax(maxVal, val);
      }
      rowMax[row] = orc_accept_f64(maxVal, nextVal, minScore, minRatio) ? argMax : -1;
   }
   orc_mutual_table(nL, colMax, rowMax, matchOfLeft);
}

int orc_find_f64(int nL, const float *L, int nR, const float *R, double minScore, double minRatio,
                 int32_t *colMax, int32_t *rowMax, int32_t *matchOfLeft, double *bestL, double *nextL)
{
   double *S = (double *)malloc(sizeof(double) * (size_t)nL * (size_t)nR + 8);
   if (!S) return -1;
   orc_scores_f64(nL, L, nR, R, S);
   orc_select_f64(nL, nR, S, minScore, minRatio, colMax, rowMax, matchOfLeft, bestL, nextL);
   free(S);
   return 0;
}

/* ------------------------------------------------------------------------ */
/* blocked multithreaded variant (the CPU baseline)                          */
/* ------------------------------------------------------------------------ */

/* Same rule as orc_find_u8 without masks, computed in left blocks so the
 * nL x nR matrix is never materialised.  Column decisions are complete per
 * block; row states are carried per thread over a contiguous, increasing
 * range of left indices and merged in thread order, which reproduces the
 * sequential scan exactly (merge(a, b) for a before b:
 *   max = b.max > a.max ? b : a (first index kept on ties);
 *   next = b.max > a.max ? max(a.max, b.next) : max(a.next, b.max)).
 * Dot products are exact int32, compared as floats (int * 2^-18). */
typedef struct
{
   float maxVal, nextVal;
   int argMax;
} orc_state;

int orc_find_u8_blocked(int nL, const uint8_t *L, int nR, const uint8_t *R, float minScore, float minRatio,
                        int nThreads, int32_t *colMax, int32_t *rowMax, int32_t *matchOfLeft)
{
#ifdef _OPENMP
   if (nThreads <= 0) nThreads = omp_get_max_threads();
#else
   nThreads = 1;
#endif
   orc_state *part = (orc_state *)malloc(sizeof(orc_state) * (size_t)nThreads * (size_t)(nR > 0 ? nR : 1));
   int16_t *R16 = (int16_t *)malloc(sizeof(int16_t) * (size_t)nR * ORC_D + 2);
   if (!part || !R16)
   {
      free(part);
      free(R16);
      return -1;
   }
   for (size_t i = 0; i < (size_t)nR * ORC_D; ++i) R16[i] = R[i];
   int usedThreads = nThreads;

#pragma omp parallel num_threads(nThreads)
   {
      int t = 0, nt = 1;
#ifdef _OPENMP
      t = omp_get_thread_num();
      nt = omp_get_num_threads();
#endif
#pragma omp single
      usedThreads = nt;
      orc_state *ps = part + (size_t)t * nR;
      for (int r = 0; r < nR; ++r)
      {
         ps[r].maxVal = ORC_INIT;
         ps[r].nextVal = ORC_INIT;
         ps[r].argMax = -1;
      }
      int chunk = (nL + nt - 1) / nt;
      int l0 = t * chunk, l1 = l0 + chunk < nL ? l0 + chunk : nL;
      int16_t a16[ORC_D];
      for (int l = l0; l < l1; ++l)
      {
         for (int k = 0; k < ORC_D; ++k) a16[k] = L[(size_t)l * ORC_D + k];
         float maxVal = ORC_INIT, nextVal = ORC_INIT;
         int argMax = -1;
         for (int r = 0; r < nR; ++r)
         {
            const int16_t *b = R16 + (size_t)r * ORC_D;
            int32_t s = 0;
            for (int k = 0; k < ORC_D; ++k) s += (int32_t)a16[k] * (int32_t)b[k];
            float val = (float)s * ORC_SCALE;
            /* column scan of left l over r */
            argMax = (val > maxVal) ? r : argMax;
            nextVal = (val > maxVal) ? maxVal : fmaxf(val, nextVal);
            maxVal = fmaxf(maxVal, val);
            /* row scan of right r over l (this thread's l range, increasing) */
            orc_state *st = ps + r;
            st->argMax = (val > st->maxVal) ? l : st->argMax;
            st->nextVal = (val > st->maxVal) ? st->maxVal : fmaxf(val, st->nextVal);
            st->maxVal = fmaxf(st->maxVal, val);
         }
         colMax[l] = orc_accept(maxVal, nextVal, minScore, minRatio) ? argMax : -1;
      }
   }

#pragma omp parallel for schedule(static)
   for (int r = 0; r < nR; ++r)
   {
      orc_state acc = part[r];
      for (int t = 1; t < usedThreads; ++t)
      {
         orc_state b = part[(size_t)t * nR + r];
         if (b.maxVal > acc.maxVal)
         {
            acc.nextVal = fmaxf(acc.maxVal, b.nextVal);
            acc.maxVal = b.maxVal;
            acc.argMax = b.argMax;
         }
         else
            acc.nextVal = fmaxf(acc.nextVal, b.maxVal);
      }
      rowMax[r] = orc_accept(acc.maxVal, acc.nextVal, minScore, minRatio) ? acc.argMax : -1;
   }
   orc_mutual_table(nL, colMax, rowMax, matchOfLeft);
   free(part);
   free(R16);
   return usedThreads;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
   return omp_get_max_threads();
#else
   return 1;
#endif
}
