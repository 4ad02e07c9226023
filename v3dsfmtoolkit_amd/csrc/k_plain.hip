// k_plain.hip -- the reference design on gfx950: an FP32 score matrix in HBM
// followed by column and row selection passes and the mutual check.
//
//   score matrix   V3D/src/CUDA/v3d_cudamatching.cpp:104-105 (cublasSgemm into
//                  S, read as row-major S[r][l], nRows = nR, nCols = nL)
//   masks          v3d_cudamatching.cu:259-325 (in place, -1e20f)
//   column pass    v3d_cudamatching.cu:117-146 (/ :17-33 without ratio test)
//   row pass       v3d_cudamatching.cu:148-194 (/ :35-69)
//   mutual check   v3d_cudamatching.cpp:121-126
//
// Every score is acc = fmaf(l[k], r[k], acc) for k = 0..127 in order, so the
// matrix is reproducible bit for bit on the CPU; for inputs u/512 it equals
// the exact integer score times 2^-18 whatever the order.  Selection states
// (max, next, arg) are merged with the first-index rule, which equals the
// reference's sequential scan for any split of the index range.
#include <hip/hip_ext.h>

#include "v3dm_internal.h"

namespace v3dm {

constexpr int kGemmT = 64;  // 64 x 64 output tile per 256-thread workgroup
constexpr int kLdsStride = kGemmT + 4;

__global__ __launch_bounds__(256) void k_sgemm_seq(int nL, int nR, const float *__restrict__ Lf,
                                                   const float *__restrict__ Rf, float *__restrict__ S)
{
   constexpr int KH = kDim / 2;          // k staged in two halves, in order
   __shared__ float As[KH][kLdsStride]; // As[k][l]
   __shared__ float Bs[KH][kLdsStride]; // Bs[k][r]
   const int tid = threadIdx.x;
   const int l0 = blockIdx.x * kGemmT, r0 = blockIdx.y * kGemmT;
   const int tx = tid & 15, ty = tid >> 4; // 4 left x 4 right outputs per thread
   float acc[4][4];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
   for (int kh = 0; kh < kDim; kh += KH)
   {
      // 64 rows x 64 floats of each operand, transposed into LDS
      for (int e = tid; e < kGemmT * KH / 4; e += 256)
      {
         const int row = e / (KH / 4), k4 = (e % (KH / 4)) * 4;
         float4 va = make_float4(0.f, 0.f, 0.f, 0.f), vb = va;
         if (l0 + row < nL) va = *reinterpret_cast<const float4 *>(Lf + (size_t)(l0 + row) * kDim + kh + k4);
         if (r0 + row < nR) vb = *reinterpret_cast<const float4 *>(Rf + (size_t)(r0 + row) * kDim + kh + k4);
         As[k4 + 0][row] = va.x;
         As[k4 + 1][row] = va.y;
         As[k4 + 2][row] = va.z;
         As[k4 + 3][row] = va.w;
         Bs[k4 + 0][row] = vb.x;
         Bs[k4 + 1][row] = vb.y;
         Bs[k4 + 2][row] = vb.z;
         Bs[k4 + 3][row] = vb.w;
      }
      __syncthreads();
      for (int k = 0; k < KH; ++k)
      {
         const float4 a = *reinterpret_cast<const float4 *>(&As[k][tx * 4]);
         const float4 b = *reinterpret_cast<const float4 *>(&Bs[k][ty * 4]);
         const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
      }
      __syncthreads();
   }
#pragma unroll
   for (int j = 0; j < 4; ++j)
   {
      const int r = r0 + ty * 4 + j;
      if (r >= nR) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i)
      {
         const int l = l0 + tx * 4 + i;
         if (l < nL) S[(size_t)r * nL + l] = acc[i][j];
      }
   }
}

// Masks S in place.  Mask arithmetic as in the fused kernel (no contraction).
__global__ void k_mask_plain(int nL, int nR, float *S, const Mask *masks, int nMasks)
{
   const int l = blockIdx.x * blockDim.x + threadIdx.x;
   if (l >= nL) return;
   for (int r = blockIdx.y; r < nR; r += gridDim.y)
   {
   bool hit = false;
   for (int m = 0; m < nMasks; ++m)
   {
      const Mask &mk = masks[m];
      const float *F = mk.m;
      const float px = mk.lpos[2 * l], py = mk.lpos[2 * l + 1], qx = mk.rpos[2 * r], qy = mk.rpos[2 * r + 1];
      if (mk.kind == 1)
      {
         float Fpx = F[0] * px + F[1] * py + F[2];
         float Fpy = F[3] * px + F[4] * py + F[5];
         float Fpz = F[6] * px + F[7] * py + F[8];
         float Ftqx = F[0] * qx + F[3] * qy + F[6];
         float Ftqy = F[1] * qx + F[4] * qy + F[7];
         float num = qx * Fpx + qy * Fpy + Fpz;
         num *= num;
         float denom = Fpx * Fpx + Fpy * Fpy + Ftqx * Ftqx + Ftqy * Ftqy;
         hit = hit || (num / denom > mk.sqrThr);
      }
      else
      {
         float Hpx = F[0] * px + F[1] * py + F[2];
         float Hpy = F[3] * px + F[4] * py + F[5];
         float Hpz = F[6] * px + F[7] * py + F[8];
         Hpx /= Hpz;
         Hpy /= Hpz;
         float d2 = (Hpx - qx) * (Hpx - qx) + (Hpy - qy) * (Hpy - qy);
         hit = hit || (d2 > mk.sqrThr);
      }
   }
   if (hit) S[(size_t)r * nL + l] = -1e20f;
   }
}

__global__ void k_fill_i32(int n, int32_t value, int32_t *out)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) out[i] = value;
}

struct SelState
{
   float maxVal, nextVal;
   int argMax;
};

// sequential update, v3d_cudamatching.cu:129-131
__device__ inline void sel_push(SelState &s, float val, int idx)
{
   const bool gt = val > s.maxVal;
   s.argMax = gt ? idx : s.argMax;
   s.nextVal = gt ? s.maxVal : fmaxf(val, s.nextVal);
   s.maxVal = fmaxf(s.maxVal, val);
}

// merge of two states over disjoint index sets (first index wins ties)
__device__ inline SelState sel_merge(const SelState &a, const SelState &b)
{
   const bool bWins = b.maxVal > a.maxVal || (b.maxVal == a.maxVal && b.argMax >= 0 && (a.argMax < 0 || b.argMax < a.argMax));
   SelState r;
   if (bWins)
   {
      r.maxVal = b.maxVal;
      r.argMax = b.argMax;
      r.nextVal = fmaxf(b.nextVal, a.maxVal);
   }
   else
   {
      r.maxVal = a.maxVal;
      r.argMax = a.argMax;
      r.nextVal = fmaxf(a.nextVal, b.maxVal);
   }
   return r;
}

__device__ inline bool sel_accept(const SelState &s, float minScore, float minRatio)
{
   if (minRatio < 0) return s.maxVal > minScore;
   float nextDist = fmaxf(0.001f, 2.0f - 2.0f * s.nextVal);
   float bestDist = 2.0f - 2.0f * s.maxVal;
   float ratio = bestDist / nextDist;
   return s.maxVal > minScore && ratio < minRatio;
}

// Column pass, stage 1: a thread takes 4 adjacent left columns (nL is a
// multiple of 32, so rows are 16-byte aligned) over rows [chunk*rowsPer, ...),
// loading 8 rows before folding them in increasing row order: 128 bytes in
// flight per thread instead of one dependent 4-byte load.
constexpr int kColRows = 8;
typedef float vf4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_colpass_part(int nL, int nR, int rowsPer, const float *__restrict__ S,
                                                      SelState *part)
{
   const int l = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
   const int chunk = blockIdx.y;
   if (l >= nL) return;
   SelState s[4];
#pragma unroll
   for (int c = 0; c < 4; ++c) s[c] = SelState{-1e30f, -1e30f, -1};
   const int r0 = chunk * rowsPer, r1 = min(nR, r0 + rowsPer);
   const vf4 *col = reinterpret_cast<const vf4 *>(S + l);
   const size_t stride = (size_t)nL / 4;
   int r = r0;
   for (; r + kColRows <= r1; r += kColRows)
   {
      vf4 v[kColRows];
#pragma unroll
      for (int i = 0; i < kColRows; ++i) v[i] = __builtin_nontemporal_load(col + (size_t)(r + i) * stride);
#pragma unroll
      for (int i = 0; i < kColRows; ++i)
      {
         sel_push(s[0], v[i].x, r + i);
         sel_push(s[1], v[i].y, r + i);
         sel_push(s[2], v[i].z, r + i);
         sel_push(s[3], v[i].w, r + i);
      }
   }
   for (; r < r1; ++r)
   {
      const vf4 v = col[(size_t)r * stride];
      sel_push(s[0], v.x, r);
      sel_push(s[1], v.y, r);
      sel_push(s[2], v.z, r);
      sel_push(s[3], v.w, r);
   }
#pragma unroll
   for (int c = 0; c < 4; ++c) part[(size_t)chunk * nL + l + c] = s[c];
}

// Column pass, stage 2: merge the chunks in increasing row order.
__global__ void k_colpass_merge(int nL, int nChunks, const SelState *part, float minScore, float minRatio,
                                int32_t *colMax)
{
   const int l = blockIdx.x * blockDim.x + threadIdx.x;
   if (l >= nL) return;
   SelState s{-1e30f, -1e30f, -1};
   for (int c = 0; c < nChunks; ++c) s = sel_merge(s, part[(size_t)c * nL + l]);
   colMax[l] = sel_accept(s, minScore, minRatio) ? s.argMax : -1;
}

// Row pass: one wave per right row; lane t scans columns 4t .. 4t+3, then
// 4t+256 .. (increasing within the lane, 16-byte loads, 4 iterations of loads
// issued before they are folded), then the 64 lane states are merged with the
// first-index rule (the lanes' column sets are disjoint).
__global__ __launch_bounds__(256) void k_rowpass(int nL, int nR, const float *__restrict__ S, float minScore,
                                                 float minRatio, int32_t *rowMax)
{
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
   if (r >= nR) return;
   SelState s{-1e30f, -1e30f, -1};
   const vf4 *row = reinterpret_cast<const vf4 *>(S + (size_t)r * nL); // nL is a multiple of 32
   const int nq = nL / 4;
   int q = lane;
   for (; q + 3 * 64 < nq; q += 4 * 64)
   {
      vf4 v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = __builtin_nontemporal_load(row + q + 64 * i);
#pragma unroll
      for (int i = 0; i < 4; ++i)
      {
         const int l = 4 * (q + 64 * i);
         sel_push(s, v[i].x, l);
         sel_push(s, v[i].y, l + 1);
         sel_push(s, v[i].z, l + 2);
         sel_push(s, v[i].w, l + 3);
      }
   }
   for (; q < nq; q += 64)
   {
      const vf4 v = row[q];
      sel_push(s, v.x, 4 * q);
      sel_push(s, v.y, 4 * q + 1);
      sel_push(s, v.z, 4 * q + 2);
      sel_push(s, v.w, 4 * q + 3);
   }
#pragma unroll
   for (int off = 32; off >= 1; off >>= 1)
   {
      SelState o;
      o.maxVal = __shfl_xor(s.maxVal, off);
      o.nextVal = __shfl_xor(s.nextVal, off);
      o.argMax = __shfl_xor(s.argMax, off);
      s = sel_merge(s, o);
   }
   if (lane == 0) rowMax[r] = sel_accept(s, minScore, minRatio) ? s.argMax : -1;
}

__global__ void k_mutual(int nL, const int32_t *colMax, const int32_t *rowMax, int32_t *out)
{
   const int k = blockIdx.x * blockDim.x + threadIdx.x;
   if (k >= nL) return;
   const int l = colMax[k];
   out[k] = (l >= 0 && rowMax[l] == k) ? l : -1;
}

// ---- descriptor preparation ------------------------------------------------

// u8 rows -> int8 (u - 128) pool rows + 128*sum(u); pad rows -> u = 0.
__global__ void k_prep_u8(const uint8_t *__restrict__ src, int n, int nPad, int8_t *__restrict__ dst,
                          int32_t *__restrict__ sum)
{
   const int row = blockIdx.x * blockDim.x + threadIdx.x;
   if (row >= nPad) return;
   uint4 *d = reinterpret_cast<uint4 *>(dst + (size_t)row * kDim);
   if (row < n)
   {
      const uint4 *s = reinterpret_cast<const uint4 *>(src + (size_t)row * kDim);
      int tot = 0;
#pragma unroll
      for (int q = 0; q < kDim / 16; ++q)
      {
         uint4 w = s[q];
         const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
         for (int h = 0; h < 4; ++h)
            tot += (ws[h] & 0xFF) + ((ws[h] >> 8) & 0xFF) + ((ws[h] >> 16) & 0xFF) + (ws[h] >> 24);
         w.x ^= 0x80808080u;
         w.y ^= 0x80808080u;
         w.z ^= 0x80808080u;
         w.w ^= 0x80808080u;
         d[q] = w;
      }
      sum[row] = 128 * tot;
   }
   else
   {
#pragma unroll
      for (int q = 0; q < kDim / 16; ++q) d[q] = make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u);
      sum[row] = kPadRowSum;
   }
}

// f32 rows: when every value is u/512 with integer u in [0, 255] the rows are
// converted like k_prep_u8; otherwise *notQuantized is set.
__global__ void k_prep_f32(const float *__restrict__ src, int n, int nPad, int8_t *__restrict__ dst,
                           int32_t *__restrict__ sum, int *notQuantized)
{
   const int row = blockIdx.x * blockDim.x + threadIdx.x;
   if (row >= nPad) return;
   int8_t *d = dst + (size_t)row * kDim;
   if (row < n)
   {
      const float4 *s = reinterpret_cast<const float4 *>(src + (size_t)row * kDim);
      int tot = 0;
      bool bad = false;
      for (int q = 0; q < kDim / 4; ++q)
      {
         const float4 v = s[q];
         const float vs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
         for (int h = 0; h < 4; ++h)
         {
            const float u = vs[h] * 512.0f; // exact scaling by a power of two
            const bool ok = u >= 0.0f && u <= 255.0f && u == truncf(u);
            bad = bad || !ok;
            const int ui = ok ? (int)u : 0;
            tot += ui;
            d[4 * q + h] = (int8_t)(ui - 128);
         }
      }
      sum[row] = 128 * tot;
      if (bad) atomicOr(notQuantized, 1);
   }
   else
   {
      for (int q = 0; q < kDim; ++q) d[q] = (int8_t)-128;
      sum[row] = kPadRowSum;
   }
}

__global__ void k_u8_to_f32(const uint8_t *__restrict__ src, int64_t count, float *__restrict__ dst)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < count) dst[i] = (float)src[i] * (1.0f / 512.0f);
}

// Memory-counter calibration: copies n16 16-byte units, each read once and
// written once (grid-stride, fully coalesced).
__global__ __launch_bounds__(256) void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16)
{
   const int64_t stride = (int64_t)gridDim.x * blockDim.x;
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// ---- launchers ------------------------------------------------------------

hipError_t launch_sgemm_seq(int nL, int nR, const float *Lf, const float *Rf, float *S, hipStream_t st,
                            hipEvent_t start, hipEvent_t stop)
{
   if (nL <= 0 || nR <= 0) return hipSuccess;
   dim3 grid((nL + kGemmT - 1) / kGemmT, (nR + kGemmT - 1) / kGemmT);
   hipExtLaunchKernelGGL(k_sgemm_seq, grid, dim3(256), 0, st, start, stop, 0, nL, nR, Lf, Rf, S);
   return hipGetLastError();
}

hipError_t launch_mask_plain(int nL, int nR, float *S, const Mask *masks, int nMasks, hipStream_t st)
{
   if (nL <= 0 || nR <= 0 || nMasks <= 0) return hipSuccess;
   dim3 grid((nL + 255) / 256, nR < 4096 ? nR : 4096);
   hipLaunchKernelGGL(k_mask_plain, grid, dim3(256), 0, st, nL, nR, S, masks, nMasks);
   return hipGetLastError();
}

hipError_t launch_fill_i32(int n, int32_t value, int32_t *out, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, n, value, out);
   return hipGetLastError();
}

size_t plain_part_bytes(int nL, int nR, int *nChunksOut, int *rowsPerOut)
{
   // enough column work items to cover the device (>= ~64K threads of 4 columns)
   int nChunks = 1;
   while ((int64_t)nChunks * (nL / 4) < 65536 && nChunks * 64 < nR) nChunks *= 2;
   int rowsPer = (nR + nChunks - 1) / nChunks;
   if (rowsPer < 1) rowsPer = 1;
   nChunks = (nR + rowsPer - 1) / rowsPer;
   if (nChunks < 1) nChunks = 1;
   if (nChunksOut) *nChunksOut = nChunks;
   if (rowsPerOut) *rowsPerOut = rowsPer;
   return sizeof(SelState) * (size_t)nChunks * (size_t)(nL > 0 ? nL : 1);
}

hipError_t launch_plain_select(int nL, int nR, const float *S, void *part, float minScore, float minRatio,
                               int32_t *colMax, int32_t *rowMax, int32_t *out, hipStream_t st)
{
   if (nL <= 0) return hipSuccess;
   if (nR <= 0) return launch_fill_i32(nL, -1, out, st);
   int nChunks = 1, rowsPer = 1;
   plain_part_bytes(nL, nR, &nChunks, &rowsPer);
   SelState *ps = static_cast<SelState *>(part);
   if (nL % 4) return hipErrorInvalidValue; // feature counts are multiples of 32
   hipLaunchKernelGGL(k_colpass_part, dim3((nL / 4 + 255) / 256, nChunks), dim3(256), 0, st, nL, nR, rowsPer, S, ps);
   hipLaunchKernelGGL(k_colpass_merge, dim3((nL + 255) / 256), dim3(256), 0, st, nL, nChunks, ps, minScore,
                      minRatio, colMax);
   hipLaunchKernelGGL(k_rowpass, dim3((nR + 3) / 4), dim3(256), 0, st, nL, nR, S, minScore, minRatio, rowMax);
   hipLaunchKernelGGL(k_mutual, dim3((nL + 255) / 256), dim3(256), 0, st, nL, colMax, rowMax, out);
   return hipGetLastError();
}

hipError_t launch_prep_u8(const uint8_t *src, int n, int nPad, int8_t *dst, int32_t *sum, hipStream_t st)
{
   if (nPad <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_prep_u8, dim3((nPad + 255) / 256), dim3(256), 0, st, src, n, nPad, dst, sum);
   return hipGetLastError();
}

hipError_t launch_prep_f32(const float *src, int n, int nPad, int8_t *dst, int32_t *sum, int *flag,
                           hipStream_t st)
{
   if (nPad <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_prep_f32, dim3((nPad + 255) / 256), dim3(256), 0, st, src, n, nPad, dst, sum, flag);
   return hipGetLastError();
}

hipError_t launch_u8_to_f32(const uint8_t *src, int64_t count, float *dst, hipStream_t st)
{
   if (count <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_u8_to_f32, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, src, count, dst);
   return hipGetLastError();
}

hipError_t launch_copy_probe(const void *src, void *dst, int64_t bytes, hipStream_t st)
{
   const int64_t n16 = bytes / 16;
   if (n16 <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16);
   return hipGetLastError();
}

} // namespace v3dm
