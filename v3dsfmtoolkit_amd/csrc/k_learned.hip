// k_learned.hip -- learned proposal path: top-k retrieval in the 64-D
// embedding space (BF16 MFMA), exact int8 rescoring of the candidates into
// the exact top-2 keys, margin flags, and row gathers for the exhaustive
// fallback on flagged queries.
//
// Every pair that reaches the keys carries its exact integer score, so the
// finalize kernel of the exact matcher decides on true scores; the learned
// part only chooses which pairs are scored.  A flagged left (right) query is
// rescanned against every right (left) by the int8 scan, after which its own
// top-2 is the exact one (by the survivor-threshold argument of k_i8.hip).
#include <hip/hip_runtime.h>

#include "v3dm_device.h"

namespace v3dm {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kTopK = 16;       // candidates per query
constexpr int kEmbBytes = 128;  // 64 BF16 per embedding row

// Descending top-16 list: insert (x, idx) where cond holds (x > list[15]).
__device__ __forceinline__ void topk_insert(float (&ts)[kTopK], int (&ti)[kTopK], float x, int idx, bool cond)
{
#pragma unroll
   for (int i = kTopK - 1; i >= 1; --i)
   {
      const bool up = x > ts[i - 1];
      const bool here = x > ts[i];
      const float ns = up ? ts[i - 1] : (here ? x : ts[i]);
      const int ni = up ? ti[i - 1] : (here ? idx : ti[i]);
      ts[i] = cond ? ns : ts[i];
      ti[i] = cond ? ni : ti[i];
   }
   const bool top = cond && x > ts[0];
   ts[0] = top ? x : ts[0];
   ti[0] = top ? idx : ti[0];
}

// grid (ceil(nQ / 128), nChunks), 256 threads.  Wave w owns queries
// q0 = 128 blockIdx.x + 32 w ... + 31 (the B operand, one query per lane
// column) and sweeps the right rows of chunk blockIdx.y in blocks of 32 (the
// A operand).  Accumulator register v of lane l is the score of right
// r0 + 8(v/4) + 4(l/32) + v%4 against query q0 + l%32, so each lane keeps a
// private top-16 for its query over the rights its half sees; the halves
// are merged through LDS at the end.  Output: per (query, chunk) a sorted
// list of 16 (score, right index), index -1 when the chunk had fewer.
__global__ __launch_bounds__(256) void k_emb_topk(int nQ, int nR, const uint16_t *__restrict__ EQ,
                                                  const uint16_t *__restrict__ ER, int chunkRows,
                                                  float *__restrict__ partS, int *__restrict__ partI)
{
   __shared__ float sS[4][64][kTopK];
   __shared__ int sI[4][64][kTopK];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int half = lane >> 5, lq = lane & 31;
   const int q0 = blockIdx.x * 128 + wave * 32;
   const int nChunks = gridDim.y, chunk = blockIdx.y;
   const int nQpad = (nQ + 31) & ~31, nRpad = (nR + 31) & ~31;
   const int rBegin = chunk * chunkRows, rEnd = min(rBegin + chunkRows, nRpad);
   float ts[kTopK];
   int ti[kTopK];
#pragma unroll
   for (int i = 0; i < kTopK; ++i)
   {
      ts[i] = -INFINITY;
      ti[i] = -1;
   }
   if (q0 < nQpad && rBegin < rEnd)
   {
      const int8_t *qrow = reinterpret_cast<const int8_t *>(EQ) + (size_t)(q0 + lq) * kEmbBytes + 64 * half;
      v4i Bq[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) Bq[s] = ld16(qrow + 16 * s);
      const int8_t *rbase = reinterpret_cast<const int8_t *>(ER) + (size_t)lq * kEmbBytes + 64 * half;
      v4i A[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) A[s] = ld16(rbase + (size_t)rBegin * kEmbBytes + 16 * s);
      const v16f zero = {};
      for (int r0 = rBegin; r0 < rEnd; r0 += 32)
      {
         v4i An[4];
         const int rn = min(r0 + 32, rEnd - 32);
#pragma unroll
         for (int s = 0; s < 4; ++s) An[s] = ld16(rbase + (size_t)rn * kEmbBytes + 16 * s);
         v16f acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, A[0]),
                                                            __builtin_bit_cast(v8bf, Bq[0]), zero, 0, 0, 0);
#pragma unroll
         for (int s = 1; s < 4; ++s)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, A[s]),
                                                          __builtin_bit_cast(v8bf, Bq[s]), acc, 0, 0, 0);
         const bool full = r0 + 32 <= nR;
         float m = -INFINITY;
#pragma unroll
         for (int v = 0; v < 16; ++v)
         {
            const int r = r0 + 8 * (v >> 2) + 4 * half + (v & 3);
            if (full || r < nR) m = fmaxf(m, acc[v]);
         }
         if (__any(m > ts[kTopK - 1]))
         {
#pragma unroll
            for (int v = 0; v < 16; ++v)
            {
               const int r = r0 + 8 * (v >> 2) + 4 * half + (v & 3);
               const bool cond = (full || r < nR) && acc[v] > ts[kTopK - 1];
               if (__any(cond)) topk_insert(ts, ti, acc[v], r, cond);
            }
         }
#pragma unroll
         for (int s = 0; s < 4; ++s) A[s] = An[s];
      }
   }
   // merge the two halves' lists of each query
#pragma unroll
   for (int i = 0; i < kTopK; ++i)
   {
      sS[wave][lane][i] = ts[i];
      sI[wave][lane][i] = ti[i];
   }
   __syncthreads();
   if (half == 0 && q0 + lq < nQ)
   {
      const float *as = sS[wave][lq], *bs = sS[wave][lq + 32];
      const int *ai = sI[wave][lq], *bi = sI[wave][lq + 32];
      const size_t base = ((size_t)(q0 + lq) * nChunks + chunk) * kTopK;
      int a = 0, b = 0;
      for (int o = 0; o < kTopK; ++o)
      {
         const bool takeA = as[a] >= bs[b]; // a, b < 16: each list holds 16 entries
         partS[base + o] = takeA ? as[a] : bs[b];
         partI[base + o] = takeA ? ai[a] : bi[b];
         a += takeA ? 1 : 0;
         b += takeA ? 0 : 1;
      }
   }
}

// One thread per query: top-16 of its nChunks partial lists.
__global__ void k_topk_merge(int nQ, int nChunks, const float *__restrict__ partS, const int *__restrict__ partI,
                             int *__restrict__ cand, float *__restrict__ candS)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   if (q >= nQ) return;
   float ts[kTopK];
   int ti[kTopK];
#pragma unroll
   for (int i = 0; i < kTopK; ++i)
   {
      ts[i] = -INFINITY;
      ti[i] = -1;
   }
   const size_t base = (size_t)q * nChunks * kTopK;
   for (int e = 0; e < nChunks * kTopK; ++e)
   {
      const int idx = partI[base + e];
      const float x = partS[base + e];
      topk_insert(ts, ti, x, idx, idx >= 0 && x > ts[kTopK - 1]);
   }
#pragma unroll
   for (int i = 0; i < kTopK; ++i)
   {
      cand[(size_t)q * kTopK + i] = ti[i];
      if (candS) candS[(size_t)q * kTopK + i] = ts[i];
   }
}

// Exact score of left k and right j from the int8 pools (bytes u ^ 0x80).
__device__ inline int32_t exact_score(const int8_t *L, const int8_t *R, int k, int j)
{
   const uint32_t *x = reinterpret_cast<const uint32_t *>(L + (size_t)k * kDim);
   const uint32_t *y = reinterpret_cast<const uint32_t *>(R + (size_t)j * kDim);
   uint32_t acc = 0;
#pragma unroll 8
   for (int w = 0; w < kDim / 4; ++w) acc = __builtin_amdgcn_udot4(x[w] ^ 0x80808080u, y[w] ^ 0x80808080u, acc, false);
   return (int32_t)acc;
}

// One thread per (query, candidate): exact score into both sides' top-2.
__global__ void k_cand_insert(int nL, int nR, int K, const int *__restrict__ cand, const int8_t *__restrict__ L,
                              const int8_t *__restrict__ R, uint64_t *top1L, uint64_t *top2L, uint64_t *top1R,
                              uint64_t *top2R, const uint32_t *epochPtr)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= (int64_t)nL * K) return;
   const int k = (int)(t / K);
   const int j = cand[t];
   if (j < 0 || j >= nR) return;
   const uint32_t epoch = *epochPtr;
   const uint32_t s = (uint32_t)exact_score(L, R, k, j);
   top2_insert(top1L + k, top2L + k, pack_key(epoch, s, (uint32_t)j));
   top2_insert(top1R + j, top2R + j, pack_key(epoch, s, (uint32_t)k));
}

// Margin flags.  Left k is flagged when best - second < margin over the pairs
// scored so far; right j also when it has exactly one scored pair (its
// second is unknown).  Flagged indices are appended to list (count in *n).
__global__ void k_flag(int n, int isRight, const uint64_t *__restrict__ top1, const uint64_t *__restrict__ top2,
                       const uint32_t *epochPtr, float margin, int *__restrict__ list, int *__restrict__ count)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const uint32_t epoch = *epochPtr;
   const uint64_t a = top1[i], b = top2[i];
   bool flag = false;
   if (key_epoch(a) == epoch)
   {
      const float best = (float)key_score(a) * (1.0f / 262144.0f);
      if (key_epoch(b) == epoch) flag = best - (float)key_score(b) * (1.0f / 262144.0f) < margin;
      else flag = isRight != 0;
   }
   if (flag) list[atomicAdd(count, 1)] = i;
}

// Gathers rows idx[0..n) of an int8 pool and its row sums into a compact
// pool of nPad rows (rows >= n zero).
__global__ void k_gather_rows(int n, int nPad, const int *__restrict__ idx, const int8_t *__restrict__ src,
                              const int32_t *__restrict__ srcSum, int8_t *__restrict__ dst, int32_t *__restrict__ dstSum)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= (int64_t)nPad * (kDim / 16)) return;
   const int row = (int)(t / (kDim / 16)), part = (int)(t % (kDim / 16));
   uint4 v = make_uint4(0, 0, 0, 0);
   int32_t s = 0;
   if (row < n)
   {
      const int r = idx[row];
      v = *reinterpret_cast<const uint4 *>(src + (size_t)r * kDim + part * 16);
      s = srcSum[r];
   }
   *reinterpret_cast<uint4 *>(dst + (size_t)row * kDim + part * 16) = v;
   if (part == 0) dstSum[row] = s;
}

// ---- launchers --------------------------------------------------------------

int emb_topk_chunks(int nR)
{
   // enough chunks for about 2048 waves on 16384 x 16384, chunks of >= 1024 rows
   int c = (nR + 4095) / 4096;
   return c < 1 ? 1 : c;
}

hipError_t launch_emb_topk(int nQ, int nR, const uint16_t *EQ, const uint16_t *ER, int nChunks, float *partS,
                           int *partI, int *cand, float *candS, hipStream_t st)
{
   if (nQ <= 0) return hipSuccess;
   const int nRpad = (nR + 31) & ~31;
   const int chunkRows = ((nRpad + nChunks - 1) / nChunks + 31) & ~31;
   dim3 grid((nQ + 127) / 128, nChunks);
   hipLaunchKernelGGL(k_emb_topk, grid, dim3(256), 0, st, nQ, nR, EQ, ER, chunkRows, partS, partI);
   hipLaunchKernelGGL(k_topk_merge, dim3((nQ + 255) / 256), dim3(256), 0, st, nQ, nChunks, partS, partI, cand, candS);
   return hipGetLastError();
}

hipError_t launch_cand_insert(int nL, int nR, int K, const int *cand, const int8_t *L, const int8_t *R,
                              uint64_t *top1L, uint64_t *top2L, uint64_t *top1R, uint64_t *top2R,
                              const uint32_t *epoch, hipStream_t st)
{
   const int64_t n = (int64_t)nL * K;
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_cand_insert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nL, nR, K, cand, L, R, top1L,
                      top2L, top1R, top2R, epoch);
   return hipGetLastError();
}

hipError_t launch_flag(int n, int isRight, const uint64_t *top1, const uint64_t *top2, const uint32_t *epoch,
                       float margin, int *list, int *count, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_flag, dim3((n + 255) / 256), dim3(256), 0, st, n, isRight, top1, top2, epoch, margin, list,
                      count);
   return hipGetLastError();
}

hipError_t launch_gather_rows(int n, int nPad, const int *idx, const int8_t *src, const int32_t *srcSum, int8_t *dst,
                              int32_t *dstSum, hipStream_t st)
{
   const int64_t t = (int64_t)nPad * (kDim / 16);
   if (t <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, st, n, nPad, idx, src, srcSum,
                      dst, dstSum);
   return hipGetLastError();
}

} // namespace v3dm
