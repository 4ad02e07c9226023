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

constexpr int kEmbBytes = 128;  // 64 BF16 per embedding row

// Descending top-K list: insert (x, idx) where cond holds (x > list[K-1]).
template <int K>
__device__ __forceinline__ void topk_insert(float (&ts)[K], int (&ti)[K], float x, int idx, bool cond)
{
#pragma unroll
   for (int i = K - 1; i >= 1; --i)
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

// Top-K retrieval in two MFMA sweeps (K = 16 or 32), no per-element sorting.
//
// Layout: grid (ceil(nQ / 128), nChunks), 256 threads.  Wave w owns queries
// q0 = 128 blockIdx.x + 32 w ... + 31 as the B operand (one query per lane
// column) and sweeps the right rows of chunk blockIdx.y in blocks of 32 as the
// A operand.  Accumulator register v of lane l scores right
// r0 + 8(v/4) + 4(l/32) + v%4 against query q0 + l%32.
//
// Groups: for a query, the rights of one chunk split into 32 groups by
// (lane half h, register v) -- the rights r with r % 32 = 8(v/4) + 4h + v%4.
// Sweep 1 keeps each group's maximum (one max per score).  If theta_q is the
// K-th largest of the query's group maxima over all chunks, at least K
// distinct rights score >= theta_q, so every member of the query's top K
// scores >= theta_q.  Sweep 2 recomputes the same scores (same instructions,
// same values) and appends every score >= theta_q to a list of kListCap
// entries per (query, chunk, half); typically a few dozen rights per query
// qualify.  The select kernel takes the top K of the lists (score
// descending, then index ascending); a query whose list overflowed is marked
// so the matcher rescans it exhaustively.
constexpr int kListCap = 32;

__device__ inline v16f emb_block(const v4i (&A)[4], const v4i (&Bq)[4])
{
   const v16f zero = {};
   v16f acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, A[0]), __builtin_bit_cast(v8bf, Bq[0]),
                                                      zero, 0, 0, 0);
#pragma unroll
   for (int s = 1; s < 4; ++s)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, A[s]), __builtin_bit_cast(v8bf, Bq[s]),
                                                    acc, 0, 0, 0);
   return acc;
}

// sweep 1 (collect == 0): group maxima to gmax[((q * nChunks + chunk) * 2 + h) * 16 + v];
// sweep 2 (collect == 1): scores >= theta[q] appended to the (q, chunk, h) list.
// A workgroup of 4 waves owns 256 queries (64 per wave: two 32-query B
// blocks) and stages the chunk's right rows through a 2 x 32 KiB LDS ring,
// 256 rows per stage, so every staged row feeds 8 MFMAs per wave.  16-byte
// unit p of a row is stored at unit p ^ (row & 7) (conflict-free fragment reads).
constexpr int kSweepStageRows = 256;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_emb_sweep(int collect, int nQ, int nR, const uint16_t *__restrict__ EQ,
                                                   const uint16_t *__restrict__ ER, int chunkRows,
                                                   float *__restrict__ gmax, const float *__restrict__ theta,
                                                   float *__restrict__ listS, int *__restrict__ listI,
                                                   int *__restrict__ listN)
{
   __shared__ __attribute__((aligned(16))) int8_t sR[2][kSweepStageRows * kEmbBytes];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int half = lane >> 5, lq = lane & 31;
   const int nChunks = gridDim.y, chunk = blockIdx.y;
   const int nQpad = (nQ + 31) & ~31, nRpad = (nR + 31) & ~31;
   const int rBegin = chunk * chunkRows, rEnd = min(rBegin + chunkRows, nRpad);
   const int nStages = rBegin < rEnd ? (rEnd - rBegin + kSweepStageRows - 1) / kSweepStageRows : 0;
   const int qw = blockIdx.x * 256 + wave * 64; // this wave's first query
   const bool active = qw < nQpad;               // wave-uniform; inactive waves still stage and sync
   v4i Bq[2][4];
   float th[2];
#pragma unroll
   for (int b = 0; b < 2; ++b)
   {
      const int q = qw + 32 * b + lq;
      const bool ok = active && q < nQpad;
      const int8_t *qrow = reinterpret_cast<const int8_t *>(EQ) + (size_t)(ok ? q : 0) * kEmbBytes + 64 * half;
#pragma unroll
      for (int s = 0; s < 4; ++s) Bq[b][s] = ok ? ld16(qrow + 16 * s) : v4i{0, 0, 0, 0};
      th[b] = collect && q < nQ ? theta[q] : INFINITY;
   }
   float gm[2][16];
#pragma unroll
   for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) gm[b][v] = -INFINITY;
   int cnt[2] = {0, 0};
   // staging: thread copies units tid + 256 j (j < 8) of a stage: row u / 8, unit u % 8
   v4i stg[8];
   auto fetch = [&](int st) {
      const int r0 = rBegin + st * kSweepStageRows;
#pragma unroll
      for (int j = 0; j < 8; ++j)
      {
         const int u = tid + 256 * j, row = u >> 3, part = u & 7;
         const int r = r0 + row;
         stg[j] = r < rEnd ? ld16(reinterpret_cast<const int8_t *>(ER) + (size_t)r * kEmbBytes + part * 16)
                           : v4i{0, 0, 0, 0};
      }
   };
   auto stash = [&](int buf) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
      {
         const int u = tid + 256 * j, row = u >> 3, part = u & 7;
         *reinterpret_cast<v4i *>(&sR[buf][row * kEmbBytes + ((part ^ (row & 7)) << 4)]) = stg[j];
      }
   };
   if (nStages > 0)
   {
      fetch(0);
      stash(0);
   }
   __syncthreads();
   for (int st = 0; st < nStages; ++st)
   {
      if (st + 1 < nStages) fetch(st + 1);
      const int r0s = rBegin + st * kSweepStageRows;
      if (active)
      {
         const int8_t *buf = sR[st & 1];
         const int nb = min(kSweepStageRows, rEnd - r0s) / 32;
         for (int blk = 0; blk < nb; ++blk)
         {
            const int row = blk * 32 + lq;
            v4i A[4];
#pragma unroll
            for (int s = 0; s < 4; ++s)
               A[s] = *reinterpret_cast<const v4i *>(buf + row * kEmbBytes + (((4 * half + s) ^ (row & 7)) << 4));
            const int r0 = r0s + blk * 32;
            const bool full = r0 + 32 <= nR;
#pragma unroll
            for (int b = 0; b < 2; ++b)
            {
               v16f acc = emb_block(A, Bq[b]);
               if (!full)
               {
                  // last block of the right set: rights >= nR score -inf
#pragma unroll
                  for (int v = 0; v < 16; ++v)
                  {
                     const int r = r0 + 8 * (v >> 2) + 4 * half + (v & 3);
                     acc[v] = r < nR ? acc[v] : -INFINITY;
                  }
               }
               if (!collect)
               {
#pragma unroll
                  for (int v = 0; v < 16; ++v) gm[b][v] = fmaxf(gm[b][v], acc[v]);
               }
               else
               {
                  float m = acc[0];
#pragma unroll
                  for (int v = 1; v < 16; ++v) m = fmaxf(m, acc[v]);
                  if (__any(m >= th[b]))
                  {
                     const size_t slot = ((size_t)(qw + 32 * b + lq) * nChunks + chunk) * 2 + half;
#pragma unroll
                     for (int v = 0; v < 16; ++v)
                     {
                        const int r = r0 + 8 * (v >> 2) + 4 * half + (v & 3);
                        if (acc[v] >= th[b])
                        {
                           if (cnt[b] < kListCap)
                           {
                              listS[slot * kListCap + cnt[b]] = acc[v];
                              listI[slot * kListCap + cnt[b]] = r;
                           }
                           ++cnt[b];
                        }
                     }
                  }
               }
            }
         }
      }
      if (st + 1 < nStages) stash((st + 1) & 1);
      __syncthreads();
   }
   if (!active) return;
#pragma unroll
   for (int b = 0; b < 2; ++b)
   {
      const int q = qw + 32 * b + lq;
      if (q >= nQ) continue;
      const size_t slot = ((size_t)q * nChunks + chunk) * 2 + half;
      if (!collect)
      {
#pragma unroll
         for (int v = 0; v < 16; ++v) gmax[slot * 16 + v] = gm[b][v];
      }
      else
         listN[slot] = cnt[b];
   }
}

// theta[q]: the K-th largest of the query's G = 8K group maxima, where group
// (i, h, v), i < K/4, is the union over chunks c = i mod (K/4) of the chunk
// groups (c, h, v) (so G does not grow with the right count).  One wave per
// query, rank counting over the G values in LDS.  theta = -inf when fewer
// than K groups exist.
template <int K>
__global__ __launch_bounds__(256) void k_emb_theta(int nQ, int nChunks, const float *__restrict__ gmax,
                                                   float *__restrict__ theta)
{
   constexpr int G = 8 * K, M = K / 4;
   __shared__ __attribute__((aligned(16))) float sv[4][G];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int q = blockIdx.x * 4 + wave;
   if (q >= nQ) return;
   const int folds = nChunks < M ? nChunks : M;
   const int nG = folds * 32;
   const float *g = gmax + (size_t)q * nChunks * 32;
   for (int e = lane; e < nG; e += 64)
   {
      const int i = e >> 5, hv = e & 31;
      float m = -INFINITY;
      for (int c = i; c < nChunks; c += M) m = fmaxf(m, g[c * 32 + hv]);
      sv[wave][e] = m;
   }
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   if (nG < K)
   {
      if (lane == 0) theta[q] = -INFINITY;
      return;
   }
   // nG is a multiple of 32; each lane ranks up to G / 64 values at once
   constexpr int PER = G / 64;
   float x[PER];
   int gt[PER], ge[PER];
#pragma unroll
   for (int i = 0; i < PER; ++i)
   {
      const int e = lane + 64 * i;
      x[i] = e < nG ? sv[wave][e] : INFINITY;
      gt[i] = ge[i] = 0;
   }
   for (int j = 0; j < nG; j += 4)
   {
      const float4 y = *reinterpret_cast<const float4 *>(&sv[wave][j]);
#pragma unroll
      for (int i = 0; i < PER; ++i)
      {
         gt[i] += (y.x > x[i]) + (y.y > x[i]) + (y.z > x[i]) + (y.w > x[i]);
         ge[i] += (y.x >= x[i]) + (y.y >= x[i]) + (y.z >= x[i]) + (y.w >= x[i]);
      }
   }
#pragma unroll
   for (int i = 0; i < PER; ++i)
      if (lane + 64 * i < nG && gt[i] < K && ge[i] >= K) theta[q] = x[i]; // all such x are equal
}

// Top K of each query's lists (score descending, index ascending), one wave
// per query: entries are gathered into LDS (at most kSelCap) and placed by
// rank.  ovf[q] = 1 when a list or the gather overflowed (the candidates may
// then miss members of the top K; the matcher rescans the query).
constexpr int kSelCap = 256;

template <int K>
__global__ __launch_bounds__(256) void k_emb_select(int nQ, int nChunks, const float *__restrict__ listS,
                                                    const int *__restrict__ listI, const int *__restrict__ listN,
                                                    int *__restrict__ cand, float *__restrict__ candS,
                                                    int *__restrict__ ovf)
{
   __shared__ float es[4][kSelCap];
   __shared__ int ei[4][kSelCap];
   __shared__ float os_[4][K];
   __shared__ int oi_[4][K];
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   const int q = blockIdx.x * 4 + wave;
   if (q >= nQ) return;
   const int nLists = nChunks * 2; // <= 64: one list per lane
   const size_t slot0 = (size_t)q * nLists;
   const int n = lane < nLists ? listN[slot0 + lane] : 0;
   int over = n > kListCap;
   const int take = min(n, kListCap);
   // exclusive prefix sum of take over the lanes
   int off = take;
#pragma unroll
   for (int d = 1; d < 64; d <<= 1)
   {
      const int y = __shfl_up(off, d, 64);
      if (lane >= d) off += y;
   }
   const int total = __shfl(off, 63, 64);
   off -= take;
   for (int e = 0; e < take; ++e)
      if (off + e < kSelCap)
      {
         es[wave][off + e] = listS[(slot0 + lane) * kListCap + e];
         ei[wave][off + e] = listI[(slot0 + lane) * kListCap + e];
      }
   over |= total > kSelCap;
   const int nE = min(total, kSelCap);
   if (lane < K)
   {
      os_[wave][lane] = -INFINITY;
      oi_[wave][lane] = -1;
   }
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   for (int e = lane; e < nE; e += 64)
   {
      const float x = es[wave][e];
      const int xi = ei[wave][e];
      int rank = 0;
      for (int j = 0; j < nE; ++j)
      {
         const float y = es[wave][j];
         rank += y > x || (y == x && ei[wave][j] < xi);
      }
      if (rank < K) // ranks are distinct: (score, index) orders strictly
      {
         os_[wave][rank] = x;
         oi_[wave][rank] = xi;
      }
   }
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   if (lane < K)
   {
      cand[(size_t)q * K + lane] = oi_[wave][lane];
      if (candS) candS[(size_t)q * K + lane] = os_[wave][lane];
   }
   over = __any(over) ? 1 : 0;
   if (lane == 0) ovf[q] = over;
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

// One thread per (query, candidate), K consecutive lanes per query (K = 16
// or 32 divides the wave): exact scores; the query's own top-2 is reduced
// across its K lanes and stored directly (nothing else writes left keys
// before the fallback scans); each right's top-2 takes the key atomically.
__global__ __launch_bounds__(256) void k_cand_insert(int nL, int nR, int K, const int *__restrict__ cand,
                                                     const int8_t *__restrict__ L, const int8_t *__restrict__ R,
                                                     uint64_t *top1L, uint64_t *top2L, uint64_t *top1R,
                                                     uint64_t *top2R, const uint32_t *epochPtr)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int k = (int)(t / K);
   const bool live = k < nL;
   const int j = live ? cand[t] : -1;
   const uint32_t epoch = *epochPtr;
   uint64_t key = 0;
   if (live && j >= 0 && j < nR)
   {
      const uint32_t s = (uint32_t)exact_score(L, R, k, j);
      key = pack_key(epoch, s, (uint32_t)j);
      top2_insert(top1R + j, top2R + j, pack_key(epoch, s, (uint32_t)k));
   }
   // keys are unique within a query (distinct candidates): max, then max of the rest
   unsigned long long best = key;
   for (int d = 1; d < K; d <<= 1)
   {
      const unsigned long long o = __shfl_xor(best, d, K);
      best = o > best ? o : best;
   }
   unsigned long long second = key == best ? 0ull : key;
   for (int d = 1; d < K; d <<= 1)
   {
      const unsigned long long o = __shfl_xor(second, d, K);
      second = o > second ? o : second;
   }
   if (live && (t % K) == 0)
   {
      top1L[k] = best;
      top2L[k] = second;
   }
}

// Margin flags.  Left k is flagged when best - second < margin over the pairs
// scored so far (or force[k]: its candidate list overflowed); right j also
// when it has exactly one scored pair (its second is unknown).  Flagged
// indices are appended to list (count in *count).
__global__ void k_flag(int n, int isRight, const uint64_t *__restrict__ top1, const uint64_t *__restrict__ top2,
                       const uint32_t *epochPtr, float margin, const int *__restrict__ force, int *__restrict__ list,
                       int *__restrict__ count)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const uint32_t epoch = *epochPtr;
   const uint64_t a = top1[i], b = top2[i];
   bool flag = force && force[i];
   if (!flag && key_epoch(a) == epoch)
   {
      const float best = (float)key_score(a) * (1.0f / 262144.0f);
      if (key_epoch(b) == epoch) flag = best - (float)key_score(b) * (1.0f / 262144.0f) < margin;
      else flag = isRight != 0;
   }
   if (flag) list[atomicAdd(count, 1)] = i;
}

// Gathers rows idx[0..n) of an int8 pool and its row sums into a compact
// pool of nPad rows (rows >= n zero, with the padding row sum).
// nDev (may be null): the row count in device memory (a flag count of this
// job); the grid then covers the largest count and the rows past the count's
// own 128-row padding are left alone.
__global__ void k_gather_rows(int n, int nPad, const int *__restrict__ nDev, const int *__restrict__ idx,
                              const int8_t *__restrict__ src, const int32_t *__restrict__ srcSum,
                              int8_t *__restrict__ dst, int32_t *__restrict__ dstSum)
{
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (nDev)
   {
      n = *nDev;
      nPad = max((n + kRowPad - 1) / kRowPad * kRowPad, kRowPad);
   }
   if (t >= (int64_t)nPad * (kDim / 16)) return;
   const int row = (int)(t / (kDim / 16)), part = (int)(t % (kDim / 16));
   uint4 v = make_uint4(0, 0, 0, 0);
   int32_t s = kPadRowSum;
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

int emb_topk_chunks(int nR, int K)
{
   // chunks of at most 2048 rights (512 workgroups on 16384 x 16384, at most
   // 32 chunks), and at least K/4 chunks (8K groups per query, so few rights
   // clear theta); every chunk keeps at least 32 rights
   int c = (nR + 2047) / 2048;
   c = c < 32 ? c : 32;
   c = c > K / 4 ? c : K / 4;
   const int most = ((nR + 31) / 32) > 0 ? (nR + 31) / 32 : 1;
   return c < 1 ? 1 : (c > most ? most : c);
}

size_t emb_topk_scratch_bytes(int nQ, int nChunks)
{
   const size_t slots = (size_t)(nQ > 0 ? nQ : 1) * nChunks * 2;
   return slots * 16 * sizeof(float) + (size_t)(nQ > 0 ? nQ : 1) * sizeof(float) +
          slots * kListCap * (sizeof(float) + sizeof(int)) + slots * sizeof(int);
}

hipError_t launch_emb_topk(int nQ, int nR, int K, const uint16_t *EQ, const uint16_t *ER, int nChunks, void *scratch,
                           int *cand, float *candS, int *ovf, hipStream_t st)
{
   if (nQ <= 0) return hipSuccess;
   if (K != 16 && K != 32) return hipErrorInvalidValue;
   const size_t slots = (size_t)nQ * nChunks * 2;
   float *gmax = reinterpret_cast<float *>(scratch);
   float *theta = gmax + slots * 16;
   float *listS = theta + nQ;
   int *listI = reinterpret_cast<int *>(listS + slots * kListCap);
   int *listN = listI + slots * kListCap;
   const int nRpad = (nR + 31) & ~31;
   const int chunkRows = ((nRpad + nChunks - 1) / nChunks + 31) & ~31;
   dim3 grid((nQ + 255) / 256, nChunks);
   hipLaunchKernelGGL(k_emb_sweep, grid, dim3(256), 0, st, 0, nQ, nR, EQ, ER, chunkRows, gmax, theta, listS, listI,
                      listN);
   dim3 wg((nQ + 3) / 4);
   if (K == 16) hipLaunchKernelGGL(k_emb_theta<16>, wg, dim3(256), 0, st, nQ, nChunks, gmax, theta);
   else hipLaunchKernelGGL(k_emb_theta<32>, wg, dim3(256), 0, st, nQ, nChunks, gmax, theta);
   hipLaunchKernelGGL(k_emb_sweep, grid, dim3(256), 0, st, 1, nQ, nR, EQ, ER, chunkRows, gmax, theta, listS, listI,
                      listN);
   if (K == 16)
      hipLaunchKernelGGL(k_emb_select<16>, wg, dim3(256), 0, st, nQ, nChunks, listS, listI, listN, cand, candS, ovf);
   else
      hipLaunchKernelGGL(k_emb_select<32>, wg, dim3(256), 0, st, nQ, nChunks, listS, listI, listN, cand, candS, ovf);
   return hipGetLastError();
}

hipError_t launch_cand_insert(int nL, int nR, int K, const int *cand, const int8_t *L, const int8_t *R,
                              uint64_t *top1L, uint64_t *top2L, uint64_t *top1R, uint64_t *top2R,
                              const uint32_t *epoch, hipStream_t st)
{
   const int64_t n = (int64_t)nL * K;
   if (n <= 0) return hipSuccess;
   if (K != 16 && K != 32) return hipErrorInvalidValue;
   hipLaunchKernelGGL(k_cand_insert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nL, nR, K, cand, L, R, top1L,
                      top2L, top1R, top2R, epoch);
   return hipGetLastError();
}

hipError_t launch_flag(int n, int isRight, const uint64_t *top1, const uint64_t *top2, const uint32_t *epoch,
                       float margin, const int *force, int *list, int *count, hipStream_t st)
{
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_flag, dim3((n + 255) / 256), dim3(256), 0, st, n, isRight, top1, top2, epoch, margin, force,
                      list, count);
   return hipGetLastError();
}

hipError_t launch_gather_rows(int n, int nPad, const int *nDev, const int *idx, const int8_t *src,
                              const int32_t *srcSum, int8_t *dst, int32_t *dstSum, hipStream_t st)
{
   const int64_t t = (int64_t)nPad * (kDim / 16);
   if (t <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, st, n, nPad, nDev, idx, src,
                      srcSum, dst, dstSum);
   return hipGetLastError();
}

// The two fallback scans' pair descriptors from this job's flag counts:
// side 0 = the flagged lefts against every right, side 1 = every left
// against the flagged rights, tiles of tileM x tileN.
__global__ void k_fallback_pairs(const int *__restrict__ count, int nL, int nR, int tileM, int tileN,
                                 PairDesc *__restrict__ out)
{
   if (threadIdx.x >= 2) return;
   const int side = threadIdx.x;
   PairDesc q{};
   q.nL = side == 0 ? count[0] : nL;
   q.nR = side == 0 ? nR : count[1];
   q.tilesM = ((max(q.nL, 1) + kRowPad - 1) / kRowPad * kRowPad + tileM - 1) / tileM;
   q.tilesN = ((max(q.nR, 1) + kRowPad - 1) / kRowPad * kRowPad + tileN - 1) / tileN;
   if (q.nL == 0 || q.nR == 0) q.tilesM = q.tilesN = 0;
   out[side] = q;
}

hipError_t launch_fallback_pairs(const int *count, int nL, int nR, int tileM, int tileN, PairDesc *out,
                                 hipStream_t st)
{
   hipLaunchKernelGGL(k_fallback_pairs, dim3(1), dim3(64), 0, st, count, nL, nR, tileM, tileN, out);
   return hipGetLastError();
}

} // namespace v3dm
