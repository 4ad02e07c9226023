// k_bf16.hip -- float descriptors on gfx950: BF16 MFMA candidate scan with
// FP32 accumulation, exact FP32 rescoring of the candidates, finalize.
//
// Replaces the reference's SGEMM + selection passes for float descriptors
// (V3D/src/CUDA/v3d_cudamatching.cpp:104-126, v3d_cudamatching.cu:117-194)
// without storing the score matrix and without FP64 anywhere.
//
// Candidate bound.  Descriptors are rounded to BF16 (round to nearest even,
// unit roundoff u = 2^-8) and multiplied on v_mfma_f32_32x32x16_bf16: the
// products of two BF16 values are exact in FP32 and the 128-term sum is
// accumulated in FP32.  For rows x, y
//   |S_bf16 - x.y| <= (2u + u^2) sum|x_k y_k| + gamma_128 sum|x_k y_k|
//                  <= (2^-7 + 2^-14) |x| |y|          (Cauchy-Schwarz),
// and the plain kernel's FP32 score (sequential fmaf, k = 0..127) is within
// 128 * 2^-24 |x||y| of x.y, inside the same 2^-14 slack.  With t the
// survivor threshold of the exact rule (see k_i8.hip: scores <= t cannot
// change a decision) and delta = (2^-7 + 2^-14) max|x| max|y| (1 + 1e-4),
// every pair whose FP32 score exceeds t has a BF16 score above
// lim = t - delta.  The scan keeps exactly those candidates (typically a
// handful per row); the rescoring kernel recomputes each candidate's score
// with the plain kernel's FP32 operation sequence and feeds the exact top-2
// keys.  The match list therefore equals the plain FP32 kernel's list, and
// the only approximation left is the reference's own FP32 arithmetic.
#include <type_traits>

#include <hip/hip_ext.h>

#include "v3dm_device.h"

namespace v3dm {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Operand packing.  A BF16 row is 256 bytes.  Byte e of a row goes to lane
// half h = e / 128 and MFMA step s = (e % 128) / 16 (16 bytes = 8 BF16 per
// lane and step), identically for A (left, lane l carries row l % 32) and B
// (right, lane l carries row l % 32), so each of the 8 MFMAs pairs the same
// 16 elements of both rows and the 8 steps cover all 128 (the sum does not
// depend on which k the hardware assigns to which operand slot, as long as
// both operands agree).  Accumulator register v of lane l is
// D[8*(v/4) + 4*(l/32) + v%4][l%32].
//
// Staging: as in the int8 scan, with 8 KiB stages of 32 right rows in a
// 3-group LDS ring; a group (one barrier) is 1 stage for 4-wave workgroups
// and 2 stages for 8-wave ones, so each thread moves two 16-byte units per
// group and the register ring stays at 16 VGPRs (the A rows take 64, the B
// fragments 2 x 32).  16-byte unit p = row * 16 + (chunk ^ swz(row)).
constexpr int kRowBytesBF = kDim * 2;                   // 256
constexpr int kFragsBF = kRowBytesBF / 32;              // 8
constexpr int kStageBytesBF = kStepCols * kRowBytesBF;  // 8 KiB

__device__ inline float fmax16(const v16f &x)
{
   float m0 = fmaxf(fmaxf(x[0], x[1]), x[2]);
   float m1 = fmaxf(fmaxf(x[3], x[4]), x[5]);
   float m2 = fmaxf(fmaxf(x[6], x[7]), x[8]);
   float m3 = fmaxf(fmaxf(x[9], x[10]), x[11]);
   float m4 = fmaxf(fmaxf(x[12], x[13]), x[14]);
   return fmaxf(fmaxf(fmaxf(m0, m1), fmaxf(m2, m3)), fmaxf(m4, x[15]));
}

__device__ inline v16f mfma_bf16(const v4i &a, const v4i &b, const v16f &c)
{
   return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                  0);
}

// One workgroup per tile, or (a.persistent) resident workgroups pulling
// tiles from a.tileCounter, exactly as k_i8_scan.
template <int WAVES, int TILE_N>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(2))) void k_bf16_scan(ScanArgsF a)
{
   touch_kernargs(a);
   constexpr int kTileM = WAVES * kWaveRows;
   constexpr int kGroupBF = WAVES >= 8 ? 2 : 1;            // steps per barrier
   constexpr int kGroupBytesBF = kGroupBF * kStageBytesBF;
   __shared__ __attribute__((aligned(16))) int8_t sB[kThirds][kGroupBytesBF];
   __shared__ int sNextTile;

   const int tid = threadIdx.x;
   const int lane = tid & 63;
   const int wave = tid >> 6;
   int tile = blockIdx.x;
   while (tile < a.nTiles)
   {
   uint32_t nextReq = 0;
   if (a.persistent && tid == 0) nextReq = atomicAdd(a.tileCounter, 1u) + gridDim.x;
   PairDesc pd = a.pair0;
   if (a.nPairs > 1) pd = a.pairs[sgpr(pair_of_tile(a.tilePair, a.pairs, a.nPairs, tile))];
   pd.offL = ((int64_t)sgpr((int)(pd.offL >> 32)) << 32) | (uint32_t)sgpr((int)pd.offL);
   pd.offR = ((int64_t)sgpr((int)(pd.offR >> 32)) << 32) | (uint32_t)sgpr((int)pd.offR);
   pd.nL = sgpr(pd.nL);
   pd.nR = sgpr(pd.nR);
   pd.tilesN = sgpr(pd.tilesN);
   pd.tileBase = sgpr(pd.tileBase);
   const int t = tile - pd.tileBase;
   const int tm = t / pd.tilesN, tn = t - tm * pd.tilesN;
   const int nLpad = (pd.nL + kRowPad - 1) / kRowPad * kRowPad;
   const int nRpad = (pd.nR + kRowPad - 1) / kRowPad * kRowPad;
   const int l0 = sgpr(tm * kTileM + wave * kWaveRows);
   const bool active = l0 < nLpad; // wave-uniform; inactive waves still load and sync
   const float lim = a.lim;
   const int half = lane >> 5, lrow = lane & 31;

   const int c0 = sgpr(tn * TILE_N);
   const int nsteps = sgpr((min(c0 + TILE_N, nRpad) - c0) / kStepCols); // multiple of 4: ngroups even
   const int ngroups = nsteps / kGroupBF;
   const int8_t *Rt = reinterpret_cast<const int8_t *>(a.R) + (size_t)(pd.offR + c0) * kRowBytesBF;
   const __amdgpu_buffer_rsrc_t rsB = __builtin_amdgcn_make_buffer_rsrc((void *)Rt, (short)0,
                                                                         nsteps * kStageBytesBF, kRsrcWord3);
   uint64_t *region = a.surv + ((size_t)tile * WAVES + wave) * a.survCap;
   int nsurv = 0;

   constexpr int kUnits = kGroupBytesBF / 16 / (WAVES * 64);
   int voff[kUnits], loff[kUnits];
#pragma unroll
   for (int j = 0; j < kUnits; ++j)
   {
      const int q = tid + WAVES * 64 * j, st = q >> 9, pu = q & 511, r = pu >> 4, x = pu & 15;
      voff[j] = st * kStageBytesBF + r * kRowBytesBF + ((x ^ swz(r)) << 4);
      loff[j] = st * kStageBytesBF + pu * 16;
   }
   v4i ring[2][kUnits];
   auto fetch = [&](auto slotc, int g) {
      constexpr int slot = decltype(slotc)::value;
      const int gc = min(g, ngroups - 1);
#pragma unroll
      for (int j = 0; j < kUnits; ++j)
         ring[slot][j] = __builtin_amdgcn_raw_buffer_load_b128(rsB, voff[j], gc * kGroupBytesBF, 0);
   };
   auto store = [&](auto slotc, int third) {
      constexpr int slot = decltype(slotc)::value;
#pragma unroll
      for (int j = 0; j < kUnits; ++j) *reinterpret_cast<v4i *>(&sB[third][loff[j]]) = ring[slot][j];
   };
   const std::integral_constant<int, 0> R0;
   const std::integral_constant<int, 1> R1;

   fetch(R0, 0);
   fetch(R1, 1);

   v4i A[2][kFragsBF];
   if (active)
   {
#pragma unroll
      for (int b = 0; b < 2; ++b)
      {
         const int8_t *ap = reinterpret_cast<const int8_t *>(a.L) +
                            (size_t)(pd.offL + l0 + 32 * b + lrow) * kRowBytesBF + 128 * half;
#pragma unroll
         for (int s = 0; s < kFragsBF; ++s) A[b][s] = ld16(ap + 16 * s);
      }
   }
   int fragoff[kFragsBF];
#pragma unroll
   for (int k = 0; k < kFragsBF; ++k) fragoff[k] = (lrow * 16 + ((8 * half + k) ^ swz(lrow))) << 4;

   store(R0, 0);
   fetch(R0, 2);

   const v16f zero = {};
   v4i Bc[kFragsBF];
   // `lastc` says at compile time whether g is the tile's last group, so the
   // next group's fragments are read, and renamed into Bc, unconditionally in
   // every other group (a run-time test there made the compiler copy all the
   // fragments at every group boundary).
   auto group = [&](auto slotc, int g, auto lastc) {
      constexpr bool kLast = decltype(lastc)::value;
      store(slotc, (g + 1) % kThirds);
      fetch(slotc, g + 3);
      lds_barrier();
      if (!active) return;
      const int8_t *sb = sB[g % kThirds];
      if (g == 0)
      {
#pragma unroll
         for (int k = 0; k < kFragsBF; ++k) Bc[k] = *reinterpret_cast<const v4i *>(sb + fragoff[k]);
      }
#pragma unroll
      for (int s = 0; s < kGroupBF; ++s)
      {
         v4i Bn[kFragsBF];
         if (s + 1 < kGroupBF)
         {
#pragma unroll
            for (int k = 0; k < kFragsBF; ++k)
               Bn[k] = *reinterpret_cast<const v4i *>(sb + (s + 1) * kStageBytesBF + fragoff[k]);
         }
         else if (!kLast)
         {
            const int8_t *sn = sB[(g + 1) % kThirds];
#pragma unroll
            for (int k = 0; k < kFragsBF; ++k) Bn[k] = *reinterpret_cast<const v4i *>(sn + fragoff[k]);
         }
         v16f acc0 = mfma_bf16(A[0][0], Bc[0], zero);
         v16f acc1 = mfma_bf16(A[1][0], Bc[0], zero);
#pragma unroll
         for (int k = 1; k < kFragsBF; ++k)
         {
            acc0 = mfma_bf16(A[0][k], Bc[k], acc0);
            acc1 = mfma_bf16(A[1][k], Bc[k], acc1);
         }
         // NaN scores (padding rows and columns) pass neither test; fmaxf skips them
         const float m = fmaxf(fmax16(acc0), fmax16(acc1));
         if (__builtin_expect(__any(m > lim), 0))
         {
            const int jrel = (g * kGroupBF + s) * kStepCols + lrow;
            // recomputed on every visit, so the entry constants below are not
            // hoisted out of the loop (and spilled)
            int half4 = 4 * half;
            asm volatile("" : "+v"(half4));
#pragma unroll
            for (int b = 0; b < 2; ++b)
            {
               const v16f &acc = b == 0 ? acc0 : acc1;
#pragma unroll
               for (int q = 0; q < 4; ++q)
               {
                  const float mq = fmaxf(fmaxf(acc[4 * q], acc[4 * q + 1]), fmaxf(acc[4 * q + 2], acc[4 * q + 3]));
                  if (!__any(mq > lim)) continue;
#pragma unroll
                  for (int e = 0; e < 4; ++e)
                  {
                     const int v = 4 * q + e;
                     const bool h = acc[v] > lim;
                     const uint64_t ball = __ballot(h);
                     if (ball == 0) continue;
                     const int irel = 32 * b + 8 * q + half4 + e;
                     const int slot = nsurv + __builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0));
                     if (h && slot < a.survCap)
                        region[slot] = ((uint64_t)__float_as_uint(acc[v]) << 32) | ((uint32_t)irel << 16) |
                                       (uint32_t)jrel;
                     nsurv += __builtin_popcountll(ball);
                  }
               }
            }
         }
         if (s + 1 < kGroupBF || !kLast)
         {
#pragma unroll
            for (int k = 0; k < kFragsBF; ++k) Bc[k] = Bn[k];
         }
      }
   };
   const std::false_type kMore;
   const std::true_type kEnd;
   int g = 0;
   for (; g + 2 < ngroups; g += 2)
   {
      group(R1, g, kMore);
      group(R0, g + 1, kMore);
   }
   if (g + 1 < ngroups)
   {
      group(R1, g, kMore);
      group(R0, g + 1, kEnd);
   }
   else if (g < ngroups)
      group(R1, g, kEnd);
   if (lane == 0) a.survCount[(size_t)tile * WAVES + wave] = nsurv;
   if (!a.persistent) break;
   __syncthreads();
   if (tid == 0) sNextTile = (int)nextReq;
   __syncthreads();
   tile = sNextTile;
   }
}

// The plain kernel's FP32 score: acc = fmaf(x[k], y[k], acc), k = 0..127,
// starting from 0 (k_sgemm_seq), so the rescored value is bit-identical.
__device__ inline float dot_seq(const float *__restrict__ x, const float *__restrict__ y)
{
   float acc = 0.0f;
#pragma unroll 8
   for (int k = 0; k < kDim; k += 4)
   {
      const float4 p = *reinterpret_cast<const float4 *>(x + k);
      const float4 q = *reinterpret_cast<const float4 *>(y + k);
      acc = fmaf(p.x, q.x, acc);
      acc = fmaf(p.y, q.y, acc);
      acc = fmaf(p.z, q.z, acc);
      acc = fmaf(p.w, q.w, acc);
   }
   return acc;
}

__device__ inline void rescore_insert(const ScanArgsF &a, const PairDesc &pd, uint32_t epoch, int il, int jr)
{
   if (il >= pd.nL || jr >= pd.nR) return;
   if (pd.guide > 0)
   {
      const Mask &mk = a.masks[pd.guide - 1];
      if (mask_hit(mk, mk.lpos[2 * il], mk.lpos[2 * il + 1], mk.rpos[2 * jr], mk.rpos[2 * jr + 1])) return;
   }
   for (int m = 0; m < a.nMasks; ++m)
   {
      const Mask &mk = a.masks[m];
      if (mask_hit(mk, mk.lpos[2 * il], mk.lpos[2 * il + 1], mk.rpos[2 * jr], mk.rpos[2 * jr + 1])) return;
   }
   const float s = dot_seq(a.Lf + (size_t)(pd.offL + il) * kDim, a.Rf + (size_t)(pd.offR + jr) * kDim);
   if (!(s > a.thr)) return;
   top2_insert_both(a.top1L + pd.slotL + il, a.top2L + pd.slotL + il, pack_keyf(epoch, s, (uint32_t)jr),
                    a.top1R + pd.slotR + jr, a.top2R + pd.slotR + jr, pack_keyf(epoch, s, (uint32_t)il));
}

// One wave per (tile, wave) region: every candidate rescored in FP32 and
// inserted; an overflowed region rescores its whole 64 x tileN block.
__global__ __launch_bounds__(256) void k_f32_rescore(ScanArgsF a)
{
   touch_kernargs(a);
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
   if (r >= a.nTiles * a.waves) return;
   // the count, the tile's pair, the epoch and the region's first 64 entries
   // (inside the region: survCap >= 64, checked on the host) are loaded before
   // any of them is waited for; stale entries past the count are not used
   const uint64_t *region = a.surv + (size_t)r * a.survCap;
   const uint64_t w0 = region[lane];
   const int tile = r / a.waves, wave = r - tile * a.waves;
   const int p = pair_of_tile(a.tilePair, a.pairs, a.nPairs, tile);
   const uint32_t epoch = job_epoch(a.epoch, a.epochValue);
   const int cnt = a.survCount[r];
   if (cnt == 0) return;
   const PairDesc pd = a.pairs[p];
   const int t = tile - pd.tileBase;
   const int tm = t / pd.tilesN, tn = t - tm * pd.tilesN;
   const int l0 = tm * a.tileM + wave * kWaveRows, c0 = tn * a.tileN;
   if (cnt <= a.survCap)
   {
      for (int e = lane; e < cnt; e += 64)
      {
         const uint64_t w = e == lane ? w0 : region[e];
         rescore_insert(a, pd, epoch, l0 + (int)((w >> 16) & 0xFFFF), c0 + (int)(w & 0xFFFF));
      }
      return;
   }
   const int nRows = min(kWaveRows, pd.nL - l0), nCols = min(a.tileN, pd.nR - c0);
   for (int e = lane; e < nRows * nCols; e += 64) rescore_insert(a, pd, epoch, l0 + e / nCols, c0 + e % nCols);
}

__device__ inline float keyf_value(uint64_t k, uint32_t epoch)
{
   return keyf_epoch(k) == epoch ? keyf_score(k) : -1e30f;
}

// As k_i8_finalize, on float-score keys.
__global__ __launch_bounds__(256) void k_f32_finalize(FinalizeArgs a)
{
   touch_kernargs(a);
   const int blk = blockIdx.x;
   PairDesc pd = a.pair0;
   if (!a.hasPair0) pd = a.pairs[pair_of_fin(a.finPair, a.pairs, a.nPairs, blk)];
   const int k = (blk - pd.finBase) * 256 + threadIdx.x;
   const uint32_t epoch = job_epoch(a.epoch, a.epochValue);
   if (k < pd.nL)
   {
      int out = -1;
      const uint64_t t1 = a.top1L[pd.slotL + k];
      const uint64_t t2 = a.top2L[pd.slotL + k];
      if (keyf_epoch(t1) == epoch)
      {
         const uint32_t j = key_index(t1);
         const uint64_t u1 = a.top1R[pd.slotR + j];
         const uint64_t u2 = a.top2R[pd.slotR + j];
         const float best = keyf_score(t1);
         const float next = keyf_value(t2, epoch);
         if (rule_accept(best, next, a.minScore, a.minRatio))
         {
            if (keyf_epoch(u1) == epoch && key_index(u1) == (uint32_t)k)
            {
               const float nextR = keyf_value(u2, epoch);
               if (rule_accept(keyf_score(u1), nextR, a.minScore, a.minRatio)) out = (int)j;
            }
         }
      }
      a.matchOfLeft[pd.slotL + k] = out;
   }
   if (!a.done) return; // the job's epoch is advanced by k_advance_epoch instead
   __syncthreads();
   if (threadIdx.x == 0)
   {
      __threadfence();
      const uint32_t prev = atomicAdd(a.done, 1u);
      if (prev == (uint32_t)gridDim.x - 1)
      {
         *a.done = 0;
         *a.tileCounter = 0;
         *a.epoch = epoch + 1;
         __threadfence();
      }
   }
}

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ inline uint32_t bf16_rne(float f)
{
   const uint32_t u = __float_as_uint(f);
   if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);
   return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// 16 threads per row, 8 floats each: BF16 row (zero rows up to nPad) and the
// largest squared row norm (atomicMax on the bits of a non-negative float).
__global__ __launch_bounds__(256) void k_prep_bf16(const float *__restrict__ src, int n, int nPad,
                                                   uint16_t *__restrict__ dst, uint32_t *norm2Max)
{
   const int row = blockIdx.x * 16 + (threadIdx.x >> 4);
   const int part = threadIdx.x & 15;
   if (row >= nPad) return;
   float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
   if (row < n)
   {
      const float4 *p = reinterpret_cast<const float4 *>(src + (size_t)row * kDim + part * 8);
      const float4 x0 = p[0], x1 = p[1];
      v[0] = x0.x; v[1] = x0.y; v[2] = x0.z; v[3] = x0.w;
      v[4] = x1.x; v[5] = x1.y; v[6] = x1.z; v[7] = x1.w;
   }
   uint32_t w[4];
   float n2 = 0.0f;
#pragma unroll
   for (int i = 0; i < 4; ++i)
   {
      // padding rows are quiet NaNs: every score with them is NaN, which no
      // survivor test passes (the scan needs no row or column bound checks)
      w[i] = row < n ? bf16_rne(v[2 * i]) | (bf16_rne(v[2 * i + 1]) << 16) : 0x7FC07FC0u;
      n2 = fmaf(v[2 * i], v[2 * i], n2);
      n2 = fmaf(v[2 * i + 1], v[2 * i + 1], n2);
   }
   *reinterpret_cast<uint4 *>(dst + (size_t)row * kDim + part * 8) = make_uint4(w[0], w[1], w[2], w[3]);
#pragma unroll
   for (int o = 8; o >= 1; o >>= 1) n2 += __shfl_xor(n2, o, 16);
   if (part == 0 && row < n) atomicMax(norm2Max, __float_as_uint(n2));
}

__global__ void k_fill_u32(uint32_t *p, int n, uint32_t v)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

// ---- launchers ------------------------------------------------------------

hipError_t launch_bf16_scan(int variant, int grid, const ScanArgsF &a, hipStream_t st, hipEvent_t start,
                            hipEvent_t stop)
{
   if (a.nTiles <= 0) return hipSuccess;
   switch (variant)
   {
   case 0: hipExtLaunchKernelGGL((k_bf16_scan<4, 512>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 1: hipExtLaunchKernelGGL((k_bf16_scan<4, 1024>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 2: hipExtLaunchKernelGGL((k_bf16_scan<8, 512>), dim3(grid), dim3(512), 0, st, start, stop, 0, a); break;
   case 3: hipExtLaunchKernelGGL((k_bf16_scan<4, 256>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   default: return hipErrorInvalidValue;
   }
   return hipGetLastError();
}

// Tile variants of the BF16 scan: the geometry of int8 scan variants 0-3.
int bf16_scan_variant_count() { return 4; }

int bf16_scan_resident(int variant, int device)
{
   const void *fn = nullptr;
   int threads = 256;
   switch (variant)
   {
   case 0: fn = reinterpret_cast<const void *>(&k_bf16_scan<4, 512>); break;
   case 1: fn = reinterpret_cast<const void *>(&k_bf16_scan<4, 1024>); break;
   case 2: fn = reinterpret_cast<const void *>(&k_bf16_scan<8, 512>); threads = 512; break;
   default: fn = reinterpret_cast<const void *>(&k_bf16_scan<4, 256>); break;
   }
   int perCU = 0, cus = 0;
   if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, threads, 0) != hipSuccess) perCU = 1;
   if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 1;
   return (perCU > 0 ? perCU : 1) * (cus > 0 ? cus : 1);
}

hipError_t launch_f32_rescore(const ScanArgsF &a, hipStream_t st)
{
   const int regions = a.nTiles * a.waves;
   if (regions <= 0) return hipSuccess;
   if (a.survCap < 64) return hipErrorInvalidValue; // each wave reads its region's first 64 entries
   hipLaunchKernelGGL(k_f32_rescore, dim3((regions + 3) / 4), dim3(256), 0, st, a);
   return hipGetLastError();
}

hipError_t launch_f32_finalize(const FinalizeArgs &a, hipStream_t st)
{
   if (a.nBlocks <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_f32_finalize, dim3(a.nBlocks), dim3(256), 0, st, a);
   return hipGetLastError();
}

hipError_t launch_prep_bf16(const float *src, int n, int nPad, uint16_t *dst, uint32_t *norm2Max, hipStream_t st,
                            bool resetNorm)
{
   if (nPad <= 0) return hipSuccess;
   if (resetNorm) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, norm2Max, 1, 0u);
   hipLaunchKernelGGL(k_prep_bf16, dim3((nPad + 15) / 16), dim3(256), 0, st, src, n, nPad, dst, norm2Max);
   return hipGetLastError();
}

} // namespace v3dm
