// k_i8.hip -- fused int8-MFMA exact matcher for uint8 descriptors (gfx950).
//
// Replaces the reference's SGEMM + column/row selection passes
// (V3D/src/CUDA/v3d_cudamatching.cpp:104-119, v3d_cudamatching.cu:117-194)
// with one scan kernel that never writes the score matrix, plus a small
// finalize kernel for the ratio test and the mutual check
// (v3d_cudamatching.cpp:121-126).
//
// Exact scores.  Descriptors u in [0,255] are stored as int8 a = u - 128 with
// the per-row term 128*sum(u).  For a left row x and right row y
//   sum_k u_x u_y = sum_k a_x a_y + 128 sum u_x + 128 sum u_y - 128^3,
// the first term from v_mfma_i32_32x32x32_i8 (int32 accumulation, exact),
// the left term entering as the MFMA's C operand, the right term added per
// lane.  The reference's FP32 score is this integer times 2^-18 exactly.
//
// Survivor filter.  With t = the caller's threshold (ScanArgs::thr, set on
// the host) every score <= t is provably unable to change a decision:
//  * ratio test off (minRatio < 0): t = floor(minScore * 2^18); a left/right
//    whose best is <= t is rejected either way, and the argmax over the
//    survivors is the argmax over all (all values equal to the best survive).
//  * ratio test on (minRatio > 0): t <= min(minScore, T) * 2^18 with
//    T = 1 - (1 - minScore) / (minRatio * (1 - 1e-5)).  If the true best b
//    survives but the true second s does not, the decision with s replaced by
//    the -1e30 initial value is "accept iff b > minScore", and the reference
//    also accepts: (2-2b)/(2-2s) < (1-minScore)/(1-T) = minRatio (1 - 1e-5),
//    and the FP32 division adds at most 2^-24 relative error.  If both
//    survive, the top-2 over survivors equals the top-2 over all scores.
//  * minRatio == 0 (or NaN): no filtering (t = -1; scores are >= 0).
// Only survivors (typically a few per row) reach the exact top-2 update.
//
// Exact top-2 per left row and per right row, order independent: each
// survivor is a unique 64-bit key (epoch | score | inverted index); a larger
// key is a larger score, then a smaller index (the reference's first-index
// rule).  Insertion: old = atomicMax(top1, key); atomicMax(top2, min(old, key)).
// Every key except the final maximum reaches top2 (directly, or as the `old`
// returned to the key that displaced it), so top2 ends as the second largest
// key, whose score is the reference's nextVal (a duplicate best counts).
// Keys of older jobs carry an older epoch and read as empty, so no clearing
// pass is needed between jobs.
#include "v3dm_internal.h"

namespace v3dm {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kScanWaves = 8;       // waves per workgroup
constexpr int kWaveRows = 64;       // left rows per wave (two 32-row MFMA blocks)
constexpr int kTileM = kScanWaves * kWaveRows; // 512 left rows per tile
constexpr int kTileN = 512;         // right rows per tile
constexpr int kBias = 1 << 21;      // 128^3

__device__ inline int find_pair_by_tile(const PairDesc *pairs, int nPairs, int tile)
{
   int lo = 0, hi = nPairs - 1;
   while (lo < hi)
   {
      int mid = (lo + hi + 1) >> 1;
      if (pairs[mid].tileBase <= tile) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

__device__ inline int find_pair_by_fin(const PairDesc *pairs, int nPairs, int blk)
{
   int lo = 0, hi = nPairs - 1;
   while (lo < hi)
   {
      int mid = (lo + hi + 1) >> 1;
      if (pairs[mid].finBase <= blk) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

// Epipolar Sampson test, v3d_cudamatching.cu:266-290 (this file is compiled
// with -ffp-contract=off so the expression is evaluated as written).
__device__ inline bool mask_hit(const Mask &mk, float px, float py, float qx, float qy)
{
   const float *F = mk.m;
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
      float sampson = num / denom;
      return sampson > mk.sqrThr;
   }
   // homography transfer, v3d_cudamatching.cu:306-322
   float Hpx = F[0] * px + F[1] * py + F[2];
   float Hpy = F[3] * px + F[4] * py + F[5];
   float Hpz = F[6] * px + F[7] * py + F[8];
   Hpx /= Hpz;
   Hpy /= Hpz;
   float sqrDist = (Hpx - qx) * (Hpx - qx) + (Hpy - qy) * (Hpy - qy);
   return sqrDist > mk.sqrThr;
}

__device__ inline void top2_insert(uint64_t *t1, uint64_t *t2, uint64_t key)
{
   unsigned long long old = atomicMax((unsigned long long *)t1, (unsigned long long)key);
   atomicMax((unsigned long long *)t2, (unsigned long long)(old < key ? old : key));
}

__device__ __forceinline__ void survivor(const ScanArgs &a, const PairDesc &pd, uint32_t epoch, int il, int jr,
                                         int32_t score)
{
   if (il >= pd.nL || jr >= pd.nR) return;
   for (int m = 0; m < a.nMasks; ++m)
   {
      const Mask &mk = a.masks[m];
      if (mask_hit(mk, mk.lpos[2 * il], mk.lpos[2 * il + 1], mk.rpos[2 * jr], mk.rpos[2 * jr + 1])) return;
   }
   top2_insert(a.top1L + pd.slotL + il, a.top2L + pd.slotL + il, pack_key(epoch, (uint32_t)score, (uint32_t)jr));
   top2_insert(a.top1R + pd.slotR + jr, a.top2R + pd.slotR + jr, pack_key(epoch, (uint32_t)score, (uint32_t)il));
}

__device__ inline v4i ld16(const int8_t *p)
{
   return *(__attribute__((address_space(1))) const v4i *)(p);
}

__device__ inline int max16(const v16i &x)
{
   int m0 = max(max(x[0], x[1]), x[2]);
   int m1 = max(max(x[3], x[4]), x[5]);
   int m2 = max(max(x[6], x[7]), x[8]);
   int m3 = max(max(x[9], x[10]), x[11]);
   int m4 = max(max(x[12], x[13]), x[14]);
   return max(max(max(m0, m1), max(m2, m3)), max(m4, x[15]));
}

// acc0/acc1 register v (0..31, block v / 16).  A variable-index vector
// extract stays in registers (indexed register moves), no scratch.
__device__ inline int select32(const v16i &a0, const v16i &a1, int v)
{
   const int x0 = a0[v & 15], x1 = a1[v & 15];
   return v < 16 ? x0 : x1;
}

__device__ inline int sgpr(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Operand packing.  Descriptor byte e of a row goes to lane half h = e / 64,
// MFMA step s = (e % 64) / 16, byte e % 16 of the lane's 16-byte operand.
// A (left) lane l carries row l % 32, B (right) lane l carries row l % 32, so
// every MFMA step pairs the same 32 descriptor bytes of both rows and four
// steps cover all 128.  Accumulator register v of lane l is
// D[8*(v/4) + 4*(l/32) + v%4][l%32].
//
// Tiling.  A workgroup of 8 waves owns 512 left rows (64 per wave, held in
// registers for the whole tile) and sweeps 512 right rows in steps of 32.
// Each step's 32 right rows (4 KiB) are loaded once per workgroup into an
// LDS stage and read by all 8 waves.  Global loads run kAhead steps ahead
// in a register ring; a 3-deep LDS ring lets one barrier per step separate
// the writes of step i+1 from the reads of step i-2 on the same stage.
// LDS stage layout: 16-byte unit p = row * 8 + (chunk ^ (row & 7)), so the
// fragment reads (32 rows, same chunk) hit distinct bank groups.

constexpr int kStepCols = 32;                  // right rows per step
constexpr int kStageBytes = kStepCols * kDim;  // 4 KiB
constexpr int kLdsStages = 3;
constexpr int kAhead = 4;                      // register-ring depth (steps)

struct StageRing
{
   uint64_t v[kAhead]; // this lane's 8 bytes of steps i+1 .. i+kAhead
   int s[kAhead];    // column sums (wave 0, lanes < 32)
};

typedef __attribute__((address_space(1))) const uint64_t gu64;
typedef __attribute__((address_space(1))) const int32_t gi32;
typedef __attribute__((address_space(1))) const v4i gv4i;

// Loads this lane's 8 bytes of step `step` and one column sum (all lanes load
// a sum so the load stream is uniform; wave 0 stores them).
__device__ inline void stage_fetch(uint64_t &v, int &sum, const int8_t *Rt, const int32_t *St, int step, int wave,
                                   int lane)
{
   const int q = wave * 64 + lane;         // 8-byte unit of the stage, 0..511
   const int p = q >> 1, r = p >> 3, x = p & 7;
   v = *(gu64 *)(Rt + (size_t)step * kStageBytes + r * kDim + ((x ^ (r & 7)) << 4) + ((q & 1) << 3));
   sum = *(gi32 *)(St + step * kStepCols + (lane & (kStepCols - 1)));
}

__device__ inline void stage_store(uint64_t v, int sum, int8_t *sB, int32_t *sC, int wave, int lane)
{
   const int q = wave * 64 + lane;
   *reinterpret_cast<uint64_t *>(sB + q * 8) = v;
   if (wave == 0 && lane < kStepCols) sC[lane] = sum;
}

__device__ inline void lds_barrier()
{
   // LDS writes of this wave done, then the workgroup barrier.  Written out
   // so no vmcnt wait is attached: the global-load ring stays in flight.
   asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__global__ __launch_bounds__(512) void k_i8_scan(ScanArgs a)
{
   __shared__ __attribute__((aligned(16))) int8_t sB[kLdsStages][kStageBytes];
   __shared__ int32_t sC[kLdsStages][kStepCols];

   const int lane = threadIdx.x & 63;
   const int wave = threadIdx.x >> 6;
   const int tile = blockIdx.x;
   const int p = sgpr(find_pair_by_tile(a.pairs, a.nPairs, tile));
   PairDesc pd = a.pairs[p];
   // tile parameters are workgroup-uniform: keep them in scalar registers
   pd.offL = ((int64_t)sgpr((int)(pd.offL >> 32)) << 32) | (uint32_t)sgpr((int)pd.offL);
   pd.offR = ((int64_t)sgpr((int)(pd.offR >> 32)) << 32) | (uint32_t)sgpr((int)pd.offR);
   pd.slotL = ((int64_t)sgpr((int)(pd.slotL >> 32)) << 32) | (uint32_t)sgpr((int)pd.slotL);
   pd.slotR = ((int64_t)sgpr((int)(pd.slotR >> 32)) << 32) | (uint32_t)sgpr((int)pd.slotR);
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
   const uint32_t epoch = (uint32_t)sgpr((int)*a.epoch);
   const int32_t thr = a.thr;
   const int half = lane >> 5, lrow = lane & 31;

   const int c0 = sgpr(tn * kTileN);
   const int nsteps = sgpr((min(c0 + kTileN, nRpad) - c0) / kStepCols); // multiple of kAhead
   const int8_t *Rt = a.R + (size_t)(pd.offR + c0) * kDim;
   const int32_t *St = a.Rsum + pd.offR + c0;

   // Every tile has a multiple of kAhead steps (images are padded to 128
   // rows), so the ring is filled and drained without branches; fetches past
   // the last step re-read the last step and are never stored to a read stage.
   StageRing ring;
#pragma unroll
   for (int j = 0; j < kAhead; ++j) stage_fetch(ring.v[j], ring.s[j], Rt, St, j, wave, lane);

   v4i A[2][4];
   v16i C[2];
   if (active)
   {
#pragma unroll
      for (int b = 0; b < 2; ++b)
      {
         const int8_t *ap = a.L + (size_t)(pd.offL + l0 + 32 * b + lrow) * kDim + 64 * half;
#pragma unroll
         for (int s = 0; s < 4; ++s) A[b][s] = ld16(ap + 16 * s);
#pragma unroll
         for (int q = 0; q < 4; ++q)
         {
            const v4i rs = *reinterpret_cast<const v4i *>(a.Lsum + pd.offL + l0 + 32 * b + 8 * q + 4 * half);
#pragma unroll
            for (int e = 0; e < 4; ++e) C[b][4 * q + e] = rs[e] - kBias;
         }
      }
   }

   // step 0 into stage 0
   stage_store(ring.v[0], ring.s[0], sB[0], sC[0], wave, lane);
   stage_fetch(ring.v[0], ring.s[0], Rt, St, min(kAhead, nsteps - 1), wave, lane);

   for (int i0 = 0; i0 < nsteps; i0 += kAhead)
   {
#pragma unroll
      for (int j = 0; j < kAhead; ++j)
      {
         const int i = i0 + j;
         {
            // step i+1 (ring slot (j+1) % kAhead) -> its LDS stage, then refill the slot.
            // Stage (i+1) % 3 was last read at step i-2, before the previous barrier.
            const int slot = (j + 1) % kAhead;
            stage_store(ring.v[slot], ring.s[slot], sB[(i + 1) % kLdsStages], sC[(i + 1) % kLdsStages], wave, lane);
            stage_fetch(ring.v[slot], ring.s[slot], Rt, St, min(i + 1 + kAhead, nsteps - 1), wave, lane);
            lds_barrier();
            if (active)
            {
               const int8_t *sb = sB[i % kLdsStages];
               v4i B[4];
#pragma unroll
               for (int s = 0; s < 4; ++s)
               {
                  const int ch = 4 * half + s;
                  B[s] = *reinterpret_cast<const v4i *>(sb + ((lrow * 8 + (ch ^ (lrow & 7))) << 4));
               }
               const int cs = sC[i % kLdsStages][lrow];
               v16i acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][0], B[0], C[0], 0, 0, 0);
               v16i acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][0], B[0], C[1], 0, 0, 0);
#pragma unroll
               for (int s = 1; s < 4; ++s)
               {
                  acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][s], B[s], acc0, 0, 0, 0);
                  acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][s], B[s], acc1, 0, 0, 0);
               }
               const int m = max(max16(acc0), max16(acc1));
               const int lim = thr - cs; // score > thr  <=>  acc > thr - 128*sum(u_right)
               if (__builtin_expect(__any(m > lim), 0))
               {
                  // rare: exact top-2 updates for every surviving score of this lane
                  uint32_t hits = 0;
#pragma unroll
                  for (int v = 0; v < 16; ++v)
                  {
                     hits |= acc0[v] > lim ? 1u << v : 0u;
                     hits |= acc1[v] > lim ? 1u << (16 + v) : 0u;
                  }
                  const int col = c0 + i * kStepCols + lrow;
                  while (hits)
                  {
                     const int v = __builtin_ctz(hits);
                     hits &= hits - 1;
                     const int il = l0 + 32 * (v >> 4) + 8 * ((v >> 2) & 3) + 4 * half + (v & 3);
                     survivor(a, pd, epoch, il, col, select32(acc0, acc1, v) + cs);
                  }
               }
            }
         }
      }
   }
}

__device__ inline bool rule_accept(float best, float next, float minScore, float minRatio)
{
   // v3d_cudamatching.cu:32 (ratio test off) and :138-145 / :187-193
   if (minRatio < 0) return best > minScore;
   float nextDist = fmaxf(0.001f, 2.0f - 2.0f * next);
   float bestDist = 2.0f - 2.0f * best;
   float ratio = bestDist / nextDist;
   return best > minScore && ratio < minRatio;
}

__device__ inline float key_value(uint64_t k, uint32_t epoch)
{
   return key_epoch(k) == epoch ? (float)key_score(k) * (1.0f / 262144.0f) : -1e30f;
}

// One thread per left feature: column decision (ratio test on the left's
// top-2), then the right side's decision for the chosen right feature and
// the mutual check.  The last block to finish advances the epoch.
__global__ __launch_bounds__(256) void k_i8_finalize(FinalizeArgs a)
{
   const int blk = blockIdx.x;
   const int p = find_pair_by_fin(a.pairs, a.nPairs, blk);
   const PairDesc pd = a.pairs[p];
   const int k = (blk - pd.finBase) * 256 + threadIdx.x;
   const uint32_t epoch = *a.epoch;
   if (k < pd.nL)
   {
      int out = -1;
      const uint64_t t1 = a.top1L[pd.slotL + k];
      if (key_epoch(t1) == epoch)
      {
         const float best = key_value(t1, epoch);
         const uint32_t j = key_index(t1);
         const float next = key_value(a.top2L[pd.slotL + k], epoch);
         if (rule_accept(best, next, a.minScore, a.minRatio))
         {
            const uint64_t u1 = a.top1R[pd.slotR + j];
            if (key_epoch(u1) == epoch && key_index(u1) == (uint32_t)k)
            {
               const float nextR = key_value(a.top2R[pd.slotR + j], epoch);
               if (rule_accept(key_value(u1, epoch), nextR, a.minScore, a.minRatio)) out = (int)j;
            }
         }
      }
      a.matchOfLeft[pd.slotL + k] = out;
   }
   __syncthreads();
   if (threadIdx.x == 0)
   {
      __threadfence();
      const uint32_t prev = atomicAdd(a.done, 1u);
      if (prev == (uint32_t)gridDim.x - 1)
      {
         *a.done = 0;
         *a.epoch = epoch + 1;
         __threadfence();
      }
   }
}

// Sets every key slot to 0 (an epoch no job uses) and the epoch to 1.
__global__ void k_clear_keys(uint64_t *keys, int64_t n, uint32_t *epoch, uint32_t *done)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) keys[i] = 0;
   if (i == 0)
   {
      *epoch = 1;
      *done = 0;
   }
}

__global__ void k_mfma_probe(int32_t *out)
{
   const int lane = threadIdx.x;
   v4i A = {0, 0, 0, 0}, B = {0, 0, 0, 0};
   if (lane < 32)
   {
      A[0] = (lane & 0xFF) | (32 << 8); // bytes (i, 32)
      B[0] = 1 | ((lane & 0xFF) << 8);  // bytes (1, j)
   }
   v16i C = {};
   v16i D = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B, C, 0, 0, 0);
#pragma unroll
   for (int v = 0; v < 16; ++v) out[lane * 16 + v] = D[v];
}

// ---- launchers ------------------------------------------------------------

hipError_t launch_i8_scan(const ScanArgs &a, hipStream_t st)
{
   if (a.nTiles <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_i8_scan, dim3(a.nTiles), dim3(kScanWaves * 64), 0, st, a);
   return hipGetLastError();
}

hipError_t launch_i8_finalize(const FinalizeArgs &a, hipStream_t st)
{
   if (a.nBlocks <= 0) return hipSuccess;
   hipLaunchKernelGGL(k_i8_finalize, dim3(a.nBlocks), dim3(256), 0, st, a);
   return hipGetLastError();
}

hipError_t launch_clear_keys(uint64_t *keys, int64_t n, uint32_t *epoch, uint32_t *done, hipStream_t st)
{
   int64_t blocks = (n + 255) / 256;
   if (blocks < 1) blocks = 1;
   hipLaunchKernelGGL(k_clear_keys, dim3((unsigned)blocks), dim3(256), 0, st, keys, n, epoch, done);
   return hipGetLastError();
}

hipError_t launch_mfma_probe(int32_t *out, hipStream_t st)
{
   hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, out);
   return hipGetLastError();
}

int scan_tile_m() { return kTileM; }
int scan_tile_n() { return kTileN; }

} // namespace v3dm
