// v3dm_device.h -- device helpers shared by the scan kernels (int8 and BF16).
#pragma once

#include "v3dm_internal.h"

namespace v3dm {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWaveRows = 64;              // left rows per wave (two 32-row MFMA blocks)
constexpr int kStepCols = 32;              // right rows per scan step
constexpr int kThirds = 3;                 // LDS staging ring depth (groups)
constexpr int kRsrcWord3 = 0x00020000;     // raw buffer resource, 32-bit data format

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

// Pair of a tile / finalize block: one load from the precomputed table when
// the batch provides it, else a binary search over the pairs' first indices.
__device__ inline int pair_of_tile(const int32_t *tilePair, const PairDesc *pairs, int nPairs, int tile)
{
   return tilePair ? tilePair[tile] : find_pair_by_tile(pairs, nPairs, tile);
}

__device__ inline int pair_of_fin(const int32_t *finPair, const PairDesc *pairs, int nPairs, int blk)
{
   return finPair ? finPair[blk] : find_pair_by_fin(pairs, nPairs, blk);
}

// Epipolar Sampson test, v3d_cudamatching.cu:266-290 (the kernels are
// compiled with -ffp-contract=off, so the expression is evaluated as written).
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

// Exact top-2 of unique keys; a key inserted again (the same pair scored by
// two passes) leaves both slots as they are.
__device__ inline void top2_insert(uint64_t *t1, uint64_t *t2, uint64_t key)
{
   unsigned long long old = atomicMax((unsigned long long *)t1, (unsigned long long)key);
   if (old != key) atomicMax((unsigned long long *)t2, (unsigned long long)(old < key ? old : key));
}

// The same for a left and a right key at once: both top-1 atomics are issued
// before either result is waited for (the two arrays are independent), so a
// survivor costs one atomic round trip, not two.
__device__ inline void top2_insert_both(uint64_t *t1a, uint64_t *t2a, uint64_t ka, uint64_t *t1b, uint64_t *t2b,
                                        uint64_t kb)
{
   const unsigned long long oa = atomicMax((unsigned long long *)t1a, (unsigned long long)ka);
   const unsigned long long ob = atomicMax((unsigned long long *)t1b, (unsigned long long)kb);
   if (oa != ka) atomicMax((unsigned long long *)t2a, (unsigned long long)(oa < ka ? oa : ka));
   if (ob != kb) atomicMax((unsigned long long *)t2b, (unsigned long long)(ob < kb ? ob : kb));
}

__device__ inline v4i ld16(const int8_t *p)
{
   return *(__attribute__((address_space(1))) const v4i *)(p);
}

__device__ inline int sgpr(int x) { return __builtin_amdgcn_readfirstlane(x); }

// The epoch of a job: from device memory (jobs whose kernel arguments are
// fixed, as in a replayed graph, advance it there) or passed by value (the
// host counts single-pair jobs, so their kernels need no dependent load of it
// and no last-block advance).
__device__ inline uint32_t job_epoch(const uint32_t *p, uint32_t value) { return p ? *p : value; }

__device__ inline int swz(int r) { return (r ^ (r >> 3)) & 7; }

// A kernel reads its argument block with scalar loads placed where each field
// is first used, and waits for each group before the branch that needs it: a
// chain of scalar-cache misses at launch (one per 64-byte line, ~0.5 us each).
// One scalar load per line of the block, issued together at kernel entry and
// waited for once, brings every line into the scalar cache so that the later
// loads hit.
template <typename T> __device__ __forceinline__ void touch_kernargs(const T &args)
{
   static_assert(sizeof(T) <= 512, "argument block of at most 8 lines");
   const uint32_t *p = reinterpret_cast<const uint32_t *>(&args);
   constexpr int kLines = (int)((sizeof(T) + 63) / 64);
   uint32_t w[8];
#pragma unroll
   for (int i = 0; i < kLines; ++i) w[i] = p[16 * i];
#pragma unroll
   for (int i = 0; i < kLines; ++i) asm volatile("" ::"s"(w[i]));
}

__device__ inline void lds_barrier()
{
   // LDS writes of this wave done, then the workgroup barrier.  Written out
   // so no vmcnt wait is attached: the register ring stays in flight.
   asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
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

} // namespace v3dm
