// k_voctree.hip -- vocabulary-tree quantisation of SIFT descriptors
// (V3D::VocabularyTree<Depth, BranchFactor, 128>::getVisualWord,
// V3D/include/Base/v3d_voctree.h:55-99) for many features at once.
//
// The reference descends the tree greedily: at every level it scores the
// key against the BranchFactor children of the current node with a
// sequential FP32 sum, score = sum_i key[i]*node[i] in ascending i with each
// product rounded before it is added (x86-64 code without FMA contraction),
// and keeps the first child with the largest score (strict '>', initial
// -1e30f).  The kernels here compute exactly that sum for every (feature,
// child) pair -- one v_mul_f32 and one v_add_f32 per term, in the same order
// (this file is compiled with -ffp-contract=off) -- so the visual words are
// bit-identical to the CPU tree's.
//
// Layout and schedule.  Nodes are stored as the tree file holds them
// (level by level, node j at floats [128 j, 128 j + 128)); node p's children
// are nodes (p+1)*B .. (p+1)*B + B-1 (the root's are 0 .. B-1).  One wave
// scores 64 features (one per lane, the key held in 128 VGPRs) against one
// common set of B children, so every node value is a wave-uniform scalar
// load and feeds the VALU as an SGPR operand.  Between levels the features
// are grouped by their current node with a counting sort (count, scan,
// scatter); a wave then takes up to 64 features of one group.
#include <hip/hip_ext.h>

#include <algorithm>

#include "v3dm_internal.h"

namespace v3dm {

namespace {

// Scores of one lane's key against B consecutive nodes; returns the
// reference's best child (first maximum, -1 if no score exceeds -1e30f).
__device__ inline int best_child(const float (&key)[kDim], const float *__restrict__ nodes, int B)
{
   float maxScore = -1e30f;
   int best = -1;
#pragma unroll 1
   for (int k = 0; k < B; ++k)
   {
      const float *nd = nodes + (size_t)k * kDim;
      float s = 0.0f;
#pragma unroll
      for (int i = 0; i < kDim; ++i) s = s + key[i] * nd[i];
      if (s > maxScore)
      {
         maxScore = s;
         best = k;
      }
   }
   return best;
}

} // namespace

// One tile (up to 64 features of one group) per wave, one wave per workgroup.  Group b of this level
// holds the features whose current node is parentBase + b (parentBase < 0:
// the root, one group).  leaf[f] receives the chosen child's node id, or the
// visual word (node id - wordOffset) on the last level; -1 marks a feature
// whose scores never exceed -1e30f (the reference then leaves the tree).
__global__ __launch_bounds__(64) void k_vt_level(const float *__restrict__ keys,
                                                           const float *__restrict__ tree,
                                                           const int32_t *__restrict__ order,
                                                           const int32_t *__restrict__ groupStart,
                                                           const int32_t *__restrict__ tileStart, int nGroups,
                                                           int parentBase, int B, int last, int wordOffset,
                                                           int32_t *__restrict__ leaf)
{
   const int lane = threadIdx.x;
   const int t = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
   if (t >= tileStart[nGroups]) return;
   int lo = 0, hi = nGroups; // tileStart[lo] <= t < tileStart[hi]
   while (hi - lo > 1)
   {
      const int mid = (lo + hi) >> 1;
      if (tileStart[mid] <= t) lo = mid;
      else hi = mid;
   }
   const int g = __builtin_amdgcn_readfirstlane(lo);
   const int first = groupStart[g], end = groupStart[g + 1];
   const int pos = first + (t - tileStart[g]) * 64 + lane;
   const bool valid = pos < end;
   const int p = valid ? pos : first;
   const int f = order ? order[p] : p;

   float key[kDim];
   const float4 *kp = reinterpret_cast<const float4 *>(keys + (size_t)f * kDim);
#pragma unroll
   for (int i = 0; i < kDim / 4; ++i)
   {
      const float4 v = kp[i];
      key[4 * i] = v.x;
      key[4 * i + 1] = v.y;
      key[4 * i + 2] = v.z;
      key[4 * i + 3] = v.w;
   }
   const int parent = parentBase < 0 ? -1 : parentBase + g;
   const int base = __builtin_amdgcn_readfirstlane((parent + 1) * B);
   const int best = best_child(key, tree + (size_t)base * kDim, B);
   if (!valid) return;
   int out = best < 0 ? -1 : base + best;
   if (last && out >= 0) out -= wordOffset;
   leaf[f] = out;
}

// ---- FP32-MFMA filter + exact rescoring (branch factor <= 64) -------------------
//
// One wave, 64 keys of one group, all children at once on
// v_mfma_f32_32x32x2_f32: children are the A rows (32 per block, NB blocks),
// keys the B columns (two blocks of 32).  In step t lane half h supplies
// element 64 h + t of its child row and of its key column, so after 64 steps
// accumulator register v of block (nb, kb) in lane l holds the score of child
// 32 nb + 8 (v/4) + 4 h + v%4 against key 32 kb + l%32.
//
// Bound.  Let s_c be the reference's sequential product-rounded FP32 sum for
// child c and t_c the MFMA sum.  Any summation of the 128 products with every
// product and addition rounded to FP32 (nearest or toward zero, relative error
// <= 2^-23) is within 129 * 2^-23 * sum|x_i y_i| of the exact dot product, and
// s_c within 129 * 2^-24 of it; with sum|x_i y_i| <= |x| |y_c| (|x|, |y_c|: upper
// bounds of the norms, vt_norm_bound and the stored node norms),
//     |t_c - s_c| <= E_c = 2^-14 |x| |y_c| + 2^-122 (|x| + |y_c|) + 2^-100
// (2.6x margin on the relative term; the middle term covers subnormal inputs
// the matrix core may flush, sum over flushed i of |x_i y_i| <= 2^-126 sqrt(128)
// (|x| + |y_c|), the last one subnormal products and partial sums).  With
// LB = max_c (t_c - E_c) <= max_c s_c, a child with t_c + E_c < LB has s_c below
// some other child's exact score and cannot be the reference's first maximum.
// One remaining candidate is therefore the answer; several are rescored with
// the exact sequential sum in ascending child order (strict '>', -1e30f start).
// Non-finite scores or norms, or LB <= -1e30f, make every child a candidate.

typedef float v16f __attribute__((ext_vector_type(16)));

// An upper bound of |x| from the FP32 sum of squares ss of its elements: the
// sum is within 2^-17 (relative) of |x|^2 plus 128 * 2^-126 for squares that
// underflow, so sqrt(ss) (1 + 2^-15) + 2^-59 >= |x| (keys too small for their
// squares to register still get a bound that covers their scores' errors).
__device__ inline float vt_norm_bound(float ss) { return sqrtf(ss) * (1.0f + 0x1p-15f) + 0x1p-59f; }

struct VtNearTie
{
   uint64_t mask; // candidate children
   int32_t key;   // key index
   int32_t base;  // node id of child 0
};

template <int NB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_vt_level_mfma(const float *__restrict__ keys,
                                                     const float4 *__restrict__ treeT,
                                                     const float *__restrict__ nodeNorm,
                                                     const int32_t *__restrict__ order,
                                                     const int32_t *__restrict__ groupStart,
                                                     const int32_t *__restrict__ tileStart, int nGroups,
                                                     int parentBase, int B, int last, int wordOffset,
                                                     int32_t *__restrict__ leaf, VtNearTie *__restrict__ nearTies,
                                                     uint32_t *__restrict__ nearTieCount)
{
   const int lane = threadIdx.x, h = lane >> 5, lj = lane & 31;
   const int t = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
   if (t >= tileStart[nGroups]) return;
   int lo = 0, hi = nGroups;
   while (hi - lo > 1)
   {
      const int mid = (lo + hi) >> 1;
      if (tileStart[mid] <= t) lo = mid;
      else hi = mid;
   }
   const int g = __builtin_amdgcn_readfirstlane(lo);
   const int tileFirst = groupStart[g] + (t - tileStart[g]) * 64;
   const int nk = min(64, groupStart[g + 1] - tileFirst); // keys in this tile, >= 1
   auto keyOf = [&](int j) {
      const int p = tileFirst + min(j, nk - 1);
      return order ? order[p] : p;
   };
   const int parent = parentBase < 0 ? -1 : parentBase + g;
   const int base = __builtin_amdgcn_readfirstlane((parent + 1) * B);

   // children in the interleaved copy: sibling group base / B, 4 elements x 64
   // children per 1 KiB block, so lanes 0-31 (and 32-63) read 512 contiguous
   // bytes per load; children >= B are zero there
   const float4 *grp = treeT + (size_t)(base / B) * (kDim / 4) * 64 + (size_t)(16 * h) * 64 + lj;

   const uint64_t all = B >= 64 ? ~0ull : ((1ull << B) - 1);
   uint64_t mask[2];

   // the two blocks of 32 keys one after the other (32 NB live accumulators)
#pragma unroll 1
   for (int kb = 0; kb < 2; ++kb)
   {
      // this lane's half of key 32 kb + lj, 16 elements (64 bytes) per segment
      const float4 *krow = reinterpret_cast<const float4 *>(keys + (size_t)keyOf(32 * kb + lj) * kDim + 64 * h);
      v16f acc[NB];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
         for (int v = 0; v < 16; ++v) acc[nb][v] = 0.0f;
      float ss = 0.0f;
      constexpr int kSeg = 2;                          // float4 per segment: 8 elements
      float4 kr[kSeg], nc[kSeg][NB];                   // segment being computed
      auto loadSeg = [&](int seg, float4(&k)[kSeg], float4(&n)[kSeg][NB]) {
#pragma unroll
         for (int j = 0; j < kSeg; ++j)
         {
            k[j] = krow[kSeg * seg + j];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) n[j][nb] = grp[(kSeg * seg + j) * 64 + 32 * nb];
         }
      };
      auto computeSeg = [&](const float4(&k)[kSeg], const float4(&n)[kSeg][NB]) {
#pragma unroll
         for (int j = 0; j < kSeg; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
               const float kv = q == 0 ? k[j].x : q == 1 ? k[j].y : q == 2 ? k[j].z : k[j].w;
               ss = fmaf(kv, kv, ss);
#pragma unroll
               for (int nb = 0; nb < NB; ++nb)
               {
                  const float nv = q == 0 ? n[j][nb].x : q == 1 ? n[j][nb].y : q == 2 ? n[j][nb].z : n[j][nb].w;
                  acc[nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(nv, kv, acc[nb], 0, 0, 0);
               }
            }
      };
      // Segments 64 h + 4 kSeg seg .., two at a time through two register sets
      // with the next segment always in flight; the load into the first set
      // targets registers already consumed, so no set is ever copied (a rolled
      // loop's run-time "is there a next segment" test made the compiler copy
      // every prefetched register on each iteration).
      constexpr int kSegs = 16 / kSeg;
      static_assert(kSegs % 2 == 0, "segments go in pairs");
      float4 kn[kSeg], nn[kSeg][NB];
      loadSeg(0, kr, nc);
#pragma unroll 1
      for (int seg = 0; seg < kSegs; seg += 2)
      {
         loadSeg(seg + 1, kn, nn);
         computeSeg(kr, nc);
         asm volatile("" : "+v"(ss)); // keeps the sum of squares of kr here, not sunk past the load below
         if (seg + 2 < kSegs) loadSeg(seg + 2, kr, nc);
         computeSeg(kn, nn);
      }

      // candidate mask of key 32 kb + lj (both lane halves end with the same mask):
      // first the lower bounds t_c - E_c, turning each accumulator into its upper
      // bound t_c + E_c on the way (-inf for padding children), then the mask
      const float nx = vt_norm_bound(ss + __shfl_xor(ss, 32, 64));
      float lbs = -INFINITY;
      bool bad = !isfinite(nx);
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
         for (int v = 0; v < 16; ++v)
         {
            const int c = 32 * nb + 8 * (v >> 2) + 4 * h + (v & 3);
            const float tc = acc[nb][v];
            if (c < B)
            {
               const float nyc = nodeNorm[base + c];
               const float e = 0x1p-14f * nx * nyc + 0x1p-122f * (nx + nyc) + 0x1p-100f;
               bad = bad || !isfinite(tc) || !isfinite(e);
               lbs = fmaxf(lbs, tc - e);
               acc[nb][v] = tc + e;
            }
            else
               acc[nb][v] = -INFINITY;
         }
      const float lb = fmaxf(lbs, __shfl_xor(lbs, 32, 64));
      uint64_t m = 0;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
         for (int v = 0; v < 16; ++v)
            if (acc[nb][v] >= lb) m |= 1ull << (32 * nb + 8 * (v >> 2) + 4 * h + (v & 3));
      const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)(m >> 32);
      m |= (uint64_t)(uint32_t)__shfl_xor((int)mlo, 32, 64) | ((uint64_t)(uint32_t)__shfl_xor((int)mhi, 32, 64) << 32);
      bad = __shfl_xor((int)bad, 32, 64) || bad;
      if (bad || !(lb > -1e30f)) m = all;
      mask[kb] = m;
   }

   // lane L now decides key L (= 32 h + lj): one candidate is the answer; the
   // others go to the near-tie list that k_vt_rescore decides exactly
   const uint64_t mine = h ? mask[1] : mask[0];
   const int f = keyOf(lane);
   const bool live = lane < nk;
   const bool unique = __popcll(mine) == 1;
   const uint64_t tied = __ballot(live && !unique);
   uint32_t slot0 = 0;
   if (tied)
   {
      if (lane == 0) slot0 = atomicAdd(nearTieCount, (uint32_t)__popcll(tied));
      slot0 = __shfl(slot0, 0, 64);
   }
   if (!live) return;
   if (!unique)
   {
      const uint32_t slot = slot0 + (uint32_t)__popcll(tied & ((1ull << lane) - 1));
      nearTies[slot] = VtNearTie{mine, f, base};
      return;
   }
   int out = base + __builtin_ctzll(mine);
   if (last) out -= wordOffset;
   leaf[f] = out;
}

// Exact decision of the keys the filter left with several candidates: the
// reference's sequential sums over the candidate children only, ascending
// child order, strict '>' from -1e30f.  Grid-stride over the list.
__global__ __launch_bounds__(64) void k_vt_rescore(const float *__restrict__ keys, const float *__restrict__ tree,
                                                  const VtNearTie *__restrict__ nearTies,
                                                  const uint32_t *__restrict__ nearTieCount, int last, int wordOffset,
                                                  int32_t *__restrict__ leaf)
{
   const uint32_t count = *nearTieCount;
   for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < count; i += gridDim.x * 64)
   {
      const VtNearTie e = nearTies[i];
      // few waves run this kernel: every load of a row is issued at once
      float4 k4[kDim / 4];
      const float4 *kp = reinterpret_cast<const float4 *>(keys + (size_t)e.key * kDim);
#pragma unroll
      for (int j = 0; j < kDim / 4; ++j) k4[j] = kp[j];
      float maxScore = -1e30f;
      int best = -1;
      for (uint64_t m = e.mask; m; m &= m - 1)
      {
         const int c = __builtin_ctzll(m);
         const float4 *np = reinterpret_cast<const float4 *>(tree + (size_t)(e.base + c) * kDim);
         float4 n4[kDim / 4];
#pragma unroll
         for (int j = 0; j < kDim / 4; ++j) n4[j] = np[j];
         float sc = 0.0f;
#pragma unroll
         for (int j = 0; j < kDim / 4; ++j)
         {
            sc = sc + k4[j].x * n4[j].x;
            sc = sc + k4[j].y * n4[j].y;
            sc = sc + k4[j].z * n4[j].z;
            sc = sc + k4[j].w * n4[j].w;
         }
         if (sc > maxScore)
         {
            maxScore = sc;
            best = c;
         }
      }
      int out = best < 0 ? -1 : e.base + best;
      if (last && out >= 0) out -= wordOffset;
      leaf[e.key] = out;
   }
}

// Group sizes of the next level: node id leaf[f] - parentBase for live features
// (direct global atomics: used when the groups outnumber the LDS bins below,
// where contention per group is low anyway).
__global__ void k_vt_count(int n, const int32_t *__restrict__ leaf, int parentBase, int nGroups,
                           int32_t *__restrict__ counts)
{
   const int f = blockIdx.x * blockDim.x + threadIdx.x;
   if (f >= n) return;
   const int g = leaf[f] - parentBase;
   if (g >= 0 && g < nGroups) atomicAdd(&counts[g], 1);
}

// Exclusive scans of the group sizes and of their tile counts (ceil(size/64));
// cursor = group starts for the scatter.  One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_vt_scan(int nGroups, const int32_t *__restrict__ counts,
                                                 int32_t *__restrict__ groupStart, int32_t *__restrict__ tileStart,
                                                 int32_t *__restrict__ cursor)
{
   __shared__ int32_t sCount[1024], sTiles[1024];
   __shared__ int32_t carryCount, carryTiles;
   if (threadIdx.x == 0)
   {
      carryCount = 0;
      carryTiles = 0;
   }
   __syncthreads();
   for (int base = 0; base < nGroups; base += 1024)
   {
      const int g = base + (int)threadIdx.x;
      const int c = g < nGroups ? counts[g] : 0;
      sCount[threadIdx.x] = c;
      sTiles[threadIdx.x] = (c + 63) / 64;
      __syncthreads();
      for (int off = 1; off < 1024; off <<= 1) // Hillis-Steele inclusive scan
      {
         const int a = threadIdx.x >= (unsigned)off ? sCount[threadIdx.x - off] : 0;
         const int b = threadIdx.x >= (unsigned)off ? sTiles[threadIdx.x - off] : 0;
         __syncthreads();
         sCount[threadIdx.x] += a;
         sTiles[threadIdx.x] += b;
         __syncthreads();
      }
      if (g < nGroups)
      {
         const int gs = carryCount + sCount[threadIdx.x] - c;
         groupStart[g] = gs;
         cursor[g] = gs;
         tileStart[g] = carryTiles + sTiles[threadIdx.x] - (c + 63) / 64;
      }
      __syncthreads();
      if (threadIdx.x == 1023)
      {
         carryCount += sCount[1023];
         carryTiles += sTiles[1023];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0)
   {
      groupStart[nGroups] = carryCount;
      tileStart[nGroups] = carryTiles;
   }
}

__global__ void k_vt_scatter(int n, const int32_t *__restrict__ leaf, int parentBase, int nGroups,
                             int32_t *__restrict__ cursor, int32_t *__restrict__ order)
{
   const int f = blockIdx.x * blockDim.x + threadIdx.x;
   if (f >= n) return;
   const int g = leaf[f] - parentBase;
   if (g >= 0 && g < nGroups) order[atomicAdd(&cursor[g], 1)] = f;
}

// The same two steps with the histogram privatised in LDS: a workgroup takes
// kVtChunk consecutive features, counts them per group in LDS and adds each
// non-empty bin to the global count once (one global atomic per non-empty
// group per workgroup instead of one per feature: the root level has only
// B groups).  The scatter reserves each bin's range with one global atomic
// and places every feature at its LDS rank inside it.
constexpr int kVtChunkThreads = 256, kVtPerThread = 8, kVtChunk = kVtChunkThreads * kVtPerThread;
constexpr int kVtLdsBins = 8192;

__global__ __launch_bounds__(kVtChunkThreads) void k_vt_count_lds(int n, const int32_t *__restrict__ leaf,
                                                                 int parentBase, int nGroups,
                                                                 int32_t *__restrict__ counts)
{
   __shared__ int32_t bins[kVtLdsBins];
   for (int b = threadIdx.x; b < nGroups; b += kVtChunkThreads) bins[b] = 0;
   __syncthreads();
   const int f0 = blockIdx.x * kVtChunk;
   for (int j = 0; j < kVtPerThread; ++j)
   {
      const int f = f0 + j * kVtChunkThreads + (int)threadIdx.x;
      if (f < n)
      {
         const int g = leaf[f] - parentBase;
         if (g >= 0 && g < nGroups) atomicAdd(&bins[g], 1);
      }
   }
   __syncthreads();
   for (int b = threadIdx.x; b < nGroups; b += kVtChunkThreads)
      if (bins[b]) atomicAdd(&counts[b], bins[b]);
}

__global__ __launch_bounds__(kVtChunkThreads) void k_vt_scatter_lds(int n, const int32_t *__restrict__ leaf,
                                                                   int parentBase, int nGroups,
                                                                   int32_t *__restrict__ cursor,
                                                                   int32_t *__restrict__ order)
{
   __shared__ int32_t bins[kVtLdsBins];
   for (int b = threadIdx.x; b < nGroups; b += kVtChunkThreads) bins[b] = 0;
   __syncthreads();
   const int f0 = blockIdx.x * kVtChunk;
   int g[kVtPerThread], rank[kVtPerThread];
#pragma unroll
   for (int j = 0; j < kVtPerThread; ++j)
   {
      const int f = f0 + j * kVtChunkThreads + (int)threadIdx.x;
      g[j] = -1;
      if (f < n)
      {
         const int gg = leaf[f] - parentBase;
         if (gg >= 0 && gg < nGroups)
         {
            g[j] = gg;
            rank[j] = atomicAdd(&bins[gg], 1);
         }
      }
   }
   __syncthreads();
   for (int b = threadIdx.x; b < nGroups; b += kVtChunkThreads)
      if (bins[b]) bins[b] = atomicAdd(&cursor[b], bins[b]); // the bin's first slot
   __syncthreads();
#pragma unroll
   for (int j = 0; j < kVtPerThread; ++j)
      if (g[j] >= 0) order[bins[g[j]] + rank[j]] = f0 + j * kVtChunkThreads + (int)threadIdx.x;
}

// Level 0's single group, and the per-level near-tie counters cleared.
__global__ void k_vt_root_groups(int n, int32_t *__restrict__ groupStart, int32_t *__restrict__ tileStart,
                                 uint32_t *__restrict__ nearTieCounts, int depth)
{
   if (threadIdx.x == 0)
   {
      groupStart[0] = 0;
      groupStart[1] = n;
      tileStart[0] = 0;
      tileStart[1] = (n + 63) / 64;
   }
   if ((int)threadIdx.x < depth) nearTieCounts[threadIdx.x] = 0;
}

__global__ void k_vt_fill(int n, int32_t *__restrict__ p, int32_t v)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

namespace {
inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
} // namespace

hipError_t launch_vt_root_groups(int n, int32_t *groupStart, int32_t *tileStart, uint32_t *nearTieCounts, int depth,
                                 hipStream_t st)
{
   hipLaunchKernelGGL(k_vt_root_groups, dim3(1), dim3(64), 0, st, n, groupStart, tileStart, nearTieCounts, depth);
   return hipGetLastError();
}

// One wave per workgroup: the finest unit the dispatcher can balance over the
// CUs (measured 3 % faster than four waves per workgroup; unrolling the node
// loop does not help, because scalar loads return out of order and every
// wait on one drains them all).
// exact: 1 = the all-VALU kernel even where the MFMA filter applies (B <= 64).
size_t vt_near_tie_bytes(int64_t n) { return sizeof(VtNearTie) * (size_t)std::max<int64_t>(n, 1); }

// nearTieCount must be zero on entry (one counter per level); nearTies holds n entries.
hipError_t launch_vt_level(int n, const float *keys, const float *tree, const float *treeT, const float *nodeNorm,
                           const int32_t *order, const int32_t *groupStart, const int32_t *tileStart, int nGroups,
                           int parentBase, int B, int last, int wordOffset, int32_t *leaf, int exact, void *nearTies,
                           uint32_t *nearTieCount, hipStream_t st, hipEvent_t start, hipEvent_t stop)
{
   // tiles <= ceil(n / 64) + non-empty groups
   const int64_t maxTiles = (n + 63) / 64 + std::min<int64_t>(nGroups, n);
   if (maxTiles <= 0) return hipSuccess;
   const dim3 grid((unsigned)maxTiles), block(64);
   VtNearTie *ties = static_cast<VtNearTie *>(nearTies);
   const float4 *tT = reinterpret_cast<const float4 *>(treeT);
   if (!exact && treeT && B <= 64)
   {
      if (B <= 32)
         hipExtLaunchKernelGGL(k_vt_level_mfma<1>, grid, block, 0, st, start, stop, 0, keys, tT, nodeNorm, order,
                               groupStart, tileStart, nGroups, parentBase, B, last, wordOffset, leaf, ties,
                               nearTieCount);
      else
         hipExtLaunchKernelGGL(k_vt_level_mfma<2>, grid, block, 0, st, start, stop, 0, keys, tT, nodeNorm, order,
                               groupStart, tileStart, nGroups, parentBase, B, last, wordOffset, leaf, ties,
                               nearTieCount);
      const unsigned rescoreBlocks = (unsigned)std::min<int64_t>((n + 63) / 64, 512);
      hipLaunchKernelGGL(k_vt_rescore, dim3(rescoreBlocks), dim3(64), 0, st, keys, tree, ties, nearTieCount, last,
                         wordOffset, leaf);
   }
   else
      hipExtLaunchKernelGGL(k_vt_level, grid, block, 0, st, start, stop, 0, keys, tree, order, groupStart, tileStart,
                            nGroups, parentBase, B, last, wordOffset, leaf);
   return hipGetLastError();
}

hipError_t launch_vt_regroup(int n, const int32_t *leaf, int parentBase, int nGroups, int32_t *counts,
                             int32_t *groupStart, int32_t *tileStart, int32_t *cursor, int32_t *order, hipStream_t st)
{
   const bool lds = nGroups <= kVtLdsBins;
   hipLaunchKernelGGL(k_vt_fill, dim3(blocks(nGroups, 256)), dim3(256), 0, st, nGroups, counts, 0);
   if (n > 0 && lds)
      hipLaunchKernelGGL(k_vt_count_lds, dim3(blocks(n, kVtChunk)), dim3(kVtChunkThreads), 0, st, n, leaf, parentBase,
                         nGroups, counts);
   else if (n > 0)
      hipLaunchKernelGGL(k_vt_count, dim3(blocks(n, 256)), dim3(256), 0, st, n, leaf, parentBase, nGroups, counts);
   hipLaunchKernelGGL(k_vt_scan, dim3(1), dim3(1024), 0, st, nGroups, counts, groupStart, tileStart, cursor);
   if (n > 0 && lds)
      hipLaunchKernelGGL(k_vt_scatter_lds, dim3(blocks(n, kVtChunk)), dim3(kVtChunkThreads), 0, st, n, leaf,
                         parentBase, nGroups, cursor, order);
   else if (n > 0)
      hipLaunchKernelGGL(k_vt_scatter, dim3(blocks(n, 256)), dim3(256), 0, st, n, leaf, parentBase, nGroups, cursor,
                         order);
   return hipGetLastError();
}

} // namespace v3dm
