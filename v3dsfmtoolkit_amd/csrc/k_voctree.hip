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

// Level 0: one group holding every feature, in index order.
__global__ void k_vt_root_groups(int n, int32_t *__restrict__ groupStart, int32_t *__restrict__ tileStart)
{
   if (threadIdx.x == 0)
   {
      groupStart[0] = 0;
      groupStart[1] = n;
      tileStart[0] = 0;
      tileStart[1] = (n + 63) / 64;
   }
}

__global__ void k_vt_fill(int n, int32_t *__restrict__ p, int32_t v)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

namespace {
inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
} // namespace

hipError_t launch_vt_root_groups(int n, int32_t *groupStart, int32_t *tileStart, hipStream_t st)
{
   hipLaunchKernelGGL(k_vt_root_groups, dim3(1), dim3(64), 0, st, n, groupStart, tileStart);
   return hipGetLastError();
}

// One wave per workgroup: the finest unit the dispatcher can balance over the
// CUs (measured 3 % faster than four waves per workgroup; unrolling the node
// loop does not help, because scalar loads return out of order and every
// wait on one drains them all).
hipError_t launch_vt_level(int n, const float *keys, const float *tree, const int32_t *order, const int32_t *groupStart,
                           const int32_t *tileStart, int nGroups, int parentBase, int B, int last, int wordOffset,
                           int32_t *leaf, hipStream_t st, hipEvent_t start, hipEvent_t stop)
{
   // tiles <= ceil(n / 64) + non-empty groups
   const int64_t maxTiles = (n + 63) / 64 + std::min<int64_t>(nGroups, n);
   if (maxTiles <= 0) return hipSuccess;
   hipExtLaunchKernelGGL(k_vt_level, dim3((unsigned)maxTiles), dim3(64), 0, st, start, stop, 0, keys, tree,
                         order, groupStart, tileStart, nGroups, parentBase, B, last, wordOffset, leaf);
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
