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
//  * minRatio == 0: the ratio (2-2b)/max(0.001, 2-2s) has a positive
//    denominator whatever s is, so it is below 0 exactly when b > 1; a row is
//    accepted iff b > max(minScore, 1), which t = floor(max(minScore, 1) 2^18)
//    keeps, and s plays no part.  minRatio (or minScore) NaN: nothing is
//    accepted, t = 2^28 (no score passes).
//  * a positive minRatio whose T is below 0: no filtering (t = -1; scores are
//    >= 0); the automatic kernel choice then takes the plain path.
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
#include <algorithm>
#include <atomic>
#include <type_traits>

#include <hip/hip_ext.h>

#include "v3dm_device.h"

namespace v3dm {

typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kBias = 1 << 21;      // 128^3

__device__ __forceinline__ void survivor(const ScanArgs &a, const PairDesc &pd, uint32_t epoch, int il, int jr,
                                         int32_t score)
{
   if (il >= pd.nL || jr >= pd.nR) return;
   // rows of a gathered pool stand for the features mapL / mapR name
   if (a.mapL) il = a.mapL[il];
   if (a.mapR) jr = a.mapR[jr];
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
   top2_insert_both(a.top1L + pd.slotL + il, a.top2L + pd.slotL + il, pack_key(epoch, (uint32_t)score, (uint32_t)jr),
                    a.top1R + pd.slotR + jr, a.top2R + pd.slotR + jr, pack_key(epoch, (uint32_t)score, (uint32_t)il));
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

// Operand packing.  Descriptor byte e of a row goes to lane half h = e / 64,
// MFMA step s = (e % 64) / 16, byte e % 16 of the lane's 16-byte operand.
// A (left) lane l carries row l % 32, B (right) lane l carries row l % 32, so
// every MFMA step pairs the same 32 descriptor bytes of both rows and four
// steps cover all 128.  Accumulator register v of lane l is
// D[8*(v/4) + 4*(l/32) + v%4][l%32].
//
// Tiling.  A workgroup of 4 (or 8) waves owns 256 (512) left rows, 64 per
// wave held in VGPRs as MFMA A operands for the whole tile, and sweeps
// TILE_N right rows in steps of 32; two waves share each SIMD (256 VGPRs
// each), so one computes while the other waits.  Ring form (k_i8_scan): the
// right rows are staged through LDS once per workgroup and read by all its
// waves, in groups of 4 steps (16 KiB) with one barrier per group; within a
// group the waves run free.  Whole-tile form (k_i8_scan_whole): both sides of
// a 512 x 512 tile are staged at once and the steps run with no barrier.  In
// both, each wave reads the fragments of step s+1 and issues its MFMAs before
// it tests the scores of step s (two accumulator sets), so the test overlaps
// the matrix pipeline.
// LDS unit layout: 16-byte unit p = row * 8 + (chunk ^ swz(row)),
// swz(r) = (r ^ r >> 3) & 7, so the 32 rows of one fragment read (same
// chunk) fall in distinct bank groups for 8- and 16-lane LDS service.

constexpr int kStageBytes = kStepCols * kDim;      // 4 KiB
constexpr int kGroup = 4;                          // steps per barrier
constexpr int kGroupBytes = kGroup * kStageBytes;  // 16 KiB

// Rare path of a scan step: every score of the step above lim goes to the
// wave's survivor region as (score, row within the wave's 64, column within
// the tile); padding rows and columns never pass, and 4-register blocks with
// no passing score are skipped whole.  Entries past cap are counted, not
// written (the region then overflows and is recomputed).  half: lane / 32;
// jrel: this lane's column within the tile.
__device__ __forceinline__ void record_survivors(const v16i &acc0, const v16i &acc1, int lim, int cs, int jrel,
                                                 int half, int cap, uint64_t *region, int &nsurv)
{
   int half4 = 4 * half;
   asm volatile("" : "+v"(half4));
#pragma unroll
   for (int bb = 0; bb < 2; ++bb)
   {
      const v16i &acc = bb == 0 ? acc0 : acc1;
#pragma unroll
      for (int q = 0; q < 4; ++q)
      {
         const int mq = max(max(acc[4 * q], acc[4 * q + 1]), max(acc[4 * q + 2], acc[4 * q + 3]));
         if (!__any(mq > lim)) continue;
#pragma unroll
         for (int e = 0; e < 4; ++e)
         {
            const int v = 4 * q + e;
            const bool h = acc[v] > lim;
            const uint64_t ball = __ballot(h);
            if (ball == 0) continue;
            const int irel = 32 * bb + 8 * q + half4 + e;
            const int slot =
               nsurv + __builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0));
            if (h && slot < cap)
               region[slot] = ((uint64_t)(uint32_t)(acc[v] + cs) << 32) | ((uint32_t)irel << 16) | (uint32_t)jrel;
            nsurv += __builtin_popcountll(ball);
         }
      }
   }
}

// Tile end of a scan that applies its own survivors: the wave's region
// (nsurv > 0 entries; the wave's 64 rows from l0, the tile's columns from c0;
// in global memory, or in LDS for kRegionInLds) is inserted into the top-2
// keys, or, when it overflowed, recomputed exactly from the descriptors with
// u8 dot products.
template <bool kRegionInLds = false>
__device__ __forceinline__ void apply_region(const ScanArgs &a, const PairDesc &pd, const uint64_t *region, int nsurv,
                                             int l0, int c0, int tileN, int lane)
{
   if constexpr (!kRegionInLds) asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); // the region's stores
   const uint32_t epoch = job_epoch(a.epoch, a.epochValue);
   if (nsurv <= a.survCap)
   {
      for (int e = lane; e < nsurv; e += 64)
      {
         const uint64_t w = region[e];
         survivor(a, pd, epoch, l0 + (int)((w >> 16) & 0xFFFF), c0 + (int)(w & 0xFFFF), (int32_t)(w >> 32));
      }
   }
   else
   {
      const int nRows = min(kWaveRows, pd.nL - l0), nCols = min(tileN, pd.nR - c0);
      for (int e = lane; e < nRows * nCols; e += 64)
      {
         const int il = l0 + e / nCols, jr = c0 + e % nCols;
         const uint32_t *x = reinterpret_cast<const uint32_t *>(a.L + (size_t)(pd.offL + il) * kDim);
         const uint32_t *y = reinterpret_cast<const uint32_t *>(a.R + (size_t)(pd.offR + jr) * kDim);
         uint32_t acc = 0;
         for (int k = 0; k < kDim / 4; ++k) // stored bytes are u ^ 0x80
            acc = __builtin_amdgcn_udot4(x[k] ^ 0x80808080u, y[k] ^ 0x80808080u, acc, false);
         if ((int32_t)acc > a.thr) survivor(a, pd, epoch, il, jr, (int32_t)acc);
      }
   }
}

// Right rows reach LDS by direct-to-LDS loads (global_load_lds_dwordx4: each
// lane's 16 bytes land at LDS base + 16 * lane), so no register ring and no
// LDS stores: the lane that loads LDS unit p fetches the global 16 bytes the
// swizzled layout puts there.  Four 16 KiB buffers per workgroup (distinct
// arrays, each named at compile time: the group loop runs in fours); group
// g+3 is loaded right after the barrier of group g into the buffer group g-1
// used.  One workgroup per tile, or (a.persistent) resident workgroups that
// take tiles blockIdx.x, then gridDim.x + atomicAdd(tileCounter) until the
// tiles run out; the next index is requested when a tile starts.
// At that barrier a wave waits until only its loads of group g+2 may be in
// flight (the group g+1 data must be there: the last step of group g issues
// the MFMAs of group g+1's first step), then the barrier makes every wave's
// loads visible.
constexpr int kBufs = 4;

template <int N> __device__ __forceinline__ void vm_wait_barrier()
{
   asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

// Direct-to-LDS loads, written out so the compiler's wait placement does not
// see them: it cannot tell, after the survivor branch rejoins, which buffer a
// pending load fills, and would wait for all of them before every LDS read.
// The waits these loads need are the explicit ones at the group barriers;
// the compiler's own vmcnt waits (for its own loads) only get stricter, as
// vector memory operations complete in issue order.  M0 holds the
// wave-uniform LDS address (one wait state after it is written) and gets its
// previous value back afterwards; lane l writes bytes [16 l, 16 l + 16)
// (dwordx4) or [4 l, 4 l + 4) (dword) there.
__device__ __forceinline__ void lds_dma16(const void *base, int voff, uint32_t ldsAddr)
{
   uint32_t saved;
   asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %3\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2\n\t"
                "s_nop 0\n\ts_mov_b32 m0, %0"
                : "=&s"(saved)
                : "v"(voff), "s"(base), "s"(ldsAddr)
                : "memory");
}
__device__ __forceinline__ void lds_dma4(const void *base, int voff, uint32_t ldsAddr)
{
   uint32_t saved;
   asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %3\n\ts_nop 0\n\tglobal_load_lds_dword %1, %2\n\t"
                "s_nop 0\n\ts_mov_b32 m0, %0"
                : "=&s"(saved)
                : "v"(voff), "s"(base), "s"(ldsAddr)
                : "memory");
}
__device__ __forceinline__ uint32_t lds_addr(const void *p)
{
   return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
// a wave-uniform pointer in scalar registers (the asm operands above need SGPRs)
__device__ __forceinline__ const void *uniform_ptr(const void *p)
{
   const uint64_t u = (uint64_t)(uintptr_t)p;
   return (const void *)(uintptr_t)(((uint64_t)(uint32_t)sgpr((int)(u >> 32)) << 32) | (uint32_t)sgpr((int)u));
}

template <int WAVES, int TILE_N, bool INLINE_INSERT>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(2))) void k_i8_scan(ScanArgs a)
{
   constexpr int kScanWaves = WAVES;
   constexpr int kTileM = WAVES * kWaveRows;
   constexpr int kTileN = TILE_N;
   constexpr int kDmaB = kGroupBytes / 1024 / WAVES;             // 16-byte loads of a group's rows per wave
   constexpr int kDmaC = kGroup * kStepCols * 4 / 256;            // 4-byte loads of a group's column terms
   constexpr int kDma = kDmaB + kDmaC;                            // loads per wave per group
   static_assert(kGroupBytes % (1024 * WAVES) == 0, "a group must split into whole 1 KiB loads per wave");
   __shared__ __attribute__((aligned(16))) int8_t sB0[kGroupBytes];
   __shared__ __attribute__((aligned(16))) int8_t sB1[kGroupBytes];
   __shared__ __attribute__((aligned(16))) int8_t sB2[kGroupBytes];
   __shared__ __attribute__((aligned(16))) int8_t sB3[kGroupBytes];
   __shared__ __attribute__((aligned(16))) int32_t sC0[kGroup * kStepCols];
   __shared__ __attribute__((aligned(16))) int32_t sC1[kGroup * kStepCols];
   __shared__ __attribute__((aligned(16))) int32_t sC2[kGroup * kStepCols];
   __shared__ __attribute__((aligned(16))) int32_t sC3[kGroup * kStepCols];
   __shared__ int sNextTile;
   touch_kernargs(a);
   auto bufB = [&](auto bc) -> int8_t * {
      constexpr int b = decltype(bc)::value;
      if constexpr (b == 0) return sB0;
      else if constexpr (b == 1) return sB1;
      else if constexpr (b == 2) return sB2;
      else return sB3;
   };
   auto bufC = [&](auto bc) -> int32_t * {
      constexpr int b = decltype(bc)::value;
      if constexpr (b == 0) return sC0;
      else if constexpr (b == 1) return sC1;
      else if constexpr (b == 2) return sC2;
      else return sC3;
   };

   const int tid = threadIdx.x;
   const int lane = tid & 63;
   const int wave = tid >> 6;
   // this lane's global offset (within a group) of each of its row loads
   int voffd[kDmaB];
#pragma unroll
   for (int j = 0; j < kDmaB; ++j)
   {
      const int p = (wave * kDmaB + j) * 64 + lane, st = p >> 8, pu = p & 255, r = pu >> 3, x = pu & 7;
      voffd[j] = st * kStageBytes + r * kDim + ((x ^ swz(r)) << 4);
   }
   int tile = blockIdx.x;
   while (tile < a.nTiles)
   {
   uint32_t nextReq = 0;
   if (a.persistent && tid == 0) nextReq = atomicAdd(a.tileCounter, 1u) + gridDim.x;
   PairDesc pd = a.pair0; // single pair: straight from the kernel arguments
   if (a.nPairs > 1) pd = a.pairs[sgpr(pair_of_tile(a.tilePair, a.pairs, a.nPairs, tile))];
   else if (a.pairsOnDevice) pd = a.pairs[0];
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
   if (a.pairsOnDevice && t >= sgpr(pd.tilesM) * pd.tilesN) break; // past the job's own grid (never persistent)
   const int tm = t / pd.tilesN, tn = t - tm * pd.tilesN;
   const int nLpad = (pd.nL + kRowPad - 1) / kRowPad * kRowPad;
   const int nRpad = (pd.nR + kRowPad - 1) / kRowPad * kRowPad;
   const int l0 = sgpr(tm * kTileM + wave * kWaveRows);
   const bool active = l0 < nLpad; // wave-uniform; inactive waves still load and sync
   const int32_t thr = a.thr;
   const int half = lane >> 5, lrow = lane & 31;

   const int c0 = sgpr(tn * kTileN);
   const int nsteps = sgpr((min(c0 + kTileN, nRpad) - c0) / kStepCols); // multiple of kGroup
   const int ngroups = nsteps / kGroup;
   const int8_t *Rt = a.R + (size_t)(pd.offR + c0) * kDim;
   const int32_t *St = a.Rsum + pd.offR + c0;
   uint64_t *region = a.surv + ((size_t)tile * kScanWaves + wave) * a.survCap;
   int nsurv = 0; // survivors this wave found in this tile (wave-uniform)

   // group g's rows and column terms into buffer bc (nothing past the tile)
   auto dma = [&](auto bc, int g) {
      if (g >= ngroups) return;
      const uint32_t db = lds_addr(bufB(bc)), dc = lds_addr(bufC(bc));
      const void *src = uniform_ptr(Rt + (size_t)g * kGroupBytes);
#pragma unroll
      for (int j = 0; j < kDmaB; ++j) lds_dma16(src, voffd[j], (uint32_t)sgpr((int)(db + (wave * kDmaB + j) * 1024)));
      const void *srcs = uniform_ptr(St + g * (kGroup * kStepCols));
#pragma unroll
      for (int j = 0; j < kDmaC; ++j) lds_dma4(srcs, 4 * (64 * j + lane), (uint32_t)sgpr((int)(dc + 256 * j)));
   };
   const std::integral_constant<int, 0> B0;
   const std::integral_constant<int, 1> B1;
   const std::integral_constant<int, 2> B2;
   const std::integral_constant<int, 3> B3;

   // groups 0 and 1 and the left rows; group 2 follows the first MFMAs (the
   // compiler's wait for the left rows before them must not include it)
   dma(B0, 0);
   dma(B1, 1);
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
            const v4i rs = *(__attribute__((address_space(1))) const v4i *)(a.Lsum + pd.offL + l0 + 32 * b + 8 * q +
                                                                            4 * half);
#pragma unroll
            for (int e = 0; e < 4; ++e) C[b][4 * q + e] = rs[e] - kBias;
         }
      }
   }
   int fragoff[4]; // this lane's fragment units within a stage
#pragma unroll
   for (int k = 0; k < 4; ++k) fragoff[k] = (lrow * 8 + ((4 * half + k) ^ swz(lrow))) << 4;

   auto mfma_step = [&](const v4i(&Bf)[4], v16i &x0, v16i &x1) {
      x0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][0], Bf[0], C[0], 0, 0, 0);
      x1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][0], Bf[0], C[1], 0, 0, 0);
#pragma unroll
      for (int k = 1; k < 4; ++k)
      {
         x0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][k], Bf[k], x0, 0, 0, 0);
         x1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][k], Bf[k], x1, 0, 0, 0);
      }
   };
   v16i p0, p1; // scores of the step in flight
   // `lastc` says at compile time whether g is the tile's last group (only its
   // last step issues no next-step MFMAs)
   auto group = [&](auto bc, int g, auto lastc) {
      constexpr int b = decltype(bc)::value;
      constexpr bool kLast = decltype(lastc)::value;
      const std::integral_constant<int, (b + 1) % kBufs> bn;
      const std::integral_constant<int, (b + 3) % kBufs> bf;
      // groups <= g+1 (and at g = 0 the left rows) in LDS / registers: only
      // group g+2's loads may still be in flight (none was issued past the
      // tile's end, and group 2 is issued after the barrier of group 0), then
      // the barrier
      if (g > 0 && g + 2 < ngroups) vm_wait_barrier<kDma>();
      else vm_wait_barrier<0>();
      const int8_t *sb = bufB(bc);
      const int32_t *sc = bufC(bc);
      if (g == 0)
      {
         if (active)
         {
            v4i Bz[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Bz[k] = *reinterpret_cast<const v4i *>(sb + fragoff[k]);
            mfma_step(Bz, p0, p1);
         }
         dma(B2, 2);
      }
      dma(bf, g + 3); // into the buffer of group g-1, which every wave has finished
      if (!active) return;
#pragma unroll
      for (int s = 0; s < kGroup; ++s)
      {
         v4i Bn[4];
         const bool more = s + 1 < kGroup || !kLast; // constant once unrolled
         if (s + 1 < kGroup)
         {
#pragma unroll
            for (int k = 0; k < 4; ++k) Bn[k] = *reinterpret_cast<const v4i *>(sb + (s + 1) * kStageBytes + fragoff[k]);
         }
         else if (more)
         {
            const int8_t *sn = bufB(bn);
#pragma unroll
            for (int k = 0; k < 4; ++k) Bn[k] = *reinterpret_cast<const v4i *>(sn + fragoff[k]);
         }
         const v16i acc0 = p0, acc1 = p1;
         if (more) mfma_step(Bn, p0, p1);
         const int cs = sc[s * kStepCols + lrow];
         const int m = max(max16(acc0), max16(acc1));
         const int lim = thr - cs; // score > thr  <=>  acc > thr - 128*sum(u_right)
         if (__builtin_expect(__any(m > lim), 0))
            record_survivors(acc0, acc1, lim, cs, (g * kGroup + s) * kStepCols + lrow, half, a.survCap, region,
                             nsurv);
      }
   };
   const std::false_type kMore;
   const std::true_type kEnd;
   // groups in fours (buffer = group mod 4 at compile time); the tile's last
   // group is peeled
   int g = 0;
   for (; g + 4 < ngroups; g += 4)
   {
      group(B0, g, kMore);
      group(B1, g + 1, kMore);
      group(B2, g + 2, kMore);
      group(B3, g + 3, kMore);
   }
   // 1 to 4 groups left, the last of them ends the tile
   const int left = ngroups - g;
   if (left == 4)
   {
      group(B0, g, kMore);
      group(B1, g + 1, kMore);
      group(B2, g + 2, kMore);
      group(B3, g + 3, kEnd);
   }
   else if (left == 3)
   {
      group(B0, g, kMore);
      group(B1, g + 1, kMore);
      group(B2, g + 2, kEnd);
   }
   else if (left == 2)
   {
      group(B0, g, kMore);
      group(B1, g + 1, kEnd);
   }
   else if (left == 1)
      group(B0, g, kEnd);
   if constexpr (!INLINE_INSERT)
   {
      if (lane == 0) a.survCount[(size_t)tile * kScanWaves + wave] = nsurv;
   }
   else if (nsurv > 0)
      apply_region(a, pd, region, nsurv, l0, c0, kTileN, lane);
   if (!a.persistent) break;
   __syncthreads(); // every wave is done with this tile's buffers (no load of it is in flight)
   if (tid == 0) sNextTile = (int)nextReq;
   __syncthreads();
   tile = sNextTile;
   }
}

// Whole-tile form: a workgroup of 8 waves scores a 512 x 512 tile with all of
// the tile's left and right rows and their row terms staged in LDS at once
// (132 KiB: one workgroup per CU).  Every load of the tile is issued at its
// start and waited for once; the 16 steps then run with no barrier between
// them, so a wave that takes the survivor path holds up no other wave, and no
// step waits for a later group's rows to arrive.  Both sides use the ring's
// swizzled 16-byte unit layout (32-row stages), so the left fragments are read
// from LDS as the right ones are.  Tiles as in k_i8_scan (one per workgroup,
// or resident workgroups pulling tiles from the queue).
constexpr int kWholeWaves = 8;
constexpr int kWholeTile = kWholeWaves * kWaveRows; // 512 rows by 512 columns

template <bool INLINE_INSERT>
__global__ __launch_bounds__(kWholeWaves * 64) __attribute__((amdgpu_waves_per_eu(2))) void k_i8_scan_whole(
   ScanArgs a)
{
   constexpr int kTileM = kWholeTile, kTileN = kWholeTile;
   constexpr int kStages = kTileN / kStepCols;                        // 16 stages of 32 rows per side
   constexpr int kLoads = kStages * kStageBytes / 1024 / kWholeWaves; // 1 KiB loads per wave per side
   static_assert(kLoads == 8 && kStages % kGroup == 0, "tile geometry");
   __shared__ __attribute__((aligned(16))) int8_t sR[kStages * kStageBytes];
   __shared__ __attribute__((aligned(16))) int8_t sL[kStages * kStageBytes];
   __shared__ __attribute__((aligned(16))) int32_t sRs[kTileN];
   __shared__ __attribute__((aligned(16))) int32_t sLs[kTileM];
   // survivor regions of a scan that applies its own survivors: kept in LDS
   // (the tile end reads them back at LDS latency); otherwise in global memory
   // for k_i8_insert
   constexpr int kLdsSurvCap = 64;
   __shared__ __attribute__((aligned(16))) uint64_t sSurv[INLINE_INSERT ? kWholeWaves : 1][kLdsSurvCap];
   __shared__ int sNextTile;
   touch_kernargs(a);

   const int tid = threadIdx.x;
   const int lane = tid & 63;
   const int wave = tid >> 6;
   const int half = lane >> 5, lrow = lane & 31;
   // LDS unit p (16 bytes) of either side holds stage p / 256, row r = (p / 8) % 32
   // of it, chunk (p % 8) ^ swz(r); load j of this wave fills units
   // (wave * kLoads + j) * 64 + lane, all in stage (wave * kLoads + j) / 4
   int voff[kLoads];
#pragma unroll
   for (int j = 0; j < kLoads; ++j)
   {
      const int p = (wave * kLoads + j) * 64 + lane, st = p >> 8, pu = p & 255, r = pu >> 3, x = pu & 7;
      voff[j] = st * kStageBytes + r * kDim + ((x ^ swz(r)) << 4);
   }
   int fragoff[4]; // this lane's fragment units within a stage: row lrow, chunk 4 * half + k
#pragma unroll
   for (int k = 0; k < 4; ++k) fragoff[k] = (lrow * 8 + ((4 * half + k) ^ swz(lrow))) << 4;

   int tile = blockIdx.x;
   while (tile < a.nTiles)
   {
   uint32_t nextReq = 0;
   if (a.persistent && tid == 0) nextReq = atomicAdd(a.tileCounter, 1u) + gridDim.x;
   PairDesc pd = a.pair0; // single pair: straight from the kernel arguments
   if (a.nPairs > 1) pd = a.pairs[sgpr(pair_of_tile(a.tilePair, a.pairs, a.nPairs, tile))];
   else if (a.pairsOnDevice) pd = a.pairs[0];
   pd.offL = ((int64_t)sgpr((int)(pd.offL >> 32)) << 32) | (uint32_t)sgpr((int)pd.offL);
   pd.offR = ((int64_t)sgpr((int)(pd.offR >> 32)) << 32) | (uint32_t)sgpr((int)pd.offR);
   pd.slotL = ((int64_t)sgpr((int)(pd.slotL >> 32)) << 32) | (uint32_t)sgpr((int)pd.slotL);
   pd.slotR = ((int64_t)sgpr((int)(pd.slotR >> 32)) << 32) | (uint32_t)sgpr((int)pd.slotR);
   pd.nL = sgpr(pd.nL);
   pd.nR = sgpr(pd.nR);
   pd.tilesN = sgpr(pd.tilesN);
   pd.tileBase = sgpr(pd.tileBase);
   const int t = tile - pd.tileBase;
   if (a.pairsOnDevice && t >= sgpr(pd.tilesM) * pd.tilesN) break; // past the job's own grid (never persistent)
   const int tm = t / pd.tilesN, tn = t - tm * pd.tilesN;
   const int nLpad = (pd.nL + kRowPad - 1) / kRowPad * kRowPad;
   const int nRpad = (pd.nR + kRowPad - 1) / kRowPad * kRowPad;
   const int r0 = sgpr(tm * kTileM), c0 = sgpr(tn * kTileN);
   const int stagesL = sgpr((min(r0 + kTileM, nLpad) - r0) / kStepCols); // stages present (multiples of 4)
   const int nsteps = sgpr((min(c0 + kTileN, nRpad) - c0) / kStepCols);
   const int l0 = sgpr(r0 + wave * kWaveRows);
   const bool active = l0 < nLpad; // wave-uniform; inactive waves still load and sync

   // every load of the tile (nothing past either side's padded end), one wait
   {
      const void *srcL = uniform_ptr(a.L + (size_t)(pd.offL + r0) * kDim);
      const void *srcR = uniform_ptr(a.R + (size_t)(pd.offR + c0) * kDim);
      const uint32_t dL = lds_addr(sL), dR = lds_addr(sR);
#pragma unroll
      for (int j = 0; j < kLoads; ++j)
      {
         const int st = (wave * kLoads + j) >> 2;
         const uint32_t off = (uint32_t)((wave * kLoads + j) * 1024);
         if (st < stagesL) lds_dma16(srcL, voff[j], (uint32_t)sgpr((int)(dL + off)));
         if (st < nsteps) lds_dma16(srcR, voff[j], (uint32_t)sgpr((int)(dR + off)));
      }
      // row terms: ints [64 wave, 64 wave + 64) of each side
      if (64 * wave < stagesL * kStepCols)
         lds_dma4(uniform_ptr(a.Lsum + pd.offL + r0 + 64 * wave), 4 * lane,
                  (uint32_t)sgpr((int)(lds_addr(sLs) + 256 * wave)));
      if (64 * wave < nsteps * kStepCols)
         lds_dma4(uniform_ptr(a.Rsum + pd.offR + c0 + 64 * wave), 4 * lane,
                  (uint32_t)sgpr((int)(lds_addr(sRs) + 256 * wave)));
      vm_wait_barrier<0>();
   }
   uint64_t *region = INLINE_INSERT ? &sSurv[INLINE_INSERT ? wave : 0][0]
                                    : a.surv + ((size_t)tile * kWholeWaves + wave) * a.survCap;
   const int survCap = INLINE_INSERT ? min(a.survCap, kLdsSurvCap) : a.survCap;
   int nsurv = 0; // survivors this wave found in this tile (wave-uniform)
   if (active)
   {
      const int32_t thr = a.thr;
      v4i A[2][4];
      v16i C[2];
#pragma unroll
      for (int b = 0; b < 2; ++b)
      {
         const int8_t *sa = sL + (2 * wave + b) * kStageBytes; // rows l0 + 32 b ..
#pragma unroll
         for (int k = 0; k < 4; ++k) A[b][k] = *reinterpret_cast<const v4i *>(sa + fragoff[k]);
#pragma unroll
         for (int q = 0; q < 4; ++q)
         {
            const v4i rs = *reinterpret_cast<const v4i *>(sLs + wave * kWaveRows + 32 * b + 8 * q + 4 * half);
#pragma unroll
            for (int e = 0; e < 4; ++e) C[b][4 * q + e] = rs[e] - kBias;
         }
      }
      auto mfma_step = [&](const v4i(&Bf)[4], v16i &x0, v16i &x1) {
         x0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][0], Bf[0], C[0], 0, 0, 0);
         x1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][0], Bf[0], C[1], 0, 0, 0);
#pragma unroll
         for (int k = 1; k < 4; ++k)
         {
            x0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][k], Bf[k], x0, 0, 0, 0);
            x1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][k], Bf[k], x1, 0, 0, 0);
         }
      };
      v16i p0, p1; // scores of the step in flight
      {
         v4i Bz[4];
#pragma unroll
         for (int k = 0; k < 4; ++k) Bz[k] = *reinterpret_cast<const v4i *>(sR + fragoff[k]);
         mfma_step(Bz, p0, p1);
      }
      // steps 4 g .. 4 g + 3; `lastc` says at compile time whether they end the
      // tile (only the last step issues no next-step MFMAs)
      auto quad = [&](int g, auto lastc) {
         constexpr bool kLast = decltype(lastc)::value;
#pragma unroll
         for (int s = 0; s < kGroup; ++s)
         {
            const int S = g * kGroup + s;
            const bool more = s + 1 < kGroup || !kLast; // constant once unrolled
            v4i Bn[4];
            if (more)
            {
#pragma unroll
               for (int k = 0; k < 4; ++k)
                  Bn[k] = *reinterpret_cast<const v4i *>(sR + (S + 1) * kStageBytes + fragoff[k]);
            }
            const v16i acc0 = p0, acc1 = p1;
            if (more) mfma_step(Bn, p0, p1);
            const int cs = sRs[S * kStepCols + lrow];
            const int m = max(max16(acc0), max16(acc1));
            const int lim = thr - cs; // score > thr  <=>  acc > thr - 128*sum(u_right)
            if (__builtin_expect(__any(m > lim), 0))
               record_survivors(acc0, acc1, lim, cs, S * kStepCols + lrow, half, survCap, region, nsurv);
         }
      };
      const std::false_type kMore;
      const std::true_type kEnd;
      const int ngroups = nsteps / kGroup; // 1 .. 4
      int g = 0;
      for (; g + 1 < ngroups; ++g) quad(g, kMore);
      quad(g, kEnd);
   }
   if constexpr (!INLINE_INSERT)
   {
      if (lane == 0) a.survCount[(size_t)tile * kWholeWaves + wave] = nsurv;
   }
   else if (nsurv > 0)
   {
      ScanArgs b = a;
      b.survCap = survCap;
      apply_region<true>(b, pd, region, nsurv, l0, c0, kTileN, lane);
   }
   if (!a.persistent) break;
   __syncthreads(); // every wave is done with this tile's LDS (no load of it is in flight)
   if (tid == 0) sNextTile = (int)nextReq;
   __syncthreads();
   tile = sNextTile;
   }
}

// Applies the survivors of one (tile, wave) region per wave: exact top-2
// insertion into the left and right key arrays (guided masks checked here).
// A region whose count exceeds its capacity (possible only when very many
// scores of one 64 x 512 block pass the filter, e.g. near-identical images)
// is recomputed here from the descriptors, so the result stays exact.
// kPerTile: one wave per tile, for all of the tile's regions (kMaxScanWaves
// at most; large grids, where the per-wave latency chains are what costs);
// otherwise one wave per region (small grids, where more waves in flight
// pay).  Every region's count and first 64 entries (inside the region:
// survCap >= 64, checked on the host), the tile's pair and the epoch are
// loaded before any of them is waited for, and the pair is looked up once per
// wave; stale entries past a count are not used.
constexpr int kMaxScanWaves = 8;
template <bool kPerTile>
__global__ __launch_bounds__(256) void k_i8_insert(ScanArgs a)
{
   touch_kernargs(a);
   constexpr int kSpan = kPerTile ? kMaxScanWaves : 1; // regions a wave may take
   const int lane = threadIdx.x & 63;
   const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
   int tile, wFirst, wCount;
   if (kPerTile)
   {
      if (wid >= a.nTiles) return;
      tile = wid;
      wFirst = 0;
      wCount = a.waves;
   }
   else
   {
      if (wid >= a.nTiles * a.waves) return;
      tile = wid / a.waves;
      wFirst = wid - tile * a.waves;
      wCount = 1;
   }
   if (a.pairsOnDevice && tile >= a.pairs[0].tilesM * a.pairs[0].tilesN) return; // no scan tile wrote it
   const size_t r0 = (size_t)tile * a.waves + wFirst;
   int cnt[kSpan];
   uint64_t w0[kSpan];
#pragma unroll
   for (int w = 0; w < kSpan; ++w)
   {
      cnt[w] = 0;
      w0[w] = 0;
      if (w < wCount)
      {
         cnt[w] = a.survCount[r0 + w];
         w0[w] = a.surv[(r0 + w) * a.survCap + lane];
      }
   }
   const int p = a.nPairs > 1 ? pair_of_tile(a.tilePair, a.pairs, a.nPairs, tile) : 0;
   const uint32_t epoch = job_epoch(a.epoch, a.epochValue);
   int any = 0;
#pragma unroll
   for (int w = 0; w < kSpan; ++w) any |= cnt[w];
   if (any == 0) return;
   const PairDesc pd = (a.nPairs > 1 || a.pairsOnDevice) ? a.pairs[p] : a.pair0; // else the pair comes by value
   const int t = tile - pd.tileBase;
   const int tm = t / pd.tilesN, tn = t - tm * pd.tilesN;
   const int c0 = tn * a.tileN;
#pragma unroll
   for (int w = 0; w < kSpan; ++w)
   {
      if (cnt[w] == 0) continue;
      const int l0 = tm * a.tileM + (wFirst + w) * kWaveRows;
      const uint64_t *region = a.surv + (r0 + w) * a.survCap;
      if (cnt[w] <= a.survCap)
      {
         for (int e = lane; e < cnt[w]; e += 64)
         {
            const uint64_t v = e == lane ? w0[w] : region[e];
            survivor(a, pd, epoch, l0 + (int)((v >> 16) & 0xFFFF), c0 + (int)(v & 0xFFFF), (int32_t)(v >> 32));
         }
         continue;
      }
      // overflowed region: every (left, right) of the block, exact u8 dot products
      const int nRows = min(kWaveRows, pd.nL - l0), nCols = min(a.tileN, pd.nR - c0);
      for (int e = lane; e < nRows * nCols; e += 64)
      {
         const int il = l0 + e / nCols, jr = c0 + e % nCols;
         const uint32_t *x = reinterpret_cast<const uint32_t *>(a.L + (size_t)(pd.offL + il) * kDim);
         const uint32_t *y = reinterpret_cast<const uint32_t *>(a.R + (size_t)(pd.offR + jr) * kDim);
         uint32_t acc = 0;
         for (int k = 0; k < kDim / 4; ++k) // stored bytes are u ^ 0x80
            acc = __builtin_amdgcn_udot4(x[k] ^ 0x80808080u, y[k] ^ 0x80808080u, acc, false);
         if ((int32_t)acc > a.thr) survivor(a, pd, epoch, il, jr, (int32_t)acc);
      }
   }
}

__device__ inline float key_value(uint64_t k, uint32_t epoch)
{
   return key_epoch(k) == epoch ? (float)key_score(k) * (1.0f / 262144.0f) : -1e30f;
}

// One thread per left feature: column decision (ratio test on the left's
// top-2), then the right side's decision for the chosen right feature and
// the mutual check.  With a block counter (single-pair jobs) the last block to
// finish advances the epoch; large grids leave that to k_advance_epoch.
__global__ __launch_bounds__(256) void k_i8_finalize(FinalizeArgs a)
{
   touch_kernargs(a);
   const int blk = blockIdx.x;
   PairDesc pd = a.pair0;
   if (!a.hasPair0) pd = a.pairs[pair_of_fin(a.finPair, a.pairs, a.nPairs, blk)];
   const int k = (blk - pd.finBase) * 256 + threadIdx.x;
   const uint32_t epoch = job_epoch(a.epoch, a.epochValue);
   if (k < pd.nL)
   {
      // two dependent rounds of loads: the left's keys, then (for a key of
      // this job, whose right index is valid) the chosen right's keys
      int out = -1;
      const uint64_t t1 = a.top1L[pd.slotL + k];
      const uint64_t t2 = a.top2L[pd.slotL + k];
      if (key_epoch(t1) == epoch)
      {
         const uint32_t j = key_index(t1);
         const uint64_t u1 = a.top1R[pd.slotR + j];
         const uint64_t u2 = a.top2R[pd.slotR + j];
         const float best = key_value(t1, epoch);
         const float next = key_value(t2, epoch);
         if (rule_accept(best, next, a.minScore, a.minRatio))
         {
            if (key_epoch(u1) == epoch && key_index(u1) == (uint32_t)k)
            {
               const float nextR = key_value(u2, epoch);
               if (rule_accept(key_value(u1, epoch), nextR, a.minScore, a.minRatio)) out = (int)j;
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

// Closes a job whose finalize grid does not count its blocks (large batches,
// where thousands of blocks would serialise on one counter): the next job's
// epoch, and the persistent scan's queue head reset.  One thread.
__global__ void k_advance_epoch(uint32_t *epoch, uint32_t *tileCounter)
{
   *tileCounter = 0;
   *epoch = *epoch + 1;
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
      done[1] = 0; // persistent scan's tile counter follows the done counter
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

// Ceiling probe of the scan's step loop: per step 8 v_mfma_i32_32x32x32_i8
// on the wave's 64 rows held in VGPRs, the previous step's max / compare
// test, the column term read from LDS and (kFromLds) the four right
// fragments read from a swizzled 16 KiB LDS stage in the scan's layout --
// with nothing else (no global loads, barriers or survivors).  The operands
// are the caller's 128 stored rows (A: this wave's 64 of them; LDS: all 128
// as four 32-row stages), so the measured rate is the loop's rate on that
// data: the int8 MFMA's clock under its power limit depends on the operand
// values.  !kFromLds keeps the fragments of one stage in registers.
template <bool kFromLds>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_mfma_ceiling(
   const int8_t *__restrict__ rows, int iters, int32_t *__restrict__ out)
{
   __shared__ __attribute__((aligned(16))) int8_t stage[4 * kStageBytes];
   __shared__ int32_t colTerm[4 * kStepCols];
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int lrow = lane & 31, half = lane >> 5;
   for (int u = tid; u < 4 * kStepCols * 8; u += 256) // 16-byte units: 128 rows x 8
   {
      const int row = u >> 3, x = u & 7, st = row / kStepCols, r = row % kStepCols;
      *reinterpret_cast<v4i *>(stage + st * kStageBytes + (r * 8 + (x ^ swz(r))) * 16) =
         ld16(rows + (size_t)row * kDim + 16 * x);
   }
   if (tid < 4 * kStepCols) colTerm[tid] = -(1 << 20);
   __syncthreads();
   v4i A[2][4];
   v16i C[2];
#pragma unroll
   for (int b = 0; b < 2; ++b)
   {
      const int8_t *ap = rows + (size_t)((wave * kWaveRows + 32 * b + lrow) & 127) * kDim + 64 * half;
#pragma unroll
      for (int s = 0; s < 4; ++s) A[b][s] = ld16(ap + 16 * s);
#pragma unroll
      for (int e = 0; e < 16; ++e) C[b][e] = (1 << 20) - kBias;
   }
   int fragoff[4];
#pragma unroll
   for (int k = 0; k < 4; ++k) fragoff[k] = (lrow * 8 + ((4 * half + k) ^ swz(lrow))) << 4;
   v4i Bk[4];
#pragma unroll
   for (int k = 0; k < 4; ++k) Bk[k] = *reinterpret_cast<const v4i *>(stage + fragoff[k]);
   v16i s0[2] = {C[0], C[0]}, s1[2] = {C[1], C[1]};
   int hits = 0;
   auto step = [&](int it, auto cur) {
      constexpr int c = decltype(cur)::value, n = c ^ 1;
      const int8_t *sb = stage + (it & 3) * kStageBytes;
      v4i B[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
      {
         if (kFromLds) B[k] = *reinterpret_cast<const v4i *>(sb + fragoff[k]);
         else B[k] = Bk[k];
      }
      if (!kFromLds) asm volatile("" : "+v"(B[0]), "+v"(B[1]), "+v"(B[2]), "+v"(B[3]));
      s0[n] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][0], B[0], C[0], 0, 0, 0);
      s1[n] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][0], B[0], C[1], 0, 0, 0);
#pragma unroll
      for (int k = 1; k < 4; ++k)
      {
         s0[n] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][k], B[k], s0[n], 0, 0, 0);
         s1[n] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][k], B[k], s1[n], 0, 0, 0);
      }
      const int cs = colTerm[(it & 3) * kStepCols + lrow];
      const int m = max(max16(s0[c]), max16(s1[c]));
      if (__builtin_expect(__any(m > (1 << 30) - cs), 0)) ++hits; // never: every score is far below 2^30
   };
   const std::integral_constant<int, 0> K0;
   const std::integral_constant<int, 1> K1;
   for (int it = 0; it + 1 < iters; it += 2)
   {
      step(it, K0);
      step(it + 1, K1);
   }
   out[(size_t)blockIdx.x * 256 + tid] = hits + s0[0][lane & 15] + s1[1][3];
}

// ---- launchers ------------------------------------------------------------

// Scan variants: (waves per workgroup, right rows per tile); the last one is
// the whole-tile form (k_i8_scan_whole).
constexpr int kMaxDevicesCached = 64;

struct ScanVariant
{
   int waves, tileN;
   const void *fn;
};
static const ScanVariant kVariants[] = {
   {4, 512, reinterpret_cast<const void *>(&k_i8_scan<4, 512, false>)},
   {4, 1024, reinterpret_cast<const void *>(&k_i8_scan<4, 1024, false>)},
   {8, 512, reinterpret_cast<const void *>(&k_i8_scan<8, 512, false>)},
   {4, 256, reinterpret_cast<const void *>(&k_i8_scan<4, 256, false>)},
   {4, 2048, reinterpret_cast<const void *>(&k_i8_scan<4, 2048, false>)},
   {4, 4096, reinterpret_cast<const void *>(&k_i8_scan<4, 4096, false>)},
   {kWholeWaves, kWholeTile, reinterpret_cast<const void *>(&k_i8_scan_whole<false>)},
};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));

int scan_variant_count() { return kNumVariants; }
int scan_variant_waves(int v) { return kVariants[v].waves; }
int scan_variant_whole(int v) { return kVariants[v].fn == reinterpret_cast<const void *>(&k_i8_scan_whole<false>); }
int scan_variant_tile_m(int v) { return kVariants[v].waves * kWaveRows; }
int scan_variant_tile_n(int v) { return kVariants[v].tileN; }
// Survivor entries per (tile, wave) region: 64 per 512 right rows, so a
// region of a wide tile overflows at the same survivor density as one of a
// 512-wide tile (and the survivor memory per scored pair stays the same).
int scan_variant_surv_cap(int v) { return 64 * std::max(1, kVariants[v].tileN / 512); }

// Workgroups of variant v resident on the whole device (for persistent launches).
int scan_variant_resident(int v, int device)
{
   // asked on every job: remembered per device and variant (0 = not yet known)
   static std::atomic<int> known[kMaxDevicesCached][kNumVariants];
   const bool cacheable = device >= 0 && device < kMaxDevicesCached && v >= 0 && v < kNumVariants;
   if (cacheable)
   {
      const int k = known[device][v].load(std::memory_order_relaxed);
      if (k > 0) return k;
   }
   int perCU = 0, cus = 0;
   if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kVariants[v].fn, kVariants[v].waves * 64, 0) !=
       hipSuccess)
      perCU = 1;
   if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 1;
   const int r = (perCU > 0 ? perCU : 1) * (cus > 0 ? cus : 1);
   if (cacheable) known[device][v].store(r, std::memory_order_relaxed);
   return r;
}

hipError_t launch_i8_scan(int variant, int grid, const ScanArgs &a, hipStream_t st, hipEvent_t start,
                          hipEvent_t stop)
{
   if (a.nTiles <= 0) return hipSuccess;
   // start / stop (may be null) take the timestamps of this dispatch itself
   const int k = 2 * variant + (a.inlineInsert ? 1 : 0);
   switch (k)
   {
   case 0: hipExtLaunchKernelGGL((k_i8_scan<4, 512, false>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 1: hipExtLaunchKernelGGL((k_i8_scan<4, 512, true>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 2: hipExtLaunchKernelGGL((k_i8_scan<4, 1024, false>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 3: hipExtLaunchKernelGGL((k_i8_scan<4, 1024, true>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 4: hipExtLaunchKernelGGL((k_i8_scan<8, 512, false>), dim3(grid), dim3(512), 0, st, start, stop, 0, a); break;
   case 5: hipExtLaunchKernelGGL((k_i8_scan<8, 512, true>), dim3(grid), dim3(512), 0, st, start, stop, 0, a); break;
   case 6: hipExtLaunchKernelGGL((k_i8_scan<4, 256, false>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 7: hipExtLaunchKernelGGL((k_i8_scan<4, 256, true>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 8: hipExtLaunchKernelGGL((k_i8_scan<4, 2048, false>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 9: hipExtLaunchKernelGGL((k_i8_scan<4, 2048, true>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 10: hipExtLaunchKernelGGL((k_i8_scan<4, 4096, false>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 11: hipExtLaunchKernelGGL((k_i8_scan<4, 4096, true>), dim3(grid), dim3(256), 0, st, start, stop, 0, a); break;
   case 12:
      hipExtLaunchKernelGGL((k_i8_scan_whole<false>), dim3(grid), dim3(kWholeWaves * 64), 0, st, start, stop, 0, a);
      break;
   case 13:
      hipExtLaunchKernelGGL((k_i8_scan_whole<true>), dim3(grid), dim3(kWholeWaves * 64), 0, st, start, stop, 0, a);
      break;
   default: return hipErrorInvalidValue;
   }
   return hipGetLastError();
}

// Insertion form for `regions` survivor regions on `device` when the caller
// leaves the choice open (form < 0): one wave per tile above 128 regions per
// CU (on 256 CUs: 32768 -- the 256-pair batch's 147 k regions ran 127 -> 116 us
// per tile, 8192^2 with 256 x 256 tiles, 4096 regions, 4.0 us per region
// against 7.3 us per tile), else one wave per region.
int insert_form_for(int64_t regions, int device, int form)
{
   if (form == 0 || form == 1) return form;
   static std::atomic<int> cusKnown[kMaxDevicesCached];
   int cus = device >= 0 && device < kMaxDevicesCached ? cusKnown[device].load(std::memory_order_relaxed) : 0;
   if (cus <= 0)
   {
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
         cus = 256;
      if (device >= 0 && device < kMaxDevicesCached) cusKnown[device].store(cus, std::memory_order_relaxed);
   }
   return regions > (int64_t)128 * cus ? 1 : 0;
}

hipError_t launch_i8_insert(const ScanArgs &a, int form, hipStream_t st)
{
   if (a.nTiles <= 0) return hipSuccess;
   if (a.survCap < 64) return hipErrorInvalidValue; // each wave reads its regions' first 64 entries
   if (a.waves < 1 || a.waves > kMaxScanWaves) return hipErrorInvalidValue;
   const int64_t regions = (int64_t)a.nTiles * a.waves;
   if (form == 1)
      hipLaunchKernelGGL(k_i8_insert<true>, dim3((a.nTiles + 3) / 4), dim3(256), 0, st, a);
   else if (form == 0)
      hipLaunchKernelGGL(k_i8_insert<false>, dim3((unsigned)((regions + 3) / 4)), dim3(256), 0, st, a);
   else
      return hipErrorInvalidValue;
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

hipError_t launch_advance_epoch(uint32_t *epoch, uint32_t *tileCounter, hipStream_t st)
{
   hipLaunchKernelGGL(k_advance_epoch, dim3(1), dim3(1), 0, st, epoch, tileCounter);
   return hipGetLastError();
}

hipError_t launch_mfma_probe(int32_t *out, hipStream_t st)
{
   hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, out);
   return hipGetLastError();
}

hipError_t launch_mfma_ceiling(bool fromLds, int grid, const int8_t *rows, int iters, int32_t *out, hipStream_t st)
{
   if (fromLds) hipLaunchKernelGGL(k_mfma_ceiling<true>, dim3(grid), dim3(256), 0, st, rows, iters, out);
   else hipLaunchKernelGGL(k_mfma_ceiling<false>, dim3(grid), dim3(256), 0, st, rows, iters, out);
   return hipGetLastError();
}

} // namespace v3dm
