// v3dm_internal.h -- types shared by the host API and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace v3dm {

constexpr int kDim = 128;          // descriptor length
constexpr int kRowPad = 128;       // descriptor pools are padded to multiples of 128 rows
// Row-sum term of a padding row of an int8 pool.  In the scan a padding left
// row's accumulator starts near -2^30 and a padding right row's threshold is
// raised by 2^30, so neither passes the survivor test (scores of real rows are
// below 2^24): the scan needs no per-step row or column bound checks.
constexpr int32_t kPadRowSum = -(1 << 30);
constexpr int kMaxMasks = 16;      // guided masks accumulated between two plain finds
constexpr int kMaxLocalIndex = (1 << 21) - 1; // per-image feature limit of the packed keys
constexpr uint32_t kEpochLimit = (1u << 20) - 2;

// One image pair of a match job.  Descriptor rows live in pools shared by
// all pairs (an image used by several pairs is stored once); the per-pair
// top-2 state lives in slots [slotL, slotL + nL) and [slotR, slotR + nR).
struct PairDesc
{
   int64_t offL, offR;   // first pool row of the left / right image
   int64_t slotL, slotR; // first top-2 slot of this pair's left / right side
   int nL, nR;           // feature counts
   int tilesM, tilesN;   // scan tiles along left / right
   int tileBase;         // index of this pair's first scan tile
   int finBase;          // index of this pair's first finalize block
   int guide;            // 0: no per-pair mask; k > 0: the batch's guided mask masks[k - 1]
};

// Guided-matching mask, applied to candidate (left p, right q) positions.
struct Mask
{
   int kind;           // 1 = epipolar (Sampson), 2 = homography transfer
   float sqrThr;       // distanceThreshold^2 computed in float
   float m[9];         // row-major 3x3, converted from double
   const float *lpos;  // interleaved (x, y) per left feature, as given to this call
   const float *rpos;  // interleaved (x, y) per right feature
};

// Packed top-2 key: [63:44] epoch | [43:21] score | [20:0] (kMaxLocalIndex - index).
// A larger key is a larger score, ties broken towards the smaller index,
// which is the reference's first-index rule.
__host__ __device__ inline uint64_t pack_key(uint32_t epoch, uint32_t score, uint32_t index)
{
   return ((uint64_t)epoch << 44) | ((uint64_t)score << 21) | (uint64_t)(kMaxLocalIndex - index);
}
__host__ __device__ inline uint32_t key_epoch(uint64_t k) { return (uint32_t)(k >> 44); }
__host__ __device__ inline uint32_t key_score(uint64_t k) { return (uint32_t)((k >> 21) & 0x7FFFFFu); }
__host__ __device__ inline uint32_t key_index(uint64_t k) { return (uint32_t)(kMaxLocalIndex - (uint32_t)(k & 0x1FFFFFu)); }

// Float-score key (BF16 candidate path): [63:53] epoch | [52:21] order-preserving
// image of the FP32 score | [20:0] (kMaxLocalIndex - index).  -0 is folded into
// +0 first, so equal scores compare equal as in the reference's `>`.
constexpr uint32_t kEpochLimitF = (1u << 11) - 2;
__host__ __device__ inline uint32_t float_order(float s)
{
   union { float f; uint32_t u; } v;
   v.f = s == 0.0f ? 0.0f : s;
   return (v.u & 0x80000000u) ? ~v.u : (v.u | 0x80000000u);
}
__host__ __device__ inline float float_unorder(uint32_t o)
{
   union { float f; uint32_t u; } v;
   v.u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
   return v.f;
}
__host__ __device__ inline uint64_t pack_keyf(uint32_t epoch, float score, uint32_t index)
{
   return ((uint64_t)epoch << 53) | ((uint64_t)float_order(score) << 21) | (uint64_t)(kMaxLocalIndex - index);
}
__host__ __device__ inline uint32_t keyf_epoch(uint64_t k) { return (uint32_t)(k >> 53); }
__host__ __device__ inline float keyf_score(uint64_t k) { return float_unorder((uint32_t)(k >> 21)); }

struct ScanArgs
{
   PairDesc pair0;              // the pair itself when nPairs == 1 (no lookup)
   const int8_t *L;     // left pool, (u - 128) as int8, rows padded to kRowPad
   const int32_t *Lsum; // 128 * sum(u) per left row (0 for pad rows)
   const int8_t *R;
   const int32_t *Rsum;
   const PairDesc *pairs;
   int nPairs;
   int nTiles;
   int32_t thr;                 // candidate survives iff exact integer score > thr
   uint64_t *top1L, *top2L;     // per left slot
   uint64_t *top1R, *top2R;     // per right slot
   const uint32_t *epoch;       // device-resident epoch of the current job; null: epochValue
   uint32_t epochValue;         // the job's epoch, passed by value (single-pair jobs)
   const Mask *masks;           // guided masks (single-pair jobs only)
   int nMasks;
   uint64_t *surv;              // survivor regions, survCap entries per (tile, wave)
   int32_t *survCount;          // entries written per region (may exceed survCap); !inlineInsert only
   int survCap;
   int inlineInsert;            // 1: each wave applies its region when its tile ends (no k_i8_insert)
   int waves, tileM, tileN;     // tile geometry of the launched scan variant
   int persistent;              // 1: resident workgroups pull tiles from *tileCounter
   uint32_t *tileCounter;       // reset to 0 by the finalize kernel
   const int32_t *mapL, *mapR;  // optional: feature index of each row of a gathered pool
   const int32_t *tilePair;     // optional: index into pairs of every tile (else a binary search)
   int pairsOnDevice;           // 1: the one pair's descriptor is pairs[0], written on the device by an
                                // earlier kernel of the job; tiles past its tile count exit at once
};

// Float descriptors: BF16 MFMA candidate scan + exact FP32 rescoring.
struct ScanArgsF
{
   PairDesc pair0;           // the pair itself when nPairs == 1
   const uint16_t *L, *R;    // BF16 pools (round-to-nearest-even), rows padded to kRowPad with zeros
   const float *Lf, *Rf;     // FP32 pools, the descriptors as given (rescoring)
   const PairDesc *pairs;
   int nPairs;
   int nTiles;
   float lim;                // candidate iff BF16 score > lim (= thr - proven BF16 error bound)
   float thr;                // a rescored FP32 score enters the top-2 iff > thr
   uint64_t *top1L, *top2L, *top1R, *top2R;
   const uint32_t *epoch; // null: epochValue
   uint32_t epochValue;
   const Mask *masks;
   int nMasks;
   uint64_t *surv;
   int32_t *survCount;
   int survCap;
   int waves, tileM, tileN;
   int persistent;
   uint32_t *tileCounter;
   const int32_t *tilePair;  // optional: index into pairs of every tile
};

struct FinalizeArgs
{
   const PairDesc *pairs;
   int nPairs;
   int nBlocks;
   const uint64_t *top1L, *top2L, *top1R, *top2R;
   uint32_t *epoch;             // null: epochValue (then done is null too)
   uint32_t epochValue;
   uint32_t *done;              // block counter of the job's last-block epoch advance; null: none
   uint32_t *tileCounter;       // persistent scan's queue head, reset at the end of the job
   float minScore, minRatio;
   int32_t *matchOfLeft; // per left slot: matched right index or -1
   const int32_t *finPair; // optional: index into pairs of every finalize block
   PairDesc pair0;         // the job's one pair, by value (hasPair0): no dependent load of it
   int hasPair0;
};

} // namespace v3dm
