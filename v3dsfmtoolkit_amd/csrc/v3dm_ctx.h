// v3dm_ctx.h -- host-side state of a matcher context (shared by the C-ABI
// translation units; not part of the public interface).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/v3dmatch.h"
#include "v3dm_internal.h"

namespace v3dm {
hipError_t launch_i8_scan(int variant, int grid, const ScanArgs &a, hipStream_t st, hipEvent_t start = nullptr,
                          hipEvent_t stop = nullptr);
int scan_variant_count();
int scan_variant_waves(int v);
int scan_variant_whole(int v); // 1: k_i8_scan_whole (whole tile staged in LDS)
int scan_variant_tile_m(int v);
int scan_variant_tile_n(int v);
int scan_variant_surv_cap(int v);
int scan_variant_resident(int v, int device);
int insert_form_for(int64_t regions, int device, int form);
hipError_t launch_i8_insert(const ScanArgs &a, int form, hipStream_t st);
hipError_t launch_i8_finalize(const FinalizeArgs &a, hipStream_t st);
hipError_t launch_bf16_scan(int variant, int grid, const ScanArgsF &a, hipStream_t st, hipEvent_t start = nullptr,
                            hipEvent_t stop = nullptr);
int bf16_scan_resident(int variant, int device);
int bf16_scan_variant_count();
hipError_t launch_f32_rescore(const ScanArgsF &a, hipStream_t st);
hipError_t launch_f32_finalize(const FinalizeArgs &a, hipStream_t st);
hipError_t launch_prep_bf16(const float *src, int n, int nPad, uint16_t *dst, uint32_t *norm2Max, hipStream_t st,
                            bool resetNorm = true);
hipError_t launch_advance_epoch(uint32_t *epoch, uint32_t *tileCounter, hipStream_t st);
hipError_t launch_clear_keys(uint64_t *keys, int64_t n, uint32_t *epoch, uint32_t *done, hipStream_t st);
hipError_t launch_mfma_probe(int32_t *out, hipStream_t st);
hipError_t launch_mfma_ceiling(bool fromLds, int grid, const int8_t *rows, int iters, int32_t *out, hipStream_t st);
hipError_t launch_copy_probe(const void *src, void *dst, int64_t bytes, hipStream_t st);
hipError_t launch_sgemm_seq(int nL, int nR, const float *Lf, const float *Rf, float *S, hipStream_t st,
                            hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_mask_plain(int nL, int nR, float *S, const Mask *masks, int nMasks, hipStream_t st);
hipError_t launch_fill_i32(int n, int32_t value, int32_t *out, hipStream_t st);
size_t plain_part_bytes(int nL, int nR, int *nChunksOut, int *rowsPerOut);
hipError_t launch_plain_select(int nL, int nR, const float *S, void *part, float minScore, float minRatio,
                               int32_t *colMax, int32_t *rowMax, int32_t *out, hipStream_t st);
hipError_t launch_prep_u8(const uint8_t *src, int n, int nPad, int8_t *dst, int32_t *sum, hipStream_t st);
hipError_t launch_prep_f32(const float *src, int n, int nPad, int8_t *dst, int32_t *sum, int *flag,
                           hipStream_t st);
hipError_t launch_u8_to_f32(const uint8_t *src, int64_t count, float *dst, hipStream_t st);
} // namespace v3dm

using namespace v3dm;

namespace v3dm_host {

extern thread_local std::string g_createError;

inline int roundUp(int n, int m) { return (n + m - 1) / m * m; }

// Survivor threshold in exact integer score units (score = int * 2^-18); see
// the proof at the top of k_i8.hip.  -1 keeps every score; kThrNone keeps none
// (no real row's score reaches 2^24, and a padding column's test limit,
// threshold + 2^30, stays below 2^31).
constexpr int32_t kThrNone = 1 << 28;
inline int32_t survivorThreshold(float minScore, float minRatio)
{
   double t;
   if (minRatio < 0)
      t = (double)minScore;
   else if (minRatio > 0)
   {
      double tr = 1.0 - (1.0 - (double)minScore) / ((double)minRatio * (1.0 - 1e-5));
      t = std::min((double)minScore, tr);
   }
   else if (minRatio == 0)
      t = std::max((double)minScore, 1.0); // accepted only with best > 1 (then the ratio is negative)
   else
      return kThrNone; // NaN: the rule accepts nothing
   if (!(t == t)) return kThrNone; // minScore NaN: likewise
   double x = std::floor(t * 262144.0);
   if (x < -1.0) return -1;
   if (x > (double)kThrNone) return kThrNone;
   return (int32_t)x;
}

// Float form of the survivor threshold (same proof): the largest float <= t,
// -inf when nothing may be filtered, +inf when the rule accepts nothing.
inline float survivorThresholdF(float minScore, float minRatio)
{
   double t;
   if (minRatio < 0)
      t = (double)minScore;
   else if (minRatio > 0)
      t = std::min((double)minScore, 1.0 - (1.0 - (double)minScore) / ((double)minRatio * (1.0 - 1e-5)));
   else if (minRatio == 0)
      t = std::max((double)minScore, 1.0);
   else
      return INFINITY;
   if (!(t == t)) return INFINITY;
   float f = (float)t;
   if ((double)f > t) f = std::nextafter(f, -INFINITY);
   return f;
}

// BF16 candidate limit: thr minus the BF16 error bound (2^-7 + 2^-14) |x| |y|
// for the largest row norms, rounded down.
inline float bf16CandidateLimit(float thr, float norm2MaxL, float norm2MaxR)
{
   if (thr == -INFINITY) return -INFINITY;
   const double delta = (std::ldexp(1.0, -7) + std::ldexp(1.0, -14)) *
                        std::sqrt((double)norm2MaxL * (double)norm2MaxR) * (1.0 + 1e-4);
   const double l = (double)thr - delta;
   if (!(l == l) || l < -3.0e38) return -INFINITY;
   float f = (float)l;
   if ((double)f > l) f = std::nextafter(f, -INFINITY);
   return f;
}

struct Side
{
   int n = 0;
   int nMax = 0;
   bool quantized = false; // int8 pool holds these features exactly
   bool isU8 = false;      // last upload was uint8
   bool f32Valid = false;  // dF32 holds the float form
   uint8_t *dRawU8 = nullptr;
   float *dF32 = nullptr; // float form (upload buffer for float input)
   int8_t *dI8 = nullptr;
   int32_t *dSum = nullptr;
   uint64_t *dTop1 = nullptr, *dTop2 = nullptr;
   float *dPos = nullptr; // kMaxMasks position sets (guided)
   void *hStage = nullptr;   // pinned host staging of the caller's descriptors (nMax x 128 floats)
   uint16_t *dBF = nullptr; // BF16 pool (float path), rows padded to kRowPad
   bool bfValid = false;
   float norm2Max = 0.0f;   // largest squared row norm (BF16 error bound)
};

} // namespace v3dm_host

namespace v3dm_host {
struct BatchState;
struct LearnedState;
}

using namespace v3dm_host;

struct v3dm_ctx
{
   int device = 0;
   hipStream_t stream = nullptr;
   std::string err;
   bool warnRounding = true;
   bool useReduction = false;
   int kernelSel = V3DM_KERNEL_AUTO;
   int lastKernel = 0;
   bool allocated = false;
   bool guided = false;
   Side L, R;

   PairDesc *dPair = nullptr;
   PairDesc hPairUploaded{};
   bool pairValid = false;
   uint32_t *dEpoch = nullptr; // [0] = epoch, [1] = done counter, [2] = tile counter
   uint32_t hostEpoch = 1;
   int32_t *dMatch = nullptr;
   int32_t *hMatch = nullptr; // pinned
   int *dFlag = nullptr;
   int *hFlag = nullptr;      // pinned

   std::vector<Mask> masks; // active since the last find
   Mask *dMasks = nullptr;

   // plain path
   float *dS = nullptr;
   size_t sBytes = 0;
   void *dPart = nullptr;
   size_t partBytes = 0;
   int32_t *dColMax = nullptr, *dRowMax = nullptr;
   bool sValid = false;     // dS holds the scores of the current descriptors
   int sMasksApplied = 0;   // masks already applied to dS
   bool lastWasPlain = false;

   bool pendingAsync = false;
   int lastNL = 0;

   // survivor regions of the fused scan
   uint64_t *dSurv = nullptr;
   int32_t *dSurvCount = nullptr;
   int survCap = 64;        // survivor entries per region of the BF16 scan (int8: per variant)
   int scanVariant = -1;                                 // single-pair int8 scan variant, -1: measured
   int lastVariant = -1;                                 // variant of the last single-pair int8 job
   int lastInline = -1;                                  // its scan applied the survivors itself (1) or not (0)
   int insertForm = -1;      // insertion pass: -1 chosen per job, 0 one wave per region, 1 one wave per tile
   int lastInsertForm = -1;  // form the last int8 job's insertion pass ran (-1: none, survivors applied in the scan)
   std::vector<std::pair<uint64_t, int>> tunedI8;       // shape bucket -> fastest variant

   // float path
   int bfVariant = 2;         // scan variant of the BF16 kernel
   uint32_t *dNorm = nullptr; // largest squared row norm, written by the BF16 prep kernel
   int keyFormat = 0;         // 0: integer-score keys in the top-2 arrays, 1: float-score keys
   float lastThr = 0.0f, lastLim = 0.0f;

   // kernel timing: three events per job (before scan, after scan, after select)
   bool profile = false;
   int profJobs = 0;
   std::vector<hipEvent_t> events;

   v3dm_host::BatchState *batch = nullptr;     // many-pair jobs (v3dm_batch_*)
   v3dm_host::LearnedState *learned = nullptr; // learned proposal path (v3dm_learned_find)
};

namespace v3dm_host {

inline int fail(v3dm_ctx *c, int code, const char *fmt, ...)
{
   char buf[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof(buf), fmt, ap);
   va_end(ap);
   if (c) c->err = buf;
   else g_createError = buf;
   return code;
}

#define HIPCHK(c, call)                                                                                  \
   do                                                                                                    \
   {                                                                                                     \
      hipError_t e_ = (call);                                                                            \
      if (e_ != hipSuccess)                                                                              \
         return fail((c), e_ == hipErrorOutOfMemory ? V3DM_ERR_NOMEM : V3DM_ERR_DEVICE, "%s: %s (%s:%d)", \
                     #call, hipGetErrorString(e_), __FILE__, __LINE__);                                  \
   } while (0)

template <class T> inline void dfree(T *&p)
{
   if (p) (void)hipFree((void *)p);
   p = nullptr;
}

int syncStream(v3dm_ctx *c);
int profMark(v3dm_ctx *c, int which);
int profKernelEvents(v3dm_ctx *c, hipEvent_t *start, hipEvent_t *stop);
void freeBatch(v3dm_ctx *c);
void freeLearned(v3dm_ctx *c);
int uploadPair(v3dm_ctx *c, int variant);
int useKeyFormat(v3dm_ctx *c, int format);
int fetch(v3dm_ctx *c, int32_t *pairs, int capacity, int *count);

} // namespace v3dm_host
