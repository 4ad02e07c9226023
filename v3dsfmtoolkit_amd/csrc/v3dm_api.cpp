// v3dm_api.cpp -- the C ABI declared in include/v3dmatch.h.
//
// Host side of V3D_CUDA::SIFT_Matching (V3D/src/CUDA/v3d_cudamatching.cpp):
// buffer ownership, the multiple-of-32 rounding and clamping of feature
// counts, the stateful guided-matching contract, kernel dispatch and the
// mutual-match list in ascending left order.
#include "v3dm_ctx.h"

namespace v3dm_host {
thread_local std::string g_createError;
}

namespace v3dm_host {

void freeSide(Side &s)
{
   dfree(s.dRawU8);
   dfree(s.dF32);
   dfree(s.dI8);
   dfree(s.dSum);
   dfree(s.dTop1);
   dfree(s.dTop2);
   dfree(s.dPos);
   dfree(s.dBF);
   if (s.hStage) (void)hipHostFree(s.hStage);
   s = Side();
}

// The caller's host descriptors (pageable) to device memory: copied into the
// side's pinned staging buffer first, then one DMA on the context stream
// (a pageable hipMemcpy goes through the runtime's own bounce buffers in
// pieces: 1.75 ms per 8192^2 pair of uint8 uploads and matches).
int uploadHost(v3dm_ctx *c, Side &s, void *dst, const void *src, size_t bytes)
{
   if (bytes == 0) return V3DM_OK;
   if (!s.hStage)
      HIPCHK(c, hipHostMalloc(&s.hStage, sizeof(float) * (size_t)std::max(s.nMax, 1) * kDim, hipHostMallocDefault));
   memcpy(s.hStage, src, bytes);
   HIPCHK(c, hipMemcpyAsync(dst, s.hStage, bytes, hipMemcpyHostToDevice, c->stream));
   return V3DM_OK;
}

constexpr int kMaxProfiledJobs = 65536;

// Events of a profiled job: 0 and 1 bracket the score kernel, 2 follows the
// selection.  Creates the events of the current job when needed.
static int profEnsure(v3dm_ctx *c)
{
   const size_t need = (size_t)c->profJobs * 3 + 3;
   while (c->events.size() < need)
   {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c->events.push_back(e);
   }
   return V3DM_OK;
}

// Records event `which` (0, 1, 2) of the current profiled job.
int profMark(v3dm_ctx *c, int which)
{
   if (!c->profile || c->profJobs >= kMaxProfiledJobs) return V3DM_OK;
   int rc = profEnsure(c);
   if (rc) return rc;
   HIPCHK(c, hipEventRecord(c->events[(size_t)c->profJobs * 3 + which], c->stream));
   if (which == 2) c->profJobs++;
   return V3DM_OK;
}

// Events 0 and 1 of the current profiled job, to be bound to the score
// kernel's dispatch (its own start and end timestamps); null when not profiling.
int profKernelEvents(v3dm_ctx *c, hipEvent_t *start, hipEvent_t *stop)
{
   *start = *stop = nullptr;
   if (!c->profile || c->profJobs >= kMaxProfiledJobs) return V3DM_OK;
   int rc = profEnsure(c);
   if (rc) return rc;
   *start = c->events[(size_t)c->profJobs * 3];
   *stop = c->events[(size_t)c->profJobs * 3 + 1];
   return V3DM_OK;
}

void freeAll(v3dm_ctx *c)
{
   freeLearned(c);
   freeSide(c->L);
   freeSide(c->R);
   dfree(c->dPair);
   dfree(c->dEpoch);
   dfree(c->dMatch);
   dfree(c->dFlag);
   dfree(c->dMasks);
   dfree(c->dS);
   dfree(c->dPart);
   dfree(c->dColMax);
   dfree(c->dRowMax);
   dfree(c->dSurv);
   dfree(c->dSurvCount);
   dfree(c->dNorm);
   if (c->hMatch) (void)hipHostFree(c->hMatch);
   if (c->hFlag) (void)hipHostFree(c->hFlag);
   c->hMatch = nullptr;
   c->hFlag = nullptr;
   c->sBytes = c->partBytes = 0;
   c->allocated = false;
   c->pairValid = false;
   c->sValid = false;
   c->masks.clear();
   c->keyFormat = 0;
}

int allocSide(v3dm_ctx *c, Side &s, int nMax, bool guided)
{
   s.nMax = nMax;
   const int nPad = std::max(roundUp(nMax, kRowPad), kRowPad);
   HIPCHK(c, hipMalloc((void **)&s.dI8, (size_t)nPad * kDim));
   HIPCHK(c, hipMalloc((void **)&s.dSum, sizeof(int32_t) * nPad));
   HIPCHK(c, hipMalloc((void **)&s.dTop1, sizeof(uint64_t) * nPad));
   HIPCHK(c, hipMalloc((void **)&s.dTop2, sizeof(uint64_t) * nPad));
   if (guided) HIPCHK(c, hipMalloc((void **)&s.dPos, sizeof(float) * 2 * (size_t)std::max(nMax, 1) * kMaxMasks));
   return V3DM_OK;
}

int clearKeys(v3dm_ctx *c)
{
   const int nPadL = std::max(roundUp(c->L.nMax, kRowPad), kRowPad);
   const int nPadR = std::max(roundUp(c->R.nMax, kRowPad), kRowPad);
   HIPCHK(c, launch_clear_keys(c->L.dTop1, nPadL, c->dEpoch, c->dEpoch + 1, c->stream));
   HIPCHK(c, launch_clear_keys(c->L.dTop2, nPadL, c->dEpoch, c->dEpoch + 1, c->stream));
   HIPCHK(c, launch_clear_keys(c->R.dTop1, nPadR, c->dEpoch, c->dEpoch + 1, c->stream));
   HIPCHK(c, launch_clear_keys(c->R.dTop2, nPadR, c->dEpoch, c->dEpoch + 1, c->stream));
   c->hostEpoch = 1;
   return V3DM_OK;
}

// The top-2 arrays hold keys of one format; a job of the other format
// starts from cleared arrays.
int useKeyFormat(v3dm_ctx *c, int format)
{
   const uint32_t limit = format == 1 ? kEpochLimitF : kEpochLimit;
   if (c->keyFormat != format || c->hostEpoch >= limit)
   {
      int rc = clearKeys(c);
      if (rc) return rc;
      c->keyFormat = format;
   }
   return V3DM_OK;
}

int roundFeatures(v3dm_ctx *c, int n, const char *who, int nMax)
{
   // v3d_cudamatching.cpp:68-77 / 84-93
   if ((n % 32) != 0)
   {
      if (c->warnRounding)
      {
         fprintf(stderr, "SIFT_Matching::%s() warning: nFeatures should be a multiple of 32\n", who);
         fprintf(stderr, "Rounding nFeatures down.\n");
      }
      n = n - (n & 31);
   }
   return std::min(n, nMax);
}

int syncStream(v3dm_ctx *c)
{
   HIPCHK(c, hipStreamSynchronize(c->stream));
   return V3DM_OK;
}

int setFeatures(v3dm_ctx *c, Side &s, int n, const void *desc, bool u8, const char *who)
{
   if (!c) return V3DM_ERR_ARG;
   if (!c->allocated) return fail(c, V3DM_ERR_STATE, "%s: allocate() has not been called", who);
   if (n < 0) return fail(c, V3DM_ERR_ARG, "%s: negative feature count %d", who, n);
   n = roundFeatures(c, n, who, s.nMax);
   if (n > 0 && !desc) return fail(c, V3DM_ERR_ARG, "%s: descriptor pointer is NULL", who);
   HIPCHK(c, hipSetDevice(c->device));
   if (c->pendingAsync)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      c->pendingAsync = false;
   }
   const int nPad = std::max(roundUp(std::max(n, 1), kRowPad), kRowPad);
   s.n = n;
   s.isU8 = u8;
   s.bfValid = false;
   if (u8)
   {
      if (!s.dRawU8) HIPCHK(c, hipMalloc((void **)&s.dRawU8, (size_t)std::max(s.nMax, 1) * kDim));
      int rc = uploadHost(c, s, s.dRawU8, desc, (size_t)n * kDim);
      if (rc) return rc;
      HIPCHK(c, launch_prep_u8(s.dRawU8, n, nPad, s.dI8, s.dSum, c->stream));
      s.quantized = true;
      s.f32Valid = false;
      rc = syncStream(c);
      if (rc) return rc;
   }
   else
   {
      if (!s.dF32) HIPCHK(c, hipMalloc((void **)&s.dF32, sizeof(float) * (size_t)std::max(s.nMax, 1) * kDim));
      int rc = uploadHost(c, s, s.dF32, desc, sizeof(float) * (size_t)n * kDim);
      if (rc) return rc;
      *c->hFlag = 0;
      HIPCHK(c, hipMemsetAsync(c->dFlag, 0, sizeof(int), c->stream));
      HIPCHK(c, launch_prep_f32(s.dF32, n, nPad, s.dI8, s.dSum, c->dFlag, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->hFlag, c->dFlag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      rc = syncStream(c);
      if (rc) return rc;
      s.quantized = (*c->hFlag == 0);
      s.f32Valid = true;
   }
   c->sValid = false;
   return V3DM_OK;
}

int ensureF32(v3dm_ctx *c, Side &s)
{
   if (s.f32Valid) return V3DM_OK;
   if (!s.dF32) HIPCHK(c, hipMalloc((void **)&s.dF32, sizeof(float) * (size_t)std::max(s.nMax, 1) * kDim));
   HIPCHK(c, launch_u8_to_f32(s.dRawU8, (int64_t)s.n * kDim, s.dF32, c->stream));
   s.f32Valid = true;
   return V3DM_OK;
}

// BF16 pool and largest squared row norm of a side (float path), from its
// FP32 form.  Synchronous: the norm feeds the candidate bound on the host.
int ensureBF16(v3dm_ctx *c, Side &s)
{
   if (s.bfValid) return V3DM_OK;
   int rc = ensureF32(c, s);
   if (rc) return rc;
   const int nPad = std::max(roundUp(std::max(s.nMax, 1), kRowPad), kRowPad);
   if (!s.dBF) HIPCHK(c, hipMalloc((void **)&s.dBF, sizeof(uint16_t) * (size_t)nPad * kDim));
   const int nPadN = std::max(roundUp(std::max(s.n, 1), kRowPad), kRowPad);
   HIPCHK(c, launch_prep_bf16(s.dF32, s.n, nPadN, s.dBF, c->dNorm, c->stream));
   uint32_t bits = 0;
   HIPCHK(c, hipMemcpyAsync(c->hFlag, c->dNorm, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
   rc = syncStream(c);
   if (rc) return rc;
   memcpy(&bits, c->hFlag, sizeof(bits));
   memcpy(&s.norm2Max, &bits, sizeof(bits));
   s.bfValid = true;
   return V3DM_OK;
}

int uploadPair(v3dm_ctx *c, int variant)
{
   PairDesc pd{};
   pd.offL = 0;
   pd.offR = 0;
   pd.slotL = 0;
   pd.slotR = 0;
   pd.nL = c->L.n;
   pd.nR = c->R.n;
   pd.tilesM = (roundUp(pd.nL, kRowPad) + scan_variant_tile_m(variant) - 1) / scan_variant_tile_m(variant);
   pd.tilesN = (roundUp(pd.nR, kRowPad) + scan_variant_tile_n(variant) - 1) / scan_variant_tile_n(variant);
   pd.tileBase = 0;
   pd.finBase = 0;
   if (c->pairValid && memcmp(&pd, &c->hPairUploaded, sizeof(pd)) == 0) return V3DM_OK;
   if (c->pendingAsync)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      c->pendingAsync = false;
   }
   HIPCHK(c, hipMemcpyAsync(c->dPair, &pd, sizeof(pd), hipMemcpyHostToDevice, c->stream));
   int rc = syncStream(c);
   if (rc) return rc;
   c->hPairUploaded = pd;
   c->pairValid = true;
   return V3DM_OK;
}

int uploadMasks(v3dm_ctx *c)
{
   if (c->masks.empty()) return V3DM_OK;
   HIPCHK(c, hipMemcpyAsync(c->dMasks, c->masks.data(), sizeof(Mask) * c->masks.size(), hipMemcpyHostToDevice,
                            c->stream));
   return syncStream(c);
}

// The automatic choice: the int8 scan for u/512 descriptors, else BF16 +
// FP32 rescoring -- unless the rule admits no survivor filter (a small
// positive minRatio: every score could be a deciding second best), where
// every region of the scan would overflow and be recomputed pair by pair;
// the plain path (score matrix + selection passes, the same list) is then
// the faster one by far.
int chooseKernel(v3dm_ctx *c, int *kernel, float minScore, float minRatio)
{
   const bool q = c->L.quantized && c->R.quantized;
   int k = c->kernelSel;
   if (k == V3DM_KERNEL_AUTO)
   {
      const bool unfiltered = q ? survivorThreshold(minScore, minRatio) < 0
                                : survivorThresholdF(minScore, minRatio) == -INFINITY;
      k = unfiltered ? V3DM_KERNEL_PLAIN : q ? V3DM_KERNEL_I8 : V3DM_KERNEL_BF16;
   }
   if (k == V3DM_KERNEL_I8 && !q)
      return fail(c, V3DM_ERR_UNSUPPORTED,
                  "int8 kernel requested but the descriptors are not all of the form u/512, u in [0,255]");
   *kernel = k;
   return V3DM_OK;
}

// Enqueue the i8 job for the current pair (masks already uploaded).
// One int8 job with scan variant v; marks: record the profiling events.
int enqueueI8Job(v3dm_ctx *c, int v, float minScore, float minRatio, bool marks)
{
   int rc = uploadPair(c, v);
   if (rc) return rc;
   rc = useKeyFormat(c, 0);
   if (rc) return rc;
   const PairDesc &pd = c->hPairUploaded;
   ScanArgs a{};
   a.pair0 = pd;
   a.L = c->L.dI8;
   a.Lsum = c->L.dSum;
   a.R = c->R.dI8;
   a.Rsum = c->R.dSum;
   a.pairs = c->dPair;
   a.nPairs = 1;
   a.nTiles = (pd.nL > 0 && pd.nR > 0) ? pd.tilesM * pd.tilesN : 0;
   a.thr = survivorThreshold(minScore, minRatio);
   a.top1L = c->L.dTop1;
   a.top2L = c->L.dTop2;
   a.top1R = c->R.dTop1;
   a.top2R = c->R.dTop2;
   a.epoch = nullptr; // by value: the host counts single-pair jobs
   a.epochValue = c->hostEpoch;
   a.masks = c->dMasks;
   a.nMasks = (int)c->masks.size();
   a.surv = c->dSurv;
   a.survCount = c->dSurvCount;
   a.survCap = scan_variant_surv_cap(v);
   // a grid the device holds at once applies its survivors in the scan; a
   // larger one leaves them to k_i8_insert (a wave's tail would hold its slot)
   a.inlineInsert = a.nTiles <= scan_variant_resident(v, c->device) ? 1 : 0;
   a.waves = scan_variant_waves(v);
   a.tileM = scan_variant_tile_m(v);
   a.tileN = scan_variant_tile_n(v);
   a.persistent = 0;
   a.tileCounter = c->dEpoch + 2;
   hipEvent_t e0 = nullptr, e1 = nullptr;
   if (marks && (rc = profKernelEvents(c, &e0, &e1))) return rc;
   if (a.nTiles <= 0 && e0)
   {
      HIPCHK(c, hipEventRecord(e0, c->stream));
      HIPCHK(c, hipEventRecord(e1, c->stream));
   }
   HIPCHK(c, launch_i8_scan(v, a.nTiles, a, c->stream, e0, e1));
   c->lastInsertForm = -1;
   if (!a.inlineInsert)
   {
      c->lastInsertForm = insert_form_for((int64_t)a.nTiles * a.waves, c->device, c->insertForm);
      HIPCHK(c, launch_i8_insert(a, c->lastInsertForm, c->stream));
   }
   FinalizeArgs f{};
   f.pairs = c->dPair;
   f.nPairs = 1;
   f.nBlocks = std::max(1, (pd.nL + 255) / 256);
   f.top1L = c->L.dTop1;
   f.top2L = c->L.dTop2;
   f.top1R = c->R.dTop1;
   f.top2R = c->R.dTop2;
   f.epoch = nullptr; // by value, and no last-block epoch advance (single-pair scans are never persistent)
   f.epochValue = c->hostEpoch;
   f.done = nullptr;
   f.tileCounter = nullptr;
   f.minScore = minScore;
   f.minRatio = minRatio;
   f.matchOfLeft = c->dMatch;
   f.pair0 = pd;
   f.hasPair0 = 1;
   HIPCHK(c, launch_i8_finalize(f, c->stream));
   if (marks && (rc = profMark(c, 2))) return rc;
   c->hostEpoch += 1;
   c->lastVariant = v;
   c->lastInline = a.inlineInsert;
   return V3DM_OK;
}

// Scan variant of the single-pair int8 path: the forced one, or the fastest
// measured (whole jobs, HIP events) for this shape bucket -- each side's
// count rounded up to a power of two -- cached in the context.
int pickI8Variant(v3dm_ctx *c, float minScore, float minRatio, int *variant)
{
   *variant = 0;
   if (c->scanVariant >= 0)
   {
      *variant = c->scanVariant;
      return V3DM_OK;
   }
   if (c->L.n == 0 || c->R.n == 0) return V3DM_OK;
   auto bucket = [](int n) {
      uint64_t b = kRowPad;
      while (b < (uint64_t)n) b <<= 1;
      return b;
   };
   const uint64_t key = (bucket(c->L.n) << 32) | bucket(c->R.n);
   for (const auto &t : c->tunedI8)
      if (t.first == key)
      {
         *variant = t.second;
         return V3DM_OK;
      }
   hipEvent_t t0, t1;
   HIPCHK(c, hipEventCreate(&t0));
   HIPCHK(c, hipEventCreate(&t1));
   float best = 0.0f;
   int rc = V3DM_OK;
   // per variant: two warm-up jobs, then the best of three trials of four jobs
   for (int v = 0; v < scan_variant_count() && rc == V3DM_OK; ++v)
   {
      for (int r = 0; r < 2 && rc == V3DM_OK; ++r) rc = enqueueI8Job(c, v, minScore, minRatio, false);
      float vbest = 0.0f;
      for (int trial = 0; trial < 3 && rc == V3DM_OK; ++trial)
      {
         if (hipEventRecord(t0, c->stream) != hipSuccess) rc = fail(c, V3DM_ERR_DEVICE, "autotune: event record");
         for (int r = 0; r < 4 && rc == V3DM_OK; ++r) rc = enqueueI8Job(c, v, minScore, minRatio, false);
         if (rc) break;
         float ms = 0.0f;
         if (hipEventRecord(t1, c->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
             hipEventElapsedTime(&ms, t0, t1) != hipSuccess)
            rc = fail(c, V3DM_ERR_DEVICE, "autotune: event timing");
         else if (trial == 0 || ms < vbest)
            vbest = ms;
      }
      if (rc == V3DM_OK && (v == 0 || vbest < best))
      {
         best = vbest;
         *variant = v;
      }
   }
   (void)hipEventDestroy(t0);
   (void)hipEventDestroy(t1);
   if (rc) return rc;
   c->tunedI8.push_back({key, *variant});
   return V3DM_OK;
}

int enqueueI8(v3dm_ctx *c, float minScore, float minRatio)
{
   int v = 0;
   int rc = pickI8Variant(c, minScore, minRatio, &v);
   if (rc) return rc;
   return enqueueI8Job(c, v, minScore, minRatio, true);
}

// Float path: BF16 candidate scan, FP32 rescoring, finalize (masks already uploaded).
int enqueueBF16(v3dm_ctx *c, float minScore, float minRatio)
{
   int rc = ensureBF16(c, c->L);
   if (rc) return rc;
   rc = ensureBF16(c, c->R);
   if (rc) return rc;
   const int v = c->bfVariant;
   rc = uploadPair(c, v);
   if (rc) return rc;
   rc = useKeyFormat(c, 1);
   if (rc) return rc;
   const PairDesc &pd = c->hPairUploaded;
   ScanArgsF a{};
   a.pair0 = pd;
   a.L = c->L.dBF;
   a.R = c->R.dBF;
   a.Lf = c->L.dF32;
   a.Rf = c->R.dF32;
   a.pairs = c->dPair;
   a.nPairs = 1;
   a.nTiles = (pd.nL > 0 && pd.nR > 0) ? pd.tilesM * pd.tilesN : 0;
   a.thr = survivorThresholdF(minScore, minRatio);
   a.lim = bf16CandidateLimit(a.thr, c->L.norm2Max, c->R.norm2Max);
   c->lastThr = a.thr;
   c->lastLim = a.lim;
   a.top1L = c->L.dTop1;
   a.top2L = c->L.dTop2;
   a.top1R = c->R.dTop1;
   a.top2R = c->R.dTop2;
   a.epoch = nullptr; // by value: the host counts single-pair jobs
   a.epochValue = c->hostEpoch;
   a.masks = c->dMasks;
   a.nMasks = (int)c->masks.size();
   a.surv = c->dSurv;
   a.survCount = c->dSurvCount;
   a.survCap = c->survCap;
   a.waves = scan_variant_waves(v);
   a.tileM = scan_variant_tile_m(v);
   a.tileN = scan_variant_tile_n(v);
   a.persistent = 0;
   a.tileCounter = c->dEpoch + 2;
   hipEvent_t e0, e1;
   if ((rc = profKernelEvents(c, &e0, &e1))) return rc;
   if (a.nTiles <= 0 && e0)
   {
      HIPCHK(c, hipEventRecord(e0, c->stream));
      HIPCHK(c, hipEventRecord(e1, c->stream));
   }
   HIPCHK(c, launch_bf16_scan(v, a.nTiles, a, c->stream, e0, e1));
   HIPCHK(c, launch_f32_rescore(a, c->stream));
   FinalizeArgs f{};
   f.pairs = c->dPair;
   f.nPairs = 1;
   f.nBlocks = std::max(1, (pd.nL + 255) / 256);
   f.top1L = c->L.dTop1;
   f.top2L = c->L.dTop2;
   f.top1R = c->R.dTop1;
   f.top2R = c->R.dTop2;
   f.epoch = nullptr; // by value, and no last-block epoch advance (single-pair scans are never persistent)
   f.epochValue = c->hostEpoch;
   f.done = nullptr;
   f.tileCounter = nullptr;
   f.minScore = minScore;
   f.minRatio = minRatio;
   f.matchOfLeft = c->dMatch;
   f.pair0 = pd;
   f.hasPair0 = 1;
   HIPCHK(c, launch_f32_finalize(f, c->stream));
   if ((rc = profMark(c, 2))) return rc;
   c->hostEpoch += 1;
   return V3DM_OK;
}

// Plain path: scores (when stale), pending masks, selection passes.
int enqueuePlain(v3dm_ctx *c, float minScore, float minRatio)
{
   const int nL = c->L.n, nR = c->R.n;
   int rc = ensureF32(c, c->L);
   if (rc) return rc;
   rc = ensureF32(c, c->R);
   if (rc) return rc;
   const size_t sNeed = sizeof(float) * (size_t)std::max(nL, 1) * (size_t)std::max(nR, 1);
   if (sNeed > c->sBytes)
   {
      if (c->pendingAsync)
      {
         rc = syncStream(c);
         if (rc) return rc;
         c->pendingAsync = false;
      }
      dfree(c->dS);
      c->sBytes = 0;
      c->sValid = false;
      HIPCHK(c, hipMalloc((void **)&c->dS, sNeed));
      c->sBytes = sNeed;
   }
   const size_t pNeed = plain_part_bytes(nL, nR, nullptr, nullptr);
   if (pNeed > c->partBytes)
   {
      dfree(c->dPart);
      c->partBytes = 0;
      HIPCHK(c, hipMalloc(&c->dPart, pNeed));
      c->partBytes = pNeed;
   }
   hipEvent_t e0, e1;
   if ((rc = profKernelEvents(c, &e0, &e1))) return rc;
   if (!c->sValid)
   {
      HIPCHK(c, launch_sgemm_seq(nL, nR, c->L.dF32, c->R.dF32, c->dS, c->stream, e0, e1));
      c->sValid = true;
      c->sMasksApplied = 0;
   }
   else if (e0)
   {
      HIPCHK(c, hipEventRecord(e0, c->stream));
      HIPCHK(c, hipEventRecord(e1, c->stream));
   }
   const int nm = (int)c->masks.size();
   if (nm > c->sMasksApplied)
   {
      HIPCHK(c, launch_mask_plain(nL, nR, c->dS, c->dMasks + c->sMasksApplied, nm - c->sMasksApplied, c->stream));
      c->sMasksApplied = nm;
   }
   HIPCHK(c, launch_plain_select(nL, nR, c->dS, c->dPart, minScore, minRatio, c->dColMax, c->dRowMax, c->dMatch,
                                 c->stream));
   return profMark(c, 2);
}

int runMatch(v3dm_ctx *c, float minScore, float minRatio, bool fromFind)
{
   if (!c->allocated) return fail(c, V3DM_ERR_STATE, "findPutativeMatches: allocate() has not been called");
   HIPCHK(c, hipSetDevice(c->device));
   int kernel = 0;
   int rc = chooseKernel(c, &kernel, minScore, minRatio);
   if (rc) return rc;
   if (fromFind)
   {
      // findPutativeMatches recomputes the score matrix: earlier guided masks
      // no longer apply (v3d_cudamatching.cpp:104-105).
      c->masks.clear();
      c->sValid = false;
   }
   if (kernel == V3DM_KERNEL_I8) rc = enqueueI8(c, minScore, minRatio);
   else if (kernel == V3DM_KERNEL_BF16) rc = enqueueBF16(c, minScore, minRatio);
   else rc = enqueuePlain(c, minScore, minRatio);
   if (rc) return rc;
   c->lastKernel = kernel;
   c->lastNL = c->L.n;
   c->pendingAsync = true;
   return V3DM_OK;
}

int fetch(v3dm_ctx *c, int32_t *pairs, int capacity, int *count)
{
   const int nL = c->lastNL;
   if (nL > 0)
      HIPCHK(c, hipMemcpyAsync(c->hMatch, c->dMatch, sizeof(int32_t) * nL, hipMemcpyDeviceToHost, c->stream));
   int rc = syncStream(c);
   if (rc) return rc;
   c->pendingAsync = false;
   int n = 0;
   for (int k = 0; k < nL; ++k)
   {
      const int l = c->hMatch[k];
      if (l < 0) continue;
      if (n < capacity && pairs)
      {
         pairs[2 * n] = k;
         pairs[2 * n + 1] = l;
      }
      ++n;
   }
   if (count) *count = n;
   if (n > capacity) return fail(c, V3DM_ERR_CAPACITY, "match list has %d pairs, capacity is %d", n, capacity);
   return V3DM_OK;
}

int guidedCommon(v3dm_ctx *c, int kind, const double M[9], float distanceThreshold, const float *leftPos,
                 const float *rightPos, float minScore, float minRatio, int32_t *pairs, int capacity, int *count)
{
   if (!c) return V3DM_ERR_ARG;
   const char *who = kind == 1 ? "runGuidedMatching" : "runGuidedHomographyMatching";
   if (!c->allocated) return fail(c, V3DM_ERR_STATE, "%s: allocate() has not been called", who);
   if (!c->guided)
      return fail(c, V3DM_ERR_STATE, "%s: allocate() was called without useGuidedMatching", who);
   if (!M) return fail(c, V3DM_ERR_ARG, "%s: matrix pointer is NULL", who);
   if ((c->L.n > 0 && !leftPos) || (c->R.n > 0 && !rightPos))
      return fail(c, V3DM_ERR_ARG, "%s: position pointer is NULL", who);
   if ((int)c->masks.size() >= kMaxMasks)
      return fail(c, V3DM_ERR_UNSUPPORTED, "%s: more than %d guided calls since the last findPutativeMatches", who,
                  kMaxMasks);
   HIPCHK(c, hipSetDevice(c->device));
   if (c->pendingAsync)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      c->pendingAsync = false;
   }
   const int slot = (int)c->masks.size();
   float *lp = c->L.dPos + (size_t)slot * 2 * std::max(c->L.nMax, 1);
   float *rp = c->R.dPos + (size_t)slot * 2 * std::max(c->R.nMax, 1);
   if (c->L.n > 0)
      HIPCHK(c, hipMemcpyAsync(lp, leftPos, sizeof(float) * 2 * c->L.n, hipMemcpyHostToDevice, c->stream));
   if (c->R.n > 0)
      HIPCHK(c, hipMemcpyAsync(rp, rightPos, sizeof(float) * 2 * c->R.n, hipMemcpyHostToDevice, c->stream));
   Mask mk{};
   mk.kind = kind;
   mk.sqrThr = distanceThreshold * distanceThreshold; // v3d_cudamatching.cu:414 (float)
   for (int i = 0; i < 9; ++i) mk.m[i] = (float)M[i];  // v3d_cudamatching.cpp:161-164
   mk.lpos = lp;
   mk.rpos = rp;
   c->masks.push_back(mk);
   int rc = uploadMasks(c);
   if (rc) return rc;
   rc = runMatch(c, minScore, minRatio, false);
   if (rc) return rc;
   return fetch(c, pairs, capacity, count);
}

} // namespace

// ============================================================================

extern "C" {

int v3dm_create(int device, v3dm_ctx **out)
{
   if (!out) return fail(nullptr, V3DM_ERR_ARG, "v3dm_create: out is NULL");
   *out = nullptr;
   int n = 0;
   hipError_t e = hipGetDeviceCount(&n);
   if (e != hipSuccess || n <= 0)
      return fail(nullptr, V3DM_ERR_DEVICE, "v3dm_create: no HIP device available (%s)",
                  e != hipSuccess ? hipGetErrorString(e) : "0 devices");
   if (device < 0 || device >= n)
      return fail(nullptr, V3DM_ERR_ARG, "v3dm_create: device %d out of range [0, %d)", device, n);
   e = hipSetDevice(device);
   if (e != hipSuccess) return fail(nullptr, V3DM_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
   v3dm_ctx *c = new v3dm_ctx();
   c->device = device;
   e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
   if (e != hipSuccess)
   {
      delete c;
      return fail(nullptr, V3DM_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
   }
   *out = c;
   return V3DM_OK;
}

void v3dm_destroy(v3dm_ctx *c)
{
   if (!c) return;
   (void)hipSetDevice(c->device);
   if (c->stream) (void)hipStreamSynchronize(c->stream);
   freeAll(c);
   freeBatch(c);
   for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
   if (c->stream) (void)hipStreamDestroy(c->stream);
   delete c;
}

const char *v3dm_last_error(const v3dm_ctx *c) { return c ? c->err.c_str() : g_createError.c_str(); }

int v3dm_set_rounding_warning(v3dm_ctx *c, int flag)
{
   if (!c) return V3DM_ERR_ARG;
   c->warnRounding = flag != 0;
   return V3DM_OK;
}

int v3dm_set_reduction_path(v3dm_ctx *c, int flag)
{
   if (!c) return V3DM_ERR_ARG;
   c->useReduction = flag != 0;
   return V3DM_OK;
}

int v3dm_allocate(v3dm_ctx *c, int nMaxLeft, int nMaxRight, int useGuided)
{
   if (!c) return V3DM_ERR_ARG;
   if (nMaxLeft < 0 || nMaxRight < 0)
      return fail(c, V3DM_ERR_ARG, "allocate: negative maximum (%d, %d)", nMaxLeft, nMaxRight);
   // v3d_cudamatching.cpp:19-35
   if (nMaxLeft % 32)
   {
      if (c->warnRounding)
      {
         fprintf(stderr, "SIFT_Matching::allocate() warning: nMaxLeftFeatures should be a multiple of 32\n");
         fprintf(stderr, "Rounding nMaxLeftFeatures down.\n");
      }
      nMaxLeft -= nMaxLeft & 31;
   }
   if (nMaxRight % 32)
   {
      if (c->warnRounding)
      {
         fprintf(stderr, "SIFT_Matching::allocate() warning: nMaxRightFeatures should be a multiple of 32\n");
         fprintf(stderr, "Rounding nMaxRightFeatures down.\n");
      }
      nMaxRight -= nMaxRight & 31;
   }
   if (nMaxLeft > kMaxLocalIndex || nMaxRight > kMaxLocalIndex)
      return fail(c, V3DM_ERR_UNSUPPORTED, "allocate: at most %d features per image", kMaxLocalIndex);
   HIPCHK(c, hipSetDevice(c->device));
   if (c->allocated)
   {
      (void)hipStreamSynchronize(c->stream);
      freeAll(c);
   }
   c->guided = useGuided != 0;
   int rc = allocSide(c, c->L, nMaxLeft, c->guided);
   if (!rc) rc = allocSide(c, c->R, nMaxRight, c->guided);
   if (rc)
   {
      freeAll(c);
      return rc;
   }
   const int nPadL = std::max(roundUp(nMaxLeft, kRowPad), kRowPad);
   HIPCHK(c, hipMalloc((void **)&c->dNorm, sizeof(uint32_t)));
   HIPCHK(c, hipMalloc((void **)&c->dPair, sizeof(PairDesc)));
   HIPCHK(c, hipMalloc((void **)&c->dEpoch, 4 * sizeof(uint32_t))); // epoch, done, tile counter
   HIPCHK(c, hipMalloc((void **)&c->dMatch, sizeof(int32_t) * nPadL));
   HIPCHK(c, hipMalloc((void **)&c->dFlag, sizeof(int)));
   HIPCHK(c, hipMalloc((void **)&c->dMasks, sizeof(Mask) * kMaxMasks));
   HIPCHK(c, hipMalloc((void **)&c->dColMax, sizeof(int32_t) * nPadL));
   HIPCHK(c, hipMalloc((void **)&c->dRowMax, sizeof(int32_t) * std::max(roundUp(nMaxRight, kRowPad), kRowPad)));
   {
      // survivor regions (and entries) for the largest count any scan variant needs
      int64_t regions = 0, entries = 0;
      for (int v = 0; v < scan_variant_count(); ++v)
      {
         const int64_t tiles =
            (int64_t)((roundUp(std::max(nMaxLeft, 1), kRowPad) + scan_variant_tile_m(v) - 1) / scan_variant_tile_m(v)) *
            ((roundUp(std::max(nMaxRight, 1), kRowPad) + scan_variant_tile_n(v) - 1) / scan_variant_tile_n(v));
         regions = std::max(regions, tiles * scan_variant_waves(v));
         entries = std::max(entries, tiles * scan_variant_waves(v) *
                                        std::max<int64_t>(scan_variant_surv_cap(v), c->survCap));
      }
      HIPCHK(c, hipMalloc((void **)&c->dSurv, sizeof(uint64_t) * (size_t)entries));
      HIPCHK(c, hipMalloc((void **)&c->dSurvCount, sizeof(int32_t) * (size_t)regions));
   }
   HIPCHK(c, hipHostMalloc((void **)&c->hMatch, sizeof(int32_t) * nPadL, hipHostMallocDefault));
   HIPCHK(c, hipHostMalloc((void **)&c->hFlag, sizeof(int), hipHostMallocDefault));
   c->allocated = true;
   int r2 = clearKeys(c);
   if (r2) return r2;
   return syncStream(c);
}

int v3dm_deallocate(v3dm_ctx *c)
{
   if (!c) return V3DM_ERR_ARG;
   (void)hipSetDevice(c->device);
   if (c->stream) (void)hipStreamSynchronize(c->stream);
   freeAll(c);
   return V3DM_OK;
}

int v3dm_set_left_f32(v3dm_ctx *c, int n, const float *d)
{
   return c ? setFeatures(c, c->L, n, d, false, "setLeftFeatures") : V3DM_ERR_ARG;
}
int v3dm_set_right_f32(v3dm_ctx *c, int n, const float *d)
{
   return c ? setFeatures(c, c->R, n, d, false, "setRightFeatures") : V3DM_ERR_ARG;
}
int v3dm_set_left_u8(v3dm_ctx *c, int n, const uint8_t *d)
{
   return c ? setFeatures(c, c->L, n, d, true, "setLeftFeatures") : V3DM_ERR_ARG;
}
int v3dm_set_right_u8(v3dm_ctx *c, int n, const uint8_t *d)
{
   return c ? setFeatures(c, c->R, n, d, true, "setRightFeatures") : V3DM_ERR_ARG;
}

int v3dm_get_counts(const v3dm_ctx *c, int *nLeft, int *nRight)
{
   if (!c) return V3DM_ERR_ARG;
   if (nLeft) *nLeft = c->L.n;
   if (nRight) *nRight = c->R.n;
   return V3DM_OK;
}

int v3dm_get_device_buffers(const v3dm_ctx *c, const float **left, const float **right, const float **scores,
                            const float **leftPos, const float **rightPos)
{
   if (!c) return V3DM_ERR_ARG;
   if (left) *left = c->L.f32Valid ? c->L.dF32 : nullptr;
   if (right) *right = c->R.f32Valid ? c->R.dF32 : nullptr;
   if (scores) *scores = c->sValid ? c->dS : nullptr;
   const bool guided = !c->masks.empty();
   if (leftPos) *leftPos = guided ? c->L.dPos : nullptr;
   if (rightPos) *rightPos = guided ? c->R.dPos : nullptr;
   return V3DM_OK;
}

int v3dm_find_putative_matches(v3dm_ctx *c, float minScore, float minRatio, int32_t *pairs, int capacity,
                               int *count)
{
   if (!c) return V3DM_ERR_ARG;
   if (capacity < 0) return fail(c, V3DM_ERR_ARG, "findPutativeMatches: negative capacity");
   int rc = runMatch(c, minScore, minRatio, true);
   if (rc) return rc;
   return fetch(c, pairs, capacity, count);
}

int v3dm_run_guided_matching(v3dm_ctx *c, const double F[9], float thr, const float *lp, const float *rp,
                             float minScore, float minRatio, int32_t *pairs, int capacity, int *count)
{
   return guidedCommon(c, 1, F, thr, lp, rp, minScore, minRatio, pairs, capacity, count);
}

int v3dm_run_guided_homography_matching(v3dm_ctx *c, const double H[9], float thr, const float *lp,
                                        const float *rp, float minScore, float minRatio, int32_t *pairs,
                                        int capacity, int *count)
{
   return guidedCommon(c, 2, H, thr, lp, rp, minScore, minRatio, pairs, capacity, count);
}

int v3dm_set_kernel(v3dm_ctx *c, int kernel)
{
   if (!c) return V3DM_ERR_ARG;
   if (kernel < V3DM_KERNEL_AUTO || kernel > V3DM_KERNEL_BF16)
      return fail(c, V3DM_ERR_ARG, "set_kernel: unknown kernel %d", kernel);
   c->kernelSel = kernel;
   return V3DM_OK;
}

int v3dm_set_scan_variant(v3dm_ctx *c, int variant)
{
   if (!c) return V3DM_ERR_ARG;
   if (variant < -1 || variant >= scan_variant_count())
      return fail(c, V3DM_ERR_ARG, "set_scan_variant: variant %d out of range", variant);
   c->scanVariant = variant;
   return V3DM_OK;
}

int v3dm_get_last_scan_variant(const v3dm_ctx *c, int *variant)
{
   if (!c || !variant) return V3DM_ERR_ARG;
   *variant = c->lastVariant;
   return V3DM_OK;
}

int v3dm_get_last_scan_inline(const v3dm_ctx *c, int *inlineInsert)
{
   if (!c || !inlineInsert) return V3DM_ERR_ARG;
   *inlineInsert = c->lastInline;
   return V3DM_OK;
}

int v3dm_set_insert_form(v3dm_ctx *c, int form)
{
   if (!c) return V3DM_ERR_ARG;
   if (form < -1 || form > 1) return fail(c, V3DM_ERR_ARG, "set_insert_form: form %d out of range", form);
   c->insertForm = form;
   return V3DM_OK;
}

int v3dm_get_last_insert_form(const v3dm_ctx *c, int *form)
{
   if (!c || !form) return V3DM_ERR_ARG;
   *form = c->lastInsertForm;
   return V3DM_OK;
}

int v3dm_get_last_kernel(const v3dm_ctx *c, int *kernel)
{
   if (!c || !kernel) return V3DM_ERR_ARG;
   *kernel = c->lastKernel;
   return V3DM_OK;
}

int v3dm_get_stream(const v3dm_ctx *c, void **stream)
{
   if (!c || !stream) return V3DM_ERR_ARG;
   *stream = (void *)c->stream;
   return V3DM_OK;
}

int v3dm_find_async(v3dm_ctx *c, float minScore, float minRatio)
{
   if (!c) return V3DM_ERR_ARG;
   return runMatch(c, minScore, minRatio, true);
}

int v3dm_fetch_matches(v3dm_ctx *c, int32_t *pairs, int capacity, int *count)
{
   if (!c) return V3DM_ERR_ARG;
   if (capacity < 0) return fail(c, V3DM_ERR_ARG, "fetch_matches: negative capacity");
   HIPCHK(c, hipSetDevice(c->device));
   return fetch(c, pairs, capacity, count);
}

int v3dm_synchronize(v3dm_ctx *c)
{
   if (!c) return V3DM_ERR_ARG;
   HIPCHK(c, hipSetDevice(c->device));
   int rc = syncStream(c);
   if (!rc) c->pendingAsync = false;
   return rc;
}

int v3dm_profile_enable(v3dm_ctx *c, int enable)
{
   if (!c) return V3DM_ERR_ARG;
   HIPCHK(c, hipSetDevice(c->device));
   int rc = syncStream(c);
   if (rc) return rc;
   c->profile = enable != 0;
   c->profJobs = 0;
   return V3DM_OK;
}

int v3dm_profile_read(v3dm_ctx *c, double *scanMs, double *selectMs, int *jobs)
{
   if (!c) return V3DM_ERR_ARG;
   HIPCHK(c, hipSetDevice(c->device));
   int rc = syncStream(c);
   if (rc) return rc;
   double a = 0.0, b = 0.0;
   for (int j = 0; j < c->profJobs; ++j)
   {
      float t01 = 0.f, t12 = 0.f;
      HIPCHK(c, hipEventElapsedTime(&t01, c->events[3 * j], c->events[3 * j + 1]));
      HIPCHK(c, hipEventElapsedTime(&t12, c->events[3 * j + 1], c->events[3 * j + 2]));
      a += t01;
      b += t12;
   }
   if (scanMs) *scanMs = a;
   if (selectMs) *selectMs = b;
   if (jobs) *jobs = c->profJobs;
   return V3DM_OK;
}

int v3dm_device_count(int *count)
{
   if (!count) return V3DM_ERR_ARG;
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
   *count = n;
   return V3DM_OK;
}

int v3dm_mfma_layout_probe(int device, int32_t *out)
{
   if (!out) return fail(nullptr, V3DM_ERR_ARG, "mfma_layout_probe: out is NULL");
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
      return fail(nullptr, V3DM_ERR_DEVICE, "mfma_layout_probe: device %d not available", device);
   hipError_t e = hipSetDevice(device);
   int32_t *d = nullptr;
   if (e == hipSuccess) e = hipMalloc((void **)&d, sizeof(int32_t) * 1024);
   if (e == hipSuccess) e = launch_mfma_probe(d, nullptr);
   if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(int32_t) * 1024, hipMemcpyDeviceToHost);
   if (d) (void)hipFree(d);
   if (e != hipSuccess) return fail(nullptr, V3DM_ERR_DEVICE, "mfma_layout_probe: %s", hipGetErrorString(e));
   return V3DM_OK;
}

int v3dm_mfma_ceiling_probe(int device, const uint8_t *rows, int fromLds, int iters, double *topsOut)
{
   if (!rows || !topsOut || iters < 2)
      return fail(nullptr, V3DM_ERR_ARG, "mfma_ceiling_probe: rows / topsOut NULL or iters < 2");
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
      return fail(nullptr, V3DM_ERR_DEVICE, "mfma_ceiling_probe: device %d not available", device);
   hipError_t e = hipSetDevice(device);
   int cus = 0;
   if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
   const int grid = 2 * std::max(cus, 1); // two 4-wave workgroups per CU: two waves per SIMD, like the scan
   int8_t *dRows = nullptr;
   int32_t *dOut = nullptr;
   hipEvent_t t0 = nullptr, t1 = nullptr;
   if (e == hipSuccess) e = hipMalloc((void **)&dRows, 128 * kDim);
   if (e == hipSuccess) e = hipMalloc((void **)&dOut, sizeof(int32_t) * 256 * (size_t)grid);
   if (e == hipSuccess) e = hipMemcpy(dRows, rows, 128 * kDim, hipMemcpyHostToDevice);
   if (e == hipSuccess) e = hipEventCreate(&t0);
   if (e == hipSuccess) e = hipEventCreate(&t1);
   if (e == hipSuccess) e = launch_mfma_ceiling(fromLds != 0, grid, dRows, iters, dOut, nullptr); // warm-up
   if (e == hipSuccess) e = hipEventRecord(t0, nullptr);
   if (e == hipSuccess) e = launch_mfma_ceiling(fromLds != 0, grid, dRows, iters, dOut, nullptr);
   if (e == hipSuccess) e = hipEventRecord(t1, nullptr);
   if (e == hipSuccess) e = hipEventSynchronize(t1);
   float ms = 0.0f;
   if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
   if (t0) (void)hipEventDestroy(t0);
   if (t1) (void)hipEventDestroy(t1);
   if (dRows) (void)hipFree(dRows);
   if (dOut) (void)hipFree(dOut);
   if (e != hipSuccess) return fail(nullptr, V3DM_ERR_DEVICE, "mfma_ceiling_probe: %s", hipGetErrorString(e));
   // 4 waves x (iters rounded down to even) steps x 8 MFMAs of 32 x 32 x 32 x 2 ops
   const double ops = (double)grid * 4 * (iters & ~1) * 8 * 65536.0;
   *topsOut = ms > 0.0f ? ops / ((double)ms * 1e-3) / 1e12 : 0.0;
   return V3DM_OK;
}

int v3dm_fetch_left_top2(v3dm_ctx *c, int n, float *best, float *second, int32_t *argBest)
{
   if (!c) return V3DM_ERR_ARG;
   if (!c->allocated) return fail(c, V3DM_ERR_STATE, "fetch_left_top2: allocate() has not been called");
   if (c->lastKernel != V3DM_KERNEL_I8 && c->lastKernel != V3DM_KERNEL_BF16)
      return fail(c, V3DM_ERR_STATE, "fetch_left_top2: the last job did not run a top-2 kernel");
   if (n < 0 || n > c->lastNL) return fail(c, V3DM_ERR_ARG, "fetch_left_top2: n = %d out of range", n);
   HIPCHK(c, hipSetDevice(c->device));
   int rc = syncStream(c);
   if (rc) return rc;
   c->pendingAsync = false;
   std::vector<uint64_t> t1(std::max(n, 1)), t2(std::max(n, 1));
   if (n > 0)
   {
      HIPCHK(c, hipMemcpy(t1.data(), c->L.dTop1, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
      HIPCHK(c, hipMemcpy(t2.data(), c->L.dTop2, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
   }
   const uint32_t ep = c->hostEpoch - 1; // epoch of the last finished job
   const bool fl = c->keyFormat == 1;
   for (int k = 0; k < n; ++k)
   {
      float b = -1e30f, s2 = -1e30f;
      int32_t arg = -1;
      if ((fl ? keyf_epoch(t1[k]) : key_epoch(t1[k])) == ep)
      {
         b = fl ? keyf_score(t1[k]) : (float)key_score(t1[k]) * (1.0f / 262144.0f);
         arg = (int32_t)key_index(t1[k]);
      }
      if ((fl ? keyf_epoch(t2[k]) : key_epoch(t2[k])) == ep)
         s2 = fl ? keyf_score(t2[k]) : (float)key_score(t2[k]) * (1.0f / 262144.0f);
      if (best) best[k] = b;
      if (second) second[k] = s2;
      if (argBest) argBest[k] = arg;
   }
   return V3DM_OK;
}

int v3dm_copy_probe(int device, int64_t bytes, int reps)
{
   if (bytes <= 0 || bytes % 16 || reps < 1) return fail(nullptr, V3DM_ERR_ARG, "copy_probe: bad size or count");
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
      return fail(nullptr, V3DM_ERR_DEVICE, "copy_probe: device %d not available", device);
   hipError_t e = hipSetDevice(device);
   void *src = nullptr, *dst = nullptr;
   if (e == hipSuccess) e = hipMalloc(&src, (size_t)bytes);
   if (e == hipSuccess) e = hipMalloc(&dst, (size_t)bytes);
   if (e == hipSuccess) e = hipMemset(src, 1, (size_t)bytes);
   for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy_probe(src, dst, bytes, nullptr);
   if (e == hipSuccess) e = hipDeviceSynchronize();
   if (src) (void)hipFree(src);
   if (dst) (void)hipFree(dst);
   if (e != hipSuccess) return fail(nullptr, V3DM_ERR_DEVICE, "copy_probe: %s", hipGetErrorString(e));
   return V3DM_OK;
}

} // extern "C"
