// v3dm_batch.cpp -- many image pairs in one job (the v3dm_batch_* C ABI).
//
// The reference matches one pair per findPutativeMatches call
// (SfM/match_sequence/src/match_sequence.cpp:157-203: setLeftFeatures once per
// view, setRightFeatures + findPutativeMatches per neighbour).  Here every
// image's descriptors are uploaded once into a pool, and any list of
// (left image, right image) pairs is matched by one scan launch (persistent
// workgroups pulling 2-D tiles of all pairs from one queue, largest pairs
// first), one insert launch and one finalize launch, optionally replayed from
// a captured hipGraph.  Each pair's result is the reference's per-pair list.
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>

#include "v3dm_ctx.h"

namespace v3dm_host {

struct BatchState
{
   // images
   int nImages = 0;
   std::vector<int> given;       // feature counts as given (rows of each image in the caller's arrays)
   std::vector<int> counts;      // feature counts after rounding to multiples of 32
   std::vector<int64_t> imgOff;  // first pool row of each image
   int64_t poolRows = 0;
   int8_t *dI8 = nullptr;
   int32_t *dSum = nullptr;
   int64_t poolCap = 0; // bytes of dI8
   int64_t sumCap = 0;  // rows of dSum
   uint8_t *dRaw = nullptr;
   int64_t rawCap = 0;
   // float images (v3dm_batch_set_images_f32): FP32 and BF16 pools at the same row offsets
   bool isFloat = false;
   float *dF32 = nullptr;
   int64_t f32Cap = 0;
   uint16_t *dBF = nullptr;
   int64_t bfCap = 0;
   float *dRawF = nullptr;
   int64_t rawFCap = 0;
   uint32_t *dNorm = nullptr;
   float norm2Max = 0.0f;

   // pairs (user order) and the tile table (largest pairs first)
   int nPairs = 0;
   std::vector<int32_t> pairImg;
   std::vector<int64_t> outOff; // nPairs + 1 offsets of the per-left result table
   std::vector<int64_t> rOff;   // nPairs + 1 offsets of the right-side slots
   std::vector<int> order;      // table position -> user pair
   std::vector<PairDesc> table;
   PairDesc *dPairs = nullptr;
   int pairsCap = 0;
   std::vector<int32_t> tilePair, finPair; // table index of every tile / finalize block
   int32_t *dTilePair = nullptr, *dFinPair = nullptr;
   int64_t tilePairCap = 0, finPairCap = 0;
   int variant = -1;            // tile variant of the current table (-1: none yet)
   int wantVariant = -1;        // -1: autotune / tuning cache
   int nTiles = 0, nFin = 0;

   uint64_t *dKeys = nullptr;   // top1L | top2L | top1R | top2R
   int64_t keyCapL = 0, keyCapR = 0;
   int32_t *dMatch = nullptr;
   int32_t *hMatch = nullptr;   // pinned
   int64_t matchCap = 0;
   uint64_t *dSurv = nullptr;
   int32_t *dSurvCount = nullptr;
   int64_t regionCap = 0;
   int64_t survEntries = 0;     // capacity of dSurv in entries
   uint32_t *dCounters = nullptr; // epoch, done, tile counter
   uint32_t hostEpoch = 1;
   bool keysFresh = false;

   // guided matching per pair: positions at the pool's row offsets (2 floats per row)
   // and one mask per user pair (PairDesc::guide = pair + 1 when the pair has one)
   float *dPos = nullptr;
   int64_t posCap = 0;
   bool hasPos = false;
   std::vector<int32_t> guideKind; // per user pair; empty: no guided masks
   Mask *dGuides = nullptr;
   int64_t guidesCap = 0;

   bool useGraph = false;
   hipGraphExec_t exec = nullptr;
   bool graphValid = false;
   float gScore = 0.f, gRatio = 0.f;
   bool pending = false;
};

void freeBatch(v3dm_ctx *c)
{
   BatchState *b = c->batch;
   if (!b) return;
   if (b->exec) (void)hipGraphExecDestroy(b->exec);
   dfree(b->dI8);
   dfree(b->dSum);
   dfree(b->dRaw);
   dfree(b->dF32);
   dfree(b->dBF);
   dfree(b->dRawF);
   dfree(b->dNorm);
   dfree(b->dPairs);
   dfree(b->dTilePair);
   dfree(b->dFinPair);
   dfree(b->dKeys);
   dfree(b->dMatch);
   dfree(b->dSurv);
   dfree(b->dSurvCount);
   dfree(b->dCounters);
   dfree(b->dPos);
   dfree(b->dGuides);
   if (b->hMatch) (void)hipHostFree(b->hMatch);
   delete b;
   c->batch = nullptr;
}

namespace {

BatchState *batchOf(v3dm_ctx *c)
{
   if (!c->batch) c->batch = new BatchState();
   return c->batch;
}

template <class T> int growDevice(v3dm_ctx *c, T *&p, int64_t &cap, int64_t need)
{
   if (need <= cap && p) return V3DM_OK;
   dfree(p);
   cap = 0;
   const int64_t n = std::max<int64_t>(need, 1);
   HIPCHK(c, hipMalloc((void **)&p, sizeof(T) * (size_t)n));
   cap = n;
   return V3DM_OK;
}

// Builds the tile table for variant v (pairs in decreasing nL * nR).
void buildTable(BatchState *b, int v)
{
   const int tm = scan_variant_tile_m(v), tn = scan_variant_tile_n(v);
   b->table.assign(b->nPairs, PairDesc{});
   b->tilePair.clear();
   b->finPair.clear();
   int tiles = 0, fin = 0;
   for (int i = 0; i < b->nPairs; ++i)
   {
      const int p = b->order[i];
      const int li = b->pairImg[2 * p], ri = b->pairImg[2 * p + 1];
      PairDesc &d = b->table[i];
      d.offL = b->imgOff[li];
      d.offR = b->imgOff[ri];
      d.slotL = b->outOff[p];
      d.slotR = b->rOff[p];
      d.nL = b->counts[li];
      d.nR = b->counts[ri];
      d.tilesM = (roundUp(d.nL, kRowPad) + tm - 1) / tm;
      d.tilesN = (roundUp(d.nR, kRowPad) + tn - 1) / tn;
      if (d.nL == 0 || d.nR == 0) d.tilesM = d.tilesN = 0;
      d.tileBase = tiles;
      d.finBase = fin;
      d.guide = (!b->guideKind.empty() && b->guideKind[p] != 0) ? p + 1 : 0;
      tiles += d.tilesM * d.tilesN;
      fin += std::max(1, (d.nL + 255) / 256);
      b->tilePair.insert(b->tilePair.end(), (size_t)d.tilesM * d.tilesN, i);
      b->finPair.insert(b->finPair.end(), (size_t)std::max(1, (d.nL + 255) / 256), i);
   }
   b->nTiles = tiles;
   b->nFin = fin;
   b->variant = v;
}

int uploadTable(v3dm_ctx *c, BatchState *b)
{
   if (b->nPairs > b->pairsCap)
   {
      dfree(b->dPairs);
      b->pairsCap = 0;
      HIPCHK(c, hipMalloc((void **)&b->dPairs, sizeof(PairDesc) * b->nPairs));
      b->pairsCap = b->nPairs;
   }
   if (b->nPairs > 0)
      HIPCHK(c, hipMemcpyAsync(b->dPairs, b->table.data(), sizeof(PairDesc) * b->nPairs, hipMemcpyHostToDevice,
                               c->stream));
   int rc0 = growDevice(c, b->dTilePair, b->tilePairCap, (int64_t)b->tilePair.size());
   if (!rc0) rc0 = growDevice(c, b->dFinPair, b->finPairCap, (int64_t)b->finPair.size());
   if (rc0) return rc0;
   if (!b->tilePair.empty())
      HIPCHK(c, hipMemcpyAsync(b->dTilePair, b->tilePair.data(), sizeof(int32_t) * b->tilePair.size(),
                               hipMemcpyHostToDevice, c->stream));
   if (!b->finPair.empty())
      HIPCHK(c, hipMemcpyAsync(b->dFinPair, b->finPair.data(), sizeof(int32_t) * b->finPair.size(),
                               hipMemcpyHostToDevice, c->stream));
   const int64_t regions = (int64_t)b->nTiles * scan_variant_waves(b->variant);
   int rc = growDevice(c, b->dSurvCount, b->regionCap, regions);
   if (rc) return rc;
   // grown only when a table needs more: fresh allocations make the next job slow (first touch)
   rc = growDevice(c, b->dSurv, b->survEntries,
                   std::max<int64_t>(regions, 1) * (b->isFloat ? c->survCap : scan_variant_surv_cap(b->variant)));
   if (rc) return rc;
   b->graphValid = false;
   return syncStream(c);
}

int clearBatchKeys(v3dm_ctx *c, BatchState *b)
{
   const int64_t n = 2 * (b->keyCapL + b->keyCapR);
   HIPCHK(c, launch_clear_keys(b->dKeys, n, b->dCounters, b->dCounters + 1, c->stream));
   b->hostEpoch = 1;
   b->graphValid = false;
   return V3DM_OK;
}

void fillArgs(v3dm_ctx *c, BatchState *b, float minScore, float minRatio, ScanArgs &a, FinalizeArgs &f, int &grid)
{
   const int v = b->variant;
   uint64_t *k = b->dKeys;
   a = ScanArgs{};
   a.pair0 = b->nPairs == 1 ? b->table[0] : PairDesc{};
   a.L = b->dI8;
   a.Lsum = b->dSum;
   a.R = b->dI8;
   a.Rsum = b->dSum;
   a.pairs = b->dPairs;
   a.nPairs = b->nPairs;
   a.nTiles = b->nTiles;
   a.thr = survivorThreshold(minScore, minRatio);
   a.top1L = k;
   a.top2L = k + b->keyCapL;
   a.top1R = k + 2 * b->keyCapL;
   a.top2R = k + 2 * b->keyCapL + b->keyCapR;
   a.epoch = b->dCounters;
   a.masks = b->guideKind.empty() ? nullptr : b->dGuides; // indexed by PairDesc::guide
   a.nMasks = 0;
   a.surv = b->dSurv;
   a.survCount = b->dSurvCount;
   a.survCap = scan_variant_surv_cap(v);
   a.waves = scan_variant_waves(v);
   a.tileM = scan_variant_tile_m(v);
   a.tileN = scan_variant_tile_n(v);
   a.tileCounter = b->dCounters + 2;
   a.tilePair = b->dTilePair;
   const int resident = scan_variant_resident(v, c->device);
   a.persistent = b->nTiles > resident ? 1 : 0;
   a.inlineInsert = a.persistent ? 0 : 1; // the whole grid resident: survivors applied in the scan
   grid = a.persistent ? resident : b->nTiles;
   f = FinalizeArgs{};
   f.pairs = b->dPairs;
   f.nPairs = b->nPairs;
   f.nBlocks = b->nFin;
   f.top1L = a.top1L;
   f.top2L = a.top2L;
   f.top1R = a.top1R;
   f.top2R = a.top2R;
   f.epoch = b->dCounters;
   f.done = nullptr; // thousands of blocks: k_advance_epoch closes the job instead
   f.tileCounter = b->dCounters + 2;
   f.minScore = minScore;
   f.minRatio = minRatio;
   f.matchOfLeft = b->dMatch;
   f.finPair = b->dFinPair;
}

// Float images: the BF16 candidate scan, FP32 rescoring, float-key finalize.
void fillArgsF(v3dm_ctx *c, BatchState *b, float minScore, float minRatio, ScanArgsF &a, int &grid)
{
   const int v = b->variant;
   uint64_t *k = b->dKeys;
   a = ScanArgsF{};
   a.pair0 = b->nPairs == 1 ? b->table[0] : PairDesc{};
   a.L = b->dBF;
   a.R = b->dBF;
   a.Lf = b->dF32;
   a.Rf = b->dF32;
   a.pairs = b->dPairs;
   a.nPairs = b->nPairs;
   a.nTiles = b->nTiles;
   a.thr = survivorThresholdF(minScore, minRatio);
   a.lim = bf16CandidateLimit(a.thr, b->norm2Max, b->norm2Max);
   a.top1L = k;
   a.top2L = k + b->keyCapL;
   a.top1R = k + 2 * b->keyCapL;
   a.top2R = k + 2 * b->keyCapL + b->keyCapR;
   a.epoch = b->dCounters;
   a.masks = b->guideKind.empty() ? nullptr : b->dGuides; // indexed by PairDesc::guide
   a.nMasks = 0;
   a.surv = b->dSurv;
   a.survCount = b->dSurvCount;
   a.survCap = c->survCap;
   a.waves = scan_variant_waves(v);
   a.tileM = scan_variant_tile_m(v);
   a.tileN = scan_variant_tile_n(v);
   a.tileCounter = b->dCounters + 2;
   a.tilePair = b->dTilePair;
   const int resident = bf16_scan_resident(v, c->device);
   a.persistent = b->nTiles > resident ? 1 : 0;
   grid = a.persistent ? resident : b->nTiles;
}

int enqueueKernels(v3dm_ctx *c, BatchState *b, float minScore, float minRatio, bool marks)
{
   if (b->isFloat)
   {
      ScanArgs unused{};
      FinalizeArgs f{};
      int grid = 0;
      fillArgs(c, b, minScore, minRatio, unused, f, grid);
      ScanArgsF a{};
      fillArgsF(c, b, minScore, minRatio, a, grid);
      // batched jobs run for milliseconds: plain event records around the scan
      int rc;
      if (b->variant >= bf16_scan_variant_count())
         return fail(c, V3DM_ERR_UNSUPPORTED, "batch: scan variant %d has no float-descriptor form (0 .. %d)",
                     b->variant, bf16_scan_variant_count() - 1);
      if (marks && (rc = profMark(c, 0))) return rc;
      HIPCHK(c, launch_bf16_scan(b->variant, grid, a, c->stream));
      if (marks && (rc = profMark(c, 1))) return rc;
      HIPCHK(c, launch_f32_rescore(a, c->stream));
      HIPCHK(c, launch_f32_finalize(f, c->stream));
      HIPCHK(c, launch_advance_epoch(b->dCounters, b->dCounters + 2, c->stream));
      if (marks && (rc = profMark(c, 2))) return rc;
      return V3DM_OK;
   }
   ScanArgs a{};
   FinalizeArgs f{};
   int grid = 0;
   fillArgs(c, b, minScore, minRatio, a, f, grid);
   int rc;
   if (marks && (rc = profMark(c, 0))) return rc;
   HIPCHK(c, launch_i8_scan(b->variant, grid, a, c->stream));
   c->lastInsertForm = -1;
   if (!a.inlineInsert)
   {
      c->lastInsertForm = insert_form_for((int64_t)a.nTiles * a.waves, c->device, c->insertForm);
      HIPCHK(c, launch_i8_insert(a, c->lastInsertForm, c->stream));
   }
   if (marks && (rc = profMark(c, 1))) return rc;
   HIPCHK(c, launch_i8_finalize(f, c->stream));
   HIPCHK(c, launch_advance_epoch(b->dCounters, b->dCounters + 2, c->stream));
   if (marks && (rc = profMark(c, 2))) return rc;
   return V3DM_OK;
}

int ensureVariant(v3dm_ctx *c, BatchState *b, int v)
{
   if (b->variant == v) return V3DM_OK;
   buildTable(b, v);
   return uploadTable(c, b);
}

std::string tuneKey(const BatchState *b)
{
   // size bucket: pair count and mean feature counts, log2-rounded
   double sl = 0, sr = 0;
   for (int p = 0; p < b->nPairs; ++p)
   {
      sl += b->counts[b->pairImg[2 * p]];
      sr += b->counts[b->pairImg[2 * p + 1]];
   }
   const int np = std::max(b->nPairs, 1);
   auto lg = [](double x) { return (int)std::lround(std::log2(std::max(x, 1.0))); };
   std::ostringstream os;
   os << (b->isFloat ? "bf16" : "i8") << " pairs=" << lg(np) << " nL=" << lg(sl / np) << " nR=" << lg(sr / np);
   return os.str();
}

std::map<std::string, int> readTuneCache(const char *path)
{
   std::map<std::string, int> m;
   std::ifstream is(path);
   std::string line;
   while (std::getline(is, line))
   {
      const size_t bar = line.rfind('|');
      if (bar == std::string::npos) continue;
      m[line.substr(0, bar)] = std::atoi(line.c_str() + bar + 1);
   }
   return m;
}

void writeTuneCache(const char *path, const std::map<std::string, int> &m)
{
   std::ofstream os(path, std::ios::trunc);
   for (auto &kv : m) os << kv.first << '|' << kv.second << '\n';
}

int enqueueJob(v3dm_ctx *c, BatchState *b, float minScore, float minRatio);

// Times every scan variant on the current batch; returns the fastest.
int autotune(v3dm_ctx *c, BatchState *b, float minScore, float minRatio, int reps, double *msOut, int *best)
{
   const int nv = b->isFloat ? bf16_scan_variant_count() : scan_variant_count();
   hipEvent_t e0, e1;
   HIPCHK(c, hipEventCreate(&e0));
   HIPCHK(c, hipEventCreate(&e1));
   double bestMs = 1e30;
   int bestV = 0;
   const bool graph = b->useGraph;
   b->useGraph = false;
   for (int v = 0; v < nv; ++v)
   {
      int rc = ensureVariant(c, b, v);
      if (!rc) rc = enqueueJob(c, b, minScore, minRatio); // warm-up
      if (rc)
      {
         b->useGraph = graph;
         return rc;
      }
      HIPCHK(c, hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r)
      {
         rc = enqueueJob(c, b, minScore, minRatio);
         if (rc)
         {
            b->useGraph = graph;
            return rc;
         }
      }
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
      const double per = ms / std::max(reps, 1);
      if (msOut) msOut[v] = per;
      if (per < bestMs)
      {
         bestMs = per;
         bestV = v;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   b->useGraph = graph;
   b->pending = false;
   if (best) *best = bestV;
   return ensureVariant(c, b, bestV);
}

int pickVariant(v3dm_ctx *c, BatchState *b, float minScore, float minRatio)
{
   if (b->wantVariant >= 0) return ensureVariant(c, b, b->wantVariant);
   if (b->variant >= 0) return V3DM_OK; // already tuned for this pair set
   const char *path = std::getenv("V3DM_TUNE_CACHE");
   const std::string key = tuneKey(b);
   std::map<std::string, int> cache;
   if (path)
   {
      cache = readTuneCache(path);
      auto it = cache.find(key);
      if (it != cache.end() && it->second >= 0 &&
          it->second < (b->isFloat ? bf16_scan_variant_count() : scan_variant_count()))
         return ensureVariant(c, b, it->second);
   }
   int best = 0;
   int rc = autotune(c, b, minScore, minRatio, 3, nullptr, &best);
   if (rc) return rc;
   if (path)
   {
      cache[key] = best;
      writeTuneCache(path, cache);
   }
   return V3DM_OK;
}

int enqueueJob(v3dm_ctx *c, BatchState *b, float minScore, float minRatio)
{
   if (b->hostEpoch >= (b->isFloat ? kEpochLimitF : kEpochLimit))
   {
      int rc = clearBatchKeys(c, b);
      if (rc) return rc;
   }
   int rc = V3DM_OK;
   if (b->useGraph)
   {
      // a replayed graph is one unit for the kernel timers: scan = whole job
      if ((rc = profMark(c, 0))) return rc;
      if (!b->graphValid || b->gScore != minScore || b->gRatio != minRatio)
      {
         if (b->exec)
         {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
         }
         hipGraph_t g = nullptr;
         HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
         rc = enqueueKernels(c, b, minScore, minRatio, false);
         hipError_t e = hipStreamEndCapture(c->stream, &g);
         if (rc) return rc;
         if (e != hipSuccess) return fail(c, V3DM_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
         e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
         (void)hipGraphDestroy(g);
         if (e != hipSuccess) return fail(c, V3DM_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
         b->graphValid = true;
         b->gScore = minScore;
         b->gRatio = minRatio;
      }
      HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
      if ((rc = profMark(c, 1)) || (rc = profMark(c, 2))) return rc;
   }
   else
   {
      rc = enqueueKernels(c, b, minScore, minRatio, true);
      if (rc) return rc;
   }
   b->hostEpoch += 1;
   b->pending = true;
   return V3DM_OK;
}

} // namespace
} // namespace v3dm_host

using namespace v3dm_host;

extern "C" {

int v3dm_batch_set_images_u8(v3dm_ctx *c, int nImages, const int32_t *counts, const uint8_t *desc)
{
   if (!c) return V3DM_ERR_ARG;
   if (nImages < 0 || (nImages > 0 && !counts)) return fail(c, V3DM_ERR_ARG, "batch_set_images: bad arguments");
   HIPCHK(c, hipSetDevice(c->device));
   BatchState *b = batchOf(c);
   if (b->pending)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      b->pending = false;
   }
   int64_t total = 0, rows = 0;
   std::vector<int> cnt(nImages);
   std::vector<int64_t> off(nImages);
   for (int i = 0; i < nImages; ++i)
   {
      if (counts[i] < 0 || counts[i] > kMaxLocalIndex)
         return fail(c, V3DM_ERR_ARG, "batch_set_images: image %d has %d features", i, counts[i]);
      int n = counts[i];
      if (n % 32)
      {
         if (c->warnRounding)
            fprintf(stderr, "SIFT_Matching batch warning: image %d has %d features, rounding down to a multiple "
                            "of 32\n", i, n);
         n -= n & 31;
      }
      cnt[i] = n;
      off[i] = rows;
      rows += std::max(roundUp(std::max(n, 1), kRowPad), kRowPad);
      total += counts[i];
   }
   if (total > 0 && !desc) return fail(c, V3DM_ERR_ARG, "batch_set_images: descriptor pointer is NULL");
   int rc = growDevice(c, b->dI8, b->poolCap, rows * kDim);
   if (rc) return rc;
   rc = growDevice(c, b->dSum, b->sumCap, rows);
   if (rc) return rc;
   rc = growDevice(c, b->dRaw, b->rawCap, total * kDim);
   if (rc) return rc;
   if (total > 0) HIPCHK(c, hipMemcpyAsync(b->dRaw, desc, (size_t)total * kDim, hipMemcpyHostToDevice, c->stream));
   int64_t src = 0;
   for (int i = 0; i < nImages; ++i)
   {
      const int nPad = std::max(roundUp(std::max(cnt[i], 1), kRowPad), kRowPad);
      HIPCHK(c, launch_prep_u8(b->dRaw + src * kDim, cnt[i], nPad, b->dI8 + off[i] * kDim, b->dSum + off[i],
                               c->stream));
      src += counts[i];
   }
   rc = syncStream(c);
   if (rc) return rc;
   b->isFloat = false;
   b->nImages = nImages;
   b->given.assign(counts, counts + nImages);
   b->hasPos = false;
   b->guideKind.clear();
   b->counts = cnt;
   b->imgOff = off;
   b->poolRows = rows;
   b->nPairs = 0; // pairs refer to images: set them again
   b->variant = -1;
   b->graphValid = false;
   return V3DM_OK;
}

int v3dm_batch_set_images_f32(v3dm_ctx *c, int nImages, const int32_t *counts, const float *desc)
{
   if (!c) return V3DM_ERR_ARG;
   if (nImages < 0 || (nImages > 0 && !counts)) return fail(c, V3DM_ERR_ARG, "batch_set_images: bad arguments");
   HIPCHK(c, hipSetDevice(c->device));
   BatchState *b = batchOf(c);
   if (b->pending)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      b->pending = false;
   }
   int64_t total = 0, rows = 0;
   std::vector<int> cnt(nImages);
   std::vector<int64_t> off(nImages);
   for (int i = 0; i < nImages; ++i)
   {
      if (counts[i] < 0 || counts[i] > kMaxLocalIndex)
         return fail(c, V3DM_ERR_ARG, "batch_set_images: image %d has %d features", i, counts[i]);
      int n = counts[i];
      if (n % 32)
      {
         if (c->warnRounding)
            fprintf(stderr, "SIFT_Matching batch warning: image %d has %d features, rounding down to a multiple "
                            "of 32\n", i, n);
         n -= n & 31;
      }
      cnt[i] = n;
      off[i] = rows;
      rows += std::max(roundUp(std::max(n, 1), kRowPad), kRowPad);
      total += counts[i];
   }
   if (total > 0 && !desc) return fail(c, V3DM_ERR_ARG, "batch_set_images: descriptor pointer is NULL");
   int rc = growDevice(c, b->dF32, b->f32Cap, rows * kDim);
   if (!rc) rc = growDevice(c, b->dBF, b->bfCap, rows * kDim);
   if (!rc) rc = growDevice(c, b->dRawF, b->rawFCap, total * kDim);
   if (rc) return rc;
   if (!b->dNorm) HIPCHK(c, hipMalloc((void **)&b->dNorm, sizeof(uint32_t)));
   if (total > 0)
      HIPCHK(c, hipMemcpyAsync(b->dRawF, desc, sizeof(float) * (size_t)total * kDim, hipMemcpyHostToDevice,
                               c->stream));
   int64_t src = 0;
   for (int i = 0; i < nImages; ++i)
   {
      const int nPad = std::max(roundUp(std::max(cnt[i], 1), kRowPad), kRowPad);
      if (cnt[i] > 0)
         HIPCHK(c, hipMemcpyAsync(b->dF32 + off[i] * kDim, b->dRawF + src * kDim, sizeof(float) * (size_t)cnt[i] * kDim,
                                  hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(c, launch_prep_bf16(b->dF32 + off[i] * kDim, cnt[i], nPad, b->dBF + off[i] * kDim, b->dNorm, c->stream,
                                 i == 0));
      src += counts[i];
   }
   uint32_t bits = 0;
   if (nImages > 0)
      HIPCHK(c, hipMemcpyAsync(c->hFlag ? (void *)c->hFlag : (void *)&bits, b->dNorm, sizeof(uint32_t),
                               hipMemcpyDeviceToHost, c->stream));
   rc = syncStream(c);
   if (rc) return rc;
   if (nImages > 0 && c->hFlag) memcpy(&bits, c->hFlag, sizeof(bits));
   memcpy(&b->norm2Max, &bits, sizeof(bits));
   b->isFloat = true;
   b->nImages = nImages;
   b->given.assign(counts, counts + nImages);
   b->hasPos = false;
   b->guideKind.clear();
   b->counts = cnt;
   b->imgOff = off;
   b->poolRows = rows;
   b->nPairs = 0; // pairs refer to images: set them again
   b->variant = -1;
   b->graphValid = false;
   return V3DM_OK;
}

int v3dm_batch_set_pairs(v3dm_ctx *c, int nPairs, const int32_t *pairImages)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = c->batch;
   if (!b || b->nImages == 0) return fail(c, V3DM_ERR_STATE, "batch_set_pairs: no images uploaded");
   if (nPairs < 0 || (nPairs > 0 && !pairImages)) return fail(c, V3DM_ERR_ARG, "batch_set_pairs: bad arguments");
   HIPCHK(c, hipSetDevice(c->device));
   if (b->pending)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      b->pending = false;
   }
   for (int p = 0; p < 2 * nPairs; ++p)
      if (pairImages[p] < 0 || pairImages[p] >= b->nImages)
         return fail(c, V3DM_ERR_ARG, "batch_set_pairs: image index %d out of range", pairImages[p]);
   b->nPairs = nPairs;
   b->pairImg.assign(pairImages, pairImages + 2 * nPairs);
   b->guideKind.clear();
   b->outOff.assign(nPairs + 1, 0);
   b->rOff.assign(nPairs + 1, 0);
   for (int p = 0; p < nPairs; ++p)
   {
      b->outOff[p + 1] = b->outOff[p] + b->counts[pairImages[2 * p]];
      b->rOff[p + 1] = b->rOff[p] + b->counts[pairImages[2 * p + 1]];
   }
   b->order.resize(nPairs);
   std::iota(b->order.begin(), b->order.end(), 0);
   std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
      const int64_t cx = (int64_t)b->counts[pairImages[2 * x]] * b->counts[pairImages[2 * x + 1]];
      const int64_t cy = (int64_t)b->counts[pairImages[2 * y]] * b->counts[pairImages[2 * y + 1]];
      return cx > cy;
   });
   const int64_t needL = std::max<int64_t>(b->outOff[nPairs], 1), needR = std::max<int64_t>(b->rOff[nPairs], 1);
   if (needL > b->keyCapL || needR > b->keyCapR || !b->dKeys)
   {
      dfree(b->dKeys);
      b->keyCapL = needL;
      b->keyCapR = needR;
      HIPCHK(c, hipMalloc((void **)&b->dKeys, sizeof(uint64_t) * 2 * (size_t)(needL + needR)));
      b->keysFresh = true;
   }
   if (!b->dCounters) HIPCHK(c, hipMalloc((void **)&b->dCounters, 4 * sizeof(uint32_t)));
   int rc = growDevice(c, b->dMatch, b->matchCap, needL);
   if (rc) return rc;
   if (b->hMatch) (void)hipHostFree(b->hMatch);
   b->hMatch = nullptr;
   HIPCHK(c, hipHostMalloc((void **)&b->hMatch, sizeof(int32_t) * (size_t)needL, hipHostMallocDefault));
   rc = clearBatchKeys(c, b);
   if (rc) return rc;
   b->variant = -1;
   if (b->wantVariant >= 0)
   {
      buildTable(b, b->wantVariant);
      return uploadTable(c, b);
   }
   return syncStream(c);
}

int v3dm_batch_set_positions(v3dm_ctx *c, const float *positions)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = c->batch;
   if (!b || b->nImages == 0) return fail(c, V3DM_ERR_STATE, "batch_set_positions: no images uploaded");
   int64_t total = 0;
   for (int g : b->given) total += g;
   if (total > 0 && !positions) return fail(c, V3DM_ERR_ARG, "batch_set_positions: position pointer is NULL");
   HIPCHK(c, hipSetDevice(c->device));
   if (b->pending)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      b->pending = false;
   }
   // the pool's layout: image i's first counts[i] positions at rows imgOff[i] (padding rows unused)
   std::vector<float> staged(2 * (size_t)std::max<int64_t>(b->poolRows, 1), 0.0f);
   int64_t src = 0;
   for (int i = 0; i < b->nImages; ++i)
   {
      std::copy(positions + 2 * src, positions + 2 * (src + b->counts[i]), staged.begin() + 2 * b->imgOff[i]);
      src += b->given[i];
   }
   int rc = growDevice(c, b->dPos, b->posCap, 2 * std::max<int64_t>(b->poolRows, 1));
   if (rc) return rc;
   HIPCHK(c, hipMemcpyAsync(b->dPos, staged.data(), sizeof(float) * staged.size(), hipMemcpyHostToDevice, c->stream));
   rc = syncStream(c);
   if (rc) return rc;
   b->hasPos = true;
   return V3DM_OK;
}

int v3dm_batch_set_guides(v3dm_ctx *c, const int32_t *kind, const double *M, const float *distanceThreshold)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = c->batch;
   if (!b || b->nImages == 0) return fail(c, V3DM_ERR_STATE, "batch_set_guides: no images uploaded");
   HIPCHK(c, hipSetDevice(c->device));
   if (b->pending)
   {
      int rc = syncStream(c);
      if (rc) return rc;
      b->pending = false;
   }
   std::vector<int32_t> kinds;
   std::vector<Mask> masks;
   if (kind && b->nPairs > 0)
   {
      kinds.assign(kind, kind + b->nPairs);
      bool any = false;
      for (int p = 0; p < b->nPairs; ++p)
      {
         if (kinds[p] < 0 || kinds[p] > 2)
            return fail(c, V3DM_ERR_ARG, "batch_set_guides: pair %d has mask kind %d (0, 1 or 2)", p, kinds[p]);
         any = any || kinds[p] != 0;
      }
      if (!any) kinds.clear();
   }
   if (!kinds.empty())
   {
      if (!M || !distanceThreshold) return fail(c, V3DM_ERR_ARG, "batch_set_guides: matrix or threshold pointer is NULL");
      if (!b->hasPos) return fail(c, V3DM_ERR_STATE, "batch_set_guides: no feature positions (v3dm_batch_set_positions)");
      masks.assign(b->nPairs, Mask{});
      for (int p = 0; p < b->nPairs; ++p)
      {
         if (!kinds[p]) continue;
         Mask &mk = masks[p];
         mk.kind = kinds[p];
         const float t = distanceThreshold[p];
         mk.sqrThr = t * t;                                           // v3d_cudamatching.cu:414 (float)
         for (int i = 0; i < 9; ++i) mk.m[i] = (float)M[9 * (size_t)p + i]; // v3d_cudamatching.cpp:161-164
         mk.lpos = b->dPos + 2 * b->imgOff[b->pairImg[2 * p]];
         mk.rpos = b->dPos + 2 * b->imgOff[b->pairImg[2 * p + 1]];
      }
      int rc = growDevice(c, b->dGuides, b->guidesCap, (int64_t)b->nPairs);
      if (rc) return rc;
      HIPCHK(c, hipMemcpyAsync(b->dGuides, masks.data(), sizeof(Mask) * masks.size(), hipMemcpyHostToDevice,
                               c->stream));
   }
   b->guideKind.swap(kinds);
   if (b->variant >= 0)
   {
      // the tile table is built: refresh its guide indices
      for (int i = 0; i < b->nPairs; ++i)
      {
         const int p = b->order[i];
         b->table[i].guide = (!b->guideKind.empty() && b->guideKind[p] != 0) ? p + 1 : 0;
      }
      if (b->nPairs > 0)
         HIPCHK(c, hipMemcpyAsync(b->dPairs, b->table.data(), sizeof(PairDesc) * b->nPairs, hipMemcpyHostToDevice,
                                  c->stream));
   }
   b->graphValid = false;
   return syncStream(c);
}

int v3dm_batch_set_variant(v3dm_ctx *c, int variant)
{
   if (!c) return V3DM_ERR_ARG;
   if (variant < -1 || variant >= scan_variant_count())
      return fail(c, V3DM_ERR_ARG, "batch_set_variant: variant %d out of range", variant);
   BatchState *b = batchOf(c);
   b->wantVariant = variant;
   if (variant >= 0 && b->nPairs > 0 && b->variant != variant)
   {
      HIPCHK(c, hipSetDevice(c->device));
      int rc = syncStream(c);
      if (rc) return rc;
      b->pending = false;
      return ensureVariant(c, b, variant);
   }
   return V3DM_OK;
}

int v3dm_batch_get_variant(const v3dm_ctx *c, int *variant)
{
   if (!c || !variant) return V3DM_ERR_ARG;
   *variant = c->batch ? c->batch->variant : -1;
   return V3DM_OK;
}

int v3dm_scan_variant_info(int variant, int *waves, int *tileRows, int *tileCols)
{
   if (variant < 0 || variant >= scan_variant_count()) return V3DM_ERR_ARG;
   if (waves) *waves = scan_variant_waves(variant);
   if (tileRows) *tileRows = scan_variant_tile_m(variant);
   if (tileCols) *tileCols = scan_variant_tile_n(variant);
   return V3DM_OK;
}

int v3dm_scan_variant_form(int variant, int *wholeTile)
{
   if (variant < 0 || variant >= scan_variant_count() || !wholeTile) return V3DM_ERR_ARG;
   *wholeTile = scan_variant_whole(variant);
   return V3DM_OK;
}

int v3dm_scan_variant_count(void) { return scan_variant_count(); }
int v3dm_scan_variant_count_f32(void) { return bf16_scan_variant_count(); }

int v3dm_batch_autotune(v3dm_ctx *c, float minScore, float minRatio, int reps, double *msPerVariant, int *best)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = c->batch;
   if (!b || b->nPairs == 0) return fail(c, V3DM_ERR_STATE, "batch_autotune: no pairs set");
   HIPCHK(c, hipSetDevice(c->device));
   int rc = syncStream(c);
   if (rc) return rc;
   int bv = 0;
   rc = autotune(c, b, minScore, minRatio, std::max(reps, 1), msPerVariant, &bv);
   if (rc) return rc;
   b->wantVariant = bv;
   if (best) *best = bv;
   return V3DM_OK;
}

int v3dm_batch_set_graph(v3dm_ctx *c, int enable)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = batchOf(c);
   b->useGraph = enable != 0;
   b->graphValid = false;
   return V3DM_OK;
}

int v3dm_batch_match_async(v3dm_ctx *c, float minScore, float minRatio)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = c->batch;
   if (!b || b->nImages == 0) return fail(c, V3DM_ERR_STATE, "batch_match: no images uploaded");
   HIPCHK(c, hipSetDevice(c->device));
   if (b->nPairs == 0)
   {
      b->pending = false;
      return V3DM_OK;
   }
   int rc = pickVariant(c, b, minScore, minRatio);
   if (rc) return rc;
   return enqueueJob(c, b, minScore, minRatio);
}

int v3dm_batch_fetch(v3dm_ctx *c, int32_t *matchOfLeft, int64_t *offsets)
{
   if (!c) return V3DM_ERR_ARG;
   BatchState *b = c->batch;
   if (!b) return fail(c, V3DM_ERR_STATE, "batch_fetch: no batch");
   HIPCHK(c, hipSetDevice(c->device));
   const int64_t n = b->nPairs > 0 ? b->outOff[b->nPairs] : 0;
   if (n > 0 && b->pending)
      HIPCHK(c, hipMemcpyAsync(b->hMatch, b->dMatch, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                               c->stream));
   int rc = syncStream(c);
   if (rc) return rc;
   if (b->pending && matchOfLeft && n > 0) memcpy(matchOfLeft, b->hMatch, sizeof(int32_t) * (size_t)n);
   b->pending = false;
   if (offsets)
      for (int p = 0; p <= b->nPairs; ++p) offsets[p] = b->outOff.empty() ? 0 : b->outOff[p];
   return V3DM_OK;
}

int v3dm_batch_match(v3dm_ctx *c, float minScore, float minRatio, int32_t *matchOfLeft, int64_t *offsets)
{
   int rc = v3dm_batch_match_async(c, minScore, minRatio);
   if (rc) return rc;
   return v3dm_batch_fetch(c, matchOfLeft, offsets);
}

int v3dm_batch_info(const v3dm_ctx *c, int *nPairs, int64_t *nLeftTotal, int *nTiles, int64_t *pairsScored)
{
   if (!c) return V3DM_ERR_ARG;
   const BatchState *b = c->batch;
   int64_t scored = 0;
   if (b)
      for (int p = 0; p < b->nPairs; ++p)
         scored += (int64_t)b->counts[b->pairImg[2 * p]] * b->counts[b->pairImg[2 * p + 1]];
   if (nPairs) *nPairs = b ? b->nPairs : 0;
   if (nLeftTotal) *nLeftTotal = b && b->nPairs > 0 ? b->outOff[b->nPairs] : 0;
   if (nTiles) *nTiles = b ? b->nTiles : 0;
   if (pairsScored) *pairsScored = scored;
   return V3DM_OK;
}

} // extern "C"
