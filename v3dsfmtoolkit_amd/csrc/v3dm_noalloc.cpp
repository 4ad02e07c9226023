// v3dm_noalloc.cpp -- the int8 findPutativeMatches job on device-resident
// uint8 descriptors without any device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The same kernels as v3dm_find_async (prep, fused scan,
// insertion when needed, finalize); their working memory -- int8 pools, row
// terms, top-2 keys, job counters, survivor regions -- lives in one
// file-scope device array per GPU, sized for kNoAllocMax features per side.
#include <mutex>

#include "v3dm_ctx.h"

namespace {

constexpr int kNoAllocMax = 8192;  // features per side
constexpr int kNoAllocVariant = 0; // scan tile variant: 4 waves, 256 x 512 tiles
constexpr int kNoAllocSurvCap = 64;
constexpr int kNoAllocTileM = 256, kNoAllocTileN = 512, kNoAllocWaves = 4; // geometry of that variant
constexpr int64_t kNoAllocRegions =
   (int64_t)(kNoAllocMax / kNoAllocTileM) * (kNoAllocMax / kNoAllocTileN) * kNoAllocWaves;

constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
// byte offsets inside the array
constexpr size_t kOffL = 0;
constexpr size_t kOffR = kOffL + align256((size_t)kNoAllocMax * kDim);
constexpr size_t kOffLsum = kOffR + align256((size_t)kNoAllocMax * kDim);
constexpr size_t kOffRsum = kOffLsum + align256(sizeof(int32_t) * kNoAllocMax);
constexpr size_t kOffKeys = kOffRsum + align256(sizeof(int32_t) * kNoAllocMax); // top1L, top2L, top1R, top2R
constexpr size_t kOffCounters = kOffKeys + align256(sizeof(uint64_t) * 4 * kNoAllocMax); // epoch, done, tile counter
constexpr size_t kOffSurv = kOffCounters + 256;
constexpr size_t kOffSurvCount = kOffSurv + align256(sizeof(uint64_t) * kNoAllocRegions * kNoAllocSurvCap);
constexpr size_t kScratchBytes = kOffSurvCount + align256(sizeof(int32_t) * kNoAllocRegions); // ~3.4 MiB

constexpr int kMaxDevices = 64;

struct DeviceScratch
{
   unsigned char *base = nullptr; // the array's address on this device
   uint32_t jobs = 0;             // epoch of the next job (the device advances its own copy)
};

std::mutex g_mu;
DeviceScratch g_scratch[kMaxDevices];

} // namespace

// Each GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

int fail0(int code, const char *fmt, ...)
{
   char buf[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof(buf), fmt, ap);
   va_end(ap);
   g_createError = buf;
   return code;
}

// Keys cleared to epoch 0 (no job uses it), epoch set to 1, counters to 0.
hipError_t clearScratchKeys(unsigned char *base, hipStream_t st)
{
   uint32_t *ctr = reinterpret_cast<uint32_t *>(base + kOffCounters);
   return launch_clear_keys(reinterpret_cast<uint64_t *>(base + kOffKeys), 4 * (int64_t)kNoAllocMax, ctr, ctr + 1,
                            st);
}

} // namespace

extern "C" int v3dm_find_u8_noalloc(const uint8_t *dLeft, int nLeft, const uint8_t *dRight, int nRight,
                                    float minScore, float minRatio, int32_t *dMatchOfLeft, void *stream)
{
   if (nLeft < 0 || nRight < 0)
      return fail0(V3DM_ERR_ARG, "find_u8_noalloc: negative feature count (%d, %d)", nLeft, nRight);
   if (nLeft > kNoAllocMax || nRight > kNoAllocMax)
      return fail0(V3DM_ERR_CAPACITY, "find_u8_noalloc: at most %d features per side (got %d x %d)", kNoAllocMax,
                   nLeft, nRight);
   if ((nLeft > 0 && (!dLeft || !dMatchOfLeft)) || (nRight > 0 && !dRight))
      return fail0(V3DM_ERR_ARG, "find_u8_noalloc: NULL device pointer");
   const hipStream_t st = static_cast<hipStream_t>(stream);
   int dev = 0;
   hipError_t e = hipGetDevice(&dev);
   if (e != hipSuccess) return fail0(V3DM_ERR_DEVICE, "find_u8_noalloc: hipGetDevice: %s", hipGetErrorString(e));
   if (dev < 0 || dev >= kMaxDevices) return fail0(V3DM_ERR_UNSUPPORTED, "find_u8_noalloc: device %d", dev);

   // host-side bookkeeping is serialised; jobs that overlap on one GPU share
   // its array and race
   std::lock_guard<std::mutex> lock(g_mu);
   DeviceScratch &d = g_scratch[dev];
   if (!d.base)
   {
      // first call on this device (which is current): the array's device
      // address; the name `scratch` itself is only a host-side stand-in
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
         return fail0(V3DM_ERR_DEVICE, "find_u8_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      e = clearScratchKeys(static_cast<unsigned char *>(p), st);
      if (e != hipSuccess) return fail0(V3DM_ERR_DEVICE, "find_u8_noalloc: %s", hipGetErrorString(e));
      d.base = static_cast<unsigned char *>(p);
      d.jobs = 1;
   }
   else if (d.jobs >= kEpochLimit)
   {
      e = clearScratchKeys(d.base, st);
      if (e != hipSuccess) return fail0(V3DM_ERR_DEVICE, "find_u8_noalloc: %s", hipGetErrorString(e));
      d.jobs = 1;
   }
   unsigned char *b = d.base;
   int8_t *poolL = reinterpret_cast<int8_t *>(b + kOffL);
   int8_t *poolR = reinterpret_cast<int8_t *>(b + kOffR);
   int32_t *sumL = reinterpret_cast<int32_t *>(b + kOffLsum);
   int32_t *sumR = reinterpret_cast<int32_t *>(b + kOffRsum);
   uint64_t *keys = reinterpret_cast<uint64_t *>(b + kOffKeys);
   uint32_t *ctr = reinterpret_cast<uint32_t *>(b + kOffCounters);

   // setLeftFeatures / setRightFeatures: counts rounded down to multiples of
   // 32 (v3d_cudamatching.cpp:68-77), rows padded to 128
   const int nL = nLeft - (nLeft & 31), nR = nRight - (nRight & 31);
   const int nPadL = std::max(roundUp(std::max(nL, 1), kRowPad), kRowPad);
   const int nPadR = std::max(roundUp(std::max(nR, 1), kRowPad), kRowPad);
   const int v = kNoAllocVariant;
   if (scan_variant_tile_m(v) != kNoAllocTileM || scan_variant_tile_n(v) != kNoAllocTileN ||
       scan_variant_waves(v) != kNoAllocWaves)
      return fail0(V3DM_ERR_STATE, "find_u8_noalloc: scan variant geometry changed");
   e = launch_prep_u8(dLeft, nL, nPadL, poolL, sumL, st);
   if (e == hipSuccess) e = launch_prep_u8(dRight, nR, nPadR, poolR, sumR, st);

   PairDesc pd{};
   pd.nL = nL;
   pd.nR = nR;
   pd.tilesM = (roundUp(nL, kRowPad) + kNoAllocTileM - 1) / kNoAllocTileM;
   pd.tilesN = (roundUp(nR, kRowPad) + kNoAllocTileN - 1) / kNoAllocTileN;
   ScanArgs a{};
   a.pair0 = pd;
   a.L = poolL;
   a.Lsum = sumL;
   a.R = poolR;
   a.Rsum = sumR;
   a.pairs = nullptr; // a single pair is passed by value
   a.nPairs = 1;
   a.nTiles = (nL > 0 && nR > 0) ? pd.tilesM * pd.tilesN : 0;
   a.thr = survivorThreshold(minScore, minRatio);
   a.top1L = keys;
   a.top2L = keys + kNoAllocMax;
   a.top1R = keys + 2 * kNoAllocMax;
   a.top2R = keys + 3 * kNoAllocMax;
   a.epoch = ctr;
   a.surv = reinterpret_cast<uint64_t *>(b + kOffSurv);
   a.survCount = reinterpret_cast<int32_t *>(b + kOffSurvCount);
   a.survCap = kNoAllocSurvCap;
   a.inlineInsert = a.nTiles <= scan_variant_resident(v, dev) ? 1 : 0;
   a.waves = kNoAllocWaves;
   a.tileM = kNoAllocTileM;
   a.tileN = kNoAllocTileN;
   a.persistent = 0;
   a.tileCounter = ctr + 2;
   if (e == hipSuccess) e = launch_i8_scan(v, a.nTiles, a, st);
   if (e == hipSuccess && !a.inlineInsert)
      e = launch_i8_insert(a, insert_form_for((int64_t)a.nTiles * a.waves, dev, -1), st);
   FinalizeArgs f{};
   f.pairs = nullptr;
   f.nPairs = 1;
   f.nBlocks = std::max(1, (nL + 255) / 256);
   f.top1L = a.top1L;
   f.top2L = a.top2L;
   f.top1R = a.top1R;
   f.top2R = a.top2R;
   f.epoch = ctr;
   f.done = ctr + 1;
   f.tileCounter = ctr + 2;
   f.minScore = minScore;
   f.minRatio = minRatio;
   f.matchOfLeft = dMatchOfLeft;
   f.pair0 = pd;
   f.hasPair0 = 1;
   if (e == hipSuccess) e = launch_i8_finalize(f, st);
   if (e != hipSuccess) return fail0(V3DM_ERR_DEVICE, "find_u8_noalloc: %s", hipGetErrorString(e));
   d.jobs += 1; // the finalize's last block advanced the device epoch
   return V3DM_OK;
}
