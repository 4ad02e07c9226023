// v3dm_voctree.cpp -- the C ABI declared in include/v3dm_voctree.h: the
// vocabulary tree on the GPU (quantisation kernels in k_voctree.hip) and the
// inverted file of V3D::LeafDocuments on the host, plus the candidate-pair
// stage of match_using_voctree.
//
// The inverted-file scores follow V3D/include/Base/v3d_voctree.h:143-289 in
// their arithmetic: the binomial score in double with the C library's log,
// the L1 distance in float in the reference's accumulation order (query
// leaves ascending), results sorted by (score, document).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/v3dm_voctree.h"
#include "v3dm_ctx.h"

namespace v3dm {
hipError_t launch_vt_root_groups(int n, int32_t *groupStart, int32_t *tileStart, uint32_t *nearTieCounts, int depth,
                                 hipStream_t st);
size_t vt_near_tie_bytes(int64_t n);
hipError_t launch_vt_level(int n, const float *keys, const float *tree, const float *treeT, const float *nodeNorm,
                           const int32_t *order, const int32_t *groupStart, const int32_t *tileStart, int nGroups,
                           int parentBase, int B, int last, int wordOffset, int32_t *leaf, int exact, void *nearTies,
                           uint32_t *nearTieCount, hipStream_t st, hipEvent_t start, hipEvent_t stop);
hipError_t launch_vt_regroup(int n, const int32_t *leaf, int parentBase, int nGroups, int32_t *counts,
                             int32_t *groupStart, int32_t *tileStart, int32_t *cursor, int32_t *order, hipStream_t st);
} // namespace v3dm

namespace {

constexpr int kMaxDepth = 8;
constexpr int64_t kMaxNodes = int64_t(1) << 26; // 32 GiB of FP32 nodes
constexpr int kMaxGroups = 1 << 24;             // regroup tables of one level

int vfail(std::string *err, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int vfail(std::string *err, int code, const char *fmt, ...)
{
   char buf[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(buf, sizeof buf, fmt, ap);
   va_end(ap);
   if (err) *err = buf;
   v3dm_host::g_createError = buf;
   return code;
}

} // namespace

struct v3dm_voctree
{
   int device = 0, depth = 0, B = 0;
   int64_t nodes = 0;
   std::vector<int64_t> levelStart; // first node of each level, levelStart[depth] = nodes
   hipStream_t stream = nullptr;
   std::string err;
   float *dTree = nullptr;
   float *dNodeNorm = nullptr; // per node an upper bound of its Euclidean norm (MFMA filter bound)
   float *dTreeT = nullptr;    // branch factor <= 64: per sibling group, 32 blocks of (64 children x 4 elements)
   int mode = 0;               // 0: MFMA filter where the branch factor allows, 1: exact VALU sums only
   // per-job buffers (grown)
   int64_t cap = 0, n = 0;
   float *dKeys = nullptr;
   int32_t *dLeaf = nullptr, *dOrder = nullptr;
   void *dNearTies = nullptr;          // keys the MFMA filter leaves with several candidates (one entry per key)
   uint32_t *dNearTieCounts = nullptr; // one counter per level
   int64_t groupCap = 0;
   int32_t *dCounts = nullptr, *dGroupStart = nullptr, *dTileStart = nullptr, *dCursor = nullptr;
   bool ran = false;
   std::vector<hipEvent_t> levelEv; // start/stop per level
   hipEvent_t runStart = nullptr, runStop = nullptr;
};

namespace {

#define VCHK(t, call)                                                                                        \
   do                                                                                                        \
   {                                                                                                         \
      hipError_t e_ = (call);                                                                                \
      if (e_ != hipSuccess) return vfail(&(t)->err, V3DM_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
   } while (0)

void freeJob(v3dm_voctree *t)
{
   for (void *p : {(void *)t->dKeys, (void *)t->dLeaf, (void *)t->dOrder, t->dNearTies})
      if (p) (void)hipFree(p);
   t->dKeys = nullptr;
   t->dLeaf = t->dOrder = nullptr;
   t->dNearTies = nullptr;
   t->cap = 0;
}

void freeGroups(v3dm_voctree *t)
{
   for (void *p : {(void *)t->dCounts, (void *)t->dGroupStart, (void *)t->dTileStart, (void *)t->dCursor})
      if (p) (void)hipFree(p);
   t->dCounts = t->dGroupStart = t->dTileStart = t->dCursor = nullptr;
   t->groupCap = 0;
}

int64_t groupsAt(const v3dm_voctree *t, int level) // groups (parents) at a level
{
   return level == 0 ? 1 : t->levelStart[level] - t->levelStart[level - 1];
}

// the largest regroup table a tree needs (levels 1 .. depth-1)
int64_t groupsAtLevelLimit(const std::vector<int64_t> &starts, int depth)
{
   int64_t g = 1;
   for (int d = 1; d < depth; ++d) g = std::max(g, starts[d] - starts[d - 1]);
   return g;
}

int makeTree(int device, int depth, int B, const float *nodes, v3dm_voctree **out)
{
   *out = nullptr;
   if (depth < 1 || depth > kMaxDepth || B < 1)
      return vfail(nullptr, V3DM_ERR_ARG, "voctree: depth %d / branch factor %d out of range", depth, B);
   std::vector<int64_t> starts(depth + 1, 0);
   int64_t len = 1, total = 0;
   for (int d = 0; d < depth; ++d)
   {
      starts[d] = total;
      len *= B;
      total += len;
      if (total > kMaxNodes) return vfail(nullptr, V3DM_ERR_UNSUPPORTED, "voctree: more than %lld nodes", (long long)kMaxNodes);
   }
   starts[depth] = total;
   if (starts[depth] - starts[depth - 1] > INT32_MAX || groupsAtLevelLimit(starts, depth) > kMaxGroups)
      return vfail(nullptr, V3DM_ERR_UNSUPPORTED, "voctree: too many leaves");
   int nDev = 0;
   if (hipGetDeviceCount(&nDev) != hipSuccess || nDev <= 0)
      return vfail(nullptr, V3DM_ERR_DEVICE, "voctree: no HIP device available");
   if (device < 0 || device >= nDev) return vfail(nullptr, V3DM_ERR_ARG, "voctree: device %d out of range", device);
   if (hipSetDevice(device) != hipSuccess) return vfail(nullptr, V3DM_ERR_DEVICE, "voctree: hipSetDevice failed");
   auto t = std::make_unique<v3dm_voctree>();
   t->device = device;
   t->depth = depth;
   t->B = B;
   t->nodes = total;
   t->levelStart = starts;
   int rc = V3DM_OK;
   auto chk = [&](hipError_t e, const char *what) {
      if (e != hipSuccess && rc == V3DM_OK)
         rc = vfail(nullptr, e == hipErrorOutOfMemory ? V3DM_ERR_NOMEM : V3DM_ERR_DEVICE, "voctree: %s: %s", what,
                    hipGetErrorString(e));
   };
   chk(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking), "hipStreamCreate");
   if (rc == V3DM_OK) chk(hipMalloc((void **)&t->dTree, sizeof(float) * kDim * (size_t)total), "hipMalloc(tree)");
   if (rc == V3DM_OK)
      chk(hipMemcpy(t->dTree, nodes, sizeof(float) * kDim * (size_t)total, hipMemcpyHostToDevice), "tree upload");
   if (rc == V3DM_OK)
   {
      // |y| in double, rounded up to float (inf for non-finite nodes: their children are then rescored exactly)
      std::vector<float> norms((size_t)total);
      for (int64_t j = 0; j < total; ++j)
      {
         double q = 0;
         for (int i = 0; i < kDim; ++i) q += (double)nodes[j * kDim + i] * (double)nodes[j * kDim + i];
         const double r = std::sqrt(q);
         float f = (float)r;
         if (std::isfinite(r) && (double)f < r) f = std::nextafter(f, INFINITY);
         norms[(size_t)j] = std::isfinite(r) ? f : INFINITY;
      }
      chk(hipMalloc((void **)&t->dNodeNorm, sizeof(float) * (size_t)total), "hipMalloc(norms)");
      if (rc == V3DM_OK)
         chk(hipMemcpy(t->dNodeNorm, norms.data(), sizeof(float) * (size_t)total, hipMemcpyHostToDevice), "norm upload");
   }
   if (rc == V3DM_OK && B <= 64)
   {
      // the MFMA filter's operand copy: sibling group g holds the children of node
      // g - 1 (nodes g B .. g B + B - 1); element e of child c at ((e / 4) 64 + c) 4 + e % 4
      const int64_t groups = starts[depth - 1] + 1;
      std::vector<float> tt((size_t)groups * kDim * 64, 0.0f);
      for (int64_t g = 0; g < groups; ++g)
         for (int c = 0; c < B; ++c)
         {
            const float *src = nodes + (size_t)(g * B + c) * kDim;
            float *dst = tt.data() + (size_t)g * kDim * 64;
            for (int e = 0; e < kDim; ++e) dst[((size_t)(e / 4) * 64 + c) * 4 + e % 4] = src[e];
         }
      chk(hipMalloc((void **)&t->dTreeT, sizeof(float) * tt.size()), "hipMalloc(tree copy)");
      if (rc == V3DM_OK)
         chk(hipMemcpy(t->dTreeT, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice), "tree copy upload");
   }
   t->levelEv.assign(2 * depth, nullptr);
   for (hipEvent_t &e : t->levelEv)
      if (rc == V3DM_OK) chk(hipEventCreate(&e), "hipEventCreate");
   if (rc == V3DM_OK) chk(hipMalloc((void **)&t->dNearTieCounts, sizeof(uint32_t) * kMaxDepth), "hipMalloc(counts)");
   if (rc == V3DM_OK) chk(hipEventCreate(&t->runStart), "hipEventCreate");
   if (rc == V3DM_OK) chk(hipEventCreate(&t->runStop), "hipEventCreate");
   if (rc != V3DM_OK)
   {
      v3dm_voctree_destroy(t.release());
      return rc;
   }
   *out = t.release();
   v3dm_host::g_createError.clear();
   return V3DM_OK;
}

int ensureJob(v3dm_voctree *t, int64_t n)
{
   if (n > t->cap)
   {
      freeJob(t);
      const int64_t c = std::max<int64_t>(n, 1024);
      VCHK(t, hipMalloc((void **)&t->dKeys, sizeof(float) * kDim * (size_t)c));
      VCHK(t, hipMalloc((void **)&t->dLeaf, sizeof(int32_t) * (size_t)c));
      VCHK(t, hipMalloc((void **)&t->dOrder, sizeof(int32_t) * (size_t)c));
      VCHK(t, hipMalloc(&t->dNearTies, vt_near_tie_bytes(c)));
      t->cap = c;
   }
   int64_t g = 2;
   for (int d = 1; d < t->depth; ++d) g = std::max(g, groupsAt(t, d));
   if (g + 1 > t->groupCap)
   {
      freeGroups(t);
      const size_t bytes = sizeof(int32_t) * (size_t)(g + 1);
      VCHK(t, hipMalloc((void **)&t->dCounts, bytes));
      VCHK(t, hipMalloc((void **)&t->dGroupStart, bytes));
      VCHK(t, hipMalloc((void **)&t->dTileStart, bytes));
      VCHK(t, hipMalloc((void **)&t->dCursor, bytes));
      t->groupCap = g + 1;
   }
   return V3DM_OK;
}

} // namespace

extern "C" {

int v3dm_voctree_create(int device, int depth, int branchFactor, const float *nodes, v3dm_voctree **out)
{
   if (!out || !nodes) return vfail(nullptr, V3DM_ERR_ARG, "voctree_create: null argument");
   return makeTree(device, depth, branchFactor, nodes, out);
}

int v3dm_voctree_load(int device, int depth, int branchFactor, const char *path, v3dm_voctree **out)
{
   if (!out || !path) return vfail(nullptr, V3DM_ERR_ARG, "voctree_load: null argument");
   *out = nullptr;
   if (depth < 1 || depth > kMaxDepth || branchFactor < 1)
      return vfail(nullptr, V3DM_ERR_ARG, "voctree_load: depth %d / branch factor %d out of range", depth, branchFactor);
   int64_t total = 0, len = 1;
   for (int d = 0; d < depth; ++d)
   {
      len *= branchFactor;
      total += len;
      if (total > kMaxNodes) return vfail(nullptr, V3DM_ERR_UNSUPPORTED, "voctree_load: tree too large");
   }
   FILE *f = std::fopen(path, "rb");
   if (!f) return vfail(nullptr, V3DM_ERR_ARG, "voctree_load: cannot open %s: %s", path, std::strerror(errno));
   std::vector<float> nodes((size_t)total * kDim);
   const size_t got = std::fread(nodes.data(), sizeof(float), nodes.size(), f);
   std::fclose(f);
   if (got != nodes.size())
      return vfail(nullptr, V3DM_ERR_ARG, "voctree_load: %s holds %zu floats, the tree needs %zu", path, got,
                   nodes.size());
   return makeTree(device, depth, branchFactor, nodes.data(), out);
}

void v3dm_voctree_destroy(v3dm_voctree *t)
{
   if (!t) return;
   (void)hipSetDevice(t->device);
   if (t->stream) (void)hipStreamSynchronize(t->stream);
   freeJob(t);
   freeGroups(t);
   if (t->dTree) (void)hipFree(t->dTree);
   if (t->dNodeNorm) (void)hipFree(t->dNodeNorm);
   if (t->dTreeT) (void)hipFree(t->dTreeT);
   if (t->dNearTieCounts) (void)hipFree(t->dNearTieCounts);
   for (hipEvent_t e : t->levelEv)
      if (e) (void)hipEventDestroy(e);
   if (t->runStart) (void)hipEventDestroy(t->runStart);
   if (t->runStop) (void)hipEventDestroy(t->runStop);
   if (t->stream) (void)hipStreamDestroy(t->stream);
   delete t;
}

const char *v3dm_voctree_last_error(const v3dm_voctree *t) { return t ? t->err.c_str() : v3dm_host::g_createError.c_str(); }

int v3dm_voctree_info(const v3dm_voctree *t, int *depth, int *branchFactor, int *nodes, int *leaves)
{
   if (!t) return vfail(nullptr, V3DM_ERR_ARG, "voctree_info: null tree");
   if (depth) *depth = t->depth;
   if (branchFactor) *branchFactor = t->B;
   if (nodes) *nodes = (int)std::min<int64_t>(t->nodes, INT32_MAX);
   if (leaves) *leaves = (int)(t->levelStart[t->depth] - t->levelStart[t->depth - 1]);
   return V3DM_OK;
}

int v3dm_voctree_set_mode(v3dm_voctree *t, int mode)
{
   if (!t || (mode != V3DM_VOCTREE_AUTO && mode != V3DM_VOCTREE_EXACT_VALU))
      return vfail(t ? &t->err : nullptr, V3DM_ERR_ARG, "voctree_set_mode: bad arguments");
   t->mode = mode;
   return V3DM_OK;
}

int v3dm_voctree_upload(v3dm_voctree *t, int64_t n, const float *keys)
{
   if (!t || n < 0 || (n > 0 && !keys)) return vfail(t ? &t->err : nullptr, V3DM_ERR_ARG, "voctree_upload: bad arguments");
   if (n > INT32_MAX / 2) return vfail(&t->err, V3DM_ERR_UNSUPPORTED, "voctree_upload: more than 2^30 keys");
   VCHK(t, hipSetDevice(t->device));
   int rc = ensureJob(t, n);
   if (rc) return rc;
   if (n) VCHK(t, hipMemcpyAsync(t->dKeys, keys, sizeof(float) * kDim * (size_t)n, hipMemcpyHostToDevice, t->stream));
   t->n = n;
   t->ran = false;
   return V3DM_OK;
}

int v3dm_voctree_run(v3dm_voctree *t)
{
   if (!t) return vfail(nullptr, V3DM_ERR_ARG, "voctree_run: null tree");
   VCHK(t, hipSetDevice(t->device));
   const int n = (int)t->n;
   VCHK(t, hipEventRecord(t->runStart, t->stream));
   VCHK(t, launch_vt_root_groups(n, t->dGroupStart, t->dTileStart, t->dNearTieCounts, t->depth, t->stream));
   for (int d = 0; d < t->depth; ++d)
   {
      const int64_t groups = groupsAt(t, d);
      const int parentBase = d == 0 ? -1 : (int)t->levelStart[d - 1];
      if (d > 0)
         VCHK(t, launch_vt_regroup(n, t->dLeaf, parentBase, (int)groups, t->dCounts, t->dGroupStart, t->dTileStart,
                                   t->dCursor, t->dOrder, t->stream));
      const bool last = d == t->depth - 1;
      VCHK(t, launch_vt_level(n, t->dKeys, t->dTree, t->dTreeT, t->dNodeNorm, d == 0 ? nullptr : t->dOrder, t->dGroupStart,
                              t->dTileStart, (int)groups, parentBase, t->B, last ? 1 : 0,
                              (int)t->levelStart[t->depth - 1], t->dLeaf, t->mode, t->dNearTies,
                              t->dNearTieCounts + d, t->stream, t->levelEv[2 * d], t->levelEv[2 * d + 1]));
   }
   VCHK(t, hipEventRecord(t->runStop, t->stream));
   t->ran = true;
   return V3DM_OK;
}

int v3dm_voctree_fetch(v3dm_voctree *t, int32_t *words)
{
   if (!t || (t->n > 0 && !words)) return vfail(t ? &t->err : nullptr, V3DM_ERR_ARG, "voctree_fetch: bad arguments");
   if (!t->ran) return vfail(&t->err, V3DM_ERR_STATE, "voctree_fetch: no run since the last upload");
   VCHK(t, hipSetDevice(t->device));
   if (t->n) VCHK(t, hipMemcpyAsync(words, t->dLeaf, sizeof(int32_t) * (size_t)t->n, hipMemcpyDeviceToHost, t->stream));
   VCHK(t, hipStreamSynchronize(t->stream));
   return V3DM_OK;
}

int v3dm_voctree_quantize(v3dm_voctree *t, int64_t n, const float *keys, int32_t *words)
{
   int rc = v3dm_voctree_upload(t, n, keys);
   if (!rc) rc = v3dm_voctree_run(t);
   if (!rc) rc = v3dm_voctree_fetch(t, words);
   return rc;
}

int v3dm_voctree_run_times(v3dm_voctree *t, double *scoringMs, double *runMs)
{
   if (!t) return vfail(nullptr, V3DM_ERR_ARG, "voctree_run_times: null tree");
   if (!t->ran) return vfail(&t->err, V3DM_ERR_STATE, "voctree_run_times: no run");
   VCHK(t, hipStreamSynchronize(t->stream));
   double s = 0;
   const int64_t maxTiles = (t->n + 63) / 64;
   for (int d = 0; d < t->depth && maxTiles > 0; ++d)
   {
      float ms = 0;
      VCHK(t, hipEventElapsedTime(&ms, t->levelEv[2 * d], t->levelEv[2 * d + 1]));
      s += ms;
   }
   float all = 0;
   VCHK(t, hipEventElapsedTime(&all, t->runStart, t->runStop));
   if (scoringMs) *scoringMs = s;
   if (runMs) *runMs = all;
   return V3DM_OK;
}

} // extern "C"

// ---- LeafDocuments -----------------------------------------------------------------

struct v3dm_leaf_documents
{
   int nLeaves = 0;
   std::vector<std::vector<std::pair<int, int>>> inverted; // per leaf: (document, count) in insertion order
   std::vector<int> nDocumentFeatures, nDocumentLeaves;
   std::vector<float> rcpNDocumentFeatures;
};

namespace {

int checkWords(int nLeaves, int n, const int32_t *words)
{
   if (n < 0 || (n > 0 && !words)) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents: bad word list");
   for (int i = 0; i < n; ++i)
      if (words[i] < 0 || words[i] >= nLeaves)
         return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents: word %d = %d outside [0, %d)", i, words[i], nLeaves);
   return V3DM_OK;
}

// getVisualWordHistogram (v3d_voctree.h:180-196): leaf -> count, ascending leaf.
std::map<int, int> wordHistogram(int n, const int32_t *words)
{
   std::map<int, int> h;
   for (int i = 0; i < n; ++i) ++h[words[i]];
   return h;
}

int emitSorted(std::vector<std::pair<float, int>> &scores, int nBest, float *outScores, int32_t *outDocs, int cap,
               int *count)
{
   // std::sort / std::partial_sort on (score, document) as the reference does
   const int m = (nBest < 0 || (int)scores.size() <= nBest) ? (int)scores.size() : nBest;
   if (m == (int)scores.size())
      std::sort(scores.begin(), scores.end());
   else
      std::partial_sort(scores.begin(), scores.begin() + m, scores.end());
   *count = m;
   if (m > cap || (m > 0 && (!outScores || !outDocs)))
      return vfail(nullptr, V3DM_ERR_CAPACITY, "leaf_documents: %d scores, capacity %d", m, cap);
   for (int i = 0; i < m; ++i)
   {
      outScores[i] = scores[(size_t)i].first;
      outDocs[i] = scores[(size_t)i].second;
   }
   return V3DM_OK;
}

void insertDoc(v3dm_leaf_documents *d, int n, const int32_t *words, int *documentId)
{
   const int id = (int)d->nDocumentFeatures.size();
   const std::map<int, int> h = wordHistogram(n, words);
   for (const auto &[leaf, count] : h) d->inverted[(size_t)leaf].push_back({id, count});
   d->nDocumentFeatures.push_back(n);
   d->nDocumentLeaves.push_back((int)h.size());
   d->rcpNDocumentFeatures.push_back(1.0f / (float)n);
   if (documentId) *documentId = id;
}

// computeBinomialScores (v3d_voctree.h:242-289) into (score, document) pairs.
void binomialScores(const v3dm_leaf_documents *d, double pMatch, int n, const int32_t *words,
                    std::vector<std::pair<float, int>> &scores)
{
   const int nDocs = (int)d->nDocumentFeatures.size();
   std::vector<int> histogram((size_t)nDocs, 0); // getDocumentHistogram (:166-178)
   for (int q = 0; q < n; ++q)
      for (const auto &e : d->inverted[(size_t)words[q]]) histogram[(size_t)e.first] += 1;
   scores.resize((size_t)nDocs);
   const double normalizer = 1.0 / (double)d->inverted.size();
   for (int i = 0; i < nDocs; ++i)
   {
      const double k = histogram[(size_t)i];
      const double nn = std::min(d->nDocumentFeatures[(size_t)i], n);
      const double p_random = std::max(d->nDocumentFeatures[(size_t)i], n) * normalizer;
      const double log_p0 = std::log(p_random);
      const double log_1_p0 = std::log(1.0 - p_random);
      const double p_matching = std::min(1.0, p_random + pMatch);
      const double log_p1 = std::log(p_matching);
      const double log_1_p1 = std::log(1.0 - p_matching);
      const double z0 = k * log_p0 + (nn - k) * log_1_p0;
      const double z1 = k * log_p1 + (nn - k) * log_1_p1;
      scores[(size_t)i] = {(float)(z0 - z1), i};
   }
}

} // namespace

extern "C" {

int v3dm_leaf_documents_create(int nLeaves, v3dm_leaf_documents **out)
{
   if (!out || nLeaves < 1) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_create: bad arguments");
   auto d = std::make_unique<v3dm_leaf_documents>();
   d->nLeaves = nLeaves;
   d->inverted.resize((size_t)nLeaves);
   *out = d.release();
   return V3DM_OK;
}

void v3dm_leaf_documents_destroy(v3dm_leaf_documents *d) { delete d; }

int v3dm_leaf_documents_insert(v3dm_leaf_documents *d, int n, const int32_t *words, int *documentId)
{
   if (!d) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_insert: null object");
   int rc = checkWords(d->nLeaves, n, words);
   if (rc) return rc;
   insertDoc(d, n, words, documentId);
   return V3DM_OK;
}

int v3dm_leaf_documents_histogram(const v3dm_leaf_documents *d, int n, const int32_t *words, int32_t *histogram,
                                  int cap)
{
   if (!d) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_histogram: null object");
   int rc = checkWords(d->nLeaves, n, words);
   if (rc) return rc;
   const int nDocs = (int)d->nDocumentFeatures.size();
   if (cap < nDocs || (nDocs > 0 && !histogram))
      return vfail(nullptr, V3DM_ERR_CAPACITY, "leaf_documents_histogram: %d documents, capacity %d", nDocs, cap);
   for (int q = 0; q < n; ++q)
      for (const auto &e : d->inverted[(size_t)words[q]]) histogram[(size_t)e.first] += 1;
   return V3DM_OK;
}

int v3dm_leaf_documents_sizes(const v3dm_leaf_documents *d, int *nLeaves, int *nDocuments, int64_t *nEntries)
{
   if (!d) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_sizes: null object");
   int64_t e = 0;
   for (const auto &l : d->inverted) e += (int64_t)l.size();
   if (nLeaves) *nLeaves = d->nLeaves;
   if (nDocuments) *nDocuments = (int)d->nDocumentFeatures.size();
   if (nEntries) *nEntries = e;
   return V3DM_OK;
}

int v3dm_leaf_documents_export(const v3dm_leaf_documents *d, int32_t *entriesPerLeaf, int32_t *entries,
                               int32_t *nDocumentFeatures, int32_t *nDocumentLeaves, float *rcpNDocumentFeatures)
{
   if (!d) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_export: null object");
   int64_t k = 0;
   for (size_t l = 0; l < d->inverted.size(); ++l)
   {
      if (entriesPerLeaf) entriesPerLeaf[l] = (int32_t)d->inverted[l].size();
      for (const auto &e : d->inverted[l])
      {
         if (entries)
         {
            entries[2 * k] = e.first;
            entries[2 * k + 1] = e.second;
         }
         ++k;
      }
   }
   for (size_t i = 0; i < d->nDocumentFeatures.size(); ++i)
   {
      if (nDocumentFeatures) nDocumentFeatures[i] = d->nDocumentFeatures[i];
      if (nDocumentLeaves) nDocumentLeaves[i] = d->nDocumentLeaves[i];
      if (rcpNDocumentFeatures) rcpNDocumentFeatures[i] = d->rcpNDocumentFeatures[i];
   }
   return V3DM_OK;
}

int v3dm_leaf_documents_import(v3dm_leaf_documents *d, int nLeaves, int nDocuments, const int32_t *entriesPerLeaf,
                               const int32_t *entries, const int32_t *nDocumentFeatures,
                               const int32_t *nDocumentLeaves, const float *rcpNDocumentFeatures)
{
   if (!d || nLeaves < 1 || nDocuments < 0 || !entriesPerLeaf ||
       (nDocuments > 0 && (!nDocumentFeatures || !nDocumentLeaves || !rcpNDocumentFeatures)))
      return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_import: bad arguments");
   std::vector<std::vector<std::pair<int, int>>> inv((size_t)nLeaves);
   int64_t k = 0;
   for (int l = 0; l < nLeaves; ++l)
   {
      if (entriesPerLeaf[l] < 0 || (entriesPerLeaf[l] > 0 && !entries))
         return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_import: leaf %d has %d entries", l, entriesPerLeaf[l]);
      for (int j = 0; j < entriesPerLeaf[l]; ++j, ++k)
      {
         const int doc = entries[2 * k], count = entries[2 * k + 1];
         if (doc < 0 || doc >= nDocuments)
            return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_import: leaf %d names document %d of %d", l, doc,
                         nDocuments);
         inv[(size_t)l].push_back({doc, count});
      }
   }
   d->nLeaves = nLeaves;
   d->inverted.swap(inv);
   d->nDocumentFeatures.assign(nDocumentFeatures, nDocumentFeatures + nDocuments);
   d->nDocumentLeaves.assign(nDocumentLeaves, nDocumentLeaves + nDocuments);
   d->rcpNDocumentFeatures.assign(rcpNDocumentFeatures, rcpNDocumentFeatures + nDocuments);
   return V3DM_OK;
}

int v3dm_leaf_documents_count(const v3dm_leaf_documents *d, int *nDocuments)
{
   if (!d || !nDocuments) return vfail(nullptr, V3DM_ERR_ARG, "leaf_documents_count: null argument");
   *nDocuments = (int)d->nDocumentFeatures.size();
   return V3DM_OK;
}

int v3dm_leaf_documents_binomial_scores(v3dm_leaf_documents *d, double pMatch, int n, const int32_t *words, int nBest,
                                        float *scores, int32_t *documents, int cap, int *count)
{
   if (!d || !count || cap < 0) return vfail(nullptr, V3DM_ERR_ARG, "binomial_scores: bad arguments");
   int rc = checkWords(d->nLeaves, n, words);
   if (rc) return rc;
   std::vector<std::pair<float, int>> s;
   binomialScores(d, pMatch, n, words, s);
   return emitSorted(s, nBest, scores, documents, cap, count);
}

int v3dm_leaf_documents_l1_scores(v3dm_leaf_documents *d, int n, const int32_t *words, int nBest, float *scores,
                                  int32_t *documents, int cap, int *count)
{
   if (!d || !count || cap < 0) return vfail(nullptr, V3DM_ERR_ARG, "l1_scores: bad arguments");
   int rc = checkWords(d->nLeaves, n, words);
   if (rc) return rc;
   const int nDocs = (int)d->nDocumentFeatures.size();
   const std::map<int, int> h = wordHistogram(n, words);
   const float rcpNVisualWords = 1.0f / (float)n;
   std::vector<float> dist((size_t)nDocs, 2.0f);
   for (const auto &[leaf, qcount] : h) // compute_L1_Scores (v3d_voctree.h:198-240)
   {
      const float qi = float(qcount) * rcpNVisualWords;
      for (const auto &[doc, dcount] : d->inverted[(size_t)leaf])
      {
         const float di = float(dcount) * d->rcpNDocumentFeatures[(size_t)doc];
         dist[(size_t)doc] += std::fabs(di - qi);
         dist[(size_t)doc] -= di;
         dist[(size_t)doc] -= qi;
      }
   }
   std::vector<std::pair<float, int>> s((size_t)nDocs);
   for (int i = 0; i < nDocs; ++i) s[(size_t)i] = {dist[(size_t)i], i};
   return emitSorted(s, nBest, scores, documents, cap, count);
}

int v3dm_voctree_candidate_pairs(int nViews, const int32_t *counts, const int32_t *words, int nLeaves, int nBest,
                                 double pMatch, int32_t *pairs, float *pairScores, int cap, int *count)
{
   if (nViews < 0 || (nViews > 0 && !counts) || nLeaves < 1 || !count || cap < 0)
      return vfail(nullptr, V3DM_ERR_ARG, "candidate_pairs: bad arguments");
   v3dm_leaf_documents d;
   d.nLeaves = nLeaves;
   d.inverted.resize((size_t)nLeaves);
   std::map<int, std::map<int, float>> chosen; // earlier view -> (view -> score), as the driver's map<int, set<int>>
   int64_t off = 0;
   std::vector<std::pair<float, int>> s;
   for (int i = 0; i < nViews; ++i)
   {
      if (counts[i] < 0) return vfail(nullptr, V3DM_ERR_ARG, "candidate_pairs: negative count");
      const int32_t *w = words + off;
      int rc = checkWords(nLeaves, counts[i], w);
      if (rc) return rc;
      binomialScores(&d, pMatch, counts[i], w, s);
      const int m = (nBest < 0 || (int)s.size() <= nBest) ? (int)s.size() : nBest;
      if (m == (int)s.size())
         std::sort(s.begin(), s.end());
      else
         std::partial_sort(s.begin(), s.begin() + m, s.end());
      for (int j = 0; j < m; ++j) chosen[s[(size_t)j].second].emplace(i, s[(size_t)j].first);
      insertDoc(&d, counts[i], w, nullptr);
      off += counts[i];
   }
   int64_t total = 0;
   for (const auto &c : chosen) total += (int64_t)c.second.size();
   if (total > INT32_MAX) return vfail(nullptr, V3DM_ERR_ARG, "candidate_pairs: too many pairs");
   *count = (int)total;
   if (total > cap || (total > 0 && !pairs)) return vfail(nullptr, V3DM_ERR_CAPACITY, "candidate_pairs: %lld pairs, capacity %d", (long long)total, cap);
   int64_t k = 0;
   for (const auto &[v1, others] : chosen)
      for (const auto &[v2, score] : others)
      {
         pairs[2 * k] = v1;
         pairs[2 * k + 1] = v2;
         if (pairScores) pairScores[k] = score;
         ++k;
      }
   v3dm_host::g_createError.clear();
   return V3DM_OK;
}

} // extern "C"
