// driver_common.h -- the part the counterpart drivers share: the
// configuration file, loading views from <entry>.features, and the batched
// putative-matching stage that writes the raw correspondences to the match
// database (SfM/match_sequence/src/match_sequence.cpp:134-229 and
// SfM/match_using_voctree/src/match_using_voctree.cpp:251-323 do the same
// one pair at a time).
#pragma once

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "v3dm_io.h"
#include "v3dmatch.h"

namespace v3dm_driver {

constexpr int kDim = 128;
constexpr float kScoreThreshold = 0.9f; // match_sequence.cpp:79, match_using_voctree.cpp:102

// V3D::ConfigurationFile (V3D/include/Base/v3d_cfgfile.h:28-117): "KEY VALUE"
// lines, '#' comments, the first occurrence of a key wins, a missing file is
// an empty configuration.
struct Config
{
   std::map<std::string, std::string> kv;
   explicit Config(const char *path)
   {
      std::ifstream is(path);
      std::string line, key, value;
      while (std::getline(is, line))
      {
         if (line.empty() || line[0] == '#') continue;
         std::istringstream ss(line);
         key.clear();
         value.clear();
         ss >> key;
         if (key.empty() || key[0] == '#') continue;
         ss >> value;
         kv.insert({key, value});
      }
   }
   template <typename T> T get(const char *key, T def) const
   {
      auto p = kv.find(key);
      if (p == kv.end() || p->second.empty()) return def;
      std::istringstream ss(p->second);
      T v;
      ss >> v;
      return v;
   }
};

struct View
{
   std::string name;
   bool used = false;
   int count = 0; // features in the file
   int n = 0;     // features matched: min(MATCHING_MAX_FEATURES, count & 0xffe0), rounded to 32
   std::vector<int32_t> ids;
   std::vector<float> pos;  // 2 per feature
   std::vector<float> desc; // count x 128 until the matching stage takes the first n
};

[[noreturn]] inline void die(const std::string &msg)
{
   std::cerr << "Exception caught: " << msg << std::endl;
   std::exit(1);
}

inline void check(int rc, const char *what, v3dm_ctx *ctx = nullptr)
{
   if (rc != V3DM_OK) die(std::string(what) + ": " + v3dm_last_error(ctx));
}

inline std::vector<View> readImageList(const char *path)
{
   std::ifstream is(path);
   if (!is) die(std::string("cannot open image list ") + path);
   std::vector<View> views;
   std::string name;
   while (is >> name)
   {
      views.emplace_back();
      views.back().name = name;
   }
   return views;
}

// serializeDataFromFile(<entry>.features, SerializableVector<SIFT_Feature>&)
// (match_sequence.cpp:22-33); keeps ids, positions and descriptors.
inline void loadView(View &v, int nMaxFeatures, bool quiet)
{
   const std::string path = v.name + ".features";
   if (!quiet) std::cout << "reading " << path << std::endl;
   int count = 0, format = 0;
   check(v3dm_features_info(path.c_str(), &count, &format), "reading features");
   v.ids.resize(count);
   v.pos.resize(2 * (size_t)count);
   v.desc.resize((size_t)count * kDim);
   check(v3dm_features_read(path.c_str(), count, v.ids.data(), v.pos.data(), nullptr, nullptr, nullptr, v.desc.data(),
                            &count),
         "reading features");
   v.count = count;
   // min(nMax, size & 0xffe0) (match_sequence.cpp:152,184), then setLeft/RightFeatures'
   // rounding to a multiple of 32, which matters only for an nMax that is not one
   v.n = v3dm_truncated_feature_count(count, nMaxFeatures) & ~31;
}

// The geometric stage after the putative matches (match_sequence.cpp:229-346,
// match_using_voctree.cpp:325-491): robust orientation through the hook, the
// inlier checks, guided matching with F = K2^-T [T]x R K1^-1 (one batched
// job over all accepted pairs), the effective-inlier check, and rows with
// the pose and normalised positions.  Off (raw putative rows, identity pose)
// when no hook is given.
struct GeometryOptions
{
   bool enabled = false;
   v3dm_orientation_hook hook{};
   std::string calibPath = "calibration_db.txt";
   float egInlierThreshold = 2.0f / 1024; // EG_INLIER_THRESHOLD / 1024, held in a float (:95)
   int nMinInlierCorrs = 100;
   double minInlierRatio = 0.5;
   bool useEffectiveInliers = true;
   bool useGuidedMatching = true;
   bool visibilityTest = false;           // match_using_voctree.cpp:397-418 (that driver also
                                          // does not print the inlier threshold)
};

// The configuration keys of the geometric stage (match_sequence.cpp:95-104;
// match_using_voctree.cpp:113-129 reads USE_EFFECTIVE_INLIERS with default false).
inline void readGeometryConfig(const Config &cf, int nMinRawCorrs, bool effectiveDefault, GeometryOptions &g)
{
   g.egInlierThreshold = (float)(cf.get("EG_INLIER_THRESHOLD", 2.0) / 1024);
   g.nMinInlierCorrs = cf.get("MINIMUM_INLIER_CORRESPONDENCES", -1);
   if (g.nMinInlierCorrs < 0) g.nMinInlierCorrs = nMinRawCorrs;
   g.minInlierRatio = cf.get("MINIMUM_INLIER_MATCHES_RATIO", 0.5);
   g.useEffectiveInliers = cf.get("USE_EFFECTIVE_INLIERS", effectiveDefault);
   g.useGuidedMatching = cf.get("USE_GUIDED_MATCHING", true);
}

// --orientation LIB: a plugin library exporting V3DM_ORIENTATION_PLUGIN_SYMBOL.
inline void loadOrientationPlugin(const std::string &path, v3dm_orientation_hook &hook)
{
   void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
   if (!h) die("cannot load orientation plugin " + path + ": " + dlerror());
   auto fn = (v3dm_orientation_plugin_fn)dlsym(h, V3DM_ORIENTATION_PLUGIN_SYMBOL);
   if (!fn) die(path + " exports no " V3DM_ORIENTATION_PLUGIN_SYMBOL);
   hook = v3dm_orientation_hook{};
   if (fn(&hook) != 0 || !hook.robust_orientation) die(path + ": the plugin did not supply robust_orientation");
}

// --poses FILE: known relative poses ("view1 view2 R (9, row-major) T (3)" per
// line, '#' comments), e.g. from an earlier reconstruction; inliers are the
// raw correspondences within the guided-matching threshold of their pose's
// epipolar geometry (v3dm_sampson_inliers).  A pair without a pose is skipped.
struct KnownPoses
{
   std::map<std::pair<int, int>, std::vector<double>> poses;

   explicit KnownPoses(const std::string &path)
   {
      std::ifstream is(path);
      if (!is) die("cannot open poses file " + path);
      std::string line;
      while (std::getline(is, line))
      {
         if (line.empty() || line[0] == '#') continue;
         std::istringstream ss(line);
         int a, b;
         std::vector<double> rt(12);
         if (!(ss >> a >> b)) continue;
         for (double &x : rt)
            if (!(ss >> x)) die("poses file " + path + ": incomplete line: " + line);
         poses[{a, b}] = rt;
      }
   }

   static int orientation(void *user, int v1, int v2, const double K1[9], const double K2[9], double thr, int n,
                          const v3dm_correspondence *corrs, double R[9], double T[3], int32_t *inliers, int *nInliers)
   {
      const KnownPoses *kp = (const KnownPoses *)user;
      auto it = kp->poses.find({v1, v2});
      if (it == kp->poses.end()) return 1;
      std::copy(it->second.begin(), it->second.begin() + 9, R);
      std::copy(it->second.begin() + 9, it->second.end(), T);
      double F[9];
      if (v3dm_fundamental_from_pose(K1, K2, R, T, F) != V3DM_OK) return 1;
      return v3dm_sampson_inliers(F, thr, n, corrs, inliers, nInliers) == V3DM_OK ? 0 : 1;
   }

   v3dm_orientation_hook hook()
   {
      v3dm_orientation_hook h{};
      h.user = this;
      h.robust_orientation = &KnownPoses::orientation;
      return h;
   }
};

struct MatchOptions
{
   float distRatio = 0.8f;
   int nMinRawCorrs = 100;
   int device = 0;
   int pairsPerJob = 4096;
   // scan tile variant of the batched jobs: 1 = 256 x 1024 tiles, the fastest on
   // large batches; -1 = time every variant on the first job (worth it only with
   // the V3DM_TUNE_CACHE file, since each run is a new process)
   int variant = 1;
   bool quiet = false;
   std::string dbPath; // empty: no database
   GeometryOptions geo;
};

namespace detail {

// PointCorrespondences of a match table (match_sequence.cpp:205-222): ids and
// pixel positions, ascending left index.
inline void tableToCorrs(const int32_t *table, int nL, int v1, int v2, const View &a, const View &b,
                         std::vector<v3dm_correspondence> &corrs)
{
   corrs.clear();
   for (int i1 = 0; i1 < nL; ++i1)
   {
      const int32_t i2 = table[i1];
      if (i2 < 0) continue;
      v3dm_correspondence c;
      c.left = {a.pos[2 * (size_t)i1], a.pos[2 * (size_t)i1 + 1], v1, a.ids[(size_t)i1]};
      c.right = {b.pos[2 * (size_t)i2], b.pos[2 * (size_t)i2 + 1], v2, b.ids[(size_t)i2]};
      corrs.push_back(c);
   }
}

// One batched job per chunk of pairs: table entries per pair (kind/F/threshold:
// per-pair guided masks, or empty for the putative lists).
template <class Sink>
inline double runJobs(v3dm_ctx *ctx, const std::vector<View> &views, const std::vector<int32_t> &pairs, int perJob,
                      float minScore, float minRatio, const std::vector<double> *F, const std::vector<float> *thr,
                      int64_t &scored, Sink sink)
{
   const int nPairs = (int)(pairs.size() / 2);
   std::vector<int32_t> table, kinds;
   std::vector<int64_t> offsets;
   double ms = 0;
   for (int p0 = 0; p0 < nPairs; p0 += perJob)
   {
      const int np = std::min(perJob, nPairs - p0);
      const int32_t *pp = pairs.data() + 2 * (size_t)p0;
      check(v3dm_batch_set_pairs(ctx, np, pp), "v3dm_batch_set_pairs", ctx);
      if (F)
      {
         kinds.assign(np, 1);
         check(v3dm_batch_set_guides(ctx, kinds.data(), F->data() + 9 * (size_t)p0, thr->data() + p0),
               "v3dm_batch_set_guides", ctx);
      }
      int64_t lefts = 0;
      for (int p = 0; p < np; ++p)
      {
         lefts += views[pp[2 * p]].n;
         scored += (int64_t)views[pp[2 * p]].n * views[pp[2 * p + 1]].n;
      }
      table.resize((size_t)std::max<int64_t>(lefts, 1));
      offsets.resize((size_t)np + 1);
      const auto m0 = std::chrono::steady_clock::now();
      check(v3dm_batch_match(ctx, minScore, minRatio, table.data(), offsets.data()), "v3dm_batch_match", ctx);
      ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0).count();
      for (int p = 0; p < np; ++p) sink(p0 + p, table.data() + offsets[p], (int)(offsets[p + 1] - offsets[p]));
   }
   return ms;
}

} // namespace detail

// Matches every (left view, right view) pair as batched jobs.  Without the
// geometric stage, the pairs with at least nMinRawCorrs raw correspondences
// are written as PairwiseMatch rows (rotation identity, translation zero,
// pixel positions); with it, the pairs that pass robust orientation, guided
// matching and the inlier checks are written with their pose and normalised
// positions, as the reference drivers store them.  OIDs 0, 1, ... in pair
// order.  Views must be loaded; their descriptors are released.  Returns the
// number of stored pairs.
inline int matchPairs(std::vector<View> &views, const std::vector<int32_t> &pairs, const MatchOptions &o)
{
   const int nViews = (int)views.size(), nPairs = (int)(pairs.size() / 2);
   const GeometryOptions &g = o.geo;
   int64_t rows = 0;
   for (const View &v : views) rows += v.n;
   std::vector<int32_t> counts(nViews);
   std::vector<float> all, allPos;
   all.reserve((size_t)rows * kDim);
   if (g.enabled) allPos.reserve(2 * (size_t)rows);
   for (int i = 0; i < nViews; ++i)
   {
      counts[i] = views[i].n;
      all.insert(all.end(), views[i].desc.begin(), views[i].desc.begin() + (size_t)views[i].n * kDim);
      if (g.enabled) allPos.insert(allPos.end(), views[i].pos.begin(), views[i].pos.begin() + 2 * (size_t)views[i].n);
      std::vector<float>().swap(views[i].desc);
   }
   std::vector<uint8_t> allU8(all.size());
   const bool useU8 = v3dm_descriptors_to_u8(rows, all.data(), allU8.data()) == V3DM_OK;

   // calibration of every view (the geometric stage only)
   std::vector<double> Ks;
   std::vector<int32_t> dims;
   if (g.enabled)
   {
      int nc = 0;
      v3dm_calibration_read(g.calibPath.c_str(), 0, nullptr, nullptr, nullptr, &nc);
      Ks.resize(9 * (size_t)std::max(nc, 1));
      dims.resize(2 * (size_t)std::max(nc, 1));
      check(v3dm_calibration_read(g.calibPath.c_str(), nc, Ks.data(), nullptr, dims.data(), &nc),
            "reading the calibration database");
      for (int32_t v : pairs)
         if (v >= nc) die("view id out of range of the calibration database " + g.calibPath);
   }

   v3dm_ctx *ctx = nullptr;
   check(v3dm_create(o.device, &ctx), "v3dm_create");
   check(v3dm_batch_set_variant(ctx, o.variant), "v3dm_batch_set_variant", ctx);
   const auto t0 = std::chrono::steady_clock::now();
   if (useU8)
      check(v3dm_batch_set_images_u8(ctx, nViews, counts.data(), allU8.data()), "upload", ctx);
   else
      check(v3dm_batch_set_images_f32(ctx, nViews, counts.data(), all.data()), "upload", ctx);
   if (g.enabled) check(v3dm_batch_set_positions(ctx, allPos.data()), "upload", ctx);
   std::vector<float>().swap(all);
   std::vector<float>().swap(allPos);
   std::vector<uint8_t>().swap(allU8);
   const auto t1 = std::chrono::steady_clock::now();

   // per pair: the log block (printed in pair order at the end), the result to store
   struct Result
   {
      bool store = false;
      double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, T[3] = {0, 0, 0};
      std::vector<v3dm_correspondence> corrs;
   };
   std::vector<Result> res(nPairs);
   std::vector<std::ostringstream> log(nPairs);
   auto say = [&](int p) -> std::ostream & { return log[p]; };

   int64_t scored = 0, totalCorrs = 0;
   std::vector<v3dm_correspondence> corrs;
   const double matchMs = detail::runJobs(
      ctx, views, pairs, o.pairsPerJob, kScoreThreshold, o.distRatio, nullptr, nullptr, scored,
      [&](int p, const int32_t *table, int nL) {
         const int v1 = pairs[2 * (size_t)p], v2 = pairs[2 * (size_t)p + 1];
         detail::tableToCorrs(table, nL, v1, v2, views[v1], views[v2], corrs);
         totalCorrs += (int64_t)corrs.size();
         say(p) << "--------------------\n"
                << "matching view " << v1 << " with view " << v2 << "\n"
                << "nFeatures1 = " << views[v1].n << ", nFeatures2 = " << views[v2].n << "\n"
                << "matches.size() = " << corrs.size() << "\n";
         if ((int)corrs.size() < o.nMinRawCorrs) return;
         res[p].store = true;
         res[p].corrs = corrs;
      });

   double guidedMs = 0;
   int nGuided = 0;
   if (g.enabled)
   {
      // robust orientation and the inlier checks (match_sequence.cpp:229-264)
      std::vector<int32_t> gPairs, inl;
      std::vector<int> gIndex;
      std::vector<double> gF;
      std::vector<float> gThr;
      for (int p = 0; p < nPairs; ++p)
      {
         Result &r = res[p];
         if (!r.store) continue;
         r.store = false;
         const int v1 = pairs[2 * (size_t)p], v2 = pairs[2 * (size_t)p + 1];
         const double *K1 = &Ks[9 * (size_t)v1], *K2 = &Ks[9 * (size_t)v2];
         const int maxWidth = std::max(std::max(dims[2 * (size_t)v1], dims[2 * (size_t)v1 + 1]),
                                       std::max(dims[2 * (size_t)v2], dims[2 * (size_t)v2 + 1]));
         const double inlierThreshold = maxWidth * g.egInlierThreshold; // int * float, as the reference
         std::ostream &os = say(p);
         os << "RANSAC for essential matrix...\n";
         if (!g.visibilityTest) os << "inlierThreshold = " << inlierThreshold << "\n";
         inl.assign(r.corrs.size(), 0);
         int nIn = 0;
         const int rc = g.hook.robust_orientation(g.hook.user, v1, v2, K1, K2, inlierThreshold, (int)r.corrs.size(),
                                                  r.corrs.data(), r.R, r.T, inl.data(), &nIn);
         os << "done.\n";
         if (rc != 0 || nIn < 0 || nIn > (int)r.corrs.size())
         {
            os << " This pair is ignored because no relative pose was found.\n";
            continue;
         }
         for (int k = 0; k < nIn; ++k)
            if (inl[k] < 0 || inl[k] >= (int)r.corrs.size()) die("orientation hook returned an inlier out of range");
         os << " this pair has " << nIn << " inliers.\n";
         if (nIn < g.nMinInlierCorrs)
         {
            os << " This pair is ignored because of too few inliers.\n";
            continue;
         }
         if (double(nIn) < g.minInlierRatio * r.corrs.size())
         {
            os << " This pair is ignored because the number of matches dropped too much.\n";
            continue;
         }
         std::vector<v3dm_correspondence> in(nIn);
         for (int k = 0; k < nIn; ++k) in[k] = r.corrs[inl[k]];
         r.corrs.swap(in);
         const int w = dims[2 * (size_t)v1], h = dims[2 * (size_t)v1 + 1];
         if (!g.useGuidedMatching)
         {
            int nEff = (int)r.corrs.size();
            if (g.useEffectiveInliers)
               check(v3dm_effective_inlier_count((int)r.corrs.size(), r.corrs.data(), w, h, &nEff),
                     "effective inlier count");
            os << "Effective inlier count is " << nEff << "\n";
            if (nEff < g.nMinInlierCorrs) continue;
            r.store = true;
            continue;
         }
         if (g.hook.refine_pose)
         {
            os << "refineRelativePose()...\n";
            if (g.hook.refine_pose(g.hook.user, v1, v2, K1, K2, (int)r.corrs.size(), r.corrs.data(), r.R, r.T) != 0)
               continue;
            os << "done.\n";
         }
         double F[9];
         check(v3dm_fundamental_from_pose(K1, K2, r.R, r.T, F), "fundamental matrix");
         gPairs.push_back(v1);
         gPairs.push_back(v2);
         gIndex.push_back(p);
         gF.insert(gF.end(), F, F + 9);
         gThr.push_back((float)inlierThreshold); // runGuidedMatching's float distanceThreshold
      }

      // guided matching of every accepted pair in batched jobs (match_sequence.cpp:275-323)
      nGuided = (int)gIndex.size();
      guidedMs = detail::runJobs(
         ctx, views, gPairs, o.pairsPerJob, kScoreThreshold, o.distRatio, &gF, &gThr, scored,
         [&](int q, const int32_t *table, int nL) {
            const int p = gIndex[q];
            Result &r = res[p];
            const int v1 = pairs[2 * (size_t)p], v2 = pairs[2 * (size_t)p + 1];
            const double *K1 = &Ks[9 * (size_t)v1], *K2 = &Ks[9 * (size_t)v2];
            std::ostream &os = say(p);
            detail::tableToCorrs(table, nL, v1, v2, views[v1], views[v2], r.corrs);
            os << "matches.size() after guided matching: " << r.corrs.size() << "\n";
            if (g.visibilityTest)
            {
               const int before = (int)r.corrs.size();
               int kept = 0;
               check(v3dm_filter_in_front(K1, K2, r.R, r.T, before, r.corrs.data(), &kept), "visibility test");
               r.corrs.resize(kept);
               os << "The number of inliers dropped from " << before << " to " << kept << " after visibility test.\n";
            }
            int nEff = (int)r.corrs.size();
            if (g.useEffectiveInliers)
               check(v3dm_effective_inlier_count((int)r.corrs.size(), r.corrs.data(), dims[2 * (size_t)v1],
                                                 dims[2 * (size_t)v1 + 1], &nEff),
                     "effective inlier count");
            os << "Effective inlier count is " << nEff << "\n";
            if (nEff < g.nMinInlierCorrs) return;
            if (g.hook.refine_pose)
            {
               os << "refineRelativePose() after guided matching:\n";
               if (g.hook.refine_pose(g.hook.user, v1, v2, K1, K2, (int)r.corrs.size(), r.corrs.data(), r.R, r.T) != 0)
                  return;
            }
            r.store = true;
         });
   }

   v3dm_match_db *db = nullptr;
   if (!o.dbPath.empty()) check(v3dm_match_db_open(o.dbPath.c_str(), 1, &db), "opening the match database");
   int stored = 0;
   for (int p = 0; p < nPairs; ++p)
   {
      if (!o.quiet) std::cout << log[p].str();
      Result &r = res[p];
      if (!r.store) continue;
      const int v1 = pairs[2 * (size_t)p], v2 = pairs[2 * (size_t)p + 1];
      if (g.enabled)
         check(v3dm_normalize_correspondences(&Ks[9 * (size_t)v1], &Ks[9 * (size_t)v2], (int)r.corrs.size(),
                                              r.corrs.data()),
               "normalising positions");
      if (db)
         check(v3dm_match_db_put(db, stored, v1, v2, r.R, r.T, (int)r.corrs.size(), r.corrs.data()),
               "writing the match database");
      ++stored;
   }
   if (db) check(v3dm_match_db_close(db), "closing the match database");
   v3dm_destroy(ctx);

   const double uploadMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
   std::cout << "matched " << nPairs << " view pairs (" << scored << " descriptor pairs, "
             << (useU8 ? "uint8" : "float") << " pool) in " << matchMs << " ms after a " << uploadMs
             << " ms upload; " << totalCorrs << " putative matches; ";
   if (g.enabled)
      std::cout << nGuided << " pairs guided-matched in " << guidedMs << " ms; " << stored
                << " pairs passed the geometric checks";
   else
      std::cout << stored << " pairs with at least " << o.nMinRawCorrs << " matches";
   std::cout << (db ? " written to " + o.dbPath : std::string()) << std::endl;
   return stored;
}

} // namespace v3dm_driver
