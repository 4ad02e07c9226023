// driver_common.h -- the part the counterpart drivers share: the
// configuration file, loading views from <entry>.features, and the batched
// putative-matching stage that writes the raw correspondences to the match
// database (SfM/match_sequence/src/match_sequence.cpp:134-229 and
// SfM/match_using_voctree/src/match_using_voctree.cpp:251-323 do the same
// one pair at a time).
#pragma once

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
};

// Matches every (left view, right view) pair as batched jobs and writes the
// pairs with at least nMinRawCorrs raw correspondences to the database as
// PairwiseMatch rows (rotation identity, translation zero, pixel positions),
// OIDs 0, 1, ... in pair order.  Views must be loaded; their descriptors are
// released.  Returns the number of stored pairs.
inline int matchPairs(std::vector<View> &views, const std::vector<int32_t> &pairs, const MatchOptions &o)
{
   const int nViews = (int)views.size(), nPairs = (int)(pairs.size() / 2);
   int64_t rows = 0;
   for (const View &v : views) rows += v.n;
   std::vector<int32_t> counts(nViews);
   std::vector<float> all;
   all.reserve((size_t)rows * kDim);
   for (int i = 0; i < nViews; ++i)
   {
      counts[i] = views[i].n;
      all.insert(all.end(), views[i].desc.begin(), views[i].desc.begin() + (size_t)views[i].n * kDim);
      std::vector<float>().swap(views[i].desc);
   }
   std::vector<uint8_t> allU8(all.size());
   const bool useU8 = v3dm_descriptors_to_u8(rows, all.data(), allU8.data()) == V3DM_OK;

   v3dm_ctx *ctx = nullptr;
   check(v3dm_create(o.device, &ctx), "v3dm_create");
   check(v3dm_batch_set_variant(ctx, o.variant), "v3dm_batch_set_variant", ctx);
   const auto t0 = std::chrono::steady_clock::now();
   if (useU8)
      check(v3dm_batch_set_images_u8(ctx, nViews, counts.data(), allU8.data()), "upload", ctx);
   else
      check(v3dm_batch_set_images_f32(ctx, nViews, counts.data(), all.data()), "upload", ctx);
   std::vector<float>().swap(all);
   std::vector<uint8_t>().swap(allU8);
   const auto t1 = std::chrono::steady_clock::now();

   v3dm_match_db *db = nullptr;
   if (!o.dbPath.empty()) check(v3dm_match_db_open(o.dbPath.c_str(), 1, &db), "opening the match database");
   const double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, T[3] = {0, 0, 0};

   double matchMs = 0;
   int64_t scored = 0, totalCorrs = 0;
   int stored = 0;
   std::vector<int32_t> table;
   std::vector<int64_t> offsets;
   std::vector<v3dm_correspondence> corrs;
   for (int p0 = 0; p0 < nPairs; p0 += o.pairsPerJob)
   {
      const int np = std::min(o.pairsPerJob, nPairs - p0);
      const int32_t *pp = pairs.data() + 2 * (size_t)p0;
      check(v3dm_batch_set_pairs(ctx, np, pp), "v3dm_batch_set_pairs", ctx);
      int64_t lefts = 0;
      for (int p = 0; p < np; ++p)
      {
         lefts += views[pp[2 * p]].n;
         scored += (int64_t)views[pp[2 * p]].n * views[pp[2 * p + 1]].n;
      }
      table.resize((size_t)std::max<int64_t>(lefts, 1));
      offsets.resize((size_t)np + 1);
      const auto m0 = std::chrono::steady_clock::now();
      check(v3dm_batch_match(ctx, kScoreThreshold, o.distRatio, table.data(), offsets.data()), "v3dm_batch_match",
            ctx);
      matchMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0).count();

      for (int p = 0; p < np; ++p)
      {
         const int v1 = pp[2 * p], v2 = pp[2 * p + 1];
         const View &a = views[v1], &b = views[v2];
         corrs.clear();
         for (int64_t k = offsets[p]; k < offsets[p + 1]; ++k)
         {
            const int32_t i2 = table[(size_t)k];
            if (i2 < 0) continue;
            const int i1 = (int)(k - offsets[p]);
            v3dm_correspondence c;
            c.left = {a.pos[2 * (size_t)i1], a.pos[2 * (size_t)i1 + 1], v1, a.ids[(size_t)i1]};
            c.right = {b.pos[2 * (size_t)i2], b.pos[2 * (size_t)i2 + 1], v2, b.ids[(size_t)i2]};
            corrs.push_back(c);
         }
         totalCorrs += (int64_t)corrs.size();
         if (!o.quiet)
            std::cout << "--------------------\n"
                      << "matching view " << v1 << " with view " << v2 << "\n"
                      << "nFeatures1 = " << a.n << ", nFeatures2 = " << b.n << "\n"
                      << "matches.size() = " << corrs.size() << "\n";
         if ((int)corrs.size() < o.nMinRawCorrs) continue;
         if (db)
            check(v3dm_match_db_put(db, stored, v1, v2, R, T, (int)corrs.size(), corrs.data()),
                  "writing the match database");
         ++stored;
      }
   }
   if (db) check(v3dm_match_db_close(db), "closing the match database");
   v3dm_destroy(ctx);

   const double uploadMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
   std::cout << "matched " << nPairs << " view pairs (" << scored << " descriptor pairs, "
             << (useU8 ? "uint8" : "float") << " pool) in " << matchMs << " ms after a " << uploadMs
             << " ms upload; " << totalCorrs << " putative matches; " << stored << " pairs with at least "
             << o.nMinRawCorrs << " matches" << (db ? " written to " + o.dbPath : std::string()) << std::endl;
   return stored;
}

} // namespace v3dm_driver
