// match_sequence.cpp -- the putative-matching stage of the reference's
// matching drivers on the MI355X engine:
//
//   v3dm_match_sequence [options] <image list file> <file.conf> <radius>
//       SfM/match_sequence/src/match_sequence.cpp:75-227: view i against views
//       i+1 .. i+2*radius-1.
//   v3dm_match_sequence [options] --pairs <pairs file> <image list file> <file.conf>
//       the matching loop of SfM/match_using_voctree/src/match_using_voctree.cpp:251-323
//       over an explicit candidate list ("a b" per line, or the "a -- b;" lines
//       of the putative_EG.dot file that driver writes, :208-219).
//
// Per view the driver reads <entry>.features (V3DBIN or V3DTXT), keeps
// min(MATCHING_MAX_FEATURES, size & 0xffe0) features (match_sequence.cpp:152,184)
// and matches every pair with the reference's rule at minScore 0.9 and
// minRatio MATCHING_SCORE_RATIO (:79,89,203).  Unlike the reference, which runs
// one setRightFeatures + findPutativeMatches per pair, all views are uploaded
// once and all pairs run as batched jobs.  Uint8-representable descriptors
// (u/512) go to the exact int8 pool, others to the BF16 + FP32-rescoring pool;
// each pair's list equals findPutativeMatches on that pair.
//
// Output: the raw correspondences the reference hands to its robust
// orientation step (:205-227, computeRobustOrientationMLE) for every pair with
// at least MINIMUM_RAW_CORRESPONDENCES of them (:229), written as PairwiseMatch
// rows (rotation identity, translation zero: that step fills them) with
// pixel positions into the tables matches_data / matches_list of an SQLite file
// (default putative_matches.db), OIDs 0, 1, ... in pair order
// (:340-346).  The per-pair log lines follow the reference's.
//
// Options: --db PATH (default putative_matches.db), --no-db, --device N,
//          --pairs-per-job N (default 4096), --quiet.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "v3dm_io.h"
#include "v3dmatch.h"

namespace {

constexpr int kDim = 128;

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
   int n = 0; // features kept
   std::vector<int32_t> ids;
   std::vector<float> pos; // 2 per feature
   std::vector<float> desc;
};

[[noreturn]] void die(const std::string &msg)
{
   std::cerr << "Exception caught: " << msg << std::endl;
   std::exit(1);
}

void check(int rc, const char *what, v3dm_ctx *ctx = nullptr)
{
   if (rc != V3DM_OK) die(std::string(what) + ": " + v3dm_last_error(ctx));
}

void usage(const char *argv0)
{
   std::cerr << "Usage: " << argv0 << " [options] <image list file> <file.conf> <radius>\n"
             << "       " << argv0 << " [options] --pairs <pairs file> <image list file> <file.conf>\n"
             << "options: --db PATH | --no-db | --device N | --pairs-per-job N | --quiet" << std::endl;
}

// "a b" or "a -- b;" per line; other lines ("graph {", "}") are skipped.
std::vector<int32_t> readPairs(const char *path, int nViews)
{
   std::ifstream is(path);
   if (!is) die(std::string("cannot open pairs file ") + path);
   std::vector<int32_t> pairs;
   std::string line;
   while (std::getline(is, line))
   {
      for (char &ch : line)
         if (ch == '-' || ch == ';') ch = ' ';
      std::istringstream ss(line);
      long a, b;
      if (!(ss >> a >> b)) continue;
      if (a < 0 || b < 0 || a >= nViews || b >= nViews)
         die("pair (" + std::to_string(a) + ", " + std::to_string(b) + ") names a view outside the image list");
      pairs.push_back((int32_t)a);
      pairs.push_back((int32_t)b);
   }
   return pairs;
}

} // namespace

int main(int argc, char **argv)
{
   std::string dbPath = "putative_matches.db", pairsFile;
   bool writeDb = true, quiet = false;
   int device = 0, pairsPerJob = 4096;
   std::vector<const char *> pos;
   for (int i = 1; i < argc; ++i)
   {
      const std::string a = argv[i];
      auto value = [&]() -> const char * {
         if (i + 1 >= argc)
         {
            usage(argv[0]);
            std::exit(2);
         }
         return argv[++i];
      };
      if (a == "--db") dbPath = value();
      else if (a == "--no-db") writeDb = false;
      else if (a == "--device") device = std::atoi(value());
      else if (a == "--pairs-per-job") pairsPerJob = std::atoi(value());
      else if (a == "--pairs") pairsFile = value();
      else if (a == "--quiet") quiet = true;
      else if (a.size() > 1 && a[0] == '-' && a[1] == '-')
      {
         usage(argv[0]);
         return 2;
      }
      else pos.push_back(argv[i]);
   }
   if (pos.size() != (pairsFile.empty() ? 3u : 2u) || pairsPerJob < 1)
   {
      usage(argv[0]);
      return 2;
   }

   const Config cf(pos[1]);
   const float scoreThreshold = 0.9f;
   const float distRatio = cf.get("MATCHING_SCORE_RATIO", 0.8f);
   const int nMinRawCorrs = cf.get("MINIMUM_RAW_CORRESPONDENCES", 100);
   const int nMaxFeatures = cf.get("MATCHING_MAX_FEATURES", 7168);

   std::vector<View> views;
   {
      std::ifstream is(pos[0]);
      if (!is) die(std::string("cannot open image list ") + pos[0]);
      std::string name;
      while (is >> name)
      {
         views.emplace_back();
         views.back().name = name;
      }
   }
   const int nViews = (int)views.size();

   std::vector<int32_t> pairs;
   if (pairsFile.empty())
   {
      const int radius = std::atoi(pos[2]);
      int count = 0;
      v3dm_sequence_pairs(nViews, radius, nullptr, 0, &count);
      pairs.resize(2 * (size_t)count);
      check(v3dm_sequence_pairs(nViews, radius, pairs.data(), count, &count), "v3dm_sequence_pairs");
   }
   else
      pairs = readPairs(pairsFile.c_str(), nViews);
   const int nPairs = (int)(pairs.size() / 2);
   for (int32_t v : pairs) views[v].used = true;

   // Load every view a pair names, once (the reference keeps an LRU cache of 500).
   bool allU8 = true;
   int64_t rows = 0;
   for (View &v : views)
   {
      if (!v.used) continue;
      const std::string path = v.name + ".features";
      if (!quiet) std::cout << "reading " << path << std::endl;
      int count = 0, format = 0;
      check(v3dm_features_info(path.c_str(), &count, &format), "reading features");
      std::vector<int32_t> ids(count);
      std::vector<float> xy(2 * (size_t)count), desc((size_t)count * kDim);
      check(v3dm_features_read(path.c_str(), count, ids.data(), xy.data(), nullptr, nullptr, nullptr, desc.data(),
                               &count),
            "reading features");
      // min(nMax, size & 0xffe0), then setLeft/RightFeatures' rounding to 32 for an odd nMax
      v.n = v3dm_truncated_feature_count(count, nMaxFeatures) & ~31;
      v.ids.assign(ids.begin(), ids.begin() + v.n);
      v.pos.assign(xy.begin(), xy.begin() + 2 * (size_t)v.n);
      v.desc.assign(desc.begin(), desc.begin() + (size_t)v.n * kDim);
      rows += v.n;
   }
   std::vector<int32_t> counts(nViews);
   std::vector<float> all;
   all.reserve((size_t)rows * kDim);
   for (int i = 0; i < nViews; ++i)
   {
      counts[i] = views[i].n;
      all.insert(all.end(), views[i].desc.begin(), views[i].desc.end());
      std::vector<float>().swap(views[i].desc);
   }
   std::vector<uint8_t> allU8Data(all.size());
   allU8 = v3dm_descriptors_to_u8(rows, all.data(), allU8Data.data()) == V3DM_OK;

   v3dm_ctx *ctx = nullptr;
   check(v3dm_create(device, &ctx), "v3dm_create");
   const auto t0 = std::chrono::steady_clock::now();
   if (allU8)
      check(v3dm_batch_set_images_u8(ctx, nViews, counts.data(), allU8Data.data()), "upload", ctx);
   else
      check(v3dm_batch_set_images_f32(ctx, nViews, counts.data(), all.data()), "upload", ctx);
   std::vector<float>().swap(all);
   std::vector<uint8_t>().swap(allU8Data);
   const auto t1 = std::chrono::steady_clock::now();

   v3dm_match_db *db = nullptr;
   if (writeDb) check(v3dm_match_db_open(dbPath.c_str(), 1, &db), "opening the match database");
   const double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, T[3] = {0, 0, 0};

   double matchMs = 0;
   int64_t scored = 0, totalCorrs = 0;
   int stored = 0;
   std::vector<int32_t> table;
   std::vector<int64_t> offsets;
   std::vector<v3dm_correspondence> corrs;
   for (int p0 = 0; p0 < nPairs; p0 += pairsPerJob)
   {
      const int np = std::min(pairsPerJob, nPairs - p0);
      check(v3dm_batch_set_pairs(ctx, np, pairs.data() + 2 * (size_t)p0), "v3dm_batch_set_pairs", ctx);
      int64_t lefts = 0;
      for (int p = 0; p < np; ++p)
      {
         lefts += views[pairs[2 * (size_t)(p0 + p)]].n;
         scored += (int64_t)views[pairs[2 * (size_t)(p0 + p)]].n * views[pairs[2 * (size_t)(p0 + p) + 1]].n;
      }
      table.resize((size_t)lefts);
      offsets.resize((size_t)np + 1);
      const auto m0 = std::chrono::steady_clock::now();
      check(v3dm_batch_match(ctx, scoreThreshold, distRatio, table.data(), offsets.data()), "v3dm_batch_match", ctx);
      matchMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - m0).count();

      for (int p = 0; p < np; ++p)
      {
         const int v1 = pairs[2 * (size_t)(p0 + p)], v2 = pairs[2 * (size_t)(p0 + p) + 1];
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
         if (!quiet)
         {
            std::cout << "--------------------\n"
                      << "matching view " << v1 << " with view " << v2 << "\n"
                      << "nFeatures1 = " << a.n << ", nFeatures2 = " << b.n << "\n"
                      << "matches.size() = " << corrs.size() << "\n";
         }
         if ((int)corrs.size() < nMinRawCorrs) continue;
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
             << (allU8 ? "uint8" : "float") << " pool) in " << matchMs << " ms after a " << uploadMs
             << " ms upload; " << totalCorrs << " putative matches; " << stored << " pairs with at least "
             << nMinRawCorrs << " matches" << (db ? " written to " + dbPath : std::string()) << std::endl;
   return 0;
}
