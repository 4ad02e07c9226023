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
// Output: without an orientation source, the raw correspondences the
// reference hands to its robust orientation step (:205-227,
// computeRobustOrientationMLE) for every pair with at least
// MINIMUM_RAW_CORRESPONDENCES of them (:229), written as PairwiseMatch rows
// (rotation identity, translation zero) with pixel positions to
// putative_matches.db.  With one (--orientation LIB: a plugin exporting
// v3dm_orientation_plugin; --poses FILE: known relative poses), the
// reference's geometric stage follows (:229-346): robust orientation, the
// inlier checks (MINIMUM_INLIER_CORRESPONDENCES, MINIMUM_INLIER_MATCHES_RATIO),
// guided matching with F = K2^-T [T]x R K1^-1 (USE_GUIDED_MATCHING; all
// accepted pairs in one batched job), the effective-inlier check
// (USE_EFFECTIVE_INLIERS), and rows with the pose and normalised positions
// in pairwise_matches.db, the reference's file.  Intrinsics and image sizes
// come from calibration_db.txt (--calibration).  Tables matches_data /
// matches_list, OIDs 0, 1, ... in pair order (:340-346).  The per-pair log
// lines follow the reference's.
//
// Options: --db PATH, --no-db, --orientation LIB, --poses FILE,
//          --calibration PATH, --device N, --pairs-per-job N (default 4096),
//          --autotune (time the scan tile variants on the first job instead
//          of 256 x 1024 tiles), --quiet.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "driver_common.h"

using namespace v3dm_driver;

namespace {

void usage(const char *argv0)
{
   std::cerr << "Usage: " << argv0 << " [options] <image list file> <file.conf> <radius>\n"
             << "       " << argv0 << " [options] --pairs <pairs file> <image list file> <file.conf>\n"
             << "options: --db PATH | --no-db | --orientation LIB | --poses FILE | --calibration PATH |\n"
             << "         --device N | --pairs-per-job N | --autotune | --quiet" << std::endl;
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
   MatchOptions o;
   std::string orientationLib, posesFile;
   bool dbGiven = false;
   std::string pairsFile;
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
      if (a == "--db") o.dbPath = value(), dbGiven = true;
      else if (a == "--no-db") o.dbPath.clear(), dbGiven = true;
      else if (a == "--orientation") orientationLib = value();
      else if (a == "--poses") posesFile = value();
      else if (a == "--calibration") o.geo.calibPath = value();
      else if (a == "--device") o.device = std::atoi(value());
      else if (a == "--pairs-per-job") o.pairsPerJob = std::atoi(value());
      else if (a == "--pairs") pairsFile = value();
      else if (a == "--quiet") o.quiet = true;
      else if (a == "--autotune") o.variant = -1;
      else if (a.size() > 1 && a[0] == '-' && a[1] == '-')
      {
         usage(argv[0]);
         return 2;
      }
      else pos.push_back(argv[i]);
   }
   if (pos.size() != (pairsFile.empty() ? 3u : 2u) || o.pairsPerJob < 1)
   {
      usage(argv[0]);
      return 2;
   }

   const Config cf(pos[1]);
   o.distRatio = cf.get("MATCHING_SCORE_RATIO", 0.8f);
   o.nMinRawCorrs = cf.get("MINIMUM_RAW_CORRESPONDENCES", 100);
   // the geometric stage: robust orientation from a plugin or from known poses
   std::unique_ptr<KnownPoses> known;
   if (!orientationLib.empty() && !posesFile.empty()) die("--orientation and --poses exclude each other");
   if (!orientationLib.empty()) loadOrientationPlugin(orientationLib, o.geo.hook);
   if (!posesFile.empty())
   {
      known.reset(new KnownPoses(posesFile));
      o.geo.hook = known->hook();
   }
   o.geo.enabled = o.geo.hook.robust_orientation != nullptr;
   o.geo.visibilityTest = false;
   readGeometryConfig(cf, o.nMinRawCorrs, true, o.geo);
   if (!dbGiven) o.dbPath = o.geo.enabled ? "pairwise_matches.db" : "putative_matches.db";
   const int nMaxFeatures = cf.get("MATCHING_MAX_FEATURES", 7168);

   std::vector<View> views = readImageList(pos[0]);
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
   for (int32_t v : pairs) views[v].used = true;

   // every view a pair names, once (the reference keeps an LRU cache of 500 views)
   for (View &v : views)
      if (v.used) loadView(v, nMaxFeatures, o.quiet);
   matchPairs(views, pairs, o);
   return 0;
}
