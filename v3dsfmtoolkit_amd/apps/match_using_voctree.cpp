// match_using_voctree.cpp -- the putative-matching stages of
// SfM/match_using_voctree on the MI355X engine:
//
//   v3dm_match_using_voctree [options] <vocabulary tree file> <image list file> <file.conf>
//
// Stage 1 (match_using_voctree.cpp:147-207): every feature of every view is
// quantised by the vocabulary tree (VocabularyTree<3, 50, 128>, :145-148) --
// here on the GPU, many views per job, bit-identical words -- and each view
// i is scored against views 0 .. i-1 through the inverted file
// (computeBinomialScores with VISUAL_WORD_MATCH_PROPABILITY, held in a float
// as the driver holds it, :137); its POTENTIAL_VIEWS_COUNT best views become
// candidate pairs (earlier view, i).  The candidate graph is written to
// putative_EG.dot (:208-219) and the per-view candidate lists are printed as
// the reference prints them (:185-190).
//
// Stage 2 (:251-323): the candidate pairs are matched as batched jobs with
// the reference's rule (minScore 0.9, MATCHING_SCORE_RATIO) and the pairs
// with at least MINIMUM_RAW_CORRESPONDENCES raw correspondences are written
// to the match database as in v3dm_match_sequence.  With an orientation
// source (--orientation LIB, --poses FILE) the geometric stage of :325-491
// follows as in v3dm_match_sequence, plus the visibility test of :397-418
// (USE_EFFECTIVE_INLIERS defaults to false here, :122); its panorama and
// homography databases (:493-560) are not produced.
//
// Options: --db PATH (default putative_matches.db, pairwise_matches.db with
//          an orientation source), --no-db, --dot PATH, --orientation LIB,
//          --poses FILE, --calibration PATH, --tree-shape DEPTH BRANCH
//          (default 3 50), --device N, --pairs-per-job N (default 4096),
//          --keys-per-job N (default 4194304), --autotune (time the scan tile
//          variants on the first job), --quiet.
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "driver_common.h"
#include "v3dm_voctree.h"

using namespace v3dm_driver;

namespace {

void usage(const char *argv0)
{
   std::cerr << "Usage: " << argv0 << " [options] <vocabulary tree file> <image list file> <file.conf>\n"
             << "options: --db PATH | --no-db | --dot PATH | --orientation LIB | --poses FILE | --calibration PATH |\n"
             << "         --tree-shape DEPTH BRANCH | --device N |\n"
             << "         --pairs-per-job N | --keys-per-job N | --autotune | --quiet" << std::endl;
}

} // namespace

int main(int argc, char **argv)
{
   MatchOptions o;
   std::string orientationLib, posesFile;
   bool dbGiven = false;
   std::string dotPath = "putative_EG.dot";
   int depth = 3, branch = 50;
   int64_t keysPerJob = 4194304;
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
      else if (a == "--dot") dotPath = value();
      else if (a == "--tree-shape")
      {
         depth = std::atoi(value());
         branch = std::atoi(value());
      }
      else if (a == "--device") o.device = std::atoi(value());
      else if (a == "--pairs-per-job") o.pairsPerJob = std::atoi(value());
      else if (a == "--keys-per-job") keysPerJob = std::atoll(value());
      else if (a == "--quiet") o.quiet = true;
      else if (a == "--autotune") o.variant = -1;
      else if (a.size() > 1 && a[0] == '-' && a[1] == '-')
      {
         usage(argv[0]);
         return 2;
      }
      else pos.push_back(argv[i]);
   }
   if (pos.size() != 3 || o.pairsPerJob < 1 || keysPerJob < 1)
   {
      usage(argv[0]);
      return 2;
   }

   const Config cf(pos[2]);
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
   o.geo.visibilityTest = true;
   readGeometryConfig(cf, o.nMinRawCorrs, false, o.geo);
   if (!dbGiven) o.dbPath = o.geo.enabled ? "pairwise_matches.db" : "putative_matches.db";
   const int nMaxFeatures = cf.get("MATCHING_MAX_FEATURES", 7168);
   const int nBestDocuments = cf.get("POTENTIAL_VIEWS_COUNT", 20);
   const float pMatch = (float)cf.get("VISUAL_WORD_MATCH_PROPABILITY", 0.1);

   v3dm_voctree *tree = nullptr;
   if (v3dm_voctree_load(o.device, depth, branch, pos[0], &tree) != V3DM_OK)
      die(std::string("loading the vocabulary tree: ") + v3dm_voctree_last_error(nullptr));
   int nLeaves = 0;
   v3dm_voctree_info(tree, nullptr, nullptr, nullptr, &nLeaves);
   if (!o.quiet) std::cout << "Loaded vocabulary tree from file." << std::endl;

   std::vector<View> views = readImageList(pos[1]);
   const int nViews = (int)views.size();
   for (View &v : views)
   {
      v.used = true;
      loadView(v, nMaxFeatures, o.quiet);
   }

   // stage 1: visual words of every feature, many views per job
   std::vector<int32_t> counts(nViews), words;
   int64_t total = 0;
   for (int i = 0; i < nViews; ++i)
   {
      counts[i] = views[i].count;
      total += views[i].count;
   }
   words.resize((size_t)std::max<int64_t>(total, 1));
   double quantMs = 0;
   {
      const auto q0 = std::chrono::steady_clock::now();
      std::vector<float> keys;
      int64_t done = 0;
      for (int i0 = 0; i0 < nViews;)
      {
         int i1 = i0;
         int64_t n = 0;
         keys.clear();
         int64_t jobKeys = 0;
         for (int i = i0; i < nViews && (i == i0 || jobKeys + views[i].count <= keysPerJob); ++i)
            jobKeys += views[i].count;
         keys.reserve((size_t)jobKeys * kDim); // one allocation for the job's keys
         while (i1 < nViews && (i1 == i0 || n + views[i1].count <= keysPerJob))
         {
            keys.insert(keys.end(), views[i1].desc.begin(), views[i1].desc.end());
            n += views[i1].count;
            ++i1;
         }
         if (n > 0)
            if (v3dm_voctree_quantize(tree, n, keys.data(), words.data() + done) != V3DM_OK)
               die(std::string("quantising: ") + v3dm_voctree_last_error(tree));
         done += n;
         i0 = i1;
      }
      quantMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - q0).count();
   }
   v3dm_voctree_destroy(tree);
   for (int64_t k = 0; k < total; ++k)
      if (words[(size_t)k] < 0) die("a descriptor with a non-finite value cannot be quantised");

   int nPairs = 0;
   int rc = v3dm_voctree_candidate_pairs(nViews, counts.data(), words.data(), nLeaves, nBestDocuments, (double)pMatch,
                                         nullptr, nullptr, 0, &nPairs);
   if (rc != V3DM_OK && rc != V3DM_ERR_CAPACITY) check(rc, "candidate pairs");
   std::vector<int32_t> pairs(2 * (size_t)nPairs);
   std::vector<float> scores((size_t)nPairs);
   check(v3dm_voctree_candidate_pairs(nViews, counts.data(), words.data(), nLeaves, nBestDocuments, (double)pMatch,
                                      pairs.data(), scores.data(), nPairs, &nPairs),
         "candidate pairs");

   if (!o.quiet)
   {
      // per view, its candidates in score order (then view order), as :185-190 prints them
      std::vector<std::vector<std::pair<float, int>>> byView((size_t)nViews);
      for (int p = 0; p < nPairs; ++p) byView[(size_t)pairs[2 * (size_t)p + 1]].push_back({scores[(size_t)p], pairs[2 * (size_t)p]});
      for (int i = 0; i < nViews; ++i)
      {
         std::sort(byView[(size_t)i].begin(), byView[(size_t)i].end());
         std::cout << "Potential matches for " << views[(size_t)i].name << ":" << std::endl;
         for (const auto &[s, j] : byView[(size_t)i])
            std::cout << " " << views[(size_t)j].name << " with score " << s << std::endl;
      }
   }
   {
      std::ofstream os(dotPath);
      os << "graph {" << std::endl;
      for (int p = 0; p < nPairs; ++p) os << pairs[2 * (size_t)p] << " -- " << pairs[2 * (size_t)p + 1] << ";" << std::endl;
      os << "}" << std::endl;
      if (!os) die("cannot write " + dotPath);
   }
   std::cout << "quantised " << total << " features of " << nViews << " views in " << quantMs << " ms; " << nPairs
             << " candidate pairs" << std::endl;

   // stage 2
   matchPairs(views, pairs, o);
   return 0;
}
