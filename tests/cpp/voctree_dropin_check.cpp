// Drives V3D::VocabularyTree<3, 50, 128> and V3D::LeafDocuments through
// include/Base/v3d_voctree.h the way SfM/match_using_voctree does
// (match_using_voctree.cpp:145-207): loadTreeFromFile, one getVisualWord per
// feature, then per view computeBinomialScores against the earlier views and
// insertDocument.  Also quantises the same keys with the batched GPU call.
// Built with fused multiply-adds available and contraction allowed
// (-march=x86-64-v3 -ffp-contract=fast): the header must still round every
// product before its add.
//
// Usage: voctree_dropin_check <tree file> <keys file> <n keys> <keys per view>
//        <pMatch> <nBest> <host words out> <gpu words out> <candidates out>
// (raw float32 nodes and keys; int32 words; candidates as "view document score"
// lines in score order per view, then L1 scores as "L1 view document score").
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include "Base/v3d_voctree.h"

int main(int argc, char **argv)
{
   if (argc != 10) return 2;
   typedef V3D::VocabularyTree<3, 50, 128> VocTree;
   VocTree tree;
   if (!tree.loadTreeFromFile(argv[1])) return 3;
   const int n = std::atoi(argv[3]), perView = std::atoi(argv[4]), nBest = std::atoi(argv[6]);
   const float pMatch = (float)std::atof(argv[5]); // the driver holds it in a float
   std::vector<float> keys((size_t)n * 128);
   FILE *f = std::fopen(argv[2], "rb");
   if (!f || std::fread(keys.data(), sizeof(float), keys.size(), f) != keys.size()) return 4;
   std::fclose(f);

   std::vector<int> host((size_t)n), gpu((size_t)n, -7);
   for (int i = 0; i < n; ++i) host[(size_t)i] = tree.getVisualWord(&keys[(size_t)i * 128]);
   try
   {
      VocTree copy(tree); // a copy quantises like the original
      copy.getVisualWords(n, keys.data(), gpu.data());
   }
   catch (std::exception const &e)
   {
      std::fprintf(stderr, "%s\n", e.what());
      return 5;
   }
   f = std::fopen(argv[7], "wb");
   if (!f || std::fwrite(host.data(), sizeof(int), host.size(), f) != host.size()) return 6;
   std::fclose(f);
   f = std::fopen(argv[8], "wb");
   if (!f || std::fwrite(gpu.data(), sizeof(int), gpu.size(), f) != gpu.size()) return 6;
   std::fclose(f);

   FILE *out = std::fopen(argv[9], "w");
   if (!out) return 7;
   V3D::LeafDocuments invertedFiles(tree.visualWordCount());
   std::vector<std::pair<float, int>> scores;
   for (int v = 0; v * perView < n; ++v)
   {
      const int b = v * perView, e = std::min(n, b + perView);
      std::vector<int> visualWords(host.begin() + b, host.begin() + e);
      // one score per earlier document; only the first nBest are ordered
      invertedFiles.computeBinomialScores(pMatch, visualWords, scores, nBest);
      if ((int)scores.size() != v) return 10;
      for (int k = 0; k < std::min(v, nBest); ++k)
         std::fprintf(out, "%d %d %.9g\n", v, scores[k].second, scores[k].first);
      invertedFiles.compute_L1_Scores(visualWords, scores, nBest);
      if ((int)scores.size() != v) return 10;
      for (int k = 0; k < std::min(v, nBest); ++k)
         std::fprintf(out, "L1 %d %d %.9g\n", v, scores[k].second, scores[k].first);
      if (invertedFiles.insertDocument(visualWords) != v) return 8;
   }
   std::fclose(out);
   try
   {
      std::vector<int> bad(3, tree.visualWordCount()); // a word outside the tree
      invertedFiles.insertDocument(bad);
      return 9;
   }
   catch (std::runtime_error const &)
   {
   }
   return 0;
}
