// leafdocs_scenario.h -- one inverted-file scenario shared by the reference
// build (oracle/ref/leafdocs_ref.cpp, against the reference's
// V3D/include/Base/v3d_voctree.h) and the drop-in build
// (tests/cpp/leafdocs_dropin.cpp, against include/Base/v3d_voctree.h): the same
// documents are inserted, the inverted file is saved in both archive formats,
// and three queries print their document histograms and scores.  Test
// infrastructure only.
#include <cstdint>
#include <cstdio>
#include <utility>
#include <vector>

namespace leafdocs_scenario
{
constexpr int kLeaves = 400;

// Words of document d: a deterministic mix of a shared vocabulary block
// (so documents overlap), repeats, and document-specific words; document 5 is
// empty of words outside one leaf.
inline std::vector<int> document(int d)
{
   std::vector<int> w;
   uint32_t s = 0x9E3779B9u * (uint32_t)(d + 1);
   const int n = d == 5 ? 7 : 40 + 13 * d;
   for (int i = 0; i < n; ++i)
   {
      s = s * 1664525u + 1013904223u;
      if (d == 5) w.push_back(17);
      else if (i % 3 == 0) w.push_back((int)(s >> 8) % 60);           // shared block
      else if (i % 5 == 1 && !w.empty()) w.push_back(w[w.size() - 1]); // repeat
      else w.push_back((int)(s >> 8) % kLeaves);
   }
   return w;
}

constexpr int kDocuments = 12;

// Query q: 30 + 20 q words, half from the shared block.
inline std::vector<int> query(int q)
{
   std::vector<int> w;
   uint32_t s = 0x85EBCA6Bu * (uint32_t)(q + 7);
   for (int i = 0; i < 30 + 20 * q; ++i)
   {
      s = s * 1664525u + 1013904223u;
      w.push_back(i % 2 ? (int)(s >> 8) % 60 : (int)(s >> 8) % kLeaves);
   }
   return w;
}

template <class Docs> void print_queries(Docs &docs, FILE *out)
{
   for (int q = 0; q < 3; ++q)
   {
      const std::vector<int> words = query(q);
      std::vector<int> hist(2, 7); // non-empty on entry: resize keeps these
      docs.getDocumentHistogram(words, hist);
      std::fprintf(out, "query %d histogram", q);
      for (int h : hist) std::fprintf(out, " %d", h);
      std::fprintf(out, "\n");
      std::vector<std::pair<float, int> > scores;
      docs.computeBinomialScores(0.1, words, scores, 5);
      std::fprintf(out, "query %d binomial", q);
      for (auto const &s : scores) std::fprintf(out, " %d:%a", s.second, s.first);
      std::fprintf(out, "\n");
      docs.compute_L1_Scores(words, scores, -1);
      std::fprintf(out, "query %d l1", q);
      for (auto const &s : scores) std::fprintf(out, " %d:%a", s.second, s.first);
      std::fprintf(out, "\n");
   }
}
} // namespace leafdocs_scenario
