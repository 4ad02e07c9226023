// Drives V3D_CUDA::SIFT_Matching through include/CUDA/v3d_cudamatching.h the
// way SfM/match_sequence does (float descriptors, findPutativeMatches, then
// runGuidedMatching with a fundamental matrix) and prints the match lists.
// Usage: wrapper_check nLeft nRight seed
// Output: "find <count>" + "k l" lines, then "guided <count>" + lines, then
// "homography <count>" + lines.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "CUDA/v3d_cudamatching.h"

namespace V3D
{
// Minimal stand-ins with the indexing interface of Math/v3d_linear.h types.
struct Matrix3x3d
{
   double m[3][3];
   double const *operator[](int i) const { return m[i]; }
   double *operator[](int i) { return m[i]; }
};
struct Vector2f
{
   float v[2];
   float operator[](int i) const { return v[i]; }
   float &operator[](int i) { return v[i]; }
};
} // namespace V3D

static void print(char const *tag, std::vector<std::pair<int, int> > const &m)
{
   std::printf("%s %d\n", tag, (int)m.size());
   for (size_t i = 0; i < m.size(); ++i) std::printf("%d %d\n", m[i].first, m[i].second);
}

int main(int argc, char **argv)
{
   if (argc < 4) return 2;
   int const nL = std::atoi(argv[1]), nR = std::atoi(argv[2]);
   unsigned long long const seed = std::strtoull(argv[3], 0, 10);
   std::vector<unsigned char> L(128 * (size_t)nL), R(128 * (size_t)nR);
   std::vector<int> src(nR);
   if (v3dm_synth_pair(seed, nL, nR, 0.4, 0.05, 0, L.data(), R.data(), 0, 0, src.data()) != V3DM_OK) return 3;
   std::vector<float> Lf(L.size()), Rf(R.size());
   for (size_t i = 0; i < L.size(); ++i) Lf[i] = L[i] / 512.0f;
   for (size_t i = 0; i < R.size(); ++i) Rf[i] = R[i] / 512.0f;

   // positions: left uniform, right copies on the same row + 30 px in x
   std::vector<V3D::Vector2f> lp(nL), rp(nR);
   unsigned long long s = seed * 2654435761ull + 1;
   for (int k = 0; k < nL; ++k)
      for (int c = 0; c < 2; ++c)
      {
         s = s * 6364136223846793005ull + 1442695040888963407ull;
         lp[k][c] = (float)((s >> 33) % 100000) / 100.0f;
      }
   for (int k = 0; k < nR; ++k)
   {
      for (int c = 0; c < 2; ++c)
      {
         s = s * 6364136223846793005ull + 1442695040888963407ull;
         rp[k][c] = (float)((s >> 33) % 100000) / 100.0f;
      }
      if (src[k] >= 0)
      {
         rp[k][0] = lp[src[k]][0] + 30.0f;
         rp[k][1] = lp[src[k]][1] + (float)((int)(k % 5) - 2);
      }
   }
   V3D::Matrix3x3d F = {{{0, 0, 0}, {0, 0, -1}, {0, 1, 0}}};
   V3D::Matrix3x3d H = {{{1, 0, 30}, {0, 1, 0}, {0, 0, 1}}};

   try
   {
      V3D_CUDA::SIFT_Matching matcher;
      matcher.allocate(nL > nR ? nL : nR, true);
      matcher.enableReductionPath(false);
      matcher.setLeftFeatures(nL, Lf.data());
      matcher.setRightFeatures(nR, Rf.data());
      std::vector<std::pair<int, int> > matches;
      matcher.findPutativeMatches(matches, 0.8f, 0.8f);
      print("find", matches);
      matcher.runGuidedMatching(F, 1.2f, &lp[0], &rp[0], matches, 0.8f, 0.8f);
      print("guided", matches);
      matcher.runGuidedHomographyMatching(H, 2.0f, &lp[0], &rp[0], matches, 0.8f, 0.8f);
      print("homography", matches);
      matcher.deallocate();
   }
   catch (std::exception const &e)
   {
      std::fprintf(stderr, "exception: %s\n", e.what());
      return 1;
   }
   return 0;
}
