// orientation_plugin.cpp -- a plugin for the driver applications' --orientation
// option, used by the GPU driver tests: it exports v3dm_orientation_plugin and
// supplies robust orientation from the known poses in the file named by
// V3DM_TEST_POSES ("view1 view2 R (9) T (3)" per line).  Its inliers are the
// raw correspondences within the threshold's Sampson distance, handed back in
// descending order (the driver must keep the hook's order), and its pose
// refinement doubles T (an exact change that leaves F's mask unchanged up to
// scale, so the stored translation shows both refinement calls).
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "v3dm_io.h"

namespace {

std::map<std::pair<int, int>, std::vector<double>> g_poses;

int orientation(void *, int v1, int v2, const double K1[9], const double K2[9], double thr, int n,
                const v3dm_correspondence *corrs, double R[9], double T[3], int32_t *inliers, int *nInliers)
{
   auto it = g_poses.find({v1, v2});
   if (it == g_poses.end()) return 1;
   std::copy(it->second.begin(), it->second.begin() + 9, R);
   std::copy(it->second.begin() + 9, it->second.end(), T);
   double F[9];
   if (v3dm_fundamental_from_pose(K1, K2, R, T, F) != V3DM_OK) return 1;
   if (v3dm_sampson_inliers(F, thr, n, corrs, inliers, nInliers) != V3DM_OK) return 1;
   std::reverse(inliers, inliers + *nInliers);
   return 0;
}

int refine(void *, int, int, const double *, const double *, int, const v3dm_correspondence *, double *, double T[3])
{
   for (int k = 0; k < 3; ++k) T[k] *= 2;
   return 0;
}

} // namespace

extern "C" int v3dm_orientation_plugin(v3dm_orientation_hook *hook)
{
   const char *path = std::getenv("V3DM_TEST_POSES");
   if (!path) return 1;
   std::ifstream is(path);
   std::string line;
   while (std::getline(is, line))
   {
      std::istringstream ss(line);
      int a, b;
      std::vector<double> rt(12);
      if (!(ss >> a >> b)) continue;
      for (double &x : rt) ss >> x;
      g_poses[{a, b}] = rt;
   }
   hook->user = nullptr;
   hook->robust_orientation = &orientation;
   hook->refine_pose = &refine;
   return 0;
}
