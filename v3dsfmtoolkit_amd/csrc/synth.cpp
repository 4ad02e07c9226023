// synth.cpp -- deterministic SIFT-like descriptor generator (host code).
//
// Workload generator for matching benchmarks and tests.  One descriptor:
//   z_k ~ N(0,1), k < 128;  v_k = |z_k|^1.5;  v /= |v|_2;  v = min(v, 0.2);
//   v /= |v|_2;  u_k = min(255, round(512 v_k))  (uint8), or v as float.
// A noisy copy of a left descriptor v_L:
//   w = v_L + N(0, sigma);  w = max(w, 0);  w /= |w|;  min(w, 0.2);  w /= |w|.
// A pair: nL independent left descriptors; the right set holds
// min(floor(rho*nR), nL) noisy copies of distinct random left rows at
// uniformly permuted positions, the rest independent draws.
//
// Every descriptor draws from its own SplitMix64 stream keyed by
// (seed, role, index), so generation is parallel and reproducible.
// Normals come from the Box-Muller transform (both outputs used).
// SIFT's 0.2 clamp and 512 quantisation scale follow Lowe's descriptor
// normalisation; the reference receives the float form u/512.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/v3dmatch.h"

namespace {

constexpr int D = 128;

enum Role : uint64_t { ROLE_LEFT = 1, ROLE_RIGHT = 2, ROLE_NOISE = 3, ROLE_PERM = 4 };

inline uint64_t splitmix64(uint64_t &s)
{
   uint64_t z = (s += 0x9E3779B97F4A7C15ull);
   z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
   z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
   return z ^ (z >> 31);
}

inline uint64_t stream_seed(uint64_t seed, uint64_t role, uint64_t index)
{
   uint64_t s = seed ^ (role << 56) ^ (0xD1B54A32D192ED03ull * (index + 1));
   splitmix64(s);
   return s;
}

// uniform in (0, 1)
inline double u01(uint64_t &s) { return ((double)(splitmix64(s) >> 11) + 0.5) * 0x1.0p-53; }

void normals(uint64_t &s, double *z, int n)
{
   const double two_pi = 6.283185307179586476925286766559;
   for (int k = 0; k < n; k += 2)
   {
      double u1 = u01(s), u2 = u01(s);
      double r = std::sqrt(-2.0 * std::log(u1));
      z[k] = r * std::cos(two_pi * u2);
      if (k + 1 < n) z[k + 1] = r * std::sin(two_pi * u2);
   }
}

void normalize(double *v)
{
   double n2 = 0.0;
   for (int k = 0; k < D; ++k) n2 += v[k] * v[k];
   if (n2 <= 0.0) return;
   double inv = 1.0 / std::sqrt(n2);
   for (int k = 0; k < D; ++k) v[k] *= inv;
}

void sift_normalize(double *v)
{
   normalize(v);
   for (int k = 0; k < D; ++k) v[k] = std::min(v[k], 0.2);
   normalize(v);
}

void base_descriptor(uint64_t seed, uint64_t role, uint64_t index, double *v)
{
   uint64_t s = stream_seed(seed, role, index);
   double z[D];
   normals(s, z, D);
   for (int k = 0; k < D; ++k)
   {
      double a = std::fabs(z[k]);
      v[k] = a * std::sqrt(a); // |z|^1.5
   }
   sift_normalize(v);
}

void noisy_copy(const double *vl, double sigma, uint64_t seed, uint64_t index, double *w)
{
   uint64_t s = stream_seed(seed, ROLE_NOISE, index);
   double z[D];
   normals(s, z, D);
   for (int k = 0; k < D; ++k) w[k] = std::max(vl[k] + sigma * z[k], 0.0);
   sift_normalize(w);
}

void emit(const double *v, uint8_t *u8, float *f32)
{
   if (u8)
      for (int k = 0; k < D; ++k)
      {
         double q = std::floor(512.0 * v[k] + 0.5);
         u8[k] = (uint8_t)std::min(255.0, std::max(0.0, q));
      }
   if (f32)
      for (int k = 0; k < D; ++k) f32[k] = (float)v[k];
}

template <class F> void parallel_for(int n, int nThreads, F &&f)
{
   if (nThreads <= 0) nThreads = (int)std::max(1u, std::thread::hardware_concurrency());
   nThreads = std::max(1, std::min(nThreads, (n + 255) / 256));
   if (nThreads == 1)
   {
      for (int i = 0; i < n; ++i) f(i);
      return;
   }
   std::vector<std::thread> pool;
   int chunk = (n + nThreads - 1) / nThreads;
   for (int t = 0; t < nThreads; ++t)
   {
      int i0 = t * chunk, i1 = std::min(n, i0 + chunk);
      if (i0 >= i1) break;
      pool.emplace_back([i0, i1, &f] {
         for (int i = i0; i < i1; ++i) f(i);
      });
   }
   for (auto &th : pool) th.join();
}

} // namespace

extern "C" int v3dm_synth_pair(uint64_t seed, int nL, int nR, double rho, double sigma, int nThreads,
                               uint8_t *leftU8, uint8_t *rightU8, float *leftF32, float *rightF32,
                               int32_t *rightSource)
{
   if (nL < 0 || nR < 0 || !(rho >= 0.0 && rho <= 1.0) || !(sigma >= 0.0)) return V3DM_ERR_ARG;

   parallel_for(nL, nThreads, [&](int i) {
      double v[D];
      base_descriptor(seed, ROLE_LEFT, (uint64_t)i, v);
      emit(v, leftU8 ? leftU8 + (size_t)i * D : nullptr, leftF32 ? leftF32 + (size_t)i * D : nullptr);
   });

   int m = (int)std::floor(rho * (double)nR);
   m = std::min(m, nL);

   // m distinct left sources: partial Fisher-Yates over [0, nL)
   std::vector<int> src(nL > 0 ? nL : 1);
   for (int i = 0; i < nL; ++i) src[i] = i;
   {
      uint64_t s = stream_seed(seed, ROLE_PERM, 0);
      for (int t = 0; t < m; ++t)
      {
         int j = t + (int)(u01(s) * (double)(nL - t));
         std::swap(src[t], src[j]);
      }
   }
   // uniformly permuted right positions
   std::vector<int> pos(nR > 0 ? nR : 1);
   for (int i = 0; i < nR; ++i) pos[i] = i;
   {
      uint64_t s = stream_seed(seed, ROLE_PERM, 1);
      for (int t = nR - 1; t > 0; --t)
      {
         int j = (int)(u01(s) * (double)(t + 1));
         std::swap(pos[t], pos[j]);
      }
   }

   parallel_for(nR, nThreads, [&](int t) {
      double v[D];
      int r = pos[t];
      if (t < m)
      {
         double vl[D];
         base_descriptor(seed, ROLE_LEFT, (uint64_t)src[t], vl);
         noisy_copy(vl, sigma, seed, (uint64_t)t, v);
      }
      else
         base_descriptor(seed, ROLE_RIGHT, (uint64_t)r, v);
      emit(v, rightU8 ? rightU8 + (size_t)r * D : nullptr, rightF32 ? rightF32 + (size_t)r * D : nullptr);
      if (rightSource) rightSource[r] = t < m ? src[t] : -1;
   });
   return V3DM_OK;
}

extern "C" uint64_t v3dm_synth_seed(int config, int pairIndex)
{
   return 0x5EED0000ull + 256ull * (uint64_t)config + (uint64_t)pairIndex;
}
