// -*- C++ -*-
// Drop-in V3D_CUDA::SIFT_Matching for the MI355X engine.
//
// Same header path, class name, public members, argument meaning and
// defaults as the reference (V3D/include/CUDA/v3d_cudamatching.h:15-67), so
// SfM/match_sequence and SfM/match_using_voctree compile unchanged apart from
// dropping their cublasInit()/cublasShutdown() calls.  Every member forwards
// to the C ABI in v3dmatch.h.
//
// Differences a caller can observe:
//  * device failures throw std::runtime_error (a std::exception; both callers
//    catch it) instead of exiting the process (v3d_cudamatching.cuh:9-24);
//  * runGuided*Matching without allocate(..., true) throws instead of using
//    unallocated buffers;
//  * the guided members take the reference's V3D::Matrix3x3d / V3D::Vector2f
//    (exact signatures, when Math/v3d_linear.h is on the include path, as it is
//    in the reference's build) and, as templates, any type indexable as
//    F[i][j] and p[0], p[1];
//  * the protected _d_* members point at this engine's device buffers, which
//    differ from the reference's: see v3dm_get_device_buffers in v3dmatch.h
//    (scores exist only after the plain kernel; positions are interleaved);
//  * the destructor releases the device state (the reference leaves it to
//    deallocate()).
//
// V3D_CUDA::VocabularyTreeTraversal<BranchFactor, Depth> (reference
// v3d_cudamatching.h:69-105, v3d_cudamatching.cpp:258-328) is here too, on the
// library's vocabulary tree (v3dm_voctree.h).  Its leaves are those of the CPU
// tree the voctree driver uses (V3D/include/Base/v3d_voctree.h:63-99,
// bit-identical); the reference's own kernel sums each score in a different
// order (v3d_cudamatching.cu:435-528), so on near-ties its leaves can differ
// from the CPU tree's, and this class follows the CPU tree.
#ifndef V3D_CUDA_MATCHING_H
#define V3D_CUDA_MATCHING_H

#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../v3dm_voctree.h"
#include "../v3dmatch.h"

#if defined(__has_include)
#if __has_include("Math/v3d_linear.h")
#include "Math/v3d_linear.h"
#define V3DM_HAVE_V3D_LINEAR 1
#endif
#endif

namespace V3D_CUDA
{

struct SIFT_Matching
{
   SIFT_Matching()
      : _warnOnRounding(true), _useReductionPath(false), _useGuidedMatching(false), _nMaxLeftFeatures(0),
        _nMaxRightFeatures(0), _nLeftFeatures(0), _nRightFeatures(0), _d_leftFeatures(0), _d_rightFeatures(0),
        _d_scores(0), _d_leftPositions(0), _d_rightPositions(0), _ctx(0)
   {
   }

   ~SIFT_Matching() { v3dm_destroy(_ctx); }

   SIFT_Matching(SIFT_Matching const &) = delete;
   SIFT_Matching &operator=(SIFT_Matching const &) = delete;

   void enableRoundingWarning(bool flag = true)
   {
      _warnOnRounding = flag;
      if (_ctx) v3dm_set_rounding_warning(_ctx, flag);
   }

   void enableReductionPath(bool flag = true)
   {
      _useReductionPath = flag;
      if (_ctx) v3dm_set_reduction_path(_ctx, flag);
   }

   void allocate(int nMaxLeftFeatures, int nMaxRightFeatures, bool useGuidedMatching = false)
   {
      ensureContext();
      check(v3dm_allocate(_ctx, nMaxLeftFeatures, nMaxRightFeatures, useGuidedMatching ? 1 : 0));
      _useGuidedMatching = useGuidedMatching;
      _nMaxLeftFeatures = nMaxLeftFeatures - (nMaxLeftFeatures & 31);
      _nMaxRightFeatures = nMaxRightFeatures - (nMaxRightFeatures & 31);
      _nLeftFeatures = _nRightFeatures = 0;
      refreshBuffers();
   }

   void allocate(int nMaxFeatures, bool useGuidedMatching = false)
   {
      this->allocate(nMaxFeatures, nMaxFeatures, useGuidedMatching);
   }

   void deallocate()
   {
      if (_ctx) check(v3dm_deallocate(_ctx));
      refreshBuffers();
   }

   void setLeftFeatures(int nFeatures, float const *descriptors)
   {
      check(v3dm_set_left_f32(context(), nFeatures, descriptors));
      refreshCounts();
   }

   void setRightFeatures(int nFeatures, float const *descriptors)
   {
      check(v3dm_set_right_f32(context(), nFeatures, descriptors));
      refreshCounts();
   }

   // Additive: uint8 descriptors u (the reference's floats u/512).
   void setLeftFeaturesU8(int nFeatures, std::uint8_t const *descriptors)
   {
      check(v3dm_set_left_u8(context(), nFeatures, descriptors));
      refreshCounts();
   }

   void setRightFeaturesU8(int nFeatures, std::uint8_t const *descriptors)
   {
      check(v3dm_set_right_u8(context(), nFeatures, descriptors));
      refreshCounts();
   }

   // Ratio test is disabled if minRatio < 0.
   void findPutativeMatches(std::vector<std::pair<int, int> > &matches, float const minScore = 0.9f,
                            float const minRatio = 0.8f)
   {
      std::vector<std::int32_t> buf(2 * (size_t)capacity());
      int count = 0;
      check(v3dm_find_putative_matches(context(), minScore, minRatio, buf.data(), capacity(), &count));
      unpack(buf, count, matches);
      refreshBuffers();
   }

#if V3DM_HAVE_V3D_LINEAR
   void runGuidedMatching(V3D::Matrix3x3d const &F, float const distanceThreshold,
                          V3D::Vector2f const *leftPositions, V3D::Vector2f const *rightPositions,
                          std::vector<std::pair<int, int> > &matches, float const minScore = 0.9f,
                          float const minRatio = 0.8f)
   {
      guided(1, F, distanceThreshold, leftPositions, rightPositions, matches, minScore, minRatio);
   }

   void runGuidedHomographyMatching(V3D::Matrix3x3d const &H, float const distanceThreshold,
                                    V3D::Vector2f const *leftPositions, V3D::Vector2f const *rightPositions,
                                    std::vector<std::pair<int, int> > &matches, float const minScore = 0.9f,
                                    float const minRatio = 0.8f)
   {
      guided(2, H, distanceThreshold, leftPositions, rightPositions, matches, minScore, minRatio);
   }
#endif

   template <class Matrix3x3, class Point2>
   void runGuidedMatching(Matrix3x3 const &F, float const distanceThreshold, Point2 const *leftPositions,
                          Point2 const *rightPositions, std::vector<std::pair<int, int> > &matches,
                          float const minScore = 0.9f, float const minRatio = 0.8f)
   {
      guided(1, F, distanceThreshold, leftPositions, rightPositions, matches, minScore, minRatio);
   }

   template <class Matrix3x3, class Point2>
   void runGuidedHomographyMatching(Matrix3x3 const &H, float const distanceThreshold, Point2 const *leftPositions,
                                    Point2 const *rightPositions, std::vector<std::pair<int, int> > &matches,
                                    float const minScore = 0.9f, float const minRatio = 0.8f)
   {
      guided(2, H, distanceThreshold, leftPositions, rightPositions, matches, minScore, minRatio);
   }

 protected:
   bool _warnOnRounding;
   bool _useReductionPath;
   bool _useGuidedMatching;
   int _nMaxLeftFeatures, _nMaxRightFeatures, _nLeftFeatures, _nRightFeatures;
   // device buffers (v3d_cudamatching.h:64-65), refreshed after every call; null
   // where this engine keeps no such buffer (v3dm_get_device_buffers)
   float *_d_leftFeatures, *_d_rightFeatures, *_d_scores;
   float *_d_leftPositions, *_d_rightPositions;
   v3dm_ctx *_ctx;

 private:
   void check(int rc) const
   {
      if (rc != V3DM_OK) throw std::runtime_error(std::string("SIFT_Matching: ") + v3dm_last_error(_ctx));
   }

   void ensureContext()
   {
      if (_ctx) return;
      if (v3dm_create(0, &_ctx) != V3DM_OK)
         throw std::runtime_error(std::string("SIFT_Matching: ") + v3dm_last_error(0));
      v3dm_set_rounding_warning(_ctx, _warnOnRounding);
      v3dm_set_reduction_path(_ctx, _useReductionPath);
   }

   v3dm_ctx *context()
   {
      if (!_ctx) throw std::runtime_error("SIFT_Matching: allocate() has not been called");
      return _ctx;
   }

   int capacity() const { return _nMaxLeftFeatures > 0 ? _nMaxLeftFeatures : 1; }

   void refreshCounts()
   {
      v3dm_get_counts(_ctx, &_nLeftFeatures, &_nRightFeatures);
      refreshBuffers();
   }

   void refreshBuffers()
   {
      const float *b[5] = {0, 0, 0, 0, 0};
      if (_ctx) v3dm_get_device_buffers(_ctx, &b[0], &b[1], &b[2], &b[3], &b[4]);
      _d_leftFeatures = const_cast<float *>(b[0]);
      _d_rightFeatures = const_cast<float *>(b[1]);
      _d_scores = const_cast<float *>(b[2]);
      _d_leftPositions = const_cast<float *>(b[3]);
      _d_rightPositions = const_cast<float *>(b[4]);
   }

   static void unpack(std::vector<std::int32_t> const &buf, int count, std::vector<std::pair<int, int> > &matches)
   {
      matches.clear();
      matches.reserve(count);
      for (int k = 0; k < count; ++k) matches.push_back(std::make_pair((int)buf[2 * k], (int)buf[2 * k + 1]));
   }

   template <class Matrix3x3, class Point2>
   void guided(int kind, Matrix3x3 const &M, float thr, Point2 const *lp, Point2 const *rp,
               std::vector<std::pair<int, int> > &matches, float minScore, float minRatio)
   {
      double m[9];
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) m[3 * i + j] = M[i][j];
      std::vector<float> l(2 * (size_t)_nLeftFeatures), r(2 * (size_t)_nRightFeatures);
      for (int k = 0; k < _nLeftFeatures; ++k)
      {
         l[2 * k] = lp[k][0];
         l[2 * k + 1] = lp[k][1];
      }
      for (int k = 0; k < _nRightFeatures; ++k)
      {
         r[2 * k] = rp[k][0];
         r[2 * k + 1] = rp[k][1];
      }
      std::vector<std::int32_t> buf(2 * (size_t)capacity());
      int count = 0;
      int rc = kind == 1 ? v3dm_run_guided_matching(context(), m, thr, l.data(), r.data(), minScore, minRatio,
                                                    buf.data(), capacity(), &count)
                         : v3dm_run_guided_homography_matching(context(), m, thr, l.data(), r.data(), minScore,
                                                               minRatio, buf.data(), capacity(), &count);
      check(rc);
      unpack(buf, count, matches);
      refreshBuffers();
   }
};

// Batch tree traversal on the GPU with the reference class's interface: node
// descriptors once (setNodes), then traverseTree for up to nMaxFeatures keys
// at a time.  leaveIds receive the leaf's node id in the whole tree (the
// reference kernel's convention: leaf index + getNodeCount() - getLeaveCount());
// a key with a non-finite value gets -1.  Device errors throw
// std::runtime_error; the destructor releases the tree.
template <int BranchFactor, int Depth>
struct VocabularyTreeTraversal
{
   VocabularyTreeTraversal() : _nMaxFeatures(0), _tree(0) {}
   ~VocabularyTreeTraversal() { v3dm_voctree_destroy(_tree); }

   VocabularyTreeTraversal(VocabularyTreeTraversal const &) = delete;
   VocabularyTreeTraversal &operator=(VocabularyTreeTraversal const &) = delete;

   void allocate(int nMaxFeatures)
   {
      if ((nMaxFeatures % 32) != 0) // v3d_cudamatching.cpp:271-276
      {
         std::cerr << "VocabularyTreeTraversal::allocate() warning: nMaxFeatures should be a multiple of 32"
                   << std::endl;
         std::cerr << "Rounding nFeatures down." << std::endl;
         nMaxFeatures = nMaxFeatures - (nMaxFeatures & 31);
      }
      _nMaxFeatures = nMaxFeatures;
   }

   void deallocate()
   {
      v3dm_voctree_destroy(_tree);
      _tree = 0;
   }

   void setNodes(float const *nodes)
   {
      v3dm_voctree_destroy(_tree);
      _tree = 0;
      if (v3dm_voctree_create(0, Depth, BranchFactor, nodes, &_tree) != V3DM_OK)
         throw std::runtime_error(std::string("VocabularyTreeTraversal: ") + v3dm_voctree_last_error(0));
   }

   void traverseTree(int nFeatures, float const *features, int *leaveIds)
   {
      static_assert(sizeof(int) == sizeof(std::int32_t), "int is 32 bits");
      if (nFeatures > _nMaxFeatures) // v3d_cudamatching.cpp:311-316
      {
         std::cerr << "VocabularyTreeTraversal::traverseTree() warning: nFeatures is larger than nMaxFeatures."
                   << std::endl;
         std::cerr << "Reducing nFeatures." << std::endl;
         nFeatures = _nMaxFeatures;
      }
      if (!_tree) throw std::runtime_error("VocabularyTreeTraversal: setNodes() has not been called");
      if (nFeatures <= 0) return;
      if (v3dm_voctree_quantize(_tree, nFeatures, features, reinterpret_cast<std::int32_t *>(leaveIds)) != V3DM_OK)
         throw std::runtime_error(std::string("VocabularyTreeTraversal: ") + v3dm_voctree_last_error(_tree));
      int const offset = getNodeCount() - getLeaveCount();
      for (int i = 0; i < nFeatures; ++i)
         if (leaveIds[i] >= 0) leaveIds[i] += offset;
   }

   static inline int getNodeCount()
   {
      int res = 0, len = 1;
      for (int d = 0; d < Depth; ++d)
      {
         len *= BranchFactor;
         res += len;
      }
      return res;
   }

   static inline int getLeaveCount()
   {
      int res = 1;
      for (int d = 0; d < Depth; ++d) res *= BranchFactor;
      return res;
   }

 protected:
   int _nMaxFeatures;
   v3dm_voctree *_tree;
};

} // end namespace V3D_CUDA

#endif
