// -*- C++ -*-
// Drop-in V3D::VocabularyTree and V3D::LeafDocuments for the MI355X engine.
//
// Same header path, class names, template parameters, public members,
// argument meaning and defaults as the reference
// (V3D/include/Base/v3d_voctree.h:19-309), so SfM/match_using_voctree
// (match_using_voctree.cpp:145-207) compiles unchanged with include/ ahead of
// V3D/include on the include path.
//
//  * VocabularyTree keeps the node descriptors on the host as the reference
//    does; getVisualWord / findBestLeafIndex / findBestNodeIndex run the same
//    greedy descent there -- each score a sequential FP32 sum with every
//    product rounded before its add (no contraction into fused multiply-adds,
//    whatever the caller's flags), strict '>' from -1e30f -- so single keys
//    cost what they cost in the reference.  getVisualWords (not in the
//    reference) quantises many keys at once on the GPU (v3dm_voctree.h) with
//    bit-identical words; the tree is uploaded on its first call and again
//    after loadTreeFromFile.
//  * LeafDocuments forwards to the library's inverted file (v3dm_voctree.h),
//    which keeps the reference's arithmetic: binomial scores in double with the
//    C library's log, L1 scores in float in the reference's order, results
//    sorted by (score, document).
//
// Differences a caller can observe: device failures and words outside
// [0, nLeaves) throw std::runtime_error (the reference indexes out of bounds);
// a key with a non-finite value gets -1 from getVisualWords (the reference's
// descent then reads outside the tree); LeafDocuments is movable, not
// copyable, and has no serialize()/getDocumentHistogram() (no driver uses
// them).
#ifndef V3D_VOCTREE_H
#define V3D_VOCTREE_H

#include <cstdint>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../v3dm_voctree.h"

namespace V3D
{

namespace v3dm_detail
{
// One product, then one add, each rounded to float: the reference's
// `score += key[i] * curNode[i]` on x86-64 without FMA contraction.
#if defined(__clang__)
inline float node_score(float const *key, float const *node, int len)
{
#pragma clang fp contract(off)
   float score = 0;
   for (int i = 0; i < len; ++i) score += key[i] * node[i];
   return score;
}
#elif defined(__GNUC__)
__attribute__((optimize("fp-contract=off"))) inline float node_score(float const *key, float const *node, int len)
{
   float score = 0;
   for (int i = 0; i < len; ++i) score += key[i] * node[i];
   return score;
}
#else
inline float node_score(float const *key, float const *node, int len)
{
   float score = 0;
   for (int i = 0; i < len; ++i)
   {
      volatile float product = key[i] * node[i]; // keeps the product rounded on its own
      score += product;
   }
   return score;
}
#endif
} // namespace v3dm_detail

template <int Depth, int BranchFactor, int DescLen = 128>
struct VocabularyTree
{
   static_assert(DescLen == 128, "the engine quantises 128-element descriptors");

   VocabularyTree() : _levelStarts(Depth + 1, 0)
   {
      int size = 0, len = 1;
      for (int d = 0; d < Depth; ++d)
      {
         len *= BranchFactor;
         size += len;
         _levelStarts[d + 1] = size;
      }
      _tree.assign((size_t)size * DescLen, 0.0f); // nodes a short tree file does not hold stay zero
   }

   VocabularyTree(VocabularyTree const &o) : _levelStarts(o._levelStarts), _tree(o._tree), _device(nullptr) {}
   VocabularyTree &operator=(VocabularyTree const &o)
   {
      if (this != &o)
      {
         release();
         _levelStarts = o._levelStarts;
         _tree = o._tree;
      }
      return *this;
   }
   ~VocabularyTree() { release(); }

   // loadTreeFromFile (v3d_voctree.h:38-52): raw float32 nodes, level by level;
   // false (with the reference's message) only when the file cannot be opened.
   bool loadTreeFromFile(char const *filename)
   {
      std::ifstream is(filename, std::ios::binary);
      if (!is)
      {
         std::cerr << "Could not open " << filename << std::endl;
         return false;
      }
      is.read(reinterpret_cast<char *>(&_tree[0]), (std::streamsize)(sizeof(float) * _tree.size()));
      release(); // the device copy is rebuilt from these nodes on the next getVisualWords
      return true;
   }

   int visualWordCount() const { return this->leafCount(); }
   int leafCount() const { return _levelStarts[Depth] - _levelStarts[Depth - 1]; }

   int getVisualWord(float const *key) const { return this->findBestLeafIndex(key); }

   // findBestLeafIndex (v3d_voctree.h:63-80): the child of node p is searched
   // among nodes (p + 1) * BranchFactor ..; the word is the leaf's offset in the
   // last level.
   int findBestLeafIndex(float const *key) const
   {
      int node = 0, next = 0;
      for (int d = 0; d < Depth; ++d)
      {
         node = next + findBestNodeIndex(key, &_tree[(size_t)next * DescLen]);
         next = (node + 1) * BranchFactor;
      }
      return node - _levelStarts[Depth - 1];
   }

   // findBestNodeIndex (v3d_voctree.h:82-99).
   static int findBestNodeIndex(float const *key, float const *nodes)
   {
      int best = -1;
      float maxScore = -1e30f;
      for (int k = 0; k < BranchFactor; ++k)
      {
         const float score = v3dm_detail::node_score(key, nodes + (size_t)k * DescLen, DescLen);
         if (score > maxScore)
         {
            maxScore = score;
            best = k;
         }
      }
      return best;
   }

   // Not in the reference: words[i] = getVisualWord(keys + 128 i) for n keys
   // (row-major), on HIP device 0, bit-identical to the host descent; -1 for a
   // key with a non-finite value.
   void getVisualWords(int n, float const *keys, int *words) const
   {
      if (n <= 0) return;
      if (!_device && v3dm_voctree_create(0, Depth, BranchFactor, _tree.data(), &_device) != V3DM_OK)
         throw std::runtime_error(std::string("VocabularyTree: ") + v3dm_voctree_last_error(nullptr));
      if (v3dm_voctree_quantize(_device, n, keys, reinterpret_cast<int32_t *>(words)) != V3DM_OK)
         throw std::runtime_error(std::string("VocabularyTree: ") + v3dm_voctree_last_error(_device));
   }

 protected:
   void release() const
   {
      if (_device) v3dm_voctree_destroy(_device);
      _device = nullptr;
   }

   std::vector<int> _levelStarts;
   std::vector<float> _tree;
   mutable v3dm_voctree *_device = nullptr;
}; // end struct VocabularyTree

struct LeafDocuments
{
   explicit LeafDocuments(int nLeaves)
   {
      if (v3dm_leaf_documents_create(nLeaves, &_docs) != V3DM_OK) fail("LeafDocuments");
   }
   LeafDocuments(LeafDocuments const &) = delete;
   LeafDocuments &operator=(LeafDocuments const &) = delete;
   LeafDocuments(LeafDocuments &&o) noexcept : _docs(o._docs) { o._docs = nullptr; }
   LeafDocuments &operator=(LeafDocuments &&o) noexcept
   {
      std::swap(_docs, o._docs);
      return *this;
   }
   ~LeafDocuments() { v3dm_leaf_documents_destroy(_docs); }

   // insertDocument (v3d_voctree.h:143-164): the new document's id, 0, 1, ...
   int insertDocument(std::vector<int> const &visualWords)
   {
      int id = -1;
      if (v3dm_leaf_documents_insert(_docs, (int)visualWords.size(), wordsOf(visualWords), &id) != V3DM_OK)
         fail("insertDocument");
      return id;
   }

   // getVisualWordHistogram (v3d_voctree.h:180-196): leaf -> count.
   void getVisualWordHistogram(std::vector<int> const &visualWords, std::map<int, int> &histogram) const
   {
      histogram.clear();
      for (int w : visualWords) ++histogram[w];
   }

   // compute_L1_Scores (v3d_voctree.h:198-240): (score, document), the best
   // (smallest) first, nBestDocuments of them (all when negative).
   void compute_L1_Scores(std::vector<int> const &visualWords, std::vector<std::pair<float, int>> &scores,
                          int nBestDocuments = -1)
   {
      fetch(visualWords, nBestDocuments, scores, [&](int cap, float *s, int32_t *d, int *count) {
         return v3dm_leaf_documents_l1_scores(_docs, (int)visualWords.size(), wordsOf(visualWords), nBestDocuments, s,
                                              d, cap, count);
      });
   }

   // computeBinomialScores (v3d_voctree.h:242-289).
   void computeBinomialScores(double const p_match, std::vector<int> const &visualWords,
                              std::vector<std::pair<float, int>> &scores, int nBestDocuments = -1)
   {
      fetch(visualWords, nBestDocuments, scores, [&](int cap, float *s, int32_t *d, int *count) {
         return v3dm_leaf_documents_binomial_scores(_docs, p_match, (int)visualWords.size(), wordsOf(visualWords),
                                                    nBestDocuments, s, d, cap, count);
      });
   }

 protected:
   static int32_t const *wordsOf(std::vector<int> const &w)
   {
      static_assert(sizeof(int) == sizeof(int32_t), "int is 32 bits");
      return reinterpret_cast<int32_t const *>(w.data());
   }
   [[noreturn]] void fail(char const *what) const
   {
      throw std::runtime_error(std::string("LeafDocuments::") + what + ": " + v3dm_last_error(nullptr));
   }
   template <class Call>
   void fetch(std::vector<int> const &, int, std::vector<std::pair<float, int>> &scores, Call call)
   {
      int nDocs = 0;
      v3dm_leaf_documents_count(_docs, &nDocs);
      std::vector<float> s((size_t)(nDocs > 0 ? nDocs : 1));
      std::vector<int32_t> d(s.size());
      int count = 0;
      if (call((int)s.size(), s.data(), d.data(), &count) != V3DM_OK) fail("scores");
      scores.resize((size_t)count);
      for (int i = 0; i < count; ++i) scores[(size_t)i] = std::make_pair(s[(size_t)i], (int)d[(size_t)i]);
   }

   v3dm_leaf_documents *_docs = nullptr;
}; // end struct LeafDocuments

} // namespace V3D

#endif
