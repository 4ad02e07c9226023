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
// copyable.  LeafDocuments::serialize (with load / save and the stream
// operators) is there when the reference's Base/v3d_serialization.h is on the
// include path, as it is in the reference's build: it reads and writes the
// reference's byte stream.
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

#if defined(__has_include)
#if __has_include("Base/v3d_serialization.h")
#include <list>

#include "Base/v3d_serialization.h"
#define V3DM_HAVE_V3D_SERIALIZATION 1
#endif
#endif

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

   // getDocumentHistogram (v3d_voctree.h:166-178): histogram is resized to the
   // document count (earlier entries kept, as resize keeps them), then every
   // word of the list adds 1 to each document in its inverted file.
   void getDocumentHistogram(std::vector<int> const &visualWords, std::vector<int> &histogram) const
   {
      int nDocs = 0;
      v3dm_leaf_documents_count(_docs, &nDocs);
      histogram.resize((size_t)nDocs, 0);
      static_assert(sizeof(int) == sizeof(int32_t), "int is 32 bits");
      if (v3dm_leaf_documents_histogram(_docs, (int)visualWords.size(), wordsOf(visualWords),
                                        reinterpret_cast<int32_t *>(histogram.data()), nDocs) != V3DM_OK)
         fail("getDocumentHistogram");
   }

   // getVisualWordHistogram (v3d_voctree.h:180-196): leaf -> count.
   void getVisualWordHistogram(std::vector<int> const &visualWords, std::map<int, int> &histogram) const
   {
      histogram.clear();
      for (int w : visualWords) ++histogram[w];
   }

   // compute_L1_Scores (v3d_voctree.h:198-240): one (score, document) per
   // document, the best (smallest) first.  The reference sorts only the first
   // nBestDocuments (partial_sort) and leaves the rest unordered; here all are
   // sorted, so the first nBestDocuments are the reference's and the size is
   // the document count, as there.
   void compute_L1_Scores(std::vector<int> const &visualWords, std::vector<std::pair<float, int>> &scores,
                          int nBestDocuments = -1)
   {
      (void)nBestDocuments;
      fetch(visualWords, -1, scores, [&](int cap, float *s, int32_t *d, int *count) {
         return v3dm_leaf_documents_l1_scores(_docs, (int)visualWords.size(), wordsOf(visualWords), -1, s, d, cap,
                                              count);
      });
   }

   // computeBinomialScores (v3d_voctree.h:242-289); ordering as compute_L1_Scores.
   void computeBinomialScores(double const p_match, std::vector<int> const &visualWords,
                              std::vector<std::pair<float, int>> &scores, int nBestDocuments = -1)
   {
      (void)nBestDocuments;
      fetch(visualWords, -1, scores, [&](int cap, float *s, int32_t *d, int *count) {
         return v3dm_leaf_documents_binomial_scores(_docs, p_match, (int)visualWords.size(), wordsOf(visualWords), -1,
                                                    s, d, cap, count);
      });
   }

#if V3DM_HAVE_V3D_SERIALIZATION
   // The reference's inverted-file entry list and its byte layout
   // (v3d_voctree.h:106-136): the entry count, then (document, count) pairs
   // inside one scope.
   struct entries_type : public std::list<std::pair<int, int> >
   {
      template <typename Archive> void serialize(Archive &ar)
      {
         unsigned int sz = (unsigned int)this->size();
         ar & sz;
         SerializationScope<Archive> scope(ar);
         if (ar.isLoading())
         {
            this->clear();
            for (unsigned int i = 0; i < sz; ++i)
            {
               int doc = 0, count = 0;
               ar & doc & count;
               this->push_back(std::make_pair(doc, count));
            }
         }
         else
            for (auto &e : *this)
            {
               int doc = e.first, count = e.second;
               ar & doc & count;
            }
      }
      V3D_DEFINE_LOAD_SAVE(entries_type);
   };

   // serialize (v3d_voctree.h:291-298): the document count, the inverted files,
   // then per document its feature count, distinct-word count and 1.0f /
   // feature count -- the reference's members in its order and types.
   template <typename Archive> void serialize(Archive &ar)
   {
      int nDocuments = 0;
      SerializableVector<entries_type> invertedFiles;
      SerializableVector<int> nDocumentFeatures, nDocumentLeaves;
      SerializableVector<float> rcpNDocumentFeatures;
      if (!ar.isLoading())
      {
         int nLeaves = 0;
         int64_t nEntries = 0;
         v3dm_leaf_documents_sizes(_docs, &nLeaves, &nDocuments, &nEntries);
         std::vector<int32_t> perLeaf((size_t)nLeaves), entries(2 * (size_t)nEntries + 2);
         nDocumentFeatures.resize((size_t)nDocuments);
         nDocumentLeaves.resize((size_t)nDocuments);
         rcpNDocumentFeatures.resize((size_t)nDocuments);
         if (v3dm_leaf_documents_export(_docs, perLeaf.data(), entries.data(),
                                        reinterpret_cast<int32_t *>(nDocumentFeatures.data()),
                                        reinterpret_cast<int32_t *>(nDocumentLeaves.data()),
                                        rcpNDocumentFeatures.data()) != V3DM_OK)
            fail("serialize");
         invertedFiles.resize((size_t)nLeaves);
         size_t k = 0;
         for (int l = 0; l < nLeaves; ++l)
            for (int j = 0; j < perLeaf[(size_t)l]; ++j, ++k)
               invertedFiles[(size_t)l].push_back(std::make_pair((int)entries[2 * k], (int)entries[2 * k + 1]));
      }
      ar & nDocuments;
      ar & invertedFiles;
      ar & nDocumentFeatures;
      ar & nDocumentLeaves;
      ar & rcpNDocumentFeatures;
      if (ar.isLoading())
      {
         std::vector<int32_t> perLeaf, entries;
         for (auto const &list : invertedFiles)
         {
            perLeaf.push_back((int32_t)list.size());
            for (auto const &e : list)
            {
               entries.push_back(e.first);
               entries.push_back(e.second);
            }
         }
         if (invertedFiles.empty() || (int)nDocumentFeatures.size() != nDocuments ||
             (int)nDocumentLeaves.size() != nDocuments || (int)rcpNDocumentFeatures.size() != nDocuments ||
             v3dm_leaf_documents_import(_docs, (int)invertedFiles.size(), nDocuments, perLeaf.data(),
                                        entries.empty() ? nullptr : entries.data(),
                                        reinterpret_cast<int32_t const *>(nDocumentFeatures.data()),
                                        reinterpret_cast<int32_t const *>(nDocumentLeaves.data()),
                                        rcpNDocumentFeatures.data()) != V3DM_OK)
            fail("serialize");
      }
   }

   V3D_DEFINE_LOAD_SAVE(LeafDocuments);
#endif

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

#if V3DM_HAVE_V3D_SERIALIZATION
V3D_DEFINE_IOSTREAM_OPS(LeafDocuments);
#endif

} // namespace V3D

#endif
