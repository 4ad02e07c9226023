/*
 * v3dm_oracle_voctree.c -- CPU oracle for vocabulary-tree quantisation.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by the test suite and the benchmark's
 * CPU-baseline leg as the checker; the product library never links it.
 *
 * Restates V3D::VocabularyTree<Depth, BranchFactor, 128>::findBestLeafIndex
 * and findBestNodeIndex (V3D/include/Base/v3d_voctree.h:63-99): a greedy
 * descent from the root; at each level the key's score against every child,
 * score = sum over i = 0..127 of key[i]*node[i] accumulated in float in
 * ascending i with the product rounded before the add (this file is built
 * with -ffp-contract=off, as x86-64 code without FMA contraction computes
 * it); the first child with the largest score wins (strict '>', initial
 * -1e30f); leafId = nextLeafId + k, nextLeafId = (leafId + 1) * BranchFactor;
 * the word is leafId - levelStart[Depth-1].
 *
 * A key none of whose scores exceeds -1e30f makes the reference return -1
 * from findBestNodeIndex and walk outside the tree; the oracle returns -1
 * for such a key (the product does the same).
 */
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VT_D 128

static int vt_best_node(const float *key, const float *nodes, int bf)
{
   int best = -1;
   float maxScore = -1e30f;
   for (int k = 0; k < bf; ++k)
   {
      const float *cur = nodes + (int64_t)k * VT_D;
      float score = 0.0f;
      for (int i = 0; i < VT_D; ++i) score += key[i] * cur[i];
      if (score > maxScore)
      {
         best = k;
         maxScore = score;
      }
   }
   return best;
}

int orc_voctree_word(int depth, int bf, const float *tree, const float *key)
{
   int64_t levelStart = 0, len = 1; /* start of the deepest level */
   for (int d = 0; d < depth - 1; ++d)
   {
      len *= bf;
      levelStart += len;
   }
   int64_t leafId = 0, nextLeafId = 0;
   for (int d = 0; d < depth; ++d)
   {
      const int k = vt_best_node(key, tree + nextLeafId * VT_D, bf);
      if (k < 0) return -1;
      leafId = nextLeafId + k;
      nextLeafId = (leafId + 1) * bf;
   }
   return (int)(leafId - levelStart);
}

void orc_voctree_quantize(int depth, int bf, const float *tree, int64_t n, const float *keys, int32_t *words)
{
#pragma omp parallel for schedule(dynamic, 64)
   for (int64_t f = 0; f < n; ++f) words[f] = orc_voctree_word(depth, bf, tree, keys + f * VT_D);
}
