/*
 * v3dm_voctree.h -- C ABI of the vocabulary-tree stage of the matching
 * drivers: GPU quantisation of SIFT descriptors into visual words and the
 * inverted-file document scoring that picks which view pairs get matched.
 *
 * Reference: V3D::VocabularyTree<Depth, BranchFactor, 128> and
 * V3D::LeafDocuments (V3D/include/Base/v3d_voctree.h:19-309), used by
 * SfM/match_using_voctree/src/match_using_voctree.cpp:147-207
 * (VocabularyTree<3, 50, 128>, computeBinomialScores, POTENTIAL_VIEWS_COUNT
 * best documents per view).  The reference's own GPU traversal
 * (_cuda_traverseTree, V3D/src/CUDA/v3d_cudamatching.cu:435-528) sums in a
 * different order and no driver calls it; the words here are those of the
 * CPU tree the driver uses.
 *
 * Status codes and v3dm_last_error(NULL) as in v3dmatch.h (the tree object
 * also keeps its own message, v3dm_voctree_last_error).
 */
#ifndef V3DM_VOCTREE_H
#define V3DM_VOCTREE_H

#include "v3dmatch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct v3dm_voctree v3dm_voctree;

/* --- vocabulary tree (GPU) ---------------------------------------------------- */

/* VocabularyTree<depth, branchFactor, 128>() with its node descriptors:
 * branchFactor + branchFactor^2 + ... + branchFactor^depth nodes of 128
 * floats, level by level, the children of node p at (p+1)*branchFactor ..
 * (p+1)*branchFactor + branchFactor-1 (v3d_voctree.h:22-36,63-79).  The tree
 * lives on HIP device `device` (65 MB for depth 3, branch factor 50). */
int v3dm_voctree_create(int device, int depth, int branchFactor, const float *nodes, v3dm_voctree **out);

/* As v3dm_voctree_create with the nodes read from a tree file:
 * loadTreeFromFile (v3d_voctree.h:38-52) -- raw float32 node descriptors, no
 * header.  A file shorter than the tree is an error here (the reference
 * leaves the missing nodes zero); longer files are read up to the tree size. */
int v3dm_voctree_load(int device, int depth, int branchFactor, const char *path, v3dm_voctree **out);

void v3dm_voctree_destroy(v3dm_voctree *tree);
const char *v3dm_voctree_last_error(const v3dm_voctree *tree);

/* depth, branch factor, node count and visualWordCount() (= leafCount(), v3d_voctree.h:54,61). */
int v3dm_voctree_info(const v3dm_voctree *tree, int *depth, int *branchFactor, int *nodes, int *leaves);

/* getVisualWord(key) for n keys (n x 128 floats, row-major): words[i] is the
 * leaf index in [0, leaves) that findBestLeafIndex (v3d_voctree.h:63-99, findBestNodeIndex :82-99)
 * returns -- bit-identical, the same FP32 sums in the same order.  A key
 * whose scores never exceed -1e30f (non-finite input) gets -1 (the reference
 * then indexes outside the tree). */
int v3dm_voctree_quantize(v3dm_voctree *tree, int64_t n, const float *keys, int32_t *words);

/* How the children are scored (the words are the same either way):
 * V3DM_VOCTREE_AUTO (default) -- for branch factors up to 64, an FP32-MFMA pass
 *   scores all children of 64 keys at once; only children within a proven
 *   rounding-error bound of the best (usually none) are rescored with the
 *   reference's sequential sum; larger branch factors use the exact kernel;
 * V3DM_VOCTREE_EXACT_VALU -- every child with the sequential sum. */
#define V3DM_VOCTREE_AUTO 0
#define V3DM_VOCTREE_EXACT_VALU 1
int v3dm_voctree_set_mode(v3dm_voctree *tree, int mode);

/* The same job in three steps, for callers that keep keys resident: upload
 * n keys, run the quantisation on the device (may be repeated; no transfer),
 * fetch the words of the last run (waits). */
int v3dm_voctree_upload(v3dm_voctree *tree, int64_t n, const float *keys);
int v3dm_voctree_run(v3dm_voctree *tree);
int v3dm_voctree_fetch(v3dm_voctree *tree, int32_t *words);

/* Waits, then returns the summed device times (ms) of the last run: the
 * scoring kernels (one per tree level, timed with events bound to their
 * dispatches) and the whole run including the regrouping between levels. */
int v3dm_voctree_run_times(v3dm_voctree *tree, double *scoringMs, double *runMs);

/* --- inverted file (host) ------------------------------------------------------ */

typedef struct v3dm_leaf_documents v3dm_leaf_documents;

/* LeafDocuments(nLeaves) (v3d_voctree.h:139-141). */
int v3dm_leaf_documents_create(int nLeaves, v3dm_leaf_documents **out);
void v3dm_leaf_documents_destroy(v3dm_leaf_documents *docs);

/* insertDocument(visualWords) (v3d_voctree.h:143-164); *documentId = 0, 1, ...
 * Words outside [0, nLeaves) are an error (the reference indexes out of bounds). */
int v3dm_leaf_documents_insert(v3dm_leaf_documents *docs, int n, const int32_t *words, int *documentId);

int v3dm_leaf_documents_count(const v3dm_leaf_documents *docs, int *nDocuments);

/* getDocumentHistogram(visualWords, histogram) (v3d_voctree.h:166-178): for every
 * word of the list (repeats included), 1 is ADDED to histogram[d] for each
 * document d in that word's inverted file; histogram holds cap >= documents
 * entries (the caller sizes and clears it, as the reference's resize does). */
int v3dm_leaf_documents_histogram(const v3dm_leaf_documents *docs, int n, const int32_t *words, int32_t *histogram,
                                  int cap);

/* The state serialize() writes (v3d_voctree.h:291-300): leaves, documents and
 * inverted-file entries, then per leaf its entry count, the (document, count)
 * entries leaf after leaf in insertion order, and per document its feature
 * count, distinct-word count and 1.0f / feature count.  Any output may be NULL. */
int v3dm_leaf_documents_sizes(const v3dm_leaf_documents *docs, int *nLeaves, int *nDocuments, int64_t *nEntries);
int v3dm_leaf_documents_export(const v3dm_leaf_documents *docs, int32_t *entriesPerLeaf, int32_t *entries,
                               int32_t *nDocumentFeatures, int32_t *nDocumentLeaves, float *rcpNDocumentFeatures);
/* Replaces the whole state with the given one (serialize() when loading). */
int v3dm_leaf_documents_import(v3dm_leaf_documents *docs, int nLeaves, int nDocuments, const int32_t *entriesPerLeaf,
                               const int32_t *entries, const int32_t *nDocumentFeatures,
                               const int32_t *nDocumentLeaves, const float *rcpNDocumentFeatures);

/* computeBinomialScores(pMatch, visualWords, scores, nBest) (v3d_voctree.h:242-289):
 * per document the double-precision log-likelihood ratio z0 - z1 of the
 * query's k shared words under p_random = max(n_doc, n_query) / nLeaves and
 * p_random + pMatch, stored as float; the best (smallest) first, ties by
 * document id.  Writes min(nBest, documents) entries (all when nBest < 0) to
 * scores / documents (cap entries; V3DM_ERR_CAPACITY with *count set). */
int v3dm_leaf_documents_binomial_scores(v3dm_leaf_documents *docs, double pMatch, int n, const int32_t *words,
                                        int nBest, float *scores, int32_t *documents, int cap, int *count);

/* compute_L1_Scores(visualWords, scores, nBest) (v3d_voctree.h:198-240): the L1
 * distance of the normalised word histograms through the inverted file, in
 * the reference's float arithmetic and order; same output convention. */
int v3dm_leaf_documents_l1_scores(v3dm_leaf_documents *docs, int n, const int32_t *words, int nBest, float *scores,
                                  int32_t *documents, int cap, int *count);

/* The first stage of match_using_voctree (match_using_voctree.cpp:147-207) over
 * already quantised views: for view i = 0, 1, ... (words of view i at
 * words[sum(counts[0..i-1]) ...]) take the nBest best binomial-score documents
 * among views 0 .. i-1 (pMatch = VISUAL_WORD_MATCH_PROPABILITY, which the driver
 * holds in a float, :137), then insert
 * view i.  Writes the candidate pairs (earlier view, view i) in the order the
 * driver matches them -- ascending earlier view, then ascending i -- with each
 * pair's score; cap in pairs, V3DM_ERR_CAPACITY with *count set. */
int v3dm_voctree_candidate_pairs(int nViews, const int32_t *counts, const int32_t *words, int nLeaves, int nBest,
                                 double pMatch, int32_t *pairs, float *pairScores, int cap, int *count);

#ifdef __cplusplus
}
#endif

#endif
