/*
 * oracle.h -- CPU restatement of RapidBoW's hot-path semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it to produce expected
 * outputs.  Nothing in the product library links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Builder-fixed rules for behaviour the reference leaves undefined or
 * toolchain-dependent are stated at each function.
 *
 * Layouts used at this interface:
 *   descriptors    uint8 [n][128], row-major (RapidBoW.h:50 SiftAttr)
 *   u8 tree (b=16) reference Node layout, 2048 B per node:
 *                  byte(dim k, child j) = (k/4)*64 + 4*j + k%4   (Vocabulary.h:42)
 *   f32 tree       float [node][b][128] (builder-defined extension, b = 16 or 32)
 *   weights        float [node][b]                                  (Vocabulary.h:64-66)
 *   nodes are breadth-first, node(level, i) = levelBeg(level) + i     (FullTree.h:37-66)
 */
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FullTree.h:37-52 -- first node index of a level: (b^level - 1) / (b - 1). */
uint64_t orc_level_beg(uint32_t b, uint32_t level);

/* kmeans.cuh:61-71 -- sum over 128 dims of (x - c)^2, exact in uint32. */
uint32_t orc_dist_u8(const uint8_t* x, const uint8_t* c);
/* Builder-defined FP32 distance for float centroids: acc = fma(d, d, acc) with
 * d = float(x_k) - c_k, k = 0..127 ascending, acc starting at +0. */
float orc_dist_f32(const uint8_t* x, const float* c);

/* vocTreeLookUp.cu:83-159 -- level-by-level descent on a b=16 uint8 tree in
 * the reference Node layout; ties go to the lowest child (packed key min,
 * :105-125); masks are never consulted. */
void orc_lookup_u8(const uint8_t* nodesRef, uint32_t nbLevels, const uint8_t* desc,
                   uint64_t nbDesc, uint32_t* leaf, int nthreads);
/* Same descent over a float tree [node][b][128] with orc_dist_f32. */
void orc_lookup_f32(const float* nodes, uint32_t b, uint32_t nbLevels, const uint8_t* desc,
                    uint64_t nbDesc, uint32_t* leaf, int nthreads);

/* kmeans.cu:120-198 one assign pass over n row-major points against 16 centers
 * (row-major [16][128]).  assign[] holds the previous assignment on entry and
 * the new one on exit.  counts[16], sums[16][128] (int64) are overwritten.
 * Returns #points whose assignment changed. Ties -> lowest center (:131-135). */
uint32_t orc_assign_u8(const uint8_t* pts, uint32_t n, const uint8_t* centers,
                       uint8_t* assign, uint32_t* counts, int64_t* sums);
/* kmeans.cu:334-375: c = uint8(roundf(float(sum)/float(count))) where count>0.
 * IEEE division (the reference's non-fast-math build). */
void orc_update_u8(uint8_t* centers, const uint32_t* counts, const int64_t* sums);
/* FP32-centroid variants (builder-defined extension): k centers [k][128]. */
uint32_t orc_assign_f32(const uint8_t* pts, uint32_t n, const float* centers, uint32_t k,
                        uint8_t* assign, uint32_t* counts, int64_t* sums);
void orc_update_f32(float* centers, uint32_t k, const uint32_t* counts, const int64_t* sums);

/* reformatDesc.cu:38-74 -- gather desc[idx[t][s]] into transposed tiles
 * tile[t][word][s] (words of 4 bytes, tileSize slots); idxInvalid -> 0xFFFFFFFF. */
void orc_reformat(const uint8_t* desc, const uint32_t* idx, uint32_t nbTiles, uint32_t tileSize,
                  uint32_t* tiles);

/* kmeans.cpp:141-168 randPick with a std::mt19937_64 engine that persists
 * across calls (the reference's thread_local engine). */
void* orc_rng_new(uint64_t seed);
void orc_rng_free(void* rng);
uint64_t orc_rng_next(void* rng);
uint32_t orc_rand_pick(void* rng, uint32_t* dst, uint32_t randSpace, uint32_t nbToPick);

/* Builder-defined k-means++ seeding over a stratified sample of at most 256
 * group members; deterministic counter hash.  posInGroup[k] receives the
 * chosen members' positions within the group (pick order); returns the number
 * of distinct picks (the valid-mask width). */
uint32_t orc_kmeanspp_pick(const uint8_t* desc, const uint32_t* grpIdx, uint32_t n, uint32_t k,
                           uint64_t seed, uint32_t level, uint32_t grp, uint32_t* posInGroup);
uint64_t orc_hash(uint64_t seed, uint64_t a, uint64_t b, uint64_t c);

/* Vocabulary.cpp:32-121 + kmeans.cpp:27-224: hierarchical k-means build.
 * centroidErr (optional): max over children with points and over dims of
 * |final centroid - binary64 mean of the child's points| (the error report).
 *   centroidType 0: uint8 centroids (b must be 16); nodes in reference layout.
 *   centroidType 1: float centroids [node][b][128]; b = 16 or 32.
 *   seedMode 0: reference randPick with a fresh std::mt19937_64(seed)
 *               (seed 5489 = the reference's default-constructed engine).
 *   seedMode 1: k-means++ (orc_kmeanspp_pick).
 *   maxIters:   Lloyd pass cap (reference 1000).
 * Outputs: masks[node] (low bits valid), weights[node][b], occ[node][b],
 * lloydPasses[node] (optional, may be NULL).  Returns 0, or -1 on the
 * reference's input errors (nbDesc < b, nbDoc == 0) or bad arguments. */
int orc_build(const uint8_t* desc, uint32_t nbDesc, uint32_t nbDoc, uint32_t b, uint32_t nbLevels,
              uint32_t maxIters, int seedMode, uint64_t seed, int centroidType, void* nodesOut,
              uint32_t* masks, float* weights, uint32_t* occ, uint32_t* lloydPasses, double* centroidErr);

/* Vocabulary.cpp:103-108 IDF weights from occurrence counts:
 * occ == 0 ? FLT_MIN : max(FLT_MIN, float(log(double(float(nbDoc) / float(occ))))). */
void orc_idf_weights(const uint32_t* occ, uint64_t n, uint32_t nbDoc, float* weights);
/* Error report along lookup paths on a float tree: for descriptor i and level
 * l, the child on leaf[i]'s path: d32[i][l] = orc_dist_f32, d64[i][l] = the
 * binary64 sum of squared differences. */
void orc_path_dist_f32(const float* nodes, uint32_t b, uint32_t nbLevels, const uint8_t* desc, uint64_t nbDesc,
                       const uint32_t* leaf, float* d32, double* d64, int nthreads);

/* DataBase.cpp:39-197 inverted-file database. lg2b = log2(branch factor). */
void* orc_db_new(uint32_t b, uint32_t nbLevels, const float* weights);
void orc_db_free(void* db);
/* returns 0, or -1 for a duplicate document (DataBase.cpp:41-43) */
int orc_db_add(void* db, uint32_t doc, const uint32_t* leaf, uint32_t n);
float orc_db_rcp(void* db, uint32_t doc);
uint64_t orc_db_nb_postings(void* db);
/* copies the sorted posting list of one leaf; returns its length */
uint32_t orc_db_postings(void* db, uint32_t leaf, uint32_t* out, uint32_t cap);
/* DataBase.cpp:71-190, reference float/double semantics with FTZ/DAZ emulation.
 * Writes up to K (doc, score) pairs ranked by (score, doc) ascending; returns
 * the count.  nbTouched (optional) receives the number of scored docs. */
uint32_t orc_db_query(void* db, const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs,
                      float* scores, uint32_t* nbTouched);
/* Same node selection, all arithmetic in double with no flushing: the FP64
 * ideal used to report score error.  Ranked by (score, doc). */
uint32_t orc_db_query_fp64(void* db, const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs,
                           double* scores);
/* Score of every touched doc (reference semantics), written densely into
 * scoresOut[nbDocsMax] (untouched docs get NaN); returns #touched. */
uint32_t orc_db_query_all(void* db, const uint32_t* leaf, uint32_t n, uint32_t K,
                          float* scoresOut, double* scores64Out, uint32_t nbDocsMax);

/* Re-creations of the reference tests' inputs (std::mt19937_64 default seed):
 * tests/KmeansAssignCenterTests.cpp:36-73 (tiles as u32 [tiles][32][32],
 * centers u32 [32][16], prevAssign u8 [nbDesc]). */
void orc_ref_assign_test_inputs(uint32_t nbDesc, uint32_t* tiles, uint32_t* centers,
                                uint8_t* prevAssign);
/* tests/ReformatTests.cpp:40-73: desc u8 [nbDesc][128], idx u32 [nbTiles][tileSize]. */
void orc_ref_reformat_test_inputs(uint32_t nbDesc, uint32_t nbTiles, uint32_t tileSize,
                                  uint8_t* desc, uint32_t* idx);
/* tests/VocTests.cpp:38-50 (which=0, 4096 descs) and :86-100 (which=1, 32 descs). */
void orc_ref_voc_test_inputs(int which, uint8_t* desc);

#ifdef __cplusplus
}
#endif
