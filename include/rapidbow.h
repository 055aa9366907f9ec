/*
 * rapidbow.h -- C ABI of the MI355X (gfx950) RapidBoW hot path.
 *
 * Drop-in boundary for the reference's public surface (RapidBoW.h:29-109):
 * vocabulary build, descriptor -> leaf lookup, serialisation, and the
 * inverted-file database.  Plain pointers and sizes only; every function
 * returns an rbow_status and never throws; rbow_last_error() returns the
 * calling thread's last error message.  Device pointers are HIP device memory
 * (hipMalloc); "host" pointers are ordinary host memory.
 *
 * The C++ facade include/RapidBoW.h rebuilds the reference's C++ names,
 * ownership and exception messages on top of these functions.
 */
#ifndef RAPIDBOW_AMD_RAPIDBOW_H
#define RAPIDBOW_AMD_RAPIDBOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* rbow_stream; /* == hipStream_t; NULL = default stream */

typedef enum rbow_status {
    RBOW_OK = 0,
    RBOW_ERR_INVALID_ARG = 1,
    RBOW_ERR_UNSUPPORTED = 2,          /* RapidBoW.cpp:37-40 (reference returns nullptr) */
    RBOW_ERR_INSUFFICIENT_DESC = 3,    /* Vocabulary.cpp:39-41 "Insufficient descriptors" */
    RBOW_ERR_NB_DOC = 4,               /* Vocabulary.cpp:42-44 "nbDoc must be positive" */
    RBOW_ERR_INVALID_BLOB = 5,         /* Vocabulary.cpp:129-131,143-145 "invalid blob" */
    RBOW_ERR_DUPLICATE_DOC = 6,        /* DataBase.cpp:41-43 "Duplicate document index" */
    RBOW_ERR_HIP = 7,                  /* replaces the reference's cudaCheck aborts */
    RBOW_ERR_OUT_OF_MEMORY = 8,
    RBOW_ERR_CAPACITY = 9              /* a fixed limit of this implementation was exceeded */
} rbow_status;

/* RapidBoW.h:31-48 */
typedef enum { RBOW_DATA_BINARY = 0, RBOW_DATA_UINT8 = 1 } rbow_data_type;
typedef enum { RBOW_DIST_HAMMING = 0, RBOW_DIST_L1 = 1, RBOW_DIST_L2 = 2 } rbow_distance_type;
typedef struct {
    int32_t data_type;     /* rbow_data_type */
    int32_t distance_type; /* rbow_distance_type */
    uint64_t nb_dims;
} rbow_descriptor_attributes;

typedef struct rbow_vocabulary rbow_vocabulary;
typedef struct rbow_database rbow_database;

/* Thread-local message of the last failing call on this thread ("" if none). */
const char* rbow_last_error(void);

/* ---------------------------------------------------------------- build -- */

/* Options of the hierarchical k-means build.  rbow_build_options_default()
 * gives the reference's behaviour (kmeans.h:112-118, kmeans.cpp:141-196). */
typedef struct {
    uint32_t max_iters;     /* Lloyd pass cap per node (reference: 1000) */
    uint32_t seed_mode;     /* 0: reference random picks; 1: k-means++ over <=256 sampled members */
    uint64_t seed;          /* engine seed; 5489 = std::mt19937_64's default (the reference) */
    uint32_t centroid_type; /* 0: uint8 centroids (reference, b = 16); 1: float32 (b = 16 or 32) */
    uint32_t kernel;        /* 0: MFMA kernels; 1: plain scalar kernels (same results) */
    uint32_t persist_max_tiles; /* MFMA kernels, uint8 centroids: a node of at most this many 64-point
                                 * tiles runs its whole Lloyd loop in one launch with its points in
                                 * LDS (clamped to what one workgroup's LDS holds); 0: every node
                                 * runs one launch per pass.  Same results either way.  Default 16. */
    uint32_t timing;        /* 1: time every assign, reduce and update launch for rbow_voc_build_stats
                             * with HIP events stamped by the dispatches (kernel start / end);
                             * 0 (default): total_ms, point and launch counts only */
} rbow_build_options;

void rbow_build_options_default(rbow_build_options* opts);

/* RapidBoW.h:73-88 impl::buildVocabularyImpl / rbow::buildVocabulary.
 * dev_desc: device uint8[nb_desc][128].  Blocks until the tree is built (the
 * reference synchronises internally).  Unsupported attributes or branch
 * factor -> RBOW_ERR_UNSUPPORTED (the reference returns nullptr). */
rbow_status rbow_build_vocabulary(rbow_descriptor_attributes attr, uint32_t branch_factor, const void* dev_desc,
                                  uint32_t nb_desc, uint32_t nb_doc, uint32_t nb_levels, rbow_stream stream,
                                  rbow_vocabulary** out);
/* Same with explicit options (NULL = defaults).  Extends the reference with
 * branch factor 32 and float32 centroids (centroid_type 1). */
rbow_status rbow_build_vocabulary_ex(rbow_descriptor_attributes attr, uint32_t branch_factor, const void* dev_desc,
                                     uint32_t nb_desc, uint32_t nb_doc, uint32_t nb_levels,
                                     const rbow_build_options* opts, rbow_stream stream, rbow_vocabulary** out);
/* Autotuner (no reference counterpart): builds the tree of these arguments with
 * each size limit of the persistent on-chip Lloyd loop, persist_max_tiles = 0,
 * 4, 8, 12 and 16 tiles of 64 points (the other options as in *opts), and
 * writes the fastest into opts->persist_max_tiles (the better of two builds
 * each; out_ms, optional, receives the five times in that order).  Every
 * candidate must build the same tree (RBOW_ERR_HIP otherwise).  Acts on uint8
 * centroids with the MFMA kernels only; other options are left unchanged.
 * Blocks. */
rbow_status rbow_tune_build_options(rbow_descriptor_attributes attr, uint32_t branch_factor, const void* dev_desc,
                                    uint32_t nb_desc, uint32_t nb_doc, uint32_t nb_levels, rbow_build_options* opts,
                                    float* out_ms, rbow_stream stream);
/* No counterpart in the reference.  A completed build leaves its device and
 * pinned work buffers (about 290 MB + 150 B per descriptor for b16 L4 MFMA
 * builds) with the current device for the next build, which then allocates only
 * the tree.  This frees them on every device; the next build allocates afresh.
 * Not to be called while a build runs. */
rbow_status rbow_release_build_workspaces(void);
void rbow_voc_destroy(rbow_vocabulary* voc);

/* ------------------------------------------------------ IVocabulary ------ */
/* RapidBoW.h:60-71 */
rbow_status rbow_voc_get_attributes(const rbow_vocabulary* voc, rbow_descriptor_attributes* attr);
uint32_t rbow_voc_branch_factor(const rbow_vocabulary* voc);
uint32_t rbow_voc_nb_levels(const rbow_vocabulary* voc);
/* 0: uint8 centroids, 1: float32 centroids */
uint32_t rbow_voc_centroid_type(const rbow_vocabulary* voc);

/* RapidBoW.h:66-68 IVocabulary::lookUp: asynchronous on `stream`; dev_desc is
 * device uint8[nb_desc][128], dev_leaf device uint32[nb_desc].  nb_desc == 0
 * is a no-op (the reference launches an empty grid and errors). */
rbow_status rbow_voc_lookup(const rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc, uint32_t* dev_leaf,
                            rbow_stream stream);

typedef enum {
    RBOW_LOOKUP_AUTO = 0,  /* plain below 12,288 descriptors (uint8 trees) / 49,152 (float b16) / 32,768 (float b32), else MFMA */
    RBOW_LOOKUP_PLAIN = 1, /* plain scalar per-descriptor descent */
    RBOW_LOOKUP_MFMA = 2   /* MFMA form: fused single launch for small batches on uint8 b16 trees,
                            * else the level-synchronous bucketed descent */
} rbow_lookup_kernel;
/* Lookup with an explicit kernel.  near_ties (optional, device uint32[1],
 * accumulated, not cleared) counts (descriptor, level) pairs whose MFMA
 * candidate set held more than one child and was re-scored in FP32. */
rbow_status rbow_voc_lookup_ex(const rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc,
                               uint32_t* dev_leaf, int32_t kernel, uint32_t* dev_near_ties, rbow_stream stream);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same lookup as rbow_voc_lookup_ex, same leaf ids, with no
 * allocation of any kind.  The MFMA form's workspace is a static 64 MiB
 * device array instead of a stream-ordered allocation: each GPU the library
 * runs on holds its own copy until the process ends.  A batch whose workspace
 * does not fit it returns RBOW_ERR_CAPACITY before anything is launched (about
 * 2 million descriptors per call on a b32 L4 tree, 4 million on b16 L4; the
 * plain form needs none).  Every call on one GPU shares that array, so calls
 * that overlap in time on one GPU race: issue them on one stream. */
rbow_status rbow_voc_lookup_noalloc(const rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc,
                                    uint32_t* dev_leaf, int32_t kernel, uint32_t* dev_near_ties, rbow_stream stream);
/* Autotuner (no reference counterpart), on prefixes of dev_desc (powers of two
 * from 4096 up to nb_desc): on uint8 b16 trees it first sets the MFMA form's
 * fused/bucketed boundary (rbow_voc_lookup_fused_max) to the largest size at
 * which the fused single-launch descent is at least as fast as the bucketed
 * one; then it sets the vocabulary's AUTO crossover to the smallest size at
 * which the MFMA form is at least as fast as the plain kernel (nb_desc + 1 if
 * never); with at least 65,536 descriptors it first picks the bucketed
 * descent's leaf item size (rbow_voc_lookup_leaf_item_rows).  Blocks; does not
 * change any leaf id.  rbow_voc_set_lookup_crossover
 * and rbow_voc_set_lookup_fused_max set them directly (0 restores the defaults). */
rbow_status rbow_voc_tune_lookup(rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc,
                                 uint32_t* out_crossover, rbow_stream stream);
rbow_status rbow_voc_set_lookup_crossover(rbow_vocabulary* voc, uint32_t crossover);
uint32_t rbow_voc_lookup_crossover(const rbow_vocabulary* voc); /* current AUTO crossover */
/* uint8 b16 trees: the MFMA form runs a fused single-launch descent for batches
 * of at most fused_max descriptors and the level-synchronous bucketed descent
 * above (same leaf ids); 0 restores the default (196,608); 0 on other trees. */
rbow_status rbow_voc_set_lookup_fused_max(rbow_vocabulary* voc, uint32_t fused_max);
uint32_t rbow_voc_lookup_fused_max(const rbow_vocabulary* voc);
/* (No reference counterpart.)  Tile shape of the bucketed MFMA descent: rows per
 * work item at the leaf level, 256 or 512 (the levels above use 256).  The
 * default is 512 on float trees, 256 on uint8 trees; rbow_voc_tune_lookup
 * measures both on the caller's descriptors (batches of at least 65,536) and
 * keeps the faster.  Leaf ids do not depend on it; 0 restores the default. */
rbow_status rbow_voc_set_lookup_leaf_item_rows(rbow_vocabulary* voc, uint32_t rows);
uint32_t rbow_voc_lookup_leaf_item_rows(const rbow_vocabulary* voc);

/* RapidBoW.h:69 IVocabulary::serialize.  Two-call form: blob may be NULL to
 * query *size.  uint8/b16 trees use the reference blob byte layout
 * (Vocabulary.cpp:175-197); float32 or b32 trees a versioned "RBW2" blob. */
rbow_status rbow_voc_serialize(const rbow_vocabulary* voc, uint8_t* blob, size_t capacity, size_t* size);
/* RapidBoW.h:90 deserializeVocabulary (Vocabulary.cpp:126-157). */
rbow_status rbow_voc_deserialize(const uint8_t* blob, size_t size, rbow_vocabulary** out);
/* RapidBoWInternal.h:29-30 getLevelWeights: host float[levelSize(level)][b]. */
const float* rbow_voc_level_weights(const rbow_vocabulary* voc, uint32_t level);
/* Host copy of one level's centroids as float32 [levelSize(level)][b][128]
 * (uint8 trees are widened); for tests and tools. */
rbow_status rbow_voc_copy_centroids(const rbow_vocabulary* voc, uint32_t level, float* host_out, size_t capacity_floats);

/* ---------------------------------------------------------- IDataBase --- */
/* RapidBoW.h:107 createDataBase: borrows the vocabulary, which must outlive
 * the database.  Postings, document norms and IDF weights live in HBM. */
rbow_status rbow_db_create(const rbow_vocabulary* voc, rbow_database** out);
void rbow_db_destroy(rbow_database* db);
/* RapidBoW.h:97 IDataBase::addDoc (DataBase.cpp:39-69); host leaf ids.
 * RBOW_ERR_DUPLICATE_DOC "Duplicate document index" for a known doc id. */
rbow_status rbow_db_add_doc(rbow_database* db, uint32_t doc, const uint32_t* host_leaf, uint32_t n);
/* RapidBoW.h:99-100 IDataBase::query (DataBase.cpp:71-190); host leaf ids.
 * Writes up to max_results (<= 1024) doc ids ranked by (score, doc id)
 * ascending and optionally their scores; *out_count = number written.
 * Queries hold a shared lock; at most 16384 leaf ids per query. */
rbow_status rbow_db_query(const rbow_database* db, const uint32_t* host_leaf, uint32_t n, uint32_t max_results,
                          uint32_t* out_docs, float* out_scores, uint32_t* out_count);
/* RapidBoW.h:102-103 IDataBase::queryAndAddDoc (DataBase.cpp:192-197). */
rbow_status rbow_db_query_and_add_doc(rbow_database* db, uint32_t doc, const uint32_t* host_leaf, uint32_t n,
                                      uint32_t max_results, uint32_t* out_docs, uint32_t* out_count);
/* Batched forms (no counterpart in the reference, which has one host query
 * at a time): doc host_doc_ids[i] owns dev_leaf[host_offsets[i] ..
 * host_offsets[i+1]) (offsets are indices into dev_leaf).  Every device leaf
 * id must be below b^L: the batch insert checks them all on the device before
 * it changes anything, and a query batch holding one out of range returns
 * RBOW_ERR_INVALID_ARG "leaf id out of range" (its other queries' results are
 * still written). */
rbow_status rbow_db_add_docs_device(rbow_database* db, const uint32_t* host_doc_ids, uint32_t nb_docs,
                                    const uint32_t* dev_leaf, const uint64_t* host_offsets, rbow_stream stream);
/* Scores nb_queries queries in one pass: results dev_docs[q][max_results],
 * dev_scores[q][max_results] (optional), dev_counts[q].  Returns when done. */
rbow_status rbow_db_query_batch(const rbow_database* db, const uint32_t* dev_leaf, const uint64_t* host_offsets,
                                uint32_t nb_queries, uint32_t max_results, uint32_t* dev_docs, float* dev_scores,
                                uint32_t* dev_counts, rbow_stream stream);
/* (No reference counterpart.)  Scoring tile shape: steps of 64 postings each
 * wave walks per iteration (its loads in flight per lane), 4 or 6 (0: the
 * default, 6).  rbow_db_tune_scoring times one query batch (arguments as
 * rbow_db_query_batch) with each, checks that they rank and score alike
 * (scores within 2e-6), keeps the faster for the database and returns it with
 * the two median times (out_ms[2], optional).  Neither changes a result. */
rbow_status rbow_db_set_walk_steps(rbow_database* db, uint32_t steps);
uint32_t rbow_db_walk_steps(const rbow_database* db);
rbow_status rbow_db_tune_scoring(rbow_database* db, const uint32_t* dev_leaf, const uint64_t* host_offsets,
                                 uint32_t nb_queries, uint32_t max_results, uint32_t* out_steps, float* out_ms,
                                 rbow_stream stream);
uint32_t rbow_db_nb_docs(const rbow_database* db);
uint64_t rbow_db_nb_postings(const rbow_database* db);
/* Incremental inserts (no counterpart in the reference, whose addDoc inserts
 * into sorted per-leaf vectors, DataBase.cpp:44-53): recent documents' postings
 * live in a delta segment beside the main CSR and are scored from both; the
 * delta is folded into the main CSR once it holds more than 2^24 postings.
 * rbow_db_nb_delta_postings = postings not yet in the main CSR;
 * rbow_db_compact folds them in now (blocks). */
uint64_t rbow_db_nb_delta_postings(const rbow_database* db);
rbow_status rbow_db_compact(rbow_database* db, rbow_stream stream);
/* Posting list of one leaf as sorted doc ids (duplicates kept), the
 * reference's InvertedFile (DataBase.h:58-59); *length = full length. */
rbow_status rbow_db_postings(const rbow_database* db, uint32_t leaf, uint32_t* host_out, uint32_t capacity,
                             uint32_t* length);
/* float(1 / L1 norm) of a document (DataBase.cpp:68). */
rbow_status rbow_db_doc_rcp(const rbow_database* db, uint32_t doc, float* rcp);

/* ------------------------------------------------------------- server --- */
/* No counterpart in the reference (one host query at a time): a continuous-
 * batching query server over a resident vocabulary + database.  Requests are
 * admitted into a batch while batch slots and the HBM scratch budget allow;
 * a worker thread runs each batch (H2D, lookup, scoring, top-K, D2H) on one of
 * nb_streams HIP streams, so batches formed from later arrivals run while
 * earlier ones are on the device.  Results do not depend on batching.  The
 * database must not receive inserts while a server exists. */
typedef struct rbow_server rbow_server;
typedef struct {
    uint32_t max_batch;    /* queries per batch (default 64) */
    uint32_t nb_streams;   /* batches in flight (default 2) */
    uint64_t mem_budget;   /* HBM scratch of admitted queries, bytes (default 4 GiB) */
    uint32_t max_results;  /* K (default 50, <= 1024) */
    int32_t lookup_kernel; /* RBOW_LOOKUP_AUTO (default), RBOW_LOOKUP_MFMA or RBOW_LOOKUP_PLAIN, per batch */
    uint64_t staging_bytes; /* pinned host ring that submit copies descriptors into (default 256 MiB,
                             * at least 2 MiB); a submit waits while it is full of queries in flight */
    uint32_t use_graphs;    /* 1 (default): each batch is one launch of a HIP graph captured per size
                             * class (lookup, scoring, top-K, result copy); 0: direct launches */
    uint32_t reserved;
    uint64_t batch_max_desc; /* descriptors per batch (0 = default min(max_batch * 16384, 2^20)); every
                              * stream's device buffers are sized for it once, at creation */
} rbow_server_options;
void rbow_server_options_default(rbow_server_options* opts);
rbow_status rbow_server_create(const rbow_vocabulary* voc, const rbow_database* db, const rbow_server_options* opts,
                               rbow_server** out);
void rbow_server_destroy(rbow_server* srv);
/* Host descriptors uint8[n][128] (copied); returns a ticket. */
rbow_status rbow_server_submit(rbow_server* srv, const uint8_t* host_desc, uint32_t n, uint64_t* ticket);
/* Blocks until the ticket completes: ranked docs / scores (K each), count,
 * latency from submit to completion. */
rbow_status rbow_server_wait(rbow_server* srv, uint64_t ticket, uint32_t* docs, float* scores, uint32_t* count,
                             double* latency_ms);
/* Load generator: submits query i (host_desc rows offsets[i]..offsets[i+1])
 * at Poisson arrival times of rate qps (seeded), then collects every result:
 * latency_ms[i], docs[i][K], scores[i][K] and counts[i] (optional),
 * batch_sizes[i] (optional), elapsed_s = first arrival to last completion. */
rbow_status rbow_server_replay(rbow_server* srv, const uint8_t* host_desc, const uint64_t* offsets,
                               uint32_t nb_queries, double qps, uint64_t seed, double* latency_ms, uint32_t* docs,
                               float* scores, uint32_t* counts, uint32_t* batch_sizes, double* elapsed_s);

/* ---------------------------------------------------------- pipeline --- */
/* No counterpart in the reference (Vocabulary.cpp:162-173 lookUp, then
 * DataBase.cpp:71-190 query, per image): a batch of query images whose
 * descriptors are already in HBM, through lookup, scoring and top-K, in device
 * buffers sized once (max_queries images, max_desc descriptors).  With
 * use_graph the batch is ONE launch of a HIP graph captured for its size class
 * (descriptors rounded up to a quarter octave, images to a power of two);
 * graphs are kept and reused, and recaptured after the database changes.
 * Results as rbow_db_query_batch; returns when done.  A pipeline runs one
 * batch at a time; the database must not change while a run is in progress. */
typedef struct rbow_query_pipeline rbow_query_pipeline;
rbow_status rbow_pipeline_create(const rbow_vocabulary* voc, const rbow_database* db, uint32_t max_queries,
                                 uint64_t max_desc, uint32_t max_results, rbow_query_pipeline** out);
void rbow_pipeline_destroy(rbow_query_pipeline* p);
/* image q owns dev_desc rows host_offsets[q] .. host_offsets[q+1] */
rbow_status rbow_pipeline_run(rbow_query_pipeline* p, const void* dev_desc, const uint64_t* host_offsets,
                              uint32_t nb_queries, int32_t lookup_kernel, int32_t use_graph, uint32_t* dev_docs,
                              float* dev_scores, uint32_t* dev_counts, rbow_stream stream);
/* graphs captured so far by this pipeline */
uint32_t rbow_pipeline_graphs(const rbow_query_pipeline* p);

/* ------------------------------------------------ kernel-level entries -- */
/* These expose the batched k-means kernels the build drives, one pass at a
 * time, for callers that run their own Lloyd loop (and for parity tests).
 * Points are device uint8 rows [sum(task_cnt)][128], tasks contiguous in
 * order; task_cnt is a HOST array of nb_tasks sizes.  b = 16 (uint8 centers
 * [task][16][128]) or b = 16/32 with float centers [task][b][128]. */

/* kmeans.cuh:102 launchAssignCenter (kmeans.cu:226-331): one assign pass.
 * dev_assign [rows] holds the previous assignment and receives the new one;
 * dev_counts [task][b], dev_sums [task][b][128] (int64) and dev_changed
 * [task] are ACCUMULATED into (zero them first, as the reference does).
 * kernel: 0 = I8 MFMA (uint8 centers only), 1 = plain scalar. */
rbow_status rbow_kmeans_assign(const uint8_t* dev_points, const uint32_t* task_cnt, uint32_t nb_tasks, uint32_t b,
                               uint32_t centroid_type, const void* dev_centers, uint8_t* dev_assign,
                               uint32_t* dev_counts, int64_t* dev_sums, uint32_t* dev_changed, int32_t kernel,
                               rbow_stream stream);
/* kmeans.cuh:103 launchKMeansUpdateCenters (kmeans.cu:334-384): for tasks
 * with dev_changed != 0, every center with count > 0 becomes
 * round(sum/count) (uint8) or sum/count (float); then the task's counts, sums
 * and changed are cleared for the next pass. */
rbow_status rbow_kmeans_update(const uint32_t* task_cnt, uint32_t nb_tasks, uint32_t b, uint32_t centroid_type,
                               void* dev_centers, uint32_t* dev_counts, int64_t* dev_sums, uint32_t* dev_changed,
                               rbow_stream stream);
/* reformatDesc.cu:78-96 launchReformatDesc / launchKMeansInitCenters:
 * dev_tiles[t][word][slot] = word of desc[idx[t][slot]] (32 words per
 * descriptor, tile_size 32 or 16); idx 0xFFFFFFFF -> 0xFFFFFFFF. */
rbow_status rbow_reformat_tiles(const void* dev_desc, const uint32_t* dev_idx, uint32_t nb_tiles, uint32_t tile_size,
                                uint32_t* dev_tiles, rbow_stream stream);
/* Vocabulary.cpp:103-108 IDF weights, as the build computes them on the device:
 * dev_weights[i] = dev_occ[i] == 0 ? FLT_MIN
 *                : max(FLT_MIN, float(log(double(float(nb_doc) / float(dev_occ[i])))))
 * (binary32 quotient, binary64 logarithm).  Asynchronous on stream. */
rbow_status rbow_idf_weights(const uint32_t* dev_occ, uint64_t n, uint32_t nb_doc, float* dev_weights,
                             rbow_stream stream);
/* Row gather used by the build: dev_out[p] = desc[idx[p]] (invalid -> 0xFF). */
rbow_status rbow_gather_rows(const void* dev_desc, const uint32_t* dev_idx, uint64_t n, uint8_t* dev_out,
                             rbow_stream stream);
/* Lloyd passes run per node by the last build of this vocabulary (host
 * uint32[nodes]); NULL when the vocabulary was deserialised. */
const uint32_t* rbow_voc_build_passes(const rbow_vocabulary* voc);

/* Timing of the last build (zeros for deserialised vocabularies).
 * assign_ops = point_passes * 2 * b * 128 (the distance contraction).  The
 * *_ms fields other than total_ms need rbow_build_options.timing = 1 (else 0). */
typedef struct {
    double total_ms;         /* host wall time of the build */
    double assign_ms;        /* summed assign-kernel time (HIP events around each assign launch) */
    double update_ms;        /* summed update-kernel time (HIP events) */
    uint64_t point_passes;   /* sum over nodes of points x Lloyd passes */
    uint64_t assign_ops;
    double reduce_ms;        /* summed time of the chunk-partial reductions (MFMA path) */
    uint32_t assign_launches; /* per-pass (streamed) MFMA assign launches */
    uint32_t reserved;
    double stream_ms;        /* time of those launches */
    uint64_t stream_point_passes; /* point-passes they processed */
    double persist_ms;       /* persistent launches of small tasks' whole Lloyd loops (in assign_ms) */
} rbow_build_stats;
rbow_status rbow_voc_build_stats(const rbow_vocabulary* voc, rbow_build_stats* out);

/* ------------------------------------------------ synthetic data -------- */
/* No counterpart in the reference: seeded SIFT-like data for benchmarks.
 * nb_centres Gaussian centres c ~ N(0, I_128) from centre_seed; descriptor i
 * (global index first+i) takes centre centre_ids[i] (or a hashed uniform pick
 * when centre_ids is NULL), x = |c + sigma n| with sigma = sigma_scale*||c||/sqrt(128),
 * L2-normalise, clip at 0.2, renormalise, x512, round, saturate to 255.
 * Host memory in and out; uses all OpenMP threads. */
void rbow_synth_centres(uint64_t centre_seed, uint32_t nb_centres, float* out);
void rbow_synth_descriptors(uint64_t centre_seed, uint32_t nb_centres, uint64_t seed, uint64_t first, uint64_t n,
                            const uint32_t* centre_ids, float sigma_scale, uint8_t* out);
/* Device forms (same distribution, generated in HBM; not bit-equal to the
 * host form).  dev_sigma[c] = sigma_scale * ||c|| / sqrt(128). */
rbow_status rbow_synth_descriptors_device(const float* dev_centres, const float* dev_sigma, uint32_t nb_centres,
                                          uint64_t seed, uint64_t first, uint64_t n, const uint32_t* dev_centre_ids,
                                          uint8_t* dev_out, rbow_stream stream);
/* Image i of [first_image, first_image + nb_images) owns descriptors
 * dev_offsets[i] .. dev_offsets[i+1]; each takes a centre from the image's
 * subset of subset_size centres (dev_ids[nb_desc]). */
rbow_status rbow_synth_image_centres_device(const uint64_t* dev_offsets, uint32_t nb_images, uint64_t first_image,
                                            uint64_t image_seed, uint64_t desc_seed, uint32_t nb_centres,
                                            uint32_t subset_size, uint32_t* dev_ids, uint64_t nb_desc,
                                            rbow_stream stream);
/* Posting lists without descriptors: posting j of image i is a leaf drawn
 * from dev_cloud[c][cloud_size], the looked-up leaves of cloud_size samples
 * around centre c, with c taken from the image's centre subset. */
rbow_status rbow_synth_image_postings_device(const uint64_t* dev_offsets, uint32_t nb_images, uint64_t first_image,
                                             uint64_t image_seed, uint64_t post_seed, uint32_t nb_centres,
                                             uint32_t subset_size, const uint32_t* dev_cloud, uint32_t cloud_size,
                                             uint32_t* dev_leaf, uint64_t nb_postings, rbow_stream stream);

#ifdef __cplusplus
}
#endif
#endif
