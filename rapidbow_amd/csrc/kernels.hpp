// kernels.hpp -- host-side launchers of the gfx950 kernels.
#pragma once
#include <vector>

#include "common.hpp"
#include "kmpartition.hpp"

namespace rbow_amd {

// ---------------------------------------------------------------- trees --
// Device-resident vocabulary tree in the library's own layout:
//   rows      [node][b][128]  centroid rows, child-major (uint8 or float32)
//   u8 trees: normU[node*b+j]  = ||c||^2            (plain dot4 form)
//             normS[node*b+j]  = ||c - 128||^2      (I8 MFMA form, offset operands)
//   f32 trees: split[t][node*b+j][128], t = 0..2: int8 terms of c - 128 =
//               A0 + A1 2^-7 + A2 2^-14 + residual (splitStride bytes per term)
//             cn2[node*b+j]   = ||c - 128||^2 (fp32 of the binary64 value)
//             e1[node*b+j]    = certified bound coefficient of ||x - 128|| (see lookup.hip)
//             e2[node*b+j]    = certified bound constant term
struct TreeView {
    uint32_t b = 16, lg2b = 4, L = 0, centroidType = 0;
    const uint8_t* u8rows = nullptr;
    const int32_t* normU = nullptr;
    const int32_t* normS = nullptr;
    const float* f32rows = nullptr;
    const int8_t* split = nullptr;
    uint64_t splitStride = 0;
    const float* cn2 = nullptr;
    const float* e1 = nullptr;
    const float* e2 = nullptr;
    uint32_t autoMin = 0;  // AUTO lookup: batches below this use the plain kernel (0: built-in default)
    uint32_t fusedMax = 0; // uint8 b16 trees: MFMA-form batches of at most this many descriptors run the
                           // fused single-launch descent, larger ones the bucketed one (0: built-in default)
    uint32_t leafItemRows = 0;  // bucketed descent: rows per work item at the leaf level, 256 or 512
                                // (0: built-in default, lookupLeafItemRowsDefault)
};
// Default leaf-level item size of the bucketed descent: 512 rows on float trees
// (a leaf node's split centroids, 12 KB at b = 32, are then read once rather
// than once per item), 256 on uint8 trees.
uint32_t lookupLeafItemRowsDefault(const TreeView& t);

// Derived per-child data for nodes [0, nbNodes): uint8 trees' norms; for float
// trees (lookups, and the float k-means assign every pass) the int8 split of
// c - 128 and its bound coefficients (split: 3 terms of splitStride bytes;
// nodes whose flag in `active`, optional, is zero are skipped).
void launchPrepU8Tree(const uint8_t* rows, uint64_t nbChildren, int32_t* normU, int32_t* normS, hipStream_t s);
void launchPrepF32Split(const float* rows, uint64_t nbNodes, uint32_t b, int8_t* split, uint64_t splitStride,
                        float* cn2, float* e1, float* e2, hipStream_t s, const uint32_t* active = nullptr);
// Children equal (bit for bit) to a lower sibling get cn2 = +inf: never a
// candidate of the float lookup's filter (the lowest of equal children wins).
void launchMarkDuplicateChildren(const float* rows, uint64_t nbNodes, uint32_t b, float* cn2, hipStream_t s);

// Plain scalar per-descriptor descent (vocTreeLookUp.cu:48-160 semantics).
void launchLookupPlain(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, hipStream_t s);
// MFMA-form descent: fused single launch (uint8 b16 trees, small batches) or
// level-synchronous bucketed (I8 for uint8 trees, I8 on the int8 split with
// certified FP32 re-score for float trees).  Allocates its workspace stream-ordered.
void launchLookupMfma(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                      hipStream_t s);
// Kernel choice (RBOW_LOOKUP_*): AUTO takes the plain kernel below the batch
// size where the bucketed MFMA form's per-level passes start to pay off
// (t.autoMin when tuned, else a default measured on MI355X).
// AUTO crossover of a tree kind when no tuned value is set (measured on MI355X)
uint32_t lookupAutoDefault(uint32_t centroidType, uint32_t b);
// Whether an MFMA-form lookup of n descriptors runs the fused single-launch
// descent (uint8 b16 trees, n <= t.fusedMax or its default), and that default.
bool lookupUsesFused(const TreeView& t, uint32_t n);
uint32_t lookupFusedDefault(const TreeView& t);
void launchLookup(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies, int kernel,
                  hipStream_t s);
// The same lookup without any allocation: the MFMA form's workspace is a static
// device array of lookupScratchBytes() per GPU, shared by every call on that GPU
// (calls that overlap in time on one GPU race).  Batches whose workspace does
// not fit it are rejected with RBOW_ERR_CAPACITY before anything is launched.
void launchLookupNoAlloc(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                         int kernel, hipStream_t s);
size_t lookupScratchBytes();
// The kernel a lookup of n descriptors runs (AUTO resolved), the workspace that
// form needs (0 for the plain one), and the lookup in a caller's workspace
// (256-byte aligned), allocating nothing.
bool lookupUsesPlain(const TreeView& t, uint32_t n, int kernel);
size_t lookupWorkspaceBytes(const TreeView& t, uint32_t n, bool plain, hipStream_t s);
void launchLookupInWs(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                      bool plain, void* ws, size_t wsBytes, hipStream_t s);
// Autotuner: on prefixes of desc (powers of two from 4096 up to n), first (uint8
// b16 trees) the largest size at which the fused MFMA descent is at least as fast
// as the bucketed one, then, with that MFMA form, the smallest size at which it
// is at least as fast as the plain kernel.  Blocks; uses `leaf` (n words) as output.
struct LookupTuning {
    uint32_t crossover;     // AUTO: the MFMA form from this many descriptors on (n + 1: never)
    uint32_t fusedMax;      // uint8 b16 trees: the fused descent up to this many descriptors (0: not applicable)
    uint32_t leafItemRows;  // bucketed descent: leaf-level item size, 256 or 512 (0: not measured)
};
LookupTuning tuneLookup(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, hipStream_t s);

// ----------------------------------------------------------------- build --
// Gather rows: out[p] = desc[idx[p]] (idx == kIdxInvalid -> 0xFF bytes).
// (ids >= nbSrc, like kIdxInvalid, give 0xFF bytes)
void launchGatherRows(const uint8_t* desc, const uint32_t* idx, uint64_t n, uint8_t* out, hipStream_t s,
                      uint64_t nbSrc = ~0ull);
// reformatDesc.cu:38-74 tile layout: tiles[t][word][slot] (tileSize 32 or 16).
void launchReformatTiles(const uint8_t* desc, const uint32_t* idx, uint32_t nbTiles, uint32_t tileSize,
                         uint32_t* tiles, hipStream_t s);

struct KmTile {        // one work item of the assign kernels: <= 64 rows of one task
    uint32_t task;
    uint32_t row;      // first row in the level's gathered array
    uint32_t cnt;      // rows in this tile (1..64)
};

// Partition of the level's streamed tiles over the resident waves of the MFMA
// assign (partitionTiles with the device's resident wave count and the
// kernel's waves per workgroup, kmAssignWavesPerBlock()).
KmPartition partitionKmTiles(const std::vector<uint32_t>& taskTiles, const std::vector<uint32_t>& tileBeg);
uint32_t kmAssignWavesPerBlock();
uint32_t kmAssignMergeSlots();  // LDS merge slots per workgroup (KmPartition::blockSlots per block)

// Partials of the MFMA assign: a task with at most kDirectParts of them is
// reduced inside the update kernel; larger ones go through launchKmReduce in
// items of at most kReduceItemParts partials.
constexpr uint32_t kDirectParts = 32;
constexpr uint32_t kReduceItemParts = 32;

struct KmLevel {       // device state of one level's batched k-means
    uint32_t b = 16, nbTasks = 0, nbTiles = 0, centroidType = 0;
    const uint8_t* X = nullptr;      // gathered rows [N][128]
    const KmTile* tiles = nullptr;
    const uint32_t* taskBeg = nullptr;
    const uint32_t* taskCnt = nullptr;
    uint8_t* centersU8 = nullptr;    // [task][16][128]
    float* centersF32 = nullptr;     // [task][b][128]
    uint8_t* assign = nullptr;       // [N]
    uint32_t* counts = nullptr;      // [task][b]
    long long* sums = nullptr;       // [task][b][128]
    uint32_t* changed = nullptr;     // [task]
    uint32_t* active = nullptr;      // [task]
    uint32_t* passes = nullptr;      // [task]
    uint32_t* nbActive = nullptr;    // [1]
    // MFMA path (uint8 centroids): per-tile swizzled operand and partial sums
    // (one per chunk, or one per run of a task's chunks merged in a workgroup)
    const int8_t* XA = nullptr;      // [tile][rb 4][kb 2][lane 64][16]  distance A fragments (x - 128)
    const KmChunk* chunks = nullptr;
    uint32_t nbChunks = 0;
    const uint32_t* waveChunkBeg = nullptr;  // [nbWaves + 1]
    uint32_t nbWaves = 0;
    const KmSlot* blockSlots = nullptr;      // [workgroup][kmAssignMergeSlots()]
    uint32_t nbParts = 0;
    const uint32_t* taskPartBeg = nullptr;   // [task]
    const uint32_t* taskPartCnt = nullptr;   // [task]
    int32_t* partSums = nullptr;     // [part][16][128]  sums of (x - 128)
    uint32_t* partCnt = nullptr;     // [part][16]
    uint32_t* partChanged = nullptr; // [part]
    // float centroids, MFMA filter: this pass's centers split into int8 terms by
    // launchPrepF32Split, with their norms and bound coefficients [task*b+j]
    const int8_t* fSplit = nullptr;  // int8 split of this pass's centers [3][task*b+j][128]
    uint64_t fSplitStride = 0;
    const float* fCn2 = nullptr;
    const float* fE1 = nullptr;
    const float* fE2 = nullptr;
    uint32_t* nearTies = nullptr;  // optional: rows re-scored in FP32
    const uint32_t* updTasks = nullptr;  // tasks the MFMA update visits (null: all); the others run elsewhere
    uint32_t nbUpdTasks = 0;
    uint32_t reverse = 0;  // MFMA assign: walk each wave's tiles backwards (odd passes)
};

// centers[task][j] = desc[seedIdx[task*b+j]] (invalid -> 0xFF / 255.0f)
void launchInitCenters(const KmLevel& k, const uint8_t* desc, const uint32_t* seedIdx, hipStream_t s);
// Builder-defined k-means++ picks: seedIdx[task*b+j], masks[task].
// desc holds nbSrc rows (order entries are row ids below it).
void launchKmeansPP(const uint8_t* desc, uint32_t nbSrc, const uint32_t* order, const uint32_t* taskBeg,
                    const uint32_t* taskCnt, uint32_t nbTasks, uint32_t b, uint64_t seed, uint32_t level,
                    uint32_t* seedIdx, uint32_t* masks, hipStream_t s);
void launchKmLevelReset(const KmLevel& k, uint64_t nbRows, hipStream_t s);
// Per-tile swizzled operand XA of the MFMA assign from desc[order[row]] (desc
// holds nbSrc rows; an order entry outside them gives a zero row).  With
// nbTilesDev, nbTiles is a bound and the device count decides.
void launchGatherTiles(const uint8_t* desc, uint32_t nbSrc, const uint32_t* order, const KmTile* tiles, uint32_t nbTiles,
                       int8_t* XA, hipStream_t s, const uint32_t* nbTilesDev = nullptr);
// Next level's tasks planned on the device from a level's histogram hist[nbTasks]
// (the child counts): taskBeg, taskCnt, tileBeg [nbTasks], *nbTiles and the
// tile list tiles[*nbTiles] (tileBound >= the tile count sizes the launch), as
// the host would derive them from the same counts; tiles == nullptr skips the
// tile list.  nbTasks <= kPlanMaxTasks.
constexpr uint32_t kPlanMaxTasks = 1u << 20;
void launchPlanLevel(const uint32_t* hist, uint32_t nbTasks, uint32_t tileBound, uint32_t* taskBeg, uint32_t* taskCnt,
                     uint32_t* tileBeg, uint32_t* nbTiles, KmTile* tiles, hipStream_t s);
// Optional timing of one launch (or of a run of launches): HIP events recorded
// by the dispatch itself when the kernel starts and when it ends, so their
// interval is the kernel's own duration.  A null event is not recorded.
struct KernelTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};

// Adds the partials of the MFMA assign into the task totals counts / sums /
// changed.  red: (task, partBeg, partEnd) items.  Returns whether a kernel ran.
bool launchKmReduce(const KmLevel& k, const KmPartRange* red, uint32_t nbRed, hipStream_t s,
                    const KernelTimer& t = {});
// Whole Lloyd loops (assign + update, up to maxIters passes, per-task stop) of
// the listed uint8 tasks, one workgroup each; leaves them inactive with their
// pass counts.  The first nbStream listed tasks (1 to kStreamTaskTiles tiles;
// blocksPerCu = 1 only) read their tiles from k.XA every pass (L2-resident after
// the first); the others have their points in LDS and 1 to maxTiles tiles
// (maxTiles <= lloydTasksTileCap(blocksPerCu), blocksPerCu = 1 or 2 workgroups
// sharing a CU).  taskTileBeg[g] = first tile of task g in k.XA.
constexpr uint32_t kStreamTaskTiles = 128;
void launchLloydTasks(const KmLevel& k, const uint32_t* taskList, uint32_t nbList, uint32_t maxTiles,
                      uint32_t blocksPerCu, const uint32_t* taskTileBeg, uint32_t maxIters, hipStream_t s,
                      const KernelTimer& t = {}, uint32_t nbStream = 0);
// Largest task (in 64-row tiles) whose points fit one workgroup's LDS when
// blocksPerCu workgroups of launchLloydTasks share a CU.
uint32_t lloydTasksTileCap(uint32_t blocksPerCu);
// Blocks of the MFMA assign kernel that the device holds at once (its grid cap).
uint32_t kmAssignResidentBlocks();
// One assign pass over all active tasks; kernel 0 = MFMA, 1 = plain.  Returns
// whether a kernel ran.
bool launchKmAssign(const KmLevel& k, int kernel, hipStream_t s, const KernelTimer& t = {});
// Update pass: converged tasks become inactive, others move their centers
// and clear their accumulators.  Counts remaining active tasks into nbActive.
// Returns whether a kernel ran.
bool launchKmUpdate(const KmLevel& k, hipStream_t s, const KernelTimer& t = {});

// Stable regroup of a level's rows by (task, child) (Vocabulary.cpp:83-98):
// orderOut = orderIn stably sorted by task * b + assignment; hist[task * b + j]
// = the occurrence counts.  Rows of the level are tasks-contiguous (k.tiles);
// orderOut == nullptr computes the histogram only.  ws: device workspace of
// regroupWorkspaceBytes(>= k.nbTiles, k.b) bytes, 256-byte aligned; no allocation.
size_t regroupWorkspaceBytes(uint64_t nbTilesMax, uint32_t b);
void launchRegroup(const KmLevel& k, uint32_t nbRows, const uint32_t* orderIn, uint32_t* orderOut, uint32_t* hist,
                   void* ws, size_t wsBytes, hipStream_t s);

// IDF weights (Vocabulary.cpp:103-108) of n children from their occurrence
// counts: w[i] = occ[i] == 0 ? FLT_MIN : max(FLT_MIN, float(log(double(float(nbDoc) / float(occ[i]))))).
void launchIdfWeights(const uint32_t* occ, uint64_t n, uint32_t nbDoc, float* w, hipStream_t s);

void launchFillU32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// p[i] = i for i < n
void launchIota(uint32_t* p, uint64_t n, hipStream_t s);

}  // namespace rbow_amd
