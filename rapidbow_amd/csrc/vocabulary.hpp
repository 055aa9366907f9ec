// vocabulary.hpp -- device-resident vocabulary tree (Vocabulary.h:33-87).
#pragma once
#include <vector>

#include "kernels.hpp"

namespace rbow_amd {

struct Vocabulary {
    uint32_t b = 16, lg2b = 4, L = 0, centroidType = 0;  // centroidType 0: uint8, 1: float32
    uint64_t nbNodes = 0;
    DevBuf<uint8_t> u8rows;  // [node][b][128]
    DevBuf<int32_t> normU, normS;
    DevBuf<float> f32rows;   // [node][b][128]
    DevBuf<int8_t> split;    // [3][node*b+j][128] int8 terms of c - 128 (float trees)
    DevBuf<float> cn2, e1, e2;
    std::vector<uint32_t> masks;   // valid-center masks, low bits (kmeans.cpp:189)
    std::vector<float> weights;    // [node][b] IDF weights (Vocabulary.cpp:103-108), host like the reference
    uint32_t lookupAutoMin = 0;    // tuned AUTO crossover (0: untuned)
    uint32_t lookupFusedMax = 0;   // tuned fused/bucketed crossover of the MFMA form (0: untuned)
    uint32_t lookupLeafItemRows = 0;  // tuned leaf-level item size of the bucketed descent (0: untuned)

    void allocate(uint32_t b_, uint32_t L_, uint32_t centroidType_);
    // derived per-child data used by the kernels (norms, int8 split, bounds)
    void prepare(hipStream_t s);
    TreeView view() const;
    uint64_t nbChildren() const { return nbNodes * b; }
};

struct BuildStats {
    std::vector<uint32_t> passes;  // Lloyd passes per node
    double totalMs = 0;            // wall time of the build (host clock)
    double assignMs = 0;           // sum of assign-kernel times (HIP events), persistent launches included
    double updateMs = 0;           // sum of update-kernel times (HIP events)
    double reduceMs = 0;           // sum of chunk-reduction times (HIP events)
    double streamMs = 0;           // per-pass MFMA assign launches only
    uint32_t streamLaunches = 0;
    uint64_t streamPointPasses = 0;
    double persistMs = 0;          // persistent small-task launches (whole Lloyd loops)
    uint64_t pointPasses = 0;      // sum over nodes of points x Lloyd passes
};

// Vocabulary.cpp:32-121 buildSiftVocabulary, generalised to b in {16, 32} and
// float32 centroids.  Blocks until done.
Vocabulary* buildVocabulary(const uint8_t* devDesc, uint32_t nbDesc, uint32_t nbDoc, uint32_t b, uint32_t L,
                            const rbow_build_options& opts, hipStream_t s, BuildStats* stats);
// Frees the work buffers that completed builds leave with each device for the next build.
void releaseBuildWorkspaces();

std::vector<uint8_t> serialize(const Vocabulary& v);
Vocabulary* deserialize(const uint8_t* blob, size_t size);

}  // namespace rbow_amd
