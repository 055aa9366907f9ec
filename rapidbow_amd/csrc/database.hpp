// database.hpp -- inverted-file database resident in HBM (DataBase.h:35-64).
//
// Postings live in a leaf-major CSR: leafOff[leaf .. leaf+1) indexes a slice
// of document slots, grouped by slot within a leaf (duplicates adjacent, one
// per descriptor, as the reference's sorted multisets DataBase.cpp:44-53).
// Because leaves of one subtree are contiguous, a node's postings are one
// contiguous slice and its posting count is leafOff[hi] - leafOff[lo], which
// makes the reference's stop threshold O(1) per node.
//
// Scoring is tiled by document range: slots [r*kRangeDocs, (r+1)*kRangeDocs)
// form range r, whose accumulators fit one workgroup's LDS.  The range
// directory dir[leaf*(R+1) + r] is the first posting of the leaf with a slot
// >= r*kRangeDocs (relative to leafOff[leaf]), so a (query, range) tile reads
// only its own postings of each query leaf.
#pragma once
#include <mutex>
#include <shared_mutex>
#include <unordered_map>
#include <vector>

#include "vocabulary.hpp"

namespace rbow_amd {

struct DbView {  // device view used by the scoring kernels
    uint32_t b, lg2b, L;
    uint64_t nbLeaves;
    uint32_t nbDocs;
    const float* W;         // [node][b] IDF weights
    const uint64_t* leafOff;  // [nbLeaves + 1]
    const uint32_t* post;   // slots
    const float* rcp;       // [slot] float(1 / L1 norm)
    const uint32_t* docId;  // [slot]
    const uint32_t* dir;    // [leaf][R + 1] range directory
    uint32_t R;             // document ranges
};

class Database {
public:
    explicit Database(const Vocabulary& voc);
    ~Database();
    const Vocabulary& vocabulary() const { return voc_; }

    // DataBase.cpp:39-69 (host leaf ids)
    void addDoc(uint32_t doc, const uint32_t* leaf, uint32_t n);
    // Batch insert from device leaf ids: docs[i] owns dLeaf[offsets[i] .. offsets[i+1]).
    void addDocsDevice(const uint32_t* docs, uint32_t nbDocs, const uint32_t* dLeaf, const uint64_t* offsets,
                       hipStream_t s);
    // DataBase.cpp:71-190 for a batch of queries (device leaf ids; offsets on the host).
    // Writes up to K ranked doc ids / scores per query and the count.  Asynchronous on s.
    void queryBatch(const uint32_t* dLeaf, const uint64_t* offsets, uint32_t nbQueries, uint32_t K, uint32_t* dDocs,
                    float* dScores, uint32_t* dCounts, hipStream_t s) const;
    // single query from host leaf ids (synchronous)
    uint32_t query(const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs, float* scores) const;

    // Serving: fold pending inserts into the CSR, then score batches
    // asynchronously (no insert may run concurrently with scoreAsync).
    void prepareForQueries(hipStream_t s) const;
    void scoreAsync(const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets, uint32_t nbQueries,
                    uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts, hipStream_t s) const;
    // HBM scratch of one query of n descriptors, top-K, during scoring
    uint64_t scratchBytes(uint32_t n, uint32_t K) const {
        return uint64_t(divUp(nbDocs(), kRangeDocs)) * (8ull * K + 4) + uint64_t(n) * L_ * 40 + 64;
    }

    uint32_t nbDocs() const;
    uint64_t nbPostings() const;
    bool docRcp(uint32_t doc, float* rcp) const;
    uint32_t postings(uint32_t leaf, uint32_t* out, uint32_t cap) const;  // sorted doc ids

    static constexpr uint32_t kMaxQueryDesc = 16384;
    static constexpr uint32_t kMaxResults = 1024;
    static constexpr uint32_t kRangeDocs = 4096;  // document slots per scoring tile

private:
    void ensureMerged(hipStream_t s) const;  // folds pending postings into the CSR
    void refreshDir(hipStream_t s) const;    // rebuilds the range directory if stale
    void mergeDevice(const uint32_t* dLeaf, const uint32_t* dSlot, uint64_t n, hipStream_t s) const;
    void growDocs(uint32_t nb, hipStream_t s) const;
    DbView view() const;

    const Vocabulary& voc_;
    uint32_t b_, lg2b_, L_;
    uint64_t nbLeaves_;
    std::vector<float> Wh_;  // IDF weights, subnormals flushed
    DevBuf<float> W_;
    std::unordered_map<uint32_t, uint32_t> slotOf_;
    std::vector<uint32_t> docIds_;
    std::vector<float> rcpHost_;
    // device state is logically const for queries but is built lazily
    mutable DevBuf<uint64_t> leafOff_;
    mutable DevBuf<uint32_t> post_;
    mutable uint64_t nbPost_ = 0;
    mutable DevBuf<float> rcp_;
    mutable DevBuf<uint32_t> docId_;
    mutable uint32_t docsOnDevice_ = 0;
    mutable DevBuf<uint32_t> dir_;
    mutable uint32_t dirR_ = 0;
    mutable bool dirStale_ = true;
    mutable std::vector<uint32_t> pendLeaf_, pendSlot_;
    mutable std::shared_mutex mtx_;
    mutable std::mutex mergeMtx_;
};

// kernels (database.hip)
void launchDocNorms(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, uint32_t nbDocs,
                    uint32_t firstSlot, float* rcp, hipStream_t s);
void launchMergePostings(const DbView& v, uint64_t nbOld, const uint32_t* dLeaf, const uint32_t* dSlot, uint64_t n,
                         uint64_t* newLeafOff, uint32_t* newPost, hipStream_t s);
// Range directory of the current CSR (v.R ranges): dir[(R+1)*leaf + r].
void launchBuildDir(const DbView& v, uint32_t* dir, hipStream_t s);
// slots[i] = firstSlot + d for the postings of doc d (offsets rebased to off[0])
void launchExpandSlots(const uint64_t* off, uint32_t nbDocs, uint32_t firstSlot, uint32_t* slots, hipStream_t s);
void launchQueryBatch(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets,
                      uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                      hipStream_t s);

}  // namespace rbow_amd
