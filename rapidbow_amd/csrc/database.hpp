// database.hpp -- inverted-file database resident in HBM (DataBase.h:35-64).
//
// Postings live in leaf-major CSRs: off[leaf .. leaf+1) indexes a slice of
// document slots, slots ascending within a leaf (duplicates adjacent, one per
// descriptor, as the reference's sorted multisets DataBase.cpp:44-53).  Because
// leaves of one subtree are contiguous, a node's postings are one contiguous
// slice and its posting count is off[hi] - off[lo], which makes the
// reference's stop threshold O(1) per node.
//
// Documents get slots in insertion order, and their postings sit in one of
// three places, by slot:
//   main CSR   slots [0, mainSlots)             the bulk of the file
//   delta CSR  slots [mainSlots, foldedSlots)   recent inserts
//   pending    slots [foldedSlots, nbDocs)      host addDoc calls not yet on the device
// Every slot of the delta is above every slot of the main CSR, so "main slice
// then delta slice" of a leaf is the slot order the merged CSR would have.
// Before a query the pending postings are folded into the delta, which costs
// O(delta + pending + leaves), independent of the main CSR's size; the delta
// is merged into the main CSR only when it holds more than kDeltaMax postings
// (or on a bulk device insert), so a queryAndAddDoc loop does not rebuild the
// main CSR per call (DataBase.cpp:39-69,192-197 insert per posting).
//
// Scoring is tiled by document range: slots [r*kRangeDocs, (r+1)*kRangeDocs)
// form range r, whose accumulators fit one workgroup's LDS.  The main CSR's
// range directory dir[leaf*(Rm+1) + r] is the first posting of the leaf with a
// slot >= r*kRangeDocs (relative to off[leaf]), so a (query, range) tile reads
// only its own postings of each query leaf; a leaf's delta slice is short and
// is cut to the range by binary search.
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
    uint32_t nbDocs;           // slots with norms on the device
    const float* W;            // [node][b] IDF weights
    const uint64_t* leafOff;   // main CSR [nbLeaves + 1]
    const uint32_t* post;      // main CSR slots
    const uint64_t* leafOffD;  // delta CSR [nbLeaves + 1]
    const uint32_t* postD;     // delta CSR slots
    uint64_t nbPostD;          // delta postings
    uint32_t slotD0;           // first slot of the delta (= slots in the main CSR)
    const float* rcp;          // [slot] float(1 / L1 norm)
    const uint32_t* docId;     // [slot]
    const uint32_t* dir;       // [leaf][Rm + 1] range directory of the main CSR
    uint32_t Rm;               // document ranges of the main CSR
    uint32_t R;                // document ranges of all device slots
    uint32_t walkSteps;        // scoring tile shape: steps of 64 postings per wave iteration (4 or 6)
};

// A leaf-major CSR of document slots on the device.
struct DevCsr {
    DevBuf<uint64_t> off;  // [nbLeaves + 1]
    DevBuf<uint32_t> post;
    uint64_t n = 0;        // postings
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
    // Writes up to K ranked doc ids / scores per query and the count.  Returns when done.
    void queryBatch(const uint32_t* dLeaf, const uint64_t* offsets, uint32_t nbQueries, uint32_t K, uint32_t* dDocs,
                    float* dScores, uint32_t* dCounts, hipStream_t s) const;
    // single query from host leaf ids (synchronous)
    uint32_t query(const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs, float* scores) const;

    // Serving: fold pending inserts into the device state, then score batches
    // asynchronously (no insert may run concurrently with scoreAsync).
    void prepareForQueries(hipStream_t s) const;
    void scoreAsync(const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets, uint32_t nbQueries,
                    uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts, hipStream_t s) const;
    // The device view the scoring kernels read (after prepareForQueries, with no
    // insert pending), and a counter that changes whenever that view does.
    DbView scoringView() const;
    uint64_t version() const;
    bool hasPendingInserts() const;  // inserts not yet folded into the device state
    // Scoring tile shape: steps of 64 postings per wave and walk iteration, 4 or 6
    // (0: the default, 6).  A change bumps version() (captured graphs bake it in).
    void setWalkSteps(uint32_t steps);
    uint32_t walkSteps() const { return walkSteps_ ? walkSteps_ : kWalkStepsDefault; }
    static constexpr uint32_t kWalkStepsDefault = 6;
    // HBM scratch of one query of n descriptors, top-K, during scoring
    uint64_t scratchBytes(uint32_t n, uint32_t K) const {
        return uint64_t(divUp(nbDocs(), kRangeDocs)) * (8ull * K + 4) + uint64_t(n) * L_ * kQNodeBytes + 80;
    }

    uint32_t nbDocs() const;
    uint64_t nbPostings() const;
    uint64_t nbDeltaPostings() const;  // postings outside the main CSR (delta + pending)
    bool docRcp(uint32_t doc, float* rcp) const;
    uint32_t postings(uint32_t leaf, uint32_t* out, uint32_t cap) const;  // sorted doc ids
    // folds the delta into the main CSR now (the periodic merge, on demand)
    void compact(hipStream_t s);

    static constexpr uint32_t kMaxQueryDesc = 16384;
    static constexpr uint32_t kMaxResults = 1024;
    static constexpr uint32_t kRangeDocs = 4096;       // document slots per scoring tile
    static constexpr uint64_t kDeltaMax = 1ull << 24;  // delta postings that trigger the merge into the main CSR
    static constexpr uint32_t kQNodeBytes = 56;        // one scored query node (database.hip QNode)

private:
    void syncDevice(hipStream_t s) const;  // norms, pending postings -> delta, merge if the delta is large
    void foldPending(hipStream_t s) const;
    void addToDelta(DevCsr& batch, hipStream_t s) const;  // batch slots above every delta slot
    void mergeDeltaIntoMain(hipStream_t s) const;
    void refreshDir(hipStream_t s) const;    // rebuilds the range directory if stale
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
    mutable DevCsr main_, delta_, spare_;  // spare_: the delta being rebuilt (capacity reused)
    mutable uint32_t mainSlots_ = 0, foldedSlots_ = 0;
    mutable DevBuf<float> rcp_;
    mutable DevBuf<uint32_t> docId_;
    mutable uint32_t docsOnDevice_ = 0;
    mutable DevBuf<uint32_t> dir_;
    mutable uint32_t dirR_ = 0;
    mutable bool dirStale_ = true;
    mutable DevBuf<uint32_t> flag_;  // device error flag of the leaf-id checks
    mutable uint64_t version_ = 0;   // changes with every change of the scoring view
    uint32_t walkSteps_ = 0;         // scoring tile shape (0: default)
    mutable std::vector<uint32_t> pendLeaf_, pendSlot_;
    mutable std::shared_mutex mtx_;
    mutable std::mutex mergeMtx_;
};

// kernels (database.hip)
void launchDocNorms(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, uint32_t nbDocs,
                    uint32_t firstSlot, float* rcp, hipStream_t s);
// CSR of (leaf, slot) pairs given in ascending slot order: out.off, out.post
// (out.post sized by the caller to n; out.off to nbLeaves + 1).
void launchCsrFromPairs(uint64_t nbLeaves, const uint32_t* dLeaf, const uint32_t* dSlot, uint64_t n, DevCsr& out,
                        hipStream_t s);
// out = a ++ b leaf by leaf (every slot of b above every slot of a); out sized by the caller.
void launchMergeCsr(uint64_t nbLeaves, const DevCsr& a, const DevCsr& b, DevCsr& out, hipStream_t s);
// Range directory of the main CSR (v.Rm ranges): dir[(Rm+1)*leaf + r].
void launchBuildDir(const DbView& v, uint32_t* dir, hipStream_t s);
// slots[i] = firstSlot + d for the postings of doc d (offsets rebased to off[0])
void launchExpandSlots(const uint64_t* off, uint32_t nbDocs, uint32_t firstSlot, uint32_t* slots, hipStream_t s);
// flag[0] += number of ids in leaf[0, n) that are >= nbLeaves (flag zeroed by the caller)
void launchCountBadLeaves(const uint32_t* leaf, uint64_t n, uint64_t nbLeaves, uint32_t* flag, hipStream_t s);
// Scoring workspace of a batch of nbQueries queries holding totalDesc leaf ids,
// and the batch launched in a caller's workspace (no allocation).
size_t queryWorkspaceBytes(const DbView& v, uint64_t totalDesc, uint32_t nbQueries, uint32_t K);
void launchQueryBatchInWs(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, uint64_t totalDesc,
                          uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                          uint32_t* badFlag, void* ws, size_t wsBytes, hipStream_t s);
// badFlag (optional, device u32, zeroed by the caller): set to nonzero when a
// query holds a leaf id >= nbLeaves; such a query scores nothing.
void launchQueryBatch(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets,
                      uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                      uint32_t* badFlag, hipStream_t s);

}  // namespace rbow_amd
