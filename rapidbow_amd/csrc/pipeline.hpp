// pipeline.hpp -- one batch of query images through descriptor lookup, query
// preparation, range scoring and top-K (Vocabulary.cpp:162-173 + DataBase.cpp:71-190
// for many images at once), in device buffers of fixed capacity.
//
// A batch runs either as direct launches or as ONE launch of a captured HIP
// graph.  A graph bakes in every kernel argument, so graphs are captured per
// size class and reused: the descriptor count is rounded up to a quarter of an
// octave (the lookup also processes the padding rows, whose leaf ids nothing
// reads) and the query count to a power of two (the padding queries are empty
// and score nothing).  The graph holds the offsets' upload, every lookup and
// scoring kernel and, for host results, the download of the result block; the
// descriptors' upload stays outside (its size varies).  Graphs are dropped when
// the database's scoring view changes.  Nothing is allocated per batch.
#pragma once
#include <map>
#include <tuple>

#include "database.hpp"

namespace rbow_amd {

class QueryPipeline {
public:
    QueryPipeline(const Vocabulary& voc, const Database& db, uint32_t maxQueries, uint64_t maxDesc, uint32_t K);
    ~QueryPipeline();
    QueryPipeline(const QueryPipeline&) = delete;
    QueryPipeline& operator=(const QueryPipeline&) = delete;

    // device uint8[maxDesc][128]: a batch's descriptors, query after query
    uint8_t* descBuffer() const { return desc_.get(); }
    // Runs the batch whose descriptors are in descBuffer(): query q owns rows
    // hOff[q] .. hOff[q+1] (hOff[0] == 0).  Asynchronous on s.  The results are in
    // deviceResults() when s reaches this point, and also in hostResults() when
    // toHost.  The pipeline may be run again once that point is reached.
    void run(const uint64_t* hOff, uint32_t nbQueries, int lookupKernel, bool useGraph, bool toHost, hipStream_t s);

    // results block [docs u32 maxQ x K][scores f32 maxQ x K][counts u32 maxQ]
    const uint8_t* deviceResults() const { return out_.get(); }
    const uint8_t* hostResults() const { return hOut_; }
    uint64_t resultBytes() const { return 8ull * maxQ_ * K_ + 4ull * maxQ_; }
    const uint32_t* docs(const uint8_t* block, uint32_t q) const {
        return reinterpret_cast<const uint32_t*>(block) + uint64_t(q) * K_;
    }
    const float* scores(const uint8_t* block, uint32_t q) const {
        return reinterpret_cast<const float*>(block + 4ull * maxQ_ * K_) + uint64_t(q) * K_;
    }
    const uint32_t* counts(const uint8_t* block) const {
        return reinterpret_cast<const uint32_t*>(block + 8ull * maxQ_ * K_);
    }
    uint32_t count(const uint8_t* hostBlock, uint32_t q) const { return counts(hostBlock)[q]; }
    uint32_t maxQueries() const { return maxQ_; }
    uint64_t maxDesc() const { return maxDesc_; }
    uint32_t K() const { return K_; }
    uint32_t graphsCaptured() const { return graphsCaptured_; }

    // size classes of the captured graphs
    static uint64_t descClass(uint64_t n, uint64_t cap);
    static uint32_t queryClass(uint32_t q, uint32_t cap);

private:
    void enqueue(const DbView& v, uint64_t nDesc, uint32_t nbQ, bool plain, bool toHost, hipStream_t s);
    void ensureScoreWs(const DbView& v);

    const Vocabulary& voc_;
    const Database& db_;
    uint32_t maxQ_, K_;
    uint64_t maxDesc_;
    DevBuf<uint8_t> desc_, out_, lookWs_, scoreWs_;
    DevBuf<uint32_t> leaf_;
    DevBuf<uint64_t> off_;
    uint64_t* hOff_ = nullptr;  // pinned [maxQ + 1]
    uint8_t* hOut_ = nullptr;   // pinned results block
    hipStream_t capture_ = nullptr;
    std::map<std::tuple<uint64_t, uint32_t, bool, bool>, hipGraphExec_t> graphs_;
    uint64_t graphVersion_ = ~0ull;
    uint32_t graphsCaptured_ = 0;
};

}  // namespace rbow_amd
