// pipeline.cpp -- fixed-capacity query batches, direct or as captured HIP graphs.
#include "pipeline.hpp"

#include <algorithm>
#include <cstring>

namespace rbow_amd {

uint64_t QueryPipeline::descClass(uint64_t n, uint64_t cap) {
    uint64_t c = 4096;
    if (n > c) {
        uint64_t p = 4096;
        while (p * 2 < n) p *= 2;
        c = p + p / 4 * ((n - p + p / 4 - 1) / (p / 4));  // p, 1.25p, 1.5p, 1.75p, 2p
    }
    return std::min(c, cap);
}

uint32_t QueryPipeline::queryClass(uint32_t q, uint32_t cap) {
    uint32_t c = 1;
    while (c < q) c *= 2;
    return std::min(c, cap);
}

QueryPipeline::QueryPipeline(const Vocabulary& voc, const Database& db, uint32_t maxQueries, uint64_t maxDesc,
                             uint32_t K)
    : voc_(voc), db_(db), maxQ_(maxQueries), K_(K), maxDesc_(maxDesc) {
    RBOW_REQUIRE(maxQueries >= 1 && maxQueries <= 65535, RBOW_ERR_INVALID_ARG, "max_queries must be in [1, 65535]");
    RBOW_REQUIRE(maxDesc >= 1 && maxDesc < (1ull << 26), RBOW_ERR_INVALID_ARG, "max_desc must be in [1, 2^26)");
    RBOW_REQUIRE(K >= 1 && K <= Database::kMaxResults, RBOW_ERR_CAPACITY, "max_results must be in [1, 1024]");
    desc_.resize(maxDesc * kDims);
    leaf_.resize(maxDesc);
    off_.resize(maxQueries + 1);
    out_.resize(resultBytes());
    const TreeView t = voc_.view();
    lookWs_.resize(std::max<size_t>(lookupWorkspaceBytes(t, uint32_t(maxDesc), false, nullptr), 256));
    RBOW_HIP(hipHostMalloc(reinterpret_cast<void**>(&hOff_), 8ull * (maxQueries + 1)));
    RBOW_HIP(hipHostMalloc(reinterpret_cast<void**>(&hOut_), resultBytes()));
    RBOW_HIP(hipStreamCreateWithFlags(&capture_, hipStreamNonBlocking));  // default priority
}

QueryPipeline::~QueryPipeline() {
    (void)hipDeviceSynchronize();
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    if (capture_) (void)hipStreamDestroy(capture_);
    if (hOff_) (void)hipHostFree(hOff_);
    if (hOut_) (void)hipHostFree(hOut_);
}

void QueryPipeline::ensureScoreWs(const DbView& v) {
    const size_t need = queryWorkspaceBytes(v, maxDesc_, maxQ_, K_);
    if (need > scoreWs_.size()) scoreWs_.resize(need);  // only ever between graph generations
}

// Everything of one batch after the descriptors' upload: nDesc rows, nbQ queries.
void QueryPipeline::enqueue(const DbView& v, uint64_t nDesc, uint32_t nbQ, bool plain, bool toHost, hipStream_t s) {
    RBOW_HIP(hipMemcpyAsync(off_.get(), hOff_, 8ull * (nbQ + 1), hipMemcpyHostToDevice, s));
    const TreeView t = voc_.view();
    if (nDesc)
        launchLookupInWs(t, desc_.get(), uint32_t(nDesc), leaf_.get(), nullptr, plain, lookWs_.get(), lookWs_.size(),
                         s);
    uint32_t* docs = reinterpret_cast<uint32_t*>(out_.get());
    float* scores = reinterpret_cast<float*>(out_.get() + 4ull * maxQ_ * K_);
    uint32_t* counts = reinterpret_cast<uint32_t*>(out_.get() + 8ull * maxQ_ * K_);
    launchQueryBatchInWs(v, leaf_.get(), off_.get(), nDesc, nbQ, K_, docs, scores, counts, nullptr, scoreWs_.get(),
                         scoreWs_.size(), s);
    if (toHost) RBOW_HIP(hipMemcpyAsync(hOut_, out_.get(), resultBytes(), hipMemcpyDeviceToHost, s));
}

void QueryPipeline::run(const uint64_t* hOff, uint32_t nbQ, int lookupKernel, bool useGraph, bool toHost,
                        hipStream_t s) {
    RBOW_REQUIRE(nbQ >= 1 && nbQ <= maxQ_, RBOW_ERR_CAPACITY, "batch above the pipeline's query capacity");
    RBOW_REQUIRE(hOff[0] == 0, RBOW_ERR_INVALID_ARG, "offsets must start at 0");
    for (uint32_t q = 0; q < nbQ; q++) {
        RBOW_REQUIRE(hOff[q + 1] >= hOff[q], RBOW_ERR_INVALID_ARG, "offsets must be non-decreasing");
        RBOW_REQUIRE(hOff[q + 1] - hOff[q] <= Database::kMaxQueryDesc, RBOW_ERR_CAPACITY,
                     "query above 16384 descriptors");
    }
    const uint64_t n = hOff[nbQ];
    RBOW_REQUIRE(n <= maxDesc_, RBOW_ERR_CAPACITY, "batch above the pipeline's descriptor capacity");
    if (db_.hasPendingInserts()) db_.prepareForQueries(s);
    const DbView v = db_.scoringView();
    const bool plain = lookupUsesPlain(voc_.view(), uint32_t(n), lookupKernel);
    if (!useGraph) {
        ensureScoreWs(v);
        std::memcpy(hOff_, hOff, 8ull * (nbQ + 1));
        enqueue(v, n, nbQ, plain, toHost, s);
        return;
    }
    const uint64_t version = db_.version();
    if (version != graphVersion_) {  // the scoring view changed: every graph is stale
        RBOW_HIP(hipDeviceSynchronize());
        for (auto& g : graphs_) RBOW_HIP(hipGraphExecDestroy(g.second));
        graphs_.clear();
        ensureScoreWs(v);
        graphVersion_ = version;
    }
    const uint64_t nc = descClass(n, maxDesc_);
    const uint32_t qc = queryClass(nbQ, maxQ_);
    // the padding queries are empty: they start and end at the batch's last row
    std::memcpy(hOff_, hOff, 8ull * (nbQ + 1));
    for (uint32_t q = nbQ + 1; q <= qc; q++) hOff_[q] = n;
    const auto key = std::make_tuple(nc, qc, plain, toHost);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        hipGraph_t g = nullptr;
        RBOW_HIP(hipStreamBeginCapture(capture_, hipStreamCaptureModeRelaxed));
        try {
            enqueue(v, nc, qc, plain, toHost, capture_);
        } catch (...) {
            (void)hipStreamEndCapture(capture_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        RBOW_HIP(hipStreamEndCapture(capture_, &g));
        hipGraphExec_t ge = nullptr;
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RBOW_HIP(e);
        it = graphs_.emplace(key, ge).first;
        graphsCaptured_++;
    }
    RBOW_HIP(hipGraphLaunch(it->second, s));
}

}  // namespace rbow_amd
