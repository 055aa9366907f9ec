// database.cpp -- host side of the HBM-resident inverted file (DataBase.cpp).
#include "database.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace rbow_amd {

namespace {
inline float ftz(float x) { return std::fabs(x) < FLT_MIN ? std::copysign(0.0f, x) : x; }

// capacity growth of reused device buffers
template <typename T>
void ensureCap(DevBuf<T>& b, size_t n) {
    if (n > b.size()) b.resize(std::max(n, b.size() + b.size() / 2));
}
}  // namespace

Database::Database(const Vocabulary& voc) : voc_(voc), b_(voc.b), lg2b_(voc.lg2b), L_(voc.L) {
    RBOW_REQUIRE(L_ >= 1, RBOW_ERR_INVALID_ARG, "vocabulary has no levels");
    nbLeaves_ = levelSize(lg2b_, L_);
    // IDF weights as the scoring sees them: subnormals read as zero (the reference's DAZ)
    Wh_ = voc.weights;
    for (auto& x : Wh_) x = ftz(x);
    W_.resize(Wh_.size());
    RBOW_HIP(hipMemcpy(W_.get(), Wh_.data(), Wh_.size() * 4, hipMemcpyHostToDevice));
    for (DevCsr* c : {&main_, &delta_}) {
        c->off.resize(nbLeaves_ + 1);
        RBOW_HIP(hipMemset(c->off.get(), 0, c->off.bytes()));
        c->post.resize(1);
    }
    spare_.off.resize(nbLeaves_ + 1);
    flag_.resize(1);
    growDocs(1024, nullptr);
    RBOW_HIP(hipDeviceSynchronize());
}

Database::~Database() { (void)hipDeviceSynchronize(); }

DbView Database::view() const {
    DbView v;
    v.b = b_;
    v.lg2b = lg2b_;
    v.L = L_;
    v.nbLeaves = nbLeaves_;
    v.nbDocs = docsOnDevice_;
    v.W = W_.get();
    v.leafOff = main_.off.get();
    v.post = main_.post.get();
    v.leafOffD = delta_.off.get();
    v.postD = delta_.post.get();
    v.nbPostD = delta_.n;
    v.slotD0 = mainSlots_;
    v.rcp = rcp_.get();
    v.docId = docId_.get();
    v.dir = dirStale_ ? nullptr : dir_.get();  // a stale directory is never handed to a kernel
    v.Rm = dirR_;
    v.R = divUp(docsOnDevice_, kRangeDocs);
    v.walkSteps = walkSteps();
    return v;
}

void Database::setWalkSteps(uint32_t steps) {
    RBOW_REQUIRE(steps == 0 || steps == 4 || steps == 6, RBOW_ERR_INVALID_ARG, "walk steps must be 4 or 6");
    std::unique_lock<std::shared_mutex> lk(mtx_);  // exclusive, as an insert
    if (steps != walkSteps_) version_++;
    walkSteps_ = steps;
}

void Database::growDocs(uint32_t nb, hipStream_t s) const {
    if (nb <= rcp_.size()) return;
    const size_t cap = std::max<size_t>(nb, rcp_.size() * 2);
    DevBuf<float> r(cap);
    DevBuf<uint32_t> d(cap);
    if (docsOnDevice_) {
        RBOW_HIP(hipMemcpyAsync(r.get(), rcp_.get(), 4ull * docsOnDevice_, hipMemcpyDeviceToDevice, s));
        RBOW_HIP(hipMemcpyAsync(d.get(), docId_.get(), 4ull * docsOnDevice_, hipMemcpyDeviceToDevice, s));
    }
    RBOW_HIP(hipStreamSynchronize(s));
    rcp_ = std::move(r);
    docId_ = std::move(d);
}

void Database::addDoc(uint32_t doc, const uint32_t* leaf, uint32_t n) {
    std::unique_lock<std::shared_mutex> lk(mtx_);  // DataBase.cpp:40: exclusive
    RBOW_REQUIRE(!slotOf_.count(doc), RBOW_ERR_DUPLICATE_DOC, "Duplicate document index");  // DataBase.cpp:41-43
    RBOW_REQUIRE(leaf != nullptr || n == 0, RBOW_ERR_INVALID_ARG, "null leaf ids");
    for (uint32_t i = 0; i < n; i++)
        RBOW_REQUIRE(leaf[i] < nbLeaves_, RBOW_ERR_INVALID_ARG, "leaf id out of range");
    // DataBase.cpp:54-68: binary64 L1 norm, level-major over the sorted leaves
    std::vector<uint32_t> nxt(leaf, leaf + n);
    std::sort(nxt.begin(), nxt.end());
    double l1 = 0;
    for (uint32_t t = 0; t < L_; t++) {
        const float* W = Wh_.data() + levelBeg(lg2b_, L_ - 1 - t) * b_;
        for (uint32_t i = 0; i < n; i++) {
            l1 += W[nxt[i]];
            nxt[i] >>= lg2b_;
        }
    }
    const uint32_t slot = uint32_t(docIds_.size());
    slotOf_[doc] = slot;
    docIds_.push_back(doc);
    rcpHost_.push_back(ftz(float(1 / l1)));
    for (uint32_t i = 0; i < n; i++) {
        pendLeaf_.push_back(leaf[i]);
        pendSlot_.push_back(slot);
    }
}

// delta <- delta ++ batch, leaf by leaf, into the spare buffers (whose capacity
// is kept), then the two swap: O(delta + batch + leaves)
void Database::addToDelta(DevCsr& batch, hipStream_t s) const {
    if (batch.n == 0) return;
    ensureCap(spare_.post, delta_.n + batch.n);
    launchMergeCsr(nbLeaves_, delta_, batch, spare_, s);
    std::swap(delta_, spare_);
}

void Database::foldPending(hipStream_t s) const {
    if (pendLeaf_.empty()) {
        foldedSlots_ = uint32_t(docIds_.size());
        return;
    }
    const uint64_t n = pendLeaf_.size();
    DevBuf<uint32_t> pairs(2 * n);
    RBOW_HIP(hipMemcpyAsync(pairs.get(), pendLeaf_.data(), 4 * n, hipMemcpyHostToDevice, s));
    RBOW_HIP(hipMemcpyAsync(pairs.get() + n, pendSlot_.data(), 4 * n, hipMemcpyHostToDevice, s));
    DevCsr batch;
    batch.off.resize(nbLeaves_ + 1);
    batch.post.resize(n);
    launchCsrFromPairs(nbLeaves_, pairs.get(), pairs.get() + n, n, batch, s);
    addToDelta(batch, s);
    RBOW_HIP(hipStreamSynchronize(s));  // the host lists and the batch buffers are released below
    pendLeaf_.clear();
    pendSlot_.clear();
    foldedSlots_ = uint32_t(docIds_.size());
}

void Database::mergeDeltaIntoMain(hipStream_t s) const {
    if (delta_.n == 0) {
        mainSlots_ = foldedSlots_;
        return;
    }
    DevCsr m;
    m.off.resize(nbLeaves_ + 1);
    m.post.resize(main_.n + delta_.n);
    launchMergeCsr(nbLeaves_, main_, delta_, m, s);
    RBOW_HIP(hipStreamSynchronize(s));  // no query still reads the old main CSR
    main_ = std::move(m);
    RBOW_HIP(hipMemsetAsync(delta_.off.get(), 0, delta_.off.bytes(), s));
    delta_.n = 0;
    mainSlots_ = foldedSlots_;
    dirStale_ = true;
}

void Database::syncDevice(hipStream_t s) const {
    std::lock_guard<std::mutex> g(mergeMtx_);
    const uint32_t nDocs = uint32_t(docIds_.size());
    if (nDocs > docsOnDevice_ || !pendLeaf_.empty()) version_++;
    if (nDocs > docsOnDevice_) {
        growDocs(nDocs, s);
        RBOW_HIP(hipMemcpyAsync(rcp_.get() + docsOnDevice_, rcpHost_.data() + docsOnDevice_,
                                4ull * (nDocs - docsOnDevice_), hipMemcpyHostToDevice, s));
        RBOW_HIP(hipMemcpyAsync(docId_.get() + docsOnDevice_, docIds_.data() + docsOnDevice_,
                                4ull * (nDocs - docsOnDevice_), hipMemcpyHostToDevice, s));
        docsOnDevice_ = nDocs;
    }
    foldPending(s);
    if (delta_.n > kDeltaMax) mergeDeltaIntoMain(s);
    if (delta_.n == 0) mainSlots_ = foldedSlots_;  // empty delta: its documents count as main ones
    refreshDir(s);
    RBOW_HIP(hipStreamSynchronize(s));
}

void Database::compact(hipStream_t s) {
    std::unique_lock<std::shared_mutex> lk(mtx_);
    syncDevice(s);
    std::lock_guard<std::mutex> g(mergeMtx_);
    version_++;
    mergeDeltaIntoMain(s);
    refreshDir(s);
    RBOW_HIP(hipStreamSynchronize(s));
}

void Database::refreshDir(hipStream_t s) const {
    const uint32_t Rm = divUp(mainSlots_, kRangeDocs);
    if (!dirStale_ && Rm == dirR_) return;
    RBOW_REQUIRE(Rm == 0 || nbLeaves_ * (Rm + 1) < (1ull << 40), RBOW_ERR_CAPACITY, "range directory too large");
    dir_.resize(std::max<uint64_t>(nbLeaves_ * (Rm + 1), 1));
    dirR_ = Rm;
    dirStale_ = false;
    if (Rm) launchBuildDir(view(), dir_.get(), s);
}

void Database::addDocsDevice(const uint32_t* docs, uint32_t nbDocs, const uint32_t* dLeaf, const uint64_t* offsets,
                             hipStream_t s) {
    std::unique_lock<std::shared_mutex> lk(mtx_);
    RBOW_REQUIRE((docs && offsets) || nbDocs == 0, RBOW_ERR_INVALID_ARG, "null argument");
    if (nbDocs == 0) return;
    {
        std::unordered_map<uint32_t, uint32_t> seen;
        for (uint32_t d = 0; d < nbDocs; d++) {
            RBOW_REQUIRE(offsets[d + 1] >= offsets[d], RBOW_ERR_INVALID_ARG, "offsets must be non-decreasing");
            RBOW_REQUIRE(!slotOf_.count(docs[d]) && seen.emplace(docs[d], d).second, RBOW_ERR_DUPLICATE_DOC,
                         "Duplicate document index");
        }
    }
    const uint64_t total = offsets[nbDocs] - offsets[0];
    RBOW_REQUIRE(dLeaf != nullptr || total == 0, RBOW_ERR_INVALID_ARG, "null leaf ids");
    // every leaf id is checked on the device before anything reads through it
    if (total) {
        RBOW_HIP(hipMemsetAsync(flag_.get(), 0, 4, s));
        launchCountBadLeaves(dLeaf + offsets[0], total, nbLeaves_, flag_.get(), s);
        uint32_t bad = 0;
        RBOW_HIP(hipMemcpyAsync(&bad, flag_.get(), 4, hipMemcpyDeviceToHost, s));
        RBOW_HIP(hipStreamSynchronize(s));
        RBOW_REQUIRE(bad == 0, RBOW_ERR_INVALID_ARG, "leaf id out of range");
    }
    syncDevice(s);  // host-added postings keep their (smaller) slots first in every leaf
    std::lock_guard<std::mutex> g(mergeMtx_);
    version_++;
    const uint32_t first = uint32_t(docIds_.size());
    for (uint32_t d = 0; d < nbDocs; d++) {
        slotOf_[docs[d]] = first + d;
        docIds_.push_back(docs[d]);
    }
    growDocs(first + nbDocs, s);
    // norms on the device, then the postings in chunks of <= 2^30 through the delta
    DevBuf<uint64_t> dOff(nbDocs + 1);
    RBOW_HIP(hipMemcpyAsync(dOff.get(), offsets, 8ull * (nbDocs + 1), hipMemcpyHostToDevice, s));
    DbView v = view();
    launchDocNorms(v, dLeaf, dOff.get(), nbDocs, first, rcp_.get(), s);
    RBOW_HIP(hipMemcpyAsync(docId_.get() + first, docs, 4ull * nbDocs, hipMemcpyHostToDevice, s));
    rcpHost_.resize(first + nbDocs);
    RBOW_HIP(hipMemcpyAsync(rcpHost_.data() + first, rcp_.get() + first, 4ull * nbDocs, hipMemcpyDeviceToHost, s));
    docsOnDevice_ = first + nbDocs;
    uint32_t d0 = 0;
    while (d0 < nbDocs) {
        uint32_t d1 = d0 + 1;
        while (d1 < nbDocs && offsets[d1 + 1] - offsets[d0] < (1ull << 30)) d1++;
        const uint64_t n = offsets[d1] - offsets[d0];
        RBOW_REQUIRE(n < (1ull << 31), RBOW_ERR_CAPACITY, "one document has too many descriptors");
        if (n) {
            DevBuf<uint32_t> slots(n);
            launchExpandSlots(dOff.get() + d0, d1 - d0, first + d0, slots.get(), s);
            DevCsr batch;
            batch.off.resize(nbLeaves_ + 1);
            batch.post.resize(n);
            launchCsrFromPairs(nbLeaves_, dLeaf + offsets[d0], slots.get(), n, batch, s);
            if (delta_.n == 0 && n > kDeltaMax) {  // a bulk chunk goes straight into the main CSR
                foldedSlots_ = first + d1;
                std::swap(delta_, batch);
                mergeDeltaIntoMain(s);
                std::swap(delta_, batch);  // the (zeroed) empty delta comes back
            } else {
                addToDelta(batch, s);
                foldedSlots_ = first + d1;
                if (delta_.n > kDeltaMax) mergeDeltaIntoMain(s);
            }
            RBOW_HIP(hipStreamSynchronize(s));  // the chunk's buffers are released here
        }
        foldedSlots_ = first + d1;
        d0 = d1;
    }
    if (delta_.n == 0) mainSlots_ = foldedSlots_;
    refreshDir(s);
    RBOW_HIP(hipStreamSynchronize(s));
}

void Database::queryBatch(const uint32_t* dLeaf, const uint64_t* offsets, uint32_t nbQueries, uint32_t K,
                          uint32_t* dDocs, float* dScores, uint32_t* dCounts, hipStream_t s) const {
    RBOW_REQUIRE(K <= kMaxResults, RBOW_ERR_CAPACITY, "maxNbResults above 1024");
    RBOW_REQUIRE(offsets != nullptr || nbQueries == 0, RBOW_ERR_INVALID_ARG, "null offsets");
    for (uint32_t q = 0; q < nbQueries; q++) {
        RBOW_REQUIRE(offsets[q + 1] >= offsets[q], RBOW_ERR_INVALID_ARG, "offsets must be non-decreasing");
        RBOW_REQUIRE(offsets[q + 1] - offsets[q] <= kMaxQueryDesc, RBOW_ERR_CAPACITY,
                     "query above 16384 descriptors");
    }
    if (nbQueries == 0) return;
    {
        std::unique_lock<std::shared_mutex> lk(mtx_);
        syncDevice(s);
    }
    std::shared_lock<std::shared_mutex> lk(mtx_);  // DataBase.cpp:73: queries share the lock
    // rebase offsets so that the kernels index dLeaf from 0
    std::vector<uint64_t> h(nbQueries + 1);
    for (uint32_t q = 0; q <= nbQueries; q++) h[q] = offsets[q] - offsets[0];
    RBOW_REQUIRE(dLeaf != nullptr || h[nbQueries] == 0, RBOW_ERR_INVALID_ARG, "null leaf ids");
    void* dOff = nullptr;
    RBOW_HIP(hipMallocAsync(&dOff, 8ull * (nbQueries + 1) + 8, s));
    uint32_t* bad = reinterpret_cast<uint32_t*>(static_cast<uint64_t*>(dOff) + nbQueries + 1);
    RBOW_HIP(hipMemcpyAsync(dOff, h.data(), 8ull * (nbQueries + 1), hipMemcpyHostToDevice, s));
    RBOW_HIP(hipMemsetAsync(bad, 0, 4, s));
    launchQueryBatch(view(), dLeaf ? dLeaf + offsets[0] : nullptr, static_cast<uint64_t*>(dOff), h.data(),
                     nbQueries, K, dDocs, dScores, dCounts, bad, s);
    uint32_t hBad = 0;
    RBOW_HIP(hipMemcpyAsync(&hBad, bad, 4, hipMemcpyDeviceToHost, s));
    RBOW_HIP(hipFreeAsync(dOff, s));
    RBOW_HIP(hipStreamSynchronize(s));  // h (host) must outlive the copy; results ready on return
    RBOW_REQUIRE(hBad == 0, RBOW_ERR_INVALID_ARG, "leaf id out of range");
}

DbView Database::scoringView() const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    RBOW_REQUIRE(pendLeaf_.empty() && docsOnDevice_ == docIds_.size(), RBOW_ERR_INVALID_ARG,
                 "database has unmerged inserts");
    return view();
}

bool Database::hasPendingInserts() const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    return !pendLeaf_.empty() || docsOnDevice_ != docIds_.size();
}

uint64_t Database::version() const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    return version_;
}

void Database::prepareForQueries(hipStream_t s) const {
    std::unique_lock<std::shared_mutex> lk(mtx_);
    syncDevice(s);
}

void Database::scoreAsync(const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets,
                          uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                          hipStream_t s) const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    RBOW_REQUIRE(pendLeaf_.empty() && docsOnDevice_ == docIds_.size(), RBOW_ERR_INVALID_ARG,
                 "database has unmerged inserts");
    launchQueryBatch(view(), dLeaf, dOffsets, hOffsets, nbQueries, K, dDocs, dScores, dCounts, nullptr, s);
}

uint32_t Database::query(const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs, float* scores) const {
    RBOW_REQUIRE(leaf != nullptr || n == 0, RBOW_ERR_INVALID_ARG, "null leaf ids");
    for (uint32_t i = 0; i < n; i++) RBOW_REQUIRE(leaf[i] < nbLeaves_, RBOW_ERR_INVALID_ARG, "leaf id out of range");
    if (K == 0) return 0;
    DevBuf<uint32_t> dl(std::max(1u, n)), dd(K), dc(1);
    DevBuf<float> ds(K);
    if (n) RBOW_HIP(hipMemcpy(dl.get(), leaf, 4ull * n, hipMemcpyHostToDevice));
    const uint64_t off[2] = {0, n};
    queryBatch(dl.get(), off, 1, K, dd.get(), ds.get(), dc.get(), nullptr);
    uint32_t cnt = 0;
    RBOW_HIP(hipMemcpy(&cnt, dc.get(), 4, hipMemcpyDeviceToHost));
    if (cnt) {
        RBOW_HIP(hipMemcpy(docs, dd.get(), 4ull * cnt, hipMemcpyDeviceToHost));
        if (scores) RBOW_HIP(hipMemcpy(scores, ds.get(), 4ull * cnt, hipMemcpyDeviceToHost));
    }
    return cnt;
}

uint32_t Database::nbDocs() const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    return uint32_t(docIds_.size());
}

uint64_t Database::nbPostings() const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    return main_.n + delta_.n + pendLeaf_.size();
}

uint64_t Database::nbDeltaPostings() const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    return delta_.n + pendLeaf_.size();
}

bool Database::docRcp(uint32_t doc, float* rcp) const {
    std::shared_lock<std::shared_mutex> lk(mtx_);
    auto it = slotOf_.find(doc);
    if (it == slotOf_.end()) return false;
    *rcp = rcpHost_[it->second];
    return true;
}

uint32_t Database::postings(uint32_t leaf, uint32_t* out, uint32_t cap) const {
    RBOW_REQUIRE(leaf < nbLeaves_, RBOW_ERR_INVALID_ARG, "leaf id out of range");
    {
        std::unique_lock<std::shared_mutex> lk(mtx_);
        syncDevice(nullptr);
    }
    std::shared_lock<std::shared_mutex> lk(mtx_);
    std::vector<uint32_t> slots;
    for (const DevCsr* c : {&main_, &delta_}) {
        uint64_t o[2];
        RBOW_HIP(hipMemcpy(o, c->off.get() + leaf, 16, hipMemcpyDeviceToHost));
        const size_t at = slots.size();
        slots.resize(at + (o[1] - o[0]));
        if (o[1] > o[0])
            RBOW_HIP(hipMemcpy(slots.data() + at, c->post.get() + o[0], 4 * (o[1] - o[0]), hipMemcpyDeviceToHost));
    }
    for (auto& x : slots) x = docIds_[x];
    std::sort(slots.begin(), slots.end());  // the reference keeps each list sorted by doc id
    if (out) std::copy_n(slots.begin(), std::min<size_t>(cap, slots.size()), out);
    return uint32_t(slots.size());
}

}  // namespace rbow_amd
