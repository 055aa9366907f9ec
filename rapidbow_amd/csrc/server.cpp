// server.cpp -- continuous-batching query server (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <cstring>

namespace rbow_amd {

Server::Server(const Vocabulary& voc, const Database& db, const ServerOptions& o) : voc_(voc), db_(db), opt_(o) {
    RBOW_REQUIRE(o.maxBatch >= 1 && o.maxBatch <= 4096, RBOW_ERR_INVALID_ARG, "max_batch must be in [1, 4096]");
    RBOW_REQUIRE(o.nbStreams >= 1 && o.nbStreams <= 8, RBOW_ERR_INVALID_ARG, "nb_streams must be in [1, 8]");
    RBOW_REQUIRE(o.K >= 1 && o.K <= Database::kMaxResults, RBOW_ERR_CAPACITY, "max_results must be in [1, 1024]");
    RBOW_REQUIRE(o.lookupKernel == RBOW_LOOKUP_PLAIN || o.lookupKernel == RBOW_LOOKUP_MFMA ||
                     o.lookupKernel == RBOW_LOOKUP_AUTO,
                 RBOW_ERR_INVALID_ARG, "bad lookup kernel");
    RBOW_REQUIRE(o.stagingBytes >= uint64_t(Database::kMaxQueryDesc) * kDims, RBOW_ERR_INVALID_ARG,
                 "staging_bytes must hold one query of 16384 descriptors (2 MiB)");
    db_.prepareForQueries(nullptr);
    RBOW_HIP(hipDeviceSynchronize());
    ringCap_ = o.stagingBytes;
    RBOW_HIP(hipHostMalloc(reinterpret_cast<void**>(&ring_), ringCap_));
    if (!opt_.batchMaxDesc)
        opt_.batchMaxDesc = std::max<uint64_t>(Database::kMaxQueryDesc,
                                               std::min<uint64_t>(uint64_t(o.maxBatch) * Database::kMaxQueryDesc, 1u << 20));
    RBOW_REQUIRE(opt_.batchMaxDesc >= Database::kMaxQueryDesc && opt_.batchMaxDesc < (1ull << 26),
                 RBOW_ERR_INVALID_ARG, "batch_max_desc must be in [16384, 2^26)");
    slots_.resize(o.nbStreams);
    for (auto& s : slots_) {
        RBOW_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));  // default priority
        RBOW_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        s.pipe.reset(new QueryPipeline(voc_, db_, o.maxBatch, opt_.batchMaxDesc, o.K));
    }
    t0_ = Clock::now();
    worker_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    cvIn_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (auto& s : slots_) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        s.pipe.reset();
    }
    if (ring_) (void)hipHostFree(ring_);
}

uint64_t Server::submit(const uint8_t* desc, uint32_t n) {
    RBOW_REQUIRE(desc != nullptr || n == 0, RBOW_ERR_INVALID_ARG, "null descriptors");
    RBOW_REQUIRE(n <= Database::kMaxQueryDesc, RBOW_ERR_CAPACITY, "query above 16384 descriptors");
    Request r;
    r.arriveMs = nowMs();
    r.n = n;
    r.bytes = db_.scratchBytes(n, opt_.K) + uint64_t(n) * (kDims + 4);
    RBOW_REQUIRE(r.bytes <= opt_.memBudget, RBOW_ERR_CAPACITY, "query scratch exceeds the server's memory budget");
    const uint64_t len = uint64_t(n) * kDims;
    std::lock_guard<std::mutex> sub(subM_);
    {
        // reserve ring bytes; a request never wraps (the bytes skipped at the
        // end of the ring belong to its span and are released with it)
        std::unique_lock<std::mutex> lk(m_);
        RBOW_REQUIRE(error_.empty(), RBOW_ERR_HIP, "server stopped: " + error_);
        uint64_t beg = ringHead_;
        if (len && beg % ringCap_ + len > ringCap_) beg += ringCap_ - beg % ringCap_;
        cvRing_.wait(lk, [&] { return beg + len - ringTail_ <= ringCap_ || !error_.empty(); });
        RBOW_REQUIRE(error_.empty(), RBOW_ERR_HIP, "server stopped: " + error_);
        ringHead_ = beg + len;
        r.ringOff = beg;
        r.span = spanBase_ + spans_.size();
        spans_.push_back(Span{beg + len, false});
    }
    if (len) std::memcpy(ring_ + r.ringOff % ringCap_, desc, len);  // outside the lock: the bytes are ours
    uint64_t t;
    {
        std::lock_guard<std::mutex> lk(m_);
        t = r.ticket = nextTicket_++;
        pending_.push_back(r);
        nPending_.fetch_add(1, std::memory_order_release);
    }
    cvIn_.notify_one();
    return t;
}

Server::Result Server::wait(uint64_t ticket) {
    std::unique_lock<std::mutex> lk(m_);
    cvOut_.wait(lk, [&] { return done_.count(ticket) || !error_.empty(); });
    auto it = done_.find(ticket);
    RBOW_REQUIRE(it != done_.end(), RBOW_ERR_HIP, "server stopped: " + error_);
    Result r = std::move(it->second);
    done_.erase(it);
    return r;
}

void Server::reap(Slot& s) {
    const double doneMs = nowMs();
    const uint32_t B = uint32_t(s.reqs.size());
    const QueryPipeline& P = *s.pipe;
    const uint8_t* res = P.hostResults();
    {
        std::lock_guard<std::mutex> lk(m_);
        for (uint32_t q = 0; q < B; q++) {
            Result r;
            const uint32_t c = P.count(res, q);
            r.docs.assign(P.docs(res, q), P.docs(res, q) + c);
            r.scores.assign(P.scores(res, q), P.scores(res, q) + c);
            r.arriveMs = s.reqs[q].arriveMs;
            r.startMs = s.startMs;
            r.doneMs = doneMs;
            r.batchSize = B;
            done_[s.reqs[q].ticket] = std::move(r);
            spans_[s.reqs[q].span - spanBase_].done = true;
        }
        inFlightBytes_ -= s.bytes;
        while (!spans_.empty() && spans_.front().done) {  // release the ring in reservation order
            ringTail_ = spans_.front().end;
            spans_.pop_front();
            spanBase_++;
        }
    }
    s.reqs.clear();
    s.busy = false;
    cvOut_.notify_all();
    cvRing_.notify_all();
}

void Server::launch(Slot& s, std::vector<Request>&& batch) {
    const uint32_t B = uint32_t(batch.size());
    s.hOff.assign(B + 1, 0);
    for (uint32_t q = 0; q < B; q++) s.hOff[q + 1] = s.hOff[q] + batch[q].n;
    hipStream_t st = s.stream;
    uint8_t* dDesc = s.pipe->descBuffer();
    // H2D straight from the staging ring: one copy per run of adjacent requests
    for (uint32_t q = 0; q < B;) {
        uint32_t e = q + 1;
        while (e < B && batch[e].ringOff == batch[e - 1].ringOff + uint64_t(batch[e - 1].n) * kDims &&
               batch[e].ringOff % ringCap_ != 0)
            e++;
        const uint64_t bytes = (s.hOff[e] - s.hOff[q]) * kDims;
        if (bytes)
            RBOW_HIP(hipMemcpyAsync(dDesc + s.hOff[q] * kDims, ring_ + batch[q].ringOff % ringCap_, bytes,
                                    hipMemcpyHostToDevice, st));
        q = e;
    }
    // lookup, scoring, top-K and the results' download: one graph launch per batch
    s.pipe->run(s.hOff.data(), B, opt_.lookupKernel, opt_.useGraphs, true, st);
    RBOW_HIP(hipEventRecord(s.done, st));
    s.reqs = std::move(batch);
    s.startMs = nowMs();
    s.busy = true;
}

void Server::loop() {
    try {
        for (;;) {
            for (auto& s : slots_) {
                if (!s.busy) continue;
                const hipError_t e = hipEventQuery(s.done);
                if (e == hipSuccess) reap(s);
                else if (e != hipErrorNotReady) RBOW_HIP(e);
            }
            Slot* free = nullptr;
            bool anyBusy = false;
            for (auto& s : slots_) {
                if (!s.busy && !free) free = &s;
                anyBusy |= s.busy;
            }
            std::vector<Request> batch;
            {
                std::unique_lock<std::mutex> lk(m_);
                if (stop_ && pending_.empty() && !anyBusy) return;
                if (free) {
                    // admission: batch slots, the slot's descriptor capacity and the HBM scratch budget
                    uint64_t desc = 0;
                    while (!pending_.empty() && batch.size() < opt_.maxBatch &&
                           desc + pending_.front().n <= opt_.batchMaxDesc &&
                           inFlightBytes_ + pending_.front().bytes <= opt_.memBudget) {
                        inFlightBytes_ += pending_.front().bytes;
                        desc += pending_.front().n;
                        batch.push_back(std::move(pending_.front()));
                        pending_.pop_front();
                        nPending_.fetch_sub(1, std::memory_order_relaxed);
                    }
                }
                if (batch.empty() && !anyBusy) {
                    // idle: spin briefly, then block until a request (or stop) arrives
                    lk.unlock();
                    const auto until = Clock::now() + std::chrono::microseconds(kIdleSpinUs);
                    while (nPending_.load(std::memory_order_acquire) == 0 && Clock::now() < until)
                        std::this_thread::yield();
                    lk.lock();
                    cvIn_.wait(lk, [&] { return stop_ || !pending_.empty(); });
                    continue;
                }
            }
            if (!batch.empty()) {
                uint64_t bytes = 0;
                for (const auto& r : batch) bytes += r.bytes;
                launch(*free, std::move(batch));
                free->bytes = bytes;
            } else {
                std::this_thread::yield();  // batches in flight: poll their events (a sleep overshoots by ~50 us)
            }
        }
    } catch (const std::exception& ex) {
        std::lock_guard<std::mutex> lk(m_);
        error_ = ex.what();
        cvOut_.notify_all();
        cvRing_.notify_all();
    }
}

}  // namespace rbow_amd
