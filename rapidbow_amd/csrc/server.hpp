// server.hpp -- continuous-batching query server over a resident vocabulary
// and inverted file.  No counterpart in the reference, which answers one host
// query at a time (DataBase.cpp:71-190).
//
// Requests (descriptors on the host) are copied by the submitting thread into
// a pinned staging ring, then admitted into a batch while batch slots and the
// HBM scratch budget allow; a worker thread launches each batch -- H2D
// straight from the ring, then lookup, scoring, top-K and D2H through the
// slot's QueryPipeline (one captured HIP graph launch per batch, or direct
// launches) -- on one of several HIP streams, so batches formed from later
// arrivals run while earlier ones are still on the device, and the worker
// itself copies no descriptors.  Every
// request's result is independent of what it was batched with (per-query
// accumulators), which the tests check.
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include <memory>

#include "pipeline.hpp"

namespace rbow_amd {

struct ServerOptions {
    uint32_t maxBatch = 64;        // queries per batch
    uint32_t nbStreams = 2;        // batches in flight
    uint64_t memBudget = 4ull << 30;  // HBM scratch of in-flight queries (bytes)
    uint32_t K = 50;
    int lookupKernel = RBOW_LOOKUP_AUTO;  // RBOW_LOOKUP_AUTO / _MFMA / _PLAIN, per batch
    uint64_t stagingBytes = 256ull << 20;  // pinned ring of submitted descriptors
    bool useGraphs = true;         // each batch as one launch of a captured graph (per size class)
    uint64_t batchMaxDesc = 0;     // descriptors per batch (0: min(maxBatch * 16384, 2^20), at least 16384)
};

class Server {
public:
    using Clock = std::chrono::steady_clock;
    struct Result {
        std::vector<uint32_t> docs;
        std::vector<float> scores;
        double arriveMs = 0, startMs = 0, doneMs = 0;  // since server creation
        uint32_t batchSize = 0;
    };

    Server(const Vocabulary& voc, const Database& db, const ServerOptions& o);
    ~Server();
    // copies the descriptors into the staging ring (waiting for room if it is
    // full of queries still in flight); returns the ticket
    uint64_t submit(const uint8_t* desc, uint32_t n);
    // blocks until the ticket's result is ready, then removes it
    Result wait(uint64_t ticket);
    double nowMs() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0_).count(); }
    const ServerOptions& options() const { return opt_; }

private:
    struct Request {
        uint64_t ticket;
        uint64_t span;     // index of its staging span
        uint64_t ringOff;  // its descriptors: ring bytes [ringOff, ringOff + n*128) (monotonic offsets)
        uint32_t n;
        double arriveMs;
        uint64_t bytes;    // HBM scratch
    };
    struct Span {          // staging ring bytes up to `end` (monotonic), owned by one request
        uint64_t end;
        bool done;
    };
    struct Slot {  // one in-flight batch
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        bool busy = false;
        std::vector<Request> reqs;
        std::vector<uint64_t> hOff;
        std::unique_ptr<QueryPipeline> pipe;  // device buffers, graphs, pinned results of this slot's batches
        uint64_t bytes = 0;
        double startMs = 0;
    };
    // the worker spins this long for a new request before it blocks (a thread
    // wake-up costs tens of microseconds)
    static constexpr int kIdleSpinUs = 500;
    void loop();
    void launch(Slot& s, std::vector<Request>&& batch);
    void reap(Slot& s);

    const Vocabulary& voc_;
    const Database& db_;
    ServerOptions opt_;
    Clock::time_point t0_;
    std::vector<Slot> slots_;
    std::mutex m_;
    std::condition_variable cvIn_, cvOut_;
    std::deque<Request> pending_;
    std::mutex subM_;             // submitters one at a time: ring order = queue order
    std::condition_variable cvRing_;
    uint8_t* ring_ = nullptr;     // pinned staging ring of ringCap_ bytes
    uint64_t ringCap_ = 0, ringHead_ = 0, ringTail_ = 0;  // monotonic offsets: next free, oldest live
    std::deque<Span> spans_;      // live spans in reservation order
    uint64_t spanBase_ = 0;       // index of spans_.front()
    std::unordered_map<uint64_t, Result> done_;
    uint64_t nextTicket_ = 1;
    std::atomic<uint32_t> nPending_{0};  // pending_.size(), read by the idle worker without the lock
    uint64_t inFlightBytes_ = 0;
    bool stop_ = false;
    std::string error_;
    std::thread worker_;
};

}  // namespace rbow_amd
