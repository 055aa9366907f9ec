// vocabulary.cpp -- hierarchical k-means build driver, serialisation.
//
// Level loop of Vocabulary.cpp:67-99: every node of a level is one k-means
// task; all tasks of the level run in the same kernel launches (no 128-task
// batches, no per-batch device sync).  Convergence is tracked per task on the
// device; the host only polls a counter when the pass cap is large.
#include "vocabulary.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <unordered_map>
#include <unordered_set>

namespace rbow_amd {

void Vocabulary::allocate(uint32_t b_, uint32_t L_, uint32_t ct) {
    b = b_;
    lg2b = b_ == 16 ? 4 : 5;
    L = L_;
    centroidType = ct;
    nbNodes = levelBeg(lg2b, L);
    const uint64_t nc = nbNodes * b;
    if (ct == 0) {
        u8rows.resize(nc * kDims);
        normU.resize(nc);
        normS.resize(nc);
    } else {
        f32rows.resize(nc * kDims);
        split.resize(3 * nc * kDims);
        cn2.resize(nc);
        e1.resize(nc);
        e2.resize(nc);
    }
    masks.assign(nbNodes, 0u);
    weights.assign(nc, 0.0f);
}

void Vocabulary::prepare(hipStream_t s) {
    if (centroidType == 0) launchPrepU8Tree(u8rows.get(), nbChildren(), normU.get(), normS.get(), s);
    else {
        launchPrepF32Split(f32rows.get(), nbNodes, b, split.get(), nbChildren() * kDims, cn2.get(), e1.get(), e2.get(), s);
        launchMarkDuplicateChildren(f32rows.get(), nbNodes, b, cn2.get(), s);
    }
}

TreeView Vocabulary::view() const {
    TreeView t;
    t.b = b;
    t.lg2b = lg2b;
    t.L = L;
    t.centroidType = centroidType;
    t.u8rows = u8rows.get();
    t.normU = normU.get();
    t.normS = normS.get();
    t.f32rows = f32rows.get();
    t.split = split.get();
    t.splitStride = nbChildren() * kDims;
    t.cn2 = cn2.get();
    t.e1 = e1.get();
    t.e2 = e2.get();
    t.autoMin = lookupAutoMin;
    t.fusedMax = lookupFusedMax;
    t.leafItemRows = lookupLeafItemRows;
    return t;
}

namespace {

uint32_t deviceCUs() {
    int d = 0, cus = 0;
    RBOW_HIP(hipGetDevice(&d));
    RBOW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d));
    return uint32_t(std::max(cus, 1));
}

// kmeans.cpp:141-168 randPick: pick min(randSpace, nbToPick) distinct members
// with the persistent engine, sorted; pad with the last pick (0 if none).
uint32_t randPick(std::mt19937_64& rng, uint32_t* dst, uint32_t randSpace, uint32_t nbToPick) {
    const uint32_t nbPick = std::min(randSpace, nbToPick);
    std::unordered_set<uint32_t> results;
    if (nbPick * 16 < randSpace) {
        while (results.size() != nbPick) results.insert(uint32_t(rng() % randSpace));
        std::copy_n(results.begin(), nbPick, dst);
        std::sort(dst, dst + nbPick);
    } else {
        std::vector<uint32_t> space(randSpace);
        std::iota(space.begin(), space.end(), 0u);
        std::shuffle(space.begin(), space.end(), rng);
        std::copy_n(space.begin(), nbPick, dst);
        std::sort(dst, dst + nbPick);
    }
    if (nbPick < nbToPick) std::fill(dst + nbPick, dst + nbToPick, nbPick != 0 ? dst[nbPick - 1u] : 0u);
    return nbPick;
}

template <typename T>
void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) RBOW_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

// Owning pinned host allocation that only grows.
class PinnedBuf {
public:
    PinnedBuf() = default;
    ~PinnedBuf() {
        if (p_) (void)hipHostFree(p_);
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    uint8_t* reserve(size_t bytes) {
        if (bytes > n_) {
            if (p_) (void)hipHostFree(p_);
            p_ = nullptr;
            n_ = 0;
            RBOW_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), bytes));
            n_ = bytes;
        }
        return p_;
    }

private:
    uint8_t* p_ = nullptr;
    size_t n_ = 0;
};

// One level's host -> device metadata, laid out in a pinned host arena and sent
// with a single asynchronous copy into a device arena of the same layout.  The
// host side is rewritten only once the previous level's copy has completed.
class UploadArena {
public:
    UploadArena(DevBuf<uint8_t>& dev, PinnedBuf& host, size_t bytes) : cap_(bytes) {
        dev.reserve(std::max<size_t>(bytes, 1));
        dev_ = dev.get();
        host_ = host.reserve(std::max<size_t>(bytes, 1));
    }
    void reset() { used_ = 0; }
    // room for n elements: host pointer to fill, *dev = the same bytes on the device
    template <typename T>
    T* take(size_t n, const T** dev) {
        const size_t off = (used_ + 255) & ~size_t(255);
        RBOW_REQUIRE(off + n * sizeof(T) <= cap_, RBOW_ERR_HIP, "level upload arena overflow");
        used_ = off + n * sizeof(T);
        *dev = reinterpret_cast<const T*>(dev_ + off);
        return reinterpret_cast<T*>(host_ + off);
    }
    template <typename T>
    const T* put(const T* src, size_t n) {
        const T* dev = nullptr;
        T* h = take<T>(n, &dev);
        if (n) std::memcpy(h, src, n * sizeof(T));
        return dev;
    }
    void upload(hipStream_t s) {
        if (used_) RBOW_HIP(hipMemcpyAsync(dev_, host_, used_, hipMemcpyHostToDevice, s));
    }
    static size_t bytesFor(std::initializer_list<size_t> regionBytes) {
        size_t t = 0;
        for (size_t r : regionBytes) t += r + 256;
        return t;
    }

private:
    uint8_t* dev_ = nullptr;
    uint8_t* host_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};

// The device and pinned buffers of a build.  After a build that completes, they
// stay with the device for the next build (grown when a build needs more), so
// a repeated build allocates nothing but the tree itself;
// releaseBuildWorkspaces() frees them.
struct BuildWs {
    DevBuf<uint32_t> order, orderNext, counts, changed, active, nbActive, seed, partCnt, partChanged, out;
    DevBuf<uint32_t> planTaskBeg, planTaskCnt, planTileBeg, planNbTiles;
    DevBuf<KmTile> planTiles;
    DevBuf<uint8_t> X, assign, sortWs, upDev;
    DevBuf<int8_t> fSplit;
    DevBuf<long long> sums;
    DevBuf<int8_t> XA;
    DevBuf<int32_t> partSums;
    DevBuf<float> fCn2, fE1, fE2;
    PinnedBuf hOut, upHost;
};

struct BuildWsCache {
    std::mutex mtx;
    std::unordered_map<int, std::unique_ptr<BuildWs>> byDevice;
};
// never destroyed: freeing device memory from a static destructor may run after the runtime's teardown
BuildWsCache& buildWsCache() {
    static BuildWsCache* c = new BuildWsCache;
    return *c;
}

// The calling thread's device's cached workspace, or a new one when there is
// none (or another build on the device holds it).
std::unique_ptr<BuildWs> takeBuildWs(int device) {
    BuildWsCache& c = buildWsCache();
    std::lock_guard<std::mutex> lk(c.mtx);
    auto it = c.byDevice.find(device);
    if (it != c.byDevice.end() && it->second) return std::move(it->second);
    return std::unique_ptr<BuildWs>(new BuildWs);
}

// Back into the cache after a build whose work on the device has completed.
void returnBuildWs(int device, std::unique_ptr<BuildWs> w) {
    BuildWsCache& c = buildWsCache();
    std::lock_guard<std::mutex> lk(c.mtx);
    std::unique_ptr<BuildWs>& slot = c.byDevice[device];
    if (!slot) slot = std::move(w);
}

}  // namespace

void releaseBuildWorkspaces() {
    std::unordered_map<int, std::unique_ptr<BuildWs>> drop;
    {
        BuildWsCache& c = buildWsCache();
        std::lock_guard<std::mutex> lk(c.mtx);
        drop.swap(c.byDevice);
    }
    int cur = 0;
    RBOW_HIP(hipGetDevice(&cur));
    for (auto& e : drop) {
        RBOW_HIP(hipSetDevice(e.first));
        e.second.reset();
    }
    RBOW_HIP(hipSetDevice(cur));
}

Vocabulary* buildVocabulary(const uint8_t* devDesc, uint32_t nbDesc, uint32_t nbDoc, uint32_t b, uint32_t L,
                            const rbow_build_options& opts, hipStream_t s, BuildStats* stats) {
    RBOW_REQUIRE(nbDesc >= b, RBOW_ERR_INSUFFICIENT_DESC, "Insufficient descriptors");  // Vocabulary.cpp:39-41
    RBOW_REQUIRE(nbDoc > 0, RBOW_ERR_NB_DOC, "nbDoc must be positive");                 // Vocabulary.cpp:42-44
    RBOW_REQUIRE(opts.max_iters > 0, RBOW_ERR_INVALID_ARG, "max_iters must be positive");
    RBOW_REQUIRE(opts.seed_mode <= 1, RBOW_ERR_INVALID_ARG, "seed_mode must be 0 or 1");
    RBOW_REQUIRE(L >= 1 && uint64_t(L) * (b == 16 ? 4 : 5) <= 28, RBOW_ERR_CAPACITY,
                 "tree too deep (at most 2^28 leaves)");
    const uint32_t ct = opts.centroid_type;
    const int kernel = int(opts.kernel);
    const auto t0 = std::chrono::steady_clock::now();
    // per-launch timing (opts.timing), with HIP events recorded by the dispatches themselves: [0, 1] the
    // start of the first persistent launch and the end of the last; 2 + 6 * pass + i: start / end of the
    // assign (from a marker before the filter operands' preparation on the float path), of the reduce,
    // of the update
    const bool timing = stats && opts.timing;
    std::vector<hipEvent_t> ev;
    struct EventsFree {
        std::vector<hipEvent_t>& v;
        ~EventsFree() {
            for (auto e : v) (void)hipEventDestroy(e);
        }
    } evGuard{ev};
    auto evAt = [&](size_t i) {
        while (ev.size() <= i) {
            hipEvent_t e;
            RBOW_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        return ev[i];
    };
    auto timer = [&](size_t i) { return timing ? KernelTimer{evAt(i), evAt(i + 1)} : KernelTimer{}; };
    std::vector<uint8_t> ranAssign, ranUpdate;  // per pass of the current level
    double assignMs = 0, updateMs = 0, reduceMs = 0, streamMs = 0, persistTotalMs = 0;
    uint32_t streamLaunches = 0;
    uint64_t streamPointPasses = 0;
    uint64_t pointPasses = 0;

    std::unique_ptr<Vocabulary> voc(new Vocabulary);
    voc->allocate(b, L, ct);
    const uint32_t lg2b = voc->lg2b;
    std::vector<uint32_t> passesAll(voc->nbNodes, 0u);

    // host grouping state: the level order lives on the device; the host keeps a
    // copy only for the reference-mode seeds, which index it
    std::vector<uint32_t> order;
    if (opts.seed_mode == 0) {
        order.resize(nbDesc);
        std::iota(order.begin(), order.end(), 0u);
    }
    std::vector<uint32_t> taskBeg{0u}, taskCnt{nbDesc};

    // ---- every device buffer sized once for the largest level ----
    const bool mfmaPath = ct == 0 && kernel == 0;
    const uint64_t nGmax = uint64_t(1) << (lg2b * (L - 1));  // tasks of the last level
    const uint64_t tilesMax = divUp(nbDesc, 64u) + nGmax;   // each task adds at most one partial tile
    const uint64_t blocksMax = mfmaPath ? uint64_t(kmAssignResidentBlocks()) : 0;
    const uint64_t wavesMax = blocksMax * kmAssignWavesPerBlock();
    const uint64_t chunksMax = mfmaPath ? wavesMax + nGmax : 0;  // a wave's range splits at task boundaries
    const uint64_t slotsMax = blocksMax * kmAssignMergeSlots();
    const uint64_t redMax = mfmaPath ? chunksMax / kReduceItemParts + nGmax : 0;
    const uint32_t persistTiles = mfmaPath ? std::min(opts.persist_max_tiles, lloydTasksTileCap(1)) : 0u;
    const uint32_t persistTiles2 = mfmaPath ? std::min(persistTiles, lloydTasksTileCap(2)) : 0u;  // 2 per CU
    int device = 0;
    RBOW_HIP(hipGetDevice(&device));
    std::unique_ptr<BuildWs> wsOwner = takeBuildWs(device);
    BuildWs& ws = *wsOwner;
    ws.order.reserve(nbDesc);
    ws.orderNext.reserve(nbDesc);
    ws.X.reserve(mfmaPath ? 0 : uint64_t(nbDesc) * kDims);
    ws.assign.reserve(nbDesc);
    ws.counts.reserve(nGmax * b);
    ws.changed.reserve(nGmax);
    ws.active.reserve(nGmax);
    ws.nbActive.reserve(1);
    ws.seed.reserve(nGmax * b);
    ws.sums.reserve(nGmax * b * kDims);
    ws.XA.reserve(mfmaPath ? tilesMax * 8192 : 0);
    ws.partSums.reserve(chunksMax * 16 * kDims);
    ws.partCnt.reserve(chunksMax * 16);
    ws.partChanged.reserve(chunksMax);
    ws.sortWs.reserve(regroupWorkspaceBytes(tilesMax, b));
    // float centroids with kernel 0: the MFMA filter's int8 split of the centers, redone every pass
    const bool f32Mfma = ct == 1 && kernel == 0;
    ws.fSplit.reserve(f32Mfma ? 3 * nGmax * b * kDims : 0);
    ws.fCn2.reserve(f32Mfma ? nGmax * b : 0);
    ws.fE1.reserve(f32Mfma ? nGmax * b : 0);
    ws.fE2.reserve(f32Mfma ? nGmax * b : 0);
    // per-level results, downloaded with one copy: [hist nG*b][passes nG][masks nG][IDF weights nG*b]
    ws.out.reserve(2 * nGmax * b + 2 * nGmax);
    uint32_t* hOut = reinterpret_cast<uint32_t*>(ws.hOut.reserve(4 * (2 * nGmax * b + 2 * nGmax + 1)));
    uint32_t* hNbActive = hOut + 2 * nGmax * b + 2 * nGmax;
    // the next level's plan on the device (k-means++ seeding only: the reference picks index the host order)
    const bool devicePlan = opts.seed_mode == 1 && L > 1 && nGmax <= kPlanMaxTasks;
    if (devicePlan) {
        ws.planTaskBeg.reserve(nGmax);
        ws.planTaskCnt.reserve(nGmax);
        ws.planTileBeg.reserve(nGmax);
        ws.planNbTiles.reserve(1);
        ws.planTiles.reserve(mfmaPath ? tilesMax : 0);
    }
    DevBuf<uint32_t>& dOrder = ws.order;
    DevBuf<uint32_t>& dOrderNext = ws.orderNext;
    UploadArena up(ws.upDev, ws.upHost,
                   UploadArena::bytesFor({4 * nGmax, 4 * nGmax, sizeof(KmTile) * tilesMax, 4 * nGmax, 4 * nGmax,
                                          4 * nGmax, 4 * (wavesMax + 1), sizeof(KmChunk) * chunksMax,
                                          sizeof(KmSlot) * slotsMax, 4 * nGmax, 4 * nGmax, sizeof(KmPartRange) * redMax,
                                          opts.seed_mode == 0 ? 4 * nGmax * b : 0}));
    hipEvent_t outReady = nullptr;  // a level's results are in hOut
    RBOW_HIP(hipEventCreateWithFlags(&outReady, hipEventDisableTiming));
    struct EventFree {
        hipEvent_t e;
        ~EventFree() { (void)hipEventDestroy(e); }
    } outGuard{outReady};
    bool preLaunched = false;  // this level's gather, seeding and reset are already on the stream

    std::vector<uint32_t> taskTiles, tileBeg, smallTasks, midTasks;
    std::vector<KmPartRange> red;
    std::mt19937_64 rng(opts.seed);
    const uint32_t cus = deviceCUs();

    for (uint32_t level = 0; level < L; level++) {
        const uint32_t nG = uint32_t(taskBeg.size());
        const uint64_t nodeBase = levelBeg(lg2b, level);
        up.reset();
        const uint32_t* dTaskBeg = up.put(taskBeg.data(), nG);
        const uint32_t* dTaskCnt = up.put(taskCnt.data(), nG);
        // ---- tiles of <= 64 rows, tasks contiguous (DescGrouping, kmeans.cpp:27-42) ----
        taskTiles.assign(nG, 0u);
        tileBeg.assign(nG, 0u);
        uint32_t nbTiles = 0;
        for (uint32_t g = 0; g < nG; g++) {
            taskTiles[g] = divUp(taskCnt[g], 64u);
            tileBeg[g] = nbTiles;
            nbTiles += taskTiles[g];
        }
        const KmTile* dTiles = nullptr;
        KmTile* hTiles = up.take<KmTile>(nbTiles, &dTiles);
        for (uint32_t g = 0, t = 0; g < nG; g++)
            for (uint32_t r = 0; r < taskCnt[g]; r += 64) hTiles[t++] = KmTile{g, taskBeg[g] + r, std::min(64u, taskCnt[g] - r)};

        KmLevel k;
        k.b = b;
        k.nbTasks = nG;
        k.nbTiles = nbTiles;
        k.centroidType = ct;
        k.X = ws.X.get();
        k.tiles = dTiles;
        k.taskBeg = dTaskBeg;
        k.taskCnt = dTaskCnt;
        if (ct == 0) k.centersU8 = voc->u8rows.get() + nodeBase * b * kDims;
        else k.centersF32 = voc->f32rows.get() + nodeBase * b * kDims;
        k.assign = ws.assign.get();
        k.counts = ws.counts.get();
        k.sums = ws.sums.get();
        k.changed = ws.changed.get();
        k.active = ws.active.get();
        uint32_t* dHist = ws.out.get();
        k.passes = dHist + uint64_t(nG) * b;
        uint32_t* dMasks = k.passes + nG;
        k.nbActive = ws.nbActive.get();
        if (f32Mfma) {
            k.fSplit = ws.fSplit.get();
            k.fSplitStride = uint64_t(nG) * b * kDims;
            k.fCn2 = ws.fCn2.get();
            k.fE1 = ws.fE1.get();
            k.fE2 = ws.fE2.get();
        }

        // Tasks of <= persistTiles tiles run their whole Lloyd loop in one launch with
        // their points in LDS (k_lloydTasks), when there are enough of them to fill
        // the device; the others stream, one launch per pass: equal tile ranges per
        // resident wave, cut into per-task chunks; reduce items of <= 32 chunks.
        smallTasks.clear();
        uint32_t smallTilesMax = 0, nbLarger = 0;  // smallTasks[0, nbLarger): more than persistTiles2 tiles
        const uint32_t* dSmall = nullptr;
        const uint32_t* dTileBeg = nullptr;
        midTasks.clear();
        if (mfmaPath) {
            if (persistTiles) {
                for (uint32_t g = 0; g < nG; g++)
                    if (taskTiles[g] && taskTiles[g] <= persistTiles) smallTasks.push_back(g);
                if (smallTasks.size() < cus) smallTasks.clear();
                // large tasks first in the list: they start early and the short ones fill in
                std::stable_sort(smallTasks.begin(), smallTasks.end(),
                                 [&](uint32_t x, uint32_t y) { return taskTiles[x] > taskTiles[y]; });
                for (uint32_t g : smallTasks) {
                    smallTilesMax = std::max(smallTilesMax, taskTiles[g]);
                    nbLarger += taskTiles[g] > persistTiles2;
                    taskTiles[g] = 0;  // not streamed
                }
                dTileBeg = up.put(tileBeg.data(), nG);
                // When every task left to stream has at most kStreamTaskTiles tiles, those tasks
                // run their whole Lloyd loop in the persistent launch as well (tiles read through
                // L2 each pass) and the level needs no launch per pass.
                midTasks.clear();
                uint32_t leftMax = 0;
                for (uint32_t g = 0; g < nG; g++) leftMax = std::max(leftMax, taskTiles[g]);
                if (leftMax && leftMax <= kStreamTaskTiles) {
                    for (uint32_t g = 0; g < nG; g++)
                        if (taskTiles[g]) midTasks.push_back(g);
                    std::stable_sort(midTasks.begin(), midTasks.end(),  // largest first: they bound the launch
                                     [&](uint32_t x, uint32_t y) { return taskTiles[x] > taskTiles[y]; });
                    for (uint32_t g : midTasks) taskTiles[g] = 0;  // not streamed per pass
                }
                // one launch list: the streamed persistent tasks, then the LDS ones
                std::vector<uint32_t> list(midTasks);
                list.insert(list.end(), smallTasks.begin(), smallTasks.end());
                dSmall = up.put(list.data(), list.size());
            }
            const KmPartition part = partitionKmTiles(taskTiles, tileBeg);
            red.clear();
            for (uint32_t g = 0; g < nG; g++)
                if (part.taskPartCnt[g] > kDirectParts)  // tasks with few partials are reduced inside the update kernel
                    for (uint32_t c = 0; c < part.taskPartCnt[g]; c += kReduceItemParts)
                        red.push_back(KmPartRange{g, part.taskPartBeg[g] + c,
                                                  part.taskPartBeg[g] + std::min(part.taskPartCnt[g], c + kReduceItemParts)});
            RBOW_REQUIRE(part.chunks.size() <= chunksMax && part.blockSlots.size() <= slotsMax && red.size() <= redMax,
                         RBOW_ERR_HIP, "k-means work partition exceeds its bound");
            k.XA = ws.XA.get();
            k.waveChunkBeg = up.put(part.waveChunkBeg.data(), part.waveChunkBeg.size());
            k.nbWaves = part.nbWaves;
            k.chunks = up.put(part.chunks.data(), part.chunks.size());
            k.nbChunks = uint32_t(part.chunks.size());
            k.blockSlots = up.put(part.blockSlots.data(), part.blockSlots.size());
            k.nbParts = part.nbParts;
            k.taskPartBeg = up.put(part.taskPartBeg.data(), nG);
            k.taskPartCnt = up.put(part.taskPartCnt.data(), nG);
            k.partSums = ws.partSums.get();
            k.partCnt = ws.partCnt.get();
            k.partChanged = ws.partChanged.get();
            if (!smallTasks.empty() || !midTasks.empty()) {  // the update visits the tasks that stream (and the empty ones)
                const uint64_t nbPersist = smallTasks.size() + midTasks.size();
                const uint32_t* dUpd = nullptr;
                uint32_t* hUpd = up.take<uint32_t>(nG - nbPersist, &dUpd);
                uint32_t nbUpd = 0;
                for (uint32_t g = 0; g < nG; g++)
                    if (taskTiles[g] || !taskCnt[g]) hUpd[nbUpd++] = g;
                RBOW_REQUIRE(nbUpd == nG - nbPersist, RBOW_ERR_HIP, "update task list size");
                k.updTasks = dUpd;
                k.nbUpdTasks = nbUpd;
            }
        }
        const KmPartRange* dRed = up.put(red.data(), red.size());

        // ---- seeds (kmeans.cpp:170-196 reference picks, or k-means++) ----
        const uint32_t* dSeedIdx = ws.seed.get();
        if (opts.seed_mode == 0) {
            uint32_t* seedIdx = up.take<uint32_t>(uint64_t(nG) * b, &dSeedIdx);
            uint32_t pos[32];
            for (uint32_t g = 0; g < nG; g++) {
                const uint32_t n = taskCnt[g];
                const uint32_t nValid = randPick(rng, pos, n, b);
                for (uint32_t j = 0; j < b; j++) {
                    uint32_t idx = kIdxInvalid;
                    if (n > 0) {
                        idx = order[taskBeg[g] + pos[j]];
                    } else {
                        // empty group: the reference reads the next group's first
                        // index (kmeans.cpp:180-186); none left -> invalid (0xFF row)
                        for (uint32_t h = g + 1; h < nG; h++)
                            if (taskCnt[h]) {
                                idx = order[taskBeg[h]];
                                break;
                            }
                    }
                    seedIdx[uint64_t(g) * b + j] = idx;
                }
                voc->masks[nodeBase + g] = nValid >= 32 ? 0xFFFFFFFFu : ((1u << nValid) - 1u);
            }
        }
        up.upload(s);
        if (!preLaunched) {
            if (level == 0) launchIota(dOrder.get(), nbDesc, s);
            if (mfmaPath) launchGatherTiles(devDesc, nbDesc, dOrder.get(), dTiles, nbTiles, ws.XA.get(), s);
            else launchGatherRows(devDesc, dOrder.get(), nbDesc, ws.X.get(), s, nbDesc);
            if (opts.seed_mode == 1)
                launchKmeansPP(devDesc, nbDesc, dOrder.get(), dTaskBeg, dTaskCnt, nG, b, opts.seed, level, ws.seed.get(), dMasks,
                               s);
            launchInitCenters(k, devDesc, dSeedIdx, s);
            launchKmLevelReset(k, nbDesc, s);
        }

        // ---- Lloyd passes (kmeans.cpp:118-133) ----
        if (!smallTasks.empty() || !midTasks.empty()) {
            // one launch holds the streamed tasks and the LDS tasks of one workgroup per CU (largest
            // first); tasks that allow two workgroups per CU run in a launch of their own; timed
            // from the start of the first launch to the end of the second
            const uint32_t nbSmaller = uint32_t(smallTasks.size()) - nbLarger, nbMid = uint32_t(midTasks.size());
            const KernelTimer t = timer(0);
            launchLloydTasks(k, dSmall, nbMid + nbLarger, nbLarger ? smallTilesMax : 0u, 1, dTileBeg, opts.max_iters, s,
                             KernelTimer{t.start, nbSmaller ? nullptr : t.stop}, nbMid);
            launchLloydTasks(k, dSmall + nbMid + nbLarger, nbSmaller, std::min(smallTilesMax, persistTiles2), 2,
                             dTileBeg, opts.max_iters, s, KernelTimer{nbLarger || nbMid ? nullptr : t.start, t.stop});
        }
        // with no streamed task left, one update pass retires the empty tasks
        const bool streamed = !mfmaPath || k.nbWaves > 0;
        const uint32_t iters = streamed ? opts.max_iters : 1u;
        const bool poll = iters > 16;
        uint32_t nbLaunched = 0;
        const bool reduced = mfmaPath && !red.empty();
        ranAssign.assign(iters, 0);
        ranUpdate.assign(iters, 0);
        for (uint32_t it = 0; it < iters; it++) {
            const size_t e = 2 + 6 * size_t(it);
            KernelTimer ta = timer(e);
            if (f32Mfma) {  // the filter's operands of this pass's centers, timed with the assign
                if (timing) RBOW_HIP(hipEventRecord(ta.start, s));
                ta.start = nullptr;
                launchPrepF32Split(k.centersF32, nG, b, ws.fSplit.get(), uint64_t(nG) * b * kDims, ws.fCn2.get(),
                                   ws.fE1.get(), ws.fE2.get(), s, k.active);
            }
            k.reverse = it & 1u;  // alternate directions: a pass starts where the last one ended
            ranAssign[it] = launchKmAssign(k, kernel, s, ta);
            if (reduced) launchKmReduce(k, dRed, uint32_t(red.size()), s, timer(e + 2));
            if (poll) launchFillU32(ws.nbActive.get(), 0u, 1, s);
            ranUpdate[it] = launchKmUpdate(k, s, timer(e + 4));
            nbLaunched = it + 1;
            if (poll && (it % 4 == 3)) {  // kmeans.cpp:125-131 polls every 4 passes
                RBOW_HIP(hipMemcpyAsync(hNbActive, ws.nbActive.get(), 4, hipMemcpyDeviceToHost, s));
                RBOW_HIP(hipStreamSynchronize(s));
                if (*hNbActive == 0) break;
            }
        }
        // ---- stable regroup + occurrence on the device (Vocabulary.cpp:83-98), and the
        // children's IDF weights from the occurrence counts (Vocabulary.cpp:103-108) ----
        // (the last level needs the occurrence counts only)
        const bool last = level + 1 == L;
        launchRegroup(k, nbDesc, dOrder.get(), last ? nullptr : dOrderNext.get(), dHist, ws.sortWs.get(), ws.sortWs.size(),
                      s);
        if (!last) std::swap(dOrder, dOrderNext);
        float* dW = reinterpret_cast<float*>(dMasks + nG);
        launchIdfWeights(dHist, uint64_t(nG) * b, nbDoc, dW, s);
        const uint64_t outWords = 2 * uint64_t(nG) * b + 2 * nG;
        RBOW_HIP(hipMemcpyAsync(hOut, ws.out.get(), 4 * outWords, hipMemcpyDeviceToHost, s));
        if (opts.seed_mode == 0 && level + 1 < L) d2h(order.data(), dOrder.get(), nbDesc, s);
        RBOW_HIP(hipEventRecord(outReady, s));
        // While the host reads the results and plans the next level, the device plans the
        // next level's tasks from the histogram itself and runs what needs no more than that:
        // the gather, k-means++ seeding, the centers and the reset.  The tree's derived data
        // for lookups follows the last level.
        preLaunched = false;
        if (devicePlan && level + 1 < L) {
            const uint32_t nG2 = nG * b;
            const uint32_t tileBound = divUp(nbDesc, 64u) + nG2;
            launchPlanLevel(dHist, nG2, tileBound, ws.planTaskBeg.get(), ws.planTaskCnt.get(), ws.planTileBeg.get(),
                            ws.planNbTiles.get(), mfmaPath ? ws.planTiles.get() : nullptr, s);
            KmLevel k2 = k;
            k2.nbTasks = nG2;
            if (ct == 0) k2.centersU8 = voc->u8rows.get() + levelBeg(lg2b, level + 1) * b * kDims;
            else k2.centersF32 = voc->f32rows.get() + levelBeg(lg2b, level + 1) * b * kDims;
            k2.passes = dHist + uint64_t(nG2) * b;
            if (mfmaPath)
                launchGatherTiles(devDesc, nbDesc, dOrder.get(), ws.planTiles.get(), tileBound, ws.XA.get(), s,
                                  ws.planNbTiles.get());
            else
                launchGatherRows(devDesc, dOrder.get(), nbDesc, ws.X.get(), s, nbDesc);
            launchKmeansPP(devDesc, nbDesc, dOrder.get(), ws.planTaskBeg.get(), ws.planTaskCnt.get(), nG2, b, opts.seed,
                           level + 1, ws.seed.get(), k2.passes + nG2, s);
            launchInitCenters(k2, devDesc, ws.seed.get(), s);
            launchKmLevelReset(k2, nbDesc, s);
            preLaunched = true;
        }
        if (level + 1 == L) voc->prepare(s);
        RBOW_HIP(hipEventSynchronize(outReady));
        const uint32_t* childCnt = hOut;
        const uint32_t* hPasses = hOut + uint64_t(nG) * b;
        const uint32_t* hMasks = hPasses + nG;
        std::memcpy(voc->weights.data() + nodeBase * b, hMasks + nG, 4ull * nG * b);
        if (opts.seed_mode == 1)
            for (uint32_t g = 0; g < nG; g++) voc->masks[nodeBase + g] = hMasks[g];
        std::vector<uint8_t> persistent(nG, 0);  // tasks of the persistent launches, not of the streamed assign
        for (uint32_t g : smallTasks) persistent[g] = 1;
        for (uint32_t g : midTasks) persistent[g] = 1;
        for (uint32_t g = 0; g < nG; g++) {
            passesAll[nodeBase + g] = hPasses[g];
            pointPasses += uint64_t(taskCnt[g]) * hPasses[g];
            if (!persistent[g]) streamPointPasses += uint64_t(taskCnt[g]) * hPasses[g];
        }
        if (streamed) streamLaunches += nbLaunched;
        if (timing) {
            float pm = 0;
            if (!smallTasks.empty() || !midTasks.empty()) RBOW_HIP(hipEventElapsedTime(&pm, ev[0], ev[1]));
            for (uint32_t it = 0; it < nbLaunched; it++) {
                const size_t e = 2 + 6 * size_t(it);
                float a = 0, r = 0, u = 0;
                if (ranAssign[it]) RBOW_HIP(hipEventElapsedTime(&a, ev[e], ev[e + 1]));
                if (reduced) RBOW_HIP(hipEventElapsedTime(&r, ev[e + 2], ev[e + 3]));
                if (ranUpdate[it]) RBOW_HIP(hipEventElapsedTime(&u, ev[e + 4], ev[e + 5]));
                assignMs += a;
                reduceMs += r;
                updateMs += u;
                if (streamed) streamMs += a;
            }
            assignMs += pm;
            persistTotalMs += pm;
        }

        std::vector<uint32_t> childBeg(uint64_t(nG) * b);
        uint32_t run = 0;
        for (uint64_t c = 0; c < childBeg.size(); c++) {
            childBeg[c] = run;
            run += childCnt[c];
        }
        RBOW_REQUIRE(run == nbDesc, RBOW_ERR_HIP, "regroup lost rows");
        taskBeg.swap(childBeg);
        taskCnt.assign(childCnt, childCnt + uint64_t(nG) * b);
    }

    RBOW_HIP(hipStreamSynchronize(s));
    returnBuildWs(device, std::move(wsOwner));  // the stream is idle: the next build may reuse it
    if (stats) {
        stats->passes = std::move(passesAll);
        stats->assignMs = assignMs;
        stats->updateMs = updateMs;
        stats->reduceMs = reduceMs;
        stats->streamMs = streamMs;
        stats->streamLaunches = streamLaunches;
        stats->streamPointPasses = streamPointPasses;
        stats->persistMs = persistTotalMs;
        stats->pointPasses = pointPasses;
        stats->totalMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return voc.release();
}

// ------------------------------------------------------------ blobs -----
namespace {
constexpr uint32_t kMagic = 0x32574252u;  // "RBW2"
struct Rbw2Header {
    uint32_t magic, version, b, L, centroidType, reserved;
};
}  // namespace

std::vector<uint8_t> serialize(const Vocabulary& v) {
    const uint64_t n = v.nbNodes, nc = v.nbChildren();
    if (v.b == 16 && v.centroidType == 0) {
        // Vocabulary.cpp:175-197: [u32 L][Node x n][BitSet<16> x n][float[16] x n],
        // Node byte(dim k, child j) = (k/4)*64 + 4j + k%4 (Vocabulary.h:42).
        std::vector<uint8_t> rows(nc * kDims);
        if (!rows.empty()) RBOW_HIP(hipMemcpy(rows.data(), v.u8rows.get(), rows.size(), hipMemcpyDeviceToHost));
        std::vector<uint8_t> out(4 + n * 2048 + n * 2 + n * 64);
        std::memcpy(out.data(), &v.L, 4);
        uint8_t* nodes = out.data() + 4;
        for (uint64_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < 16; j++)
                for (uint32_t k = 0; k < kDims; k++)
                    nodes[i * 2048 + (k >> 2) * 64 + j * 4 + (k & 3)] = rows[(i * 16 + j) * kDims + k];
        uint8_t* m = nodes + n * 2048;
        for (uint64_t i = 0; i < n; i++) {
            const uint16_t mk = uint16_t(v.masks[i]);
            std::memcpy(m + 2 * i, &mk, 2);
        }
        std::memcpy(m + 2 * n, v.weights.data(), n * 64);
        return out;
    }
    const uint64_t rowBytes = nc * kDims * (v.centroidType == 0 ? 1 : 4);
    std::vector<uint8_t> out(sizeof(Rbw2Header) + rowBytes + n * 4 + nc * 4);
    const Rbw2Header h{kMagic, 1u, v.b, v.L, v.centroidType, 0u};
    std::memcpy(out.data(), &h, sizeof h);
    uint8_t* p = out.data() + sizeof h;
    if (rowBytes)
        RBOW_HIP(hipMemcpy(p, v.centroidType == 0 ? static_cast<const void*>(v.u8rows.get()) : v.f32rows.get(),
                           rowBytes, hipMemcpyDeviceToHost));
    p += rowBytes;
    std::memcpy(p, v.masks.data(), n * 4);
    std::memcpy(p + n * 4, v.weights.data(), nc * 4);
    return out;
}

Vocabulary* deserialize(const uint8_t* blob, size_t size) {
    RBOW_REQUIRE(blob != nullptr || size == 0, RBOW_ERR_INVALID_ARG, "null blob");
    RBOW_REQUIRE(size >= 4, RBOW_ERR_INVALID_BLOB, "invalid blob");  // Vocabulary.cpp:129-131
    std::unique_ptr<Vocabulary> v(new Vocabulary);
    uint32_t first;
    std::memcpy(&first, blob, 4);
    if (first == kMagic && size >= sizeof(Rbw2Header)) {
        Rbw2Header h;
        std::memcpy(&h, blob, sizeof h);
        RBOW_REQUIRE(h.version == 1 && (h.b == 16 || h.b == 32) && h.centroidType <= 1 &&
                         (h.centroidType == 1 || h.b == 16) && uint64_t(h.L) * (h.b == 16 ? 4 : 5) <= 28,
                     RBOW_ERR_INVALID_BLOB, "invalid blob");
        const uint32_t lg2b = h.b == 16 ? 4 : 5;
        const uint64_t n = levelBeg(lg2b, h.L), nc = n * h.b;
        const uint64_t rowBytes = nc * kDims * (h.centroidType == 0 ? 1 : 4);
        RBOW_REQUIRE(size == sizeof h + rowBytes + n * 4 + nc * 4, RBOW_ERR_INVALID_BLOB, "invalid blob");
        v->allocate(h.b, h.L, h.centroidType);
        const uint8_t* p = blob + sizeof h;
        if (rowBytes)
            RBOW_HIP(hipMemcpy(h.centroidType == 0 ? static_cast<void*>(v->u8rows.get()) : v->f32rows.get(), p,
                               rowBytes, hipMemcpyHostToDevice));
        p += rowBytes;
        std::memcpy(v->masks.data(), p, n * 4);
        std::memcpy(v->weights.data(), p + n * 4, nc * 4);
    } else {
        // Vocabulary.cpp:126-157
        const uint32_t L = first;
        RBOW_REQUIRE(L <= 7, RBOW_ERR_INVALID_BLOB, "invalid blob");
        const uint64_t n = levelBeg(4, L);
        RBOW_REQUIRE(size == 4 + n * 2048 + n * 2 + n * 64, RBOW_ERR_INVALID_BLOB, "invalid blob");
        v->allocate(16, L, 0);
        const uint8_t* nodes = blob + 4;
        std::vector<uint8_t> rows(n * 16 * kDims);
        for (uint64_t i = 0; i < n; i++)
            for (uint32_t j = 0; j < 16; j++)
                for (uint32_t k = 0; k < kDims; k++)
                    rows[(i * 16 + j) * kDims + k] = nodes[i * 2048 + (k >> 2) * 64 + j * 4 + (k & 3)];
        if (!rows.empty()) RBOW_HIP(hipMemcpy(v->u8rows.get(), rows.data(), rows.size(), hipMemcpyHostToDevice));
        const uint8_t* m = nodes + n * 2048;
        for (uint64_t i = 0; i < n; i++) {
            uint16_t mk;
            std::memcpy(&mk, m + 2 * i, 2);
            v->masks[i] = mk;
        }
        std::memcpy(v->weights.data(), m + 2 * n, n * 64);
    }
    v->prepare(nullptr);
    RBOW_HIP(hipDeviceSynchronize());
    return v.release();
}

}  // namespace rbow_amd
