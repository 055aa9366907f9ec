// capi.cpp -- extern "C" boundary (include/rapidbow.h).  Every entry point
// catches internal errors and turns them into an rbow_status plus a
// thread-local message; nothing throws across the ABI.
#include <memory>
#include <new>
#include <string>
#include <vector>

#include <chrono>
#include <random>
#include <thread>

#include "database.hpp"
#include "server.hpp"
#include "vocabulary.hpp"

using namespace rbow_amd;

struct rbow_vocabulary {
    std::unique_ptr<Vocabulary> v;
    BuildStats stats;
};

struct rbow_database {
    std::unique_ptr<Database> db;
};

struct rbow_server {
    std::unique_ptr<Server> srv;
};

struct rbow_query_pipeline {
    std::unique_ptr<QueryPipeline> p;
};

namespace {
thread_local std::string g_lastError;

template <typename F>
rbow_status guarded(F&& f) {
    try {
        f();
        return RBOW_OK;
    } catch (const Error& e) {
        g_lastError = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_lastError = "host out of memory";
        return RBOW_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return RBOW_ERR_INVALID_ARG;
    }
}

hipStream_t hs(rbow_stream s) { return reinterpret_cast<hipStream_t>(s); }

// RapidBoW.cpp:29-40: only {kUInt8, kL2, 128} is supported.
void checkAttributes(const rbow_descriptor_attributes& a) {
    RBOW_REQUIRE(a.data_type == RBOW_DATA_UINT8 && a.distance_type == RBOW_DIST_L2 && a.nb_dims == 128,
                 RBOW_ERR_UNSUPPORTED,
                 "Only {DataType::kUInt8, DistanceType::kL2, 128}, i.e. quantized SIFT, is supported");
}

void checkAligned(const void* p, const char* what) {
    RBOW_REQUIRE((reinterpret_cast<uintptr_t>(p) & 15u) == 0, RBOW_ERR_INVALID_ARG,
                 std::string(what) + " must be 16-byte aligned");
}

// Task table + tiles for the kernel-level k-means entries.
struct KmScratch {
    DevBuf<uint32_t> beg, cnt, active, passes, nbActive, order, tpBeg, tpCnt, partCnt, partChanged, waveChunkBeg;
    DevBuf<KmTile> tiles;
    DevBuf<KmChunk> chunks;
    DevBuf<KmSlot> slots;
    DevBuf<KmPartRange> red;
    DevBuf<int8_t> XA;
    DevBuf<int32_t> partSums;
    uint32_t nbRed = 0;
    std::vector<KmTile> ht;
    KmLevel k;
    uint64_t rows = 0;
    KmScratch(const uint32_t* taskCnt, uint32_t nbTasks, uint32_t b, uint32_t ct, hipStream_t s) {
        RBOW_REQUIRE(taskCnt != nullptr || nbTasks == 0, RBOW_ERR_INVALID_ARG, "null task_cnt");
        RBOW_REQUIRE(ct <= 1 && (b == 16 || (b == 32 && ct == 1)), RBOW_ERR_UNSUPPORTED,
                     "b must be 16 (uint8 or float centers) or 32 (float centers)");
        std::vector<uint32_t> hb(nbTasks);
        for (uint32_t g = 0; g < nbTasks; g++) {
            hb[g] = uint32_t(rows);
            for (uint32_t r = 0; r < taskCnt[g]; r += 64)
                ht.push_back(KmTile{g, uint32_t(rows) + r, std::min(64u, taskCnt[g] - r)});
            rows += taskCnt[g];
        }
        RBOW_REQUIRE(rows < (1ull << 32), RBOW_ERR_CAPACITY, "too many points");
        beg.resize(std::max(1u, nbTasks));
        cnt.resize(std::max(1u, nbTasks));
        active.resize(std::max(1u, nbTasks));
        passes.resize(std::max(1u, nbTasks));
        nbActive.resize(1);
        tiles.resize(std::max<size_t>(1, ht.size()));
        if (nbTasks) {
            RBOW_HIP(hipMemcpyAsync(beg.get(), hb.data(), 4ull * nbTasks, hipMemcpyHostToDevice, s));
            RBOW_HIP(hipMemcpyAsync(cnt.get(), taskCnt, 4ull * nbTasks, hipMemcpyHostToDevice, s));
        }
        if (!ht.empty())
            RBOW_HIP(hipMemcpyAsync(tiles.get(), ht.data(), sizeof(KmTile) * ht.size(), hipMemcpyHostToDevice, s));
        launchFillU32(active.get(), 1u, nbTasks, s);
        launchFillU32(passes.get(), 0u, nbTasks, s);
        k.b = b;
        k.nbTasks = nbTasks;
        k.nbTiles = uint32_t(ht.size());
        k.centroidType = ct;
        k.tiles = tiles.get();
        k.taskBeg = beg.get();
        k.taskCnt = cnt.get();
        k.active = active.get();
        k.passes = passes.get();
        k.nbActive = nbActive.get();
    }
    // operands and partials of the I8 MFMA assign for row-major points
    void prepareMfma(const uint8_t* points, const uint32_t* taskCnt, hipStream_t s) {
        std::vector<uint32_t> hOrder(rows);
        for (uint64_t i = 0; i < rows; i++) hOrder[i] = uint32_t(i);
        std::vector<uint32_t> taskTiles(k.nbTasks);
        for (uint32_t g = 0; g < k.nbTasks; g++) taskTiles[g] = divUp(taskCnt[g], 64);
        KmPartition part = partitionKmTiles(taskTiles, {});
        std::vector<KmChunk>& hc = part.chunks;
        std::vector<uint32_t>& cb = part.taskPartBeg;
        std::vector<uint32_t>& cc = part.taskPartCnt;
        std::vector<KmPartRange> hr;  // every task is reduced into the caller's totals
        for (uint32_t g = 0; g < k.nbTasks; g++)
            for (uint32_t c = 0; c < cc[g]; c += kReduceItemParts)
                hr.push_back(KmPartRange{g, cb[g] + c, cb[g] + std::min(cc[g], c + kReduceItemParts)});
        waveChunkBeg.resize(part.waveChunkBeg.size());
        RBOW_HIP(hipMemcpyAsync(waveChunkBeg.get(), part.waveChunkBeg.data(), 4 * part.waveChunkBeg.size(),
                                hipMemcpyHostToDevice, s));
        k.waveChunkBeg = waveChunkBeg.get();
        k.nbWaves = part.nbWaves;
        order.resize(std::max<uint64_t>(rows, 1));
        chunks.resize(std::max<size_t>(hc.size(), 1));
        red.resize(std::max<size_t>(hr.size(), 1));
        slots.resize(std::max<size_t>(part.blockSlots.size(), 1));
        tpBeg.resize(std::max(1u, k.nbTasks));
        tpCnt.resize(std::max(1u, k.nbTasks));
        XA.resize(std::max<size_t>(ht.size(), 1) * 8192);
        partSums.resize(std::max<size_t>(part.nbParts, 1) * 16 * kDims);
        partCnt.resize(std::max<size_t>(part.nbParts, 1) * 16);
        partChanged.resize(std::max<size_t>(part.nbParts, 1));
        if (rows) RBOW_HIP(hipMemcpyAsync(order.get(), hOrder.data(), 4 * rows, hipMemcpyHostToDevice, s));
        if (!hc.empty()) RBOW_HIP(hipMemcpyAsync(chunks.get(), hc.data(), sizeof(KmChunk) * hc.size(), hipMemcpyHostToDevice, s));
        if (!hr.empty())
            RBOW_HIP(hipMemcpyAsync(red.get(), hr.data(), sizeof(KmPartRange) * hr.size(), hipMemcpyHostToDevice, s));
        if (!part.blockSlots.empty())
            RBOW_HIP(hipMemcpyAsync(slots.get(), part.blockSlots.data(), sizeof(KmSlot) * part.blockSlots.size(),
                                    hipMemcpyHostToDevice, s));
        if (k.nbTasks) {
            RBOW_HIP(hipMemcpyAsync(tpBeg.get(), cb.data(), 4ull * k.nbTasks, hipMemcpyHostToDevice, s));
            RBOW_HIP(hipMemcpyAsync(tpCnt.get(), cc.data(), 4ull * k.nbTasks, hipMemcpyHostToDevice, s));
        }
        launchGatherTiles(points, uint32_t(rows), order.get(), tiles.get(), uint32_t(ht.size()), XA.get(), s);
        k.XA = XA.get();
        k.chunks = chunks.get();
        k.nbChunks = uint32_t(hc.size());
        k.blockSlots = slots.get();
        k.nbParts = part.nbParts;
        k.taskPartBeg = tpBeg.get();
        k.taskPartCnt = tpCnt.get();
        k.partSums = partSums.get();
        k.partCnt = partCnt.get();
        k.partChanged = partChanged.get();
        nbRed = uint32_t(hr.size());
    }
};
}  // namespace

extern "C" {

const char* rbow_last_error(void) { return g_lastError.c_str(); }

void rbow_build_options_default(rbow_build_options* o) {
    if (!o) return;
    o->max_iters = 1000;  // kmeans.h:118
    o->seed_mode = 0;
    o->seed = 5489u;      // std::mt19937_64 default seed (kmeans.cpp:145)
    o->centroid_type = 0;
    o->kernel = 0;
    o->persist_max_tiles = 16;
    o->timing = 0;
}

rbow_status rbow_build_vocabulary_ex(rbow_descriptor_attributes attr, uint32_t branch_factor, const void* dev_desc,
                                     uint32_t nb_desc, uint32_t nb_doc, uint32_t nb_levels,
                                     const rbow_build_options* opts, rbow_stream stream, rbow_vocabulary** out) {
    return guarded([&] {
        RBOW_REQUIRE(out != nullptr, RBOW_ERR_INVALID_ARG, "null output handle");
        *out = nullptr;
        rbow_build_options o;
        rbow_build_options_default(&o);
        if (opts) o = *opts;
        checkAttributes(attr);
        RBOW_REQUIRE(o.centroid_type <= 1, RBOW_ERR_INVALID_ARG, "centroid_type must be 0 or 1");
        RBOW_REQUIRE(o.timing <= 1, RBOW_ERR_INVALID_ARG, "timing must be 0 or 1");
        RBOW_REQUIRE(branch_factor == 16 || (branch_factor == 32 && o.centroid_type == 1), RBOW_ERR_UNSUPPORTED,
                     "branchFactor must be 16 (or 32 with float32 centroids)");
        RBOW_REQUIRE(dev_desc != nullptr || nb_desc == 0, RBOW_ERR_INVALID_ARG, "null descriptors");
        checkAligned(dev_desc, "descriptors");
        std::unique_ptr<rbow_vocabulary> h(new rbow_vocabulary);
        h->v.reset(buildVocabulary(static_cast<const uint8_t*>(dev_desc), nb_desc, nb_doc, branch_factor, nb_levels,
                                   o, hs(stream), &h->stats));
        *out = h.release();
    });
}

rbow_status rbow_tune_build_options(rbow_descriptor_attributes attr, uint32_t branch_factor, const void* dev_desc,
                                    uint32_t nb_desc, uint32_t nb_doc, uint32_t nb_levels, rbow_build_options* opts,
                                    float* out_ms, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(opts != nullptr, RBOW_ERR_INVALID_ARG, "null options");
        static const uint32_t kCandidates[5] = {0u, 4u, 8u, 12u, 16u};
        rbow_build_options o = *opts;
        o.timing = 0;
        if (out_ms)
            for (int i = 0; i < 5; i++) out_ms[i] = 0.0f;
        if (o.centroid_type != 0 || o.kernel != 0) return;  // the option acts on uint8 MFMA builds only
        std::vector<uint8_t> first;
        double bestMs = 0;
        uint32_t best = o.persist_max_tiles;
        for (int i = 0; i < 5; i++) {
            o.persist_max_tiles = kCandidates[i];
            double ms[2];
            for (int rep = 0; rep < 2; rep++) {
                rbow_vocabulary* v = nullptr;
                const rbow_status st = rbow_build_vocabulary_ex(attr, branch_factor, dev_desc, nb_desc, nb_doc,
                                                                nb_levels, &o, stream, &v);
                if (st != RBOW_OK) throw Error(st, rbow_last_error());
                std::unique_ptr<rbow_vocabulary> h(v);
                ms[rep] = h->stats.totalMs;
                if (rep == 0) {
                    std::vector<uint8_t> blob = serialize(*h->v);
                    if (first.empty()) first = std::move(blob);
                    else RBOW_REQUIRE(blob == first, RBOW_ERR_HIP, "tuning candidates built different trees");
                }
            }
            const double m = std::min(ms[0], ms[1]);
            if (out_ms) out_ms[i] = float(m);
            if (i == 0 || m < bestMs) {
                bestMs = m;
                best = kCandidates[i];
            }
        }
        opts->persist_max_tiles = best;
    });
}

rbow_status rbow_build_vocabulary(rbow_descriptor_attributes attr, uint32_t branch_factor, const void* dev_desc,
                                  uint32_t nb_desc, uint32_t nb_doc, uint32_t nb_levels, rbow_stream stream,
                                  rbow_vocabulary** out) {
    return rbow_build_vocabulary_ex(attr, branch_factor, dev_desc, nb_desc, nb_doc, nb_levels, nullptr, stream, out);
}

void rbow_voc_destroy(rbow_vocabulary* voc) { delete voc; }

rbow_status rbow_voc_get_attributes(const rbow_vocabulary* voc, rbow_descriptor_attributes* attr) {
    return guarded([&] {
        RBOW_REQUIRE(voc && attr, RBOW_ERR_INVALID_ARG, "null argument");
        *attr = rbow_descriptor_attributes{RBOW_DATA_UINT8, RBOW_DIST_L2, 128u};
    });
}
uint32_t rbow_voc_branch_factor(const rbow_vocabulary* voc) { return voc ? voc->v->b : 0u; }
uint32_t rbow_voc_nb_levels(const rbow_vocabulary* voc) { return voc ? voc->v->L : 0u; }
uint32_t rbow_voc_centroid_type(const rbow_vocabulary* voc) { return voc ? voc->v->centroidType : 0u; }

rbow_status rbow_voc_lookup_ex(const rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc, uint32_t* dev_leaf,
                               int32_t kernel, uint32_t* dev_near_ties, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(voc != nullptr, RBOW_ERR_INVALID_ARG, "null vocabulary");
        if (nb_desc == 0) return;
        RBOW_REQUIRE(dev_desc && dev_leaf, RBOW_ERR_INVALID_ARG, "null descriptors or output");
        checkAligned(dev_desc, "descriptors");
        RBOW_REQUIRE(kernel >= RBOW_LOOKUP_AUTO && kernel <= RBOW_LOOKUP_MFMA, RBOW_ERR_INVALID_ARG, "bad kernel id");
        const TreeView t = voc->v->view();
        const auto* d = static_cast<const uint8_t*>(dev_desc);
        launchLookup(t, d, nb_desc, dev_leaf, dev_near_ties, kernel, hs(stream));
    });
}

rbow_status rbow_voc_lookup_noalloc(const rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc,
                                    uint32_t* dev_leaf, int32_t kernel, uint32_t* dev_near_ties, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(voc != nullptr, RBOW_ERR_INVALID_ARG, "null vocabulary");
        if (nb_desc == 0) return;
        RBOW_REQUIRE(dev_desc && dev_leaf, RBOW_ERR_INVALID_ARG, "null descriptors or output");
        checkAligned(dev_desc, "descriptors");
        RBOW_REQUIRE(kernel >= RBOW_LOOKUP_AUTO && kernel <= RBOW_LOOKUP_MFMA, RBOW_ERR_INVALID_ARG, "bad kernel id");
        launchLookupNoAlloc(voc->v->view(), static_cast<const uint8_t*>(dev_desc), nb_desc, dev_leaf, dev_near_ties,
                            kernel, hs(stream));
    });
}

rbow_status rbow_release_build_workspaces(void) {
    return guarded([&] { releaseBuildWorkspaces(); });
}

rbow_status rbow_voc_tune_lookup(rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc,
                                 uint32_t* out_crossover, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(voc != nullptr, RBOW_ERR_INVALID_ARG, "null vocabulary");
        RBOW_REQUIRE(dev_desc != nullptr || nb_desc == 0, RBOW_ERR_INVALID_ARG, "null descriptors");
        if (nb_desc) checkAligned(dev_desc, "descriptors");
        DevBuf<uint32_t> leaf(std::max(1u, nb_desc));
        const LookupTuning r = tuneLookup(voc->v->view(), static_cast<const uint8_t*>(dev_desc), nb_desc, leaf.get(),
                                          hs(stream));
        voc->v->lookupAutoMin = r.crossover;
        if (r.fusedMax) voc->v->lookupFusedMax = r.fusedMax;
        if (r.leafItemRows) voc->v->lookupLeafItemRows = r.leafItemRows;
        if (out_crossover) *out_crossover = r.crossover;
    });
}

rbow_status rbow_voc_set_lookup_crossover(rbow_vocabulary* voc, uint32_t crossover) {
    return guarded([&] {
        RBOW_REQUIRE(voc != nullptr, RBOW_ERR_INVALID_ARG, "null vocabulary");
        voc->v->lookupAutoMin = crossover;
    });
}

rbow_status rbow_voc_set_lookup_fused_max(rbow_vocabulary* voc, uint32_t fused_max) {
    return guarded([&] {
        RBOW_REQUIRE(voc != nullptr, RBOW_ERR_INVALID_ARG, "null vocabulary");
        voc->v->lookupFusedMax = fused_max;
    });
}

uint32_t rbow_voc_lookup_fused_max(const rbow_vocabulary* voc) {
    if (!voc) return 0;
    const TreeView t = voc->v->view();
    return t.fusedMax ? t.fusedMax : lookupFusedDefault(t);
}

rbow_status rbow_voc_set_lookup_leaf_item_rows(rbow_vocabulary* voc, uint32_t rows) {
    return guarded([&] {
        RBOW_REQUIRE(voc != nullptr, RBOW_ERR_INVALID_ARG, "null vocabulary");
        RBOW_REQUIRE(rows == 0 || rows == 256 || rows == 512, RBOW_ERR_INVALID_ARG, "leaf item rows must be 256 or 512");
        voc->v->lookupLeafItemRows = rows;
    });
}

uint32_t rbow_voc_lookup_leaf_item_rows(const rbow_vocabulary* voc) {
    if (!voc) return 0;
    const TreeView t = voc->v->view();
    return t.leafItemRows ? t.leafItemRows : lookupLeafItemRowsDefault(t);
}

uint32_t rbow_voc_lookup_crossover(const rbow_vocabulary* voc) {
    if (!voc) return 0;
    const Vocabulary& v = *voc->v;
    return v.lookupAutoMin ? v.lookupAutoMin : lookupAutoDefault(v.centroidType, v.b);
}

rbow_status rbow_voc_lookup(const rbow_vocabulary* voc, const void* dev_desc, uint32_t nb_desc, uint32_t* dev_leaf,
                            rbow_stream stream) {
    return rbow_voc_lookup_ex(voc, dev_desc, nb_desc, dev_leaf, RBOW_LOOKUP_AUTO, nullptr, stream);
}

rbow_status rbow_voc_serialize(const rbow_vocabulary* voc, uint8_t* blob, size_t capacity, size_t* size) {
    return guarded([&] {
        RBOW_REQUIRE(voc && size, RBOW_ERR_INVALID_ARG, "null argument");
        const std::vector<uint8_t> b = serialize(*voc->v);
        *size = b.size();
        if (blob) {
            RBOW_REQUIRE(capacity >= b.size(), RBOW_ERR_CAPACITY, "blob buffer too small");
            std::memcpy(blob, b.data(), b.size());
        }
    });
}

rbow_status rbow_voc_deserialize(const uint8_t* blob, size_t size, rbow_vocabulary** out) {
    return guarded([&] {
        RBOW_REQUIRE(out != nullptr, RBOW_ERR_INVALID_ARG, "null output handle");
        *out = nullptr;
        std::unique_ptr<rbow_vocabulary> h(new rbow_vocabulary);
        h->v.reset(deserialize(blob, size));
        *out = h.release();
    });
}

const float* rbow_voc_level_weights(const rbow_vocabulary* voc, uint32_t level) {
    if (!voc || level >= voc->v->L) return nullptr;
    return voc->v->weights.data() + levelBeg(voc->v->lg2b, level) * voc->v->b;
}

rbow_status rbow_voc_build_stats(const rbow_vocabulary* voc, rbow_build_stats* out) {
    return guarded([&] {
        RBOW_REQUIRE(voc && out, RBOW_ERR_INVALID_ARG, "null argument");
        const BuildStats& st = voc->stats;
        out->total_ms = st.totalMs;
        out->assign_ms = st.assignMs;
        out->update_ms = st.updateMs;
        out->point_passes = st.pointPasses;
        out->assign_ops = st.pointPasses * 2ull * voc->v->b * kDims;
        out->reduce_ms = st.reduceMs;
        out->assign_launches = st.streamLaunches;
        out->reserved = 0;
        out->stream_ms = st.streamMs;
        out->stream_point_passes = st.streamPointPasses;
        out->persist_ms = st.persistMs;
    });
}

const uint32_t* rbow_voc_build_passes(const rbow_vocabulary* voc) {
    if (!voc || voc->stats.passes.empty()) return nullptr;
    return voc->stats.passes.data();
}

rbow_status rbow_voc_copy_centroids(const rbow_vocabulary* voc, uint32_t level, float* host_out,
                                    size_t capacity_floats) {
    return guarded([&] {
        RBOW_REQUIRE(voc && host_out, RBOW_ERR_INVALID_ARG, "null argument");
        const Vocabulary& v = *voc->v;
        RBOW_REQUIRE(level < v.L, RBOW_ERR_INVALID_ARG, "level out of range");
        const uint64_t cnt = levelSize(v.lg2b, level) * v.b * kDims;
        RBOW_REQUIRE(capacity_floats >= cnt, RBOW_ERR_CAPACITY, "output too small");
        const uint64_t off = levelBeg(v.lg2b, level) * v.b * kDims;
        if (v.centroidType == 1) {
            RBOW_HIP(hipMemcpy(host_out, v.f32rows.get() + off, cnt * 4, hipMemcpyDeviceToHost));
        } else {
            std::vector<uint8_t> tmp(cnt);
            RBOW_HIP(hipMemcpy(tmp.data(), v.u8rows.get() + off, cnt, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < cnt; i++) host_out[i] = float(tmp[i]);
        }
    });
}

rbow_status rbow_kmeans_assign(const uint8_t* dev_points, const uint32_t* task_cnt, uint32_t nb_tasks, uint32_t b,
                               uint32_t centroid_type, const void* dev_centers, uint8_t* dev_assign,
                               uint32_t* dev_counts, int64_t* dev_sums, uint32_t* dev_changed, int32_t kernel,
                               rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(kernel == 0 || kernel == 1, RBOW_ERR_INVALID_ARG, "kernel must be 0 (MFMA) or 1 (plain)");
        checkAligned(dev_points, "points");
        KmScratch sc(task_cnt, nb_tasks, b, centroid_type, hs(stream));
        if (sc.rows) RBOW_REQUIRE(dev_points && dev_assign, RBOW_ERR_INVALID_ARG, "null points or assignment");
        RBOW_REQUIRE(dev_centers && dev_counts && dev_sums && dev_changed, RBOW_ERR_INVALID_ARG, "null argument");
        sc.k.X = dev_points;
        if (centroid_type == 0) sc.k.centersU8 = static_cast<uint8_t*>(const_cast<void*>(dev_centers));
        else sc.k.centersF32 = static_cast<float*>(const_cast<void*>(dev_centers));
        sc.k.assign = dev_assign;
        sc.k.counts = dev_counts;
        sc.k.sums = reinterpret_cast<long long*>(dev_sums);
        sc.k.changed = dev_changed;
        const bool mfma = kernel == 0 && centroid_type == 0;
        if (mfma) sc.prepareMfma(dev_points, task_cnt, hs(stream));
        launchKmAssign(sc.k, mfma ? 0 : 1, hs(stream));
        if (mfma) launchKmReduce(sc.k, sc.red.get(), sc.nbRed, hs(stream));
        RBOW_HIP(hipStreamSynchronize(hs(stream)));  // scratch lifetime
    });
}

rbow_status rbow_kmeans_update(const uint32_t* task_cnt, uint32_t nb_tasks, uint32_t b, uint32_t centroid_type,
                               void* dev_centers, uint32_t* dev_counts, int64_t* dev_sums, uint32_t* dev_changed,
                               rbow_stream stream) {
    return guarded([&] {
        KmScratch sc(task_cnt, nb_tasks, b, centroid_type, hs(stream));
        RBOW_REQUIRE(dev_centers && dev_counts && dev_sums && dev_changed, RBOW_ERR_INVALID_ARG, "null argument");
        if (centroid_type == 0) sc.k.centersU8 = static_cast<uint8_t*>(dev_centers);
        else sc.k.centersF32 = static_cast<float*>(dev_centers);
        sc.k.counts = dev_counts;
        sc.k.sums = reinterpret_cast<long long*>(dev_sums);
        sc.k.changed = dev_changed;
        launchKmUpdate(sc.k, hs(stream));
        RBOW_HIP(hipStreamSynchronize(hs(stream)));
    });
}

rbow_status rbow_reformat_tiles(const void* dev_desc, const uint32_t* dev_idx, uint32_t nb_tiles, uint32_t tile_size,
                                uint32_t* dev_tiles, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(tile_size == 16 || tile_size == 32, RBOW_ERR_INVALID_ARG, "tile_size must be 16 or 32");
        if (!nb_tiles) return;
        RBOW_REQUIRE(dev_desc && dev_idx && dev_tiles, RBOW_ERR_INVALID_ARG, "null argument");
        checkAligned(dev_desc, "descriptors");
        launchReformatTiles(static_cast<const uint8_t*>(dev_desc), dev_idx, nb_tiles, tile_size, dev_tiles,
                            hs(stream));
    });
}

rbow_status rbow_gather_rows(const void* dev_desc, const uint32_t* dev_idx, uint64_t n, uint8_t* dev_out,
                             rbow_stream stream) {
    return guarded([&] {
        if (!n) return;
        RBOW_REQUIRE(dev_desc && dev_idx && dev_out, RBOW_ERR_INVALID_ARG, "null argument");
        checkAligned(dev_desc, "descriptors");
        checkAligned(dev_out, "output");
        launchGatherRows(static_cast<const uint8_t*>(dev_desc), dev_idx, n, dev_out, hs(stream));
    });
}

rbow_status rbow_idf_weights(const uint32_t* dev_occ, uint64_t n, uint32_t nb_doc, float* dev_weights,
                             rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE((dev_occ && dev_weights) || n == 0, RBOW_ERR_INVALID_ARG, "null argument");
        RBOW_REQUIRE(nb_doc > 0, RBOW_ERR_NB_DOC, "nbDoc must be positive");
        launchIdfWeights(dev_occ, n, nb_doc, dev_weights, hs(stream));
    });
}

// ------------------------------------------------------------ database --

rbow_status rbow_db_create(const rbow_vocabulary* voc, rbow_database** out) {
    return guarded([&] {
        RBOW_REQUIRE(voc && out, RBOW_ERR_INVALID_ARG, "null argument");
        *out = nullptr;
        std::unique_ptr<rbow_database> h(new rbow_database);
        h->db.reset(new Database(*voc->v));
        *out = h.release();
    });
}

void rbow_db_destroy(rbow_database* db) { delete db; }

rbow_status rbow_db_add_doc(rbow_database* db, uint32_t doc, const uint32_t* host_leaf, uint32_t n) {
    return guarded([&] {
        RBOW_REQUIRE(db != nullptr, RBOW_ERR_INVALID_ARG, "null database");
        db->db->addDoc(doc, host_leaf, n);
    });
}

rbow_status rbow_db_query(const rbow_database* db, const uint32_t* host_leaf, uint32_t n, uint32_t max_results,
                          uint32_t* out_docs, float* out_scores, uint32_t* out_count) {
    return guarded([&] {
        RBOW_REQUIRE(db && out_count && (out_docs || max_results == 0), RBOW_ERR_INVALID_ARG, "null argument");
        *out_count = db->db->query(host_leaf, n, max_results, out_docs, out_scores);
    });
}

rbow_status rbow_db_query_and_add_doc(rbow_database* db, uint32_t doc, const uint32_t* host_leaf, uint32_t n,
                                      uint32_t max_results, uint32_t* out_docs, uint32_t* out_count) {
    // DataBase.cpp:192-197: query, then addDoc (two separate lock acquisitions)
    const rbow_status s = rbow_db_query(db, host_leaf, n, max_results, out_docs, nullptr, out_count);
    if (s != RBOW_OK) return s;
    return rbow_db_add_doc(db, doc, host_leaf, n);
}

rbow_status rbow_db_add_docs_device(rbow_database* db, const uint32_t* host_doc_ids, uint32_t nb_docs,
                                    const uint32_t* dev_leaf, const uint64_t* host_offsets, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(db != nullptr, RBOW_ERR_INVALID_ARG, "null database");
        if (!nb_docs) return;
        RBOW_REQUIRE(dev_leaf || host_offsets[nb_docs] == host_offsets[0], RBOW_ERR_INVALID_ARG, "null leaf ids");
        db->db->addDocsDevice(host_doc_ids, nb_docs, dev_leaf, host_offsets, hs(stream));
    });
}

rbow_status rbow_db_query_batch(const rbow_database* db, const uint32_t* dev_leaf, const uint64_t* host_offsets,
                                uint32_t nb_queries, uint32_t max_results, uint32_t* dev_docs, float* dev_scores,
                                uint32_t* dev_counts, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(db != nullptr, RBOW_ERR_INVALID_ARG, "null database");
        if (!nb_queries) return;
        RBOW_REQUIRE(dev_docs && dev_counts && max_results > 0, RBOW_ERR_INVALID_ARG, "null output");
        db->db->queryBatch(dev_leaf, host_offsets, nb_queries, max_results, dev_docs, dev_scores, dev_counts,
                           hs(stream));
    });
}

rbow_status rbow_db_set_walk_steps(rbow_database* db, uint32_t steps) {
    return guarded([&] {
        RBOW_REQUIRE(db != nullptr, RBOW_ERR_INVALID_ARG, "null database");
        db->db->setWalkSteps(steps);
    });
}

uint32_t rbow_db_walk_steps(const rbow_database* db) { return db ? db->db->walkSteps() : 0u; }

rbow_status rbow_db_tune_scoring(rbow_database* db, const uint32_t* dev_leaf, const uint64_t* host_offsets,
                                 uint32_t nb_queries, uint32_t max_results, uint32_t* out_steps, float* out_ms,
                                 rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(db != nullptr, RBOW_ERR_INVALID_ARG, "null database");
        RBOW_REQUIRE(nb_queries > 0 && max_results > 0 && host_offsets && dev_leaf, RBOW_ERR_INVALID_ARG,
                     "the tuner needs a query batch");
        static const uint32_t kCandidates[2] = {4u, 6u};
        const uint64_t slots = uint64_t(nb_queries) * max_results;
        DevBuf<uint32_t> docs(slots), counts(nb_queries), docs0(slots), counts0(nb_queries);
        DevBuf<float> scores(slots), scores0(slots);
        hipStream_t s = hs(stream);
        hipEvent_t e0, e1;
        RBOW_HIP(hipEventCreate(&e0));
        RBOW_HIP(hipEventCreate(&e1));
        const uint32_t before = db->db->walkSteps();
        float best = 0.0f;
        uint32_t bestSteps = before;
        try {
            for (int c = 0; c < 2; c++) {
                db->db->setWalkSteps(kCandidates[c]);
                // one warm run (its results kept for the comparison), then the median of three
                db->db->queryBatch(dev_leaf, host_offsets, nb_queries, max_results, c ? docs.get() : docs0.get(),
                                   c ? scores.get() : scores0.get(), c ? counts.get() : counts0.get(), s);
                float ms[3];
                for (int r = 0; r < 3; r++) {
                    RBOW_HIP(hipEventRecord(e0, s));
                    db->db->queryBatch(dev_leaf, host_offsets, nb_queries, max_results, docs.get(), scores.get(),
                                       counts.get(), s);
                    RBOW_HIP(hipEventRecord(e1, s));
                    RBOW_HIP(hipEventSynchronize(e1));
                    RBOW_HIP(hipEventElapsedTime(&ms[r], e0, e1));
                }
                std::sort(ms, ms + 3);
                if (out_ms) out_ms[c] = ms[1];
                if (c == 0 || ms[1] < best) {
                    best = ms[1];
                    bestSteps = kCandidates[c];
                }
            }
            // the tile shape changes no result: both candidates' rankings and scores are compared
            std::vector<uint32_t> a(slots), b(slots), ca(nb_queries), cb(nb_queries);
            std::vector<float> sa(slots), sb(slots);
            RBOW_HIP(hipMemcpy(a.data(), docs0.get(), 4 * slots, hipMemcpyDeviceToHost));
            RBOW_HIP(hipMemcpy(b.data(), docs.get(), 4 * slots, hipMemcpyDeviceToHost));
            RBOW_HIP(hipMemcpy(sa.data(), scores0.get(), 4 * slots, hipMemcpyDeviceToHost));
            RBOW_HIP(hipMemcpy(sb.data(), scores.get(), 4 * slots, hipMemcpyDeviceToHost));
            RBOW_HIP(hipMemcpy(ca.data(), counts0.get(), 4ull * nb_queries, hipMemcpyDeviceToHost));
            RBOW_HIP(hipMemcpy(cb.data(), counts.get(), 4ull * nb_queries, hipMemcpyDeviceToHost));
            bool same = ca == cb;
            for (uint32_t q = 0; same && q < nb_queries; q++)
                for (uint32_t i = 0; i < ca[q]; i++) {
                    const uint64_t x = uint64_t(q) * max_results + i;
                    same = same && a[x] == b[x] && std::fabs(sa[x] - sb[x]) <= 2e-6f;
                }
            RBOW_REQUIRE(same, RBOW_ERR_HIP, "scoring tile shapes gave different results");
        } catch (...) {
            db->db->setWalkSteps(before == Database::kWalkStepsDefault ? 0u : before);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        RBOW_HIP(hipEventDestroy(e0));
        RBOW_HIP(hipEventDestroy(e1));
        db->db->setWalkSteps(bestSteps);
        if (out_steps) *out_steps = bestSteps;
    });
}

uint32_t rbow_db_nb_docs(const rbow_database* db) { return db ? db->db->nbDocs() : 0u; }
uint64_t rbow_db_nb_postings(const rbow_database* db) { return db ? db->db->nbPostings() : 0u; }
uint64_t rbow_db_nb_delta_postings(const rbow_database* db) { return db ? db->db->nbDeltaPostings() : 0u; }

rbow_status rbow_db_compact(rbow_database* db, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(db != nullptr, RBOW_ERR_INVALID_ARG, "null database");
        db->db->compact(hs(stream));
    });
}

rbow_status rbow_db_postings(const rbow_database* db, uint32_t leaf, uint32_t* host_out, uint32_t capacity,
                             uint32_t* length) {
    return guarded([&] {
        RBOW_REQUIRE(db && length, RBOW_ERR_INVALID_ARG, "null argument");
        *length = db->db->postings(leaf, host_out, capacity);
    });
}

rbow_status rbow_db_doc_rcp(const rbow_database* db, uint32_t doc, float* rcp) {
    return guarded([&] {
        RBOW_REQUIRE(db && rcp, RBOW_ERR_INVALID_ARG, "null argument");
        RBOW_REQUIRE(db->db->docRcp(doc, rcp), RBOW_ERR_INVALID_ARG, "unknown document");
    });
}

// ------------------------------------------------------------- server --

void rbow_server_options_default(rbow_server_options* o) {
    if (!o) return;
    o->max_batch = 64;
    o->nb_streams = 2;
    o->mem_budget = 4ull << 30;
    o->max_results = 50;
    o->lookup_kernel = RBOW_LOOKUP_AUTO;
    o->staging_bytes = 256ull << 20;
    o->use_graphs = 1;
    o->reserved = 0;
    o->batch_max_desc = 0;
}

rbow_status rbow_server_create(const rbow_vocabulary* voc, const rbow_database* db, const rbow_server_options* opts,
                               rbow_server** out) {
    return guarded([&] {
        RBOW_REQUIRE(voc && db && out, RBOW_ERR_INVALID_ARG, "null argument");
        RBOW_REQUIRE(&db->db->vocabulary() == voc->v.get(), RBOW_ERR_INVALID_ARG,
                     "the database was not created on this vocabulary");
        *out = nullptr;
        rbow_server_options o;
        rbow_server_options_default(&o);
        if (opts) o = *opts;
        ServerOptions so;
        so.maxBatch = o.max_batch;
        so.nbStreams = o.nb_streams;
        so.memBudget = o.mem_budget;
        so.K = o.max_results;
        so.lookupKernel = o.lookup_kernel;
        so.stagingBytes = o.staging_bytes;
        so.useGraphs = o.use_graphs != 0;
        so.batchMaxDesc = o.batch_max_desc;
        std::unique_ptr<rbow_server> h(new rbow_server);
        h->srv.reset(new Server(*voc->v, *db->db, so));
        *out = h.release();
    });
}

void rbow_server_destroy(rbow_server* srv) { delete srv; }

// ------------------------------------------------------------ pipeline --

rbow_status rbow_pipeline_create(const rbow_vocabulary* voc, const rbow_database* db, uint32_t max_queries,
                                 uint64_t max_desc, uint32_t max_results, rbow_query_pipeline** out) {
    return guarded([&] {
        RBOW_REQUIRE(voc && db && out, RBOW_ERR_INVALID_ARG, "null argument");
        RBOW_REQUIRE(&db->db->vocabulary() == voc->v.get(), RBOW_ERR_INVALID_ARG,
                     "the database was not created on this vocabulary");
        *out = nullptr;
        db->db->prepareForQueries(nullptr);
        std::unique_ptr<rbow_query_pipeline> h(new rbow_query_pipeline);
        h->p.reset(new QueryPipeline(*voc->v, *db->db, max_queries, max_desc, max_results));
        *out = h.release();
    });
}

void rbow_pipeline_destroy(rbow_query_pipeline* p) { delete p; }

rbow_status rbow_pipeline_run(rbow_query_pipeline* p, const void* dev_desc, const uint64_t* host_offsets,
                              uint32_t nb_queries, int32_t lookup_kernel, int32_t use_graph, uint32_t* dev_docs,
                              float* dev_scores, uint32_t* dev_counts, rbow_stream stream) {
    return guarded([&] {
        RBOW_REQUIRE(p && host_offsets && dev_docs && dev_counts, RBOW_ERR_INVALID_ARG, "null argument");
        RBOW_REQUIRE(lookup_kernel >= RBOW_LOOKUP_AUTO && lookup_kernel <= RBOW_LOOKUP_MFMA, RBOW_ERR_INVALID_ARG,
                     "bad kernel id");
        if (!nb_queries) return;
        QueryPipeline& P = *p->p;
        RBOW_REQUIRE(nb_queries <= P.maxQueries(), RBOW_ERR_CAPACITY, "batch above the pipeline's query capacity");
        std::vector<uint64_t> off(nb_queries + 1);
        for (uint32_t q = 0; q <= nb_queries; q++) {
            RBOW_REQUIRE(host_offsets[q] >= host_offsets[0], RBOW_ERR_INVALID_ARG, "offsets must be non-decreasing");
            off[q] = host_offsets[q] - host_offsets[0];
        }
        const uint64_t n = off[nb_queries];
        RBOW_REQUIRE(n <= P.maxDesc(), RBOW_ERR_CAPACITY, "batch above the pipeline's descriptor capacity");
        RBOW_REQUIRE(dev_desc || n == 0, RBOW_ERR_INVALID_ARG, "null descriptors");
        hipStream_t s = hs(stream);
        if (n) {
            checkAligned(dev_desc, "descriptors");
            RBOW_HIP(hipMemcpyAsync(P.descBuffer(), static_cast<const uint8_t*>(dev_desc) + host_offsets[0] * kDims,
                                    n * kDims, hipMemcpyDeviceToDevice, s));
        }
        P.run(off.data(), nb_queries, lookup_kernel, use_graph != 0, false, s);
        const uint8_t* res = P.deviceResults();
        const uint64_t K = P.K();
        RBOW_HIP(hipMemcpyAsync(dev_docs, P.docs(res, 0), 4 * K * nb_queries, hipMemcpyDeviceToDevice, s));
        if (dev_scores)
            RBOW_HIP(hipMemcpyAsync(dev_scores, P.scores(res, 0), 4 * K * nb_queries, hipMemcpyDeviceToDevice, s));
        RBOW_HIP(hipMemcpyAsync(dev_counts, P.counts(res), 4ull * nb_queries, hipMemcpyDeviceToDevice, s));
        RBOW_HIP(hipStreamSynchronize(s));  // the pipeline's buffers are free again on return
    });
}

uint32_t rbow_pipeline_graphs(const rbow_query_pipeline* p) { return p ? p->p->graphsCaptured() : 0u; }

rbow_status rbow_server_submit(rbow_server* srv, const uint8_t* host_desc, uint32_t n, uint64_t* ticket) {
    return guarded([&] {
        RBOW_REQUIRE(srv && ticket, RBOW_ERR_INVALID_ARG, "null argument");
        *ticket = srv->srv->submit(host_desc, n);
    });
}

rbow_status rbow_server_wait(rbow_server* srv, uint64_t ticket, uint32_t* docs, float* scores, uint32_t* count,
                             double* latency_ms) {
    return guarded([&] {
        RBOW_REQUIRE(srv && count, RBOW_ERR_INVALID_ARG, "null argument");
        const Server::Result r = srv->srv->wait(ticket);
        *count = uint32_t(r.docs.size());
        if (docs) std::copy(r.docs.begin(), r.docs.end(), docs);
        if (scores) std::copy(r.scores.begin(), r.scores.end(), scores);
        if (latency_ms) *latency_ms = r.doneMs - r.arriveMs;
    });
}

rbow_status rbow_server_replay(rbow_server* srv, const uint8_t* host_desc, const uint64_t* offsets,
                               uint32_t nb_queries, double qps, uint64_t seed, double* latency_ms, uint32_t* docs,
                               float* scores, uint32_t* counts, uint32_t* batch_sizes, double* elapsed_s) {
    return guarded([&] {
        RBOW_REQUIRE(srv && offsets && latency_ms && qps > 0, RBOW_ERR_INVALID_ARG, "bad argument");
        Server& S = *srv->srv;
        const uint32_t K = S.options().K;
        // Poisson arrivals: exponential gaps at rate qps
        std::mt19937_64 rng(seed);
        std::exponential_distribution<double> gap(qps);
        std::vector<double> at(nb_queries);
        double t = 0;
        for (uint32_t i = 0; i < nb_queries; i++) at[i] = (t += gap(rng));
        std::vector<uint64_t> tickets(nb_queries);
        std::exception_ptr subErr;
        const auto start = std::chrono::steady_clock::now();
        std::thread submitter([&] {
            try {
                for (uint32_t i = 0; i < nb_queries; i++) {
                    std::this_thread::sleep_until(start + std::chrono::duration<double>(at[i]));
                    tickets[i] = S.submit(host_desc + offsets[i] * kDims, uint32_t(offsets[i + 1] - offsets[i]));
                }
            } catch (...) {
                subErr = std::current_exception();
            }
        });
        submitter.join();  // submissions are quick; waiting afterwards keeps the open-loop schedule
        if (subErr) std::rethrow_exception(subErr);
        double first = 1e300, last = 0;
        for (uint32_t i = 0; i < nb_queries; i++) {
            const Server::Result r = S.wait(tickets[i]);
            latency_ms[i] = r.doneMs - r.arriveMs;
            first = std::min(first, r.arriveMs);
            last = std::max(last, r.doneMs);
            if (counts) counts[i] = uint32_t(r.docs.size());
            if (docs) std::copy(r.docs.begin(), r.docs.end(), docs + uint64_t(i) * K);
            if (scores) std::copy(r.scores.begin(), r.scores.end(), scores + uint64_t(i) * K);
            if (batch_sizes) batch_sizes[i] = r.batchSize;
        }
        if (elapsed_s) *elapsed_s = nb_queries ? (last - first) * 1e-3 : 0.0;
    });
}

}  // extern "C"
