// database.hip -- inverted-file kernels: document norms, CSR merge, batched
// TF-IDF / L1 scoring and top-K (DataBase.cpp:39-190).
//
// Scoring semantics per query (DataBase.cpp:71-190):
//   qrcp = float(1 / sum_t sum_desc W[L-1-t][leaf >> t*lg2b])          (binary64 sum)
//   for every distinct node v at depth-from-leaf t of the query's leaves:
//     n_q = #query leaves under v, w = W[L-1-t][v], qE = float(n_q) * w * qrcp
//     P = postings under v; scored iff t == 0 or P < 16 + 4K
//     for every distinct doc d with m postings under v:
//       tE = float(m) * w * rcp[d];  acc[d] += double(|qE - tE| - (qE + tE))
//   score[d] = float(acc[d] + 2.0); rank by (score, doc id) ascending, keep K.
// Every binary32 operation flushes subnormal inputs and results to zero, as the
// reference's -ffast-math host build does (CMakeLists.txt:16).
#include <hipcub/hipcub.hpp>

#include "database.hpp"

namespace rbow_amd {
namespace {

constexpr uint32_t kSortCap = Database::kMaxQueryDesc;  // leaves per query sorted in LDS
constexpr uint32_t kCandCap = 2048;                     // boundary candidates kept by the top-K select

__device__ inline float ftz(float x) { return fabsf(x) < 1.17549435e-38f ? copysignf(0.0f, x) : x; }
__device__ inline float fmulz(float a, float b) { return ftz(ftz(a) * ftz(b)); }
__device__ inline float faddz(float a, float b) { return ftz(ftz(a) + ftz(b)); }
__device__ inline float fsubz(float a, float b) { return ftz(ftz(a) - ftz(b)); }

struct QNode {
    uint32_t v;      // node index within its level
    uint32_t t;      // depth from the leaf level
    uint32_t cnt;    // query leaves under the node
    float qE;
    float w;
    uint32_t pad;
    uint64_t a, e;   // posting slice
};

__device__ inline double blockSumF64(double v, double* red) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) s += red[w];
    return s;
}

// ------------------------------------------------------------- norms ---
// DataBase.cpp:54-68: one wave per document.
__global__ __launch_bounds__(256) void k_docNorms(DbView v, const uint32_t* __restrict__ leaf,
                                                 const uint64_t* __restrict__ off, uint32_t nbDocs,
                                                 uint32_t firstSlot, float* __restrict__ rcp) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t d = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (d >= nbDocs) return;
    double s = 0;
    for (uint32_t t = 0; t < v.L; t++) {
        const float* W = v.W + levelBeg(v.lg2b, v.L - 1 - t) * v.b;
        for (uint64_t i = off[d] + lane; i < off[d + 1]; i += 64) s += double(W[leaf[i] >> (t * v.lg2b)]);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) rcp[firstSlot + d] = ftz(float(1.0 / s));
}

// ------------------------------------------------------------- merge ---
__global__ void k_hist(const uint32_t* __restrict__ leaf, uint64_t n, uint64_t* __restrict__ cnt) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        atomicAdd(reinterpret_cast<unsigned long long*>(&cnt[leaf[i]]), 1ull);
}

// new CSR = for each leaf: old postings, then the batch's postings in slot order
__global__ void k_placeOld(const uint64_t* __restrict__ leafOff, uint64_t nbLeaves, const uint32_t* __restrict__ post,
                           uint64_t nbOld, const uint64_t* __restrict__ batchOff, uint32_t* __restrict__ newPost) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nbOld;
         i += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t lo = 0, hi = nbLeaves;  // leaf v with leafOff[v] <= i < leafOff[v+1]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (leafOff[mid] <= i) lo = mid;
            else hi = mid;
        }
        newPost[i + batchOff[lo]] = post[i];
    }
}
__global__ void k_placeNew(const uint64_t* __restrict__ leafOff, const uint32_t* __restrict__ sortedLeaf,
                           const uint32_t* __restrict__ sortedSlot, uint64_t n, uint32_t* __restrict__ newPost) {
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += uint64_t(gridDim.x) * blockDim.x)
        newPost[leafOff[sortedLeaf[k] + 1] + k] = sortedSlot[k];
}
__global__ void k_newOff(const uint64_t* __restrict__ leafOff, const uint64_t* __restrict__ batchOff, uint64_t n1,
                         uint64_t* __restrict__ out) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n1; i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = leafOff[i] + batchOff[i];
}
__global__ void k_fill64(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
__global__ void k_fill8(uint8_t* p, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = 0;
}

// ------------------------------------------------------------- query ---
// One 1024-thread block per query: sort the leaves, norm, scored node list.
__global__ __launch_bounds__(1024) void k_queryPrep(DbView v, const uint32_t* __restrict__ leaf,
                                                   const uint64_t* __restrict__ off, uint32_t K,
                                                   QNode* __restrict__ nodes, uint32_t* __restrict__ nodeCount) {
    __shared__ uint32_t s[kSortCap];
    __shared__ double red[16];
    __shared__ uint32_t nNodes;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint64_t beg = off[q];
    const uint32_t n = uint32_t(off[q + 1] - beg);
    uint32_t n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (uint32_t i = tid; i < n2; i += 1024) s[i] = i < n ? leaf[beg + i] : 0xFFFFFFFFu;
    if (tid == 0) nNodes = 0;
    __syncthreads();
    // bitonic sort, ascending
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < n2; i += 1024) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint32_t a = s[i], c = s[p];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        s[i] = c;
                        s[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    double part = 0;
    for (uint32_t t = 0; t < v.L; t++) {
        const float* W = v.W + levelBeg(v.lg2b, v.L - 1 - t) * v.b;
        for (uint32_t i = tid; i < n; i += 1024) part += double(W[s[i] >> (t * v.lg2b)]);
    }
    const double l1 = blockSumF64(part, red);
    const float qrcp = ftz(float(1.0 / l1));  // DataBase.cpp:77-94
    const uint64_t thr = 16ull + 4ull * K;    // DataBase.h:60-61
    QNode* out = nodes + uint64_t(v.L) * beg;
    for (uint32_t t = 0; t < v.L; t++) {
        const uint32_t shift = t * v.lg2b;
        const float* W = v.W + levelBeg(v.lg2b, v.L - 1 - t) * v.b;
        for (uint32_t i = tid; i < n; i += 1024) {
            const uint32_t x = s[i] >> shift;
            if (i != 0 && (s[i - 1] >> shift) == x) continue;
            // run end: first index whose node differs (binary search in the sorted leaves)
            const uint64_t nextLeaf = (uint64_t(x) + 1) << shift;
            uint32_t lo = i, hi = n;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if (uint64_t(s[mid]) < nextLeaf) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t cnt = lo - i;
            const uint64_t a = v.leafOff[uint64_t(x) << shift], e = v.leafOff[nextLeaf];
            if (t == 0 || e - a < thr) {
                const float w = W[x];
                QNode nd;
                nd.v = x;
                nd.t = t;
                nd.cnt = cnt;
                nd.qE = fmulz(fmulz(float(cnt), w), qrcp);  // DataBase.cpp:117
                nd.w = w;
                nd.pad = 0;
                nd.a = a;
                nd.e = e;
                out[atomicAdd(&nNodes, 1u)] = nd;
            }
        }
    }
    __syncthreads();
    if (tid == 0) nodeCount[q] = nNodes;
}

__device__ inline void addTerm(const DbView& v, double* acc, uint8_t* touched, uint32_t slot, uint32_t m, float w,
                               float qE) {
    const float tE = fmulz(fmulz(float(m), w), v.rcp[slot]);              // DataBase.cpp:128-129
    const float c = fsubz(fabsf(fsubz(qE, tE)), faddz(qE, tE));            // DataBase.cpp:130
    atomicAdd(&acc[slot], double(c));
    touched[slot] = 1;
}

// Wave per scored node; grid (x, query).
__global__ __launch_bounds__(256) void k_score(DbView v, const uint64_t* __restrict__ off,
                                              const QNode* __restrict__ nodes, const uint32_t* __restrict__ nodeCount,
                                              double* __restrict__ accAll, uint8_t* __restrict__ touchedAll) {
    const uint32_t q = blockIdx.y, lane = threadIdx.x & 63;
    const QNode* nd = nodes + uint64_t(v.L) * off[q];
    const uint32_t nn = nodeCount[q];
    double* acc = accAll + uint64_t(q) * v.nbDocs;
    uint8_t* touched = touchedAll + uint64_t(q) * v.nbDocs;
    for (uint32_t k = blockIdx.x * 4 + (threadIdx.x >> 6); k < nn; k += gridDim.x * 4) {
        const QNode node = nd[k];
        if (node.t == 0) {
            // one leaf: duplicates of a doc are adjacent (DataBase.cpp:138-149)
            for (uint64_t i = node.a + lane; i < node.e; i += 64) {
                const uint32_t slot = v.post[i];
                if (i != node.a && v.post[i - 1] == slot) continue;
                uint32_t m = 1;
                while (i + m < node.e && v.post[i + m] == slot) m++;
                addTerm(v, acc, touched, slot, m, node.w, node.qE);
            }
        } else {
            // virtual inverted file of < 16+4K postings (DataBase.cpp:150-164)
            for (uint64_t i = node.a + lane; i < node.e; i += 64) {
                const uint32_t slot = v.post[i];
                bool first = true;
                uint32_t m = 0;
                for (uint64_t j = node.a; j < node.e; j++) {
                    if (v.post[j] == slot) {
                        m++;
                        if (j < i) first = false;
                    }
                }
                if (first) addTerm(v, acc, touched, slot, m, node.w, node.qE);
            }
        }
    }
}

__device__ inline uint64_t rankKey(double acc, uint32_t docId) {
    const uint32_t u = __float_as_uint(float(acc + 2.0));  // DataBase.cpp:173
    const uint32_t o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return (uint64_t(o) << 32) | docId;
}
__device__ inline float keyScore(uint64_t k) {
    const uint32_t o = uint32_t(k >> 32);
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// One block per query: exact K smallest (score, doc id) keys by radix select
// over 12-bit digits, then a bitonic sort of the survivors.
__global__ __launch_bounds__(1024) void k_topk(DbView v, uint32_t K, const double* __restrict__ accAll,
                                              const uint8_t* __restrict__ touchedAll, uint32_t* __restrict__ outDocs,
                                              float* __restrict__ outScores, uint32_t* __restrict__ outCount) {
    __shared__ uint32_t hist[4096];
    __shared__ uint64_t cand[kCandCap];
    __shared__ uint64_t sel[Database::kMaxResults];
    __shared__ uint32_t nCand, nSel, sDigit, sBefore, sTotal;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const double* acc = accAll + uint64_t(q) * v.nbDocs;
    const uint8_t* touched = touchedAll + uint64_t(q) * v.nbDocs;
    uint64_t prefix = 0, mask = 0;
    uint32_t need = K, shiftHi = 64;
    bool all = false;
    while (shiftHi > 0) {
        const uint32_t width = shiftHi >= 12 ? 12 : shiftHi;
        const uint32_t shift = shiftHi - width;
        for (uint32_t i = tid; i < 4096; i += 1024) hist[i] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < v.nbDocs; i += 1024) {
            if (!touched[i]) continue;
            const uint64_t key = rankKey(acc[i], v.docId[i]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & ((1u << width) - 1)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t before = 0, d = 0, total = 0;
            for (uint32_t i = 0; i < (1u << width); i++) total += hist[i];
            if (total <= need) {
                d = 0xFFFFFFFFu;
            } else {
                for (d = 0; before + hist[d] < need; d++) before += hist[d];
            }
            sDigit = d;
            sBefore = before;
            sTotal = d == 0xFFFFFFFFu ? total : hist[d];
        }
        __syncthreads();
        if (sDigit == 0xFFFFFFFFu) {  // every matching key is selected
            all = true;
            break;
        }
        need -= sBefore;
        prefix |= uint64_t(sDigit) << shift;
        mask |= uint64_t((1u << width) - 1) << shift;
        shiftHi = shift;
        if (sTotal <= kCandCap) break;
        __syncthreads();
    }
    // collect: keys below the boundary prefix are selected; boundary keys are candidates
    if (tid == 0) {
        nCand = 0;
        nSel = 0;
    }
    __syncthreads();
    for (uint32_t i = tid; i < v.nbDocs; i += 1024) {
        if (!touched[i]) continue;
        const uint64_t key = rankKey(acc[i], v.docId[i]);
        const uint64_t hp = key & mask;
        if (hp == prefix) {
            const uint32_t c = atomicAdd(&nCand, 1u);
            if (c < kCandCap) cand[c] = key;
        } else if (hp < prefix) {
            const uint32_t c = atomicAdd(&nSel, 1u);
            if (c < Database::kMaxResults) sel[c] = key;
        }
    }
    __syncthreads();
    const uint32_t nc = min(nCand, kCandCap);
    uint32_t n2 = 1;
    while (n2 < nc) n2 <<= 1;
    for (uint32_t i = nc + tid; i < n2; i += 1024) cand[i] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= n2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < n2; i += 1024) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t a = cand[i], c = cand[p];
                    if ((a > c) == ((i & k) == 0)) {
                        cand[i] = c;
                        cand[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    // selected (< K keys, unsorted) + the `need` smallest candidates
    const uint32_t ns = min(nSel, K);
    const uint32_t take = all ? min(nc, K) : min(need, nc);
    for (uint32_t i = tid; i < take; i += 1024) sel[ns + i] = cand[i];
    __syncthreads();
    const uint32_t tot = ns + take;
    n2 = 1;
    while (n2 < tot) n2 <<= 1;
    for (uint32_t i = tot + tid; i < n2; i += 1024) sel[i] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= n2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < n2; i += 1024) {
                const uint32_t p = i ^ j;
                if (p > i) {
                    const uint64_t a = sel[i], c = sel[p];
                    if ((a > c) == ((i & k) == 0)) {
                        sel[i] = c;
                        sel[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    const uint32_t cnt = min(tot, K);
    for (uint32_t i = tid; i < cnt; i += 1024) {
        outDocs[uint64_t(q) * K + i] = uint32_t(sel[i]);
        if (outScores) outScores[uint64_t(q) * K + i] = keyScore(sel[i]);
    }
    if (tid == 0) outCount[q] = cnt;
}

__global__ void k_expandSlots(const uint64_t* __restrict__ off, uint32_t nbDocs, uint32_t firstSlot,
                              uint32_t* __restrict__ slots) {
    const uint32_t d = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (d >= nbDocs) return;
    for (uint64_t i = off[d] - off[0] + (threadIdx.x & 63); i < off[d + 1] - off[0]; i += 64) slots[i] = firstSlot + d;
}

uint32_t gridFor(uint64_t work, uint32_t perBlock, uint32_t cap) {
    const uint64_t g = (work + perBlock - 1) / perBlock;
    return uint32_t(g == 0 ? 1 : (g > cap ? cap : g));
}

}  // namespace

void launchExpandSlots(const uint64_t* off, uint32_t nbDocs, uint32_t firstSlot, uint32_t* slots, hipStream_t s) {
    if (!nbDocs) return;
    k_expandSlots<<<divUp(nbDocs, 4), 256, 0, s>>>(off, nbDocs, firstSlot, slots);
    RBOW_HIP(hipGetLastError());
}

void launchDocNorms(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, uint32_t nbDocs,
                    uint32_t firstSlot, float* rcp, hipStream_t s) {
    if (!nbDocs) return;
    k_docNorms<<<divUp(nbDocs, 4), 256, 0, s>>>(v, dLeaf, dOffsets, nbDocs, firstSlot, rcp);
    RBOW_HIP(hipGetLastError());
}

void launchMergePostings(const DbView& v, uint64_t nbOld, const uint32_t* dLeaf, const uint32_t* dSlot, uint64_t n,
                         uint64_t* newLeafOff, uint32_t* newPost, hipStream_t s) {
    const uint64_t n1 = v.nbLeaves + 1;
    RBOW_REQUIRE(n < (1ull << 31), RBOW_ERR_CAPACITY, "merge batch too large");
    uint32_t bits = 1;
    while ((1ull << bits) < v.nbLeaves) bits++;
    size_t tempBytes = 0;
    RBOW_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, dLeaf, (uint32_t*)nullptr, dSlot,
                                                (uint32_t*)nullptr, int(n), 0, int(bits), s));
    size_t scanBytes = 0;
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, (uint64_t*)nullptr, (uint64_t*)nullptr, int(n1), s));
    const size_t bytes = 8 * n1 * 2 + 8 * n + std::max(tempBytes, scanBytes) + 1024;
    void* ws = nullptr;
    RBOW_HIP(hipMallocAsync(&ws, bytes, s));
    uint64_t* cnt = static_cast<uint64_t*>(ws);
    uint64_t* batchOff = cnt + n1;
    uint32_t* sLeaf = reinterpret_cast<uint32_t*>(batchOff + n1);
    uint32_t* sSlot = sLeaf + n;
    void* temp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(sSlot + n) + 255) & ~uintptr_t(255));
    k_fill64<<<gridFor(n1, 256, 4096), 256, 0, s>>>(cnt, 0, n1);
    k_hist<<<gridFor(n, 256, 4096), 256, 0, s>>>(dLeaf, n, cnt);
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(temp, scanBytes, cnt, batchOff, int(n1), s));
    RBOW_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, dLeaf, sLeaf, dSlot, sSlot, int(n), 0, int(bits), s));
    k_placeOld<<<gridFor(nbOld, 256, 8192), 256, 0, s>>>(v.leafOff, v.nbLeaves, v.post, nbOld, batchOff, newPost);
    k_placeNew<<<gridFor(n, 256, 8192), 256, 0, s>>>(v.leafOff, sLeaf, sSlot, n, newPost);
    k_newOff<<<gridFor(n1, 256, 4096), 256, 0, s>>>(v.leafOff, batchOff, n1, newLeafOff);
    RBOW_HIP(hipGetLastError());
    RBOW_HIP(hipFreeAsync(ws, s));
}

void launchQueryBatch(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets,
                      uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                      hipStream_t s) {
    if (!nbQueries) return;
    uint64_t maxN = 0;
    for (uint32_t q = 0; q < nbQueries; q++) maxN = std::max(maxN, hOffsets[q + 1] - hOffsets[q]);
    const uint64_t total = hOffsets[nbQueries];
    const uint64_t accN = uint64_t(nbQueries) * std::max(v.nbDocs, 1u);
    const size_t bytes = sizeof(QNode) * std::max<uint64_t>(total * v.L, 1) + 4ull * nbQueries + 8 * accN + accN + 1024;
    void* ws = nullptr;
    RBOW_HIP(hipMallocAsync(&ws, bytes, s));
    QNode* nodes = static_cast<QNode*>(ws);
    double* acc = reinterpret_cast<double*>(nodes + std::max<uint64_t>(total * v.L, 1));
    uint8_t* touched = reinterpret_cast<uint8_t*>(acc + accN);
    uint32_t* nodeCount = reinterpret_cast<uint32_t*>((reinterpret_cast<uintptr_t>(touched + accN) + 15) & ~uintptr_t(15));
    k_fill64<<<gridFor(accN, 256, 8192), 256, 0, s>>>(reinterpret_cast<uint64_t*>(acc), 0, accN);
    k_fill8<<<gridFor(accN, 256, 8192), 256, 0, s>>>(touched, accN);
    k_queryPrep<<<nbQueries, 1024, 0, s>>>(v, dLeaf, dOffsets, K, nodes, nodeCount);
    const uint32_t gx = gridFor(maxN * v.L, 4, 256);
    k_score<<<dim3(gx, nbQueries), 256, 0, s>>>(v, dOffsets, nodes, nodeCount, acc, touched);
    k_topk<<<nbQueries, 1024, 0, s>>>(v, K, acc, touched, dDocs, dScores, dCounts);
    RBOW_HIP(hipGetLastError());
    RBOW_HIP(hipFreeAsync(ws, s));
}

}  // namespace rbow_amd
