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
constexpr uint32_t kSegBatch = 1024;                    // query leaves staged per scoring round (one per thread)
constexpr uint32_t kTopkRegs = 16;                      // candidate keys per thread kept in registers by k_topk

// This file is compiled with binary32 denormals flushed by the hardware
// (-fgpu-flush-denormals-to-zero: the inputs and the result of every binary32
// operation, sign kept), which is the reference's DAZ/FTZ host arithmetic
// (-ffast-math, CMakeLists.txt:16).  ftz() states the flush where it matters.
__device__ inline float ftz(float x) { return fabsf(x) < 1.17549435e-38f ? copysignf(0.0f, x) : x; }
__device__ inline float fmulz(float a, float b) { return ftz(ftz(a) * ftz(b)); }

// DataBase.cpp:128-130 for one (node, doc) term: |qE - tE| - (qE + tE) with
// tE = m * w * rcp, each operation rounded (no contraction into fma) and
// flushed by the hardware mode above, as the reference computes it.
__device__ inline float termNormal(uint32_t m, float w, float rcp, float qE) {
#pragma clang fp contract(off)
    const float tE = float(m) * w * rcp;
    return fabsf(qE - tE) - (qE + tE);
}

struct QNode {
    uint32_t v;      // node index within its level
    uint32_t t;      // depth from the leaf level
    uint32_t cnt;    // query leaves under the node
    float qE;
    float w;
    uint32_t pad;
    uint64_t a, e;   // posting slice in the main CSR
    uint64_t aD, eD; // posting slice in the delta CSR
};
static_assert(sizeof(QNode) == Database::kQNodeBytes, "QNode size");

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

// Block-wide scan of one value per thread pair (1024 threads): returns the
// exclusive prefix of this thread's value; *total gets the block total.
// `red` holds 16 wave totals.
__device__ inline uint32_t blockScan(uint32_t val, uint32_t* red, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t s = val;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = uint32_t(__shfl_up(int(s), d, 64));
        if (lane >= uint32_t(d)) s += t;
    }
    __syncthreads();  // red is free
    if (lane == 63) red[wave] = s;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < 16 ? red[lane] : 0u;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t t = uint32_t(__shfl_up(int(w), d, 64));
            if (lane >= uint32_t(d)) w += t;
        }
        if (lane < 16) red[lane] = w;
    }
    __syncthreads();
    *total = red[15];
    return (wave ? red[wave - 1] : 0u) + s - val;
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

// out = a ++ b, leaf by leaf (one wave per leaf, coalesced copies); offC[x] =
// offA[x] + offB[x] is written by k_newOff beforehand.
__global__ __launch_bounds__(256) void k_mergeLeaves(const uint64_t* __restrict__ offA, const uint32_t* __restrict__ postA,
                                                    const uint64_t* __restrict__ offB, const uint32_t* __restrict__ postB,
                                                    uint64_t nbLeaves, const uint64_t* __restrict__ offC,
                                                    uint32_t* __restrict__ postC) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t x = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); x < nbLeaves; x += uint64_t(gridDim.x) * 4) {
        const uint64_t a0 = offA[x], na = offA[x + 1] - a0, b0 = offB[x], nb = offB[x + 1] - b0;
        uint32_t* dst = postC + offC[x];
        for (uint64_t i = lane; i < na; i += 64) dst[i] = postA[a0 + i];
        dst += na;
        for (uint64_t i = lane; i < nb; i += 64) dst[i] = postB[b0 + i];
    }
}
__global__ void k_countBad(const uint32_t* __restrict__ leaf, uint64_t n, uint64_t nbLeaves, uint32_t* __restrict__ flag) {
    uint32_t bad = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        bad += leaf[i] >= nbLeaves;
    if (bad) atomicAdd(flag, bad);
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

// Bitonic sort, ascending, of s[0, n2), n2 = 1024 * E, by a 1024-thread block
// with E consecutive keys per thread in registers: steps whose partners i ^ j
// are in the same thread (j < E) swap registers, steps within one wave
// (j < 64 E) exchange with the partner lane, and only steps across waves pass
// through LDS (two barriers each).  s (16-byte aligned) is complete and visible
// on return.  A thread's E keys move to and from LDS as 16-byte vectors: scalar
// accesses E words apart from lane to lane hit the same banks (16-way at E = 8).
template <int E>
__device__ inline void ldsLoadKeys(const uint32_t* p, uint32_t (&r)[E]) {
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int c = 0; c < E / 4; c++) {
            const uint4 v = reinterpret_cast<const uint4*>(p)[c];
            r[4 * c] = v.x;
            r[4 * c + 1] = v.y;
            r[4 * c + 2] = v.z;
            r[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) r[e] = p[e];
    }
}
template <int E>
__device__ inline void ldsStoreKeys(uint32_t* p, const uint32_t (&r)[E]) {
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int c = 0; c < E / 4; c++)
            reinterpret_cast<uint4*>(p)[c] = make_uint4(r[4 * c], r[4 * c + 1], r[4 * c + 2], r[4 * c + 3]);
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) p[e] = r[e];
    }
}
template <int E>
__device__ void bitonicSortBlock(uint32_t* s, uint32_t n2) {
    const uint32_t tid = threadIdx.x, i0 = tid * E;
    uint32_t r[E];
    ldsLoadKeys<E>(s + i0, r);
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        uint32_t j = k >> 1;
        for (; j >= uint32_t(64 * E); j >>= 1) {  // across waves
            __syncthreads();
            ldsStoreKeys<E>(s + i0, r);
            __syncthreads();
            uint32_t o[E];  // the partners: (i0 + e) ^ j = (i0 ^ j) + e, as j >= E
            ldsLoadKeys<E>(s + (i0 ^ j), o);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const uint32_t i = i0 + e;
                r[e] = ((i & j) == 0) == ((i & k) == 0) ? min(r[e], o[e]) : max(r[e], o[e]);
            }
        }
        for (; j >= uint32_t(E); j >>= 1) {  // across lanes of one wave
            const int lm = int(j / E);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const uint32_t i = i0 + e, o = uint32_t(__shfl_xor(int(r[e]), lm, 64));
                r[e] = ((i & j) == 0) == ((i & k) == 0) ? min(r[e], o) : max(r[e], o);
            }
        }
#pragma unroll
        for (int jj = E / 2; jj > 0; jj >>= 1) {  // within the thread
            if (uint32_t(jj) > (k >> 1)) continue;
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (e & jj) continue;
                const uint32_t a = r[e], c = r[e | jj];
                const bool up = ((i0 + e) & k) == 0;
                r[e] = up ? min(a, c) : max(a, c);
                r[e | jj] = up ? max(a, c) : min(a, c);
            }
        }
    }
    __syncthreads();
    ldsStoreKeys<E>(s + i0, r);
    __syncthreads();
}

// ------------------------------------------------------------- query ---
// One 1024-thread block per query (DataBase.cpp:71-130): sort the leaves in
// LDS, binary64 L1 norm, then per level the distinct nodes with their
// multiplicities: every leaf, and inner nodes under the posting threshold
// (postings of the main and the delta CSR together).  A query holding a leaf
// id >= nbLeaves scores nothing and sets *bad (when given).
__global__ __launch_bounds__(1024) void k_queryPrep(DbView v, const uint32_t* __restrict__ leaf,
                                                   const uint64_t* __restrict__ off, uint32_t K,
                                                   QNode* __restrict__ nodes, uint32_t* __restrict__ nodeCount,
                                                   uint32_t* __restrict__ work, uint32_t* __restrict__ bad) {
    __shared__ __align__(16) uint32_t s[kSortCap];
    __shared__ uint16_t sStart[kSortCap + 1];  // run starts of one level (positions in s), then n
    __shared__ double red[16];
    __shared__ uint32_t redU[16];
    __shared__ uint32_t nNodes, nInner;
    __shared__ unsigned long long sWork;
    const uint32_t q = blockIdx.x, tid = threadIdx.x;
    const uint64_t beg = off[q];
    const uint32_t n = uint32_t(off[q + 1] - beg);
    if (n == 0) {  // an empty query (also the padding of a fixed-size batch) scores nothing
        if (tid == 0) {
            nodeCount[2 * q] = 0;
            nodeCount[2 * q + 1] = 0;
            work[q] = 0;
        }
        return;
    }
    uint32_t n2 = 1024;
    while (n2 < n) n2 <<= 1;
    {  // this thread's leaf ids, every load in flight together (a row past n reads the last id, unused)
        uint32_t lv[kSortCap / 1024];
#pragma unroll
        for (uint32_t e = 0; e < kSortCap / 1024; e++) lv[e] = leaf[beg + min(tid + e * 1024, n - 1)];
#pragma unroll
        for (uint32_t e = 0; e < kSortCap / 1024; e++) {
            const uint32_t i = tid + e * 1024;
            if (i < n2) s[i] = i < n ? lv[e] : 0xFFFFFFFFu;
        }
    }
    if (tid == 0) {
        nNodes = 0;
        nInner = 0;
        sWork = 0;
    }
    __syncthreads();
    // bitonic sort, ascending, of n2 >= 1024 keys (padded with 0xFFFFFFFF)
    switch (n2 >> 10) {
    case 1: bitonicSortBlock<1>(s, n2); break;
    case 2: bitonicSortBlock<2>(s, n2); break;
    case 4: bitonicSortBlock<4>(s, n2); break;
    case 8: bitonicSortBlock<8>(s, n2); break;
    default: bitonicSortBlock<16>(s, n2); break;
    }
    // the largest leaf id is s[n - 1]: nothing below reads W or the CSRs through an invalid one
    if (n && s[n - 1] >= v.nbLeaves) {
        if (tid == 0) {
            nodeCount[2 * q] = 0;
            nodeCount[2 * q + 1] = 0;
            work[q] = 0;
            if (bad) atomicOr(bad, 1u);
        }
        return;
    }
    // the binary64 norm, per level a sum over i = tid, tid + 1024, ...: the level's weight loads
    // of this thread are issued together (a row past n reads the last leaf's weight, unused),
    // not one round trip per leaf
    double part = 0;
    for (uint32_t t = 0; t < v.L; t++) {
        const float* W = v.W + levelBeg(v.lg2b, v.L - 1 - t) * v.b;
        float wv[kSortCap / 1024];
#pragma unroll
        for (uint32_t e = 0; e < kSortCap / 1024; e++) wv[e] = W[s[min(tid + e * 1024, n - 1)] >> (t * v.lg2b)];
#pragma unroll
        for (uint32_t e = 0; e < kSortCap / 1024; e++)
            if (tid + e * 1024 < n) part += double(wv[e]);
    }
    const double l1 = blockSumF64(part, red);
    const float qrcp = ftz(float(1.0 / l1));  // DataBase.cpp:77-94
    const uint64_t thr = 16ull + 4ull * K;    // DataBase.h:60-61
    // leaves at out[0, nLeaf), inner nodes at out[n, n + nInner)
    QNode* out = nodes + uint64_t(v.L) * beg;
    // per level: the runs of equal node ids (leaf >> shift) in the sorted leaves;
    // thread tid scans the contiguous chunk [c0, c1) for run starts
    const uint32_t E = (n + 1023) / 1024, c0 = min(n, tid * E), c1 = min(n, c0 + E);
    // the thread's chunk and the key before it, read from LDS once for all levels
    constexpr uint32_t kChunk = kSortCap / 1024;
    uint32_t ck[kChunk];
#pragma unroll
    for (uint32_t e = 0; e < kChunk; e++) ck[e] = c0 + e < c1 ? s[c0 + e] : 0u;
    const uint32_t before = c0 > 0 && c0 < c1 ? s[c0 - 1] : 0u;
    unsigned long long myWork = 0;  // postings of the query's leaves: the scoring's share of work
    for (uint32_t t = 0; t < v.L; t++) {
        const uint32_t shift = t * v.lg2b;
        const float* W = v.W + levelBeg(v.lg2b, v.L - 1 - t) * v.b;
        uint32_t startMask = 0;  // bit e: position c0 + e starts a run of node ids (leaf >> shift)
#pragma unroll
        for (uint32_t e = 0; e < kChunk; e++) {
            const uint32_t prev = e ? ck[e - 1] : before;
            const bool st = c0 + e < c1 && (c0 + e == 0 || (prev >> shift) != (ck[e] >> shift));
            startMask |= uint32_t(st) << e;
        }
        uint32_t nRuns;
        uint32_t pos = blockScan(uint32_t(__builtin_popcount(startMask)), redU, &nRuns);
        for (uint32_t m = startMask; m; m &= m - 1u) sStart[pos++] = uint16_t(c0 + __builtin_ctz(m));
        if (tid == 0) sStart[nRuns] = uint16_t(n);
        __syncthreads();
        // runs tid, tid + 1024, ... four at a time: their posting-offset loads are in flight together
        for (uint32_t k0 = tid; k0 < nRuns; k0 += 4 * 1024) {
            uint32_t x[4], cnt[4];
            uint64_t a[4], e[4], aD[4], eD[4];
            float w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const uint32_t k = k0 + u * 1024;
                x[u] = 0;
                cnt[u] = 0;
                a[u] = 0;
                e[u] = 0;
                aD[u] = 0;
                eD[u] = 0;
                w[u] = 0.0f;
                if (k < nRuns) {
                    const uint32_t st = sStart[k];
                    cnt[u] = uint32_t(sStart[k + 1]) - st;
                    x[u] = s[st] >> shift;
                    a[u] = v.leafOff[uint64_t(x[u]) << shift];
                    e[u] = v.leafOff[(uint64_t(x[u]) + 1) << shift];
                    aD[u] = v.leafOffD[uint64_t(x[u]) << shift];
                    eD[u] = v.leafOffD[(uint64_t(x[u]) + 1) << shift];
                    w[u] = W[x[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (k0 + u * 1024 >= nRuns || (t != 0 && (e[u] - a[u]) + (eD[u] - aD[u]) >= thr)) continue;
                if (t == 0) myWork += (e[u] - a[u]) + (eD[u] - aD[u]);
                QNode nd;
                nd.v = x[u];
                nd.t = t;
                nd.cnt = cnt[u];
                nd.qE = fmulz(fmulz(float(cnt[u]), w[u]), qrcp);  // DataBase.cpp:117
                nd.w = w[u];
                nd.pad = 0;
                nd.a = a[u];
                nd.e = e[u];
                nd.aD = aD[u];
                nd.eD = eD[u];
                if (t == 0) out[atomicAdd(&nNodes, 1u)] = nd;
                else out[n + atomicAdd(&nInner, 1u)] = nd;
            }
        }
        if (t == 0 && myWork) atomicAdd(&sWork, myWork);
        __syncthreads();  // sStart is rewritten by the next level
    }
    if (tid == 0) {
        nodeCount[2 * q] = nNodes;
        nodeCount[2 * q + 1] = nInner;
        work[q] = uint32_t(min(sWork, 0xFFFFFFFFull));
    }
}

// Range directory of the main CSR: dir[(Rm+1)*leaf + r] = first posting of the
// leaf whose slot is >= r*kRangeDocs, relative to leafOff[leaf] (lists are
// sorted by slot, and every main slot is below Rm*kRangeDocs, so dir[..+Rm] is
// the list length).  One wave per leaf reads the list once: posting i writes
// the entries of the ranges after its predecessor's, up to its own.
__global__ __launch_bounds__(256) void k_buildDir(DbView v, uint32_t* __restrict__ dir) {
    constexpr uint32_t S = Database::kRangeDocs;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t R1 = v.Rm + 1;
    for (uint64_t leaf = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); leaf < v.nbLeaves;
         leaf += uint64_t(gridDim.x) * 4) {
        const uint64_t a = v.leafOff[leaf], len = v.leafOff[leaf + 1] - a;
        uint32_t* d = dir + leaf * R1;
        const uint32_t* p = v.post + a;
        for (uint64_t i = lane; i < len; i += 64) {
            const uint32_t rc = p[i] / S;
            const uint32_t r0 = i ? p[i - 1] / S + 1 : 0u;
            for (uint32_t r = r0; r <= rc; r++) d[r] = uint32_t(i);
        }
        const uint32_t rl = len ? p[len - 1] / S + 1 : 0u;  // ranges past the last posting
        for (uint32_t r = rl + lane; r < R1; r += 64) d[r] = uint32_t(len);
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

// Exact selection, by 8-bit MSB-first radix passes, of the `need` smallest of
// a block's unique 64-bit keys; returns the largest selected key, so the
// selection is every key <= it.  Requires 1 <= need <= #keys.  forEach(f)
// calls f(key, valid) over each thread's share of the keys, wave-uniformly (a
// lane with nothing at a step passes valid = false).  Lanes of a wave whose
// keys fall in one bin add to the histogram once (the first passes put nearly
// every key in the same bin), and the passes stop as soon as the bin holding
// the need-th key is selected whole.  hist: 256 words, red: 3 words.
template <class ForEach>
__device__ uint64_t blockSelectKth(ForEach forEach, uint32_t need, uint32_t* hist, uint32_t* red) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    uint64_t prefix = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        __syncthreads();
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        forEach([&](uint64_t k, bool valid) {
            const bool in = valid && (shift == 56 || (k >> (shift + 8)) == (prefix >> (shift + 8)));
            const uint64_t act = __ballot(in);
            if (!act) return;
            const uint32_t bin = uint32_t(k >> shift) & 255u;
            const uint32_t first = uint32_t(__builtin_ctzll(act));
            const uint32_t fb = uint32_t(__shfl(int(bin), int(first), 64));
            const bool same = in && bin == fb;
            const uint64_t sm = __ballot(same);
            if (lane == first) atomicAdd(&hist[fb], uint32_t(__builtin_popcountll(sm)));
            else if (in && !same) atomicAdd(&hist[bin], 1u);
        });
        __syncthreads();
        if (tid < 64) {  // wave 0: the bin holding the need-th key
            const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
            const uint32_t sum = h0 + h1 + h2 + h3;
            uint32_t incl = sum;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t t = uint32_t(__shfl_up(int(incl), d, 64));
                if (lane >= uint32_t(d)) incl += t;
            }
            const uint32_t excl = incl - sum;
            if (excl < need && need <= incl) {  // one of this lane's four bins (selects, no indexed array)
                uint32_t before = excl, j = 3, hj = h3;
                if (before + h0 >= need) {
                    j = 0;
                    hj = h0;
                } else if (before += h0, before + h1 >= need) {
                    j = 1;
                    hj = h1;
                } else if (before += h1, before + h2 >= need) {
                    j = 2;
                    hj = h2;
                } else {
                    before += h2;
                }
                red[0] = 4 * lane + j;
                red[1] = need - before;
                red[2] = hj;
            }
        }
        __syncthreads();
        prefix |= uint64_t(red[0]) << shift;
        need = red[1];
        if (need == red[2]) {  // every key of the bin is selected: the lower digits do not matter
            __syncthreads();
            return prefix | ((1ull << shift) - 1ull);
        }
    }
    __syncthreads();
    return prefix;
}

// First position in p[0, n) whose slot is >= key (p sorted ascending).
__device__ inline uint32_t lowerBound(const uint32_t* p, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (p[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The launch's queries by descending work (ties by index): order[rank] = query.
// The scoring grid then starts the longest queries first, so the last
// workgroups to start are the shortest ones.
constexpr uint32_t kOrderCap = 4096;  // larger batches keep their own order
__global__ __launch_bounds__(1024) void k_orderQueries(const uint32_t* __restrict__ work, uint32_t nq,
                                                       uint32_t* __restrict__ order) {
    __shared__ uint32_t sW[kOrderCap];
    for (uint32_t i = threadIdx.x; i < nq; i += 1024) sW[i] = work[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nq; i += 1024) {
        const uint32_t wi = sW[i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < nq; j++) {
            const uint32_t wj = sW[j];
            rank += wj > wi || (wj == wi && j < i);
        }
        order[rank] = i;
    }
}

// One workgroup per (document range r, query q).  The binary64 accumulators of
// the range's kRangeDocs slots live in LDS (ds_add_f64), so a posting costs a
// coalesced 4-byte read and an LDS atomic instead of a scattered global one.
// An accumulator starts at -0.0 and every term is added as +0.0 or a negative
// double (never as -0.0), so a slot was scored iff its bits differ from -0.0.
//   Leaves, in rounds of kSegBatch: the in-range segment of each query leaf
//   comes from the main CSR's range directory, and (for ranges that hold
//   delta slots) from a binary search in the leaf's delta slice, in rounds of
//   their own.  Non-empty segments are compacted by a block scan into a table
//   over the concatenated postings: end position, address of the segment's
//   position 0 (so the walk does not care which CSR it reads), qE and w.
//   Each wave walks a contiguous share of the concatenated postings, U steps
//   of 64 per iteration (balanced however skewed the leaves are).  Per
//   iteration the last position of every segment ending inside it is marked
//   in a per-wave bitmap: a position's segment is the current one plus the
//   marks below it (mbcnt), and a segment starts after a marked position.
//   A run of one doc's duplicates (DataBase.cpp:138-149) is scored once, by
//   its first posting, with m = run length from a ballot of run starts; a
//   document's postings are all in one CSR, so no run spans two segments.
//   Inner nodes (fewer than 16+4K postings, DataBase.cpp:150-164): one wave
//   per node, multiplicity by scanning the node's slice in each CSR.
// The range then keeps its K smallest (score, doc id) keys as candidates for
// the query's final selection.  Query blockIdx.y of the launch is order[y]
// when an order is given.
constexpr uint32_t kInnerCap = 4 * kSegBatch / 16;  // per-wave list of an inner node's in-range postings (u32)

__device__ inline bool scored(double acc) { return __double_as_longlong(acc) != (long long)0x8000000000000000ull; }

__device__ inline uint64_t uniform64(uint64_t x) {
    return uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(x))))) |
           (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(x >> 32))))) << 32);
}

// A posting read through a global (not flat) pointer: a flat load also counts as
// an LDS access, so every LDS wait behind it would wait for the load as well.
typedef const uint32_t __attribute__((address_space(1)))* GlobalU32;
__device__ inline uint32_t postingAt(uint64_t vbase, uint32_t pos) {
    return *reinterpret_cast<GlobalU32>(vbase + 4ull * pos);
}

// U: steps of 64 postings per walk iteration (the loads in flight per lane), the
// scoring tile shape the database's tuner picks (4 or 6: at 8 the walk's
// registers no longer fit 64 VGPRs, the budget of 8 waves per SIMD).
template <int U>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_scoreRange(DbView v, const uint64_t* __restrict__ off, uint32_t q0,
                                                    const QNode* __restrict__ nodes,
                                                    const uint32_t* __restrict__ nodeCount,
                                                    const uint32_t* __restrict__ order, uint32_t K,
                                                    uint64_t* __restrict__ candKeys, uint32_t* __restrict__ candCount) {
    constexpr uint32_t S = Database::kRangeDocs;
    __shared__ double sAcc[S];                 // reused as the keys of the selection
    __shared__ float sRcp[S];
    // per segment: the address of its posting at concatenated position 0, then (qE, w) of its node;
    // the inner nodes reuse the table as per-wave lists of in-range postings
    __shared__ uint64_t sSegTab[2 * kSegBatch];
    uint64_t* const sSegVb = sSegTab;
    float2* const sSegQW = reinterpret_cast<float2*>(sSegTab + kSegBatch);
    __shared__ uint32_t sSegEnd[kSegBatch];    // inclusive prefix of segment lengths
    __shared__ uint64_t sMark[16 * U];         // per wave: bit 64u+b marks position base+64u+b as a segment's last
    __shared__ uint32_t sHist[256], sRed[16], sCount;
    const uint32_t r = blockIdx.x, qc = order ? order[blockIdx.y] : blockIdx.y, q = q0 + qc, tid = threadIdx.x,
                   lane = tid & 63, wave = tid >> 6;
    if (nodeCount[2 * q] == 0 && nodeCount[2 * q + 1] == 0) {  // nothing to score (an empty query)
        if (tid == 0) candCount[uint64_t(q) * v.R + r] = 0;
        return;
    }
    const uint32_t s0 = r * S, ns = min(S, v.nbDocs - s0);
    for (uint32_t i = tid; i < S; i += 1024) {
        sAcc[i] = -0.0;
        sRcp[i] = i < ns ? v.rcp[s0 + i] : 0.0f;
    }
    if (tid < 16 * U) sMark[tid] = 0;
    if (tid == 0) sCount = 0;
    __syncthreads();
    const uint64_t beg = off[qc];  // off points at the launch's first query
    const uint32_t n = uint32_t(off[qc + 1] - beg);
    const QNode* nd = nodes + uint64_t(v.L) * beg;
    const uint32_t nLeaf = nodeCount[2 * q], nInner = nodeCount[2 * q + 1];
    const bool hasMain = r < v.Rm;                                      // block-uniform
    const bool hasDelta = v.nbPostD != 0 && s0 + S > v.slotD0;
    const uint32_t nStage = (hasMain ? nLeaf : 0u) + (hasDelta ? nLeaf : 0u);
    uint64_t* mark = sMark + wave * U;
    for (uint32_t k0 = 0; k0 < nStage; k0 += kSegBatch) {
        // ---- stage this round's segments (one query leaf and CSR per thread) ----
        uint32_t len = 0;
        uint64_t sb = 0;  // address of the segment's first posting
        float qE = 0.0f, w = 0.0f;
        if (k0 + tid < nStage) {
            const bool main = hasMain && k0 + tid < nLeaf;
            const QNode node = nd[k0 + tid - (main || !hasMain ? 0u : nLeaf)];
            if (main) {
                const uint32_t* d = v.dir + uint64_t(node.v) * (v.Rm + 1) + r;
                const uint32_t a = d[0];
                len = d[1] - a;
                sb = reinterpret_cast<uint64_t>(v.post + node.a + a);
            } else {
                const uint32_t* p = v.postD + node.aD;
                const uint32_t m = uint32_t(node.eD - node.aD);
                const uint32_t lo = s0 <= v.slotD0 ? 0u : lowerBound(p, m, s0);
                len = lowerBound(p + lo, m - lo, s0 + S);
                sb = reinterpret_cast<uint64_t>(p + lo);
            }
            qE = node.qE;
            w = node.w;
        }
        uint32_t total, nc;
        const uint32_t ex = blockScan(len, sRed, &total);
        const uint32_t cx = blockScan(len ? 1u : 0u, sRed, &nc);
        if (len) {
            sSegVb[cx] = sb - 4ull * ex;  // wraps below zero as an integer; only vbase + 4 * position is read
            sSegEnd[cx] = ex + len;
            sSegQW[cx] = make_float2(qE, w);
        }
        __syncthreads();
        // ---- walk the postings: wave-contiguous shares, U steps of 64 per iteration ----
        const uint32_t per = ((total + 16 * 64 - 1) / (16 * 64)) * 64;
        const uint32_t wBeg = wave * per, wEnd = min(total, wBeg + per);
        uint32_t seg = 0, carry = 0;  // carry: the slot at position base - 1 when base is inside a segment
        if (wBeg < wEnd) {  // first segment whose end exceeds the share's start
            uint32_t lo = 0, hi = nc - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if (sSegEnd[mid] > wBeg) hi = mid;
                else lo = mid + 1;
            }
            seg = uint32_t(__builtin_amdgcn_readfirstlane(int(lo)));
            if (wBeg > (seg ? sSegEnd[seg - 1] : 0u)) carry = postingAt(sSegVb[seg], wBeg - 1);
        }
        for (uint32_t base = wBeg; base < wEnd; base += 64 * U) {
            const uint32_t itEnd = min(base + 64 * U, wEnd);
            while (sSegEnd[seg] <= base) seg++;
            seg = uint32_t(__builtin_amdgcn_readfirstlane(int(seg)));
            const bool baseStarts = base == (seg ? sSegEnd[seg - 1] : 0u);
            // mark the last position of each segment that ends inside this iteration
            // (ends ascend, so the first chunk of lanes with an end past it is the last)
            for (uint32_t j0 = seg;; j0 += 64) {
                const uint32_t j = j0 + lane;
                const uint32_t e = j < nc ? sSegEnd[j] : 0xFFFFFFFFu;
                if (e < itEnd) {
                    const uint32_t pm = e - 1 - base;  // e > base: segment seg ends after base
                    atomicOr(reinterpret_cast<uint32_t*>(mark) + (pm >> 5), 1u << (pm & 31));
                }
                if (__ballot(e >= itEnd)) break;
            }
            uint64_t M[U];
#pragma unroll
            for (int u = 0; u < U; u++) M[u] = uniform64(mark[u]);
            if (lane < U) mark[lane] = 0;  // after the reads: a wave's LDS accesses complete in order
            // step u, lane l: position base + 64u + l; sgs: its segment, bit 31 set where a segment starts
            uint32_t slots[U], sgs[U];
            uint32_t cum = 0;  // marks in the earlier steps
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t idx = base + 64 * u + lane;
                const uint32_t sg =
                    seg + cum + __builtin_amdgcn_mbcnt_hi(uint32_t(M[u] >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(M[u]), 0u));
                cum += uint32_t(__builtin_popcountll(M[u]));
                const uint64_t starts = (M[u] << 1) | (u == 0 ? uint64_t(baseStarts) : (M[u ? u - 1 : 0] >> 63));
                sgs[u] = sg | (uint32_t((starts >> lane) & 1ull) << 31);  // sg < nc for every lane: marks lie below itEnd
                slots[u] = idx < itEnd ? postingAt(sSegVb[sg], idx) : 0xFFFFFFFFu;
            }
            // run starts (a new segment or a new slot); slots[u] keeps the slot of a run start only
            uint64_t bnd[U];
            uint32_t prevLast = carry;  // the slot before the step
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t slot = slots[u];
                const bool valid = base + 64 * u + lane < itEnd;
                // wave_shr:1 -- lane l gets lane l-1's slot, lane 0 the slot before the step
                const uint32_t before = uint32_t(__builtin_amdgcn_update_dpp(int(prevLast), int(slot), 0x138, 0xF, 0xF, false));
                prevLast = uint32_t(__builtin_amdgcn_readlane(int(slot), 63));
                const bool head = valid && ((sgs[u] >> 31) != 0 || before != slot);
                bnd[u] = __ballot(head || !valid);
                slots[u] = head ? slot : 0xFFFFFFFFu;
            }
            carry = prevLast;  // position base + 64U - 1
            // steps in reverse: nxt = first boundary after step u, as a position of the
            // iteration (64 * U: none), wave-uniform, so a run length needs no per-lane loop
            uint32_t nxt = 64 * U;
#pragma unroll
            for (int u = U - 1; u >= 0; u--) {
                const uint32_t after = nxt;
                if (bnd[u]) nxt = 64 * u + uint32_t(__builtin_ctzll(bnd[u]));
                const uint32_t slot = slots[u];
                if (slot == 0xFFFFFFFFu) continue;  // not a run start
                // run length: distance to the next boundary over this iteration's steps
                const uint64_t above = lane == 63 ? 0ull : (bnd[u] >> (lane + 1));
                const bool found = above != 0 || after < 64 * U;
                uint32_t m = above ? uint32_t(__builtin_ctzll(above)) + 1 : after - (64 * u + lane);
                const uint32_t idx = base + 64 * u + lane, sg = sgs[u] & 0x7FFFFFFFu;
                if (!found || idx + m >= wEnd) {
                    // no run start follows in this iteration: the run may go on past it or past the share
                    const uint32_t segEnd = sSegEnd[sg];
                    const uint64_t vb = sSegVb[sg];
                    while (idx + m < segEnd && postingAt(vb, idx + m) == slot) m++;
                }
                const float2 qw = sSegQW[sg];
                const uint32_t j = slot - s0;
                atomicAdd(&sAcc[j], double(termNormal(m, qw.y, sRcp[j], qw.x)) + 0.0);  // + 0.0: -0.0 -> +0.0
            }
        }
        __syncthreads();  // segment arrays are reused by the next round
    }
    // Inner nodes: a node's slice is its leaves' lists one after the other, so a
    // doc's postings in it need not be adjacent.  The wave compacts the slice's
    // in-range postings (in slice order) into its LDS list, and each of them counts
    // its equals there: the first of equal slots scores the doc with m = count.
    uint32_t* inList = reinterpret_cast<uint32_t*>(sSegTab) + wave * kInnerCap;
    for (uint32_t k = wave; k < nInner; k += 16) {
        const QNode node = nd[n + k];
        for (int c = 0; c < 2; c++) {  // the node's slice in the main CSR, then in the delta
            const uint32_t* p = c ? v.postD + node.aD : v.post + node.a;
            const uint32_t len = uint32_t(c ? node.eD - node.aD : node.e - node.a);
            if (c ? !hasDelta : !hasMain) continue;
            uint32_t nIn = 0;
            for (uint32_t i0 = 0; i0 < len; i0 += 64) {
                const uint32_t slot = i0 + lane < len ? p[i0 + lane] : 0xFFFFFFFFu;
                const bool in = slot - s0 < S;  // slots are below 2^32 - 1 - S
                const uint64_t bal = __ballot(in);
                const uint32_t pos = nIn + __builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u));
                if (in && pos < kInnerCap) inList[pos] = slot;
                nIn += uint32_t(__builtin_popcountll(bal));
            }
            if (nIn <= kInnerCap) {  // the wave's own LDS writes are complete before its later reads
                for (uint32_t i = lane; i < nIn; i += 64) {
                    const uint32_t slot = inList[i];
                    bool first = true;
                    uint32_t m = 0;
                    for (uint32_t x = 0; x < nIn; x++) {
                        if (inList[x] == slot) {
                            m++;
                            if (x < i) first = false;
                        }
                    }
                    if (!first) continue;
                    const uint32_t j = slot - s0;
                    atomicAdd(&sAcc[j], double(termNormal(m, node.w, sRcp[j], node.qE)) + 0.0);  // + 0.0: -0.0 -> +0.0
                }
                continue;
            }
            // more in-range postings than the list holds (only with K above 60): count in the slice itself
            for (uint32_t i = lane; i < len; i += 64) {
                const uint32_t slot = p[i], j = slot - s0;
                if (j >= S) continue;
                bool first = true;
                uint32_t m = 0;
                for (uint32_t x = 0; x < len; x++) {
                    if (p[x] == slot) {
                        m++;
                        if (x < i) first = false;
                    }
                }
                if (!first) continue;
                atomicAdd(&sAcc[j], double(termNormal(m, node.w, sRcp[j], node.qE)) + 0.0);
            }
        }
    }
    __syncthreads();
    // keys in place of the accumulators; untouched slots get the largest key
    uint64_t* sKey = reinterpret_cast<uint64_t*>(sAcc);
    uint32_t mine = 0;
    for (uint32_t i = tid; i < S; i += 1024) {
        const bool t = i < ns && scored(sAcc[i]);
        sKey[i] = t ? rankKey(sAcc[i], v.docId[s0 + i]) : ~0ull;
        mine += t;
    }
    atomicAdd(&sCount, mine);
    __syncthreads();
    const uint32_t nT = sCount, need = min(K, nT);
    uint64_t* out = candKeys + (uint64_t(q) * v.R + r) * K;
    if (need == 0) {
        if (tid == 0) candCount[uint64_t(q) * v.R + r] = 0;
        return;
    }
    const uint64_t kth = need == nT ? ~0ull - 1
                                    : blockSelectKth(
                                          [&](auto f) {
                                              for (uint32_t i = tid; i < S; i += 1024) f(sKey[i], true);  // S = 4 x 1024
                                          },
                                          need, sHist, sRed);
    __syncthreads();
    if (tid == 0) sCount = 0;
    __syncthreads();
    for (uint32_t i = tid; i < S; i += 1024) {
        const uint64_t key = sKey[i];
        if (key <= kth) out[atomicAdd(&sCount, 1u)] = key;
    }
    if (tid == 0) candCount[uint64_t(q) * v.R + r] = need;
}

// One block per query: the K smallest of the ranges' candidate keys, sorted
// ascending ((score, doc id): DataBase.cpp:170-176 with ties by doc id).
__global__ __launch_bounds__(1024) void k_topk(DbView v, uint32_t K, const uint64_t* __restrict__ candKeys,
                                              const uint32_t* __restrict__ candCount, uint32_t* __restrict__ outDocs,
                                              float* __restrict__ outScores, uint32_t* __restrict__ outCount) {
    __shared__ uint64_t sel[Database::kMaxResults];
    __shared__ uint32_t hist[256], red[16], nSel;
    const uint32_t q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t* keys = candKeys + uint64_t(q) * v.R * K;
    const uint32_t* cnt = candCount + uint64_t(q) * v.R;
    uint32_t mine = 0;
    for (uint32_t rr = tid; rr < v.R; rr += 1024) mine += cnt[rr];
    uint32_t tot;
    (void)blockScan(mine, red, &tot);
    const uint32_t need = min(K, tot);
    if (need == 0) {
        if (tid == 0) outCount[q] = 0;
        return;
    }
    // the candidates: range rr holds cnt[rr] keys at keys[rr*K ...].  Wave w takes
    // ranges w, w + 16, ..., lane l positions l, l + 64, ... of each.  When that is
    // at most kTopkRegs keys per thread they are loaded once, all loads in flight
    // together, and the selection passes run on registers; otherwise every pass
    // walks the ranges in memory.
    const uint32_t nR = (v.R + 15) / 16, nK = (K + 63) / 64;
    uint64_t kr[kTopkRegs];
    uint32_t kv = 0;  // bit i: kr[i] holds a key
    const bool inRegs = nR * nK <= kTopkRegs;
    if (inRegs) {
#pragma unroll
        for (uint32_t i = 0; i < kTopkRegs; i++) {
            const uint32_t rr = wave + 16 * (i / nK), pos = lane + 64 * (i % nK);
            const bool valid = i < nR * nK && rr < v.R && pos < cnt[rr];
            kr[i] = valid ? keys[uint64_t(rr) * K + pos] : 0ull;
            kv |= uint32_t(valid) << i;
        }
    }
    auto forEach = [&](auto f) {
        if (inRegs) {
#pragma unroll
            for (uint32_t i = 0; i < kTopkRegs; i++) f(kr[i], ((kv >> i) & 1u) != 0);
            return;
        }
        for (uint32_t rr = wave; rr < v.R; rr += 16) {
            const uint32_t c = cnt[rr];
            for (uint32_t r0 = 0; r0 < c; r0 += 64) {
                const bool valid = r0 + lane < c;
                f(valid ? keys[uint64_t(rr) * K + r0 + lane] : 0ull, valid);
            }
        }
    };
    const uint64_t kth = need == tot ? ~0ull - 1 : blockSelectKth(forEach, need, hist, red);
    if (tid == 0) nSel = 0;
    __syncthreads();
    forEach([&](uint64_t key, bool valid) {
        if (valid && key <= kth) sel[atomicAdd(&nSel, 1u)] = key;
    });
    __syncthreads();
    uint32_t n2 = 1;
    while (n2 < need) n2 <<= 1;
    for (uint32_t i = need + tid; i < n2; i += 1024) sel[i] = ~0ull;
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
    for (uint32_t i = tid; i < need; i += 1024) {
        outDocs[uint64_t(q) * K + i] = uint32_t(sel[i]);
        if (outScores) outScores[uint64_t(q) * K + i] = keyScore(sel[i]);
    }
    if (tid == 0) outCount[q] = need;
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

void launchBuildDir(const DbView& v, uint32_t* dir, hipStream_t s) {
    if (!v.Rm) return;
    k_buildDir<<<gridFor(v.nbLeaves, 4, 16384), 256, 0, s>>>(v, dir);
    RBOW_HIP(hipGetLastError());
}

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

void launchCountBadLeaves(const uint32_t* leaf, uint64_t n, uint64_t nbLeaves, uint32_t* flag, hipStream_t s) {
    if (!n) return;
    k_countBad<<<gridFor(n, 256, 4096), 256, 0, s>>>(leaf, n, nbLeaves, flag);
    RBOW_HIP(hipGetLastError());
}

void launchCsrFromPairs(uint64_t nbLeaves, const uint32_t* dLeaf, const uint32_t* dSlot, uint64_t n, DevCsr& out,
                        hipStream_t s) {
    const uint64_t n1 = nbLeaves + 1;
    RBOW_REQUIRE(n < (1ull << 31), RBOW_ERR_CAPACITY, "insert batch too large");
    RBOW_REQUIRE(out.off.size() >= n1 && out.post.size() >= n, RBOW_ERR_HIP, "CSR buffers too small");
    uint32_t bits = 1;
    while ((1ull << bits) < nbLeaves) bits++;
    size_t tempBytes = 0;
    RBOW_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, dLeaf, (uint32_t*)nullptr, dSlot,
                                                (uint32_t*)nullptr, int(n), 0, int(bits), s));
    size_t scanBytes = 0;
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, (uint64_t*)nullptr, (uint64_t*)nullptr, int(n1), s));
    const size_t bytes = 8 * n1 + 4 * n + std::max(tempBytes, scanBytes) + 512;
    void* ws = nullptr;
    RBOW_HIP(hipMallocAsync(&ws, bytes, s));
    uint64_t* cnt = static_cast<uint64_t*>(ws);
    uint32_t* sLeaf = reinterpret_cast<uint32_t*>(cnt + n1);
    void* temp = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(sLeaf + n) + 255) & ~uintptr_t(255));
    k_fill64<<<gridFor(n1, 256, 4096), 256, 0, s>>>(cnt, 0, n1);
    if (n) k_hist<<<gridFor(n, 256, 4096), 256, 0, s>>>(dLeaf, n, cnt);
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(temp, scanBytes, cnt, out.off.get(), int(n1), s));
    // stable: the slots of one leaf keep their ascending input order
    if (n)
        RBOW_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tempBytes, dLeaf, sLeaf, dSlot, out.post.get(), int(n), 0,
                                                    int(bits), s));
    RBOW_HIP(hipGetLastError());
    RBOW_HIP(hipFreeAsync(ws, s));
    out.n = n;
}

void launchMergeCsr(uint64_t nbLeaves, const DevCsr& a, const DevCsr& b, DevCsr& out, hipStream_t s) {
    const uint64_t n1 = nbLeaves + 1;
    RBOW_REQUIRE(out.off.size() >= n1 && out.post.size() >= a.n + b.n, RBOW_ERR_HIP, "CSR buffers too small");
    k_newOff<<<gridFor(n1, 256, 4096), 256, 0, s>>>(a.off.get(), b.off.get(), n1, out.off.get());
    k_mergeLeaves<<<gridFor(nbLeaves, 4, 16384), 256, 0, s>>>(a.off.get(), a.post.get(), b.off.get(), b.post.get(),
                                                             nbLeaves, out.off.get(), out.post.get());
    RBOW_HIP(hipGetLastError());
    out.n = a.n + b.n;
}

size_t queryWorkspaceBytes(const DbView& v, uint64_t totalDesc, uint32_t nbQueries, uint32_t K) {
    const uint64_t nbCand = uint64_t(nbQueries) * v.R * K;
    const uint64_t nbNodes = std::max<uint64_t>(totalDesc * v.L, 1);
    return sizeof(QNode) * nbNodes + 8 * nbCand + 4ull * nbQueries * v.R + 16ull * nbQueries + 1024;
}

void launchQueryBatchInWs(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, uint64_t totalDesc,
                          uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                          uint32_t* badFlag, void* ws, size_t wsBytes, hipStream_t s) {
    if (!nbQueries) return;
    RBOW_REQUIRE(v.R == divUp(v.nbDocs, Database::kRangeDocs) && v.R < 65536 && v.Rm <= v.R && (v.dir || !v.Rm),
                 RBOW_ERR_HIP, "stale range directory");
    RBOW_REQUIRE(wsBytes >= queryWorkspaceBytes(v, totalDesc, nbQueries, K), RBOW_ERR_HIP, "query workspace too small");
    const uint64_t nbCand = uint64_t(nbQueries) * v.R * K;
    const uint64_t nbNodes = std::max<uint64_t>(totalDesc * v.L, 1);
    QNode* nodes = static_cast<QNode*>(ws);
    uint64_t* candKeys = reinterpret_cast<uint64_t*>(nodes + nbNodes);
    uint32_t* candCount = reinterpret_cast<uint32_t*>(candKeys + nbCand);
    uint32_t* nodeCount = candCount + uint64_t(nbQueries) * v.R;
    uint32_t* work = nodeCount + 2ull * nbQueries;
    uint32_t* order = work + nbQueries;
    k_queryPrep<<<nbQueries, 1024, 0, s>>>(v, dLeaf, dOffsets, K, nodes, nodeCount, work, badFlag);
    const bool ordered = nbQueries > 1 && nbQueries <= kOrderCap;
    if (ordered) k_orderQueries<<<1, 1024, 0, s>>>(work, nbQueries, order);
    // ranges fastest, so the tiles of one query run together and share its node list in L2
    for (uint32_t q0 = 0; v.R && q0 < nbQueries; q0 += 65535) {
        const dim3 grid(v.R, std::min(nbQueries - q0, 65535u));
        if (v.walkSteps == 4)
            k_scoreRange<4><<<grid, 1024, 0, s>>>(v, dOffsets + q0, q0, nodes, nodeCount, ordered ? order : nullptr, K,
                                                 candKeys, candCount);
        else
            k_scoreRange<6><<<grid, 1024, 0, s>>>(v, dOffsets + q0, q0, nodes, nodeCount, ordered ? order : nullptr, K,
                                                 candKeys, candCount);
    }
    k_topk<<<nbQueries, 1024, 0, s>>>(v, K, candKeys, candCount, dDocs, dScores, dCounts);
    RBOW_HIP(hipGetLastError());
}

void launchQueryBatch(const DbView& v, const uint32_t* dLeaf, const uint64_t* dOffsets, const uint64_t* hOffsets,
                      uint32_t nbQueries, uint32_t K, uint32_t* dDocs, float* dScores, uint32_t* dCounts,
                      uint32_t* badFlag, hipStream_t s) {
    if (!nbQueries) return;
    const uint64_t total = hOffsets[nbQueries] - hOffsets[0];
    const size_t bytes = queryWorkspaceBytes(v, total, nbQueries, K);
    void* ws = nullptr;
    RBOW_HIP(hipMallocAsync(&ws, bytes, s));
    launchQueryBatchInWs(v, dLeaf, dOffsets, total, nbQueries, K, dDocs, dScores, dCounts, badFlag, ws, bytes, s);
    RBOW_HIP(hipFreeAsync(ws, s));
}

}  // namespace rbow_amd
