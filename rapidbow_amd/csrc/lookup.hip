// lookup.hip -- descriptor -> leaf descent on gfx950.
//
// Semantics (vocTreeLookUp.cu:83-159): node <- 0; for each level l:
// node <- node*b + argmin_j d(x, C[levelBeg(l)+node][j]), ties to the lowest j.
//
// Two forms, identical leaf ids:
//  * plain: one group of b lanes per descriptor walks all levels (lane j owns
//    child j); uint8 trees use the exact integer identity
//    argmin ||x-c||^2 = argmin (||c||^2 - 2 x.c) with v_dot4_u32_u8; float
//    trees use the defining FP32 order acc = fma(x_k - c_k, x_k - c_k, acc).
//  * MFMA: level-synchronous.  Level 0 streams all descriptors against the
//    root; every later level first buckets descriptors by their current node
//    (histogram -> scan -> scatter), so each wave multiplies up to kTQ
//    descriptors of ONE node against that node's b children on the matrix
//    cores.  uint8 trees: I8 MFMA on offset operands (x-128, c-128), exact
//    int32.  Float trees: both operands are centred on the node's integer
//    mean mu (x' = x - mu is exact in BF16), c' = c - mu is split as
//    hi + lo + r (BF16, BF16, exact binary64 residual); split-BF16 MFMA with
//    FP32 accumulation gives an approximate score with a certified error
//    bound; every child whose interval can reach the minimum is a candidate,
//    and when more than one survives the candidates are re-scored in the plain
//    FP32 order, so the argmin equals the plain kernel's.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <utility>

#include "certfilter.hpp"
#include "kernels.hpp"

namespace rbow_amd {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kScoreBias = 1 << 25;   // makes integer scores positive before key packing
constexpr uint32_t kTQ = 256;         // descriptors per MFMA work item (16 row blocks of 16)
constexpr uint32_t kTQLeaf = 512;  // the larger leaf-level item size (lookupLeafItemRowsDefault)

// A work item: up to kTQ descriptors that sit at the same node.  Items of a
// node are consecutive; itemBase/nItems locate the node's items so that the
// per-item child counts can be laid out node-major, child-major, item-minor:
//   C[itemBase*b + j*nItems + (item - itemBase)]
// One exclusive scan over C then gives, for every (item, child j), the first
// position of its descriptors in the next level's bucketed order.
struct LkItem {
    uint32_t node;      // node index within the level
    uint32_t start;     // first position in the bucketed permutation
    uint32_t cnt;       // descriptors (1..kTQ)
    uint32_t itemBase;  // first item of this node
    uint32_t nItems;    // items of this node
    uint32_t pad;
};

// The leaf level's near-tie list: row[i] (the descriptor) and mask[i] (its
// candidate children) for i < *count.
struct LeafTies {
    uint32_t* row;
    uint32_t* mask;
    uint32_t* count;
};

__device__ inline v4i ld16(const uint8_t* p) { return *reinterpret_cast<const v4i*>(p); }
__device__ inline v4i toSigned(v4i v) { return v ^ v4i{int(0x80808080), int(0x80808080), int(0x80808080), int(0x80808080)}; }



// ------------------------------------------------------------ tree prep --
__global__ void k_prepU8(const uint8_t* rows, uint64_t nbChildren, int32_t* normU, int32_t* normS) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nbChildren) return;
    const uint8_t* r = rows + i * kDims;
    int su = 0, ss = 0;
    for (uint32_t k = 0; k < kDims; k++) {
        const int v = r[k];
        su += v * v;
        ss += (v - 128) * (v - 128);
    }
    normU[i] = su;
    normS[i] = ss;
}

// The int8 split of float centroids, one wave per child (lane l: dimensions 2l
// and 2l + 1), for float-tree lookups and, every pass, the float k-means assign.
// With c'' = c - 128 (exact), binary64 (every step exact):
//   A0 = rint(c''), r0 = c'' - A0;  A1 = rint(2^7 r0), r1 = r0 - 2^-7 A1;
//   A2 = rint(2^14 r1), r = r1 - 2^-14 A2;  c'' = Q + r,  Q = A0 + 2^-7 A1 + 2^-14 A2,
//   A0 in [-128, 127] (c in [0, 255]), A1, A2 in [-64, 64], |r| <= 2^-15.
// The MFMA products P_t = x''.A_t (x'' = x - 128) are exact int32, so the score
// s~ = fl(cn2 - 2 t~), t~ = fl(P0 + 2^-14 fl(2^7 P1 + P2)) (combineSplit), errs
// against s = ||c''||^2 - 2 x''.c'' = d(x, c) - ||x''||^2 by at most e1 ||x''|| + e2 with
//   e1 = (2||r|| + 4u (||Q|| + ||r0 - r||)) (1 + 2^-10),  e2 = 2u ||c''||^2 (1 + 2^-10),  u = 2^-24:
//   2 x''.r (Cauchy-Schwarz); the conversion of 2^7 P1 + P2 and the fma,
//   u 2^-14 |x''.(2^7 A1 + A2)| + u |x''.Q| <= u ||x''|| (||r0 - r|| + ||Q||) (as
//   2^7 A1 + A2 = 2^14 (r0 - r)), doubled by the factor -2; the final rounding
//   u |cn2 - 2 t~| and that of cn2 itself, u ||c''||^2 each (the slack covers the
//   (1 + u) factors, the binary64 norms and the roundings of e1 and e2).
// A child with a value outside [0, 255] (or not finite) has no such split: its
// terms are zero and e1 = inf, so the node's largest e1 keeps every child a
// candidate and the plain FP32 chain decides.  `active` (optional, per node):
// nodes whose flag is zero are skipped (converged k-means tasks keep their
// operands).
__global__ __launch_bounds__(256) void k_prepF32Split(const float* __restrict__ rows, uint64_t nbChildren,
                                                     uint32_t lg2b, int8_t* __restrict__ split, uint64_t stride,
                                                     float* __restrict__ cn2, float* __restrict__ e1,
                                                     float* __restrict__ e2, const uint32_t* __restrict__ active) {
    // 16 lanes per child, 8 dimensions per lane: the norms' binary64 sums take four lane
    // swaps within the child's lanes (a wave per child with 2 dimensions per lane took six,
    // for a quarter of the children per wave)
    const uint64_t child = (uint64_t(blockIdx.x) * 256 + threadIdx.x) >> 4;
    const uint32_t lane = threadIdx.x & 63, sub = lane & 15;
    if (child >= nbChildren) return;                  // the child's 16 lanes together
    if (active && !active[child >> lg2b]) return;
    const float4* cr = reinterpret_cast<const float4*>(rows + child * kDims) + 2 * sub;
    const float4 ca = cr[0], cb = cr[1];
    const float c[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
    bool out = false;
#pragma unroll
    for (int e = 0; e < 8; e++) out |= !(c[e] >= 0.0f && c[e] <= 255.0f);
    const bool bad = ((__ballot(out) >> (lane & 48)) & 0xFFFFull) != 0;  // any value of the child
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    uint64_t planes[3] = {0u, 0u, 0u};  // the eight dimensions' bytes of each term
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const double cd = bad ? 0.0 : double(c[e]) - 128.0;
        const double a0 = rint(cd), r0 = cd - a0;
        const double a1 = rint(r0 * 128.0), r1 = r0 - a1 / 128.0;
        const double a2 = rint(r1 * 16384.0), r = r1 - a2 / 16384.0;
        planes[0] |= uint64_t(uint32_t(int(a0)) & 0xFFu) << (8 * e);
        planes[1] |= uint64_t(uint32_t(int(a1)) & 0xFFu) << (8 * e);
        planes[2] |= uint64_t(uint32_t(int(a2)) & 0xFFu) << (8 * e);
        const double qv = cd - r, dr = r0 - r;
        v[0] += cd * cd;
        v[1] += r * r;
        v[2] += qv * qv;
        v[3] += dr * dr;
    }
#pragma unroll
    for (int t = 0; t < 3; t++) reinterpret_cast<uint64_t*>(split + t * stride + child * kDims)[sub] = planes[t];
#pragma unroll
    for (int d = 8; d >= 1; d >>= 1)
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] += __shfl_xor(v[q], d, 16);
    if (sub == 0) {
        const double u = 1.0 / 16777216.0, slack = 1.0 + 1.0 / 1024;
        cn2[child] = bad ? 0.0f : float(v[0]);
        e1[child] = bad ? INFINITY : float((2.0 * sqrt(v[1]) + 4.0 * u * (sqrt(v[2]) + sqrt(v[3]))) * slack);
        e2[child] = bad ? 0.0f : float(2.0 * u * v[0] * slack);
    }
}

// A child whose float row equals that of a lower sibling, bit for bit, has the
// same plain FP32 distance to every descriptor, so with ties going to the lowest
// child it is never the argmin.  Its norm is set to +inf: its filter score is
// +inf, it is never a candidate, and rows that reach a node with repeated
// children (k-means on a node with fewer points than children seeds repeated
// centres, which stay equal) are not near ties.  The lowest of equal children
// keeps its score, so the minimum and the threshold are unchanged.
__global__ __launch_bounds__(256) void k_markDuplicateChildren(const float* __restrict__ rows, uint64_t nbChildren,
                                                              uint32_t bMask, float* __restrict__ cn2) {
    for (uint64_t child = uint64_t(blockIdx.x) * 256 + threadIdx.x; child < nbChildren;
         child += uint64_t(gridDim.x) * 256) {
        const uint32_t j = uint32_t(child) & bMask;
        const uint4* rj = reinterpret_cast<const uint4*>(rows + child * kDims);
        bool dup = false;
        for (uint32_t i = 0; i < j && !dup; i++) {
            const uint4* ri = reinterpret_cast<const uint4*>(rows + (child - j + i) * kDims);
            uint32_t k = 0;
            for (; k < kDims / 4; k++) {
                const uint4 x = ri[k], y = rj[k];
                if (x.x != y.x || x.y != y.y || x.z != y.z || x.w != y.w) break;
            }
            dup = k == kDims / 4;
        }
        if (dup) cn2[child] = INFINITY;
    }
}

// -------------------------------------------------------- plain kernels --
// 16 lanes per descriptor; 16 descriptors per 256-thread block.
__global__ __launch_bounds__(256) void k_lookupPlainU8(const uint8_t* __restrict__ rows,
                                                      const int32_t* __restrict__ normU, uint32_t L,
                                                      const uint8_t* __restrict__ desc, uint32_t n,
                                                      uint32_t* __restrict__ leaf) {
    __shared__ uint32_t sx[16][33];
    const uint32_t lane = threadIdx.x & 15, slot = threadIdx.x >> 4;
    for (uint32_t w = threadIdx.x; w < 512; w += 256) {
        const uint32_t s = w >> 5, k = w & 31;
        const uint64_t q = uint64_t(blockIdx.x) * 16 + s;
        sx[s][k] = q < n ? reinterpret_cast<const uint32_t*>(desc)[q * 32 + k] : 0u;
    }
    __syncthreads();
    uint32_t node = 0;
    for (uint32_t l = 0; l < L; l++) {
        const uint64_t child = (levelBeg(4, l) + node) * 16 + lane;
        const uint4* cr = reinterpret_cast<const uint4*>(rows + child * kDims);
        uint32_t dot = 0;
#pragma unroll
        for (int w4 = 0; w4 < 8; w4++) {
            const uint4 c = cr[w4];
            dot = __builtin_amdgcn_udot4(sx[slot][4 * w4 + 0], c.x, dot, false);
            dot = __builtin_amdgcn_udot4(sx[slot][4 * w4 + 1], c.y, dot, false);
            dot = __builtin_amdgcn_udot4(sx[slot][4 * w4 + 2], c.z, dot, false);
            dot = __builtin_amdgcn_udot4(sx[slot][4 * w4 + 3], c.w, dot, false);
        }
        const int score = normU[child] - 2 * int(dot);
        const uint32_t key = min16u((uint32_t(score + kScoreBias) << 4) | lane);
        node = node * 16 + (key & 15u);
    }
    const uint64_t q = uint64_t(blockIdx.x) * 16 + slot;
    if (lane == 0 && q < n) leaf[q] = node;
}

// B lanes per descriptor (B = 16 or 32); 256/B descriptors per block.
template <int B>
__global__ __launch_bounds__(256) void k_lookupPlainF32(const float* __restrict__ rows, uint32_t lg2b, uint32_t L,
                                                       const uint8_t* __restrict__ desc, uint32_t n,
                                                       uint32_t* __restrict__ leaf) {
    constexpr int DPB = 256 / B;
    __shared__ float sx[DPB][kDims];
    const uint32_t lane = threadIdx.x % B, slot = threadIdx.x / B;
    for (uint32_t e = threadIdx.x; e < DPB * kDims; e += 256) {
        const uint32_t s = e / kDims, k = e % kDims;
        const uint64_t q = uint64_t(blockIdx.x) * DPB + s;
        sx[s][k] = q < n ? float(desc[q * kDims + k]) : 0.0f;
    }
    __syncthreads();
    uint32_t node = 0;
    for (uint32_t l = 0; l < L; l++) {
        const uint64_t child = (levelBeg(lg2b, l) + node) * B + lane;
        const float4* cr = reinterpret_cast<const float4*>(rows + child * kDims);
        float acc = 0.0f;
#pragma unroll 8
        for (int k4 = 0; k4 < 32; k4++) {
            const float4 c = cr[k4];
            float d = sx[slot][4 * k4 + 0] - c.x;
            acc = fmaf(d, d, acc);
            d = sx[slot][4 * k4 + 1] - c.y;
            acc = fmaf(d, d, acc);
            d = sx[slot][4 * k4 + 2] - c.z;
            acc = fmaf(d, d, acc);
            d = sx[slot][4 * k4 + 3] - c.w;
            acc = fmaf(d, d, acc);
        }
        float bd = acc;
        uint32_t bj = lane;
#pragma unroll
        for (int m = 1; m < B; m <<= 1) {
            const float od = __shfl_xor(bd, m, B);
            const uint32_t oj = uint32_t(__shfl_xor(int(bj), m, B));
            if (od < bd || (od == bd && oj < bj)) {
                bd = od;
                bj = oj;
            }
        }
        node = node * B + bj;
    }
    const uint64_t q = uint64_t(blockIdx.x) * DPB + slot;
    if (lane == 0 && q < n) leaf[q] = node;
}

// ---------------------------------------------------- bucketing kernels --
__device__ inline void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

__global__ void k_fillU32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Resolve an item.  kIdentity = level 0: the descriptors in input order,
// every item at the root.
template <bool kIdentity>
__device__ inline LkItem itemAt(uint32_t it, uint32_t n, const LkItem* items) {
    if (kIdentity) {
        const uint32_t start = it * kTQ;
        return LkItem{0u, start, min(kTQ, n - start), 0u, (n + kTQ - 1) / kTQ, 0u};
    }
    return items[it];
}

// Epilogue of a (non-last) level step for one item: sj[r] holds the chosen
// child of row r.  Writes code[start+r] = child | rank << 8 (rank = earlier
// rows of the item with the same child) and the item's child counts into C.
__device__ inline void itemCounts(uint32_t lane, uint32_t lg2b, uint32_t it, const LkItem& t, const uint8_t* sj,
                                  uint32_t* sc, uint32_t* __restrict__ code, uint32_t* __restrict__ C) {
    const uint32_t b = 1u << lg2b;
    if (lane < b) sc[lane] = 0;
    waveSync();
    for (uint32_t g = 0; g * 64 < t.cnt; g++) {
        const uint32_t r = g * 64 + lane;
        const bool valid = r < t.cnt;
        const uint32_t k = valid ? sj[r] : 0u;
        uint64_t m = __ballot(valid);
        for (uint32_t bit = 0; bit < lg2b; bit++) {
            const uint64_t bb = __ballot(valid && ((k >> bit) & 1u));
            m &= ((k >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t below = mbcnt(m);
        const uint32_t base = valid ? sc[k] : 0u;
        waveSync();
        if (valid && below == 0) sc[k] = base + uint32_t(__builtin_popcountll(m));
        if (valid) code[t.start + r] = k | ((base + below) << 8);
        waveSync();
    }
    if (lane < b) C[uint64_t(t.itemBase) * b + lane * t.nItems + (it - t.itemBase)] = sc[lane];
    waveSync();
}

// The scan of a level's counts runs over itemsMax items (the bound known on the
// host): the entries past the last item, and the tail entry, are zeroed here by
// the level kernel's threads once its items are done.
__device__ inline void zeroCountTail(uint32_t* __restrict__ C, uint32_t nbItems, uint32_t itemsMax, uint32_t b) {
    const uint64_t end = uint64_t(itemsMax) * b + 1, step = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(nbItems) * b + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < end; i += step)
        C[i] = 0;
}

// Items of the next level.  Child v = u*b + j of parent u occupies
// O[base + j*nI] .. O[base + (j+1)*nI) of the scanned counts (base, nI: the
// parent's items; level 0 has the single parent 0 with nItems0 items).
__device__ inline void childRange(const uint32_t* O, const uint32_t* pBase, const uint32_t* pN, uint32_t nItems0,
                                  uint32_t lg2b, uint32_t v, uint32_t& s0, uint32_t& c0) {
    const uint32_t b = 1u << lg2b, u = v >> lg2b, j = v & (b - 1);
    const uint32_t base = pBase ? pBase[u] : 0u, ni = pBase ? pN[u] : nItems0;
    s0 = O[uint64_t(base) * b + j * ni];
    c0 = O[uint64_t(base) * b + (j + 1) * ni] - s0;
}

// The next level's children and their item size, for the item counts (the
// first step of building the next level's items, done by k_scatter's threads):
// cN[v] = items of child v for v < nb, cN[nb] = 0 (the scan's tail).
struct NextItems {
    const uint32_t* pBase;  // the parents' first items and item counts (null at level 0)
    const uint32_t* pN;
    uint32_t nItems0;  // level 0: the root's items
    uint32_t nb;       // children
    uint32_t itemRows;
    uint32_t* cN;
};

// Moves every descriptor of an item to its place in the next level's order.  The
// wave first orders the item's rows by child in LDS (child c's rows start at the
// item's count of children below c), so each store instruction writes runs of
// consecutive positions instead of one or two entries in each of b runs.  The
// destination of (item, child) is O at its flat index, its count the next entry
// of O minus that one (O is the exclusive scan of the counts in the same order).
template <bool kIdentity>
__global__ __launch_bounds__(256) void k_scatter(const LkItem* __restrict__ items,
                                                const uint32_t* __restrict__ nbItemsPtr, uint32_t nbItemsImplicit,
                                                uint32_t n, uint32_t lg2b, const uint32_t* __restrict__ perm,
                                                const uint32_t* __restrict__ code, const uint32_t* __restrict__ O,
                                                uint32_t* __restrict__ permNext, NextItems nx) {
    __shared__ uint32_t sSrcAll[4][kTQ];
    __shared__ uint8_t sChildAll[4][kTQ];
    __shared__ uint32_t sOffAll[4][32], sDstAll[4][32];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, b = 1u << lg2b;
    uint32_t* sSrc = sSrcAll[wave];
    uint8_t* sChild = sChildAll[wave];
    uint32_t *sOff = sOffAll[wave], *sDst = sDstAll[wave];
    const uint32_t nbItems = kIdentity ? nbItemsImplicit : *nbItemsPtr;
    for (uint32_t it = blockIdx.x * 4 + wave; it < nbItems; it += gridDim.x * 4) {
        const LkItem t = itemAt<kIdentity>(it, n, items);
        static_assert(kTQ == 256, "an item is four groups of 64 rows");
        // the item's rows, four per lane, every load in flight together
        uint32_t c[4], src[4];
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const uint32_t r = g * 64 + lane, p = t.start + r;
            c[g] = r < t.cnt ? code[p] : 0u;
            src[g] = r < t.cnt ? (kIdentity ? p : perm[p]) : 0u;
        }
        // lane j < b: child j's count, its offset among the item's rows and its destination
        const uint64_t ci = uint64_t(t.itemBase) * b + uint64_t(lane < b ? lane : 0u) * t.nItems + (it - t.itemBase);
        const uint32_t d0 = O[ci], cnt = lane < b ? O[ci + 1] - d0 : 0u;
        uint32_t off = cnt;  // inclusive prefix over lanes 0..b-1, then exclusive
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const uint32_t o = uint32_t(__shfl_up(int(off), d, 64));
            if (lane >= uint32_t(d)) off += o;
        }
        if (lane < b) {
            sOff[lane] = off - cnt;
            sDst[lane] = d0;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int g = 0; g < 4; g++)
            if (g * 64 + lane < t.cnt) {
                const uint32_t ch = c[g] & 0xFFu, lp = sOff[ch] + (c[g] >> 8);
                sSrc[lp] = src[g];
                sChild[lp] = uint8_t(ch);
            }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const uint32_t p = g * 64 + lane;
            if (p < t.cnt) {
                const uint32_t ch = sChild[p];
                permNext[sDst[ch] + p - sOff[ch]] = sSrc[p];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // reads done before the next item's writes
        __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t v = blockIdx.x * 256 + threadIdx.x; v <= nx.nb; v += gridDim.x * 256) {
        uint32_t cnt = 0;  // v == nb: the scan's tail
        if (v < nx.nb) {
            uint32_t s0, c0;
            childRange(O, nx.pBase, nx.pN, nx.nItems0, lg2b, v, s0, c0);
            cnt = (c0 + nx.itemRows - 1) / nx.itemRows;
        }
        nx.cN[v] = cnt;
    }
}

// Exclusive scan of n <= kSmallScanMax values by one 1024-thread block: thread t
// sums its kSmallScanPer consecutive values (all loads in flight together), the
// block scans the sums, and each thread writes its values' prefixes.  One launch
// where a device-wide scan takes two (the per-child item counts of a level).
constexpr uint32_t kSmallScanPer = 33, kSmallScanMax = kSmallScanPer * 1024;
__global__ __launch_bounds__(1024) void k_scanSmall(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint32_t n) {
    __shared__ uint32_t sWave[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, i0 = tid * kSmallScanPer;
    uint32_t x[kSmallScanPer], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSmallScanPer; k++) x[k] = i0 + k < n ? in[i0 + k] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kSmallScanPer; k++) sum += x[k];
    uint32_t inc = sum;  // inclusive prefix of the sums over the wave, then the block
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = uint32_t(__shfl_up(int(inc), d, 64));
        if (lane >= uint32_t(d)) inc += o;
    }
    if (lane == 63) sWave[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; w++) base += sWave[w];
    uint32_t run = base + inc - sum;
#pragma unroll
    for (uint32_t k = 0; k < kSmallScanPer; k++) {
        if (i0 + k < n) out[i0 + k] = run;
        run += x[k];
    }
}

__global__ void k_writeItems(const uint32_t* __restrict__ O, const uint32_t* __restrict__ pBase,
                             const uint32_t* __restrict__ pN, uint32_t nItems0, uint32_t nb, uint32_t lg2b,
                             uint32_t itemRows, const uint32_t* __restrict__ cBase, LkItem* __restrict__ items,
                             uint32_t* __restrict__ nbItems, uint32_t* __restrict__ tieCount) {
    // one node per block: its items are written by the block's threads (a node
    // near the root holds thousands of items)
    for (uint32_t v = blockIdx.x; v <= nb; v += gridDim.x) {
        if (v == nb) {
            if (threadIdx.x == 0) {
                *nbItems = cBase[nb];
                *tieCount = 0;  // the leaf level's near-tie list starts empty
            }
            continue;
        }
        uint32_t s0, c0;
        childRange(O, pBase, pN, nItems0, lg2b, v, s0, c0);
        const uint32_t base = cBase[v], ni = (c0 + itemRows - 1) / itemRows;
        for (uint32_t t = threadIdx.x; t < ni; t += blockDim.x)
            items[base + t] = LkItem{v, s0 + t * itemRows, min(itemRows, c0 - t * itemRows), base, ni, 0u};
    }
}

// --------------------------------------------------- I8 MFMA level step --
// mfma_i32_16x16x64_i8: lane l feeds A row (l&15) / B column (l&15) with the
// 16 bytes of dims [16*(l>>4) + 64*kb, +16) (same placement for A and B, so
// the sum runs over all 128 dims); the result lane l holds rows 4*(l>>4)+i,
// column l&15.  The node's 16 children are A and 16 descriptor rows are B, so
// lane l holds children 4*(l>>4)+i of row (l&15): each lane owns one row, whose
// argmin is an in-lane minimum over four keys and two permlane swaps.
template <bool kIdentity, bool kLast>
__global__ __launch_bounds__(256) void k_levelI8(const uint8_t* __restrict__ rows, const int32_t* __restrict__ normS,
                                                uint64_t levelBase, const uint8_t* __restrict__ desc, uint32_t n,
                                                const uint32_t* __restrict__ perm, const LkItem* __restrict__ items,
                                                const uint32_t* __restrict__ nbItemsPtr, uint32_t nbItemsImplicit,
                                                uint32_t* __restrict__ out, uint32_t* __restrict__ C) {
    __shared__ uint8_t sjAll[4][kTQ];
    __shared__ uint32_t scAll[4][32];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    uint8_t* sj = sjAll[wave];
    const uint32_t nbItems = kIdentity ? nbItemsImplicit : *nbItemsPtr;
    for (uint32_t it = blockIdx.x * 4 + wave; it < nbItems; it += gridDim.x * 4) {
        const LkItem t = itemAt<kIdentity>(it, n, items);
        const uint64_t childBase = (levelBase + t.node) * 16;
        const uint8_t* crow = rows + (childBase + col) * kDims;
        const v4i b0 = toSigned(ld16(crow + grp * 16)), b1 = toSigned(ld16(crow + 64 + grp * 16));
        const v4i ns = *reinterpret_cast<const v4i*>(normS + childBase + 4 * grp);  // ||c - 128||^2 of children 4*grp + i
        // the lane's row of a block: its descriptor (input index) and bytes; the next
        // block's bytes and the permutation entry of the one after are in flight
        // A row past the item's end stands for its last row (a valid descriptor): its column of
        // the products is discarded and nothing is written for it.  With every index valid, the
        // loads below are unconditional: a load under a branch ends in a copy at the join that
        // waits for it, which would undo the prefetch.
        auto qAt = [&](uint32_t rb) {
            const uint32_t r = min(rb * 16 + col, t.cnt - 1u);
            return kIdentity ? t.start + r : perm[t.start + r];
        };
        // Raw bytes: the sign flip happens at use, so the loads stay in flight while the
        // current block computes.
        auto loadRows = [&](uint32_t q, v4i& x0, v4i& x1) {
            const uint8_t* x = desc + uint64_t(q) * kDims;
            x0 = ld16(x + grp * 16);
            x1 = ld16(x + 64 + grp * 16);
        };
        uint32_t qc = qAt(0), qn = qAt(1);
        v4i n0, n1;
        loadRows(qc, n0, n1);
        for (uint32_t rb = 0; rb * 16 < t.cnt; rb++) {
            const uint32_t r = rb * 16 + col, q = qc;
            const v4i a0 = toSigned(n0), a1 = toSigned(n1);
            qc = qn;
            loadRows(qc, n0, n1);  // block rb + 1 (past the end: the item's last row again)
            qn = qAt(rb + 2);
            v4i acc = {0, 0, 0, 0};
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(b0, a0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(b1, a1, acc, 0, 0, 0);
            // keys ((||c'||^2 - 2 x'.c' + bias) << 4) | child: unique, the smallest is the
            // nearest child with ties to the lowest index
            uint32_t key = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < 4; i++)
                key = min(key, (uint32_t(ns[i] - 2 * acc[i] + kScoreBias) << 4) | (4 * grp + uint32_t(i)));
            key = minAcrossGroupsU(key);
            if (grp == 0 && r < t.cnt) {
                if (kLast) out[q] = t.node * 16 + (key & 15u);
                else sj[r] = uint8_t(key & 15u);
            }
        }
        if (!kLast) itemCounts(lane, 4, it, t, sj, scAll[wave], out, C);
    }
    if (!kLast) zeroCountTail(C, nbItems, nbItemsImplicit, 16);
}

// ------------------------------------------- fused I8 descent (b = 16) --
template <int... R, class F>
__device__ inline void unrollSeq(F&& f, std::integer_sequence<int, R...>) {
    (f(std::integral_constant<int, R>{}), ...);
}
template <class F>
__device__ inline void unroll16(F&& f) {
    unrollSeq(f, std::make_integer_sequence<int, 16>{});
}

// Every level of a uint8 b16 tree in one launch, for batches too small to
// bucket.  A wave takes 16 descriptors at a time.  Level 0: the root's 16 children as the MFMA A operand against the
// 16 rows as B (as k_levelI8).  Every later level, row by row: the row's node's
// 16 children as A and the row broadcast to all 16 columns as B (DPP, from the
// registers that hold it), so every lane
// holds four children's scores of that row; the same in-lane minimum and
// permlane swaps give the row's child (ties to the lowest index).  The child
// rows of levels 0 and 1 (17 nodes, 38 KB at the 144-byte pitch) are staged in
// LDS once per workgroup; deeper levels read them from L2, the next row's
// operands in flight while one computes.  Same exact integer keys as the plain
// and bucketed kernels, hence the same leaf ids.
constexpr uint32_t kFusedPitch = 144;   // LDS bytes per 128-byte row: conflict-free 16-byte reads
constexpr uint32_t kFusedNodes = 17;    // nodes of levels 0 and 1 staged in LDS
static_assert(kFusedNodes * 16 <= 512, "the staging loads at most two child norms per thread");

__device__ inline uint32_t rowKey(const v4i& acc, const v4i& ns, uint32_t grp) {
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < 4; i++)
        key = min(key, (uint32_t(ns[i] - 2 * acc[i] + kScoreBias) << 4) | (4 * grp + uint32_t(i)));
    return minAcrossGroupsU(key);
}

// Lane r of each 16-lane row of the wave, to every lane of that row (DPP row_newbcast).
template <int R>
__device__ inline v4i rowBcast(const v4i& x) {
    v4i y;
#pragma unroll
    for (int i = 0; i < 4; i++) y[i] = __builtin_amdgcn_update_dpp(0, x[i], 0x150 + R, 0xF, 0xF, false);
    return y;
}

// One per-row level of the fused descent: row r (0..15) of the wave's block sits
// at node (lane r's value) of this level; returns, for this lane's row col, the
// index within the next level.  Row r's B operand (its bytes at dims 16*grp and
// 64 + 16*grp, held by lane r of each 16-lane group) is broadcast from registers.
// The A operands of D rows are in flight at once (a register ring; the loop is
// unrolled, so every index is a constant).
template <bool kLds, int D>
__device__ inline uint32_t fusedLevel(const uint8_t* sRows, const int32_t* sNorm, const uint8_t* __restrict__ rows,
                                      const int32_t* __restrict__ normS, uint64_t lb, const v4i& x0, const v4i& x1,
                                      uint32_t node, uint32_t col, uint32_t grp) {
    v4i a0[D], a1[D], ns[D];
    auto loadA = [&](uint32_t v, v4i& x0, v4i& x1, v4i& n4) {
        if (kLds) {  // level 1: node v is staged node 1 + v
            const uint8_t* c = sRows + ((1 + v) * 16 + col) * kFusedPitch;
            x0 = *reinterpret_cast<const v4i*>(c + grp * 16);
            x1 = *reinterpret_cast<const v4i*>(c + 64 + grp * 16);
            n4 = *reinterpret_cast<const v4i*>(sNorm + (1 + v) * 16 + 4 * grp);
        } else {
            const uint64_t child = (lb + v) * 16;
            const uint8_t* c = rows + (child + col) * kDims;
            x0 = ld16(c + grp * 16);  // raw bytes, flipped at use: the ring's loads stay in flight
            x1 = ld16(c + 64 + grp * 16);
            n4 = *reinterpret_cast<const v4i*>(normS + child + 4 * grp);
        }
    };
#pragma unroll
    for (int r = 0; r < D; r++) loadA(uint32_t(__builtin_amdgcn_readlane(int(node), r)), a0[r], a1[r], ns[r]);
    uint32_t next = 0;
    auto step = [&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const uint32_t v = uint32_t(__builtin_amdgcn_readlane(int(node), r));  // row r's node
        const v4i b0 = rowBcast<r>(x0), b1 = rowBcast<r>(x1);                   // row r in every column
        v4i d = {0, 0, 0, 0};
        d = __builtin_amdgcn_mfma_i32_16x16x64_i8(kLds ? a0[r % D] : toSigned(a0[r % D]), b0, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_i32_16x16x64_i8(kLds ? a1[r % D] : toSigned(a1[r % D]), b1, d, 0, 0, 0);
        const uint32_t key = rowKey(d, ns[r % D], grp);
        if (r + D < 16) loadA(uint32_t(__builtin_amdgcn_readlane(int(node), r + D)), a0[r % D], a1[r % D], ns[r % D]);
        if (col == uint32_t(r)) next = v * 16 + (key & 15u);
    };
    unroll16(step);
    return next;
}

__global__ __launch_bounds__(256) void k_lookupFusedI8(const uint8_t* __restrict__ rows,
                                                      const int32_t* __restrict__ normS, uint32_t L,
                                                      const uint8_t* __restrict__ desc, uint32_t n,
                                                      uint32_t* __restrict__ leaf) {
    __shared__ __align__(16) uint8_t sRows[kFusedNodes * 16 * kFusedPitch];  // children (c - 128), node-major
    __shared__ __align__(16) int32_t sNorm[kFusedNodes * 16];                // ||c - 128||^2
    // stage the children of levels 0-1 (16-byte chunks: child i / 8, chunk i % 8), every load in flight at once
    const uint32_t staged = L >= 2 ? kFusedNodes : 1u;
    constexpr int kIt = (kFusedNodes * 16 * 8 + 255) / 256;
    v4i st[kIt];
#pragma unroll
    for (int j = 0; j < kIt; j++) {
        const uint32_t i = threadIdx.x + 256 * j;
        if (i < staged * 128) st[j] = ld16(rows + uint64_t(i >> 3) * kDims + 16 * (i & 7));
    }
    const int32_t nrm0 = threadIdx.x < staged * 16 ? normS[threadIdx.x] : 0;  // 272 norms: two per thread
    const int32_t nrm1 = threadIdx.x + 256 < staged * 16 ? normS[threadIdx.x + 256] : 0;
#pragma unroll
    for (int j = 0; j < kIt; j++) {
        const uint32_t i = threadIdx.x + 256 * j;
        if (i < staged * 128) *reinterpret_cast<v4i*>(sRows + (i >> 3) * kFusedPitch + 16 * (i & 7)) = toSigned(st[j]);
    }
    if (threadIdx.x < staged * 16) sNorm[threadIdx.x] = nrm0;
    if (threadIdx.x + 256 < staged * 16) sNorm[threadIdx.x + 256] = nrm1;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    for (uint32_t blk = blockIdx.x * 4 + wave; blk * 16 < n; blk += gridDim.x * 4) {
        const uint32_t row = blk * 16 + col;  // this lane's row of the block
        v4i x0 = {0, 0, 0, 0}, x1 = {0, 0, 0, 0};  // rows past n score garbage and are not written
        if (row < n) {
            x0 = toSigned(ld16(desc + uint64_t(row) * kDims + grp * 16));
            x1 = toSigned(ld16(desc + uint64_t(row) * kDims + 64 + grp * 16));
        }
        // level 0: the root's children (A) against the 16 rows (B)
        v4i acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(*reinterpret_cast<const v4i*>(sRows + col * kFusedPitch + grp * 16),
                                                    x0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(
            *reinterpret_cast<const v4i*>(sRows + col * kFusedPitch + 64 + grp * 16), x1, acc, 0, 0, 0);
        uint32_t node = rowKey(acc, *reinterpret_cast<const v4i*>(sNorm + 4 * grp), grp) & 15u;  // in level 1
        for (uint32_t l = 1; l < L; l++) {
            if (l == 1) node = fusedLevel<true, 2>(sRows, sNorm, rows, normS, 0, x0, x1, node, col, grp);
            else node = fusedLevel<false, 4>(sRows, sNorm, rows, normS, levelBeg(4, l), x0, x1, node, col, grp);
        }
        if (grp == 0 && row < n) leaf[row] = node;
    }
}

// ------------------------------------------------ int8-split level step --
// mfma_i32_16x16x64_i8 as in k_levelI8: the node's children are A (16 per block
// cb, one int8 term t at a time) and 16 descriptor rows are B (x - 128, the
// bytes with their top bit flipped), so the accumulator of lane l holds the
// exact products P_t of children cb*16 + 4*(l>>4) + i with row (l&15).  The
// terms combine into the score's dot part (k_prepF32Split): P1 * 2^7 + P2 is
// exact in int32, then t = fl(P0 + 2^-14 fl(2^7 P1 + P2)), whose two roundings
// are within the bound k_prepF32Split derives for the three-step form.  The
// certified filter keeps every child that can be the argmin; near ties are
// queued per wave and resolved with the plain FP32 chain at the item's end.
//
template <bool kIdentity, bool kLast, int B, uint32_t NQ = kTQ>
__global__ __launch_bounds__(256, 2) void k_levelSplit(  // two waves per SIMD: at most 256 registers
    const float* __restrict__ rows32, const int8_t* __restrict__ split, uint64_t stride,
    const float* __restrict__ cn2, const float* __restrict__ e1, const float* __restrict__ e2, uint64_t levelBase,
    const uint8_t* __restrict__ desc, uint32_t n, const uint32_t* __restrict__ perm, const LkItem* __restrict__ items,
    const uint32_t* __restrict__ nbItemsPtr, uint32_t nbItemsImplicit, uint32_t* __restrict__ out,
    uint32_t* __restrict__ C, uint32_t* __restrict__ nearTies, LeafTies leafTies) {
    constexpr int CB = B / 16;
    constexpr uint32_t LG2B = B == 32 ? 5 : 4;
    constexpr int NG = int(NQ / 64);  // an item: NG groups of four 16-row blocks
    constexpr bool kTieList = kLast && !kIdentity;  // the leaf level of a tree with L >= 2 lists its near ties
    static_assert(NQ % 64 == 0 && (kLast || NQ == kTQ) && (!kIdentity || NQ == kTQ), "item size");
    __shared__ uint8_t sjAll[4][NQ];
    __shared__ uint32_t scAll[4][32];
    __shared__ TieQueueT<NQ> tqAll[4];  // a whole item's rows: resolved once, at the item's end
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    uint8_t* sj = sjAll[wave];
    TieQueueT<NQ>& tq = tqAll[wave];
    const uint32_t nbItems = kIdentity ? nbItemsImplicit : *nbItemsPtr;
    uint32_t ties = 0, nTie = 0;
    for (uint32_t it = blockIdx.x * 4 + wave; it < nbItems; it += gridDim.x * 4) {
        const LkItem t = itemAt<kIdentity>(it, n, items);
        const uint64_t node = levelBase + t.node, childBase = node * B;
        // the item's descriptors: lane l holds those of rows g*64 + l, g < NG (group g)
        // (a row past the item's end stands for its last row, a valid descriptor: its column
        // of the products is discarded and nothing is written for it)
        uint32_t qv[NG];
#pragma unroll
        for (int g = 0; g < NG; g++) {
            const uint32_t r = min(g * 64 + lane, t.cnt - 1u);
            qv[g] = kIdentity ? t.start + r : perm[t.start + r];
        }
        v4i ca[3][CB][2];
        float cn[CB][4];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int8_t* crow = split + (childBase + cb * 16 + col) * kDims + grp * 16;
#pragma unroll
            for (int tt = 0; tt < 3; tt++) {
                ca[tt][cb][0] = *reinterpret_cast<const v4i*>(crow + tt * stride);
                ca[tt][cb][1] = *reinterpret_cast<const v4i*>(crow + tt * stride + 64);
            }
            const float4 c4 = *reinterpret_cast<const float4*>(cn2 + childBase + cb * 16 + 4 * grp);
            cn[cb][0] = c4.x;
            cn[cb][1] = c4.y;
            cn[cb][2] = c4.z;
            cn[cb][3] = c4.w;
        }
        // the node's largest bound coefficients (lane l reads child l % B)
        float e1m = e1[childBase + lane % B], e2m = e2[childBase + lane % B];
#pragma unroll
        for (int d = B / 2; d >= 1; d >>= 1) {
            e1m = fmaxf(e1m, __shfl_xor(e1m, d, 64));
            e2m = fmaxf(e2m, __shfl_xor(e2m, d, 64));
        }
        // Block rb = 4 g + S of group g: its rows' bytes sit in slot S, loaded four blocks
        // ahead (the slot's previous block), so a block computes while the next four load.
        v4i rx0[4], rx1[4];
        auto rowOf = [&](uint32_t qg, int S) {  // descriptor of this lane's row of block S of a group
            return uint32_t(__shfl(int(qg), S * 16 + int(col), 64));
        };
        // Raw bytes: the sign flip happens at use, so a slot's loads stay in flight for four
        // blocks.  The loads are unconditional (every qv is a valid descriptor): a load under a
        // branch ends in a copy at the join that waits for it.
        auto loadRows = [&](uint32_t q, v4i& x0, v4i& x1) {
            const uint8_t* x = desc + uint64_t(q) * kDims;
            x0 = ld16(x + grp * 16);
            x1 = ld16(x + 64 + grp * 16);
        };
        unrollSeq([&](auto S) { loadRows(rowOf(qv[0], S), rx0[S], rx1[S]); }, std::make_integer_sequence<int, 4>{});
        for (uint32_t g = 0; g * 64 < t.cnt; g++) {
            unrollSeq(
                [&](auto S) {
                    const uint32_t rb = g * 4 + S;
                    const v4i x0 = toSigned(rx0[S]), x1 = toSigned(rx1[S]);
                    loadRows(rowOf(qv[1], S), rx0[S], rx1[S]);  // block rb + 4 (or a valid row past the end)
                    if (rb * 16 >= t.cnt) return;
                    const uint32_t r = rb * 16 + col;
                    const bool valid = r < t.cnt;
                    // ||x - 128||^2 of row col: the diagonal of the rows' Gram matrix (exact int32) on the
                    // matrix cores, which have slack here.  The rows are A and B alike, so lane l holds
                    // G[4 (l >> 4) + i][l & 15]: G[r][r] is component r & 3 of lane 16 (r >> 2) + r.
                    v4i gm = __builtin_amdgcn_mfma_i32_16x16x64_i8(x0, x0, v4i{0, 0, 0, 0}, 0, 0, 0);
                    gm = __builtin_amdgcn_mfma_i32_16x16x64_i8(x1, x1, gm, 0, 0, 0);
                    const int g01 = (col & 1) ? gm[1] : gm[0], g23 = (col & 1) ? gm[3] : gm[2];
                    const int gOwn = (col & 2) ? g23 : g01;
                    const int xs = __shfl(gOwn, int(16 * (col >> 2) + col), 64);
                    v4f acc[CB];
#pragma unroll
                    for (int cb = 0; cb < CB; cb++) {
                        v4i p[3];
#pragma unroll
                        for (int tt = 0; tt < 3; tt++) {
                            p[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ca[tt][cb][0], x0, v4i{0, 0, 0, 0}, 0, 0, 0);
                            p[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ca[tt][cb][1], x1, p[tt], 0, 0, 0);
                        }
#pragma unroll
                        for (int i = 0; i < 4; i++) acc[cb][i] = combineSplit(p[0][i], p[1][i], p[2][i]);
                    }
                    const uint32_t mask = filterSplit<CB>(acc, float(xs), cn, e1m, e2m, lane, valid);
                    const bool tie = (mask & (mask - 1u)) != 0;  // only valid rows can have several candidates
                    if (kLast) {  // a listed near tie's lowest candidate stands until k_leafTies replaces it
                        const uint32_t q = rowOf(qv[0], S);
                        if (grp == 0 && valid && (kTieList || !tie)) out[q] = t.node * B + __builtin_ctz(mask);
                        if (__ballot(tie)) ties += tiePush(tq, nTie, lane, tie, q, q, mask);
                    } else {
                        if (grp == 0 && valid && !tie) sj[r] = uint8_t(__builtin_ctz(mask));
                        if (__ballot(tie))  // rare: most blocks skip the queue's bookkeeping (and the row's index)
                            ties += tiePush(tq, nTie, lane, tie, rowOf(qv[0], S), r, mask);
                    }
                },
                std::make_integer_sequence<int, 4>{});
#pragma unroll
            for (int g2 = 0; g2 + 1 < NG; g2++) qv[g2] = qv[g2 + 1];
        }
        if (kTieList && nTie) {
            // leaf level: the item's near ties go to the level's list (one atomic per item with
            // ties), resolved by k_leafTies after this launch, so the wave moves on at once
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(leafTies.count, nTie);
            base = uint32_t(__builtin_amdgcn_readfirstlane(int(__shfl(int(base), 0, 64))));
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the queue writes of this wave are visible
            __builtin_amdgcn_wave_barrier();
            for (uint32_t i = lane; i < nTie; i += 64) {
                leafTies.row[base + i] = tq.tag[i];
                leafTies.mask[base + i] = tq.mask[i];
            }
            nTie = 0;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // reads done before the queue is reused
            __builtin_amdgcn_wave_barrier();
        } else if (nTie) {  // near ties: the plain chain over the queued rows' candidates
            resolveTies(
                tq, nTie, lane, [&](uint32_t q) { return desc + uint64_t(q) * kDims; },
                [&](uint32_t c) { return rows32 + (childBase + c) * kDims; },
                [&](uint32_t aux, uint32_t best) {
                    if (kLast) out[aux] = t.node * B + best;  // aux: the descriptor
                    else sj[aux] = uint8_t(best);             // aux: the row in the item
                });
        }
        if (!kLast) itemCounts(lane, LG2B, it, t, sj, scAll[wave], out, C);
    }
    if (!kLast) zeroCountTail(C, nbItems, nbItemsImplicit, B);
    if (nearTies && ties && lane == 0) atomicAdd(nearTies, ties);  // ties: rows of this wave (uniform)
}

// The leaf level's near ties (row, candidate mask), listed by k_levelSplit<.., kLast>.
// One lane per (row, candidate) pair: a wave takes 16 rows, spreads their
// candidates over its lanes (64 per round) and computes each pair's plain FP32
// chain; the row's lane then takes the lexicographic (distance, child) minimum
// of its candidates.  The leaf kernel wrote node * B + (lowest candidate) for
// the row, which gives the node.
constexpr uint32_t kTieRows = 16;
template <int B>
__global__ __launch_bounds__(256) void k_leafTies(const float* __restrict__ rows32, uint64_t levelBase,
                                                 const uint8_t* __restrict__ desc, LeafTies leafTies,
                                                 uint32_t* __restrict__ out) {
    __shared__ float sD[4][kTieRows * B];  // a wave's pair distances, by pair index
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t n = *leafTies.count, waves = gridDim.x * 4;
    float* dist = sD[wave];
    for (uint32_t e0 = (blockIdx.x * 4 + wave) * kTieRows; e0 < n; e0 += waves * kTieRows) {
        // lane j < 16: row e0 + j, its candidates and their offset among the chunk's pairs
        const bool live = lane < kTieRows && e0 + lane < n;
        const uint32_t q = live ? leafTies.row[e0 + lane] : 0u, mask = live ? leafTies.mask[e0 + lane] : 0u;
        const uint32_t node = live ? out[q] / B : 0u, nc = __builtin_popcount(mask);
        uint32_t off = nc;  // inclusive prefix over lanes 0..15, then exclusive
#pragma unroll
        for (int d = 1; d < int(kTieRows); d <<= 1) {
            const uint32_t o = uint32_t(__shfl_up(int(off), d, 64));
            if (lane >= uint32_t(d)) off += o;
        }
        const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(off), int(kTieRows) - 1));
        off -= nc;
        for (uint32_t p0 = 0; p0 < total; p0 += 64) {
            const uint32_t p = p0 + lane;
            // the pair's row: the last of the 16 whose offset is <= p
            uint32_t j = 0;
#pragma unroll
            for (int k = 1; k < int(kTieRows); k++)
                if (uint32_t(__builtin_amdgcn_readlane(int(off), k)) <= p) j = uint32_t(k);
            // every lane takes part in the shuffles: one from a lane that is off in this round
            // (p >= total) would read a stale value
            const uint32_t mj = uint32_t(__shfl(int(mask), int(j), 64)), qj = uint32_t(__shfl(int(q), int(j), 64));
            const uint32_t nj = uint32_t(__shfl(int(node), int(j), 64)), oj = uint32_t(__shfl(int(off), int(j), 64));
            if (p < total) {
                uint32_t mm = mj;  // the (p - oj)-th candidate of row j (p - oj < its candidate count <= B)
                for (uint32_t r = min(p - oj, uint32_t(B) - 1u); r; r--) mm &= mm - 1u;
                const uint32_t c = __builtin_ctz(mm | (1u << (B - 1)));  // masks hold bits below B only
                dist[p] = plainF32Distance(desc + uint64_t(qj) * kDims, rows32 + ((levelBase + nj) * B + c) * kDims);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the distances of this wave are visible
        __builtin_amdgcn_wave_barrier();
        if (live) {  // candidates in ascending child order: the first minimum is the lowest child
            float bd = INFINITY;
            uint32_t bj = __builtin_ctz(mask), mm = mask;
            for (uint32_t k = 0; k < nc; k++, mm &= mm - 1u) {
                const float a = dist[off + k];
                if (a < bd) {
                    bd = a;
                    bj = __builtin_ctz(mm);
                }
            }
            out[q] = node * B + bj;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // reads done before the next chunk's writes
        __builtin_amdgcn_wave_barrier();
    }
}

uint32_t gridFor(uint64_t work, uint32_t perBlock, uint32_t cap) {
    const uint64_t g = (work + perBlock - 1) / perBlock;
    return uint32_t(g == 0 ? 1 : (g > cap ? cap : g));
}

// Workgroups of a 256-thread kernel resident on the current device at once
// (occupancy x CUs), cached per kernel (one instance per kernel) and device.
template <auto kernel>
uint32_t residentBlocks() {
    static thread_local int device = -1;
    static thread_local uint32_t blocks = 0;
    int d = 0;
    RBOW_HIP(hipGetDevice(&d));
    if (d != device) {
        int perCu = 0, cus = 0;
        RBOW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, 256, 0));
        RBOW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d));
        blocks = uint32_t(std::max(1, perCu * cus));
        device = d;
    }
    return blocks;
}

template <bool kIdentity, bool kLast, int B, uint32_t NQ>
void launchSplit(const TreeView& t, uint32_t grid, uint64_t levelBase, const uint8_t* desc, uint32_t n,
                 const uint32_t* perm, const LkItem* items, const uint32_t* nbItemsPtr, uint32_t itemsMax, uint32_t* out,
                 uint32_t* C, uint32_t* nearTies, const LeafTies& lt, hipStream_t s) {
    const uint32_t g = std::min(grid, residentBlocks<k_levelSplit<kIdentity, kLast, B, NQ>>());
    k_levelSplit<kIdentity, kLast, B, NQ><<<g, 256, 0, s>>>(t.f32rows, t.split, t.splitStride, t.cn2, t.e1, t.e2,
                                                            levelBase, desc, n, perm, items, nbItemsPtr, itemsMax, out,
                                                            C, nearTies, lt);
    if (kLast && !kIdentity) {  // the leaf level's near ties
        RBOW_HIP(hipGetLastError());
        k_leafTies<B><<<std::min(grid, residentBlocks<k_leafTies<B>>()), 256, 0, s>>>(t.f32rows, levelBase, desc, lt,
                                                                                     out);
    }
}

template <bool kIdentity, bool kLast>
void launchLevel(const TreeView& t, uint32_t itemRows, uint64_t levelBase, const uint8_t* desc, uint32_t n,
                 const uint32_t* perm,
                 const LkItem* items, const uint32_t* nbItemsPtr, uint32_t itemsMax, uint32_t* out, uint32_t* C,
                 uint32_t* nearTies, const LeafTies& lt, hipStream_t s) {
    const uint32_t grid = gridFor(itemsMax, 4, 8192);
    if (t.centroidType == 0) {
        k_levelI8<kIdentity, kLast><<<grid, 256, 0, s>>>(t.u8rows, t.normS, levelBase, desc, n, perm, items, nbItemsPtr,
                                                       itemsMax, out, C);
    } else {
        // only resident workgroups, each looping over its items: 3.6 % faster on the 10M-query
        // b32 L4 lookup than one round of workgroups per 1.2 items.  The leaf level's items
        // hold up to itemRows rows (256 or 512).
        constexpr bool kLeaf = kLast && !kIdentity;
        if (kLeaf && itemRows == kTQLeaf) {
            if (t.b == 32)
                launchSplit<kIdentity, kLast, 32, kLeaf ? kTQLeaf : kTQ>(t, grid, levelBase, desc, n, perm, items,
                                                                           nbItemsPtr, itemsMax, out, C, nearTies, lt,
                                                                           s);
            else
                launchSplit<kIdentity, kLast, 16, kLeaf ? kTQLeaf : kTQ>(t, grid, levelBase, desc, n, perm, items,
                                                                           nbItemsPtr, itemsMax, out, C, nearTies, lt,
                                                                           s);
        } else if (t.b == 32) {
            launchSplit<kIdentity, kLast, 32, kTQ>(t, grid, levelBase, desc, n, perm, items, nbItemsPtr, itemsMax, out,
                                                   C, nearTies, lt, s);
        } else {
            launchSplit<kIdentity, kLast, 16, kTQ>(t, grid, levelBase, desc, n, perm, items, nbItemsPtr, itemsMax, out,
                                                   C, nearTies, lt, s);
        }
    }
    RBOW_HIP(hipGetLastError());
}

}  // namespace

void launchFillU32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    k_fillU32<<<gridFor(n, 256, 2048), 256, 0, s>>>(p, v, n);
    RBOW_HIP(hipGetLastError());
}

void launchPrepU8Tree(const uint8_t* rows, uint64_t nbChildren, int32_t* normU, int32_t* normS, hipStream_t s) {
    if (!nbChildren) return;
    k_prepU8<<<gridFor(nbChildren, 256, 1u << 30), 256, 0, s>>>(rows, nbChildren, normU, normS);
    RBOW_HIP(hipGetLastError());
}

void launchPrepF32Split(const float* rows, uint64_t nbNodes, uint32_t b, int8_t* split, uint64_t splitStride,
                        float* cn2, float* e1, float* e2, hipStream_t s, const uint32_t* active) {
    if (!nbNodes) return;
    const uint64_t nbChildren = nbNodes * b;
    k_prepF32Split<<<uint32_t((nbChildren + 15) / 16), 256, 0, s>>>(rows, nbChildren, b == 32 ? 5u : 4u, split,
                                                                  splitStride, cn2, e1, e2, active);
    RBOW_HIP(hipGetLastError());
}

void launchMarkDuplicateChildren(const float* rows, uint64_t nbNodes, uint32_t b, float* cn2, hipStream_t s) {
    if (!nbNodes) return;
    const uint64_t nbChildren = nbNodes * b;
    k_markDuplicateChildren<<<gridFor(nbChildren, 256, 1u << 20), 256, 0, s>>>(rows, nbChildren, b - 1, cn2);
    RBOW_HIP(hipGetLastError());
}

void launchLookupPlain(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, hipStream_t s) {
    if (!n) return;
    if (t.centroidType == 0) {
        k_lookupPlainU8<<<divUp(n, 16), 256, 0, s>>>(t.u8rows, t.normU, t.L, desc, n, leaf);
    } else if (t.b == 32) {
        k_lookupPlainF32<32><<<divUp(n, 8), 256, 0, s>>>(t.f32rows, t.lg2b, t.L, desc, n, leaf);
    } else {
        k_lookupPlainF32<16><<<divUp(n, 16), 256, 0, s>>>(t.f32rows, t.lg2b, t.L, desc, n, leaf);
    }
    RBOW_HIP(hipGetLastError());
}

// Workspace of one bucketed lookup of at most kChunk descriptors.
namespace {
constexpr uint32_t kChunk = 1u << 26;  // descriptors per pass (item ids stay below 2^20)

// Layout of the bucketed lookup's workspace for n descriptors (n <= kChunk, L >= 2):
// perm[2][n], code[n], C[cMax], O[cMax], node arrays [2][2][maxNodes+1], nbItems,
// then (256-aligned) items[2][itemsMax], then the scan's temporary storage.
struct LookupWs {
    uint64_t maxNodes, itemsMax, cMax;
    size_t scanBytes, bytes;
};

LookupWs lookupWsLayout(const TreeView& t, uint32_t n, hipStream_t s) {
    LookupWs w;
    const uint32_t items0 = (n + kTQ - 1) / kTQ;
    w.maxNodes = levelSize(t.lg2b, t.L - 1);
    w.itemsMax = items0 + w.maxNodes;
    w.cMax = w.itemsMax * t.b + 1;
    size_t scanBytes2 = 0;
    w.scanBytes = 0;
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, w.scanBytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                              int(w.cMax), s));
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes2, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                              int(w.maxNodes + 1), s));
    w.scanBytes = std::max(w.scanBytes, scanBytes2);
    const size_t words = 3ull * n + 2 * w.cMax + 4 * (w.maxNodes + 1) + 2;
    w.bytes = 4 * words + 256 + 2 * sizeof(LkItem) * w.itemsMax + 256 + w.scanBytes + 256;
    return w;
}

// The bucketed descent of n <= kChunk descriptors (L >= 2) in the given
// workspace (256-byte aligned, at least lookupWsLayout(t, n).bytes).  Allocates nothing.
void lookupMfmaInWs(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                    const LookupWs& w, void* ws, hipStream_t s) {
    const uint32_t b = t.b, lg2b = t.lg2b;
    const uint32_t items0 = (n + kTQ - 1) / kTQ;
    const uint64_t maxNodes = w.maxNodes, itemsMax = w.itemsMax, cMax = w.cMax;
    size_t scanBytes = w.scanBytes;
    uint32_t* perm[2] = {static_cast<uint32_t*>(ws), static_cast<uint32_t*>(ws) + n};
    uint32_t* code = perm[1] + n;
    uint32_t* C = code + n;
    uint32_t* O = C + cMax;
    uint32_t* nBase[2] = {O + cMax, O + cMax + (maxNodes + 1)};
    uint32_t* nCnt[2] = {nBase[1] + (maxNodes + 1), nBase[1] + 2 * (maxNodes + 1)};
    uint32_t* nbItems = nCnt[1] + (maxNodes + 1);
    uint32_t* tieCount = nbItems + 1;
    auto align = [](void* p) { return reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(p) + 255) & ~uintptr_t(255)); };
    LkItem* items[2];
    items[0] = static_cast<LkItem*>(align(tieCount + 1));
    items[1] = items[0] + itemsMax;
    void* temp = align(items[1] + itemsMax);

    // level 0: all descriptors against the root (the level kernels write every count entry the scan reads)
    uint64_t cLen = uint64_t(items0) * b + 1;
    launchLevel<true, false>(t, kTQ, 0, desc, n, nullptr, nullptr, nullptr, items0, code, C, nearTies, {}, s);
    RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(temp, scanBytes, C, O, int(cLen), s));
    // items of level l: kTQ rows, at the leaf level the vocabulary's leaf item size (256 or 512);
    // the scatter also writes the next level's item count per child
    const uint32_t leafRows = t.leafItemRows ? t.leafItemRows : lookupLeafItemRowsDefault(t);
    auto itemRowsForLevel = [&](uint32_t l) { return l + 1 == t.L ? leafRows : kTQ; };
    k_scatter<true><<<gridFor(items0, 4, 8192), 256, 0, s>>>(nullptr, nullptr, items0, n, lg2b, nullptr, code, O,
                                                            perm[0],
                                                            NextItems{nullptr, nullptr, items0, b, itemRowsForLevel(1),
                                                                      nCnt[0]});
    auto buildItems = [&](const uint32_t* pBase, const uint32_t* pN, uint32_t nItems0, uint32_t nbChildren, int dst,
                          uint32_t itemRows) {
        if (nbChildren + 1 <= kSmallScanMax)
            k_scanSmall<<<1, 1024, 0, s>>>(nCnt[dst], nBase[dst], nbChildren + 1);
        else
            RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(temp, scanBytes, nCnt[dst], nBase[dst], int(nbChildren + 1), s));
        k_writeItems<<<std::min(nbChildren + 1, 65536u), 64, 0, s>>>(O, pBase, pN, nItems0, nbChildren, lg2b, itemRows,
                                                                    nBase[dst], items[dst], nbItems, tieCount);
        RBOW_HIP(hipGetLastError());
    };
    buildItems(nullptr, nullptr, items0, b, 0, itemRowsForLevel(1));
    int cur = 0;
    for (uint32_t l = 1; l < t.L; l++) {
        const uint64_t nodesL = levelSize(lg2b, l);
        const uint32_t itemsL = uint32_t(items0 + nodesL);
        if (l + 1 == t.L) {
            // the leaf level's near ties are listed in the free permutation and code arrays (n entries each)
            const LeafTies lt{code, perm[cur ^ 1], tieCount};
            launchLevel<false, true>(t, leafRows, levelBeg(lg2b, l), desc, n, perm[cur], items[cur], nbItems, itemsL, leaf,
                                     nullptr, nearTies, lt, s);
            break;
        }
        cLen = uint64_t(itemsL) * b + 1;
        launchLevel<false, false>(t, kTQ, levelBeg(lg2b, l), desc, n, perm[cur], items[cur], nbItems, itemsL, code, C,
                                  nearTies, {}, s);
        RBOW_HIP(hipcub::DeviceScan::ExclusiveSum(temp, scanBytes, C, O, int(cLen), s));
        k_scatter<false><<<gridFor(itemsL, 4, 8192), 256, 0, s>>>(
            items[cur], nbItems, 0, n, lg2b, perm[cur], code, O, perm[cur ^ 1],
            NextItems{nBase[cur], nCnt[cur], 0, uint32_t(nodesL * b), itemRowsForLevel(l + 1), nCnt[cur ^ 1]});
        buildItems(nBase[cur], nCnt[cur], 0, uint32_t(nodesL * b), cur ^ 1, itemRowsForLevel(l + 1));
        cur ^= 1;
    }
}

// Static workspace of the allocation-free lookup: one copy per GPU (each device
// holds its own instance of this array for the life of the process).  Host code
// never uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) uint8_t scratch[kScratchBytes];

void* scratchOfCurrentDevice() {
    static std::mutex m;
    static void* ptr[256] = {};
    int d = 0;
    RBOW_HIP(hipGetDevice(&d));
    RBOW_REQUIRE(d >= 0 && d < 256, RBOW_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> g(m);
    if (!ptr[d]) {
        void* p = nullptr;
        RBOW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RBOW_REQUIRE(p != nullptr, RBOW_ERR_HIP, "no device address for the lookup scratch");
        ptr[d] = p;
    }
    return ptr[d];
}
}  // namespace

uint32_t lookupFusedDefault(const TreeView& t) { return t.centroidType == 0 && t.b == 16 ? 196608u : 0u; }

bool lookupUsesFused(const TreeView& t, uint32_t n) {
    if (t.centroidType != 0 || t.b != 16 || t.L == 0) return false;
    return n <= (t.fusedMax ? t.fusedMax : lookupFusedDefault(t));
}

namespace {
// Workgroups of the fused descent the device holds at once (4 per CU: 40 KB of
// LDS and at most 128 VGPRs each); the grid never exceeds them, so each
// workgroup stages the upper levels once and then walks its share of the blocks.
uint32_t fusedResidentBlocks() {
    static std::mutex m;
    static uint32_t cached[256] = {};
    int d = 0;
    RBOW_HIP(hipGetDevice(&d));
    std::lock_guard<std::mutex> g(m);
    if (d < 0 || d >= 256) return 1024;
    if (!cached[d]) {
        int cus = 0;
        RBOW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d));
        cached[d] = 4u * uint32_t(std::max(cus, 1));
    }
    return cached[d];
}

void launchLookupFused(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, hipStream_t s) {
    k_lookupFusedI8<<<gridFor(n, 64, fusedResidentBlocks()), 256, 0, s>>>(t.u8rows, t.normS, t.L, desc, n, leaf);
    RBOW_HIP(hipGetLastError());
}
}  // namespace

void launchLookupMfma(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                      hipStream_t s) {
    if (!n) return;
    if (t.L == 0) {
        launchFillU32(leaf, 0u, n, s);
        return;
    }
    if (lookupUsesFused(t, n)) {
        launchLookupFused(t, desc, n, leaf, s);
        return;
    }
    if (n > kChunk) {  // independent chunks
        for (uint64_t o = 0; o < n; o += kChunk)
            launchLookupMfma(t, desc + o * kDims, uint32_t(std::min<uint64_t>(kChunk, n - o)), leaf + o, nearTies, s);
        return;
    }
    if (t.L == 1) {
        launchLevel<true, true>(t, kTQ, 0, desc, n, nullptr, nullptr, nullptr, (n + kTQ - 1) / kTQ, leaf, nullptr,
                                nearTies, {}, s);
        return;
    }
    const LookupWs w = lookupWsLayout(t, n, s);
    void* ws = nullptr;
    RBOW_HIP(hipMallocAsync(&ws, w.bytes, s));
    lookupMfmaInWs(t, desc, n, leaf, nearTies, w, ws, s);
    RBOW_HIP(hipFreeAsync(ws, s));
}

size_t lookupScratchBytes() { return kScratchBytes; }

bool lookupUsesPlain(const TreeView& t, uint32_t n, int kernel) {
    const uint32_t autoMin = t.autoMin ? t.autoMin : lookupAutoDefault(t.centroidType, t.b);
    return kernel == RBOW_LOOKUP_PLAIN || t.L == 0 || (kernel == RBOW_LOOKUP_AUTO && n < autoMin);
}

size_t lookupWorkspaceBytes(const TreeView& t, uint32_t n, bool plain, hipStream_t s) {
    if (plain || n == 0 || t.L <= 1) return 0;
    RBOW_REQUIRE(n <= kChunk, RBOW_ERR_CAPACITY, "lookup batch above 2^26 descriptors");
    return lookupWsLayout(t, n, s).bytes;
}

void launchLookupInWs(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                      bool plain, void* ws, size_t wsBytes, hipStream_t s) {
    if (!n) return;
    if (plain || t.L <= 1 || lookupUsesFused(t, n)) {
        launchLookup(t, desc, n, leaf, nearTies, plain ? RBOW_LOOKUP_PLAIN : RBOW_LOOKUP_MFMA, s);
        return;
    }
    const LookupWs w = lookupWsLayout(t, n, s);
    RBOW_REQUIRE(n <= kChunk && w.bytes <= wsBytes, RBOW_ERR_HIP, "lookup workspace too small");
    lookupMfmaInWs(t, desc, n, leaf, nearTies, w, ws, s);
}

void launchLookupNoAlloc(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                         int kernel, hipStream_t s) {
    const uint32_t autoMin = t.autoMin ? t.autoMin : lookupAutoDefault(t.centroidType, t.b);
    const bool plain = kernel == RBOW_LOOKUP_PLAIN || t.L == 0 || (kernel == RBOW_LOOKUP_AUTO && n < autoMin);
    if (!n) return;
    if (plain || t.L == 1 || lookupUsesFused(t, n)) {  // none of these forms needs a workspace
        launchLookup(t, desc, n, leaf, nearTies, kernel, s);
        return;
    }
    RBOW_REQUIRE(n <= kChunk, RBOW_ERR_CAPACITY, "lookup batch above the allocation-free scratch");
    const LookupWs w = lookupWsLayout(t, n, s);
    RBOW_REQUIRE(w.bytes <= kScratchBytes, RBOW_ERR_CAPACITY,
                 "lookup batch needs more than the 64 MiB allocation-free scratch");
    lookupMfmaInWs(t, desc, n, leaf, nearTies, w, scratchOfCurrentDevice(), s);
}

uint32_t lookupAutoDefault(uint32_t centroidType, uint32_t b) {
    // measured crossovers on MI355X (256-descriptor items): uint8 trees ~140k
    // descriptors, float b16 ~48k, float b32 below 32k
    return centroidType == 0 ? 12288u : b == 16 ? 49152u : 32768u;
}

void launchLookup(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies, int kernel,
                  hipStream_t s) {
    // defaults measured on MI355X: ~150k descriptors for uint8 trees, ~40k for float trees
    const uint32_t autoMin = t.autoMin ? t.autoMin : lookupAutoDefault(t.centroidType, t.b);
    const bool plain = kernel == RBOW_LOOKUP_PLAIN || t.L == 0 || (kernel == RBOW_LOOKUP_AUTO && n < autoMin);
    if (plain) launchLookupPlain(t, desc, n, leaf, s);
    else launchLookupMfma(t, desc, n, leaf, nearTies, s);
}

uint32_t lookupLeafItemRowsDefault(const TreeView& t) { return t.centroidType != 0 ? kTQLeaf : kTQ; }

LookupTuning tuneLookup(const TreeView& t0, const uint8_t* desc, uint32_t n, uint32_t* leaf, hipStream_t s) {
    hipEvent_t e0, e1;
    RBOW_HIP(hipEventCreate(&e0));
    RBOW_HIP(hipEventCreate(&e1));
    // median of 3 timed runs after one warm run
    auto timeOne = [&](const TreeView& t, uint32_t m, int kernel) {
        float ms[3];
        auto run = [&]() {
            if (kernel == RBOW_LOOKUP_PLAIN) launchLookupPlain(t, desc, m, leaf, s);
            else launchLookupMfma(t, desc, m, leaf, nullptr, s);
        };
        run();
        for (int r = 0; r < 3; r++) {
            RBOW_HIP(hipEventRecord(e0, s));
            run();
            RBOW_HIP(hipEventRecord(e1, s));
            RBOW_HIP(hipEventSynchronize(e1));
            RBOW_HIP(hipEventElapsedTime(&ms[r], e0, e1));
        }
        std::sort(ms, ms + 3);
        return ms[1];
    };
    LookupTuning r{n + 1, 0u, 0u};
    try {
        TreeView t = t0;
        if (t.L >= 2 && n >= 65536) {
            // the bucketed descent's leaf-level item size, on the largest power-of-two prefix (at
            // most 2^22 descriptors): 512-row items read a leaf node's centroids once where
            // 256-row ones read them once per item, but leave more rows of a block unused
            TreeView a = t, b = t;
            a.fusedMax = b.fusedMax = 1u;  // the bucketed form
            a.leafItemRows = kTQ;
            b.leafItemRows = kTQLeaf;
            uint32_t m = 65536;
            while (uint64_t(m) * 2 <= n && m < (1u << 22)) m *= 2;
            r.leafItemRows = timeOne(b, m, RBOW_LOOKUP_MFMA) <= timeOne(a, m, RBOW_LOOKUP_MFMA) ? kTQLeaf : kTQ;
            t.leafItemRows = r.leafItemRows;
        }
        if (t.L >= 1 && lookupFusedDefault(t)) {
            // the MFMA form's fused/bucketed boundary: the largest size (powers of two from
            // 4,096) at which the fused descent is still at least as fast as the bucketed one
            TreeView fused = t, bucketed = t;
            fused.fusedMax = 0xFFFFFFFFu;
            bucketed.fusedMax = 1u;
            r.fusedMax = 4096u;
            for (uint64_t m = 4096; m <= n; m *= 2) {
                if (timeOne(fused, uint32_t(m), RBOW_LOOKUP_MFMA) > timeOne(bucketed, uint32_t(m), RBOW_LOOKUP_MFMA))
                    break;
                r.fusedMax = uint32_t(m);
            }
            t.fusedMax = r.fusedMax;
        }
        // the AUTO crossover with that MFMA form: the smallest size at which it is at least as fast as plain
        for (uint64_t m = 4096; m <= n; m *= 2) {
            if (t.L == 0) break;
            if (timeOne(t, uint32_t(m), RBOW_LOOKUP_MFMA) <= timeOne(t, uint32_t(m), RBOW_LOOKUP_PLAIN)) {
                r.crossover = uint32_t(m);
                break;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    RBOW_HIP(hipEventDestroy(e0));
    RBOW_HIP(hipEventDestroy(e1));
    return r;
}

}  // namespace rbow_amd
