// certfilter.hpp -- certified candidate filter with exact FP32 re-scoring,
// shared by the float-tree lookup (lookup.hip) and the float k-means assign
// (kmeans.hip).
//
// The defining distance between a uint8 descriptor x and a float centroid c is
// the plain FP32 chain  acc = fma(x_k - c_k, x_k - c_k, acc),  k = 0..127
// ascending (vocTreeLookUp.cu's float analogue, and the builder's float
// k-means).  Each centroid is split once (k_prepF32Split in lookup.hip) as
// c - 128 = A0 + 2^-7 A1 + 2^-14 A2 + r with int8 terms, so that I8 MFMAs on
// x - 128 give exact int32 products; their FP32 combination t approximates
// x''.(c - 128) (x'' = x - 128), and the score s = ||c - 128||^2 - 2t errs
// against d(x, c) - ||x''||^2 by at most E1 = e1 ||x''|| + e2 (per-child e1, e2
// from the split).  Adding 2^-16 (>= 131 * 2^-24) of the distance for the plain
// chain's own relative error, a child that can still be the argmin under these
// bounds is a candidate (filterSplit tests a superset with one threshold per
// row); one candidate is the plain argmin, several ("near ties") are re-scored
// with the plain chain in child order (resolveTies), so the result equals the
// plain kernel's exactly (ties to the lowest).
#pragma once
#include "common.hpp"

namespace rbow_amd {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr float kGammaPlain = 1.0f / 65536.0f;  // 2^-16 >= 131 * 2^-24: relative error of the plain FP32 chain

// The plain FP32 distance chain of a 128-byte descriptor row and a float row.
__device__ inline float plainF32Distance(const uint8_t* x, const float* c) {
    const uint4* x4 = reinterpret_cast<const uint4*>(x);
    const float4* cr = reinterpret_cast<const float4*>(c);
    float a = 0.0f;
#pragma unroll 4
    for (uint32_t w4 = 0; w4 < 8; w4++) {  // 16 dims per step, k ascending
        const uint4 xv = x4[w4];
        const uint32_t xw[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const float4 cv = cr[4 * w4 + e];
            float d = float(xw[e] & 0xFFu) - cv.x;
            a = fmaf(d, d, a);
            d = float((xw[e] >> 8) & 0xFFu) - cv.y;
            a = fmaf(d, d, a);
            d = float((xw[e] >> 16) & 0xFFu) - cv.z;
            a = fmaf(d, d, a);
            d = float(xw[e] >> 24) - cv.w;
            a = fmaf(d, d, a);
        }
    }
    return a;
}

// ---- children as the MFMA A operand, rows as B ----
//
// With the node's children as A and 16 descriptor rows as B, the accumulator
// of lane l holds child cb*16 + 4*(l>>4) + i of row (l & 15): every lane owns
// one row and 4*CB of its children, and a row's children are spread over the
// four lanes l, l^16, l^32, l^48.  Row minima then take an in-lane minimum and
// two permlane swaps instead of a 16-lane reduction per row.
//
// Per child, the computed score s_j differs from the plain chain's distance
// minus ||x''||^2 by at most E_j = (E1_j + 2^-16 (max(||x''||^2 + s_j, 0) + E1_j))
// (1 + 2^-10): E1_j = e1_j ||x''|| + e2_j against the exact distance, and 2^-16
// of the distance for the plain chain's own error.  One threshold per row
// replaces the per-child test s_j - E_j <= min_i (s_i + E_i).  With E1max =
// max_j e1_j ||x''|| + max_j e2_j >= E1_j, m = min_j s_j and D = max(||x''||^2
// + m, 0), every child j with s_j - m > tau, tau = 2 sigma (A + g D) / (1 - g
// sigma), A = E1max (1 + g), g = 2^-16, sigma = 1 + 2^-10, has
// s_j - E_j > m + E_min (E_j grows by at most g sigma per unit of s_j, and
// E_min <= (A + g D) sigma), so its plain distance exceeds that of the child
// attaining m: it cannot be the argmin.  The candidates {s_j - m <= tau} are
// therefore a superset of the per-child test's, and the argmin rule (a single
// candidate, or the plain chain over several) is unchanged.
constexpr float kTauScale = 2.0f * (1.0f + 1.0f / 1024.0f) * (1.0f + 1.0f / 16384.0f);  // 2 sigma / (1 - g sigma), rounded up

// Minimum of two finite floats: one v_min_f32 (a compare and select would take two).
// The scores are never NaN; for -0 against +0 either may come back, which changes no
// threshold (m + tau) and no test s <= thr.
__device__ inline float minF(float a, float b) { return __builtin_fminf(a, b); }

// Minimum over lanes l, l ^ 16, l ^ 32, l ^ 48, in every lane.  IEEE minimum
// (v_minimum3_f32): fminf would quiet each swapped-in value first (two more
// instructions per step), which finite scores do not need.
__device__ inline float minAcrossGroupsF(float v) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __builtin_elementwise_minimum(__uint_as_float(p[0]), __uint_as_float(p[1]));
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __builtin_elementwise_minimum(__uint_as_float(q[0]), __uint_as_float(q[1]));
}
__device__ inline uint32_t orAcrossGroups(uint32_t v) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    v = p[0] | p[1];
    const auto q = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return q[0] | q[1];
}
// Sum over lanes l, l ^ 16, l ^ 32, l ^ 48, in every lane.
__device__ inline int sumAcrossGroups(int v) {
    const auto p = __builtin_amdgcn_permlane16_swap(uint32_t(v), uint32_t(v), false, false);
    v = int(p[0]) + int(p[1]);
    const auto q = __builtin_amdgcn_permlane32_swap(uint32_t(v), uint32_t(v), false, false);
    return int(q[0]) + int(q[1]);
}

// Candidate mask of one row in the transposed layout.  t[cb][i]: the dot part of
// child cb*16 + 4*(lane>>4) + i with row (lane & 15); xn2: ||x''||^2 of that row
// (an exact integer); cn[cb][i]: ||c - 128||^2 of the same children; e1max,
// e2max: the node's largest e1 / e2; rowValid: the lane's row exists.  With s_j
// = fl(cn_j - 2 t_j) and m = min_j s_j, child j is kept when s_j <= fl(m + tau):
// rounding is monotone and s_j is a float, so s_j - m <= tau implies it, and the
// kept set contains the per-row test's above.  Returns the mask (bit j: child
// j), the same in the row's four lanes; a single bit is the plain-chain argmin.
template <int CB>
__device__ __forceinline__ uint32_t filterSplit(const v4f (&t)[CB], float xn2, const float (&cn)[CB][4], float e1max,
                                                float e2max, uint32_t lane, bool rowValid) {
    const uint32_t grp = lane >> 4;
    float sc[CB][4];
    float m = INFINITY;
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            sc[cb][i] = fmaf(-2.0f, t[cb][i], cn[cb][i]);
            m = minF(m, sc[cb][i]);
        }
    m = minAcrossGroupsF(m);
    const float xn = __builtin_amdgcn_sqrtf(xn2) * (1.0f + 1.0f / 1048576.0f);  // v_sqrt_f32: <= 1 ulp
    const float e1v = fmaf(e1max, xn, e2max);
    const float D = fmaxf(xn2 + m, 0.0f);
    const float thr = m + kTauScale * fmaf(kGammaPlain, D, e1v * (1.0f + kGammaPlain));
    uint32_t mask = 0;
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int i = 0; i < 4; i++) mask |= sc[cb][i] <= thr ? 1u << (cb * 16 + 4 * grp + i) : 0u;
    mask = orAcrossGroups(mask);
    return rowValid ? mask : 1u;
}

// The dot part of the score from the three exact products (split terms A0, A1,
// A2): t = fl(P0 + 2^-14 fl(2^7 P1 + P2)), 2^7 P1 + P2 exact in int32.
__device__ inline float combineSplit(int p0, int p1, int p2) {
    return fmaf(float((p1 << 7) + p2), 1.0f / 16384.0f, float(p0));
}

// Near ties of one wave, queued in LDS and resolved together so that the
// loads of many rows overlap (a tie resolved where it appears stalls the wave
// on one row's dependent chain).  Entry e: tag[e] (the descriptor row), aux[e]
// (the caller's handle for the result) and mask[e] (the row's candidates).
// Capacity Cap; a caller pushes one block (<= 16 rows) at a time and resolves
// the queue before it could overflow.
template <uint32_t Cap>
struct TieQueueT {
    static constexpr uint32_t kCap = Cap;
    uint32_t tag[Cap], aux[Cap], mask[Cap];
};
constexpr uint32_t kTieCap = 64;
using TieQueue = TieQueueT<kTieCap>;

// Appends the rows of this block with several candidates: the group-0 lane of
// a row with tie == true adds (tag, aux, mask).  n: the wave's queue length
// (uniform).  Returns the number of rows added.
template <class Q>
__device__ inline uint32_t tiePush(Q& tq, uint32_t& n, uint32_t lane, bool tie, uint32_t tag, uint32_t aux,
                                   uint32_t mask) {
    const uint64_t b = __ballot(tie && (lane >> 4) == 0);
    if (tie && (lane >> 4) == 0) {
        const uint32_t pos = n + __builtin_amdgcn_mbcnt_hi(uint32_t(b >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(b), 0u));
        tq.tag[pos] = tag;
        tq.aux[pos] = aux;
        tq.mask[pos] = mask;
    }
    const uint32_t added = uint32_t(__builtin_popcountll(b));
    n += added;
    return added;
}

// Resolves the n queued rows: each row's candidates are re-scored with the
// plain FP32 chain, four lanes per row taking every fourth candidate, 16 rows
// per round; done(aux, best) runs in the row's first lane with the plain
// argmin (ties to the lowest child).  rowPtr(tag): the row's 128 bytes;
// childPtr(j): centroid j's float row.  Leaves n = 0.
template <class Q, class RowPtr, class ChildPtr, class Done>
__device__ __forceinline__ void resolveTies(const Q& tq, uint32_t& n, uint32_t lane, RowPtr rowPtr,
                                            ChildPtr childPtr, Done done) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the queue writes of this wave are visible
    __builtin_amdgcn_wave_barrier();
    for (uint32_t r0 = 0; r0 < n; r0 += 16) {
        const uint32_t e = r0 + (lane >> 2), sub = lane & 3;
        const bool live = e < n;
        const uint32_t mask = live ? tq.mask[e] : 0u, tag = live ? tq.tag[e] : 0u;
        const uint32_t nc = __builtin_popcount(mask);
        float bd = INFINITY;
        uint32_t bj = 0xFFFFFFFFu;
        for (uint32_t k = sub; __ballot(live && k < nc); k += 4) {
            if (live && k < nc) {
                uint32_t mm = mask;
                for (uint32_t r = 0; r < k; r++) mm &= mm - 1u;
                const uint32_t c = __builtin_ctz(mm);
                const float a = plainF32Distance(rowPtr(tag), childPtr(c));
                if (a < bd || (a == bd && c < bj)) {
                    bd = a;
                    bj = c;
                }
            }
        }
        // lexicographic (distance, child) minimum over the row's four lanes
#pragma unroll
        for (int d = 1; d <= 2; d <<= 1) {
            const float od = __shfl_xor(bd, d, 64);
            const uint32_t oj = uint32_t(__shfl_xor(int(bj), d, 64));
            if (od < bd || (od == bd && oj < bj)) {
                bd = od;
                bj = oj;
            }
        }
        if (live && sub == 0) done(tq.aux[e], bj);
    }
    n = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // results and the emptied queue are visible
    __builtin_amdgcn_wave_barrier();
}

}  // namespace rbow_amd