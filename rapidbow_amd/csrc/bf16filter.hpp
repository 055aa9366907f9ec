// bf16filter.hpp -- certified split-BF16 candidate filter with exact FP32
// re-scoring, shared by the float-tree lookup (lookup.hip) and the float
// k-means assign (kmeans.hip).
//
// The defining distance between a uint8 descriptor x and a float centroid c is
// the plain FP32 chain  acc = fma(x_k - c_k, x_k - c_k, acc),  k = 0..127
// ascending (vocTreeLookUp.cu's float analogue, and the builder's float
// k-means).  Per node (or k-means task) of b centroids, an integer centre mu
// (0..255) is removed: x' = x - mu is exact in BF16, and c - mu is split as
// hi + lo (+ residual) in BF16 (k_prepF32 in lookup.hip).  An MFMA pass gives
//   s_j = ||c_j - mu||^2 - 2 (x'.hi_j + x'.lo_j)    (FP32 accumulation)
// whose error against d(x, c_j) - ||x'||^2 is at most
//   E_j = (E1 + 2^-16 (max(||x'||^2 + s_j, 0) + E1)) (1 + 2^-10),
//   E1  = e1_j ||x'|| + e2_j,
// e1/e2 from k_prepF32 (operand split, FP32 accumulation, stored norms) and the
// 2^-16 term (>= 131 * 2^-24) the plain chain's own relative error.  A child
// that can still be the argmin under these bounds is a candidate (filterBestT
// tests a superset with one threshold per row); one candidate is the plain
// argmin, several ("near ties") are re-scored with the plain chain in child
// order, so the result equals the plain kernel's exactly (ties to the lowest).
#pragma once
#include "common.hpp"

namespace rbow_amd {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

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
// One threshold per row replaces the per-child bounds.  With E1max =
// max_j e1_j * ||x'|| + max_j e2_j >= E1_j, m = min_j s_j and D = max(||x'||^2
// + m, 0), every child j with s_j - m > tau, tau = 2 sigma (A + g D) / (1 - g
// sigma), A = E1max (1 + g), g = 2^-16, sigma = 1 + 2^-10, has
// s_j - E_j > m + E_min (E_j as above grows by at most g sigma per unit of s_j,
// and E_min <= (A + g D) sigma), so its plain distance exceeds that of the
// child attaining m: it cannot be the argmin.  The candidates {s_j - m <= tau}
// are therefore a superset of the per-child filter's, and the argmin rule
// (a single candidate, or the plain chain over several) is unchanged.
constexpr float kTauScale = 2.0f * (1.0f + 1.0f / 1024.0f) * (1.0f + 1.0f / 16384.0f);  // 2 sigma / (1 - g sigma), rounded up

__device__ inline float minF(float a, float b) { return b < a ? b : a; }

// Minimum over lanes l, l ^ 16, l ^ 32, l ^ 48, in every lane.
__device__ inline float minAcrossGroupsF(float v) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = minF(__uint_as_float(p[0]), __uint_as_float(p[1]));
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return minF(__uint_as_float(q[0]), __uint_as_float(q[1]));
}
__device__ inline uint32_t orAcrossGroups(uint32_t v) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    v = p[0] | p[1];
    const auto q = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return q[0] | q[1];
}
// Lexicographic minimum of (distance, child) over lanes l, l ^ 16, l ^ 32, l ^ 48.
__device__ inline void argminAcrossGroups(float& d, uint32_t& j) {
    auto step = [&](float od, uint32_t oj) {
        if (od < d || (od == d && oj < j)) {
            d = od;
            j = oj;
        }
    };
    {
        const auto pd = __builtin_amdgcn_permlane16_swap(__float_as_uint(d), __float_as_uint(d), false, false);
        const auto pj = __builtin_amdgcn_permlane16_swap(j, j, false, false);
        // lane l now sees both its own value and the swapped partner's in [0] / [1]
        const float d0 = __uint_as_float(pd[0]), d1 = __uint_as_float(pd[1]);
        d = d0;
        j = pj[0];
        step(d1, pj[1]);
    }
    {
        const auto pd = __builtin_amdgcn_permlane32_swap(__float_as_uint(d), __float_as_uint(d), false, false);
        const auto pj = __builtin_amdgcn_permlane32_swap(j, j, false, false);
        const float d0 = __uint_as_float(pd[0]), d1 = __uint_as_float(pd[1]);
        d = d0;
        j = pj[0];
        step(d1, pj[1]);
    }
}

// One block of 16 rows in the transposed layout.  acc[cb][i]: dot part of
// child cb*16 + 4*(lane>>4) + i with row (lane & 15); xn2: ||x'||^2 of that row
// (an exact integer); cn[cb][i]: ||c'||^2 of the same children; e1max, e2max:
// the node's largest e1 / e2.  rowValid / rowPtr: the lane's row exists / its
// 128 bytes; childPtr(j): centroid j's float row.  Returns the plain-chain
// argmin of the lane's row (ties to the lowest child), the same in the row's
// four lanes, and counts a row that needed the plain chain in ties (once, in
// its group-0 lane).
template <int CB, class RowPtr, class ChildPtr>
__device__ __forceinline__ uint32_t filterBestT(const v4f (&acc)[CB], float xn2, const float (&cn)[CB][4],
                                                float e1max, float e2max, uint32_t lane, bool rowValid,
                                                RowPtr rowPtr, ChildPtr childPtr, uint32_t& ties) {
    const uint32_t grp = lane >> 4;
    float s[CB][4];
    float m = INFINITY;
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            s[cb][i] = fmaf(-2.0f, acc[cb][i], cn[cb][i]);
            m = minF(m, s[cb][i]);
        }
    m = minAcrossGroupsF(m);
    const float xn = __builtin_amdgcn_sqrtf(xn2) * (1.0f + 1.0f / 1048576.0f);  // v_sqrt_f32: <= 1 ulp
    const float e1v = fmaf(e1max, xn, e2max);
    const float D = fmaxf(xn2 + m, 0.0f);
    const float tau = kTauScale * fmaf(kGammaPlain, D, e1v * (1.0f + kGammaPlain));
    uint32_t mask = 0;
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (s[cb][i] - m <= tau) mask |= 1u << (cb * 16 + 4 * grp + i);
    mask = orAcrossGroups(mask);
    uint32_t best = __builtin_ctz(mask);
    const bool multi = rowValid && (mask & (mask - 1u)) != 0;
    if (__ballot(multi)) {
        // the row's four lanes take candidates grp, grp + 4, ... in turn
        const uint32_t nc = __builtin_popcount(mask);
        float bd = INFINITY;
        uint32_t bj = 0xFFFFFFFFu;
        for (uint32_t k = grp; __ballot(multi && k < nc); k += 4) {
            if (multi && k < nc) {
                uint32_t mm = mask;
                for (uint32_t r = 0; r < k; r++) mm &= mm - 1u;
                const uint32_t c = __builtin_ctz(mm);
                const float a = plainF32Distance(rowPtr(), childPtr(c));
                if (a < bd || (a == bd && c < bj)) {
                    bd = a;
                    bj = c;
                }
            }
        }
        argminAcrossGroups(bd, bj);
        if (multi) best = bj;
        if (multi && grp == 0) ties++;
    }
    return best;
}

}  // namespace rbow_amd
