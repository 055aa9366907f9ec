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
//    int32.  Float trees: split-BF16 MFMA (c ~= hi + lo) with FP32
//    accumulation gives an approximate score with a certified error bound;
//    every child whose interval can reach the minimum is a candidate, and when
//    more than one survives the candidates are re-scored in the plain FP32
//    order, so the argmin equals the plain kernel's.
#include "kernels.hpp"

namespace rbow_amd {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kScoreBias = 1 << 25;   // makes integer scores positive before key packing
constexpr uint32_t kTQ = 128;         // descriptors per MFMA work item (8 row blocks of 16)
constexpr float kRelXsq = 1.0f / 16384.0f;  // 2^-14: plain-order + rounding bound coefficient

struct LkItem {
    uint32_t node;   // node index within the level
    uint32_t start;  // first position in the bucketed permutation
    uint32_t cnt;    // descriptors (1..kTQ)
};

__device__ inline v4i ld16(const uint8_t* p) { return *reinterpret_cast<const v4i*>(p); }
__device__ inline v4i toSigned(v4i v) { return v ^ v4i{int(0x80808080), int(0x80808080), int(0x80808080), int(0x80808080)}; }

__device__ inline uint16_t bf16Rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ inline float bf16ToF(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// 8 uint8 (as two dwords) -> 8 exact bf16 values
__device__ inline v8bf bytesToBf16(uint2 w) {
    v8bf r;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        r[e] = __builtin_bit_cast(__bf16, uint16_t(__float_as_uint(float((w.x >> (8 * e)) & 0xFFu)) >> 16));
        r[4 + e] = __builtin_bit_cast(__bf16, uint16_t(__float_as_uint(float((w.y >> (8 * e)) & 0xFFu)) >> 16));
    }
    return r;
}

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

// Bound of |score_mfma - score_plain| for one child (see the header comment):
//   e1 * ||x|| + 2^-14 * ||x||^2 + e2, with
//   e1 = 2||c-hi-lo|| + 2^-15 (||hi|| + ||lo||)   (operand split + MFMA accumulation)
//   e2 = 2^-14 ||c||^2                               (plain fma-chain rounding, norms)
__global__ void k_prepF32(const float* rows, uint64_t nbChildren, uint16_t* hi, uint16_t* lo, float* cn2, float* e1,
                          float* e2) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nbChildren) return;
    double cn = 0, rn = 0, hn = 0, ln = 0;
    for (uint32_t k = 0; k < kDims; k++) {
        const float c = rows[i * kDims + k];
        const uint16_t h = bf16Rne(c);
        const float hf = bf16ToF(h);
        const uint16_t l = bf16Rne(c - hf);
        const float lf = bf16ToF(l);
        const double r = double(c) - double(hf) - double(lf);
        cn += double(c) * c;
        rn += r * r;
        hn += double(hf) * hf;
        ln += double(lf) * lf;
        hi[i * kDims + k] = h;
        lo[i * kDims + k] = l;
    }
    const double slack = 1.0 + 1.0 / 1024;
    cn2[i] = float(cn);
    e1[i] = float((2.0 * sqrt(rn) + (sqrt(hn) + sqrt(ln)) / 32768.0) * slack);
    e2[i] = float(cn / 16384.0 * slack + 1e-3);
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
        uint32_t key = (uint32_t(score + kScoreBias) << 4) | lane;
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) key = min(key, uint32_t(__shfl_xor(int(key), m, 16)));
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
__device__ inline uint32_t waveInclusiveScan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = uint32_t(__shfl_up(int(v), d, 64));
        if (lane >= uint32_t(d)) v += o;
    }
    return v;
}

// One 1024-thread block: cursor[v] = exclusive prefix of hist[v], and the
// work-item list (<= kTQ descriptors of one node per item).
__global__ __launch_bounds__(1024) void k_scanItems(const uint32_t* __restrict__ hist, uint32_t nb,
                                                   uint32_t* __restrict__ cursor, LkItem* __restrict__ items,
                                                   uint32_t* __restrict__ nbItems) {
    __shared__ uint32_t wsC[16], wsI[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t beg = min(nb, tid * per), end = min(nb, beg + per);
    uint32_t sC = 0, sI = 0;
    for (uint32_t v = beg; v < end; v++) {
        const uint32_t c = hist[v];
        sC += c;
        sI += (c + kTQ - 1) / kTQ;
    }
    const uint32_t iC = waveInclusiveScan(sC, lane), iI = waveInclusiveScan(sI, lane);
    if (lane == 63) {
        wsC[wave] = iC;
        wsI[wave] = iI;
    }
    __syncthreads();
    uint32_t offC = 0, offI = 0;
    for (uint32_t w = 0; w < wave; w++) {
        offC += wsC[w];
        offI += wsI[w];
    }
    uint32_t runC = offC + iC - sC, runI = offI + iI - sI;
    for (uint32_t v = beg; v < end; v++) {
        const uint32_t c = hist[v];
        cursor[v] = runC;
        for (uint32_t t = 0; t * kTQ < c; t++) items[runI + t] = LkItem{v, runC + t * kTQ, min(kTQ, c - t * kTQ)};
        runC += c;
        runI += (c + kTQ - 1) / kTQ;
    }
    if (tid == 1023) *nbItems = runI;
}

__global__ void k_scatter(const uint32_t* __restrict__ cur, uint32_t n, uint32_t* __restrict__ cursor,
                          uint32_t* __restrict__ perm) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const uint32_t pos = atomicAdd(&cursor[cur[q]], 1u);
        perm[pos] = q;
    }
}

__global__ void k_fillU32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Resolve an item: node, first position, count.  kIdentity = level 0 (no bucketing).
template <bool kIdentity>
__device__ inline void itemAt(uint32_t it, uint32_t n, const LkItem* items, uint32_t& node, uint32_t& start,
                              uint32_t& cnt) {
    if (kIdentity) {
        node = 0;
        start = it * kTQ;
        cnt = min(kTQ, n - start);
    } else {
        const LkItem t = items[it];
        node = t.node;
        start = t.start;
        cnt = t.cnt;
    }
}

// --------------------------------------------------- I8 MFMA level step --
// Each wave takes one item: up to kTQ descriptors that sit at the same node.
// mfma_i32_16x16x64_i8: lane l feeds A row (l&15) / B column (l&15) with the
// 16 bytes of dims [16*(l>>4) + 64*kb, +16), kb = 0,1 (same placement for A
// and B, so the sum runs over all 128 dims); the result lane l holds rows
// 4*(l>>4)+i, column l&15.
template <bool kIdentity, bool kLast>
__global__ __launch_bounds__(256) void k_levelI8(const uint8_t* __restrict__ rows, const int32_t* __restrict__ normS,
                                                uint64_t levelBase, const uint8_t* __restrict__ desc, uint32_t n,
                                                const uint32_t* __restrict__ perm, const LkItem* __restrict__ items,
                                                const uint32_t* __restrict__ nbItemsPtr, uint32_t nbItemsImplicit,
                                                uint32_t* __restrict__ out, uint32_t* __restrict__ histNext) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    const uint32_t nbItems = kIdentity ? nbItemsImplicit : *nbItemsPtr;
    for (uint32_t it = blockIdx.x * 4 + wave; it < nbItems; it += gridDim.x * 4) {
        uint32_t node, start, cnt;
        itemAt<kIdentity>(it, n, items, node, start, cnt);
        const uint64_t child = (levelBase + node) * 16 + col;
        const uint8_t* crow = rows + child * kDims;
        const v4i b0 = toSigned(ld16(crow + grp * 16)), b1 = toSigned(ld16(crow + 64 + grp * 16));
        const int ns = normS[child];
        for (uint32_t rb = 0; rb * 16 < cnt; rb++) {
            const uint32_t r = rb * 16 + col;
            v4i a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
            if (r < cnt) {
                const uint32_t q = kIdentity ? start + r : perm[start + r];
                const uint8_t* x = desc + uint64_t(q) * kDims;
                a0 = toSigned(ld16(x + grp * 16));
                a1 = toSigned(ld16(x + 64 + grp * 16));
            }
            v4i acc = {0, 0, 0, 0};
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b1, acc, 0, 0, 0);
            uint32_t key[4];
#pragma unroll
            for (int i = 0; i < 4; i++) key[i] = (uint32_t(ns - 2 * acc[i] + kScoreBias) << 4) | col;
#pragma unroll
            for (int m = 1; m < 16; m <<= 1)
#pragma unroll
                for (int i = 0; i < 4; i++) key[i] = min(key[i], uint32_t(__shfl_xor(int(key[i]), m, 16)));
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t row = rb * 16 + 4 * grp + i;
                if (col == uint32_t(i) && row < cnt) {
                    const uint32_t q = kIdentity ? start + row : perm[start + row];
                    const uint32_t nn = node * 16 + (key[i] & 15u);
                    out[q] = nn;
                    if (!kLast) atomicAdd(&histNext[nn], 1u);
                }
            }
        }
    }
}

// ----------------------------------------------- split-BF16 level step --
// mfma_f32_16x16x32_bf16: lane l feeds 8 dims [8*(l>>4) + 32*kb, +8), kb=0..3.
template <bool kIdentity, bool kLast, int B>
__global__ __launch_bounds__(256) void k_levelBf16(
    const float* __restrict__ rows32, const uint16_t* __restrict__ hi, const uint16_t* __restrict__ lo,
    const float* __restrict__ cn2, const float* __restrict__ e1, const float* __restrict__ e2, uint64_t levelBase,
    const uint8_t* __restrict__ desc, uint32_t n, const uint32_t* __restrict__ perm, const LkItem* __restrict__ items,
    const uint32_t* __restrict__ nbItemsPtr, uint32_t nbItemsImplicit, uint32_t* __restrict__ out,
    uint32_t* __restrict__ histNext, uint32_t* __restrict__ nearTies) {
    constexpr int CB = B / 16;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    const uint32_t nbItems = kIdentity ? nbItemsImplicit : *nbItemsPtr;
    uint32_t ties = 0;
    for (uint32_t it = blockIdx.x * 4 + wave; it < nbItems; it += gridDim.x * 4) {
        uint32_t node, start, cnt;
        itemAt<kIdentity>(it, n, items, node, start, cnt);
        const uint64_t childBase = (levelBase + node) * B;
        v8bf bh[CB][4], bl[CB][4];
        float cn[CB], ea[CB], eb[CB];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const uint64_t child = childBase + cb * 16 + col;
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                bh[cb][kb] = *reinterpret_cast<const v8bf*>(hi + child * kDims + kb * 32 + grp * 8);
                bl[cb][kb] = *reinterpret_cast<const v8bf*>(lo + child * kDims + kb * 32 + grp * 8);
            }
            cn[cb] = cn2[child];
            ea[cb] = e1[child];
            eb[cb] = e2[child];
        }
        for (uint32_t rb = 0; rb * 16 < cnt; rb++) {
            const uint32_t r = rb * 16 + col;
            v8bf ax[4];
            uint32_t xsq = 0;
            if (r < cnt) {
                const uint32_t q = kIdentity ? start + r : perm[start + r];
                const uint8_t* x = desc + uint64_t(q) * kDims;
#pragma unroll
                for (int kb = 0; kb < 4; kb++) {
                    const uint2 w = *reinterpret_cast<const uint2*>(x + kb * 32 + grp * 8);
                    ax[kb] = bytesToBf16(w);
                    xsq = __builtin_amdgcn_udot4(w.x, w.x, xsq, false);
                    xsq = __builtin_amdgcn_udot4(w.y, w.y, xsq, false);
                }
            } else {
#pragma unroll
                for (int kb = 0; kb < 4; kb++) ax[kb] = v8bf{};
            }
            xsq += uint32_t(__shfl_xor(int(xsq), 16, 64));
            xsq += uint32_t(__shfl_xor(int(xsq), 32, 64));
            v4f acc[CB];
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                acc[cb] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < 4; kb++) {
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[kb], bh[cb][kb], acc[cb], 0, 0, 0);
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax[kb], bl[cb][kb], acc[cb], 0, 0, 0);
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t rowLocal = 4 * grp + i;
                const float xn2 = float(uint32_t(__shfl(int(xsq), int(rowLocal), 64)));
                const float xn = sqrtf(xn2) * (1.0f + 1.0f / 1048576.0f);
                float s[CB], E[CB];
                float U = INFINITY;
#pragma unroll
                for (int cb = 0; cb < CB; cb++) {
                    s[cb] = cn[cb] - 2.0f * acc[cb][i];
                    E[cb] = ea[cb] * xn + kRelXsq * xn2 + eb[cb];
                    U = fminf(U, s[cb] + E[cb]);
                }
#pragma unroll
                for (int m = 1; m < 16; m <<= 1) U = fminf(U, __shfl_xor(U, m, 16));
                uint32_t cmask = 0;
#pragma unroll
                for (int cb = 0; cb < CB; cb++) {
                    const uint64_t bal = __ballot(s[cb] - E[cb] <= U);
                    cmask |= uint32_t((bal >> (16 * grp)) & 0xFFFFu) << (16 * cb);
                }
                const uint32_t row = rb * 16 + rowLocal;
                if (col == uint32_t(i) && row < cnt) {
                    const uint32_t q = kIdentity ? start + row : perm[start + row];
                    uint32_t j = __builtin_ctz(cmask);
                    if (__builtin_popcount(cmask) > 1) {
                        // near tie: exact plain FP32 order over the candidates, ascending j
                        ties++;
                        const uint8_t* x = desc + uint64_t(q) * kDims;
                        float best = INFINITY;
                        for (uint32_t m = cmask; m; m &= m - 1) {
                            const uint32_t c = __builtin_ctz(m);
                            const float* cr = rows32 + (childBase + c) * kDims;
                            float a = 0.0f;
                            for (uint32_t k = 0; k < kDims; k++) {
                                const float d = float(x[k]) - cr[k];
                                a = fmaf(d, d, a);
                            }
                            if (a < best) {
                                best = a;
                                j = c;
                            }
                        }
                    }
                    const uint32_t nn = node * B + j;
                    out[q] = nn;
                    if (!kLast) atomicAdd(&histNext[nn], 1u);
                }
            }
        }
    }
    if (nearTies && ties) atomicAdd(nearTies, ties);
}

uint32_t gridFor(uint64_t work, uint32_t perBlock, uint32_t cap) {
    const uint64_t g = (work + perBlock - 1) / perBlock;
    return uint32_t(g == 0 ? 1 : (g > cap ? cap : g));
}

template <bool kIdentity, bool kLast>
void launchLevel(const TreeView& t, uint64_t levelBase, const uint8_t* desc, uint32_t n, const uint32_t* perm,
                 const LkItem* items, const uint32_t* nbItemsPtr, uint32_t itemsMax, uint32_t* out,
                 uint32_t* histNext, uint32_t* nearTies, hipStream_t s) {
    const uint32_t grid = gridFor(itemsMax, 4, 4096);
    if (t.centroidType == 0) {
        k_levelI8<kIdentity, kLast><<<grid, 256, 0, s>>>(t.u8rows, t.normS, levelBase, desc, n, perm, items, nbItemsPtr,
                                                       itemsMax, out, histNext);
    } else if (t.b == 32) {
        k_levelBf16<kIdentity, kLast, 32><<<grid, 256, 0, s>>>(t.f32rows, t.hi, t.lo, t.cn2, t.e1, t.e2, levelBase, desc,
                                                             n, perm, items, nbItemsPtr, itemsMax, out, histNext,
                                                             nearTies);
    } else {
        k_levelBf16<kIdentity, kLast, 16><<<grid, 256, 0, s>>>(t.f32rows, t.hi, t.lo, t.cn2, t.e1, t.e2, levelBase, desc,
                                                             n, perm, items, nbItemsPtr, itemsMax, out, histNext,
                                                             nearTies);
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

void launchPrepF32Tree(const float* rows, uint64_t nbChildren, uint16_t* hi, uint16_t* lo, float* cn2, float* e1,
                       float* e2, hipStream_t s) {
    if (!nbChildren) return;
    k_prepF32<<<gridFor(nbChildren, 256, 1u << 30), 256, 0, s>>>(rows, nbChildren, hi, lo, cn2, e1, e2);
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

void launchLookupMfma(const TreeView& t, const uint8_t* desc, uint32_t n, uint32_t* leaf, uint32_t* nearTies,
                      hipStream_t s) {
    if (!n) return;
    if (t.L == 0) {
        launchFillU32(leaf, 0u, n, s);
        return;
    }
    const uint32_t b = t.b;
    const uint64_t maxBins = levelSize(t.lg2b, t.L - 1);
    const uint64_t itemsMax = (n + kTQ - 1) / kTQ + maxBins;
    // stream-ordered workspace: cur[n], perm[n], hist[maxBins], cursor[maxBins], items, nbItems
    void* ws = nullptr;
    const size_t bytes = sizeof(uint32_t) * (2ull * n + 2 * maxBins + 1) + sizeof(LkItem) * itemsMax + 256;
    if (t.L > 1) RBOW_HIP(hipMallocAsync(&ws, bytes, s));
    uint32_t* cur = static_cast<uint32_t*>(ws);
    uint32_t* perm = cur + n;
    uint32_t* hist = perm + n;
    uint32_t* cursor = hist + maxBins;
    uint32_t* nbItems = cursor + maxBins;
    LkItem* items = reinterpret_cast<LkItem*>((reinterpret_cast<uintptr_t>(nbItems + 1) + 255) & ~uintptr_t(255));

    const uint32_t items0 = (n + kTQ - 1) / kTQ;
    if (t.L == 1) {
        launchLevel<true, true>(t, 0, desc, n, nullptr, nullptr, nullptr, items0, leaf, nullptr, nearTies, s);
        return;
    }
    launchFillU32(hist, 0u, b, s);
    launchLevel<true, false>(t, 0, desc, n, nullptr, nullptr, nullptr, items0, cur, hist, nearTies, s);
    for (uint32_t l = 1; l < t.L; l++) {
        const uint32_t nb = uint32_t(levelSize(t.lg2b, l));
        k_scanItems<<<1, 1024, 0, s>>>(hist, nb, cursor, items, nbItems);
        RBOW_HIP(hipGetLastError());
        k_scatter<<<gridFor(n, 256, 4096), 256, 0, s>>>(cur, n, cursor, perm);
        RBOW_HIP(hipGetLastError());
        const uint32_t itemsL = uint32_t((n + kTQ - 1) / kTQ + nb);
        if (l + 1 < t.L) {
            launchFillU32(hist, 0u, uint64_t(nb) * b, s);
            launchLevel<false, false>(t, levelBeg(t.lg2b, l), desc, n, perm, items, nbItems, itemsL, cur, hist,
                                      nearTies, s);
        } else {
            launchLevel<false, true>(t, levelBeg(t.lg2b, l), desc, n, perm, items, nbItems, itemsL, leaf, nullptr,
                                     nearTies, s);
        }
    }
    RBOW_HIP(hipFreeAsync(ws, s));
}

}  // namespace rbow_amd
