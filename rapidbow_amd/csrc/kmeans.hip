// kmeans.hip -- batched hierarchical k-means kernels (one launch covers every
// task = tree node of a level, whatever their sizes).
//
// Reference semantics (kmeans.cu:120-198, 334-375; kmeans.cpp:118-133):
//   assign: nearest of the b centers, ties to the lowest index; per task the
//           counts, the per-center sums of every dimension and the number of
//           points whose assignment changed.
//   update: c = round(sum / count) (uint8 centroids, IEEE division, half away
//           from zero) or c = float(sum) / float(count) (float centroids) for
//           every center with count > 0; others keep their position.
//   loop:   a task stops after the first pass with changed == 0 (no update on
//           that pass) or after max_iters passes.
// Sums are exact: int32 per wave in LDS, flushed into int64 per task.
#include <hip/hip_ext.h>
#include <type_traits>

#include "certfilter.hpp"
#include "kernels.hpp"

namespace rbow_amd {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t kWaves = 4;  // waves per 256-thread block
// LDS merge slots per workgroup of the MFMA assign: one per wave boundary, so
// every same-task run of a workgroup leaves as one partial.  (Two slots let a
// third workgroup fit per CU by LDS, three waves per SIMD: measured no faster,
// with more partials to reduce.)
constexpr uint32_t kMergeSlots = kWaves - 1;

__device__ inline v4i ld16(const uint8_t* p) { return *reinterpret_cast<const v4i*>(p); }
__device__ inline v4i toSigned(v4i v) {
    return v ^ v4i{int(0x80808080), int(0x80808080), int(0x80808080), int(0x80808080)};
}

// Orders this wave's LDS accesses across lanes (the wave's own program order).
__device__ inline void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return mix64(mix64(mix64(seed ^ mix64(a)) ^ b) ^ c);
}

// ------------------------------------------------------------- gathers --
__global__ void k_gatherRows(const uint8_t* __restrict__ desc, const uint32_t* __restrict__ idx, uint64_t n,
                             uint8_t* __restrict__ out, uint64_t nbSrc) {
    for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n * 8;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t p = t >> 3, part = t & 7;
        const uint32_t i = idx[p];
        uint4 v = make_uint4(~0u, ~0u, ~0u, ~0u);
        if (i != kIdxInvalid && i < nbSrc) v = reinterpret_cast<const uint4*>(desc + uint64_t(i) * kDims)[part];
        reinterpret_cast<uint4*>(out + p * kDims)[part] = v;
    }
}

__global__ void k_iota(uint32_t* __restrict__ p, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = uint32_t(i);
}

// reformatDesc.cu:38-74: tiles[t][word][slot] = word of desc[idx[t][slot]].
// One workgroup per tile: the tile's rows are read whole (16-byte loads, 8 lanes
// per 128-byte row) into LDS, then written word-major, both coalesced.
__global__ __launch_bounds__(256) void k_reformatTiles(const uint8_t* __restrict__ desc, const uint32_t* __restrict__ idx,
                                                      uint32_t tileSize, uint32_t* __restrict__ tiles) {
    __shared__ uint32_t sx[32][33];  // [slot][word], tileSize <= 32
    const uint64_t t = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    for (uint32_t c = tid; c < tileSize * 8; c += 256) {
        const uint32_t slot = c >> 3, q = c & 7;
        const uint32_t i = idx[t * tileSize + slot];
        const uint4 v = i == kIdxInvalid ? make_uint4(~0u, ~0u, ~0u, ~0u)
                                         : reinterpret_cast<const uint4*>(desc + uint64_t(i) * kDims)[q];
        sx[slot][4 * q] = v.x;
        sx[slot][4 * q + 1] = v.y;
        sx[slot][4 * q + 2] = v.z;
        sx[slot][4 * q + 3] = v.w;
    }
    __syncthreads();
    uint32_t* out = tiles + t * 32 * tileSize;
    for (uint32_t e = tid; e < 32 * tileSize; e += 256) out[e] = sx[e % tileSize][e / tileSize];
}

// ----------------------------------------------------------- seeding ----
__global__ void k_initCenters(uint32_t b, uint32_t nbTasks, const uint8_t* __restrict__ desc,
                              const uint32_t* __restrict__ seedIdx, uint8_t* __restrict__ cU8,
                              float* __restrict__ cF32) {
    const uint64_t total = uint64_t(nbTasks) * b * kDims;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t i = seedIdx[e / kDims];
        const uint8_t v = i == kIdxInvalid ? uint8_t(0xFF) : desc[uint64_t(i) * kDims + e % kDims];
        if (cU8) cU8[e] = v;
        else cF32[e] = float(v);
    }
}

// Builder-defined k-means++ (same definition as the oracle's):
// candidates = one hashed member per stratum of the group (<= 256), then
// D^2-weighted picks with integer weights and hashed draws.  One wave per
// task: lane l holds candidates 4l .. 4l+3 (rows in registers, D^2 and norms);
// a chosen center's row is broadcast with readlane, the weighted draw is a wave
// prefix sum.  No LDS and no barrier, so the many small tasks of a deep level
// run side by side.
// Inclusive prefix sum over the 64 lanes of a wave: DPP row shifts within each
// row of 16 lanes (lanes with no source add 0), then the earlier rows' totals.
__device__ inline uint32_t waveInclusiveSum(uint32_t x) {
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xF, 0xF, false));  // row_shr:1
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xF, 0xF, false));  // row_shr:2
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x114, 0xF, 0xF, false));  // row_shr:4
    x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x118, 0xF, 0xF, false));  // row_shr:8
    const uint32_t r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31),
                   r2 = __builtin_amdgcn_readlane(x, 47);
    const uint32_t row = (threadIdx.x & 63) >> 4;
    return x + (row == 0 ? 0u : row == 1 ? r0 : row == 2 ? r0 + r1 : r0 + r1 + r2);
}

constexpr uint32_t kPPSlots = 4;  // candidates per lane (64 * 4 = 256)

template <uint32_t S>
__device__ inline void ppBroadcast(const uint32_t (&x)[kPPSlots][32], uint32_t owner, uint32_t (&c)[32]) {
#pragma unroll
    for (int w = 0; w < 32; w++) c[w] = __builtin_amdgcn_readlane(x[S][w], owner);
}

__global__ __launch_bounds__(256) void k_kmeansPP(const uint8_t* __restrict__ desc, const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ taskBeg,
                                                 const uint32_t* __restrict__ taskCnt, uint32_t nbTasks, uint32_t b,
                                                 uint64_t seed, uint32_t level, uint32_t* __restrict__ seedIdx,
                                                 uint32_t* __restrict__ masks, uint32_t nbSrc) {
    const uint32_t g = __builtin_amdgcn_readfirstlane(blockIdx.x * kWaves + (threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63;
    if (g >= nbTasks) return;
    const uint32_t n = taskCnt[g];
    if (n == 0) {
        if (lane < b) seedIdx[g * b + lane] = kIdxInvalid;
        if (lane == 0) masks[g] = 0;
        return;
    }
    const uint32_t m = n < 256 ? n : 256;
    uint32_t x[kPPSlots][32], nrm[kPPSlots], dd[kPPSlots], cand[kPPSlots];
#pragma unroll
    for (uint32_t i = 0; i < kPPSlots; i++) {
        const uint32_t c = lane * kPPSlots + i;
        cand[i] = kIdxInvalid;
        nrm[i] = 0;
        dd[i] = 0;  // candidates >= m weigh nothing
        if (c < m) {
            uint32_t pos = c;
            if (n > 256) {
                const uint64_t lo = uint64_t(c) * n / m, hi = uint64_t(c + 1) * n / m;
                pos = uint32_t(lo + hash4(seed, level, g, 0x100000000ull + c) % (hi - lo));
            }
            const uint32_t id = order[taskBeg[g] + pos];
            cand[i] = id < nbSrc ? id : kIdxInvalid;  // (an order entry is always a row id < nbSrc)
            dd[i] = 0xFFFFFFFFu;
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kPPSlots; i++) {
        const v4i* row = reinterpret_cast<const v4i*>(desc + uint64_t(cand[i] == kIdxInvalid ? 0u : cand[i]) * kDims);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const v4i v = lane * kPPSlots + i < m ? row[q] : v4i{0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < 4; e++) x[i][4 * q + e] = uint32_t(v[e]);
        }
#pragma unroll
        for (int w = 0; w < 32; w++) nrm[i] = __builtin_amdgcn_udot4(x[i][w], x[i][w], nrm[i], false);
    }
    // D^2 of every candidate to its nearest chosen one, after choosing candidate c
    auto relax = [&](uint32_t c) {
        c = __builtin_amdgcn_readfirstlane(c);  // the same in every lane
        const uint32_t owner = c / kPPSlots, slot = c % kPPSlots;
        uint32_t cw[32];
        switch (slot) {
            case 0: ppBroadcast<0>(x, owner, cw); break;
            case 1: ppBroadcast<1>(x, owner, cw); break;
            case 2: ppBroadcast<2>(x, owner, cw); break;
            default: ppBroadcast<3>(x, owner, cw); break;
        }
        const uint32_t nSel = slot == 0 ? nrm[0] : slot == 1 ? nrm[1] : slot == 2 ? nrm[2] : nrm[3];
        const uint32_t nc = __builtin_amdgcn_readlane(nSel, owner);
#pragma unroll
        for (uint32_t i = 0; i < kPPSlots; i++) {
            uint32_t dot = 0;
#pragma unroll
            for (int w = 0; w < 32; w++) dot = __builtin_amdgcn_udot4(x[i][w], cw[w], dot, false);
            if (lane * kPPSlots + i < m) dd[i] = min(dd[i], nrm[i] + nc - 2 * dot);
        }
    };
    uint32_t prev = uint32_t(hash4(seed, level, g, 0x200000000ull) % m);
    uint32_t myChosen = lane == 0 ? prev : 0u;  // lane r < b: the r-th chosen candidate
    // lane r: the draw of round r, hashed up front
    const uint64_t myDraw = lane < b ? hash4(seed, level, g, 0x300000000ull + lane) : 0ull;
    relax(prev);
    uint32_t nValid = 1;
    for (uint32_t r = 1; r < b; r++) {
        // inclusive prefix of D^2 in candidate order: lane-local, then across lanes.  Exact in
        // 32 bits: 256 candidates * 128 * 255^2 < 2^31.
        uint32_t inc[kPPSlots];
        uint32_t run = 0;
#pragma unroll
        for (uint32_t i = 0; i < kPPSlots; i++) inc[i] = run += dd[i];
        const uint32_t lanePre = waveInclusiveSum(run);
        const uint32_t total = __builtin_amdgcn_readlane(lanePre, 63);
        uint32_t pick = prev;
        if (total != 0) {
            // (readlane returns int: each half goes through uint32_t, never sign-extended)
            const uint64_t draw = (uint64_t(uint32_t(__builtin_amdgcn_readlane(uint32_t(myDraw >> 32), r))) << 32) |
                                  uint32_t(__builtin_amdgcn_readlane(uint32_t(myDraw), r));
            const uint32_t t = uint32_t(draw % total);
            const uint32_t excl = lanePre - run;
            uint32_t mine = 0;
            bool hit = false;
#pragma unroll
            for (uint32_t i = 0; i < kPPSlots; i++) {
                const uint32_t hiEnd = excl + inc[i];
                if (hiEnd > t && hiEnd - dd[i] <= t) {
                    hit = true;
                    mine = lane * kPPSlots + i;
                }
            }
            const uint64_t who = __ballot(hit);
            pick = __builtin_amdgcn_readlane(mine, int(__builtin_ctzll(who)));
            nValid++;
        }
        if (lane == r) myChosen = pick;
        if (total != 0) relax(pick);
        prev = pick;
    }
    // seeds: the chosen candidates' descriptor ids
    const uint32_t owner = myChosen / kPPSlots, slot = myChosen % kPPSlots;
    uint32_t id = kIdxInvalid;
#pragma unroll
    for (uint32_t i = 0; i < kPPSlots; i++) {
        const uint32_t v = __shfl(cand[i], int(owner), 64);
        if (slot == i) id = v;
    }
    if (lane < b) seedIdx[g * b + lane] = id;
    if (lane == 0) masks[g] = nValid >= 32 ? 0xFFFFFFFFu : ((1u << nValid) - 1u);
}

// ------------------------------------------------------- accumulators ----
// Per-wave LDS accumulators: sums [B][128] u32 and counts [B].  A wave owns a
// contiguous range of tiles and flushes into the task's int64 sums whenever
// the task changes.
template <int B>
struct WaveAcc {
    uint32_t* sums;   // [B][128]
    uint32_t* cnt;    // [B]
    __device__ void clear(uint32_t lane) {
        for (uint32_t e = lane; e < B * kDims; e += 64) sums[e] = 0;
        if (lane < B) cnt[lane] = 0;
        waveSync();
    }
    __device__ void flush(uint32_t lane, uint32_t task, uint32_t changedLocal, const KmLevel& k) {
        unsigned long long* gs = reinterpret_cast<unsigned long long*>(k.sums) + uint64_t(task) * B * kDims;
        waveSync();
        for (uint32_t e = lane; e < B * kDims; e += 64) {
            const uint32_t v = sums[e];
            if (v) atomicAdd(&gs[e], (unsigned long long)v);
            sums[e] = 0;
        }
        if (lane < B) {
            if (cnt[lane]) atomicAdd(&k.counts[task * B + lane], cnt[lane]);
            cnt[lane] = 0;
        }
        if (lane == 0 && changedLocal) atomicAdd(&k.changed[task], changedLocal);
        waveSync();
    }
    // As flush, for a task whose every tile this wave processed (sole): no other wave adds
    // to its totals, which hold zero since the last update, so plain stores of the centers
    // that received points replace the atomics.
    __device__ void flushTask(uint32_t lane, uint32_t task, uint32_t changedLocal, const KmLevel& k, bool sole) {
        if (!sole) {
            flush(lane, task, changedLocal, k);
            return;
        }
        unsigned long long* gs = reinterpret_cast<unsigned long long*>(k.sums) + uint64_t(task) * B * kDims;
        waveSync();
        for (uint32_t e = lane; e < B * kDims; e += 64) {
            if (cnt[e / kDims]) gs[e] = sums[e];
            sums[e] = 0;
        }
        waveSync();  // the counts are read above before they are cleared
        if (lane < B) {
            k.counts[task * B + lane] = cnt[lane];
            cnt[lane] = 0;
        }
        if (lane == 0) k.changed[task] = changedLocal;
        waveSync();
    }
};

// Epilogue shared by all assign kernels: lane r of the wave holds the new
// assignment a of tile row r (valid when r < cnt).
template <int B>
__device__ inline uint32_t tileEpilogue(const KmLevel& k, const KmTile& t, uint32_t lane, uint32_t a, WaveAcc<B>& acc) {
    const bool valid = lane < t.cnt;
    uint32_t changed = 0;
    if (valid) {
        const uint32_t prev = k.assign[t.row + lane];
        k.assign[t.row + lane] = uint8_t(a);
        changed = prev != a;
        atomicAdd(&acc.cnt[a], 1u);
    }
    const uint32_t nbChanged = __builtin_popcountll(__ballot(changed));
    // sums: lane owns dims 2*lane and 2*lane+1; 16 rows' bytes are loaded at once
    // and added with LDS atomics (no read-modify-write chain between rows)
    for (uint32_t r0 = 0; r0 < t.cnt; r0 += 16) {
        uint32_t v[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; j++)
            v[j] = r0 + j < t.cnt ? reinterpret_cast<const uint16_t*>(k.X + uint64_t(t.row + r0 + j) * kDims)[lane] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < 16; j++) {
            if (r0 + j >= t.cnt) break;
            const uint32_t ar = __builtin_amdgcn_readlane(a, r0 + j);
            atomicAdd(acc.sums + ar * kDims + 2 * lane, v[j] & 0xFFu);
            atomicAdd(acc.sums + ar * kDims + 2 * lane + 1, v[j] >> 8);
        }
    }
    waveSync();  // the sums are complete before the next tile or a flush reads them
    return nbChanged;
}

// Contiguous tile range of wave w among W waves.
__device__ inline void waveRange(uint32_t nbTiles, uint32_t& beg, uint32_t& end) {
    const uint32_t W = gridDim.x * kWaves, w = blockIdx.x * kWaves + (threadIdx.x >> 6);
    beg = uint32_t(uint64_t(nbTiles) * w / W);
    end = uint32_t(uint64_t(nbTiles) * (w + 1) / W);
}

// ----------------------------------------------- assign: uint8, I8 MFMA --
// Per 64-row tile the level gather writes one swizzled copy of x' = x - 128:
//   XA[tile][rb][kb][lane][16]: A fragments of the distance MFMA, row rb*16 + (lane&15),
//                               dims kb*64 + (lane>>4)*16 .. +16.
// Distances: mfma_i32_16x16x64_i8(XA, c') -> ||c'||^2 - 2 x'.c' (exact int32).
// Sums: mfma_i32_16x16x64_i8(onehot^T, XT) -> per (center, dim) sums of x' over
// the tile's points (exact int32): the update's centroid sums on the matrix
// cores, with no atomics.  XT is the transposed operand, byte q of lane l =
// x'[point 16*(q>>2) + 4*(l>>4) + (q&3)][dim 16*db + (l&15)] (the C layout of a
// 16x16 MFMA puts rows 4*(l>>4)+i of column l&15 in lane l, so the 4x4 rows a
// lane owns after the distance MFMA are the 16 points of its XT K slice and
// the one-hot operand is local).  XT is not read from HBM: the tile is put in
// LDS row-major (144-byte pitch) and read back with gfx950's transposing LDS
// load ds_read_b64_tr_b8, which within each 16-lane group gives output lane i,
// byte r = byte (i % 8) of the 8 bytes addressed by lane 2r + i/8.  So each
// point costs 128 bytes of HBM per pass.
__global__ __launch_bounds__(256) void k_gatherTiles(const uint8_t* __restrict__ desc, const uint32_t* __restrict__ order,
                                                    const KmTile* __restrict__ tiles, int8_t* __restrict__ XA,
                                                    const uint32_t* __restrict__ nbTilesDev, uint32_t nbSrc) {
    __shared__ __align__(16) uint8_t sx[64][144];
    const uint32_t t = blockIdx.x, tid = threadIdx.x;
    if (nbTilesDev && t >= *nbTilesDev) return;  // a grid sized by a bound: tiles past the count
    const KmTile tl = tiles[t];
    for (uint32_t c = tid; c < 512; c += 256) {
        const uint32_t r = c >> 3, part = c & 7;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);  // padding rows: x' = 0
        const uint32_t i = r < tl.cnt ? order[tl.row + r] : kIdxInvalid;
        if (i < nbSrc) {
            v = reinterpret_cast<const uint4*>(desc + uint64_t(i) * kDims)[part];
            v.x ^= 0x80808080u;
            v.y ^= 0x80808080u;
            v.z ^= 0x80808080u;
            v.w ^= 0x80808080u;
        }
        *reinterpret_cast<uint4*>(&sx[r][part * 16]) = v;
    }
    __syncthreads();
    int8_t* xa = XA + uint64_t(t) * 8192;
    for (uint32_t c = tid; c < 512; c += 256) {
        const uint32_t rb = c >> 7, kb = (c >> 6) & 1, l = c & 63;
        *reinterpret_cast<uint4*>(xa + c * 16) =
            *reinterpret_cast<const uint4*>(&sx[rb * 16 + (l & 15)][kb * 64 + (l >> 4) * 16]);
    }
}

__device__ inline void loadTileA(const int8_t* xa, uint32_t lane, v4i (&a)[4][2]) {
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int kb = 0; kb < 2; kb++) a[rb][kb] = *reinterpret_cast<const v4i*>(xa + ((rb * 2 + kb) * 64 + lane) * 16);
}

// A 64-row tile of x' in LDS, row-major with a 144-byte pitch (rows 8 words
// apart in the bank space, so the transposing reads below do not conflict).
constexpr uint32_t kRowPitch = 144;
constexpr uint32_t kTileLds = 64 * kRowPitch;

__device__ inline uint32_t fragOffset(int rb, int kb, uint32_t lane) {
    return (rb * 16 + (lane & 15)) * kRowPitch + kb * 64 + (lane >> 4) * 16;
}
// A fragments (XA layout) <-> the row-major LDS tile
__device__ inline void storeTileLds(uint8_t* t, uint32_t lane, const v4i (&a)[4][2]) {
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int kb = 0; kb < 2; kb++) *reinterpret_cast<v4i*>(t + fragOffset(rb, kb, lane)) = a[rb][kb];
}
__device__ inline void loadTileLds(const uint8_t* t, uint32_t lane, v4i (&a)[4][2]) {
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int kb = 0; kb < 2; kb++) a[rb][kb] = *reinterpret_cast<const v4i*>(t + fragOffset(rb, kb, lane));
}

typedef int v2i __attribute__((ext_vector_type(2)));
__device__ inline v2i readTr8(const uint8_t* p) {
    typedef __attribute__((address_space(3))) v2i LdsV2i;  // p points into LDS
    return __builtin_amdgcn_ds_read_tr8_b64_v2i32((LdsV2i*)(const_cast<uint8_t*>(p)));
}
// XT fragment db of an LDS tile: lane l of group grp = l >> 4 holds dim
// 16*db + (l & 15) of points 16*grp + q, q = 0..15 (byte q).  Lane t = l & 15
// addresses dims 16*db + 8*(t & 1) .. +8 of point 16*grp + (t >> 1) (first
// read, bytes 0..7) and of the point 8 rows further (second read, bytes 8..15);
// the transposing read hands lane i of the group dim 16*db + i of those points.
__device__ inline v4i loadXt(const uint8_t* t, uint32_t lane, int db) {
    const uint32_t tt = lane & 15;
    const uint8_t* base = t + (16 * (lane >> 4) + (tt >> 1)) * kRowPitch + 16 * db + 8 * (tt & 1);
    const v2i lo = readTr8(base), hi = readTr8(base + 8 * kRowPitch);
    return v4i{lo.x, lo.y, hi.x, hi.y};
}

// Center B fragments of `crow` (center `col`'s row) and ||c - 128||^2 of center col.
__device__ inline int centerOperands(const uint8_t* crow, uint32_t lane, v4i& b0, v4i& b1) {
    const uint32_t grp = lane >> 4;
    b0 = toSigned(ld16(crow + grp * 16));
    b1 = toSigned(ld16(crow + 64 + grp * 16));
    int ns = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        ns = __builtin_amdgcn_sdot4(b0[e], b0[e], ns, false);
        ns = __builtin_amdgcn_sdot4(b1[e], b1[e], ns, false);
    }
    ns += __shfl_xor(ns, 16, 64);
    ns += __shfl_xor(ns, 32, 64);
    return ns;
}

// Argmin keys ((||c'||^2 - 2 x'.c' + 2^25) << 4) | center: unique per center,
// the smallest is the nearest center with ties to the lowest index, and
// |2 x'.c'| < 2^23, ||c'||^2 <= 2^21 keep them in [0, 2^30).  base[i] is the
// part without x' of center 4*(lane >> 4) + i; a key is base - (x'.c' << 5).
struct CenterKeys {
    v4i b0, b1;         // A operand of the distance MFMA: lane l holds center l & 15
    uint32_t base[4];
};
__device__ inline void centerKeys(const uint8_t* cent, uint32_t lane, CenterKeys& ck) {
    const int ns = centerOperands(cent + (lane & 15) * kDims, lane, ck.b0, ck.b1);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t c = 4 * (lane >> 4) + i;
        ck.base[i] = (uint32_t(__shfl(ns, int(c), 64) + (1 << 25)) << 4) | c;
    }
}

// 0x01 in each byte of x that is zero, else 0x00 (exact: no carries cross bytes).
__device__ inline uint32_t zeroBytes(uint32_t x) {
    const uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return (~(t | x) & 0x80808080u) >> 7;
}

// One tile of cnt (<= 64) rows: distances by I8 MFMA with the centers as the
// A operand (lane l then holds centers 4*(l>>4) + i of point 16*rb + (l&15)),
// argmin (ties to the lowest center), the new assignments compared with `prev`
// (lane r: previous assignment of row r), and the one-hot sums of x' added
// into acc with XT from the same tile in LDS (`tile`).  sAsgWave: the wave's
// 64 bytes of LDS.  Returns the new assignment of row `lane` (meaningful for
// lane < cnt).
__device__ __forceinline__ uint32_t assignTileMfma(const v4i (&a)[4][2], const uint8_t* tile, uint32_t cnt,
                                                   uint32_t prev, const CenterKeys& ck, uint8_t* sAsgWave,
                                                   uint32_t lane, v4i (&acc)[8], uint32_t& myCnt,
                                                   uint32_t& myChanged) {
    const uint32_t grp = lane >> 4;
    const v4i zero = {0, 0, 0, 0};
    uint32_t best[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        v4i d = __builtin_amdgcn_mfma_i32_16x16x64_i8(ck.b0, a[rb][0], zero, 0, 0, 0);
        d = __builtin_amdgcn_mfma_i32_16x16x64_i8(ck.b1, a[rb][1], d, 0, 0, 0);
        const uint32_t m01 = min(ck.base[0] - (uint32_t(d[0]) << 5), ck.base[1] - (uint32_t(d[1]) << 5));
        const uint32_t m23 = min(ck.base[2] - (uint32_t(d[2]) << 5), ck.base[3] - (uint32_t(d[3]) << 5));
        best[rb] = minAcrossGroupsU(min(m01, m23));
    }
    // row `lane` is point 16*grp + (lane & 15): the block rb = grp
    const uint32_t mine = (grp == 0 ? best[0] : grp == 1 ? best[1] : grp == 2 ? best[2] : best[3]) & 15u;
    const bool valid = lane < cnt;
    if (valid) myChanged += prev != mine;
    waveSync();  // the previous tile's reads of sAsgWave are done
    sAsgWave[lane] = valid ? uint8_t(mine) : uint8_t(0xFF);  // padding rows match no center
    waveSync();
    // one-hot operand: lane l, byte q = [point 16*grp + q belongs to center l & 15]
    const v4i w = *reinterpret_cast<const v4i*>(sAsgWave + 16 * grp);
    const uint32_t rep = (lane & 15) * 0x01010101u;
    v4i oh;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        oh[q] = int(zeroBytes(uint32_t(w[q]) ^ rep));
        myCnt = __builtin_amdgcn_udot4(uint32_t(oh[q]), 0x01010101u, myCnt, false);
    }
#pragma unroll
    for (int db = 0; db < 8; db++) acc[db] = __builtin_amdgcn_mfma_i32_16x16x64_i8(oh, loadXt(tile, lane, db), acc[db], 0, 0, 0);
    return mine;
}

__global__ __launch_bounds__(256) void k_assignU8Mfma(KmLevel k) {
    __shared__ __align__(16) uint8_t sAsg[kWaves][64];          // a tile's new assignments, one byte per row
    __shared__ __align__(16) uint8_t sTile[kWaves][kTileLds];  // the tile, row-major, for the XT reads
    // merge slots: a task's chunks on neighbouring waves of this workgroup add
    // their sums, counts and #changed here and leave as one partial
    __shared__ __align__(16) int32_t sMergeSums[kMergeSlots][16 * kDims];
    __shared__ uint32_t sMergeCnt[kMergeSlots][17];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    const v4i zero = {0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < kMergeSlots * 16 * kDims / 4; i += 256)
        reinterpret_cast<v4i*>(&sMergeSums[0][0])[i] = zero;
    if (threadIdx.x < kMergeSlots * 17) (&sMergeCnt[0][0])[threadIdx.x] = 0u;
    __syncthreads();
    // one wave per resident slot, each with an equal range of tiles cut into
    // per-task chunks on the host (partitionKmTiles): no wave waits for a
    // second round and the loads balance to within one tile
    const uint32_t w = blockIdx.x * kWaves + wave;
    const uint32_t chBeg = w < k.nbWaves ? k.waveChunkBeg[w] : 0u;
    const uint32_t chEnd = w < k.nbWaves ? k.waveChunkBeg[w + 1] : 0u;
    // Every other pass walks the wave's range backwards (k.reverse): a pass then
    // starts on the tiles the previous pass read last, which the memory-side cache
    // (MALL) still holds when a level's operand is about its size.  The sums are
    // exact integers, so the order changes nothing else.
    const bool rev = k.reverse != 0;
    auto advance = [&](uint32_t c0) -> uint32_t {  // next active chunk in this pass's order (kIdxInvalid: none)
        if (!rev) {
            uint32_t c = c0 == kIdxInvalid ? chBeg : c0 + 1;
            while (c < chEnd && !k.active[k.chunks[c].task]) c++;
            return c < chEnd ? c : kIdxInvalid;
        }
        uint32_t c = c0 == kIdxInvalid ? chEnd : c0;
        while (c > chBeg) {
            c--;
            if (k.active[k.chunks[c].task]) return c;
        }
        return kIdxInvalid;
    };
    auto firstTile = [&](uint32_t c0) {  // a chunk's first tile in this pass's order (chunks are never empty)
        const KmChunk& c = k.chunks[c0];
        return rev ? c.tileEnd - 1 : c.tileBeg;
    };
    uint32_t ch = advance(kIdxInvalid);
    // the first tile of the wave's next chunk is requested while the last tile of
    // the current chunk computes
    v4i a[4][2];
    KmTile t{};
    uint32_t prev = 0;
    if (ch != kIdxInvalid) {
        const uint32_t tb = firstTile(ch);
        t = k.tiles[tb];
        loadTileA(k.XA + uint64_t(tb) * 8192, lane, a);
        prev = lane < t.cnt ? k.assign[t.row + lane] : 0u;  // lane r: previous assignment of row r
    }
    while (ch != kIdxInvalid) {
        const KmChunk c = k.chunks[ch];
        const uint32_t chNext = advance(ch);
        CenterKeys ck;
        centerKeys(k.centersU8 + uint64_t(c.task) * 16 * kDims, lane, ck);
        v4i acc[8];
#pragma unroll
        for (int db = 0; db < 8; db++) acc[db] = zero;
        uint32_t myCnt = 0, myChanged = 0;
        const uint32_t nT = c.tileEnd - c.tileBeg;
        for (uint32_t step = 0; step < nT; step++) {
            const uint32_t ti = rev ? c.tileEnd - 1 - step : c.tileBeg + step;
            // next tile's fragments and previous assignments (this chunk's, else the next
            // chunk's first), in flight while this one computes
            v4i an[4][2];
            KmTile tn = t;
            uint32_t prevN = 0;
            const uint32_t tiN = step + 1 < nT ? (rev ? ti - 1 : ti + 1)
                                                : (chNext != kIdxInvalid ? firstTile(chNext) : kIdxInvalid);
            if (tiN != kIdxInvalid) {
                tn = k.tiles[tiN];
                loadTileA(k.XA + uint64_t(tiN) * 8192, lane, an);
                prevN = lane < tn.cnt ? k.assign[tn.row + lane] : 0u;
            }
            waveSync();  // the previous tile's transposing reads are done
            storeTileLds(sTile[wave], lane, a);
            waveSync();
            const uint32_t mine =
                assignTileMfma(a, sTile[wave], t.cnt, prev, ck, sAsg[wave], lane, acc, myCnt, myChanged);
            if (lane < t.cnt) k.assign[t.row + lane] = uint8_t(mine);
            if (tiN != kIdxInvalid) {
#pragma unroll
                for (int rb = 0; rb < 4; rb++) {
                    a[rb][0] = an[rb][0];
                    a[rb][1] = an[rb][1];
                }
            }
            t = tn;
            prev = prevN;
        }
        myCnt += uint32_t(__shfl_xor(int(myCnt), 16, 64));
        myCnt += uint32_t(__shfl_xor(int(myCnt), 32, 64));
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) myChanged += uint32_t(__shfl_xor(int(myChanged), m, 64));
        if (c.slot == kNoSlot) {  // the chunk's own partial
            int32_t* ps = k.partSums + uint64_t(c.part) * 16 * kDims;
#pragma unroll
            for (int db = 0; db < 8; db++)
#pragma unroll
                for (int i = 0; i < 4; i++) ps[(4 * grp + i) * kDims + 16 * db + col] = acc[db][i];
            if (lane < 16) k.partCnt[c.part * 16 + lane] = myCnt;
            if (lane == 0) k.partChanged[c.part] = myChanged;
        } else {  // shared with a neighbouring wave: add into the workgroup's slot
            int32_t* ms = sMergeSums[c.slot];
#pragma unroll
            for (int db = 0; db < 8; db++)
#pragma unroll
                for (int i = 0; i < 4; i++) atomicAdd(&ms[(4 * grp + i) * kDims + 16 * db + col], acc[db][i]);
            if (lane < 16) atomicAdd(&sMergeCnt[c.slot][lane], myCnt);
            if (lane == 0) atomicAdd(&sMergeCnt[c.slot][16], myChanged);
        }
        ch = chNext;
    }
    __syncthreads();
    // the merged partials leave once (a slot of a converged task was never written)
    for (uint32_t sl = 0; sl < kMergeSlots; sl++) {
        const KmSlot slot = k.blockSlots[blockIdx.x * kMergeSlots + sl];
        if (slot.part == kNoSlot || !k.active[slot.task]) continue;
        v4i* ps = reinterpret_cast<v4i*>(k.partSums + uint64_t(slot.part) * 16 * kDims);
        for (uint32_t i = threadIdx.x; i < 16 * kDims / 4; i += 256) ps[i] = reinterpret_cast<const v4i*>(sMergeSums[sl])[i];
        if (threadIdx.x < 16) k.partCnt[slot.part * 16 + threadIdx.x] = sMergeCnt[sl][threadIdx.x];
        if (threadIdx.x == 16) k.partChanged[slot.part] = sMergeCnt[sl][16];
    }
}

// uint8(roundf(float(s) / float(c))), the update rule (kmeans.cu:367), for
// 0 < c < 2^16 and 0 <= s <= 255 c: the quotient q = s / c lies in [0, 255],
// where float's half-ulp is at most 2^-17, and unless q is exactly k + 1/2 it
// is at least 1/(2c) > 2^-17 away from every k + 1/2, so the float division
// never rounds q onto or across a half and roundf gives floor(q + 1/2) =
// floor((2s + c) / (2c)).  That is computed with a reciprocal estimate and one
// integer correction (the estimate is within one of the floor).
__device__ inline uint32_t roundedMean(uint32_t s, uint32_t c) {
    const uint32_t num = 2 * s + c, den = 2 * c;
    uint32_t q = uint32_t(float(num) * __builtin_amdgcn_rcpf(float(den)));
    if (q * den > num) q--;
    else if ((q + 1) * den <= num) q++;
    return q;
}

// Persistent Lloyd loops, one workgroup of W waves per listed task.  The first
// nbStream listed tasks (up to kStreamTaskTiles 64-row tiles) stream: every pass
// reads their tiles again from the gathered operand (XA), which the first pass
// leaves in L2 and the memory-side cache, and each wave puts its current tile in
// LDS for the transposing reads of the sums, as k_assignU8Mfma does.  The others
// (at most maxTiles tiles) have their tiles read from HBM once into LDS,
// row-major (dynamic: maxTiles x 9 KB, then one assignment byte per row), and
// every pass takes both MFMA operands from LDS.  One launch holds both kinds,
// largest tasks first, so short tasks fill the CUs while long ones run.  Per
// pass the waves take the tiles round robin (the next tiles in flight while one
// computes), their one-hot sums and counts meet in LDS (int32 atomics), and the
// same rule as k_updateU8Part applies: stop when no assignment changed, else
// move the centers to round(float(sum)/float(count)).  Two barriers per pass:
// each thread clears the sums it has just read, and the counters alternate
// between two buffers.  Results equal those of the one-launch-per-pass path for
// any pass cap.  At the end the assignments and centers go to HBM and the task
// is marked inactive with its pass count.
template <int W>
__global__ __launch_bounds__(64 * W) void k_lloydTasks(KmLevel k, const uint32_t* __restrict__ taskList,
                                                      const uint32_t* __restrict__ taskTileBeg, uint32_t maxTiles,
                                                      uint32_t maxIters, uint32_t nbStream) {
    constexpr uint32_t kThreads = 64 * W;
    extern __shared__ __align__(16) uint8_t dynLds[];
    __shared__ __align__(16) uint8_t sNew[W][64];
    // sums of x - 128, center rows kSumPitch words apart: the 16-lane groups of an
    // atomic add (rows 4*grp + i) fall on disjoint banks
    constexpr uint32_t kSumPitch = kDims + 4;
    __shared__ __align__(16) int32_t sSum[16 * kSumPitch];
    __shared__ uint32_t sCnt[2][16], sChanged[2];
    __shared__ __align__(16) uint8_t sCent[16 * kDims];
    const uint32_t g = taskList[blockIdx.x], tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    const uint32_t n = k.taskCnt[g], row0 = k.taskBeg[g], nt = (n + 63) / 64;
    const bool streamed = blockIdx.x < nbStream;  // block-uniform
    if (nt == 0 || nt > (streamed ? kStreamTaskTiles : maxTiles)) return;  // the host lists only tasks that fit
    // streamed: [W] wave tiles, then the assignments; resident: [nt] tiles, then the assignments
    uint8_t* sA = dynLds;
    uint8_t* sAsg = dynLds + uint64_t(streamed ? W : maxTiles) * kTileLds;
    const int8_t* xa = k.XA + uint64_t(taskTileBeg[g]) * 8192;
    if (!streamed) {
        const uint4* src = reinterpret_cast<const uint4*>(xa);
        // XA fragment (tile, rb, kb, lane) -> its rows in LDS; 8 loads in flight per thread
        constexpr uint32_t kLoadBatch = 8;
        for (uint32_t i0 = tid; i0 < nt * 512; i0 += kLoadBatch * kThreads) {
            uint4 v[kLoadBatch];
#pragma unroll
            for (uint32_t j = 0; j < kLoadBatch; j++) {
                const uint32_t i = i0 + j * kThreads;
                v[j] = i < nt * 512 ? src[i] : uint4{0, 0, 0, 0};
            }
#pragma unroll
            for (uint32_t j = 0; j < kLoadBatch; j++) {
                const uint32_t i = i0 + j * kThreads, c = i & 511;
                if (i < nt * 512)
                    *reinterpret_cast<uint4*>(sA + (i >> 9) * kTileLds + fragOffset(c >> 7, (c >> 6) & 1, c & 63)) = v[j];
            }
        }
    }
    for (uint32_t i = tid; i < nt * 64; i += kThreads) sAsg[i] = 0xFF;  // no previous assignment
    uint8_t* cent = k.centersU8 + uint64_t(g) * 16 * kDims;
    if (tid < 128) reinterpret_cast<uint4*>(sCent)[tid] = reinterpret_cast<const uint4*>(cent)[tid];
    for (uint32_t e = tid; e < 16 * kSumPitch; e += kThreads) sSum[e] = 0;
    if (tid < 32) sCnt[tid >> 4][tid & 15] = 0;
    if (tid < 2) sChanged[tid] = 0;
    const v4i zero = {0, 0, 0, 0};
    __syncthreads();
    uint32_t pass = 0;
    while (pass < maxIters) {
        const uint32_t par = pass & 1;
        CenterKeys ck;
        centerKeys(sCent, lane, ck);
        v4i acc[8];
#pragma unroll
        for (int db = 0; db < 8; db++) acc[db] = zero;
        uint32_t myCnt = 0, myChanged = 0;
        auto tileDone = [&](uint32_t i, const v4i (&a)[4][2], const uint8_t* tileLds) {
            const uint32_t cnt = min(64u, n - 64 * i);
            const uint32_t prev = sAsg[i * 64 + lane];
            const uint32_t mine = assignTileMfma(a, tileLds, cnt, prev, ck, sNew[wave], lane, acc, myCnt, myChanged);
            if (lane < cnt) sAsg[i * 64 + lane] = uint8_t(mine);
        };
        if (streamed && wave < nt) {
            // the wave's tiles wave, wave + W, ... through a ring of three (two in flight while
            // one computes), unrolled by three so that every slot index is a constant; a request
            // past the task's end repeats its last tile, so no load sits under a branch
            v4i ring[3][4][2];
            uint8_t* myTile = sA + wave * kTileLds;
            auto request = [&](auto S, uint32_t i) { loadTileA(xa + uint64_t(min(i, nt - 1)) * 8192, lane, ring[S]); };
            auto step = [&](auto S, uint32_t i) -> bool {
                if (i >= nt) return false;
                request(std::integral_constant<int, (S + 2) % 3>{}, i + 2 * W);
                waveSync();  // the previous tile's transposing reads are done
                storeTileLds(myTile, lane, ring[S]);
                waveSync();
                tileDone(i, ring[S], myTile);
                return true;
            };
            request(std::integral_constant<int, 0>{}, wave);
            request(std::integral_constant<int, 1>{}, wave + W);
            for (uint32_t i = wave;; i += 3 * W) {
                if (!step(std::integral_constant<int, 0>{}, i)) break;
                if (!step(std::integral_constant<int, 1>{}, i + W)) break;
                if (!step(std::integral_constant<int, 2>{}, i + 2 * W)) break;
            }
        } else if (wave < nt) {
            v4i a[4][2];
            loadTileLds(sA + wave * kTileLds, lane, a);
            for (uint32_t i = wave; i < nt; i += W) {
                v4i an[4][2];
                const bool more = i + W < nt;
                if (more) loadTileLds(sA + (i + W) * kTileLds, lane, an);
                tileDone(i, a, sA + i * kTileLds);
                if (more) {
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) {
                        a[rb][0] = an[rb][0];
                        a[rb][1] = an[rb][1];
                    }
                }
            }
        }
        if (wave < nt) {  // the waves' partials meet in LDS (a wave without a tile has none)
#pragma unroll
            for (int db = 0; db < 8; db++)
#pragma unroll
                for (int i = 0; i < 4; i++) atomicAdd(&sSum[(4 * grp + i) * kSumPitch + 16 * db + col], acc[db][i]);
            myCnt += uint32_t(__shfl_xor(int(myCnt), 16, 64));
            myCnt += uint32_t(__shfl_xor(int(myCnt), 32, 64));
            if (lane < 16) atomicAdd(&sCnt[par][lane], myCnt);
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) myChanged += uint32_t(__shfl_xor(int(myChanged), m, 64));
            if (lane == 0) atomicAdd(&sChanged[par], myChanged);
        }
        __syncthreads();  // every partial is in
        pass++;
        const bool stop = sChanged[par] == 0;  // converged on this pass: no update (uniform)
        // four dims of one center per thread and step
#pragma unroll 2
        for (uint32_t e4 = tid; e4 < 16 * kDims / 4; e4 += kThreads) {
            const uint32_t row = e4 / (kDims / 4), d4 = e4 % (kDims / 4);
            v4i* sp = reinterpret_cast<v4i*>(sSum + row * kSumPitch) + d4;
            const v4i sv = *sp;
            *sp = zero;  // this thread's own elements: cleared for the next pass
            const uint32_t c = sCnt[par][row];
            if (!stop && c) {
                uint32_t packed = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) packed |= roundedMean(uint32_t(sv[j] + 128 * int32_t(c)), c) << (8 * j);
                reinterpret_cast<uint32_t*>(sCent)[e4] = packed;
            }
        }
        if (tid < 16) sCnt[par ^ 1][tid] = 0;  // counters of the next pass
        if (tid == 0) sChanged[par ^ 1] = 0;
        __syncthreads();  // new centers and cleared accumulators are visible
        if (stop) break;
    }
    for (uint32_t i = tid; i < n; i += kThreads) k.assign[row0 + i] = sAsg[i];
    if (tid < 128) reinterpret_cast<uint4*>(cent)[tid] = reinterpret_cast<const uint4*>(sCent)[tid];
    if (tid == 0) {
        k.passes[g] = pass;
        k.active[g] = 0;
    }
}

// Reduce items: (task, partial range); adds the partials into the task totals.
// Grid (item, 8): block y owns elements [256y, 256y + 256) of the 16 x 128 sums.
__global__ __launch_bounds__(256) void k_reduceChunks(KmLevel k, const KmPartRange* __restrict__ red, uint32_t nbRed) {
    const KmPartRange r = red[blockIdx.x];
    const uint32_t g = r.task, tid = threadIdx.x, e = blockIdx.y * 256 + tid, center = e / kDims;
    if (!k.active[g]) return;
    if (blockIdx.y == 0 && tid <= 16) {
        uint32_t c = 0;
        for (uint32_t x = r.partBeg; x < r.partEnd; x++) c += tid < 16 ? k.partCnt[x * 16 + tid] : k.partChanged[x];
        if (c) atomicAdd(tid < 16 ? &k.counts[g * 16 + tid] : &k.changed[g], c);
    }
    long long s = 0;
    uint32_t cnt = 0;
#pragma unroll 16
    for (uint32_t x = r.partBeg; x < r.partEnd; x++) {
        s += k.partSums[uint64_t(x) * 16 * kDims + e];
        cnt += k.partCnt[x * 16 + center];
    }
    s += 128ll * cnt;  // partials are sums of x - 128
    if (s) atomicAdd(reinterpret_cast<unsigned long long*>(k.sums) + uint64_t(g) * 16 * kDims + e, (unsigned long long)s);
}

// ----------------------------------------------- assign: uint8, plain ----
// 16 lanes per point (lane j: center j), 4 points per step.
__global__ __launch_bounds__(256) void k_assignU8Plain(KmLevel k) {
    __shared__ uint32_t sSums[kWaves][16 * kDims];
    __shared__ uint32_t sCnt[kWaves][16];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t j = lane & 15, sub = lane >> 4;
    WaveAcc<16> acc{sSums[wave], sCnt[wave]};
    acc.clear(lane);
    uint32_t beg, end;
    waveRange(k.nbTiles, beg, end);
    uint32_t curTask = kIdxInvalid, changedLocal = 0;
    uint32_t cw[32];
    int nu = 0;
    for (uint32_t ti = beg; ti < end; ti++) {
        const KmTile t = k.tiles[ti];
        if (!k.active[t.task]) continue;
        if (t.task != curTask) {
            if (curTask != kIdxInvalid) acc.flush(lane, curTask, changedLocal, k);
            changedLocal = 0;
            curTask = t.task;
            const uint32_t* c = reinterpret_cast<const uint32_t*>(k.centersU8 + (uint64_t(curTask) * 16 + j) * kDims);
            uint32_t s = 0;
#pragma unroll
            for (int w = 0; w < 32; w++) {
                cw[w] = c[w];
                s = __builtin_amdgcn_udot4(cw[w], cw[w], s, false);
            }
            nu = int(s);
        }
        uint32_t myA = 0;
        for (uint32_t r0 = 0; r0 < t.cnt; r0 += 4) {
            const uint32_t r = r0 + sub;
            uint32_t dot = 0;
            if (r < t.cnt) {
                const uint32_t* x = reinterpret_cast<const uint32_t*>(k.X + uint64_t(t.row + r) * kDims);
#pragma unroll
                for (int w = 0; w < 32; w++) dot = __builtin_amdgcn_udot4(x[w], cw[w], dot, false);
            }
            const uint32_t key = min16u((uint32_t(nu - 2 * int(dot) + (1 << 25)) << 4) | j);
            const uint32_t got = uint32_t(__shfl(int(key), int(16 * (lane & 3)), 64));
            if ((lane >> 2) == (r0 >> 2)) myA = got & 15u;
        }
        changedLocal += tileEpilogue<16>(k, t, lane, myA, acc);
    }
    if (curTask != kIdxInvalid) acc.flush(lane, curTask, changedLocal, k);
}

// ----------------------------------------------- assign: float, plain ----
// B lanes per point (lane j: center j, its own fma chain over the 128 dims in
// order), 64/B points per step.  Centers are read from global (L2 resident).
template <int B>
__global__ __launch_bounds__(256) void k_assignF32Plain(KmLevel k) {
    __shared__ uint32_t sSums[kWaves][B * kDims];
    __shared__ uint32_t sCnt[kWaves][B];
    constexpr uint32_t PPS = 64 / B;  // points per step
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t j = lane % B, sub = lane / B;
    WaveAcc<B> acc{sSums[wave], sCnt[wave]};
    acc.clear(lane);
    uint32_t beg, end;
    waveRange(k.nbTiles, beg, end);
    uint32_t curTask = kIdxInvalid, changedLocal = 0;
    for (uint32_t ti = beg; ti < end; ti++) {
        const KmTile t = k.tiles[ti];
        if (!k.active[t.task]) continue;
        if (t.task != curTask) {
            if (curTask != kIdxInvalid) acc.flush(lane, curTask, changedLocal, k);
            changedLocal = 0;
            curTask = t.task;
        }
        const float4* cr = reinterpret_cast<const float4*>(k.centersF32 + (uint64_t(curTask) * B + j) * kDims);
        uint32_t myA = 0;
        for (uint32_t r0 = 0; r0 < t.cnt; r0 += PPS) {
            const uint32_t r = r0 + sub;
            float a = 0.0f;
            if (r < t.cnt) {
                const uint4* x = reinterpret_cast<const uint4*>(k.X + uint64_t(t.row + r) * kDims);
#pragma unroll 2
                for (int w4 = 0; w4 < 8; w4++) {
                    const uint4 xv = x[w4];
                    const uint32_t xw[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const float4 c = cr[w4 * 4 + q];
                        float d = float(xw[q] & 0xFFu) - c.x;
                        a = fmaf(d, d, a);
                        d = float((xw[q] >> 8) & 0xFFu) - c.y;
                        a = fmaf(d, d, a);
                        d = float((xw[q] >> 16) & 0xFFu) - c.z;
                        a = fmaf(d, d, a);
                        d = float(xw[q] >> 24) - c.w;
                        a = fmaf(d, d, a);
                    }
                }
            }
            float bd = a;
            uint32_t bj = j;
#pragma unroll
            for (int m = 1; m < B; m <<= 1) {
                const float od = __shfl_xor(bd, m, B);
                const uint32_t oj = uint32_t(__shfl_xor(int(bj), m, B));
                if (od < bd || (od == bd && oj < bj)) {
                    bd = od;
                    bj = oj;
                }
            }
            // point r0 + s lives in lanes [s*B, s*B+B)
            const uint32_t got = uint32_t(__shfl(int(bj), int(B * (lane % PPS)), 64));
            if (lane / PPS == r0 / PPS) myA = got;
        }
        changedLocal += tileEpilogue<B>(k, t, lane, myA, acc);
    }
    if (curTask != kIdxInvalid) acc.flush(lane, curTask, changedLocal, k);
}

// ------------------------------------------- assign: float, MFMA filter --
// Same results as k_assignF32Plain (the plain FP32 chain, ties to the lowest
// center): per 16-row block, the rows (x - 128, the bytes with the top bit
// flipped) against the task's centers split into three int8 terms (this pass's
// launchPrepF32Split) by 3 x 2 x B/16 I8 MFMAs with the centers as the A
// operand (each lane then owns one row), the terms combined exactly as the
// float-tree lookup does (combineSplit), the certified one-threshold-per-row
// candidate filter (filterSplit) and the plain chain for the tile's near ties,
// resolved together (resolveTies).  Work split and epilogue as the plain kernel:
// each wave a contiguous tile range, the task operands reloaded when the task
// changes, exact sums in LDS; the next block's rows are in flight while one
// computes.
template <int B>
__global__ __launch_bounds__(256) void k_assignF32Mfma(KmLevel k) {
    constexpr int CB = B / 16;
    __shared__ uint32_t sSums[kWaves][B * kDims];
    __shared__ uint32_t sCnt[kWaves][B];
    __shared__ uint8_t sBest[kWaves][64];  // a tile's assignments, one byte per row
    __shared__ TieQueue tqAll[kWaves];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t col = lane & 15, grp = lane >> 4;
    WaveAcc<B> acc{sSums[wave], sCnt[wave]};
    acc.clear(lane);
    TieQueue& tq = tqAll[wave];
    uint32_t beg, end;
    waveRange(k.nbTiles, beg, end);
    uint32_t curTask = kIdxInvalid, changedLocal = 0, ties = 0, nTie = 0;
    bool curStartsHere = false;  // the current task's first tile is in this wave's range
    float cn[CB][4], e1m = 0.0f, e2m = 0.0f;
    for (uint32_t ti = beg; ti < end; ti++) {
        const KmTile t = k.tiles[ti];
        if (!k.active[t.task]) continue;
        if (t.task != curTask) {
            // a task's tiles are contiguous: the one ending here ended inside the range
            if (curTask != kIdxInvalid) acc.flushTask(lane, curTask, changedLocal, k, curStartsHere);
            curStartsHere = ti > beg || ti == 0 || k.tiles[ti - 1].task != t.task;
            changedLocal = 0;
            curTask = t.task;
            const uint64_t childBase = uint64_t(curTask) * B;
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                const float4 c4 = *reinterpret_cast<const float4*>(k.fCn2 + childBase + cb * 16 + 4 * grp);
                cn[cb][0] = c4.x;
                cn[cb][1] = c4.y;
                cn[cb][2] = c4.z;
                cn[cb][3] = c4.w;
            }
            e1m = k.fE1[childBase + lane % B];
            e2m = k.fE2[childBase + lane % B];
#pragma unroll
            for (int d = B / 2; d >= 1; d >>= 1) {
                e1m = fmaxf(e1m, __shfl_xor(e1m, d, 64));
                e2m = fmaxf(e2m, __shfl_xor(e2m, d, 64));
            }
        }
        // the task's split centers, reloaded per tile (from L2 after the first): live only
        // while the tile's blocks compute, not during its near ties' resolution
        v4i ca[3][CB][2];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int8_t* crow = k.fSplit + (uint64_t(curTask) * B + cb * 16 + col) * kDims + grp * 16;
#pragma unroll
            for (int tt = 0; tt < 3; tt++) {
                ca[tt][cb][0] = *reinterpret_cast<const v4i*>(crow + tt * k.fSplitStride);
                ca[tt][cb][1] = *reinterpret_cast<const v4i*>(crow + tt * k.fSplitStride + 64);
            }
        }
        const float* centers = k.centersF32 + uint64_t(curTask) * B * kDims;
        // Raw bytes: the sign flip happens at use, so the next block's loads stay in flight
        // while one computes.  A row past the tile's end reads the tile's first row; its
        // column of the products is discarded (filterSplit's rowValid) and its sums come
        // from the epilogue, which skips it.
        auto loadRows = [&](uint32_t rb, v4i& x0, v4i& x1) {
            const uint32_t r = rb * 16 + col;
            const uint8_t* x = k.X + uint64_t(t.row + (r < t.cnt ? r : 0u)) * kDims;
            x0 = ld16(x + grp * 16);
            x1 = ld16(x + 64 + grp * 16);
        };
        v4i n0, n1;
        loadRows(0, n0, n1);
        for (uint32_t rb = 0; rb * 16 < t.cnt; rb++) {
            const uint32_t r = rb * 16 + col;
            const bool valid = r < t.cnt;
            const v4i x0 = toSigned(n0), x1 = toSigned(n1);
            loadRows(rb + 1, n0, n1);  // unconditional (past the end: the tile's first row), so no copy at a join waits for it
            // ||x - 128||^2 of row (lane & 15): an exact integer, summed over the row's four lanes
            int xs = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                xs = __builtin_amdgcn_sdot4(x0[e], x0[e], xs, false);
                xs = __builtin_amdgcn_sdot4(x1[e], x1[e], xs, false);
            }
            xs = sumAcrossGroups(xs);
            v4f d[CB];
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                v4i p[3];
#pragma unroll
                for (int tt = 0; tt < 3; tt++) {
                    p[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ca[tt][cb][0], x0, v4i{0, 0, 0, 0}, 0, 0, 0);
                    p[tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ca[tt][cb][1], x1, p[tt], 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < 4; i++) d[cb][i] = combineSplit(p[0][i], p[1][i], p[2][i]);
            }
            const uint32_t mask = filterSplit<CB>(d, float(xs), cn, e1m, e2m, lane, valid);
            const bool tie = (mask & (mask - 1u)) != 0;  // only valid rows can have several candidates
            if (grp == 0 && !tie) sBest[wave][r] = uint8_t(__builtin_ctz(mask));
            ties += tiePush(tq, nTie, lane, tie, t.row + r, r, mask);  // a tile's <= 64 rows fit the queue
        }
        if (nTie)  // the tile's near ties: the plain chain over their candidates
            resolveTies(
                tq, nTie, lane, [&](uint32_t row) { return k.X + uint64_t(row) * kDims; },
                [&](uint32_t c) { return centers + uint64_t(c) * kDims; },
                [&](uint32_t r, uint32_t best) { sBest[wave][r] = uint8_t(best); });
        waveSync();
        const uint32_t myA = sBest[wave][lane];  // row `lane` (meaningful for lane < cnt)
        waveSync();
        changedLocal += tileEpilogue<B>(k, t, lane, myA, acc);
    }
    if (curTask != kIdxInvalid)
        acc.flushTask(lane, curTask, changedLocal, k,
                      curStartsHere && (end == k.nbTiles || k.tiles[end].task != curTask));
    if (k.nearTies && ties && lane == 0) atomicAdd(k.nearTies, ties);  // ties: rows of this wave (uniform)
}

// ------------------------------------------------------------ update -----
template <int B>
__global__ __launch_bounds__(256) void k_update(KmLevel k) {
    const uint32_t g = blockIdx.x;
    if (!k.active[g]) return;
    __syncthreads();
    if (k.changed[g] == 0) {  // converged on this pass: no update, task stops
        if (threadIdx.x == 0) {
            k.active[g] = 0;
            k.passes[g]++;
        }
        return;
    }
    unsigned long long* sums = reinterpret_cast<unsigned long long*>(k.sums) + uint64_t(g) * B * kDims;
    for (uint32_t e = threadIdx.x; e < B * kDims; e += blockDim.x) {
        const uint32_t c = k.counts[g * B + e / kDims];
        if (c) {
            const float q = float((long long)sums[e]) / float(c);
            if (k.centersU8) k.centersU8[uint64_t(g) * B * kDims + e] = uint8_t(roundf(q));
            else k.centersF32[uint64_t(g) * B * kDims + e] = q;
        }
        sums[e] = 0;
    }
    __syncthreads();
    if (threadIdx.x < B) k.counts[g * B + threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        k.changed[g] = 0;
        k.passes[g]++;
        atomicAdd(k.nbActive, 1u);
    }
}

// Update of the MFMA path, grid (task, 8): block y owns centers 2y, 2y+1, i.e.
// elements [256y, 256y + 256) of the task's 16 x 128 sums.  Tasks with <= 32
// partials reduce them here (no atomics); larger tasks were reduced into the
// totals by k_reduceChunks.  #changed is summed from the partials' counts by
// every block, so no block depends on another's writes.
__global__ __launch_bounds__(256) void k_updateU8Part(KmLevel k) {
    __shared__ uint32_t sWave[kWaves], sCnt[2];
    const uint32_t g = k.updTasks ? k.updTasks[blockIdx.x] : blockIdx.x;
    const uint32_t y = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (!k.active[g]) return;
    const uint32_t cb = k.taskPartBeg[g], nc = k.taskPartCnt[g];
    const bool direct = nc <= kDirectParts;
    uint32_t ch = 0;
    for (uint32_t c = tid; c < nc; c += 256) ch += k.partChanged[cb + c];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) ch += uint32_t(__shfl_xor(int(ch), m, 64));
    if (lane == 0) sWave[wave] = ch;
    if (tid < 2) {
        const uint32_t center = 2 * y + tid;
        uint32_t c = 0;
        if (direct) {
#pragma unroll 8
            for (uint32_t x = cb; x < cb + nc; x++) c += k.partCnt[x * 16 + center];
        } else {
            c = k.counts[g * 16 + center];
        }
        sCnt[tid] = c;
    }
    __syncthreads();
    if (sWave[0] + sWave[1] + sWave[2] + sWave[3] == 0) {  // converged on this pass: no update, task stops
        if (y == 0 && tid == 0) {
            k.active[g] = 0;
            k.passes[g]++;
        }
        return;
    }
    const uint32_t e = y * 256 + tid, c = sCnt[tid / kDims];
    long long sum;
    if (direct) {
        sum = 128ll * c;  // partials are sums of x - 128
#pragma unroll 8
        for (uint32_t x = cb; x < cb + nc; x++) sum += k.partSums[uint64_t(x) * 16 * kDims + e];
    } else {
        unsigned long long* sums = reinterpret_cast<unsigned long long*>(k.sums) + uint64_t(g) * 16 * kDims;
        sum = (long long)sums[e];
        sums[e] = 0;
    }
    if (c) k.centersU8[uint64_t(g) * 16 * kDims + e] = uint8_t(roundf(float(sum) / float(c)));
    if (!direct && tid < 2) k.counts[g * 16 + 2 * y + tid] = 0;  // read above by this thread only
    if (y == 0 && tid == 0) {
        k.passes[g]++;
        atomicAdd(k.nbActive, 1u);
    }
}

// --------------------------------------------------------- level plan ----
// The next level's tasks from a level's (task, child) histogram, on the device
// (the host derives the same from the downloaded histogram): taskCnt = hist,
// taskBeg / tileBeg = exclusive prefix sums of the rows and of the 64-row
// tiles, *nbTiles = the tile count; then the tile list, tasks in order, each
// task's rows cut into 64-row tiles.  One workgroup scans (nbTasks <=
// kPlanMaxTasks), one thread per tile writes the list.
__global__ __launch_bounds__(1024) void k_planLevel(const uint32_t* __restrict__ cnt, uint32_t n,
                                                   uint32_t* __restrict__ taskBeg, uint32_t* __restrict__ taskCnt,
                                                   uint32_t* __restrict__ tileBeg, uint32_t* __restrict__ nbTiles) {
    __shared__ uint32_t sRows[1024], sTiles[1024];
    const uint32_t tid = threadIdx.x, per = (n + 1023) / 1024;
    const uint32_t beg = min(n, tid * per), end = min(n, beg + per);
    uint32_t rows = 0, tiles = 0;
    for (uint32_t g = beg; g < end; g++) {
        const uint32_t c = cnt[g];
        rows += c;
        tiles += (c + 63) / 64;
    }
    sRows[tid] = rows;
    sTiles[tid] = tiles;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t r = tid >= d ? sRows[tid - d] : 0u, t = tid >= d ? sTiles[tid - d] : 0u;
        __syncthreads();
        sRows[tid] += r;
        sTiles[tid] += t;
        __syncthreads();
    }
    uint32_t r0 = sRows[tid] - rows, t0 = sTiles[tid] - tiles;
    for (uint32_t g = beg; g < end; g++) {
        const uint32_t c = cnt[g];
        taskCnt[g] = c;
        taskBeg[g] = r0;
        tileBeg[g] = t0;
        r0 += c;
        t0 += (c + 63) / 64;
    }
    if (tid == 1023) *nbTiles = sTiles[1023];
}

__global__ void k_planTiles(const uint32_t* __restrict__ taskBeg, const uint32_t* __restrict__ taskCnt,
                            const uint32_t* __restrict__ tileBeg, uint32_t n, const uint32_t* __restrict__ nbTiles,
                            KmTile* __restrict__ tiles, uint32_t bound) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bound || t >= *nbTiles) return;
    // the last task whose first tile is <= t: empty tasks just before it share its first tile
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tileBeg[mid] <= t) lo = mid;
        else hi = mid;
    }
    const uint32_t r = t - tileBeg[lo];
    tiles[t] = KmTile{lo, taskBeg[lo] + 64 * r, min(64u, taskCnt[lo] - 64 * r)};
}

// ------------------------------------------------------------ regroup ----
// Stable regroup of a level's rows by (task, child) (Vocabulary.cpp:83-98 sorts
// the same keys stably) as a counting sort over the level's tiles.  A
// workgroup owns kRgTiles consecutive tiles (tiles are in row order and tasks
// are contiguous, so a task's tiles form one run):
//   count    per tile and child the rows (ballots), scanned tile by tile inside
//            the workgroup and restarted at every task boundary; the tile's
//            offsets are stored and each task run's totals added to the
//            histogram (one atomic per run and child);
//   carry    the workgroups' trailing-run totals chained across workgroups;
//   scatter  row -> taskBeg[task] + rows of the task's lower children + rows of
//            its child in earlier tiles + rows of its child earlier in its tile.
constexpr uint32_t kRgTiles = 256;  // tiles per workgroup (one per thread of the scan)

struct RgWs {
    uint32_t* tileOff;   // [tile][b] rows of the child in earlier tiles of the task within the workgroup
    uint32_t* blkTot;    // [workgroup][b] rows of the workgroup's last task in the workgroup
    uint32_t* blkCarry;  // [workgroup][b] rows of the workgroup's first task in earlier workgroups
    uint32_t* blkHead;   // [workgroup] task of the first tile
    uint32_t* blkTail;   // [workgroup] task of the last tile
};

template <uint32_t B>
__global__ __launch_bounds__(256) void k_regroupCount(KmLevel k, RgWs ws, uint32_t* __restrict__ hist) {
    constexpr uint32_t kGroups = kRgTiles / B;  // scan groups of B tiles, one thread per (group, child)
    __shared__ uint32_t sCnt[kRgTiles][B + 1], sOff[kRgTiles][B + 1];
    __shared__ uint32_t sTask[kRgTiles + 1];
    __shared__ uint32_t sGrpTot[kGroups][B], sGrpCarry[kGroups][B];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t blk = blockIdx.x, t0 = blk * kRgTiles;
    const uint32_t nT = min(kRgTiles, k.nbTiles - t0);
    for (uint32_t j = 0; j < B; j++) sCnt[tid][j] = 0;
    // lane l: tile wave * 64 + l; then every row of the wave's tiles loaded before counting
    const uint32_t myTl = wave * 64 + lane;
    const KmTile mine = myTl < nT ? k.tiles[t0 + myTl] : KmTile{kIdxInvalid, 0u, 0u};
    sTask[myTl] = mine.task;
    uint32_t a[64];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        const uint32_t row = __builtin_amdgcn_readlane(mine.row, i), cnt = __builtin_amdgcn_readlane(mine.cnt, i);
        a[i] = lane < cnt ? k.assign[row + lane] : 0xFFu;
    }
    __syncthreads();  // sCnt cleared
    // a row whose assignment is not < b counts nowhere (the caller checks the total)
#pragma unroll
    for (int i = 0; i < 64; i++)
        if (a[i] < B) atomicAdd(&sCnt[wave * 64 + i][a[i]], 1u);
    if (tid == 0) sTask[kRgTiles] = kIdxInvalid;
    __syncthreads();
    const uint32_t j = tid % B, grp = tid / B, g0 = grp * B;
    uint32_t run = 0;
    for (uint32_t i = 0; i < B; i++) {
        const uint32_t tl = g0 + i;
        if (i > 0 && sTask[tl] != sTask[tl - 1]) run = 0;
        sOff[tl][j] = run;
        run += sCnt[tl][j];
    }
    sGrpTot[grp][j] = run;
    __syncthreads();
    if (tid < B) {  // thread j: carries between the groups
        uint32_t carry = 0;
        sGrpCarry[0][j] = 0;
        for (uint32_t q = 1; q < kGroups; q++) {
            const uint32_t head = sTask[q * B], prevTail = sTask[q * B - 1], prevHead = sTask[(q - 1) * B];
            carry = prevTail == head ? (prevHead == prevTail ? carry : 0u) + sGrpTot[q - 1][j] : 0u;
            sGrpCarry[q][j] = carry;
        }
    }
    __syncthreads();
    const uint32_t carry = sGrpCarry[grp][j], head = sTask[g0];
    for (uint32_t i = 0; i < B; i++) {
        const uint32_t tl = g0 + i, task = sTask[tl];
        if (task == kIdxInvalid) break;
        const uint32_t off = sOff[tl][j] + (task == head ? carry : 0u);  // the group's leading run gets the carry
        ws.tileOff[uint64_t(t0 + tl) * B + j] = off;
        if (sTask[tl + 1] != task) {  // the task's run in this workgroup ends here
            const uint32_t tot = off + sCnt[tl][j];
            if (tot) atomicAdd(&hist[uint64_t(task) * B + j], tot);
            if (tl + 1 == nT) ws.blkTot[uint64_t(blk) * B + j] = tot;
        }
    }
    if (tid == 0) {
        ws.blkHead[blk] = sTask[0];
        ws.blkTail[blk] = sTask[nT - 1];
    }
}

// One workgroup: blkCarry[w][j] = rows of child j of workgroup w's first task
// in workgroups before w (a task's run may span many workgroups).
template <uint32_t B>
__global__ __launch_bounds__(256) void k_regroupCarry(RgWs ws, uint32_t nbBlk) {
    __shared__ uint32_t sHead[256], sTail[256], sTot[256][B];
    const uint32_t tid = threadIdx.x;
    uint32_t carry = 0, prevHead = kIdxInvalid, prevTail = kIdxInvalid, prevTot = 0;
    for (uint32_t c0 = 0; c0 < nbBlk; c0 += 256) {
        const uint32_t n = min(256u, nbBlk - c0);
        __syncthreads();
        for (uint32_t e = tid; e < n; e += 256) {
            sHead[e] = ws.blkHead[c0 + e];
            sTail[e] = ws.blkTail[c0 + e];
        }
        for (uint32_t e = tid; e < n * B; e += 256) sTot[e / B][e % B] = ws.blkTot[uint64_t(c0) * B + e];
        __syncthreads();
        if (tid < B) {
            for (uint32_t i = 0; i < n; i++) {
                const uint32_t h = sHead[i];
                carry = prevTail == h ? (prevHead == prevTail ? carry : 0u) + prevTot : 0u;
                ws.blkCarry[uint64_t(c0 + i) * B + tid] = carry;
                prevHead = h;
                prevTail = sTail[i];
                prevTot = sTot[i][tid];
            }
        }
    }
}

constexpr uint32_t kScTiles = 4;  // tiles per wave of the scatter (kRgTiles is a multiple)

template <uint32_t B>
__global__ __launch_bounds__(256) void k_regroupScatter(KmLevel k, RgWs ws, const uint32_t* __restrict__ hist,
                                                       uint32_t nbRows, const uint32_t* __restrict__ orderIn,
                                                       uint32_t* __restrict__ orderOut) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t tBeg = __builtin_amdgcn_readfirstlane((blockIdx.x * kWaves + (threadIdx.x >> 6)) * kScTiles);
    if (tBeg >= k.nbTiles) return;
    const uint32_t blk = tBeg / kRgTiles;  // the wave's tiles share a workgroup of the count pass
    const uint32_t head = ws.blkHead[blk];
    KmTile t[kScTiles];
#pragma unroll
    for (uint32_t i = 0; i < kScTiles; i++)
        t[i] = tBeg + i < k.nbTiles ? k.tiles[tBeg + i] : KmTile{kIdxInvalid, 0u, 0u};
    uint32_t a[kScTiles], src[kScTiles], h[kScTiles], tb[kScTiles], cr[kScTiles], to[kScTiles];
#pragma unroll
    for (uint32_t i = 0; i < kScTiles; i++) {
        const bool valid = lane < t[i].cnt, any = t[i].cnt != 0;
        a[i] = valid ? k.assign[t[i].row + lane] : 0xFFu;
        src[i] = valid ? orderIn[t[i].row + lane] : 0u;
        h[i] = any && lane < B ? hist[uint64_t(t[i].task) * B + lane] : 0u;
        tb[i] = any ? k.taskBeg[t[i].task] : 0u;
        cr[i] = any && t[i].task == head && lane < B ? ws.blkCarry[uint64_t(blk) * B + lane] : 0u;
        to[i] = a[i] < B ? ws.tileOff[uint64_t(tBeg + i) * B + a[i]] : 0u;
    }
#pragma unroll
    for (uint32_t i = 0; i < kScTiles; i++) {
        // lane c < B: first destination row of child c of the tile's task in this workgroup
        uint32_t inc = h[i];
#pragma unroll
        for (int d = 1; d < int(B); d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= uint32_t(d)) inc += o;
        }
        const uint32_t base = tb[i] + inc - h[i] + cr[i];
        uint64_t same = 0;
#pragma unroll
        for (uint32_t j = 0; j < B; j++) {
            const uint64_t m = __ballot(a[i] == j);
            if (a[i] == j) same = m;
        }
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi(uint32_t(same >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(same), 0u));
        const uint32_t dst = __shfl(base, int(a[i] < B ? a[i] : 0u), 64);
        if (a[i] < B) {
            const uint32_t pos = dst + to[i] + rank;
            if (pos < nbRows) orderOut[pos] = src[i];
        }
    }
}

// Vocabulary.cpp:103-108 IDF weight of one child from its occurrence count:
// occ == 0 -> FLT_MIN, else max(FLT_MIN, log(nbDoc / occ)) with the quotient
// in binary32 (IEEE division of the two binary32 conversions), its logarithm
// in binary64 and the result rounded to binary32.
__global__ void k_idfWeights(const uint32_t* __restrict__ occ, uint64_t n, uint32_t nbDoc, float* __restrict__ w) {
    constexpr float kFltMin = 1.17549435e-38f;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t o = occ[i];
        float r = kFltMin;
        if (o) r = fmaxf(kFltMin, float(log(double(__fdiv_rn(float(nbDoc), float(o))))));
        w[i] = r;
    }
}

__global__ void k_levelReset(KmLevel k, uint64_t nbRows) {
    const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x, stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = tid; i < nbRows; i += stride) k.assign[i] = 0xFF;  // "no previous assignment"
    for (uint64_t i = tid; i < uint64_t(k.nbTasks) * k.b * kDims; i += stride) k.sums[i] = 0;
    for (uint64_t i = tid; i < uint64_t(k.nbTasks) * k.b; i += stride) k.counts[i] = 0;
    for (uint64_t i = tid; i < k.nbTasks; i += stride) {
        k.changed[i] = 0;
        k.active[i] = 1;
        k.passes[i] = 0;
    }
}

uint32_t gridFor(uint64_t work, uint32_t perBlock, uint32_t cap) {
    const uint64_t g = (work + perBlock - 1) / perBlock;
    return uint32_t(g == 0 ? 1 : (g > cap ? cap : g));
}

// kernel<<<grid, block, lds, s>>>(args...), with t's events recorded by the
// dispatch (hipExtLaunchKernel) when it has any.  Args are the kernel's
// parameter types exactly: their addresses are the argument array.
template <typename... Params, typename... Args>
void launch(void (*kernel)(Params...), dim3 grid, dim3 block, uint32_t lds, hipStream_t s, const KernelTimer& t,
            Args... args) {
    static_assert(sizeof...(Params) == sizeof...(Args), "argument count");
    static_assert((std::is_same<Params, Args>::value && ...), "argument types must match the kernel's");
    if (!t.start && !t.stop) {
        kernel<<<grid, block, lds, s>>>(args...);
        RBOW_HIP(hipGetLastError());
        return;
    }
    void* a[] = {static_cast<void*>(&args)...};
    RBOW_HIP(hipExtLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, a, lds, s, t.start, t.stop, 0));
}

}  // namespace

void launchGatherRows(const uint8_t* desc, const uint32_t* idx, uint64_t n, uint8_t* out, hipStream_t s, uint64_t nbSrc) {
    if (!n) return;
    k_gatherRows<<<gridFor(n * 8, 256, 8192), 256, 0, s>>>(desc, idx, n, out, nbSrc);
    RBOW_HIP(hipGetLastError());
}

void launchIota(uint32_t* p, uint64_t n, hipStream_t s) {
    if (!n) return;
    k_iota<<<gridFor(n, 256, 4096), 256, 0, s>>>(p, n);
    RBOW_HIP(hipGetLastError());
}

void launchReformatTiles(const uint8_t* desc, const uint32_t* idx, uint32_t nbTiles, uint32_t tileSize,
                         uint32_t* tiles, hipStream_t s) {
    if (!nbTiles) return;
    RBOW_REQUIRE(tileSize >= 1 && tileSize <= 32, RBOW_ERR_INVALID_ARG, "tile size above 32");
    k_reformatTiles<<<nbTiles, 256, 0, s>>>(desc, idx, tileSize, tiles);
    RBOW_HIP(hipGetLastError());
}

void launchInitCenters(const KmLevel& k, const uint8_t* desc, const uint32_t* seedIdx, hipStream_t s) {
    if (!k.nbTasks) return;
    k_initCenters<<<gridFor(uint64_t(k.nbTasks) * k.b * kDims, 256, 8192), 256, 0, s>>>(
        k.b, k.nbTasks, desc, seedIdx, k.centroidType == 0 ? k.centersU8 : nullptr,
        k.centroidType == 0 ? nullptr : k.centersF32);
    RBOW_HIP(hipGetLastError());
}

void launchKmeansPP(const uint8_t* desc, uint32_t nbSrc, const uint32_t* order, const uint32_t* taskBeg,
                    const uint32_t* taskCnt, uint32_t nbTasks, uint32_t b, uint64_t seed, uint32_t level,
                    uint32_t* seedIdx, uint32_t* masks, hipStream_t s) {
    if (!nbTasks) return;
    k_kmeansPP<<<divUp(nbTasks, kWaves), 256, 0, s>>>(desc, order, taskBeg, taskCnt, nbTasks, b, seed, level, seedIdx,
                                                       masks, nbSrc);
    RBOW_HIP(hipGetLastError());
}

void launchKmLevelReset(const KmLevel& k, uint64_t nbRows, hipStream_t s) {
    k_levelReset<<<gridFor(std::max<uint64_t>(nbRows, uint64_t(k.nbTasks) * k.b * kDims), 256, 4096), 256, 0, s>>>(
        k, nbRows);
    RBOW_HIP(hipGetLastError());
}

void launchGatherTiles(const uint8_t* desc, uint32_t nbSrc, const uint32_t* order, const KmTile* tiles, uint32_t nbTiles,
                       int8_t* XA, hipStream_t s, const uint32_t* nbTilesDev) {
    if (!nbTiles) return;
    k_gatherTiles<<<nbTiles, 256, 0, s>>>(desc, order, tiles, XA, nbTilesDev, nbSrc);
    RBOW_HIP(hipGetLastError());
}

void launchPlanLevel(const uint32_t* hist, uint32_t nbTasks, uint32_t tileBound, uint32_t* taskBeg, uint32_t* taskCnt,
                     uint32_t* tileBeg, uint32_t* nbTiles, KmTile* tiles, hipStream_t s) {
    RBOW_REQUIRE(nbTasks >= 1 && nbTasks <= kPlanMaxTasks, RBOW_ERR_INVALID_ARG, "level plan: task count out of range");
    k_planLevel<<<1, 1024, 0, s>>>(hist, nbTasks, taskBeg, taskCnt, tileBeg, nbTiles);
    if (tiles)  // only the MFMA path's gather reads the tile list
        k_planTiles<<<gridFor(tileBound, 256, 1u << 30), 256, 0, s>>>(taskBeg, taskCnt, tileBeg, nbTasks, nbTiles,
                                                                      tiles, tileBound);
    RBOW_HIP(hipGetLastError());
}

bool launchKmReduce(const KmLevel& k, const KmPartRange* red, uint32_t nbRed, hipStream_t s, const KernelTimer& t) {
    if (!nbRed) return false;
    launch(k_reduceChunks, dim3(nbRed, 16 * kDims / 256), 256, 0, s, t, k, red, nbRed);
    return true;
}

uint32_t kmAssignResidentBlocks() {
    static thread_local int device = -1;
    static thread_local uint32_t blocks = 0;
    int d = 0;
    RBOW_HIP(hipGetDevice(&d));
    if (d != device) {
        int perCu = 0, cus = 0;
        RBOW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_assignU8Mfma, 256, 0));
        RBOW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d));
        blocks = uint32_t(std::max(1, perCu * cus));
        device = d;
    }
    return blocks;
}

KmPartition partitionKmTiles(const std::vector<uint32_t>& taskTiles, const std::vector<uint32_t>& tileBeg) {
    return partitionTiles(taskTiles, tileBeg, uint64_t(kmAssignResidentBlocks()) * kWaves, kWaves, kMergeSlots);
}

uint32_t kmAssignMergeSlots() { return kMergeSlots; }

uint32_t kmAssignWavesPerBlock() { return kWaves; }

namespace {
constexpr uint32_t kLloydTileLds = kTileLds + 64;  // LDS per tile: the rows + their assignments
}

uint32_t lloydTasksTileCap(uint32_t blocksPerCu) {
    static thread_local int device = -1;
    static thread_local int64_t perBlock = 0, perCu = 0, fixed = 0;
    int d = 0;
    RBOW_HIP(hipGetDevice(&d));
    if (d != device) {
        int maxBlock = 0, maxCu = 0;
        hipFuncAttributes f4{}, f8{};
        RBOW_HIP(hipDeviceGetAttribute(&maxBlock, hipDeviceAttributeMaxSharedMemoryPerBlock, d));
        RBOW_HIP(hipDeviceGetAttribute(&maxCu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, d));
        RBOW_HIP(hipFuncGetAttributes(&f4, reinterpret_cast<const void*>(k_lloydTasks<4>)));
        RBOW_HIP(hipFuncGetAttributes(&f8, reinterpret_cast<const void*>(k_lloydTasks<8>)));
        perBlock = maxBlock;
        perCu = maxCu;
        fixed = int64_t(std::max(f4.sharedSizeBytes, f8.sharedSizeBytes));
        device = d;
    }
    const int64_t room = std::min(perBlock, perCu / int64_t(std::max(blocksPerCu, 1u))) - fixed;
    return room > 0 ? uint32_t(room / kLloydTileLds) : 0u;
}

void launchLloydTasks(const KmLevel& k, const uint32_t* taskList, uint32_t nbList, uint32_t maxTiles,
                      uint32_t blocksPerCu, const uint32_t* taskTileBeg, uint32_t maxIters, hipStream_t s,
                      const KernelTimer& t, uint32_t nbStream) {
    if (!nbList) return;
    RBOW_REQUIRE(nbStream <= nbList && (nbStream == 0 || blocksPerCu <= 1), RBOW_ERR_INVALID_ARG,
                 "streamed persistent tasks need a workgroup per CU");
    RBOW_REQUIRE(nbStream == nbList || (maxTiles >= 1 && maxTiles <= lloydTasksTileCap(blocksPerCu)),
                 RBOW_ERR_INVALID_ARG, "persistent k-means tasks exceed the LDS capacity");
    // one workgroup per CU: 8 waves (2 per SIMD); two per CU: 4 waves each
    if (blocksPerCu <= 1) {
        const uint64_t resident = uint64_t(maxTiles) * kLloydTileLds;
        const uint64_t streamed = nbStream ? 8ull * kTileLds + kStreamTaskTiles * 64ull : 0ull;
        const int dyn = int(std::max(resident, streamed));
        RBOW_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_lloydTasks<8>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
        launch(k_lloydTasks<8>, nbList, 512, uint32_t(dyn), s, t, k, taskList, taskTileBeg, maxTiles, maxIters,
               nbStream);
    } else {
        const int dyn = int(maxTiles * kLloydTileLds);
        RBOW_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_lloydTasks<4>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
        launch(k_lloydTasks<4>, nbList, 256, uint32_t(dyn), s, t, k, taskList, taskTileBeg, maxTiles, maxIters, 0u);
    }
}

// Workgroups of a 256-thread kernel resident on the current device at once
// (occupancy x CUs), cached per kernel and device.
template <auto kernel>
uint32_t residentBlocksOf() {
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

bool launchKmAssign(const KmLevel& k, int kernel, hipStream_t s, const KernelTimer& t) {
    if (!k.nbTiles) return false;
    const uint32_t grid = gridFor(k.nbTiles, kWaves, 2048);
    if (k.centroidType == 0) {
        if (kernel == 1) launch(k_assignU8Plain, grid, 256, 0, s, t, k);
        else if (k.nbWaves) launch(k_assignU8Mfma, divUp(k.nbWaves, kWaves), 256, 0, s, t, k);
        else return false;
    } else if (kernel == 0 && k.fSplit) {
        // only resident workgroups, each wave on one contiguous tile range: a wave adds its LDS
        // sums into the task's global sums (one atomic per center and dimension) when its task
        // changes and at its end, so fewer, longer ranges mean fewer contended atomics (the b32 L4
        // float build on 2M descriptors: 17.3-17.8 ms instead of 19.1-19.4 ms with 2,048 blocks)
        if (k.b == 32)
            launch(k_assignF32Mfma<32>, std::min(grid, residentBlocksOf<k_assignF32Mfma<32>>()), 256, 0, s, t, k);
        else
            launch(k_assignF32Mfma<16>, std::min(grid, residentBlocksOf<k_assignF32Mfma<16>>()), 256, 0, s, t, k);
    } else if (k.b == 32) {
        launch(k_assignF32Plain<32>, grid, 256, 0, s, t, k);
    } else {
        launch(k_assignF32Plain<16>, grid, 256, 0, s, t, k);
    }
    return true;
}

namespace {
RgWs regroupWs(void* ws, uint64_t nbTiles, uint32_t b) {
    const uint64_t nbBlk = divUp(nbTiles, uint64_t(kRgTiles));
    auto at = [&](uint64_t off) { return reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(ws) + off); };
    auto up256 = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
    RgWs w;
    uint64_t o = 0;
    w.tileOff = at(o);
    o = up256(o + 4 * nbTiles * b);
    w.blkTot = at(o);
    o = up256(o + 4 * nbBlk * b);
    w.blkCarry = at(o);
    o = up256(o + 4 * nbBlk * b);
    w.blkHead = at(o);
    o = up256(o + 4 * nbBlk);
    w.blkTail = at(o);
    return w;
}
}  // namespace

size_t regroupWorkspaceBytes(uint64_t nbTilesMax, uint32_t b) {
    const uint64_t nbBlk = divUp(nbTilesMax, uint64_t(kRgTiles));
    return 4 * nbTilesMax * b + 8 * nbBlk * b + 8 * nbBlk + 5 * 256;
}

void launchRegroup(const KmLevel& k, uint32_t nbRows, const uint32_t* orderIn, uint32_t* orderOut, uint32_t* hist,
                   void* ws, size_t wsBytes, hipStream_t s) {
    RBOW_REQUIRE(k.b == 16 || k.b == 32, RBOW_ERR_INVALID_ARG, "branch factor must be 16 or 32");
    launchFillU32(hist, 0u, uint64_t(k.nbTasks) * k.b, s);
    if (!nbRows || !k.nbTiles) return;
    RBOW_REQUIRE(regroupWorkspaceBytes(k.nbTiles, k.b) <= wsBytes, RBOW_ERR_HIP, "regroup workspace too small");
    const RgWs w = regroupWs(ws, k.nbTiles, k.b);
    const uint32_t nbBlk = uint32_t(divUp(uint64_t(k.nbTiles), uint64_t(kRgTiles)));
    if (k.b == 16) k_regroupCount<16><<<nbBlk, 256, 0, s>>>(k, w, hist);
    else k_regroupCount<32><<<nbBlk, 256, 0, s>>>(k, w, hist);
    RBOW_HIP(hipGetLastError());
    if (!orderOut) return;  // the histogram only
    if (k.b == 16) {
        k_regroupCarry<16><<<1, 256, 0, s>>>(w, nbBlk);
        k_regroupScatter<16><<<divUp(k.nbTiles, kWaves * kScTiles), 256, 0, s>>>(k, w, hist, nbRows, orderIn, orderOut);
    } else {
        k_regroupCarry<32><<<1, 256, 0, s>>>(w, nbBlk);
        k_regroupScatter<32><<<divUp(k.nbTiles, kWaves * kScTiles), 256, 0, s>>>(k, w, hist, nbRows, orderIn, orderOut);
    }
    RBOW_HIP(hipGetLastError());
}

void launchIdfWeights(const uint32_t* occ, uint64_t n, uint32_t nbDoc, float* w, hipStream_t s) {
    if (!n) return;
    k_idfWeights<<<gridFor(n, 256, 4096), 256, 0, s>>>(occ, n, nbDoc, w);
    RBOW_HIP(hipGetLastError());
}

bool launchKmUpdate(const KmLevel& k, hipStream_t s, const KernelTimer& t) {
    if (!k.nbTasks) return false;
    if (k.partSums && k.centroidType == 0) {
        const uint32_t n = k.updTasks ? k.nbUpdTasks : k.nbTasks;
        if (!n) return false;
        launch(k_updateU8Part, dim3(n, 16 * kDims / 256), 256, 0, s, t, k);
    } else if (k.b == 32) {
        launch(k_update<32>, k.nbTasks, 256, 0, s, t, k);
    } else {
        launch(k_update<16>, k.nbTasks, 256, 0, s, t, k);
    }
    return true;
}

}  // namespace rbow_amd
