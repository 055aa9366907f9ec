/*
 * oracle.cpp -- CPU restatement of RapidBoW's hot-path semantics.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, and the CPU
 * baseline timed beside it.  The product never links or calls this file.
 *
 * Built with g++ -O3 -march=native -fopenmp -ffp-contract=off and WITHOUT
 * -ffast-math: every floating-point expression below is evaluated exactly as
 * written (IEEE binary32/binary64, round-to-nearest-even), and the flush-to-
 * zero behaviour of the reference's -ffast-math host build (CMakeLists.txt:16,
 * crtfastmath.o) is emulated explicitly where it matters (the database).
 */
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <omp.h>

namespace {

constexpr uint32_t kDims = 128;
constexpr uint32_t kIdxInvalid = 0xFFFFFFFFu;  // kmeans.cuh:31

uint64_t levelBeg(uint32_t b, uint32_t level) {  // FullTree.h:37-52 (impl 2)
    uint64_t acc = 0;
    for (uint32_t i = 0; i < level; i++) acc = acc * b + 1;
    return acc;
}
uint64_t levelSize(uint32_t b, uint32_t level) {  // FullTree.h:33-35
    uint64_t s = 1;
    for (uint32_t i = 0; i < level; i++) s *= b;
    return s;
}
uint32_t lg2(uint32_t b) { return b == 16 ? 4u : b == 32 ? 5u : 0u; }

inline uint32_t distU8(const uint8_t* x, const uint8_t* c) {  // kmeans.cuh:61-71
    uint32_t acc = 0;
    for (uint32_t k = 0; k < kDims; k++) {
        const int d = int(x[k]) - int(c[k]);
        acc += uint32_t(d * d);
    }
    return acc;
}
inline float distF32(const uint8_t* x, const float* c) {
    float acc = 0.0f;
    for (uint32_t k = 0; k < kDims; k++) {
        const float d = float(x[k]) - c[k];
        acc = std::fma(d, d, acc);
    }
    return acc;
}

// Reference Node layout (Vocabulary.h:42: KArray<uint32_t, 32, 16>) -> child-major rows.
void refNodeToRows(const uint8_t* node, uint8_t* rows /*[16][128]*/) {
    for (uint32_t j = 0; j < 16; j++)
        for (uint32_t k = 0; k < kDims; k++) rows[j * kDims + k] = node[(k >> 2) * 64 + j * 4 + (k & 3)];
}
void rowsToRefNode(const uint8_t* rows, uint8_t* node) {
    for (uint32_t j = 0; j < 16; j++)
        for (uint32_t k = 0; k < kDims; k++) node[(k >> 2) * 64 + j * 4 + (k & 3)] = rows[j * kDims + k];
}

// argmin over b children with the reference tie rule (first minimum wins).
inline uint32_t argminU8(const uint8_t* x, const uint8_t* rows, uint32_t b) {
    uint32_t best = 0, bestD = 0;
    for (uint32_t j = 0; j < b; j++) {
        const uint32_t d = distU8(x, rows + j * kDims);
        if (j == 0 || d < bestD) { bestD = d; best = j; }
    }
    return best;
}
// FP32 distances of one point to k centers held transposed [128][k]; each
// accumulator is its own sequential fma chain over k = 0..127.
inline uint32_t argminF32T(const uint8_t* x, const float* cT, uint32_t k) {
    float acc[32];
    for (uint32_t j = 0; j < k; j++) acc[j] = 0.0f;
    for (uint32_t d = 0; d < kDims; d++) {
        const float xv = float(x[d]);
        const float* row = cT + d * k;
        for (uint32_t j = 0; j < k; j++) {
            const float diff = xv - row[j];
            acc[j] = std::fma(diff, diff, acc[j]);
        }
    }
    uint32_t best = 0;
    for (uint32_t j = 1; j < k; j++)
        if (acc[j] < acc[best]) best = j;
    return best;
}
void transposeCenters(const float* c, uint32_t k, float* cT) {
    for (uint32_t j = 0; j < k; j++)
        for (uint32_t d = 0; d < kDims; d++) cT[d * k + j] = c[j * kDims + d];
}

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// kmeans.cpp:141-168 (randPick), engine passed explicitly.
uint32_t randPick(std::mt19937_64& rng, uint32_t* dst, uint32_t randSpace, uint32_t nbToPick) {
    const uint32_t nbPick = std::min(randSpace, nbToPick);
    static thread_local std::unordered_set<uint32_t> results;
    results.clear();
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

// ---- FTZ/DAZ emulation of the reference host build (CMakeLists.txt:16) ----
inline float ftz(float x) { return std::fabs(x) < FLT_MIN ? std::copysign(0.0f, x) : x; }
inline float fmulz(float a, float b) { return ftz(ftz(a) * ftz(b)); }
inline float faddz(float a, float b) { return ftz(ftz(a) + ftz(b)); }
inline float fsubz(float a, float b) { return ftz(ftz(a) - ftz(b)); }

}  // namespace

extern "C" {

uint64_t orc_level_beg(uint32_t b, uint32_t level) { return levelBeg(b, level); }
uint32_t orc_dist_u8(const uint8_t* x, const uint8_t* c) { return distU8(x, c); }
float orc_dist_f32(const uint8_t* x, const float* c) { return distF32(x, c); }
uint64_t orc_hash(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return mix64(mix64(mix64(seed ^ mix64(a)) ^ b) ^ c);
}

void orc_lookup_u8(const uint8_t* nodesRef, uint32_t nbLevels, const uint8_t* desc, uint64_t nbDesc,
                   uint32_t* leaf, int nthreads) {
    const uint64_t nbNodes = levelBeg(16, nbLevels);
    std::vector<uint8_t> rows(nbNodes * 16 * kDims);
    for (uint64_t i = 0; i < nbNodes; i++) refNodeToRows(nodesRef + i * 2048, &rows[i * 2048]);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int64_t q = 0; q < int64_t(nbDesc); q++) {
        uint32_t node = 0;  // node index within the current level
        for (uint32_t l = 0; l < nbLevels; l++) {
            const uint8_t* r = &rows[(levelBeg(16, l) + node) * 2048];
            node = node * 16 + argminU8(desc + q * kDims, r, 16);  // FullTree.h:59-61
        }
        leaf[q] = node;
    }
}

void orc_lookup_f32(const float* nodes, uint32_t b, uint32_t nbLevels, const uint8_t* desc, uint64_t nbDesc,
                    uint32_t* leaf, int nthreads) {
    const uint64_t nbNodes = levelBeg(b, nbLevels);
    std::vector<float> nodesT(nbNodes * b * kDims);
    for (uint64_t i = 0; i < nbNodes; i++) transposeCenters(nodes + i * b * kDims, b, &nodesT[i * b * kDims]);
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int64_t q = 0; q < int64_t(nbDesc); q++) {
        uint32_t node = 0;
        for (uint32_t l = 0; l < nbLevels; l++) {
            const float* cT = &nodesT[(levelBeg(b, l) + node) * b * kDims];
            node = node * b + argminF32T(desc + q * kDims, cT, b);
        }
        leaf[q] = node;
    }
}

uint32_t orc_assign_u8(const uint8_t* pts, uint32_t n, const uint8_t* centers, uint8_t* assign, uint32_t* counts,
                       int64_t* sums) {
    std::fill(counts, counts + 16, 0u);
    std::fill(sums, sums + 16 * kDims, int64_t(0));
    uint32_t changed = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint8_t* x = pts + uint64_t(i) * kDims;
        const uint32_t a = argminU8(x, centers, 16);
        if (assign[i] != a) changed++;  // kmeans.cu:159-168
        assign[i] = uint8_t(a);
        counts[a]++;  // kmeans.cu:151-157
        for (uint32_t k = 0; k < kDims; k++) sums[a * kDims + k] += x[k];  // kmeans.cu:170-197
    }
    return changed;
}

void orc_update_u8(uint8_t* centers, const uint32_t* counts, const int64_t* sums) {
    for (uint32_t j = 0; j < 16; j++) {
        if (counts[j] == 0) continue;  // kmeans.cu:355-356: empty centers keep their position
        for (uint32_t k = 0; k < kDims; k++)
            centers[j * kDims + k] = uint8_t(std::round(float(sums[j * kDims + k]) / float(counts[j])));
    }
}

uint32_t orc_assign_f32(const uint8_t* pts, uint32_t n, const float* centers, uint32_t k, uint8_t* assign,
                        uint32_t* counts, int64_t* sums) {
    std::vector<float> cT(k * kDims);
    transposeCenters(centers, k, cT.data());
    std::fill(counts, counts + k, 0u);
    std::fill(sums, sums + size_t(k) * kDims, int64_t(0));
    uint32_t changed = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint8_t* x = pts + uint64_t(i) * kDims;
        const uint32_t a = argminF32T(x, cT.data(), k);
        if (assign[i] != a) changed++;
        assign[i] = uint8_t(a);
        counts[a]++;
        for (uint32_t d = 0; d < kDims; d++) sums[a * kDims + d] += x[d];
    }
    return changed;
}

void orc_update_f32(float* centers, uint32_t k, const uint32_t* counts, const int64_t* sums) {
    for (uint32_t j = 0; j < k; j++) {
        if (counts[j] == 0) continue;
        for (uint32_t d = 0; d < kDims; d++)
            centers[j * kDims + d] = float(sums[j * kDims + d]) / float(counts[j]);
    }
}

void orc_reformat(const uint8_t* desc, const uint32_t* idx, uint32_t nbTiles, uint32_t tileSize, uint32_t* tiles) {
    for (uint32_t t = 0; t < nbTiles; t++)
        for (uint32_t w = 0; w < 32; w++)
            for (uint32_t s = 0; s < tileSize; s++) {
                const uint32_t i = idx[t * tileSize + s];
                uint32_t v = ~0u;  // reformatDesc.cu:59-64: invalid slots read as 0xFFFFFFFF
                if (i != kIdxInvalid) std::memcpy(&v, desc + uint64_t(i) * kDims + w * 4, 4);
                tiles[(uint64_t(t) * 32 + w) * tileSize + s] = v;
            }
}

void* orc_rng_new(uint64_t seed) { return new std::mt19937_64(seed); }
void orc_rng_free(void* rng) { delete static_cast<std::mt19937_64*>(rng); }
uint64_t orc_rng_next(void* rng) { return (*static_cast<std::mt19937_64*>(rng))(); }
uint32_t orc_rand_pick(void* rng, uint32_t* dst, uint32_t randSpace, uint32_t nbToPick) {
    return randPick(*static_cast<std::mt19937_64*>(rng), dst, randSpace, nbToPick);
}

uint32_t orc_kmeanspp_pick(const uint8_t* desc, const uint32_t* grpIdx, uint32_t n, uint32_t k, uint64_t seed,
                           uint32_t level, uint32_t grp, uint32_t* posInGroup) {
    if (n == 0) {
        for (uint32_t j = 0; j < k; j++) posInGroup[j] = kIdxInvalid;
        return 0;
    }
    const uint32_t m = std::min<uint32_t>(n, 256);
    std::vector<uint32_t> cand(m);
    for (uint32_t s = 0; s < m; s++) {  // stratified sample: one member per stratum
        const uint64_t lo = uint64_t(s) * n / m, hi = uint64_t(s + 1) * n / m;
        cand[s] = uint32_t(lo + orc_hash(seed, level, grp, 0x100000000ull + s) % (hi - lo));
    }
    std::vector<uint32_t> D(m, 0xFFFFFFFFu);
    std::vector<uint32_t> chosen(k);
    auto relax = [&](uint32_t c) {
        const uint8_t* cp = desc + uint64_t(grpIdx[cand[c]]) * kDims;
        for (uint32_t s = 0; s < m; s++) D[s] = std::min(D[s], distU8(desc + uint64_t(grpIdx[cand[s]]) * kDims, cp));
    };
    chosen[0] = uint32_t(orc_hash(seed, level, grp, 0x200000000ull) % m);
    relax(chosen[0]);
    uint32_t nValid = 1;
    for (uint32_t r = 1; r < k; r++) {
        uint64_t total = 0;
        for (uint32_t s = 0; s < m; s++) total += D[s];
        if (total == 0) {
            chosen[r] = chosen[r - 1];
            continue;
        }
        const uint64_t t = orc_hash(seed, level, grp, 0x300000000ull + r) % total;
        uint64_t cum = 0;
        uint32_t pick = m - 1;
        for (uint32_t s = 0; s < m; s++) {
            cum += D[s];
            if (cum > t) { pick = s; break; }
        }
        chosen[r] = pick;
        nValid++;
        relax(pick);
    }
    for (uint32_t j = 0; j < k; j++) posInGroup[j] = cand[chosen[j]];
    return nValid;
}

int orc_build(const uint8_t* desc, uint32_t nbDesc, uint32_t nbDoc, uint32_t b, uint32_t nbLevels, uint32_t maxIters,
              int seedMode, uint64_t seed, int centroidType, void* nodesOut, uint32_t* masks, float* weights,
              uint32_t* occ, uint32_t* lloydPasses, double* centroidErr) {
    if (centroidType == 0 && b != 16) return -1;
    if (centroidType == 1 && b != 16 && b != 32) return -1;
    if (nbLevels == 0 || maxIters == 0 || (seedMode != 0 && seedMode != 1)) return -1;
    if (nbDesc < b) return -1;  // Vocabulary.cpp:39-41 "Insufficient descriptors"
    if (nbDoc == 0) return -1;  // Vocabulary.cpp:42-44 "nbDoc must be positive"
    const uint64_t nbNodes = levelBeg(b, nbLevels);
    std::fill(occ, occ + nbNodes * b, 0u);
    std::mt19937_64 rng(seed);

    std::vector<std::vector<uint32_t>> groups{std::vector<uint32_t>(nbDesc)};  // Vocabulary.cpp:51-52
    double maxErr = 0;  // max |centroid - binary64 mean of its points| over children with points
    std::iota(groups[0].begin(), groups[0].end(), 0u);

    for (uint32_t level = 0; level < nbLevels; level++) {
        const uint32_t nG = uint32_t(groups.size());
        const uint64_t nodeBase = levelBeg(b, level);
        // ---- seeds, drawn sequentially in group order (kmeans.cpp:170-196) ----
        std::vector<uint32_t> seedIdx(size_t(nG) * b);
        for (uint32_t g = 0; g < nG; g++) {
            const uint32_t n = uint32_t(groups[g].size());
            uint32_t nValid;
            uint32_t pos[32];
            if (seedMode == 0) {
                nValid = randPick(rng, pos, n, b);
            } else {
                nValid = orc_kmeanspp_pick(desc, groups[g].data(), n, b, seed, level, g, pos);
            }
            for (uint32_t j = 0; j < b; j++) {
                uint32_t idx;
                if (n > 0) {
                    idx = groups[g][pos[j]];
                } else {
                    // Empty group: the reference reads the next group's first index
                    // (kmeans.cpp:180-186 via DescGrouping padding); none left -> invalid.
                    idx = kIdxInvalid;
                    if (seedMode == 0)
                        for (uint32_t h = g + 1; h < nG; h++)
                            if (!groups[h].empty()) { idx = groups[h][0]; break; }
                }
                seedIdx[size_t(g) * b + j] = idx;
            }
            masks[nodeBase + g] = nValid >= 32 ? 0xFFFFFFFFu : ((1u << nValid) - 1u);  // kmeans.cpp:189
        }
        // ---- Lloyd loop per group (kmeans.cpp:118-133 with per-task stop) ----
        std::vector<std::vector<uint8_t>> assigns(nG);
        std::vector<double> taskErr(nG, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t gi = 0; gi < int64_t(nG); gi++) {
            const uint32_t g = uint32_t(gi);
            const std::vector<uint32_t>& grp = groups[g];
            const uint32_t n = uint32_t(grp.size());
            std::vector<uint8_t> pts(size_t(n) * kDims);
            for (uint32_t i = 0; i < n; i++) std::memcpy(&pts[size_t(i) * kDims], desc + uint64_t(grp[i]) * kDims, kDims);
            std::vector<uint8_t>& assign = assigns[g];
            assign.assign(n, uint8_t(0xFF));  // builder rule: "no previous assignment"
            std::vector<uint32_t> counts(b);
            std::vector<int64_t> sums(size_t(b) * kDims);
            uint32_t passes = 0;
            if (centroidType == 0) {
                uint8_t c[16 * kDims];
                for (uint32_t j = 0; j < 16; j++) {
                    const uint32_t idx = seedIdx[size_t(g) * 16 + j];
                    if (idx == kIdxInvalid) std::memset(c + j * kDims, 0xFF, kDims);
                    else std::memcpy(c + j * kDims, desc + uint64_t(idx) * kDims, kDims);
                }
                for (uint32_t it = 0; it < maxIters; it++) {
                    passes++;
                    const uint32_t changed = orc_assign_u8(pts.data(), n, c, assign.data(), counts.data(), sums.data());
                    if (changed == 0) break;
                    orc_update_u8(c, counts.data(), sums.data());
                }
                rowsToRefNode(c, static_cast<uint8_t*>(nodesOut) + (nodeBase + g) * 2048);
                // the final centers are the update of the final assignment, whose totals the last pass left
                for (uint32_t j = 0; j < 16; j++)
                    for (uint32_t d = 0; counts[j] && d < kDims; d++)
                        taskErr[g] = std::max(taskErr[g], std::fabs(double(c[j * kDims + d]) -
                                                                    double(sums[j * kDims + d]) / counts[j]));
            } else {
                std::vector<float> c(size_t(b) * kDims);
                for (uint32_t j = 0; j < b; j++) {
                    const uint32_t idx = seedIdx[size_t(g) * b + j];
                    for (uint32_t d = 0; d < kDims; d++)
                        c[j * kDims + d] = idx == kIdxInvalid ? 255.0f : float(desc[uint64_t(idx) * kDims + d]);
                }
                for (uint32_t it = 0; it < maxIters; it++) {
                    passes++;
                    const uint32_t changed =
                        orc_assign_f32(pts.data(), n, c.data(), b, assign.data(), counts.data(), sums.data());
                    if (changed == 0) break;
                    orc_update_f32(c.data(), b, counts.data(), sums.data());
                }
                std::memcpy(static_cast<float*>(nodesOut) + (nodeBase + g) * b * kDims, c.data(),
                            sizeof(float) * b * kDims);
                for (uint32_t j = 0; j < b; j++)
                    for (uint32_t d = 0; counts[j] && d < kDims; d++)
                        taskErr[g] = std::max(taskErr[g], std::fabs(double(c[j * kDims + d]) -
                                                                    double(sums[j * kDims + d]) / counts[j]));
            }
            if (lloydPasses) lloydPasses[nodeBase + g] = passes;
        }
        for (double e : taskErr) maxErr = std::max(maxErr, e);
        // ---- stable regroup + occurrence (Vocabulary.cpp:83-98) ----
        std::vector<std::vector<uint32_t>> next(size_t(nG) * b);
        for (uint32_t g = 0; g < nG; g++) {
            for (uint32_t i = 0; i < groups[g].size(); i++) next[size_t(g) * b + assigns[g][i]].push_back(groups[g][i]);
            for (uint32_t j = 0; j < b; j++) occ[(nodeBase + g) * b + j] = uint32_t(next[size_t(g) * b + j].size());
        }
        groups.swap(next);
    }
    orc_idf_weights(occ, nbNodes * b, nbDoc, weights);
    if (centroidErr) *centroidErr = maxErr;
    return 0;
}

// Vocabulary.cpp:103-108 IDF weights; log evaluated in binary64 of the binary32
// quotient, then rounded to binary32 (builder rule for std::log(float)).
void orc_idf_weights(const uint32_t* occ, uint64_t n, uint32_t nbDoc, float* weights) {
    for (uint64_t i = 0; i < n; i++) {
        const uint32_t o = occ[i];
        if (o == 0) {
            weights[i] = FLT_MIN;
        } else {
            const float q = float(nbDoc) / float(o);
            weights[i] = std::max(FLT_MIN, float(std::log(double(q))));
        }
    }
}

// Distances along each descriptor's path: for level l, the child of the path's
// level-l node that the leaf id names; d32 = orc_dist_f32 (the FP32 chain),
// d64 = the same sum in binary64 of the binary32 centroid values.
void orc_path_dist_f32(const float* nodes, uint32_t b, uint32_t nbLevels, const uint8_t* desc, uint64_t nbDesc,
                       const uint32_t* leaf, float* d32, double* d64, int nthreads) {
    uint32_t lg2b = 0;
    while ((1u << lg2b) < b) lg2b++;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int64_t i = 0; i < int64_t(nbDesc); i++) {
        const uint8_t* x = desc + uint64_t(i) * 128;
        for (uint32_t l = 0; l < nbLevels; l++) {
            const uint64_t path = leaf[i] >> (lg2b * (nbLevels - 1 - l));  // node index within level l+1
            const uint64_t node = orc_level_beg(b, l) + (path >> lg2b);
            const float* c = nodes + (node * b + (path & (b - 1))) * 128;
            double s = 0;
            for (int k = 0; k < 128; k++) {
                const double d = double(x[k]) - double(c[k]);
                s += d * d;
            }
            d32[uint64_t(i) * nbLevels + l] = orc_dist_f32(x, c);
            d64[uint64_t(i) * nbLevels + l] = s;
        }
    }
}

}  // extern "C"

// ------------------------------------------------------------------------
// DataBase (DataBase.h:35-64, DataBase.cpp:30-197)
// ------------------------------------------------------------------------
namespace {
struct OrcDb {
    uint32_t b, lg2b, L;
    std::vector<float> W;                  // [node][b]
    std::vector<std::vector<uint32_t>> ivf;  // b^L leaf inverted files (DataBase.h:58-59)
    std::unordered_map<uint32_t, float> rcp;  // DataBase.h:62
    std::unordered_map<uint32_t, double> l1;  // binary64 L1 norm, kept for the FP64 ideal
    uint64_t nbPostings = 0;
    float w(uint32_t idxLevel, uint32_t v) const { return W[levelBeg(b, idxLevel) * b + v]; }  // DataBase.cpp:30-36
};

template <typename Acc>
void l1Norm(const OrcDb& db, const uint32_t* leaf, uint32_t n, Acc& acc) {  // DataBase.cpp:54-67
    std::vector<uint32_t> nxt(leaf, leaf + n);
    std::sort(nxt.begin(), nxt.end());
    for (uint32_t t = 0; t < db.L; t++) {
        const uint32_t idxLevel = db.L - 1u - t;
        for (uint32_t i = 0; i < n; i++) {
            acc += double(db.w(idxLevel, nxt[i]));
            nxt[i] >>= db.lg2b;
        }
    }
}

// Walks the scored (node, doc, multiplicity) triples of a query in the
// reference's order (t ascending, node ascending), DataBase.cpp:103-169.
template <typename F>
void forEachScoredTerm(const OrcDb& db, const uint32_t* leaf, uint32_t n, uint32_t K, F&& f) {
    std::vector<uint32_t> s(leaf, leaf + n);
    std::sort(s.begin(), s.end());
    const uint32_t threshold = 16u + K * 4u;  // DataBase.h:60-61
    std::unordered_map<uint32_t, uint32_t> occ;
    for (uint32_t t = 0; t < db.L; t++) {
        const uint32_t idxLevel = db.L - 1u - t;
        const uint32_t shift = t * db.lg2b;
        for (uint32_t i = 0; i < n;) {
            const uint32_t v = s[i] >> shift;
            uint32_t cnt = 1;
            while (i + cnt < n && (s[i + cnt] >> shift) == v) cnt++;
            const uint64_t lo = uint64_t(v) << shift, hi = lo + (uint64_t(1) << shift);
            uint32_t P = 0;
            for (uint64_t r = lo; r < hi; r++) P += uint32_t(db.ivf[r].size());
            if (t == 0 || P < threshold) {
                const float w = db.w(idxLevel, v);
                if (hi - lo == 1) {
                    const auto& l = db.ivf[lo];
                    for (size_t e = 0; e < l.size();) {
                        size_t m = 1;
                        while (e + m < l.size() && l[e + m] == l[e]) m++;
                        f(t, v, cnt, w, l[e], uint32_t(m));
                        e += m;
                    }
                } else {
                    occ.clear();
                    std::vector<uint32_t> order;
                    for (uint64_t r = lo; r < hi; r++)
                        for (uint32_t d : db.ivf[r]) {
                            if (occ[d]++ == 0) order.push_back(d);
                        }
                    for (uint32_t d : order) f(t, v, cnt, w, d, occ[d]);
                }
            }
            i += cnt;
        }
    }
}

struct RankItem {
    uint32_t doc;
    float score;
};
}  // namespace

extern "C" {

void* orc_db_new(uint32_t b, uint32_t nbLevels, const float* weights) {
    if ((b != 16 && b != 32) || nbLevels == 0) return nullptr;
    auto* db = new OrcDb;
    db->b = b;
    db->lg2b = lg2(b);
    db->L = nbLevels;
    const uint64_t nbNodes = levelBeg(b, nbLevels);
    db->W.assign(weights, weights + nbNodes * b);
    db->ivf.resize(levelSize(b, nbLevels));
    return db;
}
void orc_db_free(void* db) { delete static_cast<OrcDb*>(db); }

int orc_db_add(void* pdb, uint32_t doc, const uint32_t* leaf, uint32_t n) {
    OrcDb& db = *static_cast<OrcDb*>(pdb);
    if (db.rcp.count(doc)) return -1;  // DataBase.cpp:41-43
    for (uint32_t i = 0; i < n; i++) {  // DataBase.cpp:44-53: sorted insert, duplicates kept
        auto& l = db.ivf[leaf[i]];
        l.insert(std::upper_bound(l.begin(), l.end(), doc), doc);
    }
    db.nbPostings += n;
    double L1 = 0;
    l1Norm(db, leaf, n, L1);
    db.rcp[doc] = ftz(float(1 / L1));  // DataBase.cpp:68
    db.l1[doc] = L1;
    return 0;
}
float orc_db_rcp(void* db, uint32_t doc) {
    auto& m = static_cast<OrcDb*>(db)->rcp;
    auto it = m.find(doc);
    return it == m.end() ? std::nanf("") : it->second;
}
uint64_t orc_db_nb_postings(void* db) { return static_cast<OrcDb*>(db)->nbPostings; }
uint32_t orc_db_postings(void* pdb, uint32_t leaf, uint32_t* out, uint32_t cap) {
    const auto& l = static_cast<OrcDb*>(pdb)->ivf.at(leaf);
    std::copy_n(l.begin(), std::min<size_t>(cap, l.size()), out);
    return uint32_t(l.size());
}

static std::map<uint32_t, double> queryAcc(const OrcDb& db, const uint32_t* leaf, uint32_t n, uint32_t K) {
    double L1 = 0;
    l1Norm(db, leaf, n, L1);
    const float qrcp = ftz(float(1 / L1));  // DataBase.cpp:77-94
    std::map<uint32_t, double> acc;
    forEachScoredTerm(db, leaf, n, K, [&](uint32_t, uint32_t, uint32_t cnt, float w, uint32_t doc, uint32_t m) {
        const float qE = fmulz(fmulz(float(cnt), w), qrcp);  // DataBase.cpp:117
        const float tE = fmulz(fmulz(float(m), w), db.rcp.at(doc));  // :128-129
        const float c = fsubz(std::fabs(fsubz(qE, tE)), faddz(qE, tE));  // :130
        acc[doc] += double(c);
    });
    return acc;
}

uint32_t orc_db_query(void* pdb, const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs, float* scores,
                      uint32_t* nbTouched) {
    const OrcDb& db = *static_cast<OrcDb*>(pdb);
    const auto acc = queryAcc(db, leaf, n, K);
    std::vector<RankItem> items;
    items.reserve(acc.size());
    for (const auto& kv : acc) items.push_back({kv.first, float(kv.second + 2.0)});  // DataBase.cpp:170-174
    // Builder rule for equal float scores (parity unpinned in the reference): doc id ascending.
    std::sort(items.begin(), items.end(), [](const RankItem& a, const RankItem& b) {
        return a.score < b.score || (a.score == b.score && a.doc < b.doc);
    });
    const uint32_t cnt = uint32_t(std::min<size_t>(K, items.size()));
    for (uint32_t i = 0; i < cnt; i++) {
        docs[i] = items[i].doc;
        if (scores) scores[i] = items[i].score;
    }
    if (nbTouched) *nbTouched = uint32_t(items.size());
    return cnt;
}

uint32_t orc_db_query_all(void* pdb, const uint32_t* leaf, uint32_t n, uint32_t K, float* scoresOut,
                          double* scores64Out, uint32_t nbDocsMax) {
    const OrcDb& db = *static_cast<OrcDb*>(pdb);
    const auto acc = queryAcc(db, leaf, n, K);
    for (uint32_t i = 0; i < nbDocsMax; i++) {
        if (scoresOut) scoresOut[i] = std::nanf("");
        if (scores64Out) scores64Out[i] = std::nan("");
    }
    // FP64 ideal for every touched doc
    std::map<uint32_t, double> ideal;
    {
        double L1 = 0;
        l1Norm(db, leaf, n, L1);
        const double qr = 1.0 / L1;
        forEachScoredTerm(db, leaf, n, K, [&](uint32_t, uint32_t, uint32_t cnt, float w, uint32_t doc, uint32_t m) {
            const double qE = double(cnt) * double(w) * qr;
            const double tE = double(m) * double(w) / db.l1.at(doc);
            ideal[doc] += std::fabs(qE - tE) - (qE + tE);
        });
    }
    for (const auto& kv : acc)
        if (kv.first < nbDocsMax) {
            if (scoresOut) scoresOut[kv.first] = float(kv.second + 2.0);
            if (scores64Out) scores64Out[kv.first] = ideal[kv.first] + 2.0;
        }
    return uint32_t(acc.size());
}

uint32_t orc_db_query_fp64(void* pdb, const uint32_t* leaf, uint32_t n, uint32_t K, uint32_t* docs,
                           double* scores) {
    const OrcDb& db = *static_cast<OrcDb*>(pdb);
    double L1 = 0;
    l1Norm(db, leaf, n, L1);
    const double qr = 1.0 / L1;
    std::map<uint32_t, double> acc;
    forEachScoredTerm(db, leaf, n, K, [&](uint32_t, uint32_t, uint32_t cnt, float w, uint32_t doc, uint32_t m) {
        const double qE = double(cnt) * double(w) * qr;
        const double tE = double(m) * double(w) / db.l1.at(doc);
        acc[doc] += std::fabs(qE - tE) - (qE + tE);
    });
    std::vector<std::pair<double, uint32_t>> items;
    for (const auto& kv : acc) items.push_back({kv.second + 2.0, kv.first});
    std::sort(items.begin(), items.end());
    const uint32_t cnt = uint32_t(std::min<size_t>(K, items.size()));
    for (uint32_t i = 0; i < cnt; i++) {
        docs[i] = items[i].second;
        if (scores) scores[i] = items[i].first;
    }
    return cnt;
}

// ------------------------------------------------------------------------
// Reference test inputs, regenerated with the same standard engines/distributions.
// ------------------------------------------------------------------------
void orc_ref_assign_test_inputs(uint32_t nbDesc, uint32_t* tiles, uint32_t* centers, uint8_t* prevAssign) {
    // tests/KmeansAssignCenterTests.cpp:36-73
    const uint32_t nbTiles = (nbDesc + 31) / 32;
    const uint64_t descElems = 32ull * 32 * nbTiles;
    std::mt19937_64 rng{};
    std::uniform_int_distribution<uint64_t> dist{};
    for (uint64_t i = 0; i < descElems / 2; i++) {
        const uint64_t v = rng();
        std::memcpy(&tiles[2 * i], &v, 8);
    }
    for (uint32_t i = 0; i < 16; i++)
        for (uint32_t j = 0; j < 32; j++) {
            const uint32_t idxDesc = std::min(i, nbDesc - 1);
            centers[j * 16 + i] = tiles[(idxDesc / 32) * 1024 + j * 32 + idxDesc % 32];
        }
    for (uint32_t i = 0; i < nbDesc; i++) prevAssign[i] = uint8_t(dist(rng) % 16);
}

void orc_ref_reformat_test_inputs(uint32_t nbDesc, uint32_t nbTiles, uint32_t tileSize, uint8_t* desc,
                                  uint32_t* idx) {
    // tests/ReformatTests.cpp:40-73
    std::mt19937_64 rng{};
    std::uniform_int_distribution<uint64_t> dist64{};
    std::uniform_int_distribution<uint32_t> distIdx{0u, uint32_t(nbDesc - 1)};
    std::bernoulli_distribution distBin{0.125f};
    for (uint32_t d = 0; d < nbDesc; d++)
        for (uint32_t w = 0; w < 16; w++) {
            const uint64_t v = dist64(rng);
            std::memcpy(desc + uint64_t(d) * kDims + w * 8, &v, 8);
        }
    for (uint32_t t = 0; t < nbTiles; t++)
        for (uint32_t s = 0; s < tileSize; s++) idx[t * tileSize + s] = distBin(rng) ? kIdxInvalid : distIdx(rng);
}

void orc_ref_voc_test_inputs(int which, uint8_t* desc) {
    std::mt19937_64 rng{};
    if (which == 0) {  // tests/VocTests.cpp:38-50
        std::uniform_int_distribution<uint16_t> dist{};
        for (uint32_t i = 0; i < 4096; i++) {
            std::memset(desc + i * kDims, 0, kDims);
            const uint16_t v = dist(rng);
            std::memcpy(desc + i * kDims + 126, &v, 2);
        }
    } else {  // tests/VocTests.cpp:86-100
        std::uniform_int_distribution<uint64_t> dist{};
        for (uint32_t i = 0; i < 32; i++)
            for (uint32_t w = 0; w < 16; w++) {
                const uint64_t v = dist(rng);
                std::memcpy(desc + i * kDims + w * 8, &v, 8);
            }
    }
}

}  // extern "C"
