// synth.hip -- device-side synthetic workloads (no counterpart in the
// reference): SIFT-like descriptors generated straight into HBM, images as
// descriptor sets over a per-image subset of centres, and posting lists drawn
// from per-centre "leaf clouds" so that a billion-posting inverted file can be
// built without generating its descriptors.  Deterministic counter hashes;
// statistically the same distribution as the host generator, not bit-equal.
#include "../../include/rapidbow.h"
#include "common.hpp"

namespace rbow_amd {
namespace {

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ inline uint64_t h3(uint64_t seed, uint64_t a, uint64_t b) { return mix64(mix64(seed ^ mix64(a)) ^ b); }

__device__ inline uint32_t imageOf(const uint64_t* off, uint32_t nbImages, uint64_t i) {
    uint32_t lo = 0, hi = nbImages;  // off[lo] <= i < off[lo+1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One wave per descriptor, two dims per lane.
__global__ __launch_bounds__(256) void k_synthDesc(const float* __restrict__ centres, const float* __restrict__ sigma,
                                                  uint32_t nbCentres, uint64_t seed, uint64_t first, uint64_t n,
                                                  const uint32_t* __restrict__ centreIds, uint8_t* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t i = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); i < n; i += uint64_t(gridDim.x) * 4) {
        const uint64_t g = first + i;
        const uint32_t c = centreIds ? centreIds[i] % nbCentres : uint32_t(h3(seed, g, 0xFFFFull) % nbCentres);
        const uint64_t r = h3(seed, g, lane);
        const float u1 = (float(uint32_t(r >> 32)) + 1.0f) * (1.0f / 4294967296.0f);
        const float u2 = float(uint32_t(r)) * (1.0f / 4294967296.0f);
        const float rad = sqrtf(-2.0f * logf(u1));
        float sn, cs;
        sincosf(6.2831853f * u2, &sn, &cs);
        const float s = sigma[c];
        float x0 = fabsf(centres[c * 128 + 2 * lane] + s * rad * cs);
        float x1 = fabsf(centres[c * 128 + 2 * lane + 1] + s * rad * sn);
        float nrm = x0 * x0 + x1 * x1;
        for (int m = 32; m >= 1; m >>= 1) nrm += __shfl_xor(nrm, m, 64);
        nrm = nrm > 0 ? sqrtf(nrm) : 1.0f;
        x0 = fminf(x0 / nrm, 0.2f);
        x1 = fminf(x1 / nrm, 0.2f);
        float n2 = x0 * x0 + x1 * x1;
        for (int m = 32; m >= 1; m >>= 1) n2 += __shfl_xor(n2, m, 64);
        n2 = n2 > 0 ? sqrtf(n2) : 1.0f;
        const uint32_t b0 = uint32_t(fminf(rintf(512.0f * x0 / n2), 255.0f));
        const uint32_t b1 = uint32_t(fminf(rintf(512.0f * x1 / n2), 255.0f));
        reinterpret_cast<uint16_t*>(out + i * 128)[lane] = uint16_t(b0 | (b1 << 8));
    }
}

// Descriptor j of image i takes centre subset[i][hash % subsetSize], with
// subset[i][s] = hash(imageSeed, i, s) % nbCentres.
__global__ void k_imageCentres(const uint64_t* __restrict__ off, uint32_t nbImages, uint64_t firstImage,
                               uint64_t imageSeed, uint64_t descSeed, uint32_t nbCentres, uint32_t subsetSize,
                               uint32_t* __restrict__ ids) {
    const uint64_t n = off[nbImages] - off[0];
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t img = imageOf(off, nbImages, off[0] + i);
        const uint64_t gi = firstImage + img;
        const uint32_t s = uint32_t(h3(descSeed, gi, i - (off[img] - off[0])) % subsetSize);
        ids[i] = uint32_t(h3(imageSeed, gi, s) % nbCentres);
    }
}

// Posting j of image i: centre from the image's subset, leaf from the centre's cloud.
__global__ void k_imagePostings(const uint64_t* __restrict__ off, uint32_t nbImages, uint64_t firstImage,
                                uint64_t imageSeed, uint64_t postSeed, uint32_t nbCentres, uint32_t subsetSize,
                                const uint32_t* __restrict__ cloud, uint32_t cloudSize, uint32_t* __restrict__ leaf) {
    const uint64_t n = off[nbImages] - off[0];
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t img = imageOf(off, nbImages, off[0] + i);
        const uint64_t gi = firstImage + img;
        const uint64_t j = i - (off[img] - off[0]);
        const uint64_t r = h3(postSeed, gi, j);
        const uint32_t s = uint32_t((r >> 32) % subsetSize);
        const uint32_t c = uint32_t(h3(imageSeed, gi, s) % nbCentres);
        leaf[i] = cloud[uint64_t(c) * cloudSize + uint32_t(r) % cloudSize];
    }
}

uint32_t gridFor(uint64_t work, uint32_t perBlock, uint32_t cap) {
    const uint64_t g = (work + perBlock - 1) / perBlock;
    return uint32_t(g == 0 ? 1 : (g > cap ? cap : g));
}

}  // namespace
}  // namespace rbow_amd

using namespace rbow_amd;

extern "C" {

rbow_status rbow_synth_descriptors_device(const float* dev_centres, const float* dev_sigma, uint32_t nb_centres,
                                          uint64_t seed, uint64_t first, uint64_t n, const uint32_t* dev_centre_ids,
                                          uint8_t* dev_out, rbow_stream stream) {
    if (!n) return RBOW_OK;
    if (!dev_centres || !dev_sigma || !dev_out || !nb_centres) return RBOW_ERR_INVALID_ARG;
    k_synthDesc<<<gridFor(n, 4, 65536), 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
        dev_centres, dev_sigma, nb_centres, seed, first, n, dev_centre_ids, dev_out);
    return hipGetLastError() == hipSuccess ? RBOW_OK : RBOW_ERR_HIP;
}

rbow_status rbow_synth_image_centres_device(const uint64_t* dev_offsets, uint32_t nb_images, uint64_t first_image,
                                            uint64_t image_seed, uint64_t desc_seed, uint32_t nb_centres,
                                            uint32_t subset_size, uint32_t* dev_ids, uint64_t nb_desc,
                                            rbow_stream stream) {
    if (!nb_images || !nb_desc) return RBOW_OK;
    if (!dev_offsets || !dev_ids || !nb_centres || !subset_size) return RBOW_ERR_INVALID_ARG;
    k_imageCentres<<<gridFor(nb_desc, 256, 65536), 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
        dev_offsets, nb_images, first_image, image_seed, desc_seed, nb_centres, subset_size, dev_ids);
    return hipGetLastError() == hipSuccess ? RBOW_OK : RBOW_ERR_HIP;
}

rbow_status rbow_synth_image_postings_device(const uint64_t* dev_offsets, uint32_t nb_images, uint64_t first_image,
                                             uint64_t image_seed, uint64_t post_seed, uint32_t nb_centres,
                                             uint32_t subset_size, const uint32_t* dev_cloud, uint32_t cloud_size,
                                             uint32_t* dev_leaf, uint64_t nb_postings, rbow_stream stream) {
    if (!nb_images || !nb_postings) return RBOW_OK;
    if (!dev_offsets || !dev_cloud || !dev_leaf || !nb_centres || !subset_size || !cloud_size)
        return RBOW_ERR_INVALID_ARG;
    k_imagePostings<<<gridFor(nb_postings, 256, 65536), 256, 0, reinterpret_cast<hipStream_t>(stream)>>>(
        dev_offsets, nb_images, first_image, image_seed, post_seed, nb_centres, subset_size, dev_cloud, cloud_size,
        dev_leaf);
    return hipGetLastError() == hipSuccess ? RBOW_OK : RBOW_ERR_HIP;
}

}  // extern "C"
