// synth.cpp -- seeded SIFT-like descriptor generator for benchmarks and tests.
// No counterpart in the reference (its tests use raw std::mt19937_64 bytes
// or OpenCV SIFT on private images).  Deterministic: every descriptor has its
// own counter-derived SplitMix64 stream, so any slice can be generated
// independently and in parallel.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/rapidbow.h"

namespace {
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
struct SplitMix {
    uint64_t s;
    uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
    double uniform() { return (double((next() >> 11) + 1)) * (1.0 / 9007199254740992.0); }  // (0, 1]
    void normal2(double& a, double& b) {  // Box-Muller
        const double r = std::sqrt(-2.0 * std::log(uniform()));
        const double t = 6.283185307179586 * uniform();
        a = r * std::cos(t);
        b = r * std::sin(t);
    }
};
}  // namespace

extern "C" {

void rbow_synth_centres(uint64_t centre_seed, uint32_t nb_centres, float* out) {
    SplitMix g{mix64(centre_seed ^ 0xC3C3C3C3ull)};
    for (uint64_t i = 0; i < uint64_t(nb_centres) * 128; i += 2) {
        double a, b;
        g.normal2(a, b);
        out[i] = float(a);
        out[i + 1] = float(b);
    }
}

void rbow_synth_descriptors(uint64_t centre_seed, uint32_t nb_centres, uint64_t seed, uint64_t first, uint64_t n,
                            const uint32_t* centre_ids, float sigma_scale, uint8_t* out) {
    std::vector<float> c(uint64_t(nb_centres) * 128);
    rbow_synth_centres(centre_seed, nb_centres, c.data());
    std::vector<double> sigma(nb_centres);
    for (uint32_t j = 0; j < nb_centres; j++) {
        double s = 0;
        for (int k = 0; k < 128; k++) s += double(c[j * 128 + k]) * c[j * 128 + k];
        sigma[j] = double(sigma_scale) * std::sqrt(s) / std::sqrt(128.0);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(n); i++) {
        SplitMix g{mix64(seed ^ mix64(first + uint64_t(i)))};
        const uint32_t j = centre_ids ? centre_ids[i] % nb_centres : uint32_t(g.next() % nb_centres);
        double x[128];
        double nrm = 0;
        for (int k = 0; k < 128; k += 2) {
            double a, b;
            g.normal2(a, b);
            x[k] = std::fabs(double(c[j * 128 + k]) + sigma[j] * a);
            x[k + 1] = std::fabs(double(c[j * 128 + k + 1]) + sigma[j] * b);
            nrm += x[k] * x[k] + x[k + 1] * x[k + 1];
        }
        nrm = nrm > 0 ? std::sqrt(nrm) : 1.0;
        double nrm2 = 0;
        for (int k = 0; k < 128; k++) {
            x[k] = std::fmin(x[k] / nrm, 0.2);  // SIFT clip at 0.2
            nrm2 += x[k] * x[k];
        }
        nrm2 = nrm2 > 0 ? std::sqrt(nrm2) : 1.0;
        for (int k = 0; k < 128; k++) {
            const double v = std::nearbyint(512.0 * x[k] / nrm2);
            out[uint64_t(i) * 128 + k] = uint8_t(v > 255.0 ? 255.0 : v);
        }
    }
}

}  // extern "C"
