// common.hpp -- shared internals of the RapidBoW MI355X library: errors, tree
// index math, device buffers.  Host + device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/rapidbow.h"

namespace rbow_amd {

constexpr uint32_t kDims = 128;                 // SIFT dims (RapidBoW.h:50)
constexpr uint32_t kIdxInvalid = 0xFFFFFFFFu;   // kmeans.cuh:31

// Error carried from the library internals to the C ABI, where it becomes a
// status code plus a thread-local message; nothing throws across the ABI.
struct Error : std::runtime_error {
    rbow_status status;
    Error(rbow_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define RBOW_HIP(expr)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw ::rbow_amd::Error(RBOW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RBOW_REQUIRE(cond, status, msg)                      \
    do {                                                     \
        if (!(cond)) throw ::rbow_amd::Error((status), (msg)); \
    } while (0)

// FullTree.h:33-61: breadth-first full tree.  levelBeg(l) = (b^l - 1)/(b - 1).
__host__ __device__ inline uint64_t levelBeg(uint32_t lg2b, uint32_t level) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < level; i++) acc = (acc << lg2b) + 1;
    return acc;
}
__host__ __device__ inline uint64_t levelSize(uint32_t lg2b, uint32_t level) {
    return uint64_t(1) << (lg2b * level);
}

inline uint32_t divUp(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

// Minimum over each row of 16 lanes (DPP: quad swaps, half-row and row mirrors).
template <int CTRL>
__device__ inline uint32_t dppU(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, 0xF, 0xF, false));
}
__device__ inline uint32_t min16u(uint32_t v) {
    v = min(v, dppU<0xB1>(v));   // quad_perm [1,0,3,2]
    v = min(v, dppU<0x4E>(v));   // quad_perm [2,3,0,1]
    v = min(v, dppU<0x141>(v));  // row_half_mirror
    v = min(v, dppU<0x140>(v));  // row_mirror
    return v;
}

// Minimum over lanes l, l ^ 16, l ^ 32, l ^ 48 (the four 16-lane rows of a wave), in every lane.
__device__ inline uint32_t minAcrossGroupsU(uint32_t m) {
    const auto p = __builtin_amdgcn_permlane16_swap(m, m, false, false);  // rows 0<->1, 2<->3
    m = min(uint32_t(p[0]), uint32_t(p[1]));
    const auto q = __builtin_amdgcn_permlane32_swap(m, m, false, false);  // halves
    return min(uint32_t(q[0]), uint32_t(q[1]));
}

// Owning device allocation.
template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void resize(size_t n) {
        if (n == n_) return;
        reset();
        if (n) {
            hipError_t e = hipMalloc(&p_, n * sizeof(T));
            if (e != hipSuccess) {
                p_ = nullptr;
                throw Error(e == hipErrorOutOfMemory ? RBOW_ERR_OUT_OF_MEMORY : RBOW_ERR_HIP,
                            std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " B): " + hipGetErrorString(e));
            }
        }
        n_ = n;
    }
    // grows only; keeps capacity
    void reserve(size_t n) {
        if (n > n_) resize(n);
    }
    void reset() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

}  // namespace rbow_amd
