// Shared helpers for the MI355X (gfx950) FoF + unbinding library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <stdexcept>
#include <string>

namespace pyfof {

// Error codes, same meanings as the reference's ierr (fof_fortran.f90:116-119,
// py_fof/__init__.py:338-348) plus codes for conditions only a device build has.
enum : int32_t {
    IERR_OK = 0,
    IERR_INVALID = 1,
    IERR_SPAN = 2,
    IERR_ALLOC = 3,
    IERR_NONFINITE = 4,
    IERR_THREADS = 10,
    IERR_HIP = 20,      // a HIP runtime call failed
    IERR_TOO_LARGE = 21 // a size beyond what the device layout supports
};

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const char* what) : std::runtime_error(what), code(c) {}
};

#define PYFOF_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            char buf_[512];                                                          \
            snprintf(buf_, sizeof(buf_), "%s failed at %s:%d: %s", #call, __FILE__,  \
                     __LINE__, hipGetErrorString(e_));                               \
            throw ::pyfof::HipError(e_, buf_);                                       \
        }                                                                            \
    } while (0)

#define PYFOF_LAUNCH_CHECK() PYFOF_HIP(hipGetLastError())

constexpr int kWave = 64;

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// MI355X deals the workgroups of a launch round-robin to its 8 XCDs, each
// with its own L2: workgroup b runs on XCD b % 8. xcd_block(b, n) maps b to a
// logical block so that every XCD works through one contiguous range of
// logical blocks (a bijection of [0, n)); kernels whose neighbouring blocks
// share data (cell rows, z-slabs) then find it in one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    const uint32_t q = n / 8, r = n % 8, x = b % 8, k = b / 8;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Order-preserving map of doubles onto uint64 so that atomicMin/Max work.
__host__ __device__ inline unsigned long long dbl_to_ordered(double v) {
    unsigned long long u;
    memcpy(&u, &v, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double ordered_to_dbl(unsigned long long u) {
    u = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
    double v;
    memcpy(&v, &u, 8);
    return v;
}

__device__ inline int atomic_load_relaxed(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace pyfof
