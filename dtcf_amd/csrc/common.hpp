// Shared helpers for the gfx950 kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "dtcf_local.h"

namespace dtcf {

// Sum of w[0], w[stride], ..., w[(n - 1) * stride] added in that order (the
// K slices of a split product), with the loads issued eight at a time.
__device__ inline double sum_in_order(const double *w, uint32_t n, int64_t stride)
{
    double sum = 0.0;
    uint32_t i = 0;
    for (; i + 8 <= n; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = w[(int64_t)(i + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) sum += v[u];
    }
    for (; i < n; i++) sum += w[(int64_t)i * stride];
    return sum;
}

// Thread-local last-error text, returned by dtcf_last_error().
void set_error(const std::string &msg);
const char *get_error();

// Map a HIP error to DTCF_ERR_HIP and record its text.
int hip_fail(hipError_t e, const char *what);

#define DTCF_HIP(call)                                              \
    do {                                                            \
        hipError_t e_ = (call);                                     \
        if (e_ != hipSuccess) return ::dtcf::hip_fail(e_, #call);   \
    } while (0)

#define DTCF_REQUIRE(cond, msg)                                     \
    do {                                                            \
        if (!(cond)) {                                              \
            ::dtcf::set_error(msg);                                 \
            return DTCF_ERR_INVALID;                                \
        }                                                           \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Unsigned 32-bit division by a run-time invariant divisor d >= 1
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994): l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1,  n / d = (mulhi(n, m) + n) >> l,
// with the sum formed in 64 bits so it cannot overflow.
struct FastDiv {
    uint32_t d = 1, m = 1, l = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t div) : d(div) {
        l = 0;
        while ((1ull << l) < div) l++;
        m = (uint32_t)(((((uint64_t)1 << 32) * (((uint64_t)1 << l) - div)) / div) + 1);
    }
    __host__ __device__ inline uint32_t div(uint32_t n) const {
        uint64_t t = ((uint64_t)n * m) >> 32;
        return (uint32_t)((t + n) >> l);
    }
    __host__ __device__ inline uint32_t mod(uint32_t n, uint32_t q) const { return n - q * d; }
};

}  // namespace dtcf
