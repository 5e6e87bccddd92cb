// FP64 amplitude update, correlation energy and residual norm (new code:
// the reference stops at the contraction).  All three are HBM-bound
// streaming kernels.  The two reductions are deterministic: every workgroup
// covers a fixed slice of the data chosen from the problem size alone, sums
// it in a fixed tree order, and one final workgroup adds the partial sums in
// index order, so the result does not depend on timing or co-resident work.
#include <algorithm>
#include <map>
#include <mutex>

#include "common.hpp"

namespace dtcf {
namespace {

constexpr int RT = 256;          // threads per reduction workgroup
constexpr int RPARTS = 2048;     // partial sums (fixed: independent of device)

__device__ inline double block_sum(double v, double *sh)
{
    // wave64 tree with DPP-backed shuffles, then across the 4 waves via LDS.
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < RT / 64; w++) s += sh[w];
    }
    __syncthreads();
    return s;
}

// T[a,b,i,j] = R[a,b,i,j] / (eo[i] + eo[j] - ev[a] - ev[b]); layout [v,v,o,o].
__global__ __launch_bounds__(256) void denom_update_kernel(const double *__restrict__ R,
                                                           double *__restrict__ T,
                                                           const double *__restrict__ eo,
                                                           const double *__restrict__ ev,
                                                           uint32_t o, uint32_t v, FastDiv div_o,
                                                           FastDiv div_v, uint64_t total)
{
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total;
         x += (uint64_t)gridDim.x * blockDim.x) {
        // x = ((a*v + b)*o + i)*o + j ; the row index (a*v+b) may exceed 2^32
        // only for v > 65536, which the host rejects.
        const uint64_t ab = x / ((uint64_t)o * o);
        const uint32_t ij = (uint32_t)(x - ab * (uint64_t)o * o);
        const uint32_t i = div_o.div(ij), j = ij - i * o;
        const uint32_t a = div_v.div((uint32_t)ab), b = (uint32_t)ab - a * v;
        const double d = eo[i] + eo[j] - ev[a] - ev[b];
        T[x] = R[x] / d;
    }
}

// Partial sums of V[i,j,a,b] * T[a,b,i,j]: view T as P = v*v rows by
// Q = o*o columns and V as its transpose.  Workgroup w owns the 32x32 tiles
// w, w + RPARTS, ...; T tiles are read along q, V tiles along p, and V goes
// through LDS so both reads are coalesced.
__global__ __launch_bounds__(RT) void energy_partial_kernel(const double *__restrict__ V,
                                                            const double *__restrict__ T,
                                                            uint64_t P, uint64_t Q,
                                                            uint64_t tiles_q, uint64_t ntiles,
                                                            double *__restrict__ part)
{
    __shared__ double lds[32][33];
    __shared__ double sh[RT / 64];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    double acc = 0.0;
    for (uint64_t t = blockIdx.x; t < ntiles; t += RPARTS) {
        const uint64_t tp = t / tiles_q, tq = t - tp * tiles_q;
        const uint64_t p0 = tp * 32, q0 = tq * 32;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t q = q0 + ty + 8 * k, p = p0 + tx;  // V[q][p], lanes along p
            lds[ty + 8 * k][tx] = (q < Q && p < P) ? V[q * P + p] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t p = p0 + ty + 8 * k, q = q0 + tx;  // T[p][q], lanes along q
            if (p < P && q < Q) acc = fma(T[p * Q + q], lds[tx][ty + 8 * k], acc);
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RT) void sumsq_partial_kernel(const double *__restrict__ R,
                                                           uint64_t n, double *__restrict__ part)
{
    __shared__ double sh[RT / 64];
    double acc = 0.0;
    for (uint64_t x = (uint64_t)blockIdx.x * RT + threadIdx.x; x < n; x += (uint64_t)RPARTS * RT)
        acc = fma(R[x], R[x], acc);
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Final pass: one workgroup adds the RPARTS partials in a fixed order.
__global__ __launch_bounds__(RT) void finish_kernel(const double *__restrict__ part,
                                                    double scale, int do_sqrt,
                                                    double *__restrict__ out)
{
    __shared__ double sh[RT / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < RPARTS; i += RT) acc += part[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = do_sqrt ? sqrt(s) : scale * s;
}

// Partial-sum scratch, one buffer per (device, stream) so reductions on
// different streams never share it; calls on one stream are ordered by it.
std::mutex g_part_mu;
std::map<std::pair<int, hipStream_t>, double *> g_part;

int partials(hipStream_t s, double **p)
{
    int dev = 0;
    DTCF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_part_mu);
    auto key = std::make_pair(dev, s);
    auto it = g_part.find(key);
    if (it == g_part.end()) {
        double *buf = nullptr;
        DTCF_HIP(hipMalloc(&buf, RPARTS * sizeof(double)));
        it = g_part.emplace(key, buf).first;
    }
    *p = it->second;
    return DTCF_OK;
}

}  // namespace
}  // namespace dtcf

using namespace dtcf;

extern "C" int dtcf_denom_update_f64(const double *R, double *T, const double *eo,
                                     const double *ev, int64_t o, int64_t v, void *stream)
{
    DTCF_REQUIRE(o >= 0 && v >= 0, "denom: negative size");
    if (o == 0 || v == 0) return DTCF_OK;
    DTCF_REQUIRE(R && T && eo && ev, "denom: null argument");
    DTCF_REQUIRE(o < 65536 && v < 65536, "denom: o and v must be < 65536");
    const uint64_t total = (uint64_t)v * v * o * o;
    const unsigned grid = (unsigned)std::min<uint64_t>((total + 255) / 256, 256ull * 32);
    hipLaunchKernelGGL(denom_update_kernel, dim3(grid), dim3(256), 0, as_stream(stream), R, T, eo,
                       ev, (uint32_t)o, (uint32_t)v, FastDiv((uint32_t)o), FastDiv((uint32_t)v),
                       total);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

extern "C" int dtcf_energy_f64(const double *V, const double *T, int64_t o, int64_t v,
                               double *out_dev, void *stream)
{
    DTCF_REQUIRE(o >= 0 && v >= 0, "energy: negative size");
    DTCF_REQUIRE(V && T && out_dev, "energy: null argument");
    double *part = nullptr;
    int st = partials(as_stream(stream), &part);
    if (st != DTCF_OK) return st;
    const uint64_t P = (uint64_t)v * v, Q = (uint64_t)o * o;
    const uint64_t tq = (Q + 31) / 32, tp = (P + 31) / 32;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(energy_partial_kernel, dim3(RPARTS), dim3(RT), 0, s, V, T, P, Q, tq, tp * tq,
                       part);
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(RT), 0, s, part, 0.25, 0, out_dev);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

extern "C" int dtcf_norm2_f64(const double *R, int64_t n, double *out_dev, void *stream)
{
    DTCF_REQUIRE(n >= 0, "norm2: negative size");
    DTCF_REQUIRE(R && out_dev, "norm2: null argument");
    double *part = nullptr;
    int st = partials(as_stream(stream), &part);
    if (st != DTCF_OK) return st;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(sumsq_partial_kernel, dim3(RPARTS), dim3(RT), 0, s, R, (uint64_t)n, part);
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(RT), 0, s, part, 1.0, 1, out_dev);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}
