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
// A workgroup takes whole rows (a,b) of o*o elements: (a, b) once per row,
// (i, j) from a 32-bit division by o per element pair, 16-byte loads and
// stores when o*o is even.  The denominator keeps the left-to-right order
// ((eo_i + eo_j) - ev_a) - ev_b.
template <bool VEC>
__global__ __launch_bounds__(256) void denom_update_kernel(const double *__restrict__ R,
                                                           double *__restrict__ T,
                                                           const double *__restrict__ eo,
                                                           const double *__restrict__ ev,
                                                           uint32_t o, uint32_t v, FastDiv div_o,
                                                           FastDiv div_v, uint32_t rows)
{
    const uint32_t Q = o * o;
    for (uint32_t ab = blockIdx.x; ab < rows; ab += gridDim.x) {
        const uint32_t a = div_v.div(ab), b = ab - a * v;
        const double eva = ev[a], evb = ev[b];
        const double *r = R + (uint64_t)ab * Q;
        double *t = T + (uint64_t)ab * Q;
        if (VEC) {
            for (uint32_t q = 2 * threadIdx.x; q < Q; q += 512) {
                const double2 x = *reinterpret_cast<const double2 *>(r + q);
                const uint32_t i0 = div_o.div(q), j0 = q - i0 * o;
                const uint32_t i1 = div_o.div(q + 1), j1 = q + 1 - i1 * o;
                double2 y;
                y.x = x.x / (((eo[i0] + eo[j0]) - eva) - evb);
                y.y = x.y / (((eo[i1] + eo[j1]) - eva) - evb);
                *reinterpret_cast<double2 *>(t + q) = y;
            }
        } else {
            for (uint32_t q = threadIdx.x; q < Q; q += 256) {
                const uint32_t i = div_o.div(q), j = q - i * o;
                t[q] = r[q] / (((eo[i] + eo[j]) - eva) - evb);
            }
        }
    }
}

// Partial sums of V[i,j,a,b] * T[a,b,i,j]: view T as P = v*v rows by
// Q = o*o columns and V as its transpose.  Workgroup w owns the 64x64 tiles
// w, w + RPARTS, ...; the V tile goes through LDS read along p, the T tile
// is read along q, both with 16-byte loads where the extents allow.
template <bool VEC>
__global__ __launch_bounds__(RT) void energy_partial_kernel(const double *__restrict__ V,
                                                            const double *__restrict__ T,
                                                            uint64_t P, uint64_t Q,
                                                            uint64_t tiles_q, uint64_t ntiles,
                                                            double *__restrict__ part)
{
    __shared__ double lds[64][65];
    __shared__ double sh[RT / 64];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 pairs x 8 rows
    double acc = 0.0;
    for (uint64_t t = blockIdx.x; t < ntiles; t += RPARTS) {
        const uint64_t tp = t / tiles_q, tq = t - tp * tiles_q;
        const uint64_t p0 = tp * 64, q0 = tq * 64;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int ql = ty + 8 * k, pl = 2 * tx;
            const uint64_t q = q0 + ql, p = p0 + pl;  // V[q][p], lanes along p
            double x0 = 0.0, x1 = 0.0;
            if (q < Q) {
                const double *src = V + q * P + p;
                if (VEC && p + 1 < P) {
                    const double2 w = *reinterpret_cast<const double2 *>(src);
                    x0 = w.x;
                    x1 = w.y;
                } else {
                    if (p < P) x0 = src[0];
                    if (p + 1 < P) x1 = src[1];
                }
            }
            lds[ql][pl] = x0;
            lds[ql][pl + 1] = x1;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int pl = ty + 8 * k, ql = 2 * tx;
            const uint64_t p = p0 + pl, q = q0 + ql;  // T[p][q], lanes along q
            if (p >= P || q >= Q) continue;
            const double *src = T + p * Q + q;
            if (VEC && q + 1 < Q) {
                const double2 w = *reinterpret_cast<const double2 *>(src);
                acc = fma(w.x, lds[ql][pl], acc);
                acc = fma(w.y, lds[ql + 1][pl], acc);
            } else {
                acc = fma(src[0], lds[ql][pl], acc);
                if (q + 1 < Q) acc = fma(src[1], lds[ql + 1][pl], acc);
            }
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <bool VEC>
__global__ __launch_bounds__(RT) void sumsq_partial_kernel(const double *__restrict__ R,
                                                           uint64_t n, double *__restrict__ part)
{
    __shared__ double sh[RT / 64];
    double acc = 0.0;
    if (VEC) {  // 16-byte loads; an odd last element by thread 0 of workgroup 0
        const uint64_t n2 = n / 2;
        const double2 *R2 = reinterpret_cast<const double2 *>(R);
        for (uint64_t x = (uint64_t)blockIdx.x * RT + threadIdx.x; x < n2; x += (uint64_t)RPARTS * RT) {
            const double2 w = R2[x];
            acc = fma(w.x, w.x, acc);
            acc = fma(w.y, w.y, acc);
        }
        if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) acc = fma(R[n - 1], R[n - 1], acc);
    } else {
        for (uint64_t x = (uint64_t)blockIdx.x * RT + threadIdx.x; x < n; x += (uint64_t)RPARTS * RT)
            acc = fma(R[x], R[x], acc);
    }
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

static bool al16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

extern "C" int dtcf_denom_update_f64(const double *R, double *T, const double *eo,
                                     const double *ev, int64_t o, int64_t v, void *stream)
{
    DTCF_REQUIRE(o >= 0 && v >= 0, "denom: negative size");
    if (o == 0 || v == 0) return DTCF_OK;
    DTCF_REQUIRE(R && T && eo && ev, "denom: null argument");
    DTCF_REQUIRE(o < 65536 && v < 65536, "denom: o and v must be < 65536");
    const uint32_t rows = (uint32_t)(v * v);
    const unsigned grid = std::min<uint32_t>(rows, 256u * 32);
    const bool vec = (o * o) % 2 == 0 && al16(R) && al16(T);
    if (vec)
        hipLaunchKernelGGL(denom_update_kernel<true>, dim3(grid), dim3(256), 0, as_stream(stream), R,
                           T, eo, ev, (uint32_t)o, (uint32_t)v, FastDiv((uint32_t)o),
                           FastDiv((uint32_t)v), rows);
    else
        hipLaunchKernelGGL(denom_update_kernel<false>, dim3(grid), dim3(256), 0, as_stream(stream), R,
                           T, eo, ev, (uint32_t)o, (uint32_t)v, FastDiv((uint32_t)o),
                           FastDiv((uint32_t)v), rows);
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
    const uint64_t tq = (Q + 63) / 64, tp = (P + 63) / 64;
    hipStream_t s = as_stream(stream);
    if (P % 2 == 0 && Q % 2 == 0 && al16(V) && al16(T))
        hipLaunchKernelGGL(energy_partial_kernel<true>, dim3(RPARTS), dim3(RT), 0, s, V, T, P, Q, tq,
                           tp * tq, part);
    else
        hipLaunchKernelGGL(energy_partial_kernel<false>, dim3(RPARTS), dim3(RT), 0, s, V, T, P, Q, tq,
                           tp * tq, part);
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
    if (al16(R))
        hipLaunchKernelGGL(sumsq_partial_kernel<true>, dim3(RPARTS), dim3(RT), 0, s, R, (uint64_t)n, part);
    else
        hipLaunchKernelGGL(sumsq_partial_kernel<false>, dim3(RPARTS), dim3(RT), 0, s, R, (uint64_t)n, part);
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(RT), 0, s, part, 1.0, 1, out_dev);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}
