// Mixed-precision contractions with FP64 recombination, and device-side
// synthetic inputs.
//
// An FP64 operand (a dtcf_view, rows x K) is split once into low-precision
// slices of rows_padded x K_padded elements (zeros to multiples of 128, so
// the GEMM needs no bounds checks), K contiguous -- for BF16, FP16 and INT8
// in fragment-major order over the bytes of a row (i8_frag_offset):
//   F32   : one FP32 slice                              (FP32 MFMA)
//   BF16x3: x = s1 + s2 + s3 + r, s_i = bf16(rest)      (BF16 MFMA, FP32 accum)
//   F16x3 : the same in FP16 after scaling each row by a power of two so its
//           largest magnitude is near 2^14 (keeps the slices normal)
//   I8xS  : each row scaled to integers X = rint(x * 2^(8S - 2 - e)) (e the
//           exponent above the row's max |x|) and X written in base 256 with
//           balanced signed digits q_1..q_S in [-128, 127] (q_1 within +-65):
//           x = 2^(e + 2) * sum_i q_i 2^(-8i) + O(2^(e + 1 - 8S))
//           (INT8 MFMA, INT32 accum)
// A product of slices (i, j) is one low-precision GEMM whose accumulators
// are converted to FP64, scaled (2^(ea_r + eb_c + 4 - 8(i+j)) for I8,
// 2^(ea_r + eb_c) for F16) and added to the FP64 C in the epilogue.  INT8
// products are exact in INT32 for K < 2^31 / 128^2 = 131,072, so the only I8
// errors are the rounding to 8S - 2 bits, dropped levels and the FP64
// recombination (the Ozaki scheme of Ozaki,
// Ogita, Oishi and Rump, "Error-free transformations of matrix
// multiplication by using fast routines of matrix multiplication and its
// applications", Numer. Algorithms 59 (2012), with integer slices).
#include <algorithm>
#include <cmath>
#include <vector>

#include <climits>

#include "common.hpp"
#include "gemm_f64.hpp"
#include "view.hpp"

namespace dtcf {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------ synth
__device__ inline uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Same counter-based recipe as dtcf_amd/synth.py (splitmix64 -> Box-Muller).
__global__ void fill_normal_kernel(double *out, int64_t n, uint64_t seed, double sigma,
                                   int64_t start)
{
    const uint64_t base = seed << 40;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = (uint64_t)(start + i), pair = e >> 1;
        const uint64_t h1 = splitmix64(base + (pair << 1));
        const uint64_t h2 = splitmix64(base + (pair << 1) + 1);
        const double u1 = ((double)(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1)), t = 2.0 * M_PI * u2;
        out[i] = sigma * ((e & 1) ? r * sin(t) : r * cos(t));
    }
}

__global__ void fill_uniform_kernel(double *out, int64_t n, uint64_t seed, double lo, double hi)
{
    const uint64_t base = seed << 40;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t h = splitmix64(base + (uint64_t)i + (1ull << 39));
        out[i] = lo + (hi - lo) * ((double)(h >> 11) * (1.0 / 9007199254740992.0));
    }
}

// out = (x - x^{01} - x^{23} + x^{01,23}) / 2 for a [d0,d1,d2,d3] tensor.
__global__ void antisym4_kernel(const double *__restrict__ x, double *__restrict__ out, int64_t d0,
                                int64_t d1, int64_t d2, int64_t d3)
{
    const int64_t n = d0 * d1 * d2 * d3;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e;
        const int64_t l = t % d3;
        t /= d3;
        const int64_t k = t % d2;
        t /= d2;
        const int64_t j = t % d1;
        const int64_t i = t / d1;
        auto at = [&](int64_t a, int64_t b, int64_t c, int64_t d) {
            return x[((a * d1 + b) * d2 + c) * d3 + d];
        };
        out[e] = 0.5 * (at(i, j, k, l) - at(j, i, k, l) - at(i, j, l, k) + at(j, i, l, k));
    }
}

// ------------------------------------------------------------------ split
struct ViewArgs {
    const int64_t *rofs, *cofs, *blk;
    FastDiv div_rpb, div_cpb;
    int64_t ncb;
};

__device__ inline int64_t vaddr(const ViewArgs &v, int64_t r, int64_t c)
{
    int64_t a = v.rofs[r] + v.cofs[c];
    if (v.blk) a += v.blk[(int64_t)v.div_rpb.div((uint32_t)r) * v.ncb + v.div_cpb.div((uint32_t)c)];
    return a;
}

// One wave per row: max |x| (for the scaled modes), then the slices.
// INT8 digit planes are stored fragment-major: the 16 rows x 64 bytes of
// K of row block rb and K step kt form one contiguous KiB at
// ((rb * (K_p / 64) + kt) * 64 + lane) * 16, lane = row % 16 + 16 * kq
// holding bytes 16 kq .. 16 kq + 15 of that row's step -- the MFMA
// operand order, so one direct 16-byte-per-lane load moves whole 128-byte
// lines straight into an LDS fragment.
__host__ __device__ inline int64_t i8_frag_offset(int64_t row, int64_t c, int64_t K_p)
{
    const int64_t rb = row >> 4, r = row & 15, kt = c >> 6, kc = c & 63;
    return (((rb * (K_p >> 6) + kt) << 6) + r + 16 * (kc >> 4)) * 16 + (kc & 15);
}

// Balanced base-256 digits of X = rint(x 2^sh) (|X| < 2^46), least
// significant first: q = ((X + 128) mod 256) - 128 = X - 256 Y and the
// next X = Y = floor((X + 128) / 256).  Every step is exact in FP64 (X is
// an integer below 2^53; 1/256 is a power of two), which keeps the digit
// loop on full-rate FP64 instructions instead of emulated 64-bit integer
// ones.  q[s] is digit s, s = 0 the most significant.
__device__ __forceinline__ void i8_digits(double x, int sh, int nslice, int (&q)[6])
{
    double X = rint(ldexp(x, sh));
#pragma unroll
    for (int s = 5; s >= 0; s--) {
        if (s >= nslice) continue;
        const double Y = floor((X + 128.0) * (1.0 / 256.0));
        q[s] = (int)(X - 256.0 * Y);
        X = Y;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void split_kernel(const double *__restrict__ src, ViewArgs v,
                                                   int64_t rows, int64_t K, int64_t rows_p,
                                                   int64_t K_p, int nslice, void *__restrict__ dst,
                                                   int32_t *__restrict__ rexp)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows_p) return;
    const bool real = row < rows;
    int e = 0;
    if (MODE == DTCF_PREC_F16X3 || MODE == DTCF_PREC_I8) {
        double m = 0.0;
        for (int64_t c = lane; real && c < K; c += 64) m = fmax(m, fabs(src[vaddr(v, row, c)]));
        for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        if (m > 0.0) {
            int ex;
            frexp(m, &ex);  // m = f * 2^ex, f in [0.5, 1): m < 2^ex
            e = ex;
        }
        if (MODE == DTCF_PREC_F16X3) e -= 14;  // scaled rows have max |x| < 2^14
        if (lane == 0) rexp[row] = e;
    }
    const size_t plane = (size_t)rows_p * K_p;
    for (int64_t c = lane; c < K_p; c += 64) {
        const double x = (real && c < K) ? src[vaddr(v, row, c)] : 0.0;
        // F32 row-major; BF16 / F16 fragment-major over the bytes of a row
        const size_t o = MODE == DTCF_PREC_F32 ? (size_t)row * K_p + c
                                               : (size_t)(i8_frag_offset(row, 2 * c, 2 * K_p) >> 1);
        if constexpr (MODE == DTCF_PREC_F32) {
            ((float *)dst)[o] = (float)x;
        } else if constexpr (MODE == DTCF_PREC_BF16X3) {
            __bf16 *d = (__bf16 *)dst;
            double r = x;
            for (int s = 0; s < 3; s++) {
                const __bf16 q = (__bf16)(float)r;  // nearest BF16 of r, ties to even
                d[s * plane + o] = q;
                r -= (double)(float)q;
            }
        } else if constexpr (MODE == DTCF_PREC_F16X3) {
            _Float16 *d = (_Float16 *)dst;
            double r = ldexp(x, -e);
            for (int s = 0; s < 3; s++) {
                const _Float16 q = (_Float16)(float)r;
                d[s * plane + o] = q;
                r -= (double)(float)q;
            }
        } else {  // I8
            // X = rint(x * 2^(8 nslice - 2 - e)), |X| < 2^(8 nslice - 2) (exact in
            // FP64 for nslice <= 6), written in base 256 with balanced digits
            // in [-128, 127] from the least significant up; the top digit is
            // then within +-65.
            int8_t *d = (int8_t *)dst;
            const int64_t of = i8_frag_offset(row, c, K_p);
            int q[6] = {0, 0, 0, 0, 0, 0};
            i8_digits(x, 8 * nslice - 2 - e, nslice, q);
            for (int s = 0; s < nslice; s++) d[s * plane + of] = (int8_t)q[s];
        }
    }
}

// INT8 split of a table view (no block table) in two K-parallel passes.
// A row-parallel kernel that gives one workgroup all of K of a 16-row block
// leaves most of the GPU idle on an operand of a few thousand rows; here
// every workgroup takes a 16-row block by a slice of K:
//   i8_exp_kernel: 16 rows x 1024 columns; the largest frexp exponent of
//     its elements per row, merged into rexp[row] by an integer atomic max
//     (order-independent; the same exponent as frexp of the row maximum).
//     rexp starts at INT_MIN (i8_exp_init); a row that stays there is all
//     zeros and reads as exponent 0.  Eight loads in flight per lane.
//   i8_digit_kernel: 16 rows x 4 K steps, one per wave; lane l = r + 16 kq
//     converts row r's elements 16 kq .. 16 kq + 15 of its K step (its own
//     lane of the fragment) and stores 16 bytes per plane.
// PAIRS views (column pairs contiguous and 16-byte aligned) read rows along
// K; a wave stages its K step (16 rows x 64 columns) through LDS with
// 16-byte loads of 2 rows x 512 contiguous bytes each.  Other views (RUN)
// read per lane, 16 consecutive rows per load when rows run contiguously
// in 16s (rowrun16), a gather otherwise.  The source is read once per pass
// and the planes written once: both stream past the caches (nontemporal).
// Same digits as split_kernel, bit for bit.
__device__ __forceinline__ double ld_nt(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ d2 ld_nt2(const double *p) { return __builtin_nontemporal_load((const d2 *)p); }
__device__ __forceinline__ void st_nt(uint4 *p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    typedef uint32_t u4v __attribute__((ext_vector_type(4)));
    const u4v w = {a, b, c, d};
    __builtin_nontemporal_store(w, (u4v *)p);
}

// Digits of 16 elements by integer byte extraction: with X = rint(x 2^sh)
// (|X| <= 2^(8 NS - 2)) and B = sum_{s < NS} 128 * 256^s, U = X + B is an
// integer in [0, 256^NS), exact in FP64, whose base-256 digits are the
// balanced digits of X plus 128 (the balanced representation is unique).
// Digit plane p (p = 0 the most significant) of element j is byte NS-1-p
// of U, xor 0x80; wd[p][k] packs elements 4k .. 4k+3.
template <int NS>
__device__ __forceinline__ void i8_digit_words(const double (&xs)[16], int sh, uint32_t (&wd)[6][4])
{
    double B = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) B = B * 256.0 + 128.0;
    uint32_t lo[16], hi[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const double U = rint(ldexp(xs[j], sh)) + B;
        if constexpr (NS <= 4) {
            lo[j] = (uint32_t)U;
            hi[j] = 0;
        } else {
            const double h = floor(U * 0x1p-32);
            lo[j] = (uint32_t)fma(h, -0x1p32, U);
            hi[j] = (uint32_t)h;
        }
    }
#pragma unroll
    for (int p = 0; p < NS; p++) {
        const int b = NS - 1 - p;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t w = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t u = b < 4 ? lo[4 * k + j] : hi[4 * k + j];
                w |= ((u >> (8 * (b & 3))) & 255u) << (8 * j);
            }
            wd[p][k] = w ^ 0x80808080u;
        }
    }
}

__global__ __launch_bounds__(256) void i8_exp_init(int32_t *__restrict__ rexp, int64_t rows_p)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < rows_p; r += (int64_t)gridDim.x * 256)
        rexp[r] = INT_MIN;
}

template <bool PAIRS>
__global__ __launch_bounds__(256) void i8_exp_kernel(const double *__restrict__ src, ViewArgs v,
                                                     int64_t rows, int64_t K, int64_t nrb,
                                                     int32_t *__restrict__ rexp)
{
    constexpr int KC = 1024;
    const int t = threadIdx.x;
    const int64_t rb = (int64_t)blockIdx.x % nrb, ch = (int64_t)blockIdx.x / nrb;  // row block fastest
    // PAIRS: lanes along a row's columns; else lanes along the 16 rows of a column
    const int r = PAIRS ? (t >> 4) : (t & 15), cg = PAIRS ? (t & 15) : (t >> 4);
    const int64_t row = rb * 16 + r;
    const int64_t c0 = ch * KC;
    const int64_t cend = min(K, c0 + KC);
    double m = 0.0;
    if (row < rows) {
        const double *x = src + v.rofs[row];
        if (PAIRS) {  // 16-byte loads of (c, c + 1); K is even
            for (int64_t cb = c0 + 2 * cg; cb < cend; cb += 256) {
                int64_t o[8];
#pragma unroll
                for (int u = 0; u < 8; u++) o[u] = cb + 32 * u < cend ? v.cofs[cb + 32 * u] : -1;
                d2 a[8];
#pragma unroll
                for (int u = 0; u < 8; u++) a[u] = o[u] >= 0 ? ld_nt2(x + o[u]) : d2{0.0, 0.0};
#pragma unroll
                for (int u = 0; u < 8; u++) m = fmax(m, fmax(fabs(a[u].x), fabs(a[u].y)));
            }
        } else {
            for (int64_t cb = c0 + cg; cb < cend; cb += 128) {
                int64_t o[8];
#pragma unroll
                for (int u = 0; u < 8; u++) o[u] = cb + 16 * u < cend ? v.cofs[cb + 16 * u] : -1;
                double a[8];
#pragma unroll
                for (int u = 0; u < 8; u++) a[u] = o[u] >= 0 ? ld_nt(x + o[u]) : 0.0;
#pragma unroll
                for (int u = 0; u < 8; u++) m = fmax(m, fabs(a[u]));
            }
        }
    }
    // reduce over the threads of one row: 16 lanes (RUN: every 16th lane)
    if (PAIRS) {
        for (int off = 8; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    } else {
        m = fmax(m, __shfl_xor(m, 16, 64));
        m = fmax(m, __shfl_xor(m, 32, 64));
    }
    __shared__ double sm[4][16];
    if (PAIRS) {
        if (cg == 0) sm[0][r] = m;
    } else {
        if ((t & 63) < 16) sm[t >> 6][r] = m;
    }
    __syncthreads();
    if (t < 16) {
        const int64_t rr = rb * 16 + t;
        const double mm = PAIRS ? sm[0][t] : fmax(fmax(sm[0][t], sm[1][t]), fmax(sm[2][t], sm[3][t]));
        if (rr < rows && mm > 0.0) {
            int e;
            frexp(mm, &e);
            atomicMax(&rexp[rr], e);
        }
    }
}

template <bool PAIRS, int NS>
__global__ __launch_bounds__(256) void i8_digit_kernel(const double *__restrict__ src, ViewArgs v,
                                                       int64_t rows, int64_t K, int64_t rows_p,
                                                       int64_t K_p, int64_t nrb,
                                                       int8_t *__restrict__ dst,
                                                       int32_t *__restrict__ rexp)
{
    constexpr int S = 66;  // doubles per staged row: 16-byte aligned, padded
    __shared__ __attribute__((aligned(16))) double tile[PAIRS ? 4 : 1][PAIRS ? 16 * S : 2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, kq = lane >> 4;
    const int64_t rb = (int64_t)blockIdx.x % nrb, ch = (int64_t)blockIdx.x / nrb, nk = K_p >> 6;
    const int64_t kt = ch * 4 + w;
    if (kt >= nk) return;  // whole waves leave: no workgroup barrier below
    const int64_t row = rb * 16 + r;
    int e = rexp[row];
    if (e == INT_MIN) {  // all-zero row (or padding): exponent 0
        e = 0;
        if (ch == 0 && w == 0 && kq == 0) rexp[row] = 0;
    }
    const int sh = 8 * NS - 2 - e;
    double xs[16];
    if constexpr (PAIRS) {
        double *tw = tile[w];
        const int cp = lane & 31, rh = lane >> 5;
        const int64_t col = kt * 64 + 2 * cp;
        const int64_t co = col < K ? v.cofs[col] : 0;
        d2 a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int64_t rr = rb * 16 + 2 * u + rh;
            a[u] = (rr < rows && col < K) ? ld_nt2(src + v.rofs[rr] + co) : d2{0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < 8; u++) *(d2 *)(tw + (2 * u + rh) * S + 2 * cp) = a[u];
        // the tile is this wave's own: LDS accesses of one wave complete in
        // order, so only the compiler must keep the reads below the writes
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
            const d2 b = *(const d2 *)(tw + r * S + 16 * kq + j);
            xs[j] = b.x;
            xs[j + 1] = b.y;
        }
    } else {
        const int64_t c = kt * 64 + 16 * kq;
        if (row < rows) {
            const double *x = src + v.rofs[row];
#pragma unroll
            for (int j = 0; j < 16; j++) xs[j] = c + j < K ? ld_nt(x + v.cofs[c + j]) : 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) xs[j] = 0.0;
        }
    }
    uint32_t wd[6][4];
    i8_digit_words<NS>(xs, sh, wd);
    const size_t plane16 = (size_t)rows_p * K_p / 16;
    uint4 *out = (uint4 *)dst + (rb * nk + kt) * 64 + lane;
#pragma unroll
    for (int p = 0; p < NS; p++) st_nt(out + p * plane16, wd[p][0], wd[p][1], wd[p][2], wd[p][3]);
}

// Three BF16 / FP16 slices of a table view (no block table), K-parallel:
// 16 rows x 4 K steps (of 32 elements, 64 bytes) per workgroup, a wave per
// K step; lane l = r + 16 kq converts row r's elements 8 kq .. 8 kq + 7 of
// its step (its own lane of the fragment) and stores 16 bytes per slice
// plane.  Column-pair views stage the step through LDS (16-byte loads of 4
// rows x 256 contiguous bytes); other views read per lane.  FP16 rows are
// scaled by 2^-rexp[row], rexp from i8_exp_kernel and f16_rexp_finish (the
// frexp exponent of the row maximum less 14).  Same slices as split_kernel,
// bit for bit.
__global__ __launch_bounds__(256) void f16_rexp_finish(int32_t *__restrict__ rexp, int64_t rows_p)
{
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < rows_p; r += (int64_t)gridDim.x * 256) {
        const int e = rexp[r];
        rexp[r] = (e == INT_MIN ? 0 : e) - 14;  // scaled rows have max |x| < 2^14
    }
}

template <int MODE, bool PAIRS>
__global__ __launch_bounds__(256) void f16_slices_kernel(const double *__restrict__ src, ViewArgs v,
                                                         int64_t rows, int64_t K, int64_t rows_p,
                                                         int64_t K_p, int64_t nrb,
                                                         unsigned char *__restrict__ dst,
                                                         const int32_t *__restrict__ rexp)
{
    static_assert(MODE == DTCF_PREC_BF16X3 || MODE == DTCF_PREC_F16X3, "two-byte slices");
    constexpr int S = 34;  // doubles per staged row: 16-byte aligned, padded
    __shared__ __attribute__((aligned(16))) double tile[PAIRS ? 4 : 1][PAIRS ? 16 * S : 2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, kq = lane >> 4;
    const int64_t rb = (int64_t)blockIdx.x % nrb, ch = (int64_t)blockIdx.x / nrb, nk = K_p >> 5;
    const int64_t kt = ch * 4 + w;
    if (kt >= nk) return;  // whole waves leave: no workgroup barrier below
    const int64_t row = rb * 16 + r;
    double xs[8];
    if constexpr (PAIRS) {
        double *tw = tile[w];
        const int cp = lane & 15, rq = lane >> 4;
        const int64_t col = kt * 32 + 2 * cp;
        const int64_t co = col < K ? v.cofs[col] : 0;
        d2 a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t rr = rb * 16 + 4 * u + rq;
            a[u] = (rr < rows && col < K) ? ld_nt2(src + v.rofs[rr] + co) : d2{0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) *(d2 *)(tw + (4 * u + rq) * S + 2 * cp) = a[u];
        // the tile is this wave's own: LDS accesses of one wave complete in
        // order, so only the compiler must keep the reads below the writes
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            const d2 b = *(const d2 *)(tw + r * S + 8 * kq + j);
            xs[j] = b.x;
            xs[j + 1] = b.y;
        }
    } else {
        const int64_t c = kt * 32 + 8 * kq;
        if (row < rows) {
            const double *x = src + v.rofs[row];
#pragma unroll
            for (int j = 0; j < 8; j++) xs[j] = c + j < K ? ld_nt(x + v.cofs[c + j]) : 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) xs[j] = 0.0;
        }
    }
    const int e = MODE == DTCF_PREC_F16X3 ? rexp[row] : 0;
    uint32_t wd[3][4] = {};
#pragma unroll
    for (int j = 0; j < 8; j++) {
        double rr = MODE == DTCF_PREC_F16X3 ? ldexp(xs[j], -e) : xs[j];
#pragma unroll
        for (int sl = 0; sl < 3; sl++) {
            uint16_t bits;
            if constexpr (MODE == DTCF_PREC_BF16X3) {
                const __bf16 q = (__bf16)(float)rr;  // nearest BF16 of the float of rr
                bits = __builtin_bit_cast(uint16_t, q);
                rr -= (double)(float)q;
            } else {
                const _Float16 q = (_Float16)(float)rr;
                bits = __builtin_bit_cast(uint16_t, q);
                rr -= (double)(float)q;
            }
            wd[sl][j >> 1] |= (uint32_t)bits << (16 * (j & 1));
        }
    }
    const size_t plane16 = (size_t)rows_p * K_p * 2 / 16;
    uint4 *out = (uint4 *)dst + (rb * nk + kt) * 64 + lane;
#pragma unroll
    for (int sl = 0; sl < 3; sl++) st_nt(out + sl * plane16, wd[sl][0], wd[sl][1], wd[sl][2], wd[sl][3]);
}

// Wide operands (few rows, long K: the T1 statements' A operands are a few
// hundred rows by millions of K) leave the row-parallel split_kernel with a
// few hundred waves.  There the float splits (and INT8 splits of views
// with a block table) run in two K-parallel launches: one block per row
// reduces the row maximum to its exponent, then every thread writes one
// 16-byte chunk of each slice plane (16 INT8, 8 BF16 / FP16 or, row-major,
// 4 FP32 elements) -- the same values, bit for bit, as split_kernel.
template <int MODE>
__global__ __launch_bounds__(512) void row_exp_kernel(const double *__restrict__ src, ViewArgs v,
                                                      int64_t rows, int64_t K,
                                                      int32_t *__restrict__ rexp)
{
    __shared__ double part[8];
    const int64_t row = blockIdx.x;
    double m = 0.0;
    if (row < rows)
        for (int64_t c = threadIdx.x; c < K; c += 512) m = fmax(m, fabs(src[vaddr(v, row, c)]));
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 8; w++) m = fmax(m, part[w]);
        int e = 0;
        if (m > 0.0) frexp(m, &e);  // m < 2^e
        if (MODE == DTCF_PREC_F16X3) e -= 14;
        rexp[row] = e;
    }
}

// RUN (rows contiguous in runs of 16, K strided): sixteen consecutive
// threads take the 16 rows of one chunk, so each of their loads reads one
// whole cache line of 16 rows; otherwise consecutive threads take
// consecutive chunks of one row (K contiguous).
template <int MODE, bool RUN>
__global__ __launch_bounds__(256) void split_wide_kernel(const double *__restrict__ src, ViewArgs v,
                                                         int64_t rows, int64_t K, int64_t rows_p,
                                                         int64_t K_p, int nslice,
                                                         unsigned char *__restrict__ dst,
                                                         const int32_t *__restrict__ rexp)
{
    constexpr int ESZ = MODE == DTCF_PREC_F32 ? 4 : MODE == DTCF_PREC_I8 ? 1 : 2;
    constexpr int PER = 16 / ESZ;  // elements per 16-byte chunk
    const int64_t cpr = K_p / PER, total = rows_p * cpr;
    const size_t plane = (size_t)rows_p * K_p * ESZ;  // bytes
    for (int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * 256) {
        int64_t row, c0;
        if constexpr (RUN) {
            const int64_t rest = u >> 4, rb = rest / cpr;
            row = rb * 16 + (u & 15);
            c0 = (rest - rb * cpr) * PER;
        } else {
            row = u / cpr;
            c0 = (u - row * cpr) * PER;
        }
        const bool real = row < rows;
        double x[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) x[j] = (real && c0 + j < K) ? src[vaddr(v, row, c0 + j)] : 0.0;
        if constexpr (MODE == DTCF_PREC_F32) {
            float4 f = make_float4((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
            *(float4 *)(dst + ((size_t)row * K_p + c0) * 4) = f;
        } else {
            const size_t off = (size_t)i8_frag_offset(row, c0 * ESZ, K_p * ESZ);  // chunk start
            uint32_t wd[6][4] = {};
            if constexpr (MODE == DTCF_PREC_I8) {
                const int sh = 8 * nslice - 2 - rexp[row];
#pragma unroll
                for (int j = 0; j < PER; j++) {
                    int q[6] = {0, 0, 0, 0, 0, 0};
                    i8_digits(x[j], sh, nslice, q);
#pragma unroll
                    for (int sl = 0; sl < 6; sl++)
                        wd[sl][j >> 2] |= (uint32_t)(q[sl] & 255) << (8 * (j & 3));
                }
            } else {
                const int e = MODE == DTCF_PREC_F16X3 ? rexp[row] : 0;
#pragma unroll
                for (int j = 0; j < PER; j++) {
                    double r = MODE == DTCF_PREC_F16X3 ? ldexp(x[j], -e) : x[j];
#pragma unroll
                    for (int sl = 0; sl < 3; sl++) {
                        uint16_t bits;
                        if constexpr (MODE == DTCF_PREC_BF16X3) {
                            const __bf16 q = (__bf16)(float)r;  // as split_kernel
                            bits = __builtin_bit_cast(uint16_t, q);
                            r -= (double)(float)q;
                        } else {
                            const _Float16 q = (_Float16)(float)r;
                            bits = __builtin_bit_cast(uint16_t, q);
                            r -= (double)(float)q;
                        }
                        wd[sl][j >> 1] |= (uint32_t)bits << (16 * (j & 1));
                    }
                }
            }
            const int ns = MODE == DTCF_PREC_I8 ? nslice : 3;
#pragma unroll
            for (int sl = 0; sl < 6; sl++)
                if (sl < ns)
                    *(uint4 *)(dst + sl * plane + off) =
                        make_uint4(wd[sl][0], wd[sl][1], wd[sl][2], wd[sl][3]);
        }
    }
}

// ------------------------------------------------------------------ GEMM
// Tiles 128 x 128 x (128 bytes of K), 256 threads, 4 waves of 64 x 64 as
// 4 x 4 MFMA 16x16 accumulators.  Operand slices are [rows_p][K_p] with K
// contiguous (INT8 digit planes fragment-major, read as such).  LDS is
// fragment-major: [row/16][k step][lane][VB bytes], lane = row%16 + 16*kq,
// each wave fragment read one contiguous span.
template <int PREC>
struct LP;
template <>
struct LP<DTCF_PREC_F32> {
    typedef float T;
    static constexpr int VB = 4;    // bytes per lane per k step
    typedef f4 Acc;
    typedef float Frag;
    __device__ static inline Acc mfma(Frag a, Frag b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <>
struct LP<DTCF_PREC_BF16X3> {
    typedef __bf16 T;
    static constexpr int VB = 16;
    typedef f4 Acc;
    typedef bf8 Frag;
    __device__ static inline Acc mfma(Frag a, Frag b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct LP<DTCF_PREC_F16X3> {
    typedef _Float16 T;
    static constexpr int VB = 16;
    typedef f4 Acc;
    typedef h8 Frag;
    __device__ static inline Acc mfma(Frag a, Frag b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct LP<DTCF_PREC_I8> {
    typedef int8_t T;
    static constexpr int VB = 16;
    typedef i4 Acc;
    typedef i4 Frag;
    __device__ static inline Acc mfma(Frag a, Frag b, Acc c)
    {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
};

struct LPArgs {
    const void *A, *B;       // slice planes
    const int32_t *ea, *eb;  // row exponents (scaled modes) or null
    int shift;               // extra power of two subtracted (7*(i+j) for I8)
    double alpha;
    int64_t M, N, Kp;        // logical rows / cols, padded K
    int64_t Mp, Np;          // padded rows (plane strides)
    double *C;               // FP64 output
    ViewArgs cv;             // C view (when C_VIEW)
    int64_t ldc;
    uint32_t tiles_n;
};

constexpr int LBM = 128, LBN = 128, LBYTES = 128;  // K bytes per tile row

template <int PREC, bool CVIEW>
__global__ __launch_bounds__(256, 2) void lp_gemm_kernel(LPArgs g)
{
    typedef LP<PREC> P;
    constexpr int VB = P::VB;
    constexpr int NSTEP = LBYTES / (4 * VB);     // k steps per tile (4 kq lanes per step)
    constexpr int FRAG = 64 * VB;                // bytes per (group, step) fragment
    constexpr int OPB = (LBM / 16) * NSTEP * FRAG;  // bytes per operand tile = 16 KiB
    __shared__ __attribute__((aligned(16))) unsigned char smem[2 * 2 * OPB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const uint32_t tm = blockIdx.x / g.tiles_n, tn = blockIdx.x - tm * g.tiles_n;
    const int64_t m0 = (int64_t)tm * LBM, n0 = (int64_t)tn * LBN;
    const int64_t kb = g.Kp * (int64_t)sizeof(typename P::T);  // bytes per slice row
    const unsigned char *Ab = (const unsigned char *)g.A + m0 * kb;
    const unsigned char *Bb = (const unsigned char *)g.B + n0 * kb;

    // Global 16-byte chunks: pass i (0..3), wave w, lane l ->
    // row = 16*g + l%16 with g = 2*w + i/2, byte 64*(i%2) + 16*(l/16) of the
    // tile row.  Each chunk holds 16/VB fragments' worth of lanes.
    // INT8, BF16 and F16 planes are fragment-major over the bytes of a row
    // (i8_frag_offset): the same chunk is lane l's 16 bytes of fragment (row
    // block m0/16 + grp, 64-byte step 2 kt + i%2).
    uint4 ra[4], rb[4];
    auto gload = [&](int64_t kt) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int grp = 2 * wave + (i >> 1);
            if constexpr (VB == 16) {  // INT8, BF16, F16: fragment-major planes
                const int64_t nk64 = kb >> 6;
                const int64_t fa = (((m0 >> 4) + grp) * nk64 + 2 * kt + (i & 1)) * 1024 + 16 * lane;
                const int64_t fb = (((n0 >> 4) + grp) * nk64 + 2 * kt + (i & 1)) * 1024 + 16 * lane;
                ra[i] = *(const uint4 *)((const unsigned char *)g.A + fa);
                rb[i] = *(const uint4 *)((const unsigned char *)g.B + fb);
            } else {
                const int64_t off = (int64_t)(16 * grp + (lane & 15)) * kb + kt * LBYTES +
                                    64 * (i & 1) + 16 * (lane >> 4);
                ra[i] = *(const uint4 *)(Ab + off);
                rb[i] = *(const uint4 *)(Bb + off);
            }
        }
    };
    auto sstore = [&](unsigned char *sa, unsigned char *sb) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int grp = 2 * wave + (i >> 1);
            const int byte = 64 * (i & 1) + 16 * (lane >> 4);  // byte in the 128-byte row
            const int r = lane & 15;
            if constexpr (VB == 16) {
                // one chunk = one lane's fragment: step = byte/64, kq = (byte%64)/16
                const int step = byte >> 6, kq = (byte & 63) >> 4;
                const int o = ((grp * NSTEP + step) * 64 + r + 16 * kq) * VB;
                *(uint4 *)(sa + o) = ra[i];
                *(uint4 *)(sb + o) = rb[i];
            } else {
                // VB == 4: the chunk holds 4 consecutive k = one k step's 4 kq lanes
                const int step = byte >> 4;
                const uint32_t *xa = (const uint32_t *)&ra[i], *xb = (const uint32_t *)&rb[i];
#pragma unroll
                for (int kq = 0; kq < 4; kq++) {
                    const int o = ((grp * NSTEP + step) * 64 + r + 16 * kq) * VB;
                    *(uint32_t *)(sa + o) = xa[kq];
                    *(uint32_t *)(sb + o) = xb[kq];
                }
            }
        }
    };

    typename P::Acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (typename P::Acc){0, 0, 0, 0};

    const int64_t nk = g.Kp * (int64_t)sizeof(typename P::T) / LBYTES;
    gload(0);
    sstore(smem, smem + OPB);
    __syncthreads();
    for (int64_t kt = 0; kt < nk; kt++) {
        const int cur = (int)(kt & 1);
        const bool more = kt + 1 < nk;
        if (more) gload(kt + 1);
        const unsigned char *sa = smem + cur * 2 * OPB, *sb = sa + OPB;
#pragma unroll
        for (int s = 0; s < NSTEP; s++) {
            typename P::Frag af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; i++)
                af[i] = *(const typename P::Frag *)(sa + (((wm * 4 + i) * NSTEP + s) * 64 + lane) * VB);
#pragma unroll
            for (int j = 0; j < 4; j++)
                bf[j] = *(const typename P::Frag *)(sb + (((wn * 4 + j) * NSTEP + s) * 64 + lane) * VB);
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = P::mfma(af[i], bf[j], acc[i][j]);
            if (s == NSTEP / 2 - 1 && more) {
                unsigned char *na = smem + (cur ^ 1) * 2 * OPB;
                sstore(na, na + OPB);
            }
        }
        if (NSTEP == 1 && more) {
            unsigned char *na = smem + (cur ^ 1) * 2 * OPB;
            sstore(na, na + OPB);
        }
        __syncthreads();
    }
    // Epilogue: D[4*(l/16) + r][l%16] per 16x16 accumulator.
    const int fr = lane & 15, fq = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t col = n0 + wn * 64 + j * 16 + fr;
            if (col >= g.N) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * 64 + i * 16 + 4 * fq + r;
                if (row >= g.M) continue;
                double v = (double)acc[i][j][r];
                if (g.ea) v = ldexp(v, g.ea[row] + g.eb[col] - g.shift);
                int64_t addr;
                if constexpr (CVIEW) addr = vaddr(g.cv, row, col);
                else addr = row * g.ldc + col;
                g.C[addr] += g.alpha * v;
            }
        }
}

// ---------------------------------------------- fused slice products
// All slice products of one C tile in one kernel.  Per K step of 64 bytes
// the workgroup stages every slice plane of its A rows and B rows in LDS
// once and runs every product (i, j), i + j <= L, from registers.  Products
// of one level l = i + j share one scale (INT8 digits: 2^(ea + eb + 4 - 8l)),
// so they sum in one accumulator per level: L-1 accumulator sets instead
// of one FP64 read-modify-write of C per product.  INT8 levels sum exactly
// in INT32 and are flushed into C every Kc bytes of K (a multiple of 64)
// sized from the level's digit bound (|q_1| <= 65, |q_i| <= 128); BF16 and
// FP16 slices (64 bytes = 32 elements per step, the same fragment bytes and
// lane order) sum in FP32 and reach C once.
//
// Workgroup: C tile 128 x 128 (wave layout below).  LDS holds K steps as
// fragments of 1 KiB ([operand digit][16-row block][lane][16 B], the order
// of the digit planes themselves: lane = row % 16 + 16 * kq holds bytes
// 16kq..16kq+15 of the 64-byte step of its row); each fragment is one
// direct global->LDS load of 64 lanes x 16 B of one contiguous KiB.
// `ozaki_kernel` keeps two full stages (16 * NS fragments each, loaded a
// step ahead); `ozaki_ring_kernel` (NS <= 4) five half stages.
constexpr int OZ_BM = 128, OZ_BN = 128, OZ_BK = 64;

struct OzArgs {
    const unsigned char *A, *B;  // slice planes [NS][Mp][Kp bytes], [NS][Np][Kp bytes]
    const int32_t *ea, *eb;      // row exponents (INT8, F16) or null (BF16)
    double alpha;
    int64_t M, N, Kp, Mp, Np;    // Kp: bytes of K per slice row (a multiple of 64)
    int64_t kchunk_steps;        // K steps between INT32 flushes
    // K split: workgroups [0, t_full) own whole tiles; workgroup t_full + s
    // computes tile t_full + s / sk over K steps [(s % sk) * kslice, +
    // kslice) and leaves alpha * its sum in ws[s] (OZ_BM x OZ_BN);
    // oz_split_reduce adds the slices of tiles t_full.. to C in order.
    uint32_t t_full, sk;
    int64_t kslice;
    double *ws;
    double *C;
    ViewArgs cv;
    int64_t ldc;
    uint32_t tiles_m, tiles_n;
    // 1: C = alpha * A B^T (the first write of each element stores, no read
    // of C); 0: C += alpha * A B^T
    int overwrite;
};

// Workgroup of 8 waves (2 per SIMD) as 4 x 2, each owning 32 x 64 of the
// 128 x 128 tile as 2 x 4 MFMA blocks; the LDS stages take all 160 KiB (two
// stages at NS = 5, five half stages at NS = 4).  The tile is as large as
// the registers allow (L - 1 INT32 level sums per C element): every
// workgroup streams 16 * NS KiB of digit planes per K step, the loads'
// latency is the kernel's limit, and a 64 x 128 tile (1.5x the traffic per
// product) ran 15-25% slower, two 64 x 64 workgroups per CU slower still.
constexpr int OZ_WAVES = 8, OZ_WM = 32, OZ_WN = 64;
constexpr int OZ_FM = OZ_WM / 16, OZ_FN = OZ_WN / 16;  // MFMA blocks per wave

// Element traits of the fused kernels: the MFMA over one 16-byte fragment
// per lane (64 bytes of K of 16 rows) and its accumulator.  A level's sum is
// scaled back by 2^(ea + eb + 4 - 8 l) (INT8), 2^(ea + eb) (F16, rows scaled
// by a power of two) or not at all (BF16).
template <int P>
struct OzT;
template <>
struct OzT<DTCF_PREC_I8> {
    typedef i4 Acc;
    __device__ static inline Acc mfma(i4 a, i4 b, Acc c)
    {
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
    }
};
template <>
struct OzT<DTCF_PREC_BF16X3> {
    typedef f4 Acc;
    __device__ static inline Acc mfma(i4 a, i4 b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a),
                                                       __builtin_bit_cast(bf8, b), c, 0, 0, 0);
    }
};
template <>
struct OzT<DTCF_PREC_F16X3> {
    typedef f4 Acc;
    __device__ static inline Acc mfma(i4 a, i4 b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                      __builtin_bit_cast(h8, b), c, 0, 0, 0);
    }
};

template <int P, int NS, int L, bool CVIEW>
__device__ __forceinline__ void oz_flush(const OzArgs &g,
                                         typename OzT<P>::Acc (&acc)[L - 1][OZ_FM][OZ_FN],
                                         int64_t m0, int64_t n0, int wm, int wn, int lane,
                                         double *wsp, bool first)
{
    const int fr = lane & 15, fq = lane >> 4;
#pragma unroll
    for (int j = 0; j < OZ_FN; j++) {
        const int64_t col = n0 + wn * OZ_WN + j * 16 + fr;
        if (col >= g.N) continue;
        const int ebc = P == DTCF_PREC_BF16X3 ? 0 : g.eb[col];
        int64_t cof;
        if constexpr (CVIEW) cof = g.cv.cofs[col];
        else cof = col;
#pragma unroll
        for (int i = 0; i < OZ_FM; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * OZ_WM + i * 16 + 4 * fq + r;
                if (row >= g.M) continue;
                const int e = P == DTCF_PREC_BF16X3 ? 0 : g.ea[row] + ebc;
                double v = 0.0;
#pragma unroll
                for (int l = L - 2; l >= 0; l--) {  // smallest terms first
                    if constexpr (P == DTCF_PREC_I8)
                        v += ldexp((double)acc[l][i][j][r], e + 4 - 8 * (l + 2));
                    else
                        v += (double)acc[l][i][j][r];
                }
                if constexpr (P == DTCF_PREC_F16X3) v = ldexp(v, e);
                if (wsp) {  // K slice: this workgroup's own tile of the workspace
                    double *w = wsp + (row - m0) * OZ_BN + (col - n0);
                    *w = first ? g.alpha * v : *w + g.alpha * v;
                    continue;
                }
                int64_t addr;
                if constexpr (CVIEW) addr = g.cv.blk ? vaddr(g.cv, row, col) : g.cv.rofs[row] + cof;
                else addr = row * g.ldc + cof;
                if (g.overwrite && first) g.C[addr] = g.alpha * v;
                else g.C[addr] += g.alpha * v;
            }
    }
}

// One K step of a workgroup: every product (di, dj), di + dj <= L, of the
// A fragments in sa ([digit][16-row block][lane]) and the B fragments in sb,
// added to the level sums.  B is taken digit by digit (the next one read
// while this one is used); two updates of one accumulator -- (di, dj) and
// (di - 1, dj + 1) -- are about 8 (NS - 1) MFMAs apart, and the SIMD's
// second wave interleaves its own chains.
template <int P, int NS, int L>
__device__ __forceinline__ void oz_step(typename OzT<P>::Acc (&acc)[L - 1][OZ_FM][OZ_FN],
                                        const i4 *sa, const i4 *sb, int wm, int wn, int lane)
{
    i4 af[NS][OZ_FM];
#pragma unroll
    for (int d = 0; d < NS; d++)
#pragma unroll
        for (int i = 0; i < OZ_FM; i++)
            af[d][i] = sa[(d * (OZ_BM / 16) + wm * (OZ_WM / 16) + i) * 64 + lane];
    i4 bcur[OZ_FN], bnxt[OZ_FN];
#pragma unroll
    for (int j = 0; j < OZ_FN; j++) bcur[j] = sb[(wn * (OZ_WN / 16) + j) * 64 + lane];
    // Issue priority for the first B digit's MFMAs: the SIMD's two waves
    // drift apart, so one's fragment reads and barrier wait overlap the
    // other's MFMAs (1.8% on the int8x4/L6 ladder; priority while issuing
    // the loads instead gained nothing).
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int dj = 1; dj <= NS; dj++) {
        if (dj == 2) __builtin_amdgcn_s_setprio(0);
        if (dj < NS) {
#pragma unroll
            for (int j = 0; j < OZ_FN; j++)
                bnxt[j] = sb[(dj * (OZ_BN / 16) + wn * (OZ_WN / 16) + j) * 64 + lane];
        }
#pragma unroll
        for (int di = 1; di <= NS; di++) {
            if (di + dj > L) continue;
#pragma unroll
            for (int i = 0; i < OZ_FM; i++)
#pragma unroll
                for (int j = 0; j < OZ_FN; j++)
                    acc[di + dj - 2][i][j] =
                        OzT<P>::mfma(af[di - 1][i], bcur[j], acc[di + dj - 2][i][j]);
        }
#pragma unroll
        for (int j = 0; j < OZ_FN; j++) bcur[j] = bnxt[j];
    }
}

// Tile, K steps [k0, k0 + nk) and workspace tile (null: C itself) of
// workgroup blockIdx.x.  Whole tiles take an XCD-aware remap (consecutive
// tiles on one XCD and its L2); every tile index then goes through groups of
// 8 tile rows.
__device__ __forceinline__ void oz_tile(const OzArgs &g, int64_t &m0, int64_t &n0, int64_t &k0,
                                        int64_t &nk, double *&wsp)
{
    const int64_t nkall = g.Kp / OZ_BK;
    uint32_t q;
    if (blockIdx.x < g.t_full) {
        q = blockIdx.x;
        if ((g.t_full & 7u) == 0) q = (q & 7u) * (g.t_full >> 3) + (q >> 3);
        k0 = 0;
        nk = nkall;
        wsp = nullptr;
    } else {
        const uint32_t sidx = blockIdx.x - g.t_full;
        q = g.t_full + sidx / g.sk;
        k0 = (int64_t)(sidx % g.sk) * g.kslice;
        nk = min(nkall - k0, g.kslice);
        wsp = g.ws + (int64_t)sidx * OZ_BM * OZ_BN;
    }
    const uint32_t width = 8 * g.tiles_n;
    const uint32_t first = (q / width) * 8;
    const uint32_t gsz = min(g.tiles_m - first, 8u);
    const uint32_t tm = first + (q % width) % gsz, tn = (q % width) / gsz;
    m0 = (int64_t)tm * OZ_BM;
    n0 = (int64_t)tn * OZ_BN;
}

template <int NS, int L, bool CVIEW>
__global__ __launch_bounds__(64 * OZ_WAVES, 1) void ozaki_kernel(OzArgs g)
{
    constexpr int FA = NS * (OZ_BM / 16), FB = NS * (OZ_BN / 16);
    constexpr int NF = FA + FB;                          // fragments per stage
    constexpr int PER = (NF + OZ_WAVES - 1) / OZ_WAVES;  // per wave (last may be short)
    // Two stages as two objects: the direct loads into one stage provably do
    // not alias the fragment reads of the other, so no wait is placed
    // between them.
    __shared__ i4 stage0[NF * 64];
    __shared__ i4 stage1[NF * 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wv / (OZ_BN / OZ_WN), wn = wv % (OZ_BN / OZ_WN);

    int64_t m0, n0, k0, nk;
    double *wsp;
    oz_tile(g, m0, n0, k0, nk, wsp);

    const size_t pa = (size_t)g.Mp * g.Kp, pb = (size_t)g.Np * g.Kp;
    // Fragment f = wave * PER + c comes from a wave-uniform base (scalar
    // registers) plus one per-lane offset; it is loaded straight into LDS
    // (16 bytes per lane, lane l at byte 16 l of the fragment).
    // (fragment-major planes: fragment (row block, K step) is one KiB)
    const int64_t nk64 = g.Kp / OZ_BK, lane_off = 16 * lane;
    auto frag_base = [&](int f) __attribute__((always_inline)) -> const unsigned char * {
        if (f < FA) {
            const int d = f / (OZ_BM / 16), b = f % (OZ_BM / 16);
            return g.A + d * pa + ((m0 >> 4) + b) * nk64 * 1024;
        }
        const int h = f - FA, d = h / (OZ_BN / 16), b = h % (OZ_BN / 16);
        return g.B + d * pb + ((n0 >> 4) + b) * nk64 * 1024;
    };
    auto issue = [&](i4 *stage, int64_t kt) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < PER; c++) {
            const int f = wv * PER + c;
            if (NF % OZ_WAVES != 0 && f >= NF) break;  // wave-uniform
            __builtin_amdgcn_global_load_lds(
                (const void *)(frag_base(f) + lane_off + kt * 1024),
                (__attribute__((address_space(3))) void *)(stage + f * 64), 16, 0, 0);
        }
    };

    i4 acc[L - 1][OZ_FM][OZ_FN];
#pragma unroll
    for (int l = 0; l < L - 1; l++)
#pragma unroll
        for (int i = 0; i < OZ_FM; i++)
#pragma unroll
            for (int j = 0; j < OZ_FN; j++) acc[l][i][j] = (i4){0, 0, 0, 0};

    auto step = [&](const i4 *st) __attribute__((always_inline)) {
        oz_step<DTCF_PREC_I8, NS, L>(acc, st, st + FA * 64, wm, wn, lane);
    };
    bool first = true;
    int64_t since = 0;
    auto after = [&](bool more) __attribute__((always_inline)) {
        __builtin_amdgcn_s_waitcnt(0);  // this wave's direct loads have landed
        __syncthreads();                // ... and everyone's; stage reads done
        if (++since == g.kchunk_steps && more) {
            oz_flush<DTCF_PREC_I8, NS, L, CVIEW>(g, acc, m0, n0, wm, wn, lane, wsp, first);
            first = false;
#pragma unroll
            for (int l = 0; l < L - 1; l++)
#pragma unroll
                for (int i = 0; i < OZ_FM; i++)
#pragma unroll
                    for (int j = 0; j < OZ_FN; j++) acc[l][i][j] = (i4){0, 0, 0, 0};
            since = 0;
        }
    };

    issue(stage0, k0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int64_t kt = 0; kt < nk; kt += 2) {
        if (kt + 1 < nk) issue(stage1, k0 + kt + 1);
        step(stage0);
        after(kt + 1 < nk);
        if (kt + 1 >= nk) break;
        if (kt + 2 < nk) issue(stage0, k0 + kt + 2);
        step(stage1);
        after(kt + 2 < nk);
    }
    oz_flush<DTCF_PREC_I8, NS, L, CVIEW>(g, acc, m0, n0, wm, wn, lane, wsp, first);
}

// The same computation with a deeper load pipeline, for NS <= 4: LDS holds
// five half stages of NS * 8 fragments (32 KiB at NS = 4), one A or one B
// tile of a K step each, as five objects h0..h4.  Half n of the stream
// (A_k = half 2k, B_k = half 2k+1) sits in slot n mod 5; step k computes from
// slots 2k, 2k+1 and issues B_{k+1} and A_{k+2} into the two slots step k-1
// read, so A is loaded two steps ahead and B one (the two-stage kernel gives
// both one).  The end-of-step wait leaves this wave's A_{k+2} loads in
// flight (vmcnt = NS): direct loads complete in issue order.  The loop is
// unrolled by five so every slot is a fixed object and the compiler sees
// that loads into one slot never alias reads of another.
template <int P, int NS, int L, bool CVIEW>
__global__ __launch_bounds__(64 * OZ_WAVES, 1) void ozaki_ring_kernel(OzArgs g)
{
    static_assert(OZ_BM == OZ_BN, "A and B halves have the same size");
    constexpr int FH = NS * (OZ_BM / 16);  // fragments per half stage
    static_assert(FH % OZ_WAVES == 0 && NS <= 4, "five half stages fit LDS for NS <= 4");
    constexpr int PH = FH / OZ_WAVES;      // direct loads per wave per half (= NS)
    __shared__ i4 h0[FH * 64];
    __shared__ i4 h1[FH * 64];
    __shared__ i4 h2[FH * 64];
    __shared__ i4 h3[FH * 64];
    __shared__ i4 h4[FH * 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wv / (OZ_BN / OZ_WN), wn = wv % (OZ_BN / OZ_WN);
    int64_t m0, n0, k0, nk;
    double *wsp;
    oz_tile(g, m0, n0, k0, nk, wsp);

    const size_t pa = (size_t)g.Mp * g.Kp, pb = (size_t)g.Np * g.Kp;
    const int64_t nk64 = g.Kp / OZ_BK, lane_off = 16 * lane;
    // fragment f = wave * PH + c of a half: digit f / 8, 16-row block f % 8
    // (fragment-major planes: fragment (row block, K step) is one KiB)
    auto issue = [&](i4 *slot, const unsigned char *planes, size_t pstride, int64_t row0,
                     int64_t kt) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < PH; c++) {
            const int f = wv * PH + c, d = f / (OZ_BM / 16), b = f % (OZ_BM / 16);
            __builtin_amdgcn_global_load_lds(
                (const void *)(planes + d * pstride + (((row0 >> 4) + b) * nk64 + kt) * 1024 + lane_off),
                (__attribute__((address_space(3))) void *)(slot + f * 64), 16, 0, 0);
        }
    };

    typedef typename OzT<P>::Acc Acc;
    Acc acc[L - 1][OZ_FM][OZ_FN];
#pragma unroll
    for (int l = 0; l < L - 1; l++)
#pragma unroll
        for (int i = 0; i < OZ_FM; i++)
#pragma unroll
            for (int j = 0; j < OZ_FN; j++) acc[l][i][j] = (Acc){0, 0, 0, 0};

    // s_waitcnt with vmcnt = PH and no expcnt / lgkmcnt wait (GFX9 encoding:
    // vmcnt[3:0], expcnt[6:4], lgkmcnt[11:8], vmcnt[15:14])
    constexpr int WAIT_PH = (PH & 15) | (7 << 4) | (15 << 8) | ((PH >> 4) << 14);
    bool first = true;
    int64_t since = 0;
    // Workgroup barrier without __syncthreads' release fence, which would
    // wait for every outstanding load (vmcnt 0) and so for A_{k+2} too.  The
    // waits are explicit instead: before the barrier each wave has its own
    // loads for the next step landed (s_waitcnt) and has consumed its reads
    // of this step's slots (the MFMAs used them); the "memory" clobber keeps
    // the compiler from moving LDS or load instructions across it.
    auto barrier = []() __attribute__((always_inline)) { asm volatile("s_barrier" ::: "memory"); };
    auto after = [&](int64_t k) __attribute__((always_inline)) {
        if (k + 2 < nk) __builtin_amdgcn_s_waitcnt(WAIT_PH);  // A_{k+2} may stay in flight
        else __builtin_amdgcn_s_waitcnt(0);
        barrier();
        if (++since == g.kchunk_steps && k + 1 < nk) {
            oz_flush<P, NS, L, CVIEW>(g, acc, m0, n0, wm, wn, lane, wsp, first);
            first = false;
#pragma unroll
            for (int l = 0; l < L - 1; l++)
#pragma unroll
                for (int i = 0; i < OZ_FM; i++)
#pragma unroll
                    for (int j = 0; j < OZ_FN; j++) acc[l][i][j] = (Acc){0, 0, 0, 0};
            since = 0;
        }
    };

    issue(h0, g.A, pa, m0, k0);
    issue(h1, g.B, pb, n0, k0);
    if (nk > 1) {
        issue(h2, g.A, pa, m0, k0 + 1);
        __builtin_amdgcn_s_waitcnt(WAIT_PH);
    } else {
        __builtin_amdgcn_s_waitcnt(0);
    }
    barrier();
#define OZ_RING_STEP(SA, SB, NB, NA, J)                   \
    {                                                     \
        const int64_t k = kt + (J);                       \
        if (k >= nk) break;                               \
        if (k + 1 < nk) issue(NB, g.B, pb, n0, k0 + k + 1); \
        if (k + 2 < nk) issue(NA, g.A, pa, m0, k0 + k + 2); \
        oz_step<P, NS, L>(acc, SA, SB, wm, wn, lane);     \
        after(k);                                         \
    }
    for (int64_t kt = 0; kt < nk; kt += 5) {
        OZ_RING_STEP(h0, h1, h3, h4, 0)
        OZ_RING_STEP(h2, h3, h0, h1, 1)
        OZ_RING_STEP(h4, h0, h2, h3, 2)
        OZ_RING_STEP(h1, h2, h4, h0, 3)
        OZ_RING_STEP(h3, h4, h1, h2, 4)
    }
#undef OZ_RING_STEP
    oz_flush<P, NS, L, CVIEW>(g, acc, m0, n0, wm, wn, lane, wsp, first);
}

// Adds the K slices of every tile (workspace slot tile * sk + slice, alpha
// already applied) to C in slice order: deterministic.
template <bool CVIEW>
__global__ __launch_bounds__(256) void oz_split_reduce(OzArgs g)
{
    const uint32_t ti = blockIdx.y, q = g.t_full + ti;  // split tile, in oz_tile's order
    const uint32_t width = 8 * g.tiles_n;
    const uint32_t first = (q / width) * 8;
    const uint32_t gsz = min(g.tiles_m - first, 8u);
    const uint32_t tm = first + (q % width) % gsz, tn = (q % width) / gsz;
    const int64_t m0 = (int64_t)tm * OZ_BM, n0 = (int64_t)tn * OZ_BN;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < OZ_BM * OZ_BN; e += gridDim.x * 256) {
        const int64_t row = m0 + e / OZ_BN, col = n0 + e % OZ_BN;
        if (row >= g.M || col >= g.N) continue;
        const double sum = sum_in_order(g.ws + (int64_t)ti * g.sk * OZ_BM * OZ_BN + e, g.sk, OZ_BM * OZ_BN);
        int64_t addr;
        if constexpr (CVIEW)
            addr = g.cv.blk ? vaddr(g.cv, row, col) : g.cv.rofs[row] + g.cv.cofs[col];
        else addr = row * g.ldc + col;
        if (g.overwrite) g.C[addr] = sum;
        else g.C[addr] += sum;
    }
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// 1: digit products of INT8 splits run in the fused kernel where a
// specialisation exists; 0: one launch per product.
int g_fused_i8 = 1;

template <int P, int NS, int L>
int oz_launch(const OzArgs &g, bool cview, hipStream_t s)
{
    const uint32_t T = g.tiles_m * g.tiles_n;
    const unsigned grid = g.t_full + (T - g.t_full) * g.sk;
    if constexpr (NS <= 4) {
        if (cview) hipLaunchKernelGGL((ozaki_ring_kernel<P, NS, L, true>), dim3(grid), dim3(64 * OZ_WAVES), 0, s, g);
        else hipLaunchKernelGGL((ozaki_ring_kernel<P, NS, L, false>), dim3(grid), dim3(64 * OZ_WAVES), 0, s, g);
    } else {
        static_assert(NS > 4 && P == DTCF_PREC_I8, "two full stages: INT8 with five digits only");
        if (cview) hipLaunchKernelGGL((ozaki_kernel<NS, L, true>), dim3(grid), dim3(64 * OZ_WAVES), 0, s, g);
        else hipLaunchKernelGGL((ozaki_kernel<NS, L, false>), dim3(grid), dim3(64 * OZ_WAVES), 0, s, g);
    }
    DTCF_HIP(hipGetLastError());
    if (g.t_full < T) {
        const dim3 rg(OZ_BM * OZ_BN / 256, T - g.t_full);
        if (cview) hipLaunchKernelGGL(oz_split_reduce<true>, rg, dim3(256), 0, s, g);
        else hipLaunchKernelGGL(oz_split_reduce<false>, rg, dim3(256), 0, s, g);
        DTCF_HIP(hipGetLastError());
    }
    return DTCF_OK;
}

// Runs the fused kernel when (prec, ns, lev) has a specialisation; *done =
// false otherwise.
int oz_dispatch(int prec, int ns, int lev, const OzArgs &g, bool cview, hipStream_t s, bool *done)
{
    *done = true;
    if (prec == DTCF_PREC_BF16X3 && ns == 3 && lev == 4)
        return oz_launch<DTCF_PREC_BF16X3, 3, 4>(g, cview, s);
    if (prec == DTCF_PREC_F16X3 && ns == 3 && lev == 4)
        return oz_launch<DTCF_PREC_F16X3, 3, 4>(g, cview, s);
    if (prec != DTCF_PREC_I8) {
        *done = false;
        return DTCF_OK;
    }
    switch (ns * 16 + lev) {
    case 3 * 16 + 4: return oz_launch<DTCF_PREC_I8, 3, 4>(g, cview, s);
    case 4 * 16 + 4: return oz_launch<DTCF_PREC_I8, 4, 4>(g, cview, s);
    case 4 * 16 + 5: return oz_launch<DTCF_PREC_I8, 4, 5>(g, cview, s);
    case 4 * 16 + 6: return oz_launch<DTCF_PREC_I8, 4, 6>(g, cview, s);
    case 5 * 16 + 5: return oz_launch<DTCF_PREC_I8, 5, 5>(g, cview, s);
    case 5 * 16 + 6: return oz_launch<DTCF_PREC_I8, 5, 6>(g, cview, s);
    default: *done = false; return DTCF_OK;
    }
}

}  // namespace

// Whether (prec, ns, lev) has a fused slice-product kernel (the products
// then support overwrite mode).
bool gemm_split_fused(int prec, int ns, int lev)
{
    if (!g_fused_i8) return false;
    if (prec == DTCF_PREC_BF16X3 || prec == DTCF_PREC_F16X3) return ns == 3 && lev == 4;
    if (prec != DTCF_PREC_I8) return false;
    switch (ns * 16 + lev) {
    case 3 * 16 + 4: case 4 * 16 + 4: case 4 * 16 + 5: case 4 * 16 + 6: case 5 * 16 + 5:
    case 5 * 16 + 6: return true;
    default: return false;
    }
}
}  // namespace dtcf

using namespace dtcf;

static int view_args(dtcf_view v, ViewArgs &a, int64_t &rows, int64_t &cols)
{
    DTCF_REQUIRE(v, "null view");
    a.rofs = v->rofs;
    a.cofs = v->cofs;
    a.blk = v->blk;
    a.div_rpb = FastDiv((uint32_t)v->rpb);
    a.div_cpb = FastDiv((uint32_t)v->cpb);
    a.ncb = v->ncb;
    rows = v->rows;
    cols = v->cols;
    return DTCF_OK;
}

extern "C" int dtcf_fill_normal_f64(double *out, int64_t n, uint64_t seed, double sigma,
                                    int64_t start, void *stream)
{
    DTCF_REQUIRE(n >= 0 && (n == 0 || out), "fill: bad buffer");
    if (n == 0) return DTCF_OK;
    hipLaunchKernelGGL(fill_normal_kernel, dim3(2048), dim3(256), 0, as_stream(stream), out, n,
                       seed, sigma, start);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

extern "C" int dtcf_fill_uniform_f64(double *out, int64_t n, uint64_t seed, double lo, double hi,
                                     void *stream)
{
    DTCF_REQUIRE(n >= 0 && (n == 0 || out), "fill: bad buffer");
    if (n == 0) return DTCF_OK;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(256), dim3(256), 0, as_stream(stream), out, n,
                       seed, lo, hi);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

extern "C" int dtcf_antisymmetrize4_f64(const double *in, double *out, int64_t d0, int64_t d1,
                                        int64_t d2, int64_t d3, void *stream)
{
    DTCF_REQUIRE(in && out && in != out, "antisym: needs distinct buffers");
    DTCF_REQUIRE(d0 == d1 && d2 == d3 && d0 > 0 && d2 > 0, "antisym: pairs must have equal extents");
    hipLaunchKernelGGL(antisym4_kernel, dim3(4096), dim3(256), 0, as_stream(stream), in, out, d0,
                       d1, d2, d3);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

extern "C" int64_t dtcf_split_bytes(int prec, int nslice, int64_t rows, int64_t K)
{
    const int64_t rp = round_up(std::max<int64_t>(rows, 1), 128), kp = round_up(std::max<int64_t>(K, 1), 128);
    switch (prec) {
    case DTCF_PREC_F32: return rp * kp * 4;
    case DTCF_PREC_BF16X3:
    case DTCF_PREC_F16X3: return 3 * rp * kp * 2;
    case DTCF_PREC_I8: return (int64_t)nslice * rp * kp;
    default: return -1;
    }
}

extern "C" int dtcf_split_f64(int prec, int nslice, const double *src, dtcf_view v, void *dst,
                              int32_t *rexp, void *stream)
{
    DTCF_REQUIRE(src && v && dst, "split: null argument");
    DTCF_REQUIRE(prec >= DTCF_PREC_F32 && prec <= DTCF_PREC_I8, "split: unknown precision");
    if (prec == DTCF_PREC_I8) DTCF_REQUIRE(nslice >= 1 && nslice <= 6, "split: 1..6 INT8 slices");
    if (prec == DTCF_PREC_F16X3 || prec == DTCF_PREC_I8) DTCF_REQUIRE(rexp, "split: needs rexp");
    ViewArgs a{};
    int64_t rows, K;
    int st = view_args(v, a, rows, K);
    if (st) return st;
    const int64_t rp = round_up(rows, 128), kp = round_up(K, 128);
    const unsigned grid = (unsigned)((rp + 3) / 4);
    hipStream_t s = as_stream(stream);
    // BF16 / FP16 slices of table views: K-parallel (FP16: row exponents first)
    if ((prec == DTCF_PREC_BF16X3 || prec == DTCF_PREC_F16X3) && !v->blk) {
        const bool pairs = v->colpairs && (reinterpret_cast<uintptr_t>(src) & 15) == 0;
        const int64_t nrb = rp / 16, nd = nrb * ((kp / 32 + 3) / 4);
        DTCF_REQUIRE(nd < (int64_t(1) << 31), "split: operand too large");
        unsigned char *d1 = (unsigned char *)dst;
        if (prec == DTCF_PREC_F16X3) {
            const int64_t ne = nrb * ((K + 1023) / 1024);
            DTCF_REQUIRE(ne < (int64_t(1) << 31), "split: operand too large");
            const unsigned gi = (unsigned)std::min<int64_t>((rp + 255) / 256, 4096);
            hipLaunchKernelGGL(i8_exp_init, dim3(gi), dim3(256), 0, s, rexp, rp);
            if (pairs)
                hipLaunchKernelGGL(i8_exp_kernel<true>, dim3((unsigned)ne), dim3(256), 0, s, src, a, rows, K,
                                   nrb, rexp);
            else
                hipLaunchKernelGGL(i8_exp_kernel<false>, dim3((unsigned)ne), dim3(256), 0, s, src, a, rows,
                                   K, nrb, rexp);
            hipLaunchKernelGGL(f16_rexp_finish, dim3(gi), dim3(256), 0, s, rexp, rp);
            if (pairs)
                hipLaunchKernelGGL((f16_slices_kernel<DTCF_PREC_F16X3, true>), dim3((unsigned)nd), dim3(256), 0,
                                   s, src, a, rows, K, rp, kp, nrb, d1, rexp);
            else
                hipLaunchKernelGGL((f16_slices_kernel<DTCF_PREC_F16X3, false>), dim3((unsigned)nd), dim3(256),
                                   0, s, src, a, rows, K, rp, kp, nrb, d1, rexp);
        } else {
            if (pairs)
                hipLaunchKernelGGL((f16_slices_kernel<DTCF_PREC_BF16X3, true>), dim3((unsigned)nd), dim3(256),
                                   0, s, src, a, rows, K, rp, kp, nrb, d1, rexp);
            else
                hipLaunchKernelGGL((f16_slices_kernel<DTCF_PREC_BF16X3, false>), dim3((unsigned)nd),
                                   dim3(256), 0, s, src, a, rows, K, rp, kp, nrb, d1, rexp);
        }
        DTCF_HIP(hipGetLastError());
        return DTCF_OK;
    }
    // INT8 table views (no block table) split in two K-parallel passes at
    // any shape: exponents, then digits.
    if (prec == DTCF_PREC_I8 && !v->blk) {
        const bool pairs = v->colpairs && (reinterpret_cast<uintptr_t>(src) & 15) == 0;
        // one-dimensional grids, row blocks fastest
        const int64_t nrb = rp / 16, ne = nrb * ((K + 1023) / 1024), nd = nrb * ((kp / 64 + 3) / 4);
        DTCF_REQUIRE(ne < (int64_t(1) << 31) && nd < (int64_t(1) << 31), "split: operand too large");
        const dim3 ge((unsigned)ne), gd((unsigned)nd);
        int8_t *d8 = (int8_t *)dst;
        hipLaunchKernelGGL(i8_exp_init, dim3((unsigned)std::min<int64_t>((rp + 255) / 256, 4096)),
                           dim3(256), 0, s, rexp, rp);
        if (pairs)
            hipLaunchKernelGGL(i8_exp_kernel<true>, ge, dim3(256), 0, s, src, a, rows, K, nrb, rexp);
        else
            hipLaunchKernelGGL(i8_exp_kernel<false>, ge, dim3(256), 0, s, src, a, rows, K, nrb, rexp);
#define DTCF_DIGITS(NS)                                                                                \
    if (pairs) hipLaunchKernelGGL((i8_digit_kernel<true, NS>), gd, dim3(256), 0, s, src, a, rows, K, rp, \
                                  kp, nrb, d8, rexp);                                                  \
    else hipLaunchKernelGGL((i8_digit_kernel<false, NS>), gd, dim3(256), 0, s, src, a, rows, K, rp, kp, \
                            nrb, d8, rexp);
        switch (nslice) {
        case 1: DTCF_DIGITS(1) break;
        case 2: DTCF_DIGITS(2) break;
        case 3: DTCF_DIGITS(3) break;
        case 4: DTCF_DIGITS(4) break;
        case 5: DTCF_DIGITS(5) break;
        default: DTCF_DIGITS(6) break;
        }
#undef DTCF_DIGITS
        DTCF_HIP(hipGetLastError());
        return DTCF_OK;
    }
    const int64_t row_waves = rp;  // the row-parallel kernel runs a wave per row
    if (row_waves < 2048 && kp >= 8 * rp) {  // wide: K-parallel launches
        const int ns = prec == DTCF_PREC_I8 ? nslice : prec == DTCF_PREC_F32 ? 1 : 3;
        const int64_t esz = prec == DTCF_PREC_F32 ? 4 : prec == DTCF_PREC_I8 ? 1 : 2;
        const int64_t chunks = rp * kp * esz / 16;
        const unsigned g2 = (unsigned)std::min<int64_t>((chunks + 255) / 256, 1 << 16);
        unsigned char *d = (unsigned char *)dst;
        const bool run = v->rowrun16 && !v->blk;
#define DTCF_WIDE(M)                                                                               \
    if (run) hipLaunchKernelGGL((split_wide_kernel<M, true>), dim3(g2), dim3(256), 0, s, src, a,    \
                                rows, K, rp, kp, ns, d, rexp);                                      \
    else hipLaunchKernelGGL((split_wide_kernel<M, false>), dim3(g2), dim3(256), 0, s, src, a, rows, \
                            K, rp, kp, ns, d, rexp);
        switch (prec) {
        case DTCF_PREC_F32: DTCF_WIDE(DTCF_PREC_F32) break;
        case DTCF_PREC_BF16X3: DTCF_WIDE(DTCF_PREC_BF16X3) break;
        case DTCF_PREC_F16X3:
            hipLaunchKernelGGL(row_exp_kernel<DTCF_PREC_F16X3>, dim3((unsigned)rp), dim3(512), 0, s,
                               src, a, rows, K, rexp);
            DTCF_WIDE(DTCF_PREC_F16X3)
            break;
        default:
            hipLaunchKernelGGL(row_exp_kernel<DTCF_PREC_I8>, dim3((unsigned)rp), dim3(512), 0, s, src,
                               a, rows, K, rexp);
            DTCF_WIDE(DTCF_PREC_I8)
            break;
        }
#undef DTCF_WIDE
        DTCF_HIP(hipGetLastError());
        return DTCF_OK;
    }
    switch (prec) {
    case DTCF_PREC_F32:
        hipLaunchKernelGGL(split_kernel<DTCF_PREC_F32>, dim3(grid), dim3(256), 0, s, src, a, rows, K,
                           rp, kp, 1, dst, rexp);
        break;
    case DTCF_PREC_BF16X3:
        hipLaunchKernelGGL(split_kernel<DTCF_PREC_BF16X3>, dim3(grid), dim3(256), 0, s, src, a, rows,
                           K, rp, kp, 3, dst, rexp);
        break;
    case DTCF_PREC_F16X3:
        hipLaunchKernelGGL(split_kernel<DTCF_PREC_F16X3>, dim3(grid), dim3(256), 0, s, src, a, rows,
                           K, rp, kp, 3, dst, rexp);
        break;
    default:
        hipLaunchKernelGGL(split_kernel<DTCF_PREC_I8>, dim3(grid), dim3(256), 0, s, src, a, rows, K,
                           rp, kp, nslice, dst, rexp);
        break;
    }
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

template <int PREC, bool CV>
static void lp_launch(const LPArgs &g, hipStream_t s)
{
    const unsigned grid = (unsigned)(((g.M + LBM - 1) / LBM) * g.tiles_n);
    hipLaunchKernelGGL((lp_gemm_kernel<PREC, CV>), dim3(grid), dim3(256), 0, s, g);
}

extern "C" int dtcf_gemm_split_set_fused(int on)
{
    g_fused_i8 = on ? 1 : 0;
    return DTCF_OK;
}

namespace dtcf {
// Slice products C (+)= alpha * A B^T; overwrite requires a fused kernel
// (gemm_split_fused) and is refused otherwise.
int gemm_split_f64_impl(int prec, int nslice, int64_t m, int64_t n, int64_t k, double alpha,
                        const void *A_slices, const int32_t *ea, const void *B_slices,
                        const int32_t *eb, int max_level, double *C, dtcf_view cv, int64_t ldc,
                        void *stream, int overwrite)
{
    DTCF_REQUIRE(m >= 0 && n >= 0 && k >= 0, "gemm_split: negative size");
    if (m == 0 || n == 0 || k == 0 || alpha == 0.0) return DTCF_OK;
    DTCF_REQUIRE(A_slices && B_slices && C, "gemm_split: null operand");
    DTCF_REQUIRE(prec >= DTCF_PREC_F32 && prec <= DTCF_PREC_I8, "gemm_split: unknown precision");
    const bool scaled = prec == DTCF_PREC_F16X3 || prec == DTCF_PREC_I8;
    if (scaled) DTCF_REQUIRE(ea && eb, "gemm_split: scaled slices need row exponents");
    const int ns = prec == DTCF_PREC_F32 ? 1 : prec == DTCF_PREC_I8 ? nslice : 3;
    DTCF_REQUIRE(ns >= 1 && ns <= 6, "gemm_split: bad slice count");

    LPArgs g{};
    g.M = m;
    g.N = n;
    g.Mp = round_up(m, 128);
    g.Np = round_up(n, 128);
    g.Kp = round_up(k, 128);
    g.alpha = alpha;
    g.C = C;
    g.ldc = ldc;
    g.tiles_n = (uint32_t)(g.Np / LBN);
    if (cv) {
        int64_t r, c;
        int st = view_args(cv, g.cv, r, c);
        if (st) return st;
        DTCF_REQUIRE(r == m && c == n, "gemm_split: C view is not m x n");
    } else {
        DTCF_REQUIRE(ldc >= n, "gemm_split: ldc < n");
    }
    const size_t esz = prec == DTCF_PREC_F32 ? 4 : prec == DTCF_PREC_I8 ? 1 : 2;
    const size_t pa = (size_t)g.Mp * g.Kp * esz, pb = (size_t)g.Np * g.Kp * esz;
    // Slice pairs (i, j), 1-based, with i + j <= max_level (default ns + 1).
    const int lev = max_level > 0 ? max_level : ns + 1;
    if (lev < 2) return DTCF_OK;  // no pair has i + j <= 1: C is unchanged
    DTCF_REQUIRE(!overwrite || gemm_split_fused(prec, ns, lev),
                 "gemm_split: overwrite needs a fused slice-product kernel");
    hipStream_t s = as_stream(stream);
    if (prec != DTCF_PREC_F32 && g_fused_i8) {
        OzArgs o{};
        o.A = (const unsigned char *)A_slices;
        o.B = (const unsigned char *)B_slices;
        o.ea = scaled ? ea : nullptr;
        o.eb = scaled ? eb : nullptr;
        o.alpha = alpha;
        o.M = m;
        o.N = n;
        o.Kp = g.Kp * (int64_t)esz;  // bytes per slice row
        o.Mp = g.Mp;
        o.Np = g.Np;
        o.C = C;
        o.cv = g.cv;
        o.ldc = ldc;
        o.tiles_m = (uint32_t)(g.Mp / OZ_BM);
        o.tiles_n = (uint32_t)(g.Np / OZ_BN);
        o.overwrite = overwrite;
        // Level l = i + j sums the products q_i q_j, |q_1| <= 65, |q_i| <= 128
        // otherwise; INT32 stays exact while K times the largest level's
        // bound is below 2^31 (four digits to level 6: 49,408 per K, a flush
        // every 43,456 K; five: 65,792, every 32,640 K).
        // (BF16 / F16 levels sum in FP32 and are added to C at the end only.)
        int64_t worst = 0;
        for (int l = 2; l <= lev; l++) {
            int64_t s = 0;
            for (int i = std::max(1, l - ns); i <= std::min(ns, l - 1); i++)
                s += (i == 1 ? 65 : 128) * (l - i == 1 ? 65 : 128);
            worst = std::max(worst, s);
        }
        const int64_t kmax = ((int64_t(1) << 31) - 1) / worst;
        o.kchunk_steps = prec == DTCF_PREC_I8 ? std::max<int64_t>(1, kmax / OZ_BK) : INT64_MAX;
        // K split (one workgroup per CU): when the tiles fill under half the
        // CUs, every tile is cut into sk slices of >= 16 K steps with sk *
        // tiles <= CUs; when the last round of whole tiles fills at most half
        // the CUs, its tiles alone are cut so that they fill one round (the
        // 324 tiles of a 2304^2 result: 68 tail tiles x 3 slices after 256
        // whole ones, instead of a second round 27% full).  nsplit * sk <=
        // CUs, so the slices (alpha * partial C tiles) fit the stream's GEMM
        // split workspace.
        const uint32_t tiles = o.tiles_m * o.tiles_n;
        const int64_t nkall = o.Kp / OZ_BK;
        o.t_full = tiles;
        o.sk = 1;
        o.kslice = nkall;
        o.ws = nullptr;
        int cus = 0;
        int sst = gemm_device_cus(&cus);
        if (sst != DTCF_OK) return sst;
        const uint32_t rem = tiles % (uint32_t)cus;
        uint32_t nsplit = 0;
        int64_t sk = 0;
        if (tiles < (uint32_t)cus / 2 && nkall >= 32) {
            nsplit = tiles;
            sk = std::min<int64_t>(cus / tiles, nkall / 16);
        } else if (tiles > (uint32_t)cus && rem > 0 && rem <= (uint32_t)cus / 2 && nkall >= 32) {
            nsplit = rem;
            sk = std::min<int64_t>(cus / rem, nkall / 16);
        }
        if (sk >= 2) {
            const int64_t ks = (nkall + sk - 1) / sk;
            sk = (nkall + ks - 1) / ks;
            double *ws = nullptr;
            sst = gemm_workspace(s, (size_t)nsplit * sk * OZ_BM * OZ_BN * sizeof(double), &ws);
            if (sst != DTCF_OK) return sst;
            o.t_full = tiles - nsplit;
            o.sk = (uint32_t)sk;
            o.kslice = ks;
            o.ws = ws;
        }
        bool done = false;
        int st = oz_dispatch(prec, ns, lev, o, cv != nullptr, s, &done);
        if (st != DTCF_OK || done) return st;
    }
    if (prec == DTCF_PREC_I8)
        DTCF_REQUIRE(k < 131072, "gemm_split: INT8 products exact only for K < 131072");
    for (int i = 1; i <= ns; i++)
        for (int j = 1; j <= ns && i + j <= lev; j++) {
            g.A = (const unsigned char *)A_slices + (i - 1) * pa;
            g.B = (const unsigned char *)B_slices + (j - 1) * pb;
            g.ea = scaled ? ea : nullptr;
            g.eb = scaled ? eb : nullptr;
            g.shift = prec == DTCF_PREC_I8 ? 8 * (i + j) - 4 : 0;
            switch (prec) {
            case DTCF_PREC_F32:
                cv ? lp_launch<DTCF_PREC_F32, true>(g, s) : lp_launch<DTCF_PREC_F32, false>(g, s);
                break;
            case DTCF_PREC_BF16X3:
                cv ? lp_launch<DTCF_PREC_BF16X3, true>(g, s) : lp_launch<DTCF_PREC_BF16X3, false>(g, s);
                break;
            case DTCF_PREC_F16X3:
                cv ? lp_launch<DTCF_PREC_F16X3, true>(g, s) : lp_launch<DTCF_PREC_F16X3, false>(g, s);
                break;
            default:
                cv ? lp_launch<DTCF_PREC_I8, true>(g, s) : lp_launch<DTCF_PREC_I8, false>(g, s);
                break;
            }
            DTCF_HIP(hipGetLastError());
        }
    return DTCF_OK;
}
}  // namespace dtcf

extern "C" int dtcf_gemm_split_f64(int prec, int nslice, int64_t m, int64_t n, int64_t k,
                                   double alpha, const void *A_slices, const int32_t *ea,
                                   const void *B_slices, const int32_t *eb, int max_level,
                                   double *C, dtcf_view cv, int64_t ldc, void *stream)
{
    return gemm_split_f64_impl(prec, nslice, m, n, k, alpha, A_slices, ea, B_slices, eb, max_level,
                               C, cv, ldc, stream, 0);
}
