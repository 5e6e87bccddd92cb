// Strided permute-copy on gfx950: the one kernel behind RRR transpose
// (helper.cc:275-522), the DLTC transpose_<perm> family
// (DLTC/transp-kernel-seq.cc) and the create_big_matrix / revert_big_matrix
// pack and unpack (transpose_helper.cc:149-159, 393-403).
//
// The host folds dims that are adjacent in both input and output, then picks
// two dims for a 2-D tile:
//   q = the dim with the smallest output stride (output-fastest),
//   p = the dim with unit input stride (input-fastest) when it differs from q.
// TRANSPOSE tiles (p != q) go through LDS so that reads run along p and
// writes run along q, both coalesced; DIRECT tiles (q is fastest on both
// sides) copy straight through with lanes along q.  All other dims and the
// tile batch index are decoded once per workgroup.
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace dtcf {

namespace {

constexpr int kMaxOther = DTCF_MAX_DIMS;

struct PermParams {
    int32_t n_other;
    uint32_t other_ext[kMaxOther];
    FastDiv other_div[kMaxOther];
    int64_t other_istr[kMaxOther];
    int64_t other_ostr[kMaxOther];
    int64_t ext_p, ext_q;
    int64_t istr_p, istr_q, ostr_p, ostr_q;
    uint32_t ntp, ntq;          // tiles along p and q
    FastDiv div_ntq, div_ntp;
    uint64_t per_tile;          // workgroups per batch entry
    uint64_t total;             // total workgroups of work
    int64_t src_step, dst_step;
    double alpha, beta;
};

template <bool TRANSPOSE>
__global__ __launch_bounds__(256) void permute_kernel(const double *__restrict__ in,
                                                      double *__restrict__ out, PermParams P,
                                                      const int64_t *__restrict__ src_off,
                                                      const int64_t *__restrict__ dst_off)
{
    constexpr int TP = TRANSPOSE ? 32 : 16;   // tile rows along p
    constexpr int TQ = TRANSPOSE ? 32 : 64;   // tile cols along q
    __shared__ double lds[TRANSPOSE ? 32 : 1][TRANSPOSE ? 33 : 1];
    const int tid = threadIdx.x;

    for (uint64_t blk = blockIdx.x; blk < P.total; blk += gridDim.x) {
        const uint64_t t = blk / P.per_tile;
        uint32_t rem = (uint32_t)(blk - t * P.per_tile);
        const uint32_t qt_div = P.div_ntq.div(rem);
        const uint32_t qt = rem - qt_div * P.ntq;
        rem = qt_div;
        const uint32_t pt_div = P.div_ntp.div(rem);
        const uint32_t pt = rem - pt_div * P.ntp;
        rem = pt_div;
        int64_t src = src_off ? src_off[t] : (int64_t)t * P.src_step;
        int64_t dst = dst_off ? dst_off[t] : (int64_t)t * P.dst_step;
        for (int d = P.n_other - 1; d >= 0; d--) {
            const uint32_t q = P.other_div[d].div(rem);
            const uint32_t x = rem - q * P.other_ext[d];
            rem = q;
            src += (int64_t)x * P.other_istr[d];
            dst += (int64_t)x * P.other_ostr[d];
        }
        const int64_t p0 = (int64_t)pt * TP;
        const int64_t q0 = (int64_t)qt * TQ;

        if constexpr (TRANSPOSE) {
            const int tx = tid & 31, ty = tid >> 5;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t pp = p0 + tx, qq = q0 + ty + 8 * i;
                if (pp < P.ext_p && qq < P.ext_q)
                    lds[ty + 8 * i][tx] = in[src + pp * P.istr_p + qq * P.istr_q];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t qq = q0 + tx, pp = p0 + ty + 8 * i;
                if (pp < P.ext_p && qq < P.ext_q) {
                    double *o = out + dst + qq * P.ostr_q + pp * P.ostr_p;
                    const double v = lds[tx][ty + 8 * i];
                    if (P.beta == 0.0)
                        *o = (P.alpha == 1.0) ? v : P.alpha * v;
                    else
                        *o = P.alpha * v + P.beta * (*o);
                }
            }
            __syncthreads();
        } else {
            const int tx = tid & 63, ty = tid >> 6;
            double v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t qq = q0 + tx, pp = p0 + ty + 4 * i;
                if (pp < P.ext_p && qq < P.ext_q) v[i] = in[src + qq * P.istr_q + pp * P.istr_p];
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t qq = q0 + tx, pp = p0 + ty + 4 * i;
                if (pp < P.ext_p && qq < P.ext_q) {
                    double *o = out + dst + qq * P.ostr_q + pp * P.ostr_p;
                    if (P.beta == 0.0)
                        *o = (P.alpha == 1.0) ? v[i] : P.alpha * v[i];
                    else
                        *o = P.alpha * v[i] + P.beta * (*o);
                }
            }
        }
    }
}

// 16-byte forms of the two tiles, taken when the tile's fast strides are 1
// on their side, the extents and every other offset even and the buffers
// 16-byte aligned (no per-tile offset tables): TRANSPOSE 64 x 64 through
// LDS, 16 elements per thread, reads two consecutive p and writes two
// consecutive q per access; DIRECT 16 x 128, 8 elements per thread, q
// contiguous on both sides.  Same values as permute_kernel.
template <bool TRANSPOSE>
__global__ __launch_bounds__(256) void permute_vec_kernel(const double *__restrict__ in,
                                                          double *__restrict__ out, PermParams P)
{
    constexpr int TP = TRANSPOSE ? 64 : 16;
    constexpr int TQ = TRANSPOSE ? 64 : 128;
    __shared__ double lds[TRANSPOSE ? 64 : 1][TRANSPOSE ? 65 : 1];
    const int tid = threadIdx.x;
    for (uint64_t blk = blockIdx.x; blk < P.total; blk += gridDim.x) {
        const uint64_t t = blk / P.per_tile;
        uint32_t rem = (uint32_t)(blk - t * P.per_tile);
        const uint32_t qt_div = P.div_ntq.div(rem);
        const uint32_t qt = rem - qt_div * P.ntq;
        rem = qt_div;
        const uint32_t pt_div = P.div_ntp.div(rem);
        const uint32_t pt = rem - pt_div * P.ntp;
        rem = pt_div;
        int64_t src = (int64_t)t * P.src_step, dst = (int64_t)t * P.dst_step;
        for (int d = P.n_other - 1; d >= 0; d--) {
            const uint32_t q = P.other_div[d].div(rem);
            const uint32_t x = rem - q * P.other_ext[d];
            rem = q;
            src += (int64_t)x * P.other_istr[d];
            dst += (int64_t)x * P.other_ostr[d];
        }
        const int64_t p0 = (int64_t)pt * TP, q0 = (int64_t)qt * TQ;
        auto put = [&](double *o, double2 v) {
            if (P.beta == 0.0) {
                if (P.alpha != 1.0) {
                    v.x *= P.alpha;
                    v.y *= P.alpha;
                }
            } else {
                const double2 w = *reinterpret_cast<const double2 *>(o);
                v.x = P.alpha * v.x + P.beta * w.x;
                v.y = P.alpha * v.y + P.beta * w.y;
            }
            *reinterpret_cast<double2 *>(o) = v;
        };
        if constexpr (TRANSPOSE) {
            const int tx = tid & 31, ty = tid >> 5;
            double2 v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {  // rows q, pairs of p (unit input stride)
                const int64_t pp = p0 + 2 * tx, qq = q0 + ty + 8 * i;
                v[i] = (pp < P.ext_p && qq < P.ext_q)
                           ? *reinterpret_cast<const double2 *>(in + src + pp + qq * P.istr_q)
                           : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                lds[ty + 8 * i][2 * tx] = v[i].x;
                lds[ty + 8 * i][2 * tx + 1] = v[i].y;
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 8; i++) {  // rows p, pairs of q (unit output stride)
                const int64_t qq = q0 + 2 * tx, pp = p0 + ty + 8 * i;
                if (pp < P.ext_p && qq < P.ext_q)
                    put(out + dst + qq + pp * P.ostr_p,
                        make_double2(lds[2 * tx][ty + 8 * i], lds[2 * tx + 1][ty + 8 * i]));
            }
            __syncthreads();
        } else {
            const int tx = tid & 63, ty = tid >> 6;
            double2 v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t qq = q0 + 2 * tx, pp = p0 + ty + 4 * i;
                v[i] = (pp < P.ext_p && qq < P.ext_q)
                           ? *reinterpret_cast<const double2 *>(in + src + qq + pp * P.istr_p)
                           : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t qq = q0 + 2 * tx, pp = p0 + ty + 4 * i;
                if (pp < P.ext_p && qq < P.ext_q) put(out + dst + qq + pp * P.ostr_p, v[i]);
            }
        }
    }
}

struct Dim {
    int64_t ext, istr, ostr;
};

}  // namespace

// Validates and normalises a descriptor: drops unit dims and folds dims that
// are contiguous in both input and output.  Returns the folded dims in
// order of decreasing output stride.
static int fold_dims(const dtcf_strided *d, std::vector<Dim> &dims, bool &empty)
{
    DTCF_REQUIRE(d != nullptr, "permute: null descriptor");
    DTCF_REQUIRE(d->dims >= 1 && d->dims <= DTCF_MAX_DIMS, "permute: dims must be 1..8");
    empty = false;
    std::vector<Dim> v;
    for (int i = 0; i < d->dims; i++) {
        DTCF_REQUIRE(d->ext[i] >= 0, "permute: negative extent");
        if (d->ext[i] == 0) empty = true;
        if (d->ext[i] > 1) v.push_back({d->ext[i], d->istr[i], d->ostr[i]});
    }
    if (v.empty()) v.push_back({1, 1, 1});
    std::stable_sort(v.begin(), v.end(),
                     [](const Dim &a, const Dim &b) { return a.ostr > b.ostr; });
    std::vector<Dim> f;
    for (const Dim &x : v) {
        if (!f.empty()) {
            Dim &b = f.back();
            if (b.istr == x.istr * x.ext && b.ostr == x.ostr * x.ext) {
                b.ext *= x.ext;
                b.istr = x.istr;
                b.ostr = x.ostr;
                continue;
            }
        }
        f.push_back(x);
    }
    dims = f;
    return DTCF_OK;
}

}  // namespace dtcf

using namespace dtcf;

extern "C" int dtcf_permute_copy_f64(const double *in, double *out, const dtcf_strided *desc,
                                     int64_t ntiles, const int64_t *src_off, int64_t src_step,
                                     const int64_t *dst_off, int64_t dst_step, double alpha,
                                     double beta, void *stream)
{
    std::vector<Dim> dims;
    bool empty = false;
    int st = fold_dims(desc, dims, empty);
    if (st != DTCF_OK) return st;
    DTCF_REQUIRE(ntiles >= 0, "permute: negative tile count");
    if (empty || ntiles == 0) return DTCF_OK;
    DTCF_REQUIRE(in != nullptr && out != nullptr, "permute: null buffer");

    // q: output-fastest dim (folded dims are sorted, so it is the last one).
    const int q = (int)dims.size() - 1;
    // p: input-fastest dim if it is a different dim; otherwise the dim with
    // the next smallest output stride.
    int p = -1;
    int64_t best = INT64_MAX;
    for (int i = 0; i < (int)dims.size(); i++)
        if (i != q && dims[i].istr >= 0 && dims[i].istr < best) {
            best = dims[i].istr;
            p = i;
        }
    bool transpose = (p >= 0 && dims[p].istr < dims[q].istr);
    if (!transpose && dims.size() > 1) p = q - 1;

    PermParams P{};
    Dim dq = dims[q];
    Dim dp = p >= 0 ? dims[p] : Dim{1, 0, 0};
    P.ext_p = dp.ext;
    P.ext_q = dq.ext;
    P.istr_p = dp.istr;
    P.ostr_p = dp.ostr;
    P.istr_q = dq.istr;
    P.ostr_q = dq.ostr;
    // 16-byte form: unit fast strides, even extents and offsets, aligned
    // buffers, offsets from steps (not tables)
    bool vec = !src_off && !dst_off && (reinterpret_cast<uintptr_t>(in) & 15) == 0 &&
               (reinterpret_cast<uintptr_t>(out) & 15) == 0 && src_step % 2 == 0 &&
               dst_step % 2 == 0 && dq.ext % 2 == 0 && dq.ostr == 1;
    if (transpose) vec = vec && dp.istr == 1 && dp.ext % 2 == 0 && dq.istr % 2 == 0 && dp.ostr % 2 == 0;
    else vec = vec && dq.istr == 1 && (p < 0 || (dp.istr % 2 == 0 && dp.ostr % 2 == 0));
    for (int i = 0; vec && i < (int)dims.size(); i++)
        if (i != q && i != p) vec = dims[i].istr % 2 == 0 && dims[i].ostr % 2 == 0;
    const int TP = vec ? (transpose ? 64 : 16) : (transpose ? 32 : 16);
    const int TQ = vec ? (transpose ? 64 : 128) : (transpose ? 32 : 64);
    const uint64_t ntp = (dp.ext + TP - 1) / TP, ntq = (dq.ext + TQ - 1) / TQ;
    uint64_t per = ntp * ntq;
    int n_other = 0;
    for (int i = 0; i < (int)dims.size(); i++) {
        if (i == q || i == p) continue;
        P.other_ext[n_other] = (uint32_t)dims[i].ext;
        P.other_div[n_other] = FastDiv((uint32_t)dims[i].ext);
        P.other_istr[n_other] = dims[i].istr;
        P.other_ostr[n_other] = dims[i].ostr;
        per *= (uint64_t)dims[i].ext;
        n_other++;
    }
    DTCF_REQUIRE(per < (1ull << 31), "permute: a single tile is too large (> 2^31 workgroups)");
    P.n_other = n_other;
    P.ntp = (uint32_t)ntp;
    P.ntq = (uint32_t)ntq;
    P.div_ntp = FastDiv((uint32_t)ntp);
    P.div_ntq = FastDiv((uint32_t)ntq);
    P.per_tile = per;
    P.total = per * (uint64_t)ntiles;
    P.src_step = src_step;
    P.dst_step = dst_step;
    P.alpha = alpha;
    P.beta = beta;

    const uint64_t max_grid = 256ull * 64;  // plenty of waves in flight per CU
    const unsigned grid = (unsigned)std::min<uint64_t>(P.total, max_grid);
    hipStream_t s = as_stream(stream);
    if (vec && transpose)
        hipLaunchKernelGGL(permute_vec_kernel<true>, dim3(grid), dim3(256), 0, s, in, out, P);
    else if (vec)
        hipLaunchKernelGGL(permute_vec_kernel<false>, dim3(grid), dim3(256), 0, s, in, out, P);
    else if (transpose)
        hipLaunchKernelGGL(permute_kernel<true>, dim3(grid), dim3(256), 0, s, in, out, P, src_off,
                           dst_off);
    else
        hipLaunchKernelGGL(permute_kernel<false>, dim3(grid), dim3(256), 0, s, in, out, P, src_off,
                           dst_off);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

// Builds the descriptor for `ntiles` dense row-major tiles of extents w and
// output order out_src (output dim n is input dim out_src[n]).
int dtcf_build_tile_desc(int dims, const int64_t *w, const int32_t *out_src, dtcf_strided *d,
                         int64_t *block)
{
    DTCF_REQUIRE(dims >= 1 && dims <= DTCF_MAX_DIMS, "permute: dims must be 1..8");
    DTCF_REQUIRE(w != nullptr && out_src != nullptr, "permute: null extents or order");
    int seen[DTCF_MAX_DIMS] = {0};
    for (int n = 0; n < dims; n++) {
        DTCF_REQUIRE(out_src[n] >= 0 && out_src[n] < dims && !seen[out_src[n]],
                     "permute: out_src is not a permutation");
        seen[out_src[n]] = 1;
        DTCF_REQUIRE(w[n] >= 0, "permute: negative extent");
    }
    int64_t is[DTCF_MAX_DIMS];
    is[dims - 1] = 1;
    for (int i = dims - 2; i >= 0; i--) is[i] = is[i + 1] * w[i + 1];
    d->dims = dims;
    int64_t os = 1;
    for (int n = dims - 1; n >= 0; n--) {
        d->ext[n] = w[out_src[n]];
        d->istr[n] = is[out_src[n]];
        d->ostr[n] = os;
        os *= d->ext[n];
    }
    *block = os;
    return DTCF_OK;
}

extern "C" int dtcf_permute_tiles_f64(const double *in, double *out, int dims, const int64_t *w,
                                      const int32_t *out_src, int64_t ntiles, void *stream)
{
    dtcf_strided d{};
    int64_t block = 0;
    int st = dtcf_build_tile_desc(dims, w, out_src, &d, &block);
    if (st != DTCF_OK) return st;
    return dtcf_permute_copy_f64(in, out, &d, ntiles, nullptr, block, nullptr, block, 1.0, 0.0,
                                 stream);
}
