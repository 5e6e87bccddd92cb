// C-ABI layer: argument checking, view objects, GEMM dispatch, host-pointer
// staging and the gather/scatter kernels behind the big-matrix pack.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "gemm_f64.hpp"

namespace dtcf {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
const char *get_error() { return g_err.c_str(); }

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? DTCF_ERR_NOMEM : DTCF_ERR_HIP;
}

static std::atomic<int> g_cfg{-1};

int ccsd_scratch_setup(int dev);  // ccsd.hip

// Tile-shape choice when no explicit configuration is set.  The table is
// filled by the autotuner (dtcf_amd/autotune.py) from measurements on the
// device; the defaults keep 128x128 everywhere except small grids, where a
// smaller tile gives more workgroups, and thin results (64 or 96 wide), whose
// tile matches their width.
int pick_config(int64_t m, int64_t n, int64_t k)
{
    const int forced = g_cfg.load();
    if (forced >= 0) return forced;
    (void)k;
    const int64_t tiles = ((m + 127) / 128) * ((n + 127) / 128);
    if (tiles < 256) return kCfg64x64;
    // thin results: a 128-wide tile would be mostly padding
    if (n <= 64 || m <= 64) return kCfg64x64;
    if (n <= 96) return kCfg192x96;
    if (m <= 96) return kCfg96x192;
    return kCfg128x128;
}

}  // namespace dtcf

using namespace dtcf;

#include "view.hpp"

extern "C" {

const char *dtcf_last_error(void) { return get_error(); }

int dtcf_abi_version(void) { return 100; }

int dtcf_init(int device)
{
    DTCF_HIP(hipSetDevice(device));
    DTCF_HIP(hipFree(nullptr));
    return ccsd_scratch_setup(device);
}

int dtcf_gemm_set_config(int cfg)
{
    DTCF_REQUIRE(cfg >= -1 && cfg < kCfgCount, "gemm: unknown tile configuration");
    g_cfg.store(cfg);
    return DTCF_OK;
}

int dtcf_gemm_f64(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                  int64_t ldc, void *stream)
{
    DTCF_REQUIRE(m >= 0 && n >= 0 && k >= 0, "gemm: negative size");
    if (m == 0 || n == 0 || k == 0 || alpha == 0.0) return DTCF_OK;
    DTCF_REQUIRE(A && B && C, "gemm: null operand");
    DTCF_REQUIRE(ldc >= n, "gemm: ldc < n");
    DTCF_REQUIRE(transa ? lda >= m : lda >= k, "gemm: lda too small");
    DTCF_REQUIRE(transb ? ldb >= k : ldb >= n, "gemm: ldb too small");
    GemmArgs g{};
    g.m = m;
    g.n = n;
    g.k = k;
    g.alpha = alpha;
    g.A = A;
    g.B = B;
    g.C = C;
    // A view is m x k; B view is n x k.
    g.amode = transa ? kDMN : kDK;
    g.a.ld = lda;
    g.bmode = transb ? kDK : kDMN;
    g.b.ld = ldb;
    g.a.vec2 = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && (lda % 2 == 0);
    g.b.vec2 = ((reinterpret_cast<uintptr_t>(B) & 15) == 0) && (ldb % 2 == 0);
    g.c.ld = ldc;
    g.cview = 0;
    return gemm_f64_launch(g, pick_config(m, n, k), as_stream(stream));
}

int dtcf_gemm_group_f64(const dtcf_gemm_problem *probs, int nprob, double alpha, int overwrite,
                        int cfg, void *stream)
{
    DTCF_REQUIRE(nprob >= 0 && (probs || nprob == 0), "gemm group: null problem list");
    DTCF_REQUIRE(cfg >= 0 && cfg < kCfgCount, "gemm group: unknown tile configuration");
    std::vector<int> order(nprob);
    for (int i = 0; i < nprob; i++) {
        const dtcf_gemm_problem &q = probs[i];
        DTCF_REQUIRE(q.m > 0 && q.n > 0 && q.k > 0, "gemm group: empty problem");
        DTCF_REQUIRE(q.A && q.B && q.C, "gemm group: null operand");
        DTCF_REQUIRE(q.lda >= q.k && q.ldb >= q.k && q.ldc >= q.n, "gemm group: pitch too small");
        order[i] = i;
    }
    // Longest tiles first: the hardware dispatches workgroups in id order.
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return probs[x].k > probs[y].k; });
    for (int b = 0; b < nprob; b += kGroupMax) {
        GemmGroupArgs ga{};
        ga.alpha = alpha;
        ga.overwrite = overwrite ? 1 : 0;
        ga.np = std::min(kGroupMax, nprob - b);
        for (int i = 0; i < ga.np; i++) {
            const dtcf_gemm_problem &q = probs[order[b + i]];
            GemmGroupProb &P = ga.p[i];
            P.m = q.m;
            P.n = q.n;
            P.k = q.k;
            P.A = q.A;
            P.B = q.B;
            P.C = q.C;
            P.lda = q.lda;
            P.ldb = q.ldb;
            P.ldc = q.ldc;
            P.vec2a = ((reinterpret_cast<uintptr_t>(q.A) & 15) == 0) && (q.lda % 2 == 0);
            P.vec2b = ((reinterpret_cast<uintptr_t>(q.B) & 15) == 0) && (q.ldb % 2 == 0);
        }
        const int st = gemm_f64_group_launch(ga, cfg, as_stream(stream));
        if (st != DTCF_OK) return st;
    }
    return DTCF_OK;
}

int dtcf_view_create(int64_t rows, int64_t cols, const int64_t *rofs, const int64_t *cofs,
                     const int64_t *blk, int64_t rpb, int64_t cpb, dtcf_view *out)
{
    DTCF_REQUIRE(out != nullptr, "view: null output handle");
    *out = nullptr;
    DTCF_REQUIRE(rows > 0 && cols > 0, "view: empty view");
    DTCF_REQUIRE(rows < (1ll << 31) && cols < (1ll << 31), "view: extent >= 2^31");
    DTCF_REQUIRE(rofs && cofs, "view: null offset table");
    int64_t nrb = 1, ncb = 1;
    if (blk) {
        DTCF_REQUIRE(rpb > 0 && cpb > 0 && rows % rpb == 0 && cols % cpb == 0,
                     "view: rows/cols must be whole blocks");
        nrb = rows / rpb;
        ncb = cols / cpb;
    } else {
        rpb = rows;
        cpb = cols;
    }
    int64_t bmax = 0;
    for (int64_t r = 0; r < rows; r++) DTCF_REQUIRE(rofs[r] >= 0, "view: negative row offset");
    for (int64_t c = 0; c < cols; c++) DTCF_REQUIRE(cofs[c] >= 0, "view: negative column offset");
    if (blk)
        for (int64_t b = 0; b < nrb * ncb; b++) {
            DTCF_REQUIRE(blk[b] >= 0, "view: negative block offset");
            bmax = std::max(bmax, blk[b]);
        }
    // Largest addressable element: rows and columns are independent except
    // through the block table.
    int64_t span = 0;
    {
        int64_t rm = 0, cm = 0;
        for (int64_t r = 0; r < rows; r++) rm = std::max(rm, rofs[r]);
        for (int64_t c = 0; c < cols; c++) cm = std::max(cm, cofs[c]);
        span = rm + cm + bmax + 1;
    }
    // A block table that is a sum of a row-block part and a column-block part
    // (every dense tile grid) is folded into the row and column tables, so
    // the kernels read rofs[r] + cofs[c] without the per-element block lookup.
    std::vector<int64_t> rf, cf;
    if (blk) {
        bool sep = true;
        for (int64_t R = 0; sep && R < nrb; R++)
            for (int64_t C = 0; sep && C < ncb; C++)
                sep = blk[R * ncb + C] == blk[R * ncb] + blk[C] - blk[0];
        if (sep) {
            rf.assign(rofs, rofs + rows);
            cf.assign(cofs, cofs + cols);
            for (int64_t r = 0; r < rows; r++) rf[r] += blk[(r / rpb) * ncb];
            for (int64_t c = 0; c < cols; c++) cf[c] += blk[c / cpb] - blk[0];
            rofs = rf.data();
            cofs = cf.data();
            blk = nullptr;
            nrb = ncb = 1;
            rpb = rows;
            cpb = cols;
        }
    }
    // Pairs (c, c+1) for even c are adjacent inside one block, and the pair
    // start is even (16-byte aligned for an aligned buffer).
    bool reven = true, ceven = true, beven = true;
    for (int64_t r = 0; r < rows; r++) reven &= (rofs[r] % 2 == 0);
    for (int64_t c = 0; c < cols; c++) ceven &= (cofs[c] % 2 == 0);
    if (blk)
        for (int64_t b = 0; b < nrb * ncb; b++) beven &= (blk[b] % 2 == 0);
    bool colpairs = beven && reven && (cols % 2 == 0) && (cpb % 2 == 0);
    for (int64_t c = 0; colpairs && c + 1 < cols; c += 2)
        colpairs = (cofs[c + 1] == cofs[c] + 1) && (cofs[c] % 2 == 0);
    bool rowpairs = beven && ceven && (rows % 2 == 0) && (rpb % 2 == 0);
    for (int64_t r = 0; rowpairs && r + 1 < rows; r += 2)
        rowpairs = (rofs[r + 1] == rofs[r] + 1) && (rofs[r] % 2 == 0);

    dtcf_view_s *v = new dtcf_view_s();
    v->rows = rows;
    v->cols = cols;
    v->rpb = rpb;
    v->cpb = cpb;
    v->ncb = ncb;
    v->nblk = blk ? nrb * ncb : 0;
    v->span = span;
    v->colpairs = colpairs;
    v->rowpairs = rowpairs;
    v->rowrun16 = !blk;
    for (int64_t r = 0; v->rowrun16 && r < rows; r++) v->rowrun16 = rofs[r] == rofs[r & ~15ll] + (r & 15);
    hipError_t e = hipMalloc(&v->rofs, rows * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&v->cofs, cols * sizeof(int64_t));
    if (e == hipSuccess && blk) e = hipMalloc(&v->blk, v->nblk * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemcpy(v->rofs, rofs, rows * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(v->cofs, cofs, cols * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && blk)
        e = hipMemcpy(v->blk, blk, v->nblk * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dtcf_view_destroy(v);
        return hip_fail(e, "view: device tables");
    }
    *out = v;
    return DTCF_OK;
}

int dtcf_view_destroy(dtcf_view v)
{
    if (!v) return DTCF_OK;
    if (v->owner) {
        if (v->rofs) (void)hipFree(v->rofs);
        if (v->cofs) (void)hipFree(v->cofs);
        if (v->blk) (void)hipFree(v->blk);
    }
    delete v;
    return DTCF_OK;
}

int dtcf_view_sub(dtcf_view p, int64_t r0, int64_t rows, int64_t c0, int64_t cols, dtcf_view *out)
{
    DTCF_REQUIRE(out, "view_sub: null output handle");
    *out = nullptr;
    DTCF_REQUIRE(p, "view_sub: null parent");
    DTCF_REQUIRE(r0 >= 0 && rows > 0 && r0 + rows <= p->rows && c0 >= 0 && cols > 0 &&
                     c0 + cols <= p->cols,
                 "view_sub: rectangle outside the parent view");
    if (p->blk)
        DTCF_REQUIRE(r0 % p->rpb == 0 && rows % p->rpb == 0 && c0 % p->cpb == 0 &&
                         cols % p->cpb == 0,
                     "view_sub: a view with a block table is cut on block boundaries only");
    dtcf_view_s *v = new dtcf_view_s(*p);
    v->owner = false;
    v->rows = rows;
    v->cols = cols;
    v->rofs = p->rofs + r0;
    v->cofs = p->cofs + c0;
    if (p->blk) {
        v->blk = p->blk + (r0 / p->rpb) * p->ncb + c0 / p->cpb;
    } else {
        v->rpb = rows;
        v->cpb = cols;
    }
    v->colpairs = p->colpairs && c0 % 2 == 0 && cols % 2 == 0;
    v->rowpairs = p->rowpairs && r0 % 2 == 0 && rows % 2 == 0;
    v->rowrun16 = p->rowrun16 && r0 % 16 == 0;
    *out = v;
    return DTCF_OK;
}

int64_t dtcf_view_span(dtcf_view v) { return v ? v->span : -1; }

}  // extern "C"

static void fill_op(OpArgs &o, const dtcf_view_s *v)
{
    o.rofs = v->rofs;
    o.cofs = v->cofs;
    o.blk = v->blk;
    o.div_rpb = FastDiv((uint32_t)v->rpb);
    o.div_cpb = FastDiv((uint32_t)v->cpb);
    o.ncb = v->ncb;
}

static int gemm_view_impl(int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                          dtcf_view av, const double *B, dtcf_view bv, double *C, dtcf_view cv,
                          int64_t ldc, void *stream, int overwrite)
{
    DTCF_REQUIRE(m >= 0 && n >= 0 && k >= 0, "gemm_view: negative size");
    if (m == 0 || n == 0 || k == 0 || (alpha == 0.0 && !overwrite)) return DTCF_OK;
    DTCF_REQUIRE(A && B && C && av && bv, "gemm_view: null operand or view");
    DTCF_REQUIRE(av->rows == m && av->cols == k, "gemm_view: A view is not m x k");
    DTCF_REQUIRE(bv->rows == n && bv->cols == k, "gemm_view: B view is not n x k");
    GemmArgs g{};
    g.m = m;
    g.n = n;
    g.k = k;
    g.alpha = alpha;
    g.overwrite = overwrite;
    g.A = A;
    g.B = B;
    g.C = C;
    fill_op(g.a, av);
    fill_op(g.b, bv);
    const bool aal = (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    const bool bal = (reinterpret_cast<uintptr_t>(B) & 15) == 0;
    // Prefer k-contiguous loads; fall back to row-contiguous; otherwise
    // element loads through the k-mode path.
    if (av->colpairs) { g.amode = kVK; g.a.vec2 = aal; }
    else if (av->rowpairs) { g.amode = kVMN; g.a.vec2 = aal; }
    else { g.amode = kVK; g.a.vec2 = 0; }
    if (bv->colpairs) { g.bmode = kVK; g.b.vec2 = bal; }
    else if (bv->rowpairs) { g.bmode = kVMN; g.b.vec2 = bal; }
    else { g.bmode = kVK; g.b.vec2 = 0; }
    if (cv) {
        DTCF_REQUIRE(cv->rows == m && cv->cols == n, "gemm_view: C view is not m x n");
        fill_op(g.c, cv);
        g.cview = 1;
        return gemm_f64_launch(g, pick_config(m, n, k), as_stream(stream));
    }
    DTCF_REQUIRE(ldc >= n, "gemm_view: ldc < n");
    g.c.ld = ldc;
    g.cview = 0;
    return gemm_f64_launch(g, pick_config(m, n, k), as_stream(stream));
}

extern "C" int dtcf_gemm_view_f64(int64_t m, int64_t n, int64_t k, double alpha,
                                  const double *A, dtcf_view av, const double *B, dtcf_view bv,
                                  double *C, dtcf_view cv, int64_t ldc, void *stream)
{
    return gemm_view_impl(m, n, k, alpha, A, av, B, bv, C, cv, ldc, stream, 0);
}

int dtcf::gemm_view_overwrite(int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                              dtcf_view av, const double *B, dtcf_view bv, double *C, dtcf_view cv,
                              int64_t ldc, hipStream_t stream)
{
    return gemm_view_impl(m, n, k, alpha, A, av, B, bv, C, cv, ldc, stream, 1);
}

// ---------------------------------------------------------------------------
// Gather / scatter between a view and a dense row-major matrix.  32x32 tiles;
// when the view's rows are the contiguous direction the tile is transposed
// through LDS so both sides stay coalesced.
// ---------------------------------------------------------------------------
namespace {

struct GatherArgs {
    const int64_t *rofs, *cofs, *blk;
    FastDiv div_rpb, div_cpb;
    int64_t ncb, rows, cols, ld;
    uint32_t tiles_c;
};

__device__ inline int64_t view_addr(const GatherArgs &g, int64_t r, int64_t c)
{
    int64_t a = g.rofs[r] + g.cofs[c];
    if (g.blk) a += g.blk[(int64_t)g.div_rpb.div((uint32_t)r) * g.ncb + g.div_cpb.div((uint32_t)c)];
    return a;
}

// SCATTER=false: dst[r*ld+c] = src[view(r,c)];  true: dst[view(r,c)] = src[r*ld+c]
template <bool SCATTER, bool ROWMAJOR_VIEW>
__global__ __launch_bounds__(256) void view_copy_kernel(const double *__restrict__ src,
                                                        double *__restrict__ dst, GatherArgs g)
{
    __shared__ double lds[32][33];
    const uint32_t tr = blockIdx.x / g.tiles_c, tc = blockIdx.x - tr * g.tiles_c;
    const int64_t r0 = (int64_t)tr * 32, c0 = (int64_t)tc * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if constexpr (!ROWMAJOR_VIEW) {
        // Lanes along c on both sides.
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t r = r0 + ty + 8 * i, c = c0 + tx;
            if (r < g.rows && c < g.cols) {
                if constexpr (SCATTER) dst[view_addr(g, r, c)] = src[r * g.ld + c];
                else dst[r * g.ld + c] = src[view_addr(g, r, c)];
            }
        }
    } else {
        // The view is contiguous along r: move the tile through LDS.
        if constexpr (!SCATTER) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t r = r0 + tx, c = c0 + ty + 8 * i;
                if (r < g.rows && c < g.cols) lds[ty + 8 * i][tx] = src[view_addr(g, r, c)];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t r = r0 + ty + 8 * i, c = c0 + tx;
                if (r < g.rows && c < g.cols) dst[r * g.ld + c] = lds[tx][ty + 8 * i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t r = r0 + ty + 8 * i, c = c0 + tx;
                if (r < g.rows && c < g.cols) lds[tx][ty + 8 * i] = src[r * g.ld + c];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t r = r0 + tx, c = c0 + ty + 8 * i;
                if (r < g.rows && c < g.cols) dst[view_addr(g, r, c)] = lds[ty + 8 * i][tx];
            }
        }
    }
}

int view_copy(const double *src, double *dst, dtcf_view v, int64_t ld, bool scatter, void *stream)
{
    DTCF_REQUIRE(src && dst && v, "view copy: null argument");
    DTCF_REQUIRE(ld >= v->cols, "view copy: ld < cols");
    GatherArgs g{};
    g.rofs = v->rofs;
    g.cofs = v->cofs;
    g.blk = v->blk;
    g.div_rpb = FastDiv((uint32_t)v->rpb);
    g.div_cpb = FastDiv((uint32_t)v->cpb);
    g.ncb = v->ncb;
    g.rows = v->rows;
    g.cols = v->cols;
    g.ld = ld;
    g.tiles_c = (uint32_t)((v->cols + 31) / 32);
    const uint64_t tiles = (uint64_t)((v->rows + 31) / 32) * g.tiles_c;
    DTCF_REQUIRE(tiles < (1ull << 31), "view copy: too many tiles");
    const bool rowmajor = !v->colpairs && v->rowpairs;
    hipStream_t s = as_stream(stream);
    if (scatter) {
        if (rowmajor) hipLaunchKernelGGL((view_copy_kernel<true, true>), dim3(tiles), dim3(256), 0, s, src, dst, g);
        else hipLaunchKernelGGL((view_copy_kernel<true, false>), dim3(tiles), dim3(256), 0, s, src, dst, g);
    } else {
        if (rowmajor) hipLaunchKernelGGL((view_copy_kernel<false, true>), dim3(tiles), dim3(256), 0, s, src, dst, g);
        else hipLaunchKernelGGL((view_copy_kernel<false, false>), dim3(tiles), dim3(256), 0, s, src, dst, g);
    }
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

}  // namespace

extern "C" int dtcf_view_gather_f64(const double *src, dtcf_view v, double *dst, int64_t ld,
                                    void *stream)
{
    return view_copy(src, dst, v, ld, false, stream);
}

extern "C" int dtcf_view_scatter_f64(const double *src, int64_t ld, dtcf_view v, double *dst,
                                     void *stream)
{
    return view_copy(src, dst, v, ld, true, stream);
}

// ---------------------------------------------------------------------------
// Host-pointer variants: stage through device buffers on a private stream.
// ---------------------------------------------------------------------------
namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};
}  // namespace

extern "C" int dtcf_gemm_f64_host(int transa, int transb, int64_t m, int64_t n, int64_t k,
                                  double alpha, const double *A, int64_t lda, const double *B,
                                  int64_t ldb, double *C, int64_t ldc)
{
    DTCF_REQUIRE(m >= 0 && n >= 0 && k >= 0, "gemm: negative size");
    if (m == 0 || n == 0 || k == 0 || alpha == 0.0) return DTCF_OK;
    DTCF_REQUIRE(A && B && C, "gemm: null operand");
    const int64_t asz = transa ? (k - 1) * lda + m : (m - 1) * lda + k;
    const int64_t bsz = transb ? (n - 1) * ldb + k : (k - 1) * ldb + n;
    const int64_t csz = (m - 1) * ldc + n;
    DevBuf da, db, dc;
    DTCF_HIP(hipMalloc(&da.p, asz * sizeof(double)));
    DTCF_HIP(hipMalloc(&db.p, bsz * sizeof(double)));
    DTCF_HIP(hipMalloc(&dc.p, csz * sizeof(double)));
    DTCF_HIP(hipMemcpy(da.p, A, asz * sizeof(double), hipMemcpyHostToDevice));
    DTCF_HIP(hipMemcpy(db.p, B, bsz * sizeof(double), hipMemcpyHostToDevice));
    DTCF_HIP(hipMemcpy(dc.p, C, csz * sizeof(double), hipMemcpyHostToDevice));
    int st = dtcf_gemm_f64(transa, transb, m, n, k, alpha, (const double *)da.p, lda,
                           (const double *)db.p, ldb, (double *)dc.p, ldc, nullptr);
    if (st != DTCF_OK) return st;
    DTCF_HIP(hipMemcpy(C, dc.p, csz * sizeof(double), hipMemcpyDeviceToHost));
    return DTCF_OK;
}

extern "C" int dtcf_permute_tiles_f64_host(const double *in, double *out, int dims,
                                           const int64_t *w, const int32_t *out_src,
                                           int64_t ntiles)
{
    DTCF_REQUIRE(dims >= 1 && dims <= DTCF_MAX_DIMS && w, "permute: bad dims");
    int64_t bs = 1;
    for (int i = 0; i < dims; i++) {
        DTCF_REQUIRE(w[i] >= 0, "permute: negative extent");
        bs *= w[i];
    }
    const int64_t total = bs * ntiles;
    if (total == 0) return DTCF_OK;
    DTCF_REQUIRE(in && out, "permute: null buffer");
    DevBuf di, dout;
    DTCF_HIP(hipMalloc(&di.p, total * sizeof(double)));
    DTCF_HIP(hipMalloc(&dout.p, total * sizeof(double)));
    DTCF_HIP(hipMemcpy(di.p, in, total * sizeof(double), hipMemcpyHostToDevice));
    int st = dtcf_permute_tiles_f64((const double *)di.p, (double *)dout.p, dims, w, out_src,
                                    ntiles, nullptr);
    if (st != DTCF_OK) return st;
    DTCF_HIP(hipMemcpy(out, dout.p, total * sizeof(double), hipMemcpyDeviceToHost));
    return DTCF_OK;
}

extern "C" int dtcf_pack_blocks_f64_host(double *tiles, int64_t ntiles, const int64_t *order,
                                         int64_t nrb, int64_t ncb, int64_t rpb, int64_t cpb,
                                         double *big, int unpack)
{
    DTCF_REQUIRE(ntiles >= 0 && nrb >= 0 && ncb >= 0 && rpb > 0 && cpb > 0, "pack: bad sizes");
    if (nrb == 0 || ncb == 0) return DTCF_OK;
    DTCF_REQUIRE(tiles && order && big, "pack: null argument");
    const int64_t bs = rpb * cpb, rows = nrb * rpb, cols = ncb * cpb;
    std::vector<int64_t> rofs(rows), cofs(cols), blk(nrb * ncb);
    for (int64_t r = 0; r < rows; r++) rofs[r] = (r % rpb) * cpb;
    for (int64_t c = 0; c < cols; c++) cofs[c] = c % cpb;
    for (int64_t b = 0; b < nrb * ncb; b++) {
        DTCF_REQUIRE(order[b] >= 0 && order[b] < ntiles, "pack: tile number out of range");
        blk[b] = order[b] * bs;
    }
    dtcf_view v = nullptr;
    int st = dtcf_view_create(rows, cols, rofs.data(), cofs.data(), blk.data(), rpb, cpb, &v);
    if (st != DTCF_OK) return st;
    DevBuf dt, dbig;
    hipError_t e = hipMalloc(&dt.p, ntiles * bs * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dbig.p, rows * cols * sizeof(double));
    if (e == hipSuccess)
        e = unpack ? hipMemcpy(dbig.p, big, rows * cols * sizeof(double), hipMemcpyHostToDevice)
                   : hipMemcpy(dt.p, tiles, ntiles * bs * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && unpack)  // tiles outside the order keep their values
        e = hipMemcpy(dt.p, tiles, ntiles * bs * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dtcf_view_destroy(v);
        return hip_fail(e, "pack: staging");
    }
    st = unpack ? dtcf_view_scatter_f64((const double *)dbig.p, cols, v, (double *)dt.p, nullptr)
                : dtcf_view_gather_f64((const double *)dt.p, v, (double *)dbig.p, cols, nullptr);
    dtcf_view_destroy(v);
    if (st != DTCF_OK) return st;
    e = unpack ? hipMemcpy(tiles, dt.p, ntiles * bs * sizeof(double), hipMemcpyDeviceToHost)
               : hipMemcpy(big, dbig.p, rows * cols * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "pack: copy back");
    return DTCF_OK;
}
