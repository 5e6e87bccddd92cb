// C-ABI layer: argument checking, view objects, GEMM dispatch, host-pointer
// staging and the gather/scatter kernels behind the big-matrix pack.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
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
    // Measured on gfx950 (view operands, every configuration forced in turn):
    // with K <= 1024 the 64 x 64 tile wins or ties at every shape (9216 x
    // 9216 x 192: 42.6 against 30.8 TF/s; 36864 x 2304 x 768: 54.4 / 54.0):
    // a short K loop leaves the tile prologue and epilogue a large share,
    // and two workgroups per CU overlap them.  Longer K takes 128 x 128
    // unless its padding costs more than 12% over 64 x 64's (m = 192: 256
    // padded rows; 1000 x 1000 x 100000: 63.6 against 50.7 TF/s, with the
    // few tiles split over K).  192 x 96 and 96 x 192 were never the
    // fastest (9216^3: 42.8 / 47.1 against 67.0 TF/s) and are used only
    // when forced.
    // Thin results (one side at most 48) with short K stream their long
    // operand past the thin one held in LDS (thin_f64_launch): 4.2 against
    // 3.2 TB/s at 7077888 x 48 x 192, even at 442368 rows, slower than 64 x
    // 64 tiles below about 262144 (110592 x 48 x 192: 75 against 72 us).
    if (thin_shape(m, n, k, 262144)) return kCfgThin;
    // Thin results with long K: a 64-wide tile three times as deep re-reads
    // the thin operand a third as often (192 x 48 x 1769472: 0.95 against
    // 1.07 ms; 9216 x 48 x 9216: 0.25 against 0.28 ms).
    if (k >= 8192 && std::min(m, n) <= 64 && std::max(m, n) >= 192) {
        const int64_t w = std::max(m, n);
        if ((w + 191) / 192 * 192 <= 1.12 * ((w + 63) / 64 * 64)) return m >= n ? kCfg192x64 : kCfg64x192;
    }
    if (k <= 1024) return kCfg64x64;
    auto padded = [&](int64_t t) { return (double)((m + t - 1) / t * t) * (double)((n + t - 1) / t * t); };
    return padded(128) <= 1.12 * padded(64) ? kCfg128x128 : kCfg64x64;
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

int dtcf_gemm_reserve_workspace(void *stream)
{
    return gemm_reserve_workspace(as_stream(stream));
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

int dtcf_gemm_group_ex_f64(const dtcf_gemm_problem *probs, int nprob, int transa, int transb,
                           double alpha, int overwrite, int cfg, void *stream)
{
    DTCF_REQUIRE(nprob >= 0 && (probs || nprob == 0), "gemm group: null problem list");
    DTCF_REQUIRE(cfg >= 0 && cfg < kCfgGroupCount, "gemm group: unknown tile configuration");
    DTCF_REQUIRE((transa == 0 || transa == 1) && (transb == 0 || transb == 1),
                 "gemm group: transa / transb must be 0 or 1");
    std::vector<int> order(nprob);
    for (int i = 0; i < nprob; i++) {
        const dtcf_gemm_problem &q = probs[i];
        DTCF_REQUIRE(q.m > 0 && q.n > 0 && q.k > 0, "gemm group: empty problem");
        DTCF_REQUIRE(q.A && q.B && q.C, "gemm group: null operand");
        DTCF_REQUIRE(q.lda >= (transa ? q.m : q.k) && q.ldb >= (transb ? q.k : q.n) && q.ldc >= q.n,
                     "gemm group: pitch too small");
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
        const int st = gemm_f64_group_launch(ga, cfg, transa ? kDMN : kDK, transb ? kDK : kDMN,
                                             as_stream(stream));
        if (st != DTCF_OK) return st;
    }
    return DTCF_OK;
}

int dtcf_gemm_group_f64(const dtcf_gemm_problem *probs, int nprob, double alpha, int overwrite,
                        int cfg, void *stream)
{
    return dtcf_gemm_group_ex_f64(probs, nprob, 0, 1, alpha, overwrite, cfg, stream);
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
    if (cv) DTCF_REQUIRE(cv->rows == m && cv->cols == n, "gemm_view: C view is not m x n");
    else DTCF_REQUIRE(ldc >= n, "gemm_view: ldc < n");
    // A matrix-vector product (n = 1 or m = 1) whose long operand has its
    // column pairs adjacent and 16-byte aligned runs as a streaming product
    // (gemv_kernel) instead of MFMA tiles of one useful column.
    {
        const bool col = n == 1, row = !col && m == 1;
        const dtcf_view_s *mv = col ? av : bv, *xv = col ? bv : av;
        const double *M = col ? A : B, *X = col ? B : A;
        const bool aligned = (reinterpret_cast<uintptr_t>(M) & 15) == 0;
        const bool colwise = mv->colpairs && k % 2 == 0 && k <= kGemvMaxK;
        if ((col || row) && (colwise || mv->rowpairs) && !mv->blk && aligned) {
            GemvArgs q{};
            q.rows = col ? m : n;
            q.k = k;
            q.alpha = alpha;
            q.M = M;
            q.mrofs = mv->rofs;
            q.mcofs = mv->cofs;
            q.X = X;
            fill_op(q.xv, xv);
            q.Y = C;
            q.ycol = col ? 1 : 0;
            q.overwrite = overwrite;
            if (cv) {
                fill_op(q.yv, cv);
                q.yview = 1;
            } else {
                q.ldy = ldc;
            }
            return colwise ? gemv_f64_launch(q, as_stream(stream)) : gemv_rows_f64_launch(q, as_stream(stream));
        }
    }
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
        fill_op(g.c, cv);
        // column pairs of C adjacent and 16-byte aligned (the thin kernel's
        // paired stores when C runs along the long operand)
        g.c.vec2 = cv->colpairs && (reinterpret_cast<uintptr_t>(C) & 15) == 0;
        g.cview = 1;
        return gemm_f64_launch(g, pick_config(m, n, k), as_stream(stream));
    }
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
// Host-pointer variants.  Per device: device buffers kept between calls
// (grown on demand, never freed), two pinned host staging buffers and a
// private stream.  Host memory moves through the pinned buffers in chunks:
// the CPU fills (or drains) one while the DMA engine moves the other, so a
// call costs one pass of host memcpy overlapped with the PCIe transfer and
// no allocation once the buffers have grown.  Calls on one device are
// serialised (the reference calls these from one host thread per rank).
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kStageChunk = size_t(64) << 20;  // bytes per pinned buffer

struct HostStage {
    std::mutex mu;
    bool ready = false;
    hipStream_t s = nullptr;
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    void *dev[3] = {nullptr, nullptr, nullptr};
    size_t cap[3] = {0, 0, 0};

    int init()
    {
        if (ready) return DTCF_OK;
        DTCF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (int b = 0; b < 2; b++) {
            DTCF_HIP(hipHostMalloc(&pin[b], kStageChunk, hipHostMallocDefault));
            DTCF_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
            DTCF_HIP(hipEventRecord(ev[b], s));  // recorded once: waiting on it is always defined
        }
        ready = true;
        return DTCF_OK;
    }
    // device buffer `slot` of at least `bytes`
    int buf(int slot, size_t bytes, void **p)
    {
        if (cap[slot] < bytes) {
            if (dev[slot]) {
                DTCF_HIP(hipStreamSynchronize(s));
                DTCF_HIP(hipFree(dev[slot]));
                dev[slot] = nullptr;
                cap[slot] = 0;
            }
            DTCF_HIP(hipMalloc(&dev[slot], bytes));
            cap[slot] = bytes;
        }
        *p = dev[slot];
        return DTCF_OK;
    }
    int h2d(void *dst, const void *src, size_t bytes)
    {
        const char *h = static_cast<const char *>(src);
        char *d = static_cast<char *>(dst);
        for (size_t off = 0, c = 0; off < bytes; off += kStageChunk, c++) {
            const int b = (int)(c & 1);
            const size_t len = std::min(kStageChunk, bytes - off);
            DTCF_HIP(hipEventSynchronize(ev[b]));  // every earlier DMA from it is done
            std::memcpy(pin[b], h + off, len);
            DTCF_HIP(hipMemcpyAsync(d + off, pin[b], len, hipMemcpyHostToDevice, s));
            DTCF_HIP(hipEventRecord(ev[b], s));
        }
        return DTCF_OK;
    }
    int d2h(void *dst, const void *src, size_t bytes)
    {
        char *h = static_cast<char *>(dst);
        const char *d = static_cast<const char *>(src);
        size_t prev_off = 0, prev_len = 0;
        size_t c = 0;
        for (size_t off = 0; off < bytes; off += kStageChunk, c++) {
            const int b = (int)(c & 1);
            const size_t len = std::min(kStageChunk, bytes - off);
            DTCF_HIP(hipMemcpyAsync(pin[b], d + off, len, hipMemcpyDeviceToHost, s));
            DTCF_HIP(hipEventRecord(ev[b], s));
            if (c >= 1) {  // drain the previous chunk while this one moves
                DTCF_HIP(hipEventSynchronize(ev[b ^ 1]));
                std::memcpy(h + prev_off, pin[b ^ 1], prev_len);
            }
            prev_off = off;
            prev_len = len;
        }
        if (c >= 1) {
            const int b = (int)((c - 1) & 1);
            DTCF_HIP(hipEventSynchronize(ev[b]));
            std::memcpy(h + prev_off, pin[b], prev_len);
        }
        return DTCF_OK;
    }
};

std::mutex g_stage_mu;
std::map<int, HostStage *> g_stage;

int host_stage(HostStage **out)
{
    int dev = 0;
    DTCF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_stage_mu);
    auto &p = g_stage[dev];
    if (!p) p = new HostStage();  // lives until the process ends
    *out = p;
    return DTCF_OK;
}
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
    HostStage *hs = nullptr;
    int st = host_stage(&hs);
    if (st != DTCF_OK) return st;
    std::lock_guard<std::mutex> lk(hs->mu);
    void *da, *db, *dc;
    if ((st = hs->init()) || (st = hs->buf(0, asz * sizeof(double), &da)) ||
        (st = hs->buf(1, bsz * sizeof(double), &db)) || (st = hs->buf(2, csz * sizeof(double), &dc)))
        return st;
    if ((st = hs->h2d(da, A, asz * sizeof(double))) || (st = hs->h2d(db, B, bsz * sizeof(double))) ||
        (st = hs->h2d(dc, C, csz * sizeof(double))))
        return st;
    st = dtcf_gemm_f64(transa, transb, m, n, k, alpha, (const double *)da, lda, (const double *)db,
                       ldb, (double *)dc, ldc, hs->s);
    if (st != DTCF_OK) return st;
    return hs->d2h(C, dc, csz * sizeof(double));
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
    HostStage *hs = nullptr;
    int st = host_stage(&hs);
    if (st != DTCF_OK) return st;
    std::lock_guard<std::mutex> lk(hs->mu);
    void *di, *dout;
    if ((st = hs->init()) || (st = hs->buf(0, total * sizeof(double), &di)) ||
        (st = hs->buf(1, total * sizeof(double), &dout)))
        return st;
    if ((st = hs->h2d(di, in, total * sizeof(double)))) return st;
    st = dtcf_permute_tiles_f64((const double *)di, (double *)dout, dims, w, out_src, ntiles, hs->s);
    if (st != DTCF_OK) return st;
    return hs->d2h(out, dout, total * sizeof(double));
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
    HostStage *hs = nullptr;
    int st = host_stage(&hs);
    if (st != DTCF_OK) return st;
    std::lock_guard<std::mutex> lk(hs->mu);
    void *dt, *dbig;
    if ((st = hs->init()) || (st = hs->buf(0, ntiles * bs * sizeof(double), &dt)) ||
        (st = hs->buf(1, rows * cols * sizeof(double), &dbig)))
        return st;
    dtcf_view v = nullptr;
    st = dtcf_view_create(rows, cols, rofs.data(), cofs.data(), blk.data(), rpb, cpb, &v);
    if (st != DTCF_OK) return st;
    if (unpack) st = hs->h2d(dbig, big, rows * cols * sizeof(double));
    // unpack: tiles outside the order keep their values
    if (st == DTCF_OK) st = hs->h2d(dt, tiles, ntiles * bs * sizeof(double));
    if (st == DTCF_OK)
        st = unpack ? dtcf_view_scatter_f64((const double *)dbig, cols, v, (double *)dt, hs->s)
                    : dtcf_view_gather_f64((const double *)dt, v, (double *)dbig, cols, hs->s);
    if (st == DTCF_OK)
        st = unpack ? hs->d2h(tiles, dt, ntiles * bs * sizeof(double))
                    : hs->d2h(big, dbig, rows * cols * sizeof(double));
    // the view's tables are read by the kernel: release them once it is done
    const hipError_t e = hipStreamSynchronize(hs->s);
    dtcf_view_destroy(v);
    if (st != DTCF_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "pack: stream");
    return DTCF_OK;
}
