// Grouped FP64 MFMA contraction over uneven block pairs (one persistent
// launch per term).
//
// A block-sparse term is a list of C blocks, each fed by a list of (A block,
// B block) pairs.  Every operand block is a dense sub-array read through an
// affine descriptor with up to three digits per GEMM axis, so the index
// permutation of each block (the reference permutes every tile before its
// dgemm, contraction.cc:1824-1914, DLTC/task.cc:115-131) is folded into the
// operand loads.  A work item is one BM x BN tile of one C block; the
// workgroup that owns it walks all of the block's pairs and their K,
// accumulating in registers, and adds alpha * acc to C once, so results are
// deterministic (no atomics) and identical whatever else runs.  Work items
// are sorted by cost on the host and handed out round-robin over a
// persistent grid sized to the CU count.
//
// The LDS tile layout and MFMA use are those of gemm_f64.hip
// (fragment-major [row/16][s][lane], k = 4*fk + s within a K step).
#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace dtcf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BK = 16;
constexpr int NT = 256;

struct Axis {
    int32_t nd;            // 1..3 digits, digit 0 most significant
    uint32_t e[3];         // extents
    FastDiv d[3];          // division by e[i]
    int64_t s[3];          // element strides
};

struct BOp {
    int64_t base;          // element offset of the block
    Axis r, c;
    int64_t rows, cols;
};

struct Work {
    int32_t c_op, p0, p1, tm, tn, pad;
};

struct GArgs {
    const BOp *ops;
    const int32_t *pa, *pb;  // per pair: A op, B op
    const Work *work;
    int64_t nwork;
    double alpha;
    const double *A, *B;
    double *C;
};

__device__ inline int64_t axis_off(const Axis &a, uint32_t x)
{
    int64_t off = 0;
    for (int i = a.nd - 1; i > 0; i--) {
        const uint32_t q = a.d[i].div(x);
        off += (int64_t)(x - q * a.e[i]) * a.s[i];
        x = q;
    }
    return off + (int64_t)x * a.s[0];
}

// k-contiguous fragment-major loader (the kDK layout of gemm_f64.hip):
// pass i, wave w, lane l loads (row, k0) and (row, k0+1) with
// row = 16*g + l%16, g = (4i+w)/2, fk = (l/16 & 1) + 2*(w & 1), sp = l/32,
// k0 = 4*fk + 2*sp; LDS slots (g*4 + 2sp)*64 + fk*16 + r and +64.
template <int ROWS>
struct BlockLoader {
    static constexpr int NP = ROWS / 32;
    double x[NP], y[NP];
    int64_t rofs[NP];
    bool rok[NP];

    __device__ static inline void coords(int i, int w, int l, int &row, int &k0, int &slot)
    {
        const int g = (4 * i + w) >> 1;
        const int r = l & 15, fk = ((l >> 4) & 1) + 2 * (w & 1), sp = l >> 5;
        row = 16 * g + r;
        k0 = 4 * fk + 2 * sp;
        slot = (g * 4 + 2 * sp) * 64 + fk * 16 + r;
    }

    __device__ inline void rows_for(const BOp &op, int64_t row0, int w, int l)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, slot;
            coords(i, w, l, row, k0, slot);
            const int64_t r = row0 + row;
            rok[i] = r < op.rows;
            rofs[i] = rok[i] ? op.base + axis_off(op.r, (uint32_t)r) : 0;
        }
    }

    __device__ inline void load(const BOp &op, const double *__restrict__ buf, int64_t kb, int w,
                                int l)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, slot;
            coords(i, w, l, row, k0, slot);
            const int64_t c0 = kb + k0, c1 = c0 + 1;
            x[i] = (rok[i] && c0 < op.cols) ? buf[rofs[i] + axis_off(op.c, (uint32_t)c0)] : 0.0;
            y[i] = (rok[i] && c1 < op.cols) ? buf[rofs[i] + axis_off(op.c, (uint32_t)c1)] : 0.0;
        }
    }

    __device__ inline void store(double *lds, int w, int l) const
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, slot;
            coords(i, w, l, row, k0, slot);
            lds[slot] = x[i];
            lds[slot + 64] = y[i];
        }
    }
};

template <int BM, int BN>
struct GOcc {
    static constexpr int v = (BM * BN <= 64 * 64) ? 4 : 2;
};

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(NT, (GOcc<BM, BN>::v)) void grouped_kernel(GArgs g)
{
    static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDS_A = BM * BK, LDS_B = BN * BK, STAGE = LDS_A + LDS_B;
    __shared__ double smem[2 * STAGE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / (BN / WN), wn = wave % (BN / WN);
    const int gA = wm * (WM / 16), gB = wn * (WN / 16);
    const int fr = lane & 15, fq = lane >> 4;

    for (int64_t w = blockIdx.x; w < g.nwork; w += gridDim.x) {
        const Work wk = g.work[w];
        const BOp cop = g.ops[wk.c_op];
        const int64_t m0 = (int64_t)wk.tm * BM, n0 = (int64_t)wk.tn * BN;
        d4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

        for (int p = wk.p0; p < wk.p1; p++) {
            const BOp aop = g.ops[g.pa[p]];
            const BOp bop = g.ops[g.pb[p]];
            BlockLoader<BM> la;
            BlockLoader<BN> lb;
            la.rows_for(aop, m0, wave, lane);
            lb.rows_for(bop, n0, wave, lane);
            const int64_t nk = (aop.cols + BK - 1) / BK;
            la.load(aop, g.A, 0, wave, lane);
            lb.load(bop, g.B, 0, wave, lane);
            __syncthreads();  // the previous pair / work item is done with LDS
            la.store(smem, wave, lane);
            lb.store(smem + LDS_A, wave, lane);
            __syncthreads();
            for (int64_t kt = 0; kt < nk; kt++) {
                const int cur = (int)(kt & 1);
                const bool more = kt + 1 < nk;
                if (more) {
                    la.load(aop, g.A, (kt + 1) * BK, wave, lane);
                    lb.load(bop, g.B, (kt + 1) * BK, wave, lane);
                }
                const double *a_s = smem + cur * STAGE + gA * 256 + lane;
                const double *b_s = smem + cur * STAGE + LDS_A + gB * 256 + lane;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    double af[FM], bf[FN];
#pragma unroll
                    for (int i = 0; i < FM; i++) af[i] = a_s[i * 256 + s * 64];
#pragma unroll
                    for (int j = 0; j < FN; j++) bf[j] = b_s[j * 256 + s * 64];
#pragma unroll
                    for (int i = 0; i < FM; i++)
#pragma unroll
                        for (int j = 0; j < FN; j++)
                            acc[i][j] =
                                __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                    if (s == 1 && more) {
                        la.store(smem + (cur ^ 1) * STAGE, wave, lane);
                        lb.store(smem + (cur ^ 1) * STAGE + LDS_A, wave, lane);
                    }
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) {
                const int64_t col = n0 + wn * WN + j * 16 + fr;
                if (col >= cop.cols) continue;
                const int64_t coff = axis_off(cop.c, (uint32_t)col);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t row = m0 + wm * WM + i * 16 + fq + 4 * r;
                    if (row >= cop.rows) continue;
                    g.C[cop.base + axis_off(cop.r, (uint32_t)row) + coff] += g.alpha * acc[i][j][r];
                }
            }
    }
}

Axis make_axis(const dtcf_axis &a)
{
    Axis x{};
    x.nd = a.nd;
    for (int i = 0; i < a.nd; i++) {
        x.e[i] = (uint32_t)a.ext[i];
        x.d[i] = FastDiv((uint32_t)a.ext[i]);
        x.s[i] = a.stride[i];
    }
    return x;
}

int64_t axis_len(const dtcf_axis &a)
{
    int64_t n = 1;
    for (int i = 0; i < a.nd; i++) n *= a.ext[i];
    return n;
}

int64_t axis_max(const dtcf_axis &a)
{
    int64_t m = 0;
    for (int i = 0; i < a.nd; i++) m += (int64_t)(a.ext[i] - 1) * a.stride[i];
    return m;
}

}  // namespace
}  // namespace dtcf

using namespace dtcf;

struct dtcf_grouped_s {
    int cfg = 0;
    BOp *ops = nullptr;
    int32_t *pa = nullptr, *pb = nullptr;
    Work *work = nullptr;
    int64_t nwork = 0, nops = 0, npairs = 0;
    double flops = 0.0;
    int64_t span_a = 0, span_b = 0, span_c = 0;
    int grid = 0;
};

extern "C" int dtcf_grouped_destroy(dtcf_grouped g)
{
    if (!g) return DTCF_OK;
    for (void *p : {(void *)g->ops, (void *)g->pa, (void *)g->pb, (void *)g->work})
        if (p) (void)hipFree(p);
    delete g;
    return DTCF_OK;
}

extern "C" int dtcf_grouped_plan(const dtcf_block_op *ops, int64_t nops,
                                 const dtcf_block_pair *pairs, int64_t npairs, int cfg,
                                 dtcf_grouped *out)
{
    DTCF_REQUIRE(out, "grouped: null output");
    *out = nullptr;
    DTCF_REQUIRE(nops >= 0 && npairs >= 0 && (nops == 0 || ops) && (npairs == 0 || pairs),
                 "grouped: bad op or pair list");
    DTCF_REQUIRE(cfg >= 0 && cfg < 2, "grouped: tile configuration must be 0 (64x64) or 1 (128x128)");
    const int BM = cfg == 0 ? 64 : 128;
    std::vector<BOp> hops(nops);
    for (int64_t i = 0; i < nops; i++) {
        const dtcf_block_op &o = ops[i];
        DTCF_REQUIRE(o.rows.nd >= 1 && o.rows.nd <= 3 && o.cols.nd >= 1 && o.cols.nd <= 3,
                     "grouped: 1..3 digits per axis");
        for (int d = 0; d < o.rows.nd; d++) DTCF_REQUIRE(o.rows.ext[d] >= 1, "grouped: extent < 1");
        for (int d = 0; d < o.cols.nd; d++) DTCF_REQUIRE(o.cols.ext[d] >= 1, "grouped: extent < 1");
        DTCF_REQUIRE(o.base >= 0, "grouped: negative base");
        hops[i].base = o.base;
        hops[i].r = make_axis(o.rows);
        hops[i].c = make_axis(o.cols);
        hops[i].rows = axis_len(o.rows);
        hops[i].cols = axis_len(o.cols);
        DTCF_REQUIRE(hops[i].rows < (1ll << 31) && hops[i].cols < (1ll << 31),
                     "grouped: block axis too long");
    }
    // Group pairs by C op, keeping the caller's order inside a C op (that is
    // the summation order of the C block).
    std::vector<int64_t> idx(npairs);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int64_t x, int64_t y) { return pairs[x].c_op < pairs[y].c_op; });
    std::vector<int32_t> hpa(npairs), hpb(npairs);
    std::vector<Work> hwork;
    std::vector<double> cost;
    dtcf_grouped_s *g = new dtcf_grouped_s();
    g->cfg = cfg;
    for (int64_t i = 0; i < npairs;) {
        const int32_t c = pairs[idx[i]].c_op;
        if (c < 0 || c >= nops) {
            dtcf_grouped_destroy(g);
            set_error("grouped: pair refers to a missing C op");
            return DTCF_ERR_INVALID;
        }
        int64_t j = i;
        double kc = 0;
        for (; j < npairs && pairs[idx[j]].c_op == c; j++) {
            const dtcf_block_pair &p = pairs[idx[j]];
            if (p.a_op < 0 || p.a_op >= nops || p.b_op < 0 || p.b_op >= nops ||
                hops[p.a_op].rows != hops[c].rows || hops[p.b_op].rows != hops[c].cols ||
                hops[p.a_op].cols != hops[p.b_op].cols) {
                dtcf_grouped_destroy(g);
                set_error("grouped: pair shapes do not match (A rows = C rows, B rows = C cols, "
                          "A cols = B cols)");
                return DTCF_ERR_INVALID;
            }
            hpa[j] = p.a_op;
            hpb[j] = p.b_op;
            kc += (double)hops[p.a_op].cols;
            g->flops += 2.0 * hops[c].rows * hops[c].cols * hops[p.a_op].cols;
            g->span_a = std::max(g->span_a, ops[p.a_op].base + axis_max(ops[p.a_op].rows) +
                                                axis_max(ops[p.a_op].cols) + 1);
            g->span_b = std::max(g->span_b, ops[p.b_op].base + axis_max(ops[p.b_op].rows) +
                                                axis_max(ops[p.b_op].cols) + 1);
        }
        g->span_c = std::max(g->span_c, ops[c].base + axis_max(ops[c].rows) + axis_max(ops[c].cols) + 1);
        const int64_t tm = (hops[c].rows + BM - 1) / BM, tn = (hops[c].cols + BM - 1) / BM;
        for (int64_t a = 0; a < tm; a++)
            for (int64_t b = 0; b < tn; b++) {
                hwork.push_back(Work{c, (int32_t)i, (int32_t)j, (int32_t)a, (int32_t)b, 0});
                cost.push_back(kc);
            }
        i = j;
    }
    // Longest work items first: with round-robin hand-out over a persistent
    // grid this evens out the per-workgroup totals (LPT order).
    std::vector<int64_t> ord(hwork.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return cost[x] > cost[y]; });
    std::vector<Work> sorted(hwork.size());
    for (size_t i = 0; i < ord.size(); i++) sorted[i] = hwork[ord[i]];
    g->nwork = (int64_t)sorted.size();
    g->nops = nops;
    g->npairs = npairs;
    int dev = 0, cus = 1;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int per_cu = cfg == 0 ? 4 : 2;
    g->grid = (int)std::min<int64_t>(std::max<int64_t>(g->nwork, 1), (int64_t)cus * per_cu);
    if (e == hipSuccess && nops) e = hipMalloc(&g->ops, nops * sizeof(BOp));
    if (e == hipSuccess && npairs) e = hipMalloc(&g->pa, npairs * sizeof(int32_t));
    if (e == hipSuccess && npairs) e = hipMalloc(&g->pb, npairs * sizeof(int32_t));
    if (e == hipSuccess && g->nwork) e = hipMalloc(&g->work, g->nwork * sizeof(Work));
    if (e == hipSuccess && nops) e = hipMemcpy(g->ops, hops.data(), nops * sizeof(BOp), hipMemcpyHostToDevice);
    if (e == hipSuccess && npairs) e = hipMemcpy(g->pa, hpa.data(), npairs * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && npairs) e = hipMemcpy(g->pb, hpb.data(), npairs * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess && g->nwork)
        e = hipMemcpy(g->work, sorted.data(), g->nwork * sizeof(Work), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dtcf_grouped_destroy(g);
        return hip_fail(e, "grouped: plan upload");
    }
    *out = g;
    return DTCF_OK;
}

extern "C" int dtcf_grouped_info(dtcf_grouped g, double *flops, int64_t *nwork, int64_t *spans)
{
    DTCF_REQUIRE(g, "grouped: null plan");
    if (flops) *flops = g->flops;
    if (nwork) *nwork = g->nwork;
    if (spans) {
        spans[0] = g->span_a;
        spans[1] = g->span_b;
        spans[2] = g->span_c;
    }
    return DTCF_OK;
}

extern "C" int dtcf_grouped_run(dtcf_grouped g, double alpha, const double *A, const double *B,
                                double *C, void *stream)
{
    DTCF_REQUIRE(g, "grouped: null plan");
    if (g->nwork == 0) return DTCF_OK;
    DTCF_REQUIRE(A && B && C, "grouped: null buffer");
    GArgs a{g->ops, g->pa, g->pb, g->work, g->nwork, alpha, A, B, C};
    hipStream_t s = as_stream(stream);
    if (g->cfg == 0)
        hipLaunchKernelGGL((grouped_kernel<64, 64, 32, 32>), dim3(g->grid), dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL((grouped_kernel<128, 128, 64, 64>), dim3(g->grid), dim3(NT), 0, s, a);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}
