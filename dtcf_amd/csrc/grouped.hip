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
    const BOp *ops;          // C ops (and A / B ops of the unpacked path)
    const BOp *pops;         // A / B ops as packed [rows][Kp] blocks
    const int32_t *pa, *pb;  // per pair: A op, B op
    const Work *work;
    int64_t nwork;
    double alpha;
    const double *A, *B;     // operand buffers (packed copies in the packed path)
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
typedef double d2 __attribute__((ext_vector_type(2)));

// PACKED: the block was copied to a dense [rows][Kp] array (Kp a multiple of
// BK, zero padded), so the pair (row, k0), (row, k0+1) is one aligned
// 16-byte load with no K bounds check.
template <int ROWS, bool PACKED>
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
            if constexpr (PACKED) rofs[i] = rok[i] ? op.base + r * op.c.s[0] : 0;
            else rofs[i] = rok[i] ? op.base + axis_off(op.r, (uint32_t)r) : 0;
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
            if constexpr (PACKED) {
                const d2 v = rok[i] ? *reinterpret_cast<const d2 *>(buf + rofs[i] + c0) : (d2){0.0, 0.0};
                x[i] = v.x;
                y[i] = v.y;
            } else {
                x[i] = (rok[i] && c0 < op.cols) ? buf[rofs[i] + axis_off(op.c, (uint32_t)c0)] : 0.0;
                y[i] = (rok[i] && c1 < op.cols) ? buf[rofs[i] + axis_off(op.c, (uint32_t)c1)] : 0.0;
            }
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

template <int BM, int BN, int WM, int WN, bool PACKED>
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
        const BOp &cop = g.ops[wk.c_op];  // read in place: its axes are indexed at run time
        const int64_t m0 = (int64_t)wk.tm * BM, n0 = (int64_t)wk.tn * BN;
        d4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

        // The pairs of the C block form one continuous K loop: the first K
        // step of the next pair is fetched during the last step of the
        // current one, so there is no pipeline restart between pairs.
        const BOp *opsAB = PACKED ? g.pops : g.ops;
        int p = wk.p0;
        // op indices, not copies: the descriptors are read where used, so
        // the pair switch needs no per-lane storage
        int ia = g.pa[p], ib = g.pb[p];
        BlockLoader<BM, PACKED> la;
        BlockLoader<BN, PACKED> lb;
        la.rows_for(opsAB[ia], m0, wave, lane);
        lb.rows_for(opsAB[ib], n0, wave, lane);
        int64_t nk = (opsAB[ia].cols + BK - 1) / BK;  // packed: Kp / BK, no tail
        la.load(opsAB[ia], g.A, 0, wave, lane);
        lb.load(opsAB[ib], g.B, 0, wave, lane);
        __syncthreads();  // the previous work item is done with LDS
        la.store(smem, wave, lane);
        lb.store(smem + LDS_A, wave, lane);
        __syncthreads();
        int64_t kt = 0;
        int cur = 0;
        for (;;) {
            const bool last = kt + 1 == nk;
            const bool more = !last || p + 1 < wk.p1;
            if (more) {
                if (last) {
                    const int na = g.pa[p + 1], nb = g.pb[p + 1];
                    la.rows_for(opsAB[na], m0, wave, lane);
                    lb.rows_for(opsAB[nb], n0, wave, lane);
                    la.load(opsAB[na], g.A, 0, wave, lane);
                    lb.load(opsAB[nb], g.B, 0, wave, lane);
                } else {
                    la.load(opsAB[ia], g.A, (kt + 1) * BK, wave, lane);
                    lb.load(opsAB[ib], g.B, (kt + 1) * BK, wave, lane);
                }
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
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                if (s == 1 && more) {
                    la.store(smem + (cur ^ 1) * STAGE, wave, lane);
                    lb.store(smem + (cur ^ 1) * STAGE + LDS_A, wave, lane);
                }
            }
            __syncthreads();
            if (!more) break;
            if (last) {
                p++;
                ia = g.pa[p];
                ib = g.pb[p];
                nk = (opsAB[ia].cols + BK - 1) / BK;
                kt = 0;
            } else {
                kt++;
            }
            cur ^= 1;
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

// Packing: every A and B block of the term copied once into a dense
// [rows][Kp] array (Kp = cols rounded up to BK, padding zeroed).  Work item
// = 8 rows of one block; threads run along the packed row, so the writes are
// coalesced and the reads are wherever the block's column digits point.
struct PackWork {
    int32_t op, r0;
};

struct MoveWork {
    int32_t op, r0, c0;  // c0: the tile's first column (transposed moves only)
};

__global__ __launch_bounds__(256) void pack_kernel(const BOp *ops, const BOp *pops,
                                                   const PackWork *work, int64_t nwork,
                                                   const double *__restrict__ src,
                                                   double *__restrict__ dst)
{
    for (int64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const PackWork pw = work[w];
        const BOp o = ops[pw.op], q = pops[pw.op];
        const int64_t kp = q.c.s[0];
        const int r1 = (int)min<int64_t>(pw.r0 + 8, o.rows);
        for (int r = pw.r0; r < r1; r++) {
            const int64_t roff = o.base + axis_off(o.r, (uint32_t)r);
            double *d = dst + q.base + (int64_t)r * kp;
            for (int64_t c = threadIdx.x; c < kp; c += 256)
                d[c] = c < o.cols ? src[roff + axis_off(o.c, (uint32_t)c)] : 0.0;
        }
    }
}

// Block moves between block-sparse storage and dense matrices: move m takes
// block ops[m] (read through its descriptors) to dense[place.base + r * pitch
// + c] (gather), or adds alpha times that region into the block (scatter;
// blocks are disjoint, so no two moves write one element).  The dense side
// is contiguous along c.  A block whose stride-1 digit lies on its row axis
// instead (place.trans) moves as 32 x 64 tiles transposed through LDS, so
// that both sides are read and written along their contiguous runs (lanes
// along c there read one element per cache line: 0.8-1.2 TB/s).  The others
// run lanes along c in chunks 2^cwsh columns wide (64, 128 or 256, whichever
// pads the block's columns least), a work item being a band of
// 8 x 256 / 2^cwsh rows over all the columns, eight loads per thread.
struct Place {
    int64_t base, pitch;
    int32_t trans, cwsh;
};

template <bool SCATTER>
__global__ __launch_bounds__(256) void blockmove_kernel(const BOp *ops, const Place *places,
                                                        const MoveWork *work, int64_t nwork,
                                                        double alpha, const double *__restrict__ in,
                                                        double *__restrict__ out)
{
    __shared__ double t[64][33];  // [c - c0][r - r0]
    const int lo = threadIdx.x & 31, hi = threadIdx.x >> 5;
    for (int64_t w = blockIdx.x; w < nwork; w += gridDim.x) {
        const MoveWork pw = work[w];
        const BOp &o = ops[pw.op];
        const Place pl = places[pw.op];
        if (!pl.trans) {  // lanes along c on both sides
            const int cw = 1 << pl.cwsh, rr = threadIdx.x >> pl.cwsh, rp = 256 >> pl.cwsh;
            int64_t roff[8];
            bool ok[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int r = pw.r0 + rr + rp * i;
                ok[i] = r < o.rows;
                roff[i] = ok[i] ? o.base + axis_off(o.r, (uint32_t)r) : 0;
            }
            for (int c = (threadIdx.x & (cw - 1)); c < o.cols; c += cw) {
                const int64_t coff = axis_off(o.c, (uint32_t)c);
                double v[8], u[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {  // every load issued before the first store
                    const int64_t dd = pl.base + (int64_t)(pw.r0 + rr + rp * i) * pl.pitch + c;
                    if constexpr (SCATTER) {
                        v[i] = ok[i] ? in[dd] : 0.0;
                        u[i] = ok[i] ? out[roff[i] + coff] : 0.0;
                    } else {
                        v[i] = ok[i] ? in[roff[i] + coff] : 0.0;
                    }
                }
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    if (!ok[i]) break;
                    const int64_t dd = pl.base + (int64_t)(pw.r0 + rr + rp * i) * pl.pitch + c;
                    if constexpr (SCATTER) out[roff[i] + coff] = u[i] + alpha * v[i];
                    else out[dd] = v[i];
                }
            }
            continue;
        }
        // one 32 x 64 tile (rows x columns) through LDS: the block side with
        // 32 lanes along r, the dense side with 64 lanes along c (a work item
        // per tile; a band of tiles per item, one barrier pair each, ran the
        // B gathers 1.2-1.8x slower)
        const int rl = pw.r0 + lo, cl = threadIdx.x & 63, rq = threadIdx.x >> 6;
        const int64_t roff = rl < o.rows ? o.base + axis_off(o.r, (uint32_t)rl) : 0;
        const int c0 = pw.c0;
        if constexpr (!SCATTER) {
            if (rl < o.rows) {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int c = c0 + hi + 8 * i;
                    if (c < o.cols) t[hi + 8 * i][lo] = in[roff + axis_off(o.c, (uint32_t)c)];
                }
            }
            __syncthreads();
            const int c = c0 + cl;
            if (c < o.cols) {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int r = pw.r0 + rq + 4 * i;
                    if (r < o.rows) out[pl.base + (int64_t)r * pl.pitch + c] = t[cl][rq + 4 * i];
                }
            }
        } else {
            const int c = c0 + cl;
            if (c < o.cols) {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int r = pw.r0 + rq + 4 * i;
                    if (r < o.rows) t[cl][rq + 4 * i] = in[pl.base + (int64_t)r * pl.pitch + c];
                }
            }
            __syncthreads();
            if (rl < o.rows) {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int c2 = c0 + hi + 8 * i;
                    if (c2 < o.cols)
                        out[roff + axis_off(o.c, (uint32_t)c2)] += alpha * t[hi + 8 * i][lo];
                }
            }
        }
        __syncthreads();  // the tile is read before the next one is written
    }
}

// Digits i-1 and i of an axis merge when stride[i-1] == ext[i] * stride[i]
// (the pair is one run): fewer divisions per element, same addresses.
dtcf_axis merge_digits(const dtcf_axis &a)
{
    dtcf_axis m = a;
    m.nd = 1;
    m.ext[0] = a.ext[0];
    m.stride[0] = a.stride[0];
    for (int i = 1; i < a.nd; i++) {
        if (m.stride[m.nd - 1] == a.ext[i] * a.stride[i] &&
            (int64_t)m.ext[m.nd - 1] * a.ext[i] < (int64_t(1) << 31)) {
            m.ext[m.nd - 1] *= a.ext[i];
            m.stride[m.nd - 1] = a.stride[i];
        } else {
            m.ext[m.nd] = a.ext[i];
            m.stride[m.nd] = a.stride[i];
            m.nd++;
        }
    }
    return m;
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
    bool packed = false;
    BOp *ops = nullptr, *pops = nullptr;
    double *packA = nullptr, *packB = nullptr;
    PackWork *wA = nullptr, *wB = nullptr;
    int64_t nwA = 0, nwB = 0;
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
    for (void *p : {(void *)g->ops, (void *)g->pa, (void *)g->pb, (void *)g->work, (void *)g->pops,
                    (void *)g->packA, (void *)g->packB, (void *)g->wA, (void *)g->wB})
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
    DTCF_REQUIRE(cfg >= 0 && cfg < 4,
                 "grouped: configuration 0 (64x64), 1 (128x128), 2 (packed 64x64) or 3 (packed 128x128)");
    const int BM = (cfg & 1) == 0 ? 64 : 128;
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
    // Packed configurations: a dense [rows][Kp] slot per A op and per B op.
    g->packed = cfg >= 2;
    std::vector<BOp> hpops;
    std::vector<PackWork> hwA, hwB;
    int64_t lenA = 0, lenB = 0;
    if (g->packed) {
        hpops = hops;
        std::vector<char> isA(nops, 0), isB(nops, 0);
        for (int64_t j = 0; j < npairs; j++) {
            isA[hpa[j]] = 1;
            isB[hpb[j]] = 1;
        }
        for (int64_t i = 0; i < nops; i++) {
            if (!isA[i] && !isB[i]) continue;
            DTCF_REQUIRE(!(isA[i] && isB[i]), "grouped: an op used both as A and as B");
            const int64_t kp = (hops[i].cols + BK - 1) / BK * BK;
            int64_t &len = isA[i] ? lenA : lenB;
            BOp &q = hpops[i];
            q.base = len;
            q.r = Axis{};
            q.r.nd = 1;
            q.r.e[0] = (uint32_t)hops[i].rows;
            q.r.d[0] = FastDiv((uint32_t)hops[i].rows);
            q.r.s[0] = kp;
            q.c = Axis{};
            q.c.nd = 1;
            q.c.e[0] = (uint32_t)kp;
            q.c.d[0] = FastDiv((uint32_t)kp);
            q.c.s[0] = kp;  // the row pitch, read by the loader and the pack kernel
            q.cols = kp;
            len += hops[i].rows * kp;
            for (int64_t r = 0; r < hops[i].rows; r += 8)
                (isA[i] ? hwA : hwB).push_back(PackWork{(int32_t)i, (int32_t)r});
        }
        g->nwA = (int64_t)hwA.size();
        g->nwB = (int64_t)hwB.size();
    }
    int dev = 0, cus = 1;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int per_cu = (cfg & 1) == 0 ? 4 : 2;
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
    if (g->packed) {
        if (e == hipSuccess && nops) e = hipMalloc(&g->pops, nops * sizeof(BOp));
        if (e == hipSuccess && nops)
            e = hipMemcpy(g->pops, hpops.data(), nops * sizeof(BOp), hipMemcpyHostToDevice);
        if (e == hipSuccess && lenA) e = hipMalloc(&g->packA, lenA * sizeof(double));
        if (e == hipSuccess && lenB) e = hipMalloc(&g->packB, lenB * sizeof(double));
        if (e == hipSuccess && g->nwA) e = hipMalloc(&g->wA, g->nwA * sizeof(PackWork));
        if (e == hipSuccess && g->nwB) e = hipMalloc(&g->wB, g->nwB * sizeof(PackWork));
        if (e == hipSuccess && g->nwA)
            e = hipMemcpy(g->wA, hwA.data(), g->nwA * sizeof(PackWork), hipMemcpyHostToDevice);
        if (e == hipSuccess && g->nwB)
            e = hipMemcpy(g->wB, hwB.data(), g->nwB * sizeof(PackWork), hipMemcpyHostToDevice);
    }
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
    GArgs a{g->ops, g->pops, g->pa, g->pb, g->work, g->nwork, alpha, A, B, C};
    hipStream_t s = as_stream(stream);
    if (g->packed) {
        // the packed copies belong to the plan: runs of one plan on several
        // streams at once would race on them
        if (g->nwA)
            hipLaunchKernelGGL(pack_kernel, dim3((unsigned)std::min<int64_t>(g->nwA, 65536)), dim3(256), 0,
                               s, g->ops, g->pops, g->wA, g->nwA, A, g->packA);
        if (g->nwB)
            hipLaunchKernelGGL(pack_kernel, dim3((unsigned)std::min<int64_t>(g->nwB, 65536)), dim3(256), 0,
                               s, g->ops, g->pops, g->wB, g->nwB, B, g->packB);
        DTCF_HIP(hipGetLastError());
        a.A = g->packA;
        a.B = g->packB;
    }
    switch (g->cfg) {
    case 0: hipLaunchKernelGGL((grouped_kernel<64, 64, 32, 32, false>), dim3(g->grid), dim3(NT), 0, s, a); break;
    case 1: hipLaunchKernelGGL((grouped_kernel<128, 128, 64, 64, false>), dim3(g->grid), dim3(NT), 0, s, a); break;
    case 2: hipLaunchKernelGGL((grouped_kernel<64, 64, 32, 32, true>), dim3(g->grid), dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((grouped_kernel<128, 128, 64, 64, true>), dim3(g->grid), dim3(NT), 0, s, a); break;
    }
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

// ------------------------------------------------------------- block moves
struct dtcf_blockmove_s {
    BOp *ops = nullptr;
    Place *places = nullptr;
    MoveWork *work = nullptr;
    int64_t n = 0, nwork = 0, dense_span = 0;
};

extern "C" int dtcf_blockmove_destroy(dtcf_blockmove b)
{
    if (!b) return DTCF_OK;
    for (void *p : {(void *)b->ops, (void *)b->places, (void *)b->work})
        if (p) (void)hipFree(p);
    delete b;
    return DTCF_OK;
}

extern "C" int dtcf_blockmove_plan(const dtcf_block_op *ops, const int64_t *dense_base,
                                   const int64_t *dense_pitch, int64_t n, dtcf_blockmove *out)
{
    DTCF_REQUIRE(out, "blockmove: null output");
    *out = nullptr;
    DTCF_REQUIRE(n >= 0 && (n == 0 || (ops && dense_base && dense_pitch)), "blockmove: bad lists");
    std::vector<BOp> hops(n);
    std::vector<Place> hpl(n);
    std::vector<MoveWork> hw;
    int64_t span = 0;
    for (int64_t i = 0; i < n; i++) {
        const dtcf_block_op &o = ops[i];
        DTCF_REQUIRE(o.rows.nd >= 1 && o.rows.nd <= 3 && o.cols.nd >= 1 && o.cols.nd <= 3,
                     "blockmove: 1..3 digits per axis");
        for (int d = 0; d < o.rows.nd; d++) DTCF_REQUIRE(o.rows.ext[d] >= 1, "blockmove: extent < 1");
        for (int d = 0; d < o.cols.nd; d++) DTCF_REQUIRE(o.cols.ext[d] >= 1, "blockmove: extent < 1");
        DTCF_REQUIRE(o.base >= 0 && dense_base[i] >= 0, "blockmove: negative base");
        hops[i].base = o.base;
        const dtcf_axis ra = merge_digits(o.rows), ca = merge_digits(o.cols);
        hops[i].r = make_axis(ra);
        hops[i].c = make_axis(ca);
        hops[i].rows = axis_len(o.rows);
        hops[i].cols = axis_len(o.cols);
        DTCF_REQUIRE(hops[i].rows < (1ll << 31) && hops[i].cols < (1ll << 31), "blockmove: block too long");
        DTCF_REQUIRE(dense_pitch[i] >= hops[i].cols, "blockmove: pitch shorter than the block's columns");
        // transposed tiles where the block runs along its rows, not its columns
        const bool cfast = ca.stride[ca.nd - 1] == 1, rfast = ra.stride[ra.nd - 1] == 1;
        const bool trans = !cfast && rfast;
        int cwsh = 8;  // widest tile whose column padding is within 10%, else the least padded
        for (int sh = 8; sh >= 6; sh--) {
            const int64_t w = int64_t(1) << sh, pad = (hops[i].cols + w - 1) / w * w;
            const int64_t best = (hops[i].cols + (int64_t(1) << cwsh) - 1) >> cwsh << cwsh;
            if (pad * 10 <= hops[i].cols * 11) { cwsh = sh; break; }
            if (pad < best) cwsh = sh;
        }
        hpl[i] = Place{dense_base[i], dense_pitch[i], trans ? 1 : 0, cwsh};
        span = std::max(span, dense_base[i] + (hops[i].rows - 1) * dense_pitch[i] + hops[i].cols);
        // work items: 32 x 64 tiles (transposed), else bands of 8 x 256 / 2^cwsh rows
        const int64_t tr = trans ? 32 : 8 * (256 >> cwsh), tc = trans ? 64 : hops[i].cols;
        for (int64_t r = 0; r < hops[i].rows; r += tr)
            for (int64_t c = 0; c < hops[i].cols; c += tc)
                hw.push_back(MoveWork{(int32_t)i, (int32_t)r, (int32_t)c});
    }
    auto *b = new dtcf_blockmove_s();
    b->n = n;
    b->nwork = (int64_t)hw.size();
    b->dense_span = span;
    hipError_t e = hipSuccess;
    if (n) e = hipMalloc(&b->ops, n * sizeof(BOp));
    if (e == hipSuccess && n) e = hipMalloc(&b->places, n * sizeof(Place));
    if (e == hipSuccess && b->nwork) e = hipMalloc(&b->work, b->nwork * sizeof(MoveWork));
    if (e == hipSuccess && n) e = hipMemcpy(b->ops, hops.data(), n * sizeof(BOp), hipMemcpyHostToDevice);
    if (e == hipSuccess && n) e = hipMemcpy(b->places, hpl.data(), n * sizeof(Place), hipMemcpyHostToDevice);
    if (e == hipSuccess && b->nwork)
        e = hipMemcpy(b->work, hw.data(), b->nwork * sizeof(MoveWork), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dtcf_blockmove_destroy(b);
        return hip_fail(e, "blockmove: plan upload");
    }
    *out = b;
    return DTCF_OK;
}

extern "C" int dtcf_blockmove_info(dtcf_blockmove b, int64_t *dense_span)
{
    DTCF_REQUIRE(b, "blockmove: null plan");
    if (dense_span) *dense_span = b->dense_span;
    return DTCF_OK;
}

extern "C" int dtcf_blockmove_gather_f64(dtcf_blockmove b, const double *sparse, double *dense,
                                         void *stream)
{
    DTCF_REQUIRE(b, "blockmove: null plan");
    if (b->nwork == 0) return DTCF_OK;
    DTCF_REQUIRE(sparse && dense, "blockmove: null buffer");
    hipLaunchKernelGGL(blockmove_kernel<false>, dim3((unsigned)std::min<int64_t>(b->nwork, 65536)),
                       dim3(256), 0, as_stream(stream), b->ops, b->places, b->work, b->nwork, 0.0,
                       sparse, dense);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

extern "C" int dtcf_blockmove_scatter_f64(dtcf_blockmove b, double alpha, const double *dense,
                                          double *sparse, void *stream)
{
    DTCF_REQUIRE(b, "blockmove: null plan");
    if (b->nwork == 0) return DTCF_OK;
    DTCF_REQUIRE(sparse && dense, "blockmove: null buffer");
    hipLaunchKernelGGL(blockmove_kernel<true>, dim3((unsigned)std::min<int64_t>(b->nwork, 65536)),
                       dim3(256), 0, as_stream(stream), b->ops, b->places, b->work, b->nwork, alpha,
                       dense, sparse);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}
