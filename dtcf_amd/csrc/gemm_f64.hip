// FP64 GEMM on the gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Replaces the dgemm behind local_contraction (localcontract.cc:3-21) and
// DLTC compute (DLTC/compute.cc:3-36).  One kernel template serves the plain
// row-major GEMM and the fused form where operands are read straight out of
// the permuted tile storage through dtcf_view offset tables (the
// create_big_matrix pack of transpose_helper.cc:20-171 is never
// materialised).
//
// Tiling: a 256-thread workgroup (4 wave64) owns a BM x BN block of C; the
// waves form a (BM/WM) x (BN/WN) grid and each wave owns WM x WN as
// (WM/16) x (WN/16) MFMA 16x16 accumulators (4 doubles per lane each).  K is
// walked in steps of BK=16 through a double-buffered LDS stage with register
// prefetch of the next step and one barrier per step.
//
// LDS layout ("fragment-major"): an operand tile of ROWS x 16 is stored as
// [row/16][s][lane] with lane = fk*16 + row%16 and k = fk*4 + s (s, fk in
// 0..3), i.e. inside a K step the MFMA k-slices are taken in the order
// {s, s+4, s+8, s+12}.  The same permutation is used for A and B, so the sum
// over k is unchanged, and every fragment read of a wave is one contiguous
// 512-byte span: bank-conflict free whatever lane grouping the LDS uses.  The
// global->LDS stores are arranged so each 32-lane half of a wave writes 256
// contiguous bytes.
//
// MFMA operand layout (16x16x4 f64, wave64): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; the result holds D[l/16 + 4*r][l%16] in element r (checked
// on the device by tests/test_gpu_core.py::test_mfma_layout_single_block).
//
// Workgroup ids are remapped so consecutive C tiles of a row group land on
// the same XCD (its own L2), then swizzled in groups of 8 tile rows.
//
// Shapes the tiles serve badly have streaming kernels of their own, chosen
// by the launch: matrix-vector products (gemv_kernel, gemv_rows_kernel) and
// thin products with one side at most 48 (192 with short K) and short K
// (thin_kernel).  Each is bound by streaming its long operand from HBM.
#include <algorithm>
#include <map>
#include <mutex>

#include "common.hpp"
#include "gemm_f64.hpp"

namespace dtcf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int BK = 16;
constexpr int NT = 256;

// Element offset of (r, c) in an operand view.
template <int MODE>
__device__ inline int64_t op_addr(const OpArgs &o, int64_t r, int64_t c)
{
    if constexpr (MODE == kDK) return r * o.ld + c;
    else if constexpr (MODE == kDMN) return c * o.ld + r;
    else {
        int64_t a = o.rofs[r] + o.cofs[c];
        if (o.blk) {
            const uint32_t rb = o.div_rpb.div((uint32_t)r);
            const uint32_t cb = o.div_cpb.div((uint32_t)c);
            a += o.blk[(int64_t)rb * o.ncb + cb];
        }
        return a;
    }
}

// Loads one ROWS x BK operand tile into registers (a pair of doubles per
// pass) and stores it to LDS in the fragment-major layout.
//
// k-contiguous operands (kDK, kVK): in pass i, wave w, lane l the pair is
//   row  = 16*g + l%16,  g = (4i + w) / 2
//   fk   = (l/16 & 1) + 2*(w & 1),  sp = l/32,  k = 4*fk + 2*sp + {0, 1}
// so each row contributes 64 contiguous bytes per pass and each 32-lane half
// of the wave writes 256 contiguous LDS bytes.
// row-contiguous operands (kDMN, kVMN): the pair is two adjacent rows
//   r = 2*(l%8), fk = (l/8)%4, s = l/32 + 2*((4i+w) & 1), g = (4i + w) / 2
// so 8 lanes read 128 contiguous bytes of one k row and the wave writes 1 KiB
// of contiguous LDS.
template <int ROWS, int MODE>
struct TileLoader {
    static constexpr bool KC = (MODE == kDK || MODE == kVK);
    static constexpr int NP = ROWS / 32;
    static_assert(ROWS % 32 == 0, "tile rows must be a multiple of 32");
    d2 reg[NP];

    __device__ static inline void coords(int i, int w, int l, int &row, int &k0, int &k1,
                                         int &lds0, int &lds1)
    {
        const int u = 4 * i + w, g = u >> 1;
        if constexpr (KC) {
            const int r = l & 15, fk = ((l >> 4) & 1) + 2 * (w & 1), sp = l >> 5;
            row = 16 * g + r;
            k0 = 4 * fk + 2 * sp;
            k1 = k0 + 1;
            lds0 = (g * 4 + 2 * sp) * 64 + fk * 16 + r;
            lds1 = lds0 + 64;
        } else {
            const int r = 2 * (l & 7), fk = (l >> 3) & 3, s = (l >> 5) + 2 * (u & 1);
            row = 16 * g + r;
            k0 = k1 = 4 * fk + s;
            lds0 = (g * 4 + s) * 64 + fk * 16 + r;
            lds1 = lds0 + 1;
        }
    }

    // Generic path with bounds checks (edges, K tails, offset tables).
    __device__ inline void load_checked(const OpArgs &o, const double *__restrict__ base,
                                        int64_t row0, int64_t kb, int64_t rows, int64_t K, int w,
                                        int l)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, k1, l0, l1;
            coords(i, w, l, row, k0, k1, l0, l1);
            d2 v = {0.0, 0.0};
            int64_t r0 = row0 + row, r1 = KC ? r0 : r0 + 1;
            int64_t c0 = kb + k0, c1 = kb + k1;
            const bool ok0 = r0 < rows && c0 < K;
            const bool ok1 = r1 < rows && c1 < K;
            if (ok0) {
                const int64_t a0 = op_addr<MODE>(o, r0, c0);
                if (ok1 && o.vec2) {
                    v = *reinterpret_cast<const d2 *>(base + a0);
                } else {
                    v.x = base[a0];
                    if (ok1) v.y = base[op_addr<MODE>(o, r1, c1)];
                }
            }
            reg[i] = v;
        }
    }

    __device__ inline void store(double *lds, int w, int l) const
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, k1, l0, l1;
            coords(i, w, l, row, k0, k1, l0, l1);
            if constexpr (KC) {
                lds[l0] = reg[i].x;
                lds[l1] = reg[i].y;
            } else {
                *reinterpret_cast<d2 *>(lds + l0) = reg[i];
            }
        }
    }
};

// Rows past the operand's last (an edge tile) load the last row (the last
// row pair for row-contiguous operands, whose row count is then even): in
// range, and never stored, since the epilogue writes only rows < m and
// columns < n.  Each C element still sums exactly its own row's and
// column's products, so an edge tile's results equal the checked path's.
template <int MODE>
__device__ inline int64_t clamp_row(int64_t r, int64_t rows)
{
    constexpr bool KC = (MODE == kDK || MODE == kVK);
    const int64_t last = KC ? rows - 1 : rows - 2;
    return r < last ? r : last;
}

// Fast path for dense operands: per-thread pointers advanced by BK
// (k-contiguous) or BK*ld (row-contiguous) each step, no bounds checks.
template <int ROWS, int MODE>
struct FastPtrs {
    static constexpr int NP = ROWS / 32;
    const double *p[NP];
    int64_t step;

    __device__ inline void init(const OpArgs &o, const double *base, int64_t row0, int64_t rows,
                                int w, int l)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, k1, l0, l1;
            TileLoader<ROWS, MODE>::coords(i, w, l, row, k0, k1, l0, l1);
            p[i] = base + op_addr<MODE>(o, clamp_row<MODE>(row0 + row, rows), k0);
        }
        step = (MODE == kDK) ? BK : BK * o.ld;
    }

    __device__ inline void load(TileLoader<ROWS, MODE> &t, int64_t kt) const
    {
#pragma unroll
        for (int i = 0; i < NP; i++)
            t.reg[i] = *reinterpret_cast<const d2 *>(p[i] + kt * step);
    }
};

// Fast path for view operands whose block table folded away:
// (r, c) at rofs[r] + cofs[c].  Row offsets are read once; the column
// offsets of the K step after next are fetched while the next step's data
// loads are in flight, so no data load waits on a table load.
template <int ROWS, int MODE>
struct ViewFast {
    static constexpr int NP = ROWS / 32;
    int64_t rb[NP], cn[NP];
    int kk[NP];

    __device__ inline void init(const OpArgs &o, int64_t row0, int64_t rows, int w, int l)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            int row, k0, k1, l0, l1;
            TileLoader<ROWS, MODE>::coords(i, w, l, row, k0, k1, l0, l1);
            rb[i] = o.rofs[clamp_row<MODE>(row0 + row, rows)];
            kk[i] = k0;
        }
    }
    __device__ inline void fetch_cols(const OpArgs &o, int64_t kb)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) cn[i] = o.cofs[kb + kk[i]];
    }
    __device__ inline void load(TileLoader<ROWS, MODE> &t, const double *base) const
    {
#pragma unroll
        for (int i = 0; i < NP; i++)
            t.reg[i] = *reinterpret_cast<const d2 *>(base + rb[i] + cn[i]);
    }
};

template <int BM, int BN>
struct Occ {
    // Two workgroups per CU when the accumulators leave room in 256 VGPRs.
    static constexpr int v = (BM * BN <= 192 * 96) ? 2 : 1;
};

// Logical tile q -> (tm, tn) through a grouped (8 tile rows) swizzle.
__device__ inline void tile_coords(uint32_t q, uint32_t tiles_m, uint32_t tiles_n, uint32_t &tm,
                                   uint32_t &tn)
{
    const uint32_t GROUP = 8;
    const uint32_t width = GROUP * tiles_n;
    const uint32_t gid = q / width;
    const uint32_t first = gid * GROUP;
    const uint32_t gsz = min(tiles_m - first, GROUP);
    tm = first + (q % width) % gsz;
    tn = (q % width) / gsz;
}

// One workgroup's share of a GEMM: logical tile q over K steps [kt0, kt1);
// a split slice (split) writes raw partial sums to workspace slot sidx.
template <int BM, int BN, int WM, int WN, int AMODE, int BMODE, bool CVIEW>
__device__ __forceinline__ void gemm_tile(const GemmArgs &g, uint32_t q, int64_t kt0, int64_t kt1,
                                          bool split, uint32_t sidx)
{
    static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int LDS_A = BM * BK, LDS_B = BN * BK;
    constexpr int STAGE = LDS_A + LDS_B;
    extern __shared__ double smem[];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / (BN / WN), wn = wave % (BN / WN);
    const int64_t K = g.k;

    uint32_t tm, tn;
    tile_coords(q, g.tiles_m, g.tiles_n, tm, tn);
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;

    TileLoader<BM, AMODE> la;
    TileLoader<BN, BMODE> lb;
    d4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

    constexpr bool DENSE = (AMODE <= kDMN) && (BMODE <= kDMN);
    constexpr bool VIEWS = (AMODE >= kVK) && (BMODE >= kVK);
    // Edge tiles take the fast paths too when their missing rows can be
    // clamped (clamp_row): k-contiguous operands always, row-contiguous ones
    // with an even row count (whole row pairs).
    constexpr bool AKC = (AMODE == kDK || AMODE == kVK), BKC = (BMODE == kDK || BMODE == kVK);
    const bool a_in = (m0 + BM <= g.m) || AKC || (g.m % 2 == 0);
    const bool b_in = (n0 + BN <= g.n) || BKC || (g.n % 2 == 0);
    const bool interior = g.a.vec2 && g.b.vec2 && (K % BK == 0) && a_in && b_in;
    // Uniform per workgroup: every load of this tile is in range and aligned.
    const bool fast = DENSE && interior;
    const bool vfast = VIEWS && interior && !g.a.blk && !g.b.blk;
    FastPtrs<BM, AMODE> fa;
    FastPtrs<BN, BMODE> fb;
    ViewFast<BM, AMODE> va;
    ViewFast<BN, BMODE> vb;
    if (kt0 < kt1) {
        if (fast) {
            fa.init(g.a, g.A, m0, g.m, wave, lane);
            fb.init(g.b, g.B, n0, g.n, wave, lane);
            fa.load(la, kt0);
            fb.load(lb, kt0);
        } else if (vfast) {
            va.init(g.a, m0, g.m, wave, lane);
            vb.init(g.b, n0, g.n, wave, lane);
            va.fetch_cols(g.a, kt0 * BK);
            vb.fetch_cols(g.b, kt0 * BK);
            va.load(la, g.A);
            vb.load(lb, g.B);
            if (kt0 + 1 < kt1) {
                va.fetch_cols(g.a, (kt0 + 1) * BK);
                vb.fetch_cols(g.b, (kt0 + 1) * BK);
            }
        } else {
            la.load_checked(g.a, g.A, m0, kt0 * BK, g.m, K, wave, lane);
            lb.load_checked(g.b, g.B, n0, kt0 * BK, g.n, K, wave, lane);
        }
        la.store(smem, wave, lane);
        lb.store(smem + LDS_A, wave, lane);
    }
    __syncthreads();

    const int gA = wm * (WM / 16), gB = wn * (WN / 16);
    for (int64_t kt = kt0; kt < kt1; kt++) {
        const int cur = (int)((kt - kt0) & 1);
        const bool more = kt + 1 < kt1;
        if (more) {
            if (fast) {
                fa.load(la, kt + 1);
                fb.load(lb, kt + 1);
            } else if (vfast) {
                va.load(la, g.A);
                vb.load(lb, g.B);
                if (kt + 2 < kt1) {
                    va.fetch_cols(g.a, (kt + 2) * BK);
                    vb.fetch_cols(g.b, (kt + 2) * BK);
                }
            } else {
                la.load_checked(g.a, g.A, m0, (kt + 1) * BK, g.m, K, wave, lane);
                lb.load_checked(g.b, g.B, n0, (kt + 1) * BK, g.n, K, wave, lane);
            }
        }
        const double *a_s = smem + cur * STAGE + gA * 256 + lane;
        const double *b_s = smem + cur * STAGE + LDS_A + gB * 256 + lane;
        // View operands: the wave starting a K step's MFMAs takes issue
        // priority for its first k-slice, so the two workgroups sharing a
        // CU drift apart and one's loads, stores and barrier overlap the
        // other's MFMAs (ring step 63.5 -> 66.1 TF/s; priority over the
        // whole step, or for the dense modes, measured slower).
        if constexpr (VIEWS) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            if constexpr (VIEWS)
                if (s == 1) __builtin_amdgcn_s_setprio(0);
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
            // The other stage was last read before the previous barrier, so
            // the next step's tile can be stored any time in this step: here
            // after the last k-slice's MFMAs are issued, which gives its
            // global loads the whole step.  (Measured on the ring step: this
            // placement 63.4-63.6 TF/s, after slice 1 62.9, after the loop --
            // outside the unrolled body -- 61.2.)
            if (s == 3 && more) {
                la.store(smem + (cur ^ 1) * STAGE, wave, lane);
                lb.store(smem + (cur ^ 1) * STAGE + LDS_A, wave, lane);
            }
        }
        __syncthreads();
    }

    const int fr = lane & 15, fq = lane >> 4;
    if (split) {
        // K slice of a remainder tile: raw partial sums to the workspace,
        // laid out [slice index][BM][BN]; the reduce kernel adds them to C.
        double *w = g.ws + (int64_t)sidx * BM * BN;
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    w[(wm * WM + i * 16 + fq + 4 * r) * BN + wn * WN + j * 16 + fr] = acc[i][j][r];
        return;
    }

    // Epilogue: C += alpha * acc  (beta = 1, as every reference call site),
    // or C = alpha * acc when the caller asked for an overwrite.
    const double alpha = g.alpha;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) {
            const int64_t col = n0 + wn * WN + j * 16 + fr;
            if (col >= g.n) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * WM + i * 16 + fq + 4 * r;
                if (row >= g.m) continue;
                int64_t addr;
                if constexpr (CVIEW) addr = op_addr<kVK>(g.c, row, col);
                else addr = row * g.c.ld + col;
                if (g.overwrite) g.C[addr] = alpha * acc[i][j][r];
                else g.C[addr] += alpha * acc[i][j][r];
            }
        }
}

template <int BM, int BN, int WM, int WN, int AMODE, int BMODE, bool CVIEW>
__global__ __launch_bounds__(NT, (Occ<BM, BN>::v)) void gemm_f64_kernel(GemmArgs g)
{
    // Workgroups below t_full own whole tiles; their ids are first remapped
    // so consecutive tiles land on one XCD (its own L2).  Workgroups above
    // own one K slice of a remainder tile.
    const uint32_t bid = blockIdx.x;
    const int64_t nk_all = (g.k + BK - 1) / BK;
    uint32_t q, sidx = 0;
    int64_t kt0 = 0, kt1 = nk_all;
    bool split = false;
    if (bid < g.t_full) {
        q = bid;
        if ((g.t_full & 7u) == 0) q = (bid & 7u) * (g.t_full >> 3) + (bid >> 3);
    } else {
        split = true;
        sidx = bid - g.t_full;
        q = g.t_full + sidx / g.sk;
        kt0 = (int64_t)(sidx % g.sk) * g.kchunk;
        kt1 = min(nk_all, kt0 + g.kchunk);
    }
    gemm_tile<BM, BN, WM, WN, AMODE, BMODE, CVIEW>(g, q, kt0, kt1, split, sidx);
}

// Problem owning tail tile ti (tail tiles are numbered problem by problem).
__device__ inline int tail_problem(const GemmGroupArgs &ga, uint32_t ti)
{
    int p = 0;
    for (int i = 0; i < ga.np; i++) {
        const GemmGroupProb &P = ga.p[i];
        if (ti >= P.tail0 && ti < P.tail0 + (P.tiles_m * P.tiles_n - P.t_whole)) p = i;
    }
    return p;
}

// Several independent dense GEMMs C_p (+)= alpha A_p B_p^T in one launch.
// Problem p's whole tiles own workgroups [wg0_p, wg0_p + t_whole_p rounded
// up to 8); wg0_p is a multiple of 8, so a workgroup's XCD is its local id
// mod 8 and each problem gets the same XCD-contiguous tile order as a lone
// launch.  (Cutting the list of all whole tiles into eight runs of equal
// work, one per XCD, was measured slower: 8 x 2048^3 55.0 against 56.6
// TF/s, 2 x 4096 x 4096 x 2048 59.0 against 63.7.)  Workgroups past a
// problem's last whole tile exit at once.  The workgroups after wg_whole
// each own one K slice of a tail tile (TAIL: a launch without tail tiles
// compiles the slice path out, which keeps the 192 x 96 tiles' register
// spills at their level without it).
template <int BM, int BN, int WM, int WN, bool TAIL, int AM, int BMD>
__global__ __launch_bounds__(NT, (Occ<BM, BN>::v)) void gemm_f64_group_kernel(GemmGroupArgs ga)
{
    const uint32_t bid = blockIdx.x;
    int p = 0;
    uint32_t q, sidx = 0;
    bool split = false;
    if (!TAIL || bid < ga.wg_whole) {
        for (int i = 1; i < ga.np; i++)
            if (bid >= ga.p[i].wg0) p = i;
        const GemmGroupProb &P = ga.p[p];
        const uint32_t T = P.t_whole, T8 = (T + 7u) & ~7u;
        const uint32_t local = bid - P.wg0;
        q = (local & 7u) * (T8 >> 3) + (local >> 3);
        if (q >= T) return;  // whole workgroup: no barrier has been reached
    } else {
        split = true;
        sidx = bid - ga.wg_whole;
        const uint32_t ti = sidx / ga.sk;
        p = tail_problem(ga, ti);
        q = ga.p[p].t_whole + (ti - ga.p[p].tail0);
    }
    const GemmGroupProb &P = ga.p[p];
    const int64_t nk = (P.k + BK - 1) / BK;
    int64_t kt0 = 0, kt1 = nk;
    if (split) {
        kt0 = min(nk, (int64_t)(sidx % ga.sk) * P.kchunk);
        kt1 = min(nk, kt0 + P.kchunk);
    }
    GemmArgs g{};
    g.m = P.m;
    g.n = P.n;
    g.k = P.k;
    g.alpha = ga.alpha;
    g.A = P.A;
    g.B = P.B;
    g.C = P.C;
    g.a.ld = P.lda;
    g.a.vec2 = P.vec2a;
    g.b.ld = P.ldb;
    g.b.vec2 = P.vec2b;
    g.c.ld = P.ldc;
    g.tiles_m = P.tiles_m;
    g.tiles_n = P.tiles_n;
    g.overwrite = ga.overwrite;
    g.ws = ga.ws;
    gemm_tile<BM, BN, WM, WN, AM, BMD, false>(g, q, kt0, kt1, TAIL && split, sidx);
}

// Sums the K slices of each tail tile of a group in slice order and adds
// alpha * sum to (or, overwriting, stores it in) the problem's C.
template <int BM, int BN>
__global__ __launch_bounds__(256) void group_split_reduce_kernel(GemmGroupArgs ga)
{
    const uint32_t ti = blockIdx.y;
    const GemmGroupProb &P = ga.p[tail_problem(ga, ti)];
    uint32_t tm, tn;
    tile_coords(P.t_whole + (ti - P.tail0), P.tiles_m, P.tiles_n, tm, tn);
    for (int e = blockIdx.x * 256 + threadIdx.x; e < BM * BN; e += gridDim.x * 256) {
        const int64_t row = (int64_t)tm * BM + e / BN, col = (int64_t)tn * BN + e % BN;
        if (row >= P.m || col >= P.n) continue;
        const double sum = sum_in_order(ga.ws + (int64_t)ti * ga.sk * BM * BN + e, ga.sk, BM * BN);
        const int64_t addr = row * P.ldc + col;
        if (ga.overwrite) P.C[addr] = ga.alpha * sum;
        else P.C[addr] += ga.alpha * sum;
    }
}

// Sums the K slices of each remainder tile in slice order and adds
// alpha * sum to C (deterministic: the order does not depend on timing).
template <int BM, int BN, bool CVIEW>
__global__ __launch_bounds__(256) void split_reduce_kernel(GemmArgs g)
{
    const uint32_t t = blockIdx.y;  // remainder tile
    uint32_t tm, tn;
    tile_coords(g.t_full + t, g.tiles_m, g.tiles_n, tm, tn);
    for (int e = blockIdx.x * 256 + threadIdx.x; e < BM * BN; e += gridDim.x * 256) {
        const int64_t row = (int64_t)tm * BM + e / BN, col = (int64_t)tn * BN + e % BN;
        if (row >= g.m || col >= g.n) continue;
        const double sum = sum_in_order(g.ws + (int64_t)t * g.sk * BM * BN + e, g.sk, BM * BN);
        int64_t addr;
        if constexpr (CVIEW) addr = op_addr<kVK>(g.c, row, col);
        else addr = row * g.c.ld + col;
        if (g.overwrite) g.C[addr] = g.alpha * sum;
        else g.C[addr] += g.alpha * sum;
    }
}

// Matrix-vector product (GemvArgs): the product is HBM-bound (every element
// of M is used once), so instead of MFMA tiles that would run 63 of 64
// columns on padding, each wave takes one row at a time, its lanes read
// 16-byte column pairs 128 columns apart (eight loads in flight per lane),
// x sits in LDS, and the wave sums its lanes in a fixed tree: the order of
// the additions does not depend on timing.
__global__ __launch_bounds__(256) void gemv_kernel(GemvArgs g)
{
    extern __shared__ double xs[];
    for (int64_t c = threadIdx.x; c < g.k; c += 256) xs[c] = g.X[op_addr<kVK>(g.xv, 0, c)];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < g.rows; r += nw) {
        const double *row = g.M + g.mrofs[r];
        double acc[8];
#pragma unroll
        for (int u = 0; u < 8; u++) acc[u] = 0.0;
        int64_t c = 2 * lane;
        for (; c + 7 * 128 < g.k; c += 8 * 128) {
            d2 a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = *(const d2 *)(row + g.mcofs[c + u * 128]);
#pragma unroll
            for (int u = 0; u < 8; u++) {
                acc[u] = fma(a[u].x, xs[c + u * 128], acc[u]);
                acc[u] = fma(a[u].y, xs[c + u * 128 + 1], acc[u]);
            }
        }
        for (int u = 0; c < g.k; c += 128, u++) {
            const d2 a = *(const d2 *)(row + g.mcofs[c]);
            acc[u] = fma(a.x, xs[c], acc[u]);
            acc[u] = fma(a.y, xs[c + 1], acc[u]);
        }
        double v = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) {
            int64_t addr;
            if (g.yview) addr = g.ycol ? op_addr<kVK>(g.yv, r, 0) : op_addr<kVK>(g.yv, 0, r);
            else addr = g.ycol ? r * g.ldy : r;
            if (g.overwrite) g.Y[addr] = g.alpha * v;
            else g.Y[addr] += g.alpha * v;
        }
    }
}

// Matrix-vector product with M's row pairs adjacent (GemvArgs, mrofs of
// even rows point at pairs): each lane owns a pair of rows and streams one
// 16-byte load per column, the column offset and x(c) being the same for
// the whole wave (scalar loads); a workgroup's 16 waves take 64 row pairs
// each over their own sixteenth of the workgroup's columns, in column
// order, and the sixteen partial sums are added in wave order through LDS.
// With few rows the columns are also split over gridDim.y workgroups, whose
// partial sums (ws[slice][row]) gemv_rows_reduce adds in slice order: the
// order of the additions never depends on timing.
constexpr int kGemvRowsWaves = 16;
__device__ inline void gemv_store(const GemvArgs &g, int64_t r, double v)
{
    int64_t addr;
    if (g.yview) addr = g.ycol ? op_addr<kVK>(g.yv, r, 0) : op_addr<kVK>(g.yv, 0, r);
    else addr = g.ycol ? r * g.ldy : r;
    if (g.overwrite) g.Y[addr] = v;
    else g.Y[addr] += v;
}

__global__ __launch_bounds__(64 * kGemvRowsWaves) void gemv_rows_kernel(GemvArgs g, double *ws)
{
    __shared__ d2 red[kGemvRowsWaves][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t npairs = g.rows / 2;
    const int64_t pr = (int64_t)blockIdx.x * 64 + lane;
    const double *row = g.M + g.mrofs[2 * min(pr, npairs - 1)];
    const int64_t ks = (g.k + gridDim.y - 1) / gridDim.y;  // columns of this workgroup's slice
    const int64_t b0 = min((int64_t)blockIdx.y * ks, g.k), b1 = min(b0 + ks, g.k);
    const int64_t kc = (b1 - b0 + kGemvRowsWaves - 1) / kGemvRowsWaves;
    const int64_t c0 = min(b0 + (int64_t)w * kc, b1), c1 = min(c0 + kc, b1);
    d2 acc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u] = (d2){0.0, 0.0};
    int64_t c = c0;
    for (; c + 4 <= c1; c += 4) {
        d2 v[4];
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            v[u] = *reinterpret_cast<const d2 *>(row + g.mcofs[c + u]);
            x[u] = g.X[op_addr<kVK>(g.xv, 0, c + u)];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            acc[u].x = fma(v[u].x, x[u], acc[u].x);
            acc[u].y = fma(v[u].y, x[u], acc[u].y);
        }
    }
    for (int u = 0; c < c1; c++, u++) {
        const d2 v = *reinterpret_cast<const d2 *>(row + g.mcofs[c]);
        const double x = g.X[op_addr<kVK>(g.xv, 0, c)];
        acc[u].x = fma(v.x, x, acc[u].x);
        acc[u].y = fma(v.y, x, acc[u].y);
    }
    red[w][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (w != 0 || pr >= npairs) return;
    d2 sum = red[0][lane];
#pragma unroll
    for (int q = 1; q < kGemvRowsWaves; q++) sum += red[q][lane];
    if (gridDim.y > 1) {
        *reinterpret_cast<d2 *>(ws + (int64_t)blockIdx.y * g.rows + 2 * pr) = sum;
        return;
    }
    gemv_store(g, 2 * pr, g.alpha * sum.x);
    gemv_store(g, 2 * pr + 1, g.alpha * sum.y);
}

__global__ __launch_bounds__(256) void gemv_rows_reduce(GemvArgs g, const double *ws, uint32_t nslice)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= g.rows) return;
    gemv_store(g, r, g.alpha * sum_in_order(ws + r, nslice, g.rows));
}

// Thin products (ThinArgs): C (rows x nt) = alpha * T S^T with T the long
// operand (rows x k, k short) and S the thin one (nt x k; nt <= 48, or up
// to 192 when k <= 48).  Every element of T is used once, so the product
// is bound by streaming T from HBM, by the FP64 MFMA rate (about as fast at
// nt = 48) or, at wide nt and short k, by writing C; tiles of 64 x 64
// re-read S per tile and keep too few loads in flight (3 TB/s).
// Here S lives in LDS for the whole launch, pre-arranged as the MFMA
// fragments of every 8-wide K chunk (two per chunk, t = 0, 1, one read of
// 512 contiguous bytes each) and 16-wide block of S rows, and each wave
// streams units of 32 rows of T through registers: a unit's K is read in
// super chunks of kThinSC chunks (two 16-byte loads per lane per 8-wide
// chunk), the next super chunk (of this unit or the wave's next) loading
// while this one runs its MFMAs.  With nt > 48 (k <= 48: one super chunk)
// the unit's registers serve ng groups of 48 S rows in turn.
//   KP (k pairs adjacent, kDK / kVK): lane (i = l%16, q = l/16) loads
//     T[r][k0 + 2q], T[r][k0 + 2q + 1] for r = row0 + i (row block 0) and
//     r = row0 + 16 + i (block 1); component x pairs with fragment t = 0
//     (k = k0 + 2q), y with t = 1.
//   !KP (row pairs adjacent, kDMN / kVMN): lane loads T[row0 + 2i][k],
//     T[row0 + 2i + 1][k] for k = k0 + q (t = 0) and k0 + 4 + q (t = 1);
//     x feeds row block 0 (even rows), y block 1 (odd rows).
// The register fragment is the MFMA's A operand (T rows down the result,
// S rows across: lanes of a store run along S) or, TC, its B operand (the
// result transposed: lanes run along T), chosen so a store's 16 lanes
// write consecutive addresses of C: TC when C(r, j) is C's column r (the
// product of a thin A, written transposed).
// Both sum every C element over k in a fixed order, independent of timing.
// No load's result is used in the item that issues it (the compiler's
// in-order load counter would otherwise make every MFMA of that item wait
// for the prefetch as well): the row offsets of a unit are loaded one item
// ahead into one of two register sets, the T side's C offsets of the unit
// being computed are loaded at the start of each item, the S side's sit in
// LDS, and every load is unconditional (past the wave's last item and past
// K they repeat valid addresses).
// Registers: two super chunks (96), 2 x NBB accumulators and the offsets;
// NBB = 4 would spill at two waves per SIMD, so a group is 48 wide.
constexpr int kThinSC = 6;
struct ThinArgs {
    int64_t rows, nt, k;
    int32_t ng;  // groups of 16 * NBB rows of S
    double alpha;
    const double *T;
    OpArgs t;
    int32_t tdense;  // T dense (ld) or a view without block table
    const double *S;
    OpArgs s;
    int32_t smode;   // kDK, kDMN or a view mode
    double *C;
    OpArgs c;
    int32_t cview, ctr;  // ctr: element (r, j) of the product is C(j, r); views have no block table
    int32_t cpair;       // ctr: C(j, r), C(j, r + 1) adjacent and 16-byte aligned for even r
    int32_t overwrite;
};

__device__ inline double thin_s(const ThinArgs &g, int64_t j, int64_t kk)
{
    if (g.smode == kDK) return g.S[j * g.s.ld + kk];
    if (g.smode == kDMN) return g.S[kk * g.s.ld + j];
    return g.S[op_addr<kVK>(g.s, j, kk)];
}

// TD: T dense; CV: C a view; TC: transposed C (ctr), the transposed MFMA
// form.  Compile-time, so no load sits in a branch whose other side writes
// the same register, which would force a full wait.
template <bool KP, int NBB, bool TD, bool CV, bool TC>
__global__ __launch_bounds__(256, 2) void thin_kernel(ThinArgs g)
{
    extern __shared__ double smem[];
    const int64_t K = g.k;
    const int NBT = g.ng * NBB;  // 16-row blocks of S
    double *bf = smem;           // [K / 8][t][NBT][64 lanes]
    int64_t *cp = reinterpret_cast<int64_t *>(smem + K * NBT * 16);  // column part of T per k
    int64_t *cthin = cp + K;                                         // C offset per S row
    const int tid = threadIdx.x, lane = tid & 63, i16 = lane & 15, q = lane >> 4;

    for (int64_t e = tid; e < K * NBT * 16; e += 256) {
        const int l = (int)(e & 63);
        const int64_t f = e >> 6;
        const int cb = (int)(f % NBT), t = (int)((f / NBT) & 1);
        const int64_t c = f / NBT / 2;
        const int64_t j = 16 * cb + (l & 15), kk = KP ? 8 * c + 2 * (l >> 4) + t : 8 * c + 4 * t + (l >> 4);
        bf[e] = j < g.nt ? thin_s(g, j, kk) : 0.0;
    }
    for (int64_t c = tid; c < K; c += 256) {
        if constexpr (!TD) cp[c] = g.t.cofs[c];
        else cp[c] = KP ? c : c * g.t.ld;
    }
    for (int j = tid; j < 16 * NBT; j += 256) {
        const int64_t jj = min((int64_t)j, g.nt - 1);
        if constexpr (CV) cthin[j] = TC ? g.c.rofs[jj] : g.c.cofs[jj];
        else cthin[j] = TC ? jj * g.c.ld : jj;
    }
    __syncthreads();

    const int64_t U = (g.rows + 31) / 32, NW = (int64_t)gridDim.x * 4;
    int64_t cu = (int64_t)blockIdx.x * 4 + (tid >> 6);
    if (cu >= U) return;
    const int nsc = (int)((K + 8 * kThinSC - 1) / (8 * kThinSC));

    // Row parts of T for unit u in this lane (rows past the end load the
    // last row or row pair: in range, never stored).
    auto row_part = [&](int64_t r) -> int64_t {
        if constexpr (!TD) return g.t.rofs[r];
        else return KP ? r * g.t.ld : r;
    };
    auto rows_of = [&](int64_t u, int64_t (&p)[2]) __attribute__((always_inline)) {
        if constexpr (KP) {
            p[0] = row_part(min(u * 32 + i16, g.rows - 1));
            p[1] = row_part(min(u * 32 + 16 + i16, g.rows - 1));
        } else {
            p[0] = row_part(min(u * 32 + 2 * i16, g.rows - 2));
            p[1] = 0;
        }
    };
    // MFMA result: lane l, element r holds result row q + 4r, column i16.
    // The T row of element r of row block b (TC: the same for every r) and
    // the S row of element r of 16-row block jb of S.
    auto t_row = [&](int64_t u, int b, int r) -> int64_t {
        const int loc = TC ? i16 : q + 4 * r;
        return KP ? u * 32 + 16 * b + loc : u * 32 + 2 * loc + b;
    };
    auto s_row = [&](int jb, int r) -> int { return 16 * jb + (TC ? q + 4 * r : i16); };
    constexpr int NO = TC ? 1 : 4;  // distinct T rows per row block and lane
    // Rows past the end are clamped into range; with row pairs the even row
    // clamps to rows - 2 (rows is even), so a clamped pair stays a whole,
    // aligned pair.
    auto c_offsets = [&](int64_t u, int64_t (&o)[2][NO]) __attribute__((always_inline)) {
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < NO; r++) {
                const int64_t row = KP ? min(t_row(u, b, r), g.rows - 1) : min(t_row(u, 0, r), g.rows - 2) + b;
                if constexpr (CV) o[b][r] = (TC ? g.c.cofs : g.c.rofs)[row];
                else o[b][r] = TC ? row : row * g.c.ld;
            }
    };
    auto load = [&](d2 (&a)[2 * kThinSC], const int64_t (&p)[2], int sc) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kThinSC; c++) {
            const int64_t k0 = min((int64_t)(8 * kThinSC * sc + 8 * c), K - 8);
            if constexpr (KP) {
                const int64_t co = cp[k0 + 2 * q];
                a[2 * c] = *reinterpret_cast<const d2 *>(g.T + p[0] + co);
                a[2 * c + 1] = *reinterpret_cast<const d2 *>(g.T + p[1] + co);
            } else {
                a[2 * c] = *reinterpret_cast<const d2 *>(g.T + p[0] + cp[k0 + q]);
                a[2 * c + 1] = *reinterpret_cast<const d2 *>(g.T + p[0] + cp[k0 + 4 + q]);
            }
        }
    };
    d4 acc[2][NBB];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int j = 0; j < NBB; j++) acc[b][j] = (d4){0.0, 0.0, 0.0, 0.0};
    auto mma = [&](double tv, double sv, d4 c) -> d4 {
        if constexpr (TC) return __builtin_amdgcn_mfma_f64_16x16x4f64(sv, tv, c, 0, 0, 0);
        else return __builtin_amdgcn_mfma_f64_16x16x4f64(tv, sv, c, 0, 0, 0);
    };
    auto compute = [&](const d2 (&a)[2 * kThinSC], int sc, int grp) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kThinSC; c++) {
            const int64_t kc = kThinSC * sc + c;
            if (8 * kc >= K) break;
#pragma unroll
            for (int j = 0; j < NBB; j++) {
                const int jb = grp * NBB + j;
                const double b0 = bf[((kc * 2) * NBT + jb) * 64 + lane];
                const double b1 = bf[((kc * 2 + 1) * NBT + jb) * 64 + lane];
                if constexpr (KP) {
                    acc[0][j] = mma(a[2 * c].x, b0, acc[0][j]);
                    acc[1][j] = mma(a[2 * c + 1].x, b0, acc[1][j]);
                    acc[0][j] = mma(a[2 * c].y, b1, acc[0][j]);
                    acc[1][j] = mma(a[2 * c + 1].y, b1, acc[1][j]);
                } else {
                    acc[0][j] = mma(a[2 * c].x, b0, acc[0][j]);
                    acc[1][j] = mma(a[2 * c].y, b0, acc[1][j]);
                    acc[0][j] = mma(a[2 * c + 1].x, b1, acc[0][j]);
                    acc[1][j] = mma(a[2 * c + 1].y, b1, acc[1][j]);
                }
            }
        }
    };
    // Accumulating into C reads a row block's old values with one batch of
    // loads (clamped addresses, all in range) and one wait.  Row pairs in
    // the transposed form: a lane's two row blocks hold T rows 2i and 2i + 1,
    // adjacent in C, so with cpair they go out as one 16-byte store (two
    // 8-byte stores a row apart left half-written lines for L2 to write
    // back, 1.5x the bytes).
    auto store = [&](int64_t u, const int64_t (&o)[2][NO], int grp) __attribute__((always_inline)) {
        if constexpr (!KP && TC) {
            if (g.cpair) {
                const bool rin = t_row(u, 0, 0) < g.rows;  // rows are even: the pair is whole
                d2 old2[NBB][4];
                if (!g.overwrite) {
#pragma unroll
                    for (int j = 0; j < NBB; j++)
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            old2[j][r] = *reinterpret_cast<const d2 *>(
                                g.C + o[0][0] + cthin[s_row(grp * NBB + j, r)]);
                }
#pragma unroll
                for (int j = 0; j < NBB; j++) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int sr = s_row(grp * NBB + j, r);
                        if (!rin || sr >= g.nt) continue;
                        d2 v = {g.alpha * acc[0][j][r], g.alpha * acc[1][j][r]};
                        if (!g.overwrite) v = old2[j][r] + v;
                        *reinterpret_cast<d2 *>(g.C + o[0][0] + cthin[sr]) = v;
                    }
                    acc[0][j] = (d4){0.0, 0.0, 0.0, 0.0};
                    acc[1][j] = (d4){0.0, 0.0, 0.0, 0.0};
                }
                return;
            }
        }
#pragma unroll
        for (int b = 0; b < 2; b++) {
            double old[NBB][4];
            if (!g.overwrite) {
#pragma unroll
                for (int j = 0; j < NBB; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        old[j][r] = g.C[o[b][TC ? 0 : r] + cthin[s_row(grp * NBB + j, r)]];
            }
#pragma unroll
            for (int j = 0; j < NBB; j++) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int sr = s_row(grp * NBB + j, r);
                    if (t_row(u, b, r) >= g.rows || sr >= g.nt) continue;
                    const double v = g.alpha * acc[b][j][r];
                    g.C[o[b][TC ? 0 : r] + cthin[sr]] = g.overwrite ? v : old[j][r] + v;
                }
                acc[b][j] = (d4){0.0, 0.0, 0.0, 0.0};
            }
        }
    };

    // The wave's items (unit, super chunk) in order; item n + 1 loads while
    // item n computes.  p: row parts of the unit being loaded; qa / qb: those
    // of the unit after it, loaded by alternate items and used one item
    // later.  Unit indices past the end are clamped, so every item issues
    // the same loads.
    int64_t p[2], qa[2], qb[2], co[2][NO];
    rows_of(cu, p);
    rows_of(min(cu + NW, U - 1), qb);
    d2 ra[2 * kThinSC], rb[2 * kThinSC];
    load(ra, p, 0);
    int cs = 0;
    auto next = [&](d2 (&a)[2 * kThinSC], int64_t (&qload)[2], const int64_t (&quse)[2], int64_t &lu,
                    int &ls) __attribute__((always_inline)) {
        c_offsets(cu, co);
        lu = cu;
        ls = cs + 1;
        if (ls == nsc) {
            ls = 0;
            lu += NW;
        }
        const bool nu = ls == 0;
        p[0] = nu ? quse[0] : p[0];
        p[1] = nu ? quse[1] : p[1];
        rows_of(min(lu + NW, U - 1), qload);
        load(a, p, ls);
    };
    // One item: every group of S rows (several only with one super chunk).
    auto work = [&](const d2 (&a)[2 * kThinSC]) __attribute__((always_inline)) {
        for (int grp = 0; grp < g.ng; grp++) {
            compute(a, cs, grp);
            if (cs == nsc - 1) store(cu, co, grp);
        }
    };
    while (true) {
        int64_t lu;
        int ls;
        next(rb, qa, qb, lu, ls);
        work(ra);
        if (lu >= U) break;
        cu = lu;
        cs = ls;

        next(ra, qb, qa, lu, ls);
        work(rb);
        if (lu >= U) break;
        cu = lu;
        cs = ls;
    }
}

// CU count per device, and the tail-split workspace per (device, stream):
// GEMMs enqueued on different streams never share a workspace; GEMMs on one
// stream are ordered by it.
std::mutex g_dev_mu;
int g_cus[64];
std::map<std::pair<int, hipStream_t>, std::pair<double *, size_t>> g_ws;

int device_cus(int *cus)
{
    int d = 0;
    DTCF_HIP(hipGetDevice(&d));
    DTCF_REQUIRE(d >= 0 && d < 64, "gemm: device index out of range");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_cus[d] == 0) {
        int c = 0;
        DTCF_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d));
        g_cus[d] = c > 0 ? c : 1;
    }
    *cus = g_cus[d];
    return DTCF_OK;
}

int split_workspace(hipStream_t s, size_t need, double **ws)
{
    int d = 0;
    DTCF_HIP(hipGetDevice(&d));
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto &e = g_ws[std::make_pair(d, s)];
    if (e.second < need) {
        // growing allocates (and frees the old buffer): never inside a graph
        // capture, where the allocation would run outside the graph and the
        // free could release a buffer an earlier graph still uses
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        DTCF_HIP(hipStreamIsCapturing(s, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            set_error("gemm: the split-K workspace of a capturing stream must be reserved before "
                      "capture (dtcf_gemm_reserve_workspace)");
            return DTCF_ERR_INVALID;
        }
        if (e.first) DTCF_HIP(hipStreamSynchronize(s));  // the old buffer may be in use
        if (e.first) (void)hipFree(e.first);
        e.first = nullptr;
        e.second = 0;
        DTCF_HIP(hipMalloc(&e.first, need));
        e.second = need;
    }
    *ws = e.first;
    return DTCF_OK;
}

// How many of T tiles to cut into K slices (nsplit, the last ones) and
// into how many (sk >= 2; sk < 2 means no split), for `slots` resident
// workgroups and nk K steps per tile.  A last partial round of at most half
// the slots: its tiles are cut to fill one round.  Under-filled grid (at
// most half the slots): every tile is cut.  One partial round (more than
// half the slots, at most all): every tile is cut into the sk <= 8 slices
// that fill whole rounds best, within two rounds of slots, when that beats
// the unsplit fill by 3% (a 340-tile GEMM on 512 slots: 3 slices, 1020
// items in 2 rounds, 66% -> 99.6% of the slots busy).  In every case
// nsplit * sk <= 2 * slots, and every slice is non-empty.
void tail_split(uint32_t T, uint32_t slots, int64_t nk, uint32_t &nsplit, uint32_t &sk)
{
    const uint32_t rem = T % slots;
    nsplit = 0;
    sk = 0;
    if (T > slots && rem > 0 && rem <= slots / 2 && nk >= 32) {
        nsplit = rem;
        sk = std::min<uint32_t>(slots / rem, (uint32_t)(nk / 8));
    } else if (T <= slots / 2 && nk >= 16) {
        nsplit = T;
        sk = std::min<uint32_t>(slots / T, (uint32_t)(nk / 4));
    } else if (T <= slots && nk >= 32) {
        double best = (double)T / slots;
        for (uint32_t c = 2; c <= 8 && nk / c >= 8 && T * c <= 2 * slots; c++) {
            const uint32_t items = T * c;
            const double fill = (double)items / (double)(((items + slots - 1) / slots) * slots);
            if (fill > best * 1.03) {
                best = fill;
                sk = c;
            }
        }
        if (sk >= 2) nsplit = T;
    }
}

template <int BM, int BN, int WM, int WN, int AM, int BMD, bool CV>
int launch_one(GemmArgs g, hipStream_t s)
{
    auto kern = gemm_f64_kernel<BM, BN, WM, WN, AM, BMD, CV>;
    constexpr size_t lds = (size_t)2 * (BM + BN) * BK * sizeof(double);
    if constexpr (lds > 65536) {
        static std::once_flag once;
        static hipError_t attr_err = hipSuccess;
        std::call_once(once, [&] {
            attr_err = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        });
        DTCF_HIP(attr_err);
    }
    // Tail split: when the last wave of whole tiles would leave most slots
    // idle, its tiles are cut into K slices that fill one wave instead.
    int cus = 0;
    int st = device_cus(&cus);
    if (st != DTCF_OK) return st;
    const uint32_t T = g.tiles_m * g.tiles_n;
    const uint32_t slots = (uint32_t)cus * (uint32_t)Occ<BM, BN>::v;
    const int64_t nk = (g.k + BK - 1) / BK;
    g.t_full = T;
    g.sk = 1;
    g.kchunk = nk;
    uint32_t nsplit = 0, sk = 0;
    tail_split(T, slots, nk, nsplit, sk);
    if (sk >= 2) {
        const int64_t kchunk = (nk + sk - 1) / sk;
        sk = (uint32_t)((nk + kchunk - 1) / kchunk);
        double *ws = nullptr;
        st = split_workspace(s, (size_t)nsplit * sk * BM * BN * sizeof(double), &ws);
        if (st != DTCF_OK) return st;
        g.ws = ws;
        g.t_full = T - nsplit;
        g.sk = sk;
        g.kchunk = kchunk;
    }
    const unsigned grid = g.t_full + (T - g.t_full) * g.sk;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, g);
    DTCF_HIP(hipGetLastError());
    if (g.t_full < T) {
        hipLaunchKernelGGL((split_reduce_kernel<BM, BN, CV>), dim3((BM * BN + 255) / 256, T - g.t_full),
                           dim3(256), 0, s, g);
        DTCF_HIP(hipGetLastError());
    }
    return DTCF_OK;
}

template <int BM, int BN, int WM, int WN>
int launch_cfg(GemmArgs g, hipStream_t s)
{
    g.tiles_m = (uint32_t)((g.m + BM - 1) / BM);
    g.tiles_n = (uint32_t)((g.n + BN - 1) / BN);
    const int am = g.amode, bm = g.bmode;
    if (am <= kDMN && bm <= kDMN && !g.cview) {
        if (am == kDK && bm == kDK) return launch_one<BM, BN, WM, WN, kDK, kDK, false>(g, s);
        if (am == kDK && bm == kDMN) return launch_one<BM, BN, WM, WN, kDK, kDMN, false>(g, s);
        if (am == kDMN && bm == kDK) return launch_one<BM, BN, WM, WN, kDMN, kDK, false>(g, s);
        if (am == kDMN && bm == kDMN) return launch_one<BM, BN, WM, WN, kDMN, kDMN, false>(g, s);
    } else if (am >= kVK && bm >= kVK && g.cview) {
        if (am == kVK && bm == kVK) return launch_one<BM, BN, WM, WN, kVK, kVK, true>(g, s);
        if (am == kVK && bm == kVMN) return launch_one<BM, BN, WM, WN, kVK, kVMN, true>(g, s);
        if (am == kVMN && bm == kVK) return launch_one<BM, BN, WM, WN, kVMN, kVK, true>(g, s);
        if (am == kVMN && bm == kVMN) return launch_one<BM, BN, WM, WN, kVMN, kVMN, true>(g, s);
    } else if (am >= kVK && bm >= kVK) {
        if (am == kVK && bm == kVK) return launch_one<BM, BN, WM, WN, kVK, kVK, false>(g, s);
        if (am == kVK && bm == kVMN) return launch_one<BM, BN, WM, WN, kVK, kVMN, false>(g, s);
        if (am == kVMN && bm == kVK) return launch_one<BM, BN, WM, WN, kVMN, kVK, false>(g, s);
        if (am == kVMN && bm == kVMN) return launch_one<BM, BN, WM, WN, kVMN, kVMN, false>(g, s);
    }
    set_error("gemm: unsupported operand mode combination");
    return DTCF_ERR_INVALID;
}

template <int BM, int BN, int WM, int WN, int AM, int BMD>
int launch_group(GemmGroupArgs ga, hipStream_t s)
{
    auto kern = gemm_f64_group_kernel<BM, BN, WM, WN, false, AM, BMD>;
    auto kern_tail = gemm_f64_group_kernel<BM, BN, WM, WN, true, AM, BMD>;
    constexpr size_t lds = (size_t)2 * (BM + BN) * BK * sizeof(double);
    if constexpr (lds > 65536) {
        static std::once_flag once;
        static hipError_t attr_err = hipSuccess;
        std::call_once(once, [&] {
            attr_err = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (attr_err == hipSuccess)
                attr_err = hipFuncSetAttribute(reinterpret_cast<const void *>(kern_tail),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        });
        DTCF_HIP(attr_err);
    }
    uint64_t total = 0;
    int64_t nk_min = INT64_MAX;
    for (int i = 0; i < ga.np; i++) {
        GemmGroupProb &P = ga.p[i];
        P.tiles_m = (uint32_t)((P.m + BM - 1) / BM);
        P.tiles_n = (uint32_t)((P.n + BN - 1) / BN);
        total += (uint64_t)P.tiles_m * P.tiles_n;
        nk_min = std::min<int64_t>(nk_min, (P.k + BK - 1) / BK);
    }
    DTCF_REQUIRE(total + 8 * kGroupMax < (1ull << 31), "gemm group: too many tiles");
    if (total == 0) return DTCF_OK;
    // Tail split over the whole group: the last tiles (of the last
    // problems, the shortest in K) are cut into K slices as a single GEMM's
    // would be, with the shortest K of the group bounding the slice count.
    int cus = 0;
    int st = device_cus(&cus);
    if (st != DTCF_OK) return st;
    uint32_t nsplit = 0, sk = 0;
    tail_split((uint32_t)total, (uint32_t)cus * (uint32_t)Occ<BM, BN>::v, nk_min, nsplit, sk);
    if (sk < 2) nsplit = 0;
    ga.sk = sk < 2 ? 1 : sk;
    ga.ws = nullptr;
    if (nsplit) {
        st = split_workspace(s, (size_t)nsplit * sk * BM * BN * sizeof(double), &ga.ws);
        if (st != DTCF_OK) return st;
    }
    uint32_t left = nsplit;
    for (int i = ga.np - 1; i >= 0; i--) {
        GemmGroupProb &P = ga.p[i];
        const uint32_t T = P.tiles_m * P.tiles_n, take = std::min(left, T);
        P.t_whole = T - take;
        left -= take;
        const int64_t nk = (P.k + BK - 1) / BK;
        P.kchunk = (nk + ga.sk - 1) / ga.sk;
    }
    uint64_t wg = 0;
    uint32_t tail = 0;
    for (int i = 0; i < ga.np; i++) {
        GemmGroupProb &P = ga.p[i];
        P.wg0 = (uint32_t)wg;
        wg += ((uint64_t)P.t_whole + 7) & ~7ull;
        P.tail0 = tail;
        tail += P.tiles_m * P.tiles_n - P.t_whole;
    }
    ga.wg_whole = (uint32_t)wg;
    const uint64_t grid = (uint64_t)ga.wg_whole + (uint64_t)nsplit * ga.sk;
    if (nsplit) hipLaunchKernelGGL(kern_tail, dim3((unsigned)grid), dim3(NT), lds, s, ga);
    else hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NT), lds, s, ga);
    DTCF_HIP(hipGetLastError());
    if (nsplit) {
        hipLaunchKernelGGL((group_split_reduce_kernel<BM, BN>), dim3((BM * BN + 255) / 256, nsplit),
                           dim3(256), 0, s, ga);
        DTCF_HIP(hipGetLastError());
    }
    return DTCF_OK;
}

template <int BM, int BN, int WM, int WN>
int launch_group_modes(const GemmGroupArgs &ga, int amode, int bmode, hipStream_t s)
{
    if (amode == kDK && bmode == kDK) return launch_group<BM, BN, WM, WN, kDK, kDK>(ga, s);
    if (amode == kDK && bmode == kDMN) return launch_group<BM, BN, WM, WN, kDK, kDMN>(ga, s);
    if (amode == kDMN && bmode == kDK) return launch_group<BM, BN, WM, WN, kDMN, kDK>(ga, s);
    if (amode == kDMN && bmode == kDMN) return launch_group<BM, BN, WM, WN, kDMN, kDMN>(ga, s);
    set_error("gemm group: unsupported operand layout");
    return DTCF_ERR_INVALID;
}

}  // namespace

int gemm_f64_group_launch(const GemmGroupArgs &ga, int cfg, int amode, int bmode, hipStream_t s)
{
    DTCF_REQUIRE(ga.np >= 0 && ga.np <= kGroupMax, "gemm group: too many problems");
    for (int i = 0; i < ga.np; i++)
        DTCF_REQUIRE(ga.p[i].m > 0 && ga.p[i].n > 0 && ga.p[i].k > 0, "gemm group: empty problem");
    if (ga.np == 0) return DTCF_OK;
    switch (cfg) {
    case kCfg128x128: return launch_group_modes<128, 128, 64, 64>(ga, amode, bmode, s);
    case kCfg192x96: return launch_group_modes<192, 96, 96, 48>(ga, amode, bmode, s);
    case kCfg96x192: return launch_group_modes<96, 192, 48, 96>(ga, amode, bmode, s);
    case kCfg64x64: return launch_group_modes<64, 64, 32, 32>(ga, amode, bmode, s);
    default: break;
    }
    set_error("gemm group: unknown tile configuration");
    return DTCF_ERR_INVALID;
}

int gemv_f64_launch(const GemvArgs &g, hipStream_t s)
{
    DTCF_REQUIRE(g.k > 0 && g.k % 2 == 0 && g.k <= kGemvMaxK, "gemv: k must be even and <= 20480");
    if (g.rows <= 0) return DTCF_OK;
    static std::once_flag once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(once, [&] {
        attr_err = hipFuncSetAttribute(reinterpret_cast<const void *>(gemv_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(kGemvMaxK * sizeof(double)));
    });
    DTCF_HIP(attr_err);
    const unsigned grid = (unsigned)std::min<int64_t>((g.rows + 3) / 4, 2048);
    hipLaunchKernelGGL(gemv_kernel, dim3(grid), dim3(256), (size_t)g.k * sizeof(double), s, g);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

// 16-row blocks of S a thin product keeps in LDS (0: not a thin shape).
static int64_t thin_blocks(int64_t w, int64_t k)
{
    if (w <= kThinMaxWidth) return (w + 15) / 16;
    if (w <= kThinMaxWide && k <= kThinWideK) return 3 * ((w + 47) / 48);
    return 0;
}

bool thin_shape(int64_t m, int64_t n, int64_t k, int64_t min_rows)
{
    const int64_t w = std::min(m, n), rows = std::max(m, n);
    const int64_t nbt = thin_blocks(w, k);
    if (nbt == 0 || rows < min_rows || k <= 0 || k % 8 != 0) return false;
    return (k * 16 * nbt + k + 16 * nbt) * (int64_t)sizeof(double) <= kThinLdsBytes;
}

template <bool KP, int NBB, bool TD, bool CV, bool TC>
int thin_launch_t(const ThinArgs &t, hipStream_t s)
{
    auto kern = thin_kernel<KP, NBB, TD, CV, TC>;
    static std::once_flag once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(once, [&] {
        attr_err = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kThinLdsBytes);
    });
    DTCF_HIP(attr_err);
    int cus = 0;
    const int st = device_cus(&cus);
    if (st != DTCF_OK) return st;
    // Two resident workgroups per CU (LDS up to 80 KiB each), each wave
    // looping over 32-row units.
    const int64_t units = (t.rows + 31) / 32, nbt = (int64_t)t.ng * NBB;
    const unsigned grid = (unsigned)std::min<int64_t>((units + 3) / 4, (int64_t)cus * 2);
    const size_t lds = (size_t)(t.k * 16 * nbt + t.k + 16 * nbt) * sizeof(double);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, t);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

template <bool KP, bool TD, bool CV, bool TC>
int thin_dispatch_nbb(const ThinArgs &t, int nbb, hipStream_t s)
{
    if (nbb == 1) return thin_launch_t<KP, 1, TD, CV, TC>(t, s);
    if (nbb == 2) return thin_launch_t<KP, 2, TD, CV, TC>(t, s);
    return thin_launch_t<KP, 3, TD, CV, TC>(t, s);
}

template <bool KP, bool TD>
int thin_dispatch(const ThinArgs &t, int nbb, hipStream_t s)
{
    if (t.cview) {
        if (t.ctr) return thin_dispatch_nbb<KP, TD, true, true>(t, nbb, s);
        return thin_dispatch_nbb<KP, TD, true, false>(t, nbb, s);
    }
    if (t.ctr) return thin_dispatch_nbb<KP, TD, false, true>(t, nbb, s);
    return thin_dispatch_nbb<KP, TD, false, false>(t, nbb, s);
}

int thin_f64_launch(const GemmArgs &g, hipStream_t s)
{
    // The long operand: A when n is the thin side, else B (C written
    // transposed).
    const bool tall_a = g.m >= g.n;
    ThinArgs t{};
    t.rows = tall_a ? g.m : g.n;
    t.nt = tall_a ? g.n : g.m;
    t.k = g.k;
    if (!thin_shape(t.rows, t.nt, t.k, 1)) return kThinDecline;
    const OpArgs &to = tall_a ? g.a : g.b;
    const int tmode = tall_a ? g.amode : g.bmode;
    const bool kp = tmode == kDK || tmode == kVK;
    if (!to.vec2 || (tmode >= kVK && to.blk) || (!kp && t.rows % 2 != 0)) return kThinDecline;
    if (g.cview && g.c.blk) return kThinDecline;
    const int64_t nbt = thin_blocks(t.nt, t.k);
    const int nbb = (int)std::min<int64_t>(nbt, 3);
    t.ng = (int32_t)(nbt / nbb);
    t.alpha = g.alpha;
    t.T = tall_a ? g.A : g.B;
    t.t = to;
    t.tdense = tmode <= kDMN;
    t.S = tall_a ? g.B : g.A;
    t.s = tall_a ? g.b : g.a;
    t.smode = tall_a ? g.bmode : g.amode;
    t.C = g.C;
    t.c = g.c;
    t.cview = g.cview;
    t.ctr = tall_a ? 0 : 1;
    t.cpair = tall_a ? 0
                     : (g.cview ? g.c.vec2
                                : ((reinterpret_cast<uintptr_t>(g.C) & 15) == 0 && g.c.ld % 2 == 0));
    t.overwrite = g.overwrite;
    if (kp) {
        if (t.tdense) return thin_dispatch<true, true>(t, nbb, s);
        return thin_dispatch<true, false>(t, nbb, s);
    }
    if (t.tdense) return thin_dispatch<false, true>(t, nbb, s);
    return thin_dispatch<false, false>(t, nbb, s);
}

int gemv_rows_f64_launch(const GemvArgs &g, hipStream_t s)
{
    DTCF_REQUIRE(g.k > 0 && g.rows % 2 == 0, "gemv rows: k must be positive and the row count even");
    if (g.rows <= 0) return DTCF_OK;
    const int64_t blocks = (g.rows / 2 + 63) / 64;
    DTCF_REQUIRE(blocks < (1ll << 31), "gemv rows: too many rows");
    int cus = 0;
    int st = device_cus(&cus);
    if (st != DTCF_OK) return st;
    // Column slices over workgroups until the grid covers two workgroups
    // per CU, each slice keeping at least 64 columns per wave.
    int64_t nslice = 1;
    while (blocks * nslice < 2 * (int64_t)cus && g.k / (2 * nslice) >= 64 * kGemvRowsWaves) nslice *= 2;
    double *ws = nullptr;
    if (nslice > 1) {
        st = split_workspace(s, (size_t)nslice * g.rows * sizeof(double), &ws);
        if (st != DTCF_OK) return st;
    }
    hipLaunchKernelGGL(gemv_rows_kernel, dim3((unsigned)blocks, (unsigned)nslice), dim3(64 * kGemvRowsWaves), 0,
                       s, g, ws);
    DTCF_HIP(hipGetLastError());
    if (nslice > 1) {
        hipLaunchKernelGGL(gemv_rows_reduce, dim3((unsigned)((g.rows + 255) / 256)), dim3(256), 0, s, g,
                           (const double *)ws, (uint32_t)nslice);
        DTCF_HIP(hipGetLastError());
    }
    return DTCF_OK;
}

int gemm_workspace(hipStream_t s, size_t need, double **ws) { return split_workspace(s, need, ws); }

int gemm_device_cus(int *cus) { return device_cus(cus); }

int gemm_reserve_workspace(hipStream_t s)
{
    // nsplit * sk <= 2 * slots for every configuration, and slots <= 2 per
    // CU for the tiles that split (Occ), so twice the largest tile area
    // times two slots per CU bounds every request.
    int cus = 0;
    int st = device_cus(&cus);
    if (st != DTCF_OK) return st;
    const size_t tile = std::max<size_t>({128 * 128, 192 * 96, 96 * 192, 64 * 64, 192 * 64});
    double *ws = nullptr;
    return split_workspace(s, (size_t)cus * 2 * 2 * tile * sizeof(double), &ws);
}

int gemm_f64_launch(GemmArgs g, int cfg, hipStream_t s)
{
    if (g.m <= 0 || g.n <= 0 || g.k <= 0) return DTCF_OK;
    DTCF_REQUIRE((g.m + 63) / 64 * ((g.n + 63) / 64) < (1ll << 31), "gemm: too many tiles");
    switch (cfg) {
    case kCfg128x128: return launch_cfg<128, 128, 64, 64>(g, s);
    case kCfg192x96: return launch_cfg<192, 96, 96, 48>(g, s);
    case kCfg96x192: return launch_cfg<96, 192, 48, 96>(g, s);
    case kCfg64x64: return launch_cfg<64, 64, 32, 32>(g, s);
    case kCfg192x64: return launch_cfg<192, 64, 48, 64>(g, s);
    case kCfg64x192: return launch_cfg<64, 192, 64, 48>(g, s);
    case kCfgThin: {
        const int st = thin_f64_launch(g, s);
        if (st != kThinDecline) return st;
        return launch_cfg<64, 64, 32, 32>(g, s);
    }
    default: break;
    }
    set_error("gemm: unknown tile configuration");
    return DTCF_ERR_INVALID;
}

}  // namespace dtcf
