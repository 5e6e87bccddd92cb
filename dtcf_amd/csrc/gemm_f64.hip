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
// walked in steps of BK=16 doubles through a double-buffered LDS stage with
// register prefetch of the next step, one barrier per step.  LDS rows are
// padded (stride 18 doubles for k-contiguous operands, 17 for row-contiguous
// ones) so both the global->LDS stores and the fragment reads are free of
// bank conflicts.  Workgroup ids are remapped so that consecutive C tiles of
// a row group land on the same XCD (its own L2).
//
// MFMA operand layout (16x16x4 f64, wave64): lane l holds A[l%16][l/16] and
// B[l/16][l%16]; the result holds D[l/16 + 4*r][l%16] in element r (checked
// on the device by tests/test_gpu_core.py::test_mfma_layout_single_block).
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "gemm_f64.hpp"

namespace dtcf {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int BK = 16;
constexpr int NT = 256;

template <int MODE>
struct LdsStride {
    static constexpr int v = (MODE == kDK || MODE == kVK) ? 18 : 17;
};

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

// Loads one BROWS x BK operand tile (rows of the view x k) into registers
// and stores it to LDS laid out [row][k] with the padded stride.
template <int BROWS, int MODE>
struct TileLoader {
    static constexpr bool KC = (MODE == kDK || MODE == kVK);
    static constexpr int PAIRS = BROWS * BK / 2;
    static constexpr int NP = (PAIRS + NT - 1) / NT;
    static constexpr int S = LdsStride<MODE>::v;
    d2 reg[NP];

    __device__ inline void load(const OpArgs &o, const double *__restrict__ base, int64_t row0,
                                int64_t k0, int64_t rows, int64_t K, int tid)
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int p = tid + NT * i;
            d2 v = {0.0, 0.0};
            if (p < PAIRS) {
                int64_t r0, c0, r1, c1;
                if constexpr (KC) {
                    const int row = p / (BK / 2), kp = (p % (BK / 2)) * 2;
                    r0 = r1 = row0 + row;
                    c0 = k0 + kp;
                    c1 = c0 + 1;
                } else {
                    const int kk = p / (BROWS / 2), rp = (p % (BROWS / 2)) * 2;
                    c0 = c1 = k0 + kk;
                    r0 = row0 + rp;
                    r1 = r0 + 1;
                }
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
            }
            reg[i] = v;
        }
    }

    __device__ inline void store(double *lds, int tid) const
    {
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const int p = tid + NT * i;
            if (p < PAIRS) {
                if constexpr (KC) {
                    const int row = p / (BK / 2), kp = (p % (BK / 2)) * 2;
                    *reinterpret_cast<d2 *>(lds + row * S + kp) = reg[i];
                } else {
                    const int kk = p / (BROWS / 2), rp = (p % (BROWS / 2)) * 2;
                    lds[rp * S + kk] = reg[i].x;
                    lds[(rp + 1) * S + kk] = reg[i].y;
                }
            }
        }
    }
};

// Workgroups per CU the register budget is sized for: two for tiles whose
// accumulators fit 256 VGPRs, one (with AGPRs for the accumulators) above.
template <int BM, int BN>
struct Occ {
    static constexpr int v = (BM * BN <= 128 * 128) ? 2 : 1;
};

template <int BM, int BN, int WM, int WN, int AMODE, int BMODE, bool CVIEW>
__global__ __launch_bounds__(NT, (Occ<BM, BN>::v)) void gemm_f64_kernel(GemmArgs g)
{
    static_assert((BM / WM) * (BN / WN) == 4, "four waves per workgroup");
    constexpr int FM = WM / 16, FN = WN / 16;
    constexpr int SA = LdsStride<AMODE>::v, SB = LdsStride<BMODE>::v;
    constexpr int LDS_A = BM * SA, LDS_B = BN * SB;
    extern __shared__ double smem[];
    // Stage s occupies [s*(LDS_A+LDS_B), ...): A tile first, then B tile.
    constexpr int STAGE = LDS_A + LDS_B;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave / (BN / WN), wn = wave % (BN / WN);

    // XCD-aware remap then grouped (GROUP_M tile rows) swizzle.
    uint32_t pid = blockIdx.x;
    const uint32_t T = g.tiles_m * g.tiles_n;
    if ((T & 7u) == 0) pid = (pid & 7u) * (T >> 3) + (pid >> 3);
    const uint32_t GROUP = 8;
    const uint32_t width = GROUP * g.tiles_n;
    const uint32_t gid = pid / width;
    const uint32_t first = gid * GROUP;
    const uint32_t gsz = min(g.tiles_m - first, GROUP);
    const uint32_t tm = first + (pid % width) % gsz;
    const uint32_t tn = (pid % width) / gsz;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;

    TileLoader<BM, AMODE> la;
    TileLoader<BN, BMODE> lb;
    d4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

    const int64_t K = g.k;
    const int64_t nk = (K + BK - 1) / BK;
    la.load(g.a, g.A, m0, 0, g.m, K, tid);
    lb.load(g.b, g.B, n0, 0, g.n, K, tid);
    la.store(smem, tid);
    lb.store(smem + LDS_A, tid);
    __syncthreads();

    const int fr = lane & 15, fk = lane >> 4;
    for (int64_t kt = 0; kt < nk; kt++) {
        const int cur = (int)(kt & 1);
        const bool more = kt + 1 < nk;
        if (more) {
            la.load(g.a, g.A, m0, (kt + 1) * BK, g.m, K, tid);
            lb.load(g.b, g.B, n0, (kt + 1) * BK, g.n, K, tid);
        }
        const double *a_s = smem + cur * STAGE + (wm * WM + fr) * SA + fk;
        const double *b_s = smem + cur * STAGE + LDS_A + (wn * WN + fr) * SB + fk;
#pragma unroll
        for (int ks = 0; ks < BK / 4; ks++) {
            double af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; i++) af[i] = a_s[i * 16 * SA + ks * 4];
#pragma unroll
            for (int j = 0; j < FN; j++) bf[j] = b_s[j * 16 * SB + ks * 4];
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            la.store(smem + (cur ^ 1) * STAGE, tid);
            lb.store(smem + (cur ^ 1) * STAGE + LDS_A, tid);
        }
        __syncthreads();
    }

    // Epilogue: C += alpha * acc  (beta = 1, as every reference call site).
    const double alpha = g.alpha;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) {
            const int64_t col = n0 + wn * WN + j * 16 + fr;
            if (col >= g.n) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm * WM + i * 16 + fk + 4 * r;
                if (row >= g.m) continue;
                int64_t addr;
                if constexpr (CVIEW) addr = op_addr<kVK>(g.c, row, col);
                else addr = row * g.c.ld + col;
                g.C[addr] += alpha * acc[i][j][r];
            }
        }
}

template <int BM, int BN, int WM, int WN>
constexpr size_t lds_bytes(int am, int bm)
{
    return (size_t)2 * (BM * ((am == kDK || am == kVK) ? 18 : 17) +
                        BN * ((bm == kDK || bm == kVK) ? 18 : 17)) *
           sizeof(double);
}

template <int BM, int BN, int WM, int WN, int AM, int BMD, bool CV>
int launch_one(const GemmArgs &g, hipStream_t s)
{
    auto kern = gemm_f64_kernel<BM, BN, WM, WN, AM, BMD, CV>;
    const size_t lds = lds_bytes<BM, BN, WM, WN>(AM, BMD);
    static std::once_flag once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(once, [&] {
        attr_err = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    DTCF_HIP(attr_err);
    const unsigned grid = g.tiles_m * g.tiles_n;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, g);
    DTCF_HIP(hipGetLastError());
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

}  // namespace

int gemm_f64_launch(GemmArgs g, int cfg, hipStream_t s)
{
    if (g.m <= 0 || g.n <= 0 || g.k <= 0) return DTCF_OK;
    DTCF_REQUIRE((g.m + 255) / 256 * ((g.n + 255) / 256) < (1ll << 30), "gemm: too many tiles");
    switch (cfg) {
    case kCfg128x128: return launch_cfg<128, 128, 64, 64>(g, s);
    case kCfg192x96: return launch_cfg<192, 96, 96, 48>(g, s);
    case kCfg128x144: return launch_cfg<128, 144, 32, 144>(g, s);
    case kCfg64x64: return launch_cfg<64, 64, 32, 32>(g, s);
    default: break;
    }
    set_error("gemm: unknown tile configuration");
    return DTCF_ERR_INVALID;
}

}  // namespace dtcf
