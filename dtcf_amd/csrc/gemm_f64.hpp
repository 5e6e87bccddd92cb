// Internal launch interface of the FP64 MFMA GEMM (gemm_f64.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace dtcf {

// Operand addressing modes (rows of the view x k):
//   kDK : dense, (r, c) at r*ld + c   -- k contiguous
//   kDMN: dense, (r, c) at c*ld + r   -- rows contiguous
//   kVK : offset tables, pairs (c, c+1) contiguous
//   kVMN: offset tables, pairs (r, r+1) contiguous
enum { kDK = 0, kDMN = 1, kVK = 2, kVMN = 3 };

// Tile configurations (BM x BN), selectable per launch by the autotuner.
// 192x64 and 64x192 serve thin results (one side at most 64) with long K:
// a 64-wide tile of the thin side, three times as deep on the other, so
// the thin operand is re-read by a third as many tiles as with 64 x 64.
// kCfgThin is the streaming kernel for thin results (one side at most 48,
// or 192 with K <= 48) with short K (thin_f64_launch); shapes it does not
// take run 64 x 64 tiles.  The grouped launch takes the first four.
enum {
    kCfg128x128 = 0,
    kCfg192x96 = 1,
    kCfg96x192 = 2,
    kCfg64x64 = 3,
    kCfg192x64 = 4,
    kCfg64x192 = 5,
    kCfgThin = 6,
    kCfgCount = 7,
    kCfgGroupCount = 4
};
// The C tile configuration dtcf_gemm_* uses for an m x n x k GEMM.
int pick_config(int64_t m, int64_t n, int64_t k);

struct OpArgs {
    int64_t ld;
    const int64_t *rofs;
    const int64_t *cofs;
    const int64_t *blk;
    FastDiv div_rpb, div_cpb;
    int64_t ncb;
    int32_t vec2;  // 1: (r,c),(r,c+1) [k-mode] or (r,c),(r+1,c) [mn-mode] are adjacent
                   //    and 16-byte aligned wherever both are in range
};

struct GemmArgs {
    int64_t m, n, k;
    double alpha;
    const double *A;
    const double *B;
    double *C;
    OpArgs a, b, c;
    int32_t amode, bmode, cview;
    uint32_t tiles_m, tiles_n;
    // Tail split: workgroups [0, t_full) own whole tiles; the remaining
    // tiles t_full.. are split over K into `sk` slices of `kchunk` K steps
    // each, written to `ws` and summed in slice order by a second kernel.
    uint32_t t_full, sk;
    int64_t kchunk;
    double *ws;
    // 0: C += alpha * A * B (every reference call site); 1: C = alpha * A * B
    // (C is neither cleared first nor read)
    int32_t overwrite;
};

int gemm_f64_launch(GemmArgs g, int cfg, hipStream_t s);

// Up to kGroupMax independent dense GEMMs C_p (+)= alpha * op(A_p) * op(B_p)
// in one launch, every problem with the same operand layouts: A_p m x k
// (kDK) or k x m (kDMN), B_p n x k (kDK) or k x n (kDMN); C_p row-major.
// tiles_m, tiles_n, wg0, t_whole, tail0 and kchunk are filled in by the
// launch: problem p's tiles [0, t_whole) run whole, its tiles [t_whole,
// tiles) are tail tiles tail0.. of the launch, each cut into `sk` K slices
// of kchunk steps (the tail split of a single GEMM, over the group).
constexpr int kGroupMax = 16;
struct GemmGroupProb {
    int64_t m, n, k;
    const double *A;
    const double *B;
    double *C;
    int64_t lda, ldb, ldc;
    int32_t vec2a, vec2b;
    uint32_t tiles_m, tiles_n, wg0;
    uint32_t t_whole, tail0;
    int64_t kchunk;
};
struct GemmGroupArgs {
    double alpha;
    int32_t overwrite;
    int32_t np;
    uint32_t wg_whole, sk;  // workgroups of whole tiles; K slices per tail tile
    double *ws;             // tail slices, [tail tile][slice][BM][BN]
    GemmGroupProb p[kGroupMax];
};

int gemm_f64_group_launch(const GemmGroupArgs &ga, int cfg, int amode, int bmode, hipStream_t s);

// Matrix-vector products inside the view GEMM: y(r) (+)= alpha * sum_c
// M(r, c) x(c) over the rows of a view M whose column pairs are adjacent and
// 16-byte aligned (no block table), x a 1 x k view, y the single column
// (ycol = 1: y(r) = Y(r, 0)) or single row (ycol = 0: y(r) = Y(0, r)) of a
// dense Y (pitch ldy) or of a view.  k <= kGemvMaxK (x is staged in LDS).
constexpr int64_t kGemvMaxK = 20480;
struct GemvArgs {
    int64_t rows, k;
    double alpha;
    const double *M;
    const int64_t *mrofs, *mcofs;
    const double *X;
    OpArgs xv;
    double *Y;
    OpArgs yv;
    int32_t yview, ycol, overwrite;
    int64_t ldy;
};
int gemv_f64_launch(const GemvArgs &g, hipStream_t s);
// The same product when M has its row pairs adjacent instead (rofs[r + 1] =
// rofs[r] + 1 for even r, offsets even, rows even, M 16-byte aligned): any k.
int gemv_rows_f64_launch(const GemvArgs &g, hipStream_t s);

// Thin products: one side of C at most kThinMaxWidth wide (or at most
// kThinMaxWide with K <= kThinWideK), the other long (rows), K a multiple
// of 8 and short enough that the thin operand's K x width fragments and the
// offsets fit kThinLdsBytes of LDS.  thin_shape is the shape test; the
// launch also needs the long operand's pairs adjacent and 16-byte aligned
// (vec2), no block table on it or on a C view, and returns kThinDecline
// otherwise.
constexpr int64_t kThinMaxWidth = 48;
constexpr int64_t kThinMaxWide = 192;
constexpr int64_t kThinWideK = 48;
constexpr int64_t kThinLdsBytes = 80 * 1024;
constexpr int kThinDecline = 1;
bool thin_shape(int64_t m, int64_t n, int64_t k, int64_t min_rows);
int thin_f64_launch(const GemmArgs &g, hipStream_t s);

// Allocates the tail-split workspace of stream `s` at the largest size any
// GEMM on it can need, so later launches on `s` never allocate.
int gemm_reserve_workspace(hipStream_t s);
// The split workspace of stream `s` (grown to `need` bytes if smaller; work
// on one stream shares it in stream order), and the device's CU count.
int gemm_workspace(hipStream_t s, size_t need, double **ws);
int gemm_device_cus(int *cus);

// dtcf_gemm_view_f64 with C = alpha * A * B instead of C += (capi.hip).
int gemm_view_overwrite(int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                        dtcf_view av, const double *B, dtcf_view bv, double *C, dtcf_view cv,
                        int64_t ldc, hipStream_t stream);

// Slice products (mixed.hip): dtcf_gemm_split_f64 with C = alpha * A * B^T
// instead of C += when overwrite is set, which needs a fused kernel for
// (prec, nslice, level): gemm_split_fused.
bool gemm_split_fused(int prec, int ns, int lev);
int gemm_split_f64_impl(int prec, int nslice, int64_t m, int64_t n, int64_t k, double alpha,
                        const void *A_slices, const int32_t *ea, const void *B_slices,
                        const int32_t *eb, int max_level, double *C, dtcf_view cv, int64_t ldc,
                        void *stream, int overwrite);

}  // namespace dtcf
