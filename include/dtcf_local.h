/*
 * dtcf_local.h -- C ABI of the MI355X (gfx950) local-contraction engine.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * DTCF_OK (0) or a negative dtcf_status; dtcf_last_error() gives the text of
 * the last failure on the calling thread.  Device-pointer entry points take a
 * hipStream_t passed as `void *` (NULL = the default stream) and only enqueue
 * work; host-pointer entry points ("_host") stage through device buffers and
 * return after the result is back in host memory.
 *
 * Each entry point names the reference interface it replaces (file:line of
 * samyam/DTCF).  The C++ drop-in symbols with the reference's exact
 * signatures are declared in dtcf_dropin.h and forward to these.
 */
#ifndef DTCF_LOCAL_H
#define DTCF_LOCAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DTCF_OK = 0,
    DTCF_ERR_INVALID = -1,     /* bad shape, permutation, pointer or descriptor */
    DTCF_ERR_UNSUPPORTED = -2, /* e.g. DLTC transpose of a rank other than 2 or 4 */
    DTCF_ERR_HIP = -3,         /* a HIP runtime call failed */
    DTCF_ERR_NOMEM = -4,       /* device allocation failed */
    DTCF_ERR_BUDGET = -5,      /* serving: request larger than the HBM budget */
    DTCF_ERR_NOTREADY = -6     /* serving: result not complete yet */
} dtcf_status;

#define DTCF_MAX_DIMS 8

/* ------------------------------------------------------------------ */
/* Library                                                             */
/* ------------------------------------------------------------------ */
const char *dtcf_last_error(void);
/* Select the device for the calling thread and warm up the kernels. */
int dtcf_init(int device);
/* Version of this ABI (major*100 + minor). */
int dtcf_abi_version(void);

/* ------------------------------------------------------------------ */
/* 1. Strided permute-copy (device pointers)                           */
/*                                                                     */
/* Replaces the per-tile permutation of RRR transpose (helper.cc:275,  */
/* called at contraction.cc:1843,1901,2705,2825), the DLTC kernels     */
/* transpose_<perm> (DLTC/transp-kernel-seq.cc, dispatched by          */
/* DLTC/transp.cc:3-59) and, with a row pitch in ostr, the big-matrix  */
/* pack/unpack of create_big_matrix / revert_big_matrix                 */
/* (transpose_helper.cc:149-159, 393-403).                               */
/*                                                                     */
/* For every multi-index x over `dims` dims of extent ext[]:           */
/*   out[dst(t) + sum x[n]*ostr[n]] = alpha * in[src(t) + sum x[n]*istr[n]]  */
/*                                    (+ beta * old out value if beta != 0) */
/* for t in [0, ntiles); src(t) = src_off ? src_off[t] : t*src_step and  */
/* dst(t) likewise.  src_off/dst_off are device arrays.  With alpha=1,  */
/* beta=0 the copy is bit-exact.  Output elements must not alias each   */
/* other or the input.                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t dims;                 /* 1..DTCF_MAX_DIMS */
    int64_t ext[DTCF_MAX_DIMS];   /* extent of each iterated dim */
    int64_t istr[DTCF_MAX_DIMS];  /* input stride (elements) of each dim */
    int64_t ostr[DTCF_MAX_DIMS];  /* output stride (elements) of each dim */
} dtcf_strided;

int dtcf_permute_copy_f64(const double *in, double *out, const dtcf_strided *desc,
                          int64_t ntiles, const int64_t *src_off, int64_t src_step,
                          const int64_t *dst_off, int64_t dst_step, double alpha,
                          double beta, void *stream);

/* Convenience: permute `ntiles` consecutive dense row-major tiles of extents
 * w[0..dims-1].  out_src[n] = input dim that feeds output dim n (the DLTC
 * convention, s = position(y, x), DLTC/transp.h:8-17).  The RRR convention
 * pmap (input dim i -> output position pmap[i], helper.cc:289-292) converts
 * as out_src[pmap[i]] = i. */
int dtcf_permute_tiles_f64(const double *in, double *out, int dims, const int64_t *w,
                           const int32_t *out_src, int64_t ntiles, void *stream);

/* ------------------------------------------------------------------ */
/* 2. FP64 GEMM on MFMA (device pointers)                               */
/*                                                                     */
/* Replaces local_contraction (localcontract.cc:3-21) and DLTC compute  */
/* (DLTC/compute.cc:3-36): row-major C[m x n] += alpha*op(A)*op(B)      */
/* (beta is always 1, as in both reference call sites).                 */
/* op(A) = A [m x k] (lda >= k) or A^T with A [k x m] (lda >= m);       */
/* op(B) = B [k x n] (ldb >= n) or B^T with B [n x k] (ldb >= k).       */
/* These are the intended leading dimensions; the reference passes     */
/* lda = n_k and ldb = n_b whatever the flags (localcontract.cc:7-9).   */
/* ------------------------------------------------------------------ */
int dtcf_gemm_f64(int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                  int64_t ldc, void *stream);

/* A 2-D view of a tensor held as tiles: element (r, c) of a rows x cols
 * view lives at element offset
 *   rofs[r] + cofs[c] + (blk ? blk[(r / rpb) * ncb + c / cpb] : 0)
 * of the tensor's buffer.  This is how the permutation and the
 * create_big_matrix pack (transpose_helper.cc:20-171) are fused into the
 * GEMM operand loads instead of being materialised.  A view object owns
 * device copies of the tables; dtcf_view_create validates them (every
 * offset >= 0) and records which pairs of elements are adjacent so the
 * kernels can use 16-byte loads. */
typedef struct dtcf_view_s *dtcf_view;

int dtcf_view_create(int64_t rows, int64_t cols, const int64_t *rofs, const int64_t *cofs,
                     const int64_t *blk, int64_t rpb, int64_t cpb, dtcf_view *out);
int dtcf_view_destroy(dtcf_view v);
/* 1 + largest element offset the view can address (buffer size needed). */
int64_t dtcf_view_span(dtcf_view v);
/* A rectangle [r0, r0+rows) x [c0, c0+cols) of a view, sharing its device
 * tables (no allocation on the device); destroy it before its parent. */
int dtcf_view_sub(dtcf_view parent, int64_t r0, int64_t rows, int64_t c0, int64_t cols,
                  dtcf_view *out);

/* C_view[m x n] += alpha * A_view[m x k] * B_view[n x k]^T: the RRR form in
 * which both packed operands have the contraction index as columns
 * (contraction.cc:1956, kevin_dgemm(..., at=0, bt=1, 1.0)).  A NULL view
 * handle for C means C is dense row-major with leading dimension ldc. */
int dtcf_gemm_view_f64(int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                       dtcf_view av, const double *B, dtcf_view bv, double *C, dtcf_view cv,
                       int64_t ldc, void *stream);

/* Gather a view into a dense row-major matrix (create_big_matrix pack,
 * transpose_helper.cc:149-159): dst[r*ld + c] = src[view(r, c)]. */
int dtcf_view_gather_f64(const double *src, dtcf_view v, double *dst, int64_t ld, void *stream);
/* Scatter back (revert_big_matrix, transpose_helper.cc:393-403):
 * dst[view(r, c)] = src[r*ld + c]. */
int dtcf_view_scatter_f64(const double *src, int64_t ld, dtcf_view v, double *dst, void *stream);

/* GEMMs that split K keep their partial sums in a workspace of the stream
 * they run on, grown on demand (an allocation).  Under HIP graph capture it
 * cannot grow: such a GEMM returns DTCF_ERR_INVALID.  This call grows the
 * stream's workspace now to the bound any GEMM launch on this device can
 * need (151 MB on 256 CUs), so captures on that stream never allocate;
 * call it, outside capture, for every stream a graph will record GEMMs on.
 * No counterpart in the reference (its dgemm runs on the host). */
int dtcf_gemm_reserve_workspace(void *stream);

/* Tile-shape selection for the GEMM kernels: 0 = 128x128, 1 = 192x96,
 * 2 = 96x192, 3 = 64x64, 4 = 192x64, 5 = 64x192, 6 = the thin streaming
 * kernel (one side of C at most 48 wide, K a multiple of 8 and at most
 * about 200; other shapes take 64x64), -1 = automatic (autotuned table). */
int dtcf_gemm_set_config(int cfg);

/* Host-pointer variant of dtcf_gemm_f64 for the drop-in local_contraction:
 * copies A, B, C to the device, runs, copies C back. */
int dtcf_gemm_f64_host(int transa, int transb, int64_t m, int64_t n, int64_t k,
                       double alpha, const double *A, int64_t lda, const double *B,
                       int64_t ldb, double *C, int64_t ldc);

/* Host-pointer big-matrix pack / unpack (drop-in create_big_matrix /
 * revert_big_matrix, transpose_helper.cc:119-169, 372-406): the block at
 * (r, c) of the nrb x ncb block grid holds tile order[r*ncb + c], each tile
 * row-major rpb x cpb.  unpack=0: big[(r*rpb+rr)*ncb*cpb + c*cpb+cc] =
 * tiles[order*bs + rr*cpb + cc]; unpack=1 the reverse. */
int dtcf_pack_blocks_f64_host(double *tiles, int64_t ntiles, const int64_t *order, int64_t nrb,
                              int64_t ncb, int64_t rpb, int64_t cpb, double *big, int unpack);

/* Host-pointer permute of consecutive dense tiles (drop-in transpose). */
int dtcf_permute_tiles_f64_host(const double *in, double *out, int dims, const int64_t *w,
                                const int32_t *out_src, int64_t ntiles);

/* ------------------------------------------------------------------ */
/* 3. One rank's local step of an RRR contraction C += alpha * A * B    */
/*                                                                     */
/* Replaces, for one SUMMA step: contr_list (contraction.cc:2553-2572), */
/* generate_permutation_map (contraction.cc:164-282), the C pre/post   */
/* transposes and pack (contraction.cc:2692-2830), transpose_and_dgemm  */
/* (contraction.cc:1779-1976) with get_symm_permutation / compose_p_maps */
/* (contraction.cc:121-160) and create/revert_big_matrix                */
/* (transpose_helper.cc:20-409).                                        */
/* ------------------------------------------------------------------ */
/* A tensor held as RRR tiles (tensor.h:29-421): `ntiles` row-major tiles of
 * extents block_range[], tile t at tiles + t*prod(block_range), with virtual
 * tile address tile_addr[t*dims ...] (host array).  sym_group[d] is 0 for an
 * 'a'/'A' dim, 1 for 'b'/'B', 3 for a non-symmetric dim (tensor.cc:93-164);
 * tiles whose address is increasing inside a group are bounced copies whose
 * data is permuted back without a sign (contraction.cc:132-160). */
typedef struct {
    int32_t dims;
    int32_t vgrid[DTCF_MAX_DIMS];
    int32_t block_range[DTCF_MAX_DIMS];
    int32_t sym_group[DTCF_MAX_DIMS];
    int64_t ntiles;
    const int32_t *tile_addr; /* host */
    double *tiles;            /* device */
} dtcf_tileset;

enum {
    DTCF_STEP_LITERAL = 0, /* permute tiles, pack, GEMM, unpack, permute back */
    DTCF_STEP_PACKED = 1,  /* one permute+pack gather per operand, GEMM, scatter */
    DTCF_STEP_FUSED = 2    /* one GEMM reading the tiles through offset tables */
};

typedef struct dtcf_step_s *dtcf_step;

/* Index strings are comma separated, as Contraction::contract takes them
 * (e.g. "i,k,a,c").  The tilesets' device buffers must stay valid for the
 * life of the plan; the plan owns its workspaces. */
int dtcf_step_plan(const dtcf_tileset *A, const char *a_idx, const dtcf_tileset *B,
                   const char *b_idx, const dtcf_tileset *C, const char *c_idx, int mode,
                   dtcf_step *out);
/* Enqueue the step on `stream`: C tiles += alpha * contraction. */
int dtcf_step_run(dtcf_step s, double alpha, void *stream);
/* n_a, n_b, n_k per tile (generate_permutation_map) and the big-matrix
 * GEMM sizes M, N, K. */
int dtcf_step_info(dtcf_step s, int64_t *mnk_tile, int64_t *mnk_big);
int dtcf_step_destroy(dtcf_step s);
/* Measurement: record a hipEvent pair around the GEMM kernel of each of the
 * next `max_runs` runs, on the stream the kernel is launched on;
 * dtcf_step_kernel_times waits for and returns the durations (ms). */
int dtcf_step_enable_timing(dtcf_step s, int max_runs);
int dtcf_step_kernel_times(dtcf_step s, float *ms, int max_n, int *n_out);

/* ------------------------------------------------------------------ */
/* 4. DLTC tile contractions (dense tensors with uniform tiles)          */
/*                                                                     */
/* Replaces the task bodies task_mul / task_sum (DLTC/task.cc:51-137)   */
/* with the index split of Multiplication (DLTC/expr.cc:13-67), the     */
/* tile box get/acc of DLTC::Tensor (DLTC/dltc_tensor.cc:169-314), the  */
/* DLTC transposes (DLTC/transp.cc:3-59) and compute                    */
/* (DLTC/compute.cc:3-36).  Index names are DLTC IndexName values        */
/* (param.h: h1..h12 = 0..11, p1..p12 = 12..23; names below nocc are     */
/* occupied).  A tensor is dense row-major over its index ranges (O for  */
/* occupied, V for virtual), tiles of wbO / wbV per index.               */
/* ------------------------------------------------------------------ */
#define DTCF_DLTC_NIDX 24

typedef struct {
    int32_t nocc;  /* index names < nocc are occupied (12; 6 in the CHECK build) */
    int32_t O, V;  /* full ranges */
    int32_t wbO, wbV;
} dtcf_dltc_space;

typedef struct {
    int32_t dims;
    int32_t ids[DTCF_MAX_DIMS];
    double *data; /* device */
} dtcf_dltc_tensor;

typedef struct dtcf_dltc_mul_s *dtcf_dltc_mul;

/* tC += coef * tA * tB.  The plan owns device offset tables; the tensors'
 * buffers must outlive it. */
int dtcf_dltc_mul_create(const dtcf_dltc_space *space, const dtcf_dltc_tensor *tC,
                         const dtcf_dltc_tensor *tA, const dtcf_dltc_tensor *tB, double coef,
                         dtcf_dltc_mul *out);
/* GEMM sizes (M, N, K) and the m / n / k index names (IndexName order). */
int dtcf_dltc_mul_info(dtcf_dltc_mul h, int64_t *mnk, int32_t *m_ids, int32_t *n_ids,
                       int32_t *k_ids, int32_t *counts);
/* One task: the tile box [itlb[id], itub[id]) for every index name id
 * (DLTC/task.h:17-23 TaskDesc).  Whole term, one ext tile with all sum
 * tiles, or one (ext, sum) tile pair are the reference's three task
 * granularities (DLTC/exec.cc:58-110). */
int dtcf_dltc_mul_task(dtcf_dltc_mul h, const int32_t *itlb, const int32_t *itub, void *stream);
/* The whole term as one GEMM. */
int dtcf_dltc_mul_run(dtcf_dltc_mul h, void *stream);
int dtcf_dltc_mul_destroy(dtcf_dltc_mul h);
/* task_sum: tC += coef * perm(tA) over a tile box (NULL box = all tiles). */
int dtcf_dltc_sum(const dtcf_dltc_space *space, const dtcf_dltc_tensor *tC,
                  const dtcf_dltc_tensor *tA, double coef, const int32_t *itlb,
                  const int32_t *itub, void *stream);
/* DLTC::transpose of one dense tile buf_x with index names x into buf_y with
 * names y (device pointers).  DTCF_ERR_UNSUPPORTED for a rank other than 2
 * or 4, where the reference prints "transpose position error" and exits.
 * An identity permutation copies. */
int dtcf_dltc_transpose_f64(const double *buf_x, const int32_t *x, double *buf_y,
                            const int32_t *y, int dim, const dtcf_dltc_space *space,
                            void *stream);

/* ------------------------------------------------------------------ */
/* 5. Grouped contraction over uneven block pairs (one launch per term)  */
/*                                                                     */
/* Replaces, for block-sparse tensors, the per-tile loop of            */
/* transpose_and_dgemm (contraction.cc:1824-1956) and the tile-task loop */
/* of DLTC (task.cc:97-133): every (A block, B block) pair of every C    */
/* block runs inside one persistent launch, the block permutations folded */
/* into the operand loads.  Each block is addressed through an affine    */
/* descriptor: along an axis with digits d0 (most significant) .. d(nd-1) */
/* of extents ext[], index x decomposes row-major and its element offset  */
/* is sum d_i * stride[i].  A pair computes                              */
/*   C_op[rows x cols] += alpha * A_op[rows x K] * B_op[cols x K]^T      */
/* with A_op.rows = C_op.rows, B_op.rows = C_op.cols, A_op.cols = B_op.cols. */
/* Pairs of one C op are summed in the order given (deterministic).      */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t nd;          /* 1..3 */
    int32_t ext[3];
    int64_t stride[3];
} dtcf_axis;

typedef struct {
    int64_t base;        /* element offset of the block in its tensor's buffer */
    dtcf_axis rows, cols;
} dtcf_block_op;

typedef struct {
    int32_t c_op, a_op, b_op; /* indices into the op list */
} dtcf_block_pair;

typedef struct dtcf_grouped_s *dtcf_grouped;

/* cfg: 0 = 64x64 C tiles per work item, 1 = 128x128, operands read in place
 * through their descriptors; 2 = 64x64, 3 = 128x128 with every A and B block
 * first copied (one pack launch per operand, inside dtcf_grouped_run) to a
 * dense zero-padded [rows][K] slot owned by the plan, so the GEMM reads it
 * with 16-byte loads.  The packed copies make a plan unsafe to run on two
 * streams at once.  Autotuned per term. */
int dtcf_grouped_plan(const dtcf_block_op *ops, int64_t nops, const dtcf_block_pair *pairs,
                      int64_t npairs, int cfg, dtcf_grouped *out);
/* Algorithmic flops (sum of 2*m*n*k over pairs), work items, and the buffer
 * sizes (elements) the ops address in A, B and C. */
int dtcf_grouped_info(dtcf_grouped g, double *flops, int64_t *nwork, int64_t *spans);
int dtcf_grouped_run(dtcf_grouped g, double alpha, const double *A, const double *B, double *C,
                     void *stream);
int dtcf_grouped_destroy(dtcf_grouped g);

/* Block moves (new; the dense-by-irrep form of a block-sparse term): move i
 * copies the rows x cols block ops[i] of a block-sparse buffer to a dense
 * buffer at dense_base[i] + r * dense_pitch[i] + c (gather), or adds
 * alpha * that region into the block (scatter; blocks must not overlap).
 * Under an abelian group every allowed block of a term's A (B, C) sits in one
 * dense matrix per irrep of its row (column) indices, so a term becomes one
 * dense GEMM per irrep between a gather of A and B and a scatter of C. */
typedef struct dtcf_blockmove_s *dtcf_blockmove;
int dtcf_blockmove_plan(const dtcf_block_op *ops, const int64_t *dense_base,
                        const int64_t *dense_pitch, int64_t n, dtcf_blockmove *out);
/* 1 + the largest dense offset written or read. */
int dtcf_blockmove_info(dtcf_blockmove b, int64_t *dense_span);
int dtcf_blockmove_gather_f64(dtcf_blockmove b, const double *sparse, double *dense, void *stream);
int dtcf_blockmove_scatter_f64(dtcf_blockmove b, double alpha, const double *dense, double *sparse,
                               void *stream);
int dtcf_blockmove_destroy(dtcf_blockmove b);

/* Grouped dense GEMMs (new; the per-irrep GEMMs of a dense-by-irrep term in
 * one launch): for every problem p,
 *   C_p = alpha * A_p * B_p^T       (overwrite != 0: C_p is not read)
 *   C_p += alpha * A_p * B_p^T      (overwrite == 0)
 * with A_p m x k (row pitch lda), B_p n x k (ldb), C_p m x n (ldc), all
 * row-major device buffers.  Problems must be non-empty and their C regions
 * disjoint.  cfg selects the C tile (0 = 128x128, 1 = 192x96, 2 = 96x192,
 * 3 = 64x64); problems are launched in order of decreasing k, 16 at most
 * per launch. */
typedef struct {
    int64_t m, n, k;
    const double *A;
    int64_t lda;
    const double *B;
    int64_t ldb;
    double *C;
    int64_t ldc;
} dtcf_gemm_problem;
int dtcf_gemm_group_f64(const dtcf_gemm_problem *probs, int nprob, double alpha, int overwrite,
                        int cfg, void *stream);
/* The same with the operand layouts of dtcf_gemm_f64, one for the whole
 * group: C_p (+)= alpha * op(A_p) * op(B_p), op(A) = A [m x k] (transa = 0,
 * lda >= k) or A^T with A [k x m] (transa = 1, lda >= m); op(B) = B^T with
 * B [n x k] (transb = 1, ldb >= k) or B [k x n] (transb = 0, ldb >= n).
 * dtcf_gemm_group_f64 is transa = 0, transb = 1.  A block gathered from a
 * tensor whose fastest index is external lands in the k x n layout along
 * its contiguous runs (no transposed gather). */
int dtcf_gemm_group_ex_f64(const dtcf_gemm_problem *probs, int nprob, int transa, int transb,
                           double alpha, int overwrite, int cfg, void *stream);

/* ------------------------------------------------------------------ */
/* 6. Mixed precision with FP64 recombination (new; the reference is    */
/*    FP64 only) and device-side synthetic inputs                       */
/* ------------------------------------------------------------------ */
enum {
    DTCF_PREC_F32 = 1,    /* one FP32 slice, FP32 MFMA, FP32 accumulation */
    DTCF_PREC_BF16X3 = 2, /* three BF16 slices, BF16 MFMA, FP32 accumulation */
    DTCF_PREC_F16X3 = 3,  /* three FP16 slices of row-scaled values, FP32 accumulation */
    DTCF_PREC_I8 = 4      /* nslice (1..6) balanced base-256 INT8 digits of each row
                             rounded to 8*nslice-2 bits, INT8 MFMA, exact INT32
                             accumulation (Ozaki scheme) */
};
/* Device bytes needed for the slices of a rows x K operand. */
int64_t dtcf_split_bytes(int prec, int nslice, int64_t rows, int64_t K);
/* Split an FP64 operand given as a view (rows x K) into slices of
 * round_up(rows,128) x round_up(K,128) elements each, zero padded: row-major
 * (K contiguous) for F32; for BF16X3 / F16X3 / I8 fragment-major over the
 * bytes of a row (Kb = K_p times the element size), the 16 rows x 64 bytes
 * of row block rb and 64-byte K step kt forming one KiB at
 * ((rb * Kb/64 + kt) * 64 + r + 16 * (b%64 / 16)) * 16 + b%16 for byte b of
 * row r (the MFMA operand order).  rexp[round_up(rows,128)] receives the
 * per-row power-of-two exponents for F16X3 and I8. */
int dtcf_split_f64(int prec, int nslice, const double *src, dtcf_view v, void *dst, int32_t *rexp,
                   void *stream);
/* C (FP64) += alpha * sum over slice pairs (i, j), i + j <= max_level
 * (<= 0: nslice + 1), of A_i * B_j^T, each pair one low-precision GEMM whose
 * accumulators are converted to FP64 and rescaled in the epilogue.  A_slices
 * from dtcf_split_f64 over an m x k view, B_slices over an n x k view.  C is
 * a view (cv) or dense row-major with leading dimension ldc. */
/* Slice products: 1 (default) runs every product of a call in one fused
 * kernel where a specialisation exists -- I8 with (nslice, max_level) in
 * {(3,4), (4,4), (4,5), (4,6), (5,5), (5,6)} (level sums in INT32, flushed
 * to C every <= 32,640 K, a bound that relies on the top digit of
 * dtcf_split_f64 lying within +-65), BF16X3 / F16X3 at max_level 4 (the
 * default; level sums in FP32, added to C once); otherwise, or with 0, one
 * GEMM per product (I8: K < 131,072). */
int dtcf_gemm_split_set_fused(int on);
int dtcf_gemm_split_f64(int prec, int nslice, int64_t m, int64_t n, int64_t k, double alpha,
                        const void *A_slices, const int32_t *ea, const void *B_slices,
                        const int32_t *eb, int max_level, double *C, dtcf_view cv, int64_t ldc,
                        void *stream);
/* Counter-based N(0, sigma^2) (splitmix64 + Box-Muller; element index
 * start + i), uniform [lo, hi), and (X - X^{01} - X^{23} + X^{01,23}) / 2 of a
 * [d0,d1,d2,d3] tensor (d0 = d1, d2 = d3). */
int dtcf_fill_normal_f64(double *out, int64_t n, uint64_t seed, double sigma, int64_t start,
                         void *stream);
int dtcf_fill_uniform_f64(double *out, int64_t n, uint64_t seed, double lo, double hi,
                          void *stream);
int dtcf_antisymmetrize4_f64(const double *in, double *out, int64_t d0, int64_t d1, int64_t d2,
                             int64_t d3, void *stream);

/* ------------------------------------------------------------------ */
/* 7. FP64 reductions and amplitude update (new; no reference          */
/*    counterpart -- the reference stops at the contraction).          */
/*    Layouts: R, T as [v, v, o, o]; V as [o, o, v, v].                 */
/* ------------------------------------------------------------------ */
/* T[a,b,i,j] = R[a,b,i,j] / (eo[i] + eo[j] - ev[a] - ev[b]) */
int dtcf_denom_update_f64(const double *R, double *T, const double *eo, const double *ev,
                          int64_t o, int64_t v, void *stream);
/* *out = 1/4 sum_{ijab} V[i,j,a,b] * T[a,b,i,j]   (deterministic order) */
int dtcf_energy_f64(const double *V, const double *T, int64_t o, int64_t v, double *out_dev,
                    void *stream);
/* *out = sqrt(sum R^2)   (deterministic order) */
int dtcf_norm2_f64(const double *R, int64_t n, double *out_dev, void *stream);

/* ------------------------------------------------------------------ */
/* 7b. UHF CCSD amplitude update and correlation energy (new; the      */
/*     reference stops at the residuals ra_vo / rb_vo, _uhf_t1.eq:143, */
/*     242, and rab_vvoo / raa_vvoo / rbb_vvoo, _uhf_t2.eq:519,769,996)  */
/* ------------------------------------------------------------------ */
/* Dense row-major device arrays of one UHF system, the names and layouts
 * of the term lists: orbital energies eoa[Oa], eob[Ob], eva[Va], evb[Vb]
 * (the Fock diagonals), fa_ov[Oa,Va], fb_ov[Ob,Vb], vaa_oovv[Oa,Oa,Va,Va],
 * vab_oovv[Oa,Ob,Va,Vb], vbb_oovv[Ob,Ob,Vb,Vb], amplitudes ta[Va,Oa],
 * tb[Vb,Ob], taa[Va,Va,Oa,Oa], tab[Va,Vb,Oa,Ob], tbb[Vb,Vb,Ob,Ob] and
 * residuals of the same shapes. */
typedef struct {
    int64_t Oa, Ob, Va, Vb;
    const double *eoa, *eob, *eva, *evb;
    const double *fa_ov, *fb_ov;
    const double *vaa_oovv, *vab_oovv, *vbb_oovv;
    const double *ra, *rb, *raa, *rab, *rbb;
    double *ta, *tb, *taa, *tab, *tbb;
} dtcf_uhf_ccsd;

enum {
    DTCF_CC_E = 0,         /* correlation energy E1 + Eaa + Eab + Ebb */
    DTCF_CC_E1 = 1,        /* sum fa_ov.ta + sum fb_ov.tb */
    DTCF_CC_EAA = 2,       /* 1/4 sum vaa_oovv (taa + 2 ta ta) */
    DTCF_CC_EAB = 3,       /* sum vab_oovv (tab + ta tb) */
    DTCF_CC_EBB = 4,       /* 1/4 sum vbb_oovv (tbb + 2 tb tb) */
    DTCF_CC_R1 = 5,        /* |R1| over both spins */
    DTCF_CC_R2 = 6,        /* |R2| over the three blocks */
    DTCF_CC_DT = 7,        /* |T_new - T| over every amplitude */
    DTCF_CC_NOUT = 8
};

/* update != 0: t += r / D in place (D_ai = e_i - e_a, D_abij =
 * ((e_i + e_j) - e_a) - e_b with each index's spin), then the energy of the
 * updated amplitudes; update == 0: the energy of the amplitudes as they are
 * (residuals not read, norms 0).  Writes DTCF_CC_NOUT doubles to out_dev
 * (device).  Deterministic: the sums do not depend on timing. */
int dtcf_uhf_ccsd_step_f64(const dtcf_uhf_ccsd *d, int update, double *out_dev, void *stream);

/* The same step without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * partial sums live in a fixed device array of the library (147 KiB per
 * device, held until the process ends, enough for any system size the
 * ranges allow), resolved by dtcf_init or on the first call on a device.
 * Every call on one device shares that array: calls that overlap on one
 * device (other streams, other threads) race and must be serialised by the
 * caller.  Results are bitwise those of dtcf_uhf_ccsd_step_f64. */
int dtcf_uhf_ccsd_step_f64_noalloc(const dtcf_uhf_ccsd *d, int update, double *out_dev,
                                   void *stream);

/* ------------------------------------------------------------------ */
/* 8. Serving: device-resident CC iteration programs under an HBM      */
/*    budget, continuous batching over streams.  Replaces DLTC's task  */
/*    pool and tile cache (DLTC/task.cc:24-137 task_mul/task_sum and   */
/*    the Scioto pool, DLTC/cache.h:23-133 LRU tile cache,             */
/*    DLTC/exec.cc:58-110 task granularity) for one GPU.               */
/* ------------------------------------------------------------------ */
typedef struct dtcf_server_s *dtcf_server;
typedef struct dtcf_program_s *dtcf_program;

/* Opens a server on the current device: one device arena of `budget`
 * bytes (resident tensors and request blocks are carved out of it) and
 * `nslots` (1..16) non-blocking streams of default priority.  Each slot
 * also keeps a GEMM tail-split workspace (under 80 MB) outside the arena. */
int dtcf_server_open(int64_t budget, int nslots, dtcf_server *out);
/* Waits for every admitted request (queued requests that can no longer be
 * admitted fail with DTCF_ERR_BUDGET), then frees everything. */
int dtcf_server_close(dtcf_server s);
/* Resident tensor storage from the arena (zeroed).  Freed by
 * dtcf_server_free or at close. */
int dtcf_server_alloc(dtcf_server s, int64_t bytes, void **dev_ptr);
int dtcf_server_free(dtcf_server s, void *dev_ptr);
/* bytes: arena size, in use by resident tensors, reserved by admitted
 * requests; peak_reserved: high-water mark of the request reservations. */
int dtcf_server_usage(dtcf_server s, int64_t *budget, int64_t *resident, int64_t *reserved,
                      int64_t *peak_reserved);
/* Failure injection (testing the failure path; no reference counterpart):
 * after `after` more statements have been issued, the next one fails on the
 * host with DTCF_ERR_HIP without launching anything; its request issues
 * nothing more, finishes once its statements in flight complete, reports
 * the failure from dtcf_request_wait and releases its block.  -1 cancels. */
int dtcf_server_inject_failure(dtcf_server s, int64_t after);

/* A program is the statement list of one term file over one system's
 * extents (a CC residual evaluation).  Arrays are dense row-major FP64.
 * Index lists are comma-separated index names, e.g. "p1a,h2a".
 * prec: 0 = FP64 MFMA, else a DTCF_PREC_* mode (nslice, max_level as in
 * dtcf_gemm_split_f64) for every contraction C[m x n] with m n / (m + n) >=
 * mixed_min (default 256: splitting both operands costs less HBM time than
 * the slices save; thinner contractions run FP64); sums stay FP64. */
enum { DTCF_ARR_INPUT = 0, DTCF_ARR_TEMP = 1, DTCF_ARR_OUTPUT = 2 };
int dtcf_program_create(dtcf_server s, int prec, int nslice, int max_level, dtcf_program *out);
/* Before the first statement: the m n / (m + n) from which a contraction
 * takes the program's mixed precision (0: every contraction). */
int dtcf_program_set_mixed_min(dtcf_program p, double h);
int dtcf_program_array(dtcf_program p, const char *name, int ndim, const int64_t *dims, int role);
/* Mixed programs, after declaring input `name` and before the first
 * statement: on != 0 makes finalize split that input, wherever a mixed
 * contraction reads it, once (from its bound contents at that time) into the
 * server's arena, counted as resident until the program is destroyed (one
 * slice set per contraction and operand view); requests then split only the
 * other operands.  For inputs that do not change between requests (the
 * integrals of a CC iteration, not its amplitudes); such an input cannot be
 * rebound after finalize. */
int dtcf_program_set_resident(dtcf_program p, const char *name, int on);
/* Inputs are resident tensors the caller owns (e.g. from dtcf_server_alloc).
 * After finalize an input may be rebound while no request of the program is
 * in flight, unless it has resident slices. */
int dtcf_program_bind(dtcf_program p, const char *name, const double *dev_ptr);
/* C[c_idx] = alpha * sum_{contracted} A[a_idx] * B[b_idx] */
int dtcf_program_contract(dtcf_program p, const char *c, const char *c_idx, const char *a,
                          const char *a_idx, const char *b, const char *b_idx, double alpha);
/* C[c_idx] = sum_t coef[t] * X_t[idx_t]  (each idx_t a permutation of c_idx) */
int dtcf_program_sum(dtcf_program p, const char *c, const char *c_idx, int nterms,
                     const char *const *names, const char *const *idx, const double *coefs);
/* Plans every statement (operand views, sum strides), the dependency graph
 * and the static placement of temporaries inside one request block: each
 * temporary in bytes of its own (DTCF_LAYOUT_SEPARATE: only data
 * dependencies order the statements) or first fit reusing dead temporaries'
 * bytes with write-after-read dependencies (DTCF_LAYOUT_REUSE: the smallest
 * block).  DTCF_LAYOUT_AUTO (default) takes SEPARATE when its block is within
 * 1.5x (or 256 MiB) of REUSE's.  set_layout before finalize; get_layout
 * after it reports the layout used. */
enum { DTCF_LAYOUT_AUTO = 0, DTCF_LAYOUT_REUSE = 1, DTCF_LAYOUT_SEPARATE = 2 };
int dtcf_program_set_layout(dtcf_program p, int layout);
int dtcf_program_finalize(dtcf_program p);
int dtcf_program_get_layout(dtcf_program p, int *layout);
/* block_bytes: arena bytes one request reserves; flops: 2*m*n*k summed over
 * contractions plus one per summed element. */
int dtcf_program_info(dtcf_program p, int64_t *block_bytes, int64_t *nstatements, double *flops,
                      int64_t *noutputs);
/* The plan of statement i of a finalized program (no reference counterpart:
 * the reference has no planner to inspect): out[DTCF_PLAN_NOUT] receives
 * the statement kind (0 contraction, 1 sum); for a contraction the GEMM
 * m, n, k (operands possibly swapped), DTCF_PLAN_* flags (operand access:
 * column or row pairs contiguous, packed into a dense copy first, resident
 * slices; mixed precision; C written densely) and the FP64 C tile
 * configuration (0 128x128, 1 192x96, 2 96x192, 3 64x64, 4 192x64,
 * 5 64x192, 6 thin streaming kernel; -1 when mixed);
 * for a sum the number of terms and whether it stages planes in LDS. */
enum {
    DTCF_PLAN_KIND = 0, DTCF_PLAN_M = 1, DTCF_PLAN_N = 2, DTCF_PLAN_K = 3, DTCF_PLAN_FLAGS = 4,
    DTCF_PLAN_TILE = 5, DTCF_PLAN_NTERMS = 6, DTCF_PLAN_NOUT = 8
};
enum {
    DTCF_PLAN_SWAP = 1, DTCF_PLAN_MIXED = 2, DTCF_PLAN_PACK_A = 4, DTCF_PLAN_PACK_B = 8,
    DTCF_PLAN_A_COLPAIRS = 16, DTCF_PLAN_A_ROWPAIRS = 32, DTCF_PLAN_B_COLPAIRS = 64,
    DTCF_PLAN_B_ROWPAIRS = 128, DTCF_PLAN_C_DENSE = 256, DTCF_PLAN_RES_A = 512,
    DTCF_PLAN_RES_B = 1024, DTCF_PLAN_SUM_PLANE = 2048
};
int dtcf_program_statement_plan(dtcf_program p, int64_t i, int64_t *out);
int dtcf_program_destroy(dtcf_program p);

/* Requests: one evaluation of a finalized program.  `outputs` holds one
 * device buffer per DTCF_ARR_OUTPUT array, in declaration order; they must
 * stay valid until the request completes.  Requests are admitted in
 * arrival order as their block fits in the arena (later, smaller requests
 * may pass a blocked one, at most 8 times); the statements of all admitted
 * requests are issued to free slots as their inputs complete, from the
 * request with the least total work first (shortest job first, then arrival
 * order).  A request's results do not depend on what runs beside it. */
int dtcf_request_submit(dtcf_server s, dtcf_program p, double *const *outputs, int64_t *id);
/* DTCF_OK when complete, DTCF_ERR_NOTREADY on timeout, another status if
 * the request failed.  timeout_ms < 0 waits forever. */
int dtcf_request_wait(dtcf_server s, int64_t id, double timeout_ms);
typedef struct {
    int32_t state;                      /* 0 queued, 1 admitted, 2 done, 3 failed */
    int64_t ntasks;
    double t_submit, t_admit, t_done;   /* microseconds since dtcf_server_open */
    double flops;
} dtcf_request_info;
int dtcf_request_info_get(dtcf_server s, int64_t id, dtcf_request_info *info);
/* Per statement: when its inputs were complete (ready), when it was issued
 * to a slot and when its completion was observed (microseconds). */
int dtcf_request_task_times(dtcf_server s, int64_t id, double *t_ready, double *t_issue,
                            double *t_done, int64_t n);
/* Forgets a completed request's record. */
int dtcf_request_release(dtcf_server s, int64_t id);

#ifdef __cplusplus
}
#endif
#endif /* DTCF_LOCAL_H */
