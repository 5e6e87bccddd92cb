/*
 * dtcf_oracle.c -- CPU restatement of DTCF's per-rank local contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (dtcf_amd/, include/)
 * links, loads or calls this file.  It is loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and only
 * ever as the checker or as the timed CPU baseline, never as the measured GPU
 * path.
 *
 * Every function restates the reference's algorithm in plain C and cites the
 * reference file:line it follows.  How each function is pinned (against what
 * the reference's own tests/fixtures hold, against the reference kernels
 * compiled from source into oracle/_ref, or against closed forms / brute
 * force / library routines) is listed in the header comment of each block.
 *
 * Conventions: all tensors row-major, last index fastest.  Extents and
 * offsets are int64 (the reference uses int32, which overflows at o=64,
 * v=256 dense scale).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXD 8

/* ------------------------------------------------------------------------
 * 1. RRR in-tile permutation  (helper.cc:275-522)
 *
 * Convention: input dim i goes to output position pmap[i]; the output extent
 * at position p is widths[inv[p]] (helper.cc:289-320).  The reference has
 * three code paths and this restates all three:
 *   (a) every dim fixed            -> one memcpy           (helper.cc:331-334)
 *   (b) trailing dims fixed        -> memcpy runs of the
 *       trailing block over the leading odometer         (helper.cc:337-410)
 *   (c) otherwise                  -> element loop over (new-last x old-last)
 *       blocked by VECTOR_WIDTH = largest d<=32 dividing both innermost
 *       extents                                          (helper.cc:412-520)
 * Pinned by: brute-force index map on asymmetric extents (tests), composition
 * and inverse invariants.  helper.cc itself includes <mpi.h>, absent here, so
 * it is not compiled into oracle/_ref.
 * dims == 1 reads initial_offsets[-1] in the reference (helper.cc:317); here
 * it is handled as a plain copy.
 * ---------------------------------------------------------------------- */
void orc_rrr_transpose(const double *in, double *out, const int *pmap,
                       int dims, const int *w)
{
    int64_t inv[ORC_MAXD], ioff[ORC_MAXD], poff[ORC_MAXD];
    int d;
    if (dims == 1) {
        memcpy(out, in, (size_t)w[0] * sizeof(double));
        return;
    }
    for (d = 0; d < dims; d++) inv[pmap[d]] = d;
    int nlast = 0;
    for (d = dims - 1; d >= 0; d--) {
        if (pmap[d] == d) nlast++;
        else break;
    }
    ioff[dims - 1] = 1;
    ioff[dims - 2] = w[dims - 1];
    poff[dims - 1] = 1;
    poff[dims - 2] = w[inv[dims - 1]];
    for (d = dims - 3; d >= 0; d--) {
        ioff[d] = ioff[d + 1] * w[d + 1];
        poff[d] = poff[d + 1] * w[inv[d + 1]];
    }
    if (nlast == dims) { /* (a) */
        memcpy(out, in, (size_t)(ioff[0] * w[0]) * sizeof(double));
        return;
    }
    if (nlast > 0) { /* (b) */
        int nn = dims - nlast;
        int64_t total = 1;
        for (d = 0; d < nn; d++) total *= w[d];
        int64_t run = ioff[nn - 1];
        int64_t last = w[nn - 1];
        for (int64_t ci = 0; ci < total; ci += last) {
            int64_t t = ci / last;
            int64_t idx[ORC_MAXD];
            for (d = nn - 2; d >= 0; d--) {
                idx[d] = t % w[d];
                t /= w[d];
            }
            int64_t src = 0, dst = 0;
            for (d = 0; d < nn - 1; d++) {
                src += idx[d] * ioff[d];
                dst += idx[d] * poff[pmap[d]];
            }
            for (int64_t l = 0; l < last; l++) {
                memcpy(out + dst + l * poff[pmap[nn - 1]],
                       in + src + l * ioff[nn - 1], (size_t)run * sizeof(double));
            }
        }
        return;
    }
    /* (c) */
    int nonlast[ORC_MAXD];
    int cnt = 0;
    for (d = 0; d < dims - 1; d++)
        if (d != inv[dims - 1]) nonlast[cnt++] = d;
    int64_t total = 1;
    for (d = 0; d < dims - 2; d++) total *= w[nonlast[d]];
    int ind = (int)inv[dims - 1];
    int vw = 1;
    for (int v = 32; v >= 1; v--)
        if (w[dims - 1] % v == 0 && w[ind] % v == 0) { vw = v; break; }
    for (int64_t ci = 0; ci < total; ci++) {
        int64_t t = ci, src = 0, dst = 0;
        for (d = dims - 3; d >= 0; d--) {
            int k = nonlast[d];
            int64_t x = t % w[k];
            t /= w[k];
            src += x * ioff[k];
            dst += x * poff[pmap[k]];
        }
        for (int64_t nli = 0; nli < w[ind]; nli += vw)
            for (int64_t oli = 0; oli < w[dims - 1]; oli += vw)
                for (int64_t a = nli; a < nli + vw; a++) {
                    int64_t so = src + a * ioff[ind];
                    int64_t dd = dst + a * poff[pmap[ind]];
                    for (int64_t b = oli; b < oli + vw; b++)
                        out[dd + b * poff[pmap[dims - 1]]] = in[so + b * ioff[dims - 1]];
                }
    }
}

/* ------------------------------------------------------------------------
 * 2. DLTC transpose  (transp.h:8-17 position, transp.cc:3-59 dispatcher,
 *    transp-kernel-seq.cc kernels).
 *
 * position(x, y)[i] = index of x[i] inside y.  transpose(buf_x, x, buf_y, y)
 * uses s = position(y, x): output dim n is fed by input dim s[n]; widths are
 * those of x (getWB(x), define.h:59-65).  Only 2-D and 4-D are dispatched;
 * anything else is "transpose position error" + exit(0) in the reference
 * (transp.cc:48-51): here it returns -1.
 * Identity: the reference executes `buf_y = buf_x`, which leaves the caller's
 * output unwritten (transp.cc:16,24).  q2_literal=1 reproduces that; 0 copies.
 * Pinned by: the reference's own 26 kernels compiled from source into
 * oracle/_ref (tests compare all of them), and brute force.
 * ---------------------------------------------------------------------- */
void orc_dltc_position(const int *x, const int *y, int n, int *pos)
{
    int tab[64];
    for (int i = 0; i < 64; i++) tab[i] = 0;
    for (int i = 0; i < n; i++) tab[y[i]] = i;
    for (int i = 0; i < n; i++) pos[i] = tab[x[i]];
}

/* out[n-th output dim iterates input dim s[n]]; w = input extents */
void orc_permute_by_source(const double *in, double *out, const int *s,
                           int dims, const int *w)
{
    int64_t istride[ORC_MAXD], ow[ORC_MAXD];
    int64_t total = 1;
    istride[dims - 1] = 1;
    for (int d = dims - 2; d >= 0; d--) istride[d] = istride[d + 1] * w[d + 1];
    for (int n = 0; n < dims; n++) { ow[n] = w[s[n]]; total *= ow[n]; }
    for (int64_t o = 0; o < total; o++) {
        int64_t t = o, src = 0;
        for (int n = dims - 1; n >= 0; n--) {
            int64_t x = t % ow[n];
            t /= ow[n];
            src += x * istride[s[n]];
        }
        out[o] = in[src];
    }
}

int orc_dltc_transpose(const double *buf_x, const int *x, double *buf_y,
                       const int *y, int dim, const int *wx, int q2_literal)
{
    int s[ORC_MAXD];
    if (dim != 2 && dim != 4) return -1;
    orc_dltc_position(y, x, dim, s);
    int ident = 1;
    for (int i = 0; i < dim; i++) if (s[i] != i) ident = 0;
    if (ident) {
        if (!q2_literal) {
            int64_t n = 1;
            for (int i = 0; i < dim; i++) n *= wx[i];
            memcpy(buf_y, buf_x, (size_t)n * sizeof(double));
        }
        return 0;
    }
    orc_permute_by_source(buf_x, buf_y, s, dim, wx);
    return 0;
}

/* ------------------------------------------------------------------------
 * 3. Block_Grid  (block_grid.cc:22-37 rank, :52-67 address)
 * rank = sum_i addr[i] * prod_{j>i} grid[j].  Pinned by closed form.
 * ---------------------------------------------------------------------- */
int64_t orc_block_rank(const int *addr, int dims, const int *grid)
{
    int64_t x = 1, r = 0;
    for (int i = dims - 1; i >= 0; i--) {
        r += (int64_t)addr[i] * x;
        x *= grid[i];
    }
    return r;
}

void orc_block_addr(int64_t r, int dims, const int *grid, int *addr)
{
    int64_t x = 1;
    for (int i = 0; i < dims; i++) x *= grid[i];
    for (int i = 0; i < dims; i++) {
        x /= grid[i];
        addr[i] = (int)(r / x);
        r -= (int64_t)addr[i] * x;
    }
}

/* ------------------------------------------------------------------------
 * 4. Symmetry bounce permutation and composition
 *    (contraction.cc:132-160 get_symm_permutation, :121-128 compose_p_maps)
 * For i<j, addr[i] < addr[j] in the same group g<=1 (from the tensor string:
 * a/A -> 0, b/B -> 1, others non-symmetric): swap i<->j, first j only.
 * No sign is applied anywhere (DTCF symmetry is sign-free).
 * composed[i] = map2[map1[i]].  Pinned by brute force in tests.
 * ---------------------------------------------------------------------- */
void orc_symm_permutation(const int *addr, const int *sg, int dims, int *sym)
{
    for (int i = 0; i < dims; i++) sym[i] = i;
    for (int i = 0; i < dims; i++)
        for (int j = i + 1; j < dims; j++)
            if (addr[i] < addr[j] && sg[i] == sg[j] && sg[i] <= 1) {
                sym[i] = j;
                sym[j] = i;
                break;
            }
}

void orc_compose_p_maps(const int *map1, const int *map2, int dims, int *out)
{
    for (int i = 0; i < dims; i++) out[i] = map2[map1[i]];
}

/* ------------------------------------------------------------------------
 * 5. create_big_matrix / revert_big_matrix
 *    (transpose_helper.cc:20-171, :331-409)
 *
 * permuted_addr: num_tiles x dims tile addresses already in permuted order.
 * grid_T[p_map[d]] = vgrid[d].  Row key = rank(addr with dims >= nrow zeroed),
 * column key = rank(addr with dims < nrow zeroed).  Tiles are placed in
 * ascending (row key, column key) order; for a duplicate (row, col) key the
 * first tile wins (std::map insert).  Every row must hold the same number of
 * columns (assert at :126) -> return -1 here.
 * big[(r*rpb + rr)*total_cols + c*cpb + cc] = tile[rr*cpb + cc].
 * Outputs: order[r*ncols + c] = tile number placed at block (r, c).
 * Pinned by: the round-trip invariant of big_matrix_transpose_test
 * (test.cc:633-728), a closed-form index formula, and brute-force ordering.
 * ---------------------------------------------------------------------- */
typedef struct { int64_t row, col; int64_t tile; } orc_key;

static int orc_key_cmp(const void *a, const void *b)
{
    const orc_key *x = (const orc_key *)a, *y = (const orc_key *)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    if (x->tile != y->tile) return x->tile < y->tile ? -1 : 1;
    return 0;
}

int orc_big_matrix_order(int dims, const int *vgrid, const int *p_map,
                         const int *permuted_addr, int64_t num_tiles,
                         int nrow_dims, int64_t *order, int64_t *n_rows,
                         int64_t *n_cols)
{
    int grid[ORC_MAXD], tmp[ORC_MAXD];
    for (int d = 0; d < dims; d++) grid[p_map[d]] = vgrid[d];
    orc_key *k = (orc_key *)malloc(sizeof(orc_key) * (size_t)(num_tiles ? num_tiles : 1));
    for (int64_t t = 0; t < num_tiles; t++) {
        const int *a = permuted_addr + t * dims;
        for (int d = 0; d < dims; d++) tmp[d] = d < nrow_dims ? a[d] : 0;
        k[t].row = orc_block_rank(tmp, dims, grid);
        for (int d = 0; d < dims; d++) tmp[d] = d < nrow_dims ? 0 : a[d];
        k[t].col = orc_block_rank(tmp, dims, grid);
        k[t].tile = t;
    }
    qsort(k, (size_t)num_tiles, sizeof(orc_key), orc_key_cmp);
    int64_t nr = 0, nc = -1, cur = 0, w = 0;
    for (int64_t t = 0; t < num_tiles; t++) {
        if (t > 0 && k[t].row == k[t - 1].row && k[t].col == k[t - 1].col)
            continue; /* duplicate key: first insert wins */
        if (t == 0 || k[t].row != k[t - 1].row) {
            if (t > 0) {
                if (nc >= 0 && cur != nc) { free(k); return -1; }
                nc = cur;
            }
            nr++;
            cur = 0;
        }
        cur++;
        order[w++] = k[t].tile;
    }
    if (num_tiles > 0) {
        if (nc >= 0 && cur != nc) { free(k); return -1; }
        nc = cur;
    } else {
        nc = 0;
    }
    free(k);
    *n_rows = nr;
    *n_cols = nc;
    return 0;
}

void orc_big_matrix_pack(const double *tiles, double *big, const int64_t *order,
                         int64_t n_rows, int64_t n_cols, int64_t rpb, int64_t cpb,
                         int64_t block_size, int unpack)
{
    int64_t total_cols = n_cols * cpb;
    for (int64_t r = 0; r < n_rows; r++)
        for (int64_t c = 0; c < n_cols; c++) {
            const int64_t t = order[r * n_cols + c];
            for (int64_t rr = 0; rr < rpb; rr++) {
                double *b = big + (r * rpb + rr) * total_cols + c * cpb;
                double *tl = (double *)tiles + t * block_size + rr * cpb;
                if (unpack) memcpy(tl, b, (size_t)cpb * sizeof(double));
                else memcpy(b, tl, (size_t)cpb * sizeof(double));
            }
        }
}

/* ------------------------------------------------------------------------
 * 6. local_contraction  (localcontract.cc:3-21) and DLTC compute
 *    (compute.cc:3-36): row-major C[m x n] += alpha * op(A) * op(B), beta=1.
 *
 * Leading dimensions here are the intended ones: op(A) = A ([m x k], lda=k)
 * or A^T (A is [k x m], lda=m); op(B) = B ([k x n], ldb=n) or B^T (B is
 * [n x k], ldb=k).  The reference passes lda=n_k and ldb=n_b regardless of
 * the transpose flags (localcontract.cc:7-9); with bt=1 that is only a valid
 * row-major leading dimension when n_b == n_k.
 * Textbook triple loop, FP64, k innermost.  Pinned against numpy float64
 * matmul (a library routine) in tests.
 * ---------------------------------------------------------------------- */
void orc_dgemm(int at, int bt, int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb,
               double *C, int64_t ldc)
{
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++) {
            double s = 0.0;
            for (int64_t p = 0; p < k; p++) {
                double a = at ? A[p * lda + i] : A[i * lda + p];
                double b = bt ? B[j * ldb + p] : B[p * ldb + j];
                s += a * b;
            }
            C[i * ldc + j] += alpha * s;
        }
}

/* Cache-blocked, OpenMP-parallel FP64 GEMM used as the CPU baseline in place
 * of MKL (which the reference links, Makefile:15).  C[m x n] += alpha*A*B^T
 * with A [m x k] (lda) and B [n x k] (ldb): the RRR form, where both big
 * matrices have the contraction index as columns (contraction.cc:1956).
 * Blocks of 64x64 C, k-panels of 256; inner kernel is a 4x8 register block
 * written so the compiler vectorises it. */
void orc_dgemm_nt_blocked(int64_t m, int64_t n, int64_t k, double alpha,
                          const double *A, int64_t lda, const double *B,
                          int64_t ldb, double *C, int64_t ldc, int nthreads)
{
    const int64_t MB = 64, NB = 64, KB = 256;
    int64_t nmb = (m + MB - 1) / MB, nnb = (n + NB - 1) / NB;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *bp = (double *)aligned_alloc(64, (size_t)(KB * NB) * sizeof(double));
        double *ap = (double *)aligned_alloc(64, (size_t)(KB * MB) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1) collapse(2)
#endif
        for (int64_t bi = 0; bi < nmb; bi++)
            for (int64_t bj = 0; bj < nnb; bj++) {
                int64_t i0 = bi * MB, j0 = bj * NB;
                int64_t mi = (m - i0) < MB ? (m - i0) : MB;
                int64_t nj = (n - j0) < NB ? (n - j0) : NB;
                for (int64_t p0 = 0; p0 < k; p0 += KB) {
                    int64_t kk = (k - p0) < KB ? (k - p0) : KB;
                    /* pack: ap[p][i], bp[p][j] (k-major) */
                    for (int64_t i = 0; i < MB; i++)
                        for (int64_t p = 0; p < kk; p++)
                            ap[p * MB + i] = i < mi ? A[(i0 + i) * lda + p0 + p] : 0.0;
                    for (int64_t j = 0; j < NB; j++)
                        for (int64_t p = 0; p < kk; p++)
                            bp[p * NB + j] = j < nj ? B[(j0 + j) * ldb + p0 + p] : 0.0;
                    for (int64_t i = 0; i < MB; i += 4)
                        for (int64_t j = 0; j < NB; j += 8) {
                            double acc[4][8];
                            for (int a = 0; a < 4; a++)
                                for (int b = 0; b < 8; b++) acc[a][b] = 0.0;
                            for (int64_t p = 0; p < kk; p++) {
                                const double *av = ap + p * MB + i;
                                const double *bv = bp + p * NB + j;
                                for (int a = 0; a < 4; a++)
                                    for (int b = 0; b < 8; b++) acc[a][b] += av[a] * bv[b];
                            }
                            for (int a = 0; a < 4; a++) {
                                if (i + a >= mi) break;
                                for (int b = 0; b < 8; b++) {
                                    if (j + b >= nj) break;
                                    C[(i0 + i + a) * ldc + j0 + j + b] += alpha * acc[a][b];
                                }
                            }
                        }
                }
            }
        free(bp);
        free(ap);
    }
}

/* OpenMP tile-parallel RRR permutation over a batch of equal tiles, as
 * contraction.cc:1824-1857 applies transpose() tile by tile (the reference
 * parallelises inside each tile with OpenMP, helper.cc:360-364,452-456). */
void orc_rrr_transpose_tiles(const double *in, double *out, const int *pmap,
                             int dims, const int *w, int64_t ntiles,
                             int64_t block_size, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t t = 0; t < ntiles; t++)
        orc_rrr_transpose(in + t * block_size, out + t * block_size, pmap, dims, w);
}

/* ------------------------------------------------------------------------
 * 7. DLTC brute-force checkers (checkerM.cc:39-47, checkerS.cc:35-41)
 *   checkM: sC[a,b,i,j] += A[a,k,i,l] * B[b,l,j,k]   A:[V,O,O,O] B:[V,O,O,O]
 *   checkS: sC[a,b,i,j] += A[b,a,j,i]                A:[V,V,O,O]
 * The loop odometer is (P1,P2,H1,H2,H3,H4), last fastest (loops.cc:17-31).
 * Inputs in the reference are 2*drand48()-1 from glibc's default seed.
 * ---------------------------------------------------------------------- */
void orc_checkM(const double *A, const double *B, double *sC, int O, int V)
{
    for (int a = 0; a < V; a++)
        for (int b = 0; b < V; b++)
            for (int i = 0; i < O; i++)
                for (int j = 0; j < O; j++)
                    for (int k = 0; k < O; k++)
                        for (int l = 0; l < O; l++)
                            sC[(((int64_t)a * V + b) * O + i) * O + j] +=
                                A[(((int64_t)a * O + k) * O + i) * O + l] *
                                B[(((int64_t)b * O + l) * O + j) * O + k];
}

void orc_checkS(const double *A, double *sC, int O, int V)
{
    for (int a = 0; a < V; a++)
        for (int b = 0; b < V; b++)
            for (int i = 0; i < O; i++)
                for (int j = 0; j < O; j++)
                    sC[(((int64_t)a * V + b) * O + i) * O + j] +=
                        A[(((int64_t)b * V + a) * O + j) * O + i];
}

/* glibc drand48: X_{n+1} = (0x5DEECE66D * X_n + 0xB) mod 2^48, default
 * X_0 = 0x1234ABCD330E; result X / 2^48.  Pinned against libc's drand48. */
void orc_drand48_fill(double *out, int64_t n, uint64_t *state)
{
    uint64_t x = *state;
    for (int64_t i = 0; i < n; i++) {
        x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
        out[i] = (double)x / 281474976710656.0;
    }
    *state = x;
}

/* ------------------------------------------------------------------------
 * 8. Coupled-cluster closed forms (absent in the reference; new code).
 *  D[i,j,a,b] = e_i + e_j - e_a - e_b
 *  T[a,b,i,j] <- R[a,b,i,j] / D[i,j,a,b]            (layout [v,v,o,o])
 *  E = 1/4 sum_{ijab} V[i,j,a,b] * T[a,b,i,j]        (V layout [o,o,v,v])
 *  ||R||_2
 * Sums in long double.  Pinned by brute force on tiny systems and by
 * closed-form special cases (constant inputs) in tests.
 * ---------------------------------------------------------------------- */
void orc_denom_update(const double *R, double *T, const double *eo,
                      const double *ev, int o, int v)
{
    for (int a = 0; a < v; a++)
        for (int b = 0; b < v; b++)
            for (int i = 0; i < o; i++)
                for (int j = 0; j < o; j++) {
                    int64_t x = (((int64_t)a * v + b) * o + i) * o + j;
                    T[x] = R[x] / (eo[i] + eo[j] - ev[a] - ev[b]);
                }
}

double orc_energy(const double *Vijab, const double *Tabij, int o, int v)
{
    long double s = 0.0L;
    for (int i = 0; i < o; i++)
        for (int j = 0; j < o; j++)
            for (int a = 0; a < v; a++)
                for (int b = 0; b < v; b++)
                    s += (long double)Vijab[(((int64_t)i * o + j) * v + a) * v + b] *
                         (long double)Tabij[(((int64_t)a * v + b) * o + i) * o + j];
    return (double)(0.25L * s);
}

double orc_norm2(const double *R, int64_t n)
{
    long double s = 0.0L;
    for (int64_t i = 0; i < n; i++) s += (long double)R[i] * R[i];
    return (double)sqrtl(s);
}

