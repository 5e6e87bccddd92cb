/*
 * dtcf_dropin.h -- C++ drop-in replacements for the reference's local-path
 * entry points, with the reference's exact signatures (libdtcf_dropin.so).
 *
 * Each symbol forwards to the C ABI in dtcf_local.h (HIP on gfx950).  Host
 * buffers are staged through device memory on every call; reference-shaped
 * callers therefore run unchanged, and callers that keep data on the device
 * use dtcf_local.h directly.
 *
 * Errors keep the reference's behaviour: the reference aborts inside BLAS on
 * bad GEMM arguments and calls exit(0) after "transpose position error" in
 * DLTC::transpose (DLTC/transp.cc:48-51); these wrappers print the library's
 * message and do the same.
 *
 * Interfaces that take reference classes (RRR::Tensor for create_big_matrix,
 * DLTC::Multiplication for compute) are provided as header templates below
 * (dtcf_dropin_rrr.h / dtcf_dropin_dltc.h), instantiated in a translation unit
 * of the caller that also includes the reference's own class headers.
 */
#ifndef DTCF_DROPIN_H
#define DTCF_DROPIN_H

#include <list>
#include <string>
#include <vector>

#include "dtcf_local.h"

/* localcontract.h:15-16 -- row-major C[n_a x n_b] += alpha*op(A)*op(B).
 * Leading dimensions are the intended ones (A: n_k or n_a, B: n_b or n_k);
 * the reference passes ldb = n_b even for bt=1 (localcontract.cc:7-9). */
extern void local_contraction(int n_a, int n_b, int n_k, double *A, double *B, double *C, int at,
                              int bt, double alpha);

/* helper.h:105-109 -- RRR tile permutation: input dim i goes to output
 * position permutation_map[i]; block_widths are the input extents.  Unlike
 * the reference (helper.cc:317), dimension 1 is a plain copy. */
extern void transpose(double *&input_buffer, double *&output_buffer, int *permutation_map,
                      int dimension, int *block_widths);

/* DLTC/transp-kernel.h:4-29 -- out[<letters>] = in[i][j][k][l]. */
void transpose_ij(const double *in, double *out, int di, int dj);
void transpose_ji(const double *in, double *out, int di, int dj);
void transpose_ijkl(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_ijlk(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_ikjl(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_iklj(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_iljk(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_ilkj(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_jikl(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_jilk(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_jkil(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_jkli(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_jlik(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_jlki(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_kijl(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_kilj(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_kjil(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_kjli(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_klij(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_klji(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_lijk(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_likj(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_ljik(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_ljki(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_lkij(const double *in, double *out, int di, int dj, int dk, int dl);
void transpose_lkji(const double *in, double *out, int di, int dj, int dk, int dl);

namespace RRR {
/* block_grid.h:13-70.  The member list follows the reference header so the
 * object layout matches code compiled against it.  One rank: the
 * constructor records rank 0 and nblocks = product of the grid instead of
 * querying MPI (block_grid.cc:12-13). */
class Block_Grid {
  private:
    int *actual_grid;
    int *actual_addr;
    int *gdmap;
    std::list<int> *igdmap;

  public:
    int grid_dims;
    int *bgrid;
    int rank;
    int *block_addr;
    int nblocks;
    Block_Grid(int num_dims, int *block_grid);
    int get_block_rank(int *&address);
    void get_block_ranks(int n, int **&addresses, int *&ranks);
    void get_block_addr(int rank, int *&address);
    void get_block_addrs(int n, int **&addresses, int *&ranks);
    void print_block_addr(int r);
    std::string get_block_addr_str(int r);
};
}  // namespace RRR

namespace DLTC {
#ifndef DTCF_DROPIN_EXTERNAL_INDEXNAME
/* DLTC/param.h:17-22 (the non-CHECK build: NIDX 24, NOCC 12). */
enum IndexName {
    h1, h2, h3, h4, h5, h6, h7, h8, h9, h10, h11, h12,
    p1, p2, p3, p4, p5, p6, p7, p8, p9, p10, p11, p12
};
#endif
/* DLTC/transp.h:19-21.  Widths come from the caller's globals wbO / wbV
 * (define.h:59-65) when it defines them, else from dtcf_dltc_set_widths.
 * Identity permutations copy buf_x into buf_y (the reference rebinds its
 * by-value pointer and leaves buf_y unwritten, transp.cc:16,24). */
void transpose(double *buf_x, const std::vector<IndexName> &x, double *buf_y,
               const std::vector<IndexName> &y);
}  // namespace DLTC

extern "C" {
/* Tile widths for DLTC::transpose when the caller does not define the
 * globals wbO / wbV; nocc = number of occupied index names (12, or 6 for the
 * reference's CHECK build). */
void dtcf_dltc_set_widths(int wb_occ, int wb_virt, int nocc);
}

#endif /* DTCF_DROPIN_H */
