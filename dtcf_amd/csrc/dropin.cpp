// C++ drop-in symbols with the reference's signatures (include/dtcf_dropin.h),
// forwarding to the C ABI of libdtcf.so.
#include "dtcf_dropin.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <sstream>

// The reference's DLTC main.cc defines these globals (param.h:53-54); when a
// caller does not, the weak references stay null and dtcf_dltc_set_widths
// supplies the values.
extern int wbO __attribute__((weak));
extern int wbV __attribute__((weak));

namespace {

int g_wbo = 1, g_wbv = 1, g_nocc = 12;

[[noreturn]] void die(const char *where)
{
    std::fprintf(stderr, "%s: %s\n", where, dtcf_last_error());
    std::abort();
}

void permute_host(const double *in, double *out, int dims, const int64_t *w, const int32_t *src)
{
    if (dtcf_permute_tiles_f64_host(in, out, dims, w, src, 1) != DTCF_OK) die("transpose");
}

void perm4(const double *in, double *out, int di, int dj, int dk, int dl, int a, int b, int c,
           int d)
{
    const int64_t w[4] = {di, dj, dk, dl};
    const int32_t s[4] = {a, b, c, d};
    permute_host(in, out, 4, w, s);
}

}  // namespace

void local_contraction(int n_a, int n_b, int n_k, double *A, double *B, double *C, int at, int bt,
                       double alpha)
{
    const int64_t lda = at ? n_a : n_k;
    const int64_t ldb = bt ? n_k : n_b;
    if (dtcf_gemm_f64_host(at, bt, n_a, n_b, n_k, alpha, A, lda, B, ldb, C, n_b) != DTCF_OK)
        die("local_contraction");
}

void transpose(double *&input_buffer, double *&output_buffer, int *permutation_map, int dimension,
               int *block_widths)
{
    if (dimension < 1 || dimension > DTCF_MAX_DIMS) {
        std::fprintf(stderr, "transpose: unsupported dimension %d\n", dimension);
        std::abort();
    }
    int64_t w[DTCF_MAX_DIMS];
    int32_t src[DTCF_MAX_DIMS];
    for (int i = 0; i < dimension; i++) {
        w[i] = block_widths[i];
        src[i] = -1;
    }
    for (int i = 0; i < dimension; i++) {
        const int p = permutation_map[i];
        if (p < 0 || p >= dimension || src[p] != -1) {
            std::fprintf(stderr, "transpose: permutation_map is not a permutation\n");
            std::abort();
        }
        src[p] = i;
    }
    permute_host(input_buffer, output_buffer, dimension, w, src);
}

void transpose_ij(const double *in, double *out, int di, int dj)
{
    const int64_t w[2] = {di, dj};
    const int32_t s[2] = {0, 1};
    permute_host(in, out, 2, w, s);
}
void transpose_ji(const double *in, double *out, int di, int dj)
{
    const int64_t w[2] = {di, dj};
    const int32_t s[2] = {1, 0};
    permute_host(in, out, 2, w, s);
}

// Output order named by the letters: out[<letters>] = in[i][j][k][l].
#define DTCF_PERM4(name, a, b, c, d)                                                     \
    void transpose_##name(const double *in, double *out, int di, int dj, int dk, int dl) \
    {                                                                                    \
        perm4(in, out, di, dj, dk, dl, a, b, c, d);                                      \
    }
DTCF_PERM4(ijkl, 0, 1, 2, 3)
DTCF_PERM4(ijlk, 0, 1, 3, 2)
DTCF_PERM4(ikjl, 0, 2, 1, 3)
DTCF_PERM4(iklj, 0, 2, 3, 1)
DTCF_PERM4(iljk, 0, 3, 1, 2)
DTCF_PERM4(ilkj, 0, 3, 2, 1)
DTCF_PERM4(jikl, 1, 0, 2, 3)
DTCF_PERM4(jilk, 1, 0, 3, 2)
DTCF_PERM4(jkil, 1, 2, 0, 3)
DTCF_PERM4(jkli, 1, 2, 3, 0)
DTCF_PERM4(jlik, 1, 3, 0, 2)
DTCF_PERM4(jlki, 1, 3, 2, 0)
DTCF_PERM4(kijl, 2, 0, 1, 3)
DTCF_PERM4(kilj, 2, 0, 3, 1)
DTCF_PERM4(kjil, 2, 1, 0, 3)
DTCF_PERM4(kjli, 2, 1, 3, 0)
DTCF_PERM4(klij, 2, 3, 0, 1)
DTCF_PERM4(klji, 2, 3, 1, 0)
DTCF_PERM4(lijk, 3, 0, 1, 2)
DTCF_PERM4(likj, 3, 0, 2, 1)
DTCF_PERM4(ljik, 3, 1, 0, 2)
DTCF_PERM4(ljki, 3, 1, 2, 0)
DTCF_PERM4(lkij, 3, 2, 0, 1)
DTCF_PERM4(lkji, 3, 2, 1, 0)
#undef DTCF_PERM4

namespace RRR {

Block_Grid::Block_Grid(int num_dims, int *block_grid)
    : actual_grid(nullptr), actual_addr(nullptr), gdmap(nullptr), igdmap(nullptr)
{
    grid_dims = num_dims;
    bgrid = new int[grid_dims + 1];
    nblocks = 1;
    for (int i = 0; i < grid_dims; i++) {
        bgrid[i] = block_grid[i];
        nblocks *= block_grid[i];
    }
    rank = 0;
    get_block_addr(rank, block_addr);
}

int Block_Grid::get_block_rank(int *&address)
{
    int x = 1, r = 0;
    for (int i = grid_dims - 1; i >= 0; i--) {
        r += address[i] * x;
        x *= bgrid[i];
    }
    return r;
}

void Block_Grid::get_block_ranks(int n, int **&addresses, int *&ranks)
{
    ranks = new int[n];
    for (int r = 0; r < n; r++) ranks[r] = get_block_rank(addresses[r]);
}

void Block_Grid::get_block_addr(int r, int *&address)
{
    address = new int[grid_dims];
    int x = 1;
    for (int i = 0; i < grid_dims; i++) x *= bgrid[i];
    for (int i = 0; i < grid_dims; i++) {
        x /= bgrid[i];
        address[i] = r / x;
        r -= address[i] * x;
    }
}

void Block_Grid::get_block_addrs(int n, int **&addresses, int *&ranks)
{
    for (int r = 0; r < n; r++) get_block_addr(ranks[r], addresses[r]);
}

void Block_Grid::print_block_addr(int r)
{
    std::printf("node address is : %s\n", get_block_addr_str(r).c_str());
}

std::string Block_Grid::get_block_addr_str(int r)
{
    int *a = nullptr;
    get_block_addr(r, a);
    std::stringstream ss;
    ss << "[ ";
    for (int i = 0; i < grid_dims; i++) {
        ss << a[i];
        if (i < grid_dims - 1) ss << ", ";
    }
    ss << "]";
    delete[] a;
    return ss.str();
}

}  // namespace RRR

extern "C" void dtcf_dltc_set_widths(int wb_occ, int wb_virt, int nocc)
{
    g_wbo = wb_occ;
    g_wbv = wb_virt;
    g_nocc = nocc;
}

namespace DLTC {

// DLTC/transp.cc:3-59 with position() from DLTC/transp.h:8-17.
void transpose(double *buf_x, const std::vector<IndexName> &x, double *buf_y,
               const std::vector<IndexName> &y)
{
    const int dim = (int)x.size();
    if (dim != (int)y.size() || (dim != 2 && dim != 4)) {
        std::printf("transpose position error\n");
        std::exit(0);
    }
    const int wo = &wbO ? wbO : g_wbo;
    const int wv = &wbV ? wbV : g_wbv;
    int64_t w[4];
    int32_t s[4];
    for (int i = 0; i < dim; i++) w[i] = ((int)x[i] < g_nocc) ? wo : wv;
    // s[n] = position of y[n] in x: output dim n is fed by input dim s[n].
    for (int n = 0; n < dim; n++) {
        s[n] = -1;
        for (int i = 0; i < dim; i++)
            if (x[i] == y[n]) s[n] = i;
        if (s[n] < 0) {
            std::printf("transpose position error\n");
            std::exit(0);
        }
    }
    permute_host(buf_x, buf_y, dim, w, s);
}

}  // namespace DLTC
