/*
 * dtcf_dropin_rrr.h -- drop-in create_big_matrix / create_big_matrix_nl /
 * revert_big_matrix (transpose_helper.h:31-43) over the C ABI.
 *
 * These entry points take the reference's RRR::Tensor, whose layout lives in
 * the reference's tensor.h.  Include this header in one translation unit of
 * the caller after tensor.h and expand DTCF_RRR_DROPIN_DEFINE() there, in
 * place of compiling transpose_helper.cc.  The bookkeeping (tile ranks over
 * the permuted grid, std::map ordering with first-insert-wins, the equal
 * column-count assert) follows transpose_helper.cc:20-409 on the host; the
 * data movement runs on the GPU through dtcf_pack_blocks_f64_host.
 */
#ifndef DTCF_DROPIN_RRR_H
#define DTCF_DROPIN_RRR_H

#include <cassert>
#include <map>
#include <vector>

#include "dtcf_local.h"

namespace dtcf_rrr {

inline int grid_rank(const int *addr, const std::vector<int> &grid)
{
    int x = 1, r = 0;
    for (int i = (int)grid.size() - 1; i >= 0; i--) {
        r += addr[i] * x;
        x *= grid[i];
    }
    return r;
}

// TensorT needs get_dims(), get_vgrid() and get_block_size() (tensor.h).
template <class TensorT>
void create_big_matrix(TensorT *&T, double *&tensor_tiles, int *&permuted_address, int *&p_map_T,
                       int num_tiles, int row_per_block, int col_per_block, int num_row_dim,
                       std::map<int, std::map<int, int> > *&bm, double *&big,
                       std::map<int, int> *address_to_location)
{
    bm = new std::map<int, std::map<int, int> >();
    const int dims = T->get_dims();
    std::vector<int> grid(dims), tmp(dims);
    for (int d = 0; d < dims; d++) grid[p_map_T[d]] = T->get_vgrid()[d];
    for (int i = 0; i < num_tiles; i++) {
        const int *a = permuted_address + i * dims;
        for (int d = 0; d < dims; d++) tmp[d] = d < num_row_dim ? a[d] : 0;
        const int row = grid_rank(tmp.data(), grid);
        for (int d = 0; d < dims; d++) tmp[d] = d < num_row_dim ? 0 : a[d];
        const int col = grid_rank(tmp.data(), grid);
        int loc = i;
        if (address_to_location) {
            const int rank = grid_rank(a, grid);
            assert(address_to_location->find(rank) != address_to_location->end());
            loc = address_to_location->find(rank)->second;
        }
        (*bm)[row].insert(std::make_pair(col, loc));  // first insert wins
    }
    const int64_t bs = T->get_block_size();
    big = new double[(size_t)num_tiles * bs];
    if (bm->empty()) return;
    const int64_t ncb = (int64_t)bm->begin()->second.size();
    std::vector<int64_t> order;
    for (auto &row : *bm) {
        assert((int64_t)row.second.size() == ncb);  // transpose_helper.cc:126
        for (auto &c : row.second) order.push_back(c.second);
    }
    int64_t ntiles = num_tiles;
    for (int64_t t : order) ntiles = t + 1 > ntiles ? t + 1 : ntiles;
    if (dtcf_pack_blocks_f64_host(tensor_tiles, ntiles, order.data(), (int64_t)bm->size(), ncb,
                                  row_per_block, col_per_block, big, 0) != DTCF_OK)
        assert(!"dtcf_pack_blocks_f64_host failed");
}

template <class TensorT>
void revert_big_matrix(TensorT *&T, std::map<int, std::map<int, int> > *&bm, double *&big,
                       int row_per_block, int col_per_block, int *&p_map_T, double *&tensor_tiles)
{
    (void)p_map_T;
    if (bm->empty()) return;
    const int64_t ncb = (int64_t)bm->begin()->second.size();
    std::vector<int64_t> order;
    int64_t ntiles = 0;
    for (auto &row : *bm) {
        assert((int64_t)row.second.size() == ncb);
        for (auto &c : row.second) {
            order.push_back(c.second);
            ntiles = c.second + 1 > ntiles ? c.second + 1 : ntiles;
        }
    }
    (void)T;
    if (dtcf_pack_blocks_f64_host(tensor_tiles, ntiles, order.data(), (int64_t)bm->size(), ncb,
                                  row_per_block, col_per_block, big, 1) != DTCF_OK)
        assert(!"dtcf_pack_blocks_f64_host failed");
}

}  // namespace dtcf_rrr

/* Defines RRR::create_big_matrix, create_big_matrix_nl and revert_big_matrix
 * with the reference's signatures for the caller's RRR::Tensor. */
#define DTCF_RRR_DROPIN_DEFINE()                                                                  \
    namespace RRR {                                                                               \
    void create_big_matrix(Tensor *&T, double *&tiles, int *&paddr, int *&pmap, int n, int rpb,   \
                           int cpb, int nrow, std::map<int, std::map<int, int> > *&bm,            \
                           double *&big)                                                          \
    {                                                                                             \
        dtcf_rrr::create_big_matrix(T, tiles, paddr, pmap, n, rpb, cpb, nrow, bm, big, nullptr);  \
    }                                                                                             \
    void create_big_matrix_nl(Tensor *&T, double *&tiles, int *&paddr, int *&pmap, int n,         \
                              int rpb, int cpb, int nrow,                                         \
                              std::map<int, std::map<int, int> > *&bm, double *&big,              \
                              std::map<int, int> *&a2l)                                           \
    {                                                                                             \
        dtcf_rrr::create_big_matrix(T, tiles, paddr, pmap, n, rpb, cpb, nrow, bm, big, a2l);      \
    }                                                                                             \
    void revert_big_matrix(Tensor *&T, std::map<int, std::map<int, int> > *&bm, double *&big,     \
                           int rpb, int cpb, int *&pmap, double *&tiles)                          \
    {                                                                                             \
        dtcf_rrr::revert_big_matrix(T, bm, big, rpb, cpb, pmap, tiles);                           \
    }                                                                                             \
    }

#endif /* DTCF_DROPIN_RRR_H */
