// Host planner and executor for one rank's local step of an RRR contraction
// C += alpha * A * B on tiled tensors (no kernels here; it drives the
// permute, view-gather and GEMM kernels through the C ABI).
//
// Mirrors, for one rank:
//   contr_list / cntr_map           contraction.cc:2553-2572
//   generate_permutation_map        contraction.cc:164-282
//   get_symm_permutation, compose   contraction.cc:132-160, 121-128
//   C pre-transpose + pack          contraction.cc:2692-2715, 2768
//   transpose_and_dgemm             contraction.cc:1779-1976
//   revert + C inverse transpose    contraction.cc:2808, 2817-2830
//   create_big_matrix ordering      transpose_helper.cc:20-171
//   Block_Grid rank                 block_grid.cc:22-37
//
// Three execution modes:
//   LITERAL: permute every tile (per symmetry class), pack into dense big
//            matrices, GEMM, unpack, inverse-permute C -- the reference's
//            sequence of passes;
//   PACKED : permute+pack in one gather pass per operand, GEMM, one scatter;
//   FUSED  : a single GEMM that reads A, B and C straight from the tiles
//            through offset tables (no big matrices at all).
// An A or B operand of a symmetric tensor whose tiles carry several composed
// maps (bounced copies next to stored tiles) cannot be one offset table: in
// PACKED and FUSED it is packed by one permute pass per class, each writing
// its tiles straight into their blocks of the big matrix, and FUSED's GEMM
// reads that big matrix densely.  C is never bounced (one class).
#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace dtcf {
namespace {

std::vector<std::string> split_idx(const char *s)
{
    std::vector<std::string> out;
    std::string cur;
    for (const char *p = s; *p; p++) {
        if (*p == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (*p != ' ') {
            cur.push_back(*p);
        }
    }
    out.push_back(cur);
    return out;
}

int64_t grid_rank(const std::vector<int> &addr, const std::vector<int> &grid)
{
    int64_t x = 1, r = 0;
    for (int i = (int)grid.size() - 1; i >= 0; i--) {
        r += (int64_t)addr[i] * x;
        x *= grid[i];
    }
    return r;
}

// Per-operand plan: tile order in the big matrix, per-tile composed maps.
struct OperandPlan {
    int dims = 0;
    std::vector<int> w;         // tile extents (block_range)
    std::vector<int> pmap;      // input dim -> output position (no symmetry)
    int nrow_dims = 0;          // leading permuted dims that form the rows
    int64_t rpb = 1, cpb = 1, bs = 1;
    int64_t nrb = 0, ncb = 0;
    std::vector<int64_t> order; // order[R*ncb + Cb] = tile number
    std::vector<std::vector<int>> composed;  // per tile
    std::vector<int> cls;       // per tile: index into classes
    std::vector<std::vector<int>> classes;   // distinct composed maps
    double *tiles = nullptr;
    int64_t ntiles = 0;
};

// Output extents after the permutation (position p gets w[inv[p]]).
std::vector<int> permuted_widths(const std::vector<int> &w, const std::vector<int> &pm)
{
    std::vector<int> o(w.size());
    for (size_t d = 0; d < w.size(); d++) o[pm[d]] = w[d];
    return o;
}

// apply_sym: A and B tiles that are bounced copies get the symmetry swap
// composed into their permutation (contraction.cc:1836-1841, 1894-1899);
// C tiles are permuted by p_map_C alone (contraction.cc:2705).
int build_operand(const dtcf_tileset *T, const std::vector<int> &pmap, int nrow_dims,
                  bool apply_sym, OperandPlan &P)
{
    DTCF_REQUIRE(T && T->dims >= 1 && T->dims <= DTCF_MAX_DIMS, "step: bad tileset dims");
    DTCF_REQUIRE(T->ntiles >= 0 && (T->ntiles == 0 || (T->tile_addr && T->tiles)),
                 "step: tileset without tiles or addresses");
    P.dims = T->dims;
    P.w.assign(T->block_range, T->block_range + T->dims);
    P.pmap = pmap;
    P.nrow_dims = nrow_dims;
    P.tiles = T->tiles;
    P.ntiles = T->ntiles;
    P.bs = 1;
    for (int d = 0; d < P.dims; d++) {
        DTCF_REQUIRE(P.w[d] >= 1, "step: tile extent < 1");
        P.bs *= P.w[d];
    }
    std::vector<int> pw = permuted_widths(P.w, pmap);
    P.rpb = 1;
    P.cpb = 1;
    for (int p = 0; p < P.dims; p++) (p < nrow_dims ? P.rpb : P.cpb) *= pw[p];
    // Permuted tile grid: grid_T[p_map[d]] = vgrid[d]  (transpose_helper.cc:36-40)
    std::vector<int> grid(P.dims);
    for (int d = 0; d < P.dims; d++) grid[pmap[d]] = T->vgrid[d];
    struct Key { int64_t row, col, tile; };
    std::vector<Key> keys(P.ntiles);
    P.composed.resize(P.ntiles);
    P.cls.resize(P.ntiles);
    for (int64_t t = 0; t < P.ntiles; t++) {
        const int32_t *a = T->tile_addr + t * P.dims;
        for (int d = 0; d < P.dims; d++)
            DTCF_REQUIRE(a[d] >= 0 && a[d] < T->vgrid[d], "step: tile address outside vgrid");
        // get_symm_permutation: swap i<->j when addr[i] < addr[j] in one
        // symmetry group (<= 1); first j only.  No sign.
        std::vector<int> sym(P.dims);
        std::iota(sym.begin(), sym.end(), 0);
        for (int i = 0; apply_sym && i < P.dims; i++)
            for (int j = i + 1; j < P.dims; j++)
                if (a[i] < a[j] && T->sym_group[i] == T->sym_group[j] && T->sym_group[i] <= 1) {
                    sym[i] = j;
                    sym[j] = i;
                    break;
                }
        {
            // Groups of three or more dims make the reference's pairwise swap
            // produce a non-permutation (contraction.cc:146-155); refuse them.
            std::vector<int> seen(P.dims, 0);
            for (int i = 0; i < P.dims; i++) {
                DTCF_REQUIRE(!seen[sym[i]], "step: symmetry group of more than two dims");
                seen[sym[i]] = 1;
            }
        }
        std::vector<int> comp(P.dims);
        for (int i = 0; i < P.dims; i++) comp[i] = pmap[sym[i]];
        P.composed[t] = comp;
        auto it = std::find(P.classes.begin(), P.classes.end(), comp);
        if (it == P.classes.end()) {
            P.classes.push_back(comp);
            P.cls[t] = (int)P.classes.size() - 1;
        } else {
            P.cls[t] = (int)(it - P.classes.begin());
        }
        // permuted address: dest[p_map[d]] = source[d]   (contraction.cc:1846-1852)
        std::vector<int> pa(P.dims), tmp(P.dims);
        for (int d = 0; d < P.dims; d++) pa[pmap[d]] = a[d];
        for (int d = 0; d < P.dims; d++) tmp[d] = d < nrow_dims ? pa[d] : 0;
        keys[t].row = grid_rank(tmp, grid);
        for (int d = 0; d < P.dims; d++) tmp[d] = d < nrow_dims ? 0 : pa[d];
        keys[t].col = grid_rank(tmp, grid);
        keys[t].tile = t;
    }
    std::stable_sort(keys.begin(), keys.end(), [](const Key &x, const Key &y) {
        return x.row != y.row ? x.row < y.row : x.col < y.col;
    });
    // std::map insert semantics: the first tile with a (row, col) key wins.
    P.order.clear();
    P.nrb = 0;
    P.ncb = -1;
    int64_t cur = 0;
    for (int64_t i = 0; i < P.ntiles; i++) {
        if (i > 0 && keys[i].row == keys[i - 1].row && keys[i].col == keys[i - 1].col) continue;
        if (i == 0 || keys[i].row != keys[i - 1].row) {
            if (i > 0) {
                DTCF_REQUIRE(P.ncb < 0 || cur == P.ncb,
                             "step: big-matrix rows hold different column counts");
                P.ncb = cur;
            }
            P.nrb++;
            cur = 0;
        }
        cur++;
        P.order.push_back(keys[i].tile);
    }
    if (P.ntiles > 0) {
        DTCF_REQUIRE(P.ncb < 0 || cur == P.ncb, "step: big-matrix rows hold different column counts");
        P.ncb = cur;
    } else {
        P.ncb = 0;
    }
    return DTCF_OK;
}

// Intra-tile offset tables for one composed map: rows over the permuted
// positions < nrow_dims, columns over the rest.
void intra_tables(const OperandPlan &P, const std::vector<int> &comp, std::vector<int64_t> &rofs,
                  std::vector<int64_t> &cofs)
{
    const int D = P.dims;
    std::vector<int64_t> istr(D);
    istr[D - 1] = 1;
    for (int d = D - 2; d >= 0; d--) istr[d] = istr[d + 1] * P.w[d + 1];
    // position p is fed by input dim inv[p]
    std::vector<int> inv(D);
    for (int d = 0; d < D; d++) inv[comp[d]] = d;
    auto table = [&](int p_lo, int p_hi, int64_t n, std::vector<int64_t> &out) {
        out.assign(n, 0);
        for (int64_t x = 0; x < n; x++) {
            int64_t t = x, off = 0;
            for (int p = p_hi - 1; p >= p_lo; p--) {
                const int d = inv[p];
                off += (t % P.w[d]) * istr[d];
                t /= P.w[d];
            }
            out[x] = off;
        }
    };
    table(0, P.nrow_dims, P.rpb, rofs);
    table(P.nrow_dims, D, P.cpb, cofs);
}

// Full-length view tables of the big matrix (rows = nrb*rpb, cols = ncb*cpb)
// for a single-class operand.
int make_view(const OperandPlan &P, bool packed_tiles, dtcf_view *out)
{
    std::vector<int64_t> r0, c0;
    const std::vector<int> &comp = packed_tiles ? std::vector<int>() : P.classes[0];
    if (packed_tiles) {
        // tiles already permuted: row-major [rpb][cpb]
        r0.resize(P.rpb);
        c0.resize(P.cpb);
        for (int64_t i = 0; i < P.rpb; i++) r0[i] = i * P.cpb;
        for (int64_t i = 0; i < P.cpb; i++) c0[i] = i;
    } else {
        intra_tables(P, comp, r0, c0);
    }
    const int64_t rows = P.nrb * P.rpb, cols = P.ncb * P.cpb;
    std::vector<int64_t> rofs(rows), cofs(cols), blk(P.nrb * P.ncb);
    for (int64_t r = 0; r < rows; r++) rofs[r] = r0[r % P.rpb];
    for (int64_t c = 0; c < cols; c++) cofs[c] = c0[c % P.cpb];
    for (int64_t b = 0; b < P.nrb * P.ncb; b++) blk[b] = P.order[b] * P.bs;
    return dtcf_view_create(rows, cols, rofs.data(), cofs.data(), blk.data(), P.rpb, P.cpb, out);
}

// One permute pass of a symmetric operand: the tiles of one bounce class,
// each permuted by the class's composed map and written into its block of
// the row-major big matrix (row pitch ld), as create_big_matrix places
// permuted tiles (transpose_helper.cc:149-159).
struct ClassPack {
    dtcf_strided d{};
    int64_t n = 0;
    int64_t *src = nullptr, *dst = nullptr;  // device offset tables
};

int make_class_packs(const OperandPlan &P, int64_t ld, std::vector<ClassPack> &out)
{
    const int D = P.dims;
    int64_t istr[DTCF_MAX_DIMS];
    istr[D - 1] = 1;
    for (int i = D - 2; i >= 0; i--) istr[i] = istr[i + 1] * P.w[i + 1];
    for (size_t c = 0; c < P.classes.size(); c++) {
        const std::vector<int> &comp = P.classes[c];
        int64_t pw[DTCF_MAX_DIMS], pstr[DTCF_MAX_DIMS];
        for (int i = 0; i < D; i++) pw[comp[i]] = P.w[i];
        // positions < nrow_dims form the block's rows (pitch ld), the rest
        // its columns; the last position of each group is the fastest
        int64_t acc = 1;
        for (int p = D - 1; p >= P.nrow_dims; p--) {
            pstr[p] = acc;
            acc *= pw[p];
        }
        const int64_t cols = acc;
        acc = ld;
        for (int p = P.nrow_dims - 1; p >= 0; p--) {
            pstr[p] = acc;
            acc *= pw[p];
        }
        // a swap of two dims of different extents across the row/column
        // split would not fill the rpb x cpb block: refuse it
        DTCF_REQUIRE(cols == P.cpb && acc == ld * P.rpb,
                     "step: bounced dims of different extents change the block shape");
        ClassPack k;
        k.d.dims = D;
        for (int i = 0; i < D; i++) {
            k.d.ext[i] = P.w[i];
            k.d.istr[i] = istr[i];
            k.d.ostr[i] = pstr[comp[i]];
        }
        std::vector<int64_t> so, dso;
        for (int64_t b = 0; b < P.nrb * P.ncb; b++) {
            const int64_t t = P.order[b];
            if (P.cls[t] != (int)c) continue;
            so.push_back(t * P.bs);
            dso.push_back((b / P.ncb) * P.rpb * ld + (b % P.ncb) * P.cpb);
        }
        k.n = (int64_t)so.size();
        out.push_back(k);  // owned by `out` from here (freed with the step)
        if (k.n == 0) continue;
        ClassPack &o = out.back();
        hipError_t e = hipMalloc(&o.src, k.n * sizeof(int64_t));
        if (e == hipSuccess) e = hipMalloc(&o.dst, k.n * sizeof(int64_t));
        if (e == hipSuccess)
            e = hipMemcpy(o.src, so.data(), k.n * sizeof(int64_t), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(o.dst, dso.data(), k.n * sizeof(int64_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "step: class pack offsets");
    }
    return DTCF_OK;
}

// Dense row-major rows x cols view (a packed big matrix read by FUSED).
int make_dense_view(int64_t rows, int64_t cols, dtcf_view *out)
{
    std::vector<int64_t> rofs(rows), cofs(cols);
    for (int64_t r = 0; r < rows; r++) rofs[r] = r * cols;
    for (int64_t c = 0; c < cols; c++) cofs[c] = c;
    return dtcf_view_create(rows, cols, rofs.data(), cofs.data(), nullptr, 0, 0, out);
}

}  // namespace
}  // namespace dtcf

using namespace dtcf;

struct dtcf_step_s {
    int mode = 0;
    OperandPlan A, B, C;
    int64_t n_a = 0, n_b = 0, n_k = 0;      // per-tile sizes (generate_permutation_map)
    int64_t M = 0, N = 0, K = 0;            // big-matrix GEMM sizes
    std::vector<int> p_map_A, p_map_B, p_map_C, inv_p_map_C;
    dtcf_view va = nullptr, vb = nullptr, vc = nullptr;
    double *bigA = nullptr, *bigB = nullptr, *bigC = nullptr, *tmp = nullptr;
    int64_t tmp_elems = 0;
    // per-class source/destination offsets for the LITERAL permute passes
    std::vector<int64_t *> cls_off;  // device arrays
    // PACKED / FUSED: A or B packed class by class (several bounce classes)
    std::vector<ClassPack> packA, packB;
    bool empty = false;
    // Optional timing of the GEMM kernel: one event pair per run, recorded on
    // the stream the GEMM is launched on.
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
};

static int step_alloc(double **p, int64_t n)
{
    *p = nullptr;
    if (n <= 0) return DTCF_OK;
    DTCF_HIP(hipMalloc(p, n * sizeof(double)));
    return DTCF_OK;
}

extern "C" int dtcf_step_destroy(dtcf_step s)
{
    if (!s) return DTCF_OK;
    dtcf_view_destroy(s->va);
    dtcf_view_destroy(s->vb);
    dtcf_view_destroy(s->vc);
    for (double *p : {s->bigA, s->bigB, s->bigC, s->tmp})
        if (p) (void)hipFree(p);
    for (int64_t *p : s->cls_off)
        if (p) (void)hipFree(p);
    for (auto *v : {&s->packA, &s->packB})
        for (ClassPack &k : *v) {
            if (k.src) (void)hipFree(k.src);
            if (k.dst) (void)hipFree(k.dst);
        }
    for (auto &e : s->ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    delete s;
    return DTCF_OK;
}

extern "C" int dtcf_step_plan(const dtcf_tileset *A, const char *a_idx, const dtcf_tileset *B,
                              const char *b_idx, const dtcf_tileset *C, const char *c_idx,
                              int mode, dtcf_step *out)
{
    DTCF_REQUIRE(out, "step: null output");
    *out = nullptr;
    DTCF_REQUIRE(A && B && C && a_idx && b_idx && c_idx, "step: null argument");
    DTCF_REQUIRE(mode >= DTCF_STEP_LITERAL && mode <= DTCF_STEP_FUSED, "step: unknown mode");
    auto ia = split_idx(a_idx), ib = split_idx(b_idx), ic = split_idx(c_idx);
    DTCF_REQUIRE((int)ia.size() == A->dims && (int)ib.size() == B->dims && (int)ic.size() == C->dims,
                 "step: index string length does not match tensor dims");
    // contr_list: pairs (i in A, j in B) with equal names, A-major order.
    std::vector<std::pair<int, int>> contr;
    for (size_t i = 0; i < ia.size(); i++)
        for (size_t j = 0; j < ib.size(); j++)
            if (ia[i] == ib[j]) contr.push_back({(int)i, (int)j});
    std::vector<int> ca(A->dims, 0), cb(B->dims, 0);
    for (auto &p : contr) {
        ca[p.first] = 1;
        cb[p.second] = 1;
    }
    dtcf_step_s *s = new dtcf_step_s();
    s->mode = mode;
    s->p_map_A.assign(A->dims, 0);
    s->p_map_B.assign(B->dims, 0);
    s->p_map_C.assign(C->dims, -1);
    s->inv_p_map_C.assign(C->dims, 0);
    int ea = 0, eb = 0;
    int64_t na = 1, nb = 1, nka = 1, nkb = 1;
    auto fail = [&](const char *m) {
        set_error(m);
        dtcf_step_destroy(s);
        return DTCF_ERR_INVALID;
    };
    for (int i = 0; i < A->dims; i++) {
        if (!ca[i]) {
            auto it = std::find(ic.begin(), ic.end(), ia[i]);
            if (it == ic.end()) return fail("step: external index of A missing from C");
            const int ci = (int)(it - ic.begin());
            na *= A->block_range[i];
            s->p_map_A[i] = ea;
            s->p_map_C[ci] = ea;
            s->inv_p_map_C[ea] = ci;
            if (C->block_range[ci] != A->block_range[i]) return fail("step: A/C tile extents differ");
            ea++;
        } else {
            nka *= A->block_range[i];
        }
    }
    for (int j = 0; j < B->dims; j++) {
        if (!cb[j]) {
            auto it = std::find(ic.begin(), ic.end(), ib[j]);
            if (it == ic.end()) return fail("step: external index of B missing from C");
            const int ci = (int)(it - ic.begin());
            nb *= B->block_range[j];
            s->p_map_B[j] = eb;
            s->p_map_C[ci] = ea + eb;
            s->inv_p_map_C[ea + eb] = ci;
            if (C->block_range[ci] != B->block_range[j]) return fail("step: B/C tile extents differ");
            eb++;
        } else {
            nkb *= B->block_range[j];
        }
    }
    if (nka != nkb) return fail("step: contraction tile sizes of A and B differ");
    if (ea + eb != C->dims) return fail("step: C dims are not A-ext + B-ext");
    for (size_t t = 0; t < contr.size(); t++) {
        s->p_map_A[contr[t].first] = ea + (int)t;
        s->p_map_B[contr[t].second] = eb + (int)t;
        if (A->block_range[contr[t].first] != B->block_range[contr[t].second] ||
            A->vgrid[contr[t].first] != B->vgrid[contr[t].second])
            return fail("step: contracted dims differ in tiling");
    }
    s->n_a = na;
    s->n_b = nb;
    s->n_k = nka;
    int st;
    if ((st = build_operand(A, s->p_map_A, ea, true, s->A)) != DTCF_OK ||
        (st = build_operand(B, s->p_map_B, eb, true, s->B)) != DTCF_OK ||
        (st = build_operand(C, s->p_map_C, ea, false, s->C)) != DTCF_OK) {
        dtcf_step_destroy(s);
        return st;
    }
    if (A->ntiles == 0 || B->ntiles == 0 || C->ntiles == 0) {
        s->empty = true;  // transpose_and_dgemm returns early (contraction.cc:1784)
        *out = s;
        return DTCF_OK;
    }
    s->M = s->A.nrb * s->A.rpb;
    s->N = s->B.nrb * s->B.rpb;
    s->K = s->A.ncb * s->A.cpb;
    // The big matrices must agree: the contraction column order of A and B
    // and the row/column blocks of C (kevin_dgemm at contraction.cc:1956).
    if (s->B.ncb * s->B.cpb != s->K) return fail("step: A and B contraction extents differ");
    if (s->C.nrb * s->C.rpb != s->M || s->C.ncb * s->C.cpb != s->N)
        return fail("step: C big matrix is not (A rows) x (B rows)");

    if (s->C.classes.size() != 1) return fail("step: C tiles with several permutations");
    const bool multiA = s->A.classes.size() > 1, multiB = s->B.classes.size() > 1;
    if (mode != DTCF_STEP_LITERAL && (multiA || multiB)) {
        if ((multiA && ((st = make_class_packs(s->A, s->K, s->packA)) != DTCF_OK ||
                        (st = step_alloc(&s->bigA, s->M * s->K)) != DTCF_OK)) ||
            (multiB && ((st = make_class_packs(s->B, s->K, s->packB)) != DTCF_OK ||
                        (st = step_alloc(&s->bigB, s->N * s->K)) != DTCF_OK))) {
            dtcf_step_destroy(s);
            return st;
        }
    }
    if (mode == DTCF_STEP_FUSED) {
        if ((st = multiA ? make_dense_view(s->M, s->K, &s->va) : make_view(s->A, false, &s->va)) !=
                DTCF_OK ||
            (st = multiB ? make_dense_view(s->N, s->K, &s->vb) : make_view(s->B, false, &s->vb)) !=
                DTCF_OK ||
            (st = make_view(s->C, false, &s->vc)) != DTCF_OK) {
            dtcf_step_destroy(s);
            return st;
        }
    } else {
        const bool lit = mode == DTCF_STEP_LITERAL;
        if ((st = make_view(s->A, lit, &s->va)) != DTCF_OK ||
            (st = make_view(s->B, lit, &s->vb)) != DTCF_OK ||
            (st = make_view(s->C, lit, &s->vc)) != DTCF_OK ||
            (!s->bigA && (st = step_alloc(&s->bigA, s->M * s->K)) != DTCF_OK) ||
            (!s->bigB && (st = step_alloc(&s->bigB, s->N * s->K)) != DTCF_OK) ||
            (st = step_alloc(&s->bigC, s->M * s->N)) != DTCF_OK) {
            dtcf_step_destroy(s);
            return st;
        }
        if (lit) {
            s->tmp_elems = std::max({s->A.ntiles * s->A.bs, s->B.ntiles * s->B.bs,
                                     s->C.ntiles * s->C.bs});
            if ((st = step_alloc(&s->tmp, s->tmp_elems)) != DTCF_OK) {
                dtcf_step_destroy(s);
                return st;
            }
            // per-class tile offsets (device), A classes, then B, then C
            for (OperandPlan *P : {&s->A, &s->B, &s->C}) {
                for (size_t c = 0; c < P->classes.size(); c++) {
                    std::vector<int64_t> offs;
                    for (int64_t t = 0; t < P->ntiles; t++)
                        if (P->cls[t] == (int)c) offs.push_back(t * P->bs);
                    int64_t *d = nullptr;
                    hipError_t e = hipMalloc(&d, offs.size() * sizeof(int64_t));
                    if (e == hipSuccess)
                        e = hipMemcpy(d, offs.data(), offs.size() * sizeof(int64_t),
                                      hipMemcpyHostToDevice);
                    if (e != hipSuccess) {
                        dtcf_step_destroy(s);
                        return hip_fail(e, "step: class offsets");
                    }
                    s->cls_off.push_back(d);
                }
            }
        }
    }
    *out = s;
    return DTCF_OK;
}

extern "C" int dtcf_step_info(dtcf_step s, int64_t *mnk_tile, int64_t *mnk_big)
{
    DTCF_REQUIRE(s, "step: null handle");
    if (mnk_tile) {
        mnk_tile[0] = s->n_a;
        mnk_tile[1] = s->n_b;
        mnk_tile[2] = s->n_k;
    }
    if (mnk_big) {
        mnk_big[0] = s->M;
        mnk_big[1] = s->N;
        mnk_big[2] = s->K;
    }
    return DTCF_OK;
}

// LITERAL helper: permute all tiles of an operand (class by class) into the
// step's temporary buffer; `inverse` applies the inverse of each composed map
// (for the C post-transpose, contraction.cc:2817-2830) from tmp back to tiles.
static int permute_operand(dtcf_step s, OperandPlan &P, size_t off_base, bool inverse,
                           hipStream_t st)
{
    for (size_t c = 0; c < P.classes.size(); c++) {
        const std::vector<int> &comp = P.classes[c];
        int64_t n = 0;
        for (int64_t t = 0; t < P.ntiles; t++) n += (P.cls[t] == (int)c);
        if (n == 0) continue;
        dtcf_strided d{};
        int64_t istr[DTCF_MAX_DIMS];
        const int D = P.dims;
        istr[D - 1] = 1;
        for (int i = D - 2; i >= 0; i--) istr[i] = istr[i + 1] * P.w[i + 1];
        std::vector<int> pw(D);
        for (int i = 0; i < D; i++) pw[comp[i]] = P.w[i];
        int64_t ostr[DTCF_MAX_DIMS];
        ostr[D - 1] = 1;
        for (int i = D - 2; i >= 0; i--) ostr[i] = ostr[i + 1] * pw[i + 1];
        d.dims = D;
        for (int i = 0; i < D; i++) {
            // iterate over input dims: in stride istr[i], out stride ostr[comp[i]]
            d.ext[i] = P.w[i];
            d.istr[i] = inverse ? ostr[comp[i]] : istr[i];
            d.ostr[i] = inverse ? istr[i] : ostr[comp[i]];
        }
        // one class holding every tile: consecutive tiles, a step instead of
        // an offset table (which lets the permute take its 16-byte form)
        const bool all = n == P.ntiles;
        const int64_t *offs = all ? nullptr : s->cls_off[off_base + c];
        const int64_t step = all ? P.bs : 0;
        int r = inverse ? dtcf_permute_copy_f64(s->tmp, P.tiles, &d, n, offs, step, offs, step, 1.0,
                                                0.0, st)
                        : dtcf_permute_copy_f64(P.tiles, s->tmp, &d, n, offs, step, offs, step, 1.0,
                                                0.0, st);
        if (r != DTCF_OK) return r;
    }
    return DTCF_OK;
}

extern "C" int dtcf_step_enable_timing(dtcf_step s, int max_runs)
{
    DTCF_REQUIRE(s && max_runs >= 0, "step: bad timing request");
    for (auto &e : s->ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    s->ev.clear();
    s->ev_used = 0;
    for (int i = 0; i < max_runs; i++) {
        hipEvent_t a, b;
        DTCF_HIP(hipEventCreate(&a));
        DTCF_HIP(hipEventCreate(&b));
        s->ev.push_back({a, b});
    }
    return DTCF_OK;
}

extern "C" int dtcf_step_kernel_times(dtcf_step s, float *ms, int max_n, int *n_out)
{
    DTCF_REQUIRE(s && n_out, "step: null argument");
    const int n = (int)std::min<size_t>(s->ev_used, (size_t)std::max(max_n, 0));
    for (int i = 0; i < n; i++) {
        DTCF_HIP(hipEventSynchronize(s->ev[i].second));
        DTCF_HIP(hipEventElapsedTime(&ms[i], s->ev[i].first, s->ev[i].second));
    }
    *n_out = n;
    s->ev_used = 0;
    return DTCF_OK;
}

// Records the start event of the next free pair (if timing is enabled).
static hipEvent_t *timing_pair(dtcf_step s)
{
    if (s->ev_used >= s->ev.size()) return nullptr;
    return &s->ev[s->ev_used].first;
}

static int gemm_timed(dtcf_step s, hipStream_t st, double alpha)
{
    hipEvent_t *e = timing_pair(s);
    if (e) DTCF_HIP(hipEventRecord(e[0], st));
    int r;
    if (s->mode == DTCF_STEP_FUSED)
        r = dtcf_gemm_view_f64(s->M, s->N, s->K, alpha, s->packA.empty() ? s->A.tiles : s->bigA,
                               s->va, s->packB.empty() ? s->B.tiles : s->bigB, s->vb, s->C.tiles,
                               s->vc, 0, st);
    else
        r = dtcf_gemm_f64(0, 1, s->M, s->N, s->K, alpha, s->bigA, s->K, s->bigB, s->K, s->bigC,
                          s->N, st);
    if (r != DTCF_OK) return r;
    if (e) {
        DTCF_HIP(hipEventRecord(s->ev[s->ev_used].second, st));
        s->ev_used++;
    }
    return DTCF_OK;
}

extern "C" int dtcf_step_run(dtcf_step s, double alpha, void *stream)
{
    DTCF_REQUIRE(s, "step: null handle");
    if (s->empty) return DTCF_OK;
    hipStream_t st = as_stream(stream);
    int r;
    // symmetric A / B: one permute pass per bounce class into the big matrix
    auto class_packs = [&](const OperandPlan &P, std::vector<ClassPack> &packs, double *big) {
        for (const ClassPack &k : packs) {
            if (k.n == 0) continue;
            const int q = dtcf_permute_copy_f64(P.tiles, big, &k.d, k.n, k.src, 0, k.dst, 0, 1.0,
                                                0.0, st);
            if (q != DTCF_OK) return q;
        }
        return (int)DTCF_OK;
    };
    if (s->mode == DTCF_STEP_FUSED) {
        if ((r = class_packs(s->B, s->packB, s->bigB)) != DTCF_OK) return r;
        if ((r = class_packs(s->A, s->packA, s->bigA)) != DTCF_OK) return r;
        return gemm_timed(s, st, alpha);
    }
    if (s->mode == DTCF_STEP_PACKED) {
        if ((r = dtcf_view_gather_f64(s->C.tiles, s->vc, s->bigC, s->N, st)) != DTCF_OK) return r;
        if ((r = s->packB.empty() ? dtcf_view_gather_f64(s->B.tiles, s->vb, s->bigB, s->K, st)
                                  : class_packs(s->B, s->packB, s->bigB)) != DTCF_OK)
            return r;
        if ((r = s->packA.empty() ? dtcf_view_gather_f64(s->A.tiles, s->va, s->bigA, s->K, st)
                                  : class_packs(s->A, s->packA, s->bigA)) != DTCF_OK)
            return r;
        if ((r = gemm_timed(s, st, alpha)) != DTCF_OK) return r;
        return dtcf_view_scatter_f64(s->bigC, s->N, s->vc, s->C.tiles, st);
    }
    // LITERAL: the reference's pass structure.
    const size_t baseA = 0, baseB = s->A.classes.size(), baseC = baseB + s->B.classes.size();
    if ((r = permute_operand(s, s->C, baseC, false, st)) != DTCF_OK) return r;
    if ((r = dtcf_view_gather_f64(s->tmp, s->vc, s->bigC, s->N, st)) != DTCF_OK) return r;
    if ((r = permute_operand(s, s->B, baseB, false, st)) != DTCF_OK) return r;
    if ((r = dtcf_view_gather_f64(s->tmp, s->vb, s->bigB, s->K, st)) != DTCF_OK) return r;
    if ((r = permute_operand(s, s->A, baseA, false, st)) != DTCF_OK) return r;
    if ((r = dtcf_view_gather_f64(s->tmp, s->va, s->bigA, s->K, st)) != DTCF_OK) return r;
    if ((r = gemm_timed(s, st, alpha)) != DTCF_OK) return r;
    if ((r = dtcf_view_scatter_f64(s->bigC, s->N, s->vc, s->tmp, st)) != DTCF_OK) return r;
    return permute_operand(s, s->C, baseC, true, st);
}
