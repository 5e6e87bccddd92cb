// DLTC tile contractions on gfx950 (host planner; no kernels here).
//
// A DLTC term tC += coef * tA * tB over dense tensors with uniform tiles
// (DLTC/expr.cc:13-67, DLTC/dltc_tensor.cc:169-314) runs as FP64 GEMMs on
// offset-table views of the three tensors:
//   rows of A and C   = the m indices (in A only), cols of B and C = n
//   indices (in B only), the contraction = k indices (in both),
// each kept in IndexName order as the reference's Multiplication does, so
// at = m+k, bt = k+n, ct = m+n.  The views are tile-major: row
//   r = rank(tile digits) * prod(wb) + rank(intra-tile digits),
// so one tile of the reference (one GA get / transpose of
// DLTC/task.cc:105-128) is one contiguous block of the view and a task over a
// tile range (DLTC/exec.cc:42-111 at expression, ext-tile or sum-tile
// granularity) is a GEMM over contiguous sub-views.  The transposes of
// task_mul (task.cc:115,121,131) and the coef-scaled accumulate
// (dltc_tensor.cc:278-305) are folded into the operand loads and the GEMM
// epilogue.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace dtcf {
namespace {

struct Space {
    int nocc, O, V, wbO, wbV;
    bool hol(int id) const { return id < nocc; }
    int range(int id) const { return hol(id) ? O : V; }
    int wb(int id) const { return hol(id) ? wbO : wbV; }
    int sp(int id) const { return hol(id) ? O / wbO : V / wbV; }
};

int check_space(const dtcf_dltc_space *s)
{
    DTCF_REQUIRE(s, "dltc: null space");
    DTCF_REQUIRE(s->nocc >= 0 && s->nocc <= DTCF_DLTC_NIDX, "dltc: nocc out of range");
    DTCF_REQUIRE(s->O > 0 && s->V > 0 && s->wbO > 0 && s->wbV > 0, "dltc: empty range or width");
    DTCF_REQUIRE(s->O % s->wbO == 0 && s->V % s->wbV == 0,
                 "dltc: ranges must be multiples of the tile widths");
    return DTCF_OK;
}

int check_tensor(const dtcf_dltc_tensor *t)
{
    DTCF_REQUIRE(t && t->dims >= 1 && t->dims <= DTCF_MAX_DIMS, "dltc: bad tensor dims");
    for (int i = 0; i < t->dims; i++) {
        DTCF_REQUIRE(t->ids[i] >= 0 && t->ids[i] < DTCF_DLTC_NIDX, "dltc: index name out of range");
        for (int j = 0; j < i; j++)
            DTCF_REQUIRE(t->ids[i] != t->ids[j], "dltc: repeated index name in one tensor");
    }
    return DTCF_OK;
}

// Dense row-major stride of each index name in a tensor (0 if absent).
void strides_of(const Space &S, const dtcf_dltc_tensor *t, int64_t *st)
{
    for (int i = 0; i < DTCF_DLTC_NIDX; i++) st[i] = 0;
    int64_t s = 1;
    for (int d = t->dims - 1; d >= 0; d--) {
        st[t->ids[d]] = s;
        s *= S.range(t->ids[d]);
    }
}

// Tile-major offsets along a list of index names: entry r for
// r = rank(tile digits over `ids`) * prod(wb) + rank(intra digits).
std::vector<int64_t> tile_major_offsets(const Space &S, const std::vector<int> &ids,
                                        const int64_t *st)
{
    int64_t ntile = 1, per = 1;
    for (int id : ids) {
        ntile *= S.sp(id);
        per *= S.wb(id);
    }
    std::vector<int64_t> out((size_t)(ntile * per));
    for (int64_t t = 0; t < ntile; t++) {
        int64_t tt = t, tbase = 0;
        for (int i = (int)ids.size() - 1; i >= 0; i--) {
            const int id = ids[i];
            tbase += (tt % S.sp(id)) * S.wb(id) * st[id];
            tt /= S.sp(id);
        }
        for (int64_t x = 0; x < per; x++) {
            int64_t xx = x, off = tbase;
            for (int i = (int)ids.size() - 1; i >= 0; i--) {
                const int id = ids[i];
                off += (xx % S.wb(id)) * st[id];
                xx /= S.wb(id);
            }
            out[(size_t)(t * per + x)] = off;
        }
    }
    return out;
}

// Maximal runs [first, last] of consecutive tile ranks inside a box of tile
// ranges over `ids` (rank = row-major over the tile counts).  An empty id
// list (a term with no m, n or k index) has the one rank 0.
std::vector<std::pair<int64_t, int64_t>> tile_runs(const Space &S, const std::vector<int> &ids,
                                                   const int32_t *lb, const int32_t *ub)
{
    std::vector<std::pair<int64_t, int64_t>> runs;
    const int n = (int)ids.size();
    std::vector<int> cur(n);
    for (int i = 0; i < n; i++) {
        cur[i] = lb[ids[i]];
        if (lb[ids[i]] >= ub[ids[i]]) return runs;
    }
    while (true) {
        int64_t rank = 0;
        for (int i = 0; i < n; i++) rank = rank * S.sp(ids[i]) + cur[i];
        if (!runs.empty() && runs.back().second + 1 == rank) runs.back().second = rank;
        else runs.push_back({rank, rank});
        int i = n - 1;
        for (; i >= 0; i--) {
            if (++cur[i] < ub[ids[i]]) break;
            cur[i] = lb[ids[i]];
        }
        if (i < 0) break;
    }
    return runs;  // no ids: the single rank 0 (one run), as the loop leaves it
}

}  // namespace
}  // namespace dtcf

using namespace dtcf;

struct dtcf_dltc_mul_s {
    Space S;
    std::vector<int> m, n, k;  // index names, IndexName order
    double *A = nullptr, *B = nullptr, *C = nullptr;
    double coef = 1.0;
    int64_t M = 1, N = 1, K = 1, mpb = 1, npb = 1, kpb = 1;
    dtcf_view va = nullptr, vb = nullptr, vc = nullptr;
};

extern "C" int dtcf_dltc_mul_destroy(dtcf_dltc_mul h)
{
    if (!h) return DTCF_OK;
    dtcf_view_destroy(h->va);
    dtcf_view_destroy(h->vb);
    dtcf_view_destroy(h->vc);
    delete h;
    return DTCF_OK;
}

extern "C" int dtcf_dltc_mul_create(const dtcf_dltc_space *space, const dtcf_dltc_tensor *tC,
                                    const dtcf_dltc_tensor *tA, const dtcf_dltc_tensor *tB,
                                    double coef, dtcf_dltc_mul *out)
{
    DTCF_REQUIRE(out, "dltc: null output");
    *out = nullptr;
    int st;
    if ((st = check_space(space)) || (st = check_tensor(tC)) || (st = check_tensor(tA)) ||
        (st = check_tensor(tB)))
        return st;
    DTCF_REQUIRE(tA->data && tB->data && tC->data, "dltc: null tensor data");
    Space S{space->nocc, space->O, space->V, space->wbO, space->wbV};
    // Multiplication ctor (expr.cc:29-50): tab = 1 (A) + 2 (B).
    int tab[DTCF_DLTC_NIDX] = {0};
    for (int i = 0; i < tA->dims; i++) tab[tA->ids[i]] |= 1;
    for (int i = 0; i < tB->dims; i++) tab[tB->ids[i]] |= 2;
    dtcf_dltc_mul_s *h = new dtcf_dltc_mul_s();
    h->S = S;
    for (int i = 0; i < DTCF_DLTC_NIDX; i++) {
        if (tab[i] == 1) h->m.push_back(i);
        if (tab[i] == 2) h->n.push_back(i);
        if (tab[i] == 3) h->k.push_back(i);
    }
    // C must hold exactly the external indices.
    std::vector<int> ext;
    for (int i = 0; i < DTCF_DLTC_NIDX; i++)
        if (tab[i] == 1 || tab[i] == 2) ext.push_back(i);
    std::vector<int> cids(tC->ids, tC->ids + tC->dims);
    std::sort(cids.begin(), cids.end());
    if (cids != ext) {
        delete h;
        set_error("dltc: C indices are not the external indices of A*B");
        return DTCF_ERR_INVALID;
    }
    h->A = tA->data;
    h->B = tB->data;
    h->C = tC->data;
    h->coef = coef;
    int64_t sa[DTCF_DLTC_NIDX], sb[DTCF_DLTC_NIDX], sc[DTCF_DLTC_NIDX];
    strides_of(S, tA, sa);
    strides_of(S, tB, sb);
    strides_of(S, tC, sc);
    for (int id : h->m) { h->M *= S.range(id); h->mpb *= S.wb(id); }
    for (int id : h->n) { h->N *= S.range(id); h->npb *= S.wb(id); }
    for (int id : h->k) { h->K *= S.range(id); h->kpb *= S.wb(id); }
    auto am = tile_major_offsets(S, h->m, sa), ak = tile_major_offsets(S, h->k, sa);
    auto bn = tile_major_offsets(S, h->n, sb), bk = tile_major_offsets(S, h->k, sb);
    auto cm = tile_major_offsets(S, h->m, sc), cn = tile_major_offsets(S, h->n, sc);
    if ((st = dtcf_view_create(h->M, h->K, am.data(), ak.data(), nullptr, 0, 0, &h->va)) ||
        (st = dtcf_view_create(h->N, h->K, bn.data(), bk.data(), nullptr, 0, 0, &h->vb)) ||
        (st = dtcf_view_create(h->M, h->N, cm.data(), cn.data(), nullptr, 0, 0, &h->vc))) {
        dtcf_dltc_mul_destroy(h);
        return st;
    }
    *out = h;
    return DTCF_OK;
}

extern "C" int dtcf_dltc_mul_info(dtcf_dltc_mul h, int64_t *mnk, int32_t *m_ids, int32_t *n_ids,
                                  int32_t *k_ids, int32_t *counts)
{
    DTCF_REQUIRE(h, "dltc: null handle");
    if (mnk) {
        mnk[0] = h->M;
        mnk[1] = h->N;
        mnk[2] = h->K;
    }
    if (counts) {
        counts[0] = (int32_t)h->m.size();
        counts[1] = (int32_t)h->n.size();
        counts[2] = (int32_t)h->k.size();
    }
    for (size_t i = 0; m_ids && i < h->m.size(); i++) m_ids[i] = h->m[i];
    for (size_t i = 0; n_ids && i < h->n.size(); i++) n_ids[i] = h->n[i];
    for (size_t i = 0; k_ids && i < h->k.size(); i++) k_ids[i] = h->k[i];
    return DTCF_OK;
}

// task_mul (DLTC/task.cc:83-137) over the tile box [itlb, itub) given per
// index name.  Each (m run, n run) pair is one GEMM per k run, k runs in
// increasing order.
extern "C" int dtcf_dltc_mul_task(dtcf_dltc_mul h, const int32_t *itlb, const int32_t *itub,
                                  void *stream)
{
    DTCF_REQUIRE(h && itlb && itub, "dltc: null argument");
    const Space &S = h->S;
    for (const std::vector<int> *v : {&h->m, &h->n, &h->k})
        for (int id : *v)
            DTCF_REQUIRE(itlb[id] >= 0 && itub[id] <= S.sp(id), "dltc: tile range outside tensor");
    auto mr = tile_runs(S, h->m, itlb, itub);
    auto nr = tile_runs(S, h->n, itlb, itub);
    auto kr = tile_runs(S, h->k, itlb, itub);
    for (auto &a : mr)
        for (auto &b : nr) {
            dtcf_view vc = nullptr;
            int st = dtcf_view_sub(h->vc, a.first * h->mpb, (a.second - a.first + 1) * h->mpb,
                                   b.first * h->npb, (b.second - b.first + 1) * h->npb, &vc);
            if (st) return st;
            for (auto &c : kr) {
                dtcf_view va = nullptr, vb = nullptr;
                const int64_t k0 = c.first * h->kpb, kn = (c.second - c.first + 1) * h->kpb;
                if ((st = dtcf_view_sub(h->va, a.first * h->mpb, (a.second - a.first + 1) * h->mpb,
                                        k0, kn, &va)) ||
                    (st = dtcf_view_sub(h->vb, b.first * h->npb, (b.second - b.first + 1) * h->npb,
                                        k0, kn, &vb))) {
                    dtcf_view_destroy(va);
                    dtcf_view_destroy(vc);
                    return st;
                }
                st = dtcf_gemm_view_f64((a.second - a.first + 1) * h->mpb,
                                        (b.second - b.first + 1) * h->npb, kn, h->coef, h->A, va,
                                        h->B, vb, h->C, vc, 0, stream);
                dtcf_view_destroy(va);
                dtcf_view_destroy(vb);
                if (st) {
                    dtcf_view_destroy(vc);
                    return st;
                }
            }
            dtcf_view_destroy(vc);
        }
    return DTCF_OK;
}

extern "C" int dtcf_dltc_mul_run(dtcf_dltc_mul h, void *stream)
{
    DTCF_REQUIRE(h, "dltc: null handle");
    return dtcf_gemm_view_f64(h->M, h->N, h->K, h->coef, h->A, h->va, h->B, h->vb, h->C, h->vc, 0,
                              stream);
}

// task_sum (DLTC/task.cc:51-81): tC += coef * perm(tA) over the tile box.
extern "C" int dtcf_dltc_sum(const dtcf_dltc_space *space, const dtcf_dltc_tensor *tC,
                             const dtcf_dltc_tensor *tA, double coef, const int32_t *itlb,
                             const int32_t *itub, void *stream)
{
    int st;
    if ((st = check_space(space)) || (st = check_tensor(tC)) || (st = check_tensor(tA)))
        return st;
    DTCF_REQUIRE(tC->dims == tA->dims, "dltc sum: tensors differ in rank");
    DTCF_REQUIRE(tC->data && tA->data, "dltc sum: null data");
    Space S{space->nocc, space->O, space->V, space->wbO, space->wbV};
    int64_t sa[DTCF_DLTC_NIDX], sc[DTCF_DLTC_NIDX];
    strides_of(S, tA, sa);
    strides_of(S, tC, sc);
    dtcf_strided d{};
    d.dims = tC->dims;
    int64_t src = 0, dst = 0;
    for (int i = 0; i < tC->dims; i++) {
        const int id = tC->ids[i];
        bool inA = false;
        for (int j = 0; j < tA->dims; j++) inA |= tA->ids[j] == id;
        DTCF_REQUIRE(inA, "dltc sum: C and A hold different indices");
        const int lo = itlb ? itlb[id] : 0, hi = itub ? itub[id] : S.sp(id);
        DTCF_REQUIRE(lo >= 0 && hi <= S.sp(id), "dltc sum: tile range outside tensor");
        if (lo >= hi) return DTCF_OK;
        d.ext[i] = (int64_t)(hi - lo) * S.wb(id);
        d.istr[i] = sa[id];
        d.ostr[i] = sc[id];
        src += (int64_t)lo * S.wb(id) * sa[id];
        dst += (int64_t)lo * S.wb(id) * sc[id];
    }
    return dtcf_permute_copy_f64(tA->data + src, tC->data + dst, &d, 1, nullptr, 0, nullptr, 0,
                                 coef, 1.0, stream);
}

// DLTC::transpose (DLTC/transp.cc:3-59) on one dense tile: output dim n is
// fed by input dim s[n], s = position(y, x); widths of x from the space.
extern "C" int dtcf_dltc_transpose_f64(const double *buf_x, const int32_t *x, double *buf_y,
                                       const int32_t *y, int dim, const dtcf_dltc_space *space,
                                       void *stream)
{
    int st = check_space(space);
    if (st) return st;
    if (dim != 2 && dim != 4) {
        set_error("transpose position error: DLTC transposes are 2-D or 4-D");
        return DTCF_ERR_UNSUPPORTED;
    }
    Space S{space->nocc, space->O, space->V, space->wbO, space->wbV};
    int64_t w[4];
    int32_t s[4];
    for (int i = 0; i < dim; i++) w[i] = S.wb(x[i]);
    for (int n = 0; n < dim; n++) {
        s[n] = -1;
        for (int i = 0; i < dim; i++)
            if (x[i] == y[n]) s[n] = i;
        if (s[n] < 0) {
            set_error("transpose position error: y holds an index that x does not");
            return DTCF_ERR_UNSUPPORTED;
        }
    }
    return dtcf_permute_tiles_f64(buf_x, buf_y, dim, w, s, 1, stream);
}
