// C++ caller of the drop-in symbols (libdtcf_dropin.so) and of the header
// adapters, written the way the reference's own code calls them.  Each check
// compares against a brute-force loop in this file.  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "dtcf_dropin.h"
#include "dtcf_dropin_dltc.h"
#include "dtcf_dropin_rrr.h"

static int failures = 0;
#define CHECK(cond, ...)                              \
    do {                                              \
        if (!(cond)) {                                \
            std::printf("FAIL: " __VA_ARGS__);        \
            std::printf("\n");                        \
            failures++;                               \
        }                                             \
    } while (0)

static std::vector<double> rnd(size_t n, unsigned seed)
{
    std::mt19937_64 g(seed);
    std::normal_distribution<double> d(0.0, 1.0);
    std::vector<double> v(n);
    for (auto &x : v) x = d(g);
    return v;
}

static double maxrel(const std::vector<double> &a, const std::vector<double> &b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(b[i]));
    }
    return den > 0 ? num / den : num;
}

// out[n-th output dim iterates input dim s[n]] for a 4-D tile
static std::vector<double> brute_perm4(const std::vector<double> &in, const int *w, const int *s)
{
    int ow[4];
    for (int n = 0; n < 4; n++) ow[n] = w[s[n]];
    std::vector<double> out(in.size());
    int x[4];
    for (x[0] = 0; x[0] < w[0]; x[0]++)
        for (x[1] = 0; x[1] < w[1]; x[1]++)
            for (x[2] = 0; x[2] < w[2]; x[2]++)
                for (x[3] = 0; x[3] < w[3]; x[3]++) {
                    const size_t src = ((size_t)(x[0] * w[1] + x[1]) * w[2] + x[2]) * w[3] + x[3];
                    const size_t dst =
                        ((size_t)(x[s[0]] * ow[1] + x[s[1]]) * ow[2] + x[s[2]]) * ow[3] + x[s[3]];
                    out[dst] = in[src];
                }
    return out;
}

struct MockTensor {  // the accessors create_big_matrix reads (tensor.h)
    int dims, bs;
    int *vgrid;
    int get_dims() { return dims; }
    int *get_vgrid() { return vgrid; }
    int get_block_size() { return bs; }
};

struct MockMul {  // the accessors compute reads (expr.h)
    std::vector<int> m, n, k;
    std::vector<int> m_ids() { return m; }
    std::vector<int> n_ids() { return n; }
    std::vector<int> k_ids() { return k; }
};

int main()
{
    // 1. local_contraction, both call forms of the reference
    {
        const int na = 70, nb = 40, nk = 55;
        auto A = rnd(na * nk, 1), B = rnd(nb * nk, 2), C0 = rnd(na * nb, 3);
        for (int bt = 0; bt <= 1; bt++) {
            std::vector<double> C = C0, R = C0;
            local_contraction(na, nb, nk, A.data(), B.data(), C.data(), 0, bt, 0.5);
            for (int i = 0; i < na; i++)
                for (int j = 0; j < nb; j++) {
                    double s = 0;
                    for (int p = 0; p < nk; p++)
                        s += A[i * nk + p] * (bt ? B[j * nk + p] : B[p * nb + j]);
                    R[i * nb + j] += 0.5 * s;
                }
            CHECK(maxrel(C, R) < 1e-12, "local_contraction bt=%d rel %.3e", bt, maxrel(C, R));
        }
    }
    // 2. RRR transpose: input dim i -> output position pmap[i]
    {
        int w[4] = {3, 5, 4, 6};
        int pmap[4] = {2, 0, 3, 1};
        auto in = rnd(360, 4);
        std::vector<double> out(360);
        double *pi = in.data(), *po = out.data();
        transpose(pi, po, pmap, 4, w);
        int s[4];
        for (int i = 0; i < 4; i++) s[pmap[i]] = i;
        CHECK(out == brute_perm4(in, w, s), "RRR transpose");
    }
    // 3. the 26 DLTC kernels
    {
        int w[4] = {5, 6, 7, 3};
        auto in = rnd(630, 5);
        typedef void (*K4)(const double *, double *, int, int, int, int);
        struct { const char *name; K4 f; } ks[] = {
            {"ijkl", transpose_ijkl}, {"ijlk", transpose_ijlk}, {"ikjl", transpose_ikjl},
            {"iklj", transpose_iklj}, {"iljk", transpose_iljk}, {"ilkj", transpose_ilkj},
            {"jikl", transpose_jikl}, {"jilk", transpose_jilk}, {"jkil", transpose_jkil},
            {"jkli", transpose_jkli}, {"jlik", transpose_jlik}, {"jlki", transpose_jlki},
            {"kijl", transpose_kijl}, {"kilj", transpose_kilj}, {"kjil", transpose_kjil},
            {"kjli", transpose_kjli}, {"klij", transpose_klij}, {"klji", transpose_klji},
            {"lijk", transpose_lijk}, {"likj", transpose_likj}, {"ljik", transpose_ljik},
            {"ljki", transpose_ljki}, {"lkij", transpose_lkij}, {"lkji", transpose_lkji}};
        for (auto &k : ks) {
            std::vector<double> out(630);
            k.f(in.data(), out.data(), w[0], w[1], w[2], w[3]);
            int s[4];
            for (int n = 0; n < 4; n++) s[n] = k.name[n] - 'i';
            CHECK(out == brute_perm4(in, w, s), "transpose_%s", k.name);
        }
        std::vector<double> o2(42), i2 = rnd(42, 6);
        transpose_ji(i2.data(), o2.data(), 6, 7);
        bool ok = true;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 7; j++) ok &= o2[j * 6 + i] == i2[i * 7 + j];
        CHECK(ok, "transpose_ji");
        transpose_ij(i2.data(), o2.data(), 6, 7);
        CHECK(o2 == i2, "transpose_ij");
    }
    // 4. Block_Grid
    {
        int g[3] = {3, 4, 5};
        RRR::Block_Grid bg(3, g);
        bool ok = bg.nblocks == 60 && bg.rank == 0;
        for (int r = 0; r < 60; r++) {
            int *a = nullptr;
            bg.get_block_addr(r, a);
            ok &= a[0] == r / 20 && a[1] == (r / 5) % 4 && a[2] == r % 5;
            ok &= bg.get_block_rank(a) == r;
            delete[] a;
        }
        CHECK(ok, "Block_Grid rank/address");
    }
    // 5. DLTC::transpose with index-name vectors (widths from wbO / wbV)
    {
        using namespace DLTC;
        dtcf_dltc_set_widths(2, 3, 12);
        std::vector<IndexName> x = {p5, p6, h1, h2}, y = {h1, h2, p5, p6};
        auto in = rnd(36, 7);
        std::vector<double> out(36);
        DLTC::transpose(in.data(), x, out.data(), y);
        int w[4] = {3, 3, 2, 2}, s[4] = {2, 3, 0, 1};
        CHECK(out == brute_perm4(in, w, s), "DLTC::transpose klij");
        std::vector<double> out2(36, -1.0);
        DLTC::transpose(in.data(), x, out2.data(), x);
        CHECK(out2 == in, "DLTC::transpose identity copies");
    }
    // 6. create_big_matrix / revert_big_matrix adapters
    {
        int vg[4] = {2, 3, 2, 2};
        int pmap[4] = {0, 2, 1, 3};
        MockTensor mt{4, 12, vg};  // tiles of 12 elements: rpb 3 x cpb 4
        MockTensor *T = &mt;
        std::vector<int> addr;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 2; c++)
                    for (int d = 0; d < 2; d++) addr.insert(addr.end(), {a, b, c, d});
        const int nt = 24;
        std::vector<int> paddr(nt * 4);
        for (int t = 0; t < nt; t++)
            for (int d = 0; d < 4; d++) paddr[t * 4 + pmap[d]] = addr[t * 4 + d];
        auto tiles = rnd(nt * 12, 8);
        double *tp = tiles.data(), *big = nullptr;
        int *pa = paddr.data(), *pm = pmap;
        std::map<int, std::map<int, int> > *bm = nullptr;
        dtcf_rrr::create_big_matrix(T, tp, pa, pm, nt, 3, 4, 2, bm, big, nullptr);
        // brute force: row key (a, c), col key (b, d) in the permuted grid (2,2,3,2)
        bool ok = bm->size() == 4 && bm->begin()->second.size() == 6;
        int R = 0;
        for (auto &row : *bm) {
            int Cc = 0;
            for (auto &col : row.second) {
                const int t = col.second;
                for (int rr = 0; rr < 3; rr++)
                    for (int cc = 0; cc < 4; cc++)
                        ok &= big[(R * 3 + rr) * 24 + Cc * 4 + cc] == tiles[t * 12 + rr * 4 + cc];
                Cc++;
            }
            R++;
        }
        CHECK(ok, "create_big_matrix layout");
        std::vector<double> back(nt * 12, 0.0);
        double *bp = back.data();
        dtcf_rrr::revert_big_matrix(T, bm, big, 3, 4, pm, bp);
        CHECK(back == tiles, "revert_big_matrix round trip");
        delete[] big;
        delete bm;
    }
    // 7. compute adapter: buf_c[m x n] += buf_a[m x k] * buf_b[k x n]
    {
        MockMul M{{0, 1}, {12}, {13, 14}};
        auto width = [](const std::vector<int> &ids) {
            size_t s = 1;
            for (int id : ids) s *= id < 12 ? 3 : 4;
            return s;
        };
        const int m = 9, n = 4, k = 16;
        auto a = rnd(m * k, 9), b = rnd(k * n, 10), c = rnd(m * n, 11);
        std::vector<double> ref = c;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++)
                for (int p = 0; p < k; p++) ref[i * n + j] += a[i * k + p] * b[p * n + j];
        dtcf_dltc::compute(c.data(), a.data(), b.data(), M, width);
        CHECK(maxrel(c, ref) < 1e-12, "compute rel %.3e", maxrel(c, ref));
    }
    std::printf("dropin_selftest: %s (%d failures)\n", failures ? "FAIL" : "PASS", failures);
    return failures ? 1 : 0;
}
