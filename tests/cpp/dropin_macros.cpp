// Expands the header drop-ins -- DTCF_RRR_DROPIN_DEFINE (create_big_matrix,
// create_big_matrix_nl, revert_big_matrix), DTCF_DLTC_DROPIN_DEFINE (compute)
// and DTCF_TASK_DROPIN_DEFINE (TaskFunc) -- in one translation unit, the way
// a reference build would, against stand-ins that declare what the
// reference's headers declare for them: RRR::Tensor's accessors
// (tensor.h:409-414), DLTC's IndexName / ExprType / NIDX (param.h:4-25),
// getSize (define.h:87-93), Multiplication's index lists (expr.h:56-58),
// TaskDesc (task.h:17-23) and the Scioto task body.  Each result is checked
// against a brute-force loop in this file.  Exit status 0 = pass.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <random>
#include <vector>

namespace RRR {
using namespace std;
class Tensor {
  public:
    int dims = 0, block_size = 0;
    int *vgrid = nullptr;
    int get_block_size() { return block_size; }
    int get_dims() { return dims; }
    int *get_vgrid() { return vgrid; }
};
}  // namespace RRR

namespace DLTC {
using namespace std;
enum ExprType { MUL, SUM };
enum IndexType { HOL, PAR };
enum IndexName { h1, h2, h3, h4, h5, h6, h7, h8, h9, h10, h11, h12,
                 p1, p2, p3, p4, p5, p6, p7, p8, p9, p10, p11, p12 };
#define NIDX 24
static int wbO = 2, wbV = 4;
inline IndexType typeOf(IndexName x) { return x < p1 ? HOL : PAR; }
inline uint64_t getSize(const vector<IndexName> &ids)
{
    uint64_t size = 1;
    for (size_t i = 0; i < ids.size(); i++) size *= typeOf(ids[i]) == HOL ? wbO : wbV;
    return size;
}
class Multiplication {
  public:
    vector<IndexName> m_, n_, k_;
    vector<IndexName> m_ids() { return m_; }
    vector<IndexName> n_ids() { return n_; }
    vector<IndexName> k_ids() { return k_; }
};
struct TaskDesc {
    ExprType et;
    int id;
    int itlb[NIDX];
    int itub[NIDX];
    int idvm[NIDX];
};
}  // namespace DLTC

// Scioto: the task pool handle and a task whose body is a TaskDesc
typedef struct gtc_stub *gtc_t;
typedef gtc_t TaskPool;
struct task_t {
    DLTC::TaskDesc body;
};
inline void *gtc_task_body(task_t *t) { return &t->body; }

#include "dtcf_dropin_dltc.h"
#include "dtcf_dropin_rrr.h"
#include "dtcf_dropin_task.h"

DTCF_RRR_DROPIN_DEFINE()
DTCF_DLTC_DROPIN_DEFINE()
DTCF_TASK_DROPIN_DEFINE()

static int failures = 0;
#define CHECK(cond, ...)                       \
    do {                                       \
        if (!(cond)) {                         \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");                 \
            failures++;                        \
        }                                      \
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

static int rank_of(const int *a, const int *grid, int dims)
{
    int r = 0;
    for (int d = 0; d < dims; d++) r = r * grid[d] + a[d];
    return r;
}

static void check_big_matrix()
{
    // permuted tile addresses of a 4-D tensor over a permuted grid of
    // 2 x 3 x 2 x 2 tiles of 2 x 3 x 4 x 2 elements; rows = first 2 dims
    const int dims = 4, nrow = 2;
    int vgrid[4] = {2, 2, 3, 2};   // tensor grid; p_map sends dim d to position pmap[d]
    int pmap_a[4] = {0, 2, 1, 3};
    int grid[4];
    for (int d = 0; d < 4; d++) grid[pmap_a[d]] = vgrid[d];
    const int bw[4] = {2, 3, 4, 2}, bs = 48, rpb = bw[0] * bw[1], cpb = bw[2] * bw[3];
    std::vector<int> addr;  // every tile of the permuted grid, in a scrambled order
    std::vector<int> ranks;
    for (int r = 0; r < 24; r++) ranks.push_back((r * 7) % 24);
    for (int r : ranks) {
        int a[4], x = r;
        for (int d = 3; d >= 0; d--) {
            a[d] = x % grid[d];
            x /= grid[d];
        }
        addr.insert(addr.end(), a, a + 4);
    }
    const int n = (int)ranks.size();
    auto tiles = rnd((size_t)n * bs, 7);
    RRR::Tensor T;
    T.dims = dims;
    T.block_size = bs;
    T.vgrid = vgrid;
    RRR::Tensor *Tp = &T;
    double *tp = tiles.data();
    int *ap = addr.data(), *pm = pmap_a;
    std::map<int, std::map<int, int> > *bm = nullptr;
    double *big = nullptr;
    RRR::create_big_matrix(Tp, tp, ap, pm, n, rpb, cpb, nrow, bm, big);
    // brute force: row key = rank of (a0, a1, 0, 0), column key = rank of (0, 0, a2, a3)
    const int nrb = grid[0] * grid[1], ncb = grid[2] * grid[3];
    std::vector<double> ref((size_t)nrb * ncb * bs);
    for (int t = 0; t < n; t++) {
        const int *a = &addr[t * 4];
        const int R = a[0] * grid[1] + a[1], Cb = a[2] * grid[3] + a[3];
        for (int rr = 0; rr < rpb; rr++)
            for (int cc = 0; cc < cpb; cc++)
                ref[((size_t)R * rpb + rr) * ncb * cpb + (size_t)Cb * cpb + cc] =
                    tiles[(size_t)t * bs + rr * cpb + cc];
    }
    CHECK(std::vector<double>(big, big + ref.size()) == ref, "create_big_matrix placement");
    CHECK((int)bm->size() == nrb, "create_big_matrix rows %d", (int)bm->size());
    // _nl: tiles stored at location loc(t), found through the rank -> location map
    std::vector<double> shuffled(tiles.size());
    std::map<int, int> a2l;
    for (int t = 0; t < n; t++) {
        const int loc = (t * 5 + 3) % n;
        std::copy(tiles.begin() + (size_t)t * bs, tiles.begin() + (size_t)(t + 1) * bs,
                  shuffled.begin() + (size_t)loc * bs);
        a2l[rank_of(&addr[t * 4], grid, 4)] = loc;
    }
    std::map<int, std::map<int, int> > *bm2 = nullptr;
    std::map<int, int> *a2lp = &a2l;
    double *big2 = nullptr, *sp = shuffled.data();
    RRR::create_big_matrix_nl(Tp, sp, ap, pm, n, rpb, cpb, nrow, bm2, big2, a2lp);
    CHECK(std::vector<double>(big2, big2 + ref.size()) == ref, "create_big_matrix_nl placement");
    // revert: scale the big matrix, write it back into the tiles
    for (size_t i = 0; i < ref.size(); i++) big[i] *= -2.0;
    std::vector<double> back(tiles.size(), 0.0);
    double *bp = back.data();
    RRR::revert_big_matrix(Tp, bm, big, rpb, cpb, pm, bp);
    bool ok = true;
    for (size_t i = 0; i < tiles.size(); i++) ok = ok && back[i] == -2.0 * tiles[i];
    CHECK(ok, "revert_big_matrix round trip");
    delete[] big;
    delete[] big2;
    delete bm;
    delete bm2;
}

static void check_compute()
{
    // buf_c[m x n] += buf_a[m x k] buf_b[k x n], m = (h1,h2), n = (p3,p4), k = (p5,p6)
    DLTC::Multiplication M;
    M.m_ = {DLTC::h1, DLTC::h2};
    M.n_ = {DLTC::p3, DLTC::p4};
    M.k_ = {DLTC::p5, DLTC::p6};
    const int m = DLTC::wbO * DLTC::wbO, n = DLTC::wbV * DLTC::wbV, k = n;
    auto A = rnd(m * k, 11), B = rnd(k * n, 12), C = rnd(m * n, 13);
    auto R = C;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++)
            for (int p = 0; p < k; p++) R[i * n + j] += A[i * k + p] * B[p * n + j];
    DLTC::compute(C.data(), A.data(), B.data(), M);
    CHECK(maxrel(C, R) < 1e-12, "compute rel %.3e", maxrel(C, R));
}

static void check_taskfunc()
{
    // tC[p3,p4,h1,h2] += 0.5 sum_{p5,p6} tA[p5,p6,h1,h2] tB[p3,p4,p5,p6]   (MUL, id 0)
    // tD[p3,p4,h1,h2] += -1.5 tX[p4,p3,h2,h1]                             (SUM, id 0)
    const int O = 4, V = 8, wbO = DLTC::wbO, wbV = DLTC::wbV;
    const size_t nvo = (size_t)V * V * O * O, nv4 = (size_t)V * V * V * V;
    auto A = rnd(nvo, 21), B = rnd(nv4, 22), C0 = rnd(nvo, 23), X = rnd(nvo, 24), D0 = rnd(nvo, 25);
    double *dA, *dB, *dC, *dX, *dD;
    hipMalloc(&dA, nvo * 8);
    hipMalloc(&dB, nv4 * 8);
    hipMalloc(&dC, nvo * 8);
    hipMalloc(&dX, nvo * 8);
    hipMalloc(&dD, nvo * 8);
    hipMemcpy(dA, A.data(), nvo * 8, hipMemcpyHostToDevice);
    hipMemcpy(dB, B.data(), nv4 * 8, hipMemcpyHostToDevice);
    hipMemcpy(dC, C0.data(), nvo * 8, hipMemcpyHostToDevice);
    hipMemcpy(dX, X.data(), nvo * 8, hipMemcpyHostToDevice);
    hipMemcpy(dD, D0.data(), nvo * 8, hipMemcpyHostToDevice);
    dtcf_dltc_space sp = {12, O, V, wbO, wbV};
    using namespace DLTC;
    auto tens = [](std::vector<int> ids, double *d) {
        dtcf_dltc_tensor t{};
        t.dims = (int)ids.size();
        for (size_t i = 0; i < ids.size(); i++) t.ids[i] = ids[i];
        t.data = d;
        return t;
    };
    dtcf_dltc_tensor tC = tens({p3, p4, h1, h2}, dC), tA = tens({p5, p6, h1, h2}, dA),
                     tB = tens({p3, p4, p5, p6}, dB), tD = tens({p3, p4, h1, h2}, dD),
                     tX = tens({p4, p3, h2, h1}, dX);
    dtcf_dltc_mul mh = nullptr;
    CHECK(dtcf_dltc_mul_create(&sp, &tC, &tA, &tB, 0.5, &mh) == DTCF_OK, "mul_create");
    auto &reg = dtcf_task::registry();
    reg.mul = {mh};
    reg.sum.resize(1);
    reg.sum[0] = dtcf_task::Sum{sp, tD, tX, -1.5, true};
    const int so = O / wbO, sv = V / wbV;
    int ntask = 0;
    for (int a = 0; a < sv; a++)
        for (int b = 0; b < sv; b++)
            for (int i = 0; i < so; i++)
                for (int j = 0; j < so; j++) {
                    for (int c = 0; c < sv; c++)
                        for (int d = 0; d < sv; d++) {  // one (ext tile, sum tile) MUL task
                            task_t task{};
                            task.body.et = MUL;
                            task.body.id = 0;
                            const int idx[6] = {p3, p4, h1, h2, p5, p6}, val[6] = {a, b, i, j, c, d};
                            for (int q = 0; q < 6; q++) {
                                task.body.itlb[idx[q]] = val[q];
                                task.body.itub[idx[q]] = val[q] + 1;
                            }
                            DLTC::TaskFunc(nullptr, &task);
                            ntask++;
                        }
                    task_t task{};  // one ext tile SUM task
                    task.body.et = SUM;
                    task.body.id = 0;
                    const int idx[4] = {p3, p4, h1, h2}, val[4] = {a, b, i, j};
                    for (int q = 0; q < 4; q++) {
                        task.body.itlb[idx[q]] = val[q];
                        task.body.itub[idx[q]] = val[q] + 1;
                    }
                    DLTC::TaskFunc(nullptr, &task);
                }
    hipDeviceSynchronize();
    std::vector<double> C(nvo), D(nvo);
    hipMemcpy(C.data(), dC, nvo * 8, hipMemcpyDeviceToHost);
    hipMemcpy(D.data(), dD, nvo * 8, hipMemcpyDeviceToHost);
    auto Cr = C0, Dr = D0;
    for (int a = 0; a < V; a++)
        for (int b = 0; b < V; b++)
            for (int i = 0; i < O; i++)
                for (int j = 0; j < O; j++) {
                    double s = 0.0;
                    for (int c = 0; c < V; c++)
                        for (int d = 0; d < V; d++)
                            s += A[((size_t)(c * V + d) * O + i) * O + j] *
                                 B[((size_t)(a * V + b) * V + c) * V + d];
                    const size_t e = ((size_t)(a * V + b) * O + i) * O + j;
                    Cr[e] += 0.5 * s;
                    Dr[e] += -1.5 * X[((size_t)(b * V + a) * O + j) * O + i];
                }
    CHECK(maxrel(C, Cr) < 1e-12, "TaskFunc MUL (%d tasks) rel %.3e", ntask, maxrel(C, Cr));
    CHECK(maxrel(D, Dr) < 1e-15, "TaskFunc SUM rel %.3e", maxrel(D, Dr));
    dtcf_dltc_mul_destroy(mh);
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    hipFree(dX);
    hipFree(dD);
}

int main()
{
    check_big_matrix();
    check_compute();
    check_taskfunc();
    if (failures == 0) std::printf("PASS (header drop-ins expanded: RRR, DLTC compute, TaskFunc)\n");
    return failures == 0 ? 0 : 1;
}
