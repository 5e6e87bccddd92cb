// Serving runtime: CC iteration programs whose tensors stay resident in HBM,
// run as continuously batched requests under an HBM budget.
//
// Replaces, for one GPU, the DLTC task pool and tile cache: task_mul /
// task_sum (DLTC/task.cc:51-137) become statements of a program, the Scioto
// pool and the layered barriers (DLTC/task.cc:24-37, ccd-t2.cc:97-131)
// become a dependency-driven issue loop over HIP streams, and the LRU tile
// cache (DLTC/cache.h:23-133) becomes residency: inputs live in the arena for
// the server's life, temporaries live in a per-request block.
//
// Memory.  One hipMalloc'd arena of `budget` bytes, sub-allocated first-fit
// (256-byte granules) on the host.  Finalizing a program places every
// temporary (and each mixed-precision contraction's slice scratch) at a
// fixed offset of a request block by walking the statements in program
// order, either each in bytes of its own or, when that block would be much
// larger, reusing the bytes of temporaries that died earlier; wherever a
// placement reuses bytes, the statement gets an extra dependency on every
// reader of the previous occupant (write-after-read), so any
// dependency-respecting issue order is safe and a request never allocates
// after admission.  A request is admitted when its block fits in the arena.
//
// Issue.  A scheduler thread keeps `nslots` non-blocking streams.  Each
// statement of an admitted request becomes ready when all its dependencies
// have completed (observed with hipEventQuery on the host); ready statements
// go to free slots from the request with the least total work first (then
// arrival order, then program order), so short requests are not queued
// behind long ones.  A
// statement runs entirely on one stream with kernels whose reduction order
// is fixed by the shapes alone, so a request's results are bitwise the same
// alone and beside other requests.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "gemm_f64.hpp"
#include "view.hpp"

namespace dtcf {
namespace {

constexpr int kMaxSumTerms = 96;
constexpr int64_t kGranule = 256;

inline int64_t granule(int64_t b) { return (std::max<int64_t>(b, 1) + kGranule - 1) / kGranule * kGranule; }

// ---------------------------------------------------------------- kernels
__global__ __launch_bounds__(256) void zero_kernel(double *p, int64_t n)
{
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = 0.0;
}

typedef double d2v __attribute__((ext_vector_type(2)));

// out[i0,i1,i2,i3] = sum_t coef[t] * src_t[i0*s_t0 + i1*s_t1 + i2*s_t2 + i3*s_t3]
// (terms added in list order).
struct SumArgs {
    double *out;
    int64_t n, d1, d2, d3;
    int32_t nterms;
    const int64_t *strides;  // device, [nterms][4]
    const double *src[kMaxSumTerms];
    double coef[kMaxSumTerms];
    // sum_plane_kernel<true>: terms whose element pairs (i3, i3 + 1) are
    // adjacent and 16-byte aligned (unit i3 stride, even other strides,
    // aligned buffer), one bit per term
    uint64_t pairs[(kMaxSumTerms + 63) / 64];
};

__global__ __launch_bounds__(256) void sum_kernel(SumArgs a)
{
    __shared__ int64_t st[kMaxSumTerms * 4];
    __shared__ const double *src[kMaxSumTerms];
    __shared__ double coef[kMaxSumTerms];
    for (int i = threadIdx.x; i < a.nterms * 4; i += 256) st[i] = a.strides[i];
    if (threadIdx.x < a.nterms) {
        src[threadIdx.x] = a.src[threadIdx.x];
        coef[threadIdx.x] = a.coef[threadIdx.x];
    }
    __syncthreads();
    for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < a.n; e += (int64_t)gridDim.x * 256) {
        int64_t r = e;
        const int64_t i3 = r % a.d3;
        r /= a.d3;
        const int64_t i2 = r % a.d2;
        r /= a.d2;
        const int64_t i1 = r % a.d1;
        const int64_t i0 = r / a.d1;
        double s = 0.0;
        for (int t = 0; t < a.nterms; t++) {
            const int64_t *q = st + 4 * t;
            s += coef[t] * src[t][i0 * q[0] + i1 * q[1] + i2 * q[2] + i3 * q[3]];
        }
        a.out[e] = s;
    }
}

// The same sums, one (i2, i3) plane of the output per workgroup pass: the
// result element (i0, i1, i2, i3) adds its terms in list order exactly as
// sum_kernel does.  Terms go in chunks of kChunk: every term of a chunk is
// loaded before one barrier, so kChunk terms' loads are in flight at once.
// A term whose unit stride lies on i2 (the occupied pair transposed, e.g.
// X[p2,p1,h2,h1] into R[p1,p2,h1,h2]) is staged through its own LDS plane
// with lanes along i2; any other term is read directly into registers
// (lanes along i3: whole lines when its unit stride lies there).  (i2, i3)
// come from 32-bit divisions by invariants, (i0, i1) once per plane.
constexpr int kPlaneCap = 3072;  // doubles per staged plane: d3 rows of d2 + 1
constexpr int kPlaneThreads = 512;
constexpr int kPlanePer = 6;     // result elements per thread: d2 * d3 <= 3072
constexpr int kChunk = 4;        // terms per barrier

// PAIRS (d2 and d3 even): each thread owns element pairs (i3, i3 + 1) of one
// i2 -- six pairs for the 3072-element plane limit -- read with one 16-byte
// load from the terms marked in `pairs` and stored with one 16-byte store;
// every element still adds its terms in list order, with the same
// operations as the element form.
template <bool PAIRS>
__global__ __launch_bounds__(kPlaneThreads, 4) void sum_plane_kernel(SumArgs a, FastDiv div_d3,
                                                                  FastDiv div_d2, uint64_t nplanes,
                                                                  uint32_t bufstride)
{
    constexpr uint32_t T = kPlaneThreads;
    extern __shared__ double planes[];  // kChunk staged planes of bufstride doubles
    __shared__ int64_t st[kMaxSumTerms * 4];
    __shared__ const double *src[kMaxSumTerms];
    __shared__ double coef[kMaxSumTerms];
    for (int i = threadIdx.x; i < a.nterms * 4; i += T) st[i] = a.strides[i];
    if (threadIdx.x < a.nterms) {
        src[threadIdx.x] = a.src[threadIdx.x];
        coef[threadIdx.x] = a.coef[threadIdx.x];
    }
    __syncthreads();
    constexpr int W = PAIRS ? 2 : 1;          // elements per slot
    constexpr int NS = kPlanePer / W;         // slots per thread
    const uint32_t d2 = (uint32_t)a.d2, d3 = (uint32_t)a.d3, P = d2 * d3, pitch = d2 + 1;
    uint32_t i2[NS], i3[NS];
    bool in[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const uint32_t e0 = W * (threadIdx.x + T * k);
        in[k] = e0 < P;
        const uint32_t e = min<uint32_t>(e0, P - W);
        i2[k] = div_d3.div(e);
        i3[k] = e - i2[k] * d3;
    }
    // Workgroups go to the XCDs in turn; when the grid allows, consecutive
    // planes are given to workgroups of one XCD, so terms whose neighbouring
    // planes share cache lines (the plane's indices outermost in the term)
    // find those lines in that XCD's L2.
    uint64_t first = blockIdx.x;
    if ((gridDim.x & 7u) == 0) first = (uint64_t)(blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    for (uint64_t plane = first; plane < nplanes; plane += gridDim.x) {
        const uint64_t p0 = plane / (uint64_t)a.d1, p1 = plane - p0 * (uint64_t)a.d1;
        double acc[NS][W];
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
            for (int w = 0; w < W; w++) acc[k][w] = 0.0;
        for (int t0 = 0; t0 < a.nterms; t0 += kChunk) {
            double xr[kChunk][NS][W];
            bool staged[kChunk];
            __syncthreads();  // the previous chunk's staged planes have been read
#pragma unroll
            for (int j = 0; j < kChunk; j++) {
                const int t = t0 + j;
                staged[j] = false;
                if (t >= a.nterms) continue;
                const int64_t *q = st + 4 * t;
                const double *x = src[t] + (int64_t)p0 * q[0] + (int64_t)p1 * q[1];
                if (q[2] == 1 && q[3] != 1) {
                    staged[j] = true;
                    double *b = planes + j * bufstride;
                    for (uint32_t idx = threadIdx.x; idx < P; idx += T) {
                        const uint32_t r3 = div_d2.div(idx), r2 = idx - r3 * d2;
                        b[r3 * pitch + r2] = x[r2 + (int64_t)r3 * q[3]];
                    }
                } else if (PAIRS && ((a.pairs[t >> 6] >> (t & 63)) & 1)) {
#pragma unroll
                    for (int k = 0; k < NS; k++) {
                        d2v v = in[k] ? *(const d2v *)(x + (int64_t)i2[k] * q[2] + i3[k]) : d2v{0.0, 0.0};
                        xr[j][k][0] = v.x;
                        xr[j][k][W - 1] = v.y;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NS; k++)
#pragma unroll
                        for (int w = 0; w < W; w++)
                            xr[j][k][w] = in[k] ? x[(int64_t)i2[k] * q[2] + (int64_t)(i3[k] + w) * q[3]]
                                                : 0.0;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kChunk; j++) {
                const int t = t0 + j;
                if (t >= a.nterms) continue;
                const double c = coef[t];
                const double *b = planes + j * bufstride;
#pragma unroll
                for (int k = 0; k < NS; k++)
#pragma unroll
                    for (int w = 0; w < W; w++)
                        acc[k][w] += c * (staged[j] ? b[(i3[k] + w) * pitch + i2[k]] : xr[j][k][w]);
            }
        }
        double *o = a.out + plane * P;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            if (!in[k]) continue;
            const uint32_t e = W * (threadIdx.x + T * k);
            if constexpr (PAIRS) *(d2v *)(o + e) = d2v{acc[k][0], acc[k][1]};
            else o[e] = acc[k][0];
        }
    }
}

unsigned grid_for(int64_t n)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

// ---------------------------------------------------------------- arena
// First-fit over a free map (offset -> length) with coalescing.
struct Arena {
    int64_t size = 0, used = 0, high = 0;
    std::map<int64_t, int64_t> fr;
    void init(int64_t s)
    {
        size = s;
        used = high = 0;
        fr.clear();
        fr[0] = s;
    }
    int64_t alloc(int64_t bytes)
    {
        bytes = granule(bytes);
        for (auto it = fr.begin(); it != fr.end(); ++it)
            if (it->second >= bytes) {
                const int64_t off = it->first, len = it->second;
                fr.erase(it);
                if (len > bytes) fr[off + bytes] = len - bytes;
                used += bytes;
                high = std::max(high, off + bytes);
                return off;
            }
        return -1;
    }
    void release(int64_t off, int64_t bytes)
    {
        bytes = granule(bytes);
        used -= bytes;
        auto it = fr.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != fr.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fr.erase(nx);
        }
        if (it != fr.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fr.erase(it);
            }
        }
    }
};

double now_us(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
}

std::vector<std::string> split_names(const char *s)
{
    std::vector<std::string> out;
    std::string cur;
    for (const char *p = s; p && *p; p++) {
        if (*p == ',') {
            out.push_back(cur);
            cur.clear();
        } else if (*p != ' ') {
            cur.push_back(*p);
        }
    }
    if (!cur.empty() || (s && *s)) out.push_back(cur);
    return out;
}

// Adjacent pairs along `inner` (offsets of the paired direction) at even,
// 16-byte aligned positions for every offset of `outer`; the rule
// dtcf_view_create applies.
bool pairs_ok(const std::vector<int64_t> &inner, const std::vector<int64_t> &outer)
{
    if (inner.size() % 2) return false;
    for (int64_t o : outer)
        if (o % 2) return false;
    for (size_t c = 0; c + 1 < inner.size(); c += 2)
        if (inner[c + 1] != inner[c] + 1 || inner[c] % 2) return false;
    return true;
}

}  // namespace
}  // namespace dtcf

using namespace dtcf;

// ---------------------------------------------------------------- programs
struct PArr {
    std::string name;
    std::vector<int64_t> dims, stride;
    int64_t elems = 1;
    int role = DTCF_ARR_TEMP;
    const double *bound = nullptr;
    bool resident = false;  // inputs: slices made once at finalize (mixed contractions)
    int64_t off = -1;       // temporaries: offset in the request block
    int producer = -1, last_use = -1, out_index = -1;
    std::vector<int> readers;  // statements reading the array, in program order
};

struct PStmt {
    int kind = 0;  // 0 contraction, 1 sum
    int c = -1, a = -1, b = -1;
    double alpha = 1.0;
    // contraction: GEMM m x n x k over views (operands possibly swapped)
    bool swap = false;
    int64_t m = 1, n = 1, k = 1;
    dtcf_view av = nullptr, bv = nullptr, cv = nullptr;
    int64_t ldc = 0;
    int64_t scratch_off = -1, scratch_bytes = 0;
    int64_t sa_bytes = 0, sb_bytes = 0, ea_off = 0, eb_off = 0, a_off = 0, b_off = 0;
    // mixed: GEMM operand q taken from slices made once at finalize (an input
    // array, resident slices on): rs[q] slices, re[q] row exponents
    bool res[2] = {false, false};
    char *rs[2] = {nullptr, nullptr};
    int32_t *re[2] = {nullptr, nullptr};
    // mixed, input operand q: its view as (rows, K, row offsets, K offsets);
    // operands with equal keys on one array have equal slices
    std::vector<int64_t> vkey[2];
    // slices of operand q made by an earlier statement of the request:
    // share_st[q] (its operand share_q[q]), or -1
    int share_st[2] = {-1, -1}, share_q[2] = {0, 0};
    // the contraction runs in the program's mixed precision (else FP64)
    bool mixed = false;
    // FP64: operands in GEMM position 0 (A) / 1 (B) packed into scratch first
    bool pack[2] = {false, false};
    dtcf_strided pdesc[2];
    int64_t poff[2] = {0, 0};
    // sum
    std::vector<int> terms;
    std::vector<double> coefs;
    int64_t od[4] = {1, 1, 1, 1};
    int64_t *dstrides = nullptr;  // device [nterms][4]
    std::vector<int64_t> hstrides;  // host copy
    bool plane = false;           // sum_plane_kernel: (od[2] + 1) * od[3] <= kPlaneCap
    std::vector<int> deps;
    std::vector<int> succ;
    double flops = 0;
};

struct dtcf_program_s {
    dtcf_server srv = nullptr;
    int prec = 0, nslice = 0, max_level = 0;
    // mixed precision only for contractions with m n / (m + n) >= mixed_min
    double mixed_min = 256.0;
    // slices of the inputs marked resident, made once at finalize and kept
    // (server arena)
    std::vector<void *> res_allocs;
    std::vector<PArr> arrs;
    std::map<std::string, int> byname;
    std::vector<PStmt> st;
    std::vector<int> outputs;
    int64_t block = 0;
    double flops = 0;
    bool final = false;
    int layout = DTCF_LAYOUT_AUTO;  // requested placement (DTCF_LAYOUT_*)
    bool reuse = false;   // placement reuses freed temporaries (adds write-after-read edges)
    bool broken = false;  // a statement failed to plan: the program is unusable
    int refs = 0;  // live requests
};

struct Req {
    int64_t id = 0, seq = 0;
    dtcf_program p = nullptr;
    std::vector<double *> outs;
    int64_t off = -1;
    char *base = nullptr;
    int state = 0;
    int passes = 0;
    int inflight = 0, remaining = 0;
    std::vector<int> ndeps;
    std::vector<double> t_ready, t_issue, t_done;
    double t_submit = 0, t_admit = 0, t_end = 0;
    std::string err;
    int status = DTCF_OK;
};

struct Slot {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    Req *r = nullptr;
    int stmt = -1;
};

struct dtcf_server_s {
    int dev = 0;
    char *base = nullptr;
    Arena arena;
    int64_t resident = 0, reserved = 0, peak_reserved = 0;
    std::map<void *, std::pair<int64_t, int64_t>> tensors;  // ptr -> (off, bytes)
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable wake, done_cv;
    std::thread th;
    bool stop = false;
    std::deque<Req *> waiting;
    std::vector<Req *> active;
    std::map<int64_t, Req *> reqs;
    // Ready statements keyed (request flops, request id, statement):
    // requests with less work first, then arrival order (shortest job first).
    std::set<std::tuple<double, int64_t, int>> ready;
    int64_t next_id = 1, next_seq = 1;
    std::chrono::steady_clock::time_point t0;
    std::string fatal;
    int64_t inject = -1;  // statements to issue before an injected failure (-1: none)
};

namespace {

int arr_index(dtcf_program p, const char *name)
{
    auto it = p->byname.find(name ? name : "");
    return it == p->byname.end() ? -1 : it->second;
}

// Positions of `names` in an array reference's index list.
bool index_map(const std::vector<std::string> &idx, const std::string &x, int *pos)
{
    for (size_t i = 0; i < idx.size(); i++)
        if (idx[i] == x) {
            *pos = (int)i;
            return true;
        }
    return false;
}

// Offsets (in elements) of the row-major enumeration of `names` over an
// array whose index list is `idx`.
std::vector<int64_t> enum_offsets(const PArr &A, const std::vector<std::string> &idx,
                                  const std::vector<std::string> &names)
{
    std::vector<int64_t> off(1, 0);
    for (const auto &x : names) {
        int p = -1;
        index_map(idx, x, &p);
        std::vector<int64_t> nx;
        nx.reserve(off.size() * A.dims[p]);
        for (int64_t o : off)
            for (int64_t i = 0; i < A.dims[p]; i++) nx.push_back(o + i * A.stride[p]);
        off.swap(nx);
    }
    return off;
}

struct Cand {
    bool swap;
    std::vector<std::string> korder;
    double score;
};

int plan_contraction(dtcf_program p, PStmt &s, const std::vector<std::string> &ci,
                     const std::vector<std::string> &ai, const std::vector<std::string> &bi)
{
    const PArr &C = p->arrs[s.c], &A = p->arrs[s.a], &B = p->arrs[s.b];
    std::vector<std::string> ea, eb, contr;
    for (const auto &x : ci) {
        int q;
        const bool ina = index_map(ai, x, &q), inb = index_map(bi, x, &q);
        DTCF_REQUIRE(ina != inb, "program: an external index must be in exactly one operand");
        (ina ? ea : eb).push_back(x);
    }
    for (const auto &x : ai) {
        int q;
        if (index_map(bi, x, &q)) {
            DTCF_REQUIRE(!index_map(ci, x, &q), "program: batched indices are not supported");
            contr.push_back(x);
        } else {
            DTCF_REQUIRE(index_map(ci, x, &q), "program: index of A neither external nor contracted");
        }
    }
    for (const auto &x : bi) {
        int q;
        DTCF_REQUIRE(index_map(ci, x, &q) || index_map(ai, x, &q),
                     "program: index of B neither external nor contracted");
    }
    for (size_t i = 0; i < ai.size(); i++)
        for (size_t j = 0; j < bi.size(); j++)
            if (ai[i] == bi[j]) DTCF_REQUIRE(A.dims[i] == B.dims[j], "program: contracted extents differ");
    for (size_t i = 0; i < ci.size(); i++) {
        int q;
        if (index_map(ai, ci[i], &q)) DTCF_REQUIRE(A.dims[q] == C.dims[i], "program: external extents differ");
        if (index_map(bi, ci[i], &q)) DTCF_REQUIRE(B.dims[q] == C.dims[i], "program: external extents differ");
    }
    // Candidate K orders: the contracted indices in A's order and in B's.
    std::vector<std::vector<std::string>> orders{contr};
    {
        std::vector<std::string> kb;
        for (const auto &x : bi) {
            int q;
            if (index_map(contr, x, &q)) kb.push_back(x);
        }
        if (kb != contr) orders.push_back(kb);
    }
    const auto ra = enum_offsets(A, ai, ea), rb = enum_offsets(B, bi, eb);
    const auto rc = enum_offsets(C, ci, ea), cc = enum_offsets(C, ci, eb);
    auto opscore = [](const std::vector<int64_t> &rows, const std::vector<int64_t> &ks) {
        if (pairs_ok(ks, rows)) return 2.0;
        if (pairs_ok(rows, ks)) return 1.5;
        return 0.0;
    };
    Cand best{false, contr, -1.0};
    for (const auto &ko : orders) {
        const auto ka = enum_offsets(A, ai, ko), kb = enum_offsets(B, bi, ko);
        const double sa = opscore(ra, ka), sb = opscore(rb, kb);
        for (int sw = 0; sw < 2; sw++) {
            // C as written: rows ext_a, cols ext_b; swapped: rows ext_b, cols ext_a
            const double sc = sw ? (pairs_ok(rc, cc) ? 0.5 : 0.0) : (pairs_ok(cc, rc) ? 0.5 : 0.0);
            const double sc2 = sa + sb + sc + (sw ? -0.01 : 0.0);
            if (sc2 > best.score) best = Cand{sw != 0, ko, sc2};
        }
    }
    const auto ka = enum_offsets(A, ai, best.korder), kb = enum_offsets(B, bi, best.korder);
    s.swap = best.swap;
    const auto &Ar = s.swap ? rb : ra, &Ak = s.swap ? kb : ka;
    const auto &Br = s.swap ? ra : rb, &Bk = s.swap ? ka : kb;
    const auto &Cr = s.swap ? cc : rc, &Cc = s.swap ? rc : cc;
    s.m = (int64_t)Ar.size();
    s.n = (int64_t)Br.size();
    s.k = (int64_t)Ak.size();
    // Splitting both operands costs about (m + n) k (16 + S) bytes of HBM
    // traffic; the slices save about two thirds of the FP64 GEMM's 2 m n k
    // flops.  With 66 TF/s and 5 TB/s that pays off from m n / (m + n) ~ 200:
    // thinner contractions (the T1 terms, n = o) stay FP64.
    s.mixed = p->prec != 0 &&
              (double)s.m * (double)s.n >= p->mixed_min * (double)(s.m + s.n);
    // An operand with neither direction contiguous is permuted into a dense
    // [rows][K] copy first (one pass at HBM speed) unless it is small: the
    // FP64 GEMM reads it from there, and so does the mixed-precision split
    // pass, whose kernels read rows or column pairs contiguously and would
    // otherwise gather single elements.  Resident inputs are split once, at
    // bind, from the array itself.
    const int garr[2] = {s.swap ? s.b : s.a, s.swap ? s.a : s.b};  // GEMM operand q's array
    int64_t scr = 0;
    std::vector<int64_t> dense_r[2], dense_k[2];
    for (int q = 0; q < 2; q++) {
        if (s.mixed && p->arrs[garr[q]].role == DTCF_ARR_INPUT && p->arrs[garr[q]].resident) continue;
        const auto &R = q == 0 ? Ar : Br;
        const auto &Kk = q == 0 ? Ak : Bk;
        const int64_t rows = (int64_t)R.size(), K = (int64_t)Kk.size();
        if (opscore(R, Kk) > 0.0 || rows * K < (int64_t(1) << 16)) continue;
        const bool opa = (q == 0) != s.swap;  // operand A of the statement?
        const PArr &X = opa ? A : B;
        const auto &xi = opa ? ai : bi;
        const auto &rnames = opa ? ea : eb;
        std::vector<std::string> names(rnames);
        names.insert(names.end(), best.korder.begin(), best.korder.end());
        DTCF_REQUIRE(names.size() == xi.size() && names.size() <= DTCF_MAX_DIMS, "program: pack rank");
        dtcf_strided &d = s.pdesc[q];
        d = dtcf_strided{};
        d.dims = (int32_t)names.size();
        int64_t ostride = 1;
        for (int j = (int)names.size() - 1; j >= 0; j--) {
            int pos = -1;
            index_map(xi, names[j], &pos);
            d.ext[j] = X.dims[pos];
            d.istr[j] = X.stride[pos];
            d.ostr[j] = ostride;
            ostride *= X.dims[pos];
        }
        s.pack[q] = true;
        s.poff[q] = scr;
        scr += granule(rows * K * (int64_t)sizeof(double));
        dense_r[q].resize(rows);
        dense_k[q].resize(K);
        for (int64_t r = 0; r < rows; r++) dense_r[q][r] = r * K;
        for (int64_t c = 0; c < K; c++) dense_k[q][c] = c;
    }
    const auto &Ar2 = s.pack[0] ? dense_r[0] : Ar, &Ak2 = s.pack[0] ? dense_k[0] : Ak;
    const auto &Br2 = s.pack[1] ? dense_r[1] : Br, &Bk2 = s.pack[1] ? dense_k[1] : Bk;
    int st = dtcf_view_create(s.m, s.k, Ar2.data(), Ak2.data(), nullptr, s.m, s.k, &s.av);
    if (st == DTCF_OK) st = dtcf_view_create(s.n, s.k, Br2.data(), Bk2.data(), nullptr, s.n, s.k, &s.bv);
    if (st != DTCF_OK) return st;
    bool dense = true;
    for (int64_t r = 0; dense && r < s.m; r++) dense = Cr[r] == r * s.n;
    for (int64_t c = 0; dense && c < s.n; c++) dense = Cc[c] == c;
    if (dense) {
        s.ldc = s.n;
    } else {
        st = dtcf_view_create(s.m, s.n, Cr.data(), Cc.data(), nullptr, s.m, s.n, &s.cv);
        if (st != DTCF_OK) return st;
    }
    s.scratch_bytes = scr;
    s.flops = 2.0 * (double)s.m * (double)s.n * (double)s.k;
    if (s.mixed) {
        s.sa_bytes = granule(dtcf_split_bytes(p->prec, p->nslice, s.m, s.k));
        s.sb_bytes = granule(dtcf_split_bytes(p->prec, p->nslice, s.n, s.k));
        DTCF_REQUIRE(s.sa_bytes > 0 && s.sb_bytes > 0, "program: bad mixed-precision mode");
        for (int q = 0; q < 2; q++) {
            s.res[q] = p->arrs[garr[q]].role == DTCF_ARR_INPUT && p->arrs[garr[q]].resident;
            if (p->arrs[garr[q]].role != DTCF_ARR_INPUT) continue;
            const auto &R = q == 0 ? Ar : Br;
            const auto &Kk = q == 0 ? Ak : Bk;
            auto &key = s.vkey[q];
            key.reserve(2 + R.size() + Kk.size());
            key.push_back((int64_t)R.size());
            key.push_back((int64_t)Kk.size());
            key.insert(key.end(), R.begin(), R.end());
            key.insert(key.end(), Kk.begin(), Kk.end());
        }
        int64_t off = scr;  // after the dense copies, if any
        if (!s.res[0]) { s.a_off = off; off += s.sa_bytes; }
        if (!s.res[1]) { s.b_off = off; off += s.sb_bytes; }
        if (!s.res[0]) { s.ea_off = off; off += granule((s.m + 127) / 128 * 128 * 4); }
        if (!s.res[1]) { s.eb_off = off; off += granule((s.n + 127) / 128 * 128 * 4); }
        s.scratch_bytes = off;
    }
    return DTCF_OK;
}

// ------------------------------------------------------------ execution
const double *arr_ptr(const dtcf_program p, const PArr &a, const Req *r)
{
    if (a.role == DTCF_ARR_INPUT) return a.bound;
    if (a.role == DTCF_ARR_OUTPUT) return r->outs[a.out_index];
    return reinterpret_cast<const double *>(r->base + a.off);
}

int launch_stmt(dtcf_server srv, Req *r, int i, hipStream_t stream)
{
    dtcf_program p = r->p;
    const PStmt &s = p->st[i];
    const PArr &C = p->arrs[s.c];
    double *c = const_cast<double *>(arr_ptr(p, C, r));
    if (s.kind == 1) {
        SumArgs a{};
        a.out = c;
        a.n = C.elems;
        a.d1 = s.od[1];
        a.d2 = s.od[2];
        a.d3 = s.od[3];
        a.nterms = (int32_t)s.terms.size();
        a.strides = s.dstrides;
        for (size_t t = 0; t < s.terms.size(); t++) {
            a.src[t] = arr_ptr(p, p->arrs[s.terms[t]], r);
            a.coef[t] = s.coefs[t];
        }
        if (s.plane) {
            static std::once_flag once;
            static hipError_t attr = hipSuccess;
            std::call_once(once, [] {
                for (const void *k : {reinterpret_cast<const void *>(sum_plane_kernel<false>),
                                      reinterpret_cast<const void *>(sum_plane_kernel<true>)})
                    if (attr == hipSuccess)
                        attr = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)(kChunk * kPlaneCap * sizeof(double)));
            });
            DTCF_HIP(attr);
            const uint64_t nplanes = (uint64_t)(a.n / (s.od[2] * s.od[3]));
            const unsigned grid = (unsigned)std::min<uint64_t>(nplanes, 1u << 20);
            const uint32_t bufstride = (uint32_t)(s.od[3] * (s.od[2] + 1));
            // element pairs when the plane's rows are whole pairs; per term, a
            // 16-byte load where its pairs are adjacent and aligned
            const bool pairs = s.od[2] % 2 == 0 && s.od[3] % 2 == 0 &&
                               (reinterpret_cast<uintptr_t>(c) & 15) == 0;
            if (pairs)
                for (size_t t = 0; t < s.terms.size(); t++) {
                    const int64_t *q = s.hstrides.data() + 4 * t;
                    if (q[3] == 1 && q[0] % 2 == 0 && q[1] % 2 == 0 && q[2] % 2 == 0 &&
                        (reinterpret_cast<uintptr_t>(a.src[t]) & 15) == 0)
                        a.pairs[t >> 6] |= uint64_t(1) << (t & 63);
                }
            if (pairs)
                hipLaunchKernelGGL(sum_plane_kernel<true>, dim3(grid), dim3(kPlaneThreads),
                                   (size_t)kChunk * bufstride * sizeof(double), stream, a,
                                   FastDiv((uint32_t)s.od[3]), FastDiv((uint32_t)s.od[2]), nplanes,
                                   bufstride);
            else
                hipLaunchKernelGGL(sum_plane_kernel<false>, dim3(grid), dim3(kPlaneThreads),
                                   (size_t)kChunk * bufstride * sizeof(double), stream, a,
                                   FastDiv((uint32_t)s.od[3]), FastDiv((uint32_t)s.od[2]), nplanes,
                                   bufstride);
        } else {
            hipLaunchKernelGGL(sum_kernel, dim3(grid_for(a.n)), dim3(256), 0, stream, a);
        }
        DTCF_HIP(hipGetLastError());
        return DTCF_OK;
    }
    const double *A = arr_ptr(p, p->arrs[s.a], r), *B = arr_ptr(p, p->arrs[s.b], r);
    if (s.swap) std::swap(A, B);
    if (!s.mixed) {
        // FP64: the GEMM writes C = A * B directly (no clearing pass)
        const double *op[2] = {A, B};
        for (int q = 0; q < 2; q++) {
            if (!s.pack[q]) continue;
            double *dst = reinterpret_cast<double *>(r->base + s.scratch_off + s.poff[q]);
            int st = dtcf_permute_copy_f64(op[q], dst, &s.pdesc[q], 1, nullptr, 0, nullptr, 0, 1.0,
                                           0.0, stream);
            if (st != DTCF_OK) return st;
            op[q] = dst;
        }
        return gemm_view_overwrite(s.m, s.n, s.k, s.alpha, op[0], s.av, op[1], s.bv, c, s.cv, s.ldc,
                                   stream);
    }
    // mixed precision: the fused slice products write C (overwrite); the
    // per-product path accumulates into a cleared C
    const int ns_eff = p->prec == DTCF_PREC_I8 ? p->nslice : (p->prec == DTCF_PREC_F32 ? 1 : 3);
    const int lev_eff = p->max_level > 0 ? p->max_level : ns_eff + 1;
    const int overwrite = gemm_split_fused(p->prec, ns_eff, lev_eff) ? 1 : 0;
    if (!overwrite) {
        hipLaunchKernelGGL(zero_kernel, dim3(grid_for(C.elems)), dim3(256), 0, stream, c, C.elems);
        DTCF_HIP(hipGetLastError());
    }
    char *sl[2];
    int32_t *ex[2];
    bool split[2];
    for (int q = 0; q < 2; q++) {
        if (s.res[q]) {
            sl[q] = s.rs[q];
            ex[q] = s.re[q];
            split[q] = false;
            continue;
        }
        // own scratch, or the scratch of the statement whose slices it shares
        const PStmt &o = s.share_st[q] >= 0 ? p->st[s.share_st[q]] : s;
        const int oq = s.share_st[q] >= 0 ? s.share_q[q] : q;
        char *scr = r->base + o.scratch_off;
        sl[q] = scr + (oq == 0 ? o.a_off : o.b_off);
        ex[q] = reinterpret_cast<int32_t *>(scr + (oq == 0 ? o.ea_off : o.eb_off));
        split[q] = s.share_st[q] < 0;
    }
    char *sa = sl[0], *sb = sl[1];
    int32_t *ea = ex[0], *eb = ex[1];
    const double *op[2] = {A, B};
    int st = DTCF_OK;
    for (int q = 0; q < 2 && st == DTCF_OK; q++) {
        if (!split[q] || !s.pack[q]) continue;
        double *dst = reinterpret_cast<double *>(r->base + s.scratch_off + s.poff[q]);
        st = dtcf_permute_copy_f64(op[q], dst, &s.pdesc[q], 1, nullptr, 0, nullptr, 0, 1.0, 0.0, stream);
        op[q] = dst;
    }
    if (st == DTCF_OK && split[0]) st = dtcf_split_f64(p->prec, p->nslice, op[0], s.av, sa, ea, stream);
    if (st == DTCF_OK && split[1]) st = dtcf_split_f64(p->prec, p->nslice, op[1], s.bv, sb, eb, stream);
    if (st == DTCF_OK)
        st = gemm_split_f64_impl(p->prec, p->nslice, s.m, s.n, s.k, s.alpha, sa, ea, sb, eb,
                                 p->max_level, c, s.cv, s.ldc, stream, overwrite);
    return st;
}

void finish_request(dtcf_server srv, Req *r)
{
    r->t_end = now_us(srv->t0);
    r->state = r->status == DTCF_OK ? 2 : 3;
    if (r->off >= 0) {
        srv->arena.release(r->off, r->p->block);
        srv->reserved -= granule(r->p->block);
        r->off = -1;
    }
    r->p->refs--;
    srv->ready.erase(srv->ready.lower_bound({r->p->flops, r->id, -1}),
                     srv->ready.lower_bound({r->p->flops, r->id + 1, -1}));
    srv->active.erase(std::find(srv->active.begin(), srv->active.end(), r));
    srv->done_cv.notify_all();
}

// One pass of the scheduler: completions, admissions, issues.  Returns true
// if anything changed.  Called with the server lock held.
bool sched_pass(dtcf_server srv)
{
    bool changed = false;
    // completions
    for (auto &sl : srv->slots) {
        if (!sl.r) continue;
        hipError_t q = hipEventQuery(sl.ev);
        if (q == hipErrorNotReady) continue;
        Req *r = sl.r;
        const int i = sl.stmt;
        sl.r = nullptr;
        sl.stmt = -1;
        changed = true;
        const double t = now_us(srv->t0);
        r->t_done[i] = t;
        r->inflight--;
        r->remaining--;
        if (q != hipSuccess && r->status == DTCF_OK) {
            r->status = DTCF_ERR_HIP;
            r->err = hipGetErrorString(q);
        }
        if (r->status == DTCF_OK) {
            for (int j : r->p->st[i].succ)
                if (--r->ndeps[j] == 0) {
                    r->t_ready[j] = t;
                    srv->ready.insert({r->p->flops, r->id, j});
                }
        }
        if (r->remaining == 0 || (r->status != DTCF_OK && r->inflight == 0)) finish_request(srv, r);
    }
    // admissions: arrival order; a blocked head may be passed 8 times
    for (auto it = srv->waiting.begin(); it != srv->waiting.end();) {
        Req *r = *it;
        const int64_t off = srv->arena.alloc(r->p->block);
        if (off < 0) {
            if (it == srv->waiting.begin() && r->passes >= 8) break;
            ++it;
            continue;
        }
        if (it != srv->waiting.begin()) srv->waiting.front()->passes++;
        r->off = off;
        r->base = srv->base + off;
        r->state = 1;
        r->t_admit = now_us(srv->t0);
        srv->reserved += granule(r->p->block);
        srv->peak_reserved = std::max(srv->peak_reserved, srv->reserved);
        srv->active.push_back(r);
        const auto &st = r->p->st;
        for (size_t j = 0; j < st.size(); j++) {
            r->ndeps[j] = (int)st[j].deps.size();
            if (r->ndeps[j] == 0) {
                r->t_ready[j] = r->t_admit;
                srv->ready.insert({r->p->flops, r->id, (int)j});
            }
        }
        if (st.empty()) finish_request(srv, r);
        it = srv->waiting.erase(it);
        changed = true;
    }
    // issues
    for (auto &sl : srv->slots) {
        if (sl.r || srv->ready.empty()) continue;
        auto it = srv->ready.begin();
        Req *r = srv->reqs[std::get<1>(*it)];
        const int i = std::get<2>(*it);
        srv->ready.erase(it);
        changed = true;
        if (r->status != DTCF_OK) {  // a failed request issues nothing more
            r->remaining--;
            if (r->inflight == 0) finish_request(srv, r);
            continue;
        }
        r->t_issue[i] = now_us(srv->t0);
        int st;
        if (srv->inject == 0) {  // host-forced failure: nothing is launched
            srv->inject = -1;
            set_error("injected statement failure");
            st = DTCF_ERR_HIP;
        } else {
            if (srv->inject > 0) srv->inject--;
            st = launch_stmt(srv, r, i, sl.s);
        }
        hipError_t e = st == DTCF_OK ? hipEventRecord(sl.ev, sl.s) : hipSuccess;
        if (st != DTCF_OK || e != hipSuccess) {
            r->status = st != DTCF_OK ? st : DTCF_ERR_HIP;
            r->err = st != DTCF_OK ? std::string(get_error()) : std::string(hipGetErrorString(e));
            r->remaining--;
            r->t_done[i] = now_us(srv->t0);
            if (r->inflight == 0) finish_request(srv, r);
            continue;
        }
        r->inflight++;
        sl.r = r;
        sl.stmt = i;
    }
    return changed;
}

void sched_loop(dtcf_server srv)
{
    (void)hipSetDevice(srv->dev);
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        const bool changed = sched_pass(srv);
        bool busy = false;
        for (auto &sl : srv->slots) busy |= sl.r != nullptr;
        if (srv->stop && srv->active.empty() && !srv->waiting.empty() && !changed) {
            // closing: nothing runs, and what still waits cannot be admitted
            for (Req *r : srv->waiting) {
                r->status = DTCF_ERR_BUDGET;
                r->err = "server closed before the request could be admitted";
                r->state = 3;
                r->t_end = now_us(srv->t0);
                r->p->refs--;
            }
            srv->waiting.clear();
            srv->done_cv.notify_all();
        }
        if (srv->stop && !busy && srv->waiting.empty() && srv->active.empty()) break;
        if (changed) continue;
        if (busy) {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            srv->wake.wait_for(lk, std::chrono::milliseconds(2));
        }
    }
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" int dtcf_server_open(int64_t budget, int nslots, dtcf_server *out)
{
    DTCF_REQUIRE(out, "server: null output");
    DTCF_REQUIRE(budget >= kGranule, "server: budget too small");
    DTCF_REQUIRE(nslots >= 1 && nslots <= 16, "server: 1..16 slots");
    *out = nullptr;
    auto *s = new dtcf_server_s();
    hipError_t e = hipGetDevice(&s->dev);
    if (e == hipSuccess) e = hipMalloc(&s->base, (size_t)granule(budget));
    if (e != hipSuccess) {
        delete s;
        DTCF_HIP(e);
    }
    s->arena.init(granule(budget));
    s->slots.resize(nslots);
    for (auto &sl : s->slots) {
        e = hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
        if (e != hipSuccess) break;
        int st = gemm_reserve_workspace(sl.s);
        if (st != DTCF_OK) {
            e = hipErrorOutOfMemory;
            break;
        }
    }
    if (e != hipSuccess) {
        for (auto &sl : s->slots) {
            if (sl.ev) (void)hipEventDestroy(sl.ev);
            if (sl.s) (void)hipStreamDestroy(sl.s);
        }
        (void)hipFree(s->base);
        delete s;
        DTCF_HIP(e);
    }
    s->t0 = std::chrono::steady_clock::now();
    s->th = std::thread(sched_loop, s);
    *out = s;
    return DTCF_OK;
}

extern "C" int dtcf_server_close(dtcf_server s)
{
    if (!s) return DTCF_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->wake.notify_all();
    s->th.join();
    for (auto &sl : s->slots) {
        (void)hipStreamSynchronize(sl.s);
        (void)hipEventDestroy(sl.ev);
        (void)hipStreamDestroy(sl.s);
    }
    for (auto &kv : s->reqs) delete kv.second;
    (void)hipFree(s->base);
    delete s;
    return DTCF_OK;
}

extern "C" int dtcf_server_alloc(dtcf_server s, int64_t bytes, void **ptr)
{
    DTCF_REQUIRE(s && ptr && bytes > 0, "server_alloc: bad argument");
    int64_t off;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        off = s->arena.alloc(bytes);
        if (off < 0) {
            set_error("server_alloc: arena budget exhausted");
            return DTCF_ERR_BUDGET;
        }
        s->resident += granule(bytes);
        s->tensors[s->base + off] = {off, bytes};
    }
    *ptr = s->base + off;
    // null stream only: the slots are non-blocking streams and keep running
    DTCF_HIP(hipMemsetAsync(*ptr, 0, (size_t)bytes, nullptr));
    DTCF_HIP(hipStreamSynchronize(nullptr));
    return DTCF_OK;
}

extern "C" int dtcf_server_free(dtcf_server s, void *ptr)
{
    DTCF_REQUIRE(s, "server_free: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tensors.find(ptr);
    DTCF_REQUIRE(it != s->tensors.end(), "server_free: not a resident tensor of this server");
    s->arena.release(it->second.first, it->second.second);
    s->resident -= granule(it->second.second);
    s->tensors.erase(it);
    return DTCF_OK;
}

extern "C" int dtcf_server_inject_failure(dtcf_server s, int64_t after)
{
    DTCF_REQUIRE(s && after >= -1, "server_inject_failure: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->inject = after;
    return DTCF_OK;
}

extern "C" int dtcf_server_usage(dtcf_server s, int64_t *budget, int64_t *resident,
                                 int64_t *reserved, int64_t *peak)
{
    DTCF_REQUIRE(s, "server_usage: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (budget) *budget = s->arena.size;
    if (resident) *resident = s->resident;
    if (reserved) *reserved = s->reserved;
    if (peak) *peak = s->peak_reserved;
    return DTCF_OK;
}

extern "C" int dtcf_program_create(dtcf_server s, int prec, int nslice, int max_level,
                                   dtcf_program *out)
{
    DTCF_REQUIRE(s && out, "program: bad argument");
    DTCF_REQUIRE(prec == 0 || (prec >= DTCF_PREC_F32 && prec <= DTCF_PREC_I8), "program: unknown precision");
    if (prec == DTCF_PREC_I8) DTCF_REQUIRE(nslice >= 1 && nslice <= 6, "program: 1..6 INT8 slices");
    auto *p = new dtcf_program_s();
    p->srv = s;
    p->prec = prec;
    p->nslice = prec == DTCF_PREC_I8 ? nslice : (prec == DTCF_PREC_F32 ? 1 : 3);
    p->max_level = max_level;
    *out = p;
    return DTCF_OK;
}

extern "C" int dtcf_program_array(dtcf_program p, const char *name, int ndim, const int64_t *dims,
                                  int role)
{
    DTCF_REQUIRE(p && !p->final, "program: finalized or null");
    DTCF_REQUIRE(name && *name && dims, "program: bad array");
    DTCF_REQUIRE(ndim >= 1 && ndim <= 4, "program: arrays of 1..4 dims");
    DTCF_REQUIRE(role >= DTCF_ARR_INPUT && role <= DTCF_ARR_OUTPUT, "program: bad role");
    DTCF_REQUIRE(arr_index(p, name) < 0, "program: array declared twice");
    PArr a;
    a.name = name;
    a.role = role;
    a.dims.assign(dims, dims + ndim);
    a.stride.assign(ndim, 1);
    for (int d = ndim - 2; d >= 0; d--) a.stride[d] = a.stride[d + 1] * a.dims[d + 1];
    for (int d = 0; d < ndim; d++) {
        DTCF_REQUIRE(dims[d] >= 1, "program: extents must be positive");
        a.elems *= dims[d];
    }
    if (role == DTCF_ARR_OUTPUT) {
        a.out_index = (int)p->outputs.size();
        p->outputs.push_back((int)p->arrs.size());
    }
    p->byname[a.name] = (int)p->arrs.size();
    p->arrs.push_back(a);
    return DTCF_OK;
}

extern "C" int dtcf_program_bind(dtcf_program p, const char *name, const double *ptr)
{
    DTCF_REQUIRE(p, "program: null");
    const int i = arr_index(p, name);
    DTCF_REQUIRE(i >= 0, "program_bind: unknown array");
    DTCF_REQUIRE(p->arrs[i].role == DTCF_ARR_INPUT, "program_bind: only inputs are bound");
    DTCF_REQUIRE(ptr, "program_bind: null pointer");
    if (p->final) {
        // Requests read bound inputs when their statements issue; resident
        // slices were cut from the pointer bound at finalize.
        DTCF_REQUIRE(!p->arrs[i].resident,
                     "program_bind: an input with resident slices cannot be rebound after finalize");
        std::lock_guard<std::mutex> lk(p->srv->mu);
        DTCF_REQUIRE(p->refs == 0, "program_bind: requests of this program are in flight");
        p->arrs[i].bound = ptr;
        return DTCF_OK;
    }
    p->arrs[i].bound = ptr;
    return DTCF_OK;
}

extern "C" int dtcf_program_set_resident(dtcf_program p, const char *name, int on)
{
    DTCF_REQUIRE(p && !p->final && p->st.empty(), "program: set_resident before any statement");
    const int i = arr_index(p, name);
    DTCF_REQUIRE(i >= 0, "program_set_resident: unknown array");
    DTCF_REQUIRE(p->arrs[i].role == DTCF_ARR_INPUT, "program_set_resident: only inputs are resident");
    p->arrs[i].resident = on != 0;
    return DTCF_OK;
}

static int check_ref(dtcf_program p, int ai, const std::vector<std::string> &idx)
{
    DTCF_REQUIRE(ai >= 0, "program: unknown array");
    DTCF_REQUIRE(idx.size() == p->arrs[ai].dims.size(), "program: index count differs from array rank");
    for (size_t i = 0; i < idx.size(); i++)
        for (size_t j = i + 1; j < idx.size(); j++)
            DTCF_REQUIRE(idx[i] != idx[j], "program: repeated index in one reference");
    return DTCF_OK;
}

static int add_output(dtcf_program p, int c)
{
    DTCF_REQUIRE(p->arrs[c].role != DTCF_ARR_INPUT, "program: an input cannot be assigned");
    DTCF_REQUIRE(p->arrs[c].producer < 0, "program: array assigned twice");
    p->arrs[c].producer = (int)p->st.size();
    return DTCF_OK;
}

static int add_use(dtcf_program p, int a)
{
    const PArr &x = p->arrs[a];
    DTCF_REQUIRE(x.role == DTCF_ARR_INPUT || x.producer >= 0, "program: array read before assignment");
    const int i = (int)p->st.size();
    p->arrs[a].last_use = i;
    if (p->arrs[a].readers.empty() || p->arrs[a].readers.back() != i) p->arrs[a].readers.push_back(i);
    return DTCF_OK;
}

extern "C" int dtcf_program_contract(dtcf_program p, const char *c, const char *c_idx,
                                     const char *a, const char *a_idx, const char *b,
                                     const char *b_idx, double alpha)
{
    DTCF_REQUIRE(p && !p->final, "program: finalized or null");
    PStmt s;
    s.kind = 0;
    s.c = arr_index(p, c);
    s.a = arr_index(p, a);
    s.b = arr_index(p, b);
    s.alpha = alpha;
    const auto ci = split_names(c_idx), ai = split_names(a_idx), bi = split_names(b_idx);
    int st = check_ref(p, s.c, ci);
    if (!st) st = check_ref(p, s.a, ai);
    if (!st) st = check_ref(p, s.b, bi);
    DTCF_REQUIRE(s.c != s.a && s.c != s.b, "program: result aliases an operand");
    if (!st) st = add_use(p, s.a);
    if (!st) st = add_use(p, s.b);
    if (!st) st = plan_contraction(p, s, ci, ai, bi);
    if (!st) st = add_output(p, s.c);
    if (st) {
        p->broken = true;
        if (s.av) dtcf_view_destroy(s.av);
        if (s.bv) dtcf_view_destroy(s.bv);
        if (s.cv) dtcf_view_destroy(s.cv);
        return st;
    }
    p->st.push_back(std::move(s));
    return DTCF_OK;
}

extern "C" int dtcf_program_sum(dtcf_program p, const char *c, const char *c_idx, int nterms,
                                const char *const *names, const char *const *idx,
                                const double *coefs)
{
    DTCF_REQUIRE(p && !p->final, "program: finalized or null");
    DTCF_REQUIRE(nterms >= 1 && nterms <= kMaxSumTerms, "program: 1..96 sum terms");
    DTCF_REQUIRE(names && idx && coefs, "program: null term list");
    PStmt s;
    s.kind = 1;
    s.c = arr_index(p, c);
    const auto ci = split_names(c_idx);
    p->broken = true;  // cleared when the statement is added
    int st = check_ref(p, s.c, ci);
    if (st) return st;
    const PArr &C = p->arrs[s.c];
    const int nd = (int)ci.size(), pad = 4 - nd;
    for (int d = 0; d < nd; d++) s.od[pad + d] = C.dims[d];
    std::vector<int64_t> strides;
    for (int t = 0; t < nterms; t++) {
        const int ti = arr_index(p, names[t]);
        const auto ti_idx = split_names(idx[t]);
        st = check_ref(p, ti, ti_idx);
        if (st) return st;
        DTCF_REQUIRE(ti != s.c, "program: sum result aliases a term");
        const PArr &X = p->arrs[ti];
        int64_t q[4] = {0, 0, 0, 0};
        for (int d = 0; d < nd; d++) {
            int pos;
            DTCF_REQUIRE(index_map(ti_idx, ci[d], &pos), "program: sum term is not a permutation of the result");
            DTCF_REQUIRE(X.dims[pos] == C.dims[d], "program: sum term extents differ");
            q[pad + d] = X.stride[pos];
        }
        strides.insert(strides.end(), q, q + 4);
        st = add_use(p, ti);
        if (st) return st;
        s.terms.push_back(ti);
        s.coefs.push_back(coefs[t]);
    }
    st = add_output(p, s.c);
    if (st) return st;
    DTCF_HIP(hipMalloc(&s.dstrides, strides.size() * sizeof(int64_t)));
    DTCF_HIP(hipMemcpy(s.dstrides, strides.data(), strides.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    s.hstrides = strides;
    s.flops = (double)C.elems * nterms;
    s.plane = (s.od[2] + 1) * s.od[3] <= kPlaneCap && s.od[2] * s.od[3] <= kPlaneThreads * kPlanePer;
    p->st.push_back(std::move(s));
    p->broken = false;
    return DTCF_OK;
}

extern "C" int dtcf_program_finalize(dtcf_program p)
{
    DTCF_REQUIRE(p && !p->final, "program: finalized or null");
    DTCF_REQUIRE(!p->broken, "program_finalize: a statement failed to plan; destroy the program");
    for (const auto &a : p->arrs) {
        if (a.role == DTCF_ARR_INPUT) DTCF_REQUIRE(a.bound, "program_finalize: an input is not bound");
        if (a.role == DTCF_ARR_OUTPUT) DTCF_REQUIRE(a.producer >= 0, "program_finalize: an output is never assigned");
    }
    const int ns = (int)p->st.size();
    // data dependencies
    for (int i = 0; i < ns; i++) {
        PStmt &s = p->st[i];
        std::vector<int> ops = s.kind == 0 ? std::vector<int>{s.a, s.b} : s.terms;
        for (int a : ops)
            if (p->arrs[a].role != DTCF_ARR_INPUT) s.deps.push_back(p->arrs[a].producer);
    }
    // Static placement in program order, in one of two layouts.
    //  * reuse: first fit, a temporary's bytes free once its last reader is
    //    placed.  A freed region remembers every statement that touched its
    //    last occupant (all readers of a temporary, or its producer if
    //    nothing reads it; the owning statement for slice scratch):
    //    statements complete out of program order, so a placement that
    //    reuses the region waits for all of them, not only the last.
    //  * separate: every temporary and scratch region its own bytes, so no
    //    write-after-read edges -- only the data dependencies order the
    //    statements.
    // The separate layout is taken when its block is within 1.5x (or
    // 256 MiB) of the reuse layout's: small programs then run as wide as
    // their data dependencies allow, large ones keep the smaller block.
    // Input operands split per request: a later statement reading the same
    // array through the same view takes the first one's slices (after it)
    // instead of splitting again; the first one's scratch then stays placed
    // until its last sharer.
    std::vector<std::vector<int>> sharers(ns);
    {
        std::map<std::pair<int, std::vector<int64_t>>, std::pair<int, int>> first;
        for (int i = 0; i < ns; i++) {
            PStmt &s = p->st[i];
            if (s.kind != 0 || !s.mixed) continue;
            const int arr[2] = {s.swap ? s.b : s.a, s.swap ? s.a : s.b};
            std::pair<int, int> mine[2] = {{-1, 0}, {-1, 0}};
            for (int q = 0; q < 2; q++) {
                if (s.res[q] || s.vkey[q].empty()) continue;
                auto it = first.find({arr[q], s.vkey[q]});
                if (it != first.end()) {
                    s.share_st[q] = it->second.first;
                    s.share_q[q] = it->second.second;
                    s.deps.push_back(it->second.first);
                    sharers[it->second.first].push_back(i);
                } else {
                    mine[q] = {i, q};
                }
            }
            for (int q = 0; q < 2; q++)
                if (mine[q].first >= 0) first.emplace(std::make_pair(arr[q], s.vkey[q]), mine[q]);
        }
    }
    std::vector<std::vector<int>> dies(ns);
    for (int a = 0; a < (int)p->arrs.size(); a++) {
        const PArr &x = p->arrs[a];
        if (x.role == DTCF_ARR_TEMP && x.producer >= 0)
            dies[x.last_use >= 0 ? x.last_use : x.producer].push_back(a);
    }
    struct Dead {
        int64_t off, len;
        std::vector<int> waits;
    };
    // Places everything (offsets into the arrays and statements); with
    // reuse, collects the write-after-read edges into war.  Returns the
    // block size.
    auto layout = [&](bool reuse, std::vector<std::vector<int>> &war) -> int64_t {
        Arena plan;
        plan.init(int64_t(1) << 60);
        std::vector<Dead> dead;
        auto place = [&](int i, int64_t bytes) {
            const int64_t off = plan.alloc(bytes), len = granule(bytes);
            for (const auto &d : dead)
                if (d.off < off + len && off < d.off + d.len)
                    for (int w : d.waits)
                        if (w != i) war[i].push_back(w);
            return off;
        };
        std::vector<std::vector<int>> held(ns);  // scratch released after statement i
        for (int i = 0; i < ns; i++) {
            PStmt &s = p->st[i];
            PArr &C = p->arrs[s.c];
            if (C.role == DTCF_ARR_TEMP) C.off = place(i, C.elems * (int64_t)sizeof(double));
            if (s.scratch_bytes > 0) {
                s.scratch_off = place(i, s.scratch_bytes);
                if (reuse) {
                    if (sharers[i].empty()) {
                        plan.release(s.scratch_off, s.scratch_bytes);
                        dead.push_back({s.scratch_off, granule(s.scratch_bytes), {i}});
                    } else {
                        held[sharers[i].back()].push_back(i);
                    }
                }
            }
            if (!reuse) continue;
            for (int j : held[i]) {
                const PStmt &o = p->st[j];
                plan.release(o.scratch_off, o.scratch_bytes);
                std::vector<int> waits{j};
                waits.insert(waits.end(), sharers[j].begin(), sharers[j].end());
                dead.push_back({o.scratch_off, granule(o.scratch_bytes), waits});
            }
            for (int a : dies[i]) {
                PArr &x = p->arrs[a];
                plan.release(x.off, x.elems * (int64_t)sizeof(double));
                dead.push_back({x.off, granule(x.elems * (int64_t)sizeof(double)),
                                x.readers.empty() ? std::vector<int>{x.producer} : x.readers});
            }
        }
        return std::max<int64_t>(plan.high, kGranule);
    };
    std::vector<std::vector<int>> war(ns), none(ns);
    const int64_t sep_block = layout(false, none);
    const int64_t reuse_block = layout(true, war);
    const bool separate = p->layout == DTCF_LAYOUT_SEPARATE ||
                          (p->layout == DTCF_LAYOUT_AUTO &&
                           sep_block <= std::max(reuse_block + (int64_t(256) << 20),
                                                 reuse_block + reuse_block / 2));
    if (separate) {
        p->block = layout(false, none);
        p->reuse = false;
    } else {
        p->block = reuse_block;
        p->reuse = true;
        for (int i = 0; i < ns; i++)
            p->st[i].deps.insert(p->st[i].deps.end(), war[i].begin(), war[i].end());
    }
    p->flops = 0;
    for (int i = 0; i < ns; i++) {
        PStmt &s = p->st[i];
        std::sort(s.deps.begin(), s.deps.end());
        s.deps.erase(std::unique(s.deps.begin(), s.deps.end()), s.deps.end());
        for (int d : s.deps) {
            DTCF_REQUIRE(d < i, "program_finalize: dependency on a later statement");
            p->st[d].succ.push_back(i);
        }
        p->flops += s.flops;
    }
    // Resident slices: each resident input, once per distinct view, split
    // from its bound contents as they are now into the server's arena.
    std::map<std::pair<int, std::vector<int64_t>>, std::pair<char *, int32_t *>> made;
    for (auto &s : p->st) {
        if (s.kind != 0 || !s.mixed) continue;
        const int arr[2] = {s.swap ? s.b : s.a, s.swap ? s.a : s.b};
        for (int q = 0; q < 2; q++) {
            if (!s.res[q]) continue;
            auto hit = made.find({arr[q], s.vkey[q]});
            if (hit != made.end()) {
                s.rs[q] = hit->second.first;
                s.re[q] = hit->second.second;
                continue;
            }
            const int64_t rows = q == 0 ? s.m : s.n;
            const int64_t sbytes = q == 0 ? s.sa_bytes : s.sb_bytes;
            void *ptr = nullptr;
            int st = dtcf_server_alloc(p->srv, sbytes + granule((rows + 127) / 128 * 128 * 4), &ptr);
            if (st != DTCF_OK) {
                p->broken = true;  // deps are already linked: no second finalize
                return st;
            }
            p->res_allocs.push_back(ptr);
            s.rs[q] = static_cast<char *>(ptr);
            s.re[q] = reinterpret_cast<int32_t *>(s.rs[q] + sbytes);
            st = dtcf_split_f64(p->prec, p->nslice, p->arrs[arr[q]].bound, q == 0 ? s.av : s.bv, s.rs[q],
                                s.re[q], nullptr);
            if (st != DTCF_OK) {
                p->broken = true;
                return st;
            }
            made.emplace(std::make_pair(arr[q], s.vkey[q]), std::make_pair(s.rs[q], s.re[q]));
        }
    }
    if (!p->res_allocs.empty()) DTCF_HIP(hipStreamSynchronize(nullptr));
    p->final = true;
    return DTCF_OK;
}

extern "C" int dtcf_program_set_mixed_min(dtcf_program p, double h)
{
    DTCF_REQUIRE(p && !p->final && p->st.empty(), "program: set_mixed_min before any statement");
    DTCF_REQUIRE(h >= 0.0, "program: negative mixed_min");
    p->mixed_min = h;
    return DTCF_OK;
}

extern "C" int dtcf_program_set_layout(dtcf_program p, int layout)
{
    DTCF_REQUIRE(p && !p->final, "program: finalized or null");
    DTCF_REQUIRE(layout == DTCF_LAYOUT_AUTO || layout == DTCF_LAYOUT_REUSE ||
                     layout == DTCF_LAYOUT_SEPARATE,
                 "program: unknown layout");
    p->layout = layout;
    return DTCF_OK;
}

extern "C" int dtcf_program_get_layout(dtcf_program p, int *layout)
{
    DTCF_REQUIRE(p && p->final && layout, "program: not finalized or null");
    *layout = p->reuse ? DTCF_LAYOUT_REUSE : DTCF_LAYOUT_SEPARATE;
    return DTCF_OK;
}

extern "C" int dtcf_program_info(dtcf_program p, int64_t *block, int64_t *nst, double *flops,
                                 int64_t *nout)
{
    DTCF_REQUIRE(p, "program: null");
    if (block) *block = p->block;
    if (nst) *nst = (int64_t)p->st.size();
    if (flops) *flops = p->flops;
    if (nout) *nout = (int64_t)p->outputs.size();
    return DTCF_OK;
}

extern "C" int dtcf_program_statement_plan(dtcf_program p, int64_t i, int64_t *out)
{
    DTCF_REQUIRE(p && out, "program: null argument");
    DTCF_REQUIRE(p->final, "program: statement plans exist after finalize");
    DTCF_REQUIRE(i >= 0 && i < (int64_t)p->st.size(), "program: statement index out of range");
    const PStmt &s = p->st[i];
    for (int q = 0; q < DTCF_PLAN_NOUT; q++) out[q] = 0;
    out[DTCF_PLAN_KIND] = s.kind;
    if (s.kind == 1) {
        out[DTCF_PLAN_NTERMS] = (int64_t)s.terms.size();
        out[DTCF_PLAN_FLAGS] = s.plane ? DTCF_PLAN_SUM_PLANE : 0;
        return DTCF_OK;
    }
    out[DTCF_PLAN_M] = s.m;
    out[DTCF_PLAN_N] = s.n;
    out[DTCF_PLAN_K] = s.k;
    int64_t f = 0;
    if (s.swap) f |= DTCF_PLAN_SWAP;
    if (s.mixed) f |= DTCF_PLAN_MIXED;
    if (s.pack[0]) f |= DTCF_PLAN_PACK_A;
    if (s.pack[1]) f |= DTCF_PLAN_PACK_B;
    if (s.av->colpairs) f |= DTCF_PLAN_A_COLPAIRS;
    if (s.av->rowpairs) f |= DTCF_PLAN_A_ROWPAIRS;
    if (s.bv->colpairs) f |= DTCF_PLAN_B_COLPAIRS;
    if (s.bv->rowpairs) f |= DTCF_PLAN_B_ROWPAIRS;
    if (!s.cv) f |= DTCF_PLAN_C_DENSE;
    if (s.res[0]) f |= DTCF_PLAN_RES_A;
    if (s.res[1]) f |= DTCF_PLAN_RES_B;
    out[DTCF_PLAN_FLAGS] = f;
    out[DTCF_PLAN_TILE] = s.mixed ? -1 : pick_config(s.m, s.n, s.k);
    return DTCF_OK;
}

extern "C" int dtcf_program_destroy(dtcf_program p)
{
    if (!p) return DTCF_OK;
    if (p->srv) {
        std::unique_lock<std::mutex> lk(p->srv->mu);
        p->srv->done_cv.wait(lk, [&] { return p->refs == 0; });
    }
    for (auto &s : p->st) {
        if (s.av) dtcf_view_destroy(s.av);
        if (s.bv) dtcf_view_destroy(s.bv);
        if (s.cv) dtcf_view_destroy(s.cv);
        if (s.dstrides) (void)hipFree(s.dstrides);
    }
    for (void *ptr : p->res_allocs) (void)dtcf_server_free(p->srv, ptr);
    delete p;
    return DTCF_OK;
}

extern "C" int dtcf_request_submit(dtcf_server s, dtcf_program p, double *const *outputs,
                                   int64_t *id)
{
    DTCF_REQUIRE(s && p && id, "submit: bad argument");
    DTCF_REQUIRE(p->final && p->srv == s, "submit: program not finalized on this server");
    DTCF_REQUIRE(p->outputs.empty() || outputs, "submit: output buffers missing");
    for (size_t i = 0; i < p->outputs.size(); i++) DTCF_REQUIRE(outputs[i], "submit: null output buffer");
    auto *r = new Req();
    r->p = p;
    r->outs.assign(outputs, outputs + p->outputs.size());
    const size_t n = p->st.size();
    r->ndeps.assign(n, 0);
    r->t_ready.assign(n, -1.0);
    r->t_issue.assign(n, -1.0);
    r->t_done.assign(n, -1.0);
    r->remaining = (int)n;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (granule(p->block) > s->arena.size - s->resident) {
            delete r;
            set_error("submit: request block exceeds the budget left after resident tensors");
            return DTCF_ERR_BUDGET;
        }
        r->id = s->next_id++;
        r->seq = s->next_seq++;
        r->t_submit = now_us(s->t0);
        p->refs++;
        s->reqs[r->id] = r;
        s->waiting.push_back(r);
        *id = r->id;
    }
    s->wake.notify_all();
    return DTCF_OK;
}

extern "C" int dtcf_request_wait(dtcf_server s, int64_t id, double timeout_ms)
{
    DTCF_REQUIRE(s, "wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    DTCF_REQUIRE(it != s->reqs.end(), "wait: unknown request");
    Req *r = it->second;
    auto done = [&] { return r->state >= 2; };
    if (timeout_ms < 0) {
        s->done_cv.wait(lk, done);
    } else if (!s->done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done)) {
        return DTCF_ERR_NOTREADY;
    }
    if (r->state == 3) {
        set_error("request failed: " + r->err);
        return r->status;
    }
    return DTCF_OK;
}

extern "C" int dtcf_request_info_get(dtcf_server s, int64_t id, dtcf_request_info *info)
{
    DTCF_REQUIRE(s && info, "info: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    DTCF_REQUIRE(it != s->reqs.end(), "info: unknown request");
    const Req *r = it->second;
    info->state = r->state;
    info->ntasks = (int64_t)r->t_done.size();
    info->t_submit = r->t_submit;
    info->t_admit = r->t_admit;
    info->t_done = r->t_end;
    info->flops = r->p->flops;
    return DTCF_OK;
}

extern "C" int dtcf_request_task_times(dtcf_server s, int64_t id, double *t_ready,
                                       double *t_issue, double *t_done, int64_t n)
{
    DTCF_REQUIRE(s, "task_times: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    DTCF_REQUIRE(it != s->reqs.end(), "task_times: unknown request");
    const Req *r = it->second;
    DTCF_REQUIRE(n == (int64_t)r->t_done.size(), "task_times: n differs from the statement count");
    for (int64_t i = 0; i < n; i++) {
        if (t_ready) t_ready[i] = r->t_ready[i];
        if (t_issue) t_issue[i] = r->t_issue[i];
        if (t_done) t_done[i] = r->t_done[i];
    }
    return DTCF_OK;
}

extern "C" int dtcf_request_release(dtcf_server s, int64_t id)
{
    DTCF_REQUIRE(s, "release: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    DTCF_REQUIRE(it != s->reqs.end(), "release: unknown request");
    DTCF_REQUIRE(it->second->state >= 2, "release: request still running");
    delete it->second;
    s->reqs.erase(it);
    return DTCF_OK;
}
