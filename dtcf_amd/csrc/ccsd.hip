// UHF CCSD amplitude update, correlation energy and residual norms (new
// code: the reference stops at the residuals ra_vo / rb_vo, _uhf_t1.eq:143,
// 242, and rab_vvoo / raa_vvoo / rbb_vvoo, _uhf_t2.eq:519,769,996).
//
// One step, after the T1 and T2 residual programs of an iteration:
//   t_ai   += r_ai   / (e_i - e_a)
//   t_abij += r_abij / (((e_i + e_j) - e_a) - e_b)
// per spin block with its own orbital energies, then the energy of the
// updated amplitudes
//   E1  = sum fa_ov[i,a] ta[a,i] + sum fb_ov[i,a] tb[a,i]
//   Eaa = 1/4 sum vaa_oovv[i,j,a,b] (taa[a,b,i,j] + 2 ta[a,i] ta[b,j])
//   Eab =     sum vab_oovv[i,j,a,b] (tab[a,b,i,j] +   ta[a,i] tb[b,j])
//   Ebb = 1/4 sum vbb_oovv[i,j,a,b] (tbb[a,b,i,j] + 2 tb[a,i] tb[b,j])
// (the spin-orbital 1/4 <ij||ab> t_abij + 1/2 <ij||ab> t_ai t_bj resolved
// into spin blocks) and the norms of R1, R2 and of the update.
//
// HBM-bound: a T2 element costs 32 bytes (residual and amplitude read, the
// amplitude written, the integral read).  A block is viewed as P = V1*V2
// rows (a,b) by Q = O1*O2 columns (i,j); the integral [i,j,a,b] is its
// transpose.  A workgroup takes 64 x 64 tiles: the integral tile goes
// through LDS (read along p), residual and amplitude are read along q, both
// with 16-byte loads where the rows allow.  Sums are deterministic: fixed
// tiles per workgroup (from the sizes only), a fixed tree in the workgroup,
// and a final pass adding the workgroup partials in index order.
#include <algorithm>
#include <map>
#include <mutex>

#include "common.hpp"

namespace dtcf {
namespace {

constexpr int CT = 256;        // threads per workgroup
constexpr int CPARTS = 2048;   // workgroups of a T2 pass = partials per row
constexpr int TP = 64, TQ = 64;

// partial rows: [block][E, sum r^2, sum dt^2] for aa, ab, bb, then the T1
// values (E1a, r2a, dt2a, E1b, r2b, dt2b)
constexpr int NROWS = 9;
constexpr size_t PART_DOUBLES = (size_t)NROWS * CPARTS + 8;

__device__ inline double wg_sum(double v, double *sh)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < CT / 64; w++) s += sh[w];
    return s;
}

struct T1Args {
    const double *r[2];
    double *t[2];
    const double *f_ov[2];
    const double *eo[2], *ev[2];
    uint32_t O[2], V[2];
    double *out;  // 6 values
};

// Both spins in one workgroup each (blockIdx.x = spin): singles are small
// (V x O elements).  The T2 pass reads the updated singles, so it runs after.
__global__ __launch_bounds__(CT) void cc_t1_kernel(T1Args a)
{
    __shared__ double sh[CT / 64];
    const int s = blockIdx.x;
    const uint32_t O = a.O[s], V = a.V[s];
    const double *__restrict__ r = a.r[s];
    double *__restrict__ t = a.t[s];
    const double *__restrict__ f = a.f_ov[s];
    double e = 0.0, r2 = 0.0, d2 = 0.0;
    for (uint32_t x = threadIdx.x; x < V * O; x += CT) {
        const uint32_t av = x / O, i = x - av * O;
        double tv = t[x];
        if (r) {
            const double rv = r[x];
            const double dt = rv / (a.eo[s][i] - a.ev[s][av]);
            tv = tv + dt;
            t[x] = tv;
            r2 = fma(rv, rv, r2);
            d2 = fma(dt, dt, d2);
        }
        e = fma(f[(size_t)i * V + av], tv, e);
    }
    const double es = wg_sum(e, sh), rs = wg_sum(r2, sh), ds = wg_sum(d2, sh);
    if (threadIdx.x == 0) {
        a.out[3 * s + 0] = es;
        a.out[3 * s + 1] = rs;
        a.out[3 * s + 2] = ds;
    }
}

struct T2Args {
    const double *r;    // [P][Q] residual, or null (energy only)
    double *t;          // [P][Q] amplitudes (updated in place with r)
    const double *v;    // [Q][P] integral v_oovv
    const double *eo1, *eo2, *ev1, *ev2;  // energies of i, j, a, b
    const double *t1a, *t1b;              // updated singles [V1][O1], [V2][O2]
    uint32_t O1, O2, V1, V2;
    FastDiv div_v2, div_o2;
    uint64_t P, Q;
    uint32_t tiles_q;
    uint64_t ntiles;
    double g;           // weight of the singles product (2 same spin, 1 mixed)
    double *part;       // 3 rows of CPARTS
};

template <bool VEC>
__global__ __launch_bounds__(CT) void cc_t2_kernel(T2Args a)
{
    __shared__ double lds[TQ][TP + 1];
    __shared__ double sh[CT / 64];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 column pairs x 8 rows
    double e = 0.0, r2 = 0.0, d2 = 0.0;
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += CPARTS) {
        const uint64_t tp = tile / a.tiles_q, tq = tile - tp * a.tiles_q;
        const uint64_t p0 = tp * TP, q0 = tq * TQ;
        __syncthreads();
        // integral tile v[q0 .. +64][p0 .. +64] -> lds[q][p], read along p
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int ql = ty + 8 * k, pl = 2 * tx;
            const uint64_t q = q0 + ql, p = p0 + pl;
            double x0 = 0.0, x1 = 0.0;
            if (q < a.Q) {
                const double *src = a.v + q * a.P + p;
                if (VEC && p + 1 < a.P) {
                    const double2 w = *reinterpret_cast<const double2 *>(src);
                    x0 = w.x;
                    x1 = w.y;
                } else {
                    if (p < a.P) x0 = src[0];
                    if (p + 1 < a.P) x1 = src[1];
                }
            }
            lds[ql][pl] = x0;
            lds[ql][pl + 1] = x1;
        }
        __syncthreads();
        // this thread's two columns q, q+1 -> (i, j)
        const int ql = 2 * tx;
        const uint64_t qa = q0 + ql;
        uint32_t ii[2], jj[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const uint32_t q = (uint32_t)min<uint64_t>(qa + c, a.Q - 1);
            ii[c] = a.div_o2.div(q);
            jj[c] = q - ii[c] * a.O2;
        }
        const double eij[2] = {a.eo1[ii[0]] + a.eo2[jj[0]], a.eo1[ii[1]] + a.eo2[jj[1]]};
#pragma unroll 2
        for (int k = 0; k < 8; k++) {
            const int pl = ty + 8 * k;
            const uint64_t p = p0 + pl;
            if (p >= a.P || qa >= a.Q) continue;
            const uint32_t av = a.div_v2.div((uint32_t)p), bv = (uint32_t)p - av * a.V2;
            const uint64_t base = p * a.Q + qa;
            const bool two = qa + 1 < a.Q;
            double tv[2], rv[2] = {0.0, 0.0};
            if (VEC && two) {
                const double2 w = *reinterpret_cast<const double2 *>(a.t + base);
                tv[0] = w.x;
                tv[1] = w.y;
                if (a.r) {
                    const double2 z = *reinterpret_cast<const double2 *>(a.r + base);
                    rv[0] = z.x;
                    rv[1] = z.y;
                }
            } else {
                tv[0] = a.t[base];
                tv[1] = two ? a.t[base + 1] : 0.0;
                if (a.r) {
                    rv[0] = a.r[base];
                    rv[1] = two ? a.r[base + 1] : 0.0;
                }
            }
            const double eva = a.ev1[av], evb = a.ev2[bv];
#pragma unroll
            for (int c = 0; c < 2; c++) {
                if (c == 1 && !two) break;
                if (a.r) {
                    const double dt = rv[c] / ((eij[c] - eva) - evb);
                    tv[c] = tv[c] + dt;
                    r2 = fma(rv[c], rv[c], r2);
                    d2 = fma(dt, dt, d2);
                }
                const double w = tv[c] + a.g * (a.t1a[(size_t)av * a.O1 + ii[c]] *
                                                a.t1b[(size_t)bv * a.O2 + jj[c]]);
                e = fma(lds[ql + c][pl], w, e);
            }
            if (a.r) {
                if (VEC && two) {
                    *reinterpret_cast<double2 *>(a.t + base) = make_double2(tv[0], tv[1]);
                } else {
                    a.t[base] = tv[0];
                    if (two) a.t[base + 1] = tv[1];
                }
            }
        }
    }
    const double es = wg_sum(e, sh), rs = wg_sum(r2, sh), ds = wg_sum(d2, sh);
    if (threadIdx.x == 0) {
        a.part[blockIdx.x] = es;
        a.part[CPARTS + blockIdx.x] = rs;
        a.part[2 * CPARTS + blockIdx.x] = ds;
    }
}

// Adds every row of partials in index order and assembles the outputs.
__global__ __launch_bounds__(CT) void cc_finish_kernel(const double *__restrict__ part,
                                                       double *__restrict__ out)
{
    __shared__ double sh[CT / 64];
    __shared__ double row[NROWS];
    for (int q = 0; q < NROWS; q++) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < CPARTS; i += CT) acc += part[(size_t)q * CPARTS + i];
        const double s = wg_sum(acc, sh);
        if (threadIdx.x == 0) row[q] = s;
    }
    if (threadIdx.x == 0) {
        const double *t1 = part + (size_t)NROWS * CPARTS;
        const double e1 = t1[0] + t1[3];
        const double eaa = 0.25 * row[0], eab = row[3], ebb = 0.25 * row[6];
        out[0] = ((e1 + eaa) + eab) + ebb;
        out[1] = e1;
        out[2] = eaa;
        out[3] = eab;
        out[4] = ebb;
        out[5] = sqrt(t1[1] + t1[4]);
        out[6] = sqrt((row[1] + row[4]) + row[7]);
        out[7] = sqrt((((t1[2] + t1[5]) + row[2]) + row[5]) + row[8]);
    }
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

int validate(const dtcf_uhf_ccsd *d, int update)
{
    DTCF_REQUIRE(d, "uhf_ccsd: null descriptor");
    DTCF_REQUIRE(d->Oa > 0 && d->Ob > 0 && d->Va > 0 && d->Vb > 0,
                 "uhf_ccsd: every range must be positive");
    DTCF_REQUIRE(d->Oa < 65536 && d->Ob < 65536 && d->Va < 65536 && d->Vb < 65536,
                 "uhf_ccsd: ranges must be < 65536");
    DTCF_REQUIRE(d->eoa && d->eob && d->eva && d->evb, "uhf_ccsd: null orbital energies");
    DTCF_REQUIRE(d->fa_ov && d->fb_ov && d->vaa_oovv && d->vab_oovv && d->vbb_oovv,
                 "uhf_ccsd: null Fock or integral block");
    DTCF_REQUIRE(d->ta && d->tb && d->taa && d->tab && d->tbb, "uhf_ccsd: null amplitudes");
    if (update)
        DTCF_REQUIRE(d->ra && d->rb && d->raa && d->rab && d->rbb, "uhf_ccsd: null residuals");
    return DTCF_OK;
}

// Launches the step with its partial sums in `part` (PART_DOUBLES doubles).
int launch_step(const dtcf_uhf_ccsd *d, int update, double *part, double *out,
                hipStream_t s)
{
    T1Args t1{};
    const int64_t O[2] = {d->Oa, d->Ob}, V[2] = {d->Va, d->Vb};
    const double *r1[2] = {d->ra, d->rb}, *f[2] = {d->fa_ov, d->fb_ov};
    const double *eo[2] = {d->eoa, d->eob}, *ev[2] = {d->eva, d->evb};
    double *t1p[2] = {d->ta, d->tb};
    for (int q = 0; q < 2; q++) {
        t1.r[q] = update ? r1[q] : nullptr;
        t1.t[q] = t1p[q];
        t1.f_ov[q] = f[q];
        t1.eo[q] = eo[q];
        t1.ev[q] = ev[q];
        t1.O[q] = (uint32_t)O[q];
        t1.V[q] = (uint32_t)V[q];
    }
    t1.out = part + (size_t)NROWS * CPARTS;
    hipLaunchKernelGGL(cc_t1_kernel, dim3(2), dim3(CT), 0, s, t1);
    DTCF_HIP(hipGetLastError());
    struct Blk {
        const double *r;
        double *t;
        const double *v;
        int s1, s2;  // spin of (i, a) and of (j, b): 0 alpha, 1 beta
        double g;
    } blks[3] = {{d->raa, d->taa, d->vaa_oovv, 0, 0, 2.0},
                 {d->rab, d->tab, d->vab_oovv, 0, 1, 1.0},
                 {d->rbb, d->tbb, d->vbb_oovv, 1, 1, 2.0}};
    for (int b = 0; b < 3; b++) {
        const Blk &k = blks[b];
        T2Args a{};
        a.r = update ? k.r : nullptr;
        a.t = k.t;
        a.v = k.v;
        a.eo1 = eo[k.s1];
        a.eo2 = eo[k.s2];
        a.ev1 = ev[k.s1];
        a.ev2 = ev[k.s2];
        a.t1a = t1p[k.s1];
        a.t1b = t1p[k.s2];
        a.O1 = (uint32_t)O[k.s1];
        a.O2 = (uint32_t)O[k.s2];
        a.V1 = (uint32_t)V[k.s1];
        a.V2 = (uint32_t)V[k.s2];
        a.div_v2 = FastDiv(a.V2);
        a.div_o2 = FastDiv(a.O2);
        a.P = (uint64_t)a.V1 * a.V2;
        a.Q = (uint64_t)a.O1 * a.O2;
        a.tiles_q = (uint32_t)((a.Q + TQ - 1) / TQ);
        a.ntiles = ((a.P + TP - 1) / TP) * a.tiles_q;
        a.g = k.g;
        a.part = part + (size_t)3 * b * CPARTS;
        const bool vec = a.Q % 2 == 0 && a.P % 2 == 0 && aligned16(a.t) && aligned16(a.v) &&
                         (!a.r || aligned16(a.r));
        if (vec)
            hipLaunchKernelGGL(cc_t2_kernel<true>, dim3(CPARTS), dim3(CT), 0, s, a);
        else
            hipLaunchKernelGGL(cc_t2_kernel<false>, dim3(CPARTS), dim3(CT), 0, s, a);
        DTCF_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(cc_finish_kernel, dim3(1), dim3(CT), 0, s, part, out);
    DTCF_HIP(hipGetLastError());
    return DTCF_OK;
}

// Partial-sum buffer per (device, stream): steps on different streams never
// share one; steps on one stream are ordered by it.
std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, double *> g_part;

int stream_partials(hipStream_t s, double **p)
{
    int dev = 0;
    DTCF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_pair(dev, s);
    auto it = g_part.find(key);
    if (it == g_part.end()) {
        double *buf = nullptr;
        DTCF_HIP(hipMalloc(&buf, PART_DOUBLES * sizeof(double)));
        it = g_part.emplace(key, buf).first;
    }
    *p = it->second;
    return DTCF_OK;
}

// Partial sums of the allocation-free entry point: one array per device for
// the life of the process (PART_DOUBLES doubles = 147 KiB, whatever the
// system size).  Host code never passes `scratch` itself to a kernel: its
// device address comes from hipGetSymbolAddress, once per device.
__device__ __align__(256) double scratch[PART_DOUBLES];

constexpr int kMaxDev = 64;
std::mutex g_scr_mu;
double *g_scr[kMaxDev] = {};

int scratch_for_device(int dev, double **p)
{
    DTCF_REQUIRE(dev >= 0 && dev < kMaxDev, "uhf_ccsd: device index out of range");
    std::lock_guard<std::mutex> lk(g_scr_mu);
    if (!g_scr[dev]) {
        void *ptr = nullptr;
        DTCF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        DTCF_REQUIRE(ptr, "uhf_ccsd: no device address for the scratch array");
        g_scr[dev] = static_cast<double *>(ptr);
    }
    *p = g_scr[dev];
    return DTCF_OK;
}

}  // namespace

// Called by dtcf_init with `dev` current: resolves that device's scratch
// address up front, so the allocation-free entry point only looks it up.
int ccsd_scratch_setup(int dev)
{
    double *p = nullptr;
    return scratch_for_device(dev, &p);
}

}  // namespace dtcf

using namespace dtcf;

extern "C" int dtcf_uhf_ccsd_step_f64_noalloc(const dtcf_uhf_ccsd *d, int update, double *out_dev,
                                              void *stream)
{
    int st = validate(d, update);
    if (st != DTCF_OK) return st;
    DTCF_REQUIRE(out_dev, "uhf_ccsd: null output");
    int dev = 0;
    DTCF_HIP(hipGetDevice(&dev));
    double *part = nullptr;
    st = scratch_for_device(dev, &part);  // set up by dtcf_init, else on the first call here
    if (st != DTCF_OK) return st;
    // every partial the finish pass reads is written by this call's kernels,
    // so nothing in the shared scratch has to be cleared first
    return launch_step(d, update, part, out_dev, as_stream(stream));
}

extern "C" int dtcf_uhf_ccsd_step_f64(const dtcf_uhf_ccsd *d, int update, double *out_dev,
                                      void *stream)
{
    int st = validate(d, update);
    if (st != DTCF_OK) return st;
    DTCF_REQUIRE(out_dev, "uhf_ccsd: null output");
    double *part = nullptr;
    st = stream_partials(as_stream(stream), &part);
    if (st != DTCF_OK) return st;
    return launch_step(d, update, part, out_dev, as_stream(stream));
}
