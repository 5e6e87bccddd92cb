// Host-side AddressSanitizer / UndefinedBehaviorSanitizer run of the library:
// this program and libdtcf's host code are built with -fsanitize=address,
// undefined on the host side only (the GPU code is the ordinary build), and
// it drives the host planners, tables, arenas and the serving scheduler
// through the C ABI on small problems, checking the results as it goes:
// views and view GEMMs, the host-pointer GEMM and tile permutes (pinned
// staging), an RRR local step in its three modes, DLTC terms, tasks and
// sums, INT8 digit splits and their products, serving programs (both
// layouts, several requests, an injected failure, a refused budget, a mixed
// program), the UHF CCSD step and its allocation-free form, and error paths.
// Prints "host sanitize: ok" when every check passed; a sanitizer report
// ends the process with a non-zero status.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dtcf_local.h"

static int g_fail = 0;
#define CHECK(cond, ...)                                                        \
    do {                                                                        \
        if (!(cond)) {                                                          \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                    \
            std::printf(__VA_ARGS__);                                           \
            std::printf(" (%s)\n", dtcf_last_error());                          \
            g_fail++;                                                           \
        }                                                                       \
    } while (0)
#define OK(call) CHECK((call) == DTCF_OK, "%s", #call)

static void fill(std::vector<double> &v, unsigned seed)
{
    uint64_t x = 0x9E3779B97F4A7C15ull * (seed + 1);
    for (auto &e : v) {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        e = (double)((x * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}

static double *dev_copy(const std::vector<double> &h)
{
    double *d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(double)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

static std::vector<double> host_copy(const double *d, size_t n)
{
    std::vector<double> h(n);
    if (hipMemcpy(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) h.assign(n, NAN);
    return h;
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

static void views_and_gemms()
{
    const int64_t m = 37, n = 29, k = 53;
    std::vector<double> A(m * k), B(k * n), C(m * n, 0.0), ref(m * n, 0.0);
    fill(A, 1);
    fill(B, 2);
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t q = 0; q < k; q++) ref[i * n + j] += A[i * k + q] * B[q * n + j];
    double *dA = dev_copy(A), *dB = dev_copy(B), *dC = dev_copy(C);
    // A row-major (m x k); B stored k x n, read as an n x k view
    std::vector<int64_t> ar(m), ac(k), br(n), bc(k);
    for (int64_t i = 0; i < m; i++) ar[i] = i * k;
    for (int64_t q = 0; q < k; q++) ac[q] = q;
    for (int64_t j = 0; j < n; j++) br[j] = j;
    for (int64_t q = 0; q < k; q++) bc[q] = q * n;
    dtcf_view av = nullptr, bv = nullptr, sub = nullptr;
    OK(dtcf_view_create(m, k, ar.data(), ac.data(), nullptr, m, k, &av));
    OK(dtcf_view_create(n, k, br.data(), bc.data(), nullptr, n, k, &bv));
    OK(dtcf_gemm_view_f64(m, n, k, 1.0, dA, av, dB, bv, dC, nullptr, n, nullptr));
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    CHECK(maxrel(host_copy(dC, m * n), ref) < 1e-12, "view gemm");
    OK(dtcf_view_sub(av, 3, 10, 5, 20, &sub));
    CHECK(sub && dtcf_view_span(sub) > 0, "sub view");
    dtcf_view_destroy(sub);
    // gather a view into a dense matrix and back
    std::vector<double> g(m * k, 0.0);
    double *dG = dev_copy(g);
    OK(dtcf_view_gather_f64(dA, av, dG, k, nullptr));
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    CHECK(maxrel(host_copy(dG, m * k), A) == 0.0, "gather");
    // host-pointer GEMM (pinned staging): C = A * B with B transposed flag off
    std::vector<double> Ch(m * n, 0.0);
    OK(dtcf_gemm_f64_host(0, 0, m, n, k, 1.0, A.data(), k, B.data(), n, Ch.data(), n));
    CHECK(maxrel(Ch, ref) < 1e-12, "host gemm");
    // host tile permute: 3 tiles of 2 x 3 x 4, out dim order (2, 0, 1)
    const int64_t w[3] = {2, 3, 4};
    const int32_t out_src[3] = {2, 0, 1};
    std::vector<double> T(3 * 24), P(3 * 24), back(3 * 24);
    fill(T, 3);
    OK(dtcf_permute_tiles_f64_host(T.data(), P.data(), 3, w, out_src, 3));
    for (int t = 0; t < 3; t++)
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 4; c++)
                    CHECK(P[t * 24 + (c * 2 + a) * 3 + b] == T[t * 24 + (a * 3 + b) * 4 + c], "permute");
    // error paths: a negative offset, mismatched views
    std::vector<int64_t> bad(m, -1);
    dtcf_view v2 = nullptr;
    CHECK(dtcf_view_create(m, k, bad.data(), ac.data(), nullptr, m, k, &v2) != DTCF_OK && !v2, "bad view");
    CHECK(dtcf_gemm_view_f64(m, n, k + 1, 1.0, dA, av, dB, bv, dC, nullptr, n, nullptr) != DTCF_OK,
          "mismatched view");
    dtcf_view_destroy(av);
    dtcf_view_destroy(bv);
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
    hipFree(dG);
}

static void rrr_step(bool sym)
{
    // ring C[i,j,a,b] += A[i,k,a,c] B[k,j,c,b], 2 x 2 x 2 x 2 tiles of 2 x 2 x 3 x 3;
    // with `sym`, A and B are symmetric in their first two dims, so their
    // tiles with addr0 < addr1 are bounced copies (two classes per operand)
    const int dims = 4, nt = 16, tsz = 2 * 2 * 3 * 3;
    std::vector<int32_t> addr(nt * dims);
    for (int t = 0; t < nt; t++)
        for (int d = 0; d < dims; d++) addr[t * dims + d] = (t >> (dims - 1 - d)) & 1;
    std::vector<double> A(nt * tsz), B(nt * tsz), C0(nt * tsz);
    fill(A, 4);
    fill(B, 5);
    fill(C0, 6);
    double *dA = dev_copy(A), *dB = dev_copy(B);
    std::vector<std::vector<double>> out;
    for (int mode : {DTCF_STEP_FUSED, DTCF_STEP_PACKED, DTCF_STEP_LITERAL}) {
        double *dC = dev_copy(C0);
        dtcf_tileset ta{}, tb{}, tc{};
        for (dtcf_tileset *t : {&ta, &tb, &tc}) {
            t->dims = dims;
            for (int d = 0; d < dims; d++) {
                t->vgrid[d] = 2;
                t->block_range[d] = d < 2 ? 2 : 3;
                t->sym_group[d] = (sym && t != &tc && d < 2) ? 0 : 3;
            }
            t->ntiles = nt;
            t->tile_addr = addr.data();
        }
        ta.tiles = dA;
        tb.tiles = dB;
        tc.tiles = dC;
        dtcf_step st = nullptr;
        OK(dtcf_step_plan(&ta, "i,k,a,c", &tb, "k,j,c,b", &tc, "i,j,a,b", mode, &st));
        if (st) {
            OK(dtcf_step_run(st, 0.5, nullptr));
            OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
            int64_t t3[3], b3[3];
            OK(dtcf_step_info(st, t3, b3));
            dtcf_step_destroy(st);
        }
        out.push_back(host_copy(dC, C0.size()));
        hipFree(dC);
    }
    CHECK(maxrel(out[1], out[0]) < 1e-13 && maxrel(out[2], out[0]) < 1e-13, "step modes agree");
    CHECK(maxrel(out[0], C0) > 1e-3, "step changed C");
    hipFree(dA);
    hipFree(dB);
}

static void dltc_terms()
{
    // R[p3,p4,h1,h2] += 0.5 T[p5,p6,h1,h2] V[p3,p4,p5,p6], O = 4, V = 6, tiles 2 / 3
    enum { h1 = 0, h2 = 1, p3 = 14, p4 = 15, p5 = 16, p6 = 17 };
    const int O = 4, V = 6;
    dtcf_dltc_space sp = {12, O, V, 2, 3};
    std::vector<double> T(V * V * O * O), Vv(V * V * V * V), R(T.size(), 0.0), ref(T.size(), 0.0);
    fill(T, 7);
    fill(Vv, 8);
    for (int a = 0; a < V; a++)
        for (int b = 0; b < V; b++)
            for (int i = 0; i < O; i++)
                for (int j = 0; j < O; j++)
                    for (int c = 0; c < V; c++)
                        for (int d = 0; d < V; d++)
                            ref[((a * V + b) * O + i) * O + j] +=
                                0.5 * T[((c * V + d) * O + i) * O + j] * Vv[((a * V + b) * V + c) * V + d];
    double *dT = dev_copy(T), *dV = dev_copy(Vv), *dR = dev_copy(R), *dR2 = dev_copy(R);
    dtcf_dltc_tensor tR{4, {p3, p4, h1, h2}, dR}, tT{4, {p5, p6, h1, h2}, dT}, tV{4, {p3, p4, p5, p6}, dV};
    dtcf_dltc_mul h = nullptr;
    OK(dtcf_dltc_mul_create(&sp, &tR, &tT, &tV, 0.5, &h));
    OK(dtcf_dltc_mul_run(h, nullptr));
    // the same term as (output tile, summed tile) tasks into R2
    dtcf_dltc_tensor tR2{4, {p3, p4, h1, h2}, dR2};
    dtcf_dltc_mul h2m = nullptr;
    OK(dtcf_dltc_mul_create(&sp, &tR2, &tT, &tV, 0.5, &h2m));
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    for (int c = 0; c < 2; c++)
                        for (int d = 0; d < 2; d++) {
                            int32_t lb[DTCF_DLTC_NIDX] = {0}, ub[DTCF_DLTC_NIDX] = {0};
                            const int id[6] = {p3, p4, h1, h2, p5, p6}, v[6] = {a, b, i, j, c, d};
                            for (int q = 0; q < 6; q++) {
                                lb[id[q]] = v[q];
                                ub[id[q]] = v[q] + 1;
                            }
                            OK(dtcf_dltc_mul_task(h2m, lb, ub, nullptr));
                        }
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    const double e_term = maxrel(host_copy(dR, R.size()), ref), e_tasks = maxrel(host_copy(dR2, R.size()), ref);
    CHECK(e_term < 1e-12, "dltc term %.3e", e_term);
    CHECK(e_tasks < 1e-12, "dltc tasks %.3e", e_tasks);
    // summation R2[p3,p4,h1,h2] += 2 * T[p4,p3,h2,h1] over the whole box
    dtcf_dltc_tensor tTs{4, {p4, p3, h2, h1}, dT};
    int32_t lb[DTCF_DLTC_NIDX] = {0}, ub[DTCF_DLTC_NIDX] = {0};
    ub[p3] = ub[p4] = 2;
    ub[h1] = ub[h2] = 2;
    OK(dtcf_dltc_sum(&sp, &tR2, &tTs, 2.0, lb, ub, nullptr));
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    std::vector<double> got = host_copy(dR2, R.size()), want = ref;
    for (int a = 0; a < V; a++)
        for (int b = 0; b < V; b++)
            for (int i = 0; i < O; i++)
                for (int j = 0; j < O; j++) want[((a * V + b) * O + i) * O + j] += 2.0 * T[((b * V + a) * O + j) * O + i];
    CHECK(maxrel(got, want) < 1e-12, "dltc sum %.3e", maxrel(got, want));
    dtcf_dltc_mul_destroy(h);
    dtcf_dltc_mul_destroy(h2m);
    hipFree(dT);
    hipFree(dV);
    hipFree(dR);
    hipFree(dR2);
}

static void mixed_products()
{
    const int64_t m = 70, n = 90, k = 300;
    std::vector<double> A(m * k), B(n * k), ref(m * n, 0.0);
    fill(A, 9);
    fill(B, 10);
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t q = 0; q < k; q++) ref[i * n + j] += A[i * k + q] * B[j * k + q];
    double *dA = dev_copy(A), *dB = dev_copy(B);
    std::vector<double> z(m * n, 0.0);
    double *dC = dev_copy(z);
    std::vector<int64_t> ar(m), br(n), kc(k);
    for (int64_t i = 0; i < m; i++) ar[i] = i * k;
    for (int64_t j = 0; j < n; j++) br[j] = j * k;
    for (int64_t q = 0; q < k; q++) kc[q] = q;
    dtcf_view av = nullptr, bv = nullptr;
    OK(dtcf_view_create(m, k, ar.data(), kc.data(), nullptr, m, k, &av));
    OK(dtcf_view_create(n, k, br.data(), kc.data(), nullptr, n, k, &bv));
    const int ns = 5;
    const int64_t sa = dtcf_split_bytes(DTCF_PREC_I8, ns, m, k), sb = dtcf_split_bytes(DTCF_PREC_I8, ns, n, k);
    void *pa = nullptr, *pb = nullptr;
    int32_t *ea = nullptr, *eb = nullptr;
    hipMalloc(&pa, sa);
    hipMalloc(&pb, sb);
    hipMalloc(&ea, 128 * 4);
    hipMalloc(&eb, 128 * 4);
    OK(dtcf_split_f64(DTCF_PREC_I8, ns, dA, av, pa, ea, nullptr));
    OK(dtcf_split_f64(DTCF_PREC_I8, ns, dB, bv, pb, eb, nullptr));
    OK(dtcf_gemm_split_f64(DTCF_PREC_I8, ns, m, n, k, 1.0, pa, ea, pb, eb, 0, dC, nullptr, n, nullptr));
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    CHECK(maxrel(host_copy(dC, m * n), ref) < 1e-8, "int8 digit products");
    CHECK(dtcf_split_f64(DTCF_PREC_I8, 7, dA, av, pa, ea, nullptr) != DTCF_OK, "seven digits refused");
    dtcf_view_destroy(av);
    dtcf_view_destroy(bv);
    hipFree(pa);
    hipFree(pb);
    hipFree(ea);
    hipFree(eb);
    hipFree(dA);
    hipFree(dB);
    hipFree(dC);
}

static void serving()
{
    const int64_t ni = 24, nk = 16;
    std::vector<double> A(ni * nk), B(nk * ni);
    fill(A, 11);
    fill(B, 12);
    std::vector<double> X(ni * ni, 0.0), Y(ni * ni);
    for (int64_t i = 0; i < ni; i++)
        for (int64_t j = 0; j < ni; j++)
            for (int64_t q = 0; q < nk; q++) X[i * ni + j] += A[i * nk + q] * B[q * ni + j];
    for (int64_t i = 0; i < ni; i++)
        for (int64_t j = 0; j < ni; j++) Y[i * ni + j] = X[i * ni + j] - 0.5 * X[j * ni + i];
    dtcf_server srv = nullptr;
    OK(dtcf_server_open(int64_t(256) << 20, 4, &srv));
    if (!srv) return;
    double *dA = nullptr, *dB = nullptr;
    OK(dtcf_server_alloc(srv, ni * nk * 8, (void **)&dA));
    OK(dtcf_server_alloc(srv, nk * ni * 8, (void **)&dB));
    hipMemcpy(dA, A.data(), A.size() * 8, hipMemcpyHostToDevice);
    hipMemcpy(dB, B.data(), B.size() * 8, hipMemcpyHostToDevice);
    const char *names[2] = {"X", "X"}, *idx[2] = {"i,j", "j,i"};
    const double coefs[2] = {1.0, -0.5};
    auto make = [&](int prec, int layout) {
        dtcf_program p = nullptr;
        OK(dtcf_program_create(srv, prec, prec ? 5 : 0, 0, &p));
        if (!p) return p;
        if (prec) OK(dtcf_program_set_mixed_min(p, 0.0));
        const int64_t da[2] = {ni, nk}, db[2] = {nk, ni}, dx[2] = {ni, ni};
        OK(dtcf_program_array(p, "A", 2, da, DTCF_ARR_INPUT));
        OK(dtcf_program_array(p, "B", 2, db, DTCF_ARR_INPUT));
        OK(dtcf_program_array(p, "X", 2, dx, DTCF_ARR_TEMP));
        OK(dtcf_program_array(p, "Y", 2, dx, DTCF_ARR_OUTPUT));
        OK(dtcf_program_bind(p, "A", dA));
        OK(dtcf_program_bind(p, "B", dB));
        CHECK(dtcf_program_bind(p, "nope", dA) != DTCF_OK, "unknown array");
        OK(dtcf_program_contract(p, "X", "i,j", "A", "i,k", "B", "k,j", 1.0));
        OK(dtcf_program_sum(p, "Y", "i,j", 2, names, idx, coefs));
        OK(dtcf_program_set_layout(p, layout));
        OK(dtcf_program_finalize(p));
        int64_t plan[DTCF_PLAN_NOUT];
        OK(dtcf_program_statement_plan(p, 0, plan));
        CHECK(plan[DTCF_PLAN_KIND] == 0 && plan[DTCF_PLAN_K] == nk, "plan of the contraction");
        return p;
    };
    for (int layout : {DTCF_LAYOUT_REUSE, DTCF_LAYOUT_SEPARATE}) {
        dtcf_program p = make(0, layout);
        if (!p) continue;
        std::vector<double *> outs(3);
        std::vector<int64_t> ids(3);
        for (int r = 0; r < 3; r++) {
            hipMalloc(&outs[r], ni * ni * 8);
            OK(dtcf_request_submit(srv, p, &outs[r], &ids[r]));
        }
        for (int r = 0; r < 3; r++) {
            OK(dtcf_request_wait(srv, ids[r], 60000.0));
            dtcf_request_info info{};
            OK(dtcf_request_info_get(srv, ids[r], &info));
            std::vector<double> tr(info.ntasks), ti(info.ntasks), td(info.ntasks);
            OK(dtcf_request_task_times(srv, ids[r], tr.data(), ti.data(), td.data(), info.ntasks));
            OK(dtcf_request_release(srv, ids[r]));
            const double e = maxrel(host_copy(outs[r], ni * ni), Y);
            CHECK(e < 1e-12, "served result, layout %d: %.3e", layout, e);
            hipFree(outs[r]);
        }
        // injected failure: the next statement fails, the request reports it
        double *o = nullptr;
        hipMalloc(&o, ni * ni * 8);
        OK(dtcf_server_inject_failure(srv, 0));
        int64_t id = -1;
        OK(dtcf_request_submit(srv, p, &o, &id));
        CHECK(dtcf_request_wait(srv, id, 60000.0) == DTCF_ERR_HIP, "injected failure reported");
        OK(dtcf_request_release(srv, id));
        OK(dtcf_server_inject_failure(srv, -1));
        hipFree(o);
        OK(dtcf_program_destroy(p));
    }
    // with resident tensors filling the arena, a request block cannot fit
    // beside them: refused at submit
    {
        int64_t budget = 0, resident = 0, reserved = 0, peak = 0;
        OK(dtcf_server_usage(srv, &budget, &resident, &reserved, &peak));
        void *fill_all = nullptr;
        OK(dtcf_server_alloc(srv, budget - resident - 4096, &fill_all));
        dtcf_program p = make(0, DTCF_LAYOUT_SEPARATE);
        double *o = nullptr;
        hipMalloc(&o, ni * ni * 8);
        int64_t id = -1;
        if (p) CHECK(dtcf_request_submit(srv, p, &o, &id) == DTCF_ERR_BUDGET, "budget refusal");
        hipFree(o);
        if (p) OK(dtcf_program_destroy(p));
        OK(dtcf_server_free(srv, fill_all));
    }
    // mixed program (INT8, five digits) against the FP64 result
    dtcf_program mp = make(DTCF_PREC_I8, DTCF_LAYOUT_AUTO);
    if (mp) {
        double *o = nullptr;
        hipMalloc(&o, ni * ni * 8);
        int64_t id = -1;
        OK(dtcf_request_submit(srv, mp, &o, &id));
        OK(dtcf_request_wait(srv, id, 60000.0));
        OK(dtcf_request_release(srv, id));
        CHECK(maxrel(host_copy(o, ni * ni), Y) < 1e-8, "mixed served result");
        hipFree(o);
        OK(dtcf_program_destroy(mp));
    }
    CHECK(dtcf_request_wait(srv, 987654, 0.0) != DTCF_OK, "unknown request");
    int64_t budget, resident, reserved, peak;
    OK(dtcf_server_usage(srv, &budget, &resident, &reserved, &peak));
    CHECK(reserved == 0, "every block released");
    OK(dtcf_server_free(srv, dA));
    OK(dtcf_server_free(srv, dB));
    OK(dtcf_server_close(srv));
}

static void ccsd_step()
{
    const int Oa = 2, Ob = 3, Va = 4, Vb = 3;
    std::vector<std::vector<double>> host;
    auto arr = [&](size_t n, unsigned seed, double scale) {
        std::vector<double> h(n);
        fill(h, seed);
        for (auto &e : h) e *= scale;
        host.push_back(h);
        return dev_copy(h);
    };
    std::vector<double> eoh = {-1.5, -1.0, -0.9}, evh = {0.5, 0.8, 1.1, 1.7};
    double *eoa = dev_copy(std::vector<double>(eoh.begin(), eoh.begin() + Oa));
    double *eob = dev_copy(eoh), *eva = dev_copy(evh);
    double *evb = dev_copy(std::vector<double>(evh.begin(), evh.begin() + Vb));
    dtcf_uhf_ccsd d{};
    d.Oa = Oa;
    d.Ob = Ob;
    d.Va = Va;
    d.Vb = Vb;
    d.eoa = eoa;
    d.eob = eob;
    d.eva = eva;
    d.evb = evb;
    d.fa_ov = arr(Oa * Va, 20, 0.1);
    d.fb_ov = arr(Ob * Vb, 21, 0.1);
    d.vaa_oovv = arr(Oa * Oa * Va * Va, 22, 0.1);
    d.vab_oovv = arr(Oa * Ob * Va * Vb, 23, 0.1);
    d.vbb_oovv = arr(Ob * Ob * Vb * Vb, 24, 0.1);
    d.ra = arr(Va * Oa, 25, 0.01);
    d.rb = arr(Vb * Ob, 26, 0.01);
    d.raa = arr(Va * Va * Oa * Oa, 27, 0.01);
    d.rab = arr(Va * Vb * Oa * Ob, 28, 0.01);
    d.rbb = arr(Vb * Vb * Ob * Ob, 29, 0.01);
    d.ta = arr(Va * Oa, 30, 0.01);
    d.tb = arr(Vb * Ob, 31, 0.01);
    d.taa = arr(Va * Va * Oa * Oa, 32, 0.01);
    d.tab = arr(Va * Vb * Oa * Ob, 33, 0.01);
    d.tbb = arr(Vb * Vb * Ob * Ob, 34, 0.01);
    double *o1 = nullptr, *o2 = nullptr;
    hipMalloc(&o1, DTCF_CC_NOUT * 8);
    hipMalloc(&o2, DTCF_CC_NOUT * 8);
    OK(dtcf_uhf_ccsd_step_f64(&d, 0, o1, nullptr));
    OK(dtcf_uhf_ccsd_step_f64_noalloc(&d, 0, o2, nullptr));
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    std::vector<double> a = host_copy(o1, DTCF_CC_NOUT), b = host_copy(o2, DTCF_CC_NOUT);
    CHECK(std::memcmp(a.data(), b.data(), a.size() * 8) == 0, "no-alloc step bitwise");
    CHECK(std::fabs(a[DTCF_CC_E] - (a[DTCF_CC_E1] + a[DTCF_CC_EAA] + a[DTCF_CC_EAB] + a[DTCF_CC_EBB])) <
              1e-14 * (1.0 + std::fabs(a[DTCF_CC_E])),
          "energy is the sum of its blocks");
    OK(dtcf_uhf_ccsd_step_f64(&d, 1, o1, nullptr));
    OK(hipDeviceSynchronize() == hipSuccess ? DTCF_OK : DTCF_ERR_HIP);
    dtcf_uhf_ccsd bad = d;
    bad.Oa = 0;
    CHECK(dtcf_uhf_ccsd_step_f64(&bad, 1, o1, nullptr) != DTCF_OK, "bad descriptor");
    const void *bufs[] = {d.fa_ov, d.fb_ov, d.vaa_oovv, d.vab_oovv, d.vbb_oovv, d.ra, d.rb, d.raa, d.rab,
                          d.rbb, d.ta, d.tb, d.taa, d.tab, d.tbb, eoa, eob, eva, evb, o1, o2};
    for (const void *p : bufs) hipFree(const_cast<void *>(p));
}

int main()
{
    if (dtcf_init(0) != DTCF_OK) {
        std::printf("FAIL init: %s\n", dtcf_last_error());
        return 1;
    }
    views_and_gemms();
    std::printf("views and gemms done\n");
    rrr_step(false);
    rrr_step(true);
    std::printf("rrr step done\n");
    dltc_terms();
    std::printf("dltc done\n");
    mixed_products();
    std::printf("mixed done\n");
    serving();
    std::printf("serving done\n");
    ccsd_step();
    std::printf("ccsd step done\n");
    if (g_fail) {
        std::printf("host sanitize: %d failures\n", g_fail);
        return 1;
    }
    std::printf("host sanitize: ok\n");
    return 0;
}
