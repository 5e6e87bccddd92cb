// Timing of the drop-in entry points at the reference's call shapes, from
// C++ through the exported symbols (libdtcf_dropin.so) and the TaskFunc
// header adapter.  Prints one JSON object.
//
//  * local_contraction (localcontract.cc:3-21) at the ring step's shape,
//    9216 x 9216 x 9216 with host operands of 679.5 MB each, as
//    kevin_dgemm calls it (contraction.cc:1956): every call stages A, B and
//    C to the device and C back (PCIe-inclusive rate).
//  * the CCD ladder R[p3,p4,h1,h2] += 1/2 T[p5,p6,h1,h2] V[p3,p4,p5,p6] at
//    o=8, v=32 with tiles wbO=4, wbV=8 (1024 (ext tile, sum tile) tasks),
//    first as the reference's task_mul issues it per tile (task.cc:97-133):
//    tile get, transpose_klij of the A and B tiles, compute, and per ext tile
//    the transpose back and the coef-scaled accumulate, with host tiles
//    through the drop-in symbols; then as DLTC::TaskFunc tasks over
//    device-resident tensors (dtcf_dropin_task.h).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dtcf_dropin.h"
#include "dtcf_local.h"

namespace DLTC {
enum ExprType { MUL, SUM };
#define NIDX 24
struct TaskDesc {
    ExprType et;
    int id;
    int itlb[NIDX];
    int itub[NIDX];
    int idvm[NIDX];
};
}  // namespace DLTC
typedef struct gtc_stub *gtc_t;
typedef gtc_t TaskPool;
struct task_t {
    DLTC::TaskDesc body;
};
inline void *gtc_task_body(task_t *t) { return &t->body; }

#include "dtcf_dropin_task.h"
DTCF_TASK_DROPIN_DEFINE()

static double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void fill(std::vector<double> &v, unsigned s)
{
    uint64_t x = 0x9E3779B97F4A7C15ull * (s + 1);
    for (auto &e : v) {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        e = (double)((x * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}

int main()
{
    // ------------------------------------------------ local_contraction, C2 shape
    const int n = 9216;
    const size_t nn = (size_t)n * n;
    std::vector<double> A(nn), B(nn), C(nn);
    fill(A, 1);
    fill(B, 2);
    fill(C, 3);
    local_contraction(n, n, n, A.data(), B.data(), C.data(), 0, 1, 1.0);  // grows the buffers
    const int reps = 3;
    double t0 = now();
    for (int r = 0; r < reps; r++) local_contraction(n, n, n, A.data(), B.data(), C.data(), 0, 1, 1.0);
    const double lc = (now() - t0) / reps;
    A.clear();
    A.shrink_to_fit();
    B.clear();
    B.shrink_to_fit();
    C.clear();
    C.shrink_to_fit();

    // ------------------------------------------------ C1 ladder, reference-shaped tasks
    const int O = 8, V = 32, wO = 4, wV = 8, sO = O / wO, sV = V / wV;
    std::vector<double> T((size_t)V * V * O * O), Vv((size_t)V * V * V * V), R(T.size(), 0.0);
    fill(T, 4);
    fill(Vv, 5);
    std::vector<double> ta(wV * wV * wO * wO), tat(ta.size()), tb(wV * wV * wV * wV), tbt(tb.size()),
        ct(wO * wO * wV * wV), cr(ct.size());
    int ntask = 0;
    t0 = now();
    for (int a = 0; a < sV; a++)
        for (int b = 0; b < sV; b++)
            for (int i = 0; i < sO; i++)
                for (int j = 0; j < sO; j++) {
                    std::fill(ct.begin(), ct.end(), 0.0);
                    for (int c = 0; c < sV; c++)
                        for (int d = 0; d < sV; d++) {
                            // A tile T[p5,p6,h1,h2] and B tile V[p3,p4,p5,p6] (the GA gets)
                            for (int x = 0; x < wV; x++)
                                for (int y = 0; y < wV; y++)
                                    for (int u = 0; u < wO; u++)
                                        std::memcpy(&ta[((x * wV + y) * wO + u) * wO],
                                                    &T[(((size_t)(c * wV + x) * V + d * wV + y) * O +
                                                        i * wO + u) * O + j * wO],
                                                    wO * sizeof(double));
                            for (int x = 0; x < wV; x++)
                                for (int y = 0; y < wV; y++)
                                    for (int u = 0; u < wV; u++)
                                        std::memcpy(&tb[((x * wV + y) * wV + u) * wV],
                                                    &Vv[(((size_t)(a * wV + x) * V + b * wV + y) * V +
                                                         c * wV + u) * V + d * wV],
                                                    wV * sizeof(double));
                            transpose_klij(ta.data(), tat.data(), wV, wV, wO, wO);  // -> h1,h2,p5,p6
                            transpose_klij(tb.data(), tbt.data(), wV, wV, wV, wV);  // -> p5,p6,p3,p4
                            local_contraction(wO * wO, wV * wV, wV * wV, tat.data(), tbt.data(),
                                              ct.data(), 0, 0, 1.0);
                            ntask++;
                        }
                    transpose_klij(ct.data(), cr.data(), wO, wO, wV, wV);  // -> p3,p4,h1,h2
                    for (int x = 0; x < wV; x++)
                        for (int y = 0; y < wV; y++)
                            for (int u = 0; u < wO; u++)
                                for (int w = 0; w < wO; w++)
                                    R[(((size_t)(a * wV + x) * V + b * wV + y) * O + i * wO + u) * O +
                                      j * wO + w] += 0.5 * cr[((x * wV + y) * wO + u) * wO + w];
                }
    const double host_tasks = now() - t0;

    // ------------------------------------------------ the same tasks as TaskFunc
    double *dT, *dV, *dR;
    hipMalloc(&dT, T.size() * 8);
    hipMalloc(&dV, Vv.size() * 8);
    hipMalloc(&dR, R.size() * 8);
    hipMemcpy(dT, T.data(), T.size() * 8, hipMemcpyHostToDevice);
    hipMemcpy(dV, Vv.data(), Vv.size() * 8, hipMemcpyHostToDevice);
    hipMemset(dR, 0, R.size() * 8);
    enum { h1 = 0, h2 = 1, p3 = 14, p4 = 15, p5 = 16, p6 = 17 };
    dtcf_dltc_space sp = {12, O, V, wO, wV};
    dtcf_dltc_tensor tC{}, tA{}, tB{};
    const int ic[4] = {p3, p4, h1, h2}, ia[4] = {p5, p6, h1, h2}, ib[4] = {p3, p4, p5, p6};
    tC.dims = tA.dims = tB.dims = 4;
    for (int q = 0; q < 4; q++) {
        tC.ids[q] = ic[q];
        tA.ids[q] = ia[q];
        tB.ids[q] = ib[q];
    }
    tC.data = dR;
    tA.data = dT;
    tB.data = dV;
    dtcf_dltc_mul mh = nullptr;
    if (dtcf_dltc_mul_create(&sp, &tC, &tA, &tB, 0.5, &mh) != DTCF_OK) {
        std::printf("{\"error\": \"%s\"}\n", dtcf_last_error());
        return 1;
    }
    dtcf_task::registry().mul = {mh};
    auto run_tasks = [&]() {
        for (int a = 0; a < sV; a++)
            for (int b = 0; b < sV; b++)
                for (int i = 0; i < sO; i++)
                    for (int j = 0; j < sO; j++)
                        for (int c = 0; c < sV; c++)
                            for (int d = 0; d < sV; d++) {
                                task_t task{};
                                task.body.et = DLTC::MUL;
                                const int idx[6] = {p3, p4, h1, h2, p5, p6}, val[6] = {a, b, i, j, c, d};
                                for (int q = 0; q < 6; q++) {
                                    task.body.itlb[idx[q]] = val[q];
                                    task.body.itub[idx[q]] = val[q] + 1;
                                }
                                DLTC::TaskFunc(nullptr, &task);
                            }
    };
    run_tasks();  // first calls set kernel attributes
    hipDeviceSynchronize();
    hipMemset(dR, 0, R.size() * 8);
    hipDeviceSynchronize();
    t0 = now();
    run_tasks();
    hipDeviceSynchronize();
    const double dev_tasks = now() - t0;
    std::vector<double> Rd(R.size());
    hipMemcpy(Rd.data(), dR, R.size() * 8, hipMemcpyDeviceToHost);
    double num = 0, den = 0;
    for (size_t e = 0; e < R.size(); e++) {
        num = std::max(num, std::abs(Rd[e] - R[e]));
        den = std::max(den, std::abs(R[e]));
    }
    dtcf_dltc_mul_destroy(mh);
    hipFree(dT);
    hipFree(dV);
    hipFree(dR);
    const double fl = 2.0 * n * (double)n * n, bytes = 4.0 * nn * 8;
    std::printf("{\"local_contraction_9216\": {\"ms\": %.3f, \"tflops\": %.3f, \"host_bytes_moved\": %.4g, "
                "\"note\": \"host operands, staged to the device and back every call\"}, "
                "\"c1_ladder_reference_shaped_tasks\": {\"tasks\": %d, \"ms\": %.3f, \"us_per_task\": %.2f, "
                "\"note\": \"per task: host tile gets, transpose_klij x2, local_contraction; per ext tile "
                "transpose_klij back and accumulate\"}, "
                "\"c1_ladder_taskfunc\": {\"tasks\": %d, \"ms\": %.3f, \"us_per_task\": %.2f, "
                "\"max_rel_diff_vs_reference_shaped\": %.3g}}\n",
                lc * 1e3, fl / lc / 1e12, bytes, ntask, host_tasks * 1e3, host_tasks * 1e6 / ntask, ntask,
                dev_tasks * 1e3, dev_tasks * 1e6 / ntask, den > 0 ? num / den : num);
    return 0;
}
