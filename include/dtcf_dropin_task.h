/*
 * dtcf_dropin_task.h -- drop-in DLTC::TaskFunc (DLTC/task.h:29-37,
 * DLTC/task.cc:140-167) over the C ABI.
 *
 * The reference's task function reads a TaskDesc (task.h:17-23: expression
 * type, expression id, tile box itlb/itub) and runs task_mul or task_sum
 * (task.cc:51-137) of the id-th Multiplication / Summation over that box,
 * moving tiles through Global Arrays.  Here the tensors are dense device
 * arrays (dtcf_dltc_tensor) and each expression id maps to a planned
 * multiplication (dtcf_dltc_mul_create) or a summation description, held in
 * dtcf_task::registry() by the caller in place of the reference's global
 * vectors m[] and s[].  The task box runs on the GPU, enqueued on the
 * registry's stream.
 *
 * Include this header in one translation unit of the caller after the
 * reference's task.h (or headers declaring the same TaskDesc, ExprType,
 * NIDX and TaskPool / task_t) and expand DTCF_TASK_DROPIN_DEFINE() (the
 * Scioto form, TaskFunc(TaskPool, task_t *)) or
 * DTCF_TASK_DROPIN_DEFINE_TASCEL() (the TASCEL form) there, in place of
 * compiling task.cc's TaskFunc.  An expression id without an entry, or a
 * failed launch, prints the library's message and aborts: the reference
 * has no error path here.
 */
#ifndef DTCF_DROPIN_TASK_H
#define DTCF_DROPIN_TASK_H

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dtcf_local.h"

namespace dtcf_task {

struct Sum {
    dtcf_dltc_space space;
    dtcf_dltc_tensor c, a;
    double coef;
    bool set = false;
};

struct Registry {
    std::vector<dtcf_dltc_mul> mul;  // by Multiplication id (the reference's m[id])
    std::vector<Sum> sum;            // by Summation id (s[id])
    void *stream = nullptr;          // hipStream_t the task boxes are enqueued on
};

inline Registry &registry()
{
    static Registry r;
    return r;
}

inline void fail(const char *what)
{
    std::fprintf(stderr, "TaskFunc: %s: %s\n", what, dtcf_last_error());
    std::abort();
}

// One task: the box [itlb[x], itub[x]) for every index name x < nidx.
inline void run(bool is_mul, int id, const int *itlb, const int *itub, int nidx)
{
    int32_t lb[DTCF_DLTC_NIDX] = {0}, ub[DTCF_DLTC_NIDX] = {0};
    for (int x = 0; x < nidx && x < DTCF_DLTC_NIDX; x++) {
        lb[x] = itlb[x];
        ub[x] = itub[x];
    }
    Registry &r = registry();
    if (is_mul) {
        if (id < 0 || id >= (int)r.mul.size() || !r.mul[id]) {
            std::fprintf(stderr, "TaskFunc: no multiplication registered for id %d\n", id);
            std::abort();
        }
        if (dtcf_dltc_mul_task(r.mul[id], lb, ub, r.stream) != DTCF_OK) fail("task_mul");
    } else {
        if (id < 0 || id >= (int)r.sum.size() || !r.sum[id].set) {
            std::fprintf(stderr, "TaskFunc: no summation registered for id %d\n", id);
            std::abort();
        }
        const Sum &s = r.sum[id];
        if (dtcf_dltc_sum(&s.space, &s.c, &s.a, s.coef, lb, ub, r.stream) != DTCF_OK)
            fail("task_sum");
    }
}

}  // namespace dtcf_task

/* DLTC::TaskFunc, Scioto form (task.h:37). */
#define DTCF_TASK_DROPIN_DEFINE()                                                           \
    namespace DLTC {                                                                        \
    void TaskFunc(TaskPool tp, task_t *task)                                                \
    {                                                                                       \
        (void)tp;                                                                           \
        TaskDesc *t = (TaskDesc *)gtc_task_body(task);                                      \
        dtcf_task::run(t->et == MUL, t->id, t->itlb, t->itub, NIDX);                        \
    }                                                                                       \
    }

/* DLTC::TaskFunc, TASCEL form (task.h:29-32). */
#define DTCF_TASK_DROPIN_DEFINE_TASCEL()                                                    \
    namespace DLTC {                                                                        \
    void TaskFunc(TaskPool *tp, void *_bigd, int bigd_len, void *pldata, int pldata_len,    \
                  vector<void *> data_bufs, int thd)                                        \
    {                                                                                       \
        (void)tp;                                                                           \
        (void)bigd_len;                                                                     \
        (void)pldata;                                                                       \
        (void)pldata_len;                                                                   \
        (void)data_bufs;                                                                    \
        (void)thd;                                                                          \
        TaskDesc *t = (TaskDesc *)_bigd;                                                    \
        dtcf_task::run(t->et == MUL, t->id, t->itlb, t->itub, NIDX);                        \
    }                                                                                       \
    }

#endif /* DTCF_DROPIN_TASK_H */
