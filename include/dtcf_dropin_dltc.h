/*
 * dtcf_dropin_dltc.h -- drop-in DLTC::compute (DLTC/compute.h:28-29) over the
 * C ABI.  compute takes the reference's DLTC::Multiplication; include this
 * header after the reference's expr.h and expand DTCF_DLTC_DROPIN_DEFINE()
 * in one translation unit in place of compiling compute.cc.
 *
 * buf_c[m x n] += buf_a[m x k] * buf_b[k x n], row-major, alpha = beta = 1,
 * with m, n, k the products of the tile widths of M.m_ids(), n_ids(),
 * k_ids() (compute.cc:9-17), on the GPU (host buffers staged).
 */
#ifndef DTCF_DROPIN_DLTC_H
#define DTCF_DROPIN_DLTC_H

#include <cstdio>
#include <cstdlib>

#include "dtcf_local.h"

namespace dtcf_dltc {

// SizeFn maps an index-name list to the product of its tile widths
// (the reference's getSize, define.h:86-93).
template <class MulT, class SizeFn>
void compute(double *buf_c, double *buf_a, double *buf_b, MulT &M, SizeFn size)
{
    const int64_t m = (int64_t)size(M.m_ids());
    const int64_t n = (int64_t)size(M.n_ids());
    const int64_t k = (int64_t)size(M.k_ids());
    if (dtcf_gemm_f64_host(0, 0, m, n, k, 1.0, buf_a, k, buf_b, n, buf_c, n) != DTCF_OK) {
        std::fprintf(stderr, "compute: %s\n", dtcf_last_error());
        std::abort();
    }
}

}  // namespace dtcf_dltc

#define DTCF_DLTC_DROPIN_DEFINE()                                                      \
    namespace DLTC {                                                                   \
    void compute(double *buf_c, double *buf_a, double *buf_b, Multiplication &M)       \
    {                                                                                  \
        dtcf_dltc::compute(buf_c, buf_a, buf_b, M,                                     \
                           [](const vector<IndexName> &ids) { return getSize(ids); }); \
    }                                                                                  \
    }

#endif /* DTCF_DROPIN_DLTC_H */
