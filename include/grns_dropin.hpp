// grns_dropin.hpp -- the GRNS 1.1.4 host API names over libgrns_hip.so (include/grns.h).
//
// Include after the application's params.h (the moduli-set contract of src/params.h:11-39:
// RNS_MODULI_SIZE, RNS_MODULI_PRODUCT_LOG2, RNS_MODULI, ...).  rns_const_init() then builds the
// context of that moduli set on the current HIP device, and the calls below keep the reference's
// names, argument meaning and blocking behaviour:
//
//   rns_const_init()                        src/rnsbase.cuh:324   constants of the params.h set
//   mpint_const_init()                      src/mpint.cuh:20      (nothing left to initialise)
//   cuda::rns_max<g1,b1,g2,b2>(out,in,N,buf) src/dpp/rnsmax.cuh:129 arg-max over N numbers stored
//                                                                 number-major ([x1..xn][x1..xn]...);
//                                                                 the template arguments are accepted
//                                                                 and ignored (launch shapes are chosen
//                                                                 by the library); blocks until done
//
// and array forms of the per-number device functions (one call for `count` numbers, device
// pointers, number-major layout like the reference's test programs):
//
//   cuda::rns_eval_compute_array(out, x, count)       rns_eval_compute      src/rnseval.cuh:182
//   cuda::rns_eval_compute_fast_array(out, x, count)  rns_eval_compute_fast src/rnseval.cuh:266
//   cuda::rns_cmp_array(sign, x, y, count)            rns_cmp               src/rnscmp.cuh:55
//   cuda::rns_div_array(q, r, x, d, count)            rns_div               src/rnsdiv.cuh:128
//   cuda::rns_div_fast_array(q, r, x, d, count)       rns_div_fast          src/rnsdiv.cuh:180
//   cuda::rns_scale2pow_array(y, x, D, count)         rns_scale2pow         src/rnsscal.cuh:207
//
// Results are those of the reference's CPU path (bit for bit), which is also what its GPU path
// computes for comparison signs, quotients, remainders, scaled values and arg-max results.
// Errors are reported on stderr and abort (the reference prints and exits, src/common.cuh:31-37).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "grns.h"

#ifndef RNS_MODULI_SIZE
#error "include the application's params.h before grns_dropin.hpp"
#endif

namespace grns_dropin {

inline grns_ctx *&context() {
    static grns_ctx *c = nullptr;
    return c;
}

inline void check(grns_status s, const char *where) {
    if (s != GRNS_OK) {
        std::fprintf(stderr, "GRNS error %d in %s\n", (int)s, where);
        std::exit(1);
    }
}

inline grns_ctx *ctx() {
    if (!context()) check(GRNS_ERR_ARG, "rns_const_init() was not called");
    return context();
}

}  // namespace grns_dropin

inline void rns_const_init() {
    if (grns_dropin::context()) return;
    int dev = 0;
    (void)hipGetDevice(&dev);
    grns_dropin::check(grns_ctx_create(RNS_MODULI, RNS_MODULI_SIZE, dev, &grns_dropin::context()), "rns_const_init");
}

inline void mpint_const_init() {}

namespace cuda {

template <int gridSize1, int blockSize1, int gridSize2, int blockSize2, class XInterval>
void rns_max(int *out, int *in, unsigned int N, XInterval *buffer) {
    grns_ctx *c = grns_dropin::ctx();
    long long *idx = nullptr;
    grns_dropin::check(hipMalloc(&idx, sizeof(long long)) == hipSuccess ? GRNS_OK : GRNS_ERR_OOM, "rns_max");
    size_t need = grns_max_workspace_size(c, N);
    size_t have = sizeof(XInterval) * (size_t)N;
    void *ws = buffer;
    bool own = have < need;
    if (own) grns_dropin::check(hipMalloc(&ws, need) == hipSuccess ? GRNS_OK : GRNS_ERR_OOM, "rns_max");
    grns_dropin::check(grns_max(c, in, 1, RNS_MODULI_SIZE, N, out, idx, ws, own ? need : have, nullptr), "rns_max");
    grns_dropin::check(hipDeviceSynchronize() == hipSuccess ? GRNS_OK : GRNS_ERR_HIP, "rns_max");
    if (own) (void)hipFree(ws);
    (void)hipFree(idx);
}

inline void rns_eval_compute_array(grns_interval *out, const int *x, size_t count) {
    grns_dropin::check(grns_eval(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, count, out, 0, nullptr), "rns_eval_compute");
}
inline void rns_eval_compute_fast_array(grns_interval *out, const int *x, size_t count) {
    grns_dropin::check(grns_eval(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, count, out, 1, nullptr), "rns_eval_compute_fast");
}
inline void rns_cmp_array(int *sign, const int *x, const int *y, size_t count) {
    grns_dropin::check(grns_cmp(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, y, 1, RNS_MODULI_SIZE, count, sign, nullptr),
                       "rns_cmp");
}
inline void rns_div_array(int *q, int *r, const int *x, const int *d, size_t count) {
    grns_dropin::check(grns_div(grns_dropin::ctx(), x, d, 1, RNS_MODULI_SIZE, count, q, r, 0, nullptr), "rns_div");
}
inline void rns_div_fast_array(int *q, int *r, const int *x, const int *d, size_t count) {
    grns_dropin::check(grns_div(grns_dropin::ctx(), x, d, 1, RNS_MODULI_SIZE, count, q, r, 1, nullptr), "rns_div_fast");
}
inline void rns_scale2pow_array(int *y, const int *x, const unsigned *D, size_t count) {
    grns_dropin::check(grns_scale2pow(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, D, count, y, nullptr), "rns_scale2pow");
}

}  // namespace cuda
