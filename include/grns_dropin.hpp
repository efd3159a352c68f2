// grns_dropin.hpp -- the GRNS 1.1.4 host API names over libgrns_hip.so (include/grns.h).
//
// Include after the application's params.h (the moduli-set contract of src/params.h:11-39:
// RNS_MODULI_SIZE, RNS_MODULI_PRODUCT_LOG2, RNS_MODULI, ...) in place of the GRNS headers.
// rns_const_init() then builds the context of that moduli set on the current HIP device, and the
// calls below keep the reference's names, types, argument meaning and blocking behaviour:
//
//   er_float_t, interval_t, xinterval_t, mp_int_t  src/types.cuh:13-49 (same layout; checked
//                                                   against grns_er / grns_interval below)
//   rns_const_init()                        src/rnsbase.cuh:324   constants of the params.h set
//   mpint_const_init()                      src/mpint.cuh:20      (nothing left to initialise)
//   cuda::rns_max<g1,b1,g2,b2>(out,in,N,buf) src/dpp/rnsmax.cuh:129 arg-max over N numbers stored
//                                                                 number-major ([x1..xn][x1..xn]...);
//                                                                 the template arguments are accepted
//                                                                 and ignored (launch shapes are chosen
//                                                                 by the library); blocks until done
//
// and array forms of the per-number device functions (one call for `count` numbers, device
// pointers, number-major layout like the reference's test programs, blocking):
//
//   cuda::rns_eval_compute_array(out, x, count)       rns_eval_compute      src/rnseval.cuh:182
//   cuda::rns_eval_compute_fast_array(out, x, count)  rns_eval_compute_fast src/rnseval.cuh:266
//   cuda::rns_cmp_array(sign, x, y, count)            rns_cmp               src/rnscmp.cuh:55
//   cuda::rns_div_array(q, r, x, d, count)            rns_div               src/rnsdiv.cuh:128
//   cuda::rns_div_fast_array(q, r, x, d, count)       rns_div_fast          src/rnsdiv.cuh:180
//   cuda::rns_scale2pow_array(y, x, D, count)         rns_scale2pow         src/rnsscal.cuh:207
//   cuda::mpint_add_array(z, x, y, count)             mpint_add             src/mpint.cuh:336
//   cuda::mpint_sub_array(z, x, y, count)             mpint_sub             src/mpint.cuh:407
//   cuda::mpint_mul_array(z, x, y, count)             mpint_mul             src/mpint.cuh:478
//   cuda::mpint_div_array(z, x, y, count)             mpint_div             src/mpint.cuh:511
//   (mp_int_t arrays as the reference lays them out; z may be x or y)
//
// Results are those of the reference's CPU path, bit for bit, which is also what its GPU path
// computes for comparison signs, quotients, remainders, scaled values and arg-max results.
// Division: by default the pairs of canonical residues with d != 0 are divided on their binary
// forms, which returns the exact quotient and remainder; that is the reference's result except on
// the pairs where its reciprocal reduction wraps int32 (src/modular.cuh:74-79) and its loop returns
// another number (about one pair in 10^6 on uniform operands).  Define
// GRNS_DROPIN_REFERENCE_DIVISION to 1 before including this header to run the reference's loop for
// every pair instead (GRNS_DIV_REFERENCE: those results too, at about 1 us per pair at n = 64).
// Errors are reported on stderr and end the program (the reference prints and exits,
// src/common.cuh:31-37).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <cstdlib>

#include "grns.h"

#ifndef RNS_MODULI_SIZE
#error "include the application's params.h before grns_dropin.hpp"
#endif

#ifndef GRNS_DROPIN_REFERENCE_DIVISION
#define GRNS_DROPIN_REFERENCE_DIVISION 0
#endif

// src/types.cuh:13-49 (skipped when the reference's own types.cuh was included first)
#ifndef GRNS_TYPES_CUH
#define GRNS_TYPES_CUH
typedef struct {
    double frac;  // significand
    long exp;     // exponent
} er_float_t;
typedef er_float_t *er_float_ptr;

typedef struct {
    er_float_t low;
    er_float_t upp;
} interval_t;
typedef interval_t *interval_ptr;

typedef struct {
    er_float_t low;
    er_float_t upp;
    int idx;
} xinterval_t;

typedef struct {
    int digits[RNS_MODULI_SIZE];
    int sign;
    er_float_t eval[2];
} mp_int_t;
typedef mp_int_t *mp_int_ptr;
#endif

static_assert(sizeof(er_float_t) == sizeof(grns_er) && offsetof(er_float_t, exp) == offsetof(grns_er, exp),
              "er_float_t and grns_er share one layout");
static_assert(sizeof(interval_t) == sizeof(grns_interval) && offsetof(interval_t, upp) == offsetof(grns_interval, upp),
              "interval_t and grns_interval share one layout");
static_assert(offsetof(mp_int_t, sign) == 4 * RNS_MODULI_SIZE &&
                  offsetof(mp_int_t, eval) == (4 * RNS_MODULI_SIZE + 4 + 7) / 8 * 8,
              "mp_int_t: digits, sign, then eval at the next 8-byte boundary");

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

inline void sync(const char *where) { check(hipDeviceSynchronize() == hipSuccess ? GRNS_OK : GRNS_ERR_HIP, where); }

inline void mpint_op(int op, mp_int_t *z, const mp_int_t *x, const mp_int_t *y, size_t count, const char *where) {
    check(grns_mpint_aos(ctx(), op, x, y, z, count, sizeof(mp_int_t), nullptr, nullptr), where);
    sync(where);
}

}  // namespace grns_dropin

inline void rns_const_init() {
    if (grns_dropin::context()) return;
    int dev = 0;
    (void)hipGetDevice(&dev);
    grns_dropin::check(grns_ctx_create(RNS_MODULI, RNS_MODULI_SIZE, dev, &grns_dropin::context()), "rns_const_init");
    grns_dropin::check(grns_ctx_log2m(grns_dropin::context()) == RNS_MODULI_PRODUCT_LOG2 ? GRNS_OK : GRNS_ERR_ARG,
                       "rns_const_init: RNS_MODULI_PRODUCT_LOG2 is not floor(log2 M) of RNS_MODULI");
}

inline void mpint_const_init() {}

namespace cuda {

template <int gridSize1, int blockSize1, int gridSize2, int blockSize2>
void rns_max(int *out, int *in, unsigned int N, xinterval_t *buffer) {
    grns_ctx *c = grns_dropin::ctx();
    long long *idx = nullptr;
    grns_dropin::check(hipMalloc(&idx, sizeof(long long)) == hipSuccess ? GRNS_OK : GRNS_ERR_OOM, "rns_max");
    size_t need = grns_max_workspace_size(c, N);
    size_t have = sizeof(xinterval_t) * (size_t)N;
    void *ws = buffer;
    bool own = have < need;
    if (own) grns_dropin::check(hipMalloc(&ws, need) == hipSuccess ? GRNS_OK : GRNS_ERR_OOM, "rns_max");
    grns_dropin::check(grns_max(c, in, 1, RNS_MODULI_SIZE, N, out, idx, ws, own ? need : have, nullptr), "rns_max");
    grns_dropin::sync("rns_max");
    if (own) (void)hipFree(ws);
    (void)hipFree(idx);
}

inline void rns_eval_compute_array(interval_t *out, const int *x, size_t count) {
    grns_dropin::check(grns_eval(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, count, reinterpret_cast<grns_interval *>(out),
                                 0, nullptr),
                       "rns_eval_compute");
    grns_dropin::sync("rns_eval_compute");
}
inline void rns_eval_compute_fast_array(interval_t *out, const int *x, size_t count) {
    grns_dropin::check(grns_eval(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, count, reinterpret_cast<grns_interval *>(out),
                                 1, nullptr),
                       "rns_eval_compute_fast");
    grns_dropin::sync("rns_eval_compute_fast");
}
inline void rns_cmp_array(int *sign, const int *x, const int *y, size_t count) {
    grns_dropin::check(grns_cmp(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, y, 1, RNS_MODULI_SIZE, count, sign, nullptr),
                       "rns_cmp");
    grns_dropin::sync("rns_cmp");
}
inline void rns_div_array(int *q, int *r, const int *x, const int *d, size_t count) {
    const int flags = GRNS_DROPIN_REFERENCE_DIVISION ? GRNS_DIV_REFERENCE : 0;
    grns_dropin::check(grns_div(grns_dropin::ctx(), x, d, 1, RNS_MODULI_SIZE, count, q, r, flags, nullptr), "rns_div");
    grns_dropin::sync("rns_div");
}
inline void rns_div_fast_array(int *q, int *r, const int *x, const int *d, size_t count) {
    const int flags = GRNS_DIV_FAST | (GRNS_DROPIN_REFERENCE_DIVISION ? GRNS_DIV_REFERENCE : 0);
    grns_dropin::check(grns_div(grns_dropin::ctx(), x, d, 1, RNS_MODULI_SIZE, count, q, r, flags, nullptr),
                       "rns_div_fast");
    grns_dropin::sync("rns_div_fast");
}
inline void rns_scale2pow_array(int *y, const int *x, const unsigned *D, size_t count) {
    grns_dropin::check(grns_scale2pow(grns_dropin::ctx(), x, 1, RNS_MODULI_SIZE, D, count, y, nullptr), "rns_scale2pow");
    grns_dropin::sync("rns_scale2pow");
}
inline void mpint_add_array(mp_int_t *z, const mp_int_t *x, const mp_int_t *y, size_t count) {
    grns_dropin::mpint_op(GRNS_OP_MPINT_ADD, z, x, y, count, "mpint_add");
}
inline void mpint_sub_array(mp_int_t *z, const mp_int_t *x, const mp_int_t *y, size_t count) {
    grns_dropin::mpint_op(GRNS_OP_MPINT_SUB, z, x, y, count, "mpint_sub");
}
inline void mpint_mul_array(mp_int_t *z, const mp_int_t *x, const mp_int_t *y, size_t count) {
    grns_dropin::mpint_op(GRNS_OP_MPINT_MUL, z, x, y, count, "mpint_mul");
}
inline void mpint_div_array(mp_int_t *z, const mp_int_t *x, const mp_int_t *y, size_t count) {
    grns_dropin::mpint_op(GRNS_OP_MPINT_DIV, z, x, y, count, "mpint_div");
}

}  // namespace cuda
