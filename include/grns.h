/*
 * grns.h -- C ABI of libgrns_hip.so, the MI355X (gfx950) implementation of the GRNS
 * data-parallel RNS hot path.
 *
 * Plain C: pointers, sizes and status codes.  Device pointers are HIP device memory owned by
 * the caller; `stream` is a hipStream_t passed as void* (NULL = the default stream).  Calls are
 * asynchronous on the stream unless stated otherwise.
 *
 * Residue arrays are described by a base pointer and two strides (in int32 elements): the
 * residue of number j modulo m_i is at base[i * mod_stride + j * num_stride].  Modulus-major
 * (mod_stride = count or more, num_stride = 1) is the layout the kernels are tuned for;
 * number-major (the reference's [x1..xn][x1..xn]... layout: mod_stride = 1, num_stride = n) is
 * accepted by the same kernels.
 *
 * Results are bit-identical to the CPU path of GRNS 1.1.4 (the global-namespace host functions
 * named in each comment) on the same inputs.
 */
#ifndef GRNS_H
#define GRNS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GRNS_OK = 0,
    GRNS_ERR_ARG = 1,          /* bad argument (null pointer, size, moduli set)            */
    GRNS_ERR_HIP = 2,          /* a HIP runtime call failed                               */
    GRNS_ERR_OOM = 3,          /* device allocation failed                                */
    GRNS_ERR_UNSUPPORTED = 4,  /* moduli-set size not supported by this entry point       */
    GRNS_ERR_ADMISSION = 5,    /* serving: request exceeds the HBM budget                 */
    GRNS_ERR_BUSY = 6          /* serving: queue full, retry later                        */
} grns_status;

/* er_float_t (src/types.cuh:13-16): value = frac * 2^exp */
typedef struct { double frac; long long exp; } grns_er;
/* interval_t (src/types.cuh:22-25) */
typedef struct { grns_er low; grns_er upp; } grns_interval;

typedef struct grns_ctx grns_ctx;

/* Device status bits raised by kernels for inputs on which the reference itself reads outside
 * its tables or does not terminate (grns_ctx_status). */
#define GRNS_STATUS_REFINE_RANGE 1   /* refinement factor outside RNS_POW2 (rnseval.cuh:82-83) */
#define GRNS_STATUS_DIV_LOOP     2   /* division loop would not terminate (rnsdiv.cuh:34-41, 82-94) */

/* ---------------------------------------------------------------- contexts
 * Replaces rns_const_init() / mpint_const_init() (src/rnsbase.cuh:324-451, src/mpint.cuh:20-32):
 * every constant table of one moduli set, resident on `device`.  n in [1, 256]; moduli >= 2,
 * pairwise coprime, ascending (the params.h contract, src/params.h:11-39).  Synchronous. */
grns_status grns_ctx_create(const int *moduli, int n, int device, grns_ctx **out);
grns_status grns_ctx_destroy(grns_ctx *ctx);
int grns_ctx_n(const grns_ctx *ctx);
int grns_ctx_log2m(const grns_ctx *ctx);
/* Host copies of the constants for inspection: part_inv[n], recip_rd[n], recip_ru[n],
 * accuracy, ref_factor, unit/inv_unit (RNS_EVAL_UNIT / RNS_EVAL_INV_UNIT). Any pointer may be NULL. */
grns_status grns_ctx_constants(const grns_ctx *ctx, int *part_inv, double *recip_rd, double *recip_ru,
                               double *accuracy, int *ref_factor, grns_interval *unit, grns_interval *inv_unit);
/* RNS_POW2 row-major ((log2m+1) x n) and MRC_MULT_INV (n x n) host copies. */
grns_status grns_ctx_tables(const grns_ctx *ctx, int *pow2, int *mrc_inv);
/* Synchronise the device and return (and clear) the status bits raised since the last call. */
grns_status grns_ctx_status(grns_ctx *ctx, int *bits);

/* ---------------------------------------------------------------- element-wise operations */

/* rns_eval_compute (fast = 0, src/rnseval.cuh:29-102) or rns_eval_compute_fast (fast = 1,
 * :112-167) of `count` numbers; out[j] is the interval of number j. */
grns_status grns_eval(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride, size_t count,
                      grns_interval *out, int fast, void *stream);

/* rns_cmp (src/rnscmp.cuh:20-39): sign[j] = 1, 0, -1 for x_j >, =, < y_j. */
grns_status grns_cmp(grns_ctx *ctx, const int *x, long long x_mod_stride, long long x_num_stride, const int *y,
                     long long y_mod_stride, long long y_num_stride, size_t count, int *sign, void *stream);

/* rns_div (src/rnsdiv.cuh:19-61) or rns_div_fast (:67-114): q, r with the layout of x (x, d, q, r
 * share mod_stride / num_stride).  `flags` (0 and 1 keep their old meaning of `fast`):
 *   GRNS_DIV_FAST       rns_div_fast rather than rns_div;
 *   GRNS_DIV_REFERENCE  the reference's loop for every pair, so that every result is the CPU path's
 *                       bit for bit.  Without it, pairs of canonical residues with d != 0 are divided
 *                       on their binary forms: the exact quotient and remainder, which is also the
 *                       CPU path's result except on the pairs where the reference's reciprocal
 *                       reduction wraps int32 (src/modular.cuh:74-79) and its loop returns another
 *                       number (about one pair in 10^6 on uniform operands); all other pairs (and
 *                       the pairs with d = 0 or non-canonical residues) run the reference loop.
 * Other bits: GRNS_ERR_ARG. */
#define GRNS_DIV_FAST 1
#define GRNS_DIV_REFERENCE 2
grns_status grns_div(grns_ctx *ctx, const int *x, const int *d, long long mod_stride, long long num_stride,
                     size_t count, int *q, int *r, int flags, void *stream);

/* rns_scale2pow (src/rnsscal.cuh:93-119): y_j = x_j scaled by 2^D[j]. */
grns_status grns_scale2pow(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride,
                           const unsigned *D, size_t count, int *y, void *stream);

/* Mixed-radix digits (mrc, src/mrc.cuh:31-44; the reference's GPU forms src/mrc.cuh:101-181):
 * digit k of number j at d[k * d_mod_stride + j * d_num_stride], x = sum_k d_k * m_0 ... m_{k-1}. */
grns_status grns_mrc(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride, size_t count, int *d,
                     long long d_mod_stride, long long d_num_stride, void *stream);

/* Binary <-> RNS conversion.  Integers are little-endian 32-bit words: word k of number j at
 * w[j * num_stride + k * word_stride].
 * grns_from_binary: rns_from_binary (src/rnsbase.cuh:253-261, mpz_mod_ui per modulus): residue i
 * of number j is |x_j| mod m_i, and m_i minus that (when nonzero) if sign[j] != 0; sign may be
 * NULL (all nonnegative); nwords >= 1 words per number, any length.
 * grns_to_binary: rns_to_binary (src/rnsbase.cuh:277-287): the CRT value 0 <= X < M of each number
 * as grns_ctx_words(ctx) words; a negative residue contributes as the reference's unsigned long
 * conversion of it does (x_i + 2^64). */
int grns_ctx_words(const grns_ctx *ctx);
grns_status grns_from_binary(grns_ctx *ctx, const unsigned *w, long long word_stride, long long num_stride,
                             int nwords, const int *sign, size_t count, int *x, long long mod_stride,
                             long long num_stride_x, void *stream);
grns_status grns_to_binary(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride_x, size_t count,
                           unsigned *w, long long word_stride, long long num_stride, void *stream);

/* Arg-max of `count` numbers: cuda::rns_max (src/dpp/rnsmax.cuh:129-138).  out_residues[n]
 * receives the residues of the maximum, out_idx its index (the lowest index among equal
 * maxima), both in device memory.  workspace: device memory of grns_max_workspace_size bytes. */
size_t grns_max_workspace_size(const grns_ctx *ctx, size_t count);
grns_status grns_max(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride, size_t count,
                     int *out_residues, long long *out_idx, void *workspace, size_t workspace_bytes, void *stream);

/* grns_max without the workspace argument, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  The stage-1 candidates live in a file-scope device array of
 * GRNS_MAX_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends; the call allocates
 * nothing.  Same arguments (minus the workspace) and results as grns_max; GRNS_ERR_ARG when the
 * candidates of `count` numbers would not fit the array (never for n <= 64), GRNS_ERR_UNSUPPORTED
 * for moduli-set sizes other than 4, 8, 16, 32, 64.  Every call on one GPU uses the same array:
 * calls that overlap on that GPU (other streams, other threads) race and must be serialised by
 * the caller. */
#define GRNS_MAX_NOALLOC_SCRATCH_BYTES ((size_t)2293760)  /* 8192 candidates x (24 + 4 * 64) bytes */
grns_status grns_max_noalloc(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride, size_t count,
                             int *out_residues, long long *out_idx, void *stream);

/* Arg-max by mixed-radix digits: rns_max_mrc (tests/test_perf_rnsmax.cu:145-153, with
 * rns_max_mrc_compute_kernel and rns_max_mrc_tree_kernel :109-138).  Every number is converted to
 * mixed-radix digits (src/mrc.cuh:31-44) and the largest digit vector wins (mrs_cmp,
 * src/mrc.cuh:13-22).  Same arguments, workspace and results as grns_max: residues are taken mod
 * m_i, equal maxima give the lowest index. */
grns_status grns_max_mrc(grns_ctx *ctx, const int *x, long long mod_stride, long long num_stride, size_t count,
                         int *out_residues, long long *out_idx, void *workspace, size_t workspace_bytes,
                         void *stream);

/* mp_int_t (src/types.cuh:45-49) as arrays: digits modulus-major (digits[i * ld + j]),
 * sign[count], eval[2 * count] (eval[2j] lower, eval[2j+1] upper bound). */
typedef struct {
    int *digits;
    int *sign;
    grns_er *eval;
    long long ld;
} grns_mpint_array;

/* mpint_add / mpint_sub / mpint_mul (src/mpint.cuh:160-278), element-wise z = x op y.
 * overflow (optional, device, count bytes): mpint_check_overflow (src/mpint.cuh:142-152) as
 * 0 none, 1 overflow, 2 possible overflow. */
grns_status grns_mpint_add(grns_ctx *ctx, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *overflow, void *stream);
grns_status grns_mpint_sub(grns_ctx *ctx, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *overflow, void *stream);
grns_status grns_mpint_mul(grns_ctx *ctx, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *overflow, void *stream);
/* mpint_div (src/mpint.cuh:286-296): Euclidean quotient z = x / y (digits, sign, eval); the eval
 * fields of x and y are not read, and overflow (may be NULL) is not written. */
grns_status grns_mpint_div(grns_ctx *ctx, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *overflow, void *stream);

/* mpint_add / mpint_sub / mpint_mul / mpint_div (src/mpint.cuh:160-296) over arrays of the
 * reference's own mp_int_t (src/types.cuh:45-49: int digits[n]; int sign; er_float_t eval[2]; element
 * j at (char *)x + j * elem_bytes), for code that keeps that layout (the drop-in header's
 * cuda::mpint_*_array).  op: GRNS_OP_MPINT_ADD, _SUB, _MUL (overflow as in grns_mpint_add) or
 * GRNS_OP_MPINT_DIV (overflow not written).  elem_bytes must be grns_mpint_aos_bytes(ctx) =
 * sizeof(mp_int_t) for the context's n.  z may be x or y.  Results are those of the array forms. */
#define GRNS_OP_MPINT_DIV 10
size_t grns_mpint_aos_bytes(const grns_ctx *ctx);
grns_status grns_mpint_aos(grns_ctx *ctx, int op, const void *x, const void *y, void *z, size_t count,
                           size_t elem_bytes, unsigned char *overflow, void *stream);

/* ---------------------------------------------------------------- batched form (no reference
 * counterpart): many arrays of uneven length and several operations of one moduli set fused into
 * one persistent, work-stealing launch that can be captured in a hipGraph.  Each task is one of
 * the element-wise entry points above applied to its own arrays; results are identical to the
 * per-call entry points. */
enum {
    GRNS_OP_EVAL = 0,       /* grns_eval, fast = 0 : a -> out0 (grns_interval[count])        */
    GRNS_OP_EVAL_FAST = 1,  /* grns_eval, fast = 1                                           */
    GRNS_OP_CMP = 2,        /* grns_cmp            : a, b -> out0 (int[count])               */
    GRNS_OP_MPINT_ADD = 3,  /* grns_mpint_add      : mx, my -> mz (+ overflow)               */
    GRNS_OP_MPINT_SUB = 4,
    GRNS_OP_MPINT_MUL = 5,
    GRNS_OP_DIV = 6,        /* grns_div, fast = 0  : a (x), b (d) -> out0 (q), out1 (r); a, b, q, r share a's strides */
    GRNS_OP_DIV_FAST = 7,
    GRNS_OP_SCALE2POW = 8   /* grns_scale2pow      : a, D -> out0 (same strides as a)        */
};

typedef struct {
    int op;
    size_t count;
    const int *a;
    long long a_mod_stride, a_num_stride;
    const int *b;
    long long b_mod_stride, b_num_stride;
    void *out0;
    void *out1;
    const unsigned *D;
    grns_mpint_array mx, my, mz;   /* mpint operands (digits ld >= count) */
    unsigned char *overflow;       /* mpint overflow flags, optional       */
} grns_batch_task;

/* Launch shape: numbers per work tile, persistent blocks per CU, and the mapping of numbers to
 * lanes (threads_per_number = 1: one lane per number; 0: one group of 2^ceil(log2 n) lanes per
 * number).  Division and scaling always use groups. */
typedef struct {
    int tile;
    int blocks_per_cu;
    int threads_per_number;
    int use_graph;
} grns_batch_config;

typedef struct grns_batch grns_batch;
grns_status grns_batch_create(grns_ctx *ctx, const grns_batch_task *tasks, int ntasks, grns_batch **out);
grns_status grns_batch_set_config(grns_batch *b, const grns_batch_config *cfg);
grns_status grns_batch_get_config(const grns_batch *b, grns_batch_config *cfg);
/* Times the candidate launch shapes on this batch (reps launches each) and keeps the fastest. */
grns_status grns_batch_tune(grns_batch *b, void *stream, int reps, grns_batch_config *best, double *best_ms);
/* Runs every task of the batch (one persistent launch per mapping; replayed from a hipGraph when
 * use_graph is set). */
grns_status grns_batch_launch(grns_batch *b, void *stream);
grns_status grns_batch_destroy(grns_batch *b);

/* ---------------------------------------------------------------- serving form (no reference
 * counterpart): a request loop on one device.  A request is a short list of tasks (the batch task
 * of above; op GRNS_OP_MAX adds the arg-max: a -> out0 (int[n] residues), out1 (long long index))
 * whose operands and results are already resident in device memory.  A dispatcher thread admits
 * requests in arrival order while their bytes fit the HBM budget and a slot (stream) is free,
 * fuses the admitted requests of one moduli set into one persistent launch, and lets launches on
 * different slots overlap on the device.  Latency = completion seen by the dispatcher minus
 * submission. */
#define GRNS_OP_MAX 9

typedef struct {
    size_t hbm_budget;   /* bytes of operands + results of requests in flight            */
    int slots;           /* concurrent launches, one HIP stream each (default priority)  */
    int max_batch;       /* requests fused into one launch                              */
    int tile;            /* launch shape of the fused launches (see grns_batch_config)  */
    int blocks_per_cu;
} grns_serve_config;

typedef struct {
    long long submitted, completed, rejected, launches;
    double p50_ms, p99_ms, mean_ms, max_ms;
} grns_serve_stats;

typedef struct grns_server grns_server;
grns_status grns_serve_create(int device, const grns_serve_config *cfg, grns_server **out);
/* Queue a request; *ticket identifies it.  GRNS_ERR_ADMISSION if it can never fit the budget. */
grns_status grns_serve_submit(grns_server *s, grns_ctx *ctx, const grns_batch_task *tasks, int ntasks,
                              long long *ticket);
/* state: 0 queued, 1 running, 2 done, 3 failed; latency_ms valid when done. */
grns_status grns_serve_poll(grns_server *s, long long ticket, int *state, double *latency_ms);
grns_status grns_serve_wait(grns_server *s, long long ticket, double *latency_ms);
grns_status grns_serve_drain(grns_server *s);
grns_status grns_serve_stats_get(grns_server *s, grns_serve_stats *out);
grns_status grns_serve_destroy(grns_server *s);

/* Version string of the library. */
const char *grns_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GRNS_H */
