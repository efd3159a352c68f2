/*
 * grns_debug.h -- host-side test hooks of libgrns_hip.so.  These run the same __host__ __device__
 * routines the kernels use, on the CPU, so that they can be compared with a CPU restatement of
 * the reference without a GPU.  They have no counterpart in the reference.  A context made with
 * device = -1 (host-only) is enough for all of them.
 */
#ifndef GRNS_DEBUG_H
#define GRNS_DEBUG_H
#include "grns.h"
#ifdef __cplusplus
extern "C" {
#endif
int grns_debug_eval_hot(const grns_ctx *ctx, const int *x, grns_interval *out, int fast);
double grns_debug_filter(const grns_ctx *ctx, const int *x, int *regular, int *canon);
int grns_debug_cmp_lane(const grns_ctx *ctx, const int *x, const int *y);
double grns_debug_max_key(const grns_ctx *ctx, const int *x);
int grns_debug_eval_full(const grns_ctx *ctx, const int *x, grns_interval *out, int fast, int *status);
int grns_debug_cmp_full(const grns_ctx *ctx, const int *x, const int *y);
int grns_debug_ceil_log2(double s);
double grns_debug_mulmod_recip_d(double x, double y, double m, double d);
int grns_debug_mulmod_recip(int x, int y, int m, double d);
int grns_debug_addmod_recip(int x, int y, int m, double d);
double grns_debug_bn_err(double c);
double grns_debug_fast_scalbn(double x, int n);
/* The word division's quotient estimate, a three-word by two-word step (d1 with its top bit set, <u2, u1> < <d1, d0>): quotient, the
 * remainder in *r1, *r0 and the two-word reciprocal in *v. */
unsigned grns_debug_div32(unsigned u2, unsigned u1, unsigned u0, unsigned d1, unsigned d0, unsigned *r1,
                          unsigned *r0, unsigned *v);
double grns_debug_filter_tau(const grns_ctx *ctx);
double grns_debug_filter_err(const grns_ctx *ctx);
/* Kernel timing: while enabled, grns_eval, grns_cmp and grns_max record HIP events on the caller's
 * stream around their streaming kernel (op 0: the evaluation kernel, 1: the comparison kernel,
 * 2: arg-max stage 1).  grns_debug_stage_read waits for the recorded launches and returns the summed
 * kernel time in ms and the launch count since the last read, then clears them.  Not thread-safe. */
grns_status grns_debug_stage_timing(grns_ctx *ctx, int enable);
grns_status grns_debug_stage_read(grns_ctx *ctx, int op, double *ms_total, long long *launches);
/* Test hook: on != 0 makes grns_div (as its GRNS_DIV_REFERENCE flag does) and grns_scale2pow run
 * the reference's loops (src/rnsdiv.cuh, src/rnsscal.cuh) for every number instead of the word
 * forms, so that the two can be compared on the device. */
grns_status grns_debug_reference_loops(grns_ctx *ctx, int on);
/* Test hook: caps the grid of the per-number streaming kernels of grns_eval, grns_cmp and the mpint
 * operations (grid-stride loops) at blocks_per_cu blocks per compute unit; 0 restores one block per
 * 256 numbers.  Results do not depend on it; the hook lets grid shapes be timed on the device. */
grns_status grns_debug_grid_cap(grns_ctx *ctx, int blocks_per_cu);
/* Test hook: runs stage 1 of the arg-max (grns_max, grns_max_noalloc, grns_max_mrc) on nblocks blocks (at most
 * 8192, stage 2's capacity); 0 restores the default grid (from the occupancy query).
 * Results do not depend on it; the hook lets stage-1 grids be timed on the device. */
grns_status grns_debug_max_blocks(grns_ctx *ctx, int nblocks);
/* Test hook: on = 0 stops the streaming kernels of grns_eval and grns_cmp from copying the residues
 * of the numbers they queue for the rare pass (which then reads them from the inputs); 1, the
 * default, restores it.  Results do not depend on it. */
grns_status grns_debug_rare_staging(grns_ctx *ctx, int on);
/* Test hook: on = 0 puts every rare item of a call into one queue with one atomic counter for the
 * device; 1, the default, spreads them over sub-queues with a counter each, by block.  Results do
 * not depend on it. */
grns_status grns_debug_rare_spread(grns_ctx *ctx, int on);
/* Baseline of the batched form: runs every task of the batch through its per-call entry point
 * (grns_eval, grns_cmp, grns_mpint_*, grns_div, grns_scale2pow), one call per task in table order
 * on `stream`, from C++ -- what a caller without the batched form does.  Results equal those of
 * grns_batch_launch. */
grns_status grns_debug_batch_per_call(grns_batch *b, void *stream);
/* Scheduler test hook: a server whose launches run nothing and complete delay_us after issue;
 * it accepts host-only contexts (device -1), so admission and batching run without a GPU. */
grns_status grns_serve_create_sim(const grns_serve_config *cfg, int delay_us, grns_server **out);
#ifdef __cplusplus
}
#endif
#endif
