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
/* Test hook: on = 1 makes grns_eval (n = 16 and 32, modulus-major rows aligned to 16 bytes) stage
 * its residues through double-buffered LDS tiles instead of loading them straight into registers,
 * the default.  Both forms give the same results; the tiled one is slower on gfx950 (0.58 against
 * 0.49 ms for 2^24 numbers at n = 32: its 185 VGPRs halve the waves per SIMD). */
grns_status grns_debug_stream_tiles(grns_ctx *ctx, int on);
/* Test hook: caps the grid of the per-number streaming kernels of grns_eval, grns_cmp and the mpint
 * operations (grid-stride loops) at blocks_per_cu blocks per compute unit; 0 restores one block per
 * 256 numbers.  Results do not depend on it; the hook lets grid shapes be timed on the device. */
grns_status grns_debug_grid_cap(grns_ctx *ctx, int blocks_per_cu);
/* Scheduler test hook: a server whose launches run nothing and complete delay_us after issue;
 * it accepts host-only contexts (device -1), so admission and batching run without a GPU. */
grns_status grns_serve_create_sim(const grns_serve_config *cfg, int delay_us, grns_server **out);
#ifdef __cplusplus
}
#endif
#endif
