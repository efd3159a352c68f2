/*
 * grns_oracle.h -- CPU restatement of the GRNS 1.1.4 CPU path (TEST INFRASTRUCTURE ONLY).
 *
 * This is the checker that the HIP library is compared against.  It restates, operation by
 * operation, the host (global-namespace) code of the reference library with the moduli set
 * taken at run time instead of from params.h.  Nothing in the product (grns_amd/, include/)
 * may include, link or call it; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it.
 *
 * Semantics that must be kept exactly (they define the bit patterns being compared):
 *   - round-to-nearest FP64 with contraction OFF (built with -ffp-contract=off); the directed
 *     roundings are the BoundNear1 emulation of src/dinterval.cuh:37-147;
 *   - reciprocal modular reduction r - (int)((double)r * (1.0/m)) * m (src/modular.cuh:29-34,74-79),
 *     which is not always canonical;
 *   - the pairwise sum trees with zero padding of src/pairwise.cuh:121-458;
 *   - x86-64 conversions: (unsigned int)double, (int)double.
 *
 * Parity pinning of each function is listed in the header comment of grns_oracle.c.
 */
#ifndef GRNS_ORACLE_H
#define GRNS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MODULI 256
#define ORC_SCALING_THRESHOLD 30   /* RNS_P2_SCALING_THRESHOLD, src/params.h:22 */

/* er_float_t / interval_t with the binary layout of src/types.cuh:13-24 */
typedef struct { double frac; long exp; } orc_er_t;
typedef struct { orc_er_t low; orc_er_t upp; } orc_interval_t;

/* All constants of rns_const_init (src/rnsbase.cuh:324-451), for one moduli set. */
typedef struct {
    int n;
    int log2M;                       /* floor(log2 M)  == RNS_MODULI_PRODUCT_LOG2 */
    int moduli[ORC_MAX_MODULI];
    int part_inv[ORC_MAX_MODULI];    /* RNS_PART_MODULI_PRODUCT_INVERSE, w_i */
    double recip_rd[ORC_MAX_MODULI]; /* RNS_MODULI_RECIP_RD */
    double recip_ru[ORC_MAX_MODULI]; /* RNS_MODULI_RECIP_RU */
    int *pow2;                       /* RNS_POW2[(log2M+1)][n] */
    int prod_pow2_res[ORC_SCALING_THRESHOLD];                 /* M mod 2^(j+1) */
    int part_pow2_res[ORC_SCALING_THRESHOLD][ORC_MAX_MODULI]; /* M_i mod 2^(j+1) */
    int pow2_inv[ORC_SCALING_THRESHOLD][ORC_MAX_MODULI];      /* (2^(j+1))^-1 mod m_i */
    int *mrc_inv;                    /* MRC_MULT_INV[n][n], row-major */
    double accuracy;                 /* RNS_EVAL_ACCURACY */
    int ref_factor;                  /* RNS_EVAL_REF_FACTOR */
    orc_interval_t unit;             /* RNS_EVAL_UNIT  (frac in [0.5,1)) */
    orc_interval_t inv_unit;         /* RNS_EVAL_INV_UNIT */
    int one[ORC_MAX_MODULI];         /* RNS_ONE */
} orc_ctx;

/* Context; returns NULL on bad input (n out of range, modulus < 2). */
orc_ctx *orc_ctx_create(const int *moduli, int n);
void orc_ctx_destroy(orc_ctx *c);

/* Scalar primitives exposed for unit tests. */
double orc_dadd_rd(double a, double b);
double orc_dadd_ru(double a, double b);
double orc_dsub_rd(double a, double b);
double orc_dsub_ru(double a, double b);
double orc_dmul_rd(double a, double b);
double orc_dmul_ru(double a, double b);
double orc_ddiv_rd(double a, double b);
double orc_ddiv_ru(double a, double b);
double orc_fast_scalbn(double x, int n);
double orc_psum_rd(const double *x, int len);
double orc_psum_ru(const double *x, int len);
int orc_mod_add(int x, int y, int m, double d);
int orc_mod_mul(int x, int y, int m, double d);
int orc_mod_mul3(int x, int y, int m);
int orc_mod_psub(int x, int y, int m);
void orc_er_set_d(orc_er_t *r, double x);
void orc_er_adjust(orc_er_t *x);
void orc_er_add_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_add_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_sub_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_sub_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_mul_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_mul_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_div_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_div_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y);
void orc_er_md_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y, orc_er_t *z);
void orc_er_md_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y, orc_er_t *z);
int orc_er_ucmp(const orc_er_t *x, const orc_er_t *y);
int orc_er_cmp(const orc_er_t *x, const orc_er_t *y);

/* One-number operations (number-major residue vectors of length c->n). */
void orc_mrc(const orc_ctx *c, int *mr, const int *x);
int orc_mrc_compare(const orc_ctx *c, const int *x, const int *y);
void orc_eval(const orc_ctx *c, orc_interval_t *out, const int *x);
void orc_eval_fast(const orc_ctx *c, orc_interval_t *out, const int *x);
int orc_cmp(const orc_ctx *c, const int *x, const int *y);
/* Division; returns the number of main-loop iterations, or -1 when the loop cap was hit
 * (a case in which the reference would not terminate). */
int orc_div(const orc_ctx *c, int *q, int *r, const int *x, const int *d);
int orc_div_fast(const orc_ctx *c, int *q, int *r, const int *x, const int *d);
void orc_scale2pow(const orc_ctx *c, int *result, const int *x, unsigned int D);

/* mp_int_t in structure-of-arrays pieces: digits[n], sign, eval[2]. */
typedef struct { int *digits; int *sign; orc_er_t *eval; } orc_mpint_ref;
/* Overflow flag (mpint_check_overflow, src/mpint.cuh:142-152): 0 none, 1 overflow, 2 possible. */
int orc_check_overflow(const orc_ctx *c, const orc_er_t *lower, const orc_er_t *upper);
int orc_mpint_add(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                  const int *xd, int xs, const orc_er_t *xe, const int *yd, int ys, const orc_er_t *ye);
int orc_mpint_sub(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                  const int *xd, int xs, const orc_er_t *xe, const int *yd, int ys, const orc_er_t *ye);
int orc_mpint_mul(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                  const int *xd, int xs, const orc_er_t *xe, const int *yd, int ys, const orc_er_t *ye);

/* Array forms (modulus-major residues: element j, modulus i at x[i*ld + j]).
 * nthreads <= 0 uses all OpenMP threads; 1 runs serially. */
void orc_eval_array(const orc_ctx *c, const int *x, size_t ld, size_t count, orc_interval_t *out, int fast, int nthreads);
void orc_cmp_array(const orc_ctx *c, const int *x, const int *y, size_t ld, size_t count, int *sign, int nthreads);
void orc_div_array(const orc_ctx *c, const int *x, const int *d, size_t ld, size_t count, int *q, int *r, int fast, int nthreads);
void orc_scale2pow_array(const orc_ctx *c, const int *x, const unsigned int *D, size_t ld, size_t count, int *y, int nthreads);
/* Arg-max with rns_max_cmp semantics; ties resolve to the lowest index. Returns the index. */
long orc_max_array(const orc_ctx *c, const int *x, size_t ld, size_t count, int *out_residues, int nthreads);
/* The reference's rns_max pipeline (evaluate, then reduce with rns_max_cmp) on the CPU path. */
long orc_max_eval_array(const orc_ctx *c, const int *x, size_t ld, size_t count, int nthreads);
/* mpint_div (src/mpint.cuh:286-296, CPU path); -1 where the division loop would not end. */
int orc_mpint_div(const orc_ctx *c, int *rd, int *rs, orc_er_t *re, const int *xd, int xs, const int *yd, int ys);
void orc_mpint_div_array(const orc_ctx *c, size_t ld, size_t count, const int *xd, const int *xs, const int *yd,
                         const int *ys, int *rd, int *rs, orc_er_t *re, int nthreads);
/* mpint elementwise op over SoA arrays: op 0 add, 1 sub, 2 mul. Digits modulus-major with ld. */
void orc_mpint_array(const orc_ctx *c, int op, size_t ld, size_t count,
                     const int *xd, const int *xs, const orc_er_t *xe,
                     const int *yd, const int *ys, const orc_er_t *ye,
                     int *rd, int *rs, orc_er_t *re, unsigned char *ovf, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
