/*
 * grns_oracle.c -- CPU restatement of the GRNS 1.1.4 CPU path.  TEST INFRASTRUCTURE ONLY:
 * the checker for the HIP library, never part of the product.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off, no -ffast-math, no -march: round-to-nearest FP64 with every
 * product and sum rounded separately, as the reference's host code).
 *
 * Each function cites the reference lines it restates.  How each one is pinned
 * (tests/test_oracle_*.py):
 *   constants (rns_const_init)  - closed forms with exact Python integers/rationals
 *   BoundNear1 ops, psum trees  - pred/succ enclosure checks, tree shape by exact rationals,
 *                                 special values (zero -> +/-eta)
 *   er_* (extended range)       - exact rational value checks and normalisation invariants
 *   mod_* / rns_*               - exact integers (canonical cases) and the analytic
 *                                 non-canonical cases (r == m, quotient rounding)
 *   mrc                         - mixed-radix digits against Python big-int expansion,
 *                                 brute force over the {3,5,7,11} moduli set
 *   eval / eval_fast            - exact-rational enclosure low <= x/M <= upp for every x
 *                                 of the {3,5,7,11} set, random x for 8/32/64 moduli,
 *                                 2^j sweep (tests/test_perf_rnseval.cu:655-687 pattern)
 *   cmp, div, scale2pow, max    - exact integer answers (sign, q/r, floor(x/2^D), argmax)
 *   mpint add/sub/mul           - exact signed integers (KAT -55, 13 from
 *                                 tests/test_verify_mpint.cu:92-93) plus enclosure of evals
 * The bit patterns of every floating-point result (eval, eval_fast, the pairwise sums, er_*,
 * the mpint eval fields) are pinned by tests/test_oracle_bitpins.py against a second model built
 * from pred/succ (nextafter), round-to-nearest and exact integers, with closed forms for one and
 * two nonzero leaves, all-zero sums and exact extended-range operands; tests/test_oracle_mutants.py
 * builds this file with planted mistakes (ORC_MUTANT below) and shows that those pins fail.
 */
#include "grns_oracle.h"

/* Planted mistakes, for tests/test_oracle_mutants.py only: a build with -DORC_MUTANT=k changes one
 * step of the restatement so that the tests can show their pins fail on it.  The default build
 * (ORC_MUTANT 0, oracle/Makefile) has none of them.
 *   1 tree without the top-level zero pad      2 lower/upper reciprocals swapped in the evaluation
 *   3 refinement by reciprocal products        4 refinement factor without the +1
 *   5 integer part dropped with a directed subtraction
 *   6 fused fmac (std::fma) in BoundNear1      7 sequential instead of pairwise summation
 *   8 er_add without directed rounding         9 mpint_mul with the two UNIT bounds swapped
 *  10 no zero test before the sums            11 the ambiguity branch sets upp to INV_UNIT.low
 *  12 refinement shifts only the upper bound  13 refinement lower bound from the unrefined sums */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/bitwise.cuh:34-44 */
#define ORC_EPS 0x1p-53
#define ORC_ETA 0x1p-1074
static const double ORC_PHI1 = ORC_EPS * (1 + 2 * ORC_EPS);
#define ORC_SIGN_BIT ((uint64_t)1 << 63)
#define ORC_ZERO_EXP_MASK ((((uint64_t)1 << 52) - 1) | ORC_SIGN_BIT)
#define ORC_BIAS_EXP_MASK ((uint64_t)1023 << 52)

typedef union { double d; uint64_t u; } orc_du;

/* x86-64 (unsigned int)double: cvttsd2si to 64 bits, keep the low 32 (|x| < 2^63). */
static inline unsigned int orc_to_uint(double x) { return (unsigned int)(long long)x; }

/* ------------------------------------------------------------------ BoundNear1
 * src/dinterval.cuh:25-31 (fmac = x*y+z, EMPLOY_STD_FMA false), :37-147 */
#if ORC_MUTANT == 6
static inline double orc_fmac(double x, double y, double z) { return fma(x, y, z); }
#else
static inline double orc_fmac(double x, double y, double z) { return x * y + z; }
#endif
double orc_dadd_rd(double a, double b) { double c = a + b; return c - orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_dadd_ru(double a, double b) { double c = a + b; return c + orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_dsub_rd(double a, double b) { double c = a - b; return c - orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_dsub_ru(double a, double b) { double c = a - b; return c + orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_dmul_rd(double a, double b) { double c = a * b; return c - orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_dmul_ru(double a, double b) { double c = a * b; return c + orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_ddiv_rd(double a, double b) { double c = a / b; return c - orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }
double orc_ddiv_ru(double a, double b) { double c = a / b; return c + orc_fmac(ORC_PHI1, fabs(c), ORC_ETA); }

/* src/bitwise.cuh:75-80 */
double orc_fast_scalbn(double x, int n) {
    orc_du u; u.d = x;
    u.u += (uint64_t)n << 52;
    return u.d * (x != 0 && n >= -1023);
}

/* ------------------------------------------------------------------ pairwise sums
 * src/pairwise.cuh:227-327 (round-down tree), :121-223 (round-up), :384-458 (dispatch).
 * psumK<L>: s1 = psum(K/2)<L>(x); s2 = L > K/2 ? psum(K/2)<L-K/2>(x+K/2) : 0; return s1 (+) s2,
 * with psum2<L> = x0 (+) (L > 1 ? x1 : 0).  The dispatch picks K = the power of two above L
 * for 4 <= L < 4096 (8192 for L >= 4096), and special-cases L <= 3. */
static double orc_tree_rd(const double *x, int K, int L) {
    if (K == 2) return orc_dadd_rd(x[0], (L > 1) ? x[1] : 0);
    double s1 = orc_tree_rd(x, K / 2, L);
    double s2 = (L > K / 2) ? orc_tree_rd(x + K / 2, K / 2, L - K / 2) : 0;
    return orc_dadd_rd(s1, s2);
}
static double orc_tree_ru(const double *x, int K, int L) {
    if (K == 2) return orc_dadd_ru(x[0], (L > 1) ? x[1] : 0);
    double s1 = orc_tree_ru(x, K / 2, L);
    double s2 = (L > K / 2) ? orc_tree_ru(x + K / 2, K / 2, L - K / 2) : 0;
    return orc_dadd_ru(s1, s2);
}
static int orc_tree_width(int len) {
    if (len >= 4096) return 8192;
#if ORC_MUTANT == 1
    int K = 4;
    while (K < len) K <<= 1;
#else
    int K = 8;
    while (K <= len) K <<= 1;
#endif
    return K;
}
double orc_psum_rd(const double *x, int len) {
#if ORC_MUTANT == 7
    if (len >= 4) { double a = x[0]; for (int i = 1; i < len; i++) a = orc_dadd_rd(a, x[i]); return a; }
#endif
    if (len >= 4) return orc_tree_rd(x, orc_tree_width(len), len);
    if (len == 3) return orc_dadd_rd(orc_dadd_rd(x[0], x[1]), x[2]);
    if (len == 2) return orc_dadd_rd(x[0], x[1]);
    if (len == 1) return x[0];
    return 0;
}
double orc_psum_ru(const double *x, int len) {
#if ORC_MUTANT == 7
    if (len >= 4) { double a = x[0]; for (int i = 1; i < len; i++) a = orc_dadd_ru(a, x[i]); return a; }
#endif
    if (len >= 4) return orc_tree_ru(x, orc_tree_width(len), len);
    if (len == 3) return orc_dadd_ru(orc_dadd_ru(x[0], x[1]), x[2]);
    if (len == 2) return orc_dadd_ru(x[0], x[1]);
    if (len == 1) return x[0];
    return 0;
}

/* ------------------------------------------------------------------ extended range
 * src/extrange.cuh:29-42 */
void orc_er_set_d(orc_er_t *r, double x) {
    orc_du u; u.d = x;
    if (x != 0) {
        r->exp = (long)(((u.u & ~ORC_SIGN_BIT) >> 52) - 1023);
        u.u = (u.u & ORC_ZERO_EXP_MASK) | ORC_BIAS_EXP_MASK;
        r->frac = u.d;
    } else {
        r->exp = 0;
        r->frac = 0;
    }
}
/* src/extrange.cuh:116-125 (unsigned 64-bit arithmetic, then back to long) */
void orc_er_adjust(orc_er_t *x) {
    orc_du u; u.d = x->frac;
    x->exp = (long)(((uint64_t)x->exp + ((u.u & ~ORC_SIGN_BIT) >> 52) - 1023) * (uint64_t)(x->frac != 0));
    u.u = (u.u & ORC_ZERO_EXP_MASK) | ORC_BIAS_EXP_MASK;
    x->frac = u.d * (x->frac != 0);
}
/* src/extrange.cuh:153-188; the inputs' exponents are reset when their fraction is zero */
#if ORC_MUTANT == 8
#define orc_dadd_rd(a, b) ((a) + (b))
#define orc_dadd_ru(a, b) ((a) + (b))
#endif
void orc_er_add_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    x->exp *= (x->frac != 0);
    y->exp *= (y->frac != 0);
    int dexp = (int)((x->exp - y->exp) * (x->frac != 0) * (y->frac != 0));
    orc_er_t res;
    if (dexp > 0) { res.exp = x->exp; res.frac = orc_dadd_rd(x->frac, orc_fast_scalbn(y->frac, -dexp)); }
    else if (dexp < 0) { res.exp = y->exp; res.frac = orc_dadd_rd(y->frac, orc_fast_scalbn(x->frac, dexp)); }
    else { res.exp = (x->exp == 0) ? y->exp : x->exp; res.frac = orc_dadd_rd(x->frac, y->frac); }
    orc_er_adjust(&res);
    *r = res;
}
void orc_er_add_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    x->exp *= (x->frac != 0);
    y->exp *= (y->frac != 0);
    int dexp = (int)((x->exp - y->exp) * (x->frac != 0) * (y->frac != 0));
    orc_er_t res;
    if (dexp > 0) { res.exp = x->exp; res.frac = orc_dadd_ru(x->frac, orc_fast_scalbn(y->frac, -dexp)); }
    else if (dexp < 0) { res.exp = y->exp; res.frac = orc_dadd_ru(y->frac, orc_fast_scalbn(x->frac, dexp)); }
    else { res.exp = (x->exp == 0) ? y->exp : x->exp; res.frac = orc_dadd_ru(x->frac, y->frac); }
    orc_er_adjust(&res);
    *r = res;
}
#if ORC_MUTANT == 8
#undef orc_dadd_rd
#undef orc_dadd_ru
#endif
/* src/extrange.cuh:213-248 */
void orc_er_sub_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    x->exp *= (x->frac != 0);
    y->exp *= (y->frac != 0);
    int dexp = (int)((x->exp - y->exp) * (x->frac != 0) * (y->frac != 0));
    orc_er_t res;
    if (dexp > 0) { res.exp = x->exp; res.frac = orc_dsub_rd(x->frac, orc_fast_scalbn(y->frac, -dexp)); }
    else if (dexp < 0) { res.exp = y->exp; res.frac = orc_dsub_rd(orc_fast_scalbn(x->frac, dexp), y->frac); }
    else { res.exp = (x->exp == 0) ? y->exp : x->exp; res.frac = orc_dsub_rd(x->frac, y->frac); }
    orc_er_adjust(&res);
    *r = res;
}
void orc_er_sub_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    x->exp *= (x->frac != 0);
    y->exp *= (y->frac != 0);
    int dexp = (int)((x->exp - y->exp) * (x->frac != 0) * (y->frac != 0));
    orc_er_t res;
    if (dexp > 0) { res.exp = x->exp; res.frac = orc_dsub_ru(x->frac, orc_fast_scalbn(y->frac, -dexp)); }
    else if (dexp < 0) { res.exp = y->exp; res.frac = orc_dsub_ru(orc_fast_scalbn(x->frac, dexp), y->frac); }
    else { res.exp = (x->exp == 0) ? y->exp : x->exp; res.frac = orc_dsub_ru(x->frac, y->frac); }
    orc_er_adjust(&res);
    *r = res;
}
/* src/extrange.cuh:262-320 */
void orc_er_mul_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    orc_er_t res; res.exp = x->exp + y->exp; res.frac = orc_dmul_rd(x->frac, y->frac);
    orc_er_adjust(&res); *r = res;
}
void orc_er_mul_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    orc_er_t res; res.exp = x->exp + y->exp; res.frac = orc_dmul_ru(x->frac, y->frac);
    orc_er_adjust(&res); *r = res;
}
void orc_er_div_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    orc_er_t res; res.exp = x->exp - y->exp; res.frac = orc_ddiv_rd(x->frac, y->frac);
    orc_er_adjust(&res); *r = res;
}
void orc_er_div_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y) {
    orc_er_t res; res.exp = x->exp - y->exp; res.frac = orc_ddiv_ru(x->frac, y->frac);
    orc_er_adjust(&res); *r = res;
}
void orc_er_md_rd(orc_er_t *r, orc_er_t *x, orc_er_t *y, orc_er_t *z) {
    orc_er_t res; res.exp = x->exp + y->exp - z->exp;
    res.frac = orc_ddiv_rd(orc_dmul_rd(x->frac, y->frac), z->frac);
    orc_er_adjust(&res); *r = res;
}
void orc_er_md_ru(orc_er_t *r, orc_er_t *x, orc_er_t *y, orc_er_t *z) {
    orc_er_t res; res.exp = x->exp + y->exp - z->exp;
    res.frac = orc_ddiv_ru(orc_dmul_ru(x->frac, y->frac), z->frac);
    orc_er_adjust(&res); *r = res;
}
/* src/extrange.cuh:327-353 */
int orc_er_cmp(const orc_er_t *x, const orc_er_t *y) {
    int diff = (int)(x->exp - y->exp);
    if (diff == 0) return (x->frac == y->frac ? 0 : x->frac > y->frac ? 1 : -1);
    if (x->frac == 0) return (y->frac > 0 ? -1 : 1);
    if (y->frac == 0) return (x->frac > 0 ? 1 : -1);
    return (diff > 0 && x->frac > 0) || (diff < 0 && y->frac < 0) ? 1 : -1;
}
int orc_er_ucmp(const orc_er_t *x, const orc_er_t *y) {
    int diff = (int)(x->exp - y->exp);
    if (diff == 0) return (x->frac == y->frac ? 0 : x->frac > y->frac ? 1 : -1);
    return x->frac != 0 && (diff > 0 || (diff < 0 && y->frac == 0)) ? 1 : -1;
}

/* ------------------------------------------------------------------ modular ops
 * src/modular.cuh:29-79 */
int orc_mod_add(int x, int y, int m, double d) {
    long r = (long)x + (long)y;
    double quotient = (double)r * d;
    int i = (int)quotient;
    return (int)(r - (long)i * (long)m);
}
int orc_mod_mul(int x, int y, int m, double d) {
    long r = (long)x * (long)y;
    double quotient = (double)r * d;
    int i = (int)quotient;
    return (int)(r - (long)i * (long)m);
}
int orc_mod_mul3(int x, int y, int m) { return (int)(((long)x * (long)y) % (long)m); }
int orc_mod_psub(int x, int y, int m) { return (int)(((long)x - (long)y + (long)m) % (long)m); }

/* src/modular.cuh:136-143, 169-191 */
static int orc_check_zero(const orc_ctx *c, const int *x) {
    for (int i = 0; i < c->n; i++) if (x[i] != 0) return 0;
    return 1;
}
static void orc_rns_mul(const orc_ctx *c, int *r, const int *x, const int *y) {
    for (int i = 0; i < c->n; i++) r[i] = orc_mod_mul(x[i], y[i], c->moduli[i], 1.0 / c->moduli[i]);
}
static void orc_rns_add(const orc_ctx *c, int *r, const int *x, const int *y) {
    for (int i = 0; i < c->n; i++) r[i] = orc_mod_add(x[i], y[i], c->moduli[i], 1.0 / c->moduli[i]);
}
static void orc_rns_sub(const orc_ctx *c, int *r, const int *x, const int *y) {
    for (int i = 0; i < c->n; i++) r[i] = orc_mod_psub(x[i], y[i], c->moduli[i]);
}

/* ------------------------------------------------------------------ constants
 * src/rnsbase.cuh:121-150 (canonical inverse; any correct method gives the same value) */
static int orc_mod_inverse(long x, long m) {
    long a = ((x % m) + m) % m, b = m, x0 = 1, x1 = 0;
    while (b != 0) {
        long q = a / b, t = a - q * b;
        a = b; b = t;
        t = x0 - q * x1; x0 = x1; x1 = t;
    }
    return (int)(((x0 % m) + m) % m);
}

/* floor(M / 2^e) for the limb product, 53 bits at most. */
static uint64_t orc_limb_bits(const uint32_t *l, int nl, int pos, int width) {
    unsigned __int128 acc = 0;
    int w0 = pos / 32, sh = pos % 32;
    for (int k = 3; k >= 0; k--) {
        int w = w0 + k;
        acc = (acc << 32) | (unsigned __int128)(w < nl ? l[w] : 0u);
    }
    acc >>= sh;
    return (uint64_t)(acc & (((unsigned __int128)1 << width) - 1));
}

/* mpfr_get_d_2exp convention: value = frac * 2^exp, frac in [0.5,1); num is a 53-bit (or 2^53)
 * significand scaled by 2^-53. */
static orc_er_t orc_half_conv(uint64_t num, long exp) {
    orc_er_t r;
    if (num == ((uint64_t)1 << 53)) { r.frac = 0.5; r.exp = exp + 1; }
    else { r.frac = ldexp((double)num, -53); r.exp = exp; }
    return r;
}

orc_ctx *orc_ctx_create(const int *moduli, int n) {
    if (n < 1 || n > ORC_MAX_MODULI) return NULL;
    for (int i = 0; i < n; i++) if (moduli[i] < 2) return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->n = n;
    memcpy(c->moduli, moduli, sizeof(int) * n);

    /* M = prod m_i (src/rnsbase.cuh:333-335) */
    uint32_t limbs[ORC_MAX_MODULI + 2];
    int nl = 1; limbs[0] = 1;
    for (int i = 0; i < n; i++) {
        uint64_t carry = 0;
        for (int k = 0; k < nl; k++) {
            uint64_t t = (uint64_t)limbs[k] * (uint32_t)moduli[i] + carry;
            limbs[k] = (uint32_t)t; carry = t >> 32;
        }
        if (carry) limbs[nl++] = (uint32_t)carry;
    }
    int bitlen = (nl - 1) * 32 + (32 - __builtin_clz(limbs[nl - 1]));
    c->log2M = bitlen - 1;

    /* w_i = (M/m_i)^-1 mod m_i (src/rnsbase.cuh:338-346) */
    for (int i = 0; i < n; i++) {
        long p = 1 % moduli[i];
        for (int j = 0; j < n; j++) if (j != i) p = (p * (moduli[j] % moduli[i])) % moduli[i];
        c->part_inv[i] = orc_mod_inverse(p, moduli[i]);
        c->one[i] = 1;
    }
    /* RNS_POW2 (src/rnsbase.cuh:356-362) */
    c->pow2 = (int *)malloc(sizeof(int) * (size_t)(c->log2M + 1) * n);
    /* row i = 2^i mod m_j by repeated doubling; row 0 is 1 as in the reference */
    for (int j = 0; j < n; j++) {
        int v = 1;
        for (int i = 0; i <= c->log2M; i++) {
            c->pow2[(size_t)i * n + j] = v;
            v = orc_mod_mul3(v, 2, moduli[j]);
        }
    }
    /* scaling tables (src/rnsbase.cuh:366-386) */
    for (int t = 0; t < ORC_SCALING_THRESHOLD; t++) {
        uint64_t mask = ((uint64_t)1 << (t + 1)) - 1;
        uint64_t pm = 1;
        for (int i = 0; i < n; i++) pm = (pm * (uint64_t)moduli[i]) & mask;
        c->prod_pow2_res[t] = (int)pm;
        for (int j = 0; j < n; j++) {
            uint64_t pp = 1;
            for (int i = 0; i < n; i++) if (i != j) pp = (pp * (uint64_t)moduli[i]) & mask;
            c->part_pow2_res[t][j] = (int)pp;
            int pw = 1;
            for (int k = 0; k <= t; k++) pw = orc_mod_mul3(pw, 2, moduli[j]);
            c->pow2_inv[t][j] = orc_mod_inverse(pw, moduli[j]);
        }
    }
    /* accuracy constants (src/rnsbase.cuh:389-391) */
    c->accuracy = 4 * pow(2.0, 1 - 53) * n * log2((double)n) * (1 + 0.0000001 / 2) / 0.0000001;
    c->ref_factor = (int)floor(log2(1 / (2 * c->accuracy)));

    /* RNS_EVAL_UNIT / INV_UNIT (src/rnsbase.cuh:393-409): M is first rounded down to 53 bits
     * (mpfr_set_z with default precision, RNDD, :326,336); the quotients are then rounded up /
     * down to 53 bits in the mpfr_get_d_2exp convention.  The 10000-bit intermediates of the
     * reference never change a 53-bit directed rounding of these values. */
    {
        int e = bitlen > 53 ? bitlen - 53 : 0;
        uint64_t a = orc_limb_bits(limbs, nl, e, 53);
        int b = 64 - __builtin_clzll(a);
        unsigned __int128 num = (unsigned __int128)1 << (b - 1 + 53);
        uint64_t q = (uint64_t)(num / a), rem = (uint64_t)(num % a);
        long ex = 1 - b - e;
        c->unit.low = orc_half_conv(q, ex);
        c->unit.upp = orc_half_conv(q + (rem != 0), ex);
        uint64_t two53 = (uint64_t)1 << 53;
        if (bitlen <= 53) {
            /* Md = M exactly: (M-1)/M */
            unsigned __int128 nn = (unsigned __int128)(a - 1) << 53;
            uint64_t qd = (uint64_t)(nn / a), rd = (uint64_t)(nn % a);
            c->inv_unit.low = orc_half_conv(qd, 0);
            c->inv_unit.upp = orc_half_conv(qd + (rd != 0), 0);
        } else {
            int exact = (e == 1 && a == ((uint64_t)1 << 52));
            c->inv_unit.low = orc_half_conv(two53 - 1, 0);
            c->inv_unit.upp = orc_half_conv(exact ? two53 - 1 : two53, 0);
        }
    }
    /* BoundNear1 reciprocals (src/rnsbase.cuh:413-416) */
    for (int i = 0; i < n; i++) {
        c->recip_rd[i] = orc_ddiv_rd(1.0, moduli[i]);
        c->recip_ru[i] = orc_ddiv_ru(1.0, moduli[i]);
    }
    /* MRC_MULT_INV (src/rnsbase.cuh:419-428) */
    c->mrc_inv = (int *)calloc((size_t)n * n, sizeof(int));
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            c->mrc_inv[i * n + j] = orc_mod_inverse(moduli[i], moduli[j]);
    return c;
}

void orc_ctx_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->pow2);
    free(c->mrc_inv);
    free(c);
}

/* ------------------------------------------------------------------ MRC
 * src/mrc.cuh:13-22, 31-44, 66-72 */
void orc_mrc(const orc_ctx *c, int *mr, const int *x) {
    int n = c->n;
    for (int i = 0; i < n; i++) {
        mr[i] = x[i];
        for (int j = 0; j < i; j++) {
            if (mr[i] < mr[j]) {
                long tmp = (long)c->moduli[i] - (long)mr[j] + (long)mr[i];
                mr[i] = (int)tmp;
            } else {
                mr[i] = (int)((long)mr[i] - (long)mr[j]);
            }
            mr[i] = orc_mod_mul3(mr[i], c->mrc_inv[j * n + i], c->moduli[i]);
        }
    }
}
static int orc_mrs_cmp(int n, const int *x, const int *y) {
    for (int i = n - 1; i >= 0; i--) {
        if (x[i] > y[i]) return 1;
        else if (y[i] > x[i]) return -1;
    }
    return 0;
}
int orc_mrc_compare(const orc_ctx *c, const int *x, const int *y) {
    int mx[ORC_MAX_MODULI], my[ORC_MAX_MODULI];
    orc_mrc(c, mx, x);
    orc_mrc(c, my, y);
    return orc_mrs_cmp(c->n, mx, my);
}

/* ------------------------------------------------------------------ evaluation
 * src/rnseval.cuh:29-102.  Cases the reference would not survive (k outside RNS_POW2, a
 * refinement loop that does not end) stop with the NaN interval. */
#define ORC_REFINE_CAP 4096
static void orc_set_nan(orc_interval_t *out) {
    out->low.frac = NAN; out->low.exp = 0; out->upp.frac = NAN; out->upp.exp = 0;
}
static int orc_refine_k(const orc_ctx *c, double sumu) {
#if ORC_MUTANT == 4
    double a = -ceil(log2(sumu));
#else
    double a = -(ceil(log2(sumu)) + 1);
#endif
    double v = a > c->ref_factor ? a : (double)c->ref_factor;
    if (!(v >= 0 && v <= c->log2M)) return -1;
    return (int)v;
}
/* shared tail: refinement loop src/rnseval.cuh:79-101 (identical in _fast, :144-166) */
static void orc_eval_refine(const orc_ctx *c, orc_interval_t *out, int *s, double sumu) {
    int n = c->n;
    double fracl[ORC_MAX_MODULI], fracu[ORC_MAX_MODULI];
    int K = 0, it = 0;
#if ORC_MUTANT == 13
    int s0[ORC_MAX_MODULI];
    memcpy(s0, s, sizeof(int) * (size_t)n);
#endif
    while (sumu < c->accuracy) {
        int k = orc_refine_k(c, sumu);
        if (k < 0 || ++it > ORC_REFINE_CAP) { orc_set_nan(out); return; }
        orc_rns_mul(c, s, s, c->pow2 + (size_t)k * n);
#if ORC_MUTANT == 3
        for (int i = 0; i < n; i++) fracu[i] = orc_dmul_ru(s[i], c->recip_ru[i]);
#else
        for (int i = 0; i < n; i++) fracu[i] = orc_ddiv_ru((double)s[i], (double)c->moduli[i]);
#endif
        sumu = orc_psum_ru(fracu, n);
        sumu -= orc_to_uint(sumu);
        K += k;
    }
#if ORC_MUTANT == 3
    for (int i = 0; i < n; i++) fracl[i] = orc_dmul_rd(s[i], c->recip_rd[i]);
#elif ORC_MUTANT == 13
    for (int i = 0; i < n; i++) fracl[i] = orc_dmul_rd(s0[i], c->recip_rd[i]);
#else
    for (int i = 0; i < n; i++) fracl[i] = orc_ddiv_rd((double)s[i], (double)c->moduli[i]);
#endif
    double suml = orc_psum_rd(fracl, n);
    suml -= orc_to_uint(suml);
    orc_er_set_d(&out->low, suml);
    orc_er_set_d(&out->upp, sumu);
#if ORC_MUTANT != 12
    out->low.exp -= K;
#endif
    out->upp.exp -= K;
}
void orc_eval(const orc_ctx *c, orc_interval_t *out, const int *x) {
    int n = c->n;
    int s[ORC_MAX_MODULI], mrd[ORC_MAX_MODULI];
    double fracl[ORC_MAX_MODULI], fracu[ORC_MAX_MODULI];
    int mr = -1;
#if ORC_MUTANT != 10
    if (orc_check_zero(c, x)) {
        out->low.frac = 0; out->low.exp = 0; out->upp.frac = 0; out->upp.exp = 0;
        return;
    }
#endif
    orc_rns_mul(c, s, x, c->part_inv);
    for (int i = 0; i < n; i++) {
#if ORC_MUTANT == 2
        fracl[i] = orc_dmul_rd(s[i], c->recip_ru[i]);
        fracu[i] = orc_dmul_ru(s[i], c->recip_rd[i]);
#else
        fracl[i] = orc_dmul_rd(s[i], c->recip_rd[i]);
        fracu[i] = orc_dmul_ru(s[i], c->recip_ru[i]);
#endif
    }
    double suml = orc_psum_rd(fracl, n);
    double sumu = orc_psum_ru(fracu, n);
    unsigned int whl = orc_to_uint(suml);
    unsigned int whu = orc_to_uint(sumu);
#if ORC_MUTANT == 5
    suml = orc_dsub_rd(suml, whl);
    sumu = orc_dsub_ru(sumu, whu);
#else
    suml = suml - whl;
    sumu = sumu - whu;
#endif
    orc_er_set_d(&out->low, suml);
    orc_er_set_d(&out->upp, sumu);
    if (whl != whu) {
        orc_mrc(c, mrd, x);
        mr = mrd[n - 1];
    }
#if ORC_MUTANT == 11
    if (mr > 0) { out->upp = c->inv_unit.low; return; }
#else
    if (mr > 0) { out->upp = c->inv_unit.upp; return; }
#endif
    if (mr == 0) out->low = c->unit.low;
    if (sumu >= c->accuracy) return;
    orc_eval_refine(c, out, s, sumu);
}
/* src/rnseval.cuh:112-167 */
void orc_eval_fast(const orc_ctx *c, orc_interval_t *out, const int *x) {
    int n = c->n;
    int s[ORC_MAX_MODULI];
    double fracl[ORC_MAX_MODULI], fracu[ORC_MAX_MODULI];
    if (orc_check_zero(c, x)) {
        out->low.frac = 0; out->low.exp = 0; out->upp.frac = 0; out->upp.exp = 0;
        return;
    }
    orc_rns_mul(c, s, x, c->part_inv);
    for (int i = 0; i < n; i++) {
        fracl[i] = orc_dmul_rd(s[i], c->recip_rd[i]);
        fracu[i] = orc_dmul_ru(s[i], c->recip_ru[i]);
    }
    double suml = orc_psum_rd(fracl, n);
    double sumu = orc_psum_ru(fracu, n);
    suml -= orc_to_uint(suml);
    sumu -= orc_to_uint(sumu);
    if (sumu >= c->accuracy) {
        orc_er_set_d(&out->low, suml);
        orc_er_set_d(&out->upp, sumu);
        return;
    }
    orc_eval_refine(c, out, s, sumu);
}

/* ------------------------------------------------------------------ comparison
 * src/rnscmp.cuh:20-39 */
int orc_cmp(const orc_ctx *c, const int *x, const int *y) {
    orc_interval_t ex, ey;
    orc_eval(c, &ex, x);
    orc_eval(c, &ey, y);
    if (orc_er_ucmp(&ex.low, &ey.upp) > 0) return 1;
    if (orc_er_ucmp(&ey.low, &ex.upp) > 0) return -1;
    for (int i = 0; i < c->n; i++)
        if (x[i] != y[i]) return orc_mrc_compare(c, x, y);
    return 0;
}

/* ------------------------------------------------------------------ division
 * src/rnsdiv.cuh:19-61 and :67-114.  A main loop that would not end in the reference (a
 * partial quotient of zero) or an RNS_POW2 index outside the table stops with -1. */
#define ORC_DIV_CAP 20000
static void orc_div_final(const orc_ctx *c, int *q, int *r, const int *d,
                          const orc_interval_t *reval, const orc_interval_t *deval) {
    int n = c->n;
    if (orc_er_ucmp(&reval->upp, &deval->low) >= 0) {
        int dmr[ORC_MAX_MODULI], pq[ORC_MAX_MODULI];
        int rged = 1;
        orc_mrc(c, dmr, d);
        orc_mrc(c, pq, r);
        for (int i = n - 1; i >= 0; i--) {
            if (pq[i] < dmr[i]) { rged = 0; break; }
            else if (pq[i] > dmr[i]) break;
        }
        if (rged) {
            orc_rns_add(c, q, q, c->one);
            orc_rns_sub(c, r, r, d);
        }
    }
}
int orc_div(const orc_ctx *c, int *q, int *r, const int *x, const int *d) {
    int n = c->n, it = 0;
    int pq[ORC_MAX_MODULI];
    orc_er_t flq;
    orc_interval_t deval, reval;
    memset(q, 0, sizeof(int) * n);
    memcpy(r, x, sizeof(int) * n);
    orc_eval(c, &deval, d);
    orc_eval(c, &reval, r);
    while (orc_er_ucmp(&reval.low, &deval.upp) >= 0) {
        orc_er_div_rd(&flq, &reval.low, &deval.upp);
        if (flq.exp < 0 || flq.exp > c->log2M || ++it > ORC_DIV_CAP) return -1;
        memcpy(pq, c->pow2 + (size_t)flq.exp * n, sizeof(int) * n);
        orc_rns_add(c, q, q, pq);
        orc_rns_mul(c, pq, pq, d);
        orc_rns_sub(c, r, r, pq);
        orc_eval(c, &reval, r);
    }
    orc_div_final(c, q, r, d, &reval, &deval);
    return it;
}
int orc_div_fast(const orc_ctx *c, int *q, int *r, const int *x, const int *d) {
    int n = c->n, it = 0;
    int pq[ORC_MAX_MODULI];
    orc_er_t flq;
    orc_interval_t deval, reval;
    memset(q, 0, sizeof(int) * n);
    memcpy(r, x, sizeof(int) * n);
    orc_eval(c, &deval, d);
    orc_eval(c, &reval, r);
    while (orc_er_ucmp(&reval.low, &deval.upp) >= 0) {
        if (++it > ORC_DIV_CAP) return -1;
        orc_er_div_rd(&flq, &reval.low, &deval.upp);
        memset(pq, 0, sizeof(int) * n);
        int got = 0;
        while (flq.exp >= 0 && flq.frac > 0) {
            if (flq.exp > c->log2M) return -1;
            orc_rns_add(c, pq, pq, c->pow2 + (size_t)flq.exp * n);
            flq.frac = orc_dsub_rd(flq.frac, 1.0);
            orc_er_adjust(&flq);
            got = 1;
        }
        if (!got) return -1; /* pq == 0: the reference loops forever here */
        orc_rns_add(c, q, q, pq);
        orc_rns_mul(c, pq, pq, d);
        orc_rns_sub(c, r, r, pq);
        orc_eval(c, &reval, r);
    }
    orc_div_final(c, q, r, d, &reval, &deval);
    return it;
}

/* ------------------------------------------------------------------ scaling by 2^D
 * src/rnsscal.cuh:17-62 */
static int orc_rank(const orc_ctx *c, const int *x, const int *s) {
    int n = c->n;
    double fracl[ORC_MAX_MODULI], fracu[ORC_MAX_MODULI];
    int mrd[ORC_MAX_MODULI];
    for (int i = 0; i < n; i++) {
        fracl[i] = orc_dmul_rd(s[i], c->recip_rd[i]);
        fracu[i] = orc_dmul_ru(s[i], c->recip_ru[i]);
    }
    double suml = orc_psum_rd(fracl, n);
    double sumu = orc_psum_ru(fracu, n);
    unsigned int whl = orc_to_uint(suml);
    unsigned int whu = orc_to_uint(sumu);
    if (whl == whu) return (int)whl;
    orc_mrc(c, mrd, x);
    return mrd[n - 1] == 0 ? (int)whu : (int)whl;
}
static int orc_rank_fast(const orc_ctx *c, const int *s) {
    int n = c->n;
    double fracu[ORC_MAX_MODULI];
    for (int i = 0; i < n; i++) fracu[i] = orc_dmul_ru(s[i], c->recip_ru[i]);
    return (int)orc_psum_ru(fracu, n);
}
/* src/rnsscal.cuh:67-85 */
static void orc_scaling_step(const orc_ctx *c, int *y, int k, unsigned int j, int pow2j,
                             const int *x, const int *cc) {
    int n = c->n;
    long residue = 0;
    int multiple[ORC_MAX_MODULI];
    for (int i = 0; i < n; i++) residue += (long)orc_mod_mul3(c->part_pow2_res[j - 1][i], cc[i], pow2j);
    long temp = (long)k * (long)c->prod_pow2_res[j - 1];
    residue = (residue - temp) % pow2j;
    if (residue < 0) residue += pow2j;
    for (int i = 0; i < n; i++) multiple[i] = (int)(residue % c->moduli[i]);
    orc_rns_sub(c, multiple, x, multiple);
    orc_rns_mul(c, y, multiple, c->pow2_inv[j - 1]);
}
/* src/rnsscal.cuh:93-119 */
void orc_scale2pow(const orc_ctx *c, int *result, const int *x, unsigned int D) {
    int n = c->n;
    int cc[ORC_MAX_MODULI], tmp[ORC_MAX_MODULI];
    memcpy(result, x, sizeof(int) * n);
    int t = (int)(D / ORC_SCALING_THRESHOLD);
    int k;
    if (t > 0) {
        orc_rns_mul(c, cc, x, c->part_inv);
        k = orc_rank(c, x, cc);
        orc_scaling_step(c, result, k, ORC_SCALING_THRESHOLD, 1 << ORC_SCALING_THRESHOLD, x, cc);
        t -= 1;
    }
    while (t > 0) {
        orc_rns_mul(c, cc, result, c->part_inv);
        k = orc_rank_fast(c, cc);
        memcpy(tmp, result, sizeof(int) * n);
        orc_scaling_step(c, result, k, ORC_SCALING_THRESHOLD, 1 << ORC_SCALING_THRESHOLD, tmp, cc);
        t -= 1;
    }
    unsigned int d = D % ORC_SCALING_THRESHOLD;
    if (d > 0) {
        orc_rns_mul(c, cc, result, c->part_inv);
        k = d < D ? orc_rank_fast(c, cc) : orc_rank(c, result, cc);
        memcpy(tmp, result, sizeof(int) * n);
        orc_scaling_step(c, result, k, d, 1 << d, tmp, cc);
    }
}

/* ------------------------------------------------------------------ mpint
 * src/mpint.cuh:142-152 */
int orc_check_overflow(const orc_ctx *c, const orc_er_t *lower, const orc_er_t *upper) {
    if (upper->exp < -1 || orc_er_ucmp(upper, &c->inv_unit.low) <= 0) return 0;
    if (orc_er_ucmp(lower, &c->inv_unit.upp) == 1) return 1;
    return 2;
}
/* src/mpint.cuh:160-262 (add with sy = y.sign, sub with sy = y.sign ^ 1) */
static int orc_mpint_addsub(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                            const int *xd, int sx, const orc_er_t *xe,
                            const int *yd, int sy, const orc_er_t *ye) {
    int n = c->n;
    orc_er_t evx[2] = { xe[0], xe[1] };
    orc_er_t evy[2] = { ye[0], ye[1] };
    int alpha = (1 - 2 * sx);
    int beta = (1 - 2 * sy);
    for (int i = 0; i < n; i++) {
        int residue = orc_mod_add(alpha * xd[i], beta * yd[i], c->moduli[i], 1.0 / c->moduli[i]);
        rd[i] = residue < 0 ? residue + c->moduli[i] : residue;
    }
    evx[0].frac *= alpha;
    evx[1].frac *= alpha;
    evy[0].frac *= beta;
    evy[1].frac *= beta;
    orc_er_t r0, r1;
    orc_er_add_rd(&r0, &evx[sx], &evy[sy]);
    orc_er_add_ru(&r1, &evx[1 - sx], &evy[1 - sy]);
    int minus = r0.frac < 0 && r1.frac < 0;
    *rs = minus;
    if (minus) {
        for (int i = 0; i < n; i++) rd[i] = rd[i] == 0 ? 0 : (c->moduli[i] - rd[i]);
        orc_er_t tmp = r0;
        r0.frac = -1 * r1.frac;
        r0.exp = r1.exp;
        r1.frac = -1 * tmp.frac;
        r1.exp = tmp.exp;
    }
    re[0] = r0;
    re[1] = r1;
    return orc_check_overflow(c, &re[0], &re[1]);
}
int orc_mpint_add(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                  const int *xd, int xs, const orc_er_t *xe, const int *yd, int ys, const orc_er_t *ye) {
    return orc_mpint_addsub(c, rd, rs, re, xd, xs, xe, yd, ys, ye);
}
int orc_mpint_sub(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                  const int *xd, int xs, const orc_er_t *xe, const int *yd, int ys, const orc_er_t *ye) {
    return orc_mpint_addsub(c, rd, rs, re, xd, xs, xe, yd, ys ^ 1, ye);
}
/* src/mpint.cuh:270-278 */
int orc_mpint_mul(const orc_ctx *c, int *rd, int *rs, orc_er_t *re,
                  const int *xd, int xs, const orc_er_t *xe, const int *yd, int ys, const orc_er_t *ye) {
    int n = c->n;
    orc_er_t x0 = xe[0], x1 = xe[1], y0 = ye[0], y1 = ye[1];
    orc_er_t uu = c->unit.upp, ul = c->unit.low;
#if ORC_MUTANT == 9
    orc_er_md_rd(&re[0], &x0, &y0, &ul);
    orc_er_md_ru(&re[1], &x1, &y1, &uu);
#else
    orc_er_md_rd(&re[0], &x0, &y0, &uu);
    orc_er_md_ru(&re[1], &x1, &y1, &ul);
#endif
    for (int i = 0; i < n; i++) rd[i] = orc_mod_mul(xd[i], yd[i], c->moduli[i], 1.0 / c->moduli[i]);
    *rs = orc_check_zero(c, rd) ? 0 : xs ^ ys;
    return orc_check_overflow(c, &re[0], &re[1]);
}

/* ------------------------------------------------------------------ array forms */
static void orc_gather(const orc_ctx *c, int *dst, const int *src, size_t ld, size_t j) {
    for (int i = 0; i < c->n; i++) dst[i] = src[(size_t)i * ld + j];
}
static void orc_scatter(const orc_ctx *c, int *dst, const int *src, size_t ld, size_t j) {
    for (int i = 0; i < c->n; i++) dst[(size_t)i * ld + j] = src[i];
}
static int orc_threads(int nthreads) {
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

void orc_eval_array(const orc_ctx *c, const int *x, size_t ld, size_t count, orc_interval_t *out,
                    int fast, int nthreads) {
    long long cnt = (long long)count;
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (long long j = 0; j < cnt; j++) {
        int v[ORC_MAX_MODULI];
        orc_gather(c, v, x, ld, (size_t)j);
        if (fast) orc_eval_fast(c, &out[j], v);
        else orc_eval(c, &out[j], v);
    }
}
void orc_cmp_array(const orc_ctx *c, const int *x, const int *y, size_t ld, size_t count, int *sign,
                   int nthreads) {
    long long cnt = (long long)count;
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (long long j = 0; j < cnt; j++) {
        int a[ORC_MAX_MODULI], b[ORC_MAX_MODULI];
        orc_gather(c, a, x, ld, (size_t)j);
        orc_gather(c, b, y, ld, (size_t)j);
        sign[j] = orc_cmp(c, a, b);
    }
}
void orc_div_array(const orc_ctx *c, const int *x, const int *d, size_t ld, size_t count, int *q, int *r,
                   int fast, int nthreads) {
    long long cnt = (long long)count;
#pragma omp parallel for schedule(dynamic, 4) num_threads(orc_threads(nthreads))
    for (long long j = 0; j < cnt; j++) {
        int a[ORC_MAX_MODULI], b[ORC_MAX_MODULI], qq[ORC_MAX_MODULI], rr[ORC_MAX_MODULI];
        orc_gather(c, a, x, ld, (size_t)j);
        orc_gather(c, b, d, ld, (size_t)j);
        if (fast) orc_div_fast(c, qq, rr, a, b);
        else orc_div(c, qq, rr, a, b);
        orc_scatter(c, q, qq, ld, (size_t)j);
        orc_scatter(c, r, rr, ld, (size_t)j);
    }
}
void orc_scale2pow_array(const orc_ctx *c, const int *x, const unsigned int *D, size_t ld, size_t count,
                         int *y, int nthreads) {
    long long cnt = (long long)count;
#pragma omp parallel for schedule(dynamic, 16) num_threads(orc_threads(nthreads))
    for (long long j = 0; j < cnt; j++) {
        int a[ORC_MAX_MODULI], rr[ORC_MAX_MODULI];
        orc_gather(c, a, x, ld, (size_t)j);
        orc_scale2pow(c, rr, a, D[j]);
        orc_scatter(c, y, rr, ld, (size_t)j);
    }
}
/* rns_max semantics (src/dpp/rnsmax.cuh:24-47): exact magnitude order.  The reference's
 * comparison is exact whenever its interval evaluations enclose x/M, which holds for the GPU
 * evaluation it uses; here the order is taken from mixed-radix digits, which is exact.
 * Ties resolve to the lowest index. */
long orc_max_array(const orc_ctx *c, const int *x, size_t ld, size_t count, int *out_residues, int nthreads) {
    int n = c->n;
    if (count == 0) return -1;
    int nt = orc_threads(nthreads);
    long *best_idx = (long *)malloc(sizeof(long) * nt);
    int *best_mr = (int *)malloc(sizeof(int) * (size_t)nt * n);
    for (int t = 0; t < nt; t++) best_idx[t] = -1;
    long long cnt = (long long)count;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int v[ORC_MAX_MODULI], mr[ORC_MAX_MODULI];
        int *bm = best_mr + (size_t)tid * n;
#pragma omp for schedule(static)
        for (long long j = 0; j < cnt; j++) {
            orc_gather(c, v, x, ld, (size_t)j);
            orc_mrc(c, mr, v);
            if (best_idx[tid] < 0 || orc_mrs_cmp(n, mr, bm) > 0) {
                best_idx[tid] = (long)j;
                memcpy(bm, mr, sizeof(int) * n);
            }
        }
    }
    long bi = -1;
    int bt = -1;
    for (int t = 0; t < nt; t++) {
        if (best_idx[t] < 0) continue;
        if (bi < 0) { bi = best_idx[t]; bt = t; continue; }
        int r = orc_mrs_cmp(n, best_mr + (size_t)t * n, best_mr + (size_t)bt * n);
        if (r > 0 || (r == 0 && best_idx[t] < bi)) { bi = best_idx[t]; bt = t; }
    }
    if (out_residues && bi >= 0) orc_gather(c, out_residues, x, ld, (size_t)bi);
    free(best_idx);
    free(best_mr);
    return bi;
}
/* The reference's rns_max pipeline run on the CPU path (src/dpp/rnsmax.cuh:24-101 with
 * rns_eval_compute of src/rnseval.cuh:29-102): evaluate every element, then reduce with
 * rns_max_cmp (interval test, residue equality, MRC); strict ">" keeps the incumbent.
 * Used as the timed CPU baseline of the arg-max workload. */
long orc_max_eval_array(const orc_ctx *c, const int *x, size_t ld, size_t count, int nthreads) {
    int n = c->n;
    if (count == 0) return -1;
    int nt = orc_threads(nthreads);
    long *best = (long *)malloc(sizeof(long) * nt);
    orc_interval_t *bev = (orc_interval_t *)malloc(sizeof(orc_interval_t) * nt);
    long long cnt = (long long)count;
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int v[ORC_MAX_MODULI], w[ORC_MAX_MODULI];
        orc_interval_t ev;
        best[tid] = -1;
#pragma omp for schedule(static)
        for (long long j = 0; j < cnt; j++) {
            orc_gather(c, v, x, ld, (size_t)j);
            orc_eval(c, &ev, v);
            int r = 1;
            if (best[tid] >= 0) {
                if (orc_er_ucmp(&ev.low, &bev[tid].upp) > 0) r = 1;
                else if (orc_er_ucmp(&bev[tid].low, &ev.upp) > 0) r = -1;
                else {
                    orc_gather(c, w, x, ld, (size_t)best[tid]);
                    r = memcmp(v, w, sizeof(int) * n) == 0 ? 0 : orc_mrc_compare(c, v, w);
                }
            }
            if (r == 1) { best[tid] = (long)j; bev[tid] = ev; }
        }
    }
    long bi = -1;
    int bt = -1;
    for (int t = 0; t < nt; t++) {
        if (best[t] < 0) continue;
        if (bi < 0) { bi = best[t]; bt = t; continue; }
        int v[ORC_MAX_MODULI], w[ORC_MAX_MODULI];
        orc_gather(c, v, x, ld, (size_t)best[t]);
        orc_gather(c, w, x, ld, (size_t)bi);
        int r = orc_mrc_compare(c, v, w);
        if (r > 0) { bi = best[t]; bt = t; }
    }
    (void)bt;
    free(best);
    free(bev);
    return bi;
}

/* mpint_div, src/mpint.cuh:286-296 (CPU path): rns_div_fast of the magnitudes, the Euclidean +1
 * for a negative dividend with a nonzero remainder (int % on each digit), rns_eval_compute of the
 * digits, sign 0 for a zero quotient.  Returns -1 where the division loop would not end. */
int orc_mpint_div(const orc_ctx *c, int *rd, int *rs, orc_er_t *re, const int *xd, int xs, const int *yd, int ys) {
    int n = c->n;
    int rem[ORC_MAX_MODULI];
    if (orc_div_fast(c, rd, rem, xd, yd) < 0) return -1;
    int rz = 1;
    for (int i = 0; i < n; i++) rz &= rem[i] == 0;
    if (xs == 1 && !rz)
        for (int i = 0; i < n; i++) rd[i] = (rd[i] + 1) % c->moduli[i];
    orc_interval_t ev;
    orc_eval(c, &ev, rd);
    re[0] = ev.low;
    re[1] = ev.upp;
    int qz = 1;
    for (int i = 0; i < n; i++) qz &= rd[i] == 0;
    *rs = qz ? 0 : (xs ^ ys);
    return 0;
}

void orc_mpint_div_array(const orc_ctx *c, size_t ld, size_t count, const int *xd, const int *xs, const int *yd,
                         const int *ys, int *rd, int *rs, orc_er_t *re, int nthreads) {
    long long cnt = (long long)count;
#pragma omp parallel for schedule(dynamic, 64) num_threads(orc_threads(nthreads))
    for (long long j = 0; j < cnt; j++) {
        int a[ORC_MAX_MODULI], b[ORC_MAX_MODULI], r[ORC_MAX_MODULI];
        orc_gather(c, a, xd, ld, (size_t)j);
        orc_gather(c, b, yd, ld, (size_t)j);
        orc_mpint_div(c, r, &rs[j], &re[2 * j], a, xs[j], b, ys[j]);
        orc_scatter(c, rd, r, ld, (size_t)j);
    }
}

void orc_mpint_array(const orc_ctx *c, int op, size_t ld, size_t count,
                     const int *xd, const int *xs, const orc_er_t *xe,
                     const int *yd, const int *ys, const orc_er_t *ye,
                     int *rd, int *rs, orc_er_t *re, unsigned char *ovf, int nthreads) {
    long long cnt = (long long)count;
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (long long j = 0; j < cnt; j++) {
        int a[ORC_MAX_MODULI], b[ORC_MAX_MODULI], r[ORC_MAX_MODULI];
        orc_gather(c, a, xd, ld, (size_t)j);
        orc_gather(c, b, yd, ld, (size_t)j);
        int f;
        if (op == 0) f = orc_mpint_add(c, r, &rs[j], &re[2 * j], a, xs[j], &xe[2 * j], b, ys[j], &ye[2 * j]);
        else if (op == 1) f = orc_mpint_sub(c, r, &rs[j], &re[2 * j], a, xs[j], &xe[2 * j], b, ys[j], &ye[2 * j]);
        else f = orc_mpint_mul(c, r, &rs[j], &re[2 * j], a, xs[j], &xe[2 * j], b, ys[j], &ye[2 * j]);
        if (ovf) ovf[j] = (unsigned char)f;
        orc_scatter(c, rd, r, ld, (size_t)j);
    }
}
