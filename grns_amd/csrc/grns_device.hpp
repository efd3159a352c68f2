// grns_device.hpp -- device (and host-testable) primitives of the GRNS hot path for gfx950.
//
// Every floating-point routine here reproduces the operation sequence of the GRNS 1.1.4 CPU
// path bit for bit: round-to-nearest FP64 with the BoundNear1 emulation of directed rounding
// (src/dinterval.cuh:37-147), the reciprocal modular reduction of src/modular.cuh:29-79, the
// zero-padded pairwise trees of src/pairwise.cuh:121-458 and x86 conversion semantics.  This
// translation unit is compiled with -ffp-contract=off; the only fused multiply-adds are the
// explicit fma() calls below, each of which is exact or provably returns the same double as the
// unfused reference expression (comments at each site).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GRNS_HD __host__ __device__ __forceinline__
#define GRNS_D __device__ __forceinline__

namespace grns {
// Hides the value of a wave-uniform 64-bit integer from the optimiser at this point: uses after it
// cannot be hoisted above it, and the value is kept in scalar registers.  The value must be the
// same on every lane of the wave (lane 0's is taken, which also moves a value the compiler keeps
// in vector registers into scalar ones).
GRNS_HD long long uniform_point(long long v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    v = (long long)(((unsigned long long)hi << 32) | lo);
    asm volatile("" : "+s"(v));
#endif
    return v;
}

// The same for a wave-uniform pointer, so that constant table loads placed after it are issued
// there (scalar loads from the scalar cache) instead of being hoisted out of the streaming loops,
// where dozens of per-modulus constants would exceed the scalar register file and spill into
// vector registers.
template <class T>
GRNS_HD const T *reload_point(const T *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return reinterpret_cast<const T *>(uniform_point(reinterpret_cast<long long>(p)));
#else
    return p;
#endif
}

// reload_point placed after the computation of `dep` (a per-lane value): the loads through the
// returned pointer cannot be issued before `dep` is known, which bounds how far ahead of their use
// the compiler issues constant loads in long unrolled sequences.
template <class T>
GRNS_HD const T *reload_after(const T *p, int dep) {
#if defined(__HIP_DEVICE_COMPILE__)
    long long v = reinterpret_cast<long long>(p);
    asm volatile("" : "+s"(v) : "v"(dep));
    return reinterpret_cast<const T *>(v);
#else
    (void)dep;
    return p;
#endif
}

// A per-lane 32-bit value the optimiser may not widen or fold: a byte offset that stays a 32-bit
// vector offset of scalar-base global loads.
GRNS_HD unsigned lane_offset(unsigned v) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(v));
#endif
    return v;
}

// A view of a constant table (wave-uniform address) whose loads are scalar loads issued where the
// view is taken: the address goes through uniform_point, and the constant address space lets the
// compiler use the scalar cache.  On the host it is the plain pointer.  Pass an address computed
// from reload_point(ctx) so that not even the address is hoisted out of the caller's loop.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) double cdouble;
typedef const __attribute__((address_space(4))) int cint;
#else
typedef const double cdouble;
typedef const int cint;
#endif
GRNS_HD cdouble *const_view(const double *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return reinterpret_cast<cdouble *>(uniform_point(reinterpret_cast<long long>(p)));
#else
    return p;
#endif
}
GRNS_HD cint *const_view(const int *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return reinterpret_cast<cint *>(uniform_point(reinterpret_cast<long long>(p)));
#else
    return p;
#endif
}
}  // namespace grns

namespace grns {

// er_float_t / interval_t of src/types.cuh:13-24 (long is 64-bit on the LP64 targets).
struct er_t {
    double frac;
    long long exp;
};
struct interval_t {
    er_t low;
    er_t upp;
};

constexpr int MAXN = 256;      // largest moduli set a context accepts
constexpr int WORDS_BATCH = 8;  // table rows whose loads the word conversions issue together

// Pointers into global memory (address space 1): loads through them are global loads (a scalar
// base and a vector offset), not flat loads, which also wait on the LDS counter.
typedef const __attribute__((address_space(1))) int global_int;
typedef const __attribute__((address_space(1))) unsigned global_uint;
constexpr int THRESH = 30;     // RNS_P2_SCALING_THRESHOLD, src/params.h:22
constexpr int KEY_SHIFT = 26;  // arg-max keys of values near 0 or M: the key of 2^KEY_SHIFT x

// Per-moduli-set constants resident in HBM (rns_const_init, src/rnsbase.cuh:324-451, plus the
// tables this build's filtered paths need).
struct DevCtx {
    int n;
    int log2M;
    int ref_factor;
    int big_moduli;    // 1 when every m_i >= 2^22 (enables the near-tie certificate of comparisons)
    double accuracy;
    double tau;        // fraction window outside of which no evaluation needs MRC or refinement
    double err;        // bound on |approx fraction - x/M mod 1| of the filter sums
    double mr_lim;     // 1 / m_{n-1} rounded up: x/M above it makes the top mixed-radix digit positive
    int m_min;         // smallest modulus
    er_t unit_low, unit_upp, inv_low, inv_upp;
    int m[MAXN];
    int w[MAXN];                 // RNS_PART_MODULI_PRODUCT_INVERSE
    double dm[MAXN];             // (double) m_i
    double dw[MAXN];             // (double) w_i
    double dinv[MAXN];           // 1.0 / m_i (round to nearest), the reciprocal of mod_add/mod_mul
    double rd[MAXN];             // RNS_MODULI_RECIP_RD
    double ru[MAXN];             // RNS_MODULI_RECIP_RU
    double chi[MAXN];            // w_i / m_i as a double-double (filter sums)
    double clo[MAXN];
    double khi[MAXN];            // w_i / m_i truncated to 22 significant bits: x_i * khi is exact
    double klo[MAXN];            // w_i / m_i - khi (arg-max keys)
    double shi[MAXN];            // as khi, klo for (2^KEY_SHIFT w_i mod m_i) / m_i: the key of
    double slo[MAXN];            //   2^KEY_SHIFT x, for values within 2^-28 M of 0 or M
    int c32[MAXN];               // 2^32 mod m_i
    unsigned long long mi64[MAXN];  // (M / m_i) mod 2^64
    unsigned long long M64;         // M mod 2^64
    int prod_pow2_res[THRESH];               // RNS_MODULI_PRODUCT_POW2_RESIDUES
    int part_pow2_res[THRESH][MAXN];         // RNS_PART_MODULI_PRODUCT_POW2_RESIDUES
    int pow2_inv[THRESH][MAXN];              // RNS_POW2_INVERSE
    const int *pow2;                         // RNS_POW2, (log2M+1) x n, device memory
    const int *mrc_inv;                      // MRC_MULT_INV, n x n, device memory
    const double *mrc_inv_d;                 // the same as doubles (per-lane mixed-radix conversion)
    int words;                               // 32-bit words of sum_i s_i (M / m_i) < n M
    const unsigned *crt_words;               // (n + 1) x words: M / m_i for i < n, then M
};

// ---------------------------------------------------------------- BoundNear1 (src/dinterval.cuh)
constexpr double EPS = 0x1p-53;
constexpr double ETA = 0x1p-1074;
constexpr double PHI1 = EPS * (1.0 + 2.0 * EPS);   // 2^-53 (1 + 2^-52), exact

// The reference's fmac(PHI1, |c|, ETA) = RN(RN(PHI1*|c|) + ETA).  One fma returns the same
// double whenever c == 0 or |c| >= 2^-917: then ETA is below a quarter of the spacing of the
// grid around PHI1*|c| and PHI1*|c| is never within ETA of a rounding midpoint except the exact
// tie at c = 1.5 * 2^k, which both forms round upward.  For c a small multiple of ETA both give
// ETA.  Every c reaching this function on the hot paths is 0, a multiple of ETA below 2^-1000,
// or has |c| >= 2^-83 (products and sums of residue fractions, extended-range significands).
GRNS_HD double bn_err(double c) { return fma(PHI1, fabs(c), ETA); }

GRNS_HD double dadd_rd(double a, double b) { double c = a + b; return c - bn_err(c); }
GRNS_HD double dadd_ru(double a, double b) { double c = a + b; return c + bn_err(c); }
GRNS_HD double dsub_rd(double a, double b) { double c = a - b; return c - bn_err(c); }
GRNS_HD double dsub_ru(double a, double b) { double c = a - b; return c + bn_err(c); }
GRNS_HD double dmul_rd(double a, double b) { double c = a * b; return c - bn_err(c); }
GRNS_HD double dmul_ru(double a, double b) { double c = a * b; return c + bn_err(c); }
GRNS_HD double ddiv_rd(double a, double b) { double c = a / b; return c - bn_err(c); }
GRNS_HD double ddiv_ru(double a, double b) { double c = a / b; return c + bn_err(c); }

template <bool UP> GRNS_HD double dadd_dir(double a, double b) { return UP ? dadd_ru(a, b) : dadd_rd(a, b); }

// Fractional part x - floor(x) (v_fract_f64 on the device; the same value for finite x >= 0).
GRNS_HD double fract_d(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fract(v);
#else
    return v - floor(v);
#endif
}

GRNS_HD uint64_t dbits(double x) {
    union { double d; uint64_t u; } v; v.d = x; return v.u;
}
GRNS_HD double bitsd(uint64_t u) {
    union { double d; uint64_t u; } v; v.u = u; return v.d;
}

// For a positive c >= 2^-960, c - bn_err(c) is the next double below c and c + bn_err(c) the next
// above: bn_err(c) = RN(PHI1 c + ETA) lies in (ulp(c)/2, ulp(c)] (PHI1 = 2^-53 (1 + 2^-52), ETA far
// below ulp(PHI1 c)), so the rounded difference and sum are c's neighbours.  One 64-bit integer
// step on the bit pattern gives the same doubles as dsub_rd / dadd_ru of the bound.
GRNS_HD double below_pos(double c) { return bitsd(dbits(c) - 1); }
GRNS_HD double above_pos(double c) { return bitsd(dbits(c) + 1); }

// fast_scalbn, src/bitwise.cuh:75-80: add n to the exponent field, then multiply by 0 or 1.
// The multiply by zero of an infinity or NaN produced by the exponent wrap is spelled out with
// the x86 SSE results (default NaN 0xFFF8..., input NaN quieted), so that device and CPU agree.
GRNS_HD double fast_scalbn(double x, int n) {
    uint64_t u = dbits(x) + ((uint64_t)(int64_t)n << 52);
    double v = bitsd(u);
    if (x != 0 && n >= -1023) return v;
    uint64_t e = (u >> 52) & 0x7FF;
    if (e == 0x7FF) {
        if ((u & ((1ull << 52) - 1)) == 0) return bitsd(0xFFF8000000000000ull);  // inf * 0
        return bitsd(u | (1ull << 51));                                            // NaN * 0
    }
    return bitsd(u & (1ull << 63));  // signed zero
}

// The one-word reciprocal of a normalised d (top bit set): v = floor((2^64 - 1) / d) - 2^32, as in
// Moller and Granlund, "Improved division by invariant integers", IEEE Trans. Computers 60(2), 2011.
// The quotient (2^64 - 1) / d lies in (2^32, 2^33], and its correctly rounded double 2^64 / d is
// within 2^-20 of it, so its floor is off by at most one, which one product decides: (2^32 + v) d
// >= 2^64 exactly when v d >= (2^32 - d) 2^32.  (A 64-bit integer division here is a long
// sequence of scalar steps.)  Checked against the definition for every d in [2^31, 2^32).
GRNS_HD unsigned recip_preinv(unsigned d) {
    const double qe = 18446744073709551616.0 / (double)d;
    unsigned long long v = (unsigned long long)qe - (1ull << 32);  // <= 2^32
    const unsigned long long lim = (unsigned long long)(0u - d) << 32;
    const unsigned long long t = v * d;  // < 2^64
    if (t >= lim) v--;                   // (2^32 + v) d > 2^64 - 1
    else if (t + d < lim) v++;           // (2^32 + v + 1) d <= 2^64 - 1
    return (unsigned)v;
}
// Three words <u2, u1, u0> by a normalised two-word <d1, d0> (top bit of d1 set, <u2, u1> <
// <d1, d0>) with v = floor((2^96 - 1) / <d1, d0>) - 2^32: Algorithm 5 of the same paper (one product
// by v, two partial products and corrections, all integer: with wave-uniform operands it runs on
// the scalar unit).  Returns floor(<u2, u1, u0> / <d1, d0>) (< 2^32); <r1, r0> gets the remainder.
// recip32_preinv is the paper's Algorithm 6: v from the one-word reciprocal of d1, adjusted by d0.
GRNS_HD unsigned recip32_preinv(unsigned d1, unsigned d0) {
    unsigned v = recip_preinv(d1);
    unsigned p = d1 * v + d0;  // mod 2^32
    if (p < d0) {
        v--;
        if (p >= d1) {
            v--;
            p -= d1;
        }
        p -= d1;
    }
    const unsigned long long t = (unsigned long long)v * d0;
    const unsigned t1 = (unsigned)(t >> 32), t0 = (unsigned)t;
    p += t1;
    if (p < t1) {
        v--;
        if (p > d1 || (p == d1 && t0 >= d0)) v--;
    }
    return v;
}
GRNS_HD unsigned div32_preinv(unsigned u2, unsigned u1, unsigned u0, unsigned d1, unsigned d0, unsigned v,
                              unsigned &r1, unsigned &r0) {
    const unsigned long long p = (unsigned long long)v * u2 + (((unsigned long long)u2 << 32) | u1);
    unsigned q1 = (unsigned)(p >> 32);
    const unsigned q0 = (unsigned)p;
    const unsigned long long d = ((unsigned long long)d1 << 32) | d0;
    // <u1 - q1 d1, u0> - q1 d0 - d, mod 2^64
    unsigned long long rr = (((unsigned long long)(u1 - q1 * d1) << 32) | u0) - (unsigned long long)d0 * q1 - d;
    q1++;
    if ((unsigned)(rr >> 32) >= q0) {
        q1--;
        rr += d;
    }
    r1 = (unsigned)(rr >> 32);
    r0 = (unsigned)rr;
    if (r1 > d1 || (r1 == d1 && r0 >= d0)) {  // rare
        q1++;
        rr -= d;
        r1 = (unsigned)(rr >> 32);
        r0 = (unsigned)rr;
    }
    return q1;
}

// x86-64 (unsigned int)double (cvttsd2si to 64 bits, low 32 kept), for |x| < 2^31.
GRNS_HD unsigned to_uint(double x) { return (unsigned)(int)x; }

// ---------------------------------------------------------------- extended range (src/extrange.cuh)
GRNS_HD er_t er_set_d(double x) {
    er_t r;
    uint64_t u = dbits(x);
    if (x != 0) {
        r.exp = (long long)(((u & ~(1ull << 63)) >> 52) - 1023ull);
        r.frac = bitsd((u & (((1ull << 52) - 1) | (1ull << 63))) | (1023ull << 52));
    } else {
        r.exp = 0;
        r.frac = 0;
    }
    return r;
}
// src/extrange.cuh:116-125
GRNS_HD er_t er_adjust(er_t x) {
    uint64_t u = dbits(x.frac);
    bool nz = x.frac != 0;
    er_t r;
    r.exp = (long long)(((uint64_t)x.exp + ((u & ~(1ull << 63)) >> 52) - 1023ull) * (uint64_t)nz);
    double f = bitsd((u & (((1ull << 52) - 1) | (1ull << 63))) | (1023ull << 52));
    r.frac = f * (double)(nz ? 1 : 0);
    return r;
}
// src/extrange.cuh:153-248; the reference resets the inputs' exponents when their fraction is 0.
template <bool SUB, bool UP>
GRNS_HD er_t er_addsub(er_t &x, er_t &y) {
    x.exp *= (x.frac != 0);
    y.exp *= (y.frac != 0);
    int dexp = (int)((x.exp - y.exp) * (long long)(x.frac != 0) * (long long)(y.frac != 0));
    er_t r;
    if (dexp > 0) {
        r.exp = x.exp;
        double ys = fast_scalbn(y.frac, -dexp);
        r.frac = SUB ? (UP ? dsub_ru(x.frac, ys) : dsub_rd(x.frac, ys)) : (UP ? dadd_ru(x.frac, ys) : dadd_rd(x.frac, ys));
    } else if (dexp < 0) {
        r.exp = y.exp;
        double xs = fast_scalbn(x.frac, dexp);
        r.frac = SUB ? (UP ? dsub_ru(xs, y.frac) : dsub_rd(xs, y.frac)) : (UP ? dadd_ru(y.frac, xs) : dadd_rd(y.frac, xs));
    } else {
        r.exp = (x.exp == 0) ? y.exp : x.exp;
        r.frac = SUB ? (UP ? dsub_ru(x.frac, y.frac) : dsub_rd(x.frac, y.frac)) : (UP ? dadd_ru(x.frac, y.frac) : dadd_rd(x.frac, y.frac));
    }
    return er_adjust(r);
}
// src/extrange.cuh:289-320
GRNS_HD er_t er_div_rd(const er_t &x, const er_t &y) {
    er_t r; r.exp = x.exp - y.exp; r.frac = ddiv_rd(x.frac, y.frac); return er_adjust(r);
}
GRNS_HD er_t er_md_rd(const er_t &x, const er_t &y, const er_t &z) {
    er_t r; r.exp = x.exp + y.exp - z.exp; r.frac = ddiv_rd(dmul_rd(x.frac, y.frac), z.frac); return er_adjust(r);
}
GRNS_HD er_t er_md_ru(const er_t &x, const er_t &y, const er_t &z) {
    er_t r; r.exp = x.exp + y.exp - z.exp; r.frac = ddiv_ru(dmul_ru(x.frac, y.frac), z.frac); return er_adjust(r);
}
// src/extrange.cuh:347-353
GRNS_HD int er_ucmp(const er_t &x, const er_t &y) {
    int diff = (int)(x.exp - y.exp);
    if (diff == 0) return (x.frac == y.frac ? 0 : x.frac > y.frac ? 1 : -1);
    return x.frac != 0 && (diff > 0 || (diff < 0 && y.frac == 0)) ? 1 : -1;
}

// ---------------------------------------------------------------- modular (src/modular.cuh)
// mod_mul(x, y, m, 1.0/m) of the CPU path: (int)(r - (int)(RN(RN(r) * d)) * m) with r = x*y.
// Computed in FP64: RN(x*y) is one multiply, the quotient is truncated as a double, and the
// difference x*y - i*m is formed exactly by two fmas (|p - i*m| < 2^34).  The int32 wrap of
// the reference's final cast is applied to the exact value.  Requires |quotient| < 2^31, which
// holds for residues within (-2^31, 2^31) and a canonical second factor.
GRNS_HD double mulmod_recip_d(double dx, double dy, double dm, double d) {
    double p = dx * dy;
    double di = trunc(p * d);
    double lo = fma(dx, dy, -p);   // x*y - p, exact
    double t = fma(-di, dm, p);    // p - i*m, exact
    double s = t + lo;             // x*y - i*m, exact, |s| < 2^34
    // int32 wrap: subtract 2^32 * floor(s / 2^32 + 1/2) (every step exact: s has <= 34 bits)
    double k = floor(fma(s, 0x1p-32, 0.5));
    return fma(-k, 4294967296.0, s);
}
// The same before the int32 wrap: callers that treat |s| >= 2^31 (where the wrap would apply)
// as a rare case skip its three operations.
GRNS_HD double mulmod_recip_d_nowrap(double dx, double dy, double dm, double d) {
    double p = dx * dy;
    double di = trunc(p * d);
    double lo = fma(dx, dy, -p);
    double t = fma(-di, dm, p);
    return t + lo;
}

// A per-lane "any of these held" flag kept as a wave mask in scalar registers: one vector compare
// per update, the accumulation on the scalar unit.
struct LaneAny {
    unsigned long long m = 0;
    GRNS_HD void set_if(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
        m |= __ballot(p);
#else
        m |= p ? 1ull : 0ull;
#endif
    }
    GRNS_HD bool any() const {
#if defined(__HIP_DEVICE_COMPILE__)
        return (m >> (threadIdx.x & 63)) & 1ull;
#else
        return m != 0;
#endif
    }
};

GRNS_HD int mulmod_recip(int x, int y, int m, double d) {
    double q = ((double)x * (double)y) * d;
    int i = (int)q;
    return (int)((unsigned)x * (unsigned)y - (unsigned)i * (unsigned)m);
}
// mod_add(x, y, m, 1.0/m), src/modular.cuh:29-34 (x + y is exact in FP64)
GRNS_HD int addmod_recip(int x, int y, int m, double d) {
    double q = ((double)x + (double)y) * d;
    int i = (int)q;
    return (int)((unsigned)x + (unsigned)y - (unsigned)i * (unsigned)m);
}
// C's truncating remainder v % m for |v| < 2^62 (dinv = RN(1/m)).
GRNS_HD long long rem_trunc(long long v, int m, double dinv) {
    double q = trunc((double)v * dinv);
    long long r = v - (long long)q * m;
    if (v >= 0) {
        if (r < 0) r += m;
        else if (r >= m) r -= m;
    } else {
        if (r > 0) r -= m;
        else if (r <= -m) r += m;
    }
    return r;
}
// rem_trunc for a v whose double dv the caller has formed exactly (|v| < 2^32: the quotient fits
// 32 bits, so it converts as an int).  Same value as rem_trunc(v, m, dinv).
GRNS_HD long long rem_trunc_d(long long v, double dv, int m, double dinv) {
    const int q = (int)trunc(dv * dinv);
    const long long r = v - (long long)q * m;
    // the result lies in [lo, lo + m): [0, m) for v >= 0, (-m, 0] for v < 0; r is at most one m out
    const long long lo = v >= 0 ? 0 : 1 - (long long)m;
    return r < lo ? r + m : (r >= lo + m ? r - m : r);
}
// 3-argument mod_mul, src/modular.cuh:62-66
GRNS_HD int mulmod3(int x, int y, int m, double dinv) { return (int)rem_trunc((long long)x * (long long)y, m, dinv); }
// mod_psub, src/modular.cuh:52-56
GRNS_HD int psubmod(int x, int y, int m, double dinv) {
    return (int)rem_trunc((long long)x - (long long)y + (long long)m, m, dinv);
}

// ceil(log2(s)) as glibc computes it for a positive double s (the refinement factor of
// src/rnseval.cuh:82).  log2 is exact at powers of two; just above 2^e (e < 0) the correctly
// rounded result of e + log2(f) is e itself while log2(f) stays below half the spacing of the
// doubles next to |e|, so the ceiling is then e rather than e + 1.
GRNS_HD int ceil_log2_libm(double s) {
    uint64_t u = dbits(s);
    int ef = (int)((u >> 52) & 0x7FF);
    uint64_t mant = u & ((1ull << 52) - 1);
    int e;
    if (ef == 0) {  // subnormal: normalise
        int lz = 0;
        while (!(mant & (1ull << 51))) { mant <<= 1; lz++; }
        mant = (mant << 1) & ((1ull << 52) - 1);
        e = -1023 - lz;
    } else {
        e = ef - 1023;
    }
    if (mant == 0) return e;
    if (e == 0) return 1;
    if (mant > (1ull << 40)) return e + 1;  // log2(f) > 2^-12.5, far above any half-spacing
    double t = (double)mant * 0x1p-52;      // f - 1
    double l2f = t * (1.4426950408889634 - 0.7213475204444817 * t);  // log2(1 + t), t < 2^-12
    int ae = e < 0 ? -e : e;
    int p = 31 - __builtin_clz((unsigned)ae);          // 2^p <= |e| < 2^(p+1)
    double half = 0x1p-53 * (double)(1u << p);         // half-spacing above |e| (binade of |e|)
    if (e < 0 && ae == (1 << p)) half *= 0.5;          // |e| a power of two: finer grid below it
    return (l2f < half) ? e : e + 1;
}

// ---------------------------------------------------------------- thread-per-number trees
// psumK<L> of src/pairwise.cuh:227-327 unrolled at compile time over leaf functors; the leaves
// are produced on demand so only O(log N) partial sums are live.
template <int L>
struct TreeWidth {
    static constexpr int calc() {
        if (L >= 4096) return 8192;
        int K = 8;
        while (K <= L) K <<= 1;
        return K;
    }
    static constexpr int v = calc();
};

// POS: every partial sum is known to be a positive normal double (the leaves are), so the bounded
// additions are neighbour steps (below_pos / above_pos).
template <int K, int L, int OFF, bool POS = false>
struct PTree {
    GRNS_HD static double add_rd(double a, double b) { return POS ? below_pos(a + b) : dadd_rd(a, b); }
    GRNS_HD static double add_ru(double a, double b) { return POS ? above_pos(a + b) : dadd_ru(a, b); }
    template <class Leaf>
    GRNS_HD static void run(Leaf &leaf, double &lo, double &up) {
        if constexpr (K == 2) {
            double l0, u0;
            leaf.template get<OFF>(l0, u0);
            if constexpr (L > 1) {
                double l1, u1;
                leaf.template get<OFF + 1>(l1, u1);
                lo = add_rd(l0, l1);
                up = add_ru(u0, u1);
            } else {
                lo = add_rd(l0, 0.0);
                up = add_ru(u0, 0.0);
            }
        } else {
            double l1, u1, l2 = 0.0, u2 = 0.0;
            PTree<K / 2, L, OFF, POS>::run(leaf, l1, u1);
            if constexpr (L > K / 2) PTree<K / 2, L - K / 2, OFF + K / 2, POS>::run(leaf, l2, u2);
            lo = add_rd(l1, l2);
            up = add_ru(u1, u2);
        }
    }
};

// psum_rd<N> and psum_ru<N> together (src/pairwise.cuh:384-458 dispatch).  POS as for PTree.
template <int N, class Leaf, bool POS = false>
GRNS_HD void psum_pair(Leaf &leaf, double &lo, double &up) {
    if constexpr (N >= 4) {
        PTree<TreeWidth<N>::v, N, 0, POS>::run(leaf, lo, up);
    } else if constexpr (N == 3) {
        double l0, u0, l1, u1, l2, u2;
        leaf.template get<0>(l0, u0);
        leaf.template get<1>(l1, u1);
        leaf.template get<2>(l2, u2);
        lo = dadd_rd(dadd_rd(l0, l1), l2);
        up = dadd_ru(dadd_ru(u0, u1), u2);
    } else if constexpr (N == 2) {
        double l0, u0, l1, u1;
        leaf.template get<0>(l0, u0);
        leaf.template get<1>(l1, u1);
        lo = dadd_rd(l0, l1);
        up = dadd_ru(u0, u1);
    } else {
        leaf.template get<0>(lo, up);
    }
}

// One direction of psum_pair: the same tree over leaves leaf.get<I>() returning one double.
template <int K, int L, int OFF, bool UP>
struct PTree1 {
    template <class Leaf>
    GRNS_HD static double run(Leaf &leaf) {
        if constexpr (K == 2) {
            double a = leaf.template get<OFF>();
            if constexpr (L > 1) return dadd_dir<UP>(a, leaf.template get<OFF + 1>());
            else return dadd_dir<UP>(a, 0.0);
        } else {
            double a = PTree1<K / 2, L, OFF, UP>::run(leaf);
            double b = 0.0;
            if constexpr (L > K / 2) b = PTree1<K / 2, L - K / 2, OFF + K / 2, UP>::run(leaf);
            return dadd_dir<UP>(a, b);
        }
    }
};

template <int N, bool UP, class Leaf>
GRNS_HD double psum_one(Leaf &leaf) {
    if constexpr (N >= 4) {
        return PTree1<TreeWidth<N>::v, N, 0, UP>::run(leaf);
    } else if constexpr (N == 3) {
        double a = leaf.template get<0>(), b = leaf.template get<1>(), d = leaf.template get<2>();
        return dadd_dir<UP>(dadd_dir<UP>(a, b), d);
    } else if constexpr (N == 2) {
        double a = leaf.template get<0>(), b = leaf.template get<1>();
        return dadd_dir<UP>(a, b);
    } else {
        return leaf.template get<0>();
    }
}

// ---------------------------------------------------------------- wave groups (N lanes per number)
// A group is G = 2^g >= N consecutive lanes of a wave64, aligned on G; lane gl of a group holds
// modulus gl.  Cross-lane traffic stays inside the group.
struct Group {
    int gl;    // lane within group
    int base;  // first wave lane of the group
    int G;     // group width
    int L;     // number of moduli (active leaves)
};

GRNS_D unsigned long long group_mask_bits(const Group &g) {
    return (g.G == 64) ? ~0ull : (((1ull << g.G) - 1) << g.base);
}
GRNS_D bool group_all(const Group &g, bool p) {
    unsigned long long b = __ballot(p);
    unsigned long long gm = group_mask_bits(g);
    return (b & gm) == gm;
}
GRNS_D bool group_any(const Group &g, bool p) {
    return (__ballot(p) & group_mask_bits(g)) != 0;
}
GRNS_D int bcast_i(int v, int lane) { return __shfl(v, lane, 64); }
GRNS_D double bcast_d(double v, int lane) { return __shfl(v, lane, 64); }
GRNS_D long long bcast_ll(long long v, int lane) { return __shfl(v, lane, 64); }

// Pairwise tree over the group's leaves in the CPU association (xor butterfly; an absent second
// half contributes the literal 0 of src/pairwise.cuh, an absent first half makes the node absent).
template <bool UP>
GRNS_D double group_psum(const Group &g, double v) {
    const int L = g.L;
    if (L == 1) return bcast_d(v, g.base);
    if (L == 3) {
        double a = bcast_d(v, g.base), b = bcast_d(v, g.base + 1), c = bcast_d(v, g.base + 2);
        return dadd_dir<UP>(dadd_dir<UP>(a, b), c);
    }
    int K = 2;
    if (L >= 4) { K = 8; while (K <= L) K <<= 1; }
    for (int h = 1; h < K; h <<= 1) {
        double w = (h < g.G) ? __shfl_xor(v, h, 64) : 0.0;
        int b = g.gl & ~(2 * h - 1);
        bool first = b < L;
        bool second = (h < g.G) && (b + h < L);
        bool upper = (g.gl & h) != 0;
        double fv = upper ? w : v;
        double sv = second ? (upper ? v : w) : 0.0;
        double r = dadd_dir<UP>(fv, sv);
        v = first ? r : v;
    }
    return v;
}

GRNS_D long long group_sum_ll(const Group &g, long long v) {
    for (int h = 1; h < g.G; h <<= 1) v += __shfl_xor(v, h, 64);
    return v;
}

// ---------------------------------------------------------------- position certificates
// Sign of the top mixed-radix digit of a canonical x (all the reference's ambiguity branch reads of
// the conversion, src/rnseval.cuh:62-65) from its position, or -1 when the position leaves it
// open: F (|F - x/M mod 1| <= err) above 1/m_{n-1} + err without wrapping puts x above M/m_{n-1};
// a certified representative T (x = T, or M + T when T < 0, |T| < 2^62, M > 2^95) places x below
// 2^62 < M/m_{n-1} or above M - 2^62 > M/m_{n-1}.
GRNS_HD int position_digit(const DevCtx *c, double F, bool cert, long long T) {
    if (F - c->err > c->mr_lim && F + c->err < 1.0) return 1;
    if (cert) return T < 0 ? 1 : 0;
    return -1;
}

// rns_cmp of two numbers from their position classes, 2 when they do not decide it.  Class 1: the
// CPU-path evaluation is regular (a valid enclosure, F in (tau, 1 - tau)); class 2: x = M + T,
// T < 0 certified, every reduction canonical (the evaluation is valid through the reference's
// conversion branch).  Two valid enclosures make rns_cmp the exact sign; class-2 numbers lie above
// M - 2^62 > M (1 - tau + err), above every class-1 number.
GRNS_HD int class_order(int kx, long long Tx, int ky, long long Ty) {
    if (kx == 2 && ky == 2) return (Tx > Ty) - (Tx < Ty);
    if (kx == 2 && ky == 1) return 1;
    if (kx == 1 && ky == 2) return -1;
    return 2;
}

// Wave-group forms of the certificates: lane gl holds residue gl; the sums run over the group in
// any order (the certificates verify what they find, and the filter bound holds for any order).
GRNS_D double group_sum_d(const Group &g, double v) {
    for (int h = 1; h < g.G; h <<= 1) v += __shfl_xor(v, h, 64);
    return v;
}
GRNS_D unsigned long long group_sum_u64(const Group &g, unsigned long long v) {
    for (int h = 1; h < g.G; h <<= 1) v += (unsigned long long)__shfl_xor((long long)v, h, 64);
    return v;
}

// crt_rep for the group's number with canonical residue di: the representative T of x within 2^62
// of 0 or M (x = T or M + T), verified against every residue.
GRNS_D bool group_crt_rep(const DevCtx *c, const Group &g, int di, long long &Ts) {
    if (!c->big_moduli || c->log2M < 96) return false;
    const bool act = g.gl < g.L;
    double a = 0.0;
    unsigned long long t = 0;
    if (act) {
        const int i = g.gl;
        double p = (double)di * c->dw[i];
        double q = floor(p * c->dinv[i]);
        double s = fma(-q, c->dm[i], p) + fma((double)di, c->dw[i], -p);
        s += (s < 0) ? c->dm[i] : 0.0;
        s -= (s >= c->dm[i]) ? c->dm[i] : 0.0;
        a = s * c->dinv[i];
        t = (unsigned long long)(unsigned)(int)s * c->mi64[i];
    }
    const double acc = group_sum_d(g, a);
    const unsigned long long T = group_sum_u64(g, t) - (unsigned long long)(long long)rint(acc) * c->M64;
    Ts = (long long)T;
    if (Ts >= (1ll << 62) || Ts <= -(1ll << 62)) return false;
    bool ok = true;
    if (act) {
        const int m = c->m[g.gl];
        const long long v = (long long)(int)(Ts >> 32) * (long long)c->c32[g.gl] + (long long)(unsigned)Ts;
        long long r = rem_trunc(v, m, c->dinv[g.gl]);
        r += (r < 0) ? m : 0;
        ok = (r == di);
    }
    return group_all(g, ok);
}

// The comparison filter of the group's number: F (within err of x/M mod 1), whether the residues
// are canonical and whether every term frac(x_i w_i / m_i) is in (2^-20, 1 - 2^-20).
GRNS_D double group_filter(const DevCtx *c, const Group &g, int xi, bool &canon, bool &terms) {
    const bool act = g.gl < g.L;
    double fi = 0.0;
    bool cn = true, tm = true;
    if (act) {
        const int i = g.gl;
        const double dx = (double)xi;
        cn = (unsigned)xi < (unsigned)c->m[i];
        const double P = dx * c->chi[i];
        const double e = fma(dx, c->chi[i], -P);
        const double t = fma(dx, c->clo[i], e);
        fi = fract_d(P) + t;
        tm = fabs(fi - 0.5) < 0.5 - 0x1p-20;
    }
    const double F = fract_d(group_sum_d(g, fi));
    canon = group_all(g, cn);
    terms = group_all(g, tm);
    return F;
}

// Sign of the top mixed-radix digit of the group's number from its position (-1: open).
GRNS_D int group_position_digit(const DevCtx *c, const Group &g, int xi) {
    bool canon, terms;
    const double F = group_filter(c, g, xi, canon, terms);
    if (!canon) return -1;
    long long T = 0;
    bool cert = false;
    if (!(F - c->err > c->mr_lim && F + c->err < 1.0)) cert = group_crt_rep(c, g, xi, T);
    return position_digit(c, F, cert, T);
}

// Mixed-radix digits (src/mrc.cuh:31-44), pipelined: at step j the final digit j is broadcast
// and every higher lane applies the same update the serial loop applies to it.  The inverses of
// B steps are requested together, so that the table's memory latency is paid once per B steps
// rather than once per step (B = 1 keeps the registers of callers that sit in streaming kernels).
template <int B = 8>
GRNS_D int group_mrc(const DevCtx *c, const Group &g, int xi) {
    int mr = xi;
    const int n = c->n;
    const bool act = g.gl < g.L;
    const int mi = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const int *tab = c->mrc_inv;
    const int steps = g.L - 1;
    for (int j0 = 0; j0 < steps; j0 += B) {
        int inv[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int j = j0 + k;
            inv[k] = (act && j < steps && g.gl > j) ? tab[j * n + g.gl] : 0;
        }
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int j = j0 + k;
            if (j >= steps) break;  // the same on every lane of the group
            int mrj = bcast_i(mr, g.base + j);
            if (act && g.gl > j) {
                if (mr < mrj) mr = (int)((long long)mi - (long long)mrj + (long long)mr);
                else mr = (int)((unsigned)mr - (unsigned)mrj);
                mr = mulmod3(mr, inv[k], mi, dinv);
            }
        }
    }
    return mr;
}
// mrs_cmp (src/mrc.cuh:13-22) of two digit vectors held by the group.
GRNS_D int group_mrs_cmp(const Group &g, int a, int b) {
    unsigned long long diff = __ballot(g.gl < g.L && a != b) & group_mask_bits(g);
    if (diff == 0) return 0;
    int k = 63 - __clzll(diff);
    int ak = bcast_i(a, k), bk = bcast_i(b, k);
    return ak > bk ? 1 : -1;
}

// ---------------------------------------------------------------- numbers as 32-bit words
// Division and scaling by powers of two run on the binary form of the numbers: word j (bits
// 32j .. 32j+31) of a value sits in lane j of the group.  Word arithmetic is modulo 2^(32 GD); every
// value handled stays below 2^(32 GD - 1) in magnitude, so the top bit of word GD-1 is the sign.
// Carries between words are resolved with ballots: a word generates a carry, or passes an incoming
// one on (all ones for an addition, zero for a subtraction), and the carries into all words follow
// from one 64-bit addition of those bit masks.

// The group's bits of a ballot, group lane 0 at bit 0.
template <int GD>
GRNS_D unsigned long long gballot(const Group &g, bool p) {
    const unsigned long long b = __ballot(p);
    if constexpr (GD == 64) return b;
    else return (b >> g.base) & ((1ull << GD) - 1);
}
// Bit k: carry into word k, from the words that generate one (gen) or pass one on (prop);
// gen and prop are disjoint.
GRNS_D unsigned long long carry_in_bits(unsigned long long gen, unsigned long long prop) {
    const unsigned long long a = gen | prop;
    return (a + gen) ^ prop;  // a ^ gen = prop
}
// The word of the next lower lane (0 into group lane 0), by a DPP wave shift.
// A whole-wave group takes the shift's zero fill for lane 0 (bound_ctrl: no source lane reads 0).
template <int GD>
GRNS_D unsigned wlower(const Group &g, unsigned v) {
    if constexpr (GD == 64) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, true);
    const unsigned r = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);  // wave_shr:1
    return g.gl == 0 ? 0u : r;
}
// The word of the next higher lane (0 into the top lane of the group).
template <int GD>
GRNS_D unsigned whigher(const Group &g, unsigned v) {
    if constexpr (GD == 64) return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, true);
    const unsigned r = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false);  // wave_shl:1
    return g.gl == GD - 1 ? 0u : r;
}
// Bit gl of a group mask (group lane 0 at bit 0), on each lane: for a whole-wave group the mask
// is read as a lane condition (one select), otherwise shifted out.
template <int GD>
GRNS_D unsigned lane_bit(const Group &g, unsigned long long mask) {
    if constexpr (GD == 64) return (unsigned)__builtin_amdgcn_inverse_ballot_w64(mask);
    else return (unsigned)((mask >> g.gl) & 1);
}
// Word idx (uniform in the group) of the group's value, on every lane of the group.
template <int GD>
GRNS_D unsigned gword(const Group &g, unsigned v, int idx) {
    if constexpr (GD == 64) return (unsigned)__builtin_amdgcn_readlane((int)v, idx);
    else return (unsigned)__shfl((int)v, g.base + idx, 64);
}
// The value shifted up by k >= 0 words (words below k become 0).
template <int GD>
GRNS_D unsigned wshift_up(const Group &g, unsigned v, int k) {
    const unsigned r = (unsigned)__shfl((int)v, g.base + (g.gl >= k ? g.gl - k : 0), 64);
    return g.gl >= k ? r : 0u;
}
// Number of significant words (0 for zero).
template <int GD>
GRNS_D int wlen(const Group &g, unsigned v) {
    const unsigned long long b = gballot<GD>(g, v != 0);
    return b ? 64 - __clzll((long long)b) : 0;
}
template <int GD>
GRNS_D bool wneg(const Group &g, unsigned v) { return (gword<GD>(g, v, GD - 1) >> 31) != 0; }

// u += v
template <int GD>
GRNS_D void wadd(const Group &g, unsigned &u, unsigned v) {
    const unsigned long long t = (unsigned long long)u + v;
    const unsigned w = (unsigned)t;
    const unsigned long long cin = carry_in_bits(gballot<GD>(g, (t >> 32) != 0), gballot<GD>(g, w == 0xffffffffu));
    u = w + lane_bit<GD>(g, cin);
}
// u -= q v (q < 2^32, the same on every lane of the group)
template <int GD>
GRNS_D void wmul_sub(const Group &g, unsigned &u, unsigned v, unsigned q) {
    const unsigned long long p = (unsigned long long)v * q;
    const unsigned hp = wlower<GD>(g, (unsigned)(p >> 32));
    // u - low(p) - hp = w - (c0 + c1) 2^32
    unsigned a, w;
    const bool c0 = __builtin_add_overflow((unsigned)p, hp, &a);
    const bool c1 = __builtin_sub_overflow(u, a, &w);
    const unsigned bp = wlower<GD>(g, (unsigned)c0 + (unsigned)c1);  // borrow of 0, 1 or 2 from below
    const unsigned w2 = w - bp;
    unsigned long long gen, prop;
    if constexpr (GD == 64) {  // the compares straight into lane masks (ICMP_ULT, ICMP_EQ)
        gen = __builtin_amdgcn_uicmp(w, bp, 36);
        prop = __builtin_amdgcn_uicmp(w2, 0u, 32);
    } else {
        gen = gballot<GD>(g, w < bp);
        prop = gballot<GD>(g, w2 == 0);
    }
    u = w2 - lane_bit<GD>(g, carry_in_bits(gen, prop));
}

// The value of canonical residues as words, by the Chinese remainder theorem:
// x = sum_i s_i (M / m_i) - k M with s_i = x_i w_i mod m_i and k = floor(sum_i s_i / m_i).  k comes
// from a double sum and is checked against the words when that sum is near an integer.
// crt_finish: the words of sum_i s_i M_i from the lane's column sums (lo: low halves of the terms
// of word gl, hi: their high halves, which belong to word gl + 1), the carries, and k M removed.
template <int GD>
GRNS_D unsigned crt_finish(const DevCtx *c, const Group &g, unsigned long long lo, unsigned long long hi,
                           double sum) {
    const double kf = floor(sum);
    const double fr = sum - kf;
    const int WW = c->words;
    const bool wl = g.gl < WW;
    // words of sum_i s_i M_i: lo_j + hi_{j-1} (< 2^39), then the carries
    const unsigned hlo = wlower<GD>(g, (unsigned)hi), hhi = wlower<GD>(g, (unsigned)(hi >> 32));
    const unsigned long long v = lo + hlo + ((unsigned long long)hhi << 32);
    const unsigned cp = wlower<GD>(g, (unsigned)(v >> 32));
    const unsigned long long t = (unsigned long long)(unsigned)v + cp;
    const unsigned w = (unsigned)t;
    const unsigned long long cin = carry_in_bits(gballot<GD>(g, (t >> 32) != 0), gballot<GD>(g, w == 0xffffffffu));
    unsigned X = w + lane_bit<GD>(g, cin);
    const unsigned mw = wl ? c->crt_words[(size_t)g.L * WW + g.gl] : 0u;
    wmul_sub<GD>(g, X, mw, (unsigned)kf);
    if (!(fr > 0x1p-30 && fr < 1.0 - 0x1p-30)) {
        if (wneg<GD>(g, X)) {
            wadd<GD>(g, X, mw);
        } else {
            unsigned Y = X;
            wmul_sub<GD>(g, Y, mw, 1u);
            if (!wneg<GD>(g, Y)) X = Y;
        }
    }
    return X;
}

// One term of a column sum on a whole-wave group: word `lane` of sv broadcast (v_readlane), its
// product by t added into the 64-bit sum by the multiply-add itself (v_mad_u64_u32 with the sum as
// addend) and that addition's carry-out counted (v_addc_co_u32): three vector instructions where
// the same sum written in C compiles to five (a separate 64-bit add and compare: the compiler does
// not use the multiply-add's carry-out).  The wait states are those the compiler itself keeps on
// this target: one between a vector write of a VGPR and v_readlane of it (s_nop 0 first, since the
// word may have been written just before), two between a vector write of a scalar register and a
// vector read of it (s_nop 1).  The sums it writes are read only by ordinary vector arithmetic
// (the next term, the final split), which needs no wait after a vector write.
GRNS_D void term_wave(unsigned sv, int lane, unsigned t, unsigned long long &acc, unsigned &wraps) {
    unsigned sw;
    unsigned long long cy;
    asm("s_nop 0\n\t"
        "v_readlane_b32 %[sw], %[sv], %[lane]\n\t"
        "s_nop 1\n\t"
        "v_mad_u64_u32 %[acc], %[cy], %[sw], %[t], %[acc]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32 %[wr], %[cy], %[wr], 0, %[cy]"
        : [acc] "+v"(acc), [wr] "+v"(wraps), [sw] "=&s"(sw), [cy] "=&s"(cy)
        : [sv] "v"(sv), [lane] "s"(lane), [t] "v"(t));
}
// The same for two numbers sharing the table word t.
GRNS_D void terms_wave2(unsigned sa, unsigned sb, int lane, unsigned t, unsigned long long &acca, unsigned &wa,
                        unsigned long long &accb, unsigned &wb) {
    unsigned xa, xb;
    unsigned long long ca, cb;
    asm("s_nop 0\n\t"
        "v_readlane_b32 %[xa], %[sa], %[lane]\n\t"
        "v_readlane_b32 %[xb], %[sb], %[lane]\n\t"
        "s_nop 1\n\t"
        "v_mad_u64_u32 %[aa], %[ca], %[xa], %[t], %[aa]\n\t"
        "v_mad_u64_u32 %[ab], %[cb], %[xb], %[t], %[ab]\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32 %[wa], %[ca], %[wa], 0, %[ca]\n\t"
        "v_addc_co_u32 %[wb], %[cb], %[wb], 0, %[cb]"
        : [aa] "+v"(acca), [wa] "+v"(wa), [ab] "+v"(accb), [wb] "+v"(wb), [xa] "=&s"(xa), [xb] "=&s"(xb),
          [ca] "=&s"(ca), [cb] "=&s"(cb)
        : [sa] "v"(sa), [sb] "v"(sb), [lane] "s"(lane), [t] "v"(t));
}

// acc[k] += (word i0 + b of row s[k]) t[b] for the terms i0 + b < lim of one batch, with the wraps of
// each 64-bit sum counted (LIM: the batch may run past lim; otherwise all of it is in range).
template <int GD, int K, bool LIM>
GRNS_D void word_terms(const Group &g, const int (&s)[K], const unsigned (&t)[WORDS_BATCH], int i0, int lim,
                       unsigned long long (&acc)[K], unsigned (&wraps)[K]) {
#pragma unroll
    for (int b = 0; b < WORDS_BATCH; b++) {
        if (LIM && i0 + b >= lim) break;  // the same on every lane
        if constexpr (GD == 64 && K == 2) {
            terms_wave2((unsigned)s[0], (unsigned)s[1], i0 + b, t[b], acc[0], wraps[0], acc[1], wraps[1]);
            continue;
        }
        if constexpr (GD == 64 && K == 1) {
            term_wave((unsigned)s[0], i0 + b, t[b], acc[0], wraps[0]);
            continue;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const unsigned long long p = (unsigned long long)gword<GD>(g, (unsigned)s[k], i0 + b) * t[b];
            const unsigned long long r = acc[k] + p;
            wraps[k] += r < p;
            acc[k] = r;
        }
    }
}

// The column sums of K numbers at once: s[k] on lane gl < L is the k-th number's s_gl; word gl of
// M_i is loaded once for all K (8 rows of the table per batch of global loads, so that its latency
// is paid once per 8 terms).  Each column is a 64-bit sum and a count of its wraps (at most L); lanes
// at or above the word count read word 0 of the table and are cleared at the end.
template <int GD, int K>
GRNS_D void crt_columns(const DevCtx *c, const Group &g, const int (&s)[K], unsigned long long (&lo)[K],
                        unsigned long long (&hi)[K]) {
    const int WW = c->words;
    const bool wl = g.gl < WW;
    const global_uint *tw = (const global_uint *)(uintptr_t)c->crt_words + (wl ? g.gl : 0);
    unsigned long long acc[K];
    unsigned wraps[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = wraps[k] = 0;
    int i0 = 0;
    for (; i0 + WORDS_BATCH <= g.L; i0 += WORDS_BATCH) {  // whole batches: no bounds on the terms
        unsigned t[WORDS_BATCH];
#pragma unroll
        for (int b = 0; b < WORDS_BATCH; b++) t[b] = tw[(size_t)(i0 + b) * WW];
        word_terms<GD, K, false>(g, s, t, i0, g.L, acc, wraps);
    }
    if (i0 < g.L) {
        unsigned t[WORDS_BATCH];
#pragma unroll
        for (int b = 0; b < WORDS_BATCH; b++) t[b] = i0 + b < g.L ? tw[(size_t)(i0 + b) * WW] : 0u;
        word_terms<GD, K, true>(g, s, t, i0, g.L, acc, wraps);
    }
    // sum = lo + hi 2^32
#pragma unroll
    for (int k = 0; k < K; k++) {
        lo[k] = wl ? (unsigned)acc[k] : 0u;
        hi[k] = wl ? (acc[k] >> 32) + ((unsigned long long)wraps[k] << 32) : 0u;
    }
}

// Words of the group's number with canonical residue xi (lane gl < n).  The reduction of s is the
// exact one: the reciprocal one (mulmod_recip) can return s_i - m_i or wrap, and the unsigned word
// sums need 0 <= s_i < m_i.
template <int GD>
GRNS_D unsigned group_words(const DevCtx *c, const Group &g, int xi) {
    const bool act = g.gl < g.L;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const int s[1] = {act ? mulmod3(xi, c->w[g.gl], c->m[g.gl], dinv) : 0};
    const double sum = group_sum_d(g, act ? (double)s[0] * dinv : 0.0);
    unsigned long long lo[1], hi[1];
    crt_columns<GD, 1>(c, g, s, lo, hi);
    return crt_finish<GD>(c, g, lo[0], hi[0], sum);
}

// The words of two numbers (x and d of a division) in one pass over the table.
template <int GD>
GRNS_D void group_words_pair(const DevCtx *c, const Group &g, int xi, int di, unsigned &X, unsigned &D) {
    const bool act = g.gl < g.L;
    int m = 1, w = 0;
    double dinv = 1.0;
    if (act) {
        m = c->m[g.gl];
        w = c->w[g.gl];
        dinv = c->dinv[g.gl];
    }
    const int s[2] = {act ? mulmod3(xi, w, m, dinv) : 0, act ? mulmod3(di, w, m, dinv) : 0};
    const double sx = group_sum_d(g, act ? (double)s[0] * dinv : 0.0);
    const double sd = group_sum_d(g, act ? (double)s[1] * dinv : 0.0);
    unsigned long long lo[2], hi[2];
    crt_columns<GD, 2>(c, g, s, lo, hi);
    X = crt_finish<GD>(c, g, lo[0], hi[0], sx);
    D = crt_finish<GD>(c, g, lo[1], hi[1], sd);
}

// Residue modulo m_gl (lanes gl < n) of the value held in the first nw words, nw <= log2M / 32 + 1
// (RNS_POW2 has the rows 2^0 .. 2^log2M).
template <int GD>
GRNS_D int words_residue(const DevCtx *c, const Group &g, unsigned v, int nw) {
    const bool act = g.gl < g.L;
    const int n = c->n;
    // 2^(32 k) mod m_gl for 8 words at a time (loads issued together, as in group_words)
    const global_int *p2 = (const global_int *)(uintptr_t)c->pow2 + (act ? g.gl : 0);
    unsigned long long acc[1] = {0};  // the sum is acc + wraps 2^64 (lanes gl >= n read modulus 0's row)
    unsigned wraps[1] = {0};
    const int vv[1] = {(int)v};
    int k0 = 0;
    for (; k0 + WORDS_BATCH <= nw; k0 += WORDS_BATCH) {
        unsigned t[WORDS_BATCH];
#pragma unroll
        for (int k = 0; k < WORDS_BATCH; k++) t[k] = (unsigned)p2[(size_t)(32 * (k0 + k)) * n];
        word_terms<GD, 1, false>(g, vv, t, k0, nw, acc, wraps);
    }
    if (k0 < nw) {
        unsigned t[WORDS_BATCH];
#pragma unroll
        for (int k = 0; k < WORDS_BATCH; k++) t[k] = k0 + k < nw ? (unsigned)p2[(size_t)(32 * (k0 + k)) * n] : 0u;
        word_terms<GD, 1, true>(g, vv, t, k0, nw, acc, wraps);
    }
    if (!act) return 0;
    const unsigned long long lo = (unsigned)acc[0], hi = (acc[0] >> 32) + ((unsigned long long)wraps[0] << 32);
    const int m = c->m[g.gl];
    const double dinv = c->dinv[g.gl];
    const long long rl = rem_trunc((long long)lo, m, dinv), rh = rem_trunc((long long)hi, m, dinv);
    return (int)rem_trunc(rh * (long long)c->c32[g.gl] + rl, m, dinv);
}
// The same for a value below M (held in up to GD words).
template <int GD>
GRNS_D int group_words_residue(const DevCtx *c, const Group &g, unsigned v, int nw) {
    if (nw > (c->log2M >> 5) + 1) nw = (c->log2M >> 5) + 1;  // a value below M has no more words
    return words_residue<GD>(c, g, v, nw);
}

// x = q d + r for canonical residues with d != 0: long division of the words (TAOCP vol. 2,
// 4.3.1, Algorithm D), the quotient's residues from its words, r = x - q d in the residues.
// Quotient and remainder of two integers are unique, so these are the residues rns_div and
// rns_div_fast (src/rnsdiv.cuh:19-114) return for such inputs.  Needs GD >= c->words.
template <int GD>
GRNS_D void group_div_words(const DevCtx *c, const Group &g, int xi, int di, int &q, int &r) {
    const bool act = g.gl < g.L;
    unsigned X, D;
    group_words_pair<GD>(c, g, xi, di, X, D);
    const int wx = wlen<GD>(g, X), wd = wlen<GD>(g, D);
    q = 0;
    r = xi;
    if (wx < wd) return;
    const int sh = __clz((int)gword<GD>(g, D, wd - 1));
    if (sh) {
        D = (D << sh) | (wlower<GD>(g, D) >> (32 - sh));
        X = (X << sh) | (wlower<GD>(g, X) >> (32 - sh));
    }
    const unsigned v1 = gword<GD>(g, D, wd - 1);
    const unsigned v2 = wd >= 2 ? gword<GD>(g, D, wd - 2) : 0u;
    const unsigned vinv = recip32_preinv(v1, v2);  // <v1, v2> is normalised: the top bit of v1 is set
    const unsigned long long V12 = ((unsigned long long)v1 << 32) | v2;
    const int top = wx - wd;
    unsigned Ds = wshift_up<GD>(g, D, top);
    unsigned Q = 0;
    for (int j = top; j >= 0; j--) {
        const unsigned u2 = gword<GD>(g, X, j + wd), u1 = gword<GD>(g, X, j + wd - 1);
        // with one divisor word (v2 = 0) u0 does not change the estimate: any lane will do
        const unsigned u0 = gword<GD>(g, X, (j + wd - 2) & (GD - 1));
        // qh = floor(<u2, u1, u0> / <v1, v2>): Knuth's estimate of step D3 with its test on v2 made
        // exact, at most one above the quotient word.  It would reach 2^32 only for <u2, u1> =
        // <v1, v2>, where the quotient word is 2^32 - 1 or 2^32 - 2 and 2^32 - 1 is taken.
        unsigned r1, r0;
        const unsigned qe = div32_preinv(u2, u1, u0, v1, v2, vinv, r1, r0);
        const unsigned qh = (((unsigned long long)u2 << 32) | u1) < V12 ? qe : 0xffffffffu;
        wmul_sub<GD>(g, X, Ds, qh);  // qh = 0 leaves X as it is
        unsigned qd = qh;
        if (wneg<GD>(g, X)) {
            qd--;
            wadd<GD>(g, X, Ds);
        }
        if (lane_bit<GD>(g, 1ull << j)) Q = qd;
        Ds = whigher<GD>(g, Ds);
    }
    const bool qnz = gballot<GD>(g, Q != 0) != 0;
    const int qr = group_words_residue<GD>(c, g, Q, top + 1);
    if (act) {
        const int m = c->m[g.gl];
        const double dinv = c->dinv[g.gl];
        // The reference builds q by reciprocal additions (mod_add, src/modular.cuh:29-34), and one
        // whose operands sum to exactly m returns m itself when RN(m * RN(1/m)) < 1.  A nonzero
        // quotient divisible by such an m_i is always reached through such an addition (every
        // partial quotient and the final +1 are nonzero), so its residue there is m_i, not 0.
        q = (qnz && qr == 0 && (double)m * dinv < 1.0) ? m : qr;
        r = psubmod(xi, mulmod3(qr, di, m, dinv), m, dinv);
    }
}

// rns_scale2pow (src/rnsscal.cuh:93-119) replayed on the words of a canonical x.  Step s of the
// reference maps its value X_s = floor(x / 2^(30 s)) to X_{s+1} = floor(X_s / 2^j): with its
// residues y (as the reference holds them, possibly y + m or y - m) it forms c = y * w (for the
// rank), X_s mod 2^j, multiple = (y - (X_s mod 2^j) mod m + m) % m and y' = multiple * (2^j)^-1,
// the products by reciprocal reduction.  Whenever every c is canonical, the rank and the residue
// X_s mod 2^j of the reference are consistent, so X_s mod 2^j is the word bits [30 s, 30 s + j)
// of x, and y' follows from y exactly as the reference computes it.  A step whose c is not
// canonical or whose y' wraps int32 (the reference then continues with another number) sets
// `replay_ok` false and the caller runs the reference steps instead.
template <int GD>
GRNS_D int group_scale2pow_words(const DevCtx *c, const Group &g, int xi, unsigned D, bool &replay_ok) {
    const bool act = g.gl < g.L;
    const int m = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const int w = act ? c->w[g.gl] : 0;
    const unsigned X = group_words<GD>(c, g, xi);
    const int xbits = 32 * wlen<GD>(g, X);
    const unsigned t = D / THRESH, dr = D % THRESH;
    const unsigned steps = t + (dr ? 1u : 0u);
    int y = act ? xi : 0;  // lanes past n keep 0 throughout (m = 1, w = 0, inverses 0)
    bool bad = false;
    // (2^j)^-1 mod m for the full steps (j = THRESH) and the last one (j = dr), loaded once
    const int inv_full = act && t ? c->pow2_inv[THRESH - 1][g.gl] : 0;
    const int inv_last = act && dr ? c->pow2_inv[dr - 1][g.gl] : 0;
    for (unsigned st = 0; st < steps; st++) {
        const long long b = (long long)st * THRESH;
        if (b >= xbits && group_all(g, !act || y == 0)) break;  // every later step maps 0 to 0
        const int j = (st < t) ? THRESH : (int)dr;
        unsigned r = 0;  // bits [b, b + j) of x
        if (b < xbits) {
            const int w0 = (int)(b >> 5), sh = (int)(b & 31);
            const unsigned lo = gword<GD>(g, X, w0);
            const unsigned hi = (sh && w0 + 1 < GD) ? gword<GD>(g, X, w0 + 1) : 0u;
            r = (sh ? (lo >> sh) | (hi << (32 - sh)) : lo) & ((1u << j) - 1u);
        }
        const int cc = mulmod_recip(y, w, m, dinv);
        bad |= (unsigned)cc >= (unsigned)m;
        // rm = r % m and mult = (y - rm + m) % m as rem_trunc forms them: r < 2^30 and
        // |y - rm + m| < 2^32 have exact doubles, formed from 32-bit pieces
        const int rm = (int)rem_trunc_d((long long)r, (double)r, m, dinv);
        const int mult = (int)rem_trunc_d((long long)y + (m - rm), (double)y + (double)(m - rm), m, dinv);
        const int inv = st < t ? inv_full : inv_last;
        const int qi = (int)(((double)mult * (double)inv) * dinv);
        const long long v = (long long)mult * inv - (long long)qi * m;
        bad |= v < -2147483648ll || v > 2147483647ll;
        y = (int)v;
    }
    replay_ok = !group_any(g, bad);
    return y;
}

GRNS_D er_t bcast_er(er_t v, int lane) {
    er_t r;
    r.frac = bcast_d(v.frac, lane);
    r.exp = bcast_ll(v.exp, lane);
    return r;
}

// Status bits a group computation can raise (cases in which the reference itself would read
// outside its tables or not terminate).
enum : int { GST_OK = 0, GST_REFINE = 1, GST_DIVLOOP = 2 };

// Refinement loop, src/rnseval.cuh:79-101 (identical tail of rns_eval_compute_fast).
GRNS_D void group_refine(const DevCtx *c, const Group &g, int &s, double sumu, interval_t &out, int &status) {
    const bool act = g.gl < g.L;
    const int mi = act ? c->m[g.gl] : 1;
    const double dm = act ? c->dm[g.gl] : 1.0;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    long long K = 0;
    int it = 0;
    int kprev = -1, p2 = 0;
    while (sumu < c->accuracy) {
        // MAX(-(ceil(log2(sumu))+1), REF): log2 of a negative sum is NaN and selects REF; a zero
        // sum gives +inf, which the reference turns into an index outside RNS_POW2.
        double kv = (double)c->ref_factor;
        if (sumu > 0) {
            double a = (double)(-(ceil_log2_libm(sumu) + 1));
            kv = a > kv ? a : kv;
        }
        if (sumu == 0 || kv > c->log2M || ++it > 4096) {
            status |= GST_REFINE;
            out.low.frac = out.upp.frac = __builtin_nan("");
            out.low.exp = out.upp.exp = 0;
            return;
        }
        int k = (int)kv;
        if (k != kprev) {  // the scale factor repeats from step to step: its row is read once
            p2 = act ? c->pow2[(size_t)k * c->n + g.gl] : 0;
            kprev = k;
        }
        s = mulmod_recip(s, p2, mi, dinv);
        double fu = ddiv_ru((double)s, dm);
        sumu = group_psum<true>(g, fu);
        sumu -= to_uint(sumu);
        K += k;
    }
    double fl = ddiv_rd((double)s, dm);
    double suml = group_psum<false>(g, fl);
    suml -= to_uint(suml);
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    out.low.exp -= K;
    out.upp.exp -= K;
}

// rns_eval_compute (FAST=false, src/rnseval.cuh:29-102) and rns_eval_compute_fast (FAST=true,
// :112-167) for the number whose residue of modulus gl is xi.  The result is the same on all
// lanes of the group.
// mrd/has_mrd receive this lane's mixed-radix digit when the evaluation computed the digits.
template <bool FAST>
GRNS_D interval_t group_eval(const DevCtx *c, const Group &g, int xi, int &status, int &mrd, bool &has_mrd) {
    interval_t out;
    has_mrd = false;
    const bool act = g.gl < g.L;
    if (group_all(g, !act || xi == 0)) {
        out.low.frac = 0; out.low.exp = 0; out.upp.frac = 0; out.upp.exp = 0;
        return out;
    }
    int s = act ? mulmod_recip(xi, c->w[g.gl], c->m[g.gl], c->dinv[g.gl]) : 0;
    double fl = act ? dmul_rd((double)s, c->rd[g.gl]) : 0.0;
    double fu = act ? dmul_ru((double)s, c->ru[g.gl]) : 0.0;
    double suml = group_psum<false>(g, fl);
    double sumu = group_psum<true>(g, fu);
    unsigned whl = to_uint(suml), whu = to_uint(sumu);
    if (FAST) {
        suml -= whl;
        sumu -= whu;
        if (sumu >= c->accuracy) {
            out.low = er_set_d(suml);
            out.upp = er_set_d(sumu);
            return out;
        }
        group_refine(c, g, s, sumu, out, status);
        return out;
    }
    suml = suml - whl;
    sumu = sumu - whu;
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    int mr = -1;
    if (whl != whu) {
        // only the sign of the top digit is read: the position of x often fixes it
        mr = group_position_digit(c, g, xi);
        if (mr < 0) {
            int d = group_mrc(c, g, xi);
            mrd = d;
            has_mrd = true;
            mr = bcast_i(d, g.base + g.L - 1);
        }
    }
    if (mr > 0) { out.upp = c->inv_upp; return out; }
    if (mr == 0) out.low = c->unit_low;
    if (sumu >= c->accuracy) return out;
    group_refine(c, g, s, sumu, out, status);
    return out;
}

template <bool FAST>
GRNS_D interval_t group_eval(const DevCtx *c, const Group &g, int xi, int &status) {
    int d;
    bool h;
    return group_eval<FAST>(c, g, xi, status, d, h);
}

// rns_cmp decided without running the reference's evaluations, or 2: identical residues (0), a
// zero against a canonical nonzero number, two regular numbers (filter order, then the small- and
// 64-bit difference certificates), then the position classes.  Each case returns the reference's
// result (see cmp_lane and class_order).
GRNS_D int group_cmp_fast(const DevCtx *c, const Group &g, int xi, int yi) {
    const bool act = g.gl < g.L;
    if (group_all(g, !act || xi == yi)) return 0;
    bool cx, tx, cy, ty;
    const double Fx = group_filter(c, g, xi, cx, tx);
    const double Fy = group_filter(c, g, yi, cy, ty);
    if (group_all(g, !act || xi == 0) && cy) return -1;
    if (group_all(g, !act || yi == 0) && cx) return 1;
    const bool rx = cx && tx && Fx > c->tau && Fx < 1.0 - c->tau;
    const bool ry = cy && ty && Fy > c->tau && Fy < 1.0 - c->tau;
    if (rx && ry) {
        const double e2 = 2.0 * c->err;
        if (Fx - Fy > e2) return 1;
        if (Fy - Fx > e2) return -1;
        // within M/4 of each other: t = x_0 - y_0 centred, checked against every residue
        const int m0 = c->m[0];
        int t = (int)((unsigned)bcast_i(xi, g.base) - (unsigned)bcast_i(yi, g.base));
        if (t > m0 / 2) t -= m0;
        else if (t <= -((m0 + 1) / 2)) t += m0;
        bool ok = true;
        int d = 0;
        if (act) {
            const int m = c->m[g.gl];
            d = (int)((unsigned)xi - (unsigned)yi);
            d += (d < 0) ? m : 0;
            const int tm = t + ((t < 0) ? m : 0);
            ok = (d == tm) && (2 * (long long)(t < 0 ? -t : t) < (long long)m);
        }
        if (group_all(g, ok)) return (t > 0) - (t < 0);
        long long T;
        if (group_crt_rep(c, g, d, T)) return (T > 0) - (T < 0);
        return 2;
    }
    long long Tx = 0, Ty = 0;
    const int kx = rx ? 1 : (cx && tx && group_crt_rep(c, g, xi, Tx) && Tx < 0 ? 2 : 0);
    const int ky = ry ? 1 : (cy && ty && group_crt_rep(c, g, yi, Ty) && Ty < 0 ? 2 : 0);
    return class_order(kx, Tx, ky, Ty);
}

// rns_cmp, src/rnscmp.cuh:20-39 (the digits an evaluation computed are reused by the final
// mixed-radix comparison: both come from the same conversion of the same residues)
GRNS_D int group_cmp(const DevCtx *c, const Group &g, int xi, int yi, int &status) {
    const int f = group_cmp_fast(c, g, xi, yi);
    if (f != 2) return f;
    int dx = 0, dy = 0;
    bool hx, hy;
    interval_t ex = group_eval<false>(c, g, xi, status, dx, hx);
    interval_t ey = group_eval<false>(c, g, yi, status, dy, hy);
    if (er_ucmp(ex.low, ey.upp) > 0) return 1;
    if (er_ucmp(ey.low, ex.upp) > 0) return -1;
    if (group_all(g, g.gl >= g.L || xi == yi)) return 0;
    int mx = hx ? dx : group_mrc(c, g, xi);
    int my = hy ? dy : group_mrc(c, g, yi);
    return group_mrs_cmp(g, mx, my);
}

}  // namespace grns
