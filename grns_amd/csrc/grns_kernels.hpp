// grns_kernels.hpp -- HIP kernels of the GRNS hot path for gfx950 (wave64).
//
// Two mappings are used:
//  * thread-per-number for the HBM-bound element-wise work (evaluation, comparison, arg-max,
//    mpint add/sub/mul): the n residues of one number sit in one lane's registers, the loads of
//    a modulus-major array are coalesced across the wave, and the pairwise trees are unrolled
//    at compile time for n in {4, 8, 16, 32, 64};
//  * group-per-number (G = 2^ceil(log2 n) lanes, modulus i in lane i) for the long, branchy
//    loops (division, power-of-two scaling) and for the rare cases of the element-wise kernels
//    (mixed-radix conversion, refinement), which the wave resolves cooperatively after a ballot.
#pragma once

#include <utility>

#include "grns_device.hpp"

namespace grns {

constexpr int BLOCK = 256;

template <int N>
struct Pow2Ceil {
    static constexpr int v = N <= 1 ? 1 : (N <= 2 ? 2 : (N <= 4 ? 4 : (N <= 8 ? 8 : (N <= 16 ? 16 : (N <= 32 ? 32 : 64)))));
};

GRNS_D Group make_group(int G, int L) {
    int lane = threadIdx.x & 63;
    Group g;
    g.G = G;
    g.gl = lane & (G - 1);
    g.base = lane & ~(G - 1);
    g.L = L;
    return g;
}

// Block index in an order that gives each XCD a contiguous range: the hardware hands consecutive
// workgroups to the 8 XCDs in turn, and the group kernels read residue rows whose cache lines hold
// the numbers of neighbouring blocks, so neighbours should share an XCD (and its L2).  Needs a
// grid that is a multiple of 8 (else the plain order).
GRNS_D unsigned xcd_block() {
    const unsigned b = blockIdx.x, nb = gridDim.x;
    if (nb & 7u) return b;
    return (b & 7u) * (nb >> 3) + (b >> 3);
}

// The k-th set bit of mask (k >= 0), or -1.
GRNS_D int kth_bit(unsigned long long mask, int k) {
    for (int i = 0; i < k && mask; i++) mask &= mask - 1;
    return mask ? (__ffsll((long long)mask) - 1) : -1;
}
GRNS_D unsigned long long drop_low_bits(unsigned long long mask, int k) {
    for (int i = 0; i < k && mask; i++) mask &= mask - 1;
    return mask;
}

// ---------------------------------------------------------------- evaluation leaves
template <int N>
struct EvalLeaf {
    const DevCtx *c;
    const int *xr;
    // the trees ask for leaves in ascending order: the constants come in views of 4 moduli
    cdouble *dw = nullptr, *dm = nullptr, *dinv = nullptr, *rd = nullptr, *ru = nullptr;
    double smax = 0.0;  // max |s_i - 2^30|: >= 2^30 (s_i = 0, negative or wrapping) leaves the number to the rare pass
    template <int I>
    GRNS_HD void get(double &lo, double &up) {
        if constexpr (I % 4 == 0) {
            const DevCtx *cc = reload_point(c);
            dw = const_view(cc->dw + I);
            dm = const_view(cc->dm + I);
            dinv = const_view(cc->dinv + I);
            rd = const_view(cc->rd + I);
            ru = const_view(cc->ru + I);
        }
        constexpr int k = I % 4;
        int xi = xr[I];
#if defined(__HIP_DEVICE_COMPILE__)
        // Keeps the conversion after the loads of the other residues: left free, the scheduler
        // moves the first one up to its load, and the wave then waits for that load before it
        // issues the rest.
        asm volatile("" : "+v"(xi));
#endif
        double s = mulmod_recip_d_nowrap((double)xi, dw[k], dm[k], dinv[k]);
        // s in [1, 2^31): the reference's int result is s itself and both products are at least
        // 2^-32, so their bounds (and those of every partial sum) are the neighbouring doubles
        smax = fmax(smax, fabs(s - 0x1p30));
        lo = below_pos(s * rd[k]);
        up = above_pos(s * ru[k]);
    }
};

// Common path of rns_eval_compute[_fast] for one number held by this lane.  Returns false when
// the number needs mixed-radix conversion or refinement (then `out` is not final).
// The scan of all residues before the trees also makes the compiler issue every load up front.
template <int N, bool FAST>
GRNS_HD bool eval_hot(const DevCtx *c, const int (&xr)[N], interval_t &out) {
    int nz = 0;
#pragma unroll
    for (int i = 0; i < N; i++) nz |= xr[i];
    EvalLeaf<N> leaf{c, xr};
    double suml, sumu;
    psum_pair<N, EvalLeaf<N>, true>(leaf, suml, sumu);
    unsigned whl = to_uint(suml), whu = to_uint(sumu);
    suml -= whl;
    sumu -= whu;
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    if (nz == 0) {
        out.low.frac = 0; out.low.exp = 0; out.upp.frac = 0; out.upp.exp = 0;
        return true;
    }
    return leaf.smax < 0x1p30 && (FAST || whl == whu) && (sumu >= c->accuracy);
}

GRNS_D void store_interval(interval_t *p, const interval_t &v) {
    double2 *q = reinterpret_cast<double2 *>(p);
    q[0] = make_double2(v.low.frac, __longlong_as_double(v.low.exp));
    q[1] = make_double2(v.upp.frac, __longlong_as_double(v.upp.exp));
}


// Residues of number j into registers.  LD_MM: modulus-major (snum == 1) and j < 2^29, checked by
// the launcher: a wave-uniform row pointer stepped from modulus to modulus plus one 32-bit byte
// offset, so each load is a scalar-base global load (the row pointer is recomputed per modulus
// rather than N of them being held in scalar registers).  LD_ANY: any strides.  LD_AUTO: the
// modulus-major form when the number allows it.
enum LoadMode { LD_ANY = 0, LD_MM = 1, LD_AUTO = 2 };

template <int N>
GRNS_D void load_number_mm(const int *__restrict__ x, long long smod, long long j, int (&xr)[N]) {
    const unsigned off = lane_offset((unsigned)j * 4u);
    const long long step = smod * 4;
    long long row = reinterpret_cast<long long>(x);
#pragma unroll
    for (int i = 0; i < N; i++) {
        xr[i] = __builtin_nontemporal_load(reinterpret_cast<const global_int *>(row + off));
        row = uniform_point(row + step);
    }
}

// The store counterpart of load_number_mm (digit i of number j at x[i * smod + j], j < 2^30): the
// row pointer advances by a scalar add per digit instead of N row offsets i * smod computed ahead of
// the caller's loop, which would take 2N scalar registers for as long as the loop runs.
typedef __attribute__((address_space(1))) int global_int_w;
template <int N>
GRNS_D void store_number_mm(int *__restrict__ x, long long smod, long long j, const int (&xr)[N]) {
    const unsigned off = lane_offset((unsigned)j * 4u);
    const long long step = smod * 4;
    long long row = reinterpret_cast<long long>(x);
#pragma unroll
    for (int i = 0; i < N; i++) {
        __builtin_nontemporal_store(xr[i], reinterpret_cast<global_int_w *>(row + off));
        row = uniform_point(row + step);
    }
}

template <int N, int MODE = LD_AUTO>
GRNS_D void load_number(const int *__restrict__ x, long long smod, long long snum, long long j, int (&xr)[N]) {
    if (MODE == LD_MM || (MODE == LD_AUTO && snum == 1 && j < (1ll << 29))) {
        load_number_mm<N>(x, smod, j, xr);
        return;
    }
    const int *p = x + j * snum;
#pragma unroll
    for (int i = 0; i < N; i++) xr[i] = __builtin_nontemporal_load(p + (long long)i * smod);
}

// ---------------------------------------------------------------- rare-item queue
// The per-call kernels leave the numbers (or pairs) whose result needs the full reference
// algorithm to a second kernel: the streaming pass appends their indices here, and the rare pass
// spreads them over every wave group of the device, so that a few long evaluations (refinement
// loops, mixed-radix conversions) neither stall the streaming waves nor run one after another in
// the wave that met them.  The queue is split into RARE_SUBQ sub-queues, each with its own counter
// on its own 256-byte line (block b pushes to sub-queue b mod RARE_SUBQ): device-scope atomics on
// one address are served one after another, and with tens of thousands of rare lanes that alone
// outlasts a fast streaming pass.  ctr[2] is set when a sub-queue was full; then the overflowing
// items carry a marker in their output and the rare pass scans for it.  The thread-per-number
// kernels also copy a queued number's residues (a pair's, for the comparison) from registers into
// `res`, RARE_RES_WIDTH ints per slot for the first res_cap slots of each sub-queue, so that the
// rare pass reads 4n contiguous bytes per item instead of one cache line per residue row.  A
// one-wave kernel after the rare pass clears the counters for the next call on the stream.
// Evaluation, comparison, division and scaling share the queue of a (context, stream) pair, one
// call at a time.
constexpr int RARE_RES_WIDTH = 128;  // ints per staged slot: two numbers of up to 64 residues
constexpr int RARE_SUBQ = 64;        // sub-queues
constexpr int RARE_SUB_BASE = 64;    // ctr[RARE_SUB_BASE + s * RARE_SUB_STRIDE]: count of sub-queue s
constexpr int RARE_SUB_STRIDE = 64;  // 256 bytes between counters
constexpr int RARE_CTR_WORDS = RARE_SUB_BASE + RARE_SUBQ * RARE_SUB_STRIDE;
struct RareQ {
    long long *items;  // sub-queue s holds items[s * cap .. s * cap + cap)
    unsigned *ctr;
    unsigned cap;      // items per sub-queue
    int *res;          // staged residues of the first res_cap items of each sub-queue
    unsigned res_cap;
    unsigned submask;  // block b pushes to sub-queue b & submask
};

GRNS_D unsigned *rare_counter(const RareQ &q, int s) { return q.ctr + RARE_SUB_BASE + s * RARE_SUB_STRIDE; }

// Reserves slots in the block's sub-queue for the lanes with `rare` (all 64 lanes of the wave call
// it, one atomic per wave).  Returns the lane's slot within the sub-queue, or ~0u when it did not
// fit (the overflow flag is then set and the caller writes the marker); 0 for the other lanes.
GRNS_D unsigned rare_reserve(const RareQ &q, bool rare, int &sub) {
    const unsigned long long m = __ballot(rare);
    sub = (int)(blockIdx.x & q.submask);
    if (!m) return 0u;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(rare_counter(q, sub), (unsigned)__popcll(m));
    base = __shfl(base, leader, 64);
    if (!rare) return 0u;
    const unsigned slot = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    if (slot < q.cap) return slot;
    q.ctr[2] = 1u;
    return ~0u;
}

// Appends j for the lanes with `rare`.  Returns false for a lane whose item did not fit.
GRNS_D bool rare_push(const RareQ &q, bool rare, long long j) {
    int sub;
    const unsigned slot = rare_reserve(q, rare, sub);
    if (!rare) return true;
    if (slot == ~0u) return false;
    q.items[(size_t)sub * q.cap + slot] = j;
    return true;
}

// rare_push that also stages the residues of the queued numbers: xr (and yr, for a pair: PAIR) of
// the lanes whose slot is below res_cap, at res[(sub * res_cap + slot) * RARE_RES_WIDTH] (x) and
// + N (y), as 16-byte stores.  The arrays stay register arrays (references, indices known at
// compile time).
template <int N, bool PAIR>
GRNS_D bool rare_push_staged(const RareQ &q, bool rare, long long j, const int (&xr)[N], const int (&yr)[N]) {
    int sub;
    const unsigned slot = rare_reserve(q, rare, sub);
    if (!rare) return true;
    if (slot == ~0u) return false;
    if (slot < q.res_cap) {
        int4 *r = reinterpret_cast<int4 *>(q.res + ((size_t)sub * q.res_cap + slot) * RARE_RES_WIDTH);
#pragma unroll
        for (int i = 0; i < N; i += 4) r[i / 4] = make_int4(xr[i], xr[i + 1], xr[i + 2], xr[i + 3]);
        if constexpr (PAIR) {
#pragma unroll
            for (int i = 0; i < N; i += 4) r[(N + i) / 4] = make_int4(yr[i], yr[i + 1], yr[i + 2], yr[i + 3]);
        }
    }
    q.items[(size_t)sub * q.cap + slot] = j;
    return true;
}

// The rare pass's view of the queue: the items of the sub-queues laid end to end (item k of the
// pass is slot k - pre[s] of the sub-queue s with pre[s] <= k < pre[s + 1]), in LDS for the block.
struct RareView {
    unsigned pre[RARE_SUBQ + 1];
};

// Every thread of the block calls it (a barrier inside): the prefix of the sub-queue counts.
GRNS_D void rare_view(const RareQ &q, RareView &v) {
    if (threadIdx.x < 64) {
        const int s = threadIdx.x;
        const unsigned cs = *rare_counter(q, s);
        unsigned inc = cs < q.cap ? cs : q.cap;
        for (int h = 1; h < 64; h <<= 1) {
            const unsigned t = (unsigned)__shfl_up((int)inc, h, 64);
            if (s >= h) inc += t;
        }
        v.pre[s + 1] = inc;
        if (s == 0) v.pre[0] = 0u;
    }
    __syncthreads();
}

// Item k < pre[RARE_SUBQ] of the pass: its index j, and its staged residues or nullptr.
GRNS_D long long rare_locate(const RareQ &q, const RareView &v, long long k, const int *&staged) {
    int lo = 0, hi = RARE_SUBQ;  // pre[lo] <= k < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((long long)v.pre[mid] <= k) lo = mid;
        else hi = mid;
    }
    const unsigned slot = (unsigned)(k - v.pre[lo]);
    staged = slot < q.res_cap ? q.res + ((size_t)lo * q.res_cap + slot) * RARE_RES_WIDTH : nullptr;
    return q.items[(size_t)lo * q.cap + slot];
}

// Clears the counters of a rare-item queue for the next call on its stream (one wave, launched
// after the rare pass).  Clearing them from the last block of the rare pass to finish instead
// (a device-scope atomic count of finished blocks) costs more: 35 us for the 2048 blocks of a pass
// over 2^24 numbers, against 4.5 us for this kernel.
static __global__ void k_rare_reset(unsigned *ctr) {
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0u;
    if (threadIdx.x < RARE_SUBQ) ctr[RARE_SUB_BASE + threadIdx.x * RARE_SUB_STRIDE] = 0u;
}

// Marker of an evaluation left to the rare pass after a queue overflow (no result has it).
constexpr long long EVAL_MARK = 0x7ff0dead0000beefll;

// Resolve the lanes of `mask` with whole-group evaluations.  Not inlined: the rare path's
// registers then do not add to those of the streaming loop that calls it.
template <bool FAST>
__device__ __attribute__((noinline)) void coop_eval_lanes(const DevCtx *c, int G, int L, const int *__restrict__ x, long long smod, long long snum,
                            long long j, unsigned long long mask, interval_t *__restrict__ out, int *status) {
    Group g = make_group(G, L);
    const int P = 64 / G;
    const int gid = (threadIdx.x & 63) / G;
    while (mask) {
        int src = kth_bit(mask, gid);
        long long jl = __shfl(j, src < 0 ? 0 : src, 64);
        mask = drop_low_bits(mask, P);
        if (src >= 0) {
            int xi = g.gl < L ? x[jl * snum + (long long)g.gl * smod] : 0;
            int st = 0;
            interval_t r = group_eval<FAST>(c, g, xi, st);
            if (g.gl == 0) {
                store_interval(out + jl, r);
                if (st) atomicOr(status, st);
            }
        }
    }
}

// One evaluation per lane (number j, valid when j < count); all 64 lanes of the wave call it.
// MODE LD_MM: the caller guarantees snum == 1 and count <= 2^29.
template <int N, bool FAST, int MODE = LD_AUTO>
GRNS_D void eval_step(const DevCtx *__restrict__ c, const int *__restrict__ x, long long smod, long long snum,
                      long long j, long long count, interval_t *__restrict__ out, int *__restrict__ status) {
    bool rare = false;
    if (j < count) {
        int xr[N];
        load_number<N, MODE>(x, smod, snum, j, xr);
        interval_t r;
        if (eval_hot<N, FAST>(c, xr, r)) store_interval(out + j, r);
        else rare = true;
    }
    unsigned long long mask = __ballot(rare);
    if (mask) coop_eval_lanes<FAST>(c, Pow2Ceil<N>::v, N, x, smod, snum, j, mask, out, status);
}

// Plain per-call evaluation, streaming pass: one number per lane; numbers that need the
// cooperative algorithm go to the rare-item queue with their residues.  Up to n = 32 the kernel
// keeps 4 waves per SIMD (at most 128 VGPRs), as it has without the staged copy; at n = 64 it is
// held to 2 (226 VGPRs, no spill; left to itself the compiler takes 256 and one wave per SIMD,
// 0.78-0.86 ms for 2^23 numbers against 0.53 ms).
template <int N, bool FAST, int MODE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(N <= 32 ? 4 : 2))) void k_eval(const DevCtx *__restrict__ c, const int *__restrict__ x, long long smod,
                                                long long snum, long long count, interval_t *__restrict__ out,
                                                RareQ q) {
    for (long long base = (long long)blockIdx.x * BLOCK; base < count; base += (long long)gridDim.x * BLOCK) {
        const long long j = base + threadIdx.x;
        bool rare = false;
        int xr[N];
        if (j < count) {
            load_number<N, MODE>(x, smod, snum, j, xr);
            interval_t r;
            if (eval_hot<N, FAST>(c, xr, r)) store_interval(out + j, r);
            else rare = true;
        }
        if (!rare_push_staged<N, false>(q, rare, j, xr, xr)) {
            interval_t mk;
            mk.low.frac = __builtin_nan(""); mk.low.exp = EVAL_MARK;
            mk.upp.frac = __builtin_nan(""); mk.upp.exp = EVAL_MARK;
            store_interval(out + j, mk);
        }
    }
}

// Any n <= 64: one group per number throughout.
template <bool FAST>
__global__ __launch_bounds__(BLOCK) void k_eval_group(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                                      long long smod, long long snum, long long count,
                                                      interval_t *__restrict__ out, int *__restrict__ status) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        long long j = j0 + (threadIdx.x & 63) / G;
        if (j < count) {
            int xi = g.gl < g.L ? x[j * snum + (long long)g.gl * smod] : 0;
            int st = 0;
            interval_t r = group_eval<FAST>(c, g, xi, st);
            if (g.gl == 0) {
                store_interval(out + j, r);
                if (st) atomicOr(status, st);
            }
        }
    }
}

// ---------------------------------------------------------------- filter sums
// Approximation of x/M mod 1 from sum_i frac(x_i * w_i / m_i) with w_i/m_i held as a double-double;
// |F - x/M mod 1| <= n^2 2^-52 (c->err bounds it generously).  `regular` is true when the
// residues are canonical, every term frac(x_i w_i / m_i) lies in (2^-20, 1 - 2^-20) (so the CPU
// path's reciprocal reduction of x_i w_i is canonical: its quotient error is below 2^-20.4) and
// F lies in (tau, 1 - tau) (so its evaluation needs neither mixed-radix conversion nor
// refinement and is a valid enclosure of x/M).

// One term of the filter: frac(x_i w_i / m_i) within 2^-42 (x_i khi_i is exact, |x_i klo_i| < 2^9
// is added with one rounding of at most 2^-44, klo_i is within 2^-74 of w_i/m_i - khi_i).
GRNS_HD double filter_term(double dx, double kh, double kl) { return fract_d(fma(dx, kl, fract_d(dx * kh))); }

// The flags of one number: residues canonical and every term in (2^-20, 1 - 2^-20).
struct FilterAcc {
    double acc = 0.0, lo = 1.0, hi = 0.0;
    LaneAny bad;
    GRNS_HD void add(int xi, int m, double kh, double kl) {
        bad.set_if((unsigned)xi >= (unsigned)m);
        const double f = filter_term((double)xi, kh, kl);
        lo = fmin(lo, f);
        hi = fmax(hi, f);
        acc += f;
    }
    GRNS_HD double frac(const DevCtx *c, bool &regular, bool &canon, bool &terms) const {
        const double F = fract_d(acc);
        canon = !bad.any();
        terms = lo > 0x1p-20 && hi < 1.0 - 0x1p-20;
        regular = canon && terms && F > c->tau && F < 1.0 - c->tau;
        return F;
    }
};

template <int N, bool REG = true>
GRNS_HD double filter_frac(const DevCtx *c, const int (&xr)[N], bool &regular, bool &canon, bool &terms) {
    FilterAcc a;
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += 8) {
        const DevCtx *cc = reload_point(c);  // 8 moduli of constants at a time
        cdouble *kh = const_view(cc->khi + i0);
        cdouble *kl = const_view(cc->klo + i0);
        cint *cm = const_view(cc->m + i0);
#pragma unroll
        for (int k = 0; k < 8 && i0 + k < N; k++) a.add(xr[i0 + k], cm[k], kh[k], kl[k]);
    }
    return a.frac(c, regular, canon, terms);
}

template <int N, bool REG = true>
GRNS_HD double filter_frac(const DevCtx *c, const int (&xr)[N], bool &regular, bool &canon) {
    bool terms;
    return filter_frac<N, REG>(c, xr, regular, canon, terms);
}

// Approximation F of x/M mod 1 for the arg-max keys: sum_i frac(x_i khi_i) + sum_i x_i klo_i, where
// khi_i (22 significant bits) makes x_i khi_i exact and |x_i klo_i| < 2^9.  For canonical x,
// |F - x/M mod 1| <= n 2^-38 (the fma chain, |B| <= n 2^9) + n 2^-43 (klo) + 2^-37 (the last
// sums) < 2^-31 up to n = 64.  Five VALU operations per residue.
// The sum depends on the residues only mod m_i (w_i/m_i times a multiple of m_i is an integer), so
// the key holds for non-canonical int32 residues too, which then stand for their canonical value.
// SCALED: the key of 2^KEY_SHIFT x (constants shi, slo), for the arg-max keys of values near 0 or M.
template <int N, bool SCALED = false>
GRNS_HD double key_frac(const DevCtx *c, const int (&xr)[N]) {
    double A = 0.0, B = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += 8) {
        const DevCtx *cc = reload_point(c);  // 8 moduli of constants at a time
        cdouble *kh = const_view((SCALED ? cc->shi : cc->khi) + i0);
        cdouble *kl = const_view((SCALED ? cc->slo : cc->klo) + i0);
#pragma unroll
        for (int k = 0; k < 8 && i0 + k < N; k++) {
            double dx = (double)xr[i0 + k];
            A += fract_d(dx * kh[k]);
            B = fma(dx, kl[k], B);
        }
    }
    return fract_d(A + B);
}


// Exact sign of x - y for two canonical numbers known to lie within M/4 of each other when their
// difference is small: t = (x_0 - y_0) reduced into (-m_0/2, m_0/2] is x - y exactly iff
// t = x_i - y_i (mod m_i) for every i (then t = x - y mod M, and both lie within M/4 of 0).
// Returns 2 when the check fails (|x - y| >= m_0/2).
template <int N>
GRNS_HD int small_diff_sign(const DevCtx *c, const int (&xr)[N], const int (&yr)[N]) {
    const int m0 = c->m[0];
    int t = xr[0] - yr[0];  // (-m0, m0)
    if (t > m0 / 2) t -= m0;
    else if (t <= -((m0 + 1) / 2)) t += m0;
    LaneAny bad;
    bad.set_if(2 * (long long)(t < 0 ? -t : t) >= (long long)c->m_min);
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += 8) {
        const DevCtx *cc = reload_point(c);
        cint *cm = const_view(cc->m + i0);
#pragma unroll
        for (int k = (i0 == 0 ? 1 : 0); k < 8 && i0 + k < N; k++) {
            // x_i - y_i - t is in (-3m/2, 3m/2): a multiple of m_i exactly when it is -m_i, 0 or m_i
            const int u = xr[i0 + k] - yr[i0 + k] - t;
            const int m = cm[k];
            bad.set_if(u != 0 && u != m && u != -m);
        }
    }
    if (bad.any()) return 2;
    return (t > 0) - (t < 0);
}

// Exact sign of x - y for two canonical numbers known to lie within M/4 of each other, or 2 when
// this cannot be certified.  d = x - y mod M is formed residue-wise; with s_i = d_i w_i mod m_i
// and K the integer nearest to sum s_i / m_i, T = sum s_i (M/m_i) - K M equals x - y over the
// integers whenever |x - y| is small, and T is computed mod 2^64.  T is accepted only when
// |T| < 2^62 and T = d_i mod m_i for every i: then T = x - y (mod M) and |T| < M/4, so T is x - y.
// The residues are read from memory (px[i * sx], py[i * sy]) in two passes: this is a rare
// path, and keeping it out of registers leaves the callers' hot loops their occupancy.
template <int U = 4>
GRNS_HD int near_sign(const DevCtx *c, const int *px, long long sx, const int *py, long long sy) {
    if (!c->big_moduli) return 2;  // the 64-bit remainders below need m_i >= 2^22
    const int n = c->n;
    unsigned long long T = 0;
    double acc = 0.0;
#pragma unroll U
    for (int i = 0; i < n; i++) {
        int m = c->m[i];
        int d = px[i * sx] - py[i * sy];
        d += (d < 0) ? m : 0;
        // s = d * w mod m, canonical
        double p = (double)d * c->dw[i];
        double q = floor(p * c->dinv[i]);
        double s = fma(-q, c->dm[i], p) + fma((double)d, c->dw[i], -p);
        s += (s < 0) ? c->dm[i] : 0.0;
        s -= (s >= c->dm[i]) ? c->dm[i] : 0.0;
        acc = fma(s, c->dinv[i], acc);
        T += (unsigned long long)(unsigned)(int)s * c->mi64[i];
    }
    double K = rint(acc);
    T -= (unsigned long long)(long long)K * c->M64;
    long long Ts = (long long)T;
    if (Ts >= (1ll << 62) || Ts <= -(1ll << 62)) return 2;
    int th = (int)(Ts >> 32);
    unsigned tl = (unsigned)Ts;
    bool ok = true;
#pragma unroll U
    for (int i = 0; i < n; i++) {
        int m = c->m[i];
        int d = px[i * sx] - py[i * sy];
        d += (d < 0) ? m : 0;
        long long v = (long long)th * (long long)c->c32[i] + (long long)tl;  // T mod m_i, up to a multiple of m_i
        long long r = rem_trunc(v, m, c->dinv[i]);
        r += (r < 0) ? m : 0;
        ok &= (r == d);
    }
    if (!ok) return 2;
    return (Ts > 0) - (Ts < 0);
}

// The signed representative T of the number with canonical residues d when it lies within 2^62
// of 0 or of M (x = T or x = M + T), by the construction of near_sign; false when it does not.
template <int N>
GRNS_HD bool crt_rep(const DevCtx *c, const int (&d)[N], long long &Ts) {
    if (!c->big_moduli || c->log2M < 96) return false;  // unique below 2^62 only when M is large
    unsigned long long T = 0;
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        double p = (double)d[i] * c->dw[i];
        double q = floor(p * c->dinv[i]);
        double s = fma(-q, c->dm[i], p) + fma((double)d[i], c->dw[i], -p);
        s += (s < 0) ? c->dm[i] : 0.0;
        s -= (s >= c->dm[i]) ? c->dm[i] : 0.0;
        acc = fma(s, c->dinv[i], acc);
        T += (unsigned long long)(unsigned)(int)s * c->mi64[i];
    }
    double K = rint(acc);
    T -= (unsigned long long)(long long)K * c->M64;
    Ts = (long long)T;
    if (Ts >= (1ll << 62) || Ts <= -(1ll << 62)) return false;
    int th = (int)(Ts >> 32);
    unsigned tl = (unsigned)Ts;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; i++) {
        long long v = (long long)th * (long long)c->c32[i] + (long long)tl;
        long long r = rem_trunc(v, c->m[i], c->dinv[i]);
        r += (r < 0) ? c->m[i] : 0;
        ok &= (r == d[i]);
    }
    return ok;
}

// ---------------------------------------------------------------- comparison
// rns_cmp for the numbers held by this lane, or 2 when the reference algorithm itself has to run
// (a number whose CPU-path evaluation is not regular, or a near-tie the certificates cannot
// settle).  xr/yr hold the residues; px/py (sx/sy apart) are where they were read from.  NEAR
// false leaves out the 64-bit certificate (its pairs then return 2).
template <int N, bool NEAR = true>
GRNS_HD int cmp_lane(const DevCtx *c, const int (&xr)[N], const int (&yr)[N], const int *px, long long sx,
                     const int *py, long long sy) {
    bool rx, ry, cx, cy, tx, ty;
    // one pass per number, each with its own scalar loads of the constants (sharing them would keep
    // them live across both passes)
    const double fx = filter_frac<N>(c, xr, rx, cx, tx);
    const double fy = filter_frac<N>(c, yr, ry, cy, ty);
    if (!(rx && ry)) {
        // Outside the regular case the reference's evaluations have to be run, except in two cases
        // whose rns_cmp result does not depend on them: identical residues (the same interval twice,
        // low <= upp, so neither bound test holds and the residue test returns 0), and a zero
        // against a canonical nonzero number (zero evaluates to [0, 0] exactly, the other's upper
        // bound is >= 0, and either its lower bound is > 0 or the exact mixed-radix order decides).
        bool eq = true, zx = true, zy = true;
#pragma unroll
        for (int i = 0; i < N; i++) {
            eq &= (xr[i] == yr[i]);
            zx &= (xr[i] == 0);
            zy &= (yr[i] == 0);
        }
        if (eq) return 0;
        if (zx && cy) return -1;
        if (zy && cx) return 1;
        return 2;
    }
    // Both CPU-path evaluations are valid enclosures: rns_cmp returns the exact sign.
    double e2 = 2.0 * c->err;
    if (fx - fy > e2) return 1;
    if (fy - fx > e2) return -1;
    // equal values give t = 0 and every difference 0: the small-difference certificate returns 0
    int s = small_diff_sign<N>(c, xr, yr);
    if (!NEAR) return s;
    return s != 2 ? s : near_sign<2>(c, px, sx, py, sy);
}

// The lanes of `mask` run the reference comparison with whole groups.
template <int N>
GRNS_D void coop_cmp_lanes(const DevCtx *c, const int *__restrict__ x, long long xsm,
                                                         long long xsn, const int *__restrict__ y, long long ysm,
                                                         long long ysn, long long j, unsigned long long mask,
                                                         int *__restrict__ sign, int *__restrict__ status) {
    const int G = Pow2Ceil<N>::v;
    Group g = make_group(G, N);
    const int P = 64 / G;
    const int gid = (threadIdx.x & 63) / G;
    while (mask) {
        int src = kth_bit(mask, gid);
        long long jl = __shfl(j, src < 0 ? 0 : src, 64);
        mask = drop_low_bits(mask, P);
        if (src >= 0) {
            int xi = g.gl < N ? x[jl * xsn + (long long)g.gl * xsm] : 0;
            int yi = g.gl < N ? y[jl * ysn + (long long)g.gl * ysm] : 0;
            int st = 0;
            int r = group_cmp(c, g, xi, yi, st);
            if (g.gl == 0) {
                sign[jl] = r;
                if (st) atomicOr(status, st);
            }
        }
    }
}

// One comparison per lane (number j, valid when j < count); all 64 lanes of the wave call it.
// MODE LD_MM: the caller guarantees xsn == ysn == 1 and count <= 2^29.
template <int N, int MODE = LD_AUTO>
GRNS_D void cmp_step(const DevCtx *__restrict__ c, const int *__restrict__ x, long long xsm, long long xsn,
                     const int *__restrict__ y, long long ysm, long long ysn, long long j, long long count,
                     int *__restrict__ sign, int *__restrict__ status) {
    {
        bool rare = false;
        if (j < count) {
            int xr[N], yr[N];
            load_number<N, MODE>(x, xsm, xsn, j, xr);
            load_number<N, MODE>(y, ysm, ysn, j, yr);
            int res = cmp_lane<N>(c, xr, yr, x + j * xsn, xsm, y + j * ysn, ysm);
            rare = (res == 2);
            if (!rare) sign[j] = res;
        }
        unsigned long long mask = __ballot(rare);
        if (mask) coop_cmp_lanes<N>(c, x, xsm, xsn, y, ysm, ysn, j, mask, sign, status);
    }
}

// Plain per-call comparison, streaming pass: one pair per lane, no cooperative code (so the loop
// keeps its registers); pairs the lane cannot settle go to the rare-item queue.
template <int N, int MODE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(N <= 32 ? 5 : 1))) void k_cmp(const DevCtx *__restrict__ c, const int *__restrict__ x, long long xsm,
                                               long long xsn, const int *__restrict__ y, long long ysm, long long ysn,
                                               long long count, int *__restrict__ sign, RareQ q) {
    for (long long base = (long long)blockIdx.x * BLOCK; base < count; base += (long long)gridDim.x * BLOCK) {
        const long long j = base + threadIdx.x;
        int res = 0;
        int xr[N], yr[N];
        if (j < count) {
            load_number<N, MODE>(x, xsm, xsn, j, xr);
            load_number<N, MODE>(y, ysm, ysn, j, yr);
            res = cmp_lane<N, false>(c, xr, yr, x + j * xsn, xsm, y + j * ysn, ysm);
            if (res != 2) sign[j] = res;
        }
        if (!rare_push_staged<N, true>(q, res == 2, j, xr, yr)) sign[j] = 2;  // marker: left to the scan of the rare pass
    }
}

// Any n <= 64: one group per pair throughout.
static __global__ __launch_bounds__(BLOCK) void k_cmp_group(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                                     long long xsm, long long xsn, const int *__restrict__ y,
                                                     long long ysm, long long ysn, long long count,
                                                     int *__restrict__ sign, int *__restrict__ status) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        long long j = j0 + (threadIdx.x & 63) / G;
        if (j < count) {
            int xi = g.gl < g.L ? x[j * xsn + (long long)g.gl * xsm] : 0;
            int yi = g.gl < g.L ? y[j * ysn + (long long)g.gl * ysm] : 0;
            int st = 0;
            int r = group_cmp(c, g, xi, yi, st);
            if (g.gl == 0) {
                sign[j] = r;
                if (st) atomicOr(status, st);
            }
        }
    }
}

// ---------------------------------------------------------------- rare passes
// What the streaming passes leave: numbers near 0 or near M (ambiguous integer parts, refinement)
// and pairs with such numbers.  A lane first tries to settle its item with exact certificates of
// the number's position; what remains runs the reference algorithm in wave groups.

// Leaves over the lane's s_i = x_i w_i mod m_i (reciprocal reduction) for the refinement loop:
// ddiv_ru / ddiv_rd(s_i, m_i) (src/rnseval.cuh:85,93).
template <bool UP>
struct DivLeaf {
    const DevCtx *c;
    const int *s;
    template <int I>
    GRNS_HD double get() {
        return UP ? ddiv_ru((double)s[I], c->dm[I]) : ddiv_rd((double)s[I], c->dm[I]);
    }
};
// ... and for the main sums: dmul_rd / dmul_ru(s_i, RECIP_RD/RU) (src/rnseval.cuh:46-47).
struct MulLeaf {
    const DevCtx *c;
    const int *s;
    template <int I>
    GRNS_HD void get(double &lo, double &up) {
        lo = dmul_rd((double)s[I], c->rd[I]);
        up = dmul_ru((double)s[I], c->ru[I]);
    }
};

// The refinement loop of rns_eval_compute in one lane (src/rnseval.cuh:79-101), the same steps as
// group_refine.
template <int N>
GRNS_HD void refine_lane(const DevCtx *c, int (&s)[N], double sumu, interval_t &out, int &status) {
    long long K = 0;
    int it = 0;
    while (sumu < c->accuracy) {
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
        const int k = (int)kv;
        const int *row = c->pow2 + (size_t)k * c->n;
#pragma unroll
        for (int i = 0; i < N; i++) s[i] = mulmod_recip(s[i], row[i], c->m[i], c->dinv[i]);
        DivLeaf<true> lu{c, s};
        sumu = psum_one<N, true>(lu);
        sumu -= to_uint(sumu);
        K += k;
    }
    DivLeaf<false> ll{c, s};
    double suml = psum_one<N, false>(ll);
    suml -= to_uint(suml);
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    out.low.exp -= K;
    out.upp.exp -= K;
}

// rns_eval_compute[_fast] in one lane for the rare pass.  The ambiguous case (mixed-radix conversion
// in the reference, which only reads the sign of the top digit) is settled without the conversion
// when the position of x fixes that digit: x/M above 1/m_{n-1} (key) or x = M + T within 2^62 of M
// (certificate) make it positive, x = T below 2^62 makes it 0.  Returns false when the position is
// uncertain (then the cooperative algorithm runs).
template <int N, bool FAST>
GRNS_HD bool eval_full(const DevCtx *c, const int (&xr)[N], interval_t &out, int &status) {
    int nz = 0;
    bool canon = true;
#pragma unroll
    for (int i = 0; i < N; i++) {
        nz |= xr[i];
        canon &= (unsigned)xr[i] < (unsigned)c->m[i];
    }
    if (nz == 0) {
        out.low.frac = 0; out.low.exp = 0; out.upp.frac = 0; out.upp.exp = 0;
        return true;
    }
    int s[N];
#pragma unroll
    for (int i = 0; i < N; i++) s[i] = mulmod_recip(xr[i], c->w[i], c->m[i], c->dinv[i]);
    MulLeaf leaf{c, s};
    double suml, sumu;
    psum_pair<N>(leaf, suml, sumu);
    unsigned whl = to_uint(suml), whu = to_uint(sumu);
    suml -= whl;
    sumu -= whu;
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    if (!FAST && whl != whu) {
        if (!canon) return false;  // the reference's conversion of non-canonical residues is its own
        const double F = key_frac<N>(c, xr);
        long long T = 0;
        bool cert = false;
        if (!(F - c->err > c->mr_lim && F + c->err < 1.0)) cert = crt_rep<N>(c, xr, T);
        const int mr = position_digit(c, F, cert, T);
        if (mr < 0) return false;
        if (mr > 0) {
            out.upp = c->inv_upp;
            return true;
        }
        out.low = c->unit_low;
    }
    if (sumu >= c->accuracy) return true;
    refine_lane<N>(c, s, sumu, out, status);
    return true;
}

// Position class of a number for the rare comparisons: 1 when its CPU-path evaluation is regular,
// 2 when x = M + T with T < 0 certified (canonical, every reduction canonical: the evaluation is a
// valid enclosure through the reference's conversion branch), 0 otherwise.
template <int N>
GRNS_HD int edge_class(const DevCtx *c, const int (&xr)[N], long long &T) {
    bool reg, canon, terms;
    filter_frac<N>(c, xr, reg, canon, terms);
    if (reg) return 1;
    if (!canon || !terms) return 0;
    if (crt_rep<N>(c, xr, T) && T < 0) return 2;
    return 0;
}

// rns_cmp (src/rnscmp.cuh:20-39) in one lane for a pair the streaming pass left, or 2 when the
// cooperative algorithm has to run: the filters and certificates of cmp_lane, then the position
// classes (two valid enclosures make rns_cmp the exact sign), then the reference's own steps with
// in-lane evaluations, the final mixed-radix order taken from certified positions (it is exact
// for canonical residues, and so is the order of the certified values).
template <int N>
GRNS_HD int cmp_full(const DevCtx *c, const int (&xr)[N], const int (&yr)[N], const int *px, long long sx,
                    const int *py, long long sy, int &status) {
    const int r = cmp_lane<N, true>(c, xr, yr, px, sx, py, sy);
    if (r != 2) return r;
    long long Tx = 0, Ty = 0;
    const int kx = edge_class<N>(c, xr, Tx);
    const int ky = edge_class<N>(c, yr, Ty);
    const int o = class_order(kx, Tx, ky, Ty);
    if (o != 2) return o;
    interval_t ex, ey;
    if (!eval_full<N, false>(c, xr, ex, status) || !eval_full<N, false>(c, yr, ey, status)) return 2;
    if (er_ucmp(ex.low, ey.upp) > 0) return 1;
    if (er_ucmp(ey.low, ex.upp) > 0) return -1;
    bool eq = true, canon = true;
#pragma unroll
    for (int i = 0; i < N; i++) {
        eq &= xr[i] == yr[i];
        canon &= (unsigned)xr[i] < (unsigned)c->m[i] && (unsigned)yr[i] < (unsigned)c->m[i];
    }
    if (eq) return 0;
    if (!canon || !crt_rep<N>(c, xr, Tx) || !crt_rep<N>(c, yr, Ty)) return 2;
    // x = Tx or M + Tx, y likewise: order the positions
    const bool hx = Tx < 0, hy = Ty < 0;
    if (hx != hy) return hx ? 1 : -1;
    return (Tx > Ty) - (Tx < Ty);
}

// Item k of a rare pass: the queued items (k < nq = pre[RARE_SUBQ]), then (after an overflow)
// every output position k - nq, of which only the marked ones are taken.  staged: the item's
// staged residues, or nullptr.
GRNS_D long long rare_item(const RareQ &q, const RareView &v, long long k, const int *&staged) {
    staged = nullptr;
    if (k < (long long)v.pre[RARE_SUBQ]) return rare_locate(q, v, k, staged);
    return k - (long long)v.pre[RARE_SUBQ];
}

// Rare pass of the evaluation: one wave group per item (the group forms settle the ambiguous
// positions by certificates and run the refinement loop with the whole group).
template <int N, bool FAST>
__global__ __launch_bounds__(BLOCK) void k_eval_rare(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                     long long smod, long long snum, long long count,
                                                     interval_t *__restrict__ out, int *__restrict__ status, RareQ q) {
    const int G = Pow2Ceil<N>::v;
    Group g = make_group(G, N);
    const int P = 64 / G;
    const long long gid = (((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6) * P + (threadIdx.x & 63) / G;
    const long long ngroups = ((long long)gridDim.x * BLOCK >> 6) * P;
    __shared__ RareView v;
    rare_view(q, v);
    const long long nq = v.pre[RARE_SUBQ];
    const long long total = nq + (q.ctr[2] ? count : 0);
    int st = 0;
    for (long long k = gid; k < total; k += ngroups) {
        const int *src;
        const long long j = rare_item(q, v, k, src);
        if (k >= nq && reinterpret_cast<const long long *>(out + j)[1] != EVAL_MARK) continue;
        const int xi = g.gl < N ? (src ? src[g.gl] : x[j * snum + (long long)g.gl * smod]) : 0;
        interval_t r = group_eval<FAST>(c, g, xi, st);
        if (g.gl == 0) store_interval(out + j, r);
    }
    if (g.gl == 0 && st) atomicOr(status, st);
}

// Rare pass of the comparison: one wave group per item.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_cmp_rare(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                    long long xsm, long long xsn, const int *__restrict__ y,
                                                    long long ysm, long long ysn, long long count,
                                                    int *__restrict__ sign, int *__restrict__ status, RareQ q) {
    const int G = Pow2Ceil<N>::v;
    Group g = make_group(G, N);
    const int P = 64 / G;
    const long long gid = (((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6) * P + (threadIdx.x & 63) / G;
    const long long ngroups = ((long long)gridDim.x * BLOCK >> 6) * P;
    __shared__ RareView v;
    rare_view(q, v);
    const long long nq = v.pre[RARE_SUBQ];
    const long long total = nq + (q.ctr[2] ? count : 0);
    int st = 0;
    for (long long k = gid; k < total; k += ngroups) {
        const int *src;
        const long long j = rare_item(q, v, k, src);
        if (k >= nq && sign[j] != 2) continue;
        const int xi = g.gl < N ? (src ? src[g.gl] : x[j * xsn + (long long)g.gl * xsm]) : 0;
        const int yi = g.gl < N ? (src ? src[N + g.gl] : y[j * ysn + (long long)g.gl * ysm]) : 0;
        const int r = group_cmp(c, g, xi, yi, st);
        if (g.gl == 0) sign[j] = r;
    }
    if (g.gl == 0 && st) atomicOr(status, st);
}

// ---------------------------------------------------------------- arg-max (rns_max)
// Candidates carry rigorous bounds [lo, hi] on x/M (or [-1, 2] when unknown); two candidates whose
// bounds overlap are ordered exactly by mixed-radix digits, equal values by the lower index.
struct MaxCand {
    double lo;
    double hi;
    long long idx;
};

GRNS_D int canon_res(int v, int m, double dinv) {
    long long r = rem_trunc((long long)v, m, dinv);
    return (int)(r < 0 ? r + m : r);
}

// Exact comparison of numbers a and b of x by the group (value = residues taken mod m_i).
GRNS_D int group_cmp_exact(const DevCtx *c, const Group &g, const int *__restrict__ x, long long smod, long long snum,
                           long long a, long long b) {
    int xa = 0, xb = 0;
    if (g.gl < g.L) {
        int mi = c->m[g.gl];
        double dinv = c->dinv[g.gl];
        xa = canon_res(x[a * snum + (long long)g.gl * smod], mi, dinv);
        xb = canon_res(x[b * snum + (long long)g.gl * smod], mi, dinv);
    }
    if (group_all(g, g.gl >= g.L || xa == xb)) return 0;
    int ma = group_mrc<1>(c, g, xa);
    int mb = group_mrc<1>(c, g, xb);
    return group_mrs_cmp(g, ma, mb);
}

// Decide whether `e` replaces `best`: 1 yes, 0 no, 2 needs the exact order.
GRNS_D int max_decide(const MaxCand &e, const MaxCand &best) {
    if (e.idx < 0) return 0;
    if (best.idx < 0) return 1;
    if (e.lo > best.hi) return 1;
    if (best.lo > e.hi) return 0;
    return 2;
}

// Exact order (sign of x_a - x_b) settled within one lane: 0 for equal residues, +-1 from the
// difference certificates when `near` (both numbers known to lie within M/4 of each other), else
// 2, and a wave group then runs mixed-radix conversion.  pa/pb point at residue 0 of the two
// numbers, whose residues are sa/sb apart.  The residues are read from memory: this path is rare and stays out of the
// registers of the streaming loops that call it.
template <int U>
GRNS_D int lane_cmp_exact(const DevCtx *c, const int *pa, long long sa, const int *pb, long long sb, bool near) {
    const int n = c->n;
    const int m0 = c->m[0];
    const int a0 = pa[0], b0 = pb[0];
    bool eq = a0 == b0;
    bool canon = (unsigned)a0 < (unsigned)m0 && (unsigned)b0 < (unsigned)m0;
    int t = (int)((unsigned)a0 - (unsigned)b0);
    if (t > m0 / 2) t -= m0;
    else if (t <= -((m0 + 1) / 2)) t += m0;
    const long long at = t < 0 ? -(long long)t : (long long)t;
    bool small = true;
#pragma unroll U
    for (int i = 1; i < n; i++) {
        const int m = c->m[i];
        const int ai = pa[i * sa], bi = pb[i * sb];
        eq &= ai == bi;
        canon &= (unsigned)ai < (unsigned)m && (unsigned)bi < (unsigned)m;
        int d = (int)((unsigned)ai - (unsigned)bi);
        d += (d < 0) ? m : 0;
        const int tm = t + ((t < 0) ? m : 0);
        small &= (d == tm) && (2 * at < (long long)m);
    }
    if (eq) return 0;
    if (!near || !canon) return 2;
    if (small) return (t > 0) - (t < 0);
    return near_sign<U>(c, pa, sa, pb, sb);
}

// Lanes with r == 2 hand their pair (a, b) to the groups of the wave, which order it by mixed-radix
// digits; r receives the answer.  All 64 lanes of the wave must call coop_resolve; the group work
// is not inlined, so that its registers do not add to those of the streaming loops.
template <int N>
__device__ __attribute__((noinline)) int coop_resolve_lanes(const DevCtx *c, const int *__restrict__ x, long long smod,
                                                            long long snum, long long a, long long b, int r,
                                                            unsigned long long mask) {
    const int G = Pow2Ceil<N>::v;
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    const int gid = (threadIdx.x & 63) / G;
    const int lane = threadIdx.x & 63;
    while (mask) {
        int src = kth_bit(mask, gid);  // the same on every lane of a group
        long long ea = __shfl(a, src < 0 ? 0 : src, 64);
        long long eb = __shfl(b, src < 0 ? 0 : src, 64);
        mask = drop_low_bits(mask, P);
        int rr = 0;
        if (src >= 0) rr = group_cmp_exact(c, g, x, smod, snum, ea, eb);
        // hand each group's answer back to the lane that asked
        for (int k = 0; k < P; k++) {
            int s = __shfl(src, k * G, 64);
            int ans = __shfl(rr, k * G, 64);
            if (s >= 0 && lane == s) r = ans;
        }
    }
    return r;
}

template <int N>
GRNS_D void coop_resolve(const DevCtx *c, const int *__restrict__ x, long long smod, long long snum, long long a,
                         long long b, int &r) {
    unsigned long long mask = __ballot(r == 2);
    if (mask) r = coop_resolve_lanes<N>(c, x, smod, snum, a, b, r, mask);
}

// Unknown key (x/M may lie within 3E < 2^-28 of 0 or of 1, where F cannot tell the two ends
// apart) made known from the residues in memory, on the rare path of a merge: the key of
// 2^KEY_SHIFT x, H = frac(2^26 x / M) (error at most E, as F's), is below 3/16 + E for x < 3E M and
// above 13/16 - E for x > (1 - 3E) M, and gives x/M = H 2^-26 or 1 - (1 - H) 2^-26 to within
// E 2^-26.  Near 1 the doubles are 2^-53 apart, so those bounds are widened by 2^-52.  H has the
// same ambiguity at its own scale: within 2E of 0 or 1 (x within about 2^-55 M of 0 or M) the key
// stays unknown and the order is exact.  Refined keys lie near 0 or 1, so two keys that do not
// overlap may belong to x near 0 and y near M; the difference certificates of lane_cmp_exact are
// therefore only used for overlapping keys.
template <int N>
GRNS_D void refine_key(const DevCtx *c, const int *__restrict__ x, long long smod, long long snum, MaxCand &e) {
    if (e.idx < 0 || e.lo >= 0) return;
    const int *p = x + e.idx * snum;
    double A = 0.0, B = 0.0;
#pragma unroll 1
    for (int i = 0; i < N; i++) {
        const double dx = (double)p[(long long)i * smod];
        A += fract_d(dx * c->shi[i]);
        B = fma(dx, c->slo[i], B);
    }
    const double H = fract_d(A + B), E = c->err;
    const double s = 1.0 / (double)(1 << KEY_SHIFT);
    // H itself wraps: within 2E of 0 or 1 it cannot tell x < 2E 2^-26 M from x > (1 - 2E 2^-26) M
    if (H > 2.0 * E && H < 0.25) {
        e.lo = fmax(0.0, (H - E) * s);
        e.hi = (H + E) * s;
    } else if (H > 0.75 && H < 1.0 - 2.0 * E) {
        e.lo = 1.0 - ((1.0 - H) + E) * s - 0x1p-52;
        e.hi = fmin(1.0, 1.0 - ((1.0 - H) - E) * s + 0x1p-52);
    }
}

constexpr int LANE_UNROLL = 1;  // the stage-1 loop keeps its occupancy: overlaps there are rare

// Every lane proposes (e -> best): keys decide, then the exact order (lowest index among equal
// values).  All 64 lanes of the wave must call this.
template <int N>
GRNS_D void max_merge(const DevCtx *c, const int *__restrict__ x, long long smod, long long snum, MaxCand &best,
                      const MaxCand &e_in, bool valid) {
    MaxCand e = e_in;
    int d = valid ? max_decide(e, best) : 0;
    if (d == 2 && (e.lo < 0 || best.lo < 0)) {  // rare: a key near 0 or M
        refine_key<N>(c, x, smod, snum, e);
        refine_key<N>(c, x, smod, snum, best);
        d = max_decide(e, best);
    }
    int r = 0;
    if (d == 2)
        r = lane_cmp_exact<LANE_UNROLL>(c, x + e.idx * snum, smod, x + best.idx * snum, smod,
                                        e.lo >= 0 && best.lo >= 0);
    coop_resolve<N>(c, x, smod, snum, e.idx, best.idx, r);
    bool take = (d == 1) || (d == 2 && (r > 0 || (r == 0 && e.idx < best.idx)));
    if (take) best = e;
}

template <int N>
GRNS_D MaxCand max_key(const DevCtx *c, const int (&xr)[N], long long j) {
    double F = key_frac<N>(c, xr);
    MaxCand e;
    e.idx = j;
    double E = c->err;
    if (F > 2.0 * E && F < 1.0 - 2.0 * E) {
        e.lo = F - E;
        e.hi = F + E;
    } else {
        e.lo = -1.0;
        e.hi = 2.0;
    }
    return e;
}



template <int N>
GRNS_D void block_max_reduce(const DevCtx *c, const int *__restrict__ x, long long smod, long long snum,
                             MaxCand &best, MaxCand *sh) {
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int h = BLOCK / 2; h >= 1; h >>= 1) {
        MaxCand other = (threadIdx.x < (unsigned)h) ? sh[threadIdx.x + h] : best;
        MaxCand mine = sh[threadIdx.x];
        __syncthreads();
        MaxCand b = mine;
        max_merge<N>(c, x, smod, snum, b, other, threadIdx.x < (unsigned)h);
        if (threadIdx.x < (unsigned)h) sh[threadIdx.x] = b;
        __syncthreads();
    }
    best = sh[0];
}

// Stage 1: each block reduces a contiguous chunk to one candidate.
template <int N, int MODE>
__global__ __launch_bounds__(BLOCK) void k_max_partial(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                       long long smod, long long snum, long long count,
                                                       long long chunk, MaxCand *__restrict__ part,
                                                       int *__restrict__ part_res) {
    __shared__ MaxCand sh[BLOCK];
    long long start = (long long)blockIdx.x * chunk;
    long long end = start + chunk < count ? start + chunk : count;
    MaxCand best;
    best.lo = -1.0; best.hi = -1.0; best.idx = -1;
    for (long long b0 = start; b0 < end; b0 += BLOCK) {
        long long j = b0 + threadIdx.x;
        MaxCand e;
        e.idx = -1; e.lo = -1.0; e.hi = -1.0;
        bool valid = j < end;
        if (valid) {
            int xr[N];
            load_number<N, MODE>(x, smod, snum, j, xr);
            e = max_key<N>(c, xr, j);
        }
        max_merge<N>(c, x, smod, snum, best, e, valid);
    }
    block_max_reduce<N>(c, x, smod, snum, best, sh);
    if (threadIdx.x == 0) {
        refine_key<N>(c, x, smod, snum, best);  // stage 2 orders by keys where it can
        part[blockIdx.x] = best;
    }
    // a copy of the winner's residues, so that stage 2 reads them from one contiguous line
    if (threadIdx.x < (unsigned)N && best.idx >= 0)
        part_res[(long long)blockIdx.x * N + threadIdx.x] = x[best.idx * snum + (long long)threadIdx.x * smod];
}

// Pivot order for stage 2: the larger lower key bound, then the lower index.
GRNS_D bool pivot_before(const MaxCand &a, const MaxCand &b) {
    if (a.idx < 0) return false;
    if (b.idx < 0) return true;
    return a.lo > b.lo || (a.lo == b.lo && a.idx < b.idx);
}

constexpr int FBLOCK = 1024;             // threads of the stage-2 block
constexpr int MAX_FINAL_PER_THREAD = 8;  // stage 2 takes at most FBLOCK * 8 candidates
constexpr int MAX_PIVOT_ROUNDS = 3;      // then a tree over the live candidates (bounded depth)

// Stage 2 (one block): the exact maximum of the per-block candidates, by rounds.  The live
// candidate with the largest lower key bound is the pivot; every other live candidate is ordered
// against it (keys when they separate, else exactly: in-lane certificates on the residue copies
// stage 1 made, then wave groups).  Candidates below or equal to the pivot leave (equal ones lower
// the tie index); those above stay live for the next round.  When none is above, the pivot's value
// is the maximum and the tie index is the lowest index holding it; when exactly one is above, that
// candidate is the maximum (every other candidate is at most the pivot, and the pivots of earlier
// rounds are lower still) and holds its value alone.  Rounds settle keys that separate the
// candidates; when the keys do not (values within 2^-29 M of 0 or M carry no key, and every
// candidate may survive a round but the pivot), after MAX_PIVOT_ROUNDS rounds the live candidates
// are reduced by a tree of exact comparisons instead, log2(FBLOCK) levels over at most
// MAX_FINAL_PER_THREAD candidates per thread.  Every live candidate is above all earlier pivots, so
// the maximum and every candidate equal to it are still live.
template <int N>
__global__ __launch_bounds__(FBLOCK) void k_max_final(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                      long long smod, long long snum,
                                                      const MaxCand *__restrict__ part,
                                                      const int *__restrict__ part_res, int nparts,
                                                      int *__restrict__ out, long long *__restrict__ out_idx) {
    __shared__ MaxCand sh[FBLOCK];
    __shared__ int shk[FBLOCK];
    __shared__ unsigned long long s_tie;
    __shared__ int s_nabove;
    __shared__ long long s_above_idx;
    const int T = threadIdx.x;
    const int per = (nparts + FBLOCK - 1) / FBLOCK;
    unsigned live = 0;
    for (int t = 0; t < per; t++) {
        int k = T + t * FBLOCK;
        if (k < nparts && part[k].idx >= 0) live |= 1u << t;
    }
    long long win = -1;
    for (int round = 0; round < MAX_PIVOT_ROUNDS; round++) {
        MaxCand b;
        b.idx = -1; b.lo = -2.0; b.hi = -2.0;
        int bk = -1;
        for (int t = 0; t < per; t++)
            if (live >> t & 1) {
                MaxCand e = part[T + t * FBLOCK];
                if (pivot_before(e, b)) { b = e; bk = T + t * FBLOCK; }
            }
        sh[T] = b;
        shk[T] = bk;
        if (T == 0) {
            s_tie = ~0ull;
            s_nabove = 0;
        }
        __syncthreads();
        for (int h = FBLOCK / 2; h >= 1; h >>= 1) {
            if (T < h && pivot_before(sh[T + h], sh[T])) { sh[T] = sh[T + h]; shk[T] = shk[T + h]; }
            __syncthreads();
        }
        const MaxCand P = sh[0];
        const int pk = shk[0];
        if (P.idx < 0) break;  // uniform: every candidate has left
        const int *pres = part_res + (long long)pk * N;
        bool above = false;
        int nabove = 0;
        long long above_idx = -1;
        unsigned long long tie = (unsigned long long)P.idx;
        for (int t = 0; t < per; t++) {  // uniform trip count: the wave groups may be needed
            const bool mine = live >> t & 1;
            const int k = T + t * FBLOCK;
            MaxCand e;
            e.idx = -1; e.lo = -1.0; e.hi = -1.0;
            if (mine) e = part[k];
            int r = -1;
            if (mine && e.idx != P.idx && e.lo > P.hi) r = 1;  // the keys order them
            else if (mine && e.idx != P.idx && e.hi >= P.lo)  // overlapping keys: exact order
                r = lane_cmp_exact<16>(c, part_res + (long long)k * N, 1, pres, 1, e.lo >= 0 && P.lo >= 0);
            coop_resolve<N>(c, x, smod, snum, e.idx, P.idx, r);
            if (mine) {
                if (r > 0) {
                    above = true;
                    nabove++;
                    above_idx = e.idx;
                } else {
                    live &= ~(1u << t);
                    if (r == 0 && e.idx != P.idx && (unsigned long long)e.idx < tie) tie = (unsigned long long)e.idx;
                }
            }
        }
        atomicMin(&s_tie, tie);
        if (nabove) {
            atomicAdd(&s_nabove, nabove);
            s_above_idx = above_idx;  // read only when a single candidate is above (one writer)
        }
        const bool any_above = __syncthreads_or(above);
        if (!any_above) {
            win = (long long)s_tie;
            break;
        }
        if (s_nabove == 1) {
            win = s_above_idx;
            break;
        }
        __syncthreads();  // s_tie, s_nabove and sh are rewritten by the next round
    }
    if (win < 0 && __syncthreads_or(live != 0)) {  // uniform: the rounds left candidates live
        MaxCand best;
        best.idx = -1; best.lo = -1.0; best.hi = -1.0;
        for (int t = 0; t < per; t++) {  // uniform trip count: max_merge runs wave groups
            const bool mine = live >> t & 1;
            MaxCand e;
            e.idx = -1; e.lo = -1.0; e.hi = -1.0;
            if (mine) e = part[T + t * FBLOCK];
            max_merge<N>(c, x, smod, snum, best, e, mine);
        }
        sh[T] = best;
        __syncthreads();
        for (int h = FBLOCK / 2; h >= 1; h >>= 1) {
            MaxCand other = (T < h) ? sh[T + h] : best;
            MaxCand mine = sh[T];
            __syncthreads();
            max_merge<N>(c, x, smod, snum, mine, other, T < h);
            if (T < h) sh[T] = mine;
            __syncthreads();
        }
        win = sh[0].idx;
    }
    if (T < N && win >= 0) out[T] = x[win * snum + (long long)T * smod];
    if (T == 0) *out_idx = win;
}

// ---------------------------------------------------------------- arg-max by mixed-radix digits
// The mixed-radix pipeline of tests/test_perf_rnsmax.cu:95-156 (rns_max_mrc): the digits of every
// number, then the largest digit vector, compared from the most significant digit down (mrs_cmp,
// src/mrc.cuh:13-22).  Here one lane converts one number in registers and keeps its best digit
// vector in a column of LDS; the value of a number is that of its residues taken mod m_i, and
// equal values go to the lower index, so the result is the one grns_max returns.

// Mixed-radix digits of the canonical number held by this lane (mrc, src/mrc.cuh:31-44): step j
// turns residue i > j into (x_i - d_j) m_j^-1 mod m_i.  Each product is reduced exactly in FP64:
// with d in (-m_i, m_i) and the inverse in [0, m_i), hi + lo = d * inv exactly (|d * inv| < 2^62);
// floor(hi / m_i - 2^-19), computed as one fma with RN(1/m_i), is floor(d * inv / m_i) or one less
// (its error is below 2^-20), so r = d * inv - q m_i is an exact integer in [0, 2 m_i) and one
// conditional subtraction (an unsigned minimum) ends the reduction.
// One chunk of step J: residues I0 .. I0 + 3 (those below N).  The constants of the chunk are
// loaded (scalar loads) once the chunk before the previous one is computed (d2): one chunk of
// look-ahead rather than the whole table in scalar registers.
template <int N, int J, int I0>
GRNS_D void lane_mrc_chunk(const DevCtx *c, const double *tab, int (&mr)[N], int &d1, int &d2) {
    const DevCtx *cc = reload_after(c, d2);
    cdouble *inv = const_view(reload_after(tab, d2) + J * N + I0);
    cdouble *dm = const_view(cc->dm + I0);
    cdouble *dinv = const_view(cc->dinv + I0);
    cint *mi = const_view(cc->m + I0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (I0 + k < N) {
            const double d = (double)(mr[I0 + k] - mr[J]);
            const double hi = d * inv[k];
            const double lo = fma(d, inv[k], -hi);
            const double q = floor(fma(hi, dinv[k], -0x1p-19));
            const unsigned r = (unsigned)(fma(-q, dm[k], hi) + lo);
            const unsigned s = r - (unsigned)mi[k];
            mr[I0 + k] = (int)(s < r ? s : r);
        }
    }
    d2 = d1;
    d1 = mr[I0 + 3 < N ? I0 + 3 : N - 1];
}
template <int N, int J, int... C>
GRNS_D void lane_mrc_step(const DevCtx *c, const double *tab, int (&mr)[N], int &d1, int &d2,
                          std::integer_sequence<int, C...>) {
    (lane_mrc_chunk<N, J, J + 1 + 4 * C>(c, tab, mr, d1, d2), ...);
}
template <int N, int... J>
GRNS_D void lane_mrc_steps(const DevCtx *c, const double *tab, int (&mr)[N], std::integer_sequence<int, J...>) {
    int d1 = 0, d2 = 0;
    (lane_mrc_step<N, J>(c, tab, mr, d1, d2, std::make_integer_sequence<int, (N - 1 - J + 3) / 4>{}), ...);
}
// Every step, as straight-line code (all register indexes are constants).  tab: the context's
// inverse table as doubles (c->mrc_inv_d), read once by the caller.
template <int N>
GRNS_D void lane_mrc(const DevCtx *c, const double *tab, int (&mr)[N]) {
    lane_mrc_steps<N>(c, tab, mr, std::make_integer_sequence<int, N - 1>{});
}

// Order of the digit vector `a` (this lane's registers) against column `col` of `bd` below the top
// digit, which the caller found equal: 1 above, -1 below, 0 equal.
template <int N>
GRNS_D int mrc_cmp_lower(const int (&a)[N], const int (*bd)[BLOCK], int col) {
    int r = 0;
#pragma unroll
    for (int k = N - 2; k >= 0; k--) {
        const int b = bd[k][col];
        if (r == 0 && a[k] != b) r = a[k] > b ? 1 : -1;
    }
    return r;
}

// Order of columns a and b of `bd` (mrs_cmp).
template <int N>
GRNS_D int mrc_cmp_cols(const int (*bd)[BLOCK], int a, int b) {
#pragma unroll 1
    for (int k = N - 1; k >= 0; k--) {
        const int u = bd[k][a], v = bd[k][b];
        if (u != v) return u > v ? 1 : -1;
    }
    return 0;
}

// Stage 1: each block reduces a contiguous chunk to one candidate (index and digits).  Lanes walk
// their numbers in increasing index order, so a lane keeps its incumbent on equal values.  The
// load form is chosen per number (LD_AUTO): the kernel is VALU-bound and one instance per n keeps
// the straight-line conversion code compiled once.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_max_mrc_partial(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                           long long smod, long long snum, long long count,
                                                           long long chunk, long long *__restrict__ part_idx,
                                                           int *__restrict__ part_dig) {
    __shared__ int bd[N][BLOCK];
    __shared__ long long sidx[BLOCK];
    const int t = threadIdx.x;
    const long long start = (long long)blockIdx.x * chunk;
    const long long end = start + chunk < count ? start + chunk : count;
    const double *tab = c->mrc_inv_d;
    long long best = -1;
    int top = 0;
    for (long long b0 = start; b0 < end; b0 += BLOCK) {
        const long long j = b0 + t;
        if (j >= end) continue;
        int xr[N];
        load_number<N, LD_AUTO>(x, smod, snum, j, xr);
        bool canon = true;
        cint *cm = const_view(reload_point(c)->m);
#pragma unroll
        for (int i = 0; i < N; i++) canon &= (unsigned)xr[i] < (unsigned)cm[i];
        if (!canon) {  // residues taken mod m_i (constants through views: no addresses hoisted)
#pragma unroll
            for (int i0 = 0; i0 < N; i0 += 8) {
                const DevCtx *cc = reload_point(c);
                cint *m8 = const_view(cc->m + i0);
                cdouble *v8 = const_view(cc->dinv + i0);
#pragma unroll
                for (int k = 0; k < 8 && i0 + k < N; k++) xr[i0 + k] = canon_res(xr[i0 + k], m8[k], v8[k]);
            }
        }
        lane_mrc<N>(c, tab, xr);
        int r = 1;
        if (best >= 0) {
            r = (xr[N - 1] > top) - (xr[N - 1] < top);
            if (r == 0) r = mrc_cmp_lower<N>(xr, bd, t);
        }
        if (r > 0) {
            best = j;
            top = xr[N - 1];
#pragma unroll
            for (int k = 0; k < N; k++) bd[k][t] = xr[k];
        }
    }
    sidx[t] = best;
    __syncthreads();
    for (int h = BLOCK / 2; h >= 1; h >>= 1) {
        if (t < h) {
            const long long a = sidx[t], b = sidx[t + h];
            bool take = false;
            if (b >= 0) {
                if (a < 0) {
                    take = true;
                } else {
                    const int r = mrc_cmp_cols<N>(bd, t + h, t);
                    take = r > 0 || (r == 0 && b < a);
                }
            }
            if (take) {
#pragma unroll 1
                for (int k = 0; k < N; k++) bd[k][t] = bd[k][t + h];
                sidx[t] = b;
            }
        }
        __syncthreads();
    }
    if (t == 0) part_idx[blockIdx.x] = sidx[0];
    if (t < N) part_dig[(long long)blockIdx.x * N + t] = bd[t][0];
}

// Whether candidate a of stage 1 comes before candidate b: the larger digit vector, then the
// lower index.  Digits are compared 16 at a time from the top, each chunk's loads issued together
// (near-equal candidates share most digits: a digit-by-digit walk would be a chain of dependent
// loads).
template <int N>
GRNS_D bool mrc_part_before(const long long *__restrict__ part_idx, const int *__restrict__ part_dig, int a, int b) {
    const int *da = part_dig + (long long)a * N;
    const int *db = part_dig + (long long)b * N;
    constexpr int C = N < 16 ? N : 16;
    static_assert(N <= 16 || N % 16 == 0, "chunks of 16 digits must cover every digit");
#pragma unroll 1
    for (int k0 = N - C; k0 >= 0; k0 -= C) {
        int u[C], v[C];
#pragma unroll
        for (int i = 0; i < C; i++) {
            u[i] = da[k0 + i];
            v[i] = db[k0 + i];
        }
        int r = 0;
#pragma unroll
        for (int i = C - 1; i >= 0; i--)
            if (r == 0 && u[i] != v[i]) r = u[i] > v[i] ? 1 : -1;
        if (r != 0) return r > 0;
    }
    return part_idx[a] < part_idx[b];
}

// Stage 2 (one block): the best of the stage-1 candidates; the winner's residues are copied out.
// The top digit of every candidate is first loaded into LDS (independent loads); comparisons read
// it there and go to the full digit rows only when the top digits are equal.
constexpr int MRC_FINAL_MAX = FBLOCK * MAX_FINAL_PER_THREAD;  // candidates: the arg-max stage-1 grid's bound
template <int N>
GRNS_D bool mrc_cand_before(const long long *__restrict__ part_idx, const int *__restrict__ part_dig,
                            const int *top, int a, int b) {
    if (top[a] != top[b]) return top[a] > top[b];
    return mrc_part_before<N>(part_idx, part_dig, a, b);
}
template <int N>
__global__ __launch_bounds__(BLOCK) void k_max_mrc_final(const int *__restrict__ x, long long smod, long long snum,
                                                         const long long *__restrict__ part_idx,
                                                         const int *__restrict__ part_dig, int nparts,
                                                         int *__restrict__ out, long long *__restrict__ out_idx) {
    __shared__ int top[MRC_FINAL_MAX];
    __shared__ int sk[BLOCK];
    const int t = threadIdx.x;
    for (int k = t; k < nparts; k += BLOCK)
        top[k] = part_idx[k] >= 0 ? part_dig[(long long)k * N + N - 1] : 0;
    __syncthreads();
    int bk = -1;
    for (int k = t; k < nparts; k += BLOCK) {
        if (part_idx[k] < 0) continue;
        if (bk < 0 || mrc_cand_before<N>(part_idx, part_dig, top, k, bk)) bk = k;
    }
    sk[t] = bk;
    __syncthreads();
    for (int h = BLOCK / 2; h >= 1; h >>= 1) {
        if (t < h) {
            const int a = sk[t], b = sk[t + h];
            if (b >= 0 && (a < 0 || mrc_cand_before<N>(part_idx, part_dig, top, b, a))) sk[t] = b;
        }
        __syncthreads();
    }
    const int w = sk[0];
    const long long win = w >= 0 ? part_idx[w] : -1;
    if (t < N && win >= 0) out[t] = x[win * snum + (long long)t * smod];
    if (t == 0) *out_idx = win;
}

// ---------------------------------------------------------------- division (group per number)
// rns_div (FAST=false, src/rnsdiv.cuh:19-61) and rns_div_fast (FAST=true, :67-114).
template <bool FAST>
GRNS_D void group_div(const DevCtx *c, const Group &g, int xi, int di, int &q, int &r, int &status) {
    const bool act = g.gl < g.L;
    const int mi = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const int n = c->n;
    q = 0;
    r = xi;
    interval_t deval = group_eval<false>(c, g, di, status);
    interval_t reval = group_eval<false>(c, g, r, status);
    long long it = 0;
    while (er_ucmp(reval.low, deval.upp) >= 0) {
        er_t flq = er_div_rd(reval.low, deval.upp);
        int pq = 0;
        if (++it > 20000) { status |= GST_DIVLOOP; return; }
        if (FAST) {
            bool got = false;
            while (flq.exp >= 0 && flq.frac > 0) {
                if (flq.exp > c->log2M) { status |= GST_DIVLOOP; return; }
                int p2 = act ? c->pow2[(size_t)flq.exp * n + g.gl] : 0;
                pq = addmod_recip(pq, p2, mi, dinv);
                flq.frac = dsub_rd(flq.frac, 1.0);
                flq = er_adjust(flq);
                got = true;
            }
            if (!got) { status |= GST_DIVLOOP; return; }
        } else {
            if (flq.exp < 0 || flq.exp > c->log2M) { status |= GST_DIVLOOP; return; }
            pq = act ? c->pow2[(size_t)flq.exp * n + g.gl] : 0;
        }
        q = addmod_recip(q, pq, mi, dinv);
        pq = mulmod_recip(pq, di, mi, dinv);
        r = psubmod(r, pq, mi, dinv);
        reval = group_eval<false>(c, g, r, status);
    }
    if (er_ucmp(reval.upp, deval.low) >= 0) {
        int dmr = group_mrc(c, g, di);
        int rmr = group_mrc(c, g, r);
        // r >= d unless the most significant differing digit of r is smaller
        bool rged = group_mrs_cmp(g, rmr, dmr) >= 0;
        if (rged) {
            q = addmod_recip(q, 1, mi, dinv);
            r = psubmod(r, di, mi, dinv);
        }
    }
}

// Whether the pair can go through the word division: canonical residues and d != 0 (the word
// form also needs a group of at least c->words lanes).
GRNS_D bool div_words_ok(const DevCtx *c, const Group &g, int xi, int di) {
    const bool act = g.gl < g.L;
    const unsigned m = act ? (unsigned)c->m[g.gl] : 1u;
    return group_all(g, !act || ((unsigned)xi < m && (unsigned)di < m)) && group_any(g, act && di != 0);
}

// Marker of a division left to the rare pass after a queue overflow (written to q's first residue;
// the word division never writes it).
constexpr int DIV_MARK = (int)0x80000000u;

// Division of the pairs div_words_ok accepts (group of GD lanes per pair); the others go to the
// rare queue, and k_div_rare runs the reference loop for them.
template <int GD>
__global__ __launch_bounds__(BLOCK) void k_div_words(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                     const int *__restrict__ d, long long smod, long long snum,
                                                     long long count, int *__restrict__ q, int *__restrict__ r,
                                                     RareQ rq) {
    const Group g = make_group(GD, c->n);
    constexpr int P = 64 / GD;
    const long long wave = ((long long)xcd_block() * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / GD;
        const bool valid = j < count;
        bool ok = false;
        if (valid) {
            const bool act = g.gl < g.L;
            const long long off = j * snum + (long long)g.gl * smod;
            const int xi = act ? x[off] : 0;
            const int di = act ? d[off] : 0;
            ok = div_words_ok(c, g, xi, di);
            if (ok) {
                int qq, rr;
                group_div_words<GD>(c, g, xi, di, qq, rr);
                if (act) {
                    q[off] = qq;
                    r[off] = rr;
                }
            }
        }
        const bool rare = valid && !ok && g.gl == 0;
        if (!rare_push(rq, rare, j)) q[j * snum] = DIV_MARK;
    }
}

// Rare pass of the division: the reference loop, one wave group per queued pair (and, after a
// queue overflow, per marked pair).
template <bool FAST>
__global__ __launch_bounds__(BLOCK) void k_div_rare(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                                    const int *__restrict__ d, long long smod, long long snum,
                                                    long long count, int *__restrict__ q, int *__restrict__ r,
                                                    int *__restrict__ status, RareQ rq) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    const long long gid = (((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6) * P + (threadIdx.x & 63) / G;
    const long long ngroups = ((long long)gridDim.x * BLOCK >> 6) * P;
    __shared__ RareView v;
    rare_view(rq, v);
    const long long nq = v.pre[RARE_SUBQ];
    const long long total = nq + (rq.ctr[2] ? count : 0);
    int st = 0;
    for (long long k = gid; k < total; k += ngroups) {
        const int *unused;
        const long long j = rare_item(rq, v, k, unused);
        if (k >= nq && q[j * snum] != DIV_MARK) continue;
        const long long off = j * snum + (long long)g.gl * smod;
        const int xi = g.gl < g.L ? x[off] : 0;
        const int di = g.gl < g.L ? d[off] : 0;
        int qq, rr;
        group_div<FAST>(c, g, xi, di, qq, rr, st);
        if (g.gl < g.L) {
            q[off] = qq;
            r[off] = rr;
        }
    }
    if (g.gl == 0 && st) atomicOr(status, st);
}

// The reference loop for every pair.
template <bool FAST>
__global__ __launch_bounds__(BLOCK) void k_div(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                               const int *__restrict__ d, long long smod, long long snum,
                                               long long count, int *__restrict__ q, int *__restrict__ r,
                                               int *__restrict__ status) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        long long j = j0 + (threadIdx.x & 63) / G;
        if (j < count) {
            long long off = j * snum + (long long)g.gl * smod;
            int xi = g.gl < g.L ? x[off] : 0;
            int di = g.gl < g.L ? d[off] : 0;
            int qq, rr, st = 0;
            group_div<FAST>(c, g, xi, di, qq, rr, st);
            if (g.gl < g.L) {
                q[off] = qq;
                r[off] = rr;
            }
            if (g.gl == 0 && st) atomicOr(status, st);
        }
    }
}

// ---------------------------------------------------------------- binary <-> RNS conversion
// rns_from_binary (src/rnsbase.cuh:253-261): lane i of a group forms |x| mod m_i.  The words are
// taken in chunks of kc = min(GD, log2M / 32) (word k of a chunk in lane k): a chunk's residue is
// sum_k word_k (2^(32 k) mod m_i) from RNS_POW2, and the chunks, most significant first, are
// combined by Horner's rule with the factor 2^(32 kc) mod m_i.  Sets whose M has fewer than 64 bits
// take Horner's rule word by word.  Either way the result is |x| mod m_i, which is unique.
template <int GD>
__global__ __launch_bounds__(BLOCK) void k_from_binary(const DevCtx *__restrict__ c, const unsigned *__restrict__ w,
                                                       long long wst, long long wsn, int nwords,
                                                       const int *__restrict__ sign, long long count,
                                                       int *__restrict__ x, long long smod, long long snum) {
    const Group g = make_group(GD, c->n);
    constexpr int P = 64 / GD;
    const long long wave = ((long long)xcd_block() * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    const bool act = g.gl < g.L;
    const int m = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const long long c32 = act ? c->c32[g.gl] : 0;
    const int kc = (c->log2M >> 5) < GD ? (c->log2M >> 5) : GD;
    const long long F = act && kc >= 2 ? c->pow2[(size_t)(32 * kc) * c->n + g.gl] : 0;  // 2^(32 kc) mod m
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / GD;
        if (j >= count) continue;
        long long acc = 0;
        if (kc >= 2) {
            const int top = (nwords - 1) / kc;
            for (int ch = top; ch >= 0; ch--) {  // chunks of kc words, most significant first
                const int k0 = ch * kc;
                const int kn = nwords - k0 < kc ? nwords - k0 : kc;
                const unsigned wk = g.gl < kn ? w[j * wsn + (long long)(k0 + g.gl) * wst] : 0u;
                const int res = words_residue<GD>(c, g, wk, kn);
                if (act) acc = rem_trunc(acc * F + res, m, dinv);
            }
        } else {
            const int top = (nwords - 1) / GD;
            for (int ch = top; ch >= 0; ch--) {  // chunks of GD words, most significant first
                const int k0 = ch * GD;
                const int kn = nwords - k0 < GD ? nwords - k0 : GD;
                const unsigned wk = g.gl < kn ? w[j * wsn + (long long)(k0 + g.gl) * wst] : 0u;
                for (int k = kn - 1; k >= 0; k--) {
                    const unsigned v = gword<GD>(g, wk, k);
                    acc = rem_trunc(acc * c32 + (long long)v, m, dinv);
                }
            }
        }
        if (act) {
            if (sign && sign[j] && acc != 0) acc = m - acc;
            x[j * snum + (long long)g.gl * smod] = (int)acc;
        }
    }
}

// rns_to_binary (src/rnsbase.cuh:277-287): each residue reduced as the reference's mpz_mul_ui
// sees it (an int cast to unsigned long: x_i + 2^64 when negative), then the words of the CRT value.
template <int GD>
__global__ __launch_bounds__(BLOCK) void k_to_binary(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                     long long smod, long long snum, long long count,
                                                     unsigned *__restrict__ w, long long wst, long long wsn) {
    const Group g = make_group(GD, c->n);
    constexpr int P = 64 / GD;
    const long long wave = ((long long)xcd_block() * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    const bool act = g.gl < g.L;
    const int m = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const long long c32 = act ? c->c32[g.gl] : 0;
    const long long p64 = act ? rem_trunc(c32 * c32, m, dinv) : 0;  // 2^64 mod m
    const int W = c->words - 1;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / GD;
        if (j >= count) continue;
        int xi = 0;
        if (act) {
            const long long v = x[j * snum + (long long)g.gl * smod];
            long long r = rem_trunc(v, m, dinv);
            if (v < 0) r = rem_trunc(r + p64 + m, m, dinv);
            xi = (int)r;
        }
        const unsigned X = group_words<GD>(c, g, xi);
        if (g.gl < W) w[j * wsn + (long long)g.gl * wst] = X;
    }
}

// ---------------------------------------------------------------- mixed-radix digits
// mrc (src/mrc.cuh:31-44): the digits of each number, digit k in group lane k (pipelined over
// the lanes, table loads batched 8 steps at a time).
static __global__ __launch_bounds__(BLOCK) void k_mrc(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                               long long smod, long long snum, long long count,
                                               int *__restrict__ d, long long dsm, long long dsn) {
    const Group g = make_group(G, c->n);
    const int P = 64 / G;
    const long long wave = ((long long)xcd_block() * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / G;
        if (j >= count) continue;
        const bool act = g.gl < g.L;
        const int xi = act ? x[j * snum + (long long)g.gl * smod] : 0;
        const int digit = group_mrc<8>(c, g, xi);
        if (act) d[j * dsn + (long long)g.gl * dsm] = digit;
    }
}

// The numbers of the lanes in `mask` converted by the wave groups with the reference's steps
// (group_mrc), which define the digits of residues outside [0, m_i).  Not inlined: its registers
// stay out of the per-lane kernel.
template <int N>
__device__ __attribute__((noinline)) void coop_mrc_lanes(const DevCtx *c, const int *__restrict__ x, long long smod,
                                                         long long snum, long long j, unsigned long long mask,
                                                         int *__restrict__ d, long long dsm, long long dsn) {
    const int G = Pow2Ceil<N>::v;
    Group g = make_group(G, N);
    const int P = 64 / G;
    const int gid = (threadIdx.x & 63) / G;
    while (mask) {
        const int src = kth_bit(mask, gid);
        const long long jl = __shfl(j, src < 0 ? 0 : src, 64);
        mask = drop_low_bits(mask, P);
        if (src >= 0) {
            const bool act = g.gl < g.L;
            const int xi = act ? x[jl * snum + (long long)g.gl * smod] : 0;
            const int digit = group_mrc<8>(c, g, xi);
            if (act) d[jl * dsn + (long long)g.gl * dsm] = digit;
        }
    }
}

// Batched mixed-radix conversion, one number per lane: canonical numbers convert in their lane
// (lane_mrc, exact); a number with a residue outside [0, m_i) goes to the wave's groups.  Load form
// per number (LD_AUTO), as in k_max_mrc_partial.
template <int N>
__global__ __launch_bounds__(BLOCK) void k_mrc_lane(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                    long long smod, long long snum, long long count,
                                                    int *__restrict__ d, long long dsm, long long dsn) {
    const double *tab = c->mrc_inv_d;
    for (long long base = (long long)blockIdx.x * BLOCK; base < count; base += (long long)gridDim.x * BLOCK) {
        const long long j = base + threadIdx.x;
        bool coop = false;
        if (j < count) {
            int xr[N];
            load_number<N, LD_AUTO>(x, smod, snum, j, xr);
            bool canon = true;
            cint *cm = const_view(reload_point(c)->m);
#pragma unroll
            for (int i = 0; i < N; i++) canon &= (unsigned)xr[i] < (unsigned)cm[i];
            if (canon) {
                lane_mrc<N>(c, tab, xr);
                if (dsn == 1 && j < (1ll << 29)) {
                    store_number_mm<N>(d, dsm, j, xr);
                } else {
#pragma unroll
                    for (int i = 0; i < N; i++) d[j * dsn + (long long)i * dsm] = xr[i];
                }
            } else {
                coop = true;
            }
        }
        const unsigned long long mask = __ballot(coop);
        if (mask) coop_mrc_lanes<N>(c, x, smod, snum, j, mask, d, dsm, dsn);
    }
}

// ---------------------------------------------------------------- scaling by 2^D (group per number)
// rns_rank_compute / _fast, src/rnsscal.cuh:17-62
GRNS_D int group_rank(const DevCtx *c, const Group &g, int xi, int s) {
    const bool act = g.gl < g.L;
    double fl = act ? dmul_rd((double)s, c->rd[g.gl]) : 0.0;
    double fu = act ? dmul_ru((double)s, c->ru[g.gl]) : 0.0;
    double suml = group_psum<false>(g, fl);
    double sumu = group_psum<true>(g, fu);
    unsigned whl = to_uint(suml), whu = to_uint(sumu);
    if (whl == whu) return (int)whl;
    int d = group_mrc(c, g, xi);
    int mr = bcast_i(d, g.base + g.L - 1);
    return mr == 0 ? (int)whu : (int)whl;
}
GRNS_D int group_rank_fast(const DevCtx *c, const Group &g, int s) {
    const bool act = g.gl < g.L;
    double fu = act ? dmul_ru((double)s, c->ru[g.gl]) : 0.0;
    return (int)group_psum<true>(g, fu);
}
// make_scaling_step, src/rnsscal.cuh:67-85 (j in [1, 30], pow2j = 2^j)
GRNS_D int group_scaling_step(const DevCtx *c, const Group &g, int k, int j, int xi, int ci) {
    const bool act = g.gl < g.L;
    const int mi = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const long long pow2j = 1ll << j;
    long long term = 0;
    if (act) {
        long long prod = (long long)c->part_pow2_res[j - 1][g.gl] * (long long)ci;
        // C's truncating remainder by a power of two
        term = prod >= 0 ? (prod & (pow2j - 1)) : -((-prod) & (pow2j - 1));
    }
    long long residue = group_sum_ll(g, term);
    long long temp = (long long)k * (long long)c->prod_pow2_res[j - 1];
    residue = (residue - temp) % pow2j;
    if (residue < 0) residue += pow2j;
    int multiple = act ? (int)(residue % mi) : 0;
    multiple = act ? psubmod(xi, multiple, mi, dinv) : 0;
    return act ? mulmod_recip(multiple, c->pow2_inv[j - 1][g.gl], mi, dinv) : 0;
}
// rns_scale2pow, src/rnsscal.cuh:93-119
GRNS_D int group_scale2pow(const DevCtx *c, const Group &g, int xi, unsigned D) {
    const bool act = g.gl < g.L;
    const int mi = act ? c->m[g.gl] : 1;
    const double dinv = act ? c->dinv[g.gl] : 1.0;
    const int w = act ? c->w[g.gl] : 0;
    int result = xi;
    int t = (int)(D / THRESH);
    int k;
    if (t > 0) {
        int ci = act ? mulmod_recip(xi, w, mi, dinv) : 0;
        k = group_rank(c, g, xi, ci);
        result = group_scaling_step(c, g, k, THRESH, xi, ci);
        t -= 1;
    }
    while (t > 0) {
        // a step maps residues that are all 0 to all 0 (c = 0, rank 0, remainder 0): once there, the
        // remaining steps of the reference change nothing, which bounds the loop for any D
        if (group_all(g, !act || result == 0)) break;
        int ci = act ? mulmod_recip(result, w, mi, dinv) : 0;
        k = group_rank_fast(c, g, ci);
        result = group_scaling_step(c, g, k, THRESH, result, ci);
        t -= 1;
    }
    unsigned d = D % THRESH;
    if (d > 0) {
        int ci = act ? mulmod_recip(result, w, mi, dinv) : 0;
        k = d < D ? group_rank_fast(c, g, ci) : group_rank(c, g, result, ci);
        result = group_scaling_step(c, g, k, (int)d, result, ci);
    }
    return result;
}

// Scaling of canonical numbers by the replayed word form; the others (non-canonical residues, a
// replay that meets an int32 wrap) go to the rare queue for the reference steps.
template <int GD>
__global__ __launch_bounds__(BLOCK) void k_scale2pow_words(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                           long long smod, long long snum,
                                                           const unsigned *__restrict__ D, long long count,
                                                           int *__restrict__ y, RareQ rq) {
    const Group g = make_group(GD, c->n);
    constexpr int P = 64 / GD;
    const long long wave = ((long long)xcd_block() * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / GD;
        const bool valid = j < count;
        bool ok = false;
        if (valid) {
            const bool act = g.gl < g.L;
            const long long off = j * snum + (long long)g.gl * smod;
            const int xi = act ? x[off] : 0;
            if (group_all(g, !act || (unsigned)xi < (unsigned)c->m[g.gl])) {
                const int v = group_scale2pow_words<GD>(c, g, xi, D[j], ok);
                if (ok && act) y[off] = v;
            }
        }
        const bool rare = valid && !ok && g.gl == 0;
        if (!rare_push(rq, rare, j)) y[j * snum] = DIV_MARK;
    }
}

// Rare pass of the scaling: the reference steps, one wave group per queued number (and, after a
// queue overflow, per marked number).
static __global__ __launch_bounds__(BLOCK) void k_scale2pow_rare(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                                          long long smod, long long snum, const unsigned *__restrict__ D,
                                                          long long count, int *__restrict__ y, RareQ rq) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    const long long gid = (((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6) * P + (threadIdx.x & 63) / G;
    const long long ngroups = ((long long)gridDim.x * BLOCK >> 6) * P;
    __shared__ RareView v;
    rare_view(rq, v);
    const long long nq = v.pre[RARE_SUBQ];
    const long long total = nq + (rq.ctr[2] ? count : 0);
    for (long long k = gid; k < total; k += ngroups) {
        const int *unused;
        const long long j = rare_item(rq, v, k, unused);
        if (k >= nq && y[j * snum] != DIV_MARK) continue;
        const long long off = j * snum + (long long)g.gl * smod;
        const int xi = g.gl < g.L ? x[off] : 0;
        const int v = group_scale2pow(c, g, xi, D[j]);
        if (g.gl < g.L) y[off] = v;
    }
}

static __global__ __launch_bounds__(BLOCK) void k_scale2pow(const DevCtx *__restrict__ c, int G, const int *__restrict__ x,
                                                     long long smod, long long snum, const unsigned *__restrict__ D,
                                                     long long count, int *__restrict__ y) {
    Group g = make_group(G, c->n);
    const int P = 64 / G;
    long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        long long j = j0 + (threadIdx.x & 63) / G;
        if (j < count) {
            long long off = j * snum + (long long)g.gl * smod;
            int xi = g.gl < g.L ? x[off] : 0;
            int v = group_scale2pow(c, g, xi, D[j]);
            if (g.gl < g.L) y[off] = v;
        }
    }
}

// ---------------------------------------------------------------- mpint add / sub / mul
// mp_int_t split into arrays: digits (modulus-major), sign, eval[2].
// mpint_check_overflow, src/mpint.cuh:142-152 -> 0 none, 1 overflow, 2 possible overflow
GRNS_D unsigned char check_overflow(const DevCtx *c, const er_t &lower, const er_t &upper) {
    if (upper.exp < -1 || er_ucmp(upper, c->inv_low) <= 0) return 0;
    if (er_ucmp(lower, c->inv_upp) == 1) return 1;
    return 2;
}

GRNS_D er_t load_er(const er_t *p) {
    double2 v = *reinterpret_cast<const double2 *>(p);
    er_t r;
    r.frac = v.x;
    r.exp = __double_as_longlong(v.y);
    return r;
}
GRNS_D void store_er(er_t *p, const er_t &v) {
    *reinterpret_cast<double2 *>(p) = make_double2(v.frac, __longlong_as_double(v.exp));
}

// mpint operands and results as arrays (digits modulus-major with leading dimension ld).
struct MpintArgs {
    const int *xd;
    const int *xs;
    const er_t *xe;
    const int *yd;
    const int *ys;
    const er_t *ye;
    int *rd;
    int *rs;
    er_t *re;
    unsigned char *ovf;
    long long ld;
};

// Interval part of mpint_add / mpint_sub (src/mpint.cuh:160-262, OP 0/1) and mpint_mul (:270-278,
// OP 2).  For add/sub, `minus` tells whether the digits are to be negated.
template <int OP>
GRNS_D void mpint_bounds(const DevCtx *c, int sx, int sy, er_t x0, er_t x1, er_t y0, er_t y1, er_t &r0, er_t &r1,
                         int &minus) {
    if constexpr (OP == 2) {
        r0 = er_md_rd(x0, y0, c->unit_upp);
        r1 = er_md_ru(x1, y1, c->unit_low);
        minus = 0;
    } else {
        int alpha = 1 - 2 * sx, beta = 1 - 2 * sy;
        x0.frac *= alpha; x1.frac *= alpha;
        y0.frac *= beta; y1.frac *= beta;
        er_t ea = sx ? x1 : x0, eb = sy ? y1 : y0;
        er_t fa = sx ? x0 : x1, fb = sy ? y0 : y1;
        r0 = er_addsub<false, false>(ea, eb);
        r1 = er_addsub<false, true>(fa, fb);
        minus = r0.frac < 0 && r1.frac < 0;
        if (minus) {
            er_t tmp = r0;
            r0.frac = -1 * r1.frac;
            r0.exp = r1.exp;
            r1.frac = -1 * tmp.frac;
            r1.exp = tmp.exp;
        }
    }
}

// Digit of modulus i (reciprocal reductions of the CPU path).
template <int OP>
GRNS_D int mpint_digit(int a, int b, int sx, int sy, int m, double dinv) {
    if constexpr (OP == 2) {
        return mulmod_recip(a, b, m, dinv);
    } else {
        int res = addmod_recip((1 - 2 * sx) * a, (1 - 2 * sy) * b, m, dinv);
        return res < 0 ? res + m : res;
    }
}

// Thread per number: element j (j < count).  OP: 0 add, 1 sub, 2 mul.  MM: the caller guarantees
// count <= 2^29, so that every digit row is addressed by a 32-bit lane offset from a scalar row
// pointer (no per-digit 64-bit row offsets held across the caller's loop).
template <int N, int OP, bool MM = false>
GRNS_D void mpint_step(const DevCtx *__restrict__ c, const MpintArgs &A, long long j) {
    int a[N], b[N], r[N];
    load_number<N, MM ? LD_MM : LD_AUTO>(A.xd, A.ld, 1, j, a);
    load_number<N, MM ? LD_MM : LD_AUTO>(A.yd, A.ld, 1, j, b);
    int sx = A.xs[j], sy = A.ys[j];
    if (OP == 1) sy ^= 1;
    er_t x0 = load_er(A.xe + 2 * j), x1 = load_er(A.xe + 2 * j + 1);
    er_t y0 = load_er(A.ye + 2 * j), y1 = load_er(A.ye + 2 * j + 1);
    er_t r0, r1;
    int minus;
    mpint_bounds<OP>(c, sx, sy, x0, x1, y0, y1, r0, r1, minus);
    int nz = 0;
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += 8) {
        // 8 moduli of constants at a time (scalar loads here, not hoisted out of the callers' loops)
        const DevCtx *cc = reload_point(c);
        cint *m8 = const_view(cc->m + i0);
        cdouble *v8 = const_view(cc->dinv + i0);
#pragma unroll
        for (int k = 0; k < 8 && i0 + k < N; k++) {
            const int i = i0 + k;
            r[i] = mpint_digit<OP>(a[i], b[i], sx, sy, m8[k], v8[k]);
            if (OP != 2 && minus) r[i] = r[i] == 0 ? 0 : (m8[k] - r[i]);
            nz |= r[i];
        }
    }
    int sr = (OP == 2) ? ((nz == 0) ? 0 : (sx ^ sy)) : minus;
    if (MM || j < (1ll << 29)) {
        store_number_mm<N>(A.rd, A.ld, j, r);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) __builtin_nontemporal_store(r[i], A.rd + (long long)i * A.ld + j);
    }
    A.rs[j] = sr;
    store_er(A.re + 2 * j, r0);
    store_er(A.re + 2 * j + 1, r1);
    if (A.ovf) A.ovf[j] = check_overflow(c, r0, r1);
}

// Group per number: lane gl handles digit gl; the bounds are formed on every lane.
template <int OP>
GRNS_D void group_mpint(const DevCtx *__restrict__ c, const Group &g, const MpintArgs &A, long long j) {
    const bool act = g.gl < g.L;
    long long off = (long long)g.gl * A.ld + j;
    int a = act ? A.xd[off] : 0, b = act ? A.yd[off] : 0;
    int sx = A.xs[j], sy = A.ys[j];
    if (OP == 1) sy ^= 1;
    er_t r0, r1;
    int minus;
    mpint_bounds<OP>(c, sx, sy, load_er(A.xe + 2 * j), load_er(A.xe + 2 * j + 1), load_er(A.ye + 2 * j),
                     load_er(A.ye + 2 * j + 1), r0, r1, minus);
    int r = 0;
    if (act) {
        int m = c->m[g.gl];
        r = mpint_digit<OP>(a, b, sx, sy, m, c->dinv[g.gl]);
        if (OP != 2 && minus) r = r == 0 ? 0 : (m - r);
        A.rd[off] = r;
    }
    int sr = (OP == 2) ? (group_all(g, !act || r == 0) ? 0 : (sx ^ sy)) : minus;
    if (g.gl == 0) {
        A.rs[j] = sr;
        store_er(A.re + 2 * j, r0);
        store_er(A.re + 2 * j + 1, r1);
        if (A.ovf) A.ovf[j] = check_overflow(c, r0, r1);
    }
}

// mp_int_t arrays in the reference's own layout (src/types.cuh:45-49, array of structs): element j
// at base + j * stride bytes, digits[n] at offset 0, sign at 4n, eval[2] at eoff (8-byte aligned).
struct MpAos {
    const unsigned char *x;
    const unsigned char *y;
    unsigned char *z;
    unsigned char *ovf;
    long long stride;
    long long eoff;
};

GRNS_D er_t load_er8(const er_t *p) {  // 8-byte aligned (the struct's own alignment)
    er_t r;
    r.frac = p->frac;
    r.exp = p->exp;
    return r;
}

GRNS_D void store_er8(er_t *p, const er_t &v) {
    p->frac = v.frac;
    p->exp = v.exp;
}

// group_mpint on mp_int_t structs: the digits of one element are contiguous, so the lanes of a
// group read one line.  Every lane reads its operands before any lane writes, so z may be x or y.
template <int OP>
GRNS_D void group_mpint_aos(const DevCtx *__restrict__ c, const Group &g, const MpAos &A, long long j) {
    const int n = c->n;
    const bool act = g.gl < g.L;
    const int *xd = reinterpret_cast<const int *>(A.x + j * A.stride);
    const int *yd = reinterpret_cast<const int *>(A.y + j * A.stride);
    const er_t *xe = reinterpret_cast<const er_t *>(A.x + j * A.stride + A.eoff);
    const er_t *ye = reinterpret_cast<const er_t *>(A.y + j * A.stride + A.eoff);
    int a = act ? xd[g.gl] : 0, b = act ? yd[g.gl] : 0;
    int sx = xd[n], sy = yd[n];
    if (OP == 1) sy ^= 1;
    er_t r0, r1;
    int minus;
    mpint_bounds<OP>(c, sx, sy, load_er8(xe), load_er8(xe + 1), load_er8(ye), load_er8(ye + 1), r0, r1, minus);
    int r = 0;
    if (act) {
        int m = c->m[g.gl];
        r = mpint_digit<OP>(a, b, sx, sy, m, c->dinv[g.gl]);
        if (OP != 2 && minus) r = r == 0 ? 0 : (m - r);
    }
    int sr = (OP == 2) ? (group_all(g, !act || r == 0) ? 0 : (sx ^ sy)) : minus;
    int *zd = reinterpret_cast<int *>(A.z + j * A.stride);
    er_t *ze = reinterpret_cast<er_t *>(A.z + j * A.stride + A.eoff);
    if (act) zd[g.gl] = r;
    if (g.gl == 0) {
        zd[n] = sr;
        store_er8(ze, r0);
        store_er8(ze + 1, r1);
        if (A.ovf) A.ovf[j] = check_overflow(c, r0, r1);
    }
}

template <int OP>
__global__ __launch_bounds__(BLOCK) void k_mpint_aos(const DevCtx *__restrict__ c, int G, MpAos A, long long count) {
    const Group g = make_group(G, c->n);
    const int P = 64 / G;
    const long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / G;
        if (j < count) group_mpint_aos<OP>(c, g, A, j);
    }
}

template <int N, int OP, bool MM>
__global__ __launch_bounds__(BLOCK) void k_mpint(const DevCtx *__restrict__ c, MpintArgs A, long long count) {
    for (long long j = (long long)blockIdx.x * BLOCK + threadIdx.x; j < count; j += (long long)gridDim.x * BLOCK)
        mpint_step<N, OP, MM>(c, A, j);
}

// mpint_div (src/mpint.cuh:286-296, CPU path) of element j by a group: rns_div_fast of the
// magnitudes (the word division, or the reference loop for non-canonical digits and a zero
// divisor), +1 by C's int % for a negative dividend with a nonzero remainder, rns_eval_compute of
// the digits, sign 0 for a zero quotient and x.sign ^ y.sign otherwise.
template <int GD>
GRNS_D void group_mpint_div(const DevCtx *c, const Group &g, const MpintArgs &A, long long j, int &status) {
    const bool act = g.gl < g.L;
    const long long off = (long long)g.gl * A.ld + j;
    const int xi = act ? A.xd[off] : 0, yi = act ? A.yd[off] : 0;
    const int xs = A.xs[j], ys = A.ys[j];
    int q, r;
    if (GD >= c->words && div_words_ok(c, g, xi, yi)) group_div_words<GD>(c, g, xi, yi, q, r);
    else group_div<true>(c, g, xi, yi, q, r, status);
    const bool rz = group_all(g, !act || r == 0);
    if (xs == 1 && !rz && act) q = (int)rem_trunc((long long)q + 1, c->m[g.gl], c->dinv[g.gl]);
    const interval_t e = group_eval<false>(c, g, q, status);
    const bool qz = group_all(g, !act || q == 0);
    if (act) A.rd[off] = q;
    if (g.gl == 0) {
        A.rs[j] = qz ? 0 : (xs ^ ys);
        A.re[2 * j] = e.low;
        A.re[2 * j + 1] = e.upp;
    }
}

// group_mpint_div on mp_int_t structs (z may be x or y: every lane reads before any lane writes).
template <int GD>
GRNS_D void group_mpint_div_aos(const DevCtx *c, const Group &g, const MpAos &A, long long j, int &status) {
    const int n = c->n;
    const bool act = g.gl < g.L;
    const int *xd = reinterpret_cast<const int *>(A.x + j * A.stride);
    const int *yd = reinterpret_cast<const int *>(A.y + j * A.stride);
    const int xi = act ? xd[g.gl] : 0, yi = act ? yd[g.gl] : 0;
    const int xs = xd[n], ys = yd[n];
    int q, r;
    if (GD >= c->words && div_words_ok(c, g, xi, yi)) group_div_words<GD>(c, g, xi, yi, q, r);
    else group_div<true>(c, g, xi, yi, q, r, status);
    const bool rz = group_all(g, !act || r == 0);
    if (xs == 1 && !rz && act) q = (int)rem_trunc((long long)q + 1, c->m[g.gl], c->dinv[g.gl]);
    const interval_t e = group_eval<false>(c, g, q, status);
    const bool qz = group_all(g, !act || q == 0);
    int *zd = reinterpret_cast<int *>(A.z + j * A.stride);
    er_t *ze = reinterpret_cast<er_t *>(A.z + j * A.stride + A.eoff);
    if (act) zd[g.gl] = q;
    if (g.gl == 0) {
        zd[n] = qz ? 0 : (xs ^ ys);
        store_er8(ze, e.low);
        store_er8(ze + 1, e.upp);
    }
}

template <int GD>
__global__ __launch_bounds__(BLOCK) void k_mpint_div_aos(const DevCtx *__restrict__ c, MpAos A, long long count,
                                                         int *__restrict__ status) {
    const Group g = make_group(GD, c->n);
    constexpr int P = 64 / GD;
    const long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    int st = 0;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / GD;
        if (j < count) group_mpint_div_aos<GD>(c, g, A, j, st);
    }
    if (g.gl == 0 && st) atomicOr(status, st);
}

template <int GD>
__global__ __launch_bounds__(BLOCK) void k_mpint_div(const DevCtx *__restrict__ c, MpintArgs A, long long count,
                                                     int *__restrict__ status) {
    const Group g = make_group(GD, c->n);
    constexpr int P = 64 / GD;
    const long long wave = ((long long)xcd_block() * BLOCK + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * BLOCK) >> 6;
    int st = 0;
    for (long long j0 = wave * P; j0 < count; j0 += nwaves * P) {
        const long long j = j0 + (threadIdx.x & 63) / GD;
        if (j < count) group_mpint_div<GD>(c, g, A, j, st);
    }
    if (g.gl == 0 && st) atomicOr(status, st);
}

}  // namespace grns
