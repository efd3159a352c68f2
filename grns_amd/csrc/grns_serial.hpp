// grns_serial.hpp -- thread-per-number forms for moduli sets larger than a wave (n > 64, up to
// 256: the 128- and 256-moduli sets of src/params/moduli-64-4096-bit/ and moduli-100-6400-bit/).
//
// The kernels for n <= 64 keep a number's residues in registers (one lane per number) or spread
// them over a wave group (one modulus per lane); neither holds 128 or 256 residues.  Here one thread
// takes one number through the CPU path's steps with the moduli set at run time: residues,
// products and mixed-radix digits in private arrays, the pairwise sums built in slot order with a
// stack of complete subtrees.  Same results as the other forms (the CPU path's, bit for bit), at
// the cost of private-memory traffic; these kernels serve the large sets only.
#pragma once

#include "grns_kernels.hpp"

namespace grns {

constexpr int SERIAL_BLOCK = 64;  // one wave per block: private arrays of 256 entries per thread

// ---------------------------------------------------------------- pairwise sums
// psum_rd / psum_ru (src/pairwise.cuh:384-458) of leaf(0..n-1): the tree of the dispatch width
// (the power of two above n for 4 <= n < 4096), adjacent slots paired level by level; a subtree
// without elements is absent and its parent adds 0 (psum2 with one element, the top-level pad).
// Slots are visited in order and complete subtrees merged on a stack, which gives that tree.
template <bool UP, class Leaf>
GRNS_D double serial_psum(int n, const Leaf &leaf) {
    auto add = [](double a, double b) { return UP ? dadd_ru(a, b) : dadd_rd(a, b); };
    if (n == 1) return leaf(0);
    if (n == 2) return add(leaf(0), leaf(1));
    if (n == 3) return add(add(leaf(0), leaf(1)), leaf(2));
    int K = 8;
    while (K <= n && K < 8192) K <<= 1;
    double st[16];
    unsigned absent = 0;  // bit t: stack entry t is an absent subtree
    int top = 0;
    for (int i = 0; i < K; i++) {
        if (i < n) {
            st[top] = leaf(i);
            absent &= ~(1u << top);
        } else {
            st[top] = 0.0;
            absent |= 1u << top;
        }
        top++;
        for (int t = i + 1; (t & 1) == 0; t >>= 1) {  // merge the subtrees completed by slot i
            const double b = st[top - 1];
            const bool bab = absent >> (top - 1) & 1u;
            const bool aab = absent >> (top - 2) & 1u;
            top--;
            if (!aab) st[top - 1] = add(st[top - 1], bab ? 0.0 : b);
        }
    }
    return st[0];
}

// ---------------------------------------------------------------- mixed-radix digits
// mrc (src/mrc.cuh:31-44) of the number at x (residue i at x[i * sx]) into mr[0..n-1]; with
// `canon` the residues are first taken mod m_i (the arg-max orders numbers by their values).
GRNS_D void serial_mrc(const DevCtx *c, const int *x, long long sx, int *mr, bool canon = false) {
    const int n = c->n;
    for (int i = 0; i < n; i++) {
        const int m = c->m[i];
        const double dinv = c->dinv[i];
        int v = x[(long long)i * sx];
        if (canon) v = canon_res(v, m, dinv);
        for (int j = 0; j < i; j++) {
            v = v < mr[j] ? (int)((long long)m - (long long)mr[j] + (long long)v) : (int)((long long)v - (long long)mr[j]);
            v = mulmod3(v, c->mrc_inv[(size_t)j * n + i], m, dinv);
        }
        mr[i] = v;
    }
}

// mrs_cmp of the digits of x and y (src/mrc.cuh:13-22), most significant first.
GRNS_D int serial_mrc_compare(const DevCtx *c, const int *x, long long sx, const int *y, long long sy, int *mx,
                              int *my, bool canon = false) {
    serial_mrc(c, x, sx, mx, canon);
    serial_mrc(c, y, sy, my, canon);
    for (int i = c->n - 1; i >= 0; i--) {
        if (mx[i] > my[i]) return 1;
        if (my[i] > mx[i]) return -1;
    }
    return 0;
}

// ---------------------------------------------------------------- evaluation
// The refinement loop of rns_eval_compute[_fast] (src/rnseval.cuh:79-101) on s[0..n-1].
GRNS_D void serial_refine(const DevCtx *c, int *s, double sumu, interval_t &out, int &status) {
    const int n = c->n;
    long long K = 0;
    int it = 0;
    while (sumu < c->accuracy) {
        double kv = (double)c->ref_factor;
        if (sumu > 0) {
            double a = (double)(-(ceil_log2_libm(sumu) + 1));
            kv = a > kv ? a : kv;
        }
        if (sumu == 0 || kv > c->log2M || ++it > 4096) {  // the reference reads outside RNS_POW2
            status |= GST_REFINE;
            out.low.frac = out.upp.frac = __builtin_nan("");
            out.low.exp = out.upp.exp = 0;
            return;
        }
        const int k = (int)kv;
        const int *p2 = c->pow2 + (size_t)k * n;
        for (int i = 0; i < n; i++) s[i] = mulmod_recip(s[i], p2[i], c->m[i], c->dinv[i]);
        sumu = serial_psum<true>(n, [&](int i) { return ddiv_ru((double)s[i], c->dm[i]); });
        sumu -= to_uint(sumu);
        K += k;
    }
    double suml = serial_psum<false>(n, [&](int i) { return ddiv_rd((double)s[i], c->dm[i]); });
    suml -= to_uint(suml);
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    out.low.exp -= K;
    out.upp.exp -= K;
}

// rns_eval_compute (FAST = false, src/rnseval.cuh:29-102) / rns_eval_compute_fast (:112-167) of the
// number at x (residue i at x[i * sx]); s and mr are private scratch of n entries.
template <bool FAST>
GRNS_D interval_t serial_eval(const DevCtx *c, const int *x, long long sx, int *s, int *mr, int &status) {
    const int n = c->n;
    interval_t out;
    bool zero = true;
    for (int i = 0; i < n; i++) zero = zero && x[(long long)i * sx] == 0;
    if (zero) {
        out.low.frac = 0; out.low.exp = 0; out.upp.frac = 0; out.upp.exp = 0;
        return out;
    }
    for (int i = 0; i < n; i++) s[i] = mulmod_recip(x[(long long)i * sx], c->w[i], c->m[i], c->dinv[i]);
    double suml = serial_psum<false>(n, [&](int i) { return dmul_rd((double)s[i], c->rd[i]); });
    double sumu = serial_psum<true>(n, [&](int i) { return dmul_ru((double)s[i], c->ru[i]); });
    const unsigned whl = to_uint(suml), whu = to_uint(sumu);
    suml = suml - whl;
    sumu = sumu - whu;
    if (FAST) {
        if (sumu >= c->accuracy) {
            out.low = er_set_d(suml);
            out.upp = er_set_d(sumu);
            return out;
        }
        serial_refine(c, s, sumu, out, status);
        return out;
    }
    out.low = er_set_d(suml);
    out.upp = er_set_d(sumu);
    int top = -1;
    if (whl != whu) {
        serial_mrc(c, x, sx, mr);
        top = mr[n - 1];
    }
    if (top > 0) {
        out.upp = c->inv_upp;
        return out;
    }
    if (top == 0) out.low = c->unit_low;
    if (sumu >= c->accuracy) return out;
    serial_refine(c, s, sumu, out, status);
    return out;
}

// rns_cmp (src/rnscmp.cuh:20-39); s, mr, mr2 are private scratch of n entries.
GRNS_D int serial_cmp(const DevCtx *c, const int *x, long long sx, const int *y, long long sy, int *s, int *mr,
                      int *mr2, int &status) {
    const interval_t ex = serial_eval<false>(c, x, sx, s, mr, status);
    const interval_t ey = serial_eval<false>(c, y, sy, s, mr, status);
    if (er_ucmp(ex.low, ey.upp) > 0) return 1;
    if (er_ucmp(ey.low, ex.upp) > 0) return -1;
    bool equal = true;
    for (int i = 0; i < c->n; i++) equal = equal && x[(long long)i * sx] == y[(long long)i * sy];
    if (equal) return 0;
    return serial_mrc_compare(c, x, sx, y, sy, mr, mr2);
}

// ---------------------------------------------------------------- division
// Scratch of one division: the partial quotient and the evaluation / digit arrays.
struct SerialDivScratch {
    int pq[MAXN];
    int s[MAXN];
    int mr[MAXN];
    int mr2[MAXN];
};

// rns_div (FAST = false, src/rnsdiv.cuh:19-61) / rns_div_fast (:67-114): q, r (residue i at
// q[i * sq], r[i * sq]) of x / d.  A loop the reference would not leave (a zero partial quotient,
// an index outside RNS_POW2) stops with the status bit, as the other forms do.
template <bool FAST>
GRNS_D void serial_div(const DevCtx *c, const int *x, const int *d, long long sx, int *q, int *r, long long sq,
                       SerialDivScratch &w, int &status) {
    const int n = c->n;
    for (int i = 0; i < n; i++) {
        q[(long long)i * sq] = 0;
        r[(long long)i * sq] = x[(long long)i * sx];
    }
    const interval_t deval = serial_eval<false>(c, d, sx, w.s, w.mr, status);
    interval_t reval = serial_eval<false>(c, r, sq, w.s, w.mr, status);
    long long it = 0;
    while (er_ucmp(reval.low, deval.upp) >= 0) {
        er_t flq = er_div_rd(reval.low, deval.upp);
        if (++it > 20000) { status |= GST_DIVLOOP; return; }
        if (FAST) {
            for (int i = 0; i < n; i++) w.pq[i] = 0;
            bool got = false;
            while (flq.exp >= 0 && flq.frac > 0) {
                if (flq.exp > c->log2M) { status |= GST_DIVLOOP; return; }
                const int *p2 = c->pow2 + (size_t)flq.exp * n;
                for (int i = 0; i < n; i++) w.pq[i] = addmod_recip(w.pq[i], p2[i], c->m[i], c->dinv[i]);
                flq.frac = dsub_rd(flq.frac, 1.0);
                flq = er_adjust(flq);
                got = true;
            }
            if (!got) { status |= GST_DIVLOOP; return; }
        } else {
            if (flq.exp < 0 || flq.exp > c->log2M) { status |= GST_DIVLOOP; return; }
            const int *p2 = c->pow2 + (size_t)flq.exp * n;
            for (int i = 0; i < n; i++) w.pq[i] = p2[i];
        }
        for (int i = 0; i < n; i++) {
            const int m = c->m[i];
            const double dinv = c->dinv[i];
            const long long o = (long long)i * sq;
            q[o] = addmod_recip(q[o], w.pq[i], m, dinv);
            const int t = mulmod_recip(w.pq[i], d[(long long)i * sx], m, dinv);
            r[o] = psubmod(r[o], t, m, dinv);
        }
        reval = serial_eval<false>(c, r, sq, w.s, w.mr, status);
    }
    if (er_ucmp(reval.upp, deval.low) >= 0) {  // ambiguity: compare r and d by their digits
        if (serial_mrc_compare(c, r, sq, d, sx, w.mr, w.mr2) >= 0) {
            for (int i = 0; i < n; i++) {
                const int m = c->m[i];
                const double dinv = c->dinv[i];
                const long long o = (long long)i * sq;
                q[o] = addmod_recip(q[o], 1, m, dinv);
                r[o] = psubmod(r[o], d[(long long)i * sx], m, dinv);
            }
        }
    }
}

// ---------------------------------------------------------------- scaling by 2^D
// rns_rank_compute (src/rnsscal.cuh:17-45) of x with c = x w in cs.
GRNS_D int serial_rank(const DevCtx *c, const int *x, long long sx, const int *cs, int *mr) {
    const int n = c->n;
    const double suml = serial_psum<false>(n, [&](int i) { return dmul_rd((double)cs[i], c->rd[i]); });
    const double sumu = serial_psum<true>(n, [&](int i) { return dmul_ru((double)cs[i], c->ru[i]); });
    const unsigned whl = to_uint(suml), whu = to_uint(sumu);
    if (whl == whu) return (int)whl;
    serial_mrc(c, x, sx, mr);
    return mr[n - 1] == 0 ? (int)whu : (int)whl;
}
// rns_rank_compute_fast (:50-62)
GRNS_D int serial_rank_fast(const DevCtx *c, const int *cs) {
    return (int)serial_psum<true>(c->n, [&](int i) { return dmul_ru((double)cs[i], c->ru[i]); });
}
// make_scaling_step (:67-85): y = (x - (X mod 2^j)) / 2^j with X mod 2^j from the rank k.
GRNS_D void serial_scaling_step(const DevCtx *c, int k, int j, int *y, const int *cs) {
    const int n = c->n;
    const long long pow2j = 1ll << j;
    long long residue = 0;
    for (int i = 0; i < n; i++) {
        const long long prod = (long long)c->part_pow2_res[j - 1][i] * (long long)cs[i];
        residue += prod >= 0 ? (prod & (pow2j - 1)) : -((-prod) & (pow2j - 1));  // C's % by 2^j
    }
    residue = (residue - (long long)k * (long long)c->prod_pow2_res[j - 1]) % pow2j;
    if (residue < 0) residue += pow2j;
    for (int i = 0; i < n; i++) {
        const int m = c->m[i];
        const double dinv = c->dinv[i];
        const int multiple = psubmod(y[i], (int)(residue % m), m, dinv);
        y[i] = mulmod_recip(multiple, c->pow2_inv[j - 1][i], m, dinv);
    }
}
// rns_scale2pow (:93-119) of x into y (private, n entries); cs and mr are scratch.
GRNS_D void serial_scale2pow(const DevCtx *c, const int *x, long long sx, unsigned D, int *y, int *cs, int *mr) {
    const int n = c->n;
    for (int i = 0; i < n; i++) y[i] = x[(long long)i * sx];
    int t = (int)(D / THRESH);
    if (t > 0) {
        for (int i = 0; i < n; i++) cs[i] = mulmod_recip(y[i], c->w[i], c->m[i], c->dinv[i]);
        const int k = serial_rank(c, x, sx, cs, mr);
        serial_scaling_step(c, k, THRESH, y, cs);
        t -= 1;
    }
    while (t > 0) {
        bool zero = true;  // residues all 0 stay all 0: the remaining steps change nothing
        for (int i = 0; i < n; i++) zero = zero && y[i] == 0;
        if (zero) break;
        for (int i = 0; i < n; i++) cs[i] = mulmod_recip(y[i], c->w[i], c->m[i], c->dinv[i]);
        serial_scaling_step(c, serial_rank_fast(c, cs), THRESH, y, cs);
        t -= 1;
    }
    const unsigned dd = D % THRESH;
    if (dd > 0) {
        for (int i = 0; i < n; i++) cs[i] = mulmod_recip(y[i], c->w[i], c->m[i], c->dinv[i]);
        const int k = dd < D ? serial_rank_fast(c, cs) : serial_rank(c, y, 1, cs, mr);
        serial_scaling_step(c, k, (int)dd, y, cs);
    }
}

// ---------------------------------------------------------------- kernels
GRNS_D long long serial_tid() { return (long long)blockIdx.x * SERIAL_BLOCK + threadIdx.x; }
GRNS_D long long serial_threads() { return (long long)gridDim.x * SERIAL_BLOCK; }

template <bool FAST>
__global__ __launch_bounds__(SERIAL_BLOCK) void k_eval_serial(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                              long long smod, long long snum, long long count,
                                                              interval_t *__restrict__ out, int *__restrict__ status) {
    int s[MAXN], mr[MAXN];
    int st = 0;
    for (long long j = serial_tid(); j < count; j += serial_threads())
        store_interval(out + j, serial_eval<FAST>(c, x + j * snum, smod, s, mr, st));
    if (st) atomicOr(status, st);
}

__global__ __launch_bounds__(SERIAL_BLOCK) void k_cmp_serial(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                             long long xsm, long long xsn, const int *__restrict__ y,
                                                             long long ysm, long long ysn, long long count,
                                                             int *__restrict__ sign, int *__restrict__ status) {
    int s[MAXN], mr[MAXN], mr2[MAXN];
    int st = 0;
    for (long long j = serial_tid(); j < count; j += serial_threads())
        sign[j] = serial_cmp(c, x + j * xsn, xsm, y + j * ysn, ysm, s, mr, mr2, st);
    if (st) atomicOr(status, st);
}

template <bool FAST>
__global__ __launch_bounds__(SERIAL_BLOCK) void k_div_serial(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                             const int *__restrict__ d, long long smod, long long snum,
                                                             long long count, int *__restrict__ q,
                                                             int *__restrict__ r, int *__restrict__ status) {
    SerialDivScratch w;
    int st = 0;
    for (long long j = serial_tid(); j < count; j += serial_threads())
        serial_div<FAST>(c, x + j * snum, d + j * snum, smod, q + j * snum, r + j * snum, smod, w, st);
    if (st) atomicOr(status, st);
}

__global__ __launch_bounds__(SERIAL_BLOCK) void k_scale2pow_serial(const DevCtx *__restrict__ c,
                                                                   const int *__restrict__ x, long long smod,
                                                                   long long snum, const unsigned *__restrict__ D,
                                                                   long long count, int *__restrict__ y) {
    int out[MAXN], cs[MAXN], mr[MAXN];
    for (long long j = serial_tid(); j < count; j += serial_threads()) {
        serial_scale2pow(c, x + j * snum, smod, D[j], out, cs, mr);
        for (int i = 0; i < c->n; i++) y[j * snum + (long long)i * smod] = out[i];
    }
}

__global__ __launch_bounds__(SERIAL_BLOCK) void k_mrc_serial(const DevCtx *__restrict__ c, const int *__restrict__ x,
                                                             long long smod, long long snum, long long count,
                                                             int *__restrict__ d, long long dsm, long long dsn) {
    int mr[MAXN];
    for (long long j = serial_tid(); j < count; j += serial_threads()) {
        serial_mrc(c, x + j * snum, smod, mr);
        for (int i = 0; i < c->n; i++) d[j * dsn + (long long)i * dsm] = mr[i];
    }
}

// mpint operands of the serial kernels: digit i of element j at d[i * dm + j * dn], the sign at
// s[j * ss], eval[0..1] at the byte address e + j * es; SoA arrays and mp_int_t structs alike.
struct MpSerial {
    const int *d;
    long long dm, dn;
    const int *s;
    long long ss;
    const unsigned char *e;
    long long es;
};

GRNS_D const er_t *mp_eval_at(const MpSerial &A, long long j) {
    return reinterpret_cast<const er_t *>(A.e + j * A.es);
}

// mpint_add / mpint_sub / mpint_mul (OP 0 / 1 / 2, src/mpint.cuh:160-278) of element j.
template <int OP>
GRNS_D void serial_mpint(const DevCtx *c, const MpSerial &X, const MpSerial &Y, const MpSerial &Z,
                         unsigned char *ovf, long long j) {
    const int n = c->n;
    const int *xd = X.d + j * X.dn;
    const int *yd = Y.d + j * Y.dn;
    int *zd = const_cast<int *>(Z.d) + j * Z.dn;
    const int sx = X.s[j * X.ss];
    int sy = Y.s[j * Y.ss];
    if (OP == 1) sy ^= 1;
    const er_t *xe = mp_eval_at(X, j);
    const er_t *ye = mp_eval_at(Y, j);
    er_t r0, r1;
    int minus;
    mpint_bounds<OP>(c, sx, sy, load_er8(xe), load_er8(xe + 1), load_er8(ye), load_er8(ye + 1), r0, r1, minus);
    int nz = 0;
    for (int i = 0; i < n; i++) {  // every digit is read before it is written: z may be x or y
        const int m = c->m[i];
        int r = mpint_digit<OP>(xd[(long long)i * X.dm], yd[(long long)i * Y.dm], sx, sy, m, c->dinv[i]);
        if (OP != 2 && minus) r = r == 0 ? 0 : (m - r);
        nz |= r;
        zd[(long long)i * Z.dm] = r;
    }
    const int sr = (OP == 2) ? ((nz == 0) ? 0 : (sx ^ sy)) : minus;
    const_cast<int *>(Z.s)[j * Z.ss] = sr;
    er_t *ze = const_cast<er_t *>(mp_eval_at(Z, j));
    store_er8(ze, r0);
    store_er8(ze + 1, r1);
    if (ovf) ovf[j] = check_overflow(c, r0, r1);
}

template <int OP>
__global__ __launch_bounds__(SERIAL_BLOCK) void k_mpint_serial(const DevCtx *__restrict__ c, MpSerial X, MpSerial Y,
                                                               MpSerial Z, unsigned char *ovf, long long count) {
    for (long long j = serial_tid(); j < count; j += serial_threads()) serial_mpint<OP>(c, X, Y, Z, ovf, j);
}

// mpint_div (src/mpint.cuh:286-296): rns_div_fast of the magnitudes, +1 by C's int % for a negative
// dividend with a nonzero remainder, rns_eval_compute of the digits, the sign.
__global__ __launch_bounds__(SERIAL_BLOCK) void k_mpint_div_serial(const DevCtx *__restrict__ c, MpSerial X,
                                                                   MpSerial Y, MpSerial Z, long long count,
                                                                   int *__restrict__ status) {
    SerialDivScratch w;
    int q[MAXN], r[MAXN];
    int st = 0;
    const int n = c->n;
    for (long long j = serial_tid(); j < count; j += serial_threads()) {
        const int xs = X.s[j * X.ss], ys = Y.s[j * Y.ss];
        // serial_div reads d with x's stride: stage the divisor when the strides differ
        int dcopy[MAXN];
        for (int i = 0; i < n; i++) dcopy[i] = Y.d[j * Y.dn + (long long)i * Y.dm];
        int xcopy[MAXN];
        for (int i = 0; i < n; i++) xcopy[i] = X.d[j * X.dn + (long long)i * X.dm];
        serial_div<true>(c, xcopy, dcopy, 1, q, r, 1, w, st);
        bool rz = true;
        for (int i = 0; i < n; i++) rz = rz && r[i] == 0;
        if (xs == 1 && !rz)
            for (int i = 0; i < n; i++) q[i] = (int)rem_trunc((long long)q[i] + 1, c->m[i], c->dinv[i]);
        const interval_t e = serial_eval<false>(c, q, 1, w.s, w.mr, st);
        bool qz = true;
        for (int i = 0; i < n; i++) {
            qz = qz && q[i] == 0;
            const_cast<int *>(Z.d)[j * Z.dn + (long long)i * Z.dm] = q[i];
        }
        const_cast<int *>(Z.s)[j * Z.ss] = qz ? 0 : (xs ^ ys);
        er_t *ze = const_cast<er_t *>(mp_eval_at(Z, j));
        store_er8(ze, e.low);
        store_er8(ze + 1, e.upp);
    }
    if (st) atomicOr(status, st);
}

// ---------------------------------------------------------------- arg-max
// The reference's pipeline (src/dpp/rnsmax.cuh:24-138): every number's interval evaluation, then a
// reduction in which the evaluations order two numbers when they do not overlap and the mixed-
// radix digits do otherwise; equal values go to the lower index.
struct SerialCand {
    long long idx;
    interval_t iv;
};

// Whether candidate a ranks above b (b.idx < 0: empty).
GRNS_D bool serial_above(const DevCtx *c, const int *x, long long smod, long long snum, const SerialCand &a,
                         const SerialCand &b, int *mr, int *mr2) {
    if (a.idx < 0) return false;
    if (b.idx < 0) return true;
    if (er_ucmp(a.iv.low, b.iv.upp) > 0) return true;
    if (er_ucmp(b.iv.low, a.iv.upp) > 0) return false;
    const int r = serial_mrc_compare(c, x + a.idx * snum, smod, x + b.idx * snum, smod, mr, mr2, true);
    return r > 0 || (r == 0 && a.idx < b.idx);
}

constexpr int SERIAL_MAX_BLOCKS = 1024;

// Stage 1: each block reduces a strided share of the numbers to one candidate.  Non-canonical
// residues stand for their value mod m_i: the digits are those of the canonical residues, and the
// evaluations are taken of the canonical residues too.
__global__ __launch_bounds__(SERIAL_BLOCK) void k_max_serial_partial(const DevCtx *__restrict__ c,
                                                                     const int *__restrict__ x, long long smod,
                                                                     long long snum, long long count,
                                                                     SerialCand *__restrict__ part) {
    __shared__ SerialCand sh[SERIAL_BLOCK];
    int s[MAXN], mr[MAXN], mr2[MAXN], canon[MAXN];
    int st = 0;
    SerialCand best;
    best.idx = -1;
    const int n = c->n;
    for (long long j = serial_tid(); j < count; j += serial_threads()) {
        for (int i = 0; i < n; i++) canon[i] = canon_res(x[j * snum + (long long)i * smod], c->m[i], c->dinv[i]);
        SerialCand e;
        e.idx = j;
        e.iv = serial_eval<false>(c, canon, 1, s, mr, st);
        if (serial_above(c, x, smod, snum, e, best, mr, mr2)) best = e;
    }
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int h = SERIAL_BLOCK / 2; h >= 1; h >>= 1) {
        if (threadIdx.x < (unsigned)h && serial_above(c, x, smod, snum, sh[threadIdx.x + h], sh[threadIdx.x], mr, mr2))
            sh[threadIdx.x] = sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Stage 2 (one block): the best candidate; its residues and index are written out.
__global__ __launch_bounds__(SERIAL_BLOCK) void k_max_serial_final(const DevCtx *__restrict__ c,
                                                                   const int *__restrict__ x, long long smod,
                                                                   long long snum, const SerialCand *__restrict__ part,
                                                                   int nparts, int *__restrict__ out,
                                                                   long long *__restrict__ out_idx) {
    __shared__ SerialCand sh[SERIAL_BLOCK];
    int mr[MAXN], mr2[MAXN];
    SerialCand best;
    best.idx = -1;
    for (int k = threadIdx.x; k < nparts; k += SERIAL_BLOCK)
        if (serial_above(c, x, smod, snum, part[k], best, mr, mr2)) best = part[k];
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int h = SERIAL_BLOCK / 2; h >= 1; h >>= 1) {
        if (threadIdx.x < (unsigned)h && serial_above(c, x, smod, snum, sh[threadIdx.x + h], sh[threadIdx.x], mr, mr2))
            sh[threadIdx.x] = sh[threadIdx.x + h];
        __syncthreads();
    }
    const long long win = sh[0].idx;
    for (int i = threadIdx.x; i < c->n; i += SERIAL_BLOCK)
        if (win >= 0) out[i] = x[win * snum + (long long)i * smod];
    if (threadIdx.x == 0) *out_idx = win;
}

}  // namespace grns
