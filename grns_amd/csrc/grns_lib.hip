// grns_lib.hip -- host side of libgrns_hip.so: moduli-set contexts (the constants of
// rns_const_init, src/rnsbase.cuh:324-451, computed without GMP/MPFR), kernel launchers and the
// C ABI declared in include/grns.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/grns.h"
#include "../../include/grns_debug.h"
#include "grns_internal.hpp"
#include "grns_serial.hpp"

using namespace grns;


namespace {

#define HIP_OK(call)                                   \
    do {                                               \
        if ((call) != hipSuccess) return GRNS_ERR_HIP; \
    } while (0)

// Canonical inverse of x modulo m (extended Euclid); m >= 2, gcd(x, m) = 1.
int inverse_mod(long long x, long long m) {
    long long a = ((x % m) + m) % m, b = m, u = 1, v = 0;
    while (b != 0) {
        long long q = a / b, t = a - q * b;
        a = b; b = t;
        t = u - q * v; u = v; v = t;
    }
    return (int)(((u % m) + m) % m);
}

// Round-to-nearest directed bounds exactly as the reference's host code forms them
// (src/dinterval.cuh:350-362, unfused: this file is compiled with -ffp-contract=off).
double host_ddiv_rd(double a, double b) { double c = a / b; return c - (PHI1 * std::fabs(c) + ETA); }
double host_ddiv_ru(double a, double b) { double c = a / b; return c + (PHI1 * std::fabs(c) + ETA); }

// frac in [0.5, 1) (the mpfr_get_d_2exp convention of src/rnsbase.cuh:399-409) from a significand
// num in (2^52, 2^53] scaled by 2^-53.
er_t half_er(unsigned long long num, long long exp) {
    er_t r;
    if (num == (1ull << 53)) { r.frac = 0.5; r.exp = exp + 1; }
    else { r.frac = std::ldexp((double)num, -53); r.exp = exp; }
    return r;
}

grns_status build_constants(grns_ctx *c, const int *moduli, int n) {
    DevCtx &h = c->h;
    std::memset(&h, 0, sizeof(h));
    h.n = n;
    // M as base-2^32 limbs
    std::vector<unsigned> limbs(1, 1u);
    for (int i = 0; i < n; i++) {
        unsigned long long carry = 0;
        for (auto &l : limbs) {
            unsigned long long t = (unsigned long long)l * (unsigned)moduli[i] + carry;
            l = (unsigned)t;
            carry = t >> 32;
        }
        if (carry) limbs.push_back((unsigned)carry);
    }
    int bitlen = (int)(limbs.size() - 1) * 32 + (32 - __builtin_clz(limbs.back()));
    h.log2M = bitlen - 1;
    c->log2M = h.log2M;
    h.M64 = (unsigned long long)limbs[0] | (limbs.size() > 1 ? (unsigned long long)limbs[1] << 32 : 0ull);

    for (int i = 0; i < n; i++) {
        const long long m = moduli[i];
        h.m[i] = moduli[i];
        long long p = 1 % m;
        unsigned long long p64 = 1;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            p = (p * (moduli[j] % m)) % m;
            p64 *= (unsigned long long)moduli[j];
        }
        h.w[i] = inverse_mod(p, m);        // (M/m_i)^-1 mod m_i
        h.mi64[i] = p64;
        h.dm[i] = (double)moduli[i];
        h.dw[i] = (double)h.w[i];
        h.dinv[i] = 1.0 / (double)moduli[i];
        h.rd[i] = host_ddiv_rd(1.0, (double)moduli[i]);
        h.ru[i] = host_ddiv_ru(1.0, (double)moduli[i]);
        h.chi[i] = (double)h.w[i] / (double)moduli[i];
        double rem = std::fma(-h.chi[i], (double)moduli[i], (double)h.w[i]);  // exact remainder
        h.clo[i] = rem / (double)moduli[i];
        unsigned long long kb;
        std::memcpy(&kb, &h.chi[i], 8);
        kb &= ~((1ull << 31) - 1);  // keep 22 significant bits (chi is a normal double in (0, 1))
        std::memcpy(&h.khi[i], &kb, 8);
        h.klo[i] = (h.chi[i] - h.khi[i]) + h.clo[i];  // the difference is exact
        {  // the same split for (2^KEY_SHIFT w_i mod m_i) / m_i
            long long ws = h.w[i];
            for (int k = 0; k < KEY_SHIFT; k++) ws = (ws * 2) % m;
            double shi = (double)ws / (double)moduli[i];
            double srem = std::fma(-shi, (double)moduli[i], (double)ws);
            unsigned long long sb;
            std::memcpy(&sb, &shi, 8);
            sb &= ~((1ull << 31) - 1);
            double sh;
            std::memcpy(&sh, &sb, 8);
            h.shi[i] = sh;
            h.slo[i] = (shi - sh) + srem / (double)moduli[i];
        }
        h.c32[i] = (int)((1ull << 32) % (unsigned long long)m);
    }
    for (int t = 0; t < THRESH; t++) {
        unsigned long long mask = (1ull << (t + 1)) - 1;
        unsigned long long pm = 1;
        for (int i = 0; i < n; i++) pm = (pm * (unsigned long long)moduli[i]) & mask;
        h.prod_pow2_res[t] = (int)pm;
        for (int j = 0; j < n; j++) {
            unsigned long long pp = 1;
            for (int i = 0; i < n; i++)
                if (i != j) pp = (pp * (unsigned long long)moduli[i]) & mask;
            h.part_pow2_res[t][j] = (int)pp;
            long long pw = 1;
            for (int k = 0; k <= t; k++) pw = (pw * 2) % moduli[j];
            h.pow2_inv[t][j] = inverse_mod(pw, moduli[j]);
        }
    }
    // src/rnsbase.cuh:389-391 (same expression, same libm)
    h.accuracy = 4 * std::pow(2.0, 1 - 53) * n * std::log2((double)n) * (1 + 0.0000001 / 2) / 0.0000001;
    h.ref_factor = (int)std::floor(std::log2(1 / (2 * h.accuracy)));
    // Filter windows: the evaluation takes its common path when x/M lies in (tau, 1 - tau);
    // tau is a power of two at least 4x the accuracy constant.
    double tau = 0x1p-12;
    while (tau < 4 * h.accuracy) tau *= 2;
    h.tau = tau;
    h.err = 0x1p-30;
    h.mr_lim = std::nextafter(1.0 / (double)moduli[n - 1], 2.0);
    h.m_min = moduli[0];
    for (int i = 1; i < n; i++) h.m_min = moduli[i] < h.m_min ? moduli[i] : h.m_min;
    h.big_moduli = 1;
    for (int i = 0; i < n; i++) h.big_moduli &= (moduli[i] >= (1 << 22));
    // RNS_EVAL_UNIT / INV_UNIT: M rounded down to 53 bits, then 1/Md and 1 - 1/Md rounded up and
    // down to 53 bits in the [0.5, 1) convention.
    {
        int e = bitlen > 53 ? bitlen - 53 : 0;
        unsigned __int128 acc = 0;
        int w0 = e / 32, sh = e % 32;
        for (int k = 3; k >= 0; k--) {
            size_t w = (size_t)(w0 + k);
            acc = (acc << 32) | (unsigned __int128)(w < limbs.size() ? limbs[w] : 0u);
        }
        unsigned long long a = (unsigned long long)((acc >> sh) & (((unsigned __int128)1 << 53) - 1));
        int b = 64 - __builtin_clzll(a);
        unsigned __int128 num = (unsigned __int128)1 << (b - 1 + 53);
        unsigned long long q = (unsigned long long)(num / a), r = (unsigned long long)(num % a);
        long long ex = 1 - b - e;
        h.unit_low = half_er(q, ex);
        h.unit_upp = half_er(q + (r != 0), ex);
        const unsigned long long two53 = 1ull << 53;
        if (bitlen <= 53) {
            unsigned __int128 nn = (unsigned __int128)(a - 1) << 53;
            unsigned long long qd = (unsigned long long)(nn / a), rd = (unsigned long long)(nn % a);
            h.inv_low = half_er(qd, 0);
            h.inv_upp = half_er(qd + (rd != 0), 0);
        } else {
            bool exact = (e == 1 && a == (1ull << 52));
            h.inv_low = half_er(two53 - 1, 0);
            h.inv_upp = half_er(exact ? two53 - 1 : two53, 0);
        }
    }
    // RNS_POW2 and MRC_MULT_INV
    c->pow2.assign((size_t)(h.log2M + 1) * n, 0);
    for (int j = 0; j < n; j++) {
        long long v = 1;
        for (int i = 0; i <= h.log2M; i++) {
            c->pow2[(size_t)i * n + j] = (int)v;
            v = (v * 2) % moduli[j];
        }
    }
    c->mrc.assign((size_t)n * n, 0);
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) c->mrc[(size_t)i * n + j] = inverse_mod(moduli[i], moduli[j]);
    c->mrc_d.assign(c->mrc.begin(), c->mrc.end());
    // M / m_i and M as 32-bit words, one word more than M has (room for sum_i s_i M / m_i < n M)
    const int WW = (int)limbs.size() + 1;
    h.words = WW;
    c->crt.assign((size_t)(n + 1) * WW, 0u);
    for (int i = 0; i < n; i++) {
        std::vector<unsigned> mi(1, 1u);
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            unsigned long long carry = 0;
            for (auto &l : mi) {
                unsigned long long t = (unsigned long long)l * (unsigned)moduli[j] + carry;
                l = (unsigned)t;
                carry = t >> 32;
            }
            if (carry) mi.push_back((unsigned)carry);
        }
        for (size_t k = 0; k < mi.size(); k++) c->crt[(size_t)i * WW + k] = mi[k];
    }
    for (size_t k = 0; k < limbs.size(); k++) c->crt[(size_t)n * WW + k] = limbs[k];
    return GRNS_OK;
}



hipStream_t S(void *s) { return (hipStream_t)s; }

// Grid of the thread-per-number kernels of the large moduli sets (n > 64): blocks of one wave,
// at most `cap` of them (their private arrays take up to a few KiB per thread).
long long grid_serial(size_t count, long long cap = 4096) {
    long long b = (long long)((count + SERIAL_BLOCK - 1) / SERIAL_BLOCK);
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return b;
}

long long grid_elems(size_t count) {
    long long b = (long long)((count + BLOCK - 1) / BLOCK);
    if (b < 1) b = 1;
    if (b > (1ll << 20)) b = 1ll << 20;
    return b;
}
// The grid of a per-number streaming kernel (grid-stride loops): one block per BLOCK numbers, or at
// most `cap` blocks per CU (the context's test cap when it sets one).
long long grid_stream(const grns_ctx *c, size_t count, int cap = 0) {
    long long b = grid_elems(count);
    if (c->grid_cap > 0) cap = c->grid_cap;
    if (cap > 0 && c->cus > 0 && b > (long long)cap * c->cus) b = (long long)cap * c->cus;
    return b;
}
long long grid_groups(size_t count, int G) {
    long long per_wave = 64 / G;
    long long waves = (long long)((count + per_wave - 1) / per_wave);
    long long b = (waves + (BLOCK / 64) - 1) / (BLOCK / 64);
    if (b < 1) b = 1;
    if (b > (1ll << 20)) b = 1ll << 20;
    if (b >= 8) b = (b + 7) / 8 * 8;  // a multiple of the XCD count (see xcd_block)
    return b;
}

grns_status launch_check() { return hipGetLastError() == hipSuccess ? GRNS_OK : GRNS_ERR_HIP; }

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

constexpr unsigned RARE_CAP = 1u << 20;      // queued rare items per stream (8 MiB of indices)
constexpr unsigned RARE_RES_CAP = 1u << 16;  // of which staged with their residues (32 MiB)

// The queue's split: RARE_SUBQ sub-queues of RARE_CAP / RARE_SUBQ items (or, with the test hook
// off, all of it in sub-queue 0).
void rare_split(const grns_ctx *c, RareQ &q) {
    const unsigned parts = c->rare_spread ? (unsigned)RARE_SUBQ : 1u;
    q.cap = RARE_CAP / parts;
    q.res_cap = c->rare_staging ? RARE_RES_CAP / parts : 0u;
    q.submask = parts - 1u;
}

// The rare-item queue of `stream` (allocated and cleared on the first call on that stream).  The
// counters are cleared by a kernel on `stream` itself: a hipMemset is ordered on the null stream
// only, and the kernels of a non-blocking stream could otherwise read the counters before it.
grns_status rare_queue(grns_ctx *c, void *stream, RareQ &q) {
    std::lock_guard<std::mutex> lock(c->qmu);
    for (auto &e : c->queues)
        if (e.stream == stream) {
            q.items = e.items;
            q.ctr = e.ctr;
            q.res = e.res;
            rare_split(c, q);
            return GRNS_OK;
        }
    grns_ctx::RareQueue e{stream, nullptr, nullptr, nullptr};
    if (hipMalloc(&e.items, sizeof(long long) * RARE_CAP) != hipSuccess) return GRNS_ERR_OOM;
    if (hipMalloc(&e.ctr, RARE_CTR_WORDS * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&e.res, sizeof(int) * RARE_RES_WIDTH * (size_t)RARE_RES_CAP) != hipSuccess) {
        (void)hipFree(e.items);
        if (e.ctr) (void)hipFree(e.ctr);
        return GRNS_ERR_OOM;
    }
    hipLaunchKernelGGL(k_rare_reset, dim3(1), dim3(64), 0, S(stream), e.ctr);
    if (hipGetLastError() != hipSuccess) {
        (void)hipFree(e.items);
        (void)hipFree(e.ctr);
        (void)hipFree(e.res);
        return GRNS_ERR_HIP;
    }
    c->queues.push_back(e);
    q.items = e.items;
    q.ctr = e.ctr;
    q.res = e.res;
    rare_split(c, q);
    return GRNS_OK;
}

// Grid of a rare pass (one wave group per item): about one group per item up to 4096 items, then
// 2048 blocks (16384 groups of 32 lanes) that the items share.
long long rare_grid(size_t count) {
    long long b = (long long)((count + 7) / 8);
    if (b < 64) b = 64;
    if (b > 2048) b = 2048;
    return b;
}

// Grid of the arg-max stage-1 kernel on the context's device: its resident blocks (occupancy query),
// scaled where a larger grid measured faster.
int max_grid_for(const grns_ctx *c) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
        return 2048;
    hipError_t e = hipErrorInvalidValue;
    switch (c->n) {
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_max_partial<4, LD_MM>, BLOCK, 0); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_max_partial<8, LD_MM>, BLOCK, 0); break;
        case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_max_partial<16, LD_MM>, BLOCK, 0); break;
        case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_max_partial<32, LD_MM>, BLOCK, 0); break;
        case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_max_partial<64, LD_MM>, BLOCK, 0); break;
        default: break;
    }
    if (e != hipSuccess || per < 1) return 2048;
    // At n = 16 and 32 (7 resident blocks per CU) a grid of 1.75 resident waves streams faster than
    // one: 2979 instead of 1772 blocks over 2^24 numbers took 167 instead of 185 us (n = 16) and 328
    // instead of 347 us (n = 32).  At n = 64 every larger grid measured slower, so it keeps one wave.
    if (c->n == 16 || c->n == 32) return cus * per * 7 / 4;
    return cus * per;
}

// Timing of the streaming kernel of eval (op 0), cmp (op 1) and max (op 2): an event pair around
// each launch while timing is on.  A slot that still holds an unread pair is folded into the total
// before it is reused.  Not thread-safe: one caller times one context.
constexpr int TIMER_SLOTS = 64;

void fold_slot(grns_ctx::StageTimer &t, int k) {
    float ms = 0.f;
    if (!t.pending[k]) return;
    if (hipEventSynchronize((hipEvent_t)t.ev[2 * k + 1]) == hipSuccess &&
        hipEventElapsedTime(&ms, (hipEvent_t)t.ev[2 * k], (hipEvent_t)t.ev[2 * k + 1]) == hipSuccess) {
        t.ms += ms;
        t.launches++;
    }
    t.pending[k] = 0;
}

int stage_begin(grns_ctx *c, int op, void *stream) {
    if (!c->timing) return -1;
    auto &t = c->timers[op];
    int k = t.next;
    t.next = (k + 1) % TIMER_SLOTS;
    fold_slot(t, k);
    if (hipEventRecord((hipEvent_t)t.ev[2 * k], S(stream)) != hipSuccess) return -1;
    return k;
}

void stage_end(grns_ctx *c, int op, int k, void *stream) {
    if (k < 0) return;
    auto &t = c->timers[op];
    if (hipEventRecord((hipEvent_t)t.ev[2 * k + 1], S(stream)) == hipSuccess) t.pending[k] = 1;
}

// Scratch of grns_max_noalloc: the stage-1 candidates and their residue copies in a file-scope
// device array instead of a caller's workspace.  Every GPU the library runs on holds its own copy
// (GRNS_MAX_NOALLOC_SCRATCH_BYTES, until the process ends) at its own address, which
// hipGetSymbolAddress resolves once per device when the first context on it is made; host code
// uses that address, never the array's name.  Calls that overlap on one GPU share the array and race.
static_assert(GRNS_MAX_NOALLOC_SCRATCH_BYTES >= (size_t)FBLOCK * MAX_FINAL_PER_THREAD * (sizeof(MaxCand) + 64 * sizeof(int)),
              "scratch holds the largest stage-1 grid at n = 64");
static __device__ __align__(256) unsigned char scratch[GRNS_MAX_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_at[SCRATCH_DEVICES];

// Resolves this device's address of `scratch` (the device must be current).
static grns_status scratch_setup(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return GRNS_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_at[device]) return GRNS_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GRNS_ERR_HIP;
    g_scratch_at[device] = static_cast<unsigned char *>(p);
    return GRNS_OK;
}

static unsigned char *scratch_ptr(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_at[device];
}

}  // namespace

extern "C" {

const char *grns_version(void) { return "grns_amd 0.1 (gfx950)"; }

grns_status grns_ctx_create(const int *moduli, int n, int device, grns_ctx **out) {
    if (!moduli || !out || n < 1 || n > MAXN) return GRNS_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (moduli[i] < 2) return GRNS_ERR_ARG;
    grns_ctx *c = new (std::nothrow) grns_ctx();
    if (!c) return GRNS_ERR_OOM;
    c->device = device;
    c->n = n;
    grns_status st = build_constants(c, moduli, n);
    if (st != GRNS_OK) { delete c; return st; }
    if (device < 0) {  // host-only context: constants and host debug entry points, no device memory
        *out = c;
        return GRNS_OK;
    }
    DeviceGuard guard(device);
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->cus < 1)
        c->cus = 256;
    if (hipMalloc(&c->d, sizeof(DevCtx)) != hipSuccess ||
        hipMalloc(&c->d_crt, sizeof(unsigned) * c->crt.size()) != hipSuccess ||
        hipMalloc(&c->d_pow2, sizeof(int) * c->pow2.size()) != hipSuccess ||
        hipMalloc(&c->d_mrc, sizeof(int) * c->mrc.size()) != hipSuccess ||
        hipMalloc(&c->d_mrc_d, sizeof(double) * c->mrc_d.size()) != hipSuccess ||
        hipMalloc(&c->d_status, sizeof(int)) != hipSuccess) {
        grns_ctx_destroy(c);
        return GRNS_ERR_OOM;
    }
    c->h.pow2 = c->d_pow2;
    c->h.mrc_inv = c->d_mrc;
    c->h.mrc_inv_d = c->d_mrc_d;
    c->h.crt_words = c->d_crt;
    if (hipMemcpy(c->d_crt, c->crt.data(), sizeof(unsigned) * c->crt.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_pow2, c->pow2.data(), sizeof(int) * c->pow2.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_mrc, c->mrc.data(), sizeof(int) * c->mrc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_mrc_d, c->mrc_d.data(), sizeof(double) * c->mrc_d.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d, &c->h, sizeof(DevCtx), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->d_status, 0, sizeof(int)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {  // tables and status in place before any stream's kernels
        grns_ctx_destroy(c);
        return GRNS_ERR_HIP;
    }
    c->max_blocks = c->max_blocks_occ = max_grid_for(c);
    if (scratch_setup(c->device) != GRNS_OK) {
        grns_ctx_destroy(c);
        return GRNS_ERR_HIP;
    }
    *out = c;
    return GRNS_OK;
}

grns_status grns_ctx_destroy(grns_ctx *c) {
    if (!c) return GRNS_ERR_ARG;
    if (c->device >= 0) {
        DeviceGuard guard(c->device);
        if (c->d) (void)hipFree(c->d);
        if (c->d_pow2) (void)hipFree(c->d_pow2);
        if (c->d_crt) (void)hipFree(c->d_crt);
        if (c->d_mrc) (void)hipFree(c->d_mrc);
        if (c->d_mrc_d) (void)hipFree(c->d_mrc_d);
        if (c->d_status) (void)hipFree(c->d_status);
        for (auto &q : c->queues) {
            (void)hipFree(q.items);
            (void)hipFree(q.ctr);
            (void)hipFree(q.res);
        }
        for (auto &t : c->timers)
            for (void *e : t.ev) (void)hipEventDestroy((hipEvent_t)e);
    }
    delete c;
    return GRNS_OK;
}

int grns_ctx_n(const grns_ctx *c) { return c ? c->n : 0; }
int grns_ctx_log2m(const grns_ctx *c) { return c ? c->log2M : 0; }
int grns_ctx_words(const grns_ctx *c) { return c ? c->h.words - 1 : 0; }

grns_status grns_ctx_constants(const grns_ctx *c, int *part_inv, double *recip_rd, double *recip_ru,
                               double *accuracy, int *ref_factor, grns_interval *unit, grns_interval *inv_unit) {
    if (!c) return GRNS_ERR_ARG;
    for (int i = 0; i < c->n; i++) {
        if (part_inv) part_inv[i] = c->h.w[i];
        if (recip_rd) recip_rd[i] = c->h.rd[i];
        if (recip_ru) recip_ru[i] = c->h.ru[i];
    }
    if (accuracy) *accuracy = c->h.accuracy;
    if (ref_factor) *ref_factor = c->h.ref_factor;
    if (unit) {
        unit->low.frac = c->h.unit_low.frac; unit->low.exp = c->h.unit_low.exp;
        unit->upp.frac = c->h.unit_upp.frac; unit->upp.exp = c->h.unit_upp.exp;
    }
    if (inv_unit) {
        inv_unit->low.frac = c->h.inv_low.frac; inv_unit->low.exp = c->h.inv_low.exp;
        inv_unit->upp.frac = c->h.inv_upp.frac; inv_unit->upp.exp = c->h.inv_upp.exp;
    }
    return GRNS_OK;
}

grns_status grns_ctx_tables(const grns_ctx *c, int *pow2, int *mrc_inv) {
    if (!c) return GRNS_ERR_ARG;
    if (pow2) std::memcpy(pow2, c->pow2.data(), sizeof(int) * c->pow2.size());
    if (mrc_inv) std::memcpy(mrc_inv, c->mrc.data(), sizeof(int) * c->mrc.size());
    return GRNS_OK;
}

grns_status grns_ctx_status(grns_ctx *c, int *bits) {
    if (!c || !bits || c->device < 0) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy(bits, c->d_status, sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemset(c->d_status, 0, sizeof(int)));
    HIP_OK(hipDeviceSynchronize());  // cleared before any later kernel, on any stream, can raise a bit
    return GRNS_OK;
}

grns_status grns_eval(grns_ctx *c, const int *x, long long smod, long long snum, size_t count, grns_interval *out,
                      int fast, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || (!x && count) || (!out && count)) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    DeviceGuard guard(c->device);
    interval_t *o = reinterpret_cast<interval_t *>(out);
    dim3 grid((unsigned)grid_stream(c, count)), blk(BLOCK);
    long long cnt = (long long)count;
    RareQ q;
    const bool mm = snum == 1 && count <= ((size_t)1 << 29);
#define GRNS_EVAL_CASE(NN)                                                                                       \
    case NN: {                                                                                                   \
        grns_status qs = rare_queue(c, stream, q);                                                               \
        if (qs != GRNS_OK) return qs;                                                                            \
        dim3 rg((unsigned)rare_grid(count));                                                                     \
        int tk = stage_begin(c, 0, stream);                                                                      \
        if (mm) {                                                                                                \
            if (fast) hipLaunchKernelGGL((k_eval<NN, true, LD_MM>), grid, blk, 0, S(stream), c->d, x, smod, snum,  \
                                         cnt, o, q);                                                             \
            else hipLaunchKernelGGL((k_eval<NN, false, LD_MM>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, \
                                    o, q);                                                                       \
        } else {                                                                                                 \
            if (fast) hipLaunchKernelGGL((k_eval<NN, true, LD_ANY>), grid, blk, 0, S(stream), c->d, x, smod, snum, \
                                         cnt, o, q);                                                             \
            else hipLaunchKernelGGL((k_eval<NN, false, LD_ANY>), grid, blk, 0, S(stream), c->d, x, smod, snum,     \
                                    cnt, o, q);                                                                  \
        }                                                                                                        \
        stage_end(c, 0, tk, stream);                                                                             \
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;                                                \
        if (fast) hipLaunchKernelGGL((k_eval_rare<NN, true>), rg, blk, 0, S(stream), c->d, x, smod, snum, cnt, o,  \
                                     c->d_status, q);                                                            \
        else hipLaunchKernelGGL((k_eval_rare<NN, false>), rg, blk, 0, S(stream), c->d, x, smod, snum, cnt, o,      \
                                c->d_status, q);                                                                 \
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;                                                \
        hipLaunchKernelGGL(k_rare_reset, dim3(1), dim3(64), 0, S(stream), q.ctr);                                \
        return launch_check();                                                                                   \
    }
    switch (c->n) {
        GRNS_EVAL_CASE(4)
        GRNS_EVAL_CASE(8)
        GRNS_EVAL_CASE(16)
        GRNS_EVAL_CASE(32)
        GRNS_EVAL_CASE(64)
        default:
            break;
    }
#undef GRNS_EVAL_CASE
    if (c->n > 64) {
        dim3 gs((unsigned)grid_serial(count));
        if (fast) hipLaunchKernelGGL((k_eval_serial<true>), gs, dim3(SERIAL_BLOCK), 0, S(stream), c->d, x, smod, snum,
                                     cnt, reinterpret_cast<interval_t *>(out), c->d_status);
        else hipLaunchKernelGGL((k_eval_serial<false>), gs, dim3(SERIAL_BLOCK), 0, S(stream), c->d, x, smod, snum, cnt,
                                reinterpret_cast<interval_t *>(out), c->d_status);
        return launch_check();
    }
    int G = groups_for(c->n);
    dim3 g2((unsigned)grid_groups(count, G));
    if (fast) hipLaunchKernelGGL((k_eval_group<true>), g2, blk, 0, S(stream), c->d, G, x, smod, snum, cnt, o, c->d_status);
    else hipLaunchKernelGGL((k_eval_group<false>), g2, blk, 0, S(stream), c->d, G, x, smod, snum, cnt, o, c->d_status);
    return launch_check();
}

grns_status grns_cmp(grns_ctx *c, const int *x, long long xsm, long long xsn, const int *y, long long ysm,
                     long long ysn, size_t count, int *sign, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || (count && (!x || !y || !sign))) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    DeviceGuard guard(c->device);
    dim3 grid((unsigned)grid_stream(c, count)), blk(BLOCK);
    long long cnt = (long long)count;
    RareQ q;
    const bool mm = xsn == 1 && ysn == 1 && count <= ((size_t)1 << 29);
#define GRNS_CMP_CASE(NN)                                                                                          \
    case NN: {                                                                                                     \
        grns_status qs = rare_queue(c, stream, q);                                                                 \
        if (qs != GRNS_OK) return qs;                                                                              \
        dim3 rg((unsigned)rare_grid(count));                                                                       \
        int tk = stage_begin(c, 1, stream);                                                                        \
        if (mm) hipLaunchKernelGGL((k_cmp<NN, LD_MM>), grid, blk, 0, S(stream), c->d, x, xsm, xsn, y, ysm, ysn, cnt,  \
                                   sign, q);                                                                       \
        else hipLaunchKernelGGL((k_cmp<NN, LD_ANY>), grid, blk, 0, S(stream), c->d, x, xsm, xsn, y, ysm, ysn, cnt,     \
                                sign, q);                                                                          \
        stage_end(c, 1, tk, stream);                                                                               \
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;                                                  \
        hipLaunchKernelGGL((k_cmp_rare<NN>), rg, blk, 0, S(stream), c->d, x, xsm, xsn, y, ysm, ysn, cnt, sign,       \
                           c->d_status, q);                                                                        \
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;                                                  \
        hipLaunchKernelGGL(k_rare_reset, dim3(1), dim3(64), 0, S(stream), q.ctr);                                  \
        return launch_check();                                                                                     \
    }
    switch (c->n) {
        GRNS_CMP_CASE(4)
        GRNS_CMP_CASE(8)
        GRNS_CMP_CASE(16)
        GRNS_CMP_CASE(32)
        GRNS_CMP_CASE(64)
        default:
            break;
    }
#undef GRNS_CMP_CASE
    if (c->n > 64) {
        hipLaunchKernelGGL(k_cmp_serial, dim3((unsigned)grid_serial(count)), dim3(SERIAL_BLOCK), 0, S(stream), c->d, x,
                           xsm, xsn, y, ysm, ysn, cnt, sign, c->d_status);
        return launch_check();
    }
    int G = groups_for(c->n);
    dim3 g2((unsigned)grid_groups(count, G));
    hipLaunchKernelGGL(k_cmp_group, g2, blk, 0, S(stream), c->d, G, x, xsm, xsn, y, ysm, ysn, cnt, sign, c->d_status);
    return launch_check();
}

grns_status grns_div(grns_ctx *c, const int *x, const int *d, long long smod, long long snum, size_t count, int *q,
                     int *r, int flags, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || (count && (!x || !d || !q || !r))) return GRNS_ERR_ARG;
    if (flags & ~(GRNS_DIV_FAST | GRNS_DIV_REFERENCE)) return GRNS_ERR_ARG;
    const bool fast = (flags & GRNS_DIV_FAST) != 0;
    const bool reference = (flags & GRNS_DIV_REFERENCE) != 0 || c->reference_loops;
    if (count == 0) return GRNS_OK;
    DeviceGuard guard(c->device);
    if (c->n > 64) {  // the reference loop on one thread per pair (no word form this wide)
        dim3 gs((unsigned)grid_serial(count, 2048));
        if (fast) hipLaunchKernelGGL((k_div_serial<true>), gs, dim3(SERIAL_BLOCK), 0, S(stream), c->d, x, d, smod, snum,
                                     (long long)count, q, r, c->d_status);
        else hipLaunchKernelGGL((k_div_serial<false>), gs, dim3(SERIAL_BLOCK), 0, S(stream), c->d, x, d, smod, snum,
                                (long long)count, q, r, c->d_status);
        return launch_check();
    }
    int G = groups_for(c->n);
    dim3 grid((unsigned)grid_groups(count, G)), blk(BLOCK);
    long long cnt = (long long)count;
    // canonical pairs with d != 0: word division; the reference loop for the others (rare pass)
    const int GW = reference ? 0 : words_group(c);
    if (GW) {
        RareQ rq;
        grns_status qs = rare_queue(c, stream, rq);
        if (qs != GRNS_OK) return qs;
        dim3 gw((unsigned)grid_groups(count, GW)), rg((unsigned)rare_grid(count));
        switch (GW) {
            case 2: hipLaunchKernelGGL((k_div_words<2>), gw, blk, 0, S(stream), c->d, x, d, smod, snum, cnt, q, r, rq); break;
            case 4: hipLaunchKernelGGL((k_div_words<4>), gw, blk, 0, S(stream), c->d, x, d, smod, snum, cnt, q, r, rq); break;
            case 8: hipLaunchKernelGGL((k_div_words<8>), gw, blk, 0, S(stream), c->d, x, d, smod, snum, cnt, q, r, rq); break;
            case 16: hipLaunchKernelGGL((k_div_words<16>), gw, blk, 0, S(stream), c->d, x, d, smod, snum, cnt, q, r, rq); break;
            case 32: hipLaunchKernelGGL((k_div_words<32>), gw, blk, 0, S(stream), c->d, x, d, smod, snum, cnt, q, r, rq); break;
            default: hipLaunchKernelGGL((k_div_words<64>), gw, blk, 0, S(stream), c->d, x, d, smod, snum, cnt, q, r, rq); break;
        }
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;
        if (fast) hipLaunchKernelGGL((k_div_rare<true>), rg, blk, 0, S(stream), c->d, G, x, d, smod, snum, cnt, q, r, c->d_status, rq);
        else hipLaunchKernelGGL((k_div_rare<false>), rg, blk, 0, S(stream), c->d, G, x, d, smod, snum, cnt, q, r, c->d_status, rq);
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;
        hipLaunchKernelGGL(k_rare_reset, dim3(1), dim3(64), 0, S(stream), rq.ctr);
        return launch_check();
    }
    if (fast) hipLaunchKernelGGL((k_div<true>), grid, blk, 0, S(stream), c->d, G, x, d, smod, snum, cnt, q, r, c->d_status);
    else hipLaunchKernelGGL((k_div<false>), grid, blk, 0, S(stream), c->d, G, x, d, smod, snum, cnt, q, r, c->d_status);
    return launch_check();
}

grns_status grns_mrc(grns_ctx *c, const int *x, long long smod, long long snum, size_t count, int *d, long long dsm,
                     long long dsn, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || (count && (!x || !d))) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    DeviceGuard guard(c->device);
    if (c->n > 64) {
        hipLaunchKernelGGL(k_mrc_serial, dim3((unsigned)grid_serial(count)), dim3(SERIAL_BLOCK), 0, S(stream), c->d, x,
                           smod, snum, (long long)count, d, dsm, dsn);
        return launch_check();
    }
    const long long cnt = (long long)count;
    dim3 blk(BLOCK), lgrid((unsigned)grid_elems(count));
#define GRNS_MRC_CASE(NN)                                                                                         \
    case NN:                                                                                                      \
        hipLaunchKernelGGL((k_mrc_lane<NN>), lgrid, blk, 0, S(stream), c->d, x, smod, snum, cnt, d, dsm, dsn);     \
        return launch_check();
    switch (c->n) {
        GRNS_MRC_CASE(4)
        GRNS_MRC_CASE(8)
        GRNS_MRC_CASE(16)
        GRNS_MRC_CASE(32)
        GRNS_MRC_CASE(64)
        default:
            break;
    }
#undef GRNS_MRC_CASE
    const int G = groups_for(c->n);
    dim3 grid((unsigned)grid_groups(count, G));
    hipLaunchKernelGGL(k_mrc, grid, blk, 0, S(stream), c->d, G, x, smod, snum, cnt, d, dsm, dsn);
    return launch_check();
}

grns_status grns_from_binary(grns_ctx *c, const unsigned *w, long long wst, long long wsn, int nwords, const int *sign,
                             size_t count, int *x, long long smod, long long snum, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || nwords < 1 || (count && (!w || !x))) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    if (c->n > 64) return GRNS_ERR_UNSUPPORTED;
    DeviceGuard guard(c->device);
    const int G = groups_for(c->n);
    dim3 grid((unsigned)grid_groups(count, G)), blk(BLOCK);
    const long long cnt = (long long)count;
    switch (G) {
        case 1: hipLaunchKernelGGL((k_from_binary<1>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
        case 2: hipLaunchKernelGGL((k_from_binary<2>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
        case 4: hipLaunchKernelGGL((k_from_binary<4>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
        case 8: hipLaunchKernelGGL((k_from_binary<8>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
        case 16: hipLaunchKernelGGL((k_from_binary<16>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
        case 32: hipLaunchKernelGGL((k_from_binary<32>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
        default: hipLaunchKernelGGL((k_from_binary<64>), grid, blk, 0, S(stream), c->d, w, wst, wsn, nwords, sign, cnt, x, smod, snum); break;
    }
    return launch_check();
}

grns_status grns_to_binary(grns_ctx *c, const int *x, long long smod, long long snum, size_t count, unsigned *w,
                           long long wst, long long wsn, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || (count && (!w || !x))) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    const int GW = words_group(c);
    if (c->n > 64 || !GW) return GRNS_ERR_UNSUPPORTED;
    DeviceGuard guard(c->device);
    dim3 grid((unsigned)grid_groups(count, GW)), blk(BLOCK);
    const long long cnt = (long long)count;
    switch (GW) {
        case 2: hipLaunchKernelGGL((k_to_binary<2>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, w, wst, wsn); break;
        case 4: hipLaunchKernelGGL((k_to_binary<4>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, w, wst, wsn); break;
        case 8: hipLaunchKernelGGL((k_to_binary<8>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, w, wst, wsn); break;
        case 16: hipLaunchKernelGGL((k_to_binary<16>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, w, wst, wsn); break;
        case 32: hipLaunchKernelGGL((k_to_binary<32>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, w, wst, wsn); break;
        default: hipLaunchKernelGGL((k_to_binary<64>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, w, wst, wsn); break;
    }
    return launch_check();
}

grns_status grns_scale2pow(grns_ctx *c, const int *x, long long smod, long long snum, const unsigned *D, size_t count,
                           int *y, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || (count && (!x || !D || !y))) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    DeviceGuard guard(c->device);
    if (c->n > 64) {  // the reference steps on one thread per number
        hipLaunchKernelGGL(k_scale2pow_serial, dim3((unsigned)grid_serial(count, 2048)), dim3(SERIAL_BLOCK), 0,
                           S(stream), c->d, x, smod, snum, D, (long long)count, y);
        return launch_check();
    }
    int G = groups_for(c->n);
    dim3 grid((unsigned)grid_groups(count, G)), blk(BLOCK);
    const long long cnt = (long long)count;
    // canonical numbers: the reference's steps replayed on words; the reference steps for the others
    const int GW = c->reference_loops ? 0 : words_group(c);
    if (GW) {
        RareQ rq;
        grns_status qs = rare_queue(c, stream, rq);
        if (qs != GRNS_OK) return qs;
        dim3 gw((unsigned)grid_groups(count, GW)), rg((unsigned)rare_grid(count));
        switch (GW) {
            case 2: hipLaunchKernelGGL((k_scale2pow_words<2>), gw, blk, 0, S(stream), c->d, x, smod, snum, D, cnt, y, rq); break;
            case 4: hipLaunchKernelGGL((k_scale2pow_words<4>), gw, blk, 0, S(stream), c->d, x, smod, snum, D, cnt, y, rq); break;
            case 8: hipLaunchKernelGGL((k_scale2pow_words<8>), gw, blk, 0, S(stream), c->d, x, smod, snum, D, cnt, y, rq); break;
            case 16: hipLaunchKernelGGL((k_scale2pow_words<16>), gw, blk, 0, S(stream), c->d, x, smod, snum, D, cnt, y, rq); break;
            case 32: hipLaunchKernelGGL((k_scale2pow_words<32>), gw, blk, 0, S(stream), c->d, x, smod, snum, D, cnt, y, rq); break;
            default: hipLaunchKernelGGL((k_scale2pow_words<64>), gw, blk, 0, S(stream), c->d, x, smod, snum, D, cnt, y, rq); break;
        }
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;
        hipLaunchKernelGGL(k_scale2pow_rare, rg, blk, 0, S(stream), c->d, G, x, smod, snum, D, cnt, y, rq);
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;
        hipLaunchKernelGGL(k_rare_reset, dim3(1), dim3(64), 0, S(stream), rq.ctr);
        return launch_check();
    }
    hipLaunchKernelGGL(k_scale2pow, grid, blk, 0, S(stream), c->d, G, x, smod, snum, D, cnt, y);
    return launch_check();
}

// Offset of eval[2] in the reference's mp_int_t (int digits[n]; int sign; er_float_t eval[2];):
// after the 4n + 4 bytes of digits and sign, at the double's 8-byte alignment.
static size_t mpint_aos_eoff(int n) { return ((size_t)4 * n + 4 + 7) / 8 * 8; }

// Stage 1 of the arg-max runs the grid of max_grid_for (set at context creation), each block over
// a contiguous chunk that is a multiple of the block size.
static void max_shape(const grns_ctx *c, size_t count, long long &nblocks, long long &chunk) {
    long long target = c->max_blocks;
    if (target > (long long)FBLOCK * MAX_FINAL_PER_THREAD) target = (long long)FBLOCK * MAX_FINAL_PER_THREAD;
    if (target < 1) target = 1;
    long long per = (long long)((count + target - 1) / target);
    per = ((per + BLOCK - 1) / BLOCK) * BLOCK;
    if (per < BLOCK) per = BLOCK;
    chunk = per;
    nblocks = (long long)((count + per - 1) / per);
    if (nblocks < 1) nblocks = 1;
}

// Workspace of the arg-max: the per-block candidates, then a copy of each candidate's residues.
static size_t max_ws_bytes(const grns_ctx *c, long long nb) {
    size_t b = (size_t)nb * sizeof(MaxCand) + (size_t)nb * c->n * sizeof(int);
    const size_t sb = (size_t)SERIAL_MAX_BLOCKS * sizeof(SerialCand);  // the serial form's candidates
    return b > sb ? b : sb;
}

// Sizes with arg-max kernels; the others take the serial form (any n up to 256).
static bool max_kernel_size(int n) { return n == 4 || n == 8 || n == 16 || n == 32 || n == 64; }

static grns_status max_serial(grns_ctx *c, const int *x, long long smod, long long snum, size_t count,
                              int *out_residues, long long *out_idx, void *workspace, void *stream) {
    const long long nbs = grid_serial(count, SERIAL_MAX_BLOCKS);
    SerialCand *part = static_cast<SerialCand *>(workspace);
    hipLaunchKernelGGL(k_max_serial_partial, dim3((unsigned)nbs), dim3(SERIAL_BLOCK), 0, S(stream), c->d, x, smod, snum,
                       (long long)count, part);
    if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;
    hipLaunchKernelGGL(k_max_serial_final, dim3(1), dim3(SERIAL_BLOCK), 0, S(stream), c->d, x, smod, snum, part,
                       (int)nbs, out_residues, out_idx);
    return launch_check();
}

size_t grns_max_workspace_size(const grns_ctx *c, size_t count) {
    if (!c) return 0;
    long long nb, ch;
    max_shape(c, count, nb, ch);
    return max_ws_bytes(c, nb);
}

// Both arg-max stages over the candidates part[nb] and their residue copies part_res[nb * n].
static grns_status max_run(grns_ctx *c, const int *x, long long smod, long long snum, size_t count, int *out_residues,
                           long long *out_idx, MaxCand *part, int *part_res, long long nb, long long ch, void *stream,
                           bool timed) {
    const bool mm = snum == 1 && count <= ((size_t)1 << 29);
    dim3 grid((unsigned)nb), blk(BLOCK), fblk(FBLOCK);
    long long cnt = (long long)count;
    int tk = -1;
#define GRNS_MAX_CASE(NN)                                                                                    \
    case NN:                                                                                                 \
        if (timed) tk = stage_begin(c, 2, stream);                                                           \
        if (mm) hipLaunchKernelGGL((k_max_partial<NN, LD_MM>), grid, blk, 0, S(stream), c->d, x, smod, snum, \
                                   cnt, ch, part, part_res);                                                 \
        else hipLaunchKernelGGL((k_max_partial<NN, LD_ANY>), grid, blk, 0, S(stream), c->d, x, smod, snum,    \
                                cnt, ch, part, part_res);                                                    \
        if (timed) stage_end(c, 2, tk, stream);                                                              \
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;                                            \
        hipLaunchKernelGGL((k_max_final<NN>), dim3(1), fblk, 0, S(stream), c->d, x, smod, snum, part,        \
                           part_res, (int)nb, out_residues, out_idx);                                        \
        break;
    switch (c->n) {
        GRNS_MAX_CASE(4)
        GRNS_MAX_CASE(8)
        GRNS_MAX_CASE(16)
        GRNS_MAX_CASE(32)
        GRNS_MAX_CASE(64)
        default:
            return GRNS_ERR_UNSUPPORTED;
    }
#undef GRNS_MAX_CASE
    return launch_check();
}

grns_status grns_max(grns_ctx *c, const int *x, long long smod, long long snum, size_t count, int *out_residues,
                     long long *out_idx, void *workspace, size_t ws_bytes, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || !x || !out_residues || !out_idx || count == 0) return GRNS_ERR_ARG;
    long long nb, ch;
    max_shape(c, count, nb, ch);
    if (!workspace || ws_bytes < max_ws_bytes(c, nb)) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    if (!max_kernel_size(c->n)) return max_serial(c, x, smod, snum, count, out_residues, out_idx, workspace, stream);
    MaxCand *part = reinterpret_cast<MaxCand *>(workspace);
    int *part_res = reinterpret_cast<int *>(part + nb);
    return max_run(c, x, smod, snum, count, out_residues, out_idx, part, part_res, nb, ch, stream, true);
}

grns_status grns_max_noalloc(grns_ctx *c, const int *x, long long smod, long long snum, size_t count,
                             int *out_residues, long long *out_idx, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || !x || !out_residues || !out_idx || count == 0) return GRNS_ERR_ARG;
    if (c->n != 4 && c->n != 8 && c->n != 16 && c->n != 32 && c->n != 64) return GRNS_ERR_UNSUPPORTED;
    long long nb, ch;
    max_shape(c, count, nb, ch);
    if (max_ws_bytes(c, nb) > GRNS_MAX_NOALLOC_SCRATCH_BYTES) return GRNS_ERR_ARG;
    unsigned char *base = scratch_ptr(c->device);
    if (!base) return GRNS_ERR_HIP;
    DeviceGuard guard(c->device);
    MaxCand *part = reinterpret_cast<MaxCand *>(base);
    int *part_res = reinterpret_cast<int *>(part + nb);
    return max_run(c, x, smod, snum, count, out_residues, out_idx, part, part_res, nb, ch, stream, false);
}

grns_status grns_max_mrc(grns_ctx *c, const int *x, long long smod, long long snum, size_t count, int *out_residues,
                         long long *out_idx, void *workspace, size_t ws_bytes, void *stream) {
    if (c && c->device < 0) return GRNS_ERR_ARG;
    if (!c || !x || !out_residues || !out_idx || count == 0) return GRNS_ERR_ARG;
    long long nb, ch;
    max_shape(c, count, nb, ch);
    // the same workspace as grns_max: nb indices, then nb digit vectors (fewer bytes than nb
    // candidates with their residue copies)
    if (!workspace || ws_bytes < max_ws_bytes(c, nb)) return GRNS_ERR_ARG;
    if (nb > MRC_FINAL_MAX) return GRNS_ERR_ARG;  // never: max_shape caps the grid at this bound
    DeviceGuard guard(c->device);
    if (!max_kernel_size(c->n)) return max_serial(c, x, smod, snum, count, out_residues, out_idx, workspace, stream);
    long long *pidx = reinterpret_cast<long long *>(workspace);
    int *pdig = reinterpret_cast<int *>(pidx + nb);
    dim3 grid((unsigned)nb), blk(BLOCK);
    long long cnt = (long long)count;
#define GRNS_MAX_MRC_CASE(NN)                                                                                   \
    case NN:                                                                                                    \
        hipLaunchKernelGGL((k_max_mrc_partial<NN>), grid, blk, 0, S(stream), c->d, x, smod, snum, cnt, ch, pidx,   \
                           pdig);                                                                               \
        if (hipGetLastError() != hipSuccess) return GRNS_ERR_HIP;                                               \
        hipLaunchKernelGGL((k_max_mrc_final<NN>), dim3(1), blk, 0, S(stream), x, smod, snum, pidx, pdig, (int)nb, \
                           out_residues, out_idx);                                                              \
        break;
    switch (c->n) {
        GRNS_MAX_MRC_CASE(4)
        GRNS_MAX_MRC_CASE(8)
        GRNS_MAX_MRC_CASE(16)
        GRNS_MAX_MRC_CASE(32)
        GRNS_MAX_MRC_CASE(64)
        default:
            return GRNS_ERR_UNSUPPORTED;
    }
#undef GRNS_MAX_MRC_CASE
    return launch_check();
}

// mpint of the sets without a thread-per-number kernel (n not in {4, 8, 16, 32, 64}, and n > 64):
// the thread-per-number serial forms over SoA arrays (or, from grns_mpint_aos, mp_int_t structs).
static MpSerial mp_serial_soa(const grns_mpint_array *a) {
    MpSerial m;
    m.d = a->digits; m.dm = a->ld; m.dn = 1;
    m.s = a->sign; m.ss = 1;
    m.e = reinterpret_cast<const unsigned char *>(a->eval); m.es = 2 * (long long)sizeof(grns_er);
    return m;
}

static grns_status mpint_serial(grns_ctx *c, int op, const MpSerial &X, const MpSerial &Y, const MpSerial &Z,
                                size_t count, unsigned char *ovf, void *stream) {
    dim3 gs((unsigned)grid_serial(count, op == 3 ? 2048 : 4096)), blk(SERIAL_BLOCK);
    const long long cnt = (long long)count;
    switch (op) {
        case 0: hipLaunchKernelGGL((k_mpint_serial<0>), gs, blk, 0, S(stream), c->d, X, Y, Z, ovf, cnt); break;
        case 1: hipLaunchKernelGGL((k_mpint_serial<1>), gs, blk, 0, S(stream), c->d, X, Y, Z, ovf, cnt); break;
        case 2: hipLaunchKernelGGL((k_mpint_serial<2>), gs, blk, 0, S(stream), c->d, X, Y, Z, ovf, cnt); break;
        default: hipLaunchKernelGGL(k_mpint_div_serial, gs, blk, 0, S(stream), c->d, X, Y, Z, cnt, c->d_status); break;
    }
    return launch_check();
}

static grns_status mpint_launch(grns_ctx *c, int op, const grns_mpint_array *x, const grns_mpint_array *y,
                                const grns_mpint_array *z, size_t count, unsigned char *ovf, void *stream) {
    if (!c || !x || !y || !z || c->device < 0) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    if (x->ld != y->ld || x->ld != z->ld || x->ld < (long long)count) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    dim3 grid((unsigned)grid_stream(c, count)), blk(BLOCK);
    long long cnt = (long long)count;
    MpintArgs A;
    A.xd = x->digits; A.xs = x->sign; A.xe = reinterpret_cast<const er_t *>(x->eval);
    A.yd = y->digits; A.ys = y->sign; A.ye = reinterpret_cast<const er_t *>(y->eval);
    A.rd = z->digits; A.rs = z->sign; A.re = reinterpret_cast<er_t *>(z->eval);
    A.ovf = ovf; A.ld = x->ld;
    const bool mm = count <= ((size_t)1 << 29);
#define GRNS_MP_LAUNCH(NN, OPV)                                                                        \
    if (mm) hipLaunchKernelGGL((k_mpint<NN, OPV, true>), grid, blk, 0, S(stream), c->d, A, cnt);       \
    else hipLaunchKernelGGL((k_mpint<NN, OPV, false>), grid, blk, 0, S(stream), c->d, A, cnt);
#define GRNS_MP_CASE(NN)                                                                               \
    case NN:                                                                                           \
        if (op == 0) {                                                                                 \
            GRNS_MP_LAUNCH(NN, 0)                                                                      \
        } else if (op == 1) {                                                                          \
            GRNS_MP_LAUNCH(NN, 1)                                                                      \
        } else {                                                                                       \
            GRNS_MP_LAUNCH(NN, 2)                                                                      \
        }                                                                                              \
        return launch_check();
    switch (c->n) {
        GRNS_MP_CASE(4)
        GRNS_MP_CASE(8)
        GRNS_MP_CASE(16)
        GRNS_MP_CASE(32)
        GRNS_MP_CASE(64)
        default:
            return mpint_serial(c, op, mp_serial_soa(x), mp_serial_soa(y), mp_serial_soa(z), count, ovf, stream);
    }
#undef GRNS_MP_CASE
#undef GRNS_MP_LAUNCH
}

grns_status grns_mpint_add(grns_ctx *c, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *ovf, void *stream) {
    return mpint_launch(c, 0, x, y, z, count, ovf, stream);
}
grns_status grns_mpint_sub(grns_ctx *c, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *ovf, void *stream) {
    return mpint_launch(c, 1, x, y, z, count, ovf, stream);
}
grns_status grns_mpint_mul(grns_ctx *c, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *ovf, void *stream) {
    return mpint_launch(c, 2, x, y, z, count, ovf, stream);
}
grns_status grns_mpint_div(grns_ctx *c, const grns_mpint_array *x, const grns_mpint_array *y,
                           const grns_mpint_array *z, size_t count, unsigned char *ovf, void *stream) {
    (void)ovf;  // the reference checks no overflow for a division
    if (!c || !x || !y || !z || c->device < 0) return GRNS_ERR_ARG;
    if (count == 0) return GRNS_OK;
    if (x->ld != y->ld || x->ld != z->ld || x->ld < (long long)count) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    if (c->n > 64) return mpint_serial(c, 3, mp_serial_soa(x), mp_serial_soa(y), mp_serial_soa(z), count, nullptr, stream);
    MpintArgs A;
    A.xd = x->digits; A.xs = x->sign; A.xe = reinterpret_cast<const er_t *>(x->eval);
    A.yd = y->digits; A.ys = y->sign; A.ye = reinterpret_cast<const er_t *>(y->eval);
    A.rd = z->digits; A.rs = z->sign; A.re = reinterpret_cast<er_t *>(z->eval);
    A.ovf = nullptr; A.ld = x->ld;
    const int GW = words_group(c) ? words_group(c) : groups_for(c->n);
    dim3 grid((unsigned)grid_groups(count, GW)), blk(BLOCK);
    const long long cnt = (long long)count;
    switch (GW) {
        case 1: hipLaunchKernelGGL((k_mpint_div<1>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        case 2: hipLaunchKernelGGL((k_mpint_div<2>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        case 4: hipLaunchKernelGGL((k_mpint_div<4>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        case 8: hipLaunchKernelGGL((k_mpint_div<8>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        case 16: hipLaunchKernelGGL((k_mpint_div<16>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        case 32: hipLaunchKernelGGL((k_mpint_div<32>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        default: hipLaunchKernelGGL((k_mpint_div<64>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
    }
    return launch_check();
}

size_t grns_mpint_aos_bytes(const grns_ctx *c) {
    if (!c) return 0;
    return mpint_aos_eoff(c->n) + 2 * sizeof(grns_er);
}

grns_status grns_mpint_aos(grns_ctx *c, int op, const void *x, const void *y, void *z, size_t count, size_t elem_bytes,
                           unsigned char *ovf, void *stream) {
    if (!c || c->device < 0) return GRNS_ERR_ARG;
    if (op != GRNS_OP_MPINT_ADD && op != GRNS_OP_MPINT_SUB && op != GRNS_OP_MPINT_MUL && op != GRNS_OP_MPINT_DIV)
        return GRNS_ERR_ARG;
    if (elem_bytes != grns_mpint_aos_bytes(c)) return GRNS_ERR_ARG;  // a different RNS_MODULI_SIZE
    if (count == 0) return GRNS_OK;
    if (!x || !y || !z) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    if (c->n > 64) {  // the same structs through the serial forms: digits, sign, eval by strides
        auto view = [&](const void *base) {
            MpSerial m;
            m.d = static_cast<const int *>(base); m.dm = 1; m.dn = (long long)elem_bytes / 4;
            m.s = static_cast<const int *>(base) + c->n; m.ss = (long long)elem_bytes / 4;
            m.e = static_cast<const unsigned char *>(base) + mpint_aos_eoff(c->n); m.es = (long long)elem_bytes;
            return m;
        };
        const int sop = op == GRNS_OP_MPINT_ADD ? 0 : op == GRNS_OP_MPINT_SUB ? 1 : op == GRNS_OP_MPINT_MUL ? 2 : 3;
        return mpint_serial(c, sop, view(x), view(y), view(z), count, sop == 3 ? nullptr : ovf, stream);
    }
    MpAos A;
    A.x = static_cast<const unsigned char *>(x);
    A.y = static_cast<const unsigned char *>(y);
    A.z = static_cast<unsigned char *>(z);
    A.ovf = op == GRNS_OP_MPINT_DIV ? nullptr : ovf;  // the reference checks no overflow for a division
    A.stride = (long long)elem_bytes;
    A.eoff = (long long)mpint_aos_eoff(c->n);
    const long long cnt = (long long)count;
    dim3 blk(BLOCK);
    if (op == GRNS_OP_MPINT_DIV) {
        const int GW = words_group(c) ? words_group(c) : groups_for(c->n);
        dim3 grid((unsigned)grid_groups(count, GW));
        switch (GW) {
            case 1: hipLaunchKernelGGL((k_mpint_div_aos<1>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
            case 2: hipLaunchKernelGGL((k_mpint_div_aos<2>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
            case 4: hipLaunchKernelGGL((k_mpint_div_aos<4>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
            case 8: hipLaunchKernelGGL((k_mpint_div_aos<8>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
            case 16: hipLaunchKernelGGL((k_mpint_div_aos<16>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
            case 32: hipLaunchKernelGGL((k_mpint_div_aos<32>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
            default: hipLaunchKernelGGL((k_mpint_div_aos<64>), grid, blk, 0, S(stream), c->d, A, cnt, c->d_status); break;
        }
        return launch_check();
    }
    const int G = groups_for(c->n);
    dim3 grid((unsigned)grid_groups(count, G));
    switch (op) {
        case GRNS_OP_MPINT_ADD: hipLaunchKernelGGL((k_mpint_aos<0>), grid, blk, 0, S(stream), c->d, G, A, cnt); break;
        case GRNS_OP_MPINT_SUB: hipLaunchKernelGGL((k_mpint_aos<1>), grid, blk, 0, S(stream), c->d, G, A, cnt); break;
        default: hipLaunchKernelGGL((k_mpint_aos<2>), grid, blk, 0, S(stream), c->d, G, A, cnt); break;
    }
    return launch_check();
}

}  // extern "C"

// ---------------------------------------------------------------- host debug entry points
// The hot-path routines are __host__ __device__; these run the very same code on the CPU so that
// it can be checked against the CPU restatement without a GPU.
namespace {
template <int N>
int host_eval_hot_n(const grns_ctx *c, const int *x, grns_interval *out, int fast) {
    int xr[N];
    for (int i = 0; i < N; i++) xr[i] = x[i];
    interval_t r;
    bool ok = fast ? eval_hot<N, true>(&c->h, xr, r) : eval_hot<N, false>(&c->h, xr, r);
    out->low.frac = r.low.frac; out->low.exp = r.low.exp;
    out->upp.frac = r.upp.frac; out->upp.exp = r.upp.exp;
    return ok ? 1 : 0;
}
template <int N>
int host_cmp_n(const grns_ctx *c, const int *x, const int *y) {
    int xr[N], yr[N];
    for (int i = 0; i < N; i++) { xr[i] = x[i]; yr[i] = y[i]; }
    return cmp_lane<N>(&c->h, xr, yr, x, 1, y, 1);
}
template <int N>
double host_filter_n(const grns_ctx *c, const int *x, int *regular, int *canon) {
    int xr[N];
    for (int i = 0; i < N; i++) xr[i] = x[i];
    bool rg, cn;
    double F = filter_frac<N>(&c->h, xr, rg, cn);
    *regular = rg;
    *canon = cn;
    return F;
}
// The device constants with the host copies of the tables (for the host runs of device routines).
DevCtx host_view(const grns_ctx *c) {
    DevCtx h = c->h;
    h.pow2 = c->pow2.data();
    h.mrc_inv = c->mrc.data();
    h.mrc_inv_d = c->mrc_d.data();
    h.crt_words = c->crt.data();
    return h;
}
template <int N>
int host_eval_full_n(const grns_ctx *c, const int *x, grns_interval *out, int fast, int *status) {
    const DevCtx h = host_view(c);
    int xr[N];
    for (int i = 0; i < N; i++) xr[i] = x[i];
    interval_t r;
    int st = 0;
    bool ok = fast ? eval_full<N, true>(&h, xr, r, st) : eval_full<N, false>(&h, xr, r, st);
    out->low.frac = r.low.frac; out->low.exp = r.low.exp;
    out->upp.frac = r.upp.frac; out->upp.exp = r.upp.exp;
    *status = st;
    return ok ? 1 : 0;
}
template <int N>
int host_cmp_full_n(const grns_ctx *c, const int *x, const int *y) {
    const DevCtx h = host_view(c);
    int xr[N], yr[N];
    for (int i = 0; i < N; i++) { xr[i] = x[i]; yr[i] = y[i]; }
    int st = 0;
    return cmp_full<N>(&h, xr, yr, x, 1, y, 1, st);
}
template <int N>
double host_key_n(const grns_ctx *c, const int *x) {
    int xr[N];
    for (int i = 0; i < N; i++) xr[i] = x[i];
    return key_frac<N>(&c->h, xr);
}
}  // namespace

extern "C" {
/* Common path of the evaluation for one number (residues x[0..n-1]); returns 1 when the
 * common path is final, 0 when the number needs the cooperative path, -1 if n is unsupported. */
int grns_debug_eval_hot(const grns_ctx *c, const int *x, grns_interval *out, int fast) {
    switch (c->n) {
        case 4: return host_eval_hot_n<4>(c, x, out, fast);
        case 8: return host_eval_hot_n<8>(c, x, out, fast);
        case 16: return host_eval_hot_n<16>(c, x, out, fast);
        case 32: return host_eval_hot_n<32>(c, x, out, fast);
        case 64: return host_eval_hot_n<64>(c, x, out, fast);
        default: return -1;
    }
}
double grns_debug_filter(const grns_ctx *c, const int *x, int *regular, int *canon) {
    switch (c->n) {
        case 4: return host_filter_n<4>(c, x, regular, canon);
        case 8: return host_filter_n<8>(c, x, regular, canon);
        case 16: return host_filter_n<16>(c, x, regular, canon);
        case 32: return host_filter_n<32>(c, x, regular, canon);
        case 64: return host_filter_n<64>(c, x, regular, canon);
        default: *regular = 0; *canon = 0; return -1.0;
    }
}
/* Arg-max key: approximation of x/M mod 1 (within the filter error bound), -1 if n is unsupported. */
double grns_debug_max_key(const grns_ctx *c, const int *x) {
    switch (c->n) {
        case 4: return host_key_n<4>(c, x);
        case 8: return host_key_n<8>(c, x);
        case 16: return host_key_n<16>(c, x);
        case 32: return host_key_n<32>(c, x);
        case 64: return host_key_n<64>(c, x);
        default: return -1.0;
    }
}
/* Rare-pass evaluation of one number in one lane: 1 when final (out, status set), 0 when the
 * cooperative path runs, -1 if n is unsupported. */
int grns_debug_eval_full(const grns_ctx *c, const int *x, grns_interval *out, int fast, int *status) {
    switch (c->n) {
        case 4: return host_eval_full_n<4>(c, x, out, fast, status);
        case 8: return host_eval_full_n<8>(c, x, out, fast, status);
        case 16: return host_eval_full_n<16>(c, x, out, fast, status);
        case 32: return host_eval_full_n<32>(c, x, out, fast, status);
        case 64: return host_eval_full_n<64>(c, x, out, fast, status);
        default: return -1;
    }
}
/* Rare-pass comparison in one lane: 1, 0, -1, or 2 when the cooperative path runs. */
int grns_debug_cmp_full(const grns_ctx *c, const int *x, const int *y) {
    switch (c->n) {
        case 4: return host_cmp_full_n<4>(c, x, y);
        case 8: return host_cmp_full_n<8>(c, x, y);
        case 16: return host_cmp_full_n<16>(c, x, y);
        case 32: return host_cmp_full_n<32>(c, x, y);
        case 64: return host_cmp_full_n<64>(c, x, y);
        default: return -2;
    }
}
/* In-lane decision of the comparison kernel: 1, 0, -1, or 2 when the cooperative path runs. */
int grns_debug_cmp_lane(const grns_ctx *c, const int *x, const int *y) {
    switch (c->n) {
        case 4: return host_cmp_n<4>(c, x, y);
        case 8: return host_cmp_n<8>(c, x, y);
        case 16: return host_cmp_n<16>(c, x, y);
        case 32: return host_cmp_n<32>(c, x, y);
        case 64: return host_cmp_n<64>(c, x, y);
        default: return -2;
    }
}
grns_status grns_debug_stage_timing(grns_ctx *c, int enable) {
    if (!c || c->device < 0) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    for (auto &t : c->timers) {
        if (t.ev.empty()) {
            t.ev.assign(2 * TIMER_SLOTS, nullptr);
            t.pending.assign(TIMER_SLOTS, 0);
            for (auto &e : t.ev) {
                hipEvent_t h;
                if (hipEventCreate(&h) != hipSuccess) return GRNS_ERR_HIP;
                e = h;
            }
        }
        for (int k = 0; k < TIMER_SLOTS; k++) fold_slot(t, k);
        t.ms = 0.0;
        t.launches = 0;
        t.next = 0;
    }
    c->timing = enable != 0;
    return GRNS_OK;
}

grns_status grns_debug_stage_read(grns_ctx *c, int op, double *ms_total, long long *launches) {
    if (!c || c->device < 0 || op < 0 || op > 2 || !ms_total || !launches) return GRNS_ERR_ARG;
    DeviceGuard guard(c->device);
    auto &t = c->timers[op];
    for (int k = 0; k < (int)t.pending.size(); k++) fold_slot(t, k);
    *ms_total = t.ms;
    *launches = t.launches;
    t.ms = 0.0;
    t.launches = 0;
    return GRNS_OK;
}

grns_status grns_debug_rare_spread(grns_ctx *c, int on) {
    if (!c) return GRNS_ERR_ARG;
    c->rare_spread = on != 0;
    return GRNS_OK;
}

grns_status grns_debug_rare_staging(grns_ctx *c, int on) {
    if (!c) return GRNS_ERR_ARG;
    c->rare_staging = on != 0;
    return GRNS_OK;
}

grns_status grns_debug_grid_cap(grns_ctx *c, int blocks_per_cu) {
    if (!c || blocks_per_cu < 0) return GRNS_ERR_ARG;
    c->grid_cap = blocks_per_cu;
    return GRNS_OK;
}

grns_status grns_debug_max_blocks(grns_ctx *c, int nblocks) {
    if (!c || nblocks < 0) return GRNS_ERR_ARG;
    c->max_blocks = nblocks > 0 ? nblocks : c->max_blocks_occ;  // max_shape caps it at stage 2's capacity
    return GRNS_OK;
}

grns_status grns_debug_reference_loops(grns_ctx *c, int on) {
    if (!c) return GRNS_ERR_ARG;
    c->reference_loops = on != 0;
    return GRNS_OK;
}

int grns_debug_ceil_log2(double s) { return ceil_log2_libm(s); }
double grns_debug_mulmod_recip_d(double x, double y, double m, double d) { return mulmod_recip_d(x, y, m, d); }
int grns_debug_mulmod_recip(int x, int y, int m, double d) { return mulmod_recip(x, y, m, d); }
int grns_debug_addmod_recip(int x, int y, int m, double d) { return addmod_recip(x, y, m, d); }
double grns_debug_bn_err(double c) { return bn_err(c); }
double grns_debug_fast_scalbn(double x, int n) { return fast_scalbn(x, n); }
unsigned grns_debug_div32(unsigned u2, unsigned u1, unsigned u0, unsigned d1, unsigned d0, unsigned *r1,
                          unsigned *r0, unsigned *v) {
    const unsigned vv = recip32_preinv(d1, d0);
    unsigned a = 0, b = 0;
    const unsigned q = div32_preinv(u2, u1, u0, d1, d0, vv, a, b);
    if (r1) *r1 = a;
    if (r0) *r0 = b;
    if (v) *v = vv;
    return q;
}
double grns_debug_filter_tau(const grns_ctx *c) { return c->h.tau; }
double grns_debug_filter_err(const grns_ctx *c) { return c->h.err; }
}
