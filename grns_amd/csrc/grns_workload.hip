// grns_workload.hip -- libgrns_workload.so: seeded synthetic inputs for benchmarks and tests,
// generated on the device (large RNS arrays cannot be produced quickly on the host).
//
// Every value is a pure function of (seed, stream, element, digit) through a counter-based
// generator (splitmix64 finaliser), so results do not depend on the launch shape.  Uniform draws
// below a bound b use the multiply-high map (r * b) >> 64 (bias below 2^-32).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t draw(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return mix64(mix64(seed ^ mix64(stream)) + ctr);
}
__device__ __forceinline__ uint32_t below(uint64_t r, uint32_t b) { return (uint32_t)__umul64hi(r, b); }
__device__ __forceinline__ uint32_t mulmod64(uint32_t a, uint32_t b, uint32_t m) {
    return (uint32_t)(((uint64_t)a * b) % m);
}

// x uniform in [0, M): independent uniform residues.
__global__ void k_uniform(const int *__restrict__ moduli, int n, long long count, uint64_t seed, uint64_t stream,
                          int *__restrict__ out, long long ld) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    for (int i = 0; i < n; i++)
        out[(long long)i * ld + j] = (int)below(draw(seed, stream, (uint64_t)j * n + i), (uint32_t)moduli[i]);
}

// x = sum_k a_k B_k with mixed-radix digits a_k uniform in [0, m_k) (k < n-1) and a_{n-1} uniform
// in [0, top); mrb[k * n + i] = B_k mod m_i, B_k = m_0 ... m_{k-1}.
__global__ void k_mixed_radix(const int *__restrict__ moduli, int n, const int *__restrict__ mrb, int top,
                              long long count, uint64_t seed, uint64_t stream, int *__restrict__ out, long long ld) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    for (int i = 0; i < n; i++) {
        uint32_t m = (uint32_t)moduli[i];
        uint64_t acc = 0;
        for (int k = 0; k < n; k++) {
            uint32_t bound = (k == n - 1) ? (uint32_t)top : (uint32_t)moduli[k];
            uint32_t a = below(draw(seed, stream, (uint64_t)j * n + k), bound);
            acc += mulmod64(a % m, (uint32_t)mrb[k * n + i], m);
        }
        out[(long long)i * ld + j] = (int)(acc % m);
    }
}

// x an integer of bits[j] bits (bit bits[j]-1 set when top_set, else uniform below 2^bits[j]);
// p32[w * n + i] = 2^(32 w) mod m_i.
__global__ void k_bits(const int *__restrict__ moduli, int n, const int *__restrict__ p32, int words,
                       const unsigned *__restrict__ bits, int top_set, long long count, uint64_t seed, uint64_t stream,
                       int *__restrict__ out, long long ld) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    unsigned nb = bits[j];
    int nw = (int)((nb + 31) / 32);
    if (nw > words) nw = words;
    for (int i = 0; i < n; i++) {
        uint32_t m = (uint32_t)moduli[i];
        uint64_t acc = 0;
        for (int w = 0; w < nw; w++) {
            uint32_t v = (uint32_t)draw(seed, stream, (uint64_t)j * 4096 + w);
            int hi = (int)nb - 32 * w;  // bits of this word that belong to the number
            if (hi < 32) {
                v &= (hi <= 0) ? 0u : ((1u << hi) - 1u);
                if (top_set && hi > 0) v |= 1u << (hi - 1);
            }
            acc += mulmod64(v % m, (uint32_t)p32[w * n + i], m);
        }
        out[(long long)i * ld + j] = (int)(acc % m);
    }
}

// Scatter fixed residue vectors vals[k * n + i] to positions idx[k].
__global__ void k_place(int n, int nvals, const int *__restrict__ vals, const long long *__restrict__ idx,
                        int *__restrict__ out, long long ld) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nvals) return;
    for (int i = 0; i < n; i++) out[(long long)i * ld + idx[k]] = vals[k * n + i];
}

// Uniform unsigned integers in [lo, hi].
__global__ void k_uniform_u32(long long count, unsigned lo, unsigned hi, uint64_t seed, uint64_t stream,
                              unsigned *__restrict__ out) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    out[j] = lo + below(draw(seed, stream, (uint64_t)j), hi - lo + 1);
}

int blocks(long long count) { return (int)((count + 255) / 256); }

}  // namespace

extern "C" {

int gw_uniform(const int *moduli, int n, long long count, unsigned long long seed, unsigned long long stream, int *out,
               long long ld, void *s) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_uniform, dim3(blocks(count)), dim3(256), 0, (hipStream_t)s, moduli, n, count, seed, stream,
                       out, ld);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
int gw_mixed_radix(const int *moduli, int n, const int *mrb, int top, long long count, unsigned long long seed,
                   unsigned long long stream, int *out, long long ld, void *s) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_mixed_radix, dim3(blocks(count)), dim3(256), 0, (hipStream_t)s, moduli, n, mrb, top, count,
                       seed, stream, out, ld);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
int gw_bits(const int *moduli, int n, const int *p32, int words, const unsigned *bits, int top_set, long long count,
            unsigned long long seed, unsigned long long stream, int *out, long long ld, void *s) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_bits, dim3(blocks(count)), dim3(256), 0, (hipStream_t)s, moduli, n, p32, words, bits, top_set,
                       count, seed, stream, out, ld);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
int gw_place(int n, int nvals, const int *vals, const long long *idx, int *out, long long ld, void *s) {
    if (nvals <= 0) return 0;
    hipLaunchKernelGGL(k_place, dim3((nvals + 255) / 256), dim3(256), 0, (hipStream_t)s, n, nvals, vals, idx, out, ld);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
int gw_uniform_u32(long long count, unsigned lo, unsigned hi, unsigned long long seed, unsigned long long stream,
                   unsigned *out, void *s) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_uniform_u32, dim3(blocks(count)), dim3(256), 0, (hipStream_t)s, count, lo, hi, seed, stream,
                       out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
