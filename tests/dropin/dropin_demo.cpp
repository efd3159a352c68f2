// A program written against the GRNS 1.1.4 API (a params.h, the reference's types and function
// names), built against include/grns_dropin.hpp and libgrns_hip.so instead of the GRNS headers.
// tests/test_dropin.py compiles it once per moduli set (-I<dir holding that set's params.h>) and
// compares what it writes with the oracle.
//
// usage: dropin_demo <input file> <output file>
// input:  int32 N, n, Nm; x, y, d: N numbers of n int32 residues each (number-major, the layout of
//         the reference's test programs); D: N uint32 exponents; mx, my: Nm mp_int_t
// output: the arg-max residues (n int32); eval, eval_fast (N interval_t each); cmp (N int32);
//         rns_div q, r; rns_div_fast q, r; scale2pow (N numbers each); mpint add, sub, mul, div
//         (Nm mp_int_t each); mpint add computed in place (z == x, Nm mp_int_t)
#include "params.h"
#include "grns_dropin.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

template <class T>
static T *to_device(const std::vector<T> &h) {
    T *d = nullptr;
    if (hipMalloc(&d, sizeof(T) * (h.size() ? h.size() : 1)) != hipSuccess) std::exit(2);
    if (!h.empty() && hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess) std::exit(2);
    return d;
}

template <class T>
static void write_back(std::FILE *f, const T *d, size_t count) {
    std::vector<T> h(count ? count : 1);
    if (count && hipMemcpy(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost) != hipSuccess) std::exit(2);
    if (count && std::fwrite(h.data(), sizeof(T), count, f) != count) std::exit(3);
}

template <class T>
static std::vector<T> read_vec(std::FILE *f, size_t count) {
    std::vector<T> v(count);
    if (count && std::fread(v.data(), sizeof(T), count, f) != count) std::exit(4);
    return v;
}

int main(int argc, char **argv) {
    if (argc != 3) return 1;
    std::FILE *in = std::fopen(argv[1], "rb");
    if (!in) return 1;
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, in) != 3) return 4;
    const size_t N = (size_t)hdr[0], Nm = (size_t)hdr[2];
    if (hdr[1] != RNS_MODULI_SIZE) return 5;
    const size_t n = RNS_MODULI_SIZE;
    auto hx = read_vec<int>(in, N * n), hy = read_vec<int>(in, N * n), hd = read_vec<int>(in, N * n);
    auto hD = read_vec<unsigned>(in, N);
    auto hmx = read_vec<mp_int_t>(in, Nm), hmy = read_vec<mp_int_t>(in, Nm);
    std::fclose(in);

    rns_const_init();
    mpint_const_init();

    int *x = to_device(hx), *y = to_device(hy), *d = to_device(hd);
    unsigned *D = to_device(hD);
    mp_int_t *mx = to_device(hmx), *my = to_device(hmy), *mz = to_device(hmx), *mw = to_device(hmx);
    int *out = nullptr, *sign = nullptr, *q = nullptr, *r = nullptr;
    interval_t *ev = nullptr;
    xinterval_t *buffer = nullptr;
    if (hipMalloc(&out, sizeof(int) * n) != hipSuccess || hipMalloc(&sign, sizeof(int) * N) != hipSuccess ||
        hipMalloc(&q, sizeof(int) * N * n) != hipSuccess || hipMalloc(&r, sizeof(int) * N * n) != hipSuccess ||
        hipMalloc(&ev, sizeof(interval_t) * N) != hipSuccess || hipMalloc(&buffer, sizeof(xinterval_t) * N) != hipSuccess)
        return 2;

    std::FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    cuda::rns_max<64, 64, 64, 64>(out, x, (unsigned)N, buffer);  // src/dpp/rnsmax.cuh:129
    write_back(o, out, n);
    cuda::rns_eval_compute_array(ev, x, N);
    write_back(o, ev, N);
    cuda::rns_eval_compute_fast_array(ev, x, N);
    write_back(o, ev, N);
    cuda::rns_cmp_array(sign, x, y, N);
    write_back(o, sign, N);
    cuda::rns_div_array(q, r, x, d, N);
    write_back(o, q, N * n);
    write_back(o, r, N * n);
    cuda::rns_div_fast_array(q, r, x, d, N);
    write_back(o, q, N * n);
    write_back(o, r, N * n);
    cuda::rns_scale2pow_array(q, x, D, N);
    write_back(o, q, N * n);
    cuda::mpint_add_array(mz, mx, my, Nm);
    write_back(o, mz, Nm);
    cuda::mpint_sub_array(mz, mx, my, Nm);
    write_back(o, mz, Nm);
    cuda::mpint_mul_array(mz, mx, my, Nm);
    write_back(o, mz, Nm);
    cuda::mpint_div_array(mz, mx, my, Nm);
    write_back(o, mz, Nm);
    cuda::mpint_add_array(mw, mw, my, Nm);  // in place: the result overwrites x
    write_back(o, mw, Nm);
    std::fclose(o);
    return 0;
}
