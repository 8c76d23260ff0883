// Test-infrastructure harness: exposes the reference's Reed-Solomon FFT
// (OZ/src/infrastructure/RS_polynomial.cpp:46-130) and root of unity
// (OZ/src/linear_gkr/prime_field.cpp:60-73), compiled in place from
// /root/reference by oracle/Makefile, through a C ABI.  Never part of the product.
#include <cstdint>
#include <vector>
#include "infrastructure/RS_polynomial.h"

extern "C" void ref_root_of_unity(int log_order, unsigned long long out[2])
{
    prime_field::init();
    prime_field::field_element w = prime_field::get_root_of_unity(log_order);
    out[0] = w.real;
    out[1] = w.img;
}

// coef: 2^log_coef interleaved (re, im); out: 2^log_order interleaved.
extern "C" void ref_fft(const unsigned long long* coef, int log_coef, int log_order,
                        unsigned long long* out)
{
    prime_field::init();
    const long long m = 1LL << log_coef, n = 1LL << log_order;
    init_scratch_pad(n);
    std::vector<prime_field::field_element> c(m), r(n);
    for (long long i = 0; i < m; ++i) { c[i].real = coef[2 * i]; c[i].img = coef[2 * i + 1]; }
    fast_fourier_transform(c.data(), (int)m, (int)n, prime_field::get_root_of_unity(log_order), r.data());
    for (long long i = 0; i < n; ++i) { out[2 * i] = r[i].real; out[2 * i + 1] = r[i].img; }
    delete_scratch_pad();
}
