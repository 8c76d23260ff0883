// Test-infrastructure harness: exposes the reference's own degree-2 sumcheck
// update (OZ/src/VPD/fft_circuit_GKR.cpp:171-209, compiled in place from
// /root/reference by oracle/Makefile) through a C ABI, with the tables and
// challenges injected, so tests can compare it with the restated oracle and
// with the HIP kernels.  Never part of the product.
#include <cstdint>
#include <cstddef>
#include "linear_gkr/prime_field.h"
#include "linear_gkr/polynomial.h"

extern linear_poly *add_mult_sum, *V_mult_add, *addV_array;
void init_array(int max_bit_length);
quadratic_poly sumcheck_phase1_update(prime_field::field_element previous_random,
                                      int current_bit, int total_uv);

static prime_field::field_element to_fe(const unsigned long long* p)
{
    prime_field::field_element x;   // Orion's layout is {img, real}
    x.real = p[0];
    x.img = p[1];
    return x;
}
static void from_fe(const prime_field::field_element& x, unsigned long long* p)
{
    p[0] = x.real;
    p[1] = x.img;
}

static int g_cap = -1;

// f, g: 2^nv interleaved (re, im); chal: nv interleaved challenges r_0..r_{nv-1};
// polys: nv x (a, b, c) interleaved; final4: f~(r), g~(r).
extern "C" void ref_sc2(const unsigned long long* f, const unsigned long long* g, int nv,
                        const unsigned long long* chal, unsigned long long* polys,
                        unsigned long long* final4)
{
    prime_field::init();
    if (nv > g_cap) { init_array(nv); g_cap = nv; }
    const size_t n = (size_t)1 << nv;
    const prime_field::field_element zero(0);
    for (size_t i = 0; i < n; ++i) {
        V_mult_add[i] = linear_poly(to_fe(f + 2 * i));
        add_mult_sum[i] = linear_poly(to_fe(g + 2 * i));
        addV_array[i] = linear_poly(zero, zero);
    }
    prime_field::field_element prev(0);
    for (int r = 0; r < nv; ++r) {
        quadratic_poly q = sumcheck_phase1_update(prev, r, 1 << (nv - r));
        from_fe(q.a, polys + 6 * r);
        from_fe(q.b, polys + 6 * r + 2);
        from_fe(q.c, polys + 6 * r + 4);
        prev = to_fe(chal + 2 * r);
    }
    from_fe(V_mult_add[0].eval(prev), final4);
    from_fe(add_mult_sum[0].eval(prev), final4 + 2);
}
