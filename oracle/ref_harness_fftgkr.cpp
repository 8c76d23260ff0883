// Test-infrastructure harness: drives the reference's own fft_gkr
// (OZ/src/VPD/fft_circuit_GKR.cpp, compiled in place from /root/reference by
// oracle/Makefile) through its build_circuit and layer functions in the order
// of its engage_gkr (:786-825), and exposes what they leave behind: the claim
// after every layer, the final alpha / beta, the pass flag, the proof size and
// the circuit's last IFFT, scaled and sum layers.  Never part of the product.
#include <cstdint>
#include <cstdlib>
#include <vector>
#include "linear_gkr/prime_field.h"
#include "linear_gkr/polynomial.h"

class circuit
{
public:
    std::vector<prime_field::field_element*> circuit_val;
    std::vector<int> size;
};
extern circuit C;
extern prime_field::field_element alpha, beta;
extern int proof_size;
void init_array(int max_bit_length);
void build_circuit(int lg_size, prime_field::field_element* r);
void refresh_randomness(prime_field::field_element* r, prime_field::field_element* one_minus_r, int size);
int mylog(long long x);
prime_field::field_element V_output(const prime_field::field_element* one_minus_r_0,
                                    const prime_field::field_element* r_0,
                                    const prime_field::field_element* output_raw, int r_0_size,
                                    int output_size);
bool addition_layer(prime_field::field_element&, prime_field::field_element*, int, int,
                    prime_field::field_element*, prime_field::field_element*,
                    prime_field::field_element*, prime_field::field_element*,
                    prime_field::field_element*, prime_field::field_element*,
                    prime_field::field_element*, prime_field::field_element*);
bool mult_layer(prime_field::field_element&, prime_field::field_element*, int, int,
                prime_field::field_element*, prime_field::field_element*,
                prime_field::field_element*, prime_field::field_element*,
                prime_field::field_element*, prime_field::field_element*,
                prime_field::field_element*, prime_field::field_element*);
bool intermediate_layer(prime_field::field_element&, prime_field::field_element*, int,
                        prime_field::field_element*, prime_field::field_element*,
                        prime_field::field_element*, prime_field::field_element*,
                        prime_field::field_element*, prime_field::field_element*,
                        prime_field::field_element*, prime_field::field_element*);
bool ifft_gkr(prime_field::field_element&, int, prime_field::field_element*,
              prime_field::field_element*, prime_field::field_element*,
              prime_field::field_element*, prime_field::field_element*,
              prime_field::field_element*, prime_field::field_element*,
              prime_field::field_element*);
bool extension_gkr(prime_field::field_element&, int, prime_field::field_element*,
                   prime_field::field_element*, prime_field::field_element*,
                   prime_field::field_element*, prime_field::field_element*,
                   prime_field::field_element*, prime_field::field_element*,
                   prime_field::field_element*);

static void put(const prime_field::field_element& x, unsigned long long* p)
{
    p[0] = x.real;   // Orion's layout is {img, real}
    p[1] = x.img;
}

extern "C" int ref_fft_gkr(int lg, unsigned long long* claims, unsigned long long* layers,
                           int* psize)
{
    typedef prime_field::field_element fe;
    srand(1);                         // the reference never seeds rand()
    prime_field::init();
    proof_size = 0;
    for (fe* p : C.circuit_val) delete[] p;
    C.circuit_val.clear();
    C.size.clear();
    init_array(lg + 6);
    std::vector<fe> r(lg);
    for (int i = 0; i < lg; ++i) r[i] = prime_field::random();
    build_circuit(lg, r.data());
    alpha = fe(1);
    beta = fe(0);
    const int nr = lg + 10;
    std::vector<fe> r0(nr), o0(nr), r1(nr), o1(nr), ru(nr), ou(nr), rv(nr), ov(nr);
    refresh_randomness(r0.data(), o0.data(), nr);
    refresh_randomness(r1.data(), o1.data(), nr);
    const int S = (int)C.circuit_val.size();
    fe abs = V_output(o0.data(), r0.data(), C.circuit_val[S - 1], mylog(C.size[S - 1]), C.size[S - 1]);
    put(abs, claims);
    bool ok = true;
    ok &= addition_layer(abs, C.circuit_val[S - 2], C.size[S - 2] / 64, 64, r0.data(), o0.data(),
                         r1.data(), o1.data(), ru.data(), ou.data(), rv.data(), ov.data());
    put(abs, claims + 2);
    ok &= mult_layer(abs, C.circuit_val[S - 3], C.size[S - 3], 64, r0.data(), o0.data(), r1.data(),
                     o1.data(), ru.data(), ou.data(), rv.data(), ov.data());
    put(abs, claims + 4);
    ok &= intermediate_layer(abs, C.circuit_val[S - 4], C.size[S - 4], r0.data(), o0.data(),
                             r1.data(), o1.data(), ru.data(), ou.data(), rv.data(), ov.data());
    put(abs, claims + 6);
    ok &= ifft_gkr(abs, C.size[S - 4], r0.data(), o0.data(), r1.data(), o1.data(), ru.data(),
                   ou.data(), rv.data(), ov.data());
    put(abs, claims + 8);
    ok &= extension_gkr(abs, C.size[S - 4], r0.data(), o0.data(), r1.data(), o1.data(), ru.data(),
                        ou.data(), rv.data(), ov.data());
    put(alpha, claims + 10);
    put(beta, claims + 12);
    const size_t n = (size_t)1 << lg;
    if (layers) {
        for (size_t t = 0; t < n; ++t) put(C.circuit_val[S - 4][t], layers + 2 * t);
        for (size_t t = 0; t < n; ++t) put(C.circuit_val[S - 3][t], layers + 2 * n + 2 * t);
        for (int t = 0; t < 64; ++t) put(C.circuit_val[S - 1][t], layers + 4 * n + 2 * t);
    }
    *psize = proof_size;
    return ok ? 1 : 0;
}
