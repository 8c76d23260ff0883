// CPU check of the device field arithmetic in truepix_amd/csrc/fp61.hpp
// (built with plain g++, so the integer code the GPU runs is exercised as
// written): mul31 against 128-bit products, including the worst-case
// operands that set its no-overflow bounds, LazyProdSum against an exact Fp2
// dot product, the sumcheck fold against a + r (b - a), the device's limb
// multiply against 128-bit products, and the two-round quad fold against
// two single folds and the exact polynomial, and the first four sumcheck
// rounds from 256 block products (blocks.hpp) against rounds taken directly
// on folded tables.  Prints each failure; exit status 1 on any.
#include "blocks.hpp"
#include "fp61.hpp"
#include <cstdio>
#include <vector>

using namespace tp;
typedef unsigned __int128 u128;

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { if (++fails <= 20) { std::printf(__VA_ARGS__); std::printf("\n"); } } } while (0)

static u64 modp(u128 x) { return (u64)(x % P); }

static u64 rng_state = 0x9e3779b97f4a7c15ull;
static u64 rnd() {
    rng_state ^= rng_state << 13; rng_state ^= rng_state >> 7; rng_state ^= rng_state << 17;
    return rng_state;
}

// The terms of mul31 summed exactly, to see that the u64 sum does not wrap.
static u128 mul31_exact(u64 a, u64 b, u64 c) {
    const u128 al = a & 0x7fffffffu, ah = a >> 31, bl = b & 0x7fffffffu, bh = b >> 31;
    const u128 m = ah * bl + al * bh;
    return 2 * ah * bh + (m >> 30) + (m & 0x3fffffffu) * ((u128)1 << 31) + al * bl + c;
}

static void check_mul31(u64 a, u64 b, u64 c) {
    const u64 s = mul31(sp31(a), sp31(b), c);
    const u128 t = mul31_exact(a, b, c);
    CHECK(t < ((u128)1 << 64), "mul31 wraps: a=%llx b=%llx c=%llx", (unsigned long long)a,
          (unsigned long long)b, (unsigned long long)c);
    CHECK((u128)s == t, "mul31 differs from its terms: a=%llx b=%llx", (unsigned long long)a,
          (unsigned long long)b);
    CHECK(modp(s) == modp((u128)a * b + c), "mul31 not a b + c mod p: a=%llx b=%llx c=%llx",
          (unsigned long long)a, (unsigned long long)b, (unsigned long long)c);
    if (a < (1ull << 61) && b < (1ull << 61))
        CHECK(t - c < (u128)4 * P, "canonical product not below 4p: a=%llx b=%llx",
              (unsigned long long)a, (unsigned long long)b);
}

int main() {
    const u64 two61 = 1ull << 61, two62 = 1ull << 62;
    std::vector<u64> ops = {0, 1, 2, 0x7fffffffull, 0x80000000ull, 0xffffffffull, 1ull << 32,
                            (1ull << 60) - 1, P - 1, P, two61, two61 + 1, 2 * P - 1, 2 * P,
                            two62 - 2, two62 - 1, (two62 - 1) & ~0x7fffffffull,
                            (two62 - 1) ^ (1ull << 30), 0x3fffffff80000000ull, 0x3fffffff7fffffffull};
    for (int i = 0; i < 40; ++i) ops.push_back(rnd() >> 2);          // < 2^62
    for (int i = 0; i < 40; ++i) ops.push_back(rnd() % P);           // canonical
    const u64 cs[] = {0, 1, P - 1, two61 + 6, two61 + (1ull << 31)};
    for (u64 a : ops)
        for (u64 b : ops)
            for (u64 c : cs) check_mul31(a, b, c);
    for (int i = 0; i < 200000; ++i) check_mul31(rnd() >> 2, rnd() >> 2, rnd() % (two61 + (1ull << 31) + 1));

    // fold1: any u64 to at most 2^61 + 6, congruent
    const u64 xs[] = {0, P, two61, ~0ull, ~0ull - 1, 7ull << 61, (7ull << 61) | P};
    for (u64 x : xs) {
        CHECK(fold1(x) <= two61 + 6, "fold1 bound: %llx", (unsigned long long)x);
        CHECK(modp(fold1(x)) == modp(x), "fold1 value: %llx", (unsigned long long)x);
    }

    // LazyProdSum over canonical Fp2 vectors, extremes included, against the
    // exact dot product (i^2 = -1)
    for (int trial = 0; trial < 400; ++trial) {
        const int n = 1 + (int)(rnd() % 300);
        LazyProdSum z;
        z.zero();
        u64 re = 0, im = 0;
        for (int j = 0; j < n; ++j) {
            fp2 x, y;
            const int kind = (int)(rnd() % 4);
            x.re = kind == 0 ? P - 1 : rnd() % P;
            x.im = kind == 1 ? P - 1 : rnd() % P;
            y.re = kind == 2 ? P - 1 : rnd() % P;
            y.im = kind == 0 ? P - 1 : rnd() % P;
            z.mac(x, y);
            re = modp((u128)re + (u128)x.re * y.re + (u128)(P - x.im) * y.im);
            im = modp((u128)im + (u128)x.re * y.im + (u128)x.im * y.re);
        }
        const fp2 v = z.value();
        CHECK(v.re == re && v.im == im, "LazyProdSum trial %d (n=%d) differs", trial, n);
    }
    {   // every operand p - 1: the largest Karatsuba sums, many terms
        LazyProdSum z;
        z.zero();
        const fp2 x = mk2(P - 1, P - 1);
        for (int j = 0; j < 1000; ++j) z.mac(x, x);
        // (-1 - i)^2 = 2i, times 1000
        const fp2 v = z.value();
        CHECK(v.re == 0 && v.im == 2000, "LazyProdSum at p - 1 differs");
    }
    // fold2 = a + r (b - a) against the exact value, extremes included
    {
        const u64 ext[] = {0, 1, 2, P - 1, P - 2, 0x7fffffffull, 0x80000000ull, (1ull << 60) - 1,
                           1ull << 60, 0x1fffffff80000000ull, 0x1fffffff7fffffffull};
        auto check_fold = [&](fp2 a, fp2 b, fp2 r) {
            const fp2 got = fold2(a, b, r);
            // a + r (b - a), i^2 = -1
            const u64 dre = modp((u128)b.re + P - a.re), dim = modp((u128)b.im + P - a.im);
            const u64 re = modp((u128)a.re + (u128)r.re * dre + (u128)(P - r.im) * dim);
            const u64 im = modp((u128)a.im + (u128)r.re * dim + (u128)r.im * dre);
            CHECK(got.re == re && got.im == im, "fold2 differs: a=(%llx,%llx) b=(%llx,%llx) r=(%llx,%llx)",
                  (unsigned long long)a.re, (unsigned long long)a.im, (unsigned long long)b.re,
                  (unsigned long long)b.im, (unsigned long long)r.re, (unsigned long long)r.im);
        };
        for (u64 x : ext)
            for (u64 y : ext)
                for (u64 w : ext) {
                    check_fold(mk2(x, y), mk2(y, w), mk2(w, x));
                    check_fold(mk2(w, w), mk2(x, x), mk2(y, y));
                    check_fold(mk2(0, x), mk2(P - 1, y), mk2(w, P - 1));
                }
        for (int i = 0; i < 200000; ++i)
            check_fold(mk2(rnd() % P, rnd() % P), mk2(rnd() % P, rnd() % P), mk2(rnd() % P, rnd() % P));
    }
    // the device multiply from 32-bit limbs: congruent, at most 2^61 + 6, for
    // operands below 2^62 (extremes included)
    {
        auto check_limbs = [&](u64 a, u64 b) {
            const u64 got = fmul_lazy_limbs(a, b);
            CHECK(got <= two61 + 6, "fmul_lazy_limbs bound: a=%llx b=%llx", (unsigned long long)a,
                  (unsigned long long)b);
            CHECK(modp(got) == modp((u128)a * b), "fmul_lazy_limbs value: a=%llx b=%llx",
                  (unsigned long long)a, (unsigned long long)b);
        };
        for (u64 a : ops)
            for (u64 b : ops) check_limbs(a, b);
        for (int i = 0; i < 200000; ++i) check_limbs(rnd() >> 2, rnd() >> 2);
    }
    // two rounds of the fold on a quad: equal to fold2(fold2(p0, p1, r0),
    // fold2(p2, p3, r0), r1) and to the exact p0 + r0 d1 + r1 d2 + r0 r1 d3,
    // with the host multiply and with the device's limb multiply
    {
        struct LimbMul {
            u64 operator()(u64 a, u64 b) const { return fmul_lazy_limbs(a, b); }
        };
        auto exact = [](fp2 x, fp2 y) {   // x y mod p, i^2 = -1
            return mk2(modp((u128)x.re * y.re + (u128)(P - x.im) * y.im),
                       modp((u128)x.re * y.im + (u128)x.im * y.re));
        };
        auto check_quad = [&](fp2 p0, fp2 p1, fp2 p2, fp2 p3, fp2 r0, fp2 r1) {
            const fp2 r01 = mul2(r0, r1);
            const fp2 want = fold2(fold2(p0, p1, r0), fold2(p2, p3, r0), r1);
            const fp2 d1 = sub2(p1, p0), d2 = sub2(p2, p0), d3 = sub2(sub2(p3, p2), d1);
            const fp2 poly = add2(add2(p0, exact(r0, d1)), add2(exact(r1, d2), exact(r01, d3)));
            const fp2 a = fold_quad(p0, p1, p2, p3, r0, r1, r01);
            const fp2 b = fold_quad_with(p0, p1, p2, p3, r0, r1, r01, LimbMul());
            CHECK(eq2(want, poly), "two single folds differ from the quad polynomial");
            CHECK(eq2(a, want) && eq2(b, want),
                  "fold_quad differs: p0=(%llx,%llx) r0=(%llx,%llx) r1=(%llx,%llx)",
                  (unsigned long long)p0.re, (unsigned long long)p0.im, (unsigned long long)r0.re,
                  (unsigned long long)r0.im, (unsigned long long)r1.re, (unsigned long long)r1.im);
        };
        const u64 ext[] = {0, 1, P - 1, P - 2, 0x7fffffffull, 0xffffffffull, 1ull << 60,
                           0x1fffffff80000000ull, 0x1fffffffffffffffull - 1};
        for (u64 x : ext)
            for (u64 y : ext)
                for (u64 w : ext) {
                    check_quad(mk2(x, y), mk2(y, w), mk2(w, x), mk2(P - 1, 0), mk2(y, x), mk2(w, w));
                    check_quad(mk2(0, 0), mk2(x, x), mk2(P - 1, P - 1), mk2(y, w), mk2(P - 1, P - 1),
                               mk2(x, P - 1));
                }
        for (int i = 0; i < 100000; ++i) {
            auto r = [] { return mk2(rnd() % P, rnd() % P); };
            check_quad(r(), r(), r(), r(), r(), r());
        }
    }
    {
        // three products summed with one reduction, against 128-bit sums,
        // at the operand extremes and at random
        auto check3 = [&](u64 a, u64 b, u64 c, u64 d, u64 e, u64 f) {
            const u64 got = sum3_lazy(a, b, c, d, e, f);
            const u128 want = ((u128)a * b + (u128)c * d + (u128)e * f) % P;
            CHECK(got <= P + 6 && got % P == (u64)want, "sum3_lazy(%llx,%llx,%llx,%llx,%llx,%llx)",
                  (unsigned long long)a, (unsigned long long)b, (unsigned long long)c,
                  (unsigned long long)d, (unsigned long long)e, (unsigned long long)f);
        };
        const u64 ex[] = {0, 1, P - 1, P - 2, 0x7fffffffull, 0x80000000ull, (1ull << 60) - 1,
                          0x1fffffff80000000ull, 0x1fffffff7fffffffull};
        for (u64 a : ex) for (u64 b : ex) for (u64 c : ex) check3(a, b, c, b, a, c), check3(a, b, a, b, a, b);
        for (int i = 0; i < 200000; ++i) check3(rnd() % P, rnd() % P, rnd() % P, rnd() % P, rnd() % P, rnd() % P);
    }
    {
        // rounds 0..3 from the block products against the rounds on the
        // tables themselves, folded entry by entry with exact products
        auto xmul = [](fp2 x, fp2 y) {
            const u64 re = modp((u128)x.re * y.re + (u128)(P - y.im) * x.im);
            const u64 im = modp((u128)x.re * y.im + (u128)x.im * y.re);
            return mk2(re, im);
        };
        auto xfold = [&](fp2 a, fp2 b, fp2 r) { return add2(a, xmul(r, sub2(b, a))); };
        for (int trial = 0; trial < 40; ++trial) {
            const int nblk = 1 + trial % 7;
            const int n = 16 * nblk;
            std::vector<fp2> f(n), g(n);
            const u64 ext[] = {0, 1, P - 1, 0x7f7f7f7f7f7f7f7full & P, 0x0080808080808080ull};
            for (int i = 0; i < n; ++i) {
                auto v = [&]() { return trial % 3 == 0 ? ext[rnd() % 5] : rnd() % P; };
                f[i] = mk2(v(), v());
                g[i] = mk2(v(), v());
            }
            fp2 Pm[16][16];
            for (int a = 0; a < 16; ++a)
                for (int b = 0; b < 16; ++b) {
                    fp2 s = zero2();
                    for (int B = 0; B < nblk; ++B) s = add2(s, xmul(f[16 * B + a], g[16 * B + b]));
                    Pm[a][b] = s;
                }
            // challenges: special values in some trials, else drawn
            std::vector<fp2> chal(4);
            for (int k = 0; k < 4; ++k) {
                const u64 sv[] = {0, 1, P - 1};
                chal[k] = trial % 4 == 1 ? mk2(sv[(k + trial) % 3], sv[k % 3]) : mk2(rnd() % P, rnd() % P);
            }
            fp2 got[4][3];
            auto rd = [&](int k, fp2 E0, fp2 E1, fp2 EA) {
                got[k][0] = E0;
                got[k][1] = E1;
                got[k][2] = EA;
                return chal[k];
            };
            fp2 rr[4];
            rounds_from_blocks(Pm, rd, rr);
            std::vector<fp2> tf = f, tg = g;
            for (int k = 0; k < 4; ++k) {
                fp2 E0 = zero2(), E1 = zero2(), EA = zero2();
                const size_t half = tf.size() / 2;
                for (size_t j = 0; j < half; ++j) {
                    E0 = add2(E0, xmul(tf[2 * j], tg[2 * j]));
                    E1 = add2(E1, xmul(tf[2 * j + 1], tg[2 * j + 1]));
                    EA = add2(EA, xmul(sub2(tf[2 * j + 1], tf[2 * j]), sub2(tg[2 * j + 1], tg[2 * j])));
                }
                CHECK(eq2(E0, got[k][0]) && eq2(E1, got[k][1]) && eq2(EA, got[k][2]),
                      "block rounds: trial %d round %d differs", trial, k);
                CHECK(eq2(rr[k], chal[k]), "block rounds: challenge %d not returned", k);
                std::vector<fp2> nf(half), ng(half);
                for (size_t j = 0; j < half; ++j) {
                    nf[j] = xfold(tf[2 * j], tf[2 * j + 1], chal[k]);
                    ng[j] = xfold(tg[2 * j], tg[2 * j + 1], chal[k]);
                }
                tf.swap(nf);
                tg.swap(ng);
            }
        }
    }
    std::printf("field check: %d failures\n", fails);
    return fails ? 1 : 0;
}
