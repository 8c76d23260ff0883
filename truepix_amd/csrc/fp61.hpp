// fp61.hpp - arithmetic in Fp (p = 2^61 - 1) and Fp2 = Fp[i]/(i^2 + 1) for
// CDNA4 (gfx950) device code and the host side of the library.
//
// Semantics follow the reference field (TP/native_fp2/fp61.cuh:14-166): every
// public input and output limb is canonical in [0, p); i^2 = -1; products are
// exact.  The multiply is written for a GPU with 32x32->64 multiply-add
// (v_mad_u64_u32) and no 64x64 multiplier: operands are split into a 32-bit
// low limb and a high limb of at most 30 bits, and the three partial products
// are folded with 2^61 = 1 and 2^64 = 8 (mod p), so no 128-bit value is formed.
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TP_HD __host__ __device__ __forceinline__
#else
// plain C++ (the CPU check of the arithmetic below builds this header alone)
#define TP_HD inline
#endif

namespace tp {

typedef uint64_t u64;
typedef uint32_t u32;

static constexpr u64 P = (1ULL << 61) - 1;

TP_HD u64 canon(u64 x) { return x >= P ? x - P : x; }           // x < 2p
TP_HD u64 fadd(u64 a, u64 b) { return canon(a + b); }
TP_HD u64 fsub(u64 a, u64 b) { return canon(a + P - b); }
TP_HD u64 fneg(u64 a) { return a ? P - a : 0; }
TP_HD u64 fred64(u64 x) { return canon((x & P) + (x >> 61)); }  // any u64

// a * b mod p from 32-bit limbs, as the device computes it (no 64 x 64
// multiplier): a value in [0, 2^61 + 6], congruent, not yet canonical;
// a, b < 2^62 (so a canonical value plus another canonical value is allowed).
TP_HD u64 fmul_lazy_limbs(u64 a, u64 b)
{
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32);   // a1 < 2^30
    const u32 b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 lo = (u64)a0 * b0;                  // < 2^64
    const u64 mid = (u64)a0 * b1 + (u64)a1 * b0;  // < 2^63
    const u64 hi = (u64)a1 * b1;                  // < 2^60
    // a*b = hi*2^64 + mid*2^32 + lo
    //     = 8 hi + (mid >> 29) + ((mid mod 2^29) << 32) + (lo >> 61) + (lo mod 2^61)
    const u64 s = (lo & P) + (lo >> 61) + ((mid & ((1ULL << 29) - 1)) << 32)
                + (mid >> 29) + (hi << 3);        // < 2^64
    return (s & P) + (s >> 61);
}

// a * b mod p as a value in [0, 2^61 + 6] (congruent, not yet canonical);
// a, b < 2^62.  Device code has no 64 x 64 multiplier and uses the limb form
// above; host code uses the 128-bit product.
TP_HD u64 fmul_lazy(u64 a, u64 b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return fmul_lazy_limbs(a, b);
#else
    const unsigned __int128 t = (unsigned __int128)a * b;   // < 2^124
    const u64 s = ((u64)t & P) + (u64)(t >> 61);            // < 2^63 + 2^61
    return (s & P) + (s >> 61);
#endif
}

// Any u64 to its canonical residue (x mod 2^61 folded once, then < 2p).
TP_HD u64 fold61(u64 x) { return canon((x & P) + (x >> 61)); }

// One fold of any u64: at most 2^61 + 6.
TP_HD u64 fold1(u64 x) { return (x & P) + (x >> 61); }

// a, b < 2^62; result canonical.
TP_HD u64 fmul(u64 a, u64 b) { return canon(fmul_lazy(a, b)); }

struct fp2 { u64 re, im; };

TP_HD fp2 mk2(u64 r, u64 i) { fp2 x; x.re = r; x.im = i; return x; }
TP_HD fp2 zero2() { return mk2(0, 0); }
TP_HD fp2 one2() { return mk2(1, 0); }
TP_HD fp2 add2(fp2 a, fp2 b) { return mk2(fadd(a.re, b.re), fadd(a.im, b.im)); }
TP_HD fp2 sub2(fp2 a, fp2 b) { return mk2(fsub(a.re, b.re), fsub(a.im, b.im)); }
TP_HD fp2 dbl2(fp2 a) { return add2(a, a); }
TP_HD bool eq2(fp2 a, fp2 b) { return a.re == b.re && a.im == b.im; }

// Karatsuba: re = ac - bd, im = (a + b)(c + d) - ac - bd, with the three
// products left unreduced (each < 2^61 + 8) and one fold per coordinate:
// re = ac + (2p - bd) < 2^63, im = k + (4p - ac - bd) < 2^64.
TP_HD fp2 mul2(fp2 x, fp2 y)
{
    const u64 ac = fmul_lazy(x.re, y.re);
    const u64 bd = fmul_lazy(x.im, y.im);
    const u64 k = fmul_lazy(x.re + x.im, y.re + y.im);   // operands < 2^62
    return mk2(fold61(ac + (2 * P - bd)), fold61(k + (4 * P - ac - bd)));
}

// Base-field scalar times Fp2.
TP_HD fp2 smul2(u64 s, fp2 y) { return mk2(fmul(s, y.re), fmul(s, y.im)); }

// a + r (b - a): the sumcheck fold (TP/native_fp2/native_fp2_core.hpp:53-64).
// The add of a joins the Karatsuba fold, one reduction per coordinate:
// re = a.re + ac + (2p - bd) < 2^64, im = a.im + k + (4p - ac - bd) < 2^64.
TP_HD fp2 fold2(fp2 a, fp2 b, fp2 r)
{
    const fp2 d = sub2(b, a);
    const u64 ac = fmul_lazy(r.re, d.re);
    const u64 bd = fmul_lazy(r.im, d.im);
    const u64 k = fmul_lazy(r.re + r.im, d.re + d.im);
    return mk2(fold61(a.re + ac + (2 * P - bd)), fold61(a.im + k + (4 * P - ac - bd)));
}

// Two rounds of the fold on one quad: fold2(fold2(p0, p1, r0),
// fold2(p2, p3, r0), r1) = p0 + r0 d1 + r1 d2 + r01 d3 with d1 = p1 - p0,
// d2 = p2 - p0, d3 = p3 - p2 - p1 + p0 and r01 = r0 r1.  Three Karatsuba
// products (MUL: a lazy multiply, each result at most 2^61 + 6) summed
// unreduced and folded once, so one reduction per coordinate instead of
// three: with A = fold1(sum ac), B = fold1(sum bd) (each at most 2^61 + 2),
// re = p0 + A + (2p - B) < 2^63 + 3 and im = p0 + fold1(sum k) + (4p - A - B)
// < 2^64.
template <typename Mul>
TP_HD fp2 fold_quad_with(fp2 p0, fp2 p1, fp2 p2, fp2 p3, fp2 r0, fp2 r1, fp2 r01, Mul mul)
{
    const fp2 d1 = sub2(p1, p0), d2 = sub2(p2, p0), d3 = sub2(sub2(p3, p2), d1);
    const u64 ac = fold1(mul(r0.re, d1.re) + mul(r1.re, d2.re) + mul(r01.re, d3.re));
    const u64 bd = fold1(mul(r0.im, d1.im) + mul(r1.im, d2.im) + mul(r01.im, d3.im));
    const u64 k = fold1(mul(r0.re + r0.im, d1.re + d1.im) + mul(r1.re + r1.im, d2.re + d2.im) +
                        mul(r01.re + r01.im, d3.re + d3.im));
    return mk2(fold61(p0.re + ac + (2 * P - bd)), fold61(p0.im + k + (4 * P - ac - bd)));
}

struct LazyMul {
    TP_HD u64 operator()(u64 a, u64 b) const { return fmul_lazy(a, b); }
};

// x0 y0 + x1 y1 + x2 y2 (mod p) as a value at most 2^61 + 6, for canonical
// operands (below 2^61), with one reduction for the three products: each
// operand split as h 2^31 + l (h < 2^30, l < 2^31), so
//   x y = h_x h_y 2^62 + (h_x l_y + l_x h_y) 2^31 + l_x l_y
// and, with 2^62 = 2 and m 2^31 = (m >> 30) + (m mod 2^30) 2^31 (mod p),
// the three products' parts add up unreduced: L = sum l l < 3 2^62,
// M = sum (h l + l h) < 6 2^61, H = sum h h < 3 2^60 (12 multiply-adds);
// fold1(L) + 2 H + (M >> 30) + (M mod 2^30) 2^31 < 2^61 + 2 + 2^63 + 2^34 +
// 2^61 < 2^64, folded once more.
TP_HD u64 sum3_lazy(u64 x0, u64 y0, u64 x1, u64 y1, u64 x2, u64 y2)
{
    const u32 a0 = (u32)x0 & 0x7fffffffu, c0 = (u32)(x0 >> 31), b0 = (u32)y0 & 0x7fffffffu, d0 = (u32)(y0 >> 31);
    const u32 a1 = (u32)x1 & 0x7fffffffu, c1 = (u32)(x1 >> 31), b1 = (u32)y1 & 0x7fffffffu, d1 = (u32)(y1 >> 31);
    const u32 a2 = (u32)x2 & 0x7fffffffu, c2 = (u32)(x2 >> 31), b2 = (u32)y2 & 0x7fffffffu, d2 = (u32)(y2 >> 31);
    const u64 L = (u64)a0 * b0 + (u64)a1 * b1 + (u64)a2 * b2;
    const u64 M = (u64)c0 * b0 + (u64)a0 * d0 + (u64)c1 * b1 + (u64)a1 * d1 + (u64)c2 * b2 + (u64)a2 * d2;
    const u64 H = (u64)c0 * d0 + (u64)c1 * d1 + (u64)c2 * d2;
    return fold1(fold1(L) + 2 * H + (M >> 30) + ((M & ((1ULL << 30) - 1)) << 31));
}

// fold_quad_with's polynomial with each Karatsuba sum of three products
// from sum3_lazy: the (a + b) operands of the third sum are reduced to
// canonical first, so every operand is below 2^61.  Each of ac, bd and k is
// at most 2^61 + 6, within the final bounds above.
TP_HD fp2 fold_quad(fp2 p0, fp2 p1, fp2 p2, fp2 p3, fp2 r0, fp2 r1, fp2 r01)
{
    const fp2 d1 = sub2(p1, p0), d2 = sub2(p2, p0), d3 = sub2(sub2(p3, p2), d1);
    const u64 ac = sum3_lazy(r0.re, d1.re, r1.re, d2.re, r01.re, d3.re);
    const u64 bd = sum3_lazy(r0.im, d1.im, r1.im, d2.im, r01.im, d3.im);
    const u64 k = sum3_lazy(fadd(r0.re, r0.im), fadd(d1.re, d1.im), fadd(r1.re, r1.im),
                            fadd(d2.re, d2.im), fadd(r01.re, r01.im), fadd(d3.re, d3.im));
    return mk2(fold61(p0.re + ac + (2 * P - bd)), fold61(p0.im + k + (4 * P - ac - bd)));
}

// Operand of mul31: a < 2^62 as a = h 2^31 + l with l, h < 2^31.
struct Sp31 { u32 l, h; };
TP_HD Sp31 sp31(u64 a)
{
    Sp31 s;
    s.l = (u32)a & 0x7fffffffu;
    s.h = (u32)(a >> 31);
    return s;
}

// a b + c (mod p) as an unreduced u64, for a, b < 2^62 and c <= 2^61 + 2^31.
// With 2^62 = 2 and 2^61 = 1 (mod p):
//   a b = 2 h_a h_b + m 2^31 + l_a l_b,  m = h_a l_b + l_a h_b < 2^63,
//   m 2^31 = (m >> 30) + (m mod 2^30) 2^31  (mod p),
// 32 x 32 -> 64 multiply-adds and no fold.  Term by term the sum before
// c is below 7 2^61 - 2^32 (l l < 2^62, 2 h h < 2^63, m >> 30 < 2^33,
// (m mod 2^30) 2^31 < 2^61), so with c it stays below 2^64; for canonical
// a, b (< 2^61, so h < 2^30) it is below 4p.
TP_HD u64 mul31(Sp31 a, Sp31 b, u64 c)
{
    const u64 m = (u64)a.h * b.l + (u64)a.l * b.h;
    // 2 h_a < 2^32 (h_a < 2^31), so the doubling stays in 32 bits, and the
    // shift by 31 is a multiply-add by 2^31: six multiply-adds in all
    const u64 x = (u64)(2u * a.h) * b.h + (m >> 30);
    const u64 y = (u64)((u32)m & 0x3fffffffu) * 0x80000000u + x;
    return (u64)a.l * b.l + y + c;
}


// Sum of Fp2 products kept lazily: the Karatsuba products ac, bd and
// k = (a + b)(c + d) summed separately, each product added unreduced (mul31)
// and the sum folded once per add, so it stays at most 2^61 + 6; value()
// reduces once at the end.  Operands are canonical (< 2^61), so the
// Karatsuba sums are below 2^62.
struct LazyProdSum {
    u64 ac, bd, k;
    TP_HD void zero() { ac = bd = k = 0; }
    TP_HD void mac(fp2 x, fp2 y)
    {
        ac = fold1(mul31(sp31(x.re), sp31(y.re), ac));
        bd = fold1(mul31(sp31(x.im), sp31(y.im), bd));
        k = fold1(mul31(sp31(x.re + x.im), sp31(y.re + y.im), k));
    }
    TP_HD fp2 value() const
    {
        const u64 a = fold61(ac), b = fold61(bd), kk = fold61(k);
        return mk2(fsub(a, b), fsub(fsub(kk, a), b));
    }
};

// Gate functions, type indices as in TP/libTruePix/native_fp2_bridge.py:450-472
// and TP/native_fp2/native_fp2_core.hpp:32-51.
enum GateType {
    G_MUL = 0, G_ADD = 1, G_SUB = 2, G_MUXL = 3, G_MUXR = 4, G_OR = 5, G_XOR = 6,
    G_NOT = 7, G_NAND = 8, G_NOR = 9, G_NXOR = 10, G_NAAB = 11, G_PASS = 12
};

TP_HD fp2 gatefn(int tp, fp2 x, fp2 y)
{
    switch (tp) {
    case G_MUL: return mul2(x, y);
    case G_ADD: return add2(x, y);
    case G_SUB: return sub2(x, y);
    case G_MUXL: return x;
    case G_MUXR: return y;
    case G_OR: return sub2(add2(x, y), mul2(x, y));
    case G_XOR: return sub2(add2(x, y), dbl2(mul2(x, y)));
    case G_NOT: return sub2(one2(), x);
    case G_NAND: return sub2(one2(), mul2(x, y));
    case G_NOR: return sub2(add2(one2(), mul2(x, y)), add2(x, y));
    case G_NXOR: return sub2(add2(one2(), dbl2(mul2(x, y))), add2(x, y));
    case G_NAAB: return mul2(sub2(one2(), x), y);
    default: return x;   // G_PASS and unknown types pass x, as the reference does
    }
}

}  // namespace tp
