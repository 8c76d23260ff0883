// blocks.hpp - the first four sumcheck rounds from the 256 block products of
// the input, for the degree-2 prover (host side; plain C++, so the CPU check
// of the arithmetic builds it too).
//
// With variables bound LSB-first, entry e of T_k inside a 16-entry block of
// T_0 (k <= 4) is a linear combination c_e . x of the block's 16 entries:
// c_e = unit vector e for k = 0, and T_{k+1}[p] = T_k[2p] + r_k (T_k[2p+1] -
// T_k[2p]) gives c'_p = c_{2p} + r_k (c_{2p+1} - c_{2p}), supported on the
// 2^{k+1} entries [2^{k+1} p, 2^{k+1} (p + 1)).  Round k's three sums over its
// pairs (E0 = sum f_lo g_lo, E1 = sum f_hi g_hi, EA = sum (f_hi - f_lo)
// (g_hi - g_lo), as TP/native_fp2/native_fp2.cu:556-640 sums them) are
// bilinear forms in the block products P[a][b] = sum over blocks of
// f[16B + a] g[16B + b]:
//   sum over blocks of (c . f)(d . g) = sum_{a, b} c_a d_b P[a][b].
// So the 256 products give rounds 0..3 exactly, each challenge drawn (by the
// caller's transcript step) before the next round's forms are taken.
#pragma once
#include "fp61.hpp"

namespace tp {

// sum_i x_i y_i in Fp2 for canonical x, y and n <= 16: the products summed
// exactly in 128 bits (re = sum x.re y.re + x.im (p - y.im), im = sum x.re
// y.im + x.im y.re; 2n terms below 2^122 each, so below 2^127) and reduced
// once, instead of reducing every product.
inline fp2 dot2(const fp2* x, const fp2* y, int n)
{
    typedef unsigned __int128 u128;
    u128 re = 0, im = 0;
    for (int i = 0; i < n; ++i) {
        re += (u128)x[i].re * y[i].re + (u128)x[i].im * (P - y[i].im);
        im += (u128)x[i].re * y[i].im + (u128)x[i].im * y[i].re;
    }
    auto red = [](u128 v) {
        const u64 t = (u64)(v & P) + (u64)((v >> 61) & P) + (u64)(v >> 122);   // < 3p
        return canon(canon(t >= 2 * P ? t - P : t));
    };
    return mk2(red(re), red(im));
}

// Rounds 0..3 from the block products: round_done(k, E0, E1, EA) records
// round k and returns r_k; the four challenges land in r.  Per pair of T_k
// entries (2p, 2p + 1), with lo and hi their coefficients on the joint
// support S of 2 sup block entries (lo on the first sup, hi on the rest):
//   Ylo[a] = sum_b lo_b P[a][b],  Yhi[a] = sum_b hi_b P[a][b]  (a in S),
//   E0 += lo . Ylo,  E1 += hi . Yhi,  EA += (hi - lo) . (Yhi - Ylo),
// 4 sup^2 + 4 sup products a pair, about 600 for the four rounds, each dot
// product (dot2) reduced once.
template <typename RoundDone>
void rounds_from_blocks(const fp2 (&P)[16][16], RoundDone round_done, fp2 (&r)[4])
{
    // c[e][0 .. sup): the coefficients of T_k entry e on its support, which
    // starts at e * sup
    fp2 c[16][16];
    for (int e = 0; e < 16; ++e) c[e][0] = one2();
    int entries = 16, sup = 1;
    for (int k = 0; k < 4; ++k) {
        fp2 E0 = zero2(), E1 = zero2(), EA = zero2();
        for (int p = 0; p < entries / 2; ++p) {
            const int base = 2 * p * sup;
            const fp2* lo = c[2 * p];
            const fp2* hi = c[2 * p + 1];
            fp2 ylo[16], yhi[16], d[16], dy[16];
            for (int a = 0; a < 2 * sup; ++a) {
                ylo[a] = dot2(lo, &P[base + a][base], sup);
                yhi[a] = dot2(hi, &P[base + a][base + sup], sup);
                d[a] = a < sup ? sub2(zero2(), lo[a]) : hi[a - sup];
                dy[a] = sub2(yhi[a], ylo[a]);
            }
            E0 = add2(E0, dot2(lo, ylo, sup));
            E1 = add2(E1, dot2(hi, yhi + sup, sup));
            EA = add2(EA, dot2(d, dy, 2 * sup));
        }
        r[k] = round_done(k, E0, E1, EA);
        for (int p = 0; p < entries / 2; ++p) {
            fp2 nc[16];
            for (int u = 0; u < 2 * sup; ++u) {
                const fp2 a = u < sup ? c[2 * p][u] : zero2();
                const fp2 b = u < sup ? zero2() : c[2 * p + 1][u - sup];
                nc[u] = fold2(a, b, r[k]);
            }
            for (int u = 0; u < 2 * sup; ++u) c[p][u] = nc[u];
        }
        entries /= 2;
        sup *= 2;
    }
}

}  // namespace tp
