/*
 * fp2_oracle.c - CPU restatement of the reference's Fp2 / GKR / sumcheck / NTT
 * arithmetic.  THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, always as
 * the checker or as the timed CPU baseline, never as the thing measured or
 * shipped.  The product library (truepix_amd/lib/libtruepix_hip.so) does not
 * link or call anything here.
 *
 * Field: p = 2^61 - 1, Fp2 = Fp[i]/(i^2 + 1), every limb canonical in [0, p).
 * Canonical representatives are unique, so any correct evaluation order gives
 * the same bits; this file deliberately uses schoolbook Fp2 multiplication and
 * a 128-bit product, where the reference uses Karatsuba and split limbs, so it
 * is an independent restatement rather than a transliteration.
 *
 * Parity status of each function (how it is pinned; see tests/):
 *   or_fp2_*            A1  closed forms (Python big-int) + oracle/_ref build
 *   or_gatefn           A2  truth tables + oracle/_ref build
 *   or_fold/or_expand   A7  MLE definition + oracle/_ref build
 *   or_sumcheck_early   A6  oracle/_ref build + s(0)+s(1) invariant
 *   or_compute_beta     A5  closed-form product + oracle/_ref build
 *   or_mle_eval_base    A12 brute-force sum_x beta(x) in[x] + oracle/_ref
 *   or_early_claim/fold A6/A7 oracle/_ref build (session semantics)
 *   or_sc2_round/fold   A17 brute force over {0,1}^n + oracle/_ref
 *                           (sumcheck_phase1_update compiled from source)
 *   or_fft              A20 naive O(n^2) DFT + oracle/_ref
 *                           (fast_fourier_transform compiled from source)
 *   or_root_of_unity    A19 order checks + oracle/_ref
 */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned long long u64;
typedef unsigned __int128 u128;
#define P61 ((1ULL << 61) - 1)

typedef struct { u64 re, im; } fp2;

static inline u64 fcan(u64 x) { return x >= P61 ? x - P61 : x; }
static inline u64 fadd(u64 a, u64 b) { return fcan(a + b); }
static inline u64 fsub(u64 a, u64 b) { return fcan(a + P61 - b); }
static inline u64 fmul(u64 a, u64 b)
{
    u128 t = (u128)a * b;                       /* < 2^122 */
    return fcan(((u64)t & P61) + (u64)(t >> 61)); /* 2^61 == 1 (mod p) */
}
static inline fp2 mk(u64 r, u64 i) { fp2 x = {r, i}; return x; }
static inline fp2 add2(fp2 a, fp2 b) { return mk(fadd(a.re, b.re), fadd(a.im, b.im)); }
static inline fp2 sub2(fp2 a, fp2 b) { return mk(fsub(a.re, b.re), fsub(a.im, b.im)); }
/* schoolbook: (a+bi)(c+di) = (ac - bd) + (ad + bc)i, with i^2 = -1 */
static inline fp2 mul2(fp2 a, fp2 b)
{
    return mk(fsub(fmul(a.re, b.re), fmul(a.im, b.im)),
              fadd(fmul(a.re, b.im), fmul(a.im, b.re)));
}
static inline fp2 ld(const u64* re, const u64* im, size_t k) { return mk(re[k], im[k]); }
static inline void st(fp2 x, u64* re, u64* im, size_t k) { re[k] = x.re; im[k] = x.im; }

/* A1: TP/native_fp2/fp61.cuh:46-50 (mul_u64), :103-119 (mul2) */
void or_fp2_mul(u64 ar, u64 ai, u64 br, u64 bi, u64 out[2])
{
    fp2 r = mul2(mk(ar, ai), mk(br, bi));
    out[0] = r.re; out[1] = r.im;
}

/* A1: TP/native_fp2/fp61.cuh:150-157 (inverse2 via the norm) */
static u64 fpow(u64 a, u64 e)
{
    u64 r = 1;
    while (e) { if (e & 1) r = fmul(r, a); a = fmul(a, a); e >>= 1; }
    return r;
}
void or_fp2_inv(u64 ar, u64 ai, u64 out[2])
{
    u64 n = fadd(fmul(ar, ar), fmul(ai, ai));
    if (n == 0) { out[0] = out[1] = 0; return; }
    u64 ni = fpow(n, P61 - 2);
    out[0] = fmul(ar, ni);
    out[1] = fcan(P61 - fmul(ai, ni));
}

/* A2: gate functions, TP/native_fp2/native_fp2_core.hpp:32-51 and the
 * truth tables of TP/libTruePix/gateprover.py (gatefn_ of each class). */
static fp2 gatefn(int tp, fp2 x, fp2 y)
{
    const fp2 one = {1, 0};
    switch (tp) {
    case 0: return mul2(x, y);                                   /* mul  */
    case 1: return add2(x, y);                                   /* add  */
    case 2: return sub2(x, y);                                   /* sub  */
    case 3: return x;                                            /* muxL */
    case 4: return y;                                            /* muxR */
    case 5: return sub2(add2(x, y), mul2(x, y));                 /* or   */
    case 6: { fp2 xy = mul2(x, y);                               /* xor  */
              return sub2(add2(x, y), add2(xy, xy)); }
    case 7: return sub2(one, x);                                 /* not  */
    case 8: return sub2(one, mul2(x, y));                        /* nand */
    case 9: return sub2(add2(one, mul2(x, y)), add2(x, y));      /* nor  */
    case 10: { fp2 xy = mul2(x, y);                              /* nxor */
               return sub2(add2(one, add2(xy, xy)), add2(x, y)); }
    case 11: return mul2(sub2(one, x), y);                       /* naab */
    default: return x;                                           /* pass (12) */
    }
}
void or_gatefn(int tp, u64 xr, u64 xi, u64 yr, u64 yi, u64 out[2])
{
    fp2 r = gatefn(tp, mk(xr, xi), mk(yr, yi));
    out[0] = r.re; out[1] = r.im;
}

/* A7: fold, TP/native_fp2/native_fp2_core.hpp:53-64:
 * out[i] = a + r(b - a), a = in[2i], b = in[2i+1]. */
void or_fold(const u64* in_re, const u64* in_im, size_t n_out, u64 rr, u64 ri,
             u64* out_re, u64* out_im)
{
    fp2 r = mk(rr, ri);
    for (size_t i = 0; i < n_out; ++i) {
        fp2 a = ld(in_re, in_im, 2 * i), b = ld(in_re, in_im, 2 * i + 1);
        st(add2(a, mul2(r, sub2(b, a))), out_re, out_im, i);
    }
}

/* TP/native_fp2/native_fp2_core.hpp:66-80 */
void or_expand(const u64* in_re, const u64* in_im, size_t n_src, size_t nc,
               u64* out_re, u64* out_im)
{
    for (size_t i = 0; i < n_src; ++i)
        for (size_t j = 0; j < nc; ++j) { out_re[i * nc + j] = in_re[i]; out_im[i * nc + j] = in_im[i]; }
}

/* A6 (non-session): TP/native_fp2/native_fp2_core.hpp:82-137.  values are
 * [wire][copy]; fact0/fact1 are the tables evaluated at -1 and 2 on
 * [wire][copy/2]; beta, beta_f0, beta_f1 likewise. out8 = F(0),F(1),F(-1),F(2). */
void or_sumcheck_early(int n_gates, const int* gt, const int* in0, const int* in1,
                       const u64* z1r, const u64* z1i, int n_copies, int n_wires,
                       const u64* vr, const u64* vi, const u64* f0r, const u64* f0i,
                       const u64* f1r, const u64* f1i, const u64* br, const u64* bi,
                       const u64* bf0r, const u64* bf0i, const u64* bf1r, const u64* bf1i,
                       u64 out8[8])
{
    (void)n_wires;
    fp2 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const size_t nc = (size_t)n_copies, hn = nc / 2;
    for (size_t c = 0; c < nc; c += 2) {
        fp2 v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        const size_t h = c / 2;
        for (int g = 0; g < n_gates; ++g) {
            const fp2 z = mk(z1r[g], z1i[g]);
            const size_t a = (size_t)in0[g], b = (size_t)in1[g];
            v[0] = add2(v[0], mul2(gatefn(gt[g], ld(vr, vi, a * nc + c), ld(vr, vi, b * nc + c)), z));
            v[1] = add2(v[1], mul2(gatefn(gt[g], ld(vr, vi, a * nc + c + 1), ld(vr, vi, b * nc + c + 1)), z));
            v[2] = add2(v[2], mul2(gatefn(gt[g], ld(f0r, f0i, a * hn + h), ld(f0r, f0i, b * hn + h)), z));
            v[3] = add2(v[3], mul2(gatefn(gt[g], ld(f1r, f1i, a * hn + h), ld(f1r, f1i, b * hn + h)), z));
        }
        acc[0] = add2(acc[0], mul2(v[0], ld(br, bi, c)));
        acc[1] = add2(acc[1], mul2(v[1], ld(br, bi, c + 1)));
        acc[2] = add2(acc[2], mul2(v[2], ld(bf0r, bf0i, h)));
        acc[3] = add2(acc[3], mul2(v[3], ld(bf1r, bf1i, h)));
    }
    for (int t = 0; t < 4; ++t) { out8[2 * t] = acc[t].re; out8[2 * t + 1] = acc[t].im; }
}

/* A5: TP/native_fp2/native_fp2_core.hpp:139-159, LSB-first:
 * out[i] = init * prod_j (bit_j(i) ? z_j : 1 - z_j). */
void or_compute_beta(const u64* zr, const u64* zi, int n_z, u64 init_r, u64 init_i,
                     u64* out_r, u64* out_i)
{
    if (n_z <= 0) return;
    out_r[0] = init_r; out_i[0] = init_i;
    size_t cur = 1;
    for (int j = 0; j < n_z; ++j) {
        const fp2 z = mk(zr[j], zi[j]), omz = sub2(mk(1, 0), z);
        for (size_t k = 0; k < cur; ++k) {
            const fp2 b = ld(out_r, out_i, k);
            st(mul2(b, z), out_r, out_i, k + cur);
            st(mul2(b, omz), out_r, out_i, k);
        }
        cur <<= 1;
    }
}

/* A12: TP/native_fp2/native_fp2_core.hpp:300-364.  The value is the MLE of
 * the base-field vector at z (LSB-first); the square-root split of the
 * reference changes the evaluation order only, so this computes the plain
 * fold chain.  inputs has 2^n_z entries. */
void or_mle_eval_base(const u64* inputs, const u64* zr, const u64* zi, int n_z,
                      u64* out_r, u64* out_i)
{
    /* n_z == 0: the reference stores inputs[0] unreduced (core :304-307) */
    if (n_z <= 0) { *out_r = inputs ? inputs[0] : 0; *out_i = 0; return; }
    size_t n = (size_t)1 << n_z;
    u64* tr = (u64*)malloc(n / 2 * sizeof(u64));
    u64* ti = (u64*)malloc(n / 2 * sizeof(u64));
    /* first fold from the base-field vector (reduce64 of each input) */
    const fp2 z0 = mk(zr[0], zi[0]);
    for (size_t k = 0; k < n / 2; ++k) {
        u64 a0 = inputs[2 * k], a1 = inputs[2 * k + 1];
        a0 = fcan((a0 & P61) + (a0 >> 61));
        a1 = fcan((a1 & P61) + (a1 >> 61));
        fp2 a = mk(a0, 0), b = mk(a1, 0);
        st(add2(a, mul2(z0, sub2(b, a))), tr, ti, k);
    }
    for (int j = 1; j < n_z; ++j) {
        size_t m = n >> (j + 1);
        const fp2 z = mk(zr[j], zi[j]);
        for (size_t k = 0; k < m; ++k) {
            fp2 a = ld(tr, ti, 2 * k), b = ld(tr, ti, 2 * k + 1);
            st(add2(a, mul2(z, sub2(b, a))), tr, ti, k);
        }
    }
    *out_r = tr[0]; *out_i = ti[0];
    free(tr); free(ti);
}

/* A6 (session claim): TP/native_fp2/native_fp2_core.hpp:200-248.  The points
 * -1 and 2 come from the live pair: V(-1) = 2a - b, V(2) = 2b - a. */
void or_early_claim(int n_gates, const int* gt, const int* in0, const int* in1,
                    const u64* z1r, const u64* z1i, int n_copies,
                    const u64* vr, const u64* vi, const u64* br, const u64* bi,
                    u64 out8[8])
{
    fp2 acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const size_t nc = (size_t)n_copies;
    for (size_t c = 0; c < nc; c += 2) {
        fp2 v[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        for (int g = 0; g < n_gates; ++g) {
            const fp2 z = mk(z1r[g], z1i[g]);
            const size_t a = (size_t)in0[g], b = (size_t)in1[g];
            const fp2 a0 = ld(vr, vi, a * nc + c), a1 = ld(vr, vi, a * nc + c + 1);
            const fp2 b0 = ld(vr, vi, b * nc + c), b1 = ld(vr, vi, b * nc + c + 1);
            v[0] = add2(v[0], mul2(gatefn(gt[g], a0, b0), z));
            v[1] = add2(v[1], mul2(gatefn(gt[g], a1, b1), z));
            v[2] = add2(v[2], mul2(gatefn(gt[g], sub2(add2(a0, a0), a1), sub2(add2(b0, b0), b1)), z));
            v[3] = add2(v[3], mul2(gatefn(gt[g], sub2(add2(a1, a1), a0), sub2(add2(b1, b1), b0)), z));
        }
        const fp2 be0 = ld(br, bi, c), be1 = ld(br, bi, c + 1);
        acc[0] = add2(acc[0], mul2(v[0], be0));
        acc[1] = add2(acc[1], mul2(v[1], be1));
        acc[2] = add2(acc[2], mul2(v[2], sub2(add2(be0, be0), be1)));
        acc[3] = add2(acc[3], mul2(v[3], sub2(add2(be1, be1), be0)));
    }
    for (int t = 0; t < 4; ++t) { out8[2 * t] = acc[t].re; out8[2 * t + 1] = acc[t].im; }
}

/* A7 (session fold): TP/native_fp2/native_fp2_core.hpp:250-279, every wire row
 * [wire][n_copies] -> [wire][n_copies/2] and beta likewise. */
void or_early_fold(int n_wires, int n_copies, const u64* vr, const u64* vi,
                   const u64* br, const u64* bi, u64 rr, u64 ri,
                   u64* ovr, u64* ovi, u64* obr, u64* obi)
{
    const size_t nin = (size_t)n_copies, nout = nin / 2;
    for (int w = 0; w < n_wires; ++w)
        or_fold(vr + (size_t)w * nin, vi + (size_t)w * nin, nout, rr, ri,
                ovr + (size_t)w * nout, ovi + (size_t)w * nout);
    or_fold(br, bi, nout, rr, ri, obr, obi);
}

/* A17: degree-2 sumcheck round, OZ/src/VPD/fft_circuit_GKR.cpp:171-209 with the
 * addV term zero.  f and g are interleaved (re, im) tables of n Fp2 values (the
 * current hypercube after earlier folds).  Writes (a, b, c) of
 * s(X) = sum_i (f0 + X df)(g0 + X dg) = aX^2 + bX + c as out6 = a.re a.im b.re b.im c.re c.im. */
void or_sc2_round(const u64* f, const u64* g, size_t n, int nthreads, u64 out6[6])
{
    (void)nthreads;
    fp2 A = {0, 0}, B = {0, 0}, C = {0, 0};
    #pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        fp2 a = {0, 0}, b = {0, 0}, c = {0, 0};
        #pragma omp for schedule(static)
        for (long long i = 0; i < (long long)(n / 2); ++i) {
            fp2 f0 = mk(f[4 * i], f[4 * i + 1]), f1 = mk(f[4 * i + 2], f[4 * i + 3]);
            fp2 g0 = mk(g[4 * i], g[4 * i + 1]), g1 = mk(g[4 * i + 2], g[4 * i + 3]);
            fp2 df = sub2(f1, f0), dg = sub2(g1, g0);
            a = add2(a, mul2(df, dg));
            b = add2(b, add2(mul2(df, g0), mul2(f0, dg)));
            c = add2(c, mul2(f0, g0));
        }
        #pragma omp critical
        { A = add2(A, a); B = add2(B, b); C = add2(C, c); }
    }
    out6[0] = A.re; out6[1] = A.im; out6[2] = B.re; out6[3] = B.im; out6[4] = C.re; out6[5] = C.im;
}

/* A17 fold step: f[i] <- f[2i] + r (f[2i+1] - f[2i]) in place, n -> n/2. */
void or_sc2_fold(u64* f, u64* g, size_t n, u64 rr, u64 ri, int nthreads)
{
    const fp2 r = mk(rr, ri);
    u64* t[2] = {f, g};
    for (int s = 0; s < 2; ++s) {
        u64* x = t[s];
        u64* tmp = (u64*)malloc(n * sizeof(u64));
        #pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
        for (long long i = 0; i < (long long)(n / 2); ++i) {
            fp2 a = mk(x[4 * i], x[4 * i + 1]), b = mk(x[4 * i + 2], x[4 * i + 3]);
            fp2 v = add2(a, mul2(r, sub2(b, a)));
            tmp[2 * i] = v.re; tmp[2 * i + 1] = v.im;
        }
        memcpy(x, tmp, n * sizeof(u64));
        free(tmp);
    }
}

/* A19: OZ/src/linear_gkr/prime_field.cpp:60-73.  g = 2^31 + 1033321771269002680 i
 * has order 2^62; omega_k = g^(2^(62-k)). */
void or_root_of_unity(int log_order, u64 out[2])
{
    fp2 g = mk(2147483648ULL, 1033321771269002680ULL);
    for (int i = 0; i < 62 - log_order; ++i) g = mul2(g, g);
    out[0] = g.re; out[1] = g.im;
}

/* A20: OZ/src/infrastructure/RS_polynomial.cpp:46-130.  Natural-order
 * evaluation result[i] = sum_{j < coef_len} c_j w^(ij) for i < order, with
 * w = omega_{log2 order}.  c and result are interleaved (re, im).  Computed
 * here as coef_len-sized coset DFTs by an iterative radix-2 transform, an
 * algorithm different from the reference's replicate-and-Stockham network. */
static void bitrev_dft(fp2* x, size_t n, fp2 w)
{
    size_t lg = 0; while (((size_t)1 << lg) < n) ++lg;
    for (size_t i = 0, j = 0; i < n; ++i) {
        if (i < j) { fp2 t = x[i]; x[i] = x[j]; x[j] = t; }
        size_t bit = n >> 1;
        while (bit && (j & bit)) { j ^= bit; bit >>= 1; }
        j |= bit;
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        fp2 wl = w;
        for (size_t s = len; s < n; s <<= 1) wl = mul2(wl, wl);
        for (size_t st0 = 0; st0 < n; st0 += len) {
            fp2 t = {1, 0};
            for (size_t k = 0; k < len / 2; ++k) {
                fp2 u = x[st0 + k], v = mul2(t, x[st0 + k + len / 2]);
                x[st0 + k] = add2(u, v);
                x[st0 + k + len / 2] = sub2(u, v);
                t = mul2(t, wl);
            }
        }
    }
}
void or_fft(const u64* coef, int log_coef, int log_order, u64* result)
{
    size_t m = (size_t)1 << log_coef, n = (size_t)1 << log_order, q = n / m;
    u64 wv[2]; or_root_of_unity(log_order, wv);
    fp2 w = mk(wv[0], wv[1]);
    fp2 wm = w; for (size_t k = 1; k < q; k <<= 1) wm = mul2(wm, wm); /* order m */
    fp2* buf = (fp2*)malloc(m * sizeof(fp2));
    fp2 ws = {1, 0}; /* w^s */
    for (size_t s = 0; s < q; ++s) {
        fp2 tw = {1, 0};
        for (size_t j = 0; j < m; ++j) {       /* twist: c_j w^(s j) */
            buf[j] = mul2(mk(coef[2 * j], coef[2 * j + 1]), tw);
            tw = mul2(tw, ws);
        }
        bitrev_dft(buf, m, wm);
        for (size_t t = 0; t < m; ++t) {       /* result[q t + s] */
            result[2 * (q * t + s)] = buf[t].re;
            result[2 * (q * t + s) + 1] = buf[t].im;
        }
        ws = mul2(ws, w);
    }
    free(buf);
}

/* Horner evaluation of sum_j c_j x^j (c interleaved, n coefficients): used by
 * the tests to check single NTT outputs at full size. */
void or_poly_eval(const u64* coef, size_t n, u64 xr, u64 xi, u64 out[2])
{
    fp2 x = mk(xr, xi), acc = {0, 0};
    for (size_t j = n; j-- > 0;) acc = add2(mul2(acc, x), mk(coef[2 * j], coef[2 * j + 1]));
    out[0] = acc.re; out[1] = acc.im;
}

/* x^e in Fp2 */
void or_fp2_pow(u64 xr, u64 xi, u64 e, u64 out[2])
{
    fp2 r = {1, 0}, b = mk(xr, xi);
    while (e) { if (e & 1) r = mul2(r, b); b = mul2(b, b); e >>= 1; }
    out[0] = r.re; out[1] = r.im;
}
