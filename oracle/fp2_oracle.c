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
 *   or_glibc_rand       glibc rand() of this container (the library routine)
 *   or_fft_gkr          A16/A18 oracle/_ref (fft_circuit_GKR.cpp compiled from
 *                           source, driven through its own layer functions):
 *                           circuit layers, the claim after every layer, the
 *                           final alpha/beta, pass flag and proof size
 *   or_sha256           FIPS 180-4 known answers and Python's hashlib
 *   or_orion_encode     the expander code: oracle/_ref (encode of
 *                           linear_code_encode.h compiled from source) on
 *                           random columns, plus linearity and the systematic
 *                           prefix (codeword starts with the message)
 *   or_orion_commit     the signer's commit roots: oracle/_ref (the
 *                           reference's layout, mask, encode, row hash and
 *                           create_tree compiled from source)
 */
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <math.h>
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

/* ======================================================================
 * A16 / A18: the fft_gkr circuit and its layer protocols
 * (OZ/src/VPD/fft_circuit_GKR.cpp:22-845), restated.
 *
 * The reference draws every challenge with glibc rand() (never seeded, so
 * seed 1) through prime_field::random() (OZ/src/linear_gkr/prime_field.cpp:35-40:
 * real = rand() % p, then imag = rand() % p).  or_glibc_rand restates glibc's
 * TYPE_3 generator from its published definition: r[0] = seed,
 * r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31, r[i] = r[i-31] for i < 34,
 * then r[k] = r[k-31] + r[k-3] (mod 2^32); the n-th rand() is r[n + 344] >> 1.
 * ====================================================================== */

typedef struct { uint32_t r[34]; long k; } grand;

static void grand_init(grand* g, uint32_t seed)
{
    int32_t s[34];
    s[0] = (int32_t)(seed ? seed : 1);
    for (int i = 1; i < 31; ++i) {
        long long t = (16807LL * s[i - 1]) % 2147483647LL;
        if (t < 0) t += 2147483647LL;
        s[i] = (int32_t)t;
    }
    for (int i = 31; i < 34; ++i) s[i] = s[i - 31];
    for (int i = 0; i < 34; ++i) g->r[i] = (uint32_t)s[i];
    g->k = 34;
    for (int i = 0; i < 310; ++i) {           /* discarded outputs */
        uint32_t v = g->r[(g->k - 31) % 34] + g->r[(g->k - 3) % 34];
        g->r[g->k % 34] = v;
        ++g->k;
    }
}

static uint32_t grand_next(grand* g)
{
    uint32_t v = g->r[(g->k - 31) % 34] + g->r[(g->k - 3) % 34];
    g->r[g->k % 34] = v;
    ++g->k;
    return v >> 1;
}

/* The first n outputs of rand() after srand(seed). */
void or_glibc_rand(uint32_t seed, int n, int32_t* out)
{
    grand g;
    grand_init(&g, seed);
    for (int i = 0; i < n; ++i) out[i] = (int32_t)grand_next(&g);
}

static fp2 grand_fp2(grand* g)
{
    u64 re = grand_next(g) % P61;
    u64 im = grand_next(g) % P61;
    return mk(re, im);
}

static fp2 one_minus(fp2 x) { return sub2(mk(1, 0), x); }
static int eq2_(fp2 a, fp2 b) { return a.re == b.re && a.im == b.im; }

static fp2 pow2_fp2(fp2 b, u64 e)
{
    fp2 r = {1, 0};
    while (e) { if (e & 1) r = mul2(r, b); b = mul2(b, b); e >>= 1; }
    return r;
}

/* a X^2 + b X + c */
static fp2 qeval(const fp2 q[3], fp2 x) { return add2(mul2(add2(mul2(q[0], x), q[1]), x), q[2]); }

/* eq(z, idx) = prod_j (bit_j(idx) ? z_j : 1 - z_j), LSB-first, times init */
static fp2 eq_at(const fp2* z, int n, size_t idx, fp2 init)
{
    fp2 b = init;
    for (int j = 0; j < n; ++j) b = mul2(b, ((idx >> j) & 1) ? z[j] : one_minus(z[j]));
    return b;
}

typedef struct {
    fp2* polys;      /* all round polynomials (a, b, c), in order */
    int n_polys, cap;
    int proof_size;
    int ok;
} fg_state;

/* One sumcheck phase (sumcheck_phase1_update / phase2_update, :171-249) over
 * value tables f, g (and h, the addV term, or NULL) of 2^nv entries, with the
 * challenges chal[0..nv).  Round k's polynomial sums over pairs of the table
 * folded with chal[0..k).  Checks s(0) + s(1) against *claim and advances it
 * like the reference's verifier loop; returns f folded at all challenges. */
static fp2 fg_phase(fg_state* S, fp2* f, fp2* g, fp2* h, int nv, const fp2* chal, fp2* claim)
{
    size_t n = (size_t)1 << nv;
    for (int k = 0; k < nv; ++k) {
        fp2 q[3] = {{0, 0}, {0, 0}, {0, 0}};
        for (size_t i = 0; i < n / 2; ++i) {
            fp2 f0 = f[2 * i], df = sub2(f[2 * i + 1], f0);
            fp2 g0 = g[2 * i], dg = sub2(g[2 * i + 1], g0);
            q[0] = add2(q[0], mul2(dg, df));
            q[1] = add2(q[1], add2(mul2(dg, f0), mul2(g0, df)));
            q[2] = add2(q[2], mul2(g0, f0));
            if (h) {
                q[1] = add2(q[1], sub2(h[2 * i + 1], h[2 * i]));
                q[2] = add2(q[2], h[2 * i]);
            }
        }
        S->proof_size += 48;
        if (S->n_polys < S->cap) {
            S->polys[3 * S->n_polys] = q[0];
            S->polys[3 * S->n_polys + 1] = q[1];
            S->polys[3 * S->n_polys + 2] = q[2];
        }
        S->n_polys++;
        fp2 s01 = add2(add2(add2(q[2], q[2]), q[0]), q[1]);
        if (s01.re != claim->re || s01.im != claim->im) S->ok = 0;
        *claim = qeval(q, chal[k]);
        for (size_t i = 0; i < n / 2; ++i) {
            f[i] = add2(f[2 * i], mul2(chal[k], sub2(f[2 * i + 1], f[2 * i])));
            g[i] = add2(g[2 * i], mul2(chal[k], sub2(g[2 * i + 1], g[2 * i])));
            if (h) h[i] = add2(h[2 * i], mul2(chal[k], sub2(h[2 * i + 1], h[2 * i])));
        }
        n /= 2;
    }
    return f[0];
}

/* Full run of fft_gkr for lg_size (main + build_circuit + engage_gkr,
 * :22-112, :786-845).  Returns 1 when every verifier check passes.
 *   polys (cap_polys x 3 Fp2, interleaved limbs), *n_polys: round polynomials
 *   claims (7 Fp2): alpha_beta_sum after V_output, addition, mult,
 *                   intermediate and ifft, then the final alpha and beta
 *   layers: the last IFFT layer (2^lg Fp2), the scaled layer (2^lg) and the
 *           64 sums (NULL to skip) */
int or_fft_gkr(int lg, u64* polys, int cap_polys, int* n_polys, u64* claims, u64* layers,
               int* proof_size)
{
    if (lg < 1 || lg > 24) return -1;
    grand G;
    grand_init(&G, 1);
    fg_state S;
    S.polys = (fp2*)polys;
    S.n_polys = 0;
    S.cap = polys ? cap_polys : 0;
    S.proof_size = 0;
    S.ok = 1;
    const size_t n = (size_t)1 << lg;
    fp2* r = (fp2*)malloc(lg * sizeof(fp2));
    for (int i = 0; i < lg; ++i) r[i] = grand_fp2(&G);

    /* build_circuit: beta expansion of r, IFFT layers, x inv_n, 64 evaluations, sums */
    fp2** L = (fp2**)malloc((lg + 1) * sizeof(fp2*));
    for (int s = 0; s <= lg; ++s) L[s] = (fp2*)malloc(n * sizeof(fp2));
    {
        fp2* cur = L[0];
        cur[0] = mk(1, 0);
        for (int i = 0; i < lg; ++i) {          /* in place from the top: out[2j], out[2j+1] from in[j] */
            for (size_t j = ((size_t)1 << i); j-- > 0;) {
                fp2 v = cur[j];
                cur[2 * j] = mul2(v, r[i]);
                cur[2 * j + 1] = mul2(v, one_minus(r[i]));
            }
        }
    }
    u64 rv[2];
    or_root_of_unity(lg, rv);
    u64 iv[2];
    or_fp2_inv(rv[0], rv[1], iv);
    fp2 rot[62];
    rot[0] = mk(iv[0], iv[1]);
    for (int i = 1; i < 62; ++i) rot[i] = mul2(rot[i - 1], rot[i - 1]);
    for (int s = 1; s <= lg; ++s) {
        const int dep = lg - s;
        const size_t blk = (size_t)1 << (lg - dep), half = blk / 2;
        const fp2* pre = L[s - 1];
        fp2* cur = L[s];
        fp2 x = {1, 0};
        for (size_t k = 0; k < half; ++k) {
            for (size_t j = 0; j < ((size_t)1 << dep); ++j) {
                fp2 lv = pre[(k << (dep + 1)) | j];
                fp2 rvv = mul2(x, pre[(k << (dep + 1)) | ((size_t)1 << dep) | j]);
                cur[(k << dep) | j] = add2(lv, rvv);
                cur[((k + half) << dep) | j] = sub2(lv, rvv);
            }
            x = mul2(x, rot[dep]);
        }
    }
    const u64 inv_n = fpow(n % P61, P61 - 2);
    fp2* scaled = (fp2*)malloc(n * sizeof(fp2));
    for (size_t t = 0; t < n; ++t) scaled[t] = mul2(L[lg][t], mk(inv_n, 0));
    fp2 ep[64];
    fp2* ev = (fp2*)malloc(64 * n * sizeof(fp2));
    for (int i = 0; i < 64; ++i) {
        ep[i] = grand_fp2(&G);
        fp2 x = {1, 0};
        for (size_t j = 0; j < n; ++j) {
            ev[j + (size_t)i * n] = mul2(scaled[j], x);
            x = mul2(x, ep[i]);
        }
    }
    fp2 sums[64];
    for (int i = 0; i < 64; ++i) {
        fp2 s = {0, 0};
        for (size_t j = 0; j < n; ++j) s = add2(s, ev[j + (size_t)i * n]);
        sums[i] = s;
    }
    if (layers) {
        memcpy(layers, L[lg], n * sizeof(fp2));
        memcpy(layers + 2 * n, scaled, n * sizeof(fp2));
        memcpy(layers + 4 * n, sums, 64 * sizeof(fp2));
    }

    /* engage_gkr */
    fp2 alpha = {1, 0}, beta = {0, 0};
    const int nr = lg + 10;
    fp2* r0 = (fp2*)malloc(nr * sizeof(fp2));
    fp2* r1 = (fp2*)malloc(nr * sizeof(fp2));
    fp2* ru = (fp2*)malloc(nr * sizeof(fp2));
    fp2* rvv = (fp2*)malloc(nr * sizeof(fp2));
    for (int i = 0; i < nr; ++i) r0[i] = grand_fp2(&G);
    for (int i = 0; i < nr; ++i) r1[i] = grand_fp2(&G);
    fp2 claim;
    {   /* V_output: fold the 64 sums with r0[0..6) */
        fp2 o[64];
        memcpy(o, sums, sizeof(o));
        int m = 64;
        for (int i = 0; i < 6; ++i, m /= 2)
            for (int j = 0; j < m / 2; ++j)
                o[j] = add2(mul2(o[2 * j], one_minus(r0[i])), mul2(o[2 * j + 1], r0[i]));
        claim = o[0];
    }
    fp2 cl[7];
    cl[0] = claim;

    /* addition_layer (:251-367): f = the 64 x 2^lg evaluations, g[j] = tmp(j >> lg) */
    {
        const int log_uv = lg + 6;
        const size_t tot = (size_t)64 * n;
        fp2* f = (fp2*)malloc(tot * sizeof(fp2));
        fp2* g = (fp2*)malloc(tot * sizeof(fp2));
        fp2 tmp[64];
        for (int i = 0; i < 64; ++i)
            tmp[i] = add2(eq_at(r0, 6, i, alpha), eq_at(r1, 6, i, beta));
        memcpy(f, ev, tot * sizeof(fp2));
        for (size_t j = 0; j < tot; ++j) g[j] = tmp[j >> lg];
        for (int i = 0; i < log_uv; ++i) ru[i] = grand_fp2(&G);
        for (int i = 0; i < log_uv; ++i) rvv[i] = grand_fp2(&G);
        fp2 vu = fg_phase(&S, f, g, NULL, log_uv, ru, &claim);
        fp2 sv = {0, 0};
        for (int i = 0; i < 64; ++i)
            sv = add2(sv, mul2(tmp[i], eq_at(ru + (log_uv - 6), 6, i, mk(1, 0))));
        if (!(eq2_(claim, mul2(sv, vu)))) S.ok = 0;
        memcpy(r0, ru, log_uv * sizeof(fp2));
        memcpy(r1, rvv, log_uv * sizeof(fp2));
        claim = mul2(alpha, vu);
        free(f);
        free(g);
    }
    cl[1] = claim;

    /* mult_layer (:369-495): f = scaled, g[i] = sum_j tmp(j 2^lg + i) x_j^i */
    {
        const int lgg = lg + 6;
        fp2* f = (fp2*)malloc(n * sizeof(fp2));
        fp2* g = (fp2*)malloc(n * sizeof(fp2));
        memcpy(f, scaled, n * sizeof(fp2));
        fp2 xe[64];
        for (int j = 0; j < 64; ++j) xe[j] = mk(1, 0);
        for (size_t i = 0; i < n; ++i) {
            fp2 acc = {0, 0};
            for (int j = 0; j < 64; ++j) {
                size_t gid = (size_t)j * n + i;
                fp2 t = add2(eq_at(r0, lgg, gid, alpha), eq_at(r1, lgg, gid, beta));
                acc = add2(acc, mul2(t, xe[j]));
                xe[j] = mul2(xe[j], ep[j]);
            }
            g[i] = acc;
        }
        for (int i = 0; i < lg; ++i) ru[i] = grand_fp2(&G);
        for (int i = 0; i < lg; ++i) rvv[i] = grand_fp2(&G);
        fp2 vu = fg_phase(&S, f, g, NULL, lg, ru, &claim);
        fp2 sm = {0, 0};
        for (int i = 0; i < 64; ++i) {
            fp2 tg0 = alpha, tg1 = beta;
            for (int j = 0; j < 6; ++j) {
                int b = (i >> j) & 1;
                tg0 = mul2(tg0, b ? r0[lgg - 6 + j] : one_minus(r0[lgg - 6 + j]));
                tg1 = mul2(tg1, b ? r1[lgg - 6 + j] : one_minus(r1[lgg - 6 + j]));
            }
            fp2 tu0 = {1, 0}, tu1 = {1, 0}, x = ep[i];
            for (int j = 0; j < lg; ++j) {
                fp2 a0 = add2(mul2(mul2(r0[j], ru[j]), x), mul2(one_minus(r0[j]), one_minus(ru[j])));
                fp2 a1 = add2(mul2(mul2(r1[j], ru[j]), x), mul2(one_minus(r1[j]), one_minus(ru[j])));
                tu0 = mul2(tu0, a0);
                tu1 = mul2(tu1, a1);
                x = mul2(x, x);
            }
            sm = add2(sm, add2(mul2(tg0, tu0), mul2(tg1, tu1)));
        }
        if (!eq2_(claim, mul2(sm, vu))) S.ok = 0;
        memcpy(r0, ru, lg * sizeof(fp2));
        memcpy(r1, rvv, lg * sizeof(fp2));
        claim = mul2(alpha, vu);
        free(f);
        free(g);
    }
    cl[2] = claim;

    /* intermediate_layer (:497-502) */
    claim = mul2(claim, mk(n % P61, 0));
    cl[3] = claim;

    /* ifft_gkr (:504-772): layer L[lg - dep] from L[lg - dep - 1] */
    {
        fp2* f = (fp2*)malloc(n * sizeof(fp2));
        fp2* g = (fp2*)malloc(n * sizeof(fp2));
        fp2* h = (fp2*)malloc(n * sizeof(fp2));
        fp2* tg = (fp2*)malloc(n * sizeof(fp2));
        fp2* bu = (fp2*)malloc(n * sizeof(fp2));
        for (int dep = 0; dep < lg && S.ok; ++dep) {
            const size_t blk = (size_t)1 << (lg - dep), half = blk / 2;
            const fp2* pre = L[lg - dep - 1];
            for (size_t t = 0; t < n; ++t)
                tg[t] = add2(eq_at(r0, lg, t, alpha), eq_at(r1, lg, t, beta));
            memcpy(f, pre, n * sizeof(fp2));
            memset(g, 0, n * sizeof(fp2));
            memset(h, 0, n * sizeof(fp2));
            fp2 x = {1, 0};
            for (size_t k = 0; k < half; ++k) {
                for (size_t j = 0; j < ((size_t)1 << dep); ++j) {
                    size_t u = (k << (dep + 1)) | j, v = u | ((size_t)1 << dep);
                    size_t g1 = (k << dep) | j, g2 = ((k + half) << dep) | j;
                    g[u] = add2(g[u], add2(tg[g1], tg[g2]));
                    h[u] = add2(h[u], mul2(mul2(sub2(tg[g1], tg[g2]), x), pre[v]));
                }
                x = mul2(x, rot[dep]);
            }
            for (int i = 0; i < lg; ++i) ru[i] = grand_fp2(&G);
            for (int i = 0; i < lg; ++i) rvv[i] = grand_fp2(&G);
            fp2 vu = fg_phase(&S, f, g, h, lg, ru, &claim);
            for (size_t t = 0; t < n; ++t) bu[t] = eq_at(ru, lg, t, mk(1, 0));
            memcpy(f, pre, n * sizeof(fp2));
            memset(g, 0, n * sizeof(fp2));
            memset(h, 0, n * sizeof(fp2));
            x = mk(1, 0);
            for (size_t k = 0; k < half; ++k) {
                for (size_t j = 0; j < ((size_t)1 << dep); ++j) {
                    size_t u = (k << (dep + 1)) | j, v = u | ((size_t)1 << dep);
                    size_t g1 = (k << dep) | j, g2 = ((k + half) << dep) | j;
                    fp2 a = mul2(tg[g1], bu[u]), b = mul2(tg[g2], bu[u]);
                    g[v] = add2(g[v], mul2(sub2(a, b), x));
                    h[v] = add2(h[v], mul2(add2(a, b), vu));
                }
                x = mul2(x, rot[dep]);
            }
            fp2 vv = fg_phase(&S, f, g, h, lg, rvv, &claim);
            /* the verifier's closed form (:712-754) */
            const int log_k = lg - dep - 1, log_j = dep;
            const fp2 one = {1, 0};
            fp2 c0 = mul2(mul2(one_minus(ru[log_j]), rvv[log_j]), alpha);
            fp2 c1 = mul2(mul2(one_minus(ru[log_j]), rvv[log_j]), beta);
            fp2 A0u = mul2(one_minus(r0[lg - 1]), c0), A1u = mul2(one_minus(r1[lg - 1]), c1);
            fp2 A0v = A0u, A1v = A1u;
            fp2 B0u = mul2(r0[lg - 1], c0), B1u = mul2(r1[lg - 1], c1);
            fp2 B0v = B0u, B1v = B1u;
            x = rot[dep];
            for (int i = 0; i < log_k; ++i) {
                fp2 a = ru[log_j + 1 + i], b = rvv[log_j + 1 + i];
                fp2 ab = mul2(a, b), nab = mul2(one_minus(a), one_minus(b));
                fp2 p0 = mul2(r0[log_j + i], ab), q0 = mul2(one_minus(r0[log_j + i]), nab);
                fp2 p1 = mul2(r1[log_j + i], ab), q1 = mul2(one_minus(r1[log_j + i]), nab);
                A0u = mul2(A0u, add2(p0, q0));
                A1u = mul2(A1u, add2(p1, q1));
                A0v = mul2(A0v, add2(mul2(p0, x), q0));
                A1v = mul2(A1v, add2(mul2(p1, x), q1));
                B0u = mul2(B0u, add2(p0, q0));
                B1u = mul2(B1u, add2(p1, q1));
                B0v = mul2(B0v, add2(q0, mul2(p0, x)));
                B1v = mul2(B1v, add2(q1, mul2(p1, x)));
                x = mul2(x, x);
            }
            for (int i = 0; i < log_j; ++i) {
                fp2 ab = mul2(ru[i], rvv[i]), nab = mul2(one_minus(ru[i]), one_minus(rvv[i]));
                fp2 s0 = add2(mul2(r0[i], ab), mul2(one_minus(r0[i]), nab));
                fp2 s1 = add2(mul2(r1[i], ab), mul2(one_minus(r1[i]), nab));
                A0u = mul2(A0u, s0); A0v = mul2(A0v, s0); B0u = mul2(B0u, s0); B0v = mul2(B0v, s0);
                A1u = mul2(A1u, s1); A1v = mul2(A1v, s1); B1u = mul2(B1u, s1); B1v = mul2(B1v, s1);
            }
            fp2 want = add2(mul2(add2(add2(A0u, A1u), add2(B0u, B1u)), vu),
                            mul2(sub2(add2(A0v, A1v), add2(B0v, B1v)), vv));
            if (!eq2_(claim, want)) S.ok = 0;
            (void)one;
            memcpy(r0, ru, lg * sizeof(fp2));
            memcpy(r1, rvv, lg * sizeof(fp2));
            alpha = grand_fp2(&G);
            beta = grand_fp2(&G);
            claim = add2(mul2(alpha, vu), mul2(beta, vv));
        }
        free(f); free(g); free(h); free(tg); free(bu);
    }
    cl[4] = claim;
    cl[5] = alpha;
    cl[6] = beta;
    /* extension_gkr (:773-782): proof size only */
    for (int i = 1; i <= lg; ++i) S.proof_size += i * 48;
    if (claims) memcpy(claims, cl, sizeof(cl));
    if (n_polys) *n_polys = S.n_polys;
    if (proof_size) *proof_size = S.proof_size;
    for (int s = 0; s <= lg; ++s) free(L[s]);
    free(L); free(r); free(scaled); free(ev); free(r0); free(r1); free(ru); free(rvv);
    return S.ok;
}

/* ---------------------------------------------------------------- SHA-256 --
 * FIPS 180-4, plain C (the oracle's own; the product uses SHA-NI on the host
 * and its own device code). */
static const uint32_t SHA_K[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
    0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
    0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
    0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
    0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
    0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
    0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
    0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};

static uint32_t ror32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void sha_block(uint32_t h[8], const unsigned char* p)
{
    uint32_t w[64], v[8];
    for (int t = 0; t < 16; ++t)
        w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 | (uint32_t)p[4 * t + 2] << 8 |
               p[4 * t + 3];
    for (int t = 16; t < 64; ++t)
        w[t] = (ror32(w[t - 2], 17) ^ ror32(w[t - 2], 19) ^ (w[t - 2] >> 10)) + w[t - 7] +
               (ror32(w[t - 15], 7) ^ ror32(w[t - 15], 18) ^ (w[t - 15] >> 3)) + w[t - 16];
    memcpy(v, h, sizeof(v));
    for (int t = 0; t < 64; ++t) {
        const uint32_t t1 = v[7] + (ror32(v[4], 6) ^ ror32(v[4], 11) ^ ror32(v[4], 25)) +
                            ((v[4] & v[5]) ^ (~v[4] & v[6])) + SHA_K[t] + w[t];
        const uint32_t t2 = (ror32(v[0], 2) ^ ror32(v[0], 13) ^ ror32(v[0], 22)) +
                            ((v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]));
        memmove(v + 1, v, 7 * sizeof(uint32_t));
        v[4] += t1;
        v[0] = t1 + t2;
    }
    for (int i = 0; i < 8; ++i) h[i] += v[i];
}

void or_sha256(const unsigned char* msg, size_t n, unsigned char out[32])
{
    uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                     0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    size_t i = 0;
    for (; i + 64 <= n; i += 64) sha_block(h, msg + i);
    unsigned char tail[128];
    size_t m = n - i;
    memcpy(tail, msg + i, m);
    tail[m++] = 0x80;
    const size_t len = m + 8 <= 64 ? 64 : 128;
    memset(tail + m, 0, len - m);
    for (int k = 0; k < 8; ++k) tail[len - 1 - k] = (unsigned char)((u64)n * 8 >> (8 * k));
    for (size_t b = 0; b < len; b += 64) sha_block(h, tail + b);
    for (int k = 0; k < 8; ++k) {
        out[4 * k] = (unsigned char)(h[k] >> 24);
        out[4 * k + 1] = (unsigned char)(h[k] >> 16);
        out[4 * k + 2] = (unsigned char)(h[k] >> 8);
        out[4 * k + 3] = (unsigned char)h[k];
    }
}

/* ----------------------------------------------------------- Orion commit --
 * The signer's commit of a frame's inputs
 * (OZ/examples/multivariate_PC_test_commit.cpp:265-331 -> commit_pair,
 * OZ/src/VPD/linearPC.cpp:253-459):
 *   - layout copy-major-384-const128-pad512 (OZ/include/truepix/truepix_gkr.h:107-210):
 *     n = next_pow2(copies) * 512, samples at copy*512 + w, 128 at copy*512 + 384;
 *   - mask (truepix_gkr.h:331-386): real part splitmix64(seed) % p at every
 *     sample position in order, zero elsewhere; masked = coefficients + mask;
 *   - 128 columns of n/128 (column c = flat indices c*col .. c*col + col - 1),
 *     each encoded by the expander code (OZ/include/linear_code/expanders.h:19-60,
 *     linear_code_encode.h:35-110): graphs drawn from glibc rand() in
 *     expander_init's order (C at each depth on the way down, D on the way
 *     up), weights (rand() % p, rand() % p) for (real, imag);
 *     codeword(x) = x | codeword(C x) | D codeword(C x), inputs of length <= 13
 *     kept as they are;
 *   - per row r < 2 col: 64 chained SHA-256(state | col[2j][r] | col[2j+1][r])
 *     from a zero state, an element as (imag, real) little-endian
 *     (OZ/src/infrastructure/merkle_tree.cpp:15-27), zero past the codeword;
 *   - a heap-ordered binary Merkle tree over the rows, node = SHA-256(left |
 *     right) (merkle_tree.cpp:29-79), the root printed as root_hex does
 *     (truepix_gkr.h:389-402).
 */
typedef struct { long long L, R; int deg; long long* to; fp2* w; } xgraph;
static xgraph XG_C[64], XG_D[64];
static const double X_ALPHA = 0.238, X_RATE = 1.72;
static const int X_CN = 10, X_DN = 20, X_SMALL = 13, X_COLS = 128;

static void xg_make(xgraph* g, grand* rg, long long L, long long R, int deg)
{
    free(g->to);
    free(g->w);
    g->L = L;
    g->R = R;
    g->deg = deg;
    g->to = (long long*)malloc(sizeof(long long) * (size_t)(L * deg + 1));
    g->w = (fp2*)malloc(sizeof(fp2) * (size_t)(L * deg + 1));
    for (long long i = 0; i < L * deg; ++i) {
        g->to[i] = (long long)(grand_next(rg) % (uint32_t)R);
        const u64 re = grand_next(rg) % P61;
        g->w[i] = mk(re, grand_next(rg) % P61);
    }
}

static long long xg_init(grand* rg, long long n, int dep)
{
    if (n <= X_SMALL) return n;
    const long long R = (long long)(X_ALPHA * n);
    xg_make(&XG_C[dep], rg, n, R, X_CN);
    const long long L = xg_init(rg, R, dep + 1);
    const long long R2 = (long long)(n * (X_RATE - 1) - L);
    xg_make(&XG_D[dep], rg, L, R2, X_DN);
    return n + L + R2;
}

static long long x_len(long long n, int dep)
{
    if (n <= X_SMALL) return n;
    return n + x_len(XG_C[dep].R, dep + 1) + XG_D[dep].R;
}

static long long x_encode(const fp2* src, long long n, int dep, fp2* dst)
{
    if (n <= X_SMALL) {
        memcpy(dst, src, sizeof(fp2) * (size_t)n);
        return n;
    }
    const xgraph* C = &XG_C[dep];
    fp2* y = (fp2*)calloc((size_t)C->R + 1, sizeof(fp2));
    for (long long i = 0; i < n; ++i)
        for (int d = 0; d < C->deg; ++d) {
            const long long k = i * C->deg + d;
            y[C->to[k]] = add2(y[C->to[k]], mul2(C->w[k], src[i]));
        }
    const long long L = x_encode(y, C->R, dep + 1, dst + n);
    free(y);
    memcpy(dst, src, sizeof(fp2) * (size_t)n);
    const xgraph* D = &XG_D[dep];
    fp2* par = dst + n + L;
    memset(par, 0, sizeof(fp2) * (size_t)D->R);
    for (long long i = 0; i < L; ++i)
        for (int d = 0; d < D->deg; ++d) {
            const long long k = i * D->deg + d;
            par[D->to[k]] = add2(par[D->to[k]], mul2(D->w[k], dst[n + i]));
        }
    return n + L + D->R;
}

/* codeword of n Fp2 (re, im pairs) with graphs for n from glibc rand(1) */
long long or_orion_encode(long long n, const u64* src, u64* dst)
{
    grand rg;
    grand_init(&rg, 1);
    xg_init(&rg, n, 0);
    const long long len = x_len(n, 0);
    fp2* s = (fp2*)malloc(sizeof(fp2) * (size_t)n);
    fp2* d = (fp2*)malloc(sizeof(fp2) * (size_t)len);
    for (long long i = 0; i < n; ++i) s[i] = mk(src[2 * i], src[2 * i + 1]);
    x_encode(s, n, 0, d);
    for (long long i = 0; i < len; ++i) {
        dst[2 * i] = d[i].re;
        dst[2 * i + 1] = d[i].im;
    }
    free(s);
    free(d);
    return len;
}

static u64 smix(u64 z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void put_le64(unsigned char* p, u64 v)
{
    for (int k = 0; k < 8; ++k) p[k] = (unsigned char)(v >> (8 * k));
}

static void root_hex(const unsigned char d[32], char out[65])
{
    static const char hx[] = "0123456789abcdef";
    const int order[4] = {8, 0, 24, 16};   /* root_hex prints h0[1], h0[0], h1[1], h1[0] */
    for (int q = 0; q < 4; ++q) {
        u64 v = 0;
        for (int k = 7; k >= 0; --k) v = (v << 8) | d[order[q] + k];
        for (int k = 0; k < 16; ++k) out[16 * q + k] = hx[(v >> (60 - 4 * k)) & 15];
    }
    out[64] = 0;
}

/* The packing of a copy (truepix_gkr.h:107-150, from TRUEPIX_COPY_SIZE,
 * TRUEPIX_CONST_IDX, TRUEPIX_CONST_VAL in the reference's programs): copy_size
 * samples, the constant const_val at wire const_idx (none when -1), zeros up
 * to 512; the mask covers the samples only.  Default: the gray circuit's
 * 384, 384, 128. */
static int X_CS = 384, X_CI = 384, X_CV = 128;
int or_orion_set_layout(int copy_size, int const_idx, int const_val)
{
    if (copy_size < 1 || copy_size > 512 || const_idx >= 512 || (const_idx >= 0 && const_idx < copy_size))
        return -1;
    X_CS = copy_size;
    X_CI = const_idx;
    X_CV = const_val;
    return 0;
}

/* inputs [copies][copy_size]; 0 / -1 */
int or_orion_commit(int copies, const unsigned char* inputs, u64 seed, char masked_root[65],
                    char mask_root[65], long long* cw_len_out)
{
    if (copies < 1) return -1;
    long long pc = 1;
    while (pc < copies) pc <<= 1;
    const long long n = pc * 512, col = n / X_COLS, rows = 2 * col;
    grand rg;
    grand_init(&rg, 1);
    xg_init(&rg, col, 0);
    const long long cw = x_len(col, 0);
    fp2* poly[2];
    poly[0] = (fp2*)calloc((size_t)n, sizeof(fp2));   /* masked */
    poly[1] = (fp2*)calloc((size_t)n, sizeof(fp2));   /* mask */
    u64 state = seed;
    for (long long c = 0; c < copies; ++c)
        for (int w = 0; w < X_CS; ++w) {
            state += 0x9E3779B97F4A7C15ULL;
            const u64 m = smix(state) % P61;
            poly[1][c * 512 + w] = mk(m, 0);
            poly[0][c * 512 + w] = mk(fadd(inputs[c * X_CS + w], m), 0);
        }
    if (X_CI >= 0)
        for (long long c = 0; c < copies; ++c) poly[0][c * 512 + X_CI] = mk((u64)X_CV, 0);
    fp2* cwb = (fp2*)malloc(sizeof(fp2) * (size_t)X_COLS * cw);
    unsigned char* tree = (unsigned char*)malloc((size_t)(2 * rows) * 32);
    char* roots[2] = {masked_root, mask_root};
    for (int k = 0; k < 2; ++k) {
        for (int c = 0; c < X_COLS; ++c) x_encode(poly[k] + c * col, col, 0, cwb + (long long)c * cw);
        for (long long r = 0; r < rows; ++r) {
            unsigned char msg[64];
            memset(msg, 0, 32);
            for (int j = 0; j < X_COLS / 2; ++j) {
                for (int h = 0; h < 2; ++h) {
                    const fp2 e = r < cw ? cwb[(long long)(2 * j + h) * cw + r] : mk(0, 0);
                    put_le64(msg + 32 + 16 * h, e.im);
                    put_le64(msg + 40 + 16 * h, e.re);
                }
                or_sha256(msg, 64, msg);
            }
            memcpy(tree + (rows + r) * 32, msg, 32);
        }
        for (long long v = rows - 1; v >= 1; --v) or_sha256(tree + 2 * v * 32, 64, tree + v * 32);
        root_hex(tree + 32, roots[k]);
    }
    free(cwb);
    free(tree);
    free(poly[0]);
    free(poly[1]);
    if (cw_len_out) *cw_len_out = cw;
    return 0;
}

/* The opening of one committed polynomial (linearPC.cpp:700-966
 * tensor_product_protocol with r0, r1 from the point r as open_and_verify
 * forms them, :996-1034; the masked polynomial when use_group1, else the
 * mask), serialized as OpeningProof version 2 (orion_proof.h:1-22,
 * orion_proof.cpp:69-100) into a malloc'd buffer.  0 / -1. */
static void ob_put(unsigned char** o, size_t* n, size_t* cap, const void* p, size_t k)
{
    if (*n + k > *cap) {
        *cap = (*n + k) * 2;
        *o = (unsigned char*)realloc(*o, *cap);
    }
    memcpy(*o + *n, p, k);
    *n += k;
}
static void ob_fe(unsigned char** o, size_t* n, size_t* cap, fp2 x)
{
    unsigned char b[16];
    put_le64(b, x.im);
    put_le64(b + 8, x.re);
    ob_put(o, n, cap, b, 16);
}
static void ob_siblings(unsigned char** o, size_t* n, size_t* cap, const unsigned char* tree,
                        long long leaves, const long long* q, int nq)
{
    char* seen = (char*)calloc((size_t)(2 * leaves), 1);
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nq * 64 + 1));
    u64 cnt = 0;
    for (int i = 0; i < nq; ++i)
        for (long long v = q[i] + leaves; v != 1; v /= 2)
            if (!seen[v ^ 1]) {
                seen[v ^ 1] = 1;
                idx[cnt++] = (uint32_t)(v ^ 1);
            }
    ob_put(o, n, cap, &cnt, 8);
    for (u64 k = 0; k < cnt; ++k) {
        ob_put(o, n, cap, &idx[k], 4);
        ob_put(o, n, cap, tree + (size_t)idx[k] * 32, 32);
    }
    free(seen);
    free(idx);
}

int or_orion_open(int copies, const unsigned char* inputs, u64 seed, int use_group1,
                  const u64* r, int n_r, const long long* col_q, int nq, unsigned char** proof,
                  size_t* proof_len, u64 answer[2])
{
    if (copies < 1) return -1;
    long long pc = 1;
    while (pc < copies) pc <<= 1;
    const long long n = pc * 512, col = n / X_COLS, rows = 2 * col;
    int lg = 0;
    while ((1LL << lg) < n) ++lg;
    if (n_r != lg) return -1;
    grand rg;
    grand_init(&rg, 1);
    xg_init(&rg, col, 0);
    const long long cw = x_len(col, 0);
    fp2* poly = (fp2*)calloc((size_t)n, sizeof(fp2));
    u64 state = seed;
    for (long long c = 0; c < copies; ++c)
        for (int w = 0; w < X_CS; ++w) {
            state += 0x9E3779B97F4A7C15ULL;
            const u64 m = smix(state) % P61;
            poly[c * 512 + w] = use_group1 ? mk(fadd(inputs[c * X_CS + w], m), 0) : mk(m, 0);
        }
    if (use_group1 && X_CI >= 0)
        for (long long c = 0; c < copies; ++c) poly[c * 512 + X_CI] = mk((u64)X_CV, 0);
    fp2* cwb = (fp2*)malloc(sizeof(fp2) * (size_t)X_COLS * cw);
    for (int c = 0; c < X_COLS; ++c) x_encode(poly + c * col, col, 0, cwb + (long long)c * cw);
    unsigned char* tcom = (unsigned char*)malloc((size_t)(2 * rows) * 32);
    unsigned char* tcmb = (unsigned char*)malloc((size_t)(2 * rows) * 32);
    for (long long rr = 0; rr < rows; ++rr) {
        unsigned char msg[64];
        memset(msg, 0, 32);
        for (int j = 0; j < X_COLS / 2; ++j) {
            for (int h = 0; h < 2; ++h) {
                const fp2 e = rr < cw ? cwb[(long long)(2 * j + h) * cw + rr] : mk(0, 0);
                put_le64(msg + 32 + 16 * h, e.im);
                put_le64(msg + 40 + 16 * h, e.re);
            }
            or_sha256(msg, 64, msg);
        }
        memcpy(tcom + (rows + rr) * 32, msg, 32);
    }
    for (long long v = rows - 1; v >= 1; --v) or_sha256(tcom + 2 * v * 32, 64, tcom + v * 32);
    /* r0 over r[0..7), r1 over the rest, the first challenge on the top bit */
    const int lc = 7;
    fp2* r0 = (fp2*)malloc(sizeof(fp2) * X_COLS);
    fp2* r1 = (fp2*)malloc(sizeof(fp2) * (size_t)col);
    for (long long i = 0; i < X_COLS; ++i) {
        fp2 v = mk(1, 0);
        for (int d = 0; d < lc; ++d) {
            const fp2 x = mk(r[2 * d], r[2 * d + 1]);
            v = mul2(v, ((i >> (lc - 1 - d)) & 1) ? x : sub2(mk(1, 0), x));
        }
        r0[i] = v;
    }
    for (long long t = 0; t < col; ++t) {
        fp2 v = mk(1, 0);
        for (int d = 0; d < lg - lc; ++d) {
            const fp2 x = mk(r[2 * (lc + d)], r[2 * (lc + d) + 1]);
            v = mul2(v, ((t >> (lg - lc - 1 - d)) & 1) ? x : sub2(mk(1, 0), x));
        }
        r1[t] = v;
    }
    fp2* cc = (fp2*)calloc((size_t)cw, sizeof(fp2));
    fp2* cm = (fp2*)calloc((size_t)col, sizeof(fp2));
    for (int j = 0; j < X_COLS; ++j) {
        for (long long t = 0; t < cw; ++t) cc[t] = add2(cc[t], mul2(r0[j], cwb[(long long)j * cw + t]));
        for (long long t = 0; t < col; ++t) cm[t] = add2(cm[t], mul2(r0[j], poly[j * col + t]));
    }
    fp2 ans = mk(0, 0);
    for (long long t = 0; t < col; ++t) ans = add2(ans, mul2(r1[t], cm[t]));
    for (long long rr = 0; rr < rows; ++rr) {
        unsigned char msg[64];
        memset(msg, 0, 64);
        const fp2 e = rr < cw ? cc[rr] : mk(0, 0);
        put_le64(msg, e.im);
        put_le64(msg + 8, e.re);
        or_sha256(msg, 64, tcmb + (rows + rr) * 32);
    }
    for (long long v = rows - 1; v >= 1; --v) or_sha256(tcmb + 2 * v * 32, 64, tcmb + v * 32);
    unsigned char* o = NULL;
    size_t on = 0, cap = 0;
    const uint32_t version = 2, group = use_group1 ? 1 : 0, qn = (uint32_t)nq;
    const u64 hdr[4] = {(u64)n, (u64)X_COLS, (u64)col, (u64)cw};
    ob_put(&o, &on, &cap, "ORIONPV1", 8);
    ob_put(&o, &on, &cap, &version, 4);
    ob_put(&o, &on, &cap, &group, 4);
    ob_put(&o, &on, &cap, hdr, 32);
    ob_put(&o, &on, &cap, &qn, 4);
    ob_fe(&o, &on, &cap, ans);
    ob_put(&o, &on, &cap, tcom + 32, 32);
    ob_put(&o, &on, &cap, tcmb + 32, 32);
    for (int i = 0; i < nq; ++i) ob_put(&o, &on, &cap, &col_q[i], 8);
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < X_COLS; ++j) ob_fe(&o, &on, &cap, cwb[(long long)j * cw + col_q[i]]);
    for (long long t = 0; t < cw; ++t) ob_fe(&o, &on, &cap, cc[t]);
    for (long long t = 0; t < col; ++t) ob_fe(&o, &on, &cap, cm[t]);
    ob_siblings(&o, &on, &cap, tcom, rows, col_q, nq);
    ob_siblings(&o, &on, &cap, tcmb, rows, col_q, nq);
    *proof = o;
    *proof_len = on;
    answer[0] = ans.re;
    answer[1] = ans.im;
    free(poly);
    free(cwb);
    free(tcom);
    free(tcmb);
    free(r0);
    free(r1);
    free(cc);
    free(cm);
    return 0;
}

/* The Orion NIZK transcript (OZ/src/truepix/orion_fs.cpp:24-113) and the
 * opening's query draw (linearPC.cpp:711-758): reset ("truepix/orion-fs/v1"),
 * "tensor-open", the opened root, N, the group flag, r0 then r1 (real, imag
 * as separate absorbs), then query_count = (int)(-128 / log2(0.93)) = 1222
 * positions from the opened group (even positions for the masked
 * polynomial, odd for the mask). */
typedef struct { unsigned char state[32], pool[32]; int used; } ofs;
static void ofs_hash64(const unsigned char b[64], unsigned char out[32]) { or_sha256(b, 64, out); }
static void ofs_absorb(ofs* f, const void* data, size_t n)
{
    const unsigned char* p = (const unsigned char*)data;
    while (n > 0) {
        unsigned char b[64];
        memset(b, 0, 64);
        memcpy(b, f->state, 32);
        const size_t take = n < 32 ? n : 32;
        memcpy(b + 32, p, take);
        ofs_hash64(b, f->state);
        p += take;
        n -= take;
    }
    f->used = 32;
}
static void ofs_reset(ofs* f)
{
    memset(f->state, 0, 32);
    ofs_absorb(f, "truepix/orion-fs/v1", 19);
}
static void ofs_u64(ofs* f, u64 x)
{
    unsigned char b[8];
    put_le64(b, x);
    ofs_absorb(f, b, 8);
}
static u64 ofs_squeeze(ofs* f)
{
    if (f->used + 8 > 32) {
        unsigned char b[64];
        memset(b, 0, 64);
        memcpy(b, f->state, 32);
        b[32] = 1;
        ofs_hash64(b, f->pool);
        b[32] = 2;
        ofs_hash64(b, f->state);
        f->used = 0;
    }
    u64 v = 0;
    for (int k = 7; k >= 0; --k) v = (v << 8) | f->pool[f->used + k];
    f->used += 8;
    return v;
}

/* The opening's public inputs absorbed as tensor_product_protocol absorbs
 * them: "tensor-open", the opened root, N, the group flag, r0 then r1 (eq
 * weights with the first challenge on the top index bit). */
static void ofs_open_prefix(ofs* f, const unsigned char root[32], long long n, int use_group1,
                            const u64* r, int n_r)
{
    ofs_reset(f);
    ofs_absorb(f, "tensor-open", 11);
    ofs_absorb(f, root, 32);
    ofs_u64(f, (u64)n);
    ofs_u64(f, use_group1 ? 1 : 0);
    const int lc = 7;
    const long long col = n / X_COLS;
    for (long long i = 0; i < X_COLS; ++i) {
        fp2 v = mk(1, 0);
        for (int d = 0; d < lc; ++d) {
            const fp2 x = mk(r[2 * d], r[2 * d + 1]);
            v = mul2(v, ((i >> (lc - 1 - d)) & 1) ? x : sub2(mk(1, 0), x));
        }
        ofs_u64(f, v.re);
        ofs_u64(f, v.im);
    }
    for (long long t = 0; t < col; ++t) {
        fp2 v = mk(1, 0);
        for (int d = 0; d < n_r - lc; ++d) {
            const fp2 x = mk(r[2 * (lc + d)], r[2 * (lc + d) + 1]);
            v = mul2(v, ((t >> (n_r - lc - 1 - d)) & 1) ? x : sub2(mk(1, 0), x));
        }
        ofs_u64(f, v.re);
        ofs_u64(f, v.im);
    }
}

static const int X_QUERIES = 1222;   /* (int)(-128 / log2(1 - 0.07)) */

/* The positions or_orion_open_nizk opens: writes query_count positions to q
 * (capacity >= 1222) given the opened root (digest bytes), N, the group, r. */
int or_orion_draw_queries(const unsigned char root[32], long long n, int use_group1, long long cw,
                          const u64* r, int n_r, long long* q)
{
    ofs f;
    ofs_open_prefix(&f, root, n, use_group1, r, n_r);
    const int nq = (int)(-128.0 / log2(1.0 - 0.07));
    const long long g = use_group1 ? (cw + 1) / 2 : cw / 2;
    for (int i = 0; i < nq; ++i) q[i] = 2 * (long long)(ofs_squeeze(&f) % (u64)g) + (use_group1 ? 0 : 1);
    return nq;
}


/* ------------------------------------------- Virgo code-switching proof --
 * The proof that closes the NIZK opening (OZ/src/VPD/linearPC.cpp:875-905):
 * the queried entries of the combined codeword are the code of the combined
 * message, proved by the Virgo GKR prover over the expander-encoding circuit
 * and checked by its verifier, both run by the opener.
 *   - transcript: after the opening's absorbs and its 1222 group draws, 1222
 *     positions q from [0, cw) (kept sorted), "combined-mt", the combined
 *     root; every challenge after that is squeeze % p for re then im
 *     (orion_fs.cpp:74-113, verifier.cpp:936-940);
 *   - circuit (generate_circuit, linearPC.cpp:477-669): layer 0 the message
 *     (input gates), layer 1 direct relays of it, then one layer per expander
 *     step: relays of everything so far plus one custom linear combination
 *     per right vertex of C[d] (going down) or D[d] (coming up), then the
 *     relays of the queried positions; each layer's gate count comes from
 *     prepare_gates_count's arithmetic, quirks kept, rounded up to 2^bits;
 *   - evaluation gate by gate (prover.cpp:63-170);
 *   - per layer i = top .. 1 (verifier.cpp:1031-1211): r_u, r_v drawn (r_v
 *     zero at layer 1); phase 1 (prover.cpp:246-633) scatters
 *     alpha eq(r_0, g) + beta eq(r_1, g) into add_mult_sum / addV_array by
 *     gate type and runs bits(i-1) rounds of the quadratic sumcheck; phase 2
 *     (prover.cpp:634-1006) scatters the gates' terms at v_u and runs
 *     bits(i-1) more; v_u = V(r_u), v_v = V(r_v); the verifier's predicates
 *     (verifier.cpp:349-753) and direct_relay (:754-766) close the layer;
 *     alpha, beta drawn; the claim becomes alpha v_u + beta v_v (v_u at
 *     layer 1);
 *   - the input layer's claim is checked by evaluating the message's
 *     multilinear extension at the last r_u (what the reference's FRI-based
 *     VPD establishes).
 * The output words follow ref_harness_virgo.cpp's layout. */
typedef struct { fp2 a, b; } lpoly;          /* a X + b */
typedef struct {
    int bits;
    long long n;            /* 2^bits gates */
    int* ty;                /* 2 dummy, 3 input, 4 direct relay, 10 relay, 14 custom */
    long long* u;
    long long* off;         /* custom gates' sources: off[g] .. off[g + 1] */
    long long* src;
    fp2* w;
    long long n_src, cap_src;
} vlayer;

typedef struct { vlayer* L; int depth; long long cnt[128]; int n_cnt; } vcirc;
static int v_overflow;   /* a block of gates past its layer's 2^bits */
/* A planted mistake for the tests of the verifier's checks: 1 a phase-1
 * polynomial, 2 v_v, 3 the relay predicate (all of the output layer), 4 the
 * input layer's value; 0 none. */
static int v_plant;
void or_orion_virgo_plant(int k) { v_plant = k; }

static void vc_count(vcirc* c, long long key, long long val)
{
    c->cnt[key] = val;
    if (key + 1 > c->n_cnt) c->n_cnt = (int)key + 1;
}

/* prepare_enc_count (linearPC.cpp:477-488): so_far excludes the input here */
static void vc_prepare(vcirc* c, long long in, long long so_far, int dep, long long* od, long long* os)
{
    if (in <= X_SMALL) {
        *od = dep;
        *os = so_far;
        return;
    }
    vc_count(c, dep + 1, so_far + in + XG_C[dep].R);
    long long d2, s2;
    vc_prepare(c, XG_C[dep].R, so_far + in, dep + 1, &d2, &s2);
    vc_count(c, d2 + 1, s2 + XG_D[dep].R);
    *od = d2 + 1;
    *os = c->cnt[d2 + 1];
}

static int vc_log2(long long x)
{
    int b = 0;
    while ((1LL << b) < x) ++b;
    return b;
}

static void vl_alloc(vlayer* l, int bits)
{
    l->bits = bits;
    l->n = 1LL << bits;
    l->ty = (int*)malloc(sizeof(int) * (size_t)l->n);
    l->u = (long long*)calloc((size_t)l->n, sizeof(long long));
    l->off = (long long*)calloc((size_t)l->n + 1, sizeof(long long));
    for (long long g = 0; g < l->n; ++g) l->ty[g] = 2;
    l->n_src = 0;
    l->cap_src = 1024;
    l->src = (long long*)malloc(sizeof(long long) * (size_t)l->cap_src);
    l->w = (fp2*)malloc(sizeof(fp2) * (size_t)l->cap_src);
}

static void vl_free(vlayer* l)
{
    free(l->ty);
    free(l->u);
    free(l->off);
    free(l->src);
    free(l->w);
}

/* custom gates at first .. first + R - 1 of layer l: the right vertices of
 * graph g, sources offset by src_off (in the order r_neighbor lists them:
 * by left vertex, then degree) */
static void vl_custom(vlayer* l, long long first, const xgraph* g, long long src_off)
{
    if (first + g->R > l->n) {
        v_overflow = 1;
        return;
    }
    long long* cnt = (long long*)calloc((size_t)g->R + 1, sizeof(long long));
    for (long long k = 0; k < g->L * g->deg; ++k) ++cnt[g->to[k] + 1];
    for (long long t = 0; t < g->R; ++t) cnt[t + 1] += cnt[t];
    const long long base = l->n_src;
    l->n_src += g->L * g->deg;
    if (l->n_src > l->cap_src) {
        l->cap_src = l->n_src;
        l->src = (long long*)realloc(l->src, sizeof(long long) * (size_t)l->cap_src);
        l->w = (fp2*)realloc(l->w, sizeof(fp2) * (size_t)l->cap_src);
    }
    long long* fill = (long long*)malloc(sizeof(long long) * (size_t)(g->R + 1));
    memcpy(fill, cnt, sizeof(long long) * (size_t)(g->R + 1));
    for (long long i = 0; i < g->L; ++i)
        for (int d = 0; d < g->deg; ++d) {
            const long long k = i * g->deg + d;
            const long long at = base + fill[g->to[k]]++;
            l->src[at] = i + src_off;
            l->w[at] = g->w[k];
        }
    for (long long t = 0; t < g->R; ++t) {
        l->ty[first + t] = 14;
        l->off[first + t] = base + cnt[t];
    }
    /* off[] of a custom gate starts its range; the next range starts at its end */
    l->off[first + g->R] = base + cnt[g->R];
    free(cnt);
    free(fill);
}

static void vl_relays(vlayer* l, long long k)
{
    if (k > l->n) {
        v_overflow = 1;
        return;
    }
    for (long long g = 0; g < k; ++g) {
        l->ty[g] = 10;
        l->u[g] = g;
    }
}

/* generate_enc_circuit (linearPC.cpp:514-578): so_far includes the input */
static void vc_gen(vcirc* c, long long in, long long so_far, int rd, int in_depth, long long* od,
                   long long* os)
{
    if (in <= X_SMALL) {
        *od = in_depth;
        *os = so_far;
        return;
    }
    vlayer* l = &c->L[in_depth + 1];
    vl_relays(l, so_far);
    vl_custom(l, so_far, &XG_C[rd], so_far - in);
    long long d2, s2;
    vc_gen(c, XG_C[rd].R, so_far + XG_C[rd].R, rd + 1, in_depth + 1, &d2, &s2);
    const long long fd = d2 + 1;
    vlayer* m = &c->L[fd];
    vl_relays(m, s2);
    vl_custom(m, s2, &XG_D[rd], so_far);
    *od = fd;
    *os = s2 + XG_D[rd].R;
}

/* generate_circuit (linearPC.cpp:580-669) for the sorted queries q */
static void vc_build(vcirc* c, long long N, const long long* q, int nq)
{
    memset(c->cnt, 0, sizeof(c->cnt));
    c->n_cnt = 0;
    vc_count(c, 0, N);
    vc_count(c, 1, N + XG_C[0].R);
    long long d, sz;
    vc_prepare(c, XG_C[0].R, N, 1, &d, &sz);
    vc_count(c, d + 1, N + sz + XG_D[0].R);
    vc_count(c, d + 2, nq);
    c->depth = c->n_cnt + 1;
    c->L = (vlayer*)calloc((size_t)c->depth, sizeof(vlayer));
    vl_alloc(&c->L[0], vc_log2(N));
    for (int i = 0; i < c->n_cnt; ++i) vl_alloc(&c->L[i + 1], vc_log2(c->cnt[i]));
    for (long long i = 0; i < N; ++i) {
        c->L[0].ty[i] = 3;
        c->L[1].ty[i] = 4;
        c->L[1].u[i] = i;
    }
    vl_relays(&c->L[2], N);
    vl_custom(&c->L[2], N, &XG_C[0], 0);
    long long d2, s2;
    vc_gen(c, XG_C[0].R, N + XG_C[0].R, 1, 2, &d2, &s2);
    const long long fd = d2 + 1;
    vl_relays(&c->L[fd], s2);
    vl_custom(&c->L[fd], s2, &XG_D[0], N);
    if (nq > c->L[fd + 1].n) v_overflow = 1;
    for (int i = 0; i < nq && !v_overflow; ++i) {
        c->L[fd + 1].ty[i] = 10;
        c->L[fd + 1].u[i] = q[i];
    }
}

static long long vl_src_end(const vlayer* l, long long g)
{
    /* the range of a custom gate ends where the next gate's begins; the last
     * custom gate of a block has its end stored one past it */
    return l->off[g + 1];
}

/* eq(r, x) over bits with bit k of x on r[k] (beta tables of prover.cpp:262-284) */
static fp2* v_eq(const fp2* r, int bits, fp2 scale)
{
    const long long n = 1LL << bits;
    fp2* t = (fp2*)malloc(sizeof(fp2) * (size_t)n);
    t[0] = scale;
    for (int k = 0; k < bits; ++k)
        for (long long j = 0; j < (1LL << k); ++j) {
            t[j | (1LL << k)] = mul2(t[j], r[k]);
            t[j] = mul2(t[j], one_minus(r[k]));
        }
    return t;
}

/* one round of sumcheck_phase1_update / phase2_update (prover.cpp:576-633) */
static void v_round(lpoly* amul, lpoly* addv, lpoly* vm, long long* total, fp2 prev, int bit, fp2 out[3])
{
    const long long h = *total >> 1;
    for (long long i = 0; i < h; ++i) {
        lpoly* arr[3] = {vm, addv, amul};
        for (int k = 0; k < 3; ++k) {
            lpoly z = arr[k][2 * i], o = arr[k][2 * i + 1], nw;
            if (bit == 0) {
                nw.b = z.b;
                nw.a = sub2(o.b, nw.b);
            } else {
                nw.b = add2(mul2(z.a, prev), z.b);
                nw.a = sub2(add2(mul2(o.a, prev), o.b), nw.b);
            }
            arr[k][i] = nw;
        }
    }
    fp2 a = mk(0, 0), b = mk(0, 0), c = mk(0, 0);
    for (long long i = 0; i < h; ++i) {
        a = add2(a, mul2(amul[i].a, vm[i].a));
        b = add2(b, add2(add2(mul2(amul[i].a, vm[i].b), mul2(amul[i].b, vm[i].a)), addv[i].a));
        c = add2(c, add2(mul2(amul[i].b, vm[i].b), addv[i].b));
    }
    out[0] = a;
    out[1] = b;
    out[2] = c;
    *total = h;
}

static fp2 v_qeval(const fp2 q[3], fp2 x) { return add2(mul2(add2(mul2(q[0], x), q[1]), x), q[2]); }

/* V_res (prover.cpp:29-52): the multilinear value of vals at r, bit k on r[k] */
static fp2 v_res(const fp2* vals, int bits, const fp2* r)
{
    long long n = 1LL << bits;
    fp2* t = (fp2*)malloc(sizeof(fp2) * (size_t)n);
    memcpy(t, vals, sizeof(fp2) * (size_t)n);
    for (int k = 0; k < bits; ++k) {
        for (long long j = 0; j < (n >> 1); ++j)
            t[j] = add2(mul2(t[2 * j], one_minus(r[k])), mul2(t[2 * j + 1], r[k]));
        n >>= 1;
    }
    const fp2 v = t[0];
    free(t);
    return v;
}

typedef struct { u64* w; long long cap, n; } vout;
static void vo_put(vout* o, u64 x)
{
    if (o->n < o->cap) o->w[o->n] = x;
    ++o->n;
}
static void vo_fe(vout* o, fp2 x)
{
    vo_put(o, x.re);
    vo_put(o, x.im);
}
static fp2 v_squeeze_fe(ofs* f)
{
    const u64 re = ofs_squeeze(f) % P61;
    return mk(re, ofs_squeeze(f) % P61);
}
static int cmp_ll(const void* a, const void* b)
{
    const long long x = *(const long long*)a, y = *(const long long*)b;
    return x < y ? -1 : x > y;
}

long long or_orion_virgo(const unsigned char com_root[32], long long n, int use_group1, const u64* r,
                         int n_r, const unsigned char combined_root[32], const u64* msg, long long N,
                         u64* out, long long cap)
{
    if (n != N * X_COLS || N < 1 || (N & (N - 1))) return -1;
    /* a fresh process's graphs: below the threshold expander_init draws none
     * and generate_circuit reads C[0] and D[0] empty */
    for (int k = 0; k < 64; ++k) {
        xg_make(&XG_C[k], NULL, 0, 0, 0);
        xg_make(&XG_D[k], NULL, 0, 0, 0);
    }
    grand rg;
    grand_init(&rg, 1);
    xg_init(&rg, N, 0);
    const long long cw = x_len(N, 0);
    ofs f;
    ofs_open_prefix(&f, com_root, n, use_group1, r, n_r);
    const long long g = use_group1 ? (cw + 1) / 2 : cw / 2;
    for (int i = 0; i < X_QUERIES; ++i) (void)(ofs_squeeze(&f) % (u64)g);
    long long* q = (long long*)malloc(sizeof(long long) * X_QUERIES);
    for (int i = 0; i < X_QUERIES; ++i) q[i] = (long long)(ofs_squeeze(&f) % (u64)cw);
    ofs_absorb(&f, "combined-mt", 11);
    ofs_absorb(&f, combined_root, 32);
    qsort(q, X_QUERIES, sizeof(long long), cmp_ll);

    vcirc C;
    v_overflow = 0;
    vc_build(&C, N, q, X_QUERIES);
    const int D = C.depth, top = D - 1;
    if (v_overflow) {
        for (int i = 0; i < D; ++i) vl_free(&C.L[i]);
        free(C.L);
        free(q);
        return -1;
    }
    vout o = {out, cap, 0};
    vo_put(&o, (u64)D);
    for (int i = 0; i < D; ++i) vo_put(&o, (u64)C.L[i].bits);
    vo_put(&o, (u64)X_QUERIES);
    for (int i = 0; i < X_QUERIES; ++i) vo_put(&o, (u64)q[i]);

    /* evaluate (prover.cpp:63-170) */
    fp2** V = (fp2**)calloc((size_t)D, sizeof(fp2*));
    V[0] = (fp2*)calloc((size_t)C.L[0].n, sizeof(fp2));
    for (long long i = 0; i < N; ++i) V[0][i] = mk(msg[2 * i], msg[2 * i + 1]);
    for (int i = 1; i < D; ++i) {
        const vlayer* l = &C.L[i];
        V[i] = (fp2*)calloc((size_t)l->n, sizeof(fp2));
        for (long long gg = 0; gg < l->n; ++gg) {
            if (l->ty[gg] == 4 || l->ty[gg] == 10) V[i][gg] = V[i - 1][l->u[gg]];
            else if (l->ty[gg] == 14) {
                fp2 acc = mk(0, 0);
                for (long long e = l->off[gg]; e < vl_src_end(l, gg); ++e)
                    acc = add2(acc, mul2(V[i - 1][l->src[e]], l->w[e]));
                V[i][gg] = acc;
            }
        }
    }

    int ok = 1;
    fp2 alpha = mk(1, 0), beta = mk(0, 0);
    int rb = C.L[top].bits;
    fp2* r0 = (fp2*)malloc(sizeof(fp2) * (size_t)rb);
    fp2* r1 = (fp2*)malloc(sizeof(fp2) * (size_t)rb);
    for (int k = 0; k < rb; ++k) r0[k] = v_squeeze_fe(&f);
    for (int k = 0; k < rb; ++k) r1[k] = v_squeeze_fe(&f);
    for (int k = 0; k < rb; ++k) vo_fe(&o, r0[k]);
    for (int k = 0; k < rb; ++k) vo_fe(&o, r1[k]);
    fp2 claim = mul2(alpha, v_res(V[top], rb, r0));
    vo_fe(&o, claim);
    for (int i = top; i >= 1; --i) {
        const vlayer* l = &C.L[i];
        const int bl = C.L[i - 1].bits;
        const long long nu = 1LL << bl;
        fp2* ru = (fp2*)malloc(sizeof(fp2) * (size_t)bl);
        fp2* rv = (fp2*)malloc(sizeof(fp2) * (size_t)bl);
        for (int k = 0; k < bl; ++k) ru[k] = v_squeeze_fe(&f);
        for (int k = 0; k < bl; ++k) rv[k] = v_squeeze_fe(&f);
        if (i == 1)
            for (int k = 0; k < bl; ++k) rv[k] = mk(0, 0);
        for (int k = 0; k < bl; ++k) vo_fe(&o, ru[k]);
        for (int k = 0; k < bl; ++k) vo_fe(&o, rv[k]);
        /* direct_relay (verifier.cpp:754-766), over layer i's bits */
        fp2 dr0 = mk(0, 0), dr1 = mk(0, 0);
        if (i == 1) {
            dr0 = dr1 = mk(1, 0);
            for (int k = 0; k < l->bits; ++k) {
                dr0 = mul2(dr0, add2(sub2(sub2(mk(1, 0), r0[k]), ru[k]), mul2(mk(2, 0), mul2(r0[k], ru[k]))));
                dr1 = mul2(dr1, add2(sub2(sub2(mk(1, 0), r1[k]), ru[k]), mul2(mk(2, 0), mul2(r1[k], ru[k]))));
            }
        }
        const fp2 drv = add2(mul2(alpha, dr0), mul2(beta, dr1));
        /* phase 1 */
        fp2* bg0 = v_eq(r0, l->bits, alpha);
        fp2* bg1 = v_eq(r1, l->bits, beta);
        lpoly* amul = (lpoly*)calloc((size_t)nu, sizeof(lpoly));
        lpoly* addv = (lpoly*)calloc((size_t)nu, sizeof(lpoly));
        lpoly* vm = (lpoly*)calloc((size_t)nu, sizeof(lpoly));
        for (long long x = 0; x < nu; ++x) vm[x].b = V[i - 1][x];
        for (long long gg = 0; gg < l->n; ++gg) {
            const int t = l->ty[gg];
            if (t == 4) {
                const long long u = l->u[gg];
                amul[u].b = add2(amul[u].b, add2(bg0[u], bg1[u]));
            } else if (t == 10) {
                const long long u = l->u[gg];
                amul[u].b = add2(amul[u].b, add2(bg0[gg], bg1[gg]));
            } else if (t == 14) {
                const fp2 tmp = add2(bg0[gg], bg1[gg]);
                for (long long e = l->off[gg]; e < vl_src_end(l, gg); ++e)
                    amul[l->src[e]].b = add2(amul[l->src[e]].b, mul2(l->w[e], tmp));
            }
        }
        long long total = nu;
        fp2 prev = mk(0, 0);
        for (int k = 0; k < bl; ++k) {
            fp2 pq[3];
            v_round(amul, addv, vm, &total, prev, k, pq);
            if (v_plant == 1 && i == top && k == 0) pq[2] = add2(pq[2], mk(1, 0));
            vo_fe(&o, pq[0]);
            vo_fe(&o, pq[1]);
            vo_fe(&o, pq[2]);
            prev = ru[k];
            if (!eq2_(add2(v_qeval(pq, mk(0, 0)), v_qeval(pq, mk(1, 0))), claim)) ok = 0;
            claim = v_qeval(pq, ru[k]);
        }
        /* phase 2 */
        const fp2 v_u = add2(mul2(vm[0].a, prev), vm[0].b);
        fp2* bu = v_eq(ru, bl, mk(1, 0));
        memset(amul, 0, sizeof(lpoly) * (size_t)nu);
        memset(addv, 0, sizeof(lpoly) * (size_t)nu);
        for (long long x = 0; x < nu; ++x) {
            vm[x].a = mk(0, 0);
            vm[x].b = V[i - 1][x];
        }
        for (long long gg = 0; gg < l->n; ++gg) {
            const int t = l->ty[gg];
            if (t == 10) {
                const fp2 tg = add2(bg0[gg], bg1[gg]);
                addv[0].b = add2(addv[0].b, mul2(mul2(tg, bu[l->u[gg]]), v_u));
            } else if (t == 14) {
                const fp2 tgv = mul2(add2(bg0[gg], bg1[gg]), v_u);
                for (long long e = l->off[gg]; e < vl_src_end(l, gg); ++e)
                    addv[0].b = add2(addv[0].b, mul2(mul2(tgv, bu[l->src[e]]), l->w[e]));
            }
        }
        total = nu;
        prev = mk(0, 0);
        for (int k = 0; k < bl; ++k) {
            fp2 pq[3];
            v_round(amul, addv, vm, &total, prev, k, pq);
            vo_fe(&o, pq[0]);
            vo_fe(&o, pq[1]);
            vo_fe(&o, pq[2]);
            prev = rv[k];
            if (!eq2_(add2(add2(v_qeval(pq, mk(0, 0)), v_qeval(pq, mk(1, 0))), mul2(drv, v_u)), claim)) ok = 0;
            claim = add2(v_qeval(pq, rv[k]), mul2(drv, v_u));
        }
        fp2 v_v = add2(mul2(vm[0].a, prev), vm[0].b);
        if (v_plant == 2 && i == top) v_v = add2(v_v, mk(1, 0));
        vo_fe(&o, v_u);
        vo_fe(&o, v_v);
        /* predicates (verifier.cpp:349-753): zero at layer 1 */
        fp2 relay = mk(0, 0), cust = mk(0, 0);
        if (i != 1) {
            fp2* bv = v_eq(rv, bl, mk(1, 0));
            for (long long gg = 0; gg < l->n; ++gg) {
                const fp2 tg = add2(bg0[gg], bg1[gg]);
                if (l->ty[gg] == 10) relay = add2(relay, mul2(tg, bu[l->u[gg]]));
                else if (l->ty[gg] == 14)
                    for (long long e = l->off[gg]; e < vl_src_end(l, gg); ++e)
                        cust = add2(cust, mul2(mul2(mul2(tg, bv[0]), bu[l->src[e]]), l->w[e]));
            }
            relay = mul2(relay, bv[0]);
            free(bv);
            if (v_plant == 3 && i == top) relay = add2(relay, mk(1, 0));
        }
        vo_fe(&o, relay);
        vo_fe(&o, cust);
        vo_fe(&o, drv);
        if (!eq2_(claim, add2(add2(mul2(relay, v_u), mul2(cust, v_u)), mul2(drv, v_u)))) ok = 0;
        alpha = v_squeeze_fe(&f);
        beta = v_squeeze_fe(&f);
        vo_fe(&o, alpha);
        vo_fe(&o, beta);
        claim = i != 1 ? add2(mul2(alpha, v_u), mul2(beta, v_v)) : v_u;
        free(bg0);
        free(bg1);
        free(bu);
        free(amul);
        free(addv);
        free(vm);
        free(r0);
        free(r1);
        r0 = ru;
        r1 = rv;
    }
    vo_fe(&o, claim);
    fp2 vin = v_res(V[0], C.L[0].bits, r0);
    if (v_plant == 4) vin = add2(vin, mk(1, 0));
    vo_fe(&o, vin);
    if (!eq2_(vin, claim)) ok = 0;
    vo_put(&o, ok ? 1 : 0);
    free(r0);
    free(r1);
    for (int i = 0; i < D; ++i) {
        free(V[i]);
        vl_free(&C.L[i]);
    }
    free(V);
    free(C.L);
    free(q);
    return o.n;
}

void or_free(void* p) { free(p); }
