// ntt.hip - Reed-Solomon encoding by a two-pass, LDS-staged Fp2 NTT on CDNA4.
//
// Computes what OZ/src/infrastructure/RS_polynomial.cpp:46-130
// (fast_fourier_transform) computes: for each column of M = 2^c coefficients,
// the N = 2^o evaluations result[i] = sum_{j<M} c_j w^(i j), natural order,
// w = the root of unity of order N from Orion's generator
// g = 2^31 + 1033321771269002680 i (OZ/src/linear_gkr/prime_field.cpp:60-73),
// w = g^(2^(62 - o)).  The reference replicates the coefficients N/M times
// and runs c radix-2 Stockham stages over all N points; here the zero-padded
// transform is split four-step style, N = (S M2) x M1 with S = N/M,
// M = M1 M2, j = j1 + M1 j2, i = u + (S M2) q1:
//   pass A (per column j1): Y[u][j1] = w^(j1 u) * sum_{j2<M2} c[j1 + M1 j2] w_{S M2}^(j2 u)
//          computed as S cosets (u = s + S q2): twist by w_{S M2}^(s j2), then
//          an M2-point DFT over j2 giving q2;
//   pass B (per row u):    result[u + S M2 q1] = sum_{j1<M1} Y[u][j1] w_{M1}^(j1 q1).
// Each pass stages 4 rows of at most 1024 points in LDS (4 x 1025 x 16 B) and
// runs radix-8 Stockham stages (after one radix-2 or radix-4 stage when the
// length is not a power of 8); the 4-row tiles make every strided global
// load and store a 64-byte segment (4 consecutive Fp2), and two 512-thread
// workgroups share a CU, so one loads or stores its tile while the other
// computes (one 8-row, 1024-thread tile per CU left each phase idle in turn:
// 18% slower).  In both passes the DFT's first stage reads its rows from HBM
// (pass A twisting them as it loads) and its last stage writes the results
// there (pass A multiplying by w^(j1 u) as it stores), so only the stages in
// between go through LDS (10% faster than staging both ends).  Columns are
// processed a few at a time so the intermediate stays in the 256 MB MALL.
// Roots of order 4 are +-i in Fp2, so multiplying by w4 is a coordinate swap
// and one negation; roots of order 8 are 2^30 (+-1 +-i), so multiplying by w8
// is a coordinate mix and a 30-bit rotation of each 61-bit coordinate.
#include "fp61.hpp"
#include "runtime.hpp"

#include <stdint.h>
#include <mutex>

using namespace tp;

namespace {

struct alignas(16) V2 { u64 re, im; };

__device__ __forceinline__ fp2 ldv(const V2* p) { const V2 v = *p; return mk2(v.re, v.im); }
__device__ __forceinline__ void stv(V2* p, fp2 x) { V2 v; v.re = x.re; v.im = x.im; *p = v; }

constexpr int kRows = 4;          // rows per LDS tile (66 KB: two tiles per CU)
constexpr int kThreads = 512;     // threads per NTT workgroup (8 waves; 16 a CU)
constexpr int kMaxLog = 10;       // LDS DFT length up to 2^10
constexpr int kRowStride = (1 << kMaxLog) + 1;   // +1 element pads LDS banks

// Powers of the order-N root: w^e = TL[e & (2^h - 1)] * TH[e >> h].
struct Tw {
    const V2* t1024;   // w_1024^k, k < 1024
    const V2* tl;      // w^e, e < 2^h
    const V2* th;      // w^(e 2^h), e < 2^(o - h)
    int h;
    int w4_is_i;       // w_4 == +i (else -i)
    int w8_s1, w8_s2;  // w_8 = 2^30 (s1 + s2 i), s1, s2 in {+1, -1}
};

__device__ __forceinline__ fp2 wpow(const Tw& t, u64 e)
{
    return mul2(ldv(t.tl + (e & ((1ULL << t.h) - 1))), ldv(t.th + (e >> t.h)));
}

__device__ __forceinline__ fp2 mul_w4(fp2 v, int is_i)
{
    // (a + b i) * i = -b + a i;  (a + b i) * (-i) = b - a i
    return is_i ? mk2(fneg(v.im), v.re) : mk2(v.im, fneg(v.re));
}

// x 2^30 mod p for a canonical x: a 30-bit rotation of the 61-bit word
// (2^61 = 1), so the result is canonical again.
__device__ __forceinline__ u64 rot30(u64 x) { return ((x << 30) & P) | (x >> 31); }

// v * w_8 with w_8 = 2^30 (s1 + s2 i) (2^30 (1 + i) squares to 2^61 i = i):
// a coordinate mix by signs and two rotations, no field multiply.
__device__ __forceinline__ fp2 mul_w8(fp2 v, int s1, int s2)
{
    const u64 a = v.re, b = v.im;
    const u64 re = fsub(s1 > 0 ? a : fneg(a), s2 > 0 ? b : fneg(b));   // s1 a - s2 b
    const u64 im = fadd(s2 > 0 ? a : fneg(a), s1 > 0 ? b : fneg(b));   // s2 a + s1 b
    return mk2(rot30(re), rot30(im));
}

// Where a DFT's first stage reads its rows and its last stage writes them.
enum InMode {
    kInLds,    // LDS
    kInRows,   // row r, element i at in[r 2^l + i]
    kInCols,   // row r, element i at in[r + i in_step], times w^(in_tw i)
};
enum OutMode {
    kOutLds,      // LDS
    kOutCols,     // row r, element i to out[r + i out_step]
    kOutColsTw,   // the same, times w^((r0 + r)(out_e0 + out_e1 i))
};
struct GlobalIO {
    const V2* in;
    size_t in_step;
    u64 in_tw;
    V2* out;
    size_t out_step;
    u64 r0, out_e0, out_e1;
    u64 nmask;   // exponents are taken mod the order of w (a power of 2)
};

// Forward DFT of `rows` rows of length 2^l, Stockham auto-sort: stage with
// span Ns reads x[j + i L/R], multiplies by w_{Ns R}^((j mod Ns) i), applies
// DFT_R and writes y[(j / Ns) Ns R + (j mod Ns) + i Ns].  Between stages the
// rows sit in LDS (row stride kRowStride); values pass through registers with a
// barrier between the read and write halves, so one LDS buffer suffices.
// The first stage reads as kIn says, the last writes as kOut says (anything
// but kOutLds needs l >= 3, so the last stage is radix 8; then nothing is left
// in LDS).  Stages that touch rows interleaved in memory (kInCols, and the
// last stage when it stores) put the row on the fastest lane bits, so `rows`
// consecutive lanes move `rows` consecutive words.
template <int kIn, int kOut>
__device__ void lds_dft(fp2* lds, int rows, int log_rows, int l, const Tw& tw, const GlobalIO& io)
{
    const int L = 1 << l;
    int Ns = 1;
    int log_ns = 0;
    auto at = [&](int b, int lnb, int nb, bool row_fast, int& row, int& j) {
        if (row_fast) {
            row = b & (rows - 1);
            j = b >> log_rows;
        } else {
            row = b >> lnb;
            j = b & (nb - 1);
        }
    };
    // the first stage's R inputs x[j + m nb] of one row (twisted for kInCols:
    // w^(in_tw j) times ratio^m, ratio = w^(in_tw nb))
    auto first = [&](auto& a, int row, int j, int nb, fp2 ratio) {
        constexpr int R = sizeof(a) / sizeof(a[0]);
        if (kIn == kInLds) {
#pragma unroll
            for (int m = 0; m < R; ++m) a[m] = lds[row * kRowStride + j + m * nb];
        } else if (kIn == kInRows) {
#pragma unroll
            for (int m = 0; m < R; ++m) a[m] = ldv(io.in + ((size_t)row << l) + j + m * nb);
        } else {
#pragma unroll
            for (int m = 0; m < R; ++m) a[m] = ldv(io.in + row + io.in_step * (size_t)(j + m * nb));
            if (io.in_tw) {
                fp2 t = wpow(tw, (io.in_tw * (u64)j) & io.nmask);
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    a[m] = mul2(a[m], t);
                    if (m + 1 < R) t = mul2(t, ratio);
                }
            }
        }
    };
    auto in_ratio = [&](int nb) {
        return (kIn == kInCols && io.in_tw) ? wpow(tw, (io.in_tw * (u64)nb) & io.nmask) : one2();
    };
    constexpr bool kColsIn = kIn == kInCols;
    // l = 3 q + rem: one radix-2 (rem 1) or radix-4 (rem 2) stage, then radix 8
    if (l % 3 == 1) {   // radix-2 stage (twiddles of a span-1 stage are all 1)
        const int nb = L >> 1, total = rows * nb;
        constexpr int IT = (kRows << (kMaxLog - 1)) / kThreads;
        const fp2 ratio = in_ratio(nb);
        fp2 v[IT][2];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int b = threadIdx.x + it * kThreads;
            if (b < total) {
                int row, j;
                at(b, l - 1, nb, kColsIn, row, j);
                first(v[it], row, j, nb, ratio);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int b = threadIdx.x + it * kThreads;
            if (b < total) {
                int row, j;
                at(b, l - 1, nb, kColsIn, row, j);
                fp2* y = lds + row * kRowStride;
                y[2 * j] = add2(v[it][0], v[it][1]);
                y[2 * j + 1] = sub2(v[it][0], v[it][1]);
            }
        }
        __syncthreads();
        Ns = 2;
        log_ns = 1;
    } else if (l % 3 == 2) {   // radix-4 stage, span 1 (twiddles all 1)
        const int nb = L >> 2, total = rows * nb;
        const int lnb = l - 2;
        constexpr int IT = (kRows << (kMaxLog - 2)) / kThreads;
        const fp2 ratio = in_ratio(nb);
        fp2 v[IT][4];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int b = threadIdx.x + it * kThreads;
            if (b < total) {
                int row, j;
                at(b, lnb, nb, kColsIn, row, j);
                fp2 a[4];
                first(a, row, j, nb, ratio);
                const fp2 s02 = add2(a[0], a[2]), d02 = sub2(a[0], a[2]);
                const fp2 s13 = add2(a[1], a[3]), d13 = mul_w4(sub2(a[1], a[3]), tw.w4_is_i);
                v[it][0] = add2(s02, s13);
                v[it][1] = add2(d02, d13);
                v[it][2] = sub2(s02, s13);
                v[it][3] = sub2(d02, d13);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int b = threadIdx.x + it * kThreads;
            if (b < total) {
                int row, j;
                at(b, lnb, nb, kColsIn, row, j);
                fp2* y = lds + row * kRowStride + 4 * j;
                y[0] = v[it][0];
                y[1] = v[it][1];
                y[2] = v[it][2];
                y[3] = v[it][3];
            }
        }
        __syncthreads();
        Ns = 4;
        log_ns = 2;
    }
    while (Ns < L) {   // radix-8 stages: DFT_8 = two DFT_4 and w_8 rotations
        const int nb = L >> 3, total = rows * nb;
        const int lnb = l - 3;
        const int tw_shift = kMaxLog - (log_ns + 3);   // w_{8 Ns}^t = w_1024^(t << shift)
        constexpr int IT = (kRows << (kMaxLog - 3)) / kThreads;
        // the last stage (span L/8, so j < Ns and the output index is j + m Ns)
        // stores to io.out unless kOut is kOutLds
        const bool last = kOut != kOutLds && (Ns << 3) == L;
        const bool from_in = kIn != kInLds && Ns == 1;
        const bool row_fast = last || (kColsIn && Ns == 1);
        const fp2 ratio = from_in ? in_ratio(nb) : one2();
        fp2 v[IT][8];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int b = threadIdx.x + it * kThreads;
            if (b < total) {
                int row, j;
                at(b, lnb, nb, row_fast, row, j);
                fp2 a[8];
                if (from_in) {
                    first(a, row, j, nb, ratio);
                } else {
                    const fp2* x = lds + row * kRowStride;
#pragma unroll
                    for (int m = 0; m < 8; ++m) a[m] = x[j + m * nb];
                }
                const int k = j & (Ns - 1);
                if (k) {
#pragma unroll
                    for (int m = 1; m < 8; ++m)
                        a[m] = mul2(a[m], ldv(tw.t1024 + ((m * k) << tw_shift)));
                }
                fp2 E[4], O[4];
                {
                    const fp2 s02 = add2(a[0], a[4]), d02 = sub2(a[0], a[4]);
                    const fp2 s13 = add2(a[2], a[6]), d13 = mul_w4(sub2(a[2], a[6]), tw.w4_is_i);
                    E[0] = add2(s02, s13);
                    E[1] = add2(d02, d13);
                    E[2] = sub2(s02, s13);
                    E[3] = sub2(d02, d13);
                }
                {
                    const fp2 s02 = add2(a[1], a[5]), d02 = sub2(a[1], a[5]);
                    const fp2 s13 = add2(a[3], a[7]), d13 = mul_w4(sub2(a[3], a[7]), tw.w4_is_i);
                    O[0] = add2(s02, s13);
                    O[1] = mul_w8(add2(d02, d13), tw.w8_s1, tw.w8_s2);
                    O[2] = mul_w4(sub2(s02, s13), tw.w4_is_i);
                    O[3] = mul_w4(mul_w8(sub2(d02, d13), tw.w8_s1, tw.w8_s2), tw.w4_is_i);
                }
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    v[it][m] = add2(E[m], O[m]);
                    v[it][m + 4] = sub2(E[m], O[m]);
                }
            }
        }
        if (last) {
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int b = threadIdx.x + it * kThreads;
                if (b < total) {
                    const int row = b & (rows - 1), j = b >> log_rows;
                    V2* y = io.out + row + io.out_step * (size_t)j;
                    const u64 r = io.r0 + (u64)row;
                    if (kOut == kOutColsTw && r) {
                        // w^(r (e0 + e1 (j + m Ns))) = w^(r (e0 + e1 j)) (w^(r e1 Ns))^m
                        fp2 t = wpow(tw, (r * (io.out_e0 + io.out_e1 * (u64)j)) & io.nmask);
                        const fp2 step = wpow(tw, (r * io.out_e1 * (u64)Ns) & io.nmask);
#pragma unroll
                        for (int m = 0; m < 8; ++m) {
                            stv(y + io.out_step * (size_t)(m * Ns), mul2(v[it][m], t));
                            if (m < 7) t = mul2(t, step);
                        }
                    } else {
#pragma unroll
                        for (int m = 0; m < 8; ++m) stv(y + io.out_step * (size_t)(m * Ns), v[it][m]);
                    }
                }
            }
            return;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int b = threadIdx.x + it * kThreads;
            if (b < total) {
                int row, j;
                at(b, lnb, nb, row_fast, row, j);
                const int k = j & (Ns - 1);
                fp2* y = lds + row * kRowStride + ((j >> log_ns) << (log_ns + 3)) + k;
#pragma unroll
                for (int m = 0; m < 8; ++m) y[m * Ns] = v[it][m];
            }
        }
        __syncthreads();
        Ns <<= 3;
        log_ns += 3;
    }
}

// Pass A.  grid = (M1 / rows, S, columns); rows consecutive j1 per tile.
__global__ void __launch_bounds__(kThreads) k_ntt_pass_a(const V2* __restrict__ coef, size_t coef_stride,
                                                         V2* __restrict__ ybuf, size_t y_stride,
                                                         int a, int b, int log_s, int log_rows, Tw tw,
                                                         int xcd_order)
{
    extern __shared__ fp2 lds[];
    const int M1 = 1 << a, M2 = 1 << b, S = 1 << log_s;
    const int rows = 1 << log_rows;
    int s, j1_0;
    if (xcd_order) {
        // one grid dimension of tiles x cosets, workgroup ids dealt round-robin
        // to the 8 XCDs: XCD x runs, in turn, the S cosets of tile 2(8q + x)
        // and then of tile 2(8q + x) + 1, whose 64-byte segments share
        // 128-byte lines, so those lines stay in that XCD's L2 (tiles % 16 == 0)
        const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
        s = k & (S - 1);
        const int h = (k >> log_s) & 1, q = k >> (log_s + 1);
        j1_0 = (2 * (q * 8 + x) + h) * rows;
    } else {
        s = blockIdx.y;
        j1_0 = blockIdx.x * rows;
    }
    const V2* src = coef + blockIdx.z * coef_stride;
    V2* dst = ybuf + blockIdx.z * y_stride;
    const int total = rows * M2;
    const u64 nmask = ((u64)1 << (a + b + log_s)) - 1;   // exponents mod N (w^N = 1)
    if (b >= 3) {
        // the DFT's first stage loads the rows (x[j1_0 + r + M1 j2] twisted by
        // w_N^(M1 s j2)) and its last stage stores Y[u][j1] = v w_N^(j1 u),
        // u = s + S q2, at ybuf[u M1 + j1]
        GlobalIO io{};
        io.in = src + j1_0;
        io.in_step = (size_t)M1;
        io.in_tw = ((u64)M1 * s) & nmask;
        io.out = dst + (size_t)s * M1 + j1_0;
        io.out_step = (size_t)S * M1;
        io.r0 = (u64)j1_0;
        io.out_e0 = (u64)s;
        io.out_e1 = (u64)S;
        io.nmask = nmask;
        lds_dft<kInCols, kOutColsTw>(lds, rows, log_rows, b, tw, io);
        return;
    }
    // Each thread keeps one row r (kThreads is a multiple of rows) and steps
    // j2 (and q2) by kThreads / rows, so both twiddles advance by a constant
    // ratio: one multiply per element instead of two table loads and a multiply.
    const int step = kThreads >> log_rows;
    // load + coset twist w_{S M2}^(s j2) = w_N^(M1 s j2)
    {
        const int r = threadIdx.x & (rows - 1);
        fp2 cur = s ? wpow(tw, ((u64)M1 * s * (threadIdx.x >> log_rows)) & nmask) : one2();
        const fp2 ratio = s ? wpow(tw, ((u64)M1 * s * step) & nmask) : one2();
        for (int idx = threadIdx.x; idx < total; idx += kThreads) {
            const int j2 = idx >> log_rows;
            fp2 x = ldv(src + (size_t)(j1_0 + r) + (size_t)M1 * j2);
            if (s) {
                x = mul2(x, cur);
                cur = mul2(cur, ratio);
            }
            lds[r * kRowStride + j2] = x;
        }
    }
    __syncthreads();
    lds_dft<kInLds, kOutLds>(lds, rows, log_rows, b, tw, GlobalIO{});
    // twiddle w_N^(j1 u), u = s + S q2; store Y[u][j1]
    {
        const int r = threadIdx.x & (rows - 1);
        const u64 j1 = (u64)(j1_0 + r);
        const u64 u0 = (u64)s + (u64)S * (threadIdx.x >> log_rows);
        fp2 cur = j1 ? wpow(tw, (j1 * u0) & nmask) : one2();
        const fp2 ratio = j1 ? wpow(tw, (j1 * (u64)S * step) & nmask) : one2();
        for (int idx = threadIdx.x; idx < total; idx += kThreads) {
            const int q2 = idx >> log_rows;
            const u64 u = (u64)s + (u64)S * q2;
            fp2 v = lds[r * kRowStride + q2];
            if (j1) {
                v = mul2(v, cur);
                cur = mul2(cur, ratio);
            }
            stv(dst + u * M1 + j1, v);
        }
    }
}

// Pass B.  grid = (S M2 / rows, columns); rows consecutive u per tile.
__global__ void __launch_bounds__(kThreads) k_ntt_pass_b(const V2* __restrict__ ybuf, size_t y_stride,
                                                         V2* __restrict__ out, size_t out_stride,
                                                         int a, int log_sm2, int log_rows, Tw tw)
{
    extern __shared__ fp2 lds[];
    const int M1 = 1 << a;
    const int rows = 1 << log_rows;
    const size_t SM2 = (size_t)1 << log_sm2;
    const size_t u0 = (size_t)blockIdx.x * rows;
    const V2* src = ybuf + blockIdx.y * y_stride;
    V2* dst = out + blockIdx.y * out_stride;
    if (a >= 3) {   // the DFT's first stage loads the rows, its last stores them
        GlobalIO io{};
        io.in = src + u0 * M1;
        io.out = dst + u0;
        io.out_step = SM2;
        lds_dft<kInRows, kOutCols>(lds, rows, log_rows, a, tw, io);
        return;
    }
    const int total = rows * M1;
    for (int idx = threadIdx.x; idx < total; idx += kThreads) {
        const int r = idx >> a, j1 = idx & (M1 - 1);
        lds[r * kRowStride + j1] = ldv(src + (u0 + r) * M1 + j1);
    }
    __syncthreads();
    lds_dft<kInLds, kOutLds>(lds, rows, log_rows, a, tw, GlobalIO{});
    for (int idx = threadIdx.x; idx < total; idx += kThreads) {
        const int r = idx & (rows - 1), q1 = idx >> log_rows;
        stv(dst + (u0 + r) + SM2 * q1, lds[r * kRowStride + q1]);
    }
}

// Twiddle tables: thread e computes its entries by square-and-multiply.
__global__ void k_ntt_tables(fp2 w1024, fp2 w, fp2 wh, int h, int log_hi, V2* t1024, V2* tl,
                             V2* th)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    auto pw = [](fp2 base, u64 ex) {
        fp2 r = one2();
        while (ex) {
            if (ex & 1) r = mul2(r, base);
            base = mul2(base, base);
            ex >>= 1;
        }
        return r;
    };
    if (e < 1024) stv(t1024 + e, pw(w1024, e));
    if (e < (1 << h)) stv(tl + e, pw(w, e));
    if (e < (1 << log_hi)) stv(th + e, pw(wh, e));
}

struct Plan {
    int log_order = -1;
    int h = 0;
    int w4_is_i = 0;
    int w8_s1 = 1, w8_s2 = 1;
    DevBuf t1024, tl, th, ybuf;
};
Plan g_plan;

fp2 host_pow2k(fp2 x, int k)   // x^(2^k)
{
    for (int i = 0; i < k; ++i) x = mul2(x, x);
    return x;
}

bool plan_for(int log_order, hipStream_t st)
{
    Plan& P_ = g_plan;
    if (P_.log_order == log_order) return true;
    const fp2 g = mk2(2147483648ULL, 1033321771269002680ULL);   // order 2^62
    const fp2 w = host_pow2k(g, 62 - log_order);
    const int h = (log_order + 1) / 2, log_hi = log_order - h;
    const fp2 wh = host_pow2k(w, h);
    const fp2 w1024 = host_pow2k(g, 62 - 10);
    const fp2 w4 = host_pow2k(g, 60);
    P_.w4_is_i = (w4.re == 0 && w4.im == 1) ? 1 : 0;
    if (!P_.w4_is_i && !(w4.re == 0 && w4.im == P - 1))
        return fail("ntt: fourth root of unity is not +-i");
    // the generator's 8th root is 2^30 (s1 + s2 i)
    const fp2 w8 = host_pow2k(g, 59);
    const u64 c = 1ULL << 30, mc = P - c;
    if (!((w8.re == c || w8.re == mc) && (w8.im == c || w8.im == mc)))
        return fail("ntt: eighth root of unity is not 2^30 (+-1 +-i)");
    P_.w8_s1 = w8.re == c ? 1 : -1;
    P_.w8_s2 = w8.im == c ? 1 : -1;
    if (!P_.t1024.ensure(1024 * sizeof(V2)) || !P_.tl.ensure(((size_t)1 << h) * sizeof(V2)) ||
        !P_.th.ensure(((size_t)1 << log_hi) * sizeof(V2)))
        return false;
    const int n = 1 << (h > 10 ? h : 10);
    k_ntt_tables<<<(n + 255) / 256, 256, 0, st>>>(w1024, w, wh, h, log_hi, P_.t1024.as<V2>(),
                                                  P_.tl.as<V2>(), P_.th.as<V2>());
    TP_HIP_OK(hipGetLastError());
    P_.log_order = log_order;
    P_.h = h;
    return true;
}

size_t lds_bytes(int rows) { return (size_t)rows * kRowStride * sizeof(fp2); }

}  // namespace

extern "C" {

// Reed-Solomon encode n_cols columns on the device.
//   d_coef : n_cols x 2^log_coef Fp2 (interleaved re, im), column-major
//            (column k starts at k * 2^log_coef).
//   d_out  : n_cols x 2^log_order Fp2, column k at k * 2^log_order;
//            out[k][i] = sum_j coef[k][j] w^(i j), w of order 2^log_order.
//   cols_per_chunk : columns per pass-A/pass-B pair (0 = 2); the intermediate
//            holds cols_per_chunk x 2^log_order Fp2.
// 0 <= log_coef <= 20, log_coef <= log_order <= 30.  Returns 0 / -1.
int tp_rs_encode(const uint64_t* d_coef, int n_cols, int log_coef, int log_order,
                 uint64_t* d_out, int cols_per_chunk, void* stream_handle)
{
    tp::clear_error();
    if (n_cols < 1 || log_coef < 0 || log_coef > 2 * kMaxLog || log_order < log_coef ||
        log_order > 30)
        return tp::fail("tp_rs_encode: unsupported shape (cols=%d log_coef=%d log_order=%d)",
                        n_cols, log_coef, log_order), -1;
    if (!tp::device_ready()) return -1;
    hipStream_t st = stream_handle ? (hipStream_t)stream_handle : tp::stream();
    // One plan (twiddle tables, intermediate) serves every call: calls are
    // serialized on the host, and a call on another stream than the previous
    // call's first waits for that stream, so no launch still in flight reads
    // tables or an intermediate that this call rewrites.
    static std::mutex mu;
    static hipStream_t last = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (last && last != st && hipStreamSynchronize(last) != hipSuccess)
        return tp::fail("tp_rs_encode: waiting for the previous call's stream failed"), -1;
    last = st;
    if (!plan_for(log_order, st)) return -1;
    Plan& P_ = g_plan;
    const int a = log_coef / 2, b = log_coef - a, log_s = log_order - log_coef;
    if (b > kMaxLog || a > kMaxLog)
        return tp::fail("tp_rs_encode: log_coef=%d too large", log_coef), -1;
    const int chunk = cols_per_chunk > 0 ? cols_per_chunk : 2;
    const size_t N = (size_t)1 << log_order, M = (size_t)1 << log_coef;
    if (!P_.ybuf.ensure((size_t)chunk * N * sizeof(V2))) return -1;
    static bool attr_done = false;
    if (!attr_done) {
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_ntt_pass_a,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds_bytes(kRows)));
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_ntt_pass_b,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds_bytes(kRows)));
        attr_done = true;
    }
    Tw tw;
    tw.t1024 = P_.t1024.as<V2>();
    tw.tl = P_.tl.as<V2>();
    tw.th = P_.th.as<V2>();
    tw.h = P_.h;
    tw.w4_is_i = P_.w4_is_i;
    tw.w8_s1 = P_.w8_s1;
    tw.w8_s2 = P_.w8_s2;
    const int M1 = 1 << a;
    const size_t SM2 = (size_t)1 << (log_s + b);
    const int rows_a = M1 < kRows ? M1 : kRows;
    const int rows_b = SM2 < (size_t)kRows ? (int)SM2 : kRows;
    const int lra = __builtin_ctz(rows_a), lrb = __builtin_ctz(rows_b);
    const V2* coef = reinterpret_cast<const V2*>(d_coef);
    V2* out = reinterpret_cast<V2*>(d_out);
    for (int c0 = 0; c0 < n_cols; c0 += chunk) {
        const int nc = (n_cols - c0) < chunk ? (n_cols - c0) : chunk;
        const unsigned tiles = (unsigned)(M1 / rows_a);
        const int xcd_order = tiles % 16 == 0;
        dim3 ga = xcd_order ? dim3(tiles << log_s, 1u, (unsigned)nc)
                            : dim3(tiles, 1u << log_s, (unsigned)nc);
        k_ntt_pass_a<<<ga, kThreads, lds_bytes(rows_a), st>>>(coef + (size_t)c0 * M, M,
                                                               P_.ybuf.as<V2>(), N, a, b, log_s,
                                                               lra, tw, xcd_order);
        TP_HIP_OK(hipGetLastError());
        dim3 gb((unsigned)(SM2 / rows_b), (unsigned)nc);
        k_ntt_pass_b<<<gb, kThreads, lds_bytes(rows_b), st>>>(P_.ybuf.as<V2>(), N,
                                                               out + (size_t)c0 * N, N, a,
                                                               log_s + b, lrb, tw);
        TP_HIP_OK(hipGetLastError());
    }
    return 0;
}

}  // extern "C"
