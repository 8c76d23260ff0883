// layer_bodies.hpp - device bodies of the GKR layer kernels.
//
// A body takes its workgroup's position (block index and block count, or the
// tile / point it covers) as arguments instead of reading blockIdx and
// gridDim.  The single-frame kernels pass their own launch position; the
// lockstep batch kernels (batch.hip) pass the position inside one frame's
// share of a launch that covers every frame of a batch.  Same body, same
// arithmetic, so a frame's tables and sums do not depend on which driver ran
// it.
#pragma once
#include "claim_terms.hpp"
#include "fp61.hpp"
#include "reduce.hpp"

namespace tp {

// Base-field gate evaluation (ArithCircuit semantics: a MUX gate yields in0
// when its mux bit is set and in1 otherwise, as CMuxGate.run does;
// TP/libTruePix/arithcircuit.py:20-109).
__device__ __forceinline__ u64 eval_gate(int op, u64 x, u64 y, int muxbit)
{
    switch (op) {
    case G_MUL: return fmul(x, y);
    case G_ADD: return fadd(x, y);
    case G_SUB: return fsub(x, y);
    case G_MUXL: return muxbit ? x : y;
    case G_OR: return fsub(fadd(x, y), fmul(x, y));
    case G_XOR: { const u64 xy = fmul(x, y); return fsub(fadd(x, y), fadd(xy, xy)); }
    case G_NOT: return fsub(1, x);
    case G_NAND: return fsub(1, fmul(x, y));
    case G_NOR: return fsub(fadd(1, fmul(x, y)), fadd(x, y));
    case G_NXOR: { const u64 xy = fmul(x, y); return fsub(fadd(1, fadd(xy, xy)), fadd(x, y)); }
    case G_NAAB: return fmul(fsub(1, x), y);
    default: return x;   // PASS
    }
}

__device__ __forceinline__ fp2 ld2(const u64* r, const u64* i, size_t k) { return mk2(r[k], i[k]); }

// Loads the 2^nb table into T and folds it with cs[0..nf), ping-ponging
// between T and U; returns the buffer holding the result.
__device__ inline fp2* lds_fold_chain(fp2* T, fp2* U, const u64* vr, const u64* vi, int nb,
                                      const fp2* cs, int nf)
{
    const int W = 1 << nb;
    for (int w = threadIdx.x; w < W; w += blockDim.x) T[w] = ld2(vr, vi, w);
    __syncthreads();
    fp2* src = T;
    fp2* dst = U;
    for (int j = 0; j < nf; ++j) {
        const fp2 r = cs[j];
        const int n_out = W >> (j + 1);
        for (int i = threadIdx.x; i < n_out; i += blockDim.x)
            dst[i] = fold2(src[2 * i], src[2 * i + 1], r);
        __syncthreads();
        fp2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// eq table in LDS by doubling: tab[t] = prod_j (bit_j(t) ? z_j : 1 - z_j).
__device__ inline void lds_eq(fp2* tab, const fp2* z, int nb)
{
    if (threadIdx.x == 0) tab[0] = one2();
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const int cur = 1 << j;
        const fp2 zz = z[j];
        const fp2 omz = sub2(one2(), zz);
        for (int t = threadIdx.x; t < cur; t += blockDim.x) {
            const fp2 b = tab[t];
            tab[t + cur] = mul2(b, zz);
            tab[t] = mul2(b, omz);
        }
        __syncthreads();
    }
}

// Workgroup total of v written to partials[b * N + k].
template <int N>
__device__ __forceinline__ void block_partial_at(fp2 (&v)[N], fp2* partials, size_t b)
{
    block_sum<N>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) partials[b * N + k] = v[k];
    }
}

namespace body {

// Input rows [cp][2^row_log] of a circuit's input layer: private inputs where
// slot[w] >= 0, the constant cval[w] elsewhere, zero rows for padded copies
// (TP/libTruePix/util.py get_inputs + layerprover.py:113-120 padding).
__device__ __forceinline__ void input_rows(const uint8_t* src8, const u64* src64, int n_priv,
                                           int n_copies, int cp, int row_log, const int* slot,
                                           const u64* cval, u64* rows, size_t bid, size_t nb)
{
    const size_t total = (size_t)cp << row_log;
    const int mask = (1 << row_log) - 1;
    for (size_t k = bid * blockDim.x + threadIdx.x; k < total; k += nb * blockDim.x) {
        const size_t c = k >> row_log;
        const int w = (int)(k & mask);
        u64 v = 0;
        if (c < (size_t)n_copies) {
            const int s = slot[w];
            if (s >= 0) {
                const size_t at = c * n_priv + s;
                v = src8 ? (u64)src8[at] : fred64(src64[at]);
            } else {
                v = cval[w];
            }
        }
        rows[k] = v;
    }
}

// A3: out[c][g] = gate_g(in[c][in0_g], in[c][in1_g]) for c < n_copies and
// g < n_gates; zero for g >= n_gates and for padded copies < n_copies_pad.
__device__ __forceinline__ void eval_layer(const u64* in, int in_stride, u64* out, int out_log,
                                           int n_copies, int n_copies_pad, const int* op,
                                           const int* in0, const int* in1, const int* mux,
                                           int n_gates, size_t bid, size_t nb)
{
    const size_t total = (size_t)n_copies_pad << out_log;
    const int out_stride = 1 << out_log;
    for (size_t k = bid * blockDim.x + threadIdx.x; k < total; k += nb * blockDim.x) {
        const size_t c = k >> out_log;
        const int g = (int)(k & (out_stride - 1));
        u64 v = 0;
        if (g < n_gates && c < (size_t)n_copies) {
            const u64* row = in + c * in_stride;
            v = eval_gate(op[g], row[in0[g]], row[in1[g]], mux[g]);
        }
        out[k] = v;
    }
}

// A4: Fp2 SoA wire table [wire][copy] from base-field rows [copy][row_stride]
// through a 64 x 64 LDS tile (reads coalesced along wires, writes along
// copies); tile (bx, by) covers copies 64 bx.. and wires 64 by..; 256 threads.
__device__ __forceinline__ void wire_table(const u64* rows, int row_stride, int n_copies,
                                           int n_copies_pad, int n_wires, u64* tre, u64* tim,
                                           int bx, int by)
{
    __shared__ u64 tile[64][65];
    const int c0 = bx * 64, w0 = by * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4 threads
    for (int i = ty; i < 64; i += 4) {
        const int c = c0 + i, w = w0 + tx;
        tile[i][tx] = (c < n_copies && w < n_wires) ? rows[(size_t)c * row_stride + w] : 0;
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int w = w0 + i, c = c0 + tx;
        if (w < n_wires && c < n_copies_pad) {
            tre[(size_t)w * n_copies_pad + c] = tile[tx][i];
            tim[(size_t)w * n_copies_pad + c] = 0;
        }
    }
}

// eq table: out[k] = prod_j (bit_j(k) ? z_j : 1 - z_j) for k < n_out
// (TP/native_fp2/native_fp2_core.hpp:139-159, beta with init 1); entries
// bid * blockDim + tid.
__device__ __forceinline__ void eq_table(const fp2* z, int n, u64* outr, u64* outi, int n_out,
                                         size_t bid)
{
    const size_t k = bid * blockDim.x + threadIdx.x;
    if (k >= (size_t)n_out) return;
    fp2 b = one2();
    for (int j = 0; j < n; ++j) b = mul2(b, ((k >> j) & 1) ? z[j] : sub2(one2(), z[j]));
    outr[k] = b.re;
    outi[k] = b.im;
}

// A6: early-round claim over a term table (claim_terms.hpp): a term is a
// gate that is not affine in its inputs (type 0, 5, 6, 8-11, coefficient its
// z1 value), an input wire (kTermWire, coefficient the z1-weighted sum of the
// affine gates that read it, so the claim's affine part is one product per
// wire instead of four per gate) or the affine gates' constant (kTermConst).
// The affine part L(t) is linear in t, so only t = 0, 1 are formed and
// L(-1) = 2 L(0) - L(1), L(2) = 2 L(1) - L(0).
// Work items (pair p, term chunk q) with p fastest: item = bid * blockDim +
// tid < n_pairs * n_chunks, chunk q covering terms [q gpt, min(n_terms,
// (q + 1) gpt)); small rounds (few pairs) thus still fill whole workgroups.
// The workgroup total of (F(0), F(1), F(-1), F(2)) goes to partials[pidx]
// (TP/native_fp2/native_fp2.cu:232-281 semantics, SoA [wire][copy] tables;
// nc a power of two).
__device__ __forceinline__ void early_claim(int n_terms, int gpt, int n_chunks, const int* gt,
                                            const int* in0, const int* in1, const u64* tzr,
                                            const u64* tzi, int nc, const u64* vr, const u64* vi,
                                            const u64* br, const u64* bi, fp2* partials,
                                            size_t bid, size_t pidx)
{
    const int n_pairs = nc >> 1;
    const int lp = __ffs(n_pairs) - 1;
    const size_t item = bid * blockDim.x + threadIdx.x;
    fp2 acc[4] = {zero2(), zero2(), zero2(), zero2()};
    if (item < ((size_t)n_chunks << lp)) {
        const int p = (int)(item & (size_t)(n_pairs - 1));
        const int g0 = (int)(item >> lp) * gpt;
        const int g1 = min(n_terms, g0 + gpt);
        const size_t c = 2 * (size_t)p;
        fp2 v0 = zero2(), v1 = zero2(), v2 = zero2(), v3 = zero2();
        fp2 l0 = zero2(), l1 = zero2();
#pragma unroll 2
        for (int g = g0; g < g1; ++g) {
            const int tpg = gt[g];
            const fp2 z = mk2(tzr[g], tzi[g]);
            const size_t a = (size_t)in0[g] * nc + c;
            if (tpg == kTermWire) {
                l0 = add2(l0, mul2(mk2(vr[a], vi[a]), z));
                l1 = add2(l1, mul2(mk2(vr[a + 1], vi[a + 1]), z));
            } else if (tpg == kTermConst) {
                l0 = add2(l0, z);
                l1 = add2(l1, z);
            } else {
                const size_t b = (size_t)in1[g] * nc + c;
                const fp2 a0 = mk2(vr[a], vi[a]), a1 = mk2(vr[a + 1], vi[a + 1]);
                const fp2 b0 = mk2(vr[b], vi[b]), b1 = mk2(vr[b + 1], vi[b + 1]);
                v0 = add2(v0, mul2(gatefn(tpg, a0, b0), z));
                v1 = add2(v1, mul2(gatefn(tpg, a1, b1), z));
                v2 = add2(v2, mul2(gatefn(tpg, sub2(dbl2(a0), a1), sub2(dbl2(b0), b1)), z));
                v3 = add2(v3, mul2(gatefn(tpg, sub2(dbl2(a1), a0), sub2(dbl2(b1), b0)), z));
            }
        }
        v0 = add2(v0, l0);
        v1 = add2(v1, l1);
        v2 = add2(v2, sub2(dbl2(l0), l1));
        v3 = add2(v3, sub2(dbl2(l1), l0));
        const fp2 be0 = mk2(br[c], bi[c]), be1 = mk2(br[c + 1], bi[c + 1]);
        acc[0] = mul2(v0, be0);
        acc[1] = mul2(v1, be1);
        acc[2] = mul2(v2, sub2(dbl2(be0), be1));
        acc[3] = mul2(v3, sub2(dbl2(be1), be0));
    }
    block_partial_at<4>(acc, partials, pidx);
}

// The claim terms' coefficients from the gates' z1 values: term k's is the
// signed sum of z1 over its contributions [off[k], off[k + 1]) (gate src[e],
// subtracted when neg[e]).  One wave per term (a wire read by many affine
// gates has many contributions: the gray circuit's constant wire 128):
// wave k = (bid * blockDim + tid) / 64, its lanes striding the contributions.
__device__ __forceinline__ void term_coeffs(const u64* z1r, const u64* z1i, const int* off,
                                            const int* src, const int* neg, int n_terms,
                                            u64* tzr, u64* tzi, size_t bid)
{
    const size_t k = (bid * blockDim.x + threadIdx.x) >> 6;   // the same for the whole wave
    const int lane = threadIdx.x & 63;
    if (k >= (size_t)n_terms) return;
    fp2 acc[1] = {zero2()};
    for (int e = off[k] + lane; e < off[k + 1]; e += 64) {
        const fp2 z = mk2(z1r[src[e]], z1i[src[e]]);
        acc[0] = neg[e] ? sub2(acc[0], z) : add2(acc[0], z);
    }
    wave_sum<1>(acc);
    if (lane == 0) {
        tzr[k] = acc[0].re;
        tzi[k] = acc[0].im;
    }
}

// A7: fold every wire row (n_wires rows of 2^(log_out + 1) entries) and the
// beta row with r (TP/native_fp2/native_fp2_core.hpp:250-279).
__device__ __forceinline__ void fold_rows(const u64* ir, const u64* ii, int n_wires, int log_out,
                                          fp2 r, u64* orr, u64* oi, const u64* bir, const u64* bii,
                                          u64* bor, u64* boi, size_t bid, size_t nb)
{
    const size_t n_out = (size_t)1 << log_out;
    const size_t total = ((size_t)n_wires + 1) << log_out;
    for (size_t k = bid * blockDim.x + threadIdx.x; k < total; k += nb * blockDim.x) {
        const size_t row = k >> log_out, i = k & (n_out - 1);
        if (row < (size_t)n_wires) {
            const size_t src = row * 2 * n_out + 2 * i;
            const fp2 v = fold2(mk2(ir[src], ii[src]), mk2(ir[src + 1], ii[src + 1]), r);
            orr[k] = v.re;
            oi[k] = v.im;
        } else {
            const fp2 v = fold2(mk2(bir[2 * i], bii[2 * i]), mk2(bir[2 * i + 1], bii[2 * i + 1]), r);
            bor[i] = v.re;
            boi[i] = v.im;
        }
    }
}

// Start of the late rounds (TP/libTruePix/gateprover.py:51-62): each gate's
// fixed side is V[in1]; its coefficient starts at the z1 eq value; beta* is
// packed (re, im) into beta2 by the workgroup with bid 0.
__device__ __forceinline__ void late_init(const u64* vr, const u64* vi, const int* in1,
                                          int n_gates, const u64* z1cr, const u64* z1ci,
                                          u64* z1lr, u64* z1li, u64* othr, u64* othi,
                                          const u64* br, const u64* bi, u64* beta2, size_t bid)
{
    const size_t g = bid * blockDim.x + threadIdx.x;
    if (g == 0) {
        beta2[0] = br[0];
        beta2[1] = bi[0];
    }
    if (g >= (size_t)n_gates) return;
    othr[g] = vr[in1[g]];
    othi[g] = vi[in1[g]];
    z1lr[g] = z1cr[g];
    z1li[g] = z1ci[g];
}

// A9, one late round in one workgroup.  Round rd < nin binds bit rd of in0
// (w1), round nin + k binds bit k of in1 (w2).  For a gate with e = the bound
// bit of its input w and T the table folded with this pass's earlier
// challenges (kk of them):
//   t = e  : z1 * gate(cur, other)          cur = T[w >> kk]
//   t = -1 : z1 * (e ? -1 : 2) * gate(m1, other),  m1 = 2 T[2j] - T[2j+1],
//            j = w >> (kk + 1)
// (argument order swapped in the w2 pass), summed over gates and multiplied
// by beta* (TP/libTruePix/gateprover.py:102-132, layerprover.py:225-244).
// Before that, the previous round's challenge updates each gate's
// coefficient: z1 *= (bit ? r : 1 - r) (gateprover.py:135-176).  chal holds
// the 2 nin challenges (w1 then w2) drawn so far; lds has fold_lds(nin)
// bytes.  Thread 0 returns (f(0), f(1), f(-1)) * beta in out.
__device__ __forceinline__ void late_round(const int* gt, const int* in0, const int* in1, u64* z1r,
                                           u64* z1i, const u64* oth_r, const u64* oth_i,
                                           const u64* vr, const u64* vi, const fp2* chal, int nin,
                                           int n_gates, int rd, int apply, const u64* beta,
                                           fp2* lds, fp2 (&out)[3])
{
    fp2* T = lds;
    fp2* U = lds + (1 << nin);
    if (apply) {
        const int k = rd - 1;
        const fp2 r = chal[k];
        const fp2 omr = sub2(one2(), r);
        const bool w1pass = k < nin;
        const int bit = w1pass ? k : k - nin;
        for (int g = threadIdx.x; g < n_gates; g += blockDim.x) {
            const int w = w1pass ? in0[g] : in1[g];
            const fp2 nz = mul2(mk2(z1r[g], z1i[g]), ((w >> bit) & 1) ? r : omr);
            z1r[g] = nz.re;
            z1i[g] = nz.im;
        }
    }
    const bool w1pass = rd < nin;
    const int kk = w1pass ? rd : rd - nin;
    const fp2* Tk = lds_fold_chain(T, U, vr, vi, nin, chal + (w1pass ? 0 : nin), kk);
    fp2 acc[3] = {zero2(), zero2(), zero2()};
    const fp2 m_one = mk2(P - 1, 0), two = mk2(2, 0);
    for (int g = threadIdx.x; g < n_gates; g += blockDim.x) {
        const int w = w1pass ? in0[g] : in1[g];
        const int e = (w >> kk) & 1;
        const fp2 z = mk2(z1r[g], z1i[g]);
        const fp2 other = mk2(oth_r[g], oth_i[g]);
        const fp2 cur = Tk[w >> kk];
        const int j = w >> (kk + 1);
        const fp2 m1 = sub2(dbl2(Tk[2 * j]), Tk[2 * j + 1]);
        const int t = gt[g];
        const fp2 v01 = mul2(z, w1pass ? gatefn(t, cur, other) : gatefn(t, other, cur));
        const fp2 vm1 = mul2(mul2(z, e ? m_one : two),
                             w1pass ? gatefn(t, m1, other) : gatefn(t, other, m1));
        if (e) acc[1] = add2(acc[1], v01);
        else acc[0] = add2(acc[0], v01);
        acc[2] = add2(acc[2], vm1);
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        const fp2 b = mk2(beta[0], beta[1]);
        for (int t = 0; t < 3; ++t) out[t] = mul2(acc[t], b);
    }
}

// Start of the w2 pass: every gate's fixed side becomes V(w1) (the table
// folded with all nin w1 challenges; gateprover.py:150-158).
__device__ __forceinline__ void late_switch(const u64* vr, const u64* vi, int nin, const fp2* w1,
                                            int n_gates, u64* oth_r, u64* oth_i, fp2* lds)
{
    const fp2* R = lds_fold_chain(lds, lds + (1 << nin), vr, vi, nin, w1, nin);
    const fp2 vw1 = R[0];
    for (int g = threadIdx.x; g < n_gates; g += blockDim.x) {
        oth_r[g] = vw1.re;
        oth_i[g] = vw1.im;
    }
}

// A10 on the line of the late rounds: V at w1 + t (w2 - w1)
// (TP/libTruePix/layercompute.py:169-243); the point is formed in LDS past
// the two fold buffers; lds has fold_lds(nb) + nb Fp2.  Thread 0 returns it.
__device__ __forceinline__ fp2 line_value(const u64* vr, const u64* vi, int nb, const fp2* w12,
                                          int t, fp2* lds)
{
    fp2* pt = lds + ((3 << nb) >> 1);   // past the fold buffers (fold_lds)
    const fp2 tt = mk2((u64)t, 0);
    for (int j = threadIdx.x; j < nb; j += blockDim.x)
        pt[j] = add2(w12[j], mul2(tt, sub2(w12[nb + j], w12[j])));
    __syncthreads();
    const fp2* R = lds_fold_chain(lds, lds + (1 << nb), vr, vi, nb, pt, nb);
    return R[0];
}

// A11: per gate type, sum_g beta_z1[g] * beta_w1[in0_g] * beta_w2[in1_g]
// (with z1b non-null: m0 beta_z1[g] + m1 beta_z1b[g] in place of
// beta_z1[g]); eq tables built in LDS by doubling; one workgroup; thread 0
// writes out13 (TP/libTruePix/circuitverifier.py:224-290, eval_mlext).
__device__ __forceinline__ void wiring_mlext(const int* type_idx, const int* in0, const int* in1,
                                             int n_gates, int n_out_bits, int n_in_bits,
                                             const fp2* z1, const fp2* z1b, const fp2* w1,
                                             const fp2* w2, fp2 m0, fp2 m1, fp2* sm, fp2* out13)
{
    const int WO = 1 << n_out_bits, WI = 1 << n_in_bits;
    const bool two = z1b != nullptr;
    fp2* bz = sm;
    fp2* bz2 = bz + WO;
    fp2* b1 = bz2 + (two ? WO : 0);
    fp2* b2 = b1 + WI;
    lds_eq(bz, z1, n_out_bits);
    if (two) lds_eq(bz2, z1b, n_out_bits);
    lds_eq(b1, w1, n_in_bits);
    lds_eq(b2, w2, n_in_bits);
    fp2 acc[13];
#pragma unroll
    for (int t = 0; t < 13; ++t) acc[t] = zero2();
    for (int g = threadIdx.x; g < n_gates; g += blockDim.x) {
        fp2 tv = two ? add2(mul2(m0, bz[g]), mul2(m1, bz2[g])) : bz[g];
        tv = mul2(tv, mul2(b1[in0[g]], b2[in1[g]]));
        const int t = type_idx[g];
#pragma unroll
        for (int q = 0; q < 13; ++q)
            if (q == t) acc[q] = add2(acc[q], tv);
    }
    block_sum<13>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int t = 0; t < 13; ++t) out13[t] = acc[t];
    }
}

// A12 on a device-resident base-field vector of 2^n_z entries: workgroup bid
// of nb takes rows bid, bid + nb, ... of 2^lb inputs sharing an LDS low-half
// eq table; its total goes to partials[pidx].
__device__ __forceinline__ void mle_base(const u64* in, const fp2* z, int n_z, fp2* partials,
                                         size_t bid, size_t nb, size_t pidx)
{
    __shared__ fp2 lo[1024];
    __shared__ fp2 hrow;
    const int lb = n_z < 10 ? n_z : 10;
    const int row_len = 1 << lb;
    const size_t n_rows = (size_t)1 << (n_z - lb);
    lds_eq(lo, z, lb);
    fp2 acc[1] = {zero2()};
    for (size_t row = bid; row < n_rows; row += nb) {
        if (threadIdx.x == 0) {
            fp2 h = one2();
            for (int j = lb; j < n_z; ++j) {
                const fp2 zz = z[j];
                h = mul2(h, ((row >> (j - lb)) & 1) ? zz : sub2(one2(), zz));
            }
            hrow = h;
        }
        __syncthreads();
        fp2 dot = zero2();
        const u64* src = in + row * row_len;
        for (int t = threadIdx.x; t < row_len; t += blockDim.x)
            dot = add2(dot, smul2(fred64(src[t]), lo[t]));
        acc[0] = add2(acc[0], mul2(dot, hrow));
        __syncthreads();
    }
    block_partial_at<1>(acc, partials, pidx);
}

}  // namespace body
}  // namespace tp
