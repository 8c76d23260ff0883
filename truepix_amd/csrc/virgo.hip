// virgo.hip - the Virgo code-switching proof of a NIZK Orion opening on the
// device (the interface and its reference citations are in virgo.hpp).
//
// The circuit.  generate_circuit (OZ/src/VPD/linearPC.cpp:580-669) builds,
// over the message of length n: layer 0 the message, layer 1 direct relays
// of it, then one layer per expander step, each relaying everything so far
// and appending one custom linear combination per right vertex of C[d] (on
// the way down) or D[d] (on the way up), then one relay per queried position.
// Each layer's width is 2^bits, bits from prepare_gates_count's arithmetic
// (linearPC.cpp:477-512), kept as it is (it overcounts the last D layer).
// Because every layer relays all earlier values and the encoder writes each
// step's output right after its input, layer j's values are the first len_j
// entries of the codeword, zero beyond: the codeword already holds the
// evaluation of every layer.
//
// The algebra.  The encoding circuit has only relays (type 10), custom
// combinations (type 14) and direct relays (type 4).  For those the
// reference prover's phase 1 (prover.cpp:246-575) leaves addV_array zero
// and puts into add_mult_sum[u] the sum of beta(g) = alpha eq(r_0, g) +
// beta eq(r_1, g) over the gates reading u (times the weight for a
// combination): phase 1 is the quadratic sumcheck of A(u) V(u) for that
// table A.  Its phase 2 (prover.cpp:634-975) puts everything at v = 0, as
// S = v_u A~(r_u) (zero for direct relays), so its round polynomials are
// (0, -S_j, S_j) with S_{j+1} = S_j (1 - r_v[j]); v_v is V~(r_v).  The
// Fiat-Shamir challenges do not depend on the prover's messages (they are
// squeezed without absorbing, verifier.cpp:936-940), so every layer is
// proved at once.
//
// The launches.
//   k_vg_pre (256 threads per workgroup, one job each):
//     - A jobs: 256 entries of a layer's A table (a gather over the forward
//       graph, the beta weights from half-size eq tables in LDS) beside the
//       same entries of V; the first (up to) eight rounds of that layer's
//       phase 1 run on them at once, six inside each wave (partner values
//       by lane shuffles), two more over the four wave results; the round
//       sums and the folded pair go out;
//     - check jobs: the verifier's predicates (verifier.cpp:349-753) as flat
//       sums over relays, graph edges or queried positions;
//     - dot jobs: V~(r_v) of every layer, the output claim V~_top(r_0) over
//       the queried entries, the input layer's value at the last r_u;
//   k_vg_layer (1024 threads): one workgroup per layer adds the round sums
//     of its A jobs and runs the remaining rounds on their folded pairs in
//     LDS, then the phase-2 polynomials; further workgroups reduce the
//     check and dot partial sums.
// The host then walks the verifier's claim chain.
#include <algorithm>
#include <cstring>
#include <map>

#include "reduce.hpp"
#include "runtime.hpp"
#include "virgo.hpp"

namespace tp {
namespace virgo {
namespace {

constexpr int kSmall = 13;          // distance_threshold (OZ/include/linear_code/parameter.h)
constexpr int kMaxBits = 20;        // widest layer handled: 2^20 gates (a message of 2^18, 65536 copies)
constexpr int kKindDirect = 0, kKindCode = 1, kKindQuery = 2;
constexpr int kRoleA = 0, kRoleCheck = 1, kRoleDot = 2;
constexpr int kDotV = 0, kDotOut = 1, kDotIn = 2;
constexpr int kPreThreads = 256;    // k_vg_pre
constexpr int kThreads = 1024;      // k_vg_layer
constexpr int kPerThread = 4;       // entries / items per thread and step
constexpr int kJobRounds = 10;      // phase-1 rounds inside an A job (1024 entries)
constexpr int kThreadRounds = 2;    // of them inside each thread
constexpr int kWaveRounds = 6;      // then across each wave's lanes
constexpr int kCheckItems = 2048, kDotItems = 2048;
constexpr int kCheckRelays = 0, kCheckEdges = 1, kCheckQueries = 2;

struct VL {                 // sumcheck layer i: the gates of layer i over layer i-1
    int bits_g, bits_u, kind, ch;   // ch: r_0, r_1 (bits_g each), r_u, r_v (bits_u each), alpha, beta
    long long P;            // gates 0 .. P-1 relay u = g
    long long first, R;     // custom combinations first .. first+R-1
    long long src_off, L;   // their graph's left vertices are u = src_off .. src_off+L-1
    long long vlen;         // V_{i-1}: the first vlen codeword entries, zero beyond
    int deg, res;           // graph degree; results offset (Fp2)
    const int* to;
    const fp2* w;
    int r1, nwa;            // rounds inside each A job, A jobs
    long long pf_off;       // A jobs' round sums: nwa x r1 x 3 Fp2
    long long f_off;        // A jobs' folded pairs: nwa x (A, V)
};

struct Job {                // one workgroup of k_vg_pre
    int layer, role;
    long long begin, end;   // entries, items or positions
    int slot, aux;          // A job index in its layer / partial slot; dot kind
};

struct Group {              // one reduction workgroup of k_vg_layer
    int slot0, count, res, pad;
};

// t[x] = scale * prod_k (bit k of x ? r[k] : 1 - r[k]), x < 2^nb: one half of
// an eq table (the reference's first-half / second-half beta tables).
__device__ void eq_half(fp2* t, const fp2* r, int nb, fp2 scale)
{
    for (int x = threadIdx.x; x < (1 << nb); x += blockDim.x) {
        fp2 v = scale;
        for (int k = 0; k < nb; ++k) {
            const fp2 rk = r[k];
            v = mul2(v, ((x >> k) & 1) ? rk : sub2(one2(), rk));
        }
        t[x] = v;
    }
}

struct Eq {                 // eq(r, x) = lo[x mod 2^h] hi[x >> h]
    const fp2 *lo, *hi;
    int h;
    __device__ fp2 at(long long x) const { return mul2(lo[x & ((1LL << h) - 1)], hi[x >> h]); }
};

struct Beta {               // alpha eq(r_0, g) + beta eq(r_1, g)
    Eq e0, e1;
    __device__ fp2 at(long long g) const
    {
        const long long m = (1LL << e0.h) - 1;
        LazyProdSum s;
        s.zero();
        s.mac(e0.lo[g & m], e0.hi[g >> e0.h]);
        s.mac(e1.lo[g & m], e1.hi[g >> e0.h]);
        return s.value();
    }
};

// Builds eq(r, .) over nb bits into lds (2^(nb/2) + 2^(nb - nb/2) entries).
__device__ Eq make_eq(fp2* lds, int& used, const fp2* r, int nb, fp2 scale)
{
    Eq e;
    e.h = nb / 2;
    fp2* lo = lds + used;
    used += 1 << e.h;
    fp2* hi = lds + used;
    used += 1 << (nb - e.h);
    eq_half(lo, r, e.h, scale);
    eq_half(hi, r + e.h, nb - e.h, one2());
    e.lo = lo;
    e.hi = hi;
    return e;
}

__device__ Beta make_beta(fp2* lds, int& used, const VL& l, const fp2* ch)
{
    Beta b;
    const fp2* r0 = ch + l.ch;
    const fp2* r1 = r0 + l.bits_g;
    const fp2 alpha = ch[l.ch + 2 * l.bits_g + 2 * l.bits_u];
    const fp2 beta = ch[l.ch + 2 * l.bits_g + 2 * l.bits_u + 1];
    b.e0 = make_eq(lds, used, r0, l.bits_g, alpha);
    b.e1 = make_eq(lds, used, r1, l.bits_g, beta);
    return b;
}

__device__ __forceinline__ fp2 shfl_xor2(fp2 x, int s)
{
    return mk2(__shfl_xor(x.re, s, 64), __shfl_xor(x.im, s, 64));
}

// Phase-1 rounds on the entries held by the lanes of one wave (entry k in
// lane k, A in a, V in v): round j pairs lanes 2^(j+1) m and 2^(j+1) m + 2^j
// (entries 2i, 2i+1 of the table folded by the earlier rounds), adds
// (dA dV, dA V0 + A0 dV, A0 V0) over the pairs and folds by ru[j] into the
// lower lane.  Lane 0 stores round j's sums at sums + 3 j.
__device__ void wave_rounds(fp2& a, fp2& v, const fp2* ru, int n, fp2* sums)
{
    const int lane = threadIdx.x & 63;
    for (int j = 0; j < n; ++j) {
        const int s = 1 << j;
        const fp2 a1 = shfl_xor2(a, s), v1 = shfl_xor2(v, s);
        LazyProdSum x, y, z;
        x.zero();
        y.zero();
        z.zero();
        if ((lane & (2 * s - 1)) == 0) {
            const fp2 da = sub2(a1, a), dv = sub2(v1, v);
            x.mac(da, dv);
            y.mac(da, v);
            y.mac(a, dv);
            z.mac(a, v);
            const fp2 r = ru[j];
            a = fold2(a, a1, r);
            v = fold2(v, v1, r);
        }
        fp2 pr[3] = {x.value(), y.value(), z.value()};
        wave_sum<3>(pr);
        if (lane == 0) {
            sums[3 * j] = pr[0];
            sums[3 * j + 1] = pr[1];
            sums[3 * j + 2] = pr[2];
        }
    }
}

// Entry u of the A table of a code layer: beta(u) if gate u relays it, plus
// w beta(first + t) over its edges to combinations t (degree-major edge
// lists: a wave's lanes read consecutive words; five edges' loads at a time).
__device__ fp2 a_code(const VL& l, const Beta& b, long long u)
{
    fp2 a = u < l.P ? b.at(u) : zero2();
    if (u >= l.src_off && u < l.src_off + l.L) {
        const long long i = u - l.src_off;
        LazyProdSum s;
        s.zero();
        for (int d0 = 0; d0 < l.deg; d0 += 5) {
            int to[5];
            fp2 w[5];
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (d0 + k < l.deg) {
                    to[k] = l.to[(d0 + k) * l.L + i];
                    w[k] = l.w[(d0 + k) * l.L + i];
                }
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (d0 + k < l.deg) s.mac(w[k], b.at(l.first + to[k]));
        }
        a = add2(a, s.value());
    }
    return a;
}

__global__ void __launch_bounds__(kPreThreads) k_vg_pre(const VL* __restrict__ layers,
                                                     const Job* __restrict__ jobs,
                                                     const fp2* __restrict__ ch,
                                                     const fp2* __restrict__ cc,
                                                     const long long* __restrict__ q, int nq,
                                                     fp2* __restrict__ pf, fp2* __restrict__ F,
                                                     fp2* __restrict__ part)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 sw[kPreThreads / 64][3 * (kThreadRounds + kWaveRounds)];   // the waves' round sums
    __shared__ fp2 sf[kPreThreads / 64][2];                  // the waves' folded pairs
    const Job jb = jobs[blockIdx.x];
    const VL& l = layers[jb.layer];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    int used = 0;
    if (jb.role == kRoleA) {
        const Beta b = make_beta(lds, used, l, ch);
        const long long u0 = jb.begin, B = jb.end - jb.begin;
        fp2* sa = lds + used;   // the queried positions' A entries in this range
        if (l.kind == kKindQuery) {
            for (int k = t; k < B; k += blockDim.x) sa[k] = zero2();
        }
        __syncthreads();
        if (l.kind == kKindQuery) {
            // sorted positions: the first of each run adds the run's betas
            for (int k = t; k < nq; k += blockDim.x) {
                const long long x = q[k];
                if (x < u0 || x >= u0 + B || (k > 0 && q[k - 1] == x)) continue;
                fp2 s = zero2();
                for (int kk = k; kk < nq && q[kk] == x; ++kk) s = add2(s, b.at(kk));
                sa[x - u0] = s;
            }
            __syncthreads();
        }
        // entries 4t .. 4t+3 of the job in this thread
        fp2 a[kPerThread], v[kPerThread];
#pragma unroll
        for (int m = 0; m < kPerThread; ++m) {
            const long long k = (long long)kPerThread * t + m, u = u0 + k;
            a[m] = v[m] = zero2();
            if (k < B) {
                if (u < l.vlen) v[m] = cc[u];
                if (l.kind == kKindDirect) a[m] = b.at(u);
                else if (l.kind == kKindCode) a[m] = a_code(l, b, u);
                else a[m] = sa[k];
            }
        }
        const fp2* ru = ch + l.ch + 2 * l.bits_g;
        const int r1 = l.r1;
        // rounds 0 and 1 inside the thread
        const int nt = r1 < kThreadRounds ? r1 : kThreadRounds;
        for (int j = 0; j < nt; ++j) {
            const int s = 1 << j;
            LazyProdSum x, y, z;
            x.zero();
            y.zero();
            z.zero();
            const fp2 r = ru[j];
#pragma unroll
            for (int m = 0; m < kPerThread; m += 2) {
                if (m % (2 * s)) continue;
                const fp2 da = sub2(a[m + s], a[m]), dv = sub2(v[m + s], v[m]);
                x.mac(da, dv);
                y.mac(da, v[m]);
                y.mac(a[m], dv);
                z.mac(a[m], v[m]);
                a[m] = fold2(a[m], a[m + s], r);
                v[m] = fold2(v[m], v[m + s], r);
            }
            fp2 pr[3] = {x.value(), y.value(), z.value()};
            wave_sum<3>(pr);
            if (lane == 0) {
                sw[wid][3 * j] = pr[0];
                sw[wid][3 * j + 1] = pr[1];
                sw[wid][3 * j + 2] = pr[2];
            }
        }
        // rounds 2 .. 7 across each wave's lanes
        const int nwr = r1 - nt < kWaveRounds ? r1 - nt : kWaveRounds;
        wave_rounds(a[0], v[0], ru + nt, nwr, sw[wid] + 3 * nt);
        if (lane == 0) {
            sf[wid][0] = a[0];
            sf[wid][1] = v[0];
        }
        __syncthreads();
        if (wid == 0) {
            const int nw = (int)((B + 64 * kPerThread - 1) / (64 * kPerThread));
            const int nr = nt + nwr;
            fp2* out = pf + l.pf_off + (long long)jb.slot * r1 * 3;
            for (int j = 0; j < nr; ++j) {
                fp2 acc[3] = {zero2(), zero2(), zero2()};
                if (lane < nw) {
                    acc[0] = sw[lane][3 * j];
                    acc[1] = sw[lane][3 * j + 1];
                    acc[2] = sw[lane][3 * j + 2];
                }
                wave_sum<3>(acc);
                if (lane == 0) {
                    out[3 * j] = acc[0];
                    out[3 * j + 1] = acc[1];
                    out[3 * j + 2] = acc[2];
                }
            }
            // rounds 8 and 9 over the waves' folded pairs, in wave 0's lanes
            fp2 fa_ = lane < nw ? sf[lane][0] : zero2();
            fp2 fv_ = lane < nw ? sf[lane][1] : zero2();
            if (r1 > nr) wave_rounds(fa_, fv_, ru + nr, r1 - nr, out + 3 * nr);
            if (lane == 0) {
                F[l.f_off + 2LL * jb.slot] = fa_;
                F[l.f_off + 2LL * jb.slot + 1] = fv_;
            }
        }
        return;
    }
    fp2 acc[2] = {zero2(), zero2()};
    if (jb.role == kRoleCheck) {
        // relays: beta(g) eq(r_u, u_g); edges: beta(first + to) w eq(r_u, src);
        // four items per thread and step, their loads first
        const Beta b = make_beta(lds, used, l, ch);
        const Eq eu = make_eq(lds, used, ch + l.ch + 2 * l.bits_g, l.bits_u, one2());
        __syncthreads();
        LazyProdSum s;
        s.zero();
        for (long long x0 = jb.begin + kPerThread * t; x0 < jb.end; x0 += kPerThread * blockDim.x) {
            if (jb.aux == kCheckEdges) {
                int to[kPerThread];
                fp2 w[kPerThread];
#pragma unroll
                for (int m = 0; m < kPerThread; ++m)
                    if (x0 + m < jb.end) {
                        to[m] = l.to[x0 + m];
                        w[m] = l.w[x0 + m];
                    }
#pragma unroll
                for (int m = 0; m < kPerThread; ++m)
                    if (x0 + m < jb.end)   // slot e / L of left vertex e mod L
                        s.mac(b.at(l.first + to[m]), mul2(w[m], eu.at(l.src_off + (x0 + m) % l.L)));
            } else if (jb.aux == kCheckQueries) {
                long long qq[kPerThread];
#pragma unroll
                for (int m = 0; m < kPerThread; ++m)
                    if (x0 + m < jb.end) qq[m] = q[x0 + m];
#pragma unroll
                for (int m = 0; m < kPerThread; ++m)
                    if (x0 + m < jb.end) s.mac(b.at(x0 + m), eu.at(qq[m]));
            } else {
#pragma unroll
                for (int m = 0; m < kPerThread; ++m)
                    if (x0 + m < jb.end) s.mac(b.at(x0 + m), eu.at(x0 + m));
            }
        }
        acc[jb.aux == kCheckEdges ? 1 : 0] = s.value();
    } else {
        // V~(r_v) of a layer, the output claim V~_top(r_0) over the queried
        // entries, or the input layer's value V~_0(r_u of layer 1)
        const fp2* r = jb.aux == kDotOut ? ch + l.ch
                     : jb.aux == kDotV ? ch + l.ch + 2 * l.bits_g + l.bits_u
                                       : ch + l.ch + 2 * l.bits_g;
        const int nb = jb.aux == kDotOut ? l.bits_g : l.bits_u;
        const Eq e = make_eq(lds, used, r, nb, one2());
        __syncthreads();
        LazyProdSum s;
        s.zero();
        for (long long x0 = jb.begin + kPerThread * t; x0 < jb.end; x0 += kPerThread * blockDim.x) {
            fp2 val[kPerThread];
#pragma unroll
            for (int m = 0; m < kPerThread; ++m)
                if (x0 + m < jb.end) val[m] = cc[jb.aux == kDotOut ? q[x0 + m] : x0 + m];
#pragma unroll
            for (int m = 0; m < kPerThread; ++m)
                if (x0 + m < jb.end) s.mac(e.at(x0 + m), val[m]);
        }
        acc[0] = s.value();
    }
    block_sum<2>(acc);
    if (t == 0) {
        part[2 * jb.slot] = acc[0];
        part[2 * jb.slot + 1] = acc[1];
    }
}

__global__ void __launch_bounds__(kThreads) k_vg_layer(const VL* __restrict__ layers, int n_layers,
                                                       const Group* __restrict__ groups,
                                                       const fp2* __restrict__ ch,
                                                       const fp2* __restrict__ pf,
                                                       const fp2* __restrict__ F,
                                                       const fp2* __restrict__ part,
                                                       fp2* __restrict__ res)
{
    __shared__ fp2 fa0[1024], fv0[1024], fa1[512], fv1[512];   // ping-pong: halves shrink
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if ((int)blockIdx.x >= n_layers) {
        const Group gr = groups[blockIdx.x - n_layers];
        fp2 acc[2] = {zero2(), zero2()};
        for (int k = t; k < gr.count; k += blockDim.x) {
            acc[0] = add2(acc[0], part[2 * (gr.slot0 + k)]);
            acc[1] = add2(acc[1], part[2 * (gr.slot0 + k) + 1]);
        }
        block_sum<2>(acc);
        if (t == 0) {
            res[gr.res] = acc[0];
            res[gr.res + 1] = acc[1];
        }
        return;
    }
    const VL& l = layers[blockIdx.x];
    const fp2* ru = ch + l.ch + 2 * l.bits_g;
    const fp2* rv = ru + l.bits_u;
    fp2* out = res + l.res;
    const int r1 = l.r1, nwa = l.nwa;
    // rounds 0 .. r1-1: wave j adds the A jobs' sums of round j
    if (wid < r1) {
        fp2 acc[3] = {zero2(), zero2(), zero2()};
        for (int k = lane; k < nwa; k += 64) {
            const fp2* p = pf + l.pf_off + ((long long)k * r1 + wid) * 3;
            acc[0] = add2(acc[0], p[0]);
            acc[1] = add2(acc[1], p[1]);
            acc[2] = add2(acc[2], p[2]);
        }
        wave_sum<3>(acc);
        if (lane == 0) {
            out[3 * wid] = acc[0];
            out[3 * wid + 1] = acc[1];
            out[3 * wid + 2] = acc[2];
        }
    }
    // the rest over the jobs' folded pairs (nwa <= 2^(kMaxBits - kJobRounds)
    // = 1024 of them)
    for (int k = t; k < nwa; k += blockDim.x) {
        fa0[k] = F[l.f_off + 2LL * k];
        fv0[k] = F[l.f_off + 2LL * k + 1];
    }
    __syncthreads();
    int n = nwa, j = r1, cur = 0;
    while (n > 64) {
        const int half = n >> 1;
        const fp2* sa = cur ? fa1 : fa0;
        const fp2* sv = cur ? fv1 : fv0;
        fp2* da_ = cur ? fa0 : fa1;
        fp2* dv_ = cur ? fv0 : fv1;
        LazyProdSum x, y, z;
        x.zero();
        y.zero();
        z.zero();
        if (t < half) {
            const fp2 a0 = sa[2 * t], a1 = sa[2 * t + 1];
            const fp2 v0 = sv[2 * t], v1 = sv[2 * t + 1];
            const fp2 da = sub2(a1, a0), dv = sub2(v1, v0);
            x.mac(da, dv);
            y.mac(da, v0);
            y.mac(a0, dv);
            z.mac(a0, v0);
            da_[t] = fold2(a0, a1, ru[j]);
            dv_[t] = fold2(v0, v1, ru[j]);
        }
        fp2 pr[3] = {x.value(), y.value(), z.value()};
        block_sum<3>(pr);
        if (t == 0) {
            out[3 * j] = pr[0];
            out[3 * j + 1] = pr[1];
            out[3 * j + 2] = pr[2];
        }
        __syncthreads();
        cur ^= 1;
        n = half;
        ++j;
    }
    if (wid == 0) {
        fp2 a = lane < n ? (cur ? fa1 : fa0)[lane] : zero2();
        fp2 v = lane < n ? (cur ? fv1 : fv0)[lane] : zero2();
        wave_rounds(a, v, ru + j, l.bits_u - j, out + 3 * j);
        if (lane == 0) {
            const int b = l.bits_u;
            out[3 * b] = v;       // v_u
            out[3 * b + 1] = a;   // A~(r_u)
            // phase 2 (prover.cpp:634-1006): everything sits at v = 0
            fp2 s = l.kind == kKindDirect ? zero2() : mul2(v, a);
            for (int k = 0; k < b; ++k) {
                out[3 * b + 2 + 3 * k] = zero2();
                out[3 * b + 2 + 3 * k + 1] = sub2(zero2(), s);
                out[3 * b + 2 + 3 * k + 2] = s;
                s = mul2(s, sub2(one2(), rv[k]));
            }
        }
    }
}

int ceil_log2(long long x)
{
    int b = 0;
    while ((1LL << b) < x) ++b;
    return b;
}

struct HLayer {             // host view of layer j of the circuit
    int bits = 0, kind = kKindDirect;
    long long P = 0, first = 0, R = 0, src_off = 0, len = 0;
    const Graph* g = nullptr;
};

fp2 qeval(const fp2* p, fp2 x) { return add2(mul2(add2(mul2(p[0], x), p[1]), x), p[2]); }

DevBuf g_buf;
hipEvent_t g_ev[2] = {nullptr, nullptr};

}  // namespace

bool prove(const Code& code, const fp2* d_cc, std::vector<long long> q,
           const std::function<u64()>& squeeze, hipStream_t st, std::vector<u64>& words,
           bool& accepted, Stats* stats)
{
    accepted = false;
    words.clear();
    const long long N = code.n;
    if (N < 4 || (N & (N - 1))) return fail("virgo: message length %lld is not a power of two >= 4", N);
    const int nq = (int)q.size();
    std::sort(q.begin(), q.end());   // generate_circuit sorts the queries
    for (long long x : q)
        if (x < 0 || x >= code.cw) return fail("virgo: query %lld outside the codeword (%lld)", x, code.cw);
    const Graph empty;
    auto Cg = [&](int d) -> const Graph& { return d < (int)code.C.size() ? code.C[d] : empty; };
    auto Dg = [&](int d) -> const Graph& { return d < (int)code.D.size() ? code.D[d] : empty; };

    // gate counts, prepare_gates_count's arithmetic (linearPC.cpp:477-512)
    std::map<long long, long long> cnt;
    std::function<std::pair<long long, long long>(long long, long long, int)> prep =
        [&](long long in, long long so_far, int dep) -> std::pair<long long, long long> {
        if (in <= kSmall) return {dep, so_far};
        cnt[dep + 1] = so_far + in + Cg(dep).R;
        const auto o = prep(Cg(dep).R, so_far + in, dep + 1);
        cnt[o.first + 1] = o.second + Dg(dep).R;
        return {o.first + 1, cnt[o.first + 1]};
    };
    cnt[0] = N;
    cnt[1] = N + Cg(0).R;
    const auto po = prep(Cg(0).R, N, 1);
    cnt[po.first + 1] = N + po.second + Dg(0).R;
    cnt[po.first + 2] = nq;
    const int D = (int)cnt.size() + 1;
    std::vector<HLayer> H((size_t)D);
    H[0].bits = ceil_log2(N);
    H[0].len = N;
    for (int i = 0; i + 1 < D; ++i) {
        if (!cnt.count(i)) return fail("virgo: gate counts skip depth %d", i);
        H[(size_t)i + 1].bits = ceil_log2(cnt[i]);
    }
    // the layers, generate_circuit's recursion (linearPC.cpp:514-669)
    H[1].kind = kKindDirect;
    H[1].len = N;
    auto code_layer = [&](int j, long long so_far, const Graph& g, long long src_off) {
        HLayer& h = H[(size_t)j];
        h.kind = kKindCode;
        h.P = so_far;
        h.first = so_far;
        h.R = g.R;
        h.src_off = src_off;
        h.g = &g;
        h.len = so_far + g.R;
    };
    bool shape_ok = true;
    code_layer(2, N, Cg(0), 0);
    std::function<std::pair<long long, long long>(long long, long long, int, int)> gen =
        [&](long long in, long long so_far, int rd, int in_depth) -> std::pair<long long, long long> {
        if (in <= kSmall) return {in_depth, so_far};
        if (in_depth + 1 >= D) {
            shape_ok = false;
            return {in_depth, so_far};
        }
        code_layer(in_depth + 1, so_far, Cg(rd), so_far - in);
        const auto o = gen(Cg(rd).R, so_far + Cg(rd).R, rd + 1, in_depth + 1);
        const int fd = (int)o.first + 1;
        if (fd >= D) {
            shape_ok = false;
            return o;
        }
        code_layer(fd, o.second, Dg(rd), so_far);
        return {fd, o.second + Dg(rd).R};
    };
    const auto go = gen(Cg(0).R, N + Cg(0).R, 1, 2);
    const int fd = (int)go.first + 1;
    if (!shape_ok || fd + 2 != D) return fail("virgo: circuit of depth %d does not close at %d", D, fd);
    code_layer(fd, go.second, Dg(0), N);
    if (H[(size_t)fd].len != code.cw)
        return fail("virgo: last code layer holds %lld values, the codeword %lld", H[(size_t)fd].len, code.cw);
    H[(size_t)D - 1].kind = kKindQuery;
    H[(size_t)D - 1].len = nq;
    for (int j = 0; j < D; ++j) {
        if (H[(size_t)j].bits > kMaxBits) return fail("virgo: layer %d of 2^%d gates", j, H[(size_t)j].bits);
        if (H[(size_t)j].len > (1LL << H[(size_t)j].bits))
            return fail("virgo: layer %d holds %lld gates in 2^%d", j, H[(size_t)j].len, H[(size_t)j].bits);
    }

    // every challenge, in the verifier's squeeze order (verifier.cpp:1048-1195)
    auto field = [&]() {
        const u64 re = squeeze() % P;
        return mk2(re, squeeze() % P);
    };
    const int top = D - 1;
    std::vector<fp2> r0((size_t)H[(size_t)top].bits), r1(r0.size());
    for (auto& x : r0) x = field();
    for (auto& x : r1) x = field();
    struct LayerCh { std::vector<fp2> ru, rv; fp2 alpha, beta; };
    std::vector<LayerCh> lc((size_t)D);
    for (int i = top; i >= 1; --i) {
        LayerCh& c = lc[(size_t)i];
        c.ru.resize((size_t)H[(size_t)i - 1].bits);
        c.rv.resize(c.ru.size());
        for (auto& x : c.ru) x = field();
        for (auto& x : c.rv) x = field();
        if (i == 1)
            for (auto& x : c.rv) x = zero2();
        c.alpha = field();
        c.beta = field();
    }

    // device tables: layers top .. 1 in that order
    const int nl = D - 1;
    std::vector<VL> vl((size_t)nl);
    std::vector<fp2> ch;
    std::vector<Job> jobs;
    std::vector<Group> groups;
    long long pf_tot = 0, f_tot = 0;
    int res_tot = 0, slots = 0;
    // jobs over [0, n) for each (aux, n) segment, summed by one group
    auto add_jobs = [&](int layer, int role, std::initializer_list<std::pair<int, long long>> segs,
                        long long chunk) {
        Group gr{slots, 0, res_tot, 0};
        for (const auto& sg : segs)
            for (long long b = 0; b < sg.second; b += chunk) {
                jobs.push_back(Job{layer, role, b, std::min(sg.second, b + chunk), slots++, sg.first});
                ++gr.count;
            }
        res_tot += 2;
        groups.push_back(gr);
        return gr.res;
    };
    for (int k = 0; k < nl; ++k) {
        const int i = top - k;
        const HLayer& h = H[(size_t)i];
        VL& v = vl[(size_t)k];
        memset(&v, 0, sizeof(v));
        v.bits_g = h.bits;
        v.bits_u = H[(size_t)i - 1].bits;
        v.kind = h.kind;
        v.ch = (int)ch.size();
        const std::vector<fp2>& cr0 = i == top ? r0 : lc[(size_t)i + 1].ru;
        const std::vector<fp2>& cr1 = i == top ? r1 : lc[(size_t)i + 1].rv;
        ch.insert(ch.end(), cr0.begin(), cr0.end());
        ch.insert(ch.end(), cr1.begin(), cr1.end());
        ch.insert(ch.end(), lc[(size_t)i].ru.begin(), lc[(size_t)i].ru.end());
        ch.insert(ch.end(), lc[(size_t)i].rv.begin(), lc[(size_t)i].rv.end());
        ch.push_back(i == top ? one2() : lc[(size_t)i + 1].alpha);
        ch.push_back(i == top ? zero2() : lc[(size_t)i + 1].beta);
        v.P = h.P;
        v.first = h.first;
        v.R = h.R;
        v.src_off = h.src_off;
        v.vlen = H[(size_t)i - 1].len;
        if (h.g) {
            v.L = h.g->L;
            v.deg = h.g->deg;
            v.to = h.g->to;
            v.w = h.g->w;
        }
        v.r1 = std::min(kJobRounds, v.bits_u);
        v.nwa = 1 << (v.bits_u - v.r1);
        v.pf_off = pf_tot;
        pf_tot += 3LL * v.nwa * v.r1;
        v.f_off = f_tot;
        f_tot += 2LL * v.nwa;
        v.res = res_tot;
        res_tot += 6 * v.bits_u + 2;
        const long long B = 1LL << v.r1;
        for (int w = 0; w < v.nwa; ++w) jobs.push_back(Job{k, kRoleA, w * B, (w + 1) * B, w, 0});
    }
    // the verifier's predicates (layers above 1) and the dot products: V~(r_v)
    // of every layer, the output claim, the input value
    std::vector<int> chk_res((size_t)nl, -1), vv_res((size_t)nl, -1);
    for (int k = 0; k < nl; ++k) {
        const int i = top - k;
        const VL& v = vl[(size_t)k];
        if (i != 1 && v.kind == kKindQuery)
            chk_res[(size_t)k] = add_jobs(k, kRoleCheck, {{kCheckQueries, (long long)nq}}, kCheckItems);
        else if (i != 1)
            chk_res[(size_t)k] = add_jobs(k, kRoleCheck, {{kCheckRelays, v.P}, {kCheckEdges, v.L * v.deg}},
                                          kCheckItems);
        vv_res[(size_t)k] = add_jobs(k, kRoleDot, {{kDotV, v.vlen}}, kDotItems);
    }
    const int out_res = add_jobs(0, kRoleDot, {{kDotOut, (long long)nq}}, kDotItems);
    const int in_res = add_jobs(nl - 1, kRoleDot, {{kDotIn, N}}, kDotItems);

    // one device arena: layers, jobs, groups, challenges, queries, round sums,
    // folded pairs, partial sums, results
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_vl = 0, o_job = o_vl + al(vl.size() * sizeof(VL)),
                 o_grp = o_job + al(jobs.size() * sizeof(Job)),
                 o_ch = o_grp + al(groups.size() * sizeof(Group)),
                 o_q = o_ch + al(ch.size() * sizeof(fp2)),
                 o_pf = o_q + al(std::max(1, nq) * sizeof(long long)),
                 o_F = o_pf + al((size_t)pf_tot * sizeof(fp2)),
                 o_part = o_F + al((size_t)f_tot * sizeof(fp2)),
                 o_res = o_part + al((size_t)2 * slots * sizeof(fp2)),
                 total = o_res + al((size_t)res_tot * sizeof(fp2));
    if (!g_buf.ensure(total)) return false;
    char* base = g_buf.as<char>();
    std::vector<char> up(o_pf);
    memcpy(up.data() + o_vl, vl.data(), vl.size() * sizeof(VL));
    memcpy(up.data() + o_job, jobs.data(), jobs.size() * sizeof(Job));
    memcpy(up.data() + o_grp, groups.data(), groups.size() * sizeof(Group));
    memcpy(up.data() + o_ch, ch.data(), ch.size() * sizeof(fp2));
    if (nq) memcpy(up.data() + o_q, q.data(), (size_t)nq * sizeof(long long));
    if (!g_ev[0]) {
        TP_HIP_OK(hipEventCreate(&g_ev[0]));
        TP_HIP_OK(hipEventCreate(&g_ev[1]));
    }
    TP_HIP_OK(hipMemcpyAsync(base, up.data(), up.size(), hipMemcpyHostToDevice, st));
    // LDS of a job: beta (two eq tables) and eq(r_u) for a check job, beta
    // and the queried positions' entries for an A job of the output layer
    auto eq_need = [](int nb) { return (size_t)(1 << (nb / 2)) + (size_t)(1 << (nb - nb / 2)); };
    size_t lds_ent = 0;
    for (const VL& v : vl) {
        const size_t beta = 2 * eq_need(v.bits_g);
        lds_ent = std::max(lds_ent, beta + std::max(eq_need(v.bits_u), (size_t)(1 << v.r1)));
    }
    const size_t lds_pre = lds_ent * sizeof(fp2);
    if (lds_pre > 48 * 1024)
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_vg_pre, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds_pre));
    const VL* d_vl = reinterpret_cast<const VL*>(base + o_vl);
    const fp2* d_ch = reinterpret_cast<const fp2*>(base + o_ch);
    const long long* d_q = reinterpret_cast<const long long*>(base + o_q);
    fp2* d_pf = reinterpret_cast<fp2*>(base + o_pf);
    fp2* d_F = reinterpret_cast<fp2*>(base + o_F);
    fp2* d_part = reinterpret_cast<fp2*>(base + o_part);
    fp2* d_res = reinterpret_cast<fp2*>(base + o_res);
    TP_HIP_OK(hipEventRecord(g_ev[0], st));
    k_vg_pre<<<(unsigned)jobs.size(), kPreThreads, lds_pre, st>>>(
        d_vl, reinterpret_cast<const Job*>(base + o_job), d_ch, d_cc, d_q, nq, d_pf, d_F, d_part);
    TP_HIP_OK(hipGetLastError());
    k_vg_layer<<<(unsigned)(nl + groups.size()), kThreads, 0, st>>>(
        d_vl, nl, reinterpret_cast<const Group*>(base + o_grp), d_ch, d_pf, d_F, d_part, d_res);
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipEventRecord(g_ev[1], st));
    std::vector<fp2> res((size_t)res_tot);
    TP_HIP_OK(hipMemcpyAsync(res.data(), d_res, res.size() * sizeof(fp2), hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    if (stats) {
        float ms = 0;
        TP_HIP_OK(hipEventElapsedTime(&ms, g_ev[0], g_ev[1]));
        stats->ms = ms;
        stats->launches = 2;
        stats->layers = nl;
    }

    // the transcript, and the verifier's claim chain (verifier.cpp:1031-1211)
    auto put = [&](u64 x) { words.push_back(x); };
    auto fe = [&](fp2 x) {
        put(x.re);
        put(x.im);
    };
    put((u64)D);
    for (int j = 0; j < D; ++j) put((u64)H[(size_t)j].bits);
    put((u64)nq);
    for (long long x : q) put((u64)x);
    for (const fp2& x : r0) fe(x);
    for (const fp2& x : r1) fe(x);
    fp2 claim = res[(size_t)out_res];   // alpha = 1 on the output layer
    fe(claim);
    bool ok = true;
    fp2 alpha = one2(), beta = zero2();
    const std::vector<fp2>* pr0 = &r0;
    const std::vector<fp2>* pr1 = &r1;
    for (int k = 0; k < nl; ++k) {
        const int i = top - k;
        const VL& v = vl[(size_t)k];
        const LayerCh& c = lc[(size_t)i];
        const int b = v.bits_u;
        for (const fp2& x : c.ru) fe(x);
        for (const fp2& x : c.rv) fe(x);
        const fp2* p1 = &res[(size_t)v.res];
        const fp2 vu = p1[3 * b], vv = res[(size_t)vv_res[(size_t)k]];
        const fp2* p2 = p1 + 3 * b + 2;
        // direct_relay (verifier.cpp:754-766) over layer 1's bits
        fp2 drv = zero2();
        if (i == 1) {
            fp2 d0 = one2(), d1 = one2();
            for (int t = 0; t < H[1].bits; ++t) {
                const fp2 g0 = (*pr0)[(size_t)t], g1 = (*pr1)[(size_t)t], u = c.ru[(size_t)t];
                d0 = mul2(d0, add2(sub2(sub2(one2(), g0), u), mul2(mk2(2, 0), mul2(g0, u))));
                d1 = mul2(d1, add2(sub2(sub2(one2(), g1), u), mul2(mk2(2, 0), mul2(g1, u))));
            }
            drv = add2(mul2(alpha, d0), mul2(beta, d1));
        }
        for (int j = 0; j < b; ++j) {
            const fp2* p = p1 + 3 * j;
            fe(p[0]);
            fe(p[1]);
            fe(p[2]);
            if (!eq2(add2(qeval(p, zero2()), qeval(p, one2())), claim)) ok = false;
            claim = qeval(p, c.ru[(size_t)j]);
        }
        const fp2 dvu = mul2(drv, vu);
        for (int j = 0; j < b; ++j) {
            const fp2* p = p2 + 3 * j;
            fe(p[0]);
            fe(p[1]);
            fe(p[2]);
            if (!eq2(add2(add2(qeval(p, zero2()), qeval(p, one2())), dvu), claim)) ok = false;
            claim = add2(qeval(p, c.rv[(size_t)j]), dvu);
        }
        fe(vu);
        fe(vv);
        // the predicates: relay and custom combination, both times eq(r_v, 0)
        fp2 relay = zero2(), cust = zero2();
        if (i != 1) {
            fp2 bv0 = one2();
            for (const fp2& x : c.rv) bv0 = mul2(bv0, sub2(one2(), x));
            relay = mul2(res[(size_t)chk_res[(size_t)k]], bv0);
            cust = mul2(res[(size_t)chk_res[(size_t)k] + 1], bv0);
        }
        fe(relay);
        fe(cust);
        fe(drv);
        if (!eq2(claim, add2(add2(mul2(relay, vu), mul2(cust, vu)), dvu))) ok = false;
        alpha = c.alpha;
        beta = c.beta;
        fe(alpha);
        fe(beta);
        claim = i != 1 ? add2(mul2(alpha, vu), mul2(beta, vv)) : vu;
        pr0 = &c.ru;
        pr1 = &c.rv;
    }
    fe(claim);
    const fp2 vin = res[(size_t)in_res];
    fe(vin);
    if (!eq2(vin, claim)) ok = false;
    put(ok ? 1 : 0);
    accepted = ok;
    return true;
}

}  // namespace virgo
}  // namespace tp
