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
//   k_vg_pre (256 threads per workgroup, one job each): the A table of
//     every layer (a gather over the forward graph, the beta weights from
//     half-size eq tables in LDS); the verifier's sums (verifier.cpp:349-753)
//     over every layer's gates through the reversed graph: relay and
//     custom-combination predicates, the output claim V~_top(r_0) over the
//     queried entries, the input layer's value at the last r_u.
//   k_vg_layer (1024 threads): one workgroup per layer runs the phase-1
//     rounds (fold by the known challenge, round sums reduced in the
//     workgroup, tables ping-ponged through global scratch), V~(r_v) and
//     the phase-2 polynomials; further workgroups reduce the verifier's
//     partial sums.
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
constexpr int kMaxBits = 20;        // widest layer handled: 2^20 gates
constexpr int kKindDirect = 0, kKindCode = 1, kKindQuery = 2;
constexpr int kRoleA = 0, kRoleCheck = 1, kRoleOut = 2, kRoleIn = 3;
constexpr int kPreThreads = 256, kLayerThreads = 1024;

struct VL {                 // sumcheck layer i: the gates of layer i over layer i-1
    int bits_g, bits_u, kind, ch;   // ch: r_0, r_1 (bits_g each), r_u, r_v (bits_u each), alpha, beta
    long long P;            // gates 0 .. P-1 relay u = g
    long long first, R;     // custom combinations first .. first+R-1
    long long src_off, L;   // their graph's left vertices are u = src_off .. src_off+L-1
    long long vlen;         // V_{i-1}: the first vlen codeword entries, zero beyond
    int deg, res;           // graph degree; results offset (Fp2)
    const int* to;
    const fp2* w;
    const int* rowptr;
    const int* src;
    const fp2* rw;
    long long a_off, s_off; // A table (2^bits_u) and scratch (2^(bits_u+1)) offsets
};

struct Job {                // one workgroup of k_vg_pre
    int layer, role;
    long long begin, end;
    int slot, pad;          // partial-sum slot (check roles)
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

// Builds eq(r, .) over nb bits into lds (2 * 2^ceil(nb/2) entries) at *used.
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

// First position of q (sorted, n entries) holding a value >= u.
__device__ int lower_bound_q(const long long* q, int n, long long u)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (q[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ fp2 ld(const fp2* p, long long i) { return p[i]; }

__global__ void __launch_bounds__(kPreThreads) k_vg_pre(const VL* __restrict__ layers,
                                                        const Job* __restrict__ jobs,
                                                        const fp2* __restrict__ ch,
                                                        const fp2* __restrict__ cc,
                                                        const long long* __restrict__ q, int nq,
                                                        fp2* __restrict__ A, fp2* __restrict__ part)
{
    extern __shared__ fp2 lds[];
    const Job jb = jobs[blockIdx.x];
    const VL& l = layers[jb.layer];
    int used = 0;
    if (jb.role == kRoleA) {
        const Beta b = make_beta(lds, used, l, ch);
        __syncthreads();
        fp2* out = A + l.a_off;
        for (long long u = jb.begin + threadIdx.x; u < jb.end; u += blockDim.x) {
            fp2 a = zero2();
            if (l.kind == kKindDirect) {
                a = b.at(u);
            } else if (l.kind == kKindCode) {
                if (u < l.P) a = b.at(u);
                if (u >= l.src_off && u < l.src_off + l.L) {
                    const long long e0 = (u - l.src_off) * l.deg;
                    LazyProdSum s;
                    s.zero();
                    for (int d = 0; d < l.deg; ++d) s.mac(l.w[e0 + d], b.at(l.first + l.to[e0 + d]));
                    a = add2(a, s.value());
                }
            } else {
                for (int k = lower_bound_q(q, nq, u); k < nq && q[k] == u; ++k) a = add2(a, b.at(k));
            }
            out[u] = a;
        }
        return;
    }
    fp2 acc[2] = {zero2(), zero2()};
    if (jb.role == kRoleCheck) {
        // relay: sum beta(g) eq(r_u, u_g); custom: sum beta(g) sum_e w_e eq(r_u, src_e)
        const Beta b = make_beta(lds, used, l, ch);
        const Eq eu = make_eq(lds, used, ch + l.ch + 2 * l.bits_g, l.bits_u, one2());
        __syncthreads();
        LazyProdSum relay, cust;
        relay.zero();
        cust.zero();
        for (long long g = jb.begin + threadIdx.x; g < jb.end; g += blockDim.x) {
            if (l.kind == kKindQuery) {
                relay.mac(b.at(g), eu.at(q[g]));
            } else if (g < l.P) {
                relay.mac(b.at(g), eu.at(g));
            } else {
                const long long t = g - l.first;
                LazyProdSum e;
                e.zero();
                for (int k = l.rowptr[t]; k < l.rowptr[t + 1]; ++k) e.mac(l.rw[k], eu.at(l.src[k] + l.src_off));
                cust.mac(b.at(g), e.value());
            }
        }
        acc[0] = relay.value();
        acc[1] = cust.value();
    } else {
        // the output claim V~_top(r_0) over the queried entries, or the input
        // layer's value V~_0(r_u of layer 1)
        const fp2* r = jb.role == kRoleOut ? ch + l.ch : ch + l.ch + 2 * l.bits_g;
        const int nb = jb.role == kRoleOut ? l.bits_g : l.bits_u;
        const Eq e = make_eq(lds, used, r, nb, one2());
        __syncthreads();
        LazyProdSum s;
        s.zero();
        for (long long x = jb.begin + threadIdx.x; x < jb.end; x += blockDim.x)
            s.mac(e.at(x), ld(cc, jb.role == kRoleOut ? q[x] : x));
        acc[0] = s.value();
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part[2 * jb.slot] = acc[0];
        part[2 * jb.slot + 1] = acc[1];
    }
}

__global__ void __launch_bounds__(kLayerThreads) k_vg_layer(const VL* __restrict__ layers, int n_layers,
                                                            const Group* __restrict__ groups,
                                                            const fp2* __restrict__ ch,
                                                            const fp2* __restrict__ cc,
                                                            const fp2* __restrict__ A,
                                                            fp2* __restrict__ S,
                                                            const fp2* __restrict__ part,
                                                            fp2* __restrict__ res)
{
    extern __shared__ fp2 lds[];
    if ((int)blockIdx.x >= n_layers) {
        const Group gr = groups[blockIdx.x - n_layers];
        fp2 acc[2] = {zero2(), zero2()};
        for (int k = threadIdx.x; k < gr.count; k += blockDim.x) {
            acc[0] = add2(acc[0], part[2 * (gr.slot0 + k)]);
            acc[1] = add2(acc[1], part[2 * (gr.slot0 + k) + 1]);
        }
        block_sum<2>(acc);
        if (threadIdx.x == 0) {
            res[gr.res] = acc[0];
            res[gr.res + 1] = acc[1];
        }
        return;
    }
    const VL& l = layers[blockIdx.x];
    const fp2* ru = ch + l.ch + 2 * l.bits_g;
    const fp2* rv = ru + l.bits_u;
    fp2* out = res + l.res;
    long long n = 1LL << l.bits_u;
    const long long h0 = n >> 1;
    fp2* buf[2][2] = {{S + l.s_off, S + l.s_off + h0}, {S + l.s_off + 2 * h0, S + l.s_off + 3 * h0}};
    const fp2* srcA = A + l.a_off;
    const fp2* srcV = cc;
    const long long vlen = l.vlen;
    // phase 1: round j pairs entries 2i, 2i+1 of the tables folded by
    // r_u[0 .. j-1] (prover.cpp:576-633); the fold by r_u[j] is made now
    for (int j = 0; j < l.bits_u; ++j) {
        const long long half = n >> 1;
        fp2* dA = buf[j & 1][0];
        fp2* dV = buf[j & 1][1];
        const fp2 r = ru[j];
        LazyProdSum sa, sb, sc;
        sa.zero();
        sb.zero();
        sc.zero();
        for (long long i = threadIdx.x; i < half; i += blockDim.x) {
            const fp2 a0 = srcA[2 * i], a1 = srcA[2 * i + 1];
            fp2 v0, v1;
            if (j == 0) {
                v0 = 2 * i < vlen ? cc[2 * i] : zero2();
                v1 = 2 * i + 1 < vlen ? cc[2 * i + 1] : zero2();
            } else {
                v0 = srcV[2 * i];
                v1 = srcV[2 * i + 1];
            }
            const fp2 da = sub2(a1, a0), dv = sub2(v1, v0);
            sa.mac(da, dv);
            sb.mac(da, v0);
            sb.mac(a0, dv);
            sc.mac(a0, v0);
            dA[i] = fold2(a0, a1, r);
            dV[i] = fold2(v0, v1, r);
        }
        fp2 poly[3] = {sa.value(), sb.value(), sc.value()};
        block_sum<3>(poly);
        if (threadIdx.x == 0) {
            out[3 * j] = poly[0];
            out[3 * j + 1] = poly[1];
            out[3 * j + 2] = poly[2];
        }
        srcA = dA;
        srcV = dV;
        n = half;
    }
    // V~(r_v) over the layer below, the eq weights from LDS
    int used = 0;
    const Eq ev = make_eq(lds, used, rv, l.bits_u, one2());
    __syncthreads();
    LazyProdSum sv;
    sv.zero();
    for (long long x = threadIdx.x; x < vlen; x += blockDim.x) sv.mac(ev.at(x), cc[x]);
    fp2 vv[1] = {sv.value()};
    block_sum<1>(vv);
    if (threadIdx.x == 0) {
        const int b = l.bits_u;
        const fp2 af = srcA[0], vu = srcV[0];
        out[3 * b] = vu;
        out[3 * b + 1] = vv[0];
        out[3 * b + 2] = af;
        // phase 2 (prover.cpp:634-1006): everything sits at v = 0
        fp2 s = l.kind == kKindDirect ? zero2() : mul2(vu, af);
        for (int j = 0; j < b; ++j) {
            out[3 * b + 3 + 3 * j] = zero2();
            out[3 * b + 3 + 3 * j + 1] = sub2(zero2(), s);
            out[3 * b + 3 + 3 * j + 2] = s;
            s = mul2(s, sub2(one2(), rv[j]));
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

// Host Fp2 helpers for the verifier's claim chain.
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
    long long a_tot = 0, s_tot = 0;
    int res_tot = 0, slots = 0;
    auto add_jobs = [&](int layer, int role, long long n, long long chunk) {
        Group gr{slots, 0, 0, 0};
        for (long long b = 0; b < n; b += chunk) {
            jobs.push_back(Job{layer, role, b, std::min(n, b + chunk), slots++, 0});
            ++gr.count;
        }
        return gr;
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
            v.rowptr = h.g->rowptr;
            v.src = h.g->src;
            v.rw = h.g->rw;
        }
        v.a_off = a_tot;
        a_tot += 1LL << v.bits_u;
        v.s_off = s_tot;
        s_tot += 2LL << v.bits_u;
        v.res = res_tot;
        res_tot += 6 * v.bits_u + 3;
        add_jobs(k, kRoleA, 1LL << v.bits_u, 1024);
    }
    // the verifier's sums: per layer above 1, the output claim, the input value
    std::vector<int> chk_res((size_t)nl, -1);
    for (int k = 0; k < nl; ++k) {
        const int i = top - k;
        if (i == 1) continue;
        const HLayer& h = H[(size_t)i];
        Group gr = add_jobs(k, kRoleCheck, h.kind == kKindQuery ? nq : h.P + h.R, 512);
        gr.res = res_tot;
        chk_res[(size_t)k] = res_tot;
        res_tot += 2;
        groups.push_back(gr);
    }
    Group go_out = add_jobs(0, kRoleOut, nq, 1024);
    go_out.res = res_tot;
    const int out_res = res_tot;
    res_tot += 2;
    groups.push_back(go_out);
    Group go_in = add_jobs(nl - 1, kRoleIn, N, 4096);
    go_in.res = res_tot;
    const int in_res = res_tot;
    res_tot += 2;
    groups.push_back(go_in);

    // one device arena: layers, jobs, groups, challenges, queries, A, scratch,
    // partials, results
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_vl = 0, o_job = o_vl + al(vl.size() * sizeof(VL)),
                 o_grp = o_job + al(jobs.size() * sizeof(Job)),
                 o_ch = o_grp + al(groups.size() * sizeof(Group)),
                 o_q = o_ch + al(ch.size() * sizeof(fp2)),
                 o_A = o_q + al(std::max(1, nq) * sizeof(long long)),
                 o_S = o_A + al((size_t)a_tot * sizeof(fp2)),
                 o_part = o_S + al((size_t)s_tot * sizeof(fp2)),
                 o_res = o_part + al((size_t)2 * slots * sizeof(fp2)),
                 total = o_res + al((size_t)res_tot * sizeof(fp2));
    if (!g_buf.ensure(total)) return false;
    char* base = g_buf.as<char>();
    std::vector<char> up(o_A);
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
    int hmax = 0;
    for (const VL& v : vl) hmax = std::max(hmax, std::max(v.bits_g, v.bits_u) - std::max(v.bits_g, v.bits_u) / 2);
    const size_t lds_pre = (size_t)6 << hmax << 4, lds_layer = (size_t)2 << hmax << 4;
    if (lds_pre > 64 * 1024)
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_vg_pre, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds_pre));
    const VL* d_vl = reinterpret_cast<const VL*>(base + o_vl);
    const fp2* d_ch = reinterpret_cast<const fp2*>(base + o_ch);
    const long long* d_q = reinterpret_cast<const long long*>(base + o_q);
    fp2* d_A = reinterpret_cast<fp2*>(base + o_A);
    fp2* d_S = reinterpret_cast<fp2*>(base + o_S);
    fp2* d_part = reinterpret_cast<fp2*>(base + o_part);
    fp2* d_res = reinterpret_cast<fp2*>(base + o_res);
    TP_HIP_OK(hipEventRecord(g_ev[0], st));
    k_vg_pre<<<(unsigned)jobs.size(), kPreThreads, lds_pre, st>>>(
        d_vl, reinterpret_cast<const Job*>(base + o_job), d_ch, d_cc, d_q, nq, d_A, d_part);
    TP_HIP_OK(hipGetLastError());
    k_vg_layer<<<(unsigned)(nl + groups.size()), kLayerThreads, lds_layer, st>>>(
        d_vl, nl, reinterpret_cast<const Group*>(base + o_grp), d_ch, d_cc, d_A, d_S, d_part, d_res);
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipEventRecord(g_ev[1], st));
    std::vector<fp2> res((size_t)res_tot);
    TP_HIP_OK(hipMemcpyAsync(res.data(), d_res, res.size() * sizeof(fp2), hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    if (stats) {
        float t = 0;
        TP_HIP_OK(hipEventElapsedTime(&t, g_ev[0], g_ev[1]));
        stats->ms = t;
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
        const fp2 vu = p1[3 * b], vv = p1[3 * b + 1];
        const fp2* p2 = p1 + 3 * b + 3;
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
