// orion.hip - the TruePix signer's Orion commitment of a frame's inputs on the
// device (OZ/examples/multivariate_PC_test_commit.cpp:265-331, which calls
// commit_pair, OZ/src/VPD/linearPC.cpp:253-459).  The two Merkle roots it
// produces are what orion_commit.json carries and what the frame prover's
// transcript absorbs as the commitment binding.
//
// What is committed (OZ/include/truepix/truepix_gkr.h:85-386):
//   n = next_pow2(copies) * 512 coefficients, copy-major: the copy's 384
//   samples, the constant 128 at wire 384, zero elsewhere; the mask's real
//   part is splitmix64(seed) % p at the k-th sample position, in order; the
//   masked polynomial is coefficients + mask.  Each polynomial is split into
//   128 columns of col = n / 128 coefficients (column c = flat indices
//   c col .. c col + col - 1).
// Encoding (OZ/include/linear_code/expanders.h:19-60,
// linear_code_encode.h:35-110): a column x of length m > 13 becomes
//   x | codeword(C x) | D codeword(C x)
// with C (m inputs -> 0.238 m outputs, 10 edges per input) and D (the deeper
// codeword -> 0.72 m minus its length outputs, 20 edges per entry) random
// weighted bipartite graphs drawn from glibc rand() in expander_init's order;
// a column of length <= 13 is its own codeword.
// Hashing (OZ/src/infrastructure/merkle_tree.cpp:15-79): every row r of the
// 2 col rows gets 64 chained SHA-256(state | col[2j][r] | col[2j+1][r]) from a
// zero state (elements as (imag, real) little-endian, zero past the
// codeword); a heap-ordered binary tree over the rows with node =
// SHA-256(left | right) gives the root.
//
// Device layout: both polynomials side by side as rows of 256 lanes (lane c
// < 128: masked column c; lane 128 + c: mask column c), row k = codeword
// position k of every column, 16 B per element, so
//   - an expander step is a gather per target position: one workgroup per
//     target, its in-edges (reversed graph, CSR) uniform across the
//     workgroup, 256 coalesced 16 B loads per edge;
//   - the codeword of each depth sits in place: the inputs of depth d at rows
//     base_d .. base_d + m_d, its D part right after the deeper codeword;
//   - the row hash is one thread per (row, polynomial), 128 compressions in
//     sequence (independent across rows), the padding block of every 64-byte
//     message taken from a precomputed schedule;
//   - the all-zero rows past the codeword share one leaf;
//   - the tree is one launch per level, one workgroup per tree for the last
//     eleven levels.
#include "fp61.hpp"
#include "glibc_rand.hpp"
#include "runtime.hpp"
#include "../../include/truepix_hip.h"

#include <stdint.h>
#include <string.h>

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

using namespace tp;

namespace {

// OZ/include/linear_code/parameter.h
constexpr double kAlpha = 0.238, kRate = 1.72;
constexpr int kCn = 10, kDn = 20;
constexpr int kSmall = (int)(1.0 / 0.07) - 1;   // distance_threshold = 13
constexpr int kCols = 128;                      // column_size
constexpr int kLanes = 2 * kCols;               // both polynomials
constexpr u64 kGolden = 0x9E3779B97F4A7C15ULL;

struct alignas(16) V2 {
    u64 re, im;
};

// ------------------------------------------------------------- SHA-256 ----

__constant__ uint32_t c_K[64] = {
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
// K[t] + W[t] of the block that pads a 64-byte message (0x80, zeros, the bit
// length 512): the block is fixed, so its whole schedule is (set up once).
__constant__ uint32_t c_pad_kw[64];

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
// x ^ y ^ z in one v_bitop3_b32 (truth table 0x96, the 3-input parity)
__device__ __forceinline__ uint32_t xor3(uint32_t x, uint32_t y, uint32_t z)
{
    return __builtin_amdgcn_bitop3_b32(x, y, z, 0x96);
}

#define TP_SHA_ROUND(KW)                                                                   \
    do {                                                                                   \
        const uint32_t t1 = h + xor3(rotr(e, 6), rotr(e, 11), rotr(e, 25)) +               \
                            ((e & f) ^ (~e & g)) + (KW);                                   \
        const uint32_t t2 = xor3(rotr(a, 2), rotr(a, 13), rotr(a, 22)) +                   \
                            ((a & b) ^ (a & c) ^ (b & c));                                 \
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;                 \
    } while (0)

// SHA-256 of a 64-byte message given as 16 big-endian words; the digest as 8
// big-endian words.
__device__ __forceinline__ void sha64(uint32_t w[16], uint32_t out[8])
{
    uint32_t a = 0x6a09e667u, b = 0xbb67ae85u, c = 0x3c6ef372u, d = 0xa54ff53au;
    uint32_t e = 0x510e527fu, f = 0x9b05688cu, g = 0x1f83d9abu, h = 0x5be0cd19u;
#pragma unroll
    for (int t = 0; t < 64; ++t) {
        uint32_t wt;
        if (t < 16) {
            wt = w[t];
        } else {
            const uint32_t w15 = w[(t - 15) & 15], w2 = w[(t - 2) & 15];
            wt = w[t & 15] + xor3(rotr(w15, 7), rotr(w15, 18), w15 >> 3) + w[(t - 7) & 15] +
                 xor3(rotr(w2, 17), rotr(w2, 19), w2 >> 10);
            w[t & 15] = wt;
        }
        TP_SHA_ROUND(c_K[t] + wt);
    }
    const uint32_t s0 = 0x6a09e667u + a, s1 = 0xbb67ae85u + b, s2 = 0x3c6ef372u + c,
                   s3 = 0xa54ff53au + d, s4 = 0x510e527fu + e, s5 = 0x9b05688cu + f,
                   s6 = 0x1f83d9abu + g, s7 = 0x5be0cd19u + h;
    a = s0; b = s1; c = s2; d = s3; e = s4; f = s5; g = s6; h = s7;
#pragma unroll
    for (int t = 0; t < 64; ++t) TP_SHA_ROUND(c_pad_kw[t]);
    out[0] = s0 + a; out[1] = s1 + b; out[2] = s2 + c; out[3] = s3 + d;
    out[4] = s4 + e; out[5] = s5 + f; out[6] = s6 + g; out[7] = s7 + h;
}

// Host: the padding block's K[t] + W[t].
void pad_schedule(uint32_t kw[64])
{
    static const uint32_t K[64] = {
        0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u,
        0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u,
        0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u,
        0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
        0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u,
        0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u,
        0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
        0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
        0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au,
        0x5b9cca4fu, 0x682e6ff3u, 0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u,
        0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
    auto r = [](uint32_t x, int n) { return (x >> n) | (x << (32 - n)); };
    uint32_t w[64] = {0x80000000u};
    w[15] = 512;
    for (int t = 16; t < 64; ++t)
        w[t] = (r(w[t - 2], 17) ^ r(w[t - 2], 19) ^ (w[t - 2] >> 10)) + w[t - 7] +
               (r(w[t - 15], 7) ^ r(w[t - 15], 18) ^ (w[t - 15] >> 3)) + w[t - 16];
    for (int t = 0; t < 64; ++t) kw[t] = K[t] + w[t];
}

__device__ __forceinline__ uint32_t be32(uint32_t x) { return __builtin_bswap32(x); }

// ------------------------------------------------------------- kernels ----

// Both polynomials as rows of 256 lanes (rows 0 .. col - 1).
__global__ void k_or_fill(const uint8_t* __restrict__ in, int copies, u64 seed, long long col,
                          V2* __restrict__ X)
{
    const size_t total = (size_t)col * kLanes;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        const size_t k = t >> 8;
        const int lane = (int)(t & 255);
        const size_t flat = (size_t)(lane & 127) * col + k;
        const size_t copy = flat >> 9;
        const int wire = (int)(flat & 511);
        u64 re = 0;
        if (copy < (size_t)copies) {
            if (wire < 384) {
                const u64 rank = copy * 384 + wire;
                u64 z = seed + (rank + 1) * kGolden;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
                z ^= z >> 31;
                const u64 mask = fold61(z);
                re = lane < kCols ? fadd((u64)in[copy * 384 + wire], mask) : mask;
            } else if (wire == 384 && lane < kCols) {
                re = 128;
            }
        }
        V2 v;
        v.re = re;
        v.im = 0;
        X[t] = v;
    }
}

// One expander step for every lane: row out_base + t = sum over target t's
// in-edges (i, w) of w * row in_base + i; one workgroup per target.
__global__ void __launch_bounds__(kLanes) k_or_step(V2* __restrict__ X, long long in_base,
                                                    long long out_base, const int* __restrict__ rp,
                                                    const int* __restrict__ src,
                                                    const V2* __restrict__ w)
{
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    fp2 acc = zero2();
    const int e1 = rp[t + 1];
    int e = rp[t];
    // eight edges' loads in flight before their products (the gathers are
    // independent; one at a time each would wait out the L2 latency)
    for (; e + 8 <= e1; e += 8) {
        V2 x[8], we[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            we[q] = w[e + q];
            x[q] = X[(size_t)(in_base + src[e + q]) * kLanes + lane];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            acc = add2(acc, mul2(mk2(we[q].re, we[q].im), mk2(x[q].re, x[q].im)));
    }
    for (; e < e1; ++e) {
        const V2 we = w[e];
        const V2 x = X[(size_t)(in_base + src[e]) * kLanes + lane];
        acc = add2(acc, mul2(mk2(we.re, we.im), mk2(x.re, x.im)));
    }
    V2 o;
    o.re = acc.re;
    o.im = acc.im;
    X[(size_t)(out_base + t) * kLanes + lane] = o;
}

// Row leaves: one thread per (row, polynomial) for rows 0 .. cw, row cw
// standing for every all-zero row past the codeword; leaves land at the
// tree's bottom level, tree[poly][rows + r] (8 words each).
__global__ void __launch_bounds__(256) k_or_rows(const V2* __restrict__ X, long long cw,
                                                 long long rows, uint32_t* __restrict__ tree)
{
    const long long span = cw < rows ? cw + 1 : cw;
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= 2 * span) return;
    const int poly = (int)(id / span);
    const long long r = id - poly * span;
    const bool data = r < cw;
    uint32_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const V2* row = X + (size_t)(data ? r : 0) * kLanes + poly * kCols;
    V2 na, nb;
    na.re = na.im = nb.re = nb.im = 0;
    if (data) {
        na = row[0];
        nb = row[1];
    }
    for (int j = 0; j < kCols / 2; ++j) {
        // this pair, and the next one's loads in flight behind the hash
        const V2 a = na, b = nb;
        if (data && j + 1 < kCols / 2) {
            na = row[2 * j + 2];
            nb = row[2 * j + 3];
        }
        uint32_t m[16];
#pragma unroll
        for (int k = 0; k < 8; ++k) m[k] = st[k];
        m[8] = be32((uint32_t)a.im);
        m[9] = be32((uint32_t)(a.im >> 32));
        m[10] = be32((uint32_t)a.re);
        m[11] = be32((uint32_t)(a.re >> 32));
        m[12] = be32((uint32_t)b.im);
        m[13] = be32((uint32_t)(b.im >> 32));
        m[14] = be32((uint32_t)b.re);
        m[15] = be32((uint32_t)(b.re >> 32));
        sha64(m, st);
    }
    uint32_t* leaf = tree + ((size_t)poly * 2 * rows + rows + r) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) leaf[k] = st[k];
}

__device__ __forceinline__ void tree_node(uint32_t* base, long long v)
{
    uint32_t m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = base[(size_t)2 * v * 8 + k];
    uint32_t out[8];
    sha64(m, out);
#pragma unroll
    for (int k = 0; k < 8; ++k) base[(size_t)v * 8 + k] = out[k];
}

// The all-zero rows past the codeword share row cw's leaf.
__global__ void __launch_bounds__(256) k_or_zero_rows(long long cw, long long rows,
                                                      uint32_t* __restrict__ tree)
{
    const long long n = rows - cw - 1;
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= 2 * n * 8) return;
    const int poly = (int)(id / (n * 8));
    const long long k = id - poly * n * 8;
    uint32_t* base = tree + (size_t)poly * 2 * rows * 8;
    base[(size_t)(rows + cw + 1) * 8 + k] = base[(size_t)(rows + cw) * 8 + (k & 7)];
}

// One tree level, nodes lo .. 2 lo - 1 of both trees.
__global__ void __launch_bounds__(256) k_or_level(uint32_t* __restrict__ tree, long long rows,
                                                  long long lo)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= 2 * lo) return;
    const int poly = (int)(id / lo);
    tree_node(tree + (size_t)poly * 2 * rows * 8, lo + (id - poly * lo));
}

// The levels below and including lo (nodes < 2 lo) of one tree per workgroup.
__global__ void __launch_bounds__(1024) k_or_top(uint32_t* __restrict__ tree, long long rows,
                                                 long long lo)
{
    uint32_t* base = tree + (size_t)blockIdx.x * 2 * rows * 8;
    for (long long l = lo; l >= 1; l >>= 1) {
        for (long long v = l + threadIdx.x; v < 2 * l; v += blockDim.x) tree_node(base, v);
        __syncthreads();
    }
}

// -------------------------------------------------------------- plans ----

// One column length's code: the graphs drawn as the reference draws them,
// reversed into per-target edge lists, and the in-place row offsets.
struct Plan {
    long long col = 0, cw = 0;
    struct Step {
        long long in_base, out_base, targets;
        int rp_off;   // into rowptr; edges from rowptr values into src / w
    };
    std::vector<Step> down, up;   // C steps by depth, D steps from the deepest up
    DevBuf rowptr, src, w;
};

struct Fwd {
    long long L, R;
    int deg;
    std::vector<int> to;
    std::vector<fp2> w;
};

Fwd draw(GlibcRand& G, long long L, long long R, int deg)
{
    Fwd g;
    g.L = L;
    g.R = R;
    g.deg = deg;
    g.to.resize((size_t)L * deg);
    g.w.resize((size_t)L * deg);
    for (size_t k = 0; k < g.to.size(); ++k) {
        g.to[k] = (int)(G.next() % (uint32_t)R);
        g.w[k] = G.fe();
    }
    return g;
}

bool build_plan(Plan& P, long long col)
{
    GlibcRand G(1);
    std::vector<Fwd> C, D;
    std::vector<long long> m;   // input length per depth
    // expander_init order: C at this depth, the deeper levels, then D here
    std::function<long long(long long, int)> init = [&](long long n, int dep) -> long long {
        if ((int)m.size() <= dep) m.resize(dep + 1);
        m[dep] = n;
        if (n <= kSmall) return n;
        const long long R = (long long)(kAlpha * n);
        if ((int)C.size() <= dep) C.resize(dep + 1);
        C[dep] = draw(G, n, R, kCn);
        const long long L = init(R, dep + 1);
        const long long R2 = (long long)(n * (kRate - 1) - L);
        if ((int)D.size() <= dep) D.resize(dep + 1);
        D[dep] = draw(G, L, R2, kDn);
        return n + L + R2;
    };
    P.col = col;
    P.cw = init(col, 0);
    const int depth = (int)C.size();
    std::vector<long long> base(depth + 1, 0), len(depth + 1, 0);
    for (int d = 0; d < depth; ++d) base[d + 1] = base[d] + m[d];
    len[depth] = m[depth];
    for (int d = depth - 1; d >= 0; --d) len[d] = m[d] + len[d + 1] + D[d].R;
    if (len[0] != P.cw) return fail("orion plan: codeword length mismatch");
    // reversed graphs, CSR
    std::vector<int> rp, src;
    std::vector<V2> ww;
    auto reverse = [&](const Fwd& g) {
        const int off = (int)rp.size();
        std::vector<int> cnt(g.R + 1, 0);
        for (int t : g.to) ++cnt[t + 1];
        for (long long t = 0; t < g.R; ++t) cnt[t + 1] += cnt[t];
        const int e0 = (int)src.size();
        for (long long t = 0; t <= g.R; ++t) rp.push_back(e0 + cnt[t]);
        src.resize(e0 + g.to.size());
        ww.resize(e0 + g.to.size());
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (long long i = 0; i < g.L; ++i)
            for (int d = 0; d < g.deg; ++d) {
                const size_t k = (size_t)i * g.deg + d;
                const int at = e0 + fill[g.to[k]]++;
                src[at] = (int)i;
                ww[at].re = g.w[k].re;
                ww[at].im = g.w[k].im;
            }
        return off;
    };
    P.down.clear();
    P.up.clear();
    for (int d = 0; d < depth; ++d)
        P.down.push_back({base[d], base[d + 1], C[d].R, reverse(C[d])});
    for (int d = depth - 1; d >= 0; --d)
        P.up.push_back({base[d + 1], base[d] + m[d] + len[d + 1], D[d].R, reverse(D[d])});
    if (rp.empty()) rp.push_back(0);
    if (src.empty()) {
        src.push_back(0);
        ww.push_back(V2{0, 0});
    }
    hipStream_t st = stream();
    if (!P.rowptr.ensure(rp.size() * 4) || !P.src.ensure(src.size() * 4) ||
        !P.w.ensure(ww.size() * sizeof(V2)))
        return false;
    TP_HIP_OK(hipMemcpyAsync(P.rowptr.ptr, rp.data(), rp.size() * 4, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(P.src.ptr, src.data(), src.size() * 4, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(P.w.ptr, ww.data(), ww.size() * sizeof(V2), hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return true;
}

std::mutex g_mu;
std::map<long long, std::unique_ptr<Plan>> g_plans;
DevBuf g_X, g_tree, g_in;
hipEvent_t g_ev[2] = {nullptr, nullptr};

Plan* plan_for(long long col)
{
    auto it = g_plans.find(col);
    if (it != g_plans.end()) return it->second.get();
    std::unique_ptr<Plan> P(new Plan());
    if (!build_plan(*P, col)) return nullptr;
    Plan* raw = P.get();
    g_plans.emplace(col, std::move(P));
    return raw;
}

bool setup()
{
    static bool done = false;
    if (done) return true;
    uint32_t kw[64];
    pad_schedule(kw);
    TP_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_pad_kw), kw, sizeof(kw)));
    TP_HIP_OK(hipEventCreate(&g_ev[0]));
    TP_HIP_OK(hipEventCreate(&g_ev[1]));
    done = true;
    return true;
}

bool encode_all(const Plan& P, hipStream_t st)
{
    const int* rp = P.rowptr.as<int>();
    for (const auto& s : P.down) {
        if (!s.targets) continue;
        k_or_step<<<(unsigned)s.targets, kLanes, 0, st>>>(g_X.as<V2>(), s.in_base, s.out_base,
                                                          rp + s.rp_off, P.src.as<int>(),
                                                          P.w.as<V2>());
        TP_HIP_OK(hipGetLastError());
    }
    for (const auto& s : P.up) {
        if (!s.targets) continue;
        k_or_step<<<(unsigned)s.targets, kLanes, 0, st>>>(g_X.as<V2>(), s.in_base, s.out_base,
                                                          rp + s.rp_off, P.src.as<int>(),
                                                          P.w.as<V2>());
        TP_HIP_OK(hipGetLastError());
    }
    return true;
}

void hex_root(const uint32_t st[8], char out[65])
{
    uint8_t d[32];
    for (int k = 0; k < 8; ++k) {
        d[4 * k] = (uint8_t)(st[k] >> 24);
        d[4 * k + 1] = (uint8_t)(st[k] >> 16);
        d[4 * k + 2] = (uint8_t)(st[k] >> 8);
        d[4 * k + 3] = (uint8_t)st[k];
    }
    // root_hex prints the digest's two 16-byte halves as (high, low) u64 pairs
    static const char hx[] = "0123456789abcdef";
    const int order[4] = {8, 0, 24, 16};
    for (int q = 0; q < 4; ++q) {
        u64 v = 0;
        for (int k = 7; k >= 0; --k) v = (v << 8) | d[order[q] + k];
        for (int k = 0; k < 16; ++k) out[16 * q + k] = hx[(v >> (60 - 4 * k)) & 15];
    }
    out[64] = 0;
}

}  // namespace

extern "C" {

// The signer's Orion commitment of one frame: inputs [copies][384] sample
// values (the lines of input.txt), the mask seed; writes the masked and mask
// Merkle roots as orion_commit.json carries them (64 hex digits + NUL) and,
// if ms is not null, the device time in milliseconds.  0 / -1.
int tp_orion_commit(int copies, const uint8_t* inputs, uint64_t seed, char masked_root[65],
                    char mask_root[65], double* ms)
{
    clear_error();
    if (!device_ready()) return -1;
    if (copies < 1 || copies > (1 << 16) || !inputs || !masked_root || !mask_root)
        return fail("tp_orion_commit: bad arguments (copies=%d)", copies), -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!setup()) return -1;
    long long pc = 1;
    while (pc < copies) pc <<= 1;
    const long long n = pc * 512, col = n / kCols, rows = 2 * col;
    Plan* P = plan_for(col);
    if (!P) return -1;
    hipStream_t st = stream();
    const size_t in_bytes = (size_t)copies * 384;
    if (!g_in.ensure(in_bytes) || !g_X.ensure((size_t)P->cw * kLanes * sizeof(V2)) ||
        !g_tree.ensure((size_t)2 * 2 * rows * 8 * 4))
        return -1;
    TP_HIP_OK(hipMemcpyAsync(g_in.ptr, inputs, in_bytes, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipEventRecord(g_ev[0], st));
    k_or_fill<<<grid_for((size_t)col * kLanes, kBlock, 8192), kBlock, 0, st>>>(
        g_in.as<uint8_t>(), copies, seed, col, g_X.as<V2>());
    TP_HIP_OK(hipGetLastError());
    if (!encode_all(*P, st)) return -1;
    const long long span = P->cw < rows ? P->cw + 1 : P->cw;
    k_or_rows<<<(unsigned)((2 * span + 255) / 256), 256, 0, st>>>(g_X.as<V2>(), P->cw, rows,
                                                                  g_tree.as<uint32_t>());
    TP_HIP_OK(hipGetLastError());
    if (rows - P->cw - 1 > 0) {
        const long long words = 2 * (rows - P->cw - 1) * 8;
        k_or_zero_rows<<<(unsigned)((words + 255) / 256), 256, 0, st>>>(P->cw, rows,
                                                                       g_tree.as<uint32_t>());
        TP_HIP_OK(hipGetLastError());
    }
    long long lo = rows / 2;
    for (; lo > 1024; lo >>= 1) {
        k_or_level<<<(unsigned)((2 * lo + 255) / 256), 256, 0, st>>>(g_tree.as<uint32_t>(), rows, lo);
        TP_HIP_OK(hipGetLastError());
    }
    if (lo >= 1) {
        k_or_top<<<2, 1024, 0, st>>>(g_tree.as<uint32_t>(), rows, lo);
        TP_HIP_OK(hipGetLastError());
    }
    TP_HIP_OK(hipEventRecord(g_ev[1], st));
    uint32_t roots[2][8];
    for (int k = 0; k < 2; ++k)
        TP_HIP_OK(hipMemcpyAsync(roots[k], g_tree.as<uint32_t>() + ((size_t)k * 2 * rows + 1) * 8, 32,
                                 hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    if (ms) {
        float t = 0;
        TP_HIP_OK(hipEventElapsedTime(&t, g_ev[0], g_ev[1]));
        *ms = t;
    }
    hex_root(roots[0], masked_root);
    hex_root(roots[1], mask_root);
    return 0;
}

// The expander code alone: n_cols <= 256 columns of n Fp2 each, src
// [n_cols][n] (re, im), dst [n_cols][cw] (re, im).  Returns cw, or -1.
long long tp_orion_encode(long long n, int n_cols, const uint64_t* src, uint64_t* dst)
{
    clear_error();
    if (!device_ready()) return -1;
    if (n < 1 || n > (1LL << 16) || n_cols < 1 || n_cols > kLanes || !src || !dst)
        return fail("tp_orion_encode: bad arguments"), -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!setup()) return -1;
    Plan* P = plan_for(n);
    if (!P) return -1;
    hipStream_t st = stream();
    const long long cw = P->cw;
    std::vector<V2> X((size_t)cw * kLanes);
    memset(X.data(), 0, X.size() * sizeof(V2));
    for (int c = 0; c < n_cols; ++c)
        for (long long k = 0; k < n; ++k) {
            X[(size_t)k * kLanes + c].re = src[2 * ((size_t)c * n + k)];
            X[(size_t)k * kLanes + c].im = src[2 * ((size_t)c * n + k) + 1];
        }
    if (!g_X.ensure(X.size() * sizeof(V2))) return -1;
    TP_HIP_OK(hipMemcpyAsync(g_X.ptr, X.data(), X.size() * sizeof(V2), hipMemcpyHostToDevice, st));
    if (!encode_all(*P, st)) return -1;
    TP_HIP_OK(hipMemcpyAsync(X.data(), g_X.ptr, X.size() * sizeof(V2), hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    for (int c = 0; c < n_cols; ++c)
        for (long long k = 0; k < cw; ++k) {
            dst[2 * ((size_t)c * cw + k)] = X[(size_t)k * kLanes + c].re;
            dst[2 * ((size_t)c * cw + k) + 1] = X[(size_t)k * kLanes + c].im;
        }
    return cw;
}

}  // extern "C"
