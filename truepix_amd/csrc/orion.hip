// orion.hip - the TruePix signer's Orion commitment of a frame's inputs on the
// device (OZ/examples/multivariate_PC_test_commit.cpp:265-331, which calls
// commit_pair, OZ/src/VPD/linearPC.cpp:253-459).  The two Merkle roots it
// produces are what orion_commit.json carries and what the frame prover's
// transcript absorbs as the commitment binding.
//
// What is committed (OZ/include/truepix/truepix_gkr.h:85-386):
//   n = next_pow2(copies) * 512 coefficients, copy-major: the copy's
//   copy_size samples, a constant at wire const_idx (none when -1), zero
//   elsewhere (the circuit's packing: gray 384 samples and 128 at 384,
//   invert 255 at 384, mask 384 samples and no constant, crop 385 samples and
//   no constant); the mask's real part is splitmix64(seed) % p at the k-th
//   sample position, in order; the masked polynomial is coefficients + mask.  Each polynomial is split into
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
#include "host_sha.hpp"
#include "reduce.hpp"
#include "runtime.hpp"
#include "virgo.hpp"
#include "../../include/truepix_hip.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
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
__global__ void k_or_fill(const uint8_t* __restrict__ in, int copies, int copy_size, int const_idx,
                          u64 const_val, u64 seed, long long col, V2* __restrict__ X)
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
            if (wire < copy_size) {
                const u64 rank = copy * copy_size + wire;
                u64 z = seed + (rank + 1) * kGolden;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
                z ^= z >> 31;
                const u64 mask = fold61(z);
                re = lane < kCols ? fadd((u64)in[copy * copy_size + wire], mask) : mask;
            } else if (wire == const_idx && lane < kCols) {
                re = const_val;
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

// One tree level, nodes lo .. 2 lo - 1 of ntrees trees (2 rows nodes each).
__global__ void __launch_bounds__(256) k_or_level(uint32_t* __restrict__ tree, long long rows,
                                                  long long lo, int ntrees)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= ntrees * lo) return;
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

// ------------------------------------------------------------ opening ----

// The combined codeword of one polynomial (lanes off .. off + 127):
// cc[t] = sum_j r0[j] X[t][off + j], one wave per codeword row.
__global__ void __launch_bounds__(256) k_or_combine(const V2* __restrict__ X, long long cw,
                                                    int off, const V2* __restrict__ r0,
                                                    V2* __restrict__ cc)
{
    const long long row = blockIdx.x * 4LL + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= cw) return;   // whole waves: row is uniform across a wave
    const V2* x = X + (size_t)row * kLanes + off;
    LazyProdSum z;
    z.zero();
    z.mac(mk2(r0[lane].re, r0[lane].im), mk2(x[lane].re, x[lane].im));
    z.mac(mk2(r0[lane + 64].re, r0[lane + 64].im), mk2(x[lane + 64].re, x[lane + 64].im));
    fp2 e[1] = {z.value()};
    wave_sum<1>(e);
    if (lane == 0) {
        V2 o;
        o.re = e[0].re;
        o.im = e[0].im;
        cc[row] = o;
    }
}

// sum_t r1[t] cm[t] over n terms in one workgroup.
__global__ void __launch_bounds__(1024) k_or_answer(const V2* __restrict__ cm,
                                                    const V2* __restrict__ r1, long long n,
                                                    V2* __restrict__ out)
{
    LazyProdSum z;
    z.zero();
    for (long long t = threadIdx.x; t < n; t += blockDim.x)
        z.mac(mk2(r1[t].re, r1[t].im), mk2(cm[t].re, cm[t].im));
    fp2 e[1] = {z.value()};
    block_sum<1>(e);
    if (threadIdx.x == 0) {
        V2 o;
        o.re = e[0].re;
        o.im = e[0].im;
        *out = o;
    }
}

// The queried columns: out[i][j] = X[q[i]][off + j].
__global__ void __launch_bounds__(256) k_or_gather(const V2* __restrict__ X, int off,
                                                   const long long* __restrict__ q, int nq,
                                                   V2* __restrict__ out)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)nq * kCols) return;
    const long long i = id / kCols;
    const int j = (int)(id - i * kCols);
    out[id] = X[(size_t)q[i] * kLanes + off + j];
}

// Leaves of the combined codeword's tree (OZ/src/infrastructure/merkle_tree.cpp:
// hash_single_field_element): SHA-256 of the element (imag, real, little-
// endian) and 48 zero bytes; the zero element past the codeword.
__global__ void __launch_bounds__(256) k_or_cleaves(const V2* __restrict__ cc, long long cw,
                                                    long long rows, uint32_t* __restrict__ tree)
{
    const long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (r >= rows) return;
    V2 x;
    x.re = x.im = 0;
    if (r < cw) x = cc[r];
    uint32_t m[16];
    m[0] = be32((uint32_t)x.im);
    m[1] = be32((uint32_t)(x.im >> 32));
    m[2] = be32((uint32_t)x.re);
    m[3] = be32((uint32_t)(x.re >> 32));
#pragma unroll
    for (int k = 4; k < 16; ++k) m[k] = 0;
    uint32_t out[8];
    sha64(m, out);
#pragma unroll
    for (int k = 0; k < 8; ++k) tree[(size_t)(rows + r) * 8 + k] = out[k];
}

// -------------------------------------------------------------- plans ----

// One column length's code: the graphs drawn as the reference draws them,
// reversed into per-target edge lists, and the in-place row offsets; the
// forward edge lists too, for the code-switching proof's gathers.
struct Plan {
    long long col = 0, cw = 0;
    struct Step {
        long long in_base, out_base, targets;
        int rp_off;   // into rowptr; edges from rowptr values into src / w
    };
    std::vector<Step> down, up;   // C steps by depth, D steps from the deepest up
    DevBuf rowptr, src, w;
    DevBuf fto, fw;               // forward edges of every graph, by depth
    virgo::Code code;             // the graphs as the code-switching proof reads them
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
    // forward edges, degree-major (edge slot e of every left vertex, then
    // slot e + 1, ...), so consecutive left vertices read consecutive words:
    // C[0], D[0], C[1], D[1], ...
    std::vector<int> fto;
    std::vector<V2> fw;
    std::vector<size_t> offC((size_t)depth), offD((size_t)depth);
    for (int d = 0; d < depth; ++d)
        for (int k = 0; k < 2; ++k) {
            const Fwd& g = k ? D[d] : C[d];
            (k ? offD : offC)[(size_t)d] = fto.size();
            for (int e = 0; e < g.deg; ++e)
                for (long long i = 0; i < g.L; ++i) {
                    const size_t at = (size_t)i * g.deg + e;
                    fto.push_back(g.to[at]);
                    fw.push_back(V2{g.w[at].re, g.w[at].im});
                }
        }
    if (fto.empty()) {
        fto.push_back(0);
        fw.push_back(V2{0, 0});
    }
    hipStream_t st = stream();
    if (!P.rowptr.ensure(rp.size() * 4) || !P.src.ensure(src.size() * 4) ||
        !P.w.ensure(ww.size() * sizeof(V2)) || !P.fto.ensure(fto.size() * 4) ||
        !P.fw.ensure(fw.size() * sizeof(V2)))
        return false;
    TP_HIP_OK(hipMemcpyAsync(P.rowptr.ptr, rp.data(), rp.size() * 4, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(P.src.ptr, src.data(), src.size() * 4, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(P.w.ptr, ww.data(), ww.size() * sizeof(V2), hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(P.fto.ptr, fto.data(), fto.size() * 4, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(P.fw.ptr, fw.data(), fw.size() * sizeof(V2), hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    P.code.n = col;
    P.code.cw = P.cw;
    P.code.C.assign((size_t)depth, virgo::Graph());
    P.code.D.assign((size_t)depth, virgo::Graph());
    for (int d = 0; d < depth; ++d)
        for (int k = 0; k < 2; ++k) {
            const Fwd& g = k ? D[d] : C[d];
            virgo::Graph& v = (k ? P.code.D : P.code.C)[(size_t)d];
            v.L = g.L;
            v.R = g.R;
            v.deg = g.deg;
            v.to = P.fto.as<int>() + (k ? offD : offC)[(size_t)d];
            v.w = P.fw.as<fp2>() + (k ? offD : offC)[(size_t)d];
        }
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

DevBuf g_open, g_ctree;   // the opening's vectors and the combined codeword's tree

// The packing of a copy (truepix_gkr.h:107-150; the reference's programs
// read it from TRUEPIX_COPY_SIZE, TRUEPIX_CONST_IDX, TRUEPIX_CONST_VAL).
struct Packing {
    int copy_size = 384, const_idx = 384;
    u64 const_val = 128;
};
Packing g_pack;

// Fills, encodes and hashes both polynomials of a frame and builds both
// trees (g_X, g_tree); the plan and the row count come back.
bool commit_device(int copies, const uint8_t* inputs, uint64_t seed, hipStream_t st,
                   const Plan*& P_out, long long& rows_out)
{
    long long pc = 1;
    while (pc < copies) pc <<= 1;
    const long long n = pc * 512, col = n / kCols, rows = 2 * col;
    Plan* P = plan_for(col);
    if (!P) return false;
    const size_t in_bytes = (size_t)copies * g_pack.copy_size;
    if (!g_in.ensure(in_bytes) || !g_X.ensure((size_t)P->cw * kLanes * sizeof(V2)) ||
        !g_tree.ensure((size_t)2 * 2 * rows * 8 * 4))
        return false;
    TP_HIP_OK(hipMemcpyAsync(g_in.ptr, inputs, in_bytes, hipMemcpyHostToDevice, st));
    k_or_fill<<<grid_for((size_t)col * kLanes, kBlock, 8192), kBlock, 0, st>>>(
        g_in.as<uint8_t>(), copies, g_pack.copy_size, g_pack.const_idx, g_pack.const_val, seed, col,
        g_X.as<V2>());
    TP_HIP_OK(hipGetLastError());
    if (!encode_all(*P, st)) return false;
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
        k_or_level<<<(unsigned)((2 * lo + 255) / 256), 256, 0, st>>>(g_tree.as<uint32_t>(), rows, lo, 2);
        TP_HIP_OK(hipGetLastError());
    }
    if (lo >= 1) {
        k_or_top<<<2, 1024, 0, st>>>(g_tree.as<uint32_t>(), rows, lo);
        TP_HIP_OK(hipGetLastError());
    }
    P_out = P;
    rows_out = rows;
    return true;
}

// The Orion NIZK transcript (OZ/src/truepix/orion_fs.cpp): a 32-byte state,
// each absorbed piece taken 32 bytes at a time as state = SHA-256(state |
// piece bytes, zero-padded to 32); squeezes come from a 32-byte pool,
// SHA-256(state | 0x01 | zeros), the state moving to SHA-256(state | 0x02 |
// zeros) at each refill; an absorb empties the pool.
struct OrionFs {
    uint8_t state[32], pool[32];
    int used = 32;
    // SHA-256 of one 64-byte message: the message block and the constant
    // padding block (0x80, zeros, bit length 512) in one compression call
    static void hash64(const uint8_t b[64], uint8_t out[32])
    {
        alignas(16) uint8_t m[128];
        memcpy(m, b, 64);
        memset(m + 64, 0, 64);
        m[64] = 0x80;
        m[126] = 0x02;   // 512 bits, big-endian
        uint32_t st[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                          0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
        sha256_blocks(st, m, 2);
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 4; ++j) out[4 * k + j] = (uint8_t)(st[k] >> (24 - 8 * j));
    }
    void reset()
    {
        memset(state, 0, 32);
        absorb("truepix/orion-fs/v1", 19);
        used = 32;
    }
    void absorb(const void* data, size_t n)
    {
        const uint8_t* p = static_cast<const uint8_t*>(data);
        while (n > 0) {
            uint8_t block[64] = {0};
            memcpy(block, state, 32);
            const size_t take = n < 32 ? n : 32;
            memcpy(block + 32, p, take);
            hash64(block, state);
            p += take;
            n -= take;
        }
        used = 32;
    }
    void absorb_u64(uint64_t x)
    {
        uint8_t b[8];
        for (int k = 0; k < 8; ++k) b[k] = (uint8_t)(x >> (8 * k));
        absorb(b, 8);
    }
    void absorb_field(u64 re, u64 im)   // real then imag, each its own absorb
    {
        absorb_u64(re);
        absorb_u64(im);
    }
    uint64_t squeeze_u64()
    {
        if (used + 8 > 32) {
            uint8_t block[64] = {0};
            memcpy(block, state, 32);
            block[32] = 0x01;
            hash64(block, pool);
            block[32] = 0x02;
            hash64(block, state);
            used = 0;
        }
        uint64_t v = 0;
        for (int k = 7; k >= 0; --k) v = (v << 8) | pool[used + k];
        used += 8;
        return v;
    }
    int index(int bound) { return bound <= 0 ? 0 : (int)(squeeze_u64() % (uint64_t)bound); }
};

// Little-endian appends in the opening proof's layout.
template <class T>
void put_le(std::vector<uint8_t>& o, T v)
{
    for (size_t k = 0; k < sizeof(T); ++k) o.push_back((uint8_t)((uint64_t)v >> (8 * k)));
}
void put_fe(std::vector<uint8_t>& o, u64 re, u64 im)   // the element's (imag, real) words
{
    put_le<u64>(o, im);
    put_le<u64>(o, re);
}
void put_digest(std::vector<uint8_t>& o, const uint32_t* w)   // the digest's bytes
{
    for (int k = 0; k < 8; ++k)
        for (int b = 3; b >= 0; --b) o.push_back((uint8_t)(w[k] >> (8 * b)));
}

// The sibling nodes every query's path reads, each once, in the order the
// paths first read them (collect_merkle_siblings, OZ/src/truepix/
// orion_proof.cpp:153-174), as (index u32, digest).
void put_siblings(std::vector<uint8_t>& o, const std::vector<uint32_t>& tree, long long leaves,
                  const int64_t* q, int nq)
{
    std::vector<char> seen((size_t)2 * leaves, 0);
    std::vector<uint32_t> idx;
    for (int i = 0; i < nq; ++i) {
        long long v = q[i] + leaves;
        while (v != 1) {
            const long long sib = v ^ 1;
            if (!seen[(size_t)sib]) {
                seen[(size_t)sib] = 1;
                idx.push_back((uint32_t)sib);
            }
            v >>= 1;
        }
    }
    put_le<u64>(o, (u64)idx.size());
    for (uint32_t s : idx) {
        put_le<uint32_t>(o, s);
        put_digest(o, tree.data() + (size_t)s * 8);
    }
}

// OpeningProof version 2 (OZ/include/truepix/orion_proof.h:1-22): magic,
// version, group, N, column_size, msg_len, cw_len, query_count, answer, the
// two roots, the queries, the queried columns, the combined codeword and
// message, then both trees' siblings.
void serialize_opening(std::vector<uint8_t>& o, bool group1, long long n, long long msg,
                       long long cw, V2 ans, const std::vector<uint32_t>& tcom,
                       const std::vector<uint32_t>& tcmb, long long rows, const int64_t* q, int nq,
                       const std::vector<V2>& qc, const std::vector<V2>& cc)
{
    static const char magic[8] = {'O', 'R', 'I', 'O', 'N', 'P', 'V', '1'};
    o.insert(o.end(), magic, magic + 8);
    put_le<uint32_t>(o, 2);
    put_le<uint32_t>(o, group1 ? 1 : 0);
    put_le<u64>(o, (u64)n);
    put_le<u64>(o, (u64)kCols);
    put_le<u64>(o, (u64)msg);
    put_le<u64>(o, (u64)cw);
    put_le<uint32_t>(o, (uint32_t)nq);
    put_fe(o, ans.re, ans.im);
    put_digest(o, tcom.data() + 8);
    put_digest(o, tcmb.data() + 8);
    for (int i = 0; i < nq; ++i) put_le<int64_t>(o, q[i]);
    for (const V2& x : qc) put_fe(o, x.re, x.im);
    for (const V2& x : cc) put_fe(o, x.re, x.im);
    for (long long t = 0; t < msg; ++t) put_fe(o, cc[(size_t)t].re, cc[(size_t)t].im);
    put_siblings(o, tcom, rows, q, nq);
    put_siblings(o, tcmb, rows, q, nq);
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

// The packing every later commit and opening uses: copy_size samples per
// copy (1..512), the constant const_val at wire const_idx (-1: none; outside
// the samples, below 512).  Validated as make_layout does
// (truepix_gkr.h:107-150).  0 / -1.
int tp_orion_set_layout(int copy_size, int const_idx, int const_val)
{
    clear_error();
    if (copy_size < 1 || copy_size > 512 || const_idx < -1 || const_idx >= 512 ||
        (const_idx >= 0 && const_idx < copy_size) || const_val < 0)
        return fail("tp_orion_set_layout: copy_size=%d const_idx=%d const_val=%d is not a packing "
                    "(samples within 512, the constant outside them)",
                    copy_size, const_idx, const_val), -1;
    std::lock_guard<std::mutex> lk(g_mu);
    g_pack.copy_size = copy_size;
    g_pack.const_idx = const_idx;
    g_pack.const_val = (u64)const_val;
    return 0;
}

// The signer's Orion commitment of one frame: inputs [copies][copy_size] sample
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
    hipStream_t st = stream();
    const Plan* P = nullptr;
    long long rows = 0;
    TP_HIP_OK(hipEventRecord(g_ev[0], st));
    if (!commit_device(copies, inputs, seed, st, P, rows)) return -1;
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

// The opening of one committed polynomial (OZ/src/VPD/linearPC.cpp:700-966,
// tensor_product_protocol, with r0 and r1 from the point r as open_and_verify
// forms them, :1013-1034): the masked polynomial when use_group1 (opened on
// the even columns of the code), else the mask.  r: n_r = lg n challenges
// (re, im); col_q: the n_queries code positions the transcript drew.  Writes
// the opening proof in the reference's binary layout (ORIONPV1 version 2,
// OZ/include/truepix/orion_proof.h:1-22) to *proof (free with tp_free) and
// the claimed value to answer.  The Virgo code-switching proof that follows
// in the reference is not part of it.  0 / -1.
virgo::Stats g_virgo_stats;

static int open_impl(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                     const uint64_t* r, int n_r, const int64_t* col_q_in, int n_queries, bool nizk,
                     uint8_t** proof, size_t* proof_len, uint64_t answer[2],
                     std::vector<u64>* virgo_words = nullptr)
{
    clear_error();
    if (!device_ready()) return -1;
    if (copies < 1 || copies > (1 << 16) || !inputs || !r || !proof || !proof_len || !answer ||
        (!nizk && (n_queries < 0 || (n_queries && !col_q_in))))
        return fail("tp_orion_open: bad arguments (copies=%d)", copies), -1;
    long long pc = 1;
    while (pc < copies) pc <<= 1;
    const long long n = pc * 512, msg = n / kCols;
    int lg = 0;
    while ((1LL << lg) < n) ++lg;
    if (n_r != lg) return fail("tp_orion_open: %d challenges for n = 2^%d", n_r, lg), -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!setup()) return -1;
    hipStream_t st = stream();
    const Plan* P = nullptr;
    long long rows = 0;
    if (!commit_device(copies, inputs, seed, st, P, rows)) return -1;
    const long long cw = P->cw;
    std::vector<int64_t> drawn;
    std::vector<long long> vq;   // the code-switching proof's positions
    const int64_t* col_q = col_q_in;
    OrionFs fs;
    // r0 over r[0..7), r1 over the rest: eq tensors, the first challenge on
    // the highest bit of the index (dfs in linearPC.cpp:996-1007)
    auto tensor = [&](const uint64_t* rr, int k, std::vector<fp2>& out) {
        out.assign((size_t)1 << k, one2());
        for (int d = 0; d < k; ++d) {
            const fp2 x = mk2(rr[2 * d], rr[2 * d + 1]);
            const fp2 nx = sub2(one2(), x);
            const size_t half = (size_t)1 << (k - 1 - d);
            for (size_t i = 0; i < out.size(); ++i) out[i] = mul2(out[i], (i & half) ? x : nx);
        }
    };
    std::vector<fp2> r0, r1;
    int lc = 0;
    while ((1 << lc) < kCols) ++lc;
    tensor(r, lc, r0);
    tensor(r + 2 * lc, lg - lc, r1);
    if (nizk) {
        // the opening's public inputs, then the queries from the opened
        // group (even code positions for the masked polynomial, odd for the
        // mask), as tensor_product_protocol draws them (linearPC.cpp:711-758)
        uint32_t root[8];
        TP_HIP_OK(hipMemcpyAsync(root, g_tree.as<uint32_t>() + ((size_t)(use_group1 ? 0 : 1) * 2 * rows + 1) * 8,
                                 32, hipMemcpyDeviceToHost, st));
        TP_HIP_OK(hipStreamSynchronize(st));
        uint8_t rb[32];
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 4; ++b) rb[4 * k + b] = (uint8_t)(root[k] >> (24 - 8 * b));
        fs.reset();
        fs.absorb("tensor-open", 11);
        fs.absorb(rb, 32);
        fs.absorb_u64((uint64_t)n);
        fs.absorb_u64(use_group1 ? 1 : 0);
        for (const fp2& x : r0) fs.absorb_field(x.re, x.im);
        for (const fp2& x : r1) fs.absorb_field(x.re, x.im);
        n_queries = (int)(-128.0 / log2(1.0 - 0.07));   // the reference's query_count
        const long long gsize = use_group1 ? (cw + 1) / 2 : cw / 2;
        drawn.resize((size_t)n_queries);
        for (int i = 0; i < n_queries; ++i)
            drawn[(size_t)i] = 2LL * fs.index((int)gsize) + (use_group1 ? 0 : 1);
        col_q = drawn.data();
        // then the code-switching proof's positions over the whole codeword
        // (linearPC.cpp:879-881)
        vq.resize((size_t)n_queries);
        for (int i = 0; i < n_queries; ++i) vq[(size_t)i] = fs.index((int)cw);
    }
    for (int i = 0; i < n_queries; ++i)
        if (col_q[i] < 0 || col_q[i] >= cw)
            return fail("tp_orion_open: query %d at %lld outside the codeword (%lld)", i,
                        (long long)col_q[i], cw), -1;
    const int off = use_group1 ? 0 : kCols;
    const size_t nq = (size_t)n_queries;
    if (!g_open.ensure(((size_t)kCols + (size_t)msg + (size_t)cw + 1 + nq * kCols) * sizeof(V2) +
                       nq * sizeof(long long)) ||
        !g_ctree.ensure((size_t)2 * rows * 8 * 4))
        return -1;
    V2* d_r0 = g_open.as<V2>();
    V2* d_r1 = d_r0 + kCols;
    V2* d_cc = d_r1 + msg;
    V2* d_ans = d_cc + cw;
    V2* d_qc = d_ans + 1;
    long long* d_q = reinterpret_cast<long long*>(d_qc + nq * kCols);
    std::vector<V2> h_r((size_t)kCols + msg);
    for (int j = 0; j < kCols; ++j) h_r[j].re = r0[j].re, h_r[j].im = r0[j].im;
    for (long long t = 0; t < msg; ++t) h_r[kCols + t].re = r1[t].re, h_r[kCols + t].im = r1[t].im;
    TP_HIP_OK(hipMemcpyAsync(d_r0, h_r.data(), h_r.size() * sizeof(V2), hipMemcpyHostToDevice, st));
    if (nq) {
        std::vector<long long> hq(col_q, col_q + nq);
        TP_HIP_OK(hipMemcpyAsync(d_q, hq.data(), nq * sizeof(long long), hipMemcpyHostToDevice, st));
        TP_HIP_OK(hipStreamSynchronize(st));   // hq leaves scope
    }
    k_or_combine<<<(unsigned)((cw + 3) / 4), 256, 0, st>>>(g_X.as<V2>(), cw, off, d_r0, d_cc);
    TP_HIP_OK(hipGetLastError());
    k_or_answer<<<1, 1024, 0, st>>>(d_cc, d_r1, msg, d_ans);   // the code is systematic
    TP_HIP_OK(hipGetLastError());
    if (nq) {
        k_or_gather<<<(unsigned)((nq * kCols + 255) / 256), 256, 0, st>>>(g_X.as<V2>(), off, d_q,
                                                                       n_queries, d_qc);
        TP_HIP_OK(hipGetLastError());
    }
    uint32_t* ct = g_ctree.as<uint32_t>();
    k_or_cleaves<<<(unsigned)((rows + 255) / 256), 256, 0, st>>>(d_cc, cw, rows, ct);
    TP_HIP_OK(hipGetLastError());
    long long lo = rows / 2;
    for (; lo > 1024; lo >>= 1) {
        k_or_level<<<(unsigned)((lo + 255) / 256), 256, 0, st>>>(ct, rows, lo, 1);
        TP_HIP_OK(hipGetLastError());
    }
    if (lo >= 1) {
        k_or_top<<<1, 1024, 0, st>>>(ct, rows, lo);
        TP_HIP_OK(hipGetLastError());
    }
    // results: the combined codeword (its first msg entries are the combined
    // message), the answer, the queried columns and both trees
    std::vector<V2> cc((size_t)cw), qc(nq * kCols);
    V2 ans;
    std::vector<uint32_t> tcom((size_t)2 * rows * 8), tcmb((size_t)2 * rows * 8);
    TP_HIP_OK(hipMemcpyAsync(cc.data(), d_cc, cc.size() * sizeof(V2), hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipMemcpyAsync(&ans, d_ans, sizeof(V2), hipMemcpyDeviceToHost, st));
    if (nq)
        TP_HIP_OK(hipMemcpyAsync(qc.data(), d_qc, qc.size() * sizeof(V2), hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipMemcpyAsync(tcom.data(), g_tree.as<uint32_t>() + (size_t)(use_group1 ? 0 : 1) * 2 * rows * 8,
                             tcom.size() * 4, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipMemcpyAsync(tcmb.data(), ct, tcmb.size() * 4, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    answer[0] = ans.re;
    answer[1] = ans.im;
    *proof = nullptr;
    *proof_len = 0;
    if (nizk) {
        // "combined-mt" and the combined root, then the Virgo proof that the
        // queried entries are the code of the combined message; the proof is
        // written only if it checks (linearPC.cpp:883-928)
        uint8_t cb[32];
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 4; ++b) cb[4 * k + b] = (uint8_t)(tcmb[8 + k] >> (24 - 8 * b));
        fs.absorb("combined-mt", 11);
        fs.absorb(cb, 32);
        std::vector<u64> words;
        bool accepted = false;
        if (!virgo::prove(P->code, reinterpret_cast<const fp2*>(d_cc), vq,
                          [&fs]() { return fs.squeeze_u64(); }, st, words, accepted, &g_virgo_stats))
            return -1;
        if (virgo_words) virgo_words->swap(words);
        if (!accepted) return fail("tp_orion_open_nizk: the code-switching proof does not check"), -1;
    }
    std::vector<uint8_t> out;
    serialize_opening(out, use_group1 != 0, n, msg, cw, ans, tcom, tcmb, rows, col_q, n_queries,
                      qc, cc);
    uint8_t* buf = static_cast<uint8_t*>(malloc(out.size()));
    if (!buf) return fail("tp_orion_open: out of host memory"), -1;
    memcpy(buf, out.data(), out.size());
    *proof = buf;
    *proof_len = out.size();
    return 0;
}

int tp_orion_open(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                  const uint64_t* r, int n_r, const int64_t* col_q, int n_queries,
                  uint8_t** proof, size_t* proof_len, uint64_t answer[2])
{
    return open_impl(copies, inputs, seed, use_group1, r, n_r, col_q, n_queries, false, proof,
                     proof_len, answer);
}

// tp_orion_open with the queries drawn as the reference's NIZK mode draws
// them (OZ/src/truepix/orion_fs.cpp: the transcript reset, the label
// "tensor-open", the opened root, N, the group flag, r0 and r1, then 1222
// positions from the opened group; linearPC.cpp:711-758).  0 / -1.
int tp_orion_open_nizk(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                       const uint64_t* r, int n_r, uint8_t** proof, size_t* proof_len,
                       uint64_t answer[2])
{
    return open_impl(copies, inputs, seed, use_group1, r, n_r, nullptr, 0, true, proof, proof_len,
                     answer);
}

// tp_orion_open_nizk, with the words of its code-switching transcript.
int tp_orion_open_nizk_virgo(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                             const uint64_t* r, int n_r, uint8_t** proof, size_t* proof_len,
                             uint64_t answer[2], uint64_t* words, long long cap, long long* n_words)
{
    if (cap < 0 || (cap && !words) || !n_words)
        return fail("tp_orion_open_nizk_virgo: bad transcript buffer"), -1;
    std::vector<u64> w;
    const int rc = open_impl(copies, inputs, seed, use_group1, r, n_r, nullptr, 0, true, proof,
                             proof_len, answer, &w);
    *n_words = (long long)w.size();
    if (cap && !w.empty())
        memcpy(words, w.data(), (size_t)std::min<long long>(cap, (long long)w.size()) * sizeof(u64));
    return rc;
}

// Device time (ms), launches and sumcheck layers of the last code-switching
// proof.
int tp_orion_virgo_stats(double* ms, int* launches, int* layers)
{
    if (ms) *ms = g_virgo_stats.ms;
    if (launches) *launches = g_virgo_stats.launches;
    if (layers) *layers = g_virgo_stats.layers;
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
