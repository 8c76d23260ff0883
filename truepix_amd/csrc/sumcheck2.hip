// sumcheck2.hip - degree-2 sumcheck over two resident multilinear tables: the
// round-sum + fold-with-challenge pattern of OZ/src/VPD/fft_circuit_GKR.cpp:171-209
// (sumcheck_phase1_update, addV term zero), HBM-streaming on CDNA4.
//
// Round k polynomial: s_k(X) = sum_i (F0 + X dF)(G0 + X dG) = a X^2 + b X + c,
// variables bound LSB-first, pairs (2i, 2i+1).  The reference stores linear
// polynomials and folds them in place; here a table is kept as values and each
// round's launch fuses the fold of round k-1 with the round-k sum:
//   read T_{k-1} (m values per table), write T_k (m/2), accumulate
//   E0 = sum F0 G0, E1 = sum F1 G1, EA = sum dF dG  ->  a = EA, c = E0,
//   b = E1 - EA - E0.
// Bytes per fused launch: 32 m read + 16 m written (both tables).  Tables are
// AoS {re, im} so every element is one 128-bit load; a thread owns one output
// pair (64 B per table).
//
// Rounds 0 and 1 share one pass over the input (nv >= 3).  T_1 = A + r_0 B is
// linear in r_0, so round 1's sums are quadratics in r_0 whose coefficients
// are sums over the input:
//   E0_1(r) = S00 + r (S11 - S00 - SAA) + r^2 SAA   (even round-0 pairs)
//   E1_1(r) = S22 + r (S33 - S22 - SBB) + r^2 SBB   (odd round-0 pairs)
//   EA_1(r) = X0  + r (X1  - X0  - X2 ) + r^2 X2
// with, per input quad f0..f3 / g0..g3 and d = f1 - f0, d' = f3 - f2 (e, e'
// for g): S00 = sum f0 g0, S11 = sum f1 g1, SAA = sum d e, S22, S33, SBB the
// same over (f2, f3), X0 = sum (f2 - f0)(g2 - g0), X1 = sum (f3 - f1)(g3 - g1),
// X2 = sum (d' - d)(e' - e).  Round 0's sums are S00 + S22, S11 + S33 and
// SAA + SBB.  The first pass reads the input once (32 n bytes) and yields both
// polynomials; the second pass folds the input by r_0 and r_1 straight into
// T_2 (read 32 n, write 8 n), so T_1 is never written or read.
//
// Every challenge is a SHA-256 draw of the host transcript, so each round
// goes to the host and back.  The schedule, by the size of the round's input:
//   - above kMidGroups * kTailIn entries (streaming rounds): a fused launch
//     whose last workgroup writes the sums, tagged with the round, straight
//     into mapped pinned host memory (ticket_publish), the host polling them,
//     running the transcript and launching the next round with the challenge
//     as a kernel argument;
//   - down to kMidGroups * kMidEnd entries (k_sc2_mid): one launch of kMidGroups
//     workgroups, each folding its own slice in LDS, the sums gathered on the
//     device and the challenges coming back through mapped memory;
//   - the rest (k_sc2_tail_host): one workgroup, the same exchange without
//     the gather.
// A launch plus its dependent-launch gap costs ~6 us here, more than a
// small round's work; the last two phases pay it once for all their rounds.
// They also take two rounds per host exchange, as the first pass does: the
// nine sums of T_k's quads give round k's sums and round k+1's as
// quadratics in r_k, so the host draws r_k and r_{k+1} from one exchange and
// the next exchange folds by both (fold_quad).  Their tables are small, so
// the exchange is shaped for latency (exchange_work): each thread folds its
// own entries, then wave s adds sum s alone.
// A one-lane SHA-256 on the device was measured at 4.8 us per compression
// (~26 us per transcript round), against ~3.4 us for the host round trip, so
// the transcript stays on the host.
#include "blocks.hpp"
#include "fp61.hpp"
#include "host_sha.hpp"
#include "reduce.hpp"
#include "relay.hpp"
#include "runtime.hpp"
#include "transcript.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <stdint.h>
#include <thread>

using namespace tp;

namespace {

struct alignas(16) V2 { u64 re, im; };

__device__ __forceinline__ fp2 ldv(const V2* p) { const V2 v = *p; return mk2(v.re, v.im); }
__device__ __forceinline__ void stv(V2* p, fp2 x) { V2 v; v.re = x.re; v.im = x.im; *p = v; }


// A value every lane holds alike, moved to scalar registers.
__device__ __forceinline__ u64 uniform64(u64 x)
{
    return ((u64)(u32)__builtin_amdgcn_readfirstlane((int)(x >> 32)) << 32) |
           (u32)__builtin_amdgcn_readfirstlane((int)(u32)x);
}
__device__ __forceinline__ fp2 uniform2(fp2 x) { return mk2(uniform64(x.re), uniform64(x.im)); }

// Arguments of a launch that publishes its own round sums (ticket_publish).
struct TicketPub {
    u64* slots;          // [gridDim.x][N][2] tagged words; all ones between launches
    unsigned* ticket;    // 0 between launches
    fp2* host_vals;      // mapped: the N totals
    unsigned* host_seq;  // mapped: then seq, or kSeqAbort if a wait expired
    unsigned seq;
    long long spin_ticks;
    // A launch queued ahead of its transcript step takes its n_ans challenge
    // words from the mapped answer slot, tagged with sequence number `want`
    // (launch_challenges); n_ans = 0: they are kernel arguments.
    const u64* answer;
    unsigned want;
    int n_ans;
    u64* relay;          // device: workgroup 0's copy of the answer for the others; all ones when idle
};

// The challenge words of a streaming launch: the arguments in w already, or,
// for a launch queued ahead of its transcript step, n_ans words the host
// writes into the mapped answer slot tagged (want mod 8) << 61 (relay.hpp).
// Only workgroup 0 reads the host's slot (a read of host memory crosses
// PCIe; a thousand workgroups polling it slowed the host's own accesses):
// it copies the tagged words into the device relay, which the other
// workgroups poll (agent scope).  The launch's last workgroup empties the
// relay (ticket_publish), so the next launch starts from an idle one.  The
// words are shared through LDS.  False in every thread if a wait expired or
// the host marked the prove failed (kSeqHostAbort in the sequence word): the
// launch then ends without taking a ticket or publishing.
constexpr int kMaxAns = 4;   // answer words a launch may take (two challenges)

template <int W>
__device__ __forceinline__ bool launch_challenges(const TicketPub& a, u64 (&w)[W])
{
    static_assert(W <= kMaxAns, "answer words");
    if (a.n_ans == 0) return true;
    __shared__ u64 s_w[kMaxAns];
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        const u64 tag = (u64)(a.want & 7) << 61;
        const bool first = blockIdx.x == 0;
        const u64* src = first ? a.answer : a.relay;
        const long long t0 = wall_clock64();
        int ok = 0;
        for (unsigned it = 0;; ++it) {
            u64 v[kMaxAns];
            bool all = true;
            for (int i = 0; i < a.n_ans; ++i) {
                v[i] = first ? __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
                             : __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                all = all && (v[i] & ~P) == tag && (v[i] & P) != P;
            }
            if (all) {
                for (int i = 0; i < a.n_ans; ++i) {
                    s_w[i] = v[i] & P;
                    if (first)
                        __hip_atomic_store(a.relay + i, v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                ok = 1;
                break;
            }
            if ((first || (it & 255) == 255) &&
                __hip_atomic_load(a.host_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) ==
                    kSeqHostAbort)
                break;
            if (wall_clock64() - t0 > a.spin_ticks) {
                __hip_atomic_store(a.host_seq, kSeqAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        s_ok = ok;
    }
    __syncthreads();
    if (!s_ok) return false;
    // uniform values: scalar registers for the launch's products
    for (int i = 0; i < W; ++i)
        if (i < a.n_ans) w[i] = uniform64(s_w[i]);
    return true;
}

// Slots and tickets of the launches that publish their own sums: the last
// workgroup of a launch gathers every workgroup's slot, so a launch runs at
// most kMaxPubGroups workgroups.
constexpr unsigned kMaxPubGroups = 1024;
constexpr size_t kTicketBytes = 256;

// Adds `count` slots of N tagged sums at `base` (slot j at base + 2 N j) in
// this workgroup once every word carries `tag` (spinning, bounded by the
// wall clock, on any still in flight), then clears them to all ones, so no
// word of an earlier round, whatever its tag, is ever read as a later one's.
// The totals land in s_tot; the return value (every thread) is 0 if a wait
// expired.
// With EXT the per-thread partials go to `ext` (LDS of at least 1024 Fp2
// that the caller no longer uses) instead of a buffer of their own.
template <int N, bool EXT = false>
__device__ __forceinline__ int gather_slots(u64* base, unsigned count, u64 tag, long long spin_ticks,
                                            fp2 (&s_tot)[N], fp2* ext = nullptr)
{
    constexpr unsigned kRow = 1024 / N;
    fp2* s_g;   // [N][kRow]
    if constexpr (EXT) {
        s_g = ext;
    } else {
        __shared__ fp2 own[N * kRow];
        s_g = own;
    }
    const unsigned t = threadIdx.x;
    // thread t adds sum k = t mod N over slots g, g + G, ...
    const unsigned G = blockDim.x / N, k = t % N, g = t / N;
    const long long t0 = wall_clock64();
    int ok = 1;
    fp2 acc = zero2();
    auto word = [&](unsigned j, int c) -> u64 {   // slot j, coordinate c of sum k
        return j < count ? __hip_atomic_load(base + ((size_t)j * N + k) * 2 + c, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)
                         : tag;                    // past the last slot: 0 under this tag
    };
    if (g < G) {
        // four slots' words in flight at once (eight held the launches'
        // registers above their loops' needs): one or two rounds of loads
        // cover every launch here (count <= 8 G)
        for (unsigned j0 = g; j0 < count; j0 += 4 * G) {
            u64 w[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) w[i] = word(j0 + (i >> 1) * G, i & 1);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                while (ok && ((w[i] & ~P) != tag || (w[i] & P) == P)) {
                    if (wall_clock64() - t0 > spin_ticks) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    w[i] = word(j0 + (i >> 1) * G, i & 1);
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) acc = add2(acc, mk2(w[2 * i] & P, w[2 * i + 1] & P));
        }
    }
    ok = __syncthreads_and(ok);   // every word is read before the slots are cleared
    for (unsigned w = t; w < count * 2 * N; w += blockDim.x)
        __hip_atomic_store(base + w, kSlotEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g < G) s_g[k * kRow + g] = acc;
    __syncthreads();
    const unsigned wid = t >> 6, lane = t & 63, nwv = blockDim.x >> 6;
    // wave w adds the G partials of sums w, w + nwv, ... (any workgroup size)
    for (unsigned q = wid; q < (unsigned)N; q += nwv) {
        fp2 sv[1] = {zero2()};
        for (unsigned x = lane; x < G; x += 64) sv[0] = add2(sv[0], s_g[q * kRow + x]);
        wave_sum<1>(sv);
        if (lane == 0) s_tot[q] = sv[0];
    }
    __syncthreads();
    return ok;
}

// Thread 0 stores this workgroup's N sums to slot `slot` as tagged words,
// value | (seq mod 8) << 61 (values are canonical, below 2^61), and takes a
// ticket from *ticket; true (every thread) in the workgroup that took the
// last of `of` tickets.
template <int N>
__device__ __forceinline__ bool slot_and_ticket(const fp2 (&v)[N], u64* slot, unsigned* ticket,
                                                unsigned of, u64 tag)
{
    __shared__ int s_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            __hip_atomic_store(slot + 2 * k, v[k].re | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(slot + 2 * k + 1, v[k].im | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_last = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == of - 1;
    }
    __syncthreads();
    return s_last != 0;
}

// End of a streaming launch, in place of a separate publish launch (which
// cost its own launch and a dependent-launch gap per round).  Each
// workgroup's N sums go to its slot and it takes a ticket (slot_and_ticket);
// the workgroup taking the last ticket adds the slots (gather_slots), resets
// the ticket for the next launch and publishes the totals as tagged words
// (publish_tagged).
template <int N, bool EXT = false>
__device__ __forceinline__ void ticket_publish(fp2 (&v)[N], const TicketPub& a, fp2* ext = nullptr)
{
    __shared__ fp2 s_tot[N];
    block_sum<N>(v);
    const unsigned nb = gridDim.x;
    const u64 tag = (u64)(a.seq & 7) << 61;
    if (!slot_and_ticket<N>(v, a.slots + (size_t)blockIdx.x * 2 * N, a.ticket, nb, tag)) return;
    const int ok = gather_slots<N, EXT>(a.slots, nb, tag, a.spin_ticks, s_tot, ext);
    // every workgroup took its ticket after its challenges: the relay is idle
    // again (the first pass of a prove, which waits for none, clears what an
    // earlier prove may have left)
    const unsigned t = threadIdx.x;
    if (a.relay && t < (unsigned)kMaxAns)
        __hip_atomic_store(a.relay + t, kSlotEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t == 0) {
        __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fp2 tot[N];
#pragma unroll
        for (int q = 0; q < N; ++q) tot[q] = s_tot[q];
        if (ok)
            publish_tagged<N>(a.host_vals, tot, a.seq);
        else
            __hip_atomic_store(a.host_seq, kSeqAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ticket_publish for a workgroup whose N totals thread 0 holds already.
template <int N>
__device__ __forceinline__ void ticket_publish_wg(const fp2 (&v)[N], const TicketPub& a)
{
    __shared__ fp2 s_tot[N];
    const unsigned nb = gridDim.x;
    const u64 tag = (u64)(a.seq & 7) << 61;
    if (!slot_and_ticket<N>(v, a.slots + (size_t)blockIdx.x * 2 * N, a.ticket, nb, tag)) return;
    const int ok = gather_slots<N>(a.slots, nb, tag, a.spin_ticks, s_tot);
    const unsigned t = threadIdx.x;
    if (a.relay && t < (unsigned)kMaxAns)
        __hip_atomic_store(a.relay + t, kSlotEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t == 0) {
        __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fp2 tot[N];
#pragma unroll
        for (int q = 0; q < N; ++q) tot[q] = s_tot[q];
        if (ok)
            publish_tagged<N>(a.host_vals, tot, a.seq);
        else
            __hip_atomic_store(a.host_seq, kSeqAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The totals of N lazy sums published by ticket_publish.
template <int N>
__device__ __forceinline__ void lazy_publish(const LazyProdSum (&z)[N], const TicketPub& a)
{
    fp2 e[N];
#pragma unroll
    for (int k = 0; k < N; ++k) e[k] = z[k].value();
    ticket_publish<N>(e, a);
}

// Round sums of one pair of T_k, lazily: E0 += f0 g0, E1 += f1 g1,
// EA += (f1 - f0)(g1 - g0).
__device__ __forceinline__ void accum_pair(LazyProdSum (&z)[3], fp2 f0, fp2 f1, fp2 g0, fp2 g1)
{
    z[0].mac(f0, g0);
    z[1].mac(f1, g1);
    z[2].mac(sub2(f1, f0), sub2(g1, g0));
}

template <int N>
__device__ __forceinline__ void lazy_zero(LazyProdSum (&z)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k) z[k].zero();
}

// The workgroup's totals of N lazy sums into partials[block * N + k].
template <int N>
__device__ __forceinline__ void lazy_partial(const LazyProdSum (&z)[N], fp2* partials)
{
    fp2 e[N];
#pragma unroll
    for (int k = 0; k < N; ++k) e[k] = z[k].value();
    block_partial<N>(e, partials);
}

// Round 0: no fold, sum over the pairs of the input tables; a thread owns two
// adjacent pairs (64 contiguous bytes per table), as the fused rounds do.
__global__ void __launch_bounds__(256) k_sc2_first(const V2* __restrict__ f,
                                                   const V2* __restrict__ g, size_t n_pairs,
                                                   fp2* partials)
{
    LazyProdSum z[3];
    lazy_zero(z);
    const size_t n_quads = n_pairs / 2;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n_quads;
         q += (size_t)gridDim.x * blockDim.x) {
        const V2* fp = f + 4 * q;
        const V2* gp = g + 4 * q;
        const fp2 fa = ldv(fp), fb = ldv(fp + 1), fc = ldv(fp + 2), fd = ldv(fp + 3);
        const fp2 ga = ldv(gp), gb = ldv(gp + 1), gc = ldv(gp + 2), gd = ldv(gp + 3);
        accum_pair(z, fa, fb, ga, gb);
        accum_pair(z, fc, fd, gc, gd);
    }
    if (n_pairs & 1) {   // one pair (n = 2) left over
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const size_t p = n_pairs - 1;
            accum_pair(z, ldv(f + 2 * p), ldv(f + 2 * p + 1), ldv(g + 2 * p), ldv(g + 2 * p + 1));
        }
    }
    lazy_partial<3>(z, partials);
}

// Rounds 0 and 1 in one pass (see the header): per quad the 9 sums S00, S11,
// SAA, S22, S33, SBB, X0, X1, X2 (9 products per quad against 6 for round 0
// alone), each kept lazily.
__device__ __forceinline__ void accum_quad01(LazyProdSum (&z)[9], const V2 (&f)[4], const V2 (&g)[4])
{
    const fp2 f0 = mk2(f[0].re, f[0].im), f1 = mk2(f[1].re, f[1].im);
    const fp2 f2 = mk2(f[2].re, f[2].im), f3 = mk2(f[3].re, f[3].im);
    const fp2 g0 = mk2(g[0].re, g[0].im), g1 = mk2(g[1].re, g[1].im);
    const fp2 g2 = mk2(g[2].re, g[2].im), g3 = mk2(g[3].re, g[3].im);
    const fp2 d = sub2(f1, f0), dd = sub2(f3, f2), c = sub2(g1, g0), cc = sub2(g3, g2);
    z[0].mac(f0, g0);
    z[1].mac(f1, g1);
    z[2].mac(d, c);
    z[3].mac(f2, g2);
    z[4].mac(f3, g3);
    z[5].mac(dd, cc);
    z[6].mac(sub2(f2, f0), sub2(g2, g0));
    z[7].mac(sub2(f3, f1), sub2(g3, g1));
    z[8].mac(sub2(dd, d), sub2(cc, c));
}

// 116 VGPRs hold four waves per SIMD, so one 16-wave workgroup per CU (256
// workgroups of 1024 threads) is the same occupancy as more, smaller ones,
// with an eighth of the partial sums for the publish launch to add.  A
// prefetch of the next quad (768 threads, 165 VGPRs) and other schedules of
// the nine products measured the same: at four waves per SIMD the pass
// overlaps its loads with its products only in part.
constexpr int kFirst01Threads = 1024;
constexpr unsigned kFirst01Groups = 256;

__global__ void __launch_bounds__(kFirst01Threads) k_sc2_first01(const V2* __restrict__ f,
                                                     const V2* __restrict__ g, size_t n_quads,
                                                     TicketPub pub)
{
    LazyProdSum z[9];
    lazy_zero(z);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n_quads; q += stride) {
        V2 cf[4], cg[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cf[k] = f[4 * q + k];
            cg[k] = g[4 * q + k];
        }
        accum_quad01(z, cf, cg);
    }
    lazy_publish<9>(z, pub);
}

// Rounds 0 and 1 by int8 matrix cores.  The nine sums are exact integer
// bilinear forms over the quads before any reduction mod p, so they can be
// taken on MFMA with the bytes of each value as its limbs:
//   sum_q x_q y_q = sum_{i,j} 2^{8(i+j)} sum_q x_q[i] y_q[j]
// (x_q[i]: byte i of the 64-bit word), and the inner sums over the quads are
// the K dimension of v_mfma_i32_16x16x64_i8.  Per wave and chunk of 64
// quads, for every pair (a, b) of quad entries (f_a, g_b, 16 bytes each: re
// then im) one MFMA adds the 16 x 16 byte products summed over the chunk:
//   A[i][k] = byte i of f_a of quad k,  B[k][j] = byte j of g_b of quad k.
// MFMA operands hold 16 consecutive k of one row per lane, while a load
// gives a lane 16 consecutive bytes of one quad; ds_read_b64_tr_b8 turns one
// into the other: within 16 lanes, lane 8h + j receives byte j of the eight
// 8-byte words that lanes 2b + h address (b = 0..7), so pointing lanes 2b and
// 2b + 1 at the re and im words of one quad gives a lane its row for eight
// quads per read.  A and B are read with the same lane pattern, so their k
// pair up whatever the hardware's k order (checked on the device: C[i][j] =
// sum over (lane group, byte) of A[16G + i][.] B[16G + j][.], lane l holding
// C[4 (l / 16) + r][l % 16]).  The MFMA takes signed bytes, so every byte is
// stored as u = b - 128 (b xor 0x80) and the offset restored at the end:
//   sum_q X Y = sum_q X_u Y_u + 128 K1 (sum_q Y_u + sum_q X_u) + 128^2 K1^2 N
// per pair of words, K1 = sum_{i<8} 2^{8i}, N the quads, X_u = sum_i u_i 2^{8i}
// (sum_q X_u from the bytes' column sums, v_dot4 against ones).  The 16
// integer tiles of a workgroup are added in LDS (|entry| < 2^31 for up to
// 2^24 quads), weighted by 2^{8(i+j)} mod p, and combined into the nine
// sums: S00 = P00, SAA = P11 - P10 - P01 + P00, ..., X2 = sum of s_a s_b P_ab
// with s = (1, -1, -1, 1), P_ab = sum_q f_a g_b in Fp2.  Loads are coalesced
// (a wave instruction covers 1 KB of the chunk) and non-temporal, and each
// wave has the next chunk's loads in flight while it stages, transposes and
// multiplies the current one.  About 100 VALU instructions per quad and
// lane instead of the VALU pass's 690.  The same kernel with temporal loads
// ran 129 us at 2 x 2^24 against 100 us non-temporal; round 2's pass after
// it then ran 117 rather than 105 us (less of the tables left in the
// memory-side cache), 217 against 234 us for the two.
constexpr int kMfThreads = 768;                    // 12 waves: ~156 VGPRs with the prefetch
constexpr int kMfWaves = kMfThreads / 64;
constexpr unsigned kMfGroups = 256;
constexpr int kMfStage = 8192;                     // LDS per wave: a chunk of f, then of g
constexpr size_t kMfLds = (size_t)kMfWaves * kMfStage;
constexpr size_t kMfMaxQuads = (size_t)1 << 24;    // int32 tile sums stay exact up to here
typedef int mf_v4i __attribute__((ext_vector_type(4)));
typedef int mf_v2i __attribute__((ext_vector_type(2)));
typedef unsigned mf_v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mf_v2i lds_tr8(const uint8_t* p)
{
    return __builtin_amdgcn_ds_read_tr8_b64_v2i32((__attribute__((address_space(3))) mf_v2i*)p);
}

// A signed int32 as its residue mod p.
__device__ __forceinline__ u64 mod_i32(int v) { return v < 0 ? P - (u64)(-(long long)v) : (u64)v; }

// x 2^e (mod p) for canonical x and 0 <= e <= 60, as a value below
// 1.5 x 2^61: x 2^e = (x >> (61 - e)) 2^61 + ((x << e) mod 2^61), and
// 2^61 = 1 (mod p) - a rotation of the 61-bit word, no multiply.  (In the
// block pass's closing weights it replaced a multiply per tile entry; the
// pass measured the same, its closing being a small part of it.)
__device__ __forceinline__ u64 rot61(u64 x, int e) { return ((x << e) & P) + (x >> (61 - e)); }

__global__ void __launch_bounds__(kMfThreads) k_sc2_first01_mfma(const V2* __restrict__ f,
                                                                const V2* __restrict__ g,
                                                                size_t n_quads, TicketPub pub)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t mf_lds[];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    uint8_t* stage = mf_lds + wid * kMfStage;
    mf_v4i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mf_v4i{0, 0, 0, 0};
    int ucol[4] = {0, 0, 0, 0}, vcol[4] = {0, 0, 0, 0};   // column sums of the u bytes
    int nq = 0;
    // transposed-read addresses: lane 16 G + m reads word m & 1 of quads
    // 16 G + m / 2 and 16 G + 8 + m / 2
    const int grp = lane >> 4, m = lane & 15;
    const int r1 = 64 * (16 * grp + (m >> 1)) + (m & 1) * 8, r2 = r1 + 64 * 8;
    const size_t n_chunks = n_quads / 64;
    const size_t cstride = (size_t)gridDim.x * kMfWaves;
    uint4 fv[4], gv[4];
    // load k covers bytes [1024 k, 1024 k + 1024) of the chunk: lane l its 16
    // bytes at 1024 k + 16 l, staged at the same offset
    auto load = [&](size_t c) {
        const mf_v4u* pf = reinterpret_cast<const mf_v4u*>(f) + 256 * c + lane;
        const mf_v4u* pg = reinterpret_cast<const mf_v4u*>(g) + 256 * c + lane;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const mf_v4u x = __builtin_nontemporal_load(pf + 64 * k);
            const mf_v4u y = __builtin_nontemporal_load(pg + 64 * k);
            fv[k] = make_uint4(x.x, x.y, x.z, x.w);
            gv[k] = make_uint4(y.x, y.y, y.z, y.w);
        }
    };
    size_t c = (size_t)blockIdx.x * kMfWaves + wid;
    if (c < n_chunks) load(c);
    for (; c < n_chunks; c += cstride) {
        uint4* sf = reinterpret_cast<uint4*>(stage + 16 * lane);
        uint4* sg = reinterpret_cast<uint4*>(stage + 4096 + 16 * lane);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 x = fv[k], y = gv[k];
            sf[64 * k] = make_uint4(x.x ^ 0x80808080u, x.y ^ 0x80808080u, x.z ^ 0x80808080u, x.w ^ 0x80808080u);
            sg[64 * k] = make_uint4(y.x ^ 0x80808080u, y.y ^ 0x80808080u, y.z ^ 0x80808080u, y.w ^ 0x80808080u);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (c + cstride < n_chunks) load(c + cstride);   // in flight while this chunk runs
        __builtin_amdgcn_sched_barrier(0);
        mf_v4i A[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const mf_v2i lo = lds_tr8(stage + r1 + 16 * a), hi = lds_tr8(stage + r2 + 16 * a);
            A[a] = mf_v4i{lo.x, lo.y, hi.x, hi.y};
#pragma unroll
            for (int w = 0; w < 4; ++w) ucol[a] = __builtin_amdgcn_sdot4(A[a][w], 0x01010101, ucol[a], false);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const mf_v2i lo = lds_tr8(stage + 4096 + r1 + 16 * b), hi = lds_tr8(stage + 4096 + r2 + 16 * b);
            const mf_v4i Bb = mf_v4i{lo.x, lo.y, hi.x, hi.y};
#pragma unroll
            for (int w = 0; w < 4; ++w) vcol[b] = __builtin_amdgcn_sdot4(Bb[w], 0x01010101, vcol[b], false);
#pragma unroll
            for (int a = 0; a < 4; ++a)
                acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[a], Bb, acc[a][b], 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // reads done before the next stage
        nq += 1;
    }
    // ---- the workgroup's integer tiles and column sums, added in LDS
    __syncthreads();
    int* Ts = reinterpret_cast<int*>(mf_lds);      // [16 (a, b)][16 i][16 j]
    int* Us = Ts + 4096;                           // [4 a][16 i]
    int* Vs = Us + 64;                             // [4 b][16 j]
    int* Ns = Vs + 64;                             // chunks
    u64* Qs = reinterpret_cast<u64*>(Ns + 4);      // [16 (a, b)][4 word pairs]
    u64* Ws = Qs + 64;                             // [4 a][2] weighted U, then [4 b][2] V
    for (int i = t; i < 4096 + 64 + 64 + 4; i += kMfThreads) Ts[i] = 0;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                atomicAdd(&Ts[(a * 4 + b) * 256 + (4 * (lane >> 4) + r) * 16 + (lane & 15)], acc[a][b][r]);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        atomicAdd(&Us[a * 16 + (lane & 15)], ucol[a]);
        atomicAdd(&Vs[a * 16 + (lane & 15)], vcol[a]);
    }
    if (lane == 0) atomicAdd(Ns, nq);
    __syncthreads();
    // a wave per tile (a, b); a row of 16 lanes per word pair (hf, hg); a lane
    // adds 4 of the pair's 64 (i, j) entries weighted by 2^{8(i+j)} mod p
    for (int tile = wid; tile < 16; tile += kMfWaves) {
        const int qd = lane >> 4, hf = qd >> 1, hg = qd & 1, sub = lane & 15;
        const int bi = sub >> 1, bj0 = (sub & 1) * 4;
        u64 sacc = 0;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int bj = bj0 + jj;
            sacc += fmul_lazy(mod_i32(Ts[tile * 256 + (8 * hf + bi) * 16 + (8 * hg + bj)]),
                              1ull << ((8 * (bi + bj)) % 61));
        }
        u64 w[1] = {fold1(sacc)};
        dpp_add_step<0xb1, 0xf>(w);    // the 16 lanes of the row
        dpp_add_step<0x4e, 0xf>(w);
        fold_words(w);
        dpp_add_step<0x141, 0xf>(w);
        dpp_add_step<0x140, 0xf>(w);
        if (sub == 0) Qs[tile * 4 + qd] = canon(fold1(w[0]));
    }
    if (t < 16) {
        // t < 8: U of entry t / 2, word t % 2; then V: sum_i 2^{8i} U[8 h + i]
        const int* src = (t < 8 ? Us : Vs) + ((t >> 1) & 3) * 16 + 8 * (t & 1);
        u64 wsum = 0;
        for (int i = 0; i < 8; ++i) wsum += fmul_lazy(mod_i32(src[i]), 1ull << (8 * i));
        Ws[t] = fold61(wsum);
    }
    __syncthreads();
    fp2 e[9];
    if (t == 0) {
        u64 k1 = 0;
        for (int i = 0; i < 8; ++i) k1 = fadd(k1, 1ull << (8 * i));
        const u64 c128 = fmul(128, k1);                                // 128 K1
        const u64 cnn = fmul(fmul(c128, c128), (u64)*Ns * 64);         // 128^2 K1^2 N
        fp2 Pm[4][4];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                u64 Q[2][2];
                for (int hf = 0; hf < 2; ++hf)
                    for (int hg = 0; hg < 2; ++hg)
                        Q[hf][hg] = fadd(fadd(Qs[(a * 4 + b) * 4 + hf * 2 + hg],
                                              fmul(c128, fadd(Ws[a * 2 + hf], Ws[8 + b * 2 + hg]))),
                                         cnn);
                Pm[a][b] = mk2(fsub(Q[0][0], Q[1][1]), fadd(Q[0][1], Q[1][0]));
            }
        // sum of sa[a] sb[b] P_ab
        auto comb = [&](const int (&sa)[4], const int (&sb)[4]) {
            fp2 r = zero2();
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    const int sg = sa[a] * sb[b];
                    if (sg > 0) r = add2(r, Pm[a][b]);
                    else if (sg < 0) r = sub2(r, Pm[a][b]);
                }
            return r;
        };
        const int e0[4] = {1, 0, 0, 0}, e1[4] = {0, 1, 0, 0}, e2[4] = {0, 0, 1, 0}, e3[4] = {0, 0, 0, 1};
        const int dA[4] = {-1, 1, 0, 0}, dB[4] = {0, 0, -1, 1}, x0[4] = {-1, 0, 1, 0},
                  x1[4] = {0, -1, 0, 1}, x2[4] = {1, -1, -1, 1};
        e[0] = comb(e0, e0); e[1] = comb(e1, e1); e[2] = comb(dA, dA);
        e[3] = comb(e2, e2); e[4] = comb(e3, e3); e[5] = comb(dB, dB);
        e[6] = comb(x0, x0); e[7] = comb(x1, x1); e[8] = comb(x2, x2);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) e[i] = zero2();
    }
    ticket_publish_wg<9>(e, pub);
}

// Two rounds of folding in one pass: T_{k-2} (4 n_out entries per table)
// folded by r_{k-2} and r_{k-1} into T_k, with T_k's round sums (round 2
// from the input, and the later passes of the streaming rounds).  A wave
// takes a chunk of 64 outputs (256 inputs, 4 KB per table): it reads each
// table's chunk with four coalesced non-temporal 16-byte loads per lane
// (1 KB a load instruction), stages it in its own LDS slice and reads back
// the quad of the output its lane owns; outputs are stored contiguous
// across the wave.  With each lane loading its own 64 contiguous bytes
// instead, one load instruction touched 64 lines 64 B apart and the kernel
// ran at 5.2 TB/s from a cold cache, against 5.7 TB/s staged (which is also
// what a staged kernel without the arithmetic reaches).
constexpr int kFold01Waves = 8;
// A wave's staging slice holds 256 entries with one entry of padding after
// every 16 (entry e at slice_pos(e)): a lane reading its quad back (64 B at
// lane x 64 B) then meets no other lane of its cycle on a bank; unpadded,
// lanes four apart did (11% of the T_0 -> T_4 pass's time in bank
// conflicts, 0% padded, the pass's time the same: it waits on its
// multiplies).  Entry lane + 64 k sits at slice_pos(lane) + 68 k.
constexpr int kSlice = 272;
__host__ __device__ constexpr int slice_pos(int e) { return e + (e >> 4); }
__device__ __forceinline__ fp2 shfl_xor1(fp2 x)
{
    return mk2(__shfl_xor(x.re, 1, 64), __shfl_xor(x.im, 1, 64));
}

// x from lane (quad_perm CTRL) of this lane's group of four.
template <int CTRL>
__device__ __forceinline__ fp2 quad_pick(fp2 x)
{
    return mk2(dpp_u64<CTRL, 0xf>(x.re), dpp_u64<CTRL, 0xf>(x.im));
}

// The streaming launches run 512-thread workgroups, at most kFold01Groups /
// kFoldGroups of them.  Round 2 measured the same at 2048 workgroups (116.5
// against 118.0 us) and the ticket publish takes at most kMaxPubGroups;
// for the streaming rounds grids of 2048 to 8192 workgroups measured the
// same or slower: with fewer outputs per thread, each wave's closing
// reduction costs more per output.
constexpr int kFoldThreads = 512;
constexpr unsigned kFold01Groups = 1024;
// The quad-sum pass publishes nine sums a workgroup: 512 workgroups, as many
// as are resident at once (two per CU at 128 VGPRs).
constexpr unsigned kFold2QuadGroups = 512;
constexpr unsigned kFoldGroups = 768;      // 76 VGPRs: 3 workgroups per CU

// The two-round pass's grid: a wave per 64-output chunk, at most `cap`
// workgroups (kFold01Groups with pair sums, kFold2QuadGroups with quad sums).
inline unsigned fold01_groups(size_t n_out, unsigned cap = kFold01Groups)
{
    const size_t chunks = (n_out + 63) / 64;
    return (unsigned)std::min<size_t>(cap, (chunks + kFold01Waves - 1) / kFold01Waves);
}

// The sums of T_k:
// - NS = 3, round k's pair sums from two products per lane: the lane
//   holding output j adds f_j g_j (E0 over the even outputs, E1 over the
//   odd) and f_j (g_j - g_j') with j' its pair neighbour, whose sum over both
//   lanes of a pair is f0 g0 + f1 g1 - f0 g1 - f1 g0 = (f1 - f0)(g1 - g0)
//   (EA): every lane busy, where one lane of each pair adding all three left
//   its neighbour idle.
// - NS = 9, the nine quad sums (rounds k and k + 1, the second as
//   quadratics in r_k), from three products per lane: the four lanes of a
//   quad take their neighbours' values by DPP quad permutes; lane i of the
//   quad adds f_i g_i (S00, S11, S22, S33), then one difference product
//   (lane 0: SAA from p1 - p0, lane 1: X1 from p3 - p1, lane 2: X0 from
//   p2 - p0, lane 3: SBB from p3 - p2), and lane 0 also X2 from
//   (p3 - p2) - (p1 - p0) (the others add a product that is dropped).  Each
//   wave's sums meet in lanes 0-3 of each 16-lane row by two row rotations,
//   the workgroup's in LDS; needs n_out a multiple of 64.
// The sweep runs over the chunks from the top down when `desc`, else from
// the bottom up: a pass starts where the pass before it ended, whose reads
// or writes there are the likeliest to be still in the memory-side cache.
// NTL: non-temporal loads (an input read once, from HBM) or cached ones.
template <int NS, bool NTL = true>
__device__ __forceinline__ void fold2_body(const V2* __restrict__ fin, const V2* __restrict__ gin,
                                           size_t n_out, u64 r0r, u64 r0i, u64 r1r, u64 r1i,
                                           V2* __restrict__ fout, V2* __restrict__ gout,
                                           const TicketPub& pub, int desc)
{
    static_assert(kFoldThreads == 64 * kFold01Waves, "one LDS slice per wave");
    static_assert(NS == 3 || NS == 9, "pair or quad sums");
    __shared__ __attribute__((aligned(16))) V2 stage[kFold01Waves][kSlice];
    u64 rw[4] = {r0r, r0i, r1r, r1i};
    if (!launch_challenges(pub, rw)) return;
    const fp2 r0 = mk2(rw[0], rw[1]), r1 = mk2(rw[2], rw[3]);
    const fp2 r01 = uniform2(mul2(r0, r1));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    mf_v4u* s_st = reinterpret_cast<mf_v4u*>(stage[wid]) + slice_pos(lane);   // staging stores
    const V2* s_ld = stage[wid] + slice_pos(4 * lane);                       // and its quad
    // n_out is a power of two: either a multiple of 64 or one short chunk
    // of an even length m.  In a short chunk the loads past its 4m inputs
    // repeat its last one (in bounds), and only lanes below m, both lanes of
    // each pair, store and add.
    const size_t n_chunks = (n_out + 63) / 64;
    const unsigned m = n_out < 64 ? (unsigned)n_out : 64u;
    constexpr int NZ = NS == 9 ? 3 : 2;
    LazyProdSum z[NZ];
    lazy_zero(z);
    for (size_t c0 = (size_t)blockIdx.x * kFold01Waves + wid; c0 < n_chunks;
         c0 += (size_t)gridDim.x * kFold01Waves) {
        const size_t c = desc ? n_chunks - 1 - c0 : c0;
        const mf_v4u* pf = reinterpret_cast<const mf_v4u*>(fin + 256 * c);
        const mf_v4u* pg = reinterpret_cast<const mf_v4u*>(gin + 256 * c);
        mf_v4u xf[4], xg[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned i = min((unsigned)(lane + 64 * k), 4 * m - 1);
            if (NTL) {
                xf[k] = __builtin_nontemporal_load(pf + i);
                xg[k] = __builtin_nontemporal_load(pg + i);
            } else {
                xf[k] = pf[i];
                xg[k] = pg[i];
            }
        }
        // one slice per wave: f staged and its quads read back, then g; the
        // wave-scope fences keep the compiler from moving an LDS store
        // across the reads of the data it replaces
#pragma unroll
        for (int k = 0; k < 4; ++k) s_st[68 * k] = xf[k];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const fp2 f0 = ldv(s_ld), f1 = ldv(s_ld + 1), f2 = ldv(s_ld + 2), f3 = ldv(s_ld + 3);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int k = 0; k < 4; ++k) s_st[68 * k] = xg[k];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const fp2 g0 = ldv(s_ld), g1 = ldv(s_ld + 1), g2 = ldv(s_ld + 2), g3 = ldv(s_ld + 3);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const fp2 f = fold_quad(f0, f1, f2, f3, r0, r1, r01);
        const fp2 g = fold_quad(g0, g1, g2, g3, r0, r1, r01);
        if constexpr (NS == 3) {
            const fp2 go = shfl_xor1(g);
            if ((unsigned)lane < m) {
                const size_t j = 64 * c + lane;
                stv(fout + j, f);
                stv(gout + j, g);
                z[0].mac(f, g);
                z[1].mac(f, sub2(g, go));
            }
        } else {
            // difference operands: lane i of the quad takes p_hi(i) - p_lo(i)
            // with hi = (1, 3, 2, 3) and lo = (0, 1, 0, 2); lane 0 then
            // (p3 - p2) - (p1 - p0) from lane 3's difference and its own
            const fp2 df = sub2(quad_pick<0xED>(f), quad_pick<0x84>(f));
            const fp2 dg = sub2(quad_pick<0xED>(g), quad_pick<0x84>(g));
            const fp2 xf2 = sub2(quad_pick<0xFF>(df), df);
            const fp2 xg2 = sub2(quad_pick<0xFF>(dg), dg);
            const size_t j = 64 * c + lane;
            stv(fout + j, f);
            stv(gout + j, g);
            z[0].mac(f, g);
            z[1].mac(df, dg);
            z[2].mac(xf2, xg2);
        }
    }
    if constexpr (NS == 3) {
        const fp2 self = z[0].value();
        const bool odd_out = (lane & 1) != 0;
        fp2 e[3] = {odd_out ? zero2() : self, odd_out ? self : zero2(), z[1].value()};
        ticket_publish<3>(e, pub);
    } else {
        // lane l of a row adds the four lanes l mod 4 of the row (row_ror 8,
        // then 4); lanes 0-3 of each row store their three sums
        __shared__ fp2 s_part[kFold01Waves * 4][4][3];
        __shared__ fp2 s_tot9[9];
        u64 w[6];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const fp2 v = z[q].value();
            w[2 * q] = v.re;
            w[2 * q + 1] = v.im;
        }
        dpp_add_step<0x128, 0xf>(w);   // row_ror:8
        dpp_add_step<0x124, 0xf>(w);   // row_ror:4
        if ((lane & 15) < 4) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
                s_part[wid * 4 + (lane >> 4)][lane & 3][q] =
                    mk2(canon(fold1(w[2 * q])), canon(fold1(w[2 * q + 1])));
        }
        __syncthreads();
        if (threadIdx.x < 9) {
            // sum s: (product, quad lane) = S00 (0, 0), S11 (0, 1), SAA (1, 0),
            // S22 (0, 2), S33 (0, 3), SBB (1, 3), X0 (1, 2), X1 (1, 1), X2 (2, 0),
            // packed as nibble s = 4 product + lane
            const int sidx = threadIdx.x;
            const int code = (int)((0x856732410ull >> (4 * sidx)) & 0xf);
            fp2 acc = zero2();
            for (int r = 0; r < kFold01Waves * 4; ++r) acc = add2(acc, s_part[r][code & 3][code >> 2]);
            s_tot9[sidx] = acc;
        }
        __syncthreads();
        fp2 e[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) e[q] = s_tot9[q];
        ticket_publish_wg<9>(e, pub);
    }
}

__global__ void __launch_bounds__(kFoldThreads) k_sc2_fold2_pairs(const V2* __restrict__ fin,
                                                         const V2* __restrict__ gin,
                                                         size_t n_out, u64 r0r, u64 r0i,
                                                         u64 r1r, u64 r1i,
                                                         V2* __restrict__ fout,
                                                         V2* __restrict__ gout, TicketPub pub,
                                                         int desc)
{
    fold2_body<3>(fin, gin, n_out, r0r, r0i, r1r, r1i, fout, gout, pub, desc);
}

// The same with plain (cached) loads, for an input the pass before has just
// written: 43 us against 46 us for 2 x 2^22 entries (the memory-side cache
// still holds much of it).
__global__ void __launch_bounds__(kFoldThreads) k_sc2_fold2_pairs_cached(const V2* __restrict__ fin,
                                                         const V2* __restrict__ gin,
                                                         size_t n_out, u64 r0r, u64 r0i,
                                                         u64 r1r, u64 r1i,
                                                         V2* __restrict__ fout,
                                                         V2* __restrict__ gout, TicketPub pub,
                                                         int desc)
{
    fold2_body<3, false>(fin, gin, n_out, r0r, r0i, r1r, r1i, fout, gout, pub, desc);
}

__global__ void __launch_bounds__(kFoldThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_sc2_fold2_quads(const V2* __restrict__ fin,
                                                         const V2* __restrict__ gin,
                                                         size_t n_out, u64 r0r, u64 r0i,
                                                         u64 r1r, u64 r1i,
                                                         V2* __restrict__ fout,
                                                         V2* __restrict__ gout, TicketPub pub,
                                                         int desc)
{
    fold2_body<9>(fin, gin, n_out, r0r, r0i, r1r, r1i, fout, gout, pub, desc);
}

// Rounds 0 to 3 from one pass over the input on the int8 matrix cores: the
// 256 block products P[a][b] = sum over 16-entry blocks B of
// f[16B + a] g[16B + b] (a, b < 16), from which the host takes the four
// rounds (blocks.hpp).  The byte-limb method of k_sc2_first01_mfma, with a
// block of 16 entries in place of a quad: per chunk of 64 blocks (16 KB a
// table) one MFMA per pair (a, b) adds the 16 x 16 byte products summed
// over the chunk.  The 256 pairs need 256 MFMAs a chunk, so the chunk is
// staged once per workgroup and shared: 16 waves, wave w taking the pairs
// with a in 4 (w / 4) .. + 3 and b in 4 (w % 4) .. + 3 (16 accumulator
// tiles), every wave reading its operands from the same staged chunk.
// Staging is double buffered (one barrier a chunk).  A
// block takes 272 bytes of LDS (16 of padding), so the eight blocks a
// transposed read touches sit on distinct banks.  Each workgroup writes its
// 256 products to its slots; k_sc2_blocks_reduce adds them.
constexpr int kBkThreads = 1024;
constexpr int kBkWaves = kBkThreads / 64;
constexpr unsigned kBkGroups = 256;
constexpr int kBkStride = 272;
// block b of a staged chunk sits at kBkStride b + kBkGroupPad (b / 16): the
// extra 128 B puts the second 16-lane group of a transposed read (blocks
// 16 G + ...) on the other half of the banks from the first.  Bank conflicts
// went from 13% of the pass's time to 0%, its time the same: one chunk per
// CU in flight (32 KB) at the 1.7 us a chunk takes is the 4.7 TB/s it
// reaches, and a second chunk in flight only lengthened that latency.
constexpr int kBkGroupPad = 128;
__host__ __device__ constexpr int bk_block_off(int b) { return kBkStride * b + kBkGroupPad * (b >> 4); }
constexpr int kBkTable = bk_block_off(64);   // one table's chunk in LDS
constexpr int kBkBuf = 2 * kBkTable;         // f and g
constexpr size_t kBkLds = 2 * (size_t)kBkBuf;
// |tile entry| <= (number of blocks) x 2^14 below 2^31: at most 131071
// blocks a workgroup; the host keeps n / 16 <= 2^16 x kBkGroups
constexpr size_t kBkMaxEntries = (size_t)16 * 65536 * kBkGroups;
// Inputs from 2^14 entries (16 chunks) take the block pass.
constexpr int kBlocksMinVars = 14;

__global__ void __launch_bounds__(kBkThreads) k_sc2_blocks_mfma(const V2* __restrict__ f,
                                                               const V2* __restrict__ g,
                                                               size_t n_blocks, fp2* __restrict__ slots)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t bk_lds[];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int a0 = 4 * (wid >> 2), b0 = 4 * (wid & 3);
    mf_v4i acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mf_v4i{0, 0, 0, 0};
    // column sums of the u bytes: wave w adds those of f entry w (its A
    // operand wid % 4) and of g entry 4 (w % 4) + w / 4 (its B operand w / 4),
    // so that every entry's are added once
    int ucol = 0, vcol = 0;
    const int ua = wid & 3, vb = wid >> 2;
    // transposed-read addresses: lane 16 G + m reads word m & 1 of blocks
    // 16 G + m / 2 and 16 G + 8 + m / 2
    const int grp = lane >> 4, m = lane & 15;
    const int r1 = bk_block_off(16 * grp + (m >> 1)) + (m & 1) * 8, r2 = r1 + kBkStride * 8;
    // thread t stages entry t of the chunk's 1024: block t / 16, entry t % 16
    const int st_off = bk_block_off(t >> 4) + 16 * (t & 15);
    const size_t n_chunks = n_blocks / 64;
    // one chunk of loads in flight while a chunk is staged and multiplied
    // (two or three, in registers of a loop unrolled by as many, measured
    // 125 and 143-147 us against 117-122 us at 2 x 2^24 entries)
    mf_v4u fv = {0, 0, 0, 0}, gv = {0, 0, 0, 0};
    const mf_v4u* pf = reinterpret_cast<const mf_v4u*>(f) + t;
    const mf_v4u* pg = reinterpret_cast<const mf_v4u*>(g) + t;
    if (blockIdx.x < n_chunks) {
        fv = __builtin_nontemporal_load(pf + 1024 * (size_t)blockIdx.x);
        gv = __builtin_nontemporal_load(pg + 1024 * (size_t)blockIdx.x);
    }
    unsigned nb = 0;
    const mf_v4u x80 = {0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u};
    for (size_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        uint8_t* sb = bk_lds + ((nb & 1) ? kBkBuf : 0);
        *reinterpret_cast<mf_v4u*>(sb + st_off) = fv ^ x80;
        *reinterpret_cast<mf_v4u*>(sb + kBkTable + st_off) = gv ^ x80;
        if (c + gridDim.x < n_chunks) {   // the next chunk, in flight while this one runs
            fv = __builtin_nontemporal_load(pf + 1024 * (c + gridDim.x));
            gv = __builtin_nontemporal_load(pg + 1024 * (c + gridDim.x));
        }
        __syncthreads();   // the chunk is staged; every read of the buffer two chunks back is done
        mf_v4i A[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const mf_v2i lo = lds_tr8(sb + r1 + 16 * (a0 + a)), hi = lds_tr8(sb + r2 + 16 * (a0 + a));
            A[a] = mf_v4i{lo.x, lo.y, hi.x, hi.y};
            if (a == ua) {
#pragma unroll
                for (int w = 0; w < 4; ++w) ucol = __builtin_amdgcn_sdot4(A[a][w], 0x01010101, ucol, false);
            }
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const mf_v2i lo = lds_tr8(sb + kBkTable + r1 + 16 * (b0 + b));
            const mf_v2i hi = lds_tr8(sb + kBkTable + r2 + 16 * (b0 + b));
            const mf_v4i Bb = mf_v4i{lo.x, lo.y, hi.x, hi.y};
            if (b == vb) {
#pragma unroll
                for (int w = 0; w < 4; ++w) vcol = __builtin_amdgcn_sdot4(Bb[w], 0x01010101, vcol, false);
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
                acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[a], Bb, acc[a][b], 0, 0, 0);
        }
        ++nb;
    }
    // ---- column sums in LDS (one wave of each row and column of the wave
    // grid adds them), their weighted words, then each wave's 16 products
    __syncthreads();
    int* Us = reinterpret_cast<int*>(bk_lds);      // [16 a][16 i]
    int* Vs = Us + 256;                            // [16 b][16 j]
    u64* Ws = reinterpret_cast<u64*>(Vs + 256);    // [16 a][2] weighted U, then [16 b][2] V
    fp2* Ps = reinterpret_cast<fp2*>(Ws + 64);     // [16 a][16 b]
    if (t < 512) Us[t] = 0;
    __syncthreads();
    atomicAdd(&Us[(a0 + ua) * 16 + m], ucol);
    atomicAdd(&Vs[(b0 + vb) * 16 + m], vcol);
    __syncthreads();
    const u64 k1 = 0x0101010101010101ull;                           // sum_{i<8} 2^{8i}, below p
    const u64 c128 = fmul(128, k1);                                 // 128 K1
    const u64 cnn = fmul(fmul(c128, c128), fold61((u64)nb * 64));   // 128^2 K1^2 N
    if (t < 64) {
        // t < 32: U of entry t / 2, word t % 2; then V: 128 K1 sum_i 2^{8i} U[8 h + i]
        const int* src = (t < 32 ? Us : Vs) + ((t >> 1) & 15) * 16 + 8 * (t & 1);
        u64 wsum = 0;
        for (int i = 0; i < 8; ++i) wsum = fold1(wsum + rot61(mod_i32(src[i]), 8 * i));   // < 2^64
        Ws[t] = fmul(c128, fold61(wsum));
    }
    __syncthreads();
    // lane l holds C[4 (l / 16) + r][l % 16]: word pair (hf, hg) = (l / 32,
    // (l / 8) % 2), bytes (bi, bj) = ((4 (l / 16) + r) % 8, l % 8)
    const int hf = lane >> 5, hg = (lane >> 3) & 1, bj = lane & 7;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            u64 w[1] = {0};   // four rotations, each below 1.5 x 2^61
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int bi = (4 * (lane >> 4) + r) & 7;
                w[0] += rot61(mod_i32(acc[a][b][r]), (8 * (bi + bj)) % 61);
            }
            w[0] = fold1(w[0]);
            // the 16 lanes of each word pair: the 8 of a half row, then the
            // other row of the pair
            dpp_add_step<0xb1, 0xf>(w);    // quad_perm [1,0,3,2]
            dpp_add_step<0x4e, 0xf>(w);    // quad_perm [2,3,0,1]
            fold_words(w);
            dpp_add_step<0x141, 0xf>(w);   // row_half_mirror
            w[0] = fold1(w[0]);
            w[0] += __shfl_xor(w[0], 16, 64);
            const u64 q = fold61(w[0]);
            // Q[hf][hg] with the byte offsets restored
            const u64 qt = fadd(fadd(q, fadd(Ws[(a0 + a) * 2 + hf], Ws[32 + (b0 + b) * 2 + hg])), cnn);
            const u64 q00 = __shfl(qt, 0, 64), q01 = __shfl(qt, 8, 64);
            const u64 q10 = __shfl(qt, 32, 64), q11 = __shfl(qt, 40, 64);
            if (lane == 0) Ps[(a0 + a) * 16 + b0 + b] = mk2(fsub(q00, q11), fadd(q01, q10));
        }
    __syncthreads();
    if (t < 256) {   // whole lines from one workgroup (product-major slots, each line
                     // written by eight workgroups, made the reduction 6.2 us against 4.3)
        const fp2 v = Ps[t];
        V2 x;
        x.re = v.re;
        x.im = v.im;
        reinterpret_cast<V2*>(slots)[(size_t)blockIdx.x * 256 + t] = x;
    }
}

// The 256 block products over the groups' slots (256 a group): workgroup s
// adds product s and publishes it to the host as two tagged words
// (value | (seq mod 8) << 61).
__global__ void __launch_bounds__(256) k_sc2_blocks_reduce(const fp2* __restrict__ slots,
                                                           unsigned groups, fp2* host_vals,
                                                           unsigned seq)
{
    const unsigned s = blockIdx.x, t = threadIdx.x;
    fp2 v[1] = {zero2()};
    for (unsigned g = t; g < groups; g += blockDim.x) {
        const V2 x = reinterpret_cast<const V2*>(slots)[(size_t)g * 256 + s];
        v[0] = add2(v[0], mk2(x.re, x.im));
    }
    block_sum<1>(v);
    if (t == 0) {
        const u64 tag = (u64)(seq & 7) << 61;
        u64* w = reinterpret_cast<u64*>(host_vals) + 2 * s;
        __hip_atomic_store(w, v[0].re | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(w + 1, v[0].im | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Rounds 0 to 3 applied in one pass: T_0 (16 n_out entries a table) folded
// by r_0 .. r_3 into T_4, with round 4's pair sums (published as
// k_sc2_fold2_pairs does).  A wave takes 64 outputs (1024 inputs, 16 KB a
// table), one table at a time, in four 4 KB sub-chunks staged through its
// LDS slice as in k_sc2_fold2_pairs: lane l folds quad l of sub-chunk s by
// r_0, r_1 into T_2 entry 64 s + l; the four T_2 entries of each lane then
// go through the slice again so that lane j reads entries 4j .. 4j + 3 and
// folds them by r_2, r_3 into output j.  The sweep runs from the top down.
// The pass is bound by its integer arithmetic, not by HBM: per wave and 64
// outputs (2048 input entries) its loop issued 3078 VALU instructions, 474
// of them 64-bit multiply-adds, with a reduction per product; summing each
// Karatsuba sum's three products before one reduction (sum3_lazy) took it to
// 2327 and the pass from 119-122 to 112-118 us (79 VGPRs, 6 waves a SIMD;
// forced to 7 waves it spills and ran 118-120 against 115-117 us).  These
// measured the same or slower (before that change): issuing
// each sub-chunk's loads one step ahead (127 VGPRs: 127-132 us against
// 121-123 us), 8-wave workgroups, and 4, 5 or 6 waves per SIMD (6 spills).
// Workgroups of 4 waves; the last one's ticket gather reuses the slices.
constexpr int kFold16Waves = 4;
constexpr int kFold16Threads = 64 * kFold16Waves;
inline unsigned fold16_groups(size_t n_out)
{
    return (unsigned)std::min<size_t>(kMaxPubGroups, (n_out / 64 + kFold16Waves - 1) / kFold16Waves);
}

__global__ void __launch_bounds__(kFold16Threads) __attribute__((amdgpu_waves_per_eu(5))) k_sc2_fold16(const V2* __restrict__ fin,
                                                    const V2* __restrict__ gin, size_t n_out,
                                                    u64 r0r, u64 r0i, u64 r1r, u64 r1i,
                                                    u64 r2r, u64 r2i, u64 r3r, u64 r3i,
                                                    V2* __restrict__ fout, V2* __restrict__ gout,
                                                    TicketPub pub)
{
    // one 4 KB slice per wave; the last workgroup's ticket gather reuses them
    static_assert(kFold16Waves * kSlice >= 1024, "gather partials fit the slices");
    __shared__ __attribute__((aligned(16))) V2 stage[kFold16Waves][kSlice];
    const fp2 r0 = uniform2(mk2(r0r, r0i)), r1 = uniform2(mk2(r1r, r1i));
    const fp2 r2 = uniform2(mk2(r2r, r2i)), r3 = uniform2(mk2(r3r, r3i));
    const fp2 r01 = uniform2(mul2(r0, r1)), r23 = uniform2(mul2(r2, r3));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    mf_v4u* s_st = reinterpret_cast<mf_v4u*>(stage[wid]) + slice_pos(lane);
    const V2* s_ld = stage[wid] + slice_pos(4 * lane);
    const size_t n_super = n_out / 64;   // n_out is a multiple of 64
    LazyProdSum z[2];   // f g, f (g - g')
    lazy_zero(z);
    for (size_t c0 = (size_t)blockIdx.x * kFold16Waves + wid; c0 < n_super;
         c0 += (size_t)gridDim.x * kFold16Waves) {
        const size_t c = n_super - 1 - c0;
        fp2 out[2];
#pragma unroll
        for (int tab = 0; tab < 2; ++tab) {
            const mf_v4u* src = reinterpret_cast<const mf_v4u*>((tab ? gin : fin) + 1024 * c) + lane;
            fp2 t2[4];
#pragma unroll
            for (int sub = 0; sub < 4; ++sub) {
                mf_v4u x[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) x[k] = __builtin_nontemporal_load(src + 256 * sub + 64 * k);
#pragma unroll
                for (int k = 0; k < 4; ++k) s_st[68 * k] = x[k];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                t2[sub] = fold_quad(ldv(s_ld), ldv(s_ld + 1), ldv(s_ld + 2), ldv(s_ld + 3), r0, r1, r01);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
#pragma unroll
            for (int sub = 0; sub < 4; ++sub) stv(stage[wid] + slice_pos(lane) + 68 * sub, t2[sub]);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            out[tab] = fold_quad(ldv(s_ld), ldv(s_ld + 1), ldv(s_ld + 2), ldv(s_ld + 3), r2, r3, r23);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        const size_t j = 64 * c + lane;
        stv(fout + j, out[0]);
        stv(gout + j, out[1]);
        const fp2 go = shfl_xor1(out[1]);
        z[0].mac(out[0], out[1]);
        z[1].mac(out[0], sub2(out[1], go));
    }
    const fp2 self = z[0].value();
    const bool odd_out = (lane & 1) != 0;
    fp2 e[3] = {odd_out ? zero2() : self, odd_out ? self : zero2(), z[1].value()};
    ticket_publish<3, true>(e, pub, reinterpret_cast<fp2*>(&stage[0][0]));
}

// Round k >= 1: fold T_{k-1} (length 4 * n_quads) with r into T_k and sum
// the round-k pairs.  Quad q = inputs 4q..4q+3 -> outputs 2q, 2q+1.  With
// `desc` the sweep runs from the top quad down, so a pass starts where the
// previous pass (which ran the other way) ended: there its input is still
// in the memory-side cache.
__global__ void __launch_bounds__(kFoldThreads) k_sc2_fold_round(const V2* __restrict__ fin,
                                                        const V2* __restrict__ gin,
                                                        size_t n_quads, u64 rr, u64 ri,
                                                        V2* __restrict__ fout,
                                                        V2* __restrict__ gout, TicketPub pub,
                                                        int desc)
{
    u64 rw[4] = {rr, ri, 0, 0};
    if (!launch_challenges(pub, rw)) return;
    const fp2 r = mk2(rw[0], rw[1]);
    LazyProdSum z[3];
    lazy_zero(z);
    for (size_t q0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q0 < n_quads;
         q0 += (size_t)gridDim.x * blockDim.x) {
        const size_t q = desc ? n_quads - 1 - q0 : q0;
        const V2* fp = fin + 4 * q;
        const V2* gp = gin + 4 * q;
        const fp2 fa = ldv(fp), fb = ldv(fp + 1), fc = ldv(fp + 2), fd = ldv(fp + 3);
        const fp2 ga = ldv(gp), gb = ldv(gp + 1), gc = ldv(gp + 2), gd = ldv(gp + 3);
        const fp2 f0 = fold2(fa, fb, r), f1 = fold2(fc, fd, r);
        const fp2 g0 = fold2(ga, gb, r), g1 = fold2(gc, gd, r);
        stv(fout + 2 * q, f0);
        stv(fout + 2 * q + 1, f1);
        stv(gout + 2 * q, g0);
        stv(gout + 2 * q + 1, g1);
        accum_pair(z, f0, f1, g0, g1);
    }
    lazy_publish<3>(z, pub);
}

// The last rounds in one workgroup, two rounds per exchange with the host.
// Entered at round k0 with T_{k0-1} (len_in <= kTailIn entries per table) in
// global memory and r_{k0-1}.  Each exchange applies the challenges not yet
// applied (r_{k-1}, or r_{k-2} and r_{k-1}) to the table, which gives T_k (in
// LDS, in place: f at [0, kTailIn/2), g after it), and lane 0 publishes,
// tagged with seq0 + e like the other launches:
//   - the nine sums of T_k's quads (accum9), from which the host draws r_k
//     and r_{k+1} (round k+1's sums are quadratics in r_k, as after the
//     first pass), or
//   - for a last single round (an odd count), the three sums of T_k's one
//     pair;
// then waits for the host to write the challenges, tagged the same, into the
// answer slot.  One exchange per two rounds: the host round trip and the
// wave reductions are paid half as often.  After the last exchange the
// kernel applies the last challenges and publishes the two final values
// tagged with seq0 + ne (ne = tail_exchanges(nv - k0)).  Every wait is
// bounded by the wall clock (spin_ticks): on expiry the kernel publishes
// kSeqAbort and exits, so it ends whatever the host does.
constexpr int kTailThreads = 1024;
constexpr unsigned kTailIn = 8192;
constexpr size_t kTailLds = (size_t)kTailIn * sizeof(V2);   // both tables, T_{k0}

// Exchanges for nr rounds: pairs, then a single one if nr is odd.
__host__ __device__ constexpr int tail_exchanges(int nr) { return (nr + 1) / 2; }

// Entry 0 of the next table from the first 2^np entries at s.
__device__ __forceinline__ fp2 fold_head(const V2* s, int np, fp2 ra, fp2 rb, fp2 rab)
{
    return np == 1 ? fold2(ldv(s), ldv(s + 1), ra)
                   : fold_quad(ldv(s), ldv(s + 1), ldv(s + 2), ldv(s + 3), ra, rb, rab);
}

// Operands of sum s (0..8) of quad q of T_k in LDS (see the file header:
// S00, S11, SAA, S22, S33, SBB, X0, X1, X2).
__device__ __forceinline__ fp2 quad_operand(const V2* t, unsigned q, int s)
{
    const V2* p = t + 4 * q;
    switch (s) {
    case 0: return ldv(p);
    case 1: return ldv(p + 1);
    case 2: return sub2(ldv(p + 1), ldv(p));
    case 3: return ldv(p + 2);
    case 4: return ldv(p + 3);
    case 5: return sub2(ldv(p + 3), ldv(p + 2));
    case 6: return sub2(ldv(p + 2), ldv(p));
    case 7: return sub2(ldv(p + 3), ldv(p + 1));
    default: return sub2(sub2(ldv(p + 3), ldv(p + 2)), sub2(ldv(p + 1), ldv(p)));
    }
}

// Operands of sum s (0..2: E0, E1, EA) of pair j of T_k in LDS.
__device__ __forceinline__ fp2 pair_operand(const V2* t, unsigned j, int s)
{
    const V2* p = t + 2 * j;
    return s == 0 ? ldv(p) : s == 1 ? ldv(p + 1) : sub2(ldv(p + 1), ldv(p));
}

constexpr int kFoldPer = 4;   // entries of T_k a thread folds (kTailIn / 2 / kTailThreads)

// The device work of one exchange on a workgroup's table, shaped for the
// latency-bound sizes of the last rounds: T_k = T_{k-np} (len entries per
// table at sf / sg: global memory, or LDS when src_lds) folded by the
// challenges not yet applied (np = 1: ra; np = 2: ra, rb, rab = ra rb) into
// tf / tg, each thread folding its own entries; then wave s < NSUM adds sum s
// over T_k's quads (NSUM = 9: two rounds at once) or pairs (NSUM = 3: one
// round), a product a lane per quad or pair, and reduces its own two words.
// The sums land in s_sum; every thread of the workgroup calls it.
template <int NSUM>
__device__ __forceinline__ void exchange_work(const V2* sf, const V2* sg, bool src_lds,
                                              unsigned len, int np, fp2 ra, fp2 rb, fp2 rab,
                                              V2* tf, V2* tg, fp2* s_sum)
{
    const unsigned lk = len >> np;   // entries of T_k, at most kFoldPer * blockDim.x
    const unsigned t = threadIdx.x, nt = blockDim.x;
    fp2 of[kFoldPer], og[kFoldPer];
#pragma unroll
    for (int i = 0; i < kFoldPer; ++i) {
        const unsigned j = t + i * nt;
        if (j < lk) {
            of[i] = fold_head(sf + ((size_t)j << np), np, ra, rb, rab);
            og[i] = fold_head(sg + ((size_t)j << np), np, ra, rb, rab);
        }
    }
    if (src_lds) __syncthreads();   // every read of T_{k-np} is done before T_k overwrites it
#pragma unroll
    for (int i = 0; i < kFoldPer; ++i) {
        const unsigned j = t + i * nt;
        if (j < lk) {
            stv(tf + j, of[i]);
            stv(tg + j, og[i]);
        }
    }
    __syncthreads();
    const int wid = (int)(t >> 6);
    const unsigned lane = t & 63;
    if (wid < NSUM) {
        LazyProdSum z;
        z.zero();
        const unsigned nu = NSUM == 9 ? lk / 4 : lk / 2;
        for (unsigned u = lane; u < nu; u += 64) {
            if (NSUM == 9)
                z.mac(quad_operand(tf, u, wid), quad_operand(tg, u, wid));
            else
                z.mac(pair_operand(tf, u, wid), pair_operand(tg, u, wid));
        }
        fp2 v[1] = {z.value()};
        wave_sum_upto<1>(v, nu);
        if (lane == 0) s_sum[wid] = v[0];
    }
    __syncthreads();
}

// Lane 0 waits for the host's NA answer words to the exchange published as
// `want`; returns 0 (and publishes kSeqAbort) if the wait expires.
template <int NA>
__device__ __forceinline__ int tail_wait(unsigned want, unsigned* host_seq, u64* answer,
                                         long long spin_ticks, long long* trace, int i, u64 (&rv)[NA])
{
    if (!take_answer_n<NA>(answer, want, wall_clock64(), spin_ticks, rv)) {
        __hip_atomic_store(host_seq, kSeqAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return 0;
    }
    if (trace) trace[3 * i + 2] = wall_clock64();
    return 1;
}

// Fold of quad q of T_{k-1} (src) by r: the two T_k entries per table.
__device__ __forceinline__ void tail_fold(const V2* sf, const V2* sg, unsigned q, fp2 r, fp2& f0,
                                          fp2& f1, fp2& g0, fp2& g1)
{
    f0 = fold2(ldv(sf + 4 * q), ldv(sf + 4 * q + 1), r);
    f1 = fold2(ldv(sf + 4 * q + 2), ldv(sf + 4 * q + 3), r);
    g0 = fold2(ldv(sg + 4 * q), ldv(sg + 4 * q + 1), r);
    g1 = fold2(ldv(sg + 4 * q + 2), ldv(sg + 4 * q + 3), r);
}

// Rounds on a table too large for one workgroup but small enough that the
// launches, not the bytes, set the time: kMidGroups workgroups, all resident
// at once (checked by an arrival gate at the start), each keep a contiguous
// slice of the tables in LDS.  Variables bind LSB-first, so a slice of S entries folds
// into its own S/2 entries and its pairs stay inside it for log2(S) - 1
// rounds.  Per exchange (one or two rounds, mid_exchange_rounds) every
// workgroup writes its three or nine partial sums to its words of `parts`
// (device memory) as agent-scope stores, each value (61 bits) tagged with the
// exchange mod 8 in its top three bits; waves of workgroup 0 gather them
// (wave s: sum s; lane l waits for the words of workgroups l, l + 64 to carry
// the tag),
// adds them, and publishes the total to the host like the one-workgroup tail
// (tagged words, publish_tagged).  The words start as all ones (value
// bits p, never canonical), so round 0 cannot match a stale word, and later
// rounds find the previous round's tag until the new word lands: no fence and
// no extra launch.  The host runs the transcript and answers through the
// answer slot; workgroup 0 relays the challenges to the others through two or
// four tagged words the same way.  Every wait is bounded by the wall clock; an
// expired wait publishes kSeqAbort and relays p, and every workgroup leaves.
// 128 measured 1% faster a prove than 64: one streaming launch fewer (2^20
// entries go to the slices), for a dearer gather (two workgroups per lane)
constexpr int kMidGroups = 128;
constexpr unsigned kGatePending = 0, kGateGo = 1, kGateAbort = 2;
// The mid launch's last table per workgroup: kMidGroups of them make the
// tail's input.  With two rounds an exchange, handing 2048 entries to the
// tail measured 150 us for the mid and tail rounds of a 2^24 prove, against
// 161 us for 4096 and 164 us for 8192 (one workgroup folding the larger
// tables costs more than the mid exchanges it saves).
constexpr unsigned kMidEnd = 2048 / kMidGroups;
static_assert(kMidEnd >= 2 && kMidEnd * kMidGroups <= kTailIn, "mid hand-over");
constexpr unsigned kSeqGateAbort = 0xfffffffeu;   // the mid launch did not start whole
constexpr int kMidWords = 9 * 2;   // a workgroup's tagged words per exchange (up to nine sums)

// Fresh state for a mid launch: the tagged partial words and the relay all
// ones (value bits p under tag 7: no round's word), the gate at no arrivals
// and no decision.  One launch instead of three fills.
__global__ void __launch_bounds__(256) k_sc2_mid_reset(u64* parts, u64* relay, unsigned* gate)
{
    for (int j = threadIdx.x; j < kMidGroups * kMidWords; j += blockDim.x) parts[j] = ~0ull;
    if (threadIdx.x == 0) {
        relay[0] = relay[1] = relay[2] = relay[3] = ~0ull;
        gate[0] = 0;
        gate[1] = kGatePending;
    }
}

// Exchanges of a mid launch of nr rounds: the last is a single round (its
// challenge goes to the next launch as an argument), the first too when nr is
// even, every other one two rounds (the nine quad sums, as in the tail).
__host__ __device__ constexpr int mid_exchanges(int nr) { return (nr & 1) ? (nr + 1) / 2 : nr / 2 + 1; }
__host__ __device__ inline int mid_exchange_rounds(int nr, int e)
{
    return (e == mid_exchanges(nr) - 1 || (e == 0 && !(nr & 1))) ? 1 : 2;
}

// End of a mid exchange, run by every thread of every workgroup with this
// workgroup's N sums in s_sum: thread 0 stores them to its words of `parts`,
// tagged with the exchange mod 8; in workgroup 0, wave s < N adds sum s over
// every workgroup's words (lane l: workgroups l, l + 64, each word waited for
// until it carries the tag) and thread 0 publishes the totals to the host
// and, unless `last`, waits for the host's NA challenge words and relays
// them; the other workgroups' thread 0 waits for the relay.  Returns 0 (in
// every thread) when a wait expired: every workgroup then leaves.  The
// challenges land in s_r.
template <int N, int NA>
__device__ __forceinline__ int mid_exchange(const fp2* s_sum, fp2* s_tot, int* s_flag, u64* s_r,
                                            int e, bool last, unsigned b, u64* parts,
                                            fp2* host_vals, unsigned* host_seq, unsigned seq0,
                                            u64* answer, long long spin_ticks, u64* relay,
                                            long long* trace)
{
    const unsigned t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const u64 tag = (u64)(e & 7) << 61;
    if (t == 0) {
        if (trace) trace[((size_t)e * kMidGroups + b) * 2] = wall_clock64();
        u64* part = parts + (size_t)b * kMidWords;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            __hip_atomic_store(part + 2 * j, s_sum[j].re | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(part + 2 * j + 1, s_sum[j].im | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        *s_flag = 1;
    }
    const long long t0 = wall_clock64();
    if (b == 0) {
        __syncthreads();   // s_flag set before any wave can clear it
        if (wid < (unsigned)N) {
            int ok = 1;
            fp2 acc = zero2();
            for (unsigned g = lane; g < (unsigned)kMidGroups; g += 64) {
                const u64* w = parts + (size_t)g * kMidWords + 2 * wid;
                u64 x = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                u64 y = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (ok && ((x & ~P) != tag || (x & P) == P || (y & ~P) != tag || (y & P) == P)) {
                    if (wall_clock64() - t0 > spin_ticks) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    x = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    y = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                acc = add2(acc, mk2(x & P, y & P));
            }
            fp2 v[1] = {acc};
            wave_sum<1>(v);
            if (lane == 0) s_tot[wid] = v[0];
            if (!__all(ok) && lane == 0) *s_flag = 0;
        }
        __syncthreads();
        if (t == 0) {
            int ok = *s_flag;
            u64 rv[NA];
#pragma unroll
            for (int i = 0; i < NA; ++i) rv[i] = P;   // p: the wait expired
            if (ok) {
                fp2 v[N];
#pragma unroll
                for (int j = 0; j < N; ++j) v[j] = s_tot[j];
                publish_tagged<N>(host_vals, v, seq0 + (unsigned)e);
                if (!last) ok = take_answer_n<NA>(answer, seq0 + (unsigned)e, t0, spin_ticks, rv);
            }
            if (!ok) __hip_atomic_store(host_seq, kSeqAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (!last || !ok) {
                // relay to the other workgroups (p under the tag: leave)
#pragma unroll
                for (int i = 0; i < NA; ++i)
                    __hip_atomic_store(relay + i, (ok ? rv[i] : P) | tag, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            if (trace) trace[((size_t)e * kMidGroups) * 2 + 1] = wall_clock64();
#pragma unroll
            for (int i = 0; i < NA; ++i) s_r[i] = rv[i];
            *s_flag = ok;
        }
    } else if (t == 0) {
        int ok = 1;
        if (!last) {
            // wait for workgroup 0's relay
            for (;;) {
                u64 x[NA];
                bool all = true;
#pragma unroll
                for (int i = 0; i < NA; ++i) {
                    x[i] = __hip_atomic_load(relay + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    all = all && (x[i] & ~P) == tag;
                }
                if (all) {
#pragma unroll
                    for (int i = 0; i < NA; ++i) {
                        s_r[i] = x[i] & P;
                        if (s_r[i] == P) ok = 0;
                    }
                    break;
                }
                if (wall_clock64() - t0 > spin_ticks) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (trace) trace[((size_t)e * kMidGroups + b) * 2 + 1] = wall_clock64();
        }
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag;
}

__global__ void __launch_bounds__(kTailThreads) k_sc2_mid(
    const V2* __restrict__ fin, const V2* __restrict__ gin, unsigned slice_in, int n_rounds,
    u64 rr, u64 ri, u64* parts, fp2* host_vals, unsigned* host_seq, unsigned seq0,
    u64* answer, long long spin_ticks, V2* __restrict__ fout,
    V2* __restrict__ gout, u64* relay, long long* trace, unsigned* gate, long long gate_ticks,
    unsigned r_want, int r_from_slot, u64* arelay)
{
    extern __shared__ V2 tl[];
    __shared__ u64 s_r[4];
    __shared__ int s_ok;
    __shared__ fp2 s_sum[9], s_tot[9];
    const unsigned half = slice_in / 2;
    V2* tf = tl;
    V2* tg = tl + half;
    const int t = threadIdx.x, lane = t & 63;
    const unsigned b = blockIdx.x;
    // Arrival gate: the rounds need every workgroup resident at once.  Each
    // counts itself into gate[0]; workgroup 0 waits up to gate_ticks for all
    // of them and sets gate[1] to kGateGo, or to kGateAbort when some are
    // missing (a GPU shared with other work); a workgroup that sees no
    // decision within 2 gate_ticks sets kGateAbort itself.  Both settings are
    // compare-and-swaps from kGatePending, so there is one decision; on abort
    // every workgroup leaves and the host proves these rounds with streaming
    // launches instead.
    if (t == 0) {
        __hip_atomic_fetch_add(gate, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = wall_clock64();
        unsigned d = kGatePending;
        if (b == 0) {
            while (__hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x &&
                   wall_clock64() - t0 <= gate_ticks)
                __builtin_amdgcn_s_sleep(1);
            const unsigned want =
                __hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= gridDim.x
                    ? kGateGo : kGateAbort;
            unsigned e = kGatePending;   // on failure e holds the decision already made
            d = __hip_atomic_compare_exchange_strong(gate + 1, &e, want, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    ? want : e;
        } else {
            while ((d = __hip_atomic_load(gate + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ==
                       kGatePending &&
                   wall_clock64() - t0 <= 2 * gate_ticks)
                __builtin_amdgcn_s_sleep(1);
            if (d == kGatePending) {
                unsigned e = kGatePending;
                if (__hip_atomic_compare_exchange_strong(gate + 1, &e, kGateAbort, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    d = kGateAbort;
                else
                    d = e;
            }
        }
        if (d != kGateGo && b == 0)
            __hip_atomic_store(host_seq, kSeqGateAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        s_ok = d == kGateGo;
    }
    __syncthreads();
    if (!s_ok) return;
    // queued ahead of its transcript step: the challenge of the round before
    // the first comes through the answer slot, as a streaming launch's does
    if (r_from_slot) {
        TicketPub a;
        a.host_seq = host_seq;
        a.spin_ticks = spin_ticks;
        a.answer = answer;
        a.want = r_want;
        a.n_ans = 2;
        a.relay = arelay;
        u64 rw[4] = {0, 0, 0, 0};
        if (!launch_challenges(a, rw)) return;
        rr = rw[0];
        ri = rw[1];
    }
    const V2* sf = fin + (size_t)b * slice_in;   // T_{k-np}
    const V2* sg = gin + (size_t)b * slice_in;
    fp2 ra = mk2(rr, ri), rb = zero2(), rab = zero2();
    int np = 1;
    unsigned len = slice_in;
    const int ne = mid_exchanges(n_rounds);
    for (int e = 0; e < ne; ++e) {
        const int nre = mid_exchange_rounds(n_rounds, e);
        const bool last = e == ne - 1;
        int ok;
        if (nre == 2) {
            exchange_work<9>(sf, sg, e > 0, len, np, ra, rb, rab, tf, tg, s_sum);
            ok = mid_exchange<9, 4>(s_sum, s_tot, &s_ok, s_r, e, last, b, parts, host_vals,
                                    host_seq, seq0, answer, spin_ticks, relay, trace);
        } else {
            exchange_work<3>(sf, sg, e > 0, len, np, ra, rb, rab, tf, tg, s_sum);
            ok = mid_exchange<3, 2>(s_sum, s_tot, &s_ok, s_r, e, last, b, parts, host_vals,
                                    host_seq, seq0, answer, spin_ticks, relay, trace);
        }
        if (!ok) return;
        len >>= np;   // the table at tf / tg is T_k now
        np = nre;
        sf = tf;
        sg = tg;
        if (!last) {
            ra = uniform2(mk2(s_r[0], s_r[1]));
            rb = uniform2(mk2(s_r[2], s_r[3]));
            rab = uniform2(mul2(ra, rb));
        }
        __syncthreads();   // s_r and s_sum are rewritten by the next exchange
    }
    // the slice's last table goes out (its round's challenge is the next launch's)
    for (unsigned j = t; j < len; j += kTailThreads) {
        fout[(size_t)b * len + j] = tf[j];
        gout[(size_t)b * len + j] = tg[j];
    }
    if (b == 0 && t < 64) {
        // every exchange done: workgroup 0 has gathered every workgroup's last
        // sums, so all of them are past the gate and every relay.  It leaves
        // the words as k_sc2_mid_reset would, and the next mid launch needs
        // no reset launch of its own.
        for (int j = lane; j < kMidGroups * kMidWords; j += 64)
            __hip_atomic_store(parts + j, ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lane < kMaxAns && r_from_slot)   // every workgroup took the first challenge long ago
            __hip_atomic_store(arelay + lane, ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lane < 4) __hip_atomic_store(relay + lane, ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (lane == 0) {
            __hip_atomic_store(gate, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(gate + 1, kGatePending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void __launch_bounds__(kTailThreads) k_sc2_tail_host(
    const V2* __restrict__ fin, const V2* __restrict__ gin, unsigned len_in, int k0, int nv,
    u64 rr, u64 ri, fp2* host_vals, unsigned* host_seq, unsigned seq0, u64* answer,
    long long spin_ticks, long long* trace, unsigned r_want, int r_from_slot)
{
    extern __shared__ V2 tl[];
    __shared__ u64 s_r[4];
    __shared__ int s_ok;
    __shared__ fp2 s_sum[9];
    V2* tf = tl;
    V2* tg = tl + kTailIn / 2;
    const int t = threadIdx.x;
    if (r_from_slot) {
        // queued before the transcript step of r = r_{k0-1}: the host's answer
        // under sequence number r_want, unless the host gives up first
        if (t == 0) {
            const u64 tag = (u64)(r_want & 7) << 61;
            const long long t0 = wall_clock64();
            int ok = 0;
            for (;;) {
                const u64 a0 = __hip_atomic_load(answer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                const u64 a1 = __hip_atomic_load(answer + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                if ((a0 & ~P) == tag && (a0 & P) != P && (a1 & ~P) == tag && (a1 & P) != P) {
                    s_r[0] = a0 & P;
                    s_r[1] = a1 & P;
                    ok = 1;
                    break;
                }
                if (__hip_atomic_load(host_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == kSeqHostAbort)
                    break;
                if (wall_clock64() - t0 > spin_ticks) {
                    __hip_atomic_store(host_seq, kSeqAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            s_ok = ok;
        }
        __syncthreads();
        if (!s_ok) return;
        rr = s_r[0];
        ri = s_r[1];
        __syncthreads();   // s_r and s_ok are rewritten by the exchanges
    }
    fp2 ra = mk2(rr, ri), rb = zero2(), rab = zero2();
    int np = 1;                 // challenges not yet applied to the table at sf / sg
    const V2* sf = fin;         // T_{k-np}
    const V2* sg = gin;
    unsigned len = len_in;      // its entries per table
    for (int k = k0, e = 0; k < nv; ++e) {
        if (trace && t == 0) trace[3 * e] = wall_clock64();
        const bool pair = nv - k >= 2;
        if (pair)
            exchange_work<9>(sf, sg, sf != fin, len, np, ra, rb, rab, tf, tg, s_sum);
        else
            exchange_work<3>(sf, sg, sf != fin, len, np, ra, rb, rab, tf, tg, s_sum);
        if (t == 0) {
            // lane 0 publishes the sums and waits for the challenges
            if (trace) trace[3 * e + 1] = wall_clock64();
            const unsigned want = seq0 + (unsigned)e;
            u64 rv[4] = {0, 0, 0, 0};
            if (pair) {
                fp2 v[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) v[i] = s_sum[i];
                publish_tagged<9>(host_vals, v, want);
                s_ok = tail_wait<4>(want, host_seq, answer, spin_ticks, trace, e, rv);
            } else {
                const fp2 v[3] = {s_sum[0], s_sum[1], s_sum[2]};
                publish_tagged<3>(host_vals, v, want);
                u64 r2[2] = {0, 0};
                s_ok = tail_wait<2>(want, host_seq, answer, spin_ticks, trace, e, r2);
                rv[0] = r2[0];
                rv[1] = r2[1];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) s_r[i] = rv[i];
        }
        __syncthreads();   // the challenges visible to every thread
        if (!s_ok) return;
        ra = uniform2(mk2(s_r[0], s_r[1]));
        rb = uniform2(mk2(s_r[2], s_r[3]));
        rab = uniform2(mul2(ra, rb));
        len >>= np;   // the table at sf / sg is T_k now
        np = pair ? 2 : 1;
        k += np;
        sf = tf;
        sg = tg;
        if (k == nv && t == 0) {
            // T_{nv-np} (2^np entries) folded by the last challenges
            const fp2 out[2] = {fold_head(sf, np, ra, rb, rab), fold_head(sg, np, ra, rb, rab)};
            publish_tagged<2>(host_vals, out, seq0 + (unsigned)(e + 1));
        }
    }
    if (k0 == nv && t == 0) {
        // no round here: the input's two entries folded by r_{nv-1}
        const fp2 out[2] = {fold_head(fin, 1, ra, rb, rab), fold_head(gin, 1, ra, rb, rab)};
        publish_tagged<2>(host_vals, out, seq0);
    }
}


// Event-timed launch classes of one prove (timing on).
enum Sc2Class { SC_FIRST, SC_FOLD01, SC_FOLD, SC_TAIL, SC_MID, SC_PAIR, SC_NCLASS };

struct Sc2State {
    DevBuf work;                 // T_k ping-pong: A (n/2) and B (n/4), for f and g
    DevBuf partials;             // up to 9 fp2 per workgroup
    // mapped pinned memory: up to 16 published values at 0 (tagged words), the
    // sequence word (abort marks; k_publish_* launches) at 256, the answer (a
    // challenge, two tagged words) at 512
    fp2* host_vals = nullptr;
    unsigned* host_seq = nullptr;
    u64* host_answer = nullptr;         // mapped answer slot: 2 tagged words
    fp2* dev_vals = nullptr;     // device aliases of the mapped memory
    unsigned* dev_seq = nullptr;
    u64* dev_answer = nullptr;
    long long spin_ticks = 0;    // the tail's wait bound in wall-clock ticks
    double clock_mhz = 100;      // wall-clock rate
    long long* host_trace = nullptr;   // mapped: 3 stamps per tail round (timing on)
    long long* dev_trace = nullptr;
    int tail_rounds = 0;         // tail profile of the last timed prove
    double tail_work_us = 0, tail_wait_us = 0, tail_host_us = 0;
    double tail_round_us[40] = {};   // device us of each tail round
    int mid_rounds = 0;          // mid profile of the last timed prove
    double mid_host_us = 0, mid_relay_us = 0, mid_work_us = 0, mid_spread_us = 0;
    unsigned seq = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t ev_round = nullptr;   // end of a mid launch
    bool timed = false;          // per-launch event timing
    double ms[SC_NCLASS] = {};
    double bytes[SC_NCLASS] = {};
    int launches[SC_NCLASS] = {};
    DevBuf mid_parts;            // the mid rounds' tagged partials, [kMidGroups][6]
    DevBuf relay;                // the mid rounds' challenge relay (2 words)
    DevBuf arelay;               // queued launches' answer relay (kMaxAns words, launch_challenges)
    bool ahead = false;          // untimed proves queue their launches ahead (tp_sc2_set_ahead)
    bool blocks = true;          // rounds 0 to 3 from the block products (tp_sc2_set_block_pass)
    DevBuf gate;                 // the mid launch's arrival gate (2 words)
    fp2 pub[16];                 // the last published round's values, read out of the slot
    DevBuf slots;                // the streaming launches' tagged sums (ticket_publish)
    DevBuf ticket;               // their ticket counter
    bool slots_clean = false;    // slots all ones and ticket 0 (the last prove ended cleanly)
    bool mid_clean = false;      // mid words reset by the last mid launch, which finished
    long long gate_ticks = 0;    // its wait bound in wall-clock ticks
    int mid_fallbacks = 0;       // mid launches that did not start whole (last prove)
    long long* host_mtrace = nullptr;   // mapped: [40][kMidGroups][2] stamps (timing on)
    long long* dev_mtrace = nullptr;
    DevBuf blk_slots;            // the block products of each k_sc2_blocks_mfma workgroup
    fp2* host_blk = nullptr;     // mapped: the 256 block products as tagged words
    fp2* dev_blk = nullptr;
};
Sc2State g_sc2;

// Tagged words of the streaming launches, two words a sum: up to
// kMaxPubGroups workgroups x 3 sums, or kFold2QuadGroups x 9.
constexpr size_t kSlotWords = (size_t)kFold2QuadGroups * 9 * 2;
static_assert(kFirst01Groups * 9 * 2 <= kSlotWords && kMfGroups * 9 * 2 <= kSlotWords &&
                  kMaxPubGroups * 3 * 2 <= kSlotWords && kFold01Groups <= kMaxPubGroups &&
                  kFoldGroups <= kMaxPubGroups && kFold2QuadGroups <= kMaxPubGroups,
              "ticket slots");

bool sc2_setup(size_t n)
{
    Sc2State& S = g_sc2;
    if (!S.host_vals) {
        void* hv = nullptr;
        TP_HIP_OK(hipHostMalloc(&hv, 4096, hipHostMallocMapped | hipHostMallocCoherent));
        char* hc = static_cast<char*>(hv);
        S.host_vals = reinterpret_cast<fp2*>(hc);
        S.host_seq = reinterpret_cast<unsigned*>(hc + 256);
        S.host_answer = reinterpret_cast<u64*>(hc + 512);
        S.host_trace = reinterpret_cast<long long*>(hc + 1024);   // 3 x 40 stamps
        *S.host_seq = 0;
        S.host_answer[0] = S.host_answer[1] = kSlotEmpty;
        void* dv = nullptr;
        TP_HIP_OK(hipHostGetDevicePointer(&dv, hv, 0));
        char* dc = static_cast<char*>(dv);
        S.dev_vals = reinterpret_cast<fp2*>(dc);
        S.dev_seq = reinterpret_cast<unsigned*>(dc + 256);
        S.dev_answer = reinterpret_cast<u64*>(dc + 512);
        S.dev_trace = reinterpret_cast<long long*>(dc + 1024);
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_sc2_tail_host,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLds));
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_sc2_first01_mfma,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfLds));
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_sc2_blocks_mfma,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBkLds));
        void* hb = nullptr;
        TP_HIP_OK(hipHostMalloc(&hb, 256 * sizeof(fp2), hipHostMallocMapped | hipHostMallocCoherent));
        S.host_blk = static_cast<fp2*>(hb);
        for (int i = 0; i < 512; ++i) reinterpret_cast<u64*>(hb)[i] = kSlotEmpty;
        void* db = nullptr;
        TP_HIP_OK(hipHostGetDevicePointer(&db, hb, 0));
        S.dev_blk = static_cast<fp2*>(db);
        TP_HIP_OK(hipFuncSetAttribute((const void*)k_sc2_mid,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTailLds));
        void* tp_ = nullptr;
        TP_HIP_OK(hipHostMalloc(&tp_, 40 * kMidGroups * 2 * sizeof(long long),
                                hipHostMallocMapped | hipHostMallocCoherent));
        S.host_mtrace = static_cast<long long*>(tp_);
        void* dtp = nullptr;
        TP_HIP_OK(hipHostGetDevicePointer(&dtp, tp_, 0));
        S.dev_mtrace = static_cast<long long*>(dtp);
        int khz = 0, dev = 0;
        TP_HIP_OK(hipGetDevice(&dev));
        TP_HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        S.spin_ticks = (long long)(khz > 0 ? khz : 100000) * 1000LL * 10;   // 10 s
        S.clock_mhz = (khz > 0 ? khz : 100000) / 1000.0;
        S.gate_ticks = (long long)(khz > 0 ? khz : 100000) * 20;   // 20 ms
    }
    if (!S.ev[0]) {
        TP_HIP_OK(hipEventCreate(&S.ev[0]));
        TP_HIP_OK(hipEventCreate(&S.ev[1]));
        TP_HIP_OK(hipEventCreateWithFlags(&S.ev_round, hipEventDisableTiming));
    }
    if (!S.work.ensure(2 * (n / 2 + n / 4 + 2) * sizeof(V2))) return false;
    if (!S.partials.ensure(2048 * 9 * sizeof(fp2))) return false;
    if (!S.relay.ensure(4 * sizeof(u64))) return false;
    if (!S.arelay.ensure(kMaxAns * sizeof(u64))) return false;
    if (!S.mid_parts.ensure(kMidGroups * kMidWords * sizeof(u64))) return false;
    if (!S.gate.ensure(2 * sizeof(unsigned))) return false;
    if (!S.slots.ensure(kSlotWords * sizeof(u64))) return false;
    if (!S.ticket.ensure(kTicketBytes)) return false;
    if (!S.blk_slots.ensure((size_t)kBkGroups * 256 * sizeof(fp2))) return false;
    return true;
}

unsigned blocks_for(size_t items)
{
    // 2048 workgroups (8 per CU on 256 CUs) saturate HBM; fewer for small rounds.
    size_t b = (items + kBlock - 1) / kBlock;
    if (b > 2048) b = 2048;
    return (unsigned)(b ? b : 1);
}

// Waits until a k_publish_* launch (round 0 of tables under 8 entries) has
// published sequence number `seq`; checks the stream for errors while waiting
// and gives up after a generous timeout.
bool wait_seq(hipStream_t st, unsigned seq)
{
    volatile unsigned* hs = g_sc2.host_seq;
    const auto t0 = std::chrono::steady_clock::now();
    long spins = 0;
    while (*hs != seq) {
        if (*hs == kSeqAbort) return fail("sumcheck: a kernel gave up waiting (round %u)", seq);
        if (++spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail("sumcheck stream error: %s", hipGetErrorString(q));
            if (q == hipSuccess && *hs != seq)
                return fail("sumcheck: stream idle but round %u not published", seq);
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30))
                return fail("sumcheck: timed out waiting for round %u", seq);
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return true;
}

// Waits until the device has published round `seq` as n tagged Fp2 values
// (publish_tagged), copies them to pub[] and empties the slot.  Returns 1;
// 2 when `gate` and the mid launch gave way at its arrival gate instead; 0
// on failure (a kernel's abort mark, a stream error, the stream idle without
// the values, or 30 s).
int wait_pub(hipStream_t st, unsigned seq, int n, bool gate)
{
    Sc2State& S = g_sc2;
    volatile u64* w = reinterpret_cast<volatile u64*>(S.host_vals);
    volatile unsigned* hs = S.host_seq;
    const u64 tag = (u64)(seq & 7) << 61;
    auto ready = [&]() {
        for (int i = 0; i < 2 * n; ++i) {
            const u64 x = w[i];
            if ((x & ~P) != tag || (x & P) == P) return false;
        }
        return true;
    };
    const auto t0 = std::chrono::steady_clock::now();
    for (long spins = 1; !ready(); ++spins) {
        const unsigned v = *hs;
        if (gate && v == kSeqGateAbort) return 2;
        if (v == kSeqAbort) return fail("sumcheck: a kernel gave up waiting (round %u)", seq), 0;
        if (spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail("sumcheck stream error: %s", hipGetErrorString(q)), 0;
            if (q == hipSuccess && !ready() && !(gate && *hs == kSeqGateAbort))
                return fail("sumcheck: stream idle but round %u not published", seq), 0;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30))
                return fail("sumcheck: timed out waiting for round %u", seq), 0;
        }
    }
    for (int i = 0; i < n; ++i) S.pub[i] = mk2(w[2 * i] & P, w[2 * i + 1] & P);
    for (int i = 0; i < 2 * n; ++i) w[i] = kSlotEmpty;
    return 1;
}

// Waits until k_sc2_blocks_reduce has published the 256 block products as
// sequence number `seq` (tagged words) and copies them to prod; the caller
// empties the words (empty_blocks) once the next launch is queued.  Returns
// false on a kernel abort mark, a stream error, the stream idle without
// them, or 30 s.
bool wait_blocks(hipStream_t st, unsigned seq, fp2 (&prod)[16][16])
{
    Sc2State& S = g_sc2;
    volatile u64* w = reinterpret_cast<volatile u64*>(S.host_blk);
    const u64 tag = (u64)(seq & 7) << 61;
    int next = 0;   // words before this one already carry the tag
    auto ready = [&]() {
        for (; next < 512; ++next) {
            const u64 x = w[next];
            if ((x & ~P) != tag || (x & P) == P) return false;
        }
        return true;
    };
    const auto t0 = std::chrono::steady_clock::now();
    for (long spins = 1; !ready(); ++spins) {
        if (*S.host_seq == kSeqAbort) return fail("sumcheck: a kernel gave up waiting (round %u)", seq);
        if (spins % 4096 == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail("sumcheck stream error: %s", hipGetErrorString(q));
            if (q == hipSuccess && !ready())
                return fail("sumcheck: stream idle but block products %u not published", seq);
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30))
                return fail("sumcheck: timed out waiting for block products %u", seq);
        }
    }
    for (int i = 0; i < 256; ++i) prod[i / 16][i % 16] = mk2(w[2 * i] & P, w[2 * i + 1] & P);
    return true;
}

// All ones in the block products' words (value bits p: no sequence's word).
void empty_blocks()
{
    volatile u64* w = reinterpret_cast<volatile u64*>(g_sc2.host_blk);
    for (int i = 0; i < 512; ++i) w[i] = kSlotEmpty;
}

// The n values a k_publish_* launch left under sequence number seq (read
// after wait_seq), copied to pub[]; the slot is emptied like wait_pub's.
void take_seq_vals(int n)
{
    Sc2State& S = g_sc2;
    volatile u64* w = reinterpret_cast<volatile u64*>(S.host_vals);
    for (int i = 0; i < n; ++i) S.pub[i] = mk2(w[2 * i], w[2 * i + 1]);
    for (int i = 0; i < 2 * n; ++i) w[i] = kSlotEmpty;
}

// c0 + r (c1 - c0 - c2) + r^2 c2: a round's sum at challenge r of the round
// before, from the three sums over the quads of that round's table whose
// products give its coefficients (see the file header).
fp2 quad_at(fp2 r, fp2 c0, fp2 c1, fp2 c2)
{
    return add2(c0, mul2(r, add2(sub2(sub2(c1, c0), c2), mul2(r, c2))));
}

// Waits until the launches before `ev` on its stream are done.
bool wait_event(hipEvent_t ev)
{
    const auto t0 = std::chrono::steady_clock::now();
    for (long spins = 1;; ++spins) {
        const hipError_t q = hipEventQuery(ev);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) return fail("sumcheck stream error: %s", hipGetErrorString(q));
        if (spins % 4096 == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30))
            return fail("sumcheck: timed out waiting for a round");
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return true;
}


}  // namespace

extern "C" {

// Degree-2 sumcheck prover over resident device tables.
//   d_f, d_g : 2^nv Fp2 values each, interleaved (re, im), canonical; read only.
//   seed     : the transcript is FiatShamir("truepix") after put(["c2", seed]);
//              round k absorbs put([a, b, c]) and draws r_k = rand_ext("c2-<k>").
//   fixed    : optional 2*nv challenges to use instead of the transcript.
//   polys    : out, nv x (a, b, c) interleaved (6 u64 per round).
//   chals    : out, nv x (re, im).
//   final4   : out, f~(r) and g~(r).
//   stream   : hipStream_t or NULL for the library stream.
// Returns 0 on success, -1 on failure (tp_last_error()).
int tp_sc2_prove(const uint64_t* d_f, const uint64_t* d_g, int nv, uint64_t seed,
                 const uint64_t* fixed, uint64_t* polys, uint64_t* chals, uint64_t* final4,
                 void* stream_handle)
{
    tp::clear_error();
    if (nv < 1 || nv > 34) return tp::fail("tp_sc2_prove: nv=%d out of range", nv), -1;
    if (!tp::device_ready()) return -1;
    hipStream_t st = stream_handle ? (hipStream_t)stream_handle : tp::stream();
    const size_t n = (size_t)1 << nv;
    if (!sc2_setup(n)) return -1;
    // its mid and tail launches wait on this thread's answers: no frees until
    // the prove returns
    tp::ResidentPhase resident;
    Sc2State& S = g_sc2;
    for (int c = 0; c < SC_NCLASS; ++c) S.ms[c] = S.bytes[c] = 0, S.launches[c] = 0;
    if (!S.slots_clean) {
        // first prove, or the last one stopped part way: fresh slots and ticket
        TP_HIP_OK(hipMemsetAsync(S.slots.ptr, 0xff, kSlotWords * sizeof(u64), st));
        TP_HIP_OK(hipMemsetAsync(S.ticket.ptr, 0, kTicketBytes, st));
        TP_HIP_OK(hipMemsetAsync(S.arelay.ptr, 0xff, kMaxAns * sizeof(u64), st));
        empty_blocks();
    }
    S.slots_clean = false;   // until this prove returns 0
    {
        // no answer pending, no values published, no abort mark
        volatile u64* ans = S.host_answer;
        for (int i = 0; i < kMaxAns; ++i) ans[i] = kSlotEmpty;
        volatile u64* w = reinterpret_cast<volatile u64*>(S.host_vals);
        for (int i = 0; i < 32; ++i) w[i] = kSlotEmpty;
        *reinterpret_cast<volatile unsigned*>(S.host_seq) = 0;
    }
    auto pub_for = [&](unsigned seq, unsigned want = 0, int n_ans = 0) {
        TicketPub a;
        a.slots = S.slots.as<u64>();
        a.ticket = S.ticket.as<unsigned>();
        a.host_vals = S.dev_vals;
        a.host_seq = S.dev_seq;
        a.seq = seq;
        a.spin_ticks = S.spin_ticks;
        a.answer = S.dev_answer;
        a.want = want;
        a.n_ans = n_ans;
        a.relay = S.arelay.as<u64>();
        return a;
    };
    // Rounds 0 and 1 from one pass over the input: on the matrix cores when
    // the quads come in whole chunks of 64 and their tile sums fit int32,
    // else by the VALU pass.
    auto launch_first01 = [&](unsigned seq) {
        const V2* cf = reinterpret_cast<const V2*>(d_f);
        const V2* cg = reinterpret_cast<const V2*>(d_g);
        const size_t nq = n / 4;
        if (nq % 64 == 0 && nq >= (size_t)64 * kMfWaves && nq <= kMfMaxQuads) {
            const unsigned nbm = (unsigned)std::min<size_t>(kMfGroups, nq / 64 / kMfWaves);
            k_sc2_first01_mfma<<<nbm, kMfThreads, kMfLds, st>>>(cf, cg, nq, pub_for(seq));
        } else {
            const unsigned nb1 = (unsigned)std::min<size_t>(
                kFirst01Groups, (nq + kFirst01Threads - 1) / kFirst01Threads);
            k_sc2_first01<<<nb1, kFirst01Threads, 0, st>>>(cf, cg, nq, pub_for(seq));
        }
    };
    // The host's answer for the launch waiting on sequence number `want`:
    // n Fp2 challenges as tagged words (the tags written with the values, so
    // a launch never takes a half-written answer for this one).
    auto answer = [&](unsigned want, const fp2* rs, int n) {
        volatile u64* ans = S.host_answer;
        const u64 tag = (u64)(want & 7) << 61;
        for (int i = 0; i < n; ++i) {
            ans[2 * i] = rs[i].re | tag;
            ans[2 * i + 1] = rs[i].im | tag;
        }
    };

    fp2* partials = S.partials.as<fp2>();
    V2* A = S.work.as<V2>();             // f: n/2, then g: n/2
    V2* B = A + 2 * (n / 2 + 1);         // f: n/4, then g: n/4
    const size_t capA = n / 2 + 1, capB = n / 4 + 1;

    FsTranscriptT<HostSha256> fs;
    fs.init("truepix");
    fs.item_str("c2", 2);   // put(["c2", seed])
    fs.item_int(seed);

    // Records round k's polynomial from its sums and returns r_k.
    auto round_done = [&](int k, fp2 E0, fp2 E1, fp2 EA) {
        const fp2 a = EA, c = E0, b = sub2(sub2(E1, EA), E0);
        polys[6 * k + 0] = a.re; polys[6 * k + 1] = a.im;
        polys[6 * k + 2] = b.re; polys[6 * k + 3] = b.im;
        polys[6 * k + 4] = c.re; polys[6 * k + 5] = c.im;
        fp2 r;
        if (fixed) {
            r = mk2(fixed[2 * k], fixed[2 * k + 1]);
        } else {
            fs.item_fp2(a);   // put([a, b, c])
            fs.item_fp2(b);
            fs.item_fp2(c);
            char lab[24];
            const uint32_t ln = label_num(lab, "c2-", k);
            r = fs.rand_ext(lab, ln);
        }
        chals[2 * k] = r.re;
        chals[2 * k + 1] = r.im;
        return r;
    };

    // The host side of a mid launch's exchanges (mid_exchange_rounds): from
    // round k_first, sequence numbers seq0m + e; the first exchange's values
    // are already in S.pub when first_done.  Every exchange but the last is
    // answered (the last round's challenge goes to the next launch); r is the
    // last round's challenge.  False on failure.
    auto mid_host = [&](int k_first, unsigned seq0m, int nrm, bool first_done, fp2& r_out,
                        bool answer_last = false) -> bool {
        const int ne = mid_exchanges(nrm);
        int kk = k_first;
        for (int e = 0; e < ne; ++e) {
            const int nre = mid_exchange_rounds(nrm, e);
            if ((e > 0 || !first_done) && wait_pub(st, seq0m + e, nre == 2 ? 9 : 3, false) != 1)
                return false;
            const bool last = e + 1 == ne;
            if (nre == 2) {
                fp2 v[9];
                for (int i = 0; i < 9; ++i) v[i] = S.pub[i];
                fp2 rs[2];
                rs[0] = round_done(kk, add2(v[0], v[3]), add2(v[1], v[4]), add2(v[2], v[5]));
                rs[1] = round_done(kk + 1, quad_at(rs[0], v[0], v[1], v[2]),
                                   quad_at(rs[0], v[3], v[4], v[5]), quad_at(rs[0], v[6], v[7], v[8]));
                if (!last) answer(seq0m + e, rs, 2);
                r_out = rs[1];
            } else {
                r_out = round_done(kk, S.pub[0], S.pub[1], S.pub[2]);
                if (!last || answer_last) answer(seq0m + e, &r_out, 1);
            }
            kk += nre;
        }
        return true;
    };

    // The sliced mid launch over the input T_{k-1} (2m entries a table) with
    // r = r_{k-1}, or, when from_slot, r taken from the answer slot under
    // sequence number `want` (a launch queued before that transcript step).
    struct MidLaunch { unsigned seq0 = 0; int nr = 0; V2* of = nullptr; V2* og = nullptr; unsigned slice = 0; };
    auto launch_mid = [&](const V2* in_f, const V2* in_g, size_t m2, fp2 rr, unsigned want,
                          int from_slot, MidLaunch& ml) -> bool {
        ml.slice = (unsigned)(m2 / kMidGroups);   // > kMidEnd here
        ml.nr = 0;
        while ((ml.slice >> ml.nr) > kMidEnd) ++ml.nr;
        ml.of = (in_f >= A && in_f < A + 2 * capA) ? B : A;
        ml.og = ml.of + (ml.of == A ? capA : capB);
        ml.seq0 = S.seq + 1;
        S.seq += (unsigned)mid_exchanges(ml.nr);
        const V2* a_f = in_f;
        const V2* a_g = in_g;
        unsigned a_slice = ml.slice;
        int a_nr = ml.nr;
        u64 a_rr = rr.re, a_ri = rr.im;
        u64* a_parts = S.mid_parts.as<u64>();
        fp2* a_vals = S.dev_vals;
        unsigned* a_seq = S.dev_seq;
        unsigned a_seq0 = ml.seq0;
        u64* a_ans = S.dev_answer;
        long long a_spin = S.spin_ticks;
        V2* a_of = ml.of;
        V2* a_og = ml.og;
        u64* a_relay = S.relay.as<u64>();
        long long* a_trace = S.timed && mid_exchanges(ml.nr) <= 40 ? S.dev_mtrace : nullptr;
        unsigned* a_gate = S.gate.as<unsigned>();
        long long a_gate_ticks = S.gate_ticks;
        unsigned a_want = want;
        int a_from = from_slot;
        u64* a_arelay = S.arelay.as<u64>();
        void* args[] = {&a_f, &a_g, &a_slice, &a_nr, &a_rr, &a_ri, &a_parts, &a_vals, &a_seq,
                        &a_seq0, &a_ans, &a_spin, &a_of, &a_og, &a_relay, &a_trace,
                        &a_gate, &a_gate_ticks, &a_want, &a_from, &a_arelay};
        if (!S.mid_clean) {   // first mid launch, or the last one did not finish
            k_sc2_mid_reset<<<1, 256, 0, st>>>(a_parts, a_relay, a_gate);
            TP_HIP_OK(hipGetLastError());
        }
        S.mid_clean = false;
        TP_HIP_OK(hipLaunchKernel((const void*)k_sc2_mid, dim3(kMidGroups), dim3(kTailThreads),
                                  args, (unsigned)(ml.slice * sizeof(V2)), st));
        return true;
    };
    // A mid launch queued behind the pass before it (untimed proves), for the
    // loop's round mid_pre_k.
    MidLaunch mid_pre;
    int mid_pre_k = -1;
    // The one-workgroup tail over T_{k-1} (2m entries) from round k, r =
    // r_{k-1} or from the answer slot (from_slot: under sequence number want).
    auto launch_tail = [&](const V2* in_f, const V2* in_g, size_t m2, int k, fp2 rr, unsigned seq0t,
                           unsigned want, int from_slot) -> bool {
        k_sc2_tail_host<<<1, kTailThreads, kTailLds, st>>>(
            in_f, in_g, (unsigned)m2, k, nv, rr.re, rr.im, S.dev_vals, S.dev_seq, seq0t,
            S.dev_answer, S.spin_ticks, S.timed ? S.dev_trace : nullptr, want, from_slot);
        TP_HIP_OK(hipGetLastError());
        return true;
    };
    // A tail launch queued behind the mid launch (untimed proves), for the
    // loop's round tail_pre_k.
    int tail_pre_k = -1;
    unsigned tail_pre_seq0 = 0;

    // Launch bracket: events around one launch of class c (timing on).
    auto t_begin = [&]() -> bool {
        if (S.timed) TP_HIP_OK(hipEventRecord(S.ev[0], st));
        return true;
    };
    auto t_end = [&](int c, double bytes) -> bool {
        if (S.timed) TP_HIP_OK(hipEventRecord(S.ev[1], st));
        S.bytes[c] += bytes;
        S.launches[c] += 1;
        return true;
    };
    auto t_collect = [&](int c) -> bool {   // after the round's wait
        if (!S.timed) return true;
        TP_HIP_OK(hipEventSynchronize(S.ev[1]));
        float ms = 0;
        TP_HIP_OK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
        S.ms[c] += ms;
        return true;
    };

    const V2* cf = reinterpret_cast<const V2*>(d_f);
    const V2* cg = reinterpret_cast<const V2*>(d_g);
    size_t m;       // length of T_k at the top of the loop below
    int k0;         // first round the loop handles
    fp2 r;          // challenge of the round before k0
    bool use_mid = true;   // cleared when the mid launch did not start whole
    S.mid_fallbacks = 0;
    // With `ahead` set, an untimed prove queues every launch down to the mid
    // launch at once, each taking its challenges from the answer slot, so
    // only the transcript steps (not a launch each) sit between the rounds.
    // Measured on MI355X it is no faster (0.571-0.574 against 0.568-0.572 ms
    // a prove, the same runs): a launch behind the host's transcript step
    // starts about as soon as a queued one sees its answer.  So by default
    // each launch is made after the step before it; timed proves always are,
    // so that each launch's events hold its kernel alone.
    const bool ahead = !S.timed && S.ahead && nv >= 3;
    // If the host stops while launches queued ahead still wait for their
    // challenges, it marks the prove failed in the sequence word (they
    // leave at once) and waits for the stream, so no launch of this prove
    // outlives it to take a later prove's answers.
    struct AheadGuard {
        bool armed = false;
        volatile unsigned* seq = nullptr;
        hipStream_t st = nullptr;
        ~AheadGuard()
        {
            if (!armed) return;
            *seq = kSeqHostAbort;
            (void)hipStreamSynchronize(st);
        }
    } guard;
    guard.seq = S.host_seq;
    guard.st = st;
    if (ahead) {
        // rounds 0 and 1 from one pass, round 2 folds the input by r_0, r_1,
        // then the streaming rounds and the mid launch, all queued now
        const unsigned seq_a = ++S.seq, seq_b = ++S.seq;
        launch_first01(seq_a);
        TP_HIP_OK(hipGetLastError());
        S.bytes[SC_FIRST] += 32.0 * (double)n;
        S.launches[SC_FIRST] += 1;
        const unsigned nb2 = fold01_groups(n / 4);
        k_sc2_fold2_pairs<<<nb2, kFoldThreads, 0, st>>>(cf, cg, n / 4, 0, 0, 0, 0, B, B + capB,
                                                        pub_for(seq_b, seq_a, 4), 1);
        TP_HIP_OK(hipGetLastError());
        S.bytes[SC_FOLD01] += 40.0 * (double)n;
        S.launches[SC_FOLD01] += 1;
        struct Queued { int k; unsigned seq; };
        std::vector<Queued> qs;
        unsigned prev = seq_b;
        const V2* qf = B;
        const V2* qg = B + capB;
        size_t qm = n / 8;
        int qk = 3;
        while (2 * qm > kTailIn && 2 * qm > (size_t)kMidGroups * kTailIn) {
            const unsigned seq = ++S.seq;
            const size_t quads = qm / 2;
            V2* of = (qk & 1) ? A : B;
            V2* og = of + ((qk & 1) ? capA : capB);
            const unsigned nb =
                (unsigned)std::min<size_t>(kFoldGroups, (quads + kFoldThreads - 1) / kFoldThreads);
            k_sc2_fold_round<<<nb, kFoldThreads, 0, st>>>(qf, qg, quads, 0, 0, of, og,
                                                          pub_for(seq, prev, 2), (qk & 1) == 0);
            TP_HIP_OK(hipGetLastError());
            S.bytes[SC_FOLD] += 48.0 * (double)(2 * qm);
            S.launches[SC_FOLD] += 1;
            qs.push_back({qk, seq});
            prev = seq;
            qf = of;
            qg = og;
            qm /= 2;
            ++qk;
        }
        // the mid launch, if these sizes reach it (as the loop below decides)
        const bool mid_q = 2 * qm > kTailIn && 2 * qm <= (size_t)kMidGroups * kTailIn;
        unsigned mid_seq0 = 0;
        int mid_nr = 0;
        V2* mid_of = nullptr;
        V2* mid_og = nullptr;
        if (mid_q) {
            const unsigned slice = (unsigned)(2 * qm / kMidGroups);
            while ((slice >> mid_nr) > kMidEnd) ++mid_nr;
            mid_of = (qf >= A && qf < A + 2 * capA) ? B : A;
            mid_og = mid_of + (mid_of == A ? capA : capB);
            mid_seq0 = S.seq + 1;
            S.seq += (unsigned)mid_exchanges(mid_nr);
            const V2* a_f = qf;
            const V2* a_g = qg;
            unsigned a_slice = slice;
            int a_nr = mid_nr;
            u64 a_rr = 0, a_ri = 0;
            u64* a_parts = S.mid_parts.as<u64>();
            fp2* a_vals = S.dev_vals;
            unsigned* a_seq = S.dev_seq;
            unsigned a_seq0 = mid_seq0;
            u64* a_ans = S.dev_answer;
            long long a_spin = S.spin_ticks;
            V2* a_of = mid_of;
            V2* a_og = mid_og;
            u64* a_relay = S.relay.as<u64>();
            long long* a_trace = nullptr;
            unsigned* a_gate = S.gate.as<unsigned>();
            long long a_gate_ticks = S.gate_ticks;
            unsigned a_want = prev;
            int a_from = 1;
            u64* a_arelay = S.arelay.as<u64>();
            void* args[] = {&a_f, &a_g, &a_slice, &a_nr, &a_rr, &a_ri, &a_parts, &a_vals, &a_seq,
                            &a_seq0, &a_ans, &a_spin, &a_of, &a_og, &a_relay, &a_trace,
                            &a_gate, &a_gate_ticks, &a_want, &a_from, &a_arelay};
            if (!S.mid_clean) {   // first mid launch, or the last one did not finish
                k_sc2_mid_reset<<<1, 256, 0, st>>>(a_parts, a_relay, a_gate);
                TP_HIP_OK(hipGetLastError());
            }
            S.mid_clean = false;
            TP_HIP_OK(hipLaunchKernel((const void*)k_sc2_mid, dim3(kMidGroups), dim3(kTailThreads),
                                      args, (unsigned)(slice * sizeof(V2)), st));
            S.bytes[SC_MID] += 32.0 * (double)(2 * qm) + 64.0 * kMidGroups;
            S.launches[SC_MID] += 1;
        }
        guard.armed = true;
        // the transcript steps, each answering the launch that waits on it
        if (wait_pub(st, seq_a, 9, false) != 1) return -1;
        fp2 v[9];
        for (int i = 0; i < 9; ++i) v[i] = S.pub[i];
        fp2 rs[2];
        rs[0] = round_done(0, add2(v[0], v[3]), add2(v[1], v[4]), add2(v[2], v[5]));
        rs[1] = round_done(1, quad_at(rs[0], v[0], v[1], v[2]), quad_at(rs[0], v[3], v[4], v[5]),
                           quad_at(rs[0], v[6], v[7], v[8]));
        answer(seq_a, rs, 2);
        if (wait_pub(st, seq_b, 3, false) != 1) return -1;
        r = round_done(2, S.pub[0], S.pub[1], S.pub[2]);
        answer(seq_b, &r, 1);
        for (const Queued& q : qs) {
            if (wait_pub(st, q.seq, 3, false) != 1) return -1;
            r = round_done(q.k, S.pub[0], S.pub[1], S.pub[2]);
            answer(q.seq, &r, 1);
        }
        cf = qf;
        cg = qg;
        m = qm;
        k0 = qk;
        if (mid_q) {
            const int first =
                wait_pub(st, mid_seq0, mid_exchange_rounds(mid_nr, 0) == 2 ? 9 : 3, true);
            if (first == 0) return -1;
            if (first == 2) {
                // not every workgroup was resident: the loop below proves
                // these rounds with streaming launches (the tables were only read)
                TP_HIP_OK(hipEventRecord(S.ev_round, st));
                if (!wait_event(S.ev_round)) return -1;
                S.launches[SC_MID] -= 1;
                S.bytes[SC_MID] -= 32.0 * (double)(2 * qm) + 64.0 * kMidGroups;
                S.seq += 1;   // past the abort mark
                S.mid_fallbacks += 1;
                use_mid = false;
            } else {
                if (!mid_host(qk, mid_seq0, mid_nr, true, r)) return -1;
                // every round is published; the launch resets its words as it ends
                S.mid_clean = true;
                cf = mid_of;
                cg = mid_og;
                k0 = qk + mid_nr;
                m = (size_t)kMidGroups * kMidEnd / 2;
            }
        }
        guard.armed = false;   // every queued launch has had its answers
    } else if (S.blocks && nv >= kBlocksMinVars && n % 1024 == 0 && n <= kBkMaxEntries) {
        // rounds 0 to 3 from the 256 block products (one pass over the input
        // and a 256-workgroup reduction), then one pass folds the input by
        // r_0 .. r_3 into T_4 with round 4's sums
        unsigned seq = ++S.seq;
        const unsigned nbk = (unsigned)std::min<size_t>(kBkGroups, n / 1024);
        if (!t_begin()) return -1;
        k_sc2_blocks_mfma<<<nbk, kBkThreads, kBkLds, st>>>(cf, cg, n / 16, S.blk_slots.as<fp2>());
        k_sc2_blocks_reduce<<<256, 256, 0, st>>>(S.blk_slots.as<fp2>(), nbk, S.dev_blk, seq);
        if (!t_end(SC_FIRST, 32.0 * (double)n)) return -1;
        TP_HIP_OK(hipGetLastError());
        fp2 prod[16][16];
        if (!wait_blocks(st, seq, prod) || !t_collect(SC_FIRST)) return -1;
        fp2 rk[4];
        rounds_from_blocks(prod, round_done, rk);
        // launched after the host's four rounds: queued ahead of them (its
        // 1024 workgroups waiting for the four challenges through the answer
        // slot) the prove took 0.455-0.461 ms against 0.435-0.438 ms
        seq = ++S.seq;
        V2* of = A;
        V2* og = A + capA;
        const size_t out = n / 16;
        if (!t_begin()) return -1;
        k_sc2_fold16<<<fold16_groups(out), kFold16Threads, 0, st>>>(
            cf, cg, out, rk[0].re, rk[0].im, rk[1].re, rk[1].im, rk[2].re, rk[2].im, rk[3].re,
            rk[3].im, of, og, pub_for(seq));
        TP_HIP_OK(hipGetLastError());
        if (!t_end(SC_FOLD01, 34.0 * (double)n)) return -1;
        // an untimed prove queues the sliced mid launch (rounds from 5) now,
        // taking r_4 from the answer slot: its launch latency then overlaps
        // this pass and the round-4 transcript step (0.425-0.428 ms a prove
        // against 0.436-0.447 ms launching it after that step)
        const bool queue_mid = !S.timed && out > kTailIn && out <= (size_t)kMidGroups * kTailIn;
        if (queue_mid) {
            if (!launch_mid(of, og, out, zero2(), seq, 1, mid_pre)) return -1;
            if (!t_end(SC_MID, 32.0 * (double)out + 64.0 * kMidGroups)) return -1;
            mid_pre_k = 5;
            guard.armed = true;   // it waits for r_4 until answered or aborted
        }
        empty_blocks();   // while the fold runs
        if (wait_pub(st, seq, 3, false) != 1 || !t_collect(SC_FOLD01)) return -1;
        r = round_done(4, S.pub[0], S.pub[1], S.pub[2]);
        if (queue_mid) answer(seq, &r, 1);
        cf = of;
        cg = og;
        k0 = 5;
        m = out / 2;
    } else if (nv >= 3) {
        // rounds 0 and 1 from one pass, round 2 folds the input by r_0, r_1
        unsigned seq = ++S.seq;
        unsigned nb = 0;
        if (!t_begin()) return -1;
        launch_first01(seq);
        if (!t_end(SC_FIRST, 32.0 * (double)n)) return -1;
        TP_HIP_OK(hipGetLastError());
        if (wait_pub(st, seq, 9, false) != 1 || !t_collect(SC_FIRST)) return -1;
        fp2 v[9];
        for (int i = 0; i < 9; ++i) v[i] = S.pub[i];
        const fp2 r0 = round_done(0, add2(v[0], v[3]), add2(v[1], v[4]), add2(v[2], v[5]));
        // the round-1 quadratics at r_0
        const fp2 r1 = round_done(1, quad_at(r0, v[0], v[1], v[2]), quad_at(r0, v[3], v[4], v[5]),
                                  quad_at(r0, v[6], v[7], v[8]));
        // From round 2, passes that fold by two challenges: T_{k-2} into T_k
        // by r_{k-2}, r_{k-1}.  While T_k is larger than the sliced mid
        // rounds take, the pass publishes T_k's nine quad sums (rounds k and
        // k + 1) and the next pass folds by both; the last one publishes
        // round k's pair sums and the rounds below go on from r_k.  Each
        // pass writes the work buffer that does not hold its input and
        // sweeps the other way from the pass before it.
        fp2 ra = r0, rb = r1;
        size_t len = n;   // entries per table of T_{k-2}
        int cls = SC_FOLD01;
        int desc = 1;     // the first pass read the input bottom up
        for (int kp = 2;; kp += 2) {
            const size_t out = len / 4;   // entries per table of T_kp
            const bool quads = out > (size_t)kMidGroups * kTailIn && out % 64 == 0;
            seq = ++S.seq;
            V2* of = (cf >= A && cf < A + 2 * capA) ? B : A;
            V2* og = of + (of == A ? capA : capB);
            nb = fold01_groups(out, quads ? kFold2QuadGroups : kFold01Groups);
            if ((size_t)nb * (quads ? 9 : 3) * 2 > kSlotWords) {
                fail("sumcheck: %u workgroups exceed the ticket slots", nb);
                return -1;
            }
            if (!t_begin()) return -1;
            if (quads)
                k_sc2_fold2_quads<<<nb, kFoldThreads, 0, st>>>(cf, cg, out, ra.re, ra.im, rb.re,
                                                               rb.im, of, og, pub_for(seq), desc);
            else if (cls == SC_PAIR)   // its input was just written: cached loads
                k_sc2_fold2_pairs_cached<<<nb, kFoldThreads, 0, st>>>(
                    cf, cg, out, ra.re, ra.im, rb.re, rb.im, of, og, pub_for(seq), desc);
            else
                k_sc2_fold2_pairs<<<nb, kFoldThreads, 0, st>>>(cf, cg, out, ra.re, ra.im, rb.re,
                                                               rb.im, of, og, pub_for(seq), desc);
            if (!t_end(cls, 40.0 * (double)len)) return -1;
            TP_HIP_OK(hipGetLastError());
            if (wait_pub(st, seq, quads ? 9 : 3, false) != 1 || !t_collect(cls)) return -1;
            cf = of;
            cg = og;
            if (!quads) {
                r = round_done(kp, S.pub[0], S.pub[1], S.pub[2]);
                k0 = kp + 1;
                m = out / 2;
                break;
            }
            for (int i = 0; i < 9; ++i) v[i] = S.pub[i];
            ra = round_done(kp, add2(v[0], v[3]), add2(v[1], v[4]), add2(v[2], v[5]));
            rb = round_done(kp + 1, quad_at(ra, v[0], v[1], v[2]), quad_at(ra, v[3], v[4], v[5]),
                            quad_at(ra, v[6], v[7], v[8]));
            len = out;
            cls = SC_PAIR;
            desc ^= 1;
        }
    } else {
        const unsigned seq = ++S.seq;
        const unsigned nb = blocks_for((n / 2 + 1) / 2);
        if (!t_begin()) return -1;
        k_sc2_first<<<nb, kBlock, 0, st>>>(cf, cg, n / 2, partials);
        if (!t_end(SC_FIRST, 32.0 * (double)n)) return -1;
        TP_HIP_OK(hipGetLastError());
        k_publish_wide<3><<<1, 1024, 0, st>>>(partials, nb, S.dev_vals, S.dev_seq, seq);
        TP_HIP_OK(hipGetLastError());
        if (!wait_seq(st, seq) || !t_collect(SC_FIRST)) return -1;
        take_seq_vals(3);
        r = round_done(0, S.pub[0], S.pub[1], S.pub[2]);
        k0 = 1;
        m = n / 2;
    }
    for (int k = k0;; ++k) {
        if (2 * m <= kTailIn) {
            // rounds k..nv-1 and the final fold in one workgroup, two rounds
            // per exchange; the host answers through the mapped answer slot
            const int nr = nv - k;
            const int ne = tail_exchanges(nr);
            unsigned seq0;
            if (tail_pre_k == k) {
                seq0 = tail_pre_seq0;   // queued already, taking r from the answer slot
                tail_pre_k = -1;
            } else {
                seq0 = S.seq + 1;
                S.seq += (unsigned)ne + 1;
                if (!t_begin()) return -1;
                if (!launch_tail(cf, cg, 2 * m, k, r, seq0, 0, 0)) return -1;
                if (!t_end(SC_TAIL, 32.0 * (double)(2 * m))) return -1;
            }
            volatile u64* ans = S.host_answer;
            double host_us = 0;
            for (int e = 0; e < ne; ++e) {
                const int kk = k + 2 * e;
                const bool pair = nv - kk >= 2;
                if (wait_pub(st, seq0 + e, pair ? 9 : 3, false) != 1) return -1;
                const auto h0 = std::chrono::steady_clock::now();
                const u64 tag = (u64)((seq0 + e) & 7) << 61;   // marks the words as this exchange's
                if (pair) {
                    // rounds kk and kk + 1 from the nine quad sums of T_kk
                    fp2 v[9];
                    for (int i = 0; i < 9; ++i) v[i] = S.pub[i];
                    const fp2 ra = round_done(kk, add2(v[0], v[3]), add2(v[1], v[4]), add2(v[2], v[5]));
                    const fp2 rb = round_done(kk + 1, quad_at(ra, v[0], v[1], v[2]),
                                              quad_at(ra, v[3], v[4], v[5]),
                                              quad_at(ra, v[6], v[7], v[8]));
                    ans[0] = ra.re | tag;
                    ans[1] = ra.im | tag;
                    ans[2] = rb.re | tag;
                    ans[3] = rb.im | tag;
                } else {
                    const fp2 rk = round_done(kk, S.pub[0], S.pub[1], S.pub[2]);
                    ans[0] = rk.re | tag;
                    ans[1] = rk.im | tag;
                }
                host_us += std::chrono::duration<double, std::micro>(
                               std::chrono::steady_clock::now() - h0).count();
            }
            if (wait_pub(st, seq0 + ne, 2, false) != 1 || !t_collect(SC_TAIL)) return -1;
            if (S.timed && ne > 0) {
                const volatile long long* tr = S.host_trace;
                double work = 0, wait = 0;
                for (int i = 0; i < ne; ++i) {
                    work += (double)(tr[3 * i + 1] - tr[3 * i]);
                    wait += (double)(tr[3 * i + 2] - tr[3 * i + 1]);
                    if (i < 40) S.tail_round_us[i] = (double)(tr[3 * i + 1] - tr[3 * i]) / S.clock_mhz;
                }
                S.tail_rounds = ne;
                S.tail_work_us = work / S.clock_mhz / ne;
                S.tail_wait_us = wait / S.clock_mhz / ne;
                S.tail_host_us = host_us / ne;
            }
            final4[0] = S.pub[0].re; final4[1] = S.pub[0].im;
            final4[2] = S.pub[1].re; final4[3] = S.pub[1].im;
            S.slots_clean = true;   // every streaming launch cleared its slots and ticket
            guard.armed = false;    // a queued tail has had every answer
            return 0;
        }
        if (use_mid && 2 * m <= (size_t)kMidGroups * kTailIn) {
            // rounds k .. k + nr - 1 on slices in kMidGroups workgroups, down
            // to kMidEnd entries per workgroup; that table goes to the tail
            MidLaunch ml;
            if (mid_pre_k == k) {
                ml = mid_pre;   // queued already, taking r from the answer slot
                mid_pre_k = -1;
            } else {
                if (!t_begin()) return -1;
                if (!launch_mid(cf, cg, 2 * m, r, 0, 0, ml)) return -1;
                if (!t_end(SC_MID, 32.0 * (double)(2 * m) + 64.0 * kMidGroups)) return -1;
            }
            const unsigned seq0 = ml.seq0;
            const int nr = ml.nr;
            V2* of = ml.of;
            V2* og = ml.og;
            const int first = wait_pub(st, seq0, mid_exchange_rounds(nr, 0) == 2 ? 9 : 3, true);
            if (first == 0) return -1;
            if (first == 2) {
                // not every workgroup was resident: these rounds go to the
                // streaming launches (the tables were only read); an answer
                // left for a queued launch goes (it was never taken)
                TP_HIP_OK(hipEventRecord(S.ev_round, st));
                if (!wait_event(S.ev_round)) return -1;
                {
                    volatile u64* ans = S.host_answer;
                    for (int i = 0; i < kMaxAns; ++i) ans[i] = kSlotEmpty;
                }
                guard.armed = false;
                S.launches[SC_MID] -= 1;
                S.bytes[SC_MID] -= 32.0 * (double)(2 * m) + 64.0 * kMidGroups;
                S.seq += 1;   // past the abort mark
                S.mid_fallbacks += 1;
                use_mid = false;
                --k;          // the loop's ++k returns to this round
                continue;
            }
            // an untimed prove queues the tail now, behind the mid launch, to
            // take the mid's last challenge from the answer slot (0.428-0.429 ms
            // a prove against 0.433-0.439 ms launching it after that step)
            const int tail_k = k + nr;
            const bool queue_tail = !S.timed && (size_t)kMidGroups * kMidEnd <= kTailIn && tail_k <= nv &&
                                    mid_exchange_rounds(nr, mid_exchanges(nr) - 1) == 1;
            if (queue_tail) {
                const int ne_t = tail_exchanges(nv - tail_k);
                tail_pre_seq0 = S.seq + 1;
                S.seq += (unsigned)ne_t + 1;
                if (!launch_tail(of, og, (size_t)kMidGroups * kMidEnd, tail_k, zero2(), tail_pre_seq0,
                                 seq0 + (unsigned)mid_exchanges(nr) - 1, 1))
                    return -1;
                if (!t_end(SC_TAIL, 32.0 * (double)kMidGroups * kMidEnd)) return -1;
                tail_pre_k = tail_k;
                guard.armed = true;   // it waits for the mid's last challenge
            }
            if (!mid_host(k, seq0, nr, true, r, queue_tail)) return -1;
            if (!queue_tail) guard.armed = false;   // a queued mid launch has had its first challenge
            // Every round is published; the launch resets its words as it ends.
            // The next launches queue behind it on the stream, so the host
            // waits for its end only to read its timing (timing on).
            S.mid_clean = true;
            if (S.timed) {
                TP_HIP_OK(hipEventRecord(S.ev_round, st));
                if (!wait_event(S.ev_round) || !t_collect(SC_MID)) return -1;
            }
            const int ne = mid_exchanges(nr);
            if (S.timed && ne > 1 && ne <= 40) {
                // per exchange: out = partials written, in = challenges held
                const volatile long long* tr = S.host_mtrace;
                auto out_t = [&](int i, int b) { return (double)tr[((size_t)i * kMidGroups + b) * 2]; };
                auto in_t = [&](int i, int b) { return (double)tr[((size_t)i * kMidGroups + b) * 2 + 1]; };
                double host = 0, relay = 0, work = 0, spread = 0;
                for (int i = 0; i + 1 < ne; ++i) {
                    double last_out = out_t(i, 0), first_out = out_t(i, 0), last_in = in_t(i, 0);
                    double w = 0;
                    for (int b = 0; b < kMidGroups; ++b) {
                        last_out = std::max(last_out, out_t(i, b));
                        first_out = std::min(first_out, out_t(i, b));
                        last_in = std::max(last_in, in_t(i, b));
                        w += out_t(i + 1, b) - in_t(i, b);
                    }
                    host += in_t(i, 0) - last_out;      // last partials -> challenge at workgroup 0
                    relay += last_in - in_t(i, 0);      // workgroup 0 -> the last to hold it
                    work += w / kMidGroups;             // challenge -> next partials
                    spread += last_out - first_out;     // partials: first to last
                }
                const double d = S.clock_mhz * (ne - 1);
                S.mid_rounds = ne;
                S.mid_host_us = host / d;
                S.mid_relay_us = relay / d;
                S.mid_work_us = work / d;
                S.mid_spread_us = spread / d;
            }
            cf = of;
            cg = og;
            k += nr - 1;          // the loop's ++k moves to the first round after these
            m = (size_t)kMidGroups * kMidEnd / 2;   // its input has kMidGroups x kMidEnd entries
            continue;
        }
        const unsigned seq = ++S.seq;
        const size_t quads = m / 2;   // input T_{k-1} has length 2m
        // the work buffer not holding the input (A holds up to n/2 entries
        // per table, B n/4: T_1 goes to A, and any later table fits either)
        V2* of = (cf >= A && cf < A + 2 * capA) ? B : A;
        V2* og = of + (of == A ? capA : capB);
        const unsigned nb =
            (unsigned)std::min<size_t>(kFoldGroups, (quads + kFoldThreads - 1) / kFoldThreads);
        if (!t_begin()) return -1;
        // round 2 ran top down, so round 3 runs bottom up, round 4 top down, ...
        k_sc2_fold_round<<<nb, kFoldThreads, 0, st>>>(cf, cg, quads, r.re, r.im, of, og,
                                                      pub_for(seq), (k & 1) == 0);
        if (!t_end(SC_FOLD, 48.0 * (double)(2 * m))) return -1;
        TP_HIP_OK(hipGetLastError());
        if (wait_pub(st, seq, 3, false) != 1 || !t_collect(SC_FOLD)) return -1;
        cf = of;
        cg = og;
        r = round_done(k, S.pub[0], S.pub[1], S.pub[2]);
        m /= 2;
    }
}

// Enables per-launch HIP-event timing of the fused fold+round kernel.
void tp_sc2_set_timing(int on) { g_sc2.timed = on != 0; }

// 1 (default): proves of 2^14 to 2^28 entries take rounds 0 to 3 from the
// block products and fold the input into T_4 in one pass; 0: rounds 0 and 1
// from the quad products and the passes folding by two challenges.  The
// proof is the same either way.
void tp_sc2_set_block_pass(int on) { g_sc2.blocks = on != 0; }

// 1: an untimed prove queues its launches down to the mid launch at once,
// each taking its challenges from the answer slot; 0 (default): each launch
// is made after the transcript step before it (as timed proves always are).
// The proof is the same either way.
void tp_sc2_set_ahead(int on) { g_sc2.ahead = on != 0; }

// Statistics of the last tp_sc2_prove: summed event time and algorithmic
// bytes of the streaming fused fold+round launches (k_sc2_fold_round), and
// their count.
void tp_sc2_stats(double* fold_ms, double* fold_bytes, int* fold_launches)
{
    *fold_ms = g_sc2.ms[SC_FOLD];
    *fold_bytes = g_sc2.bytes[SC_FOLD];
    *fold_launches = g_sc2.launches[SC_FOLD];
}

// Summed event time and count of the one-workgroup tail rounds of the last
// tp_sc2_prove (timing on).
void tp_sc2_tail_stats(double* tail_ms, int* tail_launches)
{
    *tail_ms = g_sc2.ms[SC_TAIL];
    *tail_launches = g_sc2.launches[SC_TAIL];
}

// Per-round latency of the one-workgroup tail in the last timed
// tp_sc2_prove: out[0] rounds, out[1] mean device time from a round's start
// to its sums being ready (us), out[2] mean time the workgroup waited for the
// host's challenge (us), out[3] mean host time from seeing the sums to
// writing the challenge (transcript, us), out[4 + i] the device time of
// round i of the tail (us, while 4 + i < n).  n >= 4.
void tp_sc2_tail_profile(double* out, int n)
{
    if (n < 4) return;
    out[0] = g_sc2.tail_rounds;
    out[1] = g_sc2.tail_work_us;
    out[2] = g_sc2.tail_wait_us;
    out[3] = g_sc2.tail_host_us;
    for (int i = 0; i < g_sc2.tail_rounds && i < 40 && 4 + i < n; ++i) out[4 + i] = g_sc2.tail_round_us[i];
}

// Per-round latency of the sliced mid rounds in the last timed tp_sc2_prove
// (means over the rounds that wait for a challenge): out[0] rounds, out[1] us
// from the last workgroup's partials to workgroup 0 holding the challenge
// (host sum, transcript, answer), out[2] us until the last workgroup holds it
// (the relay), out[3] us from a workgroup's challenge to its next partials,
// out[4] us between the first and the last workgroup's partials.  n >= 5.
void tp_sc2_mid_profile(double* out, int n)
{
    if (n < 5) return;
    out[0] = g_sc2.mid_rounds;
    out[1] = g_sc2.mid_host_us;
    out[2] = g_sc2.mid_relay_us;
    out[3] = g_sc2.mid_work_us;
    out[4] = g_sc2.mid_spread_us;
}

// Bound of the mid launch's arrival gate in wall-clock ticks (default 20 ms;
// a negative value restores it).  0 makes the launch give way unless every
// workgroup has arrived when workgroup 0 looks, for tests of the fallback.
void tp_sc2_set_gate_ticks(long long ticks)
{
    if (!sc2_setup(2)) return;
    g_sc2.gate_ticks = ticks >= 0 ? ticks : (long long)(g_sc2.clock_mhz * 1000.0) * 20;
}

// Mid launches of the last tp_sc2_prove that gave way at the arrival gate.
int tp_sc2_mid_fallbacks(void) { return g_sc2.mid_fallbacks; }

// Bound of every device-side wait of a prove (the tail's and mid launch's
// answers, the mid relay, the streaming gathers) in wall-clock ticks; a
// negative value restores the default 10 s.  A small bound makes a prove
// fail with every kernel leaving on its own, for tests of the error path.
void tp_sc2_set_spin_ticks(long long ticks)
{
    if (!sc2_setup(2)) return;
    g_sc2.spin_ticks = ticks >= 0 ? ticks : (long long)(g_sc2.clock_mhz * 1000.0) * 1000LL * 10;
}

// The schedule's sizes: the workgroups of the sliced mid rounds, the
// largest input (entries per table) the one-workgroup tail takes, and the
// input the mid rounds hand it.  The streaming launches run while the input
// exceeds mid_groups * tail_in, the mid rounds while it exceeds tail_in.
void tp_sc2_schedule(int* mid_groups, int* tail_in, int* mid_end)
{
    *mid_groups = kMidGroups;
    *tail_in = (int)kTailIn;
    *mid_end = (int)(kMidGroups * kMidEnd);
}

// Per-class statistics of the last tp_sc2_prove (timing on): for each class
// (0 the first pass over the input, 1 the round-2 pass folding the input by
// two challenges, 2 the streaming fold+round launches, 3 the one-workgroup
// tail rounds, 4 the sliced mid rounds, 5 the later passes folding by two
// challenges) the summed event ms, the bytes the launches read and write,
// and the launch count: out[3 c + 0..2].  n >= 18.
void tp_sc2_class_stats(double* out, int n)
{
    for (int c = 0; c < SC_NCLASS && 3 * c + 2 < n; ++c) {
        out[3 * c] = g_sc2.ms[c];
        out[3 * c + 1] = g_sc2.bytes[c];
        out[3 * c + 2] = g_sc2.launches[c];
    }
}

}  // extern "C"
