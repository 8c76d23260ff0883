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
// Bytes per fused launch: 32 m read + 16 m written (both tables), the first
// launch reads 32 n and writes nothing.  Tables are AoS {re, im} so every
// element is one 128-bit load; a thread owns one output pair (64 B per table).
//
// Round trip per round: fused launch -> one-workgroup finalize launch that
// writes (E0, E1, EA) and a sequence number straight into mapped pinned host
// memory -> the host polls the sequence number, runs the transcript, and
// launches the next round with the challenge as a kernel argument.
#include "fp61.hpp"
#include "host_sha.hpp"
#include "reduce.hpp"
#include "runtime.hpp"
#include "transcript.hpp"

#include <atomic>
#include <chrono>
#include <stdint.h>
#include <thread>

using namespace tp;

namespace {

struct alignas(16) V2 { u64 re, im; };

__device__ __forceinline__ fp2 ldv(const V2* p) { const V2 v = *p; return mk2(v.re, v.im); }
__device__ __forceinline__ void stv(V2* p, fp2 x) { V2 v; v.re = x.re; v.im = x.im; *p = v; }

__device__ __forceinline__ void accum_pair(fp2 (&e)[3], fp2 f0, fp2 f1, fp2 g0, fp2 g1)
{
    e[0] = add2(e[0], mul2(f0, g0));
    e[1] = add2(e[1], mul2(f1, g1));
    e[2] = add2(e[2], mul2(sub2(f1, f0), sub2(g1, g0)));
}

// Round 0: no fold, sum over the pairs of the input tables; a thread owns two
// adjacent pairs (64 contiguous bytes per table), as the fused rounds do.
__global__ void __launch_bounds__(256) k_sc2_first(const V2* __restrict__ f,
                                                   const V2* __restrict__ g, size_t n_pairs,
                                                   fp2* partials)
{
    fp2 e[3] = {zero2(), zero2(), zero2()};
    const size_t n_quads = n_pairs / 2;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n_quads;
         q += (size_t)gridDim.x * blockDim.x) {
        const V2* fp = f + 4 * q;
        const V2* gp = g + 4 * q;
        const fp2 fa = ldv(fp), fb = ldv(fp + 1), fc = ldv(fp + 2), fd = ldv(fp + 3);
        const fp2 ga = ldv(gp), gb = ldv(gp + 1), gc = ldv(gp + 2), gd = ldv(gp + 3);
        accum_pair(e, fa, fb, ga, gb);
        accum_pair(e, fc, fd, gc, gd);
    }
    if (n_pairs & 1) {   // one pair (n = 2) left over
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const size_t p = n_pairs - 1;
            accum_pair(e, ldv(f + 2 * p), ldv(f + 2 * p + 1), ldv(g + 2 * p), ldv(g + 2 * p + 1));
        }
    }
    block_partial<3>(e, partials);
}

// Round k >= 1: fold T_{k-1} (length 4 * n_quads) with r into T_k and sum
// the round-k pairs.  Quad q = inputs 4q..4q+3 -> outputs 2q, 2q+1.
__global__ void __launch_bounds__(256) k_sc2_fold_round(const V2* __restrict__ fin,
                                                        const V2* __restrict__ gin,
                                                        size_t n_quads, u64 rr, u64 ri,
                                                        V2* __restrict__ fout,
                                                        V2* __restrict__ gout, fp2* partials)
{
    const fp2 r = mk2(rr, ri);
    fp2 e[3] = {zero2(), zero2(), zero2()};
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n_quads;
         q += (size_t)gridDim.x * blockDim.x) {
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
        accum_pair(e, f0, f1, g0, g1);
    }
    block_partial<3>(e, partials);
}

// Rounds whose input table is at most one workgroup's quads (kTailQuads): fold,
// round sum and publish in one launch, without the partials round trip.
constexpr int kTailThreads = 256;
constexpr size_t kTailQuads = 256;
__global__ void __launch_bounds__(kTailThreads) k_sc2_tail_round(
    const V2* __restrict__ fin, const V2* __restrict__ gin, size_t n_quads, u64 rr, u64 ri,
    V2* __restrict__ fout, V2* __restrict__ gout, fp2* host_vals, volatile unsigned* host_seq,
    unsigned seq)
{
    const fp2 r = mk2(rr, ri);
    fp2 e[3] = {zero2(), zero2(), zero2()};
    for (size_t q = threadIdx.x; q < n_quads; q += blockDim.x) {
        const V2* fp = fin + 4 * q;
        const V2* gp = gin + 4 * q;
        const fp2 f0 = fold2(ldv(fp), ldv(fp + 1), r), f1 = fold2(ldv(fp + 2), ldv(fp + 3), r);
        const fp2 g0 = fold2(ldv(gp), ldv(gp + 1), r), g1 = fold2(ldv(gp + 2), ldv(gp + 3), r);
        stv(fout + 2 * q, f0);
        stv(fout + 2 * q + 1, f1);
        stv(gout + 2 * q, g0);
        stv(gout + 2 * q + 1, g1);
        accum_pair(e, f0, f1, g0, g1);
    }
    block_sum<3>(e);
    if (threadIdx.x == 0) {
        host_vals[0] = e[0];
        host_vals[1] = e[1];
        host_vals[2] = e[2];
        __threadfence_system();
        *host_seq = seq;
    }
}

// One workgroup: sum the partials and publish (E0, E1, EA) plus the sequence
// number `seq` in host-mapped memory (values first, then a system-scope fence,
// then the sequence number).
__global__ void __launch_bounds__(256) k_sc2_publish(const fp2* partials, unsigned n_blocks,
                                                     fp2* host_vals, volatile unsigned* host_seq,
                                                     unsigned seq)
{
    fp2 acc[3];
    sum_partials<3>(partials, n_blocks, acc);
    if (threadIdx.x == 0) {
        host_vals[0] = acc[0];
        host_vals[1] = acc[1];
        host_vals[2] = acc[2];
        __threadfence_system();
        *host_seq = seq;
    }
}

// Final fold of the last pair: f~(r), g~(r), published like the round sums.
__global__ void k_sc2_final(const V2* __restrict__ f, const V2* __restrict__ g, u64 rr, u64 ri,
                            fp2* host_vals, volatile unsigned* host_seq, unsigned seq)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const fp2 r = mk2(rr, ri);
        host_vals[0] = fold2(ldv(f), ldv(f + 1), r);
        host_vals[1] = fold2(ldv(g), ldv(g + 1), r);
        __threadfence_system();
        *host_seq = seq;
    }
}

struct Sc2State {
    DevBuf work;                 // T_k ping-pong: A (n/2) and B (n/4), for f and g
    DevBuf partials;             // 3 fp2 per workgroup
    fp2* host_vals = nullptr;    // mapped pinned: 3 values
    unsigned* host_seq = nullptr;
    fp2* dev_vals = nullptr;     // device aliases of the mapped memory
    unsigned* dev_seq = nullptr;
    unsigned seq = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool timed = false;          // per-launch event timing of the fused kernel
    double fold_ms = 0, fold_bytes = 0;
    int fold_launches = 0;
    double tail_ms = 0;          // the one-workgroup tail rounds
    int tail_launches = 0;
};
Sc2State g_sc2;

bool sc2_setup(size_t n)
{
    Sc2State& S = g_sc2;
    if (!S.host_vals) {
        void* hv = nullptr;
        TP_HIP_OK(hipHostMalloc(&hv, 256, hipHostMallocMapped | hipHostMallocCoherent));
        S.host_vals = static_cast<fp2*>(hv);
        S.host_seq = reinterpret_cast<unsigned*>(static_cast<char*>(hv) + 128);
        *S.host_seq = 0;
        void* dv = nullptr;
        TP_HIP_OK(hipHostGetDevicePointer(&dv, hv, 0));
        S.dev_vals = static_cast<fp2*>(dv);
        S.dev_seq = reinterpret_cast<unsigned*>(static_cast<char*>(dv) + 128);
    }
    if (!S.ev[0]) {
        TP_HIP_OK(hipEventCreate(&S.ev[0]));
        TP_HIP_OK(hipEventCreate(&S.ev[1]));
    }
    if (!S.work.ensure(2 * (n / 2 + n / 4 + 2) * sizeof(V2))) return false;
    if (!S.partials.ensure(2048 * 3 * sizeof(fp2))) return false;
    return true;
}

unsigned blocks_for(size_t items)
{
    // 2048 workgroups (8 per CU on 256 CUs) saturate HBM; fewer for small rounds.
    size_t b = (items + kBlock - 1) / kBlock;
    if (b > 2048) b = 2048;
    return (unsigned)(b ? b : 1);
}

// Waits until the device has published sequence number `seq`; checks the
// stream for errors while waiting and gives up after a generous timeout.
bool wait_seq(hipStream_t st, unsigned seq)
{
    volatile unsigned* hs = g_sc2.host_seq;
    const auto t0 = std::chrono::steady_clock::now();
    long spins = 0;
    while (*hs != seq) {
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
    Sc2State& S = g_sc2;
    S.fold_ms = 0;
    S.fold_bytes = 0;
    S.fold_launches = 0;
    S.tail_ms = 0;
    S.tail_launches = 0;

    fp2* partials = S.partials.as<fp2>();
    V2* A = S.work.as<V2>();             // f: n/2, then g: n/2
    V2* B = A + 2 * (n / 2 + 1);         // f: n/4, then g: n/4
    const size_t capA = n / 2 + 1, capB = n / 4 + 1;

    FsTranscriptT<HostSha256> fs;
    fs.init("truepix");
    fs.item_str("c2", 2);   // put(["c2", seed])
    fs.item_int(seed);

    const V2* cf = reinterpret_cast<const V2*>(d_f);
    const V2* cg = reinterpret_cast<const V2*>(d_g);
    size_t m = n;   // length of the current table T_k
    u64 rr = 0, ri = 0;
    for (int k = 0; k < nv; ++k) {
        unsigned nb;
        bool published = false;
        const unsigned seq = ++S.seq;
        if (k == 0) {
            nb = blocks_for((m / 2 + 1) / 2);
            k_sc2_first<<<nb, kBlock, 0, st>>>(cf, cg, m / 2, partials);
        } else if (m / 2 <= kTailQuads) {
            const size_t quads = m / 2;
            V2* of = (k & 1) ? A : B;
            V2* og = of + ((k & 1) ? capA : capB);
            nb = 1;
            if (S.timed) TP_HIP_OK(hipEventRecord(S.ev[0], st));
            k_sc2_tail_round<<<1, kTailThreads, 0, st>>>(cf, cg, quads, rr, ri, of, og, S.dev_vals,
                                                         S.dev_seq, seq);
            if (S.timed) TP_HIP_OK(hipEventRecord(S.ev[1], st));
            cf = of;
            cg = og;
            published = true;
        } else {
            const size_t quads = m / 2;   // input T_{k-1} has length 2m
            V2* of = (k & 1) ? A : B;
            V2* og = of + ((k & 1) ? capA : capB);
            nb = blocks_for(quads);
            if (S.timed) TP_HIP_OK(hipEventRecord(S.ev[0], st));
            k_sc2_fold_round<<<nb, kBlock, 0, st>>>(cf, cg, quads, rr, ri, of, og, partials);
            if (S.timed) TP_HIP_OK(hipEventRecord(S.ev[1], st));
            cf = of;
            cg = og;
            S.fold_bytes += 48.0 * (double)(2 * m);
            S.fold_launches += 1;
        }
        TP_HIP_OK(hipGetLastError());
        if (!published) {
            k_sc2_publish<<<1, kBlock, 0, st>>>(partials, nb, S.dev_vals, S.dev_seq, seq);
            TP_HIP_OK(hipGetLastError());
        }
        if (!wait_seq(st, seq)) return -1;
        if (S.timed && k > 0) {
            TP_HIP_OK(hipEventSynchronize(S.ev[1]));
            float ms = 0;
            TP_HIP_OK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
            if (published) {
                S.tail_ms += ms;
                S.tail_launches += 1;
            } else {
                S.fold_ms += ms;
            }
        }
        const fp2 E0 = S.host_vals[0], E1 = S.host_vals[1], EA = S.host_vals[2];
        const fp2 a = EA, c = E0, b = sub2(sub2(E1, EA), E0);
        polys[6 * k + 0] = a.re; polys[6 * k + 1] = a.im;
        polys[6 * k + 2] = b.re; polys[6 * k + 3] = b.im;
        polys[6 * k + 4] = c.re; polys[6 * k + 5] = c.im;
        if (fixed) {
            rr = fixed[2 * k];
            ri = fixed[2 * k + 1];
        } else {
            fs.item_fp2(a);   // put([a, b, c])
            fs.item_fp2(b);
            fs.item_fp2(c);
            char lab[24];
            const uint32_t ln = label_num(lab, "c2-", k);
            const fp2 r = fs.rand_ext(lab, ln);
            rr = r.re;
            ri = r.im;
        }
        chals[2 * k] = rr;
        chals[2 * k + 1] = ri;
        m /= 2;
    }
    // T_{nv-1} has two entries: fold them with r_{nv-1}.
    const unsigned seq = ++S.seq;
    k_sc2_final<<<1, 64, 0, st>>>(cf, cg, rr, ri, S.dev_vals, S.dev_seq, seq);
    TP_HIP_OK(hipGetLastError());
    if (!wait_seq(st, seq)) return -1;
    final4[0] = S.host_vals[0].re; final4[1] = S.host_vals[0].im;
    final4[2] = S.host_vals[1].re; final4[3] = S.host_vals[1].im;
    return 0;
}

// Enables per-launch HIP-event timing of the fused fold+round kernel.
void tp_sc2_set_timing(int on) { g_sc2.timed = on != 0; }

// Statistics of the last tp_sc2_prove: summed event time and algorithmic
// bytes of the streaming fused fold+round launches (k_sc2_fold_round), and
// their count.
void tp_sc2_stats(double* fold_ms, double* fold_bytes, int* fold_launches)
{
    *fold_ms = g_sc2.fold_ms;
    *fold_bytes = g_sc2.fold_bytes;
    *fold_launches = g_sc2.fold_launches;
}

// Summed event time and count of the one-workgroup tail rounds of the last
// tp_sc2_prove (timing on).
void tp_sc2_tail_stats(double* tail_ms, int* tail_launches)
{
    *tail_ms = g_sc2.tail_ms;
    *tail_launches = g_sc2.tail_launches;
}

}  // extern "C"
