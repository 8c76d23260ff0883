// batch.hip - lockstep batch driver of the frame prover.
//
// Every frame of a batch advances through the protocol of
// CircuitProverNIZK.run (TP/libTruePix/circuitnizkvec_vole.py:265-501, one
// layer) together: each protocol step is ONE launch covering every frame
// that has that step, instead of one launch per frame per step.  Frames
// differ in circuit (gates, input / output bits) and in copies, so a launch
// carries
//   - a per-frame descriptor table (FDesc: the circuit arrays and the frame's
//     tables inside one device arena), and
//   - a per-launch item table: one (frame, first workgroup, two
//     launch-specific ints) item per frame in the launch, sorted by first
//     workgroup, with a sentinel; a workgroup finds its frame by binary
//     search over the first-workgroup column and runs the same kernel body
//     as the single-frame driver (layer_bodies.hpp) at its position inside
//     that frame's share of the launch.
// All item tables depend only on the frames' shapes, so the whole schedule is
// built on the host once per batch and uploaded with the descriptors in one
// copy.
//
// Host side per step: the driver thread waits for every frame's published
// sums (mapped host memory, one sequence word per frame), then the frames'
// transcript steps (FrameProto) run in parallel on the task pool, each
// writing its challenge into a mapped challenge table ([slot][frame], so one
// step's challenges are contiguous); a one-workgroup pull copies them to the
// device before the next step's launch.  Per batch that is one host round
// trip per protocol step (about 2 nInBits + nCopyBits + 3 for the largest
// frame) instead of one per frame per step.
//
// Outputs come down in groups of frames with an event per group, and each
// frame's transcript head (the SHA-256 of its outputs, the bulk of the host
// work) starts as soon as its group has landed, overlapping the rest of the
// download.  They come down as one byte per output (the edit circuits output
// pixel values) with a per-frame flag set on the device when any output is
// above 255, in which case that frame's u64 outputs are fetched instead: an
// eighth of the PCIe and host-memory traffic, and the transcript widens the
// bytes back to u64 in L1-sized chunks as it absorbs them (same bytes).  The proof prefix (q, io) is pickled by background tasks while
// the rounds run.
#include "frame.hpp"
#include "layer_bodies.hpp"
#include "reduce.hpp"
#include "../../include/truepix_hip.h"

#include <stdint.h>
#include <string.h>

#include <algorithm>

namespace tp {
namespace frame {
namespace {

// Challenge slots per frame in the [slot][frame] challenge table.
constexpr int SZ1 = 0;    // z1, nOutBits <= kMaxNin
constexpr int SZ2 = 11;   // z2, nCopyBits <= 16
constexpr int SER = 27;   // early-round challenges (w3)
constexpr int SLC = 43;   // late-round challenges (w1 then w2)
constexpr int SPT = 65;   // the final input point, n_z <= kMaxZ
constexpr int kSlots = 112;
static_assert(SZ2 - SZ1 >= kMaxNin && SER - SZ2 >= 16 && SLC - SER >= 16 &&
                  SPT - SLC >= 2 * kMaxNin && kSlots - SPT >= kMaxZ,
              "challenge slots overlap");
constexpr int kVals = 32;        // published Fp2 values per frame
constexpr int kSmall = 64;       // Fp2 of each frame's small device buffer
constexpr int kElemCap = 1024;   // workgroups per frame of an elementwise launch
// Autotuned schedule candidates (the sums are exact, so every choice gives the
// same bits): the early claim's (pair, gate chunk) items per frame, and the
// workgroups per frame of the fold launches.
constexpr long long kClaimCand[3] = {1 << 14, 1 << 16, 1 << 18};
constexpr int kFoldCand[3] = {256, 1024, 4096};
constexpr long long kTuneMinItems = 1 << 21;   // batches with fewer claim terms are not trials
constexpr int kMleCap = 256;     // workgroups per frame of the input MLE

struct FDesc {
    const int *op, *mux, *gt, *in0, *in1, *slot;
    const u64* cval;
    int n_gates, nin, nout, n_priv;
    int n_copies, cp, ncb, in_bytes;
    const void* raw;
    u64 *rows_in, *rows_out;
    uint8_t* out8;    // the outputs as bytes (valid when no output is above 255)
    unsigned* big;    // mapped flag: set when an output is above 255
    u64 *vr[2], *vi[2], *br[2], *bi[2];
    u64 *z1cr, *z1ci, *z1lr, *z1li, *othr, *othi;
    // the early claim's terms (claim_terms.hpp) and their coefficients
    const int *tt, *tin0, *tin1, *toff, *tsrc, *tneg;
    int n_terms;
    u64 *tzr, *tzi;
    fp2* partials;
    fp2* small;   // beta* (as two limbs), then line values at 4, wiring sums at 20
};

// f = frame, b0 = first workgroup; a, b per launch (see the kernels).
struct Item {
    int f, b0, a, b;
};

__device__ __forceinline__ int find_item(const Item* it, int n, int blk)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (it[mid].b0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ kernels

__global__ void k_b_pull(const fp2* __restrict__ src, fp2* __restrict__ dst, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// Input rows of every frame (body::input_rows).
__global__ void __launch_bounds__(kBlock) k_b_rows(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int n)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const bool b8 = d.in_bytes == 1;
    body::input_rows(b8 ? (const uint8_t*)d.raw : nullptr, b8 ? nullptr : (const u64*)d.raw,
                     d.n_priv, d.n_copies, d.cp, d.nin, d.slot, d.cval, d.rows_in,
                     blockIdx.x - m.b0, it[i + 1].b0 - m.b0);
}

// A3 for every frame (body::eval_layer).
__global__ void __launch_bounds__(kBlock) k_b_eval(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int n)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    body::eval_layer(d.rows_in, 1 << d.nin, d.rows_out, d.nout, d.n_copies, d.cp, d.op, d.in0,
                     d.in1, d.mux, d.n_gates, blockIdx.x - m.b0, it[i + 1].b0 - m.b0);
}

// The outputs of every frame as bytes, flagging a frame with any output above
// 255 (its transcript then takes the u64 outputs instead).
__global__ void __launch_bounds__(kBlock) k_b_pack(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int n)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const size_t total = (size_t)d.cp << d.nout;
    const size_t nb = (size_t)(it[i + 1].b0 - m.b0);
    bool big = false;
    for (size_t k = (size_t)(blockIdx.x - m.b0) * blockDim.x + threadIdx.x; k < total;
         k += nb * blockDim.x) {
        const u64 v = d.rows_out[k];
        d.out8[k] = (uint8_t)v;
        big |= v > 255;
    }
    if (big) {
        *reinterpret_cast<volatile unsigned*>(d.big) = 1;
        __threadfence_system();
    }
}

// A4 for every frame; a = tiles along copies (body::wire_table).
__global__ void __launch_bounds__(kBlock) k_b_wire(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int n)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const int local = blockIdx.x - m.b0;
    const int NW = 1 << d.nin;
    body::wire_table(d.rows_in, NW, d.n_copies, d.cp, NW, d.vr[0], d.vi[0], local % m.a,
                     local / m.a);
}

// eq tables: the first a workgroups of a frame build beta over its copies
// from z2, the rest the gate coefficients from z1 (body::eq_table).
__global__ void __launch_bounds__(kBlock) k_b_eq(const FDesc* __restrict__ D,
                                                 const Item* __restrict__ it, int n,
                                                 const fp2* __restrict__ ch, int nf)
{
    __shared__ fp2 zs[16];
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const int local = blockIdx.x - m.b0;
    const bool beta = local < m.a;
    const int nz = beta ? d.ncb : d.nout;
    const int s0 = beta ? SZ2 : SZ1;
    for (int j = threadIdx.x; j < nz; j += blockDim.x) zs[j] = ch[(size_t)(s0 + j) * nf + m.f];
    __syncthreads();
    if (beta) body::eq_table(zs, d.ncb, d.br[0], d.bi[0], d.cp, local);
    else body::eq_table(zs, d.nout, d.z1cr, d.z1ci, d.n_gates, local - m.a);
}

// The claim terms' coefficients of every frame from its z1 values
// (body::term_coeffs).
__global__ void __launch_bounds__(kBlock) k_b_terms(const FDesc* __restrict__ D,
                                                    const Item* __restrict__ it, int n)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    body::term_coeffs(d.z1cr, d.z1ci, d.toff, d.tsrc, d.tneg, d.n_terms, d.tzr, d.tzi,
                      blockIdx.x - m.b0);
}

// A6, early round rd for every frame that has it; a = term chunks per pair,
// b = terms per chunk (body::early_claim).
__global__ void __launch_bounds__(kBlock) k_b_claim(const FDesc* __restrict__ D,
                                                    const Item* __restrict__ it, int n, int rd)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const int cur = rd & 1;
    const int local = blockIdx.x - m.b0;
    body::early_claim(d.n_terms, m.b, m.a, d.tt, d.tin0, d.tin1, d.tzr, d.tzi, d.cp >> rd,
                      d.vr[cur], d.vi[cur], d.br[cur], d.bi[cur], d.partials, local, local);
}

// One workgroup per item: the frame's a partials of N sums to its published
// values, then its sequence word.
template <int N>
__global__ void __launch_bounds__(kBlock) k_b_pub(const FDesc* __restrict__ D,
                                                  const Item* __restrict__ it, fp2* hv,
                                                  unsigned* hseq, unsigned seq)
{
    const Item m = it[blockIdx.x];
    fp2 acc[N];
    sum_partials<N>(D[m.f].partials, (unsigned)m.a, acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) hv[(size_t)m.f * kVals + k] = acc[k];
        __threadfence_system();
        reinterpret_cast<volatile unsigned*>(hseq)[m.f] = seq;
    }
}

// A7, fold of early round rd with the frame's challenge (body::fold_rows).
__global__ void __launch_bounds__(kBlock) k_b_fold(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int n, int rd,
                                                   const fp2* __restrict__ ch, int nf)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const fp2 r = ch[(size_t)(SER + rd) * nf + m.f];
    const int cur = rd & 1;
    body::fold_rows(d.vr[cur], d.vi[cur], 1 << d.nin, d.ncb - rd - 1, r, d.vr[cur ^ 1],
                    d.vi[cur ^ 1], d.br[cur], d.bi[cur], d.br[cur ^ 1], d.bi[cur ^ 1],
                    blockIdx.x - m.b0, it[i + 1].b0 - m.b0);
}

// Start of the late rounds for every frame (body::late_init).
__global__ void __launch_bounds__(kBlock) k_b_linit(const FDesc* __restrict__ D,
                                                    const Item* __restrict__ it, int n)
{
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const int cur = d.ncb & 1;
    body::late_init(d.vr[cur], d.vi[cur], d.in1, d.n_gates, d.z1cr, d.z1ci, d.z1lr, d.z1li, d.othr,
                    d.othi, d.br[cur], d.bi[cur], reinterpret_cast<u64*>(d.small),
                    blockIdx.x - m.b0);
}

// A9, late round k: one workgroup per frame that has it; at k = nin the w2
// pass starts in the same workgroup (body::late_switch, body::late_round).
__global__ void __launch_bounds__(kOneWg) k_b_late(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int k,
                                                   const fp2* __restrict__ ch, int nf, fp2* hv,
                                                   unsigned* hseq, unsigned seq)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 cs[2 * kMaxNin];
    const Item m = it[blockIdx.x];
    const FDesc& d = D[m.f];
    const int nin = d.nin;
    for (int j = threadIdx.x; j < 2 * nin; j += blockDim.x)
        cs[j] = j < k ? ch[(size_t)(SLC + j) * nf + m.f] : zero2();
    __syncthreads();
    const int cur = d.ncb & 1;
    if (k == nin) {
        body::late_switch(d.vr[cur], d.vi[cur], nin, cs, d.n_gates, d.othr, d.othi, lds);
        __syncthreads();
    }
    fp2 out[3];
    body::late_round(d.gt, d.in0, d.in1, d.z1lr, d.z1li, d.othr, d.othi, d.vr[cur], d.vi[cur], cs,
                     nin, d.n_gates, k, k > 0 ? 1 : 0, reinterpret_cast<const u64*>(d.small), lds,
                     out);
    if (threadIdx.x == 0) {
        for (int t = 0; t < 3; ++t) hv[(size_t)m.f * kVals + t] = out[t];
        __threadfence_system();
        reinterpret_cast<volatile unsigned*>(hseq)[m.f] = seq;
    }
}

// A10 and A11: workgroups t <= nin of a frame evaluate V on the line at t
// (a = nin + 1 of them), the last one the wiring predicate sums.
__global__ void __launch_bounds__(kOneWg) k_b_line(const FDesc* __restrict__ D,
                                                   const Item* __restrict__ it, int n,
                                                   const fp2* __restrict__ ch, int nf)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 cs[2 * kMaxNin];
    __shared__ fp2 zs[kMaxNin];
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const int nin = d.nin;
    for (int j = threadIdx.x; j < 2 * nin; j += blockDim.x) cs[j] = ch[(size_t)(SLC + j) * nf + m.f];
    for (int j = threadIdx.x; j < d.nout; j += blockDim.x) zs[j] = ch[(size_t)(SZ1 + j) * nf + m.f];
    __syncthreads();
    const int t = blockIdx.x - m.b0;
    const int cur = d.ncb & 1;
    fp2* out = d.small + 4;
    if (t < m.a) {
        const fp2 v = body::line_value(d.vr[cur], d.vi[cur], nin, cs, t, lds);
        if (threadIdx.x == 0) out[t] = v;
    } else {
        body::wiring_mlext(d.gt, d.in0, d.in1, d.n_gates, d.nout, nin, zs, nullptr, cs, cs + nin,
                           zero2(), zero2(), lds, out + 16);
    }
}

// One workgroup per frame: its line values (16 slots) and wiring sums (13) to
// the published values, then its sequence word.
__global__ void k_b_pubcopy(const FDesc* __restrict__ D, const Item* __restrict__ it, fp2* hv,
                            unsigned* hseq, unsigned seq)
{
    const Item m = it[blockIdx.x];
    const fp2* src = D[m.f].small + 4;
    for (int k = threadIdx.x; k < 29; k += blockDim.x) hv[(size_t)m.f * kVals + k] = src[k];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) reinterpret_cast<volatile unsigned*>(hseq)[m.f] = seq;
}

// A12 for every frame at its input point (body::mle_base).
__global__ void __launch_bounds__(kBlock) k_b_mle(const FDesc* __restrict__ D,
                                                  const Item* __restrict__ it, int n,
                                                  const fp2* __restrict__ ch, int nf)
{
    __shared__ fp2 zs[kMaxZ];
    const int i = find_item(it, n, blockIdx.x);
    const Item m = it[i];
    const FDesc& d = D[m.f];
    const int n_z = d.nin + d.ncb;
    for (int j = threadIdx.x; j < n_z; j += blockDim.x) zs[j] = ch[(size_t)(SPT + j) * nf + m.f];
    __syncthreads();
    const int local = blockIdx.x - m.b0;
    body::mle_base(d.rows_in, zs, n_z, d.partials, local, it[i + 1].b0 - m.b0, local);
}

// ------------------------------------------------------------ host layout

size_t al(size_t n) { return (n + 255) & ~(size_t)255; }

// Early claim work of round rd with `items` (pair, chunk) items per frame.
ClaimShape claim_of(const Shape& S, int rd, long long items)
{
    return claim_shape(S.n_terms, (S.cp >> rd) >> 1, items);
}

int mle_blocks(const Shape& S)
{
    const int lb = S.n_z < 10 ? S.n_z : 10;
    const size_t rows = (size_t)1 << (S.n_z - lb);
    return (int)(rows < (size_t)kMleCap ? rows : (size_t)kMleCap);
}

size_t partial_fp2(const Shape& S)
{
    size_t n = (size_t)mle_blocks(S);
    // sized for the largest claim candidate, so a frame's footprint does not
    // depend on the schedule the tuner picks
    for (int rd = 0; rd < S.ncb; ++rd)
        n = std::max(n, (size_t)claim_of(S, rd, kClaimCand[2]).blocks * 4);
    return n;
}

// Bytes of a frame's tables region, and the pointers inside it from base.
size_t tables_layout(const Shape& S, char* base, FDesc* d)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += al(bytes);
        return p;
    };
    const size_t tab = (size_t)S.NW * S.cp, half = (size_t)S.NW * (S.cp / 2);
    const size_t g8 = (size_t)S.n_gates * 8;
    char* rows_in = take(S.rows_n * 8);
    char* vr0 = take(tab * 8);
    char* vi0 = take(tab * 8);
    char* vr1 = take(half * 8);
    char* vi1 = take(half * 8);
    char* br0 = take((size_t)S.cp * 8);
    char* bi0 = take((size_t)S.cp * 8);
    char* br1 = take(((size_t)S.cp / 2 + 1) * 8);
    char* bi1 = take(((size_t)S.cp / 2 + 1) * 8);
    char* z1cr = take(g8);
    char* z1ci = take(g8);
    char* z1lr = take(g8);
    char* z1li = take(g8);
    char* othr = take(g8);
    char* othi = take(g8);
    char* tzr = take((size_t)S.n_terms * 8);
    char* tzi = take((size_t)S.n_terms * 8);
    char* parts = take(partial_fp2(S) * sizeof(fp2));
    char* small = take(kSmall * sizeof(fp2));
    if (d) {
        d->rows_in = (u64*)rows_in;
        d->vr[0] = (u64*)vr0;
        d->vi[0] = (u64*)vi0;
        d->vr[1] = (u64*)vr1;
        d->vi[1] = (u64*)vi1;
        d->br[0] = (u64*)br0;
        d->bi[0] = (u64*)bi0;
        d->br[1] = (u64*)br1;
        d->bi[1] = (u64*)bi1;
        d->z1cr = (u64*)z1cr;
        d->z1ci = (u64*)z1ci;
        d->z1lr = (u64*)z1lr;
        d->z1li = (u64*)z1li;
        d->othr = (u64*)othr;
        d->othi = (u64*)othi;
        d->tzr = (u64*)tzr;
        d->tzi = (u64*)tzi;
        d->partials = (fp2*)parts;
        d->small = (fp2*)small;
    }
    return off;
}

// A launch's items in the schedule: [first, first + n) plus a sentinel whose
// b0 is the launch's workgroup count.
struct Launch {
    int first = 0, n = 0, blocks = 0;
};

struct Sched {
    std::vector<Item> items;
    Launch begin() const
    {
        Launch l;
        l.first = (int)items.size();
        return l;
    }
    void add(Launch& l, int f, int blocks, int a = 0, int b = 0)
    {
        if (blocks <= 0) return;
        items.push_back({f, l.blocks, a, b});
        l.blocks += blocks;
        ++l.n;
    }
    void end(Launch& l) { items.push_back({-1, l.blocks, 0, 0}); }
};

int elem_blocks(size_t n, int cap = kElemCap)
{
    const size_t b = (n + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > (size_t)cap ? (size_t)cap : b));
}

}  // namespace

size_t Shape::footprint() const
{
    return al(in_bytes ? in_bytes : 8) + al(out_n * 8) + al(out_n) +
           tables_layout(*this, nullptr, nullptr);
}

// The schedule's two knobs and their autotuner: coordinate descent over the
// candidates on real batches, cost = device time of the claim launches per
// (pair, gate) term and of the fold launches per algorithmic byte, each
// measured with events around those launches; first the claim candidates
// (fold at its default), then the fold candidates with the best claim.
struct Knobs {
    long long claim_items = kClaimItems;
    int fold_cap = kElemCap;
};

struct Tuner {
    int phase = 0;          // 0: claim candidates, 1: fold candidates, 2: settled
    int trial = 0;          // candidate under trial in this phase
    int best_c = 1, best_f = 1;
    double cost_c[3] = {0, 0, 0}, cost_f[3] = {0, 0, 0};
    bool fixed = false;     // knobs set by the caller
    Knobs fixed_knobs;

    Knobs pick() const
    {
        if (fixed) return fixed_knobs;
        Knobs k;
        k.claim_items = kClaimCand[phase == 0 ? trial : best_c];
        k.fold_cap = kFoldCand[phase == 1 ? trial : best_f];
        return k;
    }
    // one batch's costs under knobs k (ns per claim term, ns per fold byte)
    void record(double claim_cost, double fold_cost)
    {
        if (fixed || phase == 2) return;
        if (phase == 0) {
            cost_c[trial] = claim_cost;
            if (++trial == 3) {
                best_c = (int)(std::min_element(cost_c, cost_c + 3) - cost_c);
                phase = 1;
                trial = 0;
            }
        } else {
            cost_f[trial] = fold_cost;
            if (++trial == 3) {
                best_f = (int)(std::min_element(cost_f, cost_f + 3) - cost_f);
                phase = 2;
            }
        }
    }
};

struct BatchCtx {
    hipStream_t st = nullptr;
    Tuner tuner;
    std::vector<hipEvent_t> tev;   // timing events around claim and fold launches
    DevBuf arena;
    uint8_t* h_in = nullptr;
    size_t h_in_bytes = 0;
    uint8_t* h_out8 = nullptr;    // outputs as bytes
    size_t h_out8_bytes = 0;
    uint8_t* h_stage = nullptr;   // descriptors and items, copied up per batch
    size_t h_stage_bytes = 0;
    void* h_map = nullptr;        // challenge table, published values, sequence words, flags
    void* d_map = nullptr;
    int map_frames = 0;
    std::vector<hipEvent_t> ev;
    unsigned seq = 0;

    fp2* hch() const { return static_cast<fp2*>(h_map); }
    fp2* dch_mapped() const { return static_cast<fp2*>(d_map); }   // device view of hch
    fp2* hv(int nf) const { return static_cast<fp2*>(h_map) + (size_t)kSlots * nf; }
    fp2* dv(int nf) const { return static_cast<fp2*>(d_map) + (size_t)kSlots * nf; }
    volatile unsigned* hseq(int nf) const
    {
        return reinterpret_cast<volatile unsigned*>(hv(nf) + (size_t)kVals * nf);
    }
    unsigned* dseq(int nf) const { return reinterpret_cast<unsigned*>(dv(nf) + (size_t)kVals * nf); }
    volatile unsigned* hflag(int nf) const { return hseq(nf) + nf; }
    unsigned* dflag(int nf) const { return dseq(nf) + nf; }

    bool pinned(uint8_t*& p, size_t& have, size_t need)
    {
        if (need <= have && p) return true;
        free_host(p);
        p = nullptr;
        have = 0;
        TP_HIP_OK(hipHostMalloc((void**)&p, need ? need : 64));
        have = need;
        return true;
    }
    bool mapped(int nf)
    {
        if (nf <= map_frames && h_map) return true;
        free_host(h_map);
        h_map = d_map = nullptr;
        map_frames = 0;
        const size_t bytes = ((size_t)kSlots + kVals) * nf * sizeof(fp2) + (size_t)nf * 8;
        TP_HIP_OK(hipHostMalloc(&h_map, bytes, hipHostMallocMapped | hipHostMallocCoherent));
        TP_HIP_OK(hipHostGetDevicePointer(&d_map, h_map, 0));
        memset(h_map, 0, bytes);
        map_frames = nf;
        return true;
    }
    bool timing_events(int n)
    {
        while ((int)tev.size() < n) {
            hipEvent_t e;
            TP_HIP_OK(hipEventCreate(&e));
            tev.push_back(e);
        }
        return true;
    }
    bool events(int n)
    {
        while ((int)ev.size() < n) {
            hipEvent_t e;
            TP_HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventBlockingSync));
            ev.push_back(e);
        }
        return true;
    }
};

void batch_tuning(const BatchCtx& B, double state[9])
{
    const Tuner& t = B.tuner;
    const Knobs k = t.pick();
    state[0] = t.fixed ? 2 : t.phase;
    state[1] = (double)k.claim_items;
    state[2] = (double)k.fold_cap;
    for (int i = 0; i < 3; ++i) {
        state[3 + i] = t.cost_c[i];
        state[6 + i] = t.cost_f[i];
    }
}

void batch_fix_knobs(BatchCtx& B, long long claim_items, int fold_cap)
{
    B.tuner.fixed = true;
    B.tuner.fixed_knobs.claim_items = claim_items;
    B.tuner.fixed_knobs.fold_cap = fold_cap;
}

namespace {

// Waits for the sequence word of every frame of a launch; fails on a stream
// error or a stream that drained without publishing.
bool wait_frames(BatchCtx& B, int nf, const Item* items, int n_items, unsigned s, const char* what)
{
    volatile unsigned* seq = B.hseq(nf);
    for (int i = 0; i < n_items; ++i) {
        const int f = items[i].f;
        long spins = 0;
        while (seq[f] != s) {
            if (++spins % 2048 == 0) {
                const hipError_t q = hipStreamQuery(B.st);
                if (q != hipSuccess && q != hipErrorNotReady)
                    return fail("batch %s: stream error %s", what, hipGetErrorString(q));
                if (q == hipSuccess && seq[f] != s)
                    return fail("batch %s: frame %d not published", what, f);
            } else {
                __builtin_ia32_pause();
            }
        }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return true;
}

// Waits for spawned tasks and drains the stream on every exit path of a batch.
struct BatchExit {
    TaskPool& tasks;
    TaskPool::Group& g;
    hipStream_t st;
    ~BatchExit()
    {
        tasks.wait(g);
        (void)hipStreamSynchronize(st);
    }
};

}  // namespace

BatchCtx* batch_ctx_create()
{
    auto* B = new BatchCtx();
    bool ok = hipSetDevice(0) == hipSuccess &&
              hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) == hipSuccess;
    if (ok) {
        const int lds_late = (int)fold_lds(kMaxNin);
        const int lds_line = (int)std::max(fold_lds(kMaxNin) + kMaxNin * sizeof(fp2),
                                           (((size_t)1 << kMaxNin) + ((size_t)2 << kMaxNin)) *
                                               sizeof(fp2));
        ok = hipFuncSetAttribute((const void*)k_b_late, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 lds_late) == hipSuccess &&
             hipFuncSetAttribute((const void*)k_b_line, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 lds_line) == hipSuccess;
    }
    if (!ok) {
        fail("batch driver: stream or LDS limit setup failed");
        if (B->st) (void)hipStreamDestroy(B->st);
        delete B;
        return nullptr;
    }
    return B;
}

void batch_ctx_destroy(BatchCtx* B)
{
    if (!B) return;
    if (B->st) (void)hipStreamSynchronize(B->st);
    B->arena.release();
    free_host(B->h_in);
    free_host(B->h_out8);
    free_host(B->h_stage);
    free_host(B->h_map);
    for (hipEvent_t e : B->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : B->tev) (void)hipEventDestroy(e);
    if (B->st) (void)hipStreamDestroy(B->st);
    delete B;
}

namespace {
double ms_since(Clock::time_point t)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}
// Adds milliseconds to an atomic microsecond counter.
void add_us(std::atomic<long long>& acc, Clock::time_point t0, Clock::time_point t1)
{
    acc.fetch_add((long long)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
}
}  // namespace

bool prove_batch(BatchCtx& B, std::vector<Job*>& jobs_in, TaskPool& tasks, double* prof)
{
    // largest outputs first: their downloads land and their transcript heads
    // (the longest host tasks) start first, so the heads finish together
    std::vector<std::pair<Job*, Shape>> order;
    for (Job* J : jobs_in) {
        Shape s;
        if (!s.set(*J)) {
            job_fail(*J);
            continue;
        }
        order.push_back({J, s});
    }
    std::stable_sort(order.begin(), order.end(), [](const std::pair<Job*, Shape>& a,
                                                    const std::pair<Job*, Shape>& b) {
        return a.second.out_n > b.second.out_n;
    });
    std::vector<Job*> jobs;
    std::vector<Shape> S;
    for (auto& o : order) {
        jobs.push_back(o.first);
        S.push_back(o.second);
    }
    const int nf = (int)jobs.size();
    if (nf == 0) return true;
    hipStream_t st = B.st;
    auto t_mark = Clock::now();
    double ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto lap = [&](int k) {
        const auto t = Clock::now();
        ph[k] = std::chrono::duration<double, std::milli>(t - t_mark).count();
        t_mark = t;
    };

    // 1. layout: inputs, outputs and tables regions of one arena
    std::vector<size_t> in_off(nf), out_off(nf), out8_off(nf), tab_off(nf);
    size_t in_total = 0, out_total = 0, out8_total = 0, tab_total = 0;
    int max_ncb = 0, max_nin = 0, max_nz = 0;
    for (int f = 0; f < nf; ++f) {
        in_off[f] = in_total;
        in_total += al(S[f].in_bytes ? S[f].in_bytes : 8);
        out_off[f] = out_total;
        out_total += al(S[f].out_n * 8);
        out8_off[f] = out8_total;
        out8_total += al(S[f].out_n);
        tab_off[f] = tab_total;
        tab_total += tables_layout(S[f], nullptr, nullptr);
        max_ncb = std::max(max_ncb, S[f].ncb);
        max_nin = std::max(max_nin, S[f].nin);
        max_nz = std::max(max_nz, S[f].n_z);
    }

    // 2. the schedule (every launch's items) from the shapes and the knobs
    const Knobs kn = B.tuner.pick();
    Sched sc;
    Launch L_rows = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_rows, f, elem_blocks(S[f].rows_n));
    sc.end(L_rows);
    Launch L_eval = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_eval, f, elem_blocks(S[f].out_n));
    sc.end(L_eval);
    Launch L_pack = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_pack, f, elem_blocks(S[f].out_n));
    sc.end(L_pack);
    Launch L_wire = sc.begin();
    for (int f = 0; f < nf; ++f) {
        const int gx = (S[f].cp + 63) / 64, gy = (S[f].NW + 63) / 64;
        sc.add(L_wire, f, gx * gy, gx);
    }
    sc.end(L_wire);
    Launch L_eq = sc.begin();
    for (int f = 0; f < nf; ++f) {
        const int bb = (S[f].cp + kBlock - 1) / kBlock, bz = (S[f].n_gates + kBlock - 1) / kBlock;
        sc.add(L_eq, f, bb + bz, bb);
    }
    sc.end(L_eq);
    Launch L_terms = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_terms, f, (S[f].n_terms * 64 + kBlock - 1) / kBlock);
    sc.end(L_terms);
    std::vector<Launch> L_claim(max_ncb), L_pub4(max_ncb), L_fold(max_ncb);
    double claim_terms = 0, fold_bytes = 0;
    for (int rd = 0; rd < max_ncb; ++rd) {
        L_claim[rd] = sc.begin();
        for (int f = 0; f < nf; ++f) {
            if (rd >= S[f].ncb) continue;
            const ClaimShape c = claim_of(S[f], rd, kn.claim_items);
            sc.add(L_claim[rd], f, (int)c.blocks, c.n_chunks, c.gpt);
            claim_terms += (double)((S[f].cp >> rd) >> 1) * S[f].n_terms;
        }
        sc.end(L_claim[rd]);
        L_pub4[rd] = sc.begin();
        for (int f = 0; f < nf; ++f) {
            if (rd >= S[f].ncb) continue;
            sc.add(L_pub4[rd], f, 1, (int)claim_of(S[f], rd, kn.claim_items).blocks);
        }
        sc.end(L_pub4[rd]);
        L_fold[rd] = sc.begin();
        for (int f = 0; f < nf; ++f) {
            if (rd >= S[f].ncb) continue;
            const size_t total = ((size_t)S[f].NW + 1) * (size_t)((S[f].cp >> rd) >> 1);
            sc.add(L_fold[rd], f, elem_blocks(total, kn.fold_cap));
            fold_bytes += 48.0 * (double)total;
        }
        sc.end(L_fold[rd]);
    }
    Launch L_linit = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_linit, f, (S[f].n_gates + kBlock - 1) / kBlock);
    sc.end(L_linit);
    std::vector<Launch> L_late(2 * max_nin);
    for (int k = 0; k < 2 * max_nin; ++k) {
        L_late[k] = sc.begin();
        for (int f = 0; f < nf; ++f)
            if (k < 2 * S[f].nin) sc.add(L_late[k], f, 1);
        sc.end(L_late[k]);
    }
    Launch L_line = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_line, f, S[f].nin + 2, S[f].nin + 1);
    sc.end(L_line);
    Launch L_all = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_all, f, 1);
    sc.end(L_all);
    Launch L_mle = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_mle, f, mle_blocks(S[f]));
    sc.end(L_mle);
    Launch L_pub1 = sc.begin();
    for (int f = 0; f < nf; ++f) sc.add(L_pub1, f, 1, mle_blocks(S[f]));
    sc.end(L_pub1);

    // 3. memory: device arena [inputs | outputs | outputs as bytes | tables |
    // descriptors | items | challenges], pinned staging, the mapped table
    const size_t desc_bytes = al((size_t)nf * sizeof(FDesc));
    const size_t item_bytes = al(sc.items.size() * sizeof(Item));
    const size_t ch_bytes = al((size_t)kSlots * nf * sizeof(fp2));
    const size_t o_out = in_total, o_out8 = o_out + out_total, o_tab = o_out8 + out8_total;
    const size_t o_desc = o_tab + tab_total;
    const size_t o_item = o_desc + desc_bytes, o_ch = o_item + item_bytes;
    const size_t arena_bytes = o_ch + ch_bytes;
    if (!B.arena.ensure(arena_bytes) || !B.pinned(B.h_in, B.h_in_bytes, in_total) ||
        !B.pinned(B.h_out8, B.h_out8_bytes, out8_total) ||
        !B.pinned(B.h_stage, B.h_stage_bytes, desc_bytes + item_bytes) || !B.mapped(nf))
        return false;
    char* base = B.arena.as<char>();
    FDesc* hd = reinterpret_cast<FDesc*>(B.h_stage);
    for (int f = 0; f < nf; ++f) {
        const Job& J = *jobs[f];
        const Circuit& C = *J.circ;
        FDesc& d = hd[f];
        memset(&d, 0, sizeof(d));
        d.op = C.d_op.as<int>();
        d.mux = C.d_mux.as<int>();
        d.gt = C.d_type.as<int>();
        d.tt = C.d_ttype.as<int>();
        d.tin0 = C.d_tin0.as<int>();
        d.tin1 = C.d_tin1.as<int>();
        d.toff = C.d_toff.as<int>();
        d.tsrc = C.d_tsrc.as<int>();
        d.tneg = C.d_tneg.as<int>();
        d.n_terms = C.n_terms;
        d.in0 = C.d_in0.as<int>();
        d.in1 = C.d_in1.as<int>();
        d.slot = C.d_slot.as<int>();
        d.cval = C.d_cval.as<u64>();
        d.n_gates = C.n_gates;
        d.nin = S[f].nin;
        d.nout = S[f].nout;
        d.n_priv = J.n_priv;
        d.n_copies = J.n_copies;
        d.cp = S[f].cp;
        d.ncb = S[f].ncb;
        d.in_bytes = J.in_bytes;
        d.raw = base + in_off[f];
        d.rows_out = reinterpret_cast<u64*>(base + o_out + out_off[f]);
        d.out8 = reinterpret_cast<uint8_t*>(base + o_out8 + out8_off[f]);
        d.big = B.dflag(nf) + f;
        tables_layout(S[f], base + o_tab + tab_off[f], &d);
    }
    memcpy(B.h_stage + desc_bytes, sc.items.data(), sc.items.size() * sizeof(Item));
    const FDesc* dD = reinterpret_cast<const FDesc*>(base + o_desc);
    const Item* dI = reinterpret_cast<const Item*>(base + o_item);
    fp2* dch = reinterpret_cast<fp2*>(base + o_ch);
    const Item* hI = sc.items.data();
    fp2* hch = B.hch();
    // this batch's sequence words start clear (the table's layout depends on nf)
    for (int f = 0; f < nf; ++f) B.hseq(nf)[f] = B.hflag(nf)[f] = 0;
    fp2* hv = B.hv(nf);
    fp2* dv = B.dv(nf);
    unsigned* dseq = B.dseq(nf);

    TaskPool::Group prefixes;
    std::vector<std::vector<u64>> wide(nf);   // u64 outputs of frames flagged big
    std::vector<std::unique_ptr<FrameProto>> pr(nf);
    BatchExit exit_guard{tasks, prefixes, st};
    std::vector<char> alive(nf, 1);
    auto frame_fail = [&](int f) {
        job_fail(*jobs[f]);
        alive[f] = 0;
    };

    // 4. inputs up, rows, evaluation, wire tables (A3, A4), outputs down
    auto t_stage = Clock::now();
    tasks.parallel_for(nf, [&](int f) {
        const Job& J = *jobs[f];
        if (S[f].in_bytes && !J.dev_in.ptr)
            memcpy(B.h_in + in_off[f],
                   J.in_bytes == 1 ? (const void*)J.in8.data() : (const void*)J.in64.data(),
                   S[f].in_bytes);
    });
    prof[BP_STAGE_WALL] += ms_since(t_stage);
    auto t_launch = Clock::now();
    TP_HIP_OK(hipMemcpyAsync(base, B.h_in, in_total, hipMemcpyHostToDevice, st));
    // frames submitted on the device: their inputs device to device, after
    // the staged host inputs (which hold stale bytes in their places)
    for (int f = 0; f < nf; ++f)
        if (jobs[f]->dev_in.ptr && S[f].in_bytes)
            TP_HIP_OK(hipMemcpyAsync(base + in_off[f], jobs[f]->dev_in.ptr, S[f].in_bytes,
                                     hipMemcpyDeviceToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(base + o_desc, B.h_stage, desc_bytes + item_bytes,
                             hipMemcpyHostToDevice, st));
    k_b_rows<<<L_rows.blocks, kBlock, 0, st>>>(dD, dI + L_rows.first, L_rows.n);
    k_b_eval<<<L_eval.blocks, kBlock, 0, st>>>(dD, dI + L_eval.first, L_eval.n);
    k_b_pack<<<L_pack.blocks, kBlock, 0, st>>>(dD, dI + L_pack.first, L_pack.n);
    TP_HIP_OK(hipGetLastError());
    // outputs (as bytes) in groups of frames, an event per group
    std::vector<int> grp(nf);
    {
        const size_t target = std::max((size_t)4 << 20, out8_total / 48 + 1);
        int g = 0;
        size_t acc = 0;
        int first = 0;
        std::vector<std::pair<int, int>> groups;
        for (int f = 0; f < nf; ++f) {
            grp[f] = g;
            acc += al(S[f].out_n);
            if (acc >= target || f == nf - 1) {
                groups.push_back({first, f + 1});
                first = f + 1;
                acc = 0;
                ++g;
            }
        }
        if (!B.events((int)groups.size())) return false;
        for (size_t k = 0; k < groups.size(); ++k) {
            const size_t a = out8_off[groups[k].first];
            const size_t e = groups[k].second < nf ? out8_off[groups[k].second] : out8_total;
            TP_HIP_OK(hipMemcpyAsync(B.h_out8 + a, base + o_out8 + a, e - a, hipMemcpyDeviceToHost,
                                     st));
            TP_HIP_OK(hipEventRecord(B.ev[k], st));
        }
    }
    // wire tables behind the copies, so the first outputs land as early as possible
    k_b_wire<<<L_wire.blocks, kBlock, 0, st>>>(dD, dI + L_wire.first, L_wire.n);
    TP_HIP_OK(hipGetLastError());
    prof[BP_LAUNCH_WALL] += ms_since(t_launch);
    lap(0);

    // 5. transcript heads: each frame as soon as its outputs have landed
    std::atomic<int> dev_err{0};
    std::atomic<long long> us_wait{0}, us_head{0}, us_prefix{0};
    auto t_heads = Clock::now();
    tasks.parallel_for(nf, [&](int f) {
        const auto t0 = Clock::now();
        if (hipEventSynchronize(B.ev[grp[f]]) != hipSuccess) {
            dev_err.store(1);
            return;
        }
        const auto t1 = Clock::now();
        pr[f].reset(new FrameProto(*jobs[f], S[f]));
        FrameProto* p = pr[f].get();
        if (B.hflag(nf)[f]) {
            // an output above 255: this frame's transcript takes the u64 outputs
            wide[f].resize(S[f].out_n);
            if (hipMemcpy(wide[f].data(), base + o_out + out_off[f], S[f].out_n * 8,
                          hipMemcpyDeviceToHost) != hipSuccess) {
                dev_err.store(1);
                return;
            }
            p->prepare(wide[f].data());
        } else {
            p->prepare8(B.h_out8 + out8_off[f]);
        }
        tasks.spawn(prefixes, [p, &us_prefix] {
            const auto a = Clock::now();
            p->write_prefix();
            add_us(us_prefix, a, Clock::now());
        });
        p->head();
        for (int j = 0; j < S[f].nout; ++j) hch[(size_t)(SZ1 + j) * nf + f] = p->z1[j];
        for (int j = 0; j < S[f].ncb; ++j) hch[(size_t)(SZ2 + j) * nf + f] = p->z2[j];
        add_us(us_wait, t0, t1);
        add_us(us_head, t1, Clock::now());
    });
    if (dev_err.load()) return fail("batch: outputs download failed");
    prof[BP_HEAD_WALL] += ms_since(t_heads);
    lap(1);
    auto t_rounds = Clock::now();
    double step_wait = 0, step_host = 0;
    auto timed_wait = [&](const Item* items, int n, unsigned s, const char* what) {
        const auto t = Clock::now();
        const bool ok = wait_frames(B, nf, items, n, s, what);
        step_wait += ms_since(t);
        return ok;
    };
    auto timed_for = [&](int n, const std::function<void(int)>& fn) {
        const auto t = Clock::now();
        tasks.parallel_for(n, fn);
        step_host += ms_since(t);
    };

    auto pull = [&](int s0, int ns) -> bool {
        const int n = ns * nf;
        k_b_pull<<<grid_for((size_t)n, kBlock, 64), kBlock, 0, st>>>(
            B.dch_mapped() + (size_t)s0 * nf, dch + (size_t)s0 * nf, n);
        TP_HIP_OK(hipGetLastError());
        return true;
    };

    // 6. eq tables (A5)
    if (!pull(SZ1, SER - SZ1)) return false;
    k_b_eq<<<L_eq.blocks, kBlock, 0, st>>>(dD, dI + L_eq.first, L_eq.n, dch, nf);
    TP_HIP_OK(hipGetLastError());
    k_b_terms<<<L_terms.blocks, kBlock, 0, st>>>(dD, dI + L_terms.first, L_terms.n);
    TP_HIP_OK(hipGetLastError());
    lap(2);

    // 7. early rounds in lockstep (A6, A7); the claim and fold launches are
    // timed for the schedule tuner
    if (!B.timing_events(4 * max_ncb)) return false;
    for (int rd = 0; rd < max_ncb; ++rd) {
        const Launch& lc = L_claim[rd];
        const Launch& lp = L_pub4[rd];
        const Launch& lf = L_fold[rd];
        TP_HIP_OK(hipEventRecord(B.tev[4 * rd], st));
        k_b_claim<<<lc.blocks, kBlock, 0, st>>>(dD, dI + lc.first, lc.n, rd);
        TP_HIP_OK(hipEventRecord(B.tev[4 * rd + 1], st));
        const unsigned s = ++B.seq;
        k_b_pub<4><<<lp.n, kBlock, 0, st>>>(dD, dI + lp.first, dv, dseq, s);
        TP_HIP_OK(hipGetLastError());
        if (!timed_wait(hI + lp.first, lp.n, s, "early claim")) return false;
        timed_for(lp.n, [&](int i) {
            const int f = hI[lp.first + i].f;
            fp2 r = zero2();
            if (alive[f]) {
                fp2 fv[4];
                for (int t = 0; t < 4; ++t) fv[t] = hv[(size_t)f * kVals + t];
                if (!pr[f]->early(rd, fv, r)) frame_fail(f);
            }
            hch[(size_t)(SER + rd) * nf + f] = r;
        });
        if (!pull(SER + rd, 1)) return false;
        TP_HIP_OK(hipEventRecord(B.tev[4 * rd + 2], st));
        k_b_fold<<<lf.blocks, kBlock, 0, st>>>(dD, dI + lf.first, lf.n, rd, dch, nf);
        TP_HIP_OK(hipEventRecord(B.tev[4 * rd + 3], st));
        TP_HIP_OK(hipGetLastError());
    }
    lap(3);

    // 8. late rounds in lockstep (A8, A9)
    k_b_linit<<<L_linit.blocks, kBlock, 0, st>>>(dD, dI + L_linit.first, L_linit.n);
    TP_HIP_OK(hipGetLastError());
    const size_t lds_late = fold_lds(max_nin);
    for (int k = 0; k < 2 * max_nin; ++k) {
        const Launch& ll = L_late[k];
        const unsigned s = ++B.seq;
        k_b_late<<<ll.n, kOneWg, lds_late, st>>>(dD, dI + ll.first, k, dch, nf, dv, dseq, s);
        TP_HIP_OK(hipGetLastError());
        if (!timed_wait(hI + ll.first, ll.n, s, "late round")) return false;
        timed_for(ll.n, [&](int i) {
            const int f = hI[ll.first + i].f;
            fp2 r = zero2();
            if (alive[f]) {
                fp2 fv[3];
                for (int t = 0; t < 3; ++t) fv[t] = hv[(size_t)f * kVals + t];
                if (!pr[f]->late(k, fv, r)) frame_fail(f);
            }
            hch[(size_t)(SLC + k) * nf + f] = r;
        });
        if (!pull(SLC + k, 1)) return false;
    }
    lap(4);

    // 9. line values and wiring sums (A10, A11), the end of the layer
    {
        size_t lds = 0;
        for (int f = 0; f < nf; ++f) {
            lds = std::max(lds, fold_lds(S[f].nin) + (size_t)S[f].nin * sizeof(fp2));
            lds = std::max(lds, (((size_t)1 << S[f].nout) + ((size_t)2 << S[f].nin)) * sizeof(fp2));
        }
        k_b_line<<<L_line.blocks, kOneWg, lds, st>>>(dD, dI + L_line.first, L_line.n, dch, nf);
        const unsigned s = ++B.seq;
        k_b_pubcopy<<<L_all.n, 64, 0, st>>>(dD, dI + L_all.first, dv, dseq, s);
        TP_HIP_OK(hipGetLastError());
        if (!timed_wait(hI + L_all.first, L_all.n, s, "line values")) return false;
        timed_for(nf, [&](int f) {
            if (alive[f]) {
                const fp2* v = hv + (size_t)f * kVals;
                if (!pr[f]->finish_layer(v, v + 16)) frame_fail(f);
            }
            for (int j = 0; j < S[f].n_z; ++j)
                hch[(size_t)(SPT + j) * nf + f] = alive[f] ? pr[f]->point.z[j] : zero2();
        });
        if (!pull(SPT, max_nz)) return false;
    }
    lap(5);

    // 10. input MLE at each frame's point (A12) and its value proof
    {
        k_b_mle<<<L_mle.blocks, kBlock, 0, st>>>(dD, dI + L_mle.first, L_mle.n, dch, nf);
        const unsigned s = ++B.seq;
        k_b_pub<1><<<L_pub1.n, kBlock, 0, st>>>(dD, dI + L_pub1.first, dv, dseq, s);
        TP_HIP_OK(hipGetLastError());
        if (!timed_wait(hI + L_pub1.first, L_pub1.n, s, "input MLE")) return false;
        timed_for(nf, [&](int f) {
            if (alive[f]) pr[f]->input_value(hv[(size_t)f * kVals]);
        });
    }
    lap(6);
    prof[BP_ROUNDS_WALL] += ms_since(t_rounds);
    {   // the schedule tuner's costs of this batch (its early rounds are done)
        double claim_ms = 0, fold_ms = 0;
        for (int rd = 0; rd < max_ncb; ++rd) {
            float a = 0, b = 0;
            TP_HIP_OK(hipEventElapsedTime(&a, B.tev[4 * rd], B.tev[4 * rd + 1]));
            TP_HIP_OK(hipEventElapsedTime(&b, B.tev[4 * rd + 2], B.tev[4 * rd + 3]));
            claim_ms += a;
            fold_ms += b;
        }
        prof[BP_CLAIM_DEV] += claim_ms;
        prof[BP_FOLD_DEV] += fold_ms;
        if (claim_terms >= (double)kTuneMinItems)
            B.tuner.record(claim_ms * 1e6 / claim_terms, fold_ms * 1e6 / fold_bytes);
    }
    prof[BP_STEP_WAIT] += step_wait;
    prof[BP_STEP_HOST_WALL] += step_host;
    auto t_rest = Clock::now();

    // 11. proof bytes: the prefixes are pickled by now or soon
    tasks.wait(prefixes);
    tasks.parallel_for(nf, [&](int f) {
        if (alive[f]) pr[f]->write_rest();
    });
    lap(7);
    prof[BP_REST_WALL] += ms_since(t_rest);
    prof[BP_HEAD_WAIT_SUM] += us_wait.load() / 1e3;
    prof[BP_HEAD_CPU_SUM] += us_head.load() / 1e3;
    prof[BP_PREFIX_CPU_SUM] += us_prefix.load() / 1e3;
    prof[BP_BATCHES] += 1;
    prof[BP_FRAMES] += nf;
    for (int f = 0; f < nf; ++f)
        for (int k = 0; k < 8; ++k) jobs[f]->phase[k] = ph[k];
    return true;
}

}  // namespace frame
}  // namespace tp
