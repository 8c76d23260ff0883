// frame.hip - device-resident prover of whole TruePix frames, and a pool of
// such provers for batched and served operation.
//
// One frame is one run of CircuitProverNIZK.run
// (TP/libTruePix/circuitnizkvec_vole.py:265-501) over a one-layer circuit,
// such as the four edit circuits (TP/pws/{gray,inv,Crop,mask}.pws): circuit
// evaluation, the nCopyBits early (cubic) rounds, the 2 nInBits late
// (quadratic) rounds, the line projection, the final product proof, the
// wiring predicate, the input MLE and the proof bytes.  Every protocol step,
// transcript absorb and challenge draw happens in the reference's order
// (FrameProto below), so the proof bytes equal the reference prover's for the
// same inputs, VOLE correlations, commitment binding and pid.
//
// Device residency: inputs go up once (the private bytes), outputs come down
// once (they are absorbed into the transcript and pickled into the proof);
// between those, tables stay in HBM, challenges travel as kernel arguments
// and each round's sums come back through a mapped host slot that the host
// polls (no stream synchronization per round).
//
// The pool has two kinds of drivers taking frames from one FIFO queue:
//  - stream workers: one host thread per worker, each with its own HIP
//    stream, mapped slot and grow-only buffers, proving one frame at a time
//    (prove_frame); kernels of different workers overlap on the GPU;
//  - lockstep batch drivers (batch.hip): each takes every queued frame that
//    fits its batch size and the pool's device-memory budget, and proves them
//    together, one launch per protocol step across the whole batch, with the
//    per-frame transcript work spread over a shared pool of host threads.
// A batch is a queue filled at once; a served stream is a queue filled as
// requests arrive (continuous batching: frames that arrive while a batch runs
// form the next one).
#include "frame.hpp"
#include "layer_bodies.hpp"
#include "reduce.hpp"
#include "relay.hpp"
#include "../../include/truepix_hip.h"

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <unordered_map>

using namespace tp;
using namespace tp::frame;

namespace {

// ------------------------------------------------------------------ kernels

__global__ void k_fe_rows(const uint8_t* __restrict__ src8, const u64* __restrict__ src64,
                          int n_priv, int n_copies, int cp, int row_log,
                          const int* __restrict__ slot, const u64* __restrict__ cval,
                          u64* __restrict__ rows)
{
    body::input_rows(src8, src64, n_priv, n_copies, cp, row_log, slot, cval, rows, blockIdx.x,
                     gridDim.x);
}

__global__ void k_fe_eq(PointArg z, int n, u64* __restrict__ outr, u64* __restrict__ outi,
                        int n_out)
{
    __shared__ fp2 zs[kMaxZ];
    for (int j = threadIdx.x; j < n; j += blockDim.x) zs[j] = z.z[j];
    __syncthreads();
    body::eq_table(zs, n, outr, outi, n_out, blockIdx.x);
}

__global__ void k_fe_late_init(const u64* __restrict__ vr, const u64* __restrict__ vi,
                               const int* __restrict__ in1, int n_gates,
                               const u64* __restrict__ z1cr, const u64* __restrict__ z1ci,
                               u64* __restrict__ z1lr, u64* __restrict__ z1li,
                               u64* __restrict__ othr, u64* __restrict__ othi,
                               const u64* __restrict__ br, const u64* __restrict__ bi,
                               u64* __restrict__ beta2)
{
    body::late_init(vr, vi, in1, n_gates, z1cr, z1ci, z1lr, z1li, othr, othi, br, bi, beta2,
                    blockIdx.x);
}

// The 2 nin late rounds of a frame in one resident workgroup (A9:
// TP/libTruePix/layerprover.py:225-244,282-285, gateprover.py:102-176): each
// round's three sums go to the host as tagged words (relay.hpp, sequence
// number seq0 + round); the host's transcript answers with the round's
// challenge in the answer slot, which the workgroup waits for (bounded by
// spin_ticks) before the next round applies it; the w2 pass starts with
// body::late_switch.  One launch instead of one per round plus a switch.
__global__ void __launch_bounds__(kOneWg) k_fe_late_resident(
    const int* __restrict__ gt, const int* __restrict__ in0, const int* __restrict__ in1,
    u64* __restrict__ z1r, u64* __restrict__ z1i, u64* __restrict__ oth_r,
    u64* __restrict__ oth_i, const u64* __restrict__ vr, const u64* __restrict__ vi, int nin,
    int n_gates, const u64* __restrict__ beta, fp2* host_vals, unsigned* host_seq,
    unsigned seq0, const u64* answer, long long spin_ticks)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 cs[2 * kMaxNin];
    __shared__ int s_ok;
    for (int rd = 0; rd < 2 * nin; ++rd) {
        if (rd == nin) {
            body::late_switch(vr, vi, nin, cs, n_gates, oth_r, oth_i, lds);
            __syncthreads();   // the fold buffers are reused below
        }
        fp2 out[3];
        body::late_round(gt, in0, in1, z1r, z1i, oth_r, oth_i, vr, vi, cs, nin, n_gates, rd,
                         rd > 0 ? 1 : 0, beta, lds, out);
        if (threadIdx.x == 0) {
            publish_tagged<3>(host_vals, out, seq0 + (unsigned)rd);
            int ok = 1;
            if (rd + 1 < 2 * nin) {
                u64 a[2];
                ok = take_answer(answer, seq0 + (unsigned)rd, wall_clock64(), spin_ticks, a);
                if (ok) cs[rd] = mk2(a[0], a[1]);
                else __hip_atomic_store(host_seq, kSeqAbort, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_SYSTEM);
            }
            s_ok = ok;
        }
        __syncthreads();   // cs[rd] (and every thread's LDS use of this round) settled
        if (!s_ok) return;
    }
}

// Copies n <= 64 device values to the mapped slot, then the sequence number.
__global__ void k_publish_copy(const fp2* __restrict__ src, int n, fp2* host_vals,
                               volatile unsigned* host_seq, unsigned seq)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        for (int k = 0; k < n; ++k) host_vals[k] = src[k];
        __threadfence_system();
        *host_seq = seq;
    }
}

// --------------------------------------------------------------- host math

const u64 kHalf = 1ULL << 60;                  // 1/2 mod p
const u64 kThird = (2 * P + 1) / 3;            // 1/3 mod p

fp2 sc(u64 s, fp2 x) { return smul2(s, x); }

u64 powmod(u64 a, u64 e)
{
    u64 r = 1;
    while (e) {
        if (e & 1) r = fmul(r, a);
        a = fmul(a, a);
        e >>= 1;
    }
    return r;
}

// TP/libTruePix/util.py:117-137: coefficients of the cubic through
// (0, f0), (1, f1), (-1, fm1), (2, f2).
void interp_cubic(const fp2 f[4], fp2 c[4])
{
    const fp2 f0 = f[0], f1 = f[1], fm1 = f[2], f2 = f[3];
    const fp2 b = sc(kHalf, add2(f1, fm1));
    const fp2 cc = sc(kHalf, sub2(f1, fm1));
    const fp2 d = sc(kThird, sub2(f2, fm1));
    const fp2 k2 = sub2(b, f0);
    const fp2 k3 = sc(kHalf, sub2(sub2(d, cc), k2));
    c[0] = f0;
    c[1] = sub2(cc, k3);
    c[2] = k2;
    c[3] = k3;
}

// TP/libTruePix/util.py:91-115: the quadratic through (0, f0), (1, f1), (-1, fm1).
void interp_quadratic(const fp2 f[3], fp2 c[3])
{
    const fp2 k2 = sub2(sc(kHalf, add2(f[1], f[2])), f[0]);
    c[0] = f[0];
    c[1] = sub2(sub2(f[1], k2), f[0]);
    c[2] = k2;
}

// TP/libTruePix/util.py:279-312: monomial coefficients of the interpolant
// through (t, y_t), t = 0..n-1, as out[i] = sum_j L[j][i] y_j with the
// base-field Lagrange basis coefficients L (cached per n).
const std::vector<u64>& lagrange_basis(int n)
{
    static std::mutex mu;
    static std::unordered_map<int, std::vector<u64>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    std::vector<u64> L((size_t)n * n, 0);
    for (int j = 0; j < n; ++j) {
        std::vector<u64> num(1, 1);
        u64 den = 1;
        for (int m = 0; m < n; ++m) {
            if (m == j) continue;
            std::vector<u64> nxt(num.size() + 1, 0);
            for (size_t i = 0; i < num.size(); ++i) {   // num * (x - m)
                nxt[i] = fsub(nxt[i], fmul((u64)m, num[i]));
                nxt[i + 1] = fadd(nxt[i + 1], num[i]);
            }
            num.swap(nxt);
            den = fmul(den, j >= m ? (u64)(j - m) : P - (u64)(m - j));
        }
        const u64 inv = powmod(den, P - 2);
        for (int i = 0; i < n; ++i) L[(size_t)j * n + i] = fmul(num[i], inv);
    }
    return cache.emplace(n, std::move(L)).first->second;
}

fp2 horner(const std::vector<fp2>& c, fp2 x)
{
    fp2 out = c.back();
    for (size_t k = c.size() - 1; k-- > 0;) out = add2(mul2(out, x), c[k]);
    return out;
}

// Commitment-side gate forms f(x, y, xy) of the prover
// (TP/libTruePix/gateprover_vole.py:706-720, GateFunctionsPC).
fp2 gate_com(int t, fp2 x, fp2 y, fp2 xy)
{
    const fp2 one = one2();
    switch (t) {
    case G_MUL: return xy;
    case G_ADD: return add2(x, y);
    case G_SUB: return sub2(x, y);
    case G_MUXL: return x;
    case G_MUXR: return y;
    case G_OR: return sub2(add2(x, y), xy);
    case G_XOR: return sub2(add2(x, y), dbl2(xy));
    case G_NOT: return sub2(one, x);
    case G_NAND: return sub2(one, xy);
    case G_NOR: return sub2(add2(one, xy), add2(x, y));
    case G_NXOR: return sub2(add2(one, dbl2(xy)), add2(x, y));
    case G_NAAB: return sub2(y, xy);
    default: return x;
    }
}

// Joins its thread when it goes out of scope (error returns included).
struct Joiner {
    std::thread t;
    explicit Joiner(std::thread&& th) : t(std::move(th)) {}
    void join()
    {
        if (t.joinable()) t.join();
    }
    ~Joiner() { join(); }
};

}  // namespace

namespace tp {
namespace frame {

int clog2(long long v)
{
    int b = 0;
    while ((1LL << b) < v) ++b;
    return b;
}

double now_s(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Waits for the n Fp2 values a resident kernel publishes as tagged words
// under sequence number seq (relay.hpp) in the slot's value words, copies
// them to out and empties the words.  Fails if the kernel gave up waiting
// (kSeqAbort) or after 30 s.  It makes no runtime call: the kernel is waiting
// for this thread, and a runtime call can queue behind another thread's
// device-wide synchronization, which waits for the kernel.
bool wait_tagged(MappedSlot& slot, unsigned seq, int n, fp2* out, const char* what)
{
    volatile u64* w = reinterpret_cast<volatile u64*>(slot.vals);
    const u64 tag = (u64)(seq & 7) << 61;
    auto ready = [&]() {
        for (int i = 0; i < 2 * n; ++i) {
            const u64 x = w[i];
            if ((x & ~P) != tag || (x & P) == P) return false;
        }
        return true;
    };
    const auto t0 = Clock::now();
    for (long spins = 1; !ready(); ++spins) {
        if (*slot.seq == kSeqAbort) return fail("%s: the device gave up waiting for a challenge", what);
        if (spins % 4096 == 0) {
            if (Clock::now() - t0 > std::chrono::seconds(30))
                return fail("%s: timed out waiting for the device", what);
        } else {
            __builtin_ia32_pause();
        }
    }
    for (int i = 0; i < n; ++i) out[i] = mk2(w[2 * i] & P, w[2 * i + 1] & P);
    for (int i = 0; i < 2 * n; ++i) w[i] = kSlotEmpty;
    return true;
}

// The resident late rounds' bound on each challenge wait: 10 s of the
// device's wall clock.
long long late_spin_ticks()
{
    static long long ticks = [] {
        int khz = 0, dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
            khz = 0;
        return (long long)(khz > 0 ? khz : 100000) * 1000LL * 10;
    }();
    return ticks;
}

bool Shape::set(const Job& J)
{
    const Circuit& C = *J.circ;
    nin = C.in_bits;
    nout = C.out_bits;
    n_gates = C.n_gates;
    n_terms = C.n_terms;
    if (J.n_copies < 2) return fail("frame: nCopies must be at least 2");
    ncb = clog2(J.n_copies);
    cp = 1 << ncb;
    NW = 1 << nin;
    NO = 1 << nout;
    n_z = nin + ncb;
    if (ncb > 16 || n_z > kMaxZ) return fail("frame: %d copies is beyond this build's limit", J.n_copies);
    rows_n = (size_t)cp << nin;
    out_n = (size_t)cp << nout;
    in_bytes = (size_t)J.n_copies * J.n_priv * (J.in_bytes == 1 ? 1 : 8);
    return true;
}

bool job_fail(Job& J)
{
    J.err = tp_last_error();
    if (J.err.empty()) J.err = "frame failed";
    return false;
}

// ------------------------------------------------------------ FrameProto

FrameProto::FrameProto(Job& job, const Shape& shape) : J(job), S(shape), pw(job.proof)
{
    memset(&point, 0, sizeof(point));
}

fp2 FrameProto::chal(const char* label) { return fs.rand_ext(label, (uint32_t)strlen(label)); }

fp2 FrameProto::chal_num(const char* prefix, int k)
{
    char buf[64];
    const uint32_t n = label_num(buf, prefix, k);
    return fs.rand_ext(buf, n);
}

// put(x) of an Fp2: the entry [x]
void FrameProto::put(fp2 x)
{
    fs.item_fp2(x);
    PNode l = PNode::list();
    l.kids.push_back(PNode::fp2v(x.re, x.im));
    trans.push_back(std::move(l));
}

// put((a, b)): the entry [(a, b)]
void FrameProto::put2(fp2 a, fp2 b)
{
    fs.item_fp2(a);
    fs.item_fp2(b);
    PNode t = PNode::tuple();
    t.kids.push_back(PNode::fp2v(a.re, a.im));
    t.kids.push_back(PNode::fp2v(b.re, b.im));
    PNode l = PNode::list();
    l.kids.push_back(std::move(t));
    trans.push_back(std::move(l));
}

// commit(m): the next VOLE correlation (TP/libTruePix/ExtVOLECommit.py:180-298,
// one security repetition); tag = V, correction d = m - U.
bool FrameProto::commit(fp2 m, fp2& tag, fp2& d)
{
    if (used >= (int)J.ur.size()) return fail("VOLE instance index out of range");
    tag = mk2(J.vr[used], J.vi[used]);
    d = sub2(m, mk2(J.ur[used], J.ui[used]));
    ++used;
    return true;
}

// Commit to a round polynomial's coefficients (create_pok_temp,
// circuitnizkvec_vole.py:112-133) and accumulate them for the final pok.
bool FrameProto::pok_temp(const fp2* outs, int n, std::vector<fp2>& rv)
{
    rv.resize(n);
    for (int k = 0; k < n; ++k) {
        fp2 d;
        if (!commit(outs[k], rv[k], d)) return false;
        put(d);
        if ((int)v_elem.size() < n) {
            v_elem.push_back(zero2());
            out_elem.push_back(zero2());
        }
    }
    for (int k = 0; k < n; ++k) {
        v_elem[k] = add2(v_elem[k], rv[k]);
        out_elem[k] = add2(out_elem[k], outs[k]);
    }
    return true;
}

// create_eq_proof (circuitnizkvec_vole.py:220-233)
void FrameProto::eq_proof(fp2 v2)
{
    const fp2 c = chal("eq");
    put(have_prev ? mul2(sub2(prev, v2), c) : mul2(v2, c));
}

bool FrameProto::round_tail(const fp2* outs, int n, int rd, fp2& nrand)
{
    std::vector<fp2> rv;
    if (!pok_temp(outs, n, rv)) return false;
    fp2 zp1 = rv[0];
    for (int k = 0; k < n; ++k) zp1 = add2(zp1, rv[k]);
    eq_proof(zp1);
    nrand = chal_num("sc-0-", rd);
    prev = horner(rv, nrand);
    have_prev = true;
    return true;
}

// Proof head (circuitnizkvec_vole.py:265-363): the io list of the proof
// refers to the outputs buffer, which must outlive write_prefix.
void FrameProto::prepare(const u64* outputs)
{
    outs = outputs;
    outs8 = nullptr;
    begin_io(PNode::int_array(outputs, S.out_n));
}

// The same with the outputs as bytes (every output below 256): the
// transcript absorbs them widened to u64, the proof lists the same ints.
void FrameProto::prepare8(const uint8_t* outputs)
{
    outs = nullptr;
    outs8 = outputs;
    begin_io(PNode::int_array8(outputs, S.out_n));
}

void FrameProto::begin_io(PNode&& values)
{
    const u64 n_inv = (u64)1 << S.n_z;
    io = PNode::list();
    PNode l = PNode::list();
    l.kids.push_back(PNode::none());
    io.kids.push_back(std::move(l));
    PNode m = PNode::list();
    m.kids.push_back(PNode::str("private_input_omitted"));
    m.kids.push_back(PNode::integer(n_inv));
    io.kids.push_back(std::move(m));
    io.kids.push_back(std::move(values));
    J.proof.clear();
    J.proof.reserve(S.out_n * 2 + (S.out_n / 1000 + 64) * 2 + (size_t)(S.out_n / 32768 + 8) * 9 +
                    65536);
}

void FrameProto::write_prefix()
{
    pw.start();
    pw.open_tuple();
    pw.save(PNode::integer(P));
    pw.save(io);
}

void FrameProto::head()
{
    const u64 n_inv = (u64)1 << S.n_z;
    fs.init("truepix");
    fs.int_list(&J.pid, 1);                                      // put(pid)
    {
        PNode l = PNode::list();
        l.kids.push_back(PNode::integer(J.pid));
        trans.push_back(std::move(l));
    }
    fs.item_none();                                              // put(muxbits=None, io)
    fs.item_str("private_input_omitted");                        // put([..omitted, n], io)
    fs.item_int(n_inv);
    {                                                            // put(outvals, io)
        const uint8_t tag = 0x01;
        fs.raw(&tag, 1);
        fs.le32((uint32_t)S.out_n);
        if (outs8) {
            u64 wide[512];   // little-endian host: the u64 array is the absorbed bytes
            for (size_t i = 0; i < S.out_n; i += 512) {
                const size_t m = S.out_n - i < 512 ? S.out_n - i : 512;
                for (size_t k = 0; k < m; ++k) wide[k] = outs8[i + k];
                fs.raw(wide, m * 8);
            }
        } else {
            fs.raw(outs, S.out_n * 8);
        }
    }
    {                                                            // put(commit_binding)
        PNode l = PNode::list();
        size_t si = 0, ii = 0;
        for (int kind : J.bkind) {
            if (kind == 1) {
                const std::string& s = J.bstr[si++];
                fs.item_str(s.data(), (uint32_t)s.size());
                l.kids.push_back(PNode::str(s));
            } else {
                const u64 v = J.bint[ii++];
                fs.item_int(v);
                l.kids.push_back(PNode::integer(v));
            }
        }
        trans.push_back(std::move(l));
    }
    if (J.keep_outputs) {
        if (outs8) J.outputs.assign(outs8, outs8 + S.out_n);
        else J.outputs.assign(outs, outs + S.out_n);
    }
    z1.resize(S.nout);
    z2.resize(S.ncb);
    for (int i = 0; i < S.nout; ++i) z1[i] = chal_num("z1-", i);
    for (int i = 0; i < S.ncb; ++i) z2[i] = chal_num("z2-", i);
}

// Early round rd from F(0), F(1), F(-1), F(2) (layerprover.py:195-224).
bool FrameProto::early(int rd, const fp2 f[4], fp2& r)
{
    fp2 c[4];
    interp_cubic(f, c);
    if (!round_tail(c, 4, rd, r)) return false;
    w3.push_back(r);
    return true;
}

// Late round k from f(0), f(1), f(-1) (layerprover.py:225-244).
bool FrameProto::late(int k, const fp2 f[3], fp2& r)
{
    fp2 c[3];
    interp_quadratic(f, c);
    if (!round_tail(c, 3, S.ncb + k, r)) return false;
    (k < S.nin ? w1 : w2).push_back(r);
    return true;
}

WiringArgs FrameProto::wiring_args() const
{
    WiringArgs a;
    memset(&a, 0, sizeof(a));
    for (int j = 0; j < S.nout; ++j) a.z1[j] = z1[j];
    for (int j = 0; j < S.nin; ++j) {
        a.w1[j] = w1[j];
        a.w2[j] = w2[j];
    }
    return a;
}

// The end of the layer: line projection H from V on the line
// (layercompute.py:169-243), create_final_prod_pok
// (circuitnizkvec_vole.py:175-218, called at :435), the wiring-predicate
// equality proof (:441-449), tau (:456-460), create_pok_final (:135-147,
// called at :485) and the input point (:496).
bool FrameProto::finish_layer(const fp2* line_vals, const fp2* wiring13)
{
    const int nin = S.nin;
    std::vector<fp2> H(nin + 1, zero2());
    {
        const int n = nin + 1;
        const std::vector<u64>& L = lagrange_basis(n);
        for (int j = 0; j < n; ++j) {
            const fp2 y = line_vals[j];
            for (int i = 0; i < n; ++i) H[i] = add2(H[i], smul2(L[(size_t)j * n + i], y));
        }
    }
    std::vector<fp2> vvals;
    fp2 prv[3];
    {
        fp2 v1, d1, v3, d3;
        fp2 sval = zero2();
        for (const fp2& h : H) sval = add2(sval, h);
        const fp2 prod = mul2(H[0], sval);
        if (!commit(H[0], v1, d1) || !commit(prod, v3, d3)) return false;
        put2(d1, d3);
        std::vector<fp2> rest_v;
        fp2 rest_m = zero2(), rest_t = zero2();
        for (size_t k = 1; k < H.size(); ++k) {
            fp2 v, d;
            if (!commit(H[k], v, d)) return false;
            rest_v.push_back(v);
            put(d);
            rest_m = add2(rest_m, H[k]);
            rest_t = add2(rest_t, v);
        }
        const fp2 c = chal("pok-vec");
        put2(mul2(rest_m, c), mul2(rest_t, c));
        vvals.push_back(v1);
        vvals.insert(vvals.end(), rest_v.begin(), rest_v.end());
        fp2 vs = zero2();
        for (const fp2& v : vvals) vs = add2(vs, v);
        prv[0] = v1;
        prv[1] = vs;
        prv[2] = v3;
        // prod_finish((outs[0], sval, prod), (v1, vs, v3))
        put2(mul2(prv[0], prv[1]), sub2(add2(mul2(H[0], prv[1]), mul2(sval, prv[0])), prv[2]));
    }
    {                                                            // wiring predicate tV (A11)
        fp2 mlx = one2();
        for (int j = 0; j < S.ncb; ++j) {
            const fp2 w = w3[j], z = z2[j];
            mlx = mul2(mlx, sub2(add2(dbl2(mul2(w, z)), one2()), add2(w, z)));
        }
        fp2 tv = zero2();
        const fp2 xy = prv[2];
        for (int t = 0; t < 13; ++t) tv = add2(tv, mul2(wiring13[t], gate_com(t, prv[0], prv[1], xy)));
        tv = mul2(tv, mlx);
        eq_proof(tv);
    }
    const fp2 tau = chal("tau-0");
    prev = horner(vvals, tau);
    std::vector<fp2> zline(nin);
    for (int j = 0; j < nin; ++j) zline[j] = add2(w1[j], mul2(sub2(w2[j], w1[j]), tau));
    {                                                            // create_pok_final
        const fp2 c = chal("pok-final");
        fp2 so = zero2(), sv = zero2();
        for (const fp2& x : out_elem) so = add2(so, x);
        for (const fp2& x : v_elem) sv = add2(sv, x);
        put2(mul2(so, c), mul2(sv, c));
    }
    // final input point = line point (nInBits) then w3 (nCopyBits)
    J.point.clear();
    memset(&point, 0, sizeof(point));
    PNode l = PNode::list();
    l.kids.push_back(PNode::str("gkr_input_point"));
    fs.item_str("gkr_input_point");
    PNode pts = PNode::list();
    for (int j = 0; j < S.n_z; ++j) {
        const fp2 e = j < nin ? zline[j] : w3[j - nin];
        point.z[j] = e;
        J.point.push_back(e.re);
        J.point.push_back(e.im);
        fs.item_int(e.re);
        fs.item_int(e.im);
        PNode xy = PNode::list();
        xy.kids.push_back(PNode::integer(e.re));
        xy.kids.push_back(PNode::integer(e.im));
        pts.kids.push_back(std::move(xy));
    }
    l.kids.push_back(std::move(pts));
    trans.push_back(std::move(l));
    return true;
}

// The input MLE value proof (create_input_value_proof,
// circuitnizkvec_vole.py:235-242, called at :498).
void FrameProto::input_value(fp2 v)
{
    J.value = v;
    const fp2 c = chal("input-val");
    put(mul2(prev, c));
}

// pickle((q, io, trans, ndb, rvstart, rvend, wDiv, 4), 5) after (q, io).
void FrameProto::write_rest()
{
    PNode tr = PNode::list();
    tr.kids = std::move(trans);
    pw.save(tr);
    for (int k = 0; k < 4; ++k) pw.save(PNode::none());
    pw.save(PNode::integer(4));
    pw.close_tuple();
    pw.finish();
}

// ------------------------------------------------------------ TaskPool

TaskPool::TaskPool(int n_threads)
{
    if (n_threads < 1) n_threads = 1;
    for (int i = 0; i < n_threads; ++i) th.emplace_back([this] { run(); });
}

TaskPool::~TaskPool()
{
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto& t : th) t.join();
}

void TaskPool::run()
{
    for (;;) {
        std::function<void()> f;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !hi.empty() || !lo.empty(); });
            if (!hi.empty()) {
                f = std::move(hi.front());
                hi.pop_front();
            } else if (!lo.empty()) {
                f = std::move(lo.front());
                lo.pop_front();
            } else {
                return;
            }
        }
        f();
    }
}

namespace {
struct ForState {
    std::atomic<int> next{0}, done{0};
    int n = 0;
    std::function<void(int)> fn;
    std::mutex mu;
    std::condition_variable cv;
};
}  // namespace

void TaskPool::parallel_for(int n, const std::function<void(int)>& fn)
{
    if (n <= 0) return;
    auto st = std::make_shared<ForState>();
    st->n = n;
    st->fn = fn;
    auto work = [st] {
        int k;
        while ((k = st->next.fetch_add(1)) < st->n) {
            st->fn(k);
            if (st->done.fetch_add(1) + 1 == st->n) {
                std::lock_guard<std::mutex> lk(st->mu);
                st->cv.notify_all();
            }
        }
    };
    const int helpers = std::min(n - 1, (int)th.size());
    if (helpers > 0) {
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int i = 0; i < helpers; ++i) hi.push_back(work);
        }
        if (helpers == 1) cv.notify_one();
        else cv.notify_all();
    }
    work();
    std::unique_lock<std::mutex> lk(st->mu);
    st->cv.wait(lk, [&] { return st->done.load() == n; });
}

void TaskPool::spawn(Group& g, std::function<void()> fn)
{
    {
        std::lock_guard<std::mutex> lk(g.mu);
        ++g.pending;
    }
    Group* gp = &g;
    auto task = [gp, fn = std::move(fn)] {
        fn();
        std::lock_guard<std::mutex> lk(gp->mu);
        if (--gp->pending == 0) gp->cv.notify_all();
    };
    {
        std::lock_guard<std::mutex> lk(mu);
        lo.push_back(std::move(task));
    }
    cv.notify_one();
}

void TaskPool::wait(Group& g)
{
    std::unique_lock<std::mutex> lk(g.mu);
    g.cv.wait(lk, [&] { return g.pending == 0; });
}

}  // namespace frame
}  // namespace tp

namespace {

// ----------------------------------------------------------- stream worker

// Per-worker device and host resources (grow-only, reused across frames).
struct Worker {
    hipStream_t st = nullptr;
    MappedSlot slot;
    DevBuf raw_in, rows_in, rows_out, vr[2], vi[2], br[2], bi[2];
    DevBuf z1cr, z1ci, z1lr, z1li, othr, othi, partials, small;
    DevBuf tzr, tzi;   // the claim terms' coefficients
    uint8_t* h_in = nullptr;
    size_t h_in_bytes = 0;
    u64* h_out = nullptr;
    size_t h_out_bytes = 0;

    bool init()
    {
        if (hipSetDevice(0) != hipSuccess) return fail("hipSetDevice(0) failed");
        TP_HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return slot.ready() && gkr_lds_limits();
    }
    bool host_in(size_t n)
    {
        if (n <= h_in_bytes) return true;
        free_host(h_in);
        h_in = nullptr;
        TP_HIP_OK(hipHostMalloc((void**)&h_in, n));
        h_in_bytes = n;
        return true;
    }
    bool host_out(size_t n)
    {
        if (n <= h_out_bytes) return true;
        free_host(h_out);
        h_out = nullptr;
        TP_HIP_OK(hipHostMalloc((void**)&h_out, n));
        h_out_bytes = n;
        return true;
    }
    void release()
    {
        DevBuf* all[] = {&raw_in, &rows_in, &rows_out, &vr[0], &vr[1], &vi[0], &vi[1], &br[0],
                         &br[1], &bi[0], &bi[1], &z1cr, &z1ci, &z1lr, &z1li, &othr, &othi,
                         &partials, &small, &tzr, &tzi};
        for (DevBuf* b : all) b->release();
        free_host(h_in);
        free_host(h_out);
        h_in = nullptr;
        h_out = nullptr;
        slot.release();
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
    }
};

// One frame on the worker's stream.
bool prove_frame(Worker& W, Job& J)
{
    Shape S;
    if (!S.set(J)) return false;
    const Circuit& C = *J.circ;
    const int nin = S.nin, nout = S.nout, ncb = S.ncb, cp = S.cp, NW = S.NW;
    hipStream_t st = W.st;
    auto t_mark = Clock::now();
    auto lap = [&](int k) {
        const auto t = Clock::now();
        J.phase[k] = std::chrono::duration<double, std::milli>(t - t_mark).count();
        t_mark = t;
    };

    // 1. inputs -> rows, evaluate the layer, outputs to the host (A3, A4)
    const size_t in_bytes = S.in_bytes, rows_n = S.rows_n, out_n = S.out_n;
    if (!W.host_in(in_bytes ? in_bytes : 8) || !W.host_out(out_n * 8)) return false;
    if (!W.raw_in.ensure(in_bytes ? in_bytes : 8) || !W.rows_in.ensure(rows_n * 8) ||
        !W.rows_out.ensure(out_n * 8))
        return false;
    // inputs submitted on the device are read where they are
    const void* raw = J.dev_in.ptr ? J.dev_in.ptr : W.raw_in.ptr;
    if (!J.dev_in.ptr && in_bytes) {
        memcpy(W.h_in, J.in_bytes == 1 ? (const void*)J.in8.data() : (const void*)J.in64.data(), in_bytes);
        TP_HIP_OK(hipMemcpyAsync(W.raw_in.ptr, W.h_in, in_bytes, hipMemcpyHostToDevice, st));
    }
    k_fe_rows<<<grid_for(rows_n, kBlock, 8192), kBlock, 0, st>>>(
        J.in_bytes == 1 ? static_cast<const uint8_t*>(raw) : nullptr,
        J.in_bytes == 1 ? nullptr : static_cast<const u64*>(raw), J.n_priv, J.n_copies, cp, nin,
        C.d_slot.as<int>(), C.d_cval.as<u64>(), W.rows_in.as<u64>());
    TP_HIP_OK(hipGetLastError());
    kern::k_eval_layer<<<grid_for(out_n, kBlock, 8192), kBlock, 0, st>>>(
        W.rows_in.as<u64>(), NW, W.rows_out.as<u64>(), nout, J.n_copies, cp, C.d_op.as<int>(),
        C.d_in0.as<int>(), C.d_in1.as<int>(), C.d_mux.as<int>(), C.n_gates);
    TP_HIP_OK(hipGetLastError());
    // wire table [wire][copy] for the early rounds, while the outputs travel
    const size_t tab_n = (size_t)NW * cp;
    if (!W.vr[0].ensure(tab_n * 8) || !W.vi[0].ensure(tab_n * 8) || !W.vr[1].ensure(tab_n * 4) ||
        !W.vi[1].ensure(tab_n * 4) || !W.br[0].ensure((size_t)cp * 8) ||
        !W.bi[0].ensure((size_t)cp * 8) || !W.br[1].ensure((size_t)cp * 4 + 8) ||
        !W.bi[1].ensure((size_t)cp * 4 + 8))
        return false;
    {
        dim3 grid((unsigned)((cp + 63) / 64), (unsigned)((NW + 63) / 64));
        kern::k_wire_table<<<grid, 256, 0, st>>>(W.rows_in.as<u64>(), NW, J.n_copies, cp, NW,
                                                W.vr[0].as<u64>(), W.vi[0].as<u64>());
        TP_HIP_OK(hipGetLastError());
    }
    TP_HIP_OK(hipMemcpyAsync(W.h_out, W.rows_out.ptr, out_n * 8, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    lap(0);

    // 2. transcript head.  The proof's leading items (q, io) are known now, so
    // their bytes are written on a helper thread while this one hashes the
    // outputs and runs the rounds.
    FrameProto pr(J, S);
    pr.prepare(W.h_out);
    Joiner prefix(std::thread([&pr] { pr.write_prefix(); }));
    pr.head();
    lap(1);

    // 3. eq tables of z2 (beta over copies) and z1 (gate coefficients) (A5)
    PointArg z1a, z2a;
    memset(&z1a, 0, sizeof(z1a));
    memset(&z2a, 0, sizeof(z2a));
    for (int i = 0; i < nout; ++i) z1a.z[i] = pr.z1[i];
    for (int i = 0; i < ncb; ++i) z2a.z[i] = pr.z2[i];
    if (!W.z1cr.ensure((size_t)C.n_gates * 8) || !W.z1ci.ensure((size_t)C.n_gates * 8) ||
        !W.z1lr.ensure((size_t)C.n_gates * 8) || !W.z1li.ensure((size_t)C.n_gates * 8) ||
        !W.othr.ensure((size_t)C.n_gates * 8) || !W.othi.ensure((size_t)C.n_gates * 8) ||
        !W.small.ensure(64 * sizeof(fp2)))
        return false;
    k_fe_eq<<<(cp + kBlock - 1) / kBlock, kBlock, 0, st>>>(z2a, ncb, W.br[0].as<u64>(),
                                                          W.bi[0].as<u64>(), cp);
    k_fe_eq<<<(C.n_gates + kBlock - 1) / kBlock, kBlock, 0, st>>>(z1a, nout, W.z1cr.as<u64>(),
                                                                 W.z1ci.as<u64>(), C.n_gates);
    TP_HIP_OK(hipGetLastError());
    if (!W.tzr.ensure((size_t)C.n_terms * 8) || !W.tzi.ensure((size_t)C.n_terms * 8)) return false;
    kern::k_term_coeffs<<<(C.n_terms * 64 + kBlock - 1) / kBlock, kBlock, 0, st>>>(
        W.z1cr.as<u64>(), W.z1ci.as<u64>(), C.d_toff.as<int>(), C.d_tsrc.as<int>(),
        C.d_tneg.as<int>(), C.n_terms, W.tzr.as<u64>(), W.tzi.as<u64>());
    TP_HIP_OK(hipGetLastError());
    lap(2);

    // 4. the rounds (layer loop of circuitnizkvec_vole.py:318-372, one layer)
    const fp2* host = static_cast<const fp2*>(W.slot.vals);
    int cur = 0, nc = cp;
    for (int rd = 0; rd < ncb; ++rd) {                           // early rounds (A6, A7)
        const ClaimShape cs = claim_shape(C.n_terms, nc >> 1, kClaimItems);
        const unsigned nblk = cs.blocks;
        if (!W.partials.ensure((size_t)nblk * 4 * sizeof(fp2))) return false;
        fp2* parts = W.partials.as<fp2>();
        kern::k_early_claim<<<nblk, kBlock, 0, st>>>(
            C.n_terms, cs.gpt, cs.n_chunks, C.d_ttype.as<int>(), C.d_tin0.as<int>(),
            C.d_tin1.as<int>(), W.tzr.as<u64>(), W.tzi.as<u64>(), nc, W.vr[cur].as<u64>(),
            W.vi[cur].as<u64>(), W.br[cur].as<u64>(), W.bi[cur].as<u64>(), parts);
        const unsigned seq = W.slot.advance();
        k_publish_sum<4><<<1, kBlock, 0, st>>>(parts, nblk, static_cast<fp2*>(W.slot.dvals),
                                                W.slot.dseq, seq);
        TP_HIP_OK(hipGetLastError());
        if (!W.slot.wait(st, seq, "early claim")) return false;
        fp2 f[4];
        for (int t = 0; t < 4; ++t) f[t] = host[t];
        fp2 r;
        if (!pr.early(rd, f, r)) return false;
        const int n_out = nc >> 1;
        const int log_out = clog2(n_out);
        const size_t total = ((size_t)NW + 1) * n_out;
        kern::k_fold_rows<<<grid_for(total), kBlock, 0, st>>>(
            W.vr[cur].as<u64>(), W.vi[cur].as<u64>(), NW, log_out, r.re, r.im,
            W.vr[cur ^ 1].as<u64>(), W.vi[cur ^ 1].as<u64>(), W.br[cur].as<u64>(),
            W.bi[cur].as<u64>(), W.br[cur ^ 1].as<u64>(), W.bi[cur ^ 1].as<u64>());
        TP_HIP_OK(hipGetLastError());
        cur ^= 1;
        nc = n_out;
    }
    lap(3);

    // late rounds over V = the folded wire values, [wire][1] (A8, A9)
    const u64* Vr = W.vr[cur].as<u64>();
    const u64* Vi = W.vi[cur].as<u64>();
    u64* beta2 = W.small.as<u64>();
    fp2* dev_vals = W.small.as<fp2>() + 4;
    k_fe_late_init<<<(C.n_gates + kBlock - 1) / kBlock, kBlock, 0, st>>>(
        Vr, Vi, C.d_in1.as<int>(), C.n_gates, W.z1cr.as<u64>(), W.z1ci.as<u64>(),
        W.z1lr.as<u64>(), W.z1li.as<u64>(), W.othr.as<u64>(), W.othi.as<u64>(),
        W.br[cur].as<u64>(), W.bi[cur].as<u64>(), beta2);
    TP_HIP_OK(hipGetLastError());
    ChalSet ch;
    memset(&ch, 0, sizeof(ch));
    const size_t lds = fold_lds(nin);
    {
        // the late rounds in one resident workgroup, the host answering each
        // round's challenge through the mapped answer slot (relay.hpp)
        volatile u64* hw = reinterpret_cast<volatile u64*>(W.slot.vals);
        for (int i = 0; i < 6; ++i) hw[i] = kSlotEmpty;
        W.slot.ans[0] = W.slot.ans[1] = kSlotEmpty;
        *W.slot.seq = 0;
        const unsigned seq0 = W.slot.next + 1;
        W.slot.next += 2 * nin;
        ResidentPhase resident;   // no frees from here to the last answer
        k_fe_late_resident<<<1, kOneWg, lds, st>>>(
            C.d_type.as<int>(), C.d_in0.as<int>(), C.d_in1.as<int>(), W.z1lr.as<u64>(),
            W.z1li.as<u64>(), W.othr.as<u64>(), W.othi.as<u64>(), Vr, Vi, nin, C.n_gates, beta2,
            static_cast<fp2*>(W.slot.dvals), W.slot.dseq, seq0, W.slot.dans, late_spin_ticks());
        TP_HIP_OK(hipGetLastError());
        auto answer = [&](int k, fp2 r) {
            if (k + 1 >= 2 * nin) return;   // the last round's sums end the kernel
            const u64 tag = (u64)((seq0 + (unsigned)k) & 7) << 61;
            W.slot.ans[0] = r.re | tag;
            W.slot.ans[1] = r.im | tag;
        };
        for (int k = 0; k < 2 * nin; ++k) {
            fp2 f[3];
            if (!wait_tagged(W.slot, seq0 + (unsigned)k, 3, f, "late round")) return false;
            fp2 r;
            if (!pr.late(k, f, r)) {
                // answer the kernel's remaining rounds with zeros so that it
                // ends now rather than at its wait bound, then report
                const std::string err = tp_last_error();
                answer(k, zero2());
                for (int j = k + 1; j < 2 * nin; ++j) {
                    if (!wait_tagged(W.slot, seq0 + (unsigned)j, 3, f, "late round")) break;
                    answer(j, zero2());
                }
                (void)hipStreamSynchronize(st);
                return fail("%s", err.c_str());
            }
            ch.c[k] = r;
            answer(k, r);
        }
    }
    lap(4);

    // 5. line values V(w1 + t (w2 - w1)) (A10) and the wiring predicate sums
    // (A11) in one publish: both need only z1, w1 and w2
    {
        const int K = nin + 1;
        kern::k_line_values<<<K, kOneWg, lds + (size_t)nin * sizeof(fp2), st>>>(Vr, Vi, nin, ch,
                                                                               dev_vals);
        const size_t wl = (((size_t)1 << nout) + ((size_t)2 << nin)) * sizeof(fp2);
        kern::k_wiring_mlext<<<1, kOneWg, wl, st>>>(C.d_type.as<int>(), C.d_in0.as<int>(),
                                                    C.d_in1.as<int>(), C.n_gates, nout, nin, 0,
                                                    pr.wiring_args(), dev_vals + 16);
        const unsigned seq = W.slot.advance();
        k_publish_copy<<<1, 64, 0, st>>>(dev_vals, 16 + 13, static_cast<fp2*>(W.slot.dvals),
                                         W.slot.dseq, seq);
        TP_HIP_OK(hipGetLastError());
        if (!W.slot.wait(st, seq, "line values and wiring predicate")) return false;
    }
    if (!pr.finish_layer(host, host + 16)) return false;
    lap(5);

    // 6. input MLE at the point (A12) and its value proof
    {
        const int lb = S.n_z < 10 ? S.n_z : 10;
        const size_t nrows = (size_t)1 << (S.n_z - lb);
        const unsigned blocks = (unsigned)(nrows < 1024 ? nrows : 1024);
        if (!W.partials.ensure((size_t)blocks * sizeof(fp2) * 4)) return false;
        kern::k_mle_base_dev<<<blocks, kBlock, 0, st>>>(W.rows_in.as<u64>(), pr.point, S.n_z,
                                                       W.partials.as<fp2>());
        const unsigned seq = W.slot.advance();
        k_publish_sum<1><<<1, kBlock, 0, st>>>(W.partials.as<fp2>(), blocks,
                                                static_cast<fp2*>(W.slot.dvals), W.slot.dseq, seq);
        TP_HIP_OK(hipGetLastError());
        if (!W.slot.wait(st, seq, "input MLE")) return false;
        pr.input_value(host[0]);
    }
    lap(6);

    // 7. proof bytes; (q, io) were written by the helper thread
    prefix.join();
    pr.write_rest();
    lap(7);
    return true;
}

// ------------------------------------------------------------------- pool

struct Pool {
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv_job, cv_done;
    std::deque<Job*> queue;
    std::unordered_map<int64_t, std::unique_ptr<Job>> jobs;
    int64_t next_id = 1;
    bool stop = false;
    Clock::time_point t0 = Clock::now();
    std::string init_err;
    // lockstep batch drivers
    bool batched = false;
    int max_batch = 1;
    size_t budget = 0, in_use = 0, peak = 0;
    uint64_t batches = 0, batched_frames = 0;
    double prof[BP_COUNT] = {};
    double tuning[9] = {};
    long long fix_claim = 0;   // > 0: the schedule is fixed by the caller
    int fix_fold = 0;
    std::unique_ptr<TaskPool> tasks;
};

void worker_main(Pool* pool)
{
    Worker W;
    const bool ok = device_ready() && W.init();
    if (!ok) {
        std::lock_guard<std::mutex> lk(pool->mu);
        pool->init_err = std::string("worker init failed: ") + tp_last_error();
    }
    for (;;) {
        Job* J = nullptr;
        {
            std::unique_lock<std::mutex> lk(pool->mu);
            pool->cv_job.wait(lk, [&] { return pool->stop || !pool->queue.empty(); });
            if (pool->queue.empty()) break;
            J = pool->queue.front();
            pool->queue.pop_front();
        }
        J->t_start = now_s(pool->t0);
        J->batch = 1;
        J->state.store(1);
        clear_error();
        bool done = ok && prove_frame(W, *J);
        if (!done) {
            // drain the stream so no kernel of this frame still runs
            if (W.st) (void)hipStreamSynchronize(W.st);
        }
        J->t_end = now_s(pool->t0);
        {
            std::lock_guard<std::mutex> lk(pool->mu);
            if (!done) J->err = ok ? std::string(tp_last_error()) : pool->init_err;
            J->state.store(done ? 2 : 3);
        }
        pool->cv_done.notify_all();
    }
    W.release();
}

// A lockstep batch driver: admits queued frames in arrival order while the
// batch has room and their table bytes fit the pool's budget beside the other
// drivers' batches, then proves them together.
void batch_main(Pool* pool)
{
    BatchCtx* B = device_ready() ? batch_ctx_create() : nullptr;
    if (!B) {
        std::lock_guard<std::mutex> lk(pool->mu);
        pool->init_err = std::string("batch driver init failed: ") + tp_last_error();
    }
    for (;;) {
        std::vector<Job*> batch;
        size_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(pool->mu);
            pool->cv_job.wait(lk, [&] {
                if (pool->stop || pool->queue.empty()) return pool->stop;
                const Job* h = pool->queue.front();
                return h->footprint > pool->budget || pool->in_use + h->footprint <= pool->budget;
            });
            if (pool->queue.empty()) break;
            while (!pool->queue.empty() && (int)batch.size() < pool->max_batch) {
                Job* J = pool->queue.front();
                if (J->footprint > pool->budget) {
                    if (!batch.empty()) break;   // fail it when it reaches the head alone
                    pool->queue.pop_front();
                    char msg[160];
                    snprintf(msg, sizeof(msg),
                             "frame needs %zu device bytes, over the pool's budget of %zu",
                             J->footprint, pool->budget);
                    J->err = msg;
                    J->t_start = J->t_end = now_s(pool->t0);
                    J->state.store(3);
                    pool->cv_done.notify_all();
                    continue;
                }
                if (pool->in_use + bytes + J->footprint > pool->budget) break;
                bytes += J->footprint;
                batch.push_back(J);
                pool->queue.pop_front();
            }
            pool->in_use += bytes;
            if (pool->in_use > pool->peak) pool->peak = pool->in_use;
            if (!batch.empty()) {
                ++pool->batches;
                pool->batched_frames += batch.size();
            }
        }
        if (batch.empty()) continue;
        const double t = now_s(pool->t0);
        for (Job* J : batch) {
            J->t_start = t;
            J->batch = (int)batch.size();
            J->state.store(1);
        }
        clear_error();
        if (B) {
            std::lock_guard<std::mutex> lk(pool->mu);
            if (pool->fix_claim > 0) batch_fix_knobs(*B, pool->fix_claim, pool->fix_fold);
        }
        double prof[BP_COUNT] = {};
        const bool dev_ok = B && prove_batch(*B, batch, *pool->tasks, prof);
        const std::string dev_err = dev_ok ? std::string() : (B ? std::string(tp_last_error())
                                                                : pool->init_err);
        const double t1 = now_s(pool->t0);
        {
            std::lock_guard<std::mutex> lk(pool->mu);
            pool->in_use -= bytes;
            for (int k = 0; k < BP_COUNT; ++k) pool->prof[k] += prof[k];
            if (B) batch_tuning(*B, pool->tuning);
            for (Job* J : batch) {
                J->t_end = t1;
                if (!dev_ok && J->err.empty()) J->err = dev_err;
                J->state.store(J->err.empty() ? 2 : 3);
            }
        }
        pool->cv_done.notify_all();
        pool->cv_job.notify_all();
    }
    if (B) batch_ctx_destroy(B);
}

Job* find(Pool* p, int64_t id)
{
    auto it = p->jobs.find(id);
    return it == p->jobs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

// A one-layer circuit for the frame prover: per gate the evaluation op and mux
// bit (as tp_eval_layer takes them), the sumcheck type index, in0 / in1; the
// input layer of n_inputs wires, input_slot[w] = index of the private input
// feeding wire w or -1, input_const[w] the constant of a non-private wire.
void* tp_circuit_create(int n_gates, const int* op, const int* muxbit, const int* type_idx,
                        const int* in0, const int* in1, int n_inputs, const int* input_slot,
                        const uint64_t* input_const)
{
    clear_error();
    if (!device_ready()) return nullptr;
    const int in_bits = clog2(n_inputs), out_bits = clog2(n_gates);
    if (n_gates < 1 || n_inputs < 1 || in_bits > kMaxNin || out_bits > kMaxNin || in_bits < 1) {
        fail("tp_circuit_create: bad shape (n_gates=%d n_inputs=%d)", n_gates, n_inputs);
        return nullptr;
    }
    int n_priv = 0;
    for (int w = 0; w < n_inputs; ++w)
        if (input_slot[w] >= 0) n_priv = input_slot[w] + 1 > n_priv ? input_slot[w] + 1 : n_priv;
    for (int g = 0; g < n_gates; ++g)
        if (in0[g] < 0 || in0[g] >= n_inputs || in1[g] < 0 || in1[g] >= n_inputs ||
            type_idx[g] < 0 || type_idx[g] > 12 || op[g] < 0 || op[g] > 12) {
            fail("tp_circuit_create: gate %d out of range", g);
            return nullptr;
        }
    auto* C = new Circuit();
    C->n_gates = n_gates;
    C->in_bits = in_bits;
    C->out_bits = out_bits;
    C->n_inputs = n_inputs;
    C->n_priv = n_priv;
    C->type.assign(type_idx, type_idx + n_gates);
    C->in0.assign(in0, in0 + n_gates);
    C->in1.assign(in1, in1 + n_gates);
    const int NW = 1 << in_bits;
    std::vector<int> slot(NW, -1);
    std::vector<u64> cval(NW, 0);
    for (int w = 0; w < n_inputs; ++w) {
        slot[w] = input_slot[w];
        cval[w] = input_slot[w] >= 0 ? 0 : input_const[w] % P;
    }
    hipStream_t st = stream();
    auto up = [&](DevBuf& b, const void* src, size_t n) {
        return b.ensure(n) && hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    const size_t gb = (size_t)n_gates * 4;
    const ClaimTerms T = build_claim_terms(n_gates, type_idx, in0, in1);
    C->n_terms = T.n_terms();
    const size_t tb = (size_t)C->n_terms * 4;
    bool ok = up(C->d_op, op, gb) && up(C->d_mux, muxbit, gb) && up(C->d_type, type_idx, gb) &&
              up(C->d_in0, in0, gb) && up(C->d_in1, in1, gb) &&
              up(C->d_slot, slot.data(), (size_t)NW * 4) && up(C->d_cval, cval.data(), (size_t)NW * 8) &&
              up(C->d_ttype, T.type.data(), tb) && up(C->d_tin0, T.in0.data(), tb) &&
              up(C->d_tin1, T.in1.data(), tb) && up(C->d_toff, T.off.data(), tb + 4) &&
              up(C->d_tsrc, T.src.data(), T.src.size() * 4) &&
              up(C->d_tneg, T.neg.data(), T.neg.size() * 4) &&
              hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        if (!tp_last_error_set()) fail("tp_circuit_create: upload failed");
        delete C;
        return nullptr;
    }
    return C;
}

// Drops the creator's reference: the circuit is freed now, or by the last
// queued or finished job that still holds it (tp_pool_release, tp_pool_destroy).
void tp_circuit_destroy(void* c) { circuit_unref(static_cast<const Circuit*>(c)); }

void* tp_pool_create(int n_workers)
{
    clear_error();
    if (!device_ready()) return nullptr;
    if (n_workers < 1 || n_workers > 64) {
        fail("tp_pool_create: n_workers=%d outside [1, 64]", n_workers);
        return nullptr;
    }
    auto* p = new Pool();
    p->budget = SIZE_MAX;
    for (int i = 0; i < n_workers; ++i) p->threads.emplace_back(worker_main, p);
    return p;
}

// A pool of n_drivers lockstep batch drivers sharing host_threads transcript
// threads.  Each driver takes up to max_batch queued frames (in arrival
// order) whose table bytes fit budget_bytes beside the other drivers'
// batches (0: three quarters of the device memory free now) and proves them
// with one launch per protocol step across the batch.
void* tp_pool_create_batched(int n_drivers, int max_batch, uint64_t budget_bytes, int host_threads)
{
    clear_error();
    if (!device_ready()) return nullptr;
    if (n_drivers < 1 || n_drivers > 8 || max_batch < 1 || max_batch > 1024 || host_threads < 1 ||
        host_threads > 256) {
        fail("tp_pool_create_batched: drivers=%d max_batch=%d host_threads=%d out of range",
             n_drivers, max_batch, host_threads);
        return nullptr;
    }
    size_t budget = (size_t)budget_bytes;
    if (!budget) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            fail("tp_pool_create_batched: hipMemGetInfo failed");
            return nullptr;
        }
        budget = free_b / 4 * 3;
    }
    auto* p = new Pool();
    p->batched = true;
    p->max_batch = max_batch;
    p->budget = budget;
    p->tasks.reset(new TaskPool(host_threads));
    for (int i = 0; i < n_drivers; ++i) p->threads.emplace_back(batch_main, p);
    return p;
}

void tp_pool_destroy(void* pool)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        p->stop = true;
    }
    p->cv_job.notify_all();
    for (auto& t : p->threads) t.join();
    delete p;
}

// Seconds since the pool was created (the clock of the job times).
double tp_pool_now(void* pool)
{
    auto* p = static_cast<Pool*>(pool);
    return p ? now_s(p->t0) : 0.0;
}

}  // extern "C"

namespace tp {
namespace frame {

void circuit_ref(const Circuit* c)
{
    if (c) const_cast<Circuit*>(c)->refs.fetch_add(1, std::memory_order_relaxed);
}

void circuit_unref(const Circuit* c)
{
    if (!c) return;
    auto* C = const_cast<Circuit*>(c);
    if (C->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    DevBuf* all[] = {&C->d_op,    &C->d_mux,   &C->d_type,  &C->d_in0,  &C->d_in1,
                     &C->d_slot,  &C->d_cval,  &C->d_ttype, &C->d_tin0, &C->d_tin1,
                     &C->d_toff,  &C->d_tsrc,  &C->d_tneg};
    for (DevBuf* b : all) b->release();
    delete C;
}

}  // namespace frame
}  // namespace tp

namespace {

// A job from its description (everything it reads is copied, the circuit is
// referenced); null with the error recorded on a bad shape.
std::unique_ptr<Job> make_job(const tp_frame_job* d)
{
    if (!d || !d->circuit) return fail("tp_pool_submit: null argument"), nullptr;
    const auto* C = static_cast<const Circuit*>(d->circuit);
    if (d->n_copies < 2 || d->n_private != C->n_priv || (d->input_bytes != 1 && d->input_bytes != 8))
        return fail("tp_pool_submit: bad frame shape (copies=%d private=%d bytes=%d)", d->n_copies,
                    d->n_private, d->input_bytes), nullptr;
    if (d->n_vole < 0 || d->n_binding < 0) return fail("tp_pool_submit: bad counts"), nullptr;
    auto J = std::make_unique<Job>();
    circuit_ref(C);
    J->circ = C;
    J->n_copies = d->n_copies;
    J->n_priv = d->n_private;
    J->in_bytes = d->input_bytes;
    Shape S;
    if (!S.set(*J)) return nullptr;
    J->footprint = S.footprint();
    const size_t n = (size_t)d->n_copies * d->n_private;
    if (d->inputs_on_device) {
        // device to device into the job's own buffer, so the caller may free
        // or reuse its memory as soon as submit returns
        const size_t bytes = n * (size_t)d->input_bytes;
        if (!d->inputs && bytes) return fail("tp_pool_submit: null device inputs"), nullptr;
        if (!J->dev_in.ensure(bytes ? bytes : 8)) return nullptr;
        if (bytes) {
            hipStream_t st = stream();
            if (hipMemcpyAsync(J->dev_in.ptr, d->inputs, bytes, hipMemcpyDeviceToDevice, st) !=
                    hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return fail("tp_pool_submit: copy of the device inputs failed"), nullptr;
        }
    } else if (d->input_bytes == 1) {
        const auto* s = static_cast<const uint8_t*>(d->inputs);
        J->in8.assign(s, s + n);
    } else {
        const auto* s = static_cast<const u64*>(d->inputs);
        J->in64.assign(s, s + n);
    }
    J->pid = d->pid;
    J->ur.assign(d->vole_u_re, d->vole_u_re + d->n_vole);
    J->ui.assign(d->vole_u_im, d->vole_u_im + d->n_vole);
    J->vr.assign(d->vole_v_re, d->vole_v_re + d->n_vole);
    J->vi.assign(d->vole_v_im, d->vole_v_im + d->n_vole);
    for (int k = 0; k < d->n_binding; ++k) {
        J->bkind.push_back(d->binding_kind[k]);
        if (d->binding_kind[k] == 1) J->bstr.emplace_back(d->binding_str[k]);
        else J->bint.push_back(d->binding_int[k]);
    }
    J->keep_outputs = d->keep_outputs != 0;
    return J;
}

// Queues jobs in order under one lock (a batch driver sees them together).
void enqueue(Pool* p, std::vector<std::unique_ptr<Job>>& js, int64_t* ids)
{
    {
        std::lock_guard<std::mutex> lk(p->mu);
        const double t = now_s(p->t0);
        for (size_t k = 0; k < js.size(); ++k) {
            Job* raw = js[k].get();
            raw->id = p->next_id++;
            raw->t_submit = t;
            ids[k] = raw->id;
            p->queue.push_back(raw);
            p->jobs.emplace(raw->id, std::move(js[k]));
        }
    }
    if (p->batched || js.size() > 1) p->cv_job.notify_all();
    else p->cv_job.notify_one();
}

}  // namespace

extern "C" {

// Queues one frame; the job copies everything it reads.  Returns the job id
// (> 0) or -1.
int64_t tp_pool_submit(void* pool, const tp_frame_job* d)
{
    clear_error();
    auto* p = static_cast<Pool*>(pool);
    if (!p) return fail("tp_pool_submit: null pool"), -1;
    std::vector<std::unique_ptr<Job>> js;
    js.push_back(make_job(d));
    if (!js[0]) return -1;
    int64_t id = -1;
    enqueue(p, js, &id);
    return id;
}

// Queues n frames at once, in order (all or none): ids[k] receives the job id
// of jobs[k].  0 / -1.
int tp_pool_submit_many(void* pool, const tp_frame_job* jobs, int n, int64_t* ids)
{
    clear_error();
    auto* p = static_cast<Pool*>(pool);
    if (!p || n < 0 || (n && (!jobs || !ids))) return fail("tp_pool_submit_many: bad arguments"), -1;
    std::vector<std::unique_ptr<Job>> js;
    js.reserve(n);
    for (int k = 0; k < n; ++k) {
        js.push_back(make_job(&jobs[k]));
        if (!js.back()) return -1;
    }
    enqueue(p, js, ids);
    return 0;
}

// 1 = done, 0 = still running after timeout_s (< 0: wait forever), -1 = failed
// or unknown job.
int tp_pool_wait(void* pool, int64_t job, double timeout_s)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::unique_lock<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (!J) return -1;
    auto finished = [&] { return J->state.load() >= 2; };
    if (timeout_s < 0) p->cv_done.wait(lk, finished);
    else if (!p->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) return 0;
    return J->state.load() == 2 ? 1 : -1;
}

// Message of a failed job ("" otherwise).
const char* tp_pool_error(void* pool, int64_t job)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return "null pool";
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    return J ? J->err.c_str() : "unknown job";
}

// Proof bytes of a finished job: *len = size; copies when cap suffices.
// 0 / -1.
int tp_pool_proof(void* pool, int64_t job, uint8_t* buf, size_t cap, size_t* len)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (!J || J->state.load() != 2) return -1;
    if (len) *len = J->proof.size();
    if (buf && cap >= J->proof.size()) memcpy(buf, J->proof.data(), J->proof.size());
    return 0;
}

// Proof bytes of a finished job in place: valid until the job is released.
// 0 / -1.
int tp_pool_proof_view(void* pool, int64_t job, const uint8_t** data, size_t* len)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p || !data || !len) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (!J || J->state.load() != 2) return -1;
    *data = J->proof.data();
    *len = J->proof.size();
    return 0;
}

// Final input point (2 limbs per variable, n_point variables written), input
// MLE value, times (submit, start, end in pool seconds) and phase times (ms).
int tp_pool_info(void* pool, int64_t job, uint64_t* point, int point_cap, int* n_point,
                 uint64_t value[2], double times[3], double phases[8])
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (!J || J->state.load() < 2) return -1;
    const int np = (int)J->point.size() / 2;
    if (n_point) *n_point = np;
    if (point)
        for (int k = 0; k < np && k < point_cap; ++k) {
            point[2 * k] = J->point[2 * k];
            point[2 * k + 1] = J->point[2 * k + 1];
        }
    if (value) {
        value[0] = J->value.re;
        value[1] = J->value.im;
    }
    if (times) {
        times[0] = J->t_submit;
        times[1] = J->t_start;
        times[2] = J->t_end;
    }
    if (phases)
        for (int k = 0; k < 8; ++k) phases[k] = J->phase[k];
    return 0;
}

// Size of the launch set that proved a finished job (1 for a stream worker)
// and its table bytes.  0 / -1.
int tp_pool_job_stats(void* pool, int64_t job, int* batch, uint64_t* footprint)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (!J || J->state.load() < 2) return -1;
    if (batch) *batch = J->batch;
    if (footprint) *footprint = J->footprint;
    return 0;
}

// Device-memory admission of a batched pool: bytes admitted now, the budget,
// the most admitted at once; batches run and frames they held.  0 / -1.
int tp_pool_memory(void* pool, uint64_t* in_use, uint64_t* budget, uint64_t* peak,
                   uint64_t* batches, uint64_t* frames)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    if (in_use) *in_use = p->in_use;
    if (budget) *budget = p->budget;
    if (peak) *peak = p->peak;
    if (batches) *batches = p->batches;
    if (frames) *frames = p->batched_frames;
    return 0;
}

// Where a batched pool's batches spent their time, accumulated since the
// pool was created (batches, frames, then milliseconds in the order of
// tp::frame::BatchProf); writes min(n, count) values, returns count.
int tp_pool_batch_profile(void* pool, double* out, int n)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    for (int k = 0; k < n && k < BP_COUNT; ++k) out[k] = p->prof[k];
    return BP_COUNT;
}

// The lockstep schedule tuner of a batched pool (the driver that ran the
// last batch): phase (0 trying claim sizes, 1 trying fold sizes, 2 settled),
// the claim items per frame and fold workgroups per frame in use, the claim
// costs (ns per term) and fold costs (ns per byte) of the three candidates
// each.  Writes min(n, 9) values; returns 9, -1 on a null pool.
int tp_pool_tuning(void* pool, double* out, int n)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    for (int k = 0; k < n && k < 9; ++k) out[k] = p->tuning[k];
    return 9;
}

// Fixes the lockstep schedule of a batched pool from the next batch on
// (claim items per frame in [1024, 2^18], fold workgroups per frame in
// [16, 65536]); the tuner stops.  0 / -1.
int tp_pool_fix_schedule(void* pool, long long claim_items, int fold_cap)
{
    clear_error();
    auto* p = static_cast<Pool*>(pool);
    if (!p || !p->batched) return fail("tp_pool_fix_schedule: not a batched pool"), -1;
    if (claim_items < 1024 || claim_items > (1 << 18) || fold_cap < 16 || fold_cap > 65536)
        return fail("tp_pool_fix_schedule: claim_items=%lld fold_cap=%d out of range",
                    claim_items, fold_cap), -1;
    std::lock_guard<std::mutex> lk(p->mu);
    p->fix_claim = claim_items;
    p->fix_fold = fold_cap;
    return 0;
}

// Circuit outputs [cp][2^out_bits] of a finished job submitted with
// keep_outputs; *n = count, copied when cap suffices.  0 / -1.
int tp_pool_outputs(void* pool, int64_t job, uint64_t* out, size_t cap, size_t* n)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return -1;
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (!J || J->state.load() != 2) return -1;
    if (n) *n = J->outputs.size();
    if (out && cap >= J->outputs.size()) memcpy(out, J->outputs.data(), J->outputs.size() * 8);
    return 0;
}

// Forgets a finished job (a queued or running job is kept).
void tp_pool_release(void* pool, int64_t job)
{
    auto* p = static_cast<Pool*>(pool);
    if (!p) return;
    std::lock_guard<std::mutex> lk(p->mu);
    Job* J = find(p, job);
    if (J && J->state.load() >= 2) p->jobs.erase(job);
}

}  // extern "C"
