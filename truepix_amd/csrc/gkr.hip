// gkr.hip - device side of one GKR layer proof beyond the early rounds:
//   A3 layer evaluation     TP/libTruePix/arithcircuit.py:20-145 (C*Gate.run,
//                           ArithCircuitLayer.run): base-field gates over copies
//   A4 wire table           TP/libTruePix/layerprover.py:113-120 (transpose
//                           [copy][wire] -> [wire][copy], zero-padded copies)
//   A9 late rounds          TP/libTruePix/gateprover.py:102-176 and
//                           layerprover.py:225-244,282-285: the quadratic rounds
//                           over w1 then w2 with the per-gate wiring predicate
//   A10 line projection     TP/libTruePix/layercompute.py:169-243: V at points
//   A11 wiring MLE          TP/libTruePix/circuitverifier.py:224-290 (eval_mlext)
//   A12 input MLE           native_fp2_mle_eval_base on a device-resident vector
//
// Layouts: circuit values are u64 base-field rows [copy][row_stride] with
// row_stride = 2^bits (copy-major, as the proof's input vector is laid out);
// Fp2 tables are SoA (re[], im[]) like the early-round session.
#include "fp61.hpp"
#include "kernels.hpp"
#include "layer_bodies.hpp"
#include "reduce.hpp"
#include "runtime.hpp"

#include <stdint.h>
#include <string.h>

#include <mutex>
#include <vector>

using namespace tp;

namespace tp {
namespace kern {

// A3: one layer over copies (body::eval_layer).
__global__ void k_eval_layer(const u64* __restrict__ in, int in_stride, u64* __restrict__ out,
                             int out_log, int n_copies, int n_copies_pad,
                             const int* __restrict__ op, const int* __restrict__ in0,
                             const int* __restrict__ in1, const int* __restrict__ mux, int n_gates)
{
    body::eval_layer(in, in_stride, out, out_log, n_copies, n_copies_pad, op, in0, in1, mux,
                     n_gates, blockIdx.x, gridDim.x);
}

// A4: [wire][copy] table, one 64 x 64 tile per workgroup (body::wire_table).
__global__ void k_wire_table(const u64* __restrict__ rows, int row_stride, int n_copies,
                             int n_copies_pad, int n_wires, u64* __restrict__ tre,
                             u64* __restrict__ tim)
{
    body::wire_table(rows, row_stride, n_copies, n_copies_pad, n_wires, tre, tim, blockIdx.x,
                     blockIdx.y);
}

// ---------------------------------------------------------------- late rounds
// One workgroup per round (body::late_round); the challenges are staged from
// the kernel argument into LDS, the result goes to mapped host memory.
__global__ void __launch_bounds__(kOneWg) k_late_round(
    const int* __restrict__ gt, const int* __restrict__ in0, const int* __restrict__ in1,
    u64* __restrict__ z1r, u64* __restrict__ z1i, const u64* __restrict__ oth_r,
    const u64* __restrict__ oth_i, const u64* __restrict__ vr, const u64* __restrict__ vi,
    ChalSet chal, int nin, int n_gates, int rd, int apply, const u64* __restrict__ beta,
    fp2* host_vals, volatile unsigned* host_seq, unsigned seq)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 cs[2 * kMaxNin];
    for (int j = threadIdx.x; j < 2 * nin; j += blockDim.x) cs[j] = chal.c[j];
    __syncthreads();
    fp2 out[3];
    body::late_round(gt, in0, in1, z1r, z1i, oth_r, oth_i, vr, vi, cs, nin, n_gates, rd, apply,
                     beta, lds, out);
    if (threadIdx.x == 0) {
        for (int t = 0; t < 3; ++t) host_vals[t] = out[t];
        __threadfence_system();
        *host_seq = seq;
    }
}

// Start of the w2 pass (body::late_switch).
__global__ void __launch_bounds__(kOneWg) k_late_switch(const u64* __restrict__ vr,
                                                        const u64* __restrict__ vi, int nin,
                                                        ChalSet chal, int n_gates,
                                                        u64* __restrict__ oth_r,
                                                        u64* __restrict__ oth_i)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 cs[kMaxNin];
    for (int j = threadIdx.x; j < nin; j += blockDim.x) cs[j] = chal.c[j];
    __syncthreads();
    body::late_switch(vr, vi, nin, cs, n_gates, oth_r, oth_i, lds);
}

// A10: MLE of a 2^nb Fp2 table at K points (one workgroup per point).
__global__ void __launch_bounds__(kOneWg) k_mle_points(const u64* __restrict__ vr,
                                                       const u64* __restrict__ vi, int nb,
                                                       const fp2* __restrict__ pts,
                                                       fp2* __restrict__ out)
{
    extern __shared__ fp2 lds[];
    const fp2* R = lds_fold_chain(lds, lds + (1 << nb), vr, vi, nb,
                                  pts + (size_t)blockIdx.x * nb, nb);
    if (threadIdx.x == 0) out[blockIdx.x] = R[0];
}

// A10 on the line of the late rounds: workgroup t evaluates V at
// w1 + t (w2 - w1) (body::line_value); dynamic LDS fold_lds(nb) + nb Fp2.
__global__ void __launch_bounds__(kOneWg) k_line_values(const u64* __restrict__ vr,
                                                        const u64* __restrict__ vi, int nb,
                                                        ChalSet w12, fp2* __restrict__ out)
{
    extern __shared__ fp2 lds[];
    __shared__ fp2 cs[2 * kMaxNin];
    for (int j = threadIdx.x; j < 2 * nb; j += blockDim.x) cs[j] = w12.c[j];
    __syncthreads();
    const fp2 v = body::line_value(vr, vi, nb, cs, (int)blockIdx.x, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// A11: 13 per-type sums of the wiring predicate (body::wiring_mlext), one
// workgroup; the arguments are staged into LDS past the eq tables.
__global__ void __launch_bounds__(kOneWg) k_wiring_mlext(
    const int* __restrict__ type_idx, const int* __restrict__ in0, const int* __restrict__ in1,
    int n_gates, int n_out_bits, int n_in_bits, int use_z1_2, WiringArgs a, fp2* __restrict__ out13)
{
    extern __shared__ fp2 sm[];
    __shared__ fp2 args[4 * kMaxNin];
    for (int j = threadIdx.x; j < kMaxNin; j += blockDim.x) {
        args[j] = a.z1[j];
        args[kMaxNin + j] = a.z1b[j];
        args[2 * kMaxNin + j] = a.w1[j];
        args[3 * kMaxNin + j] = a.w2[j];
    }
    __syncthreads();
    body::wiring_mlext(type_idx, in0, in1, n_gates, n_out_bits, n_in_bits, args,
                       use_z1_2 ? args + kMaxNin : nullptr, args + 2 * kMaxNin,
                       args + 3 * kMaxNin, a.m0, a.m1, sm, out13);
}

// A12 on a device-resident base-field vector of 2^n_z entries (body::mle_base).
__global__ void k_mle_base_dev(const u64* __restrict__ in, PointArg z, int n_z, fp2* partials)
{
    __shared__ fp2 zs[kMaxZ];
    for (int j = threadIdx.x; j < n_z; j += blockDim.x) zs[j] = z.z[j];
    __syncthreads();
    body::mle_base(in, zs, n_z, partials, blockIdx.x, gridDim.x, blockIdx.x);
}

}  // namespace kern

size_t fold_lds(int nb) { return ((size_t)3 << nb) / 2 * sizeof(fp2); }

bool gkr_lds_limits()
{
    static std::once_flag once;
    static bool ok = false;
    std::call_once(once, [] {
        const int need = (int)(((size_t)4 << kMaxNin) * sizeof(fp2) / 2);
        ok = hipFuncSetAttribute((const void*)kern::k_late_round,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, need) == hipSuccess &&
             hipFuncSetAttribute((const void*)kern::k_late_switch,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, need) == hipSuccess &&
             hipFuncSetAttribute((const void*)kern::k_mle_points,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, need) == hipSuccess &&
             hipFuncSetAttribute((const void*)kern::k_line_values,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 need + kMaxNin * (int)sizeof(fp2)) == hipSuccess &&
             hipFuncSetAttribute((const void*)kern::k_wiring_mlext,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 2 * need) == hipSuccess;
    });
    return ok ? true : fail("hipFuncSetAttribute (dynamic LDS) failed");
}

}  // namespace tp

namespace {

// ------------------------------------------------------------------ host side

MappedSlot g_map;   // results of the ABI's late rounds

bool h2d(DevBuf& b, const void* src, size_t bytes, hipStream_t st)
{
    if (!b.ensure(bytes ? bytes : 8)) return false;
    if (bytes) TP_HIP_OK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st));
    return true;
}

struct LateSession {
    int n_gates = 0, nin = 0, next_round = 0;
    ChalSet chal;                         // challenges received so far
    DevBuf gt, i0, i1, z1r, z1i, othr, othi, vr, vi, beta;
    ~LateSession()
    {
        DevBuf* all[] = {&gt, &i0, &i1, &z1r, &z1i, &othr, &othi, &vr, &vi, &beta};
        for (DevBuf* b : all) b->release();
    }
};

}  // namespace

extern "C" {

// A3: evaluate one layer.  d_in: [n_copies_pad][in_stride] u64 (canonical);
// d_out: [n_copies_pad][2^out_log] u64.  op: gate type index per gate
// (0 MUL, 1 ADD, 2 SUB, 3 MUX, 5 OR, 6 XOR, 7 NOT, 8 NAND, 9 NOR, 10 NXOR,
// 11 NAAB, 12 PASS); muxbit: the mux bit value of each MUX gate.
int tp_eval_layer(const uint64_t* d_in, int in_stride, uint64_t* d_out, int out_log,
                  int n_copies, int n_copies_pad, const int* op, const int* in0, const int* in1,
                  const int* muxbit, int n_gates, void* stream_handle)
{
    tp::clear_error();
    if (!tp::device_ready()) return -1;
    if (n_gates < 0 || out_log < 0 || out_log > 24 || n_gates > (1 << out_log) ||
        n_copies > n_copies_pad)
        return tp::fail("tp_eval_layer: bad shape"), -1;
    for (int g = 0; g < n_gates; ++g)
        if (in0[g] < 0 || in0[g] >= in_stride || in1[g] < 0 || in1[g] >= in_stride)
            return tp::fail("tp_eval_layer: gate %d reads outside the row", g), -1;
    hipStream_t st = stream_handle ? (hipStream_t)stream_handle : tp::stream();
    static DevBuf b_op, b_i0, b_i1, b_mx;
    const size_t gb = (size_t)n_gates * 4;
    if (!h2d(b_op, op, gb, st) || !h2d(b_i0, in0, gb, st) || !h2d(b_i1, in1, gb, st) ||
        !h2d(b_mx, muxbit, gb, st))
        return -1;
    const size_t total = (size_t)n_copies_pad << out_log;
    kern::k_eval_layer<<<grid_for(total, kBlock, 8192), kBlock, 0, st>>>(
        (const u64*)d_in, in_stride, (u64*)d_out, out_log, n_copies, n_copies_pad,
        b_op.as<int>(), b_i0.as<int>(), b_i1.as<int>(), b_mx.as<int>(), n_gates);
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipStreamSynchronize(st));   // gate arrays are reused by the next call
    return 0;
}

// A4: SoA Fp2 wire table [n_wires][n_copies_pad] from rows [n_copies][row_stride].
int tp_wire_table(const uint64_t* d_rows, int row_stride, int n_copies, int n_copies_pad,
                  int n_wires, uint64_t* d_re, uint64_t* d_im, void* stream_handle)
{
    tp::clear_error();
    if (!tp::device_ready()) return -1;
    if (n_wires > row_stride || n_copies > n_copies_pad || n_wires <= 0 || n_copies_pad <= 0)
        return tp::fail("tp_wire_table: bad shape"), -1;
    hipStream_t st = stream_handle ? (hipStream_t)stream_handle : tp::stream();
    dim3 grid((unsigned)((n_copies_pad + 63) / 64), (unsigned)((n_wires + 63) / 64));
    kern::k_wire_table<<<grid, 256, 0, st>>>((const u64*)d_rows, row_stride, n_copies, n_copies_pad,
                                       n_wires, (u64*)d_re, (u64*)d_im);
    TP_HIP_OK(hipGetLastError());
    return 0;
}

// A9: late-phase session over the 2^nin wire values V (host SoA, the early
// rounds' result), per-gate type index, inputs, wiring coefficient z1 and
// beta*.  NULL on a bad shape or device failure.
void* tp_late_create(int n_gates, const int* type_idx, const int* in0, const int* in1,
                     const uint64_t* z1_re, const uint64_t* z1_im, int nin,
                     const uint64_t* v_re, const uint64_t* v_im, uint64_t beta_re,
                     uint64_t beta_im)
{
    tp::clear_error();
    if (!tp::device_ready() || !g_map.ready() || !gkr_lds_limits()) return nullptr;
    if (n_gates <= 0 || nin < 1 || nin > kMaxNin) {
        tp::fail("tp_late_create: bad shape (n_gates=%d nin=%d)", n_gates, nin);
        return nullptr;
    }
    for (int g = 0; g < n_gates; ++g)
        if (in0[g] < 0 || in0[g] >= (1 << nin) || in1[g] < 0 || in1[g] >= (1 << nin)) {
            tp::fail("tp_late_create: gate %d input outside 2^nin", g);
            return nullptr;
        }
    hipStream_t st = stream();
    auto* s = new LateSession();
    s->n_gates = n_gates;
    s->nin = nin;
    const u64 beta2[2] = {beta_re, beta_im};
    const size_t W = (size_t)1 << nin;
    // fixed side during the w1 pass: V[in1] (gateprover.py:51-62 reset())
    std::vector<u64> oth_r(n_gates), oth_i(n_gates);
    for (int g = 0; g < n_gates; ++g) { oth_r[g] = v_re[in1[g]]; oth_i[g] = v_im[in1[g]]; }
    bool ok = h2d(s->gt, type_idx, n_gates * 4, st) && h2d(s->i0, in0, n_gates * 4, st) &&
              h2d(s->i1, in1, n_gates * 4, st) && h2d(s->z1r, z1_re, n_gates * 8, st) &&
              h2d(s->z1i, z1_im, n_gates * 8, st) && h2d(s->othr, oth_r.data(), n_gates * 8, st) &&
              h2d(s->othi, oth_i.data(), n_gates * 8, st) && h2d(s->vr, v_re, W * 8, st) &&
              h2d(s->vi, v_im, W * 8, st) && h2d(s->beta, beta2, 16, st);
    if (ok && hipStreamSynchronize(st) != hipSuccess) ok = tp::fail("tp_late_create: upload");
    if (!ok) {
        delete s;
        return nullptr;
    }
    return s;
}

// Output of late round `round` (0 .. 2 nin - 1, in order): out6 =
// (f(0), f(1), f(-1)) times beta*.  r_prev (2 limbs) is the previous round's
// challenge, required for round > 0.  0 / -1.
int tp_late_round(void* handle, int round, const uint64_t* r_prev, uint64_t out6[6])
{
    tp::clear_error();
    auto* s = static_cast<LateSession*>(handle);
    if (!s) return -1;
    if (round != s->next_round || round >= 2 * s->nin)
        return tp::fail("tp_late_round: round %d out of order", round), -1;
    if (round > 0 && !r_prev) return tp::fail("tp_late_round: missing challenge"), -1;
    hipStream_t st = stream();
    if (round > 0) s->chal.c[round - 1] = mk2(r_prev[0], r_prev[1]);
    const size_t lds = fold_lds(s->nin);
    if (round == s->nin) {
        kern::k_late_switch<<<1, kOneWg, lds, st>>>(s->vr.as<u64>(), s->vi.as<u64>(), s->nin,
                                                    s->chal, s->n_gates, s->othr.as<u64>(),
                                                    s->othi.as<u64>());
        TP_HIP_OK(hipGetLastError());
    }
    const unsigned seq = g_map.advance();
    kern::k_late_round<<<1, kOneWg, lds, st>>>(
        s->gt.as<int>(), s->i0.as<int>(), s->i1.as<int>(), s->z1r.as<u64>(), s->z1i.as<u64>(),
        s->othr.as<u64>(), s->othi.as<u64>(), s->vr.as<u64>(), s->vi.as<u64>(), s->chal, s->nin,
        s->n_gates, round, round > 0 ? 1 : 0, s->beta.as<u64>(), static_cast<fp2*>(g_map.dvals),
        g_map.dseq, seq);
    TP_HIP_OK(hipGetLastError());
    if (!g_map.wait(st, seq, "tp_late_round")) return -1;
    const fp2* vals = static_cast<const fp2*>(g_map.vals);
    for (int t = 0; t < 3; ++t) {
        out6[2 * t] = vals[t].re;
        out6[2 * t + 1] = vals[t].im;
    }
    s->next_round = round + 1;
    return 0;
}

// A10: MLE of the session's V table at K points (K x nin x 2 limbs) -> out (K x 2).
int tp_late_values(void* handle, const uint64_t* pts, int K, uint64_t* out)
{
    tp::clear_error();
    auto* s = static_cast<LateSession*>(handle);
    if (!s || K <= 0) return tp::fail("tp_late_values: bad arguments"), -1;
    hipStream_t st = stream();
    static DevBuf b_pts, b_out;
    if (!h2d(b_pts, pts, (size_t)K * s->nin * 16, st) || !b_out.ensure((size_t)K * 16)) return -1;
    kern::k_mle_points<<<K, kOneWg, fold_lds(s->nin), st>>>(s->vr.as<u64>(), s->vi.as<u64>(),
                                                             s->nin, b_pts.as<fp2>(), b_out.as<fp2>());
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipMemcpyAsync(out, b_out.ptr, (size_t)K * 16, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return 0;
}

void tp_late_destroy(void* handle) { delete static_cast<LateSession*>(handle); }

// A11: wiring-predicate MLE per gate type.  z1 (n_out_bits x 2 limbs),
// optional z1_2 with mul0 / mul1 (use_z1_2), w1 / w2 (n_in_bits x 2 limbs);
// out26 = 13 Fp2 sums indexed by gate type.  0 / -1.
int tp_wiring_mlext(int n_gates, const int* type_idx, const int* in0, const int* in1,
                    const uint64_t* z1, const uint64_t* z1_2, int n_out_bits, const uint64_t* w1,
                    const uint64_t* w2, int n_in_bits, const uint64_t* mul0, const uint64_t* mul1,
                    int use_z1_2, uint64_t out26[26])
{
    tp::clear_error();
    if (!tp::device_ready() || !gkr_lds_limits()) return -1;
    if (n_gates <= 0 || n_gates > (1 << n_out_bits) || n_out_bits > kMaxNin || n_out_bits < 0 ||
        n_in_bits > kMaxNin || n_in_bits < 0)
        return tp::fail("tp_wiring_mlext: bad shape"), -1;
    for (int g = 0; g < n_gates; ++g)
        if (in0[g] < 0 || in0[g] >= (1 << n_in_bits) || in1[g] < 0 || in1[g] >= (1 << n_in_bits) ||
            type_idx[g] < 0 || type_idx[g] > 12)
            return tp::fail("tp_wiring_mlext: gate %d out of range", g), -1;
    hipStream_t st = stream();
    static DevBuf b_t, b_i0, b_i1, b_out;
    WiringArgs a;
    memset(&a, 0, sizeof(a));
    for (int j = 0; j < n_out_bits; ++j) {
        a.z1[j] = mk2(z1[2 * j], z1[2 * j + 1]);
        if (use_z1_2) a.z1b[j] = mk2(z1_2[2 * j], z1_2[2 * j + 1]);
    }
    for (int j = 0; j < n_in_bits; ++j) {
        a.w1[j] = mk2(w1[2 * j], w1[2 * j + 1]);
        a.w2[j] = mk2(w2[2 * j], w2[2 * j + 1]);
    }
    if (use_z1_2) {
        a.m0 = mk2(mul0[0], mul0[1]);
        a.m1 = mk2(mul1[0], mul1[1]);
    }
    if (!h2d(b_t, type_idx, n_gates * 4, st) || !h2d(b_i0, in0, n_gates * 4, st) ||
        !h2d(b_i1, in1, n_gates * 4, st) || !b_out.ensure(13 * 16))
        return -1;
    const size_t lds = (((size_t)(use_z1_2 ? 2 : 1) << n_out_bits) + ((size_t)2 << n_in_bits)) * 16;
    kern::k_wiring_mlext<<<1, kOneWg, lds, st>>>(b_t.as<int>(), b_i0.as<int>(), b_i1.as<int>(),
                                                 n_gates, n_out_bits, n_in_bits, use_z1_2, a,
                                                 b_out.as<fp2>());
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipMemcpyAsync(out26, b_out.ptr, 13 * 16, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// A12: MLE of a device-resident base-field vector of 2^n_z entries at z
// (n_z x 2 limbs, LSB-first); out2 = the Fp2 value.  0 / -1.
int tp_mle_eval_base_dev(const uint64_t* d_inputs, int n_z, const uint64_t* z, uint64_t out2[2],
                         void* stream_handle)
{
    tp::clear_error();
    if (!tp::device_ready()) return -1;
    if (n_z < 1 || n_z > 40) return tp::fail("tp_mle_eval_base_dev: n_z=%d", n_z), -1;
    hipStream_t st = stream_handle ? (hipStream_t)stream_handle : tp::stream();
    static DevBuf b_part;
    PointArg pz;
    memset(&pz, 0, sizeof(pz));
    for (int j = 0; j < n_z; ++j) pz.z[j] = mk2(z[2 * j], z[2 * j + 1]);
    const int lb = n_z < 10 ? n_z : 10;
    const size_t rows = (size_t)1 << (n_z - lb);
    const unsigned blocks = (unsigned)(rows < 1024 ? rows : 1024);
    if (!b_part.ensure(256 + (size_t)blocks * 16)) return -1;
    fp2* result = b_part.as<fp2>();
    fp2* partials = reinterpret_cast<fp2*>(b_part.as<char>() + 256);
    kern::k_mle_base_dev<<<blocks, kBlock, 0, st>>>((const u64*)d_inputs, pz, n_z, partials);
    TP_HIP_OK(hipGetLastError());
    k_finalize_sum<1><<<1, kBlock, 0, st>>>(partials, blocks, result);
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipMemcpyAsync(out2, result, 16, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return 0;
}

// tp_mle_eval_base_dev with its partial sums and result in the library's
// static device scratch (tp::scratch_ptr) instead of a library allocation:
// no hipMalloc / hipFree on any call.  The same kernels; every word of the
// scratch it reads is written by the same call first, so nothing needs
// clearing.  Calls that overlap on one device share the scratch and race.
int tp_mle_eval_base_dev_noalloc(const uint64_t* d_inputs, int n_z, const uint64_t* z,
                                 uint64_t out2[2], void* stream_handle)
{
    tp::clear_error();
    if (n_z < 1 || n_z > 40) return tp::fail("tp_mle_eval_base_dev_noalloc: n_z=%d", n_z), -1;
    char* scr = static_cast<char*>(tp::scratch_ptr());
    if (!scr) return -1;
    hipStream_t st = stream_handle ? (hipStream_t)stream_handle : tp::stream();
    PointArg pz;
    memset(&pz, 0, sizeof(pz));
    for (int j = 0; j < n_z; ++j) pz.z[j] = mk2(z[2 * j], z[2 * j + 1]);
    const int lb = n_z < 10 ? n_z : 10;
    const size_t rows = (size_t)1 << (n_z - lb);
    const unsigned blocks = (unsigned)(rows < 1024 ? rows : 1024);
    const size_t need = 256 + (size_t)blocks * 16;
    if (need > tp::kScratchBytes)
        return tp::fail("tp_mle_eval_base_dev_noalloc: needs %zu scratch bytes, has %zu", need,
                        tp::kScratchBytes), -1;
    fp2* result = reinterpret_cast<fp2*>(scr);
    fp2* partials = reinterpret_cast<fp2*>(scr + 256);
    kern::k_mle_base_dev<<<blocks, kBlock, 0, st>>>((const u64*)d_inputs, pz, n_z, partials);
    TP_HIP_OK(hipGetLastError());
    k_finalize_sum<1><<<1, kBlock, 0, st>>>(partials, blocks, result);
    TP_HIP_OK(hipGetLastError());
    TP_HIP_OK(hipMemcpyAsync(out2, result, 16, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
