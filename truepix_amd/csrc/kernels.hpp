// kernels.hpp - GKR layer kernels shared by the drop-in ABI (native_fp2.hip,
// gkr.hip) and the device-resident frame prover (frame.hip).  Each kernel is
// defined once, in the file named beside it, and launched from any of them.
//
// Challenges reach the one-workgroup kernels by value (kernel arguments), so
// a round costs one launch and no host-to-device copy.
#pragma once
#include "fp61.hpp"

namespace tp {

constexpr int kOneWg = 512;   // threads of the one-workgroup kernels
constexpr int kMaxNin = 11;   // 2^11 wires: two 32 KiB LDS fold buffers
constexpr int kMaxZ = 40;     // variables of an input-MLE point

// Late-round challenges: w1 (bits of in0) in c[0, nin), w2 (bits of in1) in
// c[nin, 2 nin).
struct ChalSet { fp2 c[2 * kMaxNin]; };
// Wiring-predicate MLE arguments: z1 (and z1b with the mul0 / mul1 weights
// when use_z1_2) over the output bits, w1 / w2 over the input bits.
struct WiringArgs { fp2 z1[kMaxNin], z1b[kMaxNin], w1[kMaxNin], w2[kMaxNin]; fp2 m0, m1; };
// A point of up to kMaxZ variables, LSB-first.
struct PointArg { fp2 z[kMaxZ]; };

namespace kern {

// gkr.hip ---------------------------------------------------------------------
// A3: out[c][g] = gate_g(in[c][in0_g], in[c][in1_g]); zero past n_gates and
// for padded copies.
__global__ void k_eval_layer(const u64* in, int in_stride, u64* out, int out_log, int n_copies,
                             int n_copies_pad, const int* op, const int* in0, const int* in1,
                             const int* mux, int n_gates);
// A4: Fp2 SoA wire table [wire][copy] from base-field rows [copy][row_stride].
__global__ void k_wire_table(const u64* rows, int row_stride, int n_copies, int n_copies_pad,
                             int n_wires, u64* tre, u64* tim);
// A9: late round rd; writes (f(0), f(1), f(-1)) * beta[0] to host_vals, then
// host_seq = seq (mapped host memory).
__global__ void k_late_round(const int* gt, const int* in0, const int* in1, u64* z1r, u64* z1i,
                             const u64* oth_r, const u64* oth_i, const u64* vr, const u64* vi,
                             ChalSet chal, int nin, int n_gates, int rd, int apply,
                             const u64* beta, fp2* host_vals, volatile unsigned* host_seq,
                             unsigned seq);
// A9: start of the w2 pass: every gate's fixed side becomes V(w1).
__global__ void k_late_switch(const u64* vr, const u64* vi, int nin, ChalSet chal, int n_gates,
                              u64* oth_r, u64* oth_i);
// A10: V (2^nb table) at the points pts[k] (k = blockIdx.x), nb x Fp2 each.
__global__ void k_mle_points(const u64* vr, const u64* vi, int nb, const fp2* pts, fp2* out);
// A10: V at the line points w1 + t (w2 - w1), t = blockIdx.x.
__global__ void k_line_values(const u64* vr, const u64* vi, int nb, ChalSet w12, fp2* out);
// A11: 13 per-type sums of the wiring predicate.
__global__ void k_wiring_mlext(const int* type_idx, const int* in0, const int* in1, int n_gates,
                               int n_out_bits, int n_in_bits, int use_z1_2, WiringArgs a,
                               fp2* out13);
// A12: partial sums of beta_z(x) * in[x] over a device base-field vector.
__global__ void k_mle_base_dev(const u64* in, PointArg z, int n_z, fp2* partials);

// native_fp2.hip --------------------------------------------------------------
// A6: early-round claim partials (4 accumulators per workgroup) over the
// layer's claim terms (claim_terms.hpp), 1-D grid of ClaimShape::blocks
// workgroups.
__global__ void k_early_claim(int n_gates, int gpt, int n_chunks, const int* gt, const int* in0,
                              const int* in1, const u64* z1r, const u64* z1i, int nc,
                              const u64* vr, const u64* vi, const u64* br, const u64* bi,
                              fp2* partials);
// Claim-term coefficients from z1 (one wave per term, body::term_coeffs).
__global__ void k_term_coeffs(const u64* z1r, const u64* z1i, const int* off, const int* src,
                              const int* neg, int n_terms, u64* tzr, u64* tzi);
// A7: fold every wire row and beta with r.
__global__ void k_fold_rows(const u64* ir, const u64* ii, int n_wires, int log_out, u64 rr, u64 ri,
                            u64* orr, u64* oi, const u64* bir, const u64* bii, u64* bor, u64* boi);

}  // namespace kern

// Dynamic LDS of a one-workgroup fold over a 2^nb table (gkr.hip), and the
// one-time raise of the dynamic LDS limit of the late-round kernels.
size_t fold_lds(int nb);
bool gkr_lds_limits();

// Gate-chunk size for the early claim so that pairs x chunks gives about 64K
// threads (native_fp2.hip).
int gates_per_thread(int n_gates, int n_pairs);

// Work of one early claim: gate chunks of gpt gates, n_chunks per pair, about
// `target` (pair, chunk) items in `blocks` workgroups of kBlock threads.
struct ClaimShape {
    int gpt, n_chunks;
    unsigned blocks;
};
ClaimShape claim_shape(int n_gates, int n_pairs, long long target);
constexpr long long kClaimItems = 1 << 16;   // default target

}  // namespace tp
