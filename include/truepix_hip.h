/*
 * truepix_hip.h - C ABI of libtruepix_hip.so, the MI355X (gfx950) GKR prover
 * library for TruePix's fft_gkr / native_fp2 path.
 *
 * Conventions (same as the reference ABI):
 *   - Fp2 values over p = 2^61 - 1 with i^2 = -1, every limb canonical [0, p).
 *   - The native_fp2_* symbols take SoA arrays (re[], im[]) in host memory,
 *     are synchronous, and serve single-threaded callers.  Wire tables are
 *     [wire][copy] row-major; variables bind LSB-first (pairs 2i, 2i+1).
 *   - tp_* symbols are this library's own; device pointers are marked d_.
 *   - No host fallback: a device failure makes void entry points leave their
 *     outputs unwritten and record a message for tp_last_error(); int entry
 *     points return a negative value.
 */
#ifndef TRUEPIX_HIP_H
#define TRUEPIX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- drop-in: TP/native_fp2/native_fp2.cu:824-1022 (CUDA ABI) and
 *      TP/native_fp2/native_fp2.cpp:11-160 (CPU ABI), bound by
 *      TP/libTruePix/native_fp2_bridge.py:46-113 -------------------------- */

/* native_fp2.cu:826-829: init, idempotent; 0 on success (here: -1 without a device). */
int native_fp2_init(void);

/* native_fp2.cu:831-834: 1 when the device path is active. */
int native_fp2_cuda_available(void);

/* native_fp2.cu:836-848 / core.hpp:53-64: out[i] = in[2i] + r (in[2i+1] - in[2i]). */
void native_fp2_fold(const uint64_t* in_real, const uint64_t* in_imag, size_t n_out,
                     uint64_t r_real, uint64_t r_imag, uint64_t* out_real, uint64_t* out_imag);

/* native_fp2.cu:850-862 / core.hpp:66-80: out[i*ncopies + j] = in[i]. */
void native_fp2_expand(const uint64_t* in_real, const uint64_t* in_imag, size_t n_src,
                       size_t ncopies, uint64_t* out_real, uint64_t* out_imag);

/* native_fp2.cu:864-893 / core.hpp:82-137: early-round claim with explicit
 * evaluation tables at -1 (fact0) and 2 (fact1); out8 = F(0), F(1), F(-1), F(2). */
void native_fp2_sumcheck_early(int n_gates, const int* gate_type, const int* in0, const int* in1,
                               const uint64_t* z1_real, const uint64_t* z1_imag, int n_copies,
                               int n_wires, const uint64_t* values_real,
                               const uint64_t* values_imag, const uint64_t* fact0_real,
                               const uint64_t* fact0_imag, const uint64_t* fact1_real,
                               const uint64_t* fact1_imag, const uint64_t* beta_real,
                               const uint64_t* beta_imag, const uint64_t* beta_f0_real,
                               const uint64_t* beta_f0_imag, const uint64_t* beta_f1_real,
                               const uint64_t* beta_f1_imag, uint64_t out8[8]);

/* native_fp2.cu:895-907 / core.hpp:139-159: beta[i] = init prod_j (bit_j(i) ? z_j : 1 - z_j). */
void native_fp2_compute_beta(const uint64_t* z_real, const uint64_t* z_imag, int n_z,
                             uint64_t init_real, uint64_t init_imag, uint64_t* out_real,
                             uint64_t* out_imag);

/* native_fp2.cu:909-921 / core.hpp:300-364: MLE of 2^n_z base-field inputs at z. */
void native_fp2_mle_eval_base(const uint64_t* inputs, const uint64_t* z_real,
                              const uint64_t* z_imag, int n_z, uint64_t* out_real,
                              uint64_t* out_imag);

/* native_fp2.cu:923-952 / core.hpp:173-198: resident early-round session;
 * NULL on a bad shape (n_gates <= 0, n_wires <= 0, n_copies < 2 or not a
 * power of two) or on device failure. */
void* native_fp2_early_create(int n_gates, const int* gate_type, const int* in0, const int* in1,
                              const uint64_t* z1_real, const uint64_t* z1_imag, int n_copies,
                              int n_wires, const uint64_t* values_real,
                              const uint64_t* values_imag, const uint64_t* beta_real,
                              const uint64_t* beta_imag);

/* native_fp2.cu:954-963 / core.hpp:200-248: claim at t = 0, 1, -1, 2; 0 or -1. */
int native_fp2_early_claim(void* handle, uint64_t out8[8]);

/* native_fp2.cu:965-974 / core.hpp:250-279: fold every wire row and beta with r. */
int native_fp2_early_fold(void* handle, uint64_t r_real, uint64_t r_imag);

/* native_fp2.cu:976-989 / core.hpp:281-296: element 0 of each wire row and of beta. */
int native_fp2_early_finish(void* handle, uint64_t* wire_real, uint64_t* wire_imag,
                            uint64_t* beta_real, uint64_t* beta_imag);

/* native_fp2.cu:991-1005: current copy count / wire count (0 for NULL). */
int native_fp2_early_n_copies(void* handle);
int native_fp2_early_n_wires(void* handle);

/* native_fp2.cu:1007-1020: frees the session (NULL is a no-op). */
void native_fp2_early_destroy(void* handle);

/* ---- library status (no reference counterpart) -------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* tp_last_error(void);
void tp_clear_error(void);
/* Number of visible HIP devices (does not initialize the library). */
int tp_device_count(void);

/* ---- degree-2 sumcheck: OZ/src/VPD/fft_circuit_GKR.cpp:171-209
 *      (sumcheck_phase1_update; the addV term is zero) --------------------- */

/* Prover over two resident device tables of 2^nv Fp2 values (interleaved re,
 * im; read only).  Challenges come from the TruePix transcript
 * (TP/libTruePix/fiatshamir.py) after put(["c2", seed]); round k absorbs
 * put([a, b, c]) and draws r_k = rand_ext("c2-<k>"), unless `fixed` (2*nv
 * limbs) is given.  Outputs: polys (nv x 6 limbs, (a, b, c) of
 * s_k(X) = a X^2 + b X + c as in the reference's quadratic_poly), chals
 * (nv x 2), final4 = f~(r), g~(r).  stream: hipStream_t or NULL.  0 / -1.
 * One prove at a time per process: the mapped answer slot, the launches'
 * tagged sum slots and the work tables are the library's own, shared by
 * every call. */
int tp_sc2_prove(const uint64_t* d_f, const uint64_t* d_g, int nv, uint64_t seed,
                 const uint64_t* fixed, uint64_t* polys, uint64_t* chals, uint64_t* final4,
                 void* stream);
/* Per-launch HIP-event timing of the fused fold+round kernel (off by default). */
void tp_sc2_set_timing(int on);

/* 1: an untimed prove queues its launches down to the sliced mid launch at
 * once, each taking its challenges from the mapped answer slot, so only the
 * host's transcript steps sit between rounds; 0 (default, measured as fast):
 * each launch is made after the transcript step before it (timed proves
 * always are).  The proof bytes do not depend on it. */
void tp_sc2_set_ahead(int on);
/* Summed event milliseconds, algorithmic bytes and launch count of the
 * streaming fused fold+round launches in the last tp_sc2_prove (the rounds
 * before the sliced ones; see tp_sc2_schedule). */
void tp_sc2_stats(double* fold_ms, double* fold_bytes, int* fold_launches);
/* Summed event milliseconds and count of the one-workgroup tail rounds
 * (fold + round sum + publish in one launch) of the last tp_sc2_prove. */
void tp_sc2_tail_stats(double* tail_ms, int* tail_launches);
/* Bound of the sliced mid launch's arrival gate in wall-clock ticks (default
 * 20 ms; negative restores it): the launch needs all its workgroups resident
 * at once and otherwise gives way to streaming launches for those rounds.  0
 * makes it give way unless every workgroup has already arrived (tests). */
void tp_sc2_set_gate_ticks(long long ticks);
/* Bound of every device-side wait of a prove in wall-clock ticks (default
 * 10 s; negative restores it).  A small bound makes a prove fail with every
 * kernel leaving on its own (tests of the error path); the next prove starts
 * from fresh slots. */
void tp_sc2_set_spin_ticks(long long ticks);
/* Mid launches of the last tp_sc2_prove that gave way at the gate. */
int tp_sc2_mid_fallbacks(void);
/* The prove schedule's sizes: workgroups of the sliced mid rounds, the
 * largest input (entries per table) of the one-workgroup tail, and the input
 * the mid rounds hand to the tail.  Streaming launches run while the input
 * exceeds mid_groups * tail_in; the mid rounds run from there while it
 * exceeds tail_in, down to mid_end entries. */
void tp_sc2_schedule(int* mid_groups, int* tail_in, int* mid_end);
/* Per-class statistics of the last tp_sc2_prove (timing on): for class c in
 * 0 = the first pass over the input (rounds 0 and 1), 1 = round 2 folding
 * the input by r_0 and r_1 (with round 3's sums too on inputs above 2^22
 * entries), 2 = the streaming fold+round launches, 3 = the one-workgroup
 * tail rounds, 4 = the rounds on per-workgroup slices, 5 = the later passes
 * folding by two challenges: out[3c] summed event ms, out[3c+1] bytes the
 * launches read and write, out[3c+2] launch count.  Fills the classes that
 * fit in n (all six for n >= 18). */
void tp_sc2_class_stats(double* out, int n);
/* 1 (default): proves of 2^14 to 2^28 entries per table take rounds 0 to 3
 * from the 256 block products of the input (one pass on the matrix cores)
 * and fold the input into T_4 in one pass; 0: rounds 0 and 1 from the quad
 * products, then passes folding by two challenges.  The same proof either
 * way.  No counterpart in the reference (a schedule switch of this
 * library's prover, for tests and measurement). */
void tp_sc2_set_block_pass(int on);
/* Per-exchange latency of the one-workgroup tail of the last timed
 * tp_sc2_prove (an exchange: one host round trip for one or two rounds):
 * out[0] exchanges, out[1] device us per exchange until its sums are ready,
 * out[2] us the workgroup waited for the host's challenges, out[3] host us
 * per exchange (transcript), out[4 + i] device us of exchange i (while
 * 4 + i < n).  n >= 4. */
void tp_sc2_tail_profile(double* out, int n);
/* Per-exchange latency of the sliced mid rounds of the last timed
 * tp_sc2_prove (an exchange: one host round trip for one or two rounds):
 * out[0] exchanges, out[1] us from the last workgroup's partials to workgroup
 * 0 holding the challenges, out[2] us of the relay to the last workgroup,
 * out[3] us from the challenges to the next partials, out[4] us between the
 * first and the last workgroup's partials.  n >= 5. */
void tp_sc2_mid_profile(double* out, int n);

/* ---- Reed-Solomon encoding: OZ/src/infrastructure/RS_polynomial.cpp:46-130
 *      (fast_fourier_transform, with Orion's root of unity,
 *      OZ/src/linear_gkr/prime_field.cpp:60-73) ------------------------------ */

/* n_cols columns of 2^log_coef coefficients (device, interleaved re/im, column
 * k at k*2^log_coef) -> 2^log_order natural-order evaluations per column
 * (column k at k*2^log_order): out[k][i] = sum_j coef[k][j] w^(i j).
 * cols_per_chunk: columns per pass pair (0 = 2).  log_coef <= 20,
 * log_coef <= log_order <= 30.  stream: hipStream_t or NULL.  0 / -1.
 * One set of twiddle tables and one intermediate serve every call: calls are
 * serialized on the host, and a call on another stream than the previous
 * call's waits for that stream before it enqueues its work. */
int tp_rs_encode(const uint64_t* d_coef, int n_cols, int log_coef, int log_order,
                 uint64_t* d_out, int cols_per_chunk, void* stream);

/* ---- device-resident GKR layer (the engine under the NIZK prover of
 *      TP/libTruePix/circuitnizkvec_vole.py:265-501) --------------------------- */

/* A3, TP/libTruePix/arithcircuit.py:20-145: one layer over copies.  d_in
 * [n_copies_pad][in_stride], d_out [n_copies_pad][2^out_log] (u64 base field,
 * zero padding).  op: 0 MUL 1 ADD 2 SUB 3 MUX 5 OR 6 XOR 7 NOT 8 NAND 9 NOR
 * 10 NXOR 11 NAAB 12 PASS; muxbit: each gate's mux bit value.  0 / -1. */
int tp_eval_layer(const uint64_t* d_in, int in_stride, uint64_t* d_out, int out_log,
                  int n_copies, int n_copies_pad, const int* op, const int* in0, const int* in1,
                  const int* muxbit, int n_gates, void* stream);

/* A4, TP/libTruePix/layerprover.py:113-120 + layercompute.py:37-56: SoA Fp2
 * table [n_wires][n_copies_pad] (im = 0, padded copies zero) from rows. */
int tp_wire_table(const uint64_t* d_rows, int row_stride, int n_copies, int n_copies_pad,
                  int n_wires, uint64_t* d_re, uint64_t* d_im, void* stream);

/* A6-A8 on device tables: native_fp2_early_create (native_fp2.cu:923-952) with
 * the [wire][copy] SoA tables already in device memory. */
void* tp_early_create_dev(int n_gates, const int* gate_type, const int* in0, const int* in1,
                          const uint64_t* z1_real, const uint64_t* z1_imag, int n_copies,
                          int n_wires, const uint64_t* d_values_real,
                          const uint64_t* d_values_imag, const uint64_t* beta_real,
                          const uint64_t* beta_imag);

/* A9, TP/libTruePix/gateprover.py:102-176 + layerprover.py:225-244,282-285:
 * late-phase session over the 2^nin folded wire values (host SoA), gate type
 * indices, inputs, wiring coefficients z1 and beta*.  NULL on failure. */
void* tp_late_create(int n_gates, const int* type_idx, const int* in0, const int* in1,
                     const uint64_t* z1_re, const uint64_t* z1_im, int nin,
                     const uint64_t* v_re, const uint64_t* v_im, uint64_t beta_re,
                     uint64_t beta_im);
/* Round `round` (0 .. 2 nin - 1, in order); r_prev = previous challenge (2
 * limbs, NULL for round 0); out6 = (f(0), f(1), f(-1)) x beta*.  0 / -1. */
int tp_late_round(void* handle, int round, const uint64_t* r_prev, uint64_t out6[6]);
/* A10, TP/libTruePix/layercompute.py:169-243: V at K points (K x nin x 2). */
int tp_late_values(void* handle, const uint64_t* pts, int K, uint64_t* out);
void tp_late_destroy(void* handle);

/* A11, TP/libTruePix/circuitverifier.py:224-290 (eval_mlext): per gate type,
 * sum_g eq(z1, g) eq(w1, in0_g) eq(w2, in1_g) (with use_z1_2: mul0 eq(z1, g) +
 * mul1 eq(z1_2, g)); out26 = 13 Fp2 sums.  Points are interleaved limbs. */
int tp_wiring_mlext(int n_gates, const int* type_idx, const int* in0, const int* in1,
                    const uint64_t* z1, const uint64_t* z1_2, int n_out_bits, const uint64_t* w1,
                    const uint64_t* w2, int n_in_bits, const uint64_t* mul0, const uint64_t* mul1,
                    int use_z1_2, uint64_t out26[26]);

/* A12, native_fp2_mle_eval_base (native_fp2.cu:909-921) on a device-resident
 * base-field vector of 2^n_z entries; z: n_z x 2 limbs; out2 = the value. */
int tp_mle_eval_base_dev(const uint64_t* d_inputs, int n_z, const uint64_t* z, uint64_t out2[2],
                         void* stream);

/* No counterpart in the reference: tp_mle_eval_base_dev for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result; its
 * partial sums and result live in the library's static device scratch
 * (64 KiB per device, of which a call uses 256 + 16 * min(1024, 2^(n_z-10))
 * bytes), so no call allocates or frees anything.  n_z in [1, 40]; larger
 * inputs fail with an error.  Calls that overlap on one device share the
 * scratch and race: serialize them.  0 / -1. */
int tp_mle_eval_base_dev_noalloc(const uint64_t* d_inputs, int n_z, const uint64_t* z,
                                 uint64_t out2[2], void* stream);

/* ---- whole frames: CircuitProverNIZK.run (TP/libTruePix/circuitnizkvec_vole.py:265-501)
 *      for a one-layer circuit, as run_truepix_prover.run_truepix calls it
 *      (TP/run_truepix_prover.py:218-423), device-resident, batched and served --- */

/* A one-layer circuit (the TP/pws files via TP/libTruePix/parse_pws.py:8-64): per
 * gate the evaluation op and mux bit (as tp_eval_layer), the sumcheck type
 * index, in0 / in1; input_slot[w] (w < n_inputs) = the private input feeding
 * input wire w, or -1 for a constant wire whose value is input_const[w].
 * NULL on a bad shape.  Circuits are shared by any number of jobs: each
 * queued job holds a reference, so tp_circuit_destroy while jobs still use the
 * circuit drops only the creator's reference and the last of those jobs frees
 * it (when tp_pool_release or tp_pool_destroy forgets the job). */
void* tp_circuit_create(int n_gates, const int* op, const int* muxbit, const int* type_idx,
                        const int* in0, const int* in1, int n_inputs, const int* input_slot,
                        const uint64_t* input_const);
void tp_circuit_destroy(void* circuit);

/* One frame to prove.  Inputs [n_copies][n_private] (uint8 or uint64);
 * pid is the first transcript value (os.getpid() in the reference);
 * the prover's VOLE correlations U, V (TP/libTruePix/ExtVOLEGen.py:130-187,
 * one security repetition) and the commitment binding list
 * (TP/libTruePix/fp2_link.py:487-504) as typed items (kind 0 int, 1 str). */
typedef struct tp_frame_job {
    void* circuit;
    int n_copies;
    int n_private;
    int input_bytes;                /* 1 or 8 */
    const void* inputs;
    uint64_t pid;
    int n_vole;
    const uint64_t* vole_u_re;
    const uint64_t* vole_u_im;
    const uint64_t* vole_v_re;
    const uint64_t* vole_v_im;
    int n_binding;
    const int* binding_kind;
    const uint64_t* binding_int;    /* read where kind == 0 */
    const char* const* binding_str; /* read where kind == 1 */
    int keep_outputs;               /* keep the circuit outputs for tp_pool_outputs */
    int inputs_on_device;           /* 1: inputs is device memory (HIP), e.g. rows
                                     * that tp_yuv_pack built; complete when the job
                                     * is submitted, and copied device to device */
} tp_frame_job;

/* A pool of n_workers provers (host threads, each with its own HIP stream and
 * buffers) taking frames from one FIFO queue.  NULL on failure.  The
 * reference runs frames as separate prover processes (TP/TruePix.sh --threads
 * with TP/run_truepix_prover.py); a pool is that fan-out inside one process. */
void* tp_pool_create(int n_workers);
/* A pool of n_drivers lockstep batch drivers: each takes up to max_batch queued
 * frames in arrival order whose device table bytes fit budget_bytes beside the
 * other drivers' batches (0 = three quarters of the device memory free now),
 * and proves them together with one launch per protocol step across the batch
 * (per-frame descriptor and item tables), the frames' transcript work spread
 * over host_threads threads.  Frames arriving while a batch runs form the
 * next one.  Same job API and the same proof bytes as tp_pool_create.  No
 * reference counterpart beyond the process fan-out named above. */
void* tp_pool_create_batched(int n_drivers, int max_batch, uint64_t budget_bytes,
                             int host_threads);
void tp_pool_destroy(void* pool);
/* Seconds since the pool was created (the clock of tp_pool_info's times). */
double tp_pool_now(void* pool);
/* Queues a frame (everything it reads is copied); job id > 0, or -1. */
int64_t tp_pool_submit(void* pool, const tp_frame_job* job);
/* Queues n frames at once and in order, all or none (a batch driver admits
 * them together); ids[k] = the job id of jobs[k].  0 / -1. */
int tp_pool_submit_many(void* pool, const tp_frame_job* jobs, int n, int64_t* ids);
/* 1 done, 0 not done within timeout_s (< 0 waits forever), -1 failed / unknown. */
int tp_pool_wait(void* pool, int64_t job, double timeout_s);
/* Message of a failed job. */
const char* tp_pool_error(void* pool, int64_t job);
/* Proof bytes (pickle of the transcript, fiatshamir.py:234-237): *len = size,
 * copied to buf when cap suffices.  0 / -1. */
int tp_pool_proof(void* pool, int64_t job, uint8_t* buf, size_t cap, size_t* len);
/* The same bytes in place (no copy), valid until tp_pool_release.  0 / -1. */
int tp_pool_proof_view(void* pool, int64_t job, const uint8_t** data, size_t* len);
/* Final GKR input point (nInBits + nCopyBits variables, 2 limbs each), the
 * input MLE there, times (submit, start, end) and phase times in ms
 * (eval+download, output absorb, tables, early, late, tail, input MLE, proof
 * bytes).  0 / -1. */
int tp_pool_info(void* pool, int64_t job, uint64_t* point, int point_cap, int* n_point,
                 uint64_t value[2], double times[3], double phases[8]);
/* Frames in the launch set that proved a finished job (1 for a stream worker)
 * and the job's device table bytes (what admission counts).  0 / -1. */
int tp_pool_job_stats(void* pool, int64_t job, int* batch, uint64_t* footprint);
/* Admission state of a batched pool: bytes admitted now, the budget, the most
 * admitted at once, batches run and the frames they held.  0 / -1. */
int tp_pool_memory(void* pool, uint64_t* in_use, uint64_t* budget, uint64_t* peak,
                   uint64_t* batches, uint64_t* frames);
/* Where a batched pool's batches spent their time since it was created: batches,
 * frames, then ms: input staging, launch queueing, transcript heads (wall),
 * output waits (sum over frames), heads (sum), proof prefixes (sum), rounds
 * (wall), step waits, step transcript work (wall), proof ends (wall), device
 * time of the early-claim and early-fold launches.  Writes min(n, count)
 * values and returns the count (14), -1 on a null pool. */
int tp_pool_batch_profile(void* pool, double* out, int n);
/* The lockstep schedule autotuner of a batched pool: phase (0 timing the
 * early-claim sizes, 1 timing the fold sizes, 2 settled), the claim items per
 * frame and fold workgroups per frame in use, then the three claim costs (ns
 * per pair-gate term) and three fold costs (ns per algorithmic byte) measured
 * with events on real batches.  Writes min(n, 9) values; returns 9. */
int tp_pool_tuning(void* pool, double* out, int n);
/* Fixes the lockstep schedule (claim items per frame in [1024, 2^18], fold
 * workgroups per frame in [16, 65536]) from the next batch on; the tuner
 * stops.  Every schedule gives the same proof bytes.  0 / -1. */
int tp_pool_fix_schedule(void* pool, long long claim_items, int fold_cap);
/* Circuit outputs [copies_pad][2^out_bits] of a job submitted with keep_outputs. */
int tp_pool_outputs(void* pool, int64_t job, uint64_t* out, size_t cap, size_t* n);
/* Forgets a finished job. */
void tp_pool_release(void* pool, int64_t job);

/* ---- the signer's Orion commitment of a frame: commit_pair
 *      (OZ/src/VPD/linearPC.cpp:253-459) as OZ/examples/multivariate_PC_test_commit.cpp:265-331
 *      runs it (layout and mask of OZ/include/truepix/truepix_gkr.h) ------------ */

/* The packing of a copy that every later commit and opening uses
 * (OZ/include/truepix/truepix_gkr.h:107-150, make_layout, where the
 * reference's programs read it from TRUEPIX_COPY_SIZE, TRUEPIX_CONST_IDX and
 * TRUEPIX_CONST_VAL): copy_size samples (1..512), the constant const_val at
 * wire const_idx (-1: none; outside the samples), zeros up to 512.  The
 * circuits use gray 384/384/128 (the default), invert 384/384/255, mask
 * 384/-1/0 and crop 385/-1/0 (TP/libTruePix/fp2_link.py:533-563).  0 / -1. */
int tp_orion_set_layout(int copy_size, int const_idx, int const_val);
/* inputs: [copies][copy_size] sample values (the lines of input.txt); seed: the
 * mask seed (orion_mask_seed.txt).  Writes the masked and mask Merkle roots
 * as orion_commit.json carries them (64 hex digits + NUL); ms (optional):
 * device time.  copies <= 65536.  0 / -1. */
int tp_orion_commit(int copies, const uint8_t* inputs, uint64_t seed, char masked_root[65],
                    char mask_root[65], double* ms);
/* OZ/src/VPD/linearPC.cpp:700-966 (tensor_product_protocol; r0, r1 formed
 * from the point r as open_and_verify does, :996-1034): the opening of the
 * committed masked polynomial (use_group1) or the mask at r (n_r = lg n
 * challenges, (re, im) each) at the code positions col_q[n_queries] the
 * transcript drew, serialized in the reference's OpeningProof version-2
 * layout (OZ/include/truepix/orion_proof.h:1-22) into *proof (free with
 * tp_free); the claimed value in answer.  The Virgo code-switching proof that
 * follows in the reference is not part of it.  0 / -1. */
int tp_orion_open(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                  const uint64_t* r, int n_r, const int64_t* col_q, int n_queries,
                  uint8_t** proof, size_t* proof_len, uint64_t answer[2]);
/* tp_orion_open with the code positions drawn as the reference's NIZK mode
 * draws them (OZ/src/truepix/orion_fs.cpp: "tensor-open", the opened root, N,
 * the group flag, r0 and r1, then 1222 positions from the opened group;
 * linearPC.cpp:711-758), followed, as in the reference (linearPC.cpp:875-928),
 * by the Virgo code-switching proof over the expander-encoding circuit: 1222
 * more positions from the whole codeword, "combined-mt" and the combined
 * root absorbed, then generate_circuit's circuit proved by the Virgo GKR
 * prover (OZ/src/linear_gkr/prover.cpp) and checked by its verifier's
 * per-layer checks (OZ/src/linear_gkr/verifier.cpp:1031-1211).  Its input
 * layer is checked by evaluating the combined message, which the proof
 * carries, instead of the reference's FRI-based VPD.  The proof is written
 * only if every check passes (the reference writes it only then): -1 and
 * tp_last_error() otherwise.  0 / -1. */
int tp_orion_open_nizk(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                       const uint64_t* r, int n_r, uint8_t** proof, size_t* proof_len,
                       uint64_t answer[2]);
/* tp_orion_open_nizk that also writes the code-switching transcript: up to
 * cap words into words, the full count into *n_words.  Layout (Fp2 as re, im;
 * D layers, bits(j) of layer j, layers i = D-1 .. 1):
 *   D, bits(0) .. bits(D-1), Q = 1222, the Q positions sorted, r_0 and r_1
 *   of the output layer (bits(D-1) each), the output claim;
 *   per layer i: r_u, r_v (bits(i-1) each; r_v zero at layer 1), the
 *   bits(i-1) phase-1 polynomials (a, b, c of a X^2 + b X + c), the bits(i-1)
 *   phase-2 polynomials, v_u, v_v, the relay, custom-combination and
 *   direct-relay predicate values, the next alpha and beta;
 *   the input-layer claim, the combined message's multilinear value at the
 *   last r_u, and the verdict (1 / 0).
 * The words are written when the proof does not check, too. */
int tp_orion_open_nizk_virgo(int copies, const uint8_t* inputs, uint64_t seed, int use_group1,
                             const uint64_t* r, int n_r, uint8_t** proof, size_t* proof_len,
                             uint64_t answer[2], uint64_t* words, long long cap,
                             long long* n_words);
/* Device time (ms), launches and sumcheck layers of the last code-switching
 * proof (no counterpart in the reference, which prints its prover time).
 * Returns 0. */
int tp_orion_virgo_stats(double* ms, int* launches, int* layers);
/* The expander code alone (encode, OZ/include/linear_code/linear_code_encode.h:35-110,
 * graphs from glibc's initial rand() state as expander_init draws them):
 * n_cols <= 256 columns of n <= 65536 Fp2, src [n_cols][n] and dst
 * [n_cols][cw] as (re, im) pairs.  Returns the codeword length cw, or -1. */
long long tp_orion_encode(long long n, int n_cols, const uint64_t* src, uint64_t* dst);

/* ---- fft_gkr: OZ/src/VPD/fft_circuit_GKR.cpp (main :827-845), the program
 *      vpd_verifier.cpp:80 runs as "./fft_gkr <lg> <logfile>" ------------------ */

/* A16 + A18: builds the circuit for lg_size and runs every layer protocol on the
 * device with the reference's challenges (glibc rand(), seed 1), then the
 * verifier's checks on the host.  Returns 1 when every check passes (the
 * reference prints "Error, fft gkr failed" otherwise), 0 on a failed check,
 * -1 on error.  v_time / p_time in seconds, proof_size in bytes as the
 * reference counts them (what it writes to its log file). */
int tp_fft_gkr(int lg_size, double* v_time, int* proof_size, double* p_time);
/* The same run, exposing what it computes: polys (cap_polys x (a, b, c) Fp2,
 * interleaved limbs) and *n_polys = (lg+6) + lg + 2 lg^2 round polynomials in
 * protocol order; claims (7 Fp2) = alpha_beta_sum after V_output, addition,
 * mult, intermediate and all IFFT layers, then the final alpha and beta;
 * layers (optional): last IFFT layer (2^lg Fp2), the 1/n-scaled layer (2^lg)
 * and the 64 sums.  Return as tp_fft_gkr. */
int tp_fft_gkr_trace(int lg_size, uint64_t* polys, int cap_polys, int* n_polys, uint64_t* claims,
                     uint64_t* layers, int* proof_size);
/* On (default): a run's launch sequence, which depends only on lg_size and
 * the buffer addresses, is captured once into a hipGraph and replayed; off:
 * every launch is enqueued directly. */
void tp_fft_gkr_set_graph(int on);

/* ---- the frame's macroblock layout: VideoEdition/Frame_Block_Extraction.py
 *      :67-94 (process_macroblocks) and :220-240 (one input.txt line per
 *      macroblock), VideoEdition/Crop.py:79-85 (the crop bit appended),
 *      VideoEdition/Encode.py:19-42 (reconstruct_frame).  The reference does
 *      these steps in numpy on the host and hands the prover and the signer a
 *      text file; these build the same rows on the device. */

/* YUV 4:2:0 planes (device) -> copies x copy_bytes rows (device): Y 16x16,
 * Cb 8x8, Cr 8x8 per macroblock in raster order, partial edge macroblocks
 * dropped; copy_bytes 385 appends extra[copy] (the crop bit), 384 takes no
 * extra.  Even width and height >= 16; strides >= plane widths.  Returns the
 * copies ((width/16)*(height/16)) or -1. */
long long tp_yuv_pack(const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int width,
                      int height, long long y_stride, long long c_stride, const uint8_t* extra,
                      int copy_bytes, uint8_t* out, void* stream);
/* copies x copy_stride rows (device, first 384 bytes used) -> planes
 * (device), zeros outside the macroblock area.  0 / -1. */
int tp_yuv_unpack(const uint8_t* mb, long long copy_stride, int width, int height,
                  long long y_stride, long long c_stride, uint8_t* y, uint8_t* cb, uint8_t* cr,
                  void* stream);
/* The same from and to host buffers (tightly packed planes), staged through
 * the device; synchronous. */
long long tp_yuv_pack_host(const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int width,
                           int height, const uint8_t* extra, int copy_bytes, uint8_t* out);
int tp_yuv_unpack_host(const uint8_t* mb, long long copy_stride, int width, int height, uint8_t* y,
                       uint8_t* cb, uint8_t* cr);

/* ---- host pieces of the frame prover, exported for their CPU tests ------- */

/* A15, TP/libTruePix/fiatshamir.py:234-237: pickle.dumps(tree, 5) of a
 * token-encoded tree ('i' u64 | 's' u32 len bytes | 'n' | 'f' u64 u64 |
 * 'l' u32 n items | 't' u32 n items | 'a' u32 n u64[n]); *out is freed with
 * tp_free.  0 / -1. */
int tp_pickle_tree(const uint8_t* tokens, size_t n, uint8_t** out, size_t* out_len);
void tp_free(void* p);
/* A14: SHA-256 of the transcript's host hash (SHA-NI when present), fed in
 * pieces of `chunk` bytes (0 = one piece). */
void tp_sha256(const uint8_t* data, size_t n, size_t chunk, uint8_t out[32]);
int tp_sha256_has_ni(void);

#ifdef __cplusplus
}
#endif

#endif /* TRUEPIX_HIP_H */
