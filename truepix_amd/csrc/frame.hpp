// frame.hpp - host side of the whole-frame prover, shared by its two drivers:
// frame.hip proves one frame at a time on a worker's stream; batch.hip proves
// a batch of frames in lockstep, one launch per protocol step covering every
// frame of the batch.
//
// FrameProto is the transcript of one frame (CircuitProverNIZK.run,
// TP/libTruePix/circuitnizkvec_vole.py:265-501, over a one-layer circuit)
// split at the points where device results arrive.  Both drivers call the
// same steps in the same order, so a frame's proof bytes do not depend on the
// driver, the batch it was in or the frames beside it.
#pragma once
#include "claim_terms.hpp"
#include "fp61.hpp"
#include "host_sha.hpp"
#include "kernels.hpp"
#include "proof_pickle.hpp"
#include "runtime.hpp"
#include "transcript.hpp"

#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace tp {
namespace frame {

using Clock = std::chrono::steady_clock;

int clog2(long long v);
double now_s(Clock::time_point t0);

// A one-layer circuit resident on the device (gate arrays, input slots).
// Reference-counted: tp_circuit_create holds one reference and every queued
// job one more, so tp_circuit_destroy while jobs still use the circuit only
// drops the creator's reference and the last job frees it.
struct Circuit {
    int n_gates = 0, in_bits = 0, out_bits = 0, n_inputs = 0, n_priv = 0;
    std::vector<int> type, in0, in1;
    DevBuf d_op, d_mux, d_type, d_in0, d_in1, d_slot, d_cval;
    // the early claim's terms (claim_terms.hpp): type, wires, and the CSR of
    // their coefficients' z1 contributions
    int n_terms = 0;
    DevBuf d_ttype, d_tin0, d_tin1, d_toff, d_tsrc, d_tneg;
    std::atomic<int> refs{1};
};

// Takes / drops a reference; the last drop frees the device arrays.
void circuit_ref(const Circuit* c);
void circuit_unref(const Circuit* c);

// One frame to prove: inputs, VOLE correlations, commitment binding, pid; and
// its results.
struct Job {
    Job() = default;
    Job(const Job&) = delete;
    Job& operator=(const Job&) = delete;
    ~Job()
    {
        circuit_unref(circ);   // the job's reference on its circuit
        dev_in.release();
    }
    int64_t id = 0;
    const Circuit* circ = nullptr;     // referenced (circuit_ref) while the job lives
    int n_copies = 0, n_priv = 0, in_bytes = 1;
    std::vector<uint8_t> in8;
    std::vector<u64> in64;
    DevBuf dev_in;                     // the inputs, when they were submitted on the device
    u64 pid = 0;
    std::vector<u64> ur, ui, vr, vi;
    std::vector<int> bkind;
    std::vector<u64> bint;
    std::vector<std::string> bstr;
    bool keep_outputs = false;
    // results
    std::atomic<int> state{0};   // 0 queued, 1 running, 2 done, 3 failed
    std::string err;
    std::vector<uint8_t> proof;
    std::vector<u64> point;      // (re, im) per variable: nInBits line point, then w3
    fp2 value = {0, 0};
    std::vector<u64> outputs;    // [cp][2^out_bits] when keep_outputs
    double t_submit = 0, t_start = 0, t_end = 0;
    double phase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int batch = 0;               // frames in the launch set that proved it (1 = alone)
    size_t footprint = 0;        // device bytes of its tables (Shape::footprint)
};

// Sizes of one frame.
struct Shape {
    int nin = 0, nout = 0, ncb = 0, cp = 0, NW = 0, NO = 0, n_z = 0, n_gates = 0;
    int n_terms = 0;   // early-claim terms (claim_terms.hpp)
    size_t rows_n = 0, out_n = 0, in_bytes = 0;
    // false (with the error recorded) for frames this build does not take
    bool set(const Job& J);
    // device bytes of the frame's tables in a batch (inputs, rows, outputs,
    // wire tables, beta, gate coefficients, partials)
    size_t footprint() const;
};

// Records the thread's error message on the job; returns false.
bool job_fail(Job& J);

// The transcript of one frame, driven step by step:
//   prepare(outputs) / prepare8     the proof's io item over the outputs (u64,
//                                   or one byte each when all are below 256)
//   write_prefix()                  (q, io) pickled (any thread, after prepare)
//   head()                          transcript head, z1 / z2 drawn
//   early(rd, f)   rd < ncb         cubic round, challenge r -> w3
//   late(k, f)     k < 2 nin        quadratic round, challenge -> w1 / w2
//   finish_layer(line, wiring13)    H, product proof, wiring equality, tau,
//                                   pok-final, the input point
//   input_value(v)                  input MLE value proof
//   write_rest()                    after write_prefix: rest of the proof
class FrameProto {
public:
    FrameProto(Job& J, const Shape& S);
    void prepare(const u64* outputs);
    void prepare8(const uint8_t* outputs);
    void write_prefix();
    void head();
    bool early(int rd, const fp2 f[4], fp2& r);
    bool late(int k, const fp2 f[3], fp2& r);
    WiringArgs wiring_args() const;   // z1, w1, w2 for the wiring kernel
    bool finish_layer(const fp2* line_vals, const fp2* wiring13);
    void input_value(fp2 v);
    void write_rest();

    std::vector<fp2> z1, z2, w1, w2, w3;
    PointArg point;   // n_z variables after finish_layer

private:
    Job& J;
    Shape S;
    const u64* outs = nullptr;
    const uint8_t* outs8 = nullptr;
    FsTranscriptT<HostSha256> fs;
    std::vector<PNode> trans;
    int used = 0;
    PickleWriter pw;
    PNode io;
    std::vector<fp2> v_elem, out_elem;
    fp2 prev = {0, 0};
    bool have_prev = false;

    void begin_io(PNode&& values);
    fp2 chal(const char* label);
    fp2 chal_num(const char* prefix, int k);
    void put(fp2 x);
    void put2(fp2 a, fp2 b);
    bool commit(fp2 m, fp2& tag, fp2& d);
    bool pok_temp(const fp2* outs, int n, std::vector<fp2>& rv);
    void eq_proof(fp2 v2);
    bool round_tail(const fp2* outs, int n, int rd, fp2& nrand);
};

// Host threads for the per-frame transcript work of a batch.  parallel_for
// runs on the pool and on the calling thread and returns when every index is
// done; spawn queues background work behind it (taken only when no
// parallel_for work waits) and wait joins a group of spawned tasks.
class TaskPool {
public:
    struct Group {
        std::mutex mu;
        std::condition_variable cv;
        int pending = 0;
    };
    explicit TaskPool(int n_threads);
    ~TaskPool();
    int threads() const { return (int)th.size(); }
    void parallel_for(int n, const std::function<void(int)>& fn);
    void spawn(Group& g, std::function<void()> fn);
    void wait(Group& g);

private:
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> hi, lo;
    bool stop = false;
    void run();
};

// Device and host resources of one lockstep batch driver (batch.hip).
struct BatchCtx;
BatchCtx* batch_ctx_create();
void batch_ctx_destroy(BatchCtx* b);

// Where a lockstep batch spends its time, accumulated over batches (ms unless
// named otherwise; "sum" = summed over the batch's frames, "wall" = elapsed).
enum BatchProf {
    BP_BATCHES, BP_FRAMES,
    BP_STAGE_WALL,        // inputs copied into pinned staging
    BP_LAUNCH_WALL,       // uploads, evaluation launches, output copies queued
    BP_HEAD_WALL,         // every frame's transcript head done
    BP_HEAD_WAIT_SUM,     // waiting for a frame's outputs to land
    BP_HEAD_CPU_SUM,      // transcript head (SHA-256 of the outputs) per frame
    BP_PREFIX_CPU_SUM,    // proof prefix (q, io) pickled per frame
    BP_ROUNDS_WALL,       // eq tables through the input MLE value
    BP_STEP_WAIT,         // driver waiting for published sums in those steps
    BP_STEP_HOST_WALL,    // per-step transcript work (parallel) in those steps
    BP_REST_WALL,         // waiting for the prefixes and writing the proof ends
    BP_CLAIM_DEV,         // device time of the early-claim launches (events)
    BP_FOLD_DEV,          // device time of the early-fold launches (events)
    BP_COUNT
};

// Proves the frames of `jobs` together; a frame that fails (its error in
// J->err, state 3) does not stop the others.  false only on a device failure,
// after which every frame of the batch is failed.  prof (BP_COUNT doubles)
// accumulates the batch's timings.
bool prove_batch(BatchCtx& B, std::vector<Job*>& jobs, TaskPool& tasks, double* prof);
// The schedule tuner of a driver: state[0] phase (2 = settled), [1] claim
// items per frame, [2] fold workgroups per frame (the knobs in use next),
// [3..5] claim costs (ns per term) and [6..8] fold costs (ns per byte) of the
// candidates tried.  Fixing the knobs (claim_items > 0) stops the tuning.
void batch_tuning(const BatchCtx& B, double state[9]);
void batch_fix_knobs(BatchCtx& B, long long claim_items, int fold_cap);

}  // namespace frame
}  // namespace tp
