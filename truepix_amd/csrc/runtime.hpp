// runtime.hpp - device bring-up, error reporting and grow-only device buffers
// shared by every translation unit of libtruepix_hip.so.
//
// Error model: the library never falls back to host arithmetic.  Entry points
// that return int report failure with a negative value; entry points whose
// reference signature returns void record the failure in a thread-local
// message that tp_last_error() returns (the Python bridge checks it after
// every call and raises).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace tp {

// Records a formatted error for tp_last_error(); returns false for chaining.
bool fail(const char* fmt, ...);
void clear_error();
// True if an error message is recorded on this thread.
bool tp_last_error_set();

// One-time device bring-up on device 0 (the reference uses device 0 only,
// TP/native_fp2/native_fp2.cu:88-93).  Returns false if no HIP device.
bool device_ready();

// Device scratch for the entry points that allocate nothing: a file-scope
// __device__ array of kScratchBytes in runtime.hip, whose address on device d
// is looked up once, at bring-up, with hipGetSymbolAddress.  Null when the
// lookup failed (the no-alloc entry points then fail with that error).  Every
// call on a device shares that device's copy: calls that overlap race.
constexpr size_t kScratchBytes = 64 * 1024;
void* scratch_ptr();

// Library stream (default priority).  Every drop-in entry point enqueues on
// it and synchronizes before returning, as the reference ABI is synchronous.
hipStream_t stream();

// Resident phases and device-wide frees.  A resident kernel (a workgroup that
// takes each round's challenge from mapped host memory) ends only while its
// host thread keeps answering.  hipFree and hipHostFree synchronize the whole
// device, so one that starts while such a kernel runs waits for it; if the
// answering thread then queues behind that free inside the runtime, neither
// moves until the kernel's wait bound expires.  So a thread answering a
// resident kernel makes no runtime calls between the launch and the last
// answer, it holds a ResidentPhase over that span, and the library's frees
// (free_device, free_host) wait until no phase is in flight; new phases wait
// while a free is pending.  A thread must not free while it holds a phase.
void resident_enter();
void resident_leave();
struct ResidentPhase {
    ResidentPhase() { resident_enter(); }
    ~ResidentPhase() { resident_leave(); }
    ResidentPhase(const ResidentPhase&) = delete;
    ResidentPhase& operator=(const ResidentPhase&) = delete;
};
void free_device(void* p);   // hipFree outside every resident phase
void free_host(void* p);     // hipHostFree outside every resident phase

// Grow-only device allocation (malloc on grow, never shrinks).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool ensure(size_t need);
    void release();
    template <typename T> T* as() const { return reinterpret_cast<T*>(ptr); }
};

// A result slot in mapped, coherent host memory: a kernel writes up to 64
// Fp2 values, fences system scope and then stores the sequence number; the
// host spins on the sequence number instead of synchronizing the stream.
struct MappedSlot {
    void* vals = nullptr;                 // host view of the values
    volatile unsigned* seq = nullptr;     // host view of the sequence number
    void* dvals = nullptr;                // device views of the same memory
    unsigned* dseq = nullptr;
    volatile uint64_t* ans = nullptr;     // host view of a two-word answer slot
    uint64_t* dans = nullptr;             // (relay.hpp), and its device view
    unsigned next = 0;

    bool ready();                         // allocates on first use
    unsigned advance() { return ++next; }
    // Waits until *seq == s; fails if the stream reports an error or has
    // drained without publishing.
    bool wait(hipStream_t st, unsigned s, const char* what);
    void release();
};

// Kernel launch shapes: 256-thread workgroups (four wave64s per workgroup).
static constexpr int kBlock = 256;
inline unsigned grid_for(size_t n, size_t per_block = kBlock, unsigned cap = 1u << 20)
{
    size_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace tp

#define TP_HIP_OK(call)                                                        \
    do {                                                                       \
        hipError_t tp_err_ = (call);                                           \
        if (tp_err_ != hipSuccess)                                             \
            return tp::fail("%s failed: %s (%s:%d)", #call,                    \
                            hipGetErrorString(tp_err_), __FILE__, __LINE__);   \
    } while (0)
