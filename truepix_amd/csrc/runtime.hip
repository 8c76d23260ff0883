// runtime.hip - device bring-up, error state and buffers for libtruepix_hip.so.
#include "runtime.hpp"

#include <condition_variable>
#include <mutex>
#include <stdarg.h>
#include <stdio.h>

namespace tp {

static thread_local char g_err[512] = {0};
static std::once_flag g_once;
static bool g_ok = false;
static hipStream_t g_stream = nullptr;

// The no-alloc scratch.  In host code `scratch` names a host-side stand-in,
// not device memory: host code uses only the address looked up below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices] = {nullptr};
static char g_scratch_err[256] = {0};

bool fail(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return false;
}

void clear_error() { g_err[0] = 0; }

bool tp_last_error_set() { return g_err[0] != 0; }

static void init_once()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        fail("no HIP device visible");
        return;
    }
    if (hipSetDevice(0) != hipSuccess) { fail("hipSetDevice(0) failed"); return; }
    if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
        fail("hipStreamCreate failed");
        return;
    }
    // this device's copy of the scratch array (device 0 is current here)
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && ptr)
        g_scratch[0] = ptr;
    else
        snprintf(g_scratch_err, sizeof(g_scratch_err),
                 "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_ok = true;
}

void* scratch_ptr()
{
    if (!device_ready()) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices || !g_scratch[dev]) {
        fail("no-alloc scratch unavailable on device %d%s%s", dev, g_scratch_err[0] ? ": " : "",
             g_scratch_err);
        return nullptr;
    }
    return g_scratch[dev];
}

bool device_ready()
{
    std::call_once(g_once, init_once);
    if (!g_ok) return fail("HIP device unavailable: truepix_hip has no host fallback");
    return true;
}

hipStream_t stream() { return g_stream; }

// The gate between resident phases and frees: writer-preferring, so a free
// is not starved by phases that keep overlapping.
static std::mutex g_gate_mu;
static std::condition_variable g_gate_cv;
static int g_resident = 0;        // phases in flight
static int g_free_waiting = 0;    // frees waiting for them to end
static bool g_freeing = false;

void resident_enter()
{
    std::unique_lock<std::mutex> lk(g_gate_mu);
    g_gate_cv.wait(lk, [] { return !g_freeing && g_free_waiting == 0; });
    ++g_resident;
}

void resident_leave()
{
    std::lock_guard<std::mutex> lk(g_gate_mu);
    if (--g_resident == 0) g_gate_cv.notify_all();
}

template <typename F> static void outside_resident(F&& f)
{
    {
        std::unique_lock<std::mutex> lk(g_gate_mu);
        ++g_free_waiting;
        g_gate_cv.wait(lk, [] { return g_resident == 0 && !g_freeing; });
        --g_free_waiting;
        g_freeing = true;
    }
    f();
    {
        std::lock_guard<std::mutex> lk(g_gate_mu);
        g_freeing = false;
    }
    g_gate_cv.notify_all();
}

void free_device(void* p)
{
    if (p) outside_resident([p] { (void)hipFree(p); });
}

void free_host(void* p)
{
    if (p) outside_resident([p] { (void)hipHostFree(p); });
}

bool DevBuf::ensure(size_t need)
{
    if (need <= bytes && ptr) return true;
    void* fresh = nullptr;
    if (hipMalloc(&fresh, need ? need : 16) != hipSuccess)
        return fail("hipMalloc(%zu) failed", need);
    free_device(ptr);
    ptr = fresh;
    bytes = need;
    return true;
}

void DevBuf::release()
{
    free_device(ptr);
    ptr = nullptr;
    bytes = 0;
}

bool MappedSlot::ready()
{
    if (vals) return true;
    void* hv = nullptr;
    TP_HIP_OK(hipHostMalloc(&hv, 2048, hipHostMallocMapped | hipHostMallocCoherent));
    void* dv = nullptr;
    if (hipHostGetDevicePointer(&dv, hv, 0) != hipSuccess) {
        free_host(hv);
        return fail("hipHostGetDevicePointer failed");
    }
    vals = hv;
    seq = reinterpret_cast<volatile unsigned*>(static_cast<char*>(hv) + 1024);
    *seq = 0;
    dvals = dv;
    dseq = reinterpret_cast<unsigned*>(static_cast<char*>(dv) + 1024);
    ans = reinterpret_cast<volatile uint64_t*>(static_cast<char*>(hv) + 1536);
    dans = reinterpret_cast<uint64_t*>(static_cast<char*>(dv) + 1536);
    ans[0] = ans[1] = ~0ull;
    next = 0;
    return true;
}

bool MappedSlot::wait(hipStream_t st, unsigned s, const char* what)
{
    long spins = 0;
    while (*seq != s) {
        if (++spins % 2048 == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail("%s: stream error %s", what, hipGetErrorString(q));
            if (q == hipSuccess && *seq != s) return fail("%s: result not published", what);
        } else {
            __builtin_ia32_pause();
        }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return true;
}

void MappedSlot::release()
{
    free_host(vals);
    vals = dvals = nullptr;
    seq = nullptr;
    dseq = nullptr;
    ans = nullptr;
    dans = nullptr;
}

}  // namespace tp

extern "C" {

const char* tp_last_error(void) { return tp::g_err; }

void tp_clear_error(void) { tp::clear_error(); }

int tp_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
