// runtime.hip - device bring-up, error state and buffers for libtruepix_hip.so.
#include "runtime.hpp"

#include <mutex>
#include <stdarg.h>
#include <stdio.h>

namespace tp {

static thread_local char g_err[512] = {0};
static std::once_flag g_once;
static bool g_ok = false;
static hipStream_t g_stream = nullptr;

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
    g_ok = true;
}

bool device_ready()
{
    std::call_once(g_once, init_once);
    if (!g_ok) return fail("HIP device unavailable: truepix_hip has no host fallback");
    return true;
}

hipStream_t stream() { return g_stream; }

bool DevBuf::ensure(size_t need)
{
    if (need <= bytes && ptr) return true;
    void* fresh = nullptr;
    if (hipMalloc(&fresh, need ? need : 16) != hipSuccess)
        return fail("hipMalloc(%zu) failed", need);
    if (ptr) (void)hipFree(ptr);
    ptr = fresh;
    bytes = need;
    return true;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
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
