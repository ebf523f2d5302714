// Library runtime: error channel, thread-local device scratch, device memory
// and event helpers exported through the C ABI (declared in include/pq3d.h).
#include "common.hpp"

#include <mutex>

namespace pq3d {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
void clear_error() { g_err.clear(); }

Scratch::~Scratch() {
    // Thread-local destructors may run after the HIP runtime is torn down at
    // process exit; the allocation is then reclaimed with the context.
}

int scratch_get(Scratch& s, size_t bytes, void** out) {
    int dev = 0;
    PQ_HIP(hipGetDevice(&dev));
    if (s.p && (s.bytes < bytes || s.device != dev)) {
        if (s.device == dev) hipFree(s.p);
        s.p = nullptr;
        s.bytes = 0;
    }
    if (!s.p) {
        size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        PQ_HIP(hipMalloc(&s.p, want));
        s.bytes = want;
        s.device = dev;
    }
    *out = s.p;
    return 0;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

const char* pq3d_last_error(void) { return g_err.c_str(); }

int pq3d_device_count(int* n) {
    clear_error();
    PQ_HIP(hipGetDeviceCount(n));
    return 0;
}

int pq3d_device_info(int dev, char* name, int name_len, int* cu_count, double* hbm_bytes, int* arch_major,
                     int* arch_minor) {
    clear_error();
    hipDeviceProp_t p;
    PQ_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (double)p.totalGlobalMem;
    if (arch_major) *arch_major = p.major;
    if (arch_minor) *arch_minor = p.minor;
    return 0;
}

int pq3d_malloc(void** p, size_t bytes) {
    clear_error();
    PQ_HIP(hipMalloc(p, bytes ? bytes : 8));
    return 0;
}

int pq3d_free(void* p) {
    clear_error();
    if (p) PQ_HIP(hipFree(p));
    return 0;
}

int pq3d_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    clear_error();
    PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int pq3d_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    clear_error();
    PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int pq3d_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    clear_error();
    PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return 0;
}

int pq3d_memset(void* dst, int value, size_t bytes) {
    clear_error();
    PQ_HIP(hipMemset(dst, value, bytes));
    return 0;
}

int pq3d_synchronize(void) {
    clear_error();
    PQ_HIP(hipDeviceSynchronize());
    return 0;
}

// Events on the library's stream (the null stream).
int pq3d_event_create(void** ev) {
    clear_error();
    hipEvent_t e;
    PQ_HIP(hipEventCreate(&e));
    *ev = (void*)e;
    return 0;
}
int pq3d_event_destroy(void* ev) {
    clear_error();
    PQ_HIP(hipEventDestroy((hipEvent_t)ev));
    return 0;
}
int pq3d_event_record(void* ev) {
    clear_error();
    PQ_HIP(hipEventRecord((hipEvent_t)ev, 0));
    return 0;
}
int pq3d_event_elapsed_ms(void* start, void* stop, float* ms) {
    clear_error();
    PQ_HIP(hipEventSynchronize((hipEvent_t)stop));
    PQ_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return 0;
}

}  // extern "C"
