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
        if (s.device == dev) (void)hipFree(s.p);
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

// flag[0] = 1 if any 16-byte word of p[0, nwords) has a bit set outside the
// sign bits described by `keep` (the per-32-bit-lane mask of magnitude bits).
__global__ void nonzero_words_kernel(const uint4* __restrict__ p, long long nwords, uint4 keep, int* flag) {
    bool nz = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
         i += (long long)gridDim.x * blockDim.x) {
        const uint4 w = p[i];
        nz |= ((w.x & keep.x) | (w.y & keep.y) | (w.z & keep.z) | (w.w & keep.w)) != 0u;
    }
    if (__any(nz) && (threadIdx.x & 63) == 0) flag[0] = 1;  // every writer stores 1
}

int scan_zero_matrices(void* const* mats, int nmat, int mat_type, long long ld, long long n, int* mask) {
    PQ_CHECK(mat_type == MAT_F64 || mat_type == MAT_F32 || mat_type == MAT_BF16, "scan: bad matrix type %d",
             mat_type);
    const size_t bytes = (size_t)ld * (size_t)n * mat_elem_bytes(mat_type);
    PQ_CHECK(bytes % 16 == 0, "scan: matrix bytes must be a multiple of 16");
    uint4 keep;
    if (mat_type == MAT_F32) keep = make_uint4(0x7fffffffu, 0x7fffffffu, 0x7fffffffu, 0x7fffffffu);
    else if (mat_type == MAT_BF16) keep = make_uint4(0x7fff7fffu, 0x7fff7fffu, 0x7fff7fffu, 0x7fff7fffu);
    else keep = make_uint4(0xffffffffu, 0x7fffffffu, 0xffffffffu, 0x7fffffffu);  // low word, high word
    // flags in a per-thread scratch kept across calls: no hipMalloc / hipFree
    // (a hipFree waits for the whole device, work of other streams included)
    static thread_local Scratch flags;
    int* d = nullptr;
    if (int rc = scratch_get(flags, 8 * sizeof(int), (void**)&d)) return rc;
    hipError_t e = hipMemset(d, 0, 8 * sizeof(int));
    const long long nwords = (long long)(bytes / 16);
    for (int k = 0; k < nmat && e == hipSuccess; ++k) {
        if (!mats[k]) continue;
        hipLaunchKernelGGL(nonzero_words_kernel, dim3(2048), dim3(256), 0, 0, (const uint4*)mats[k], nwords, keep,
                           d + k);
        e = hipGetLastError();
    }
    int h[8] = {0};
    if (e == hipSuccess) e = hipMemcpy(h, d, 8 * sizeof(int), hipMemcpyDeviceToHost);
    PQ_HIP(e);
    *mask = 0;
    for (int k = 0; k < nmat; ++k)
        if (mats[k] && !h[k]) *mask |= 1 << k;
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
