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

// ---------------------------------------------------------------------------
// Measurement probes: what this device sustains for the two resources the
// hot path is bound by, to quote beside the published peaks.
// ---------------------------------------------------------------------------

// Read-only stream: each workgroup reads one contiguous share of the buffer,
// its lanes 16-byte vectors 4 KB apart per instruction, eight loads in flight
// per lane (the fastest of the forms measured here: grid-stride and blocked,
// 4-16 loads, 2-32 workgroups per CU); a word is written only if the XOR of
// everything read hits a constant, so the loads cannot be dropped.
__global__ __launch_bounds__(256) void probe_stream_kernel(const uint4* __restrict__ p, long long n4,
                                                           unsigned* __restrict__ out) {
    const long long per = (n4 + gridDim.x - 1) / gridDim.x;
    const long long b0 = (long long)blockIdx.x * per, b1 = b0 + per < n4 ? b0 + per : n4;
    long long i = b0 + threadIdx.x;
    unsigned x = 0;
    for (; i + 7 * 256 < b1; i += 8 * 256) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[i + u * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u) x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    for (; i < b1; i += 256) {
        const uint4 v = p[i];
        x ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (x == 0x9e3779b9u) out[blockIdx.x] = x;
}

// FP64 FMA throughput: eight independent chains per lane; the result is
// written only if it hits a constant.
__global__ __launch_bounds__(256) void probe_fma64_kernel(int iters, double seed, double* __restrict__ out) {
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = seed + 1e-9 * threadIdx.x + k;
    const double m = 1.0000000001, c = 1e-12;
#pragma unroll 4
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = fma(a[k], m, c);
    }
    double s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    if (s == 12345.678) out[blockIdx.x] = s;
}

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

// Sustained read bandwidth over a device buffer of `bytes` (allocated here,
// rounded down to 16 bytes): the fastest of `reps` timed passes after one
// untimed pass; GB/s (1e9 bytes) and its ms.  Workgroups: 32 per CU.
int pq3d_probe_stream(long long bytes, int reps, float* ms_per, double* gbs) {
    clear_error();
    PQ_CHECK(bytes >= 16 && reps > 0 && ms_per && gbs, "pq3d_probe_stream: bad arguments");
    int dev = 0;
    PQ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PQ_HIP(hipGetDeviceProperties(&p, dev));
    const long long n4 = bytes / 16;
    const unsigned blocks = (unsigned)p.multiProcessorCount * 32;
    void* buf = nullptr;
    unsigned* out = nullptr;
    PQ_HIP(hipMalloc(&buf, (size_t)n4 * 16));
    hipError_t err = hipMalloc((void**)&out, blocks * sizeof(unsigned));
    hipEvent_t a = nullptr, b = nullptr;
    if (err == hipSuccess) err = hipMemset(buf, 0, (size_t)n4 * 16);
    if (err == hipSuccess) err = hipEventCreate(&a);
    if (err == hipSuccess) err = hipEventCreate(&b);
    float best = 0;
    for (int r = -1; r < reps && err == hipSuccess; ++r) {
        err = hipEventRecord(a, 0);
        hipLaunchKernelGGL(probe_stream_kernel, dim3(blocks), dim3(256), 0, 0, (const uint4*)buf, n4, out);
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(b, 0);
        if (err == hipSuccess) err = hipEventSynchronize(b);
        float ms = 0;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
        if (r >= 0 && (best == 0 || ms < best)) best = ms;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(buf);
    if (out) (void)hipFree(out);
    PQ_HIP(err);
    *ms_per = best;
    *gbs = (double)n4 * 16 / (best * 1e-3) / 1e9;
    return 0;
}

// Sustained FP64 FMA rate: `iters` FMAs in each of eight chains per lane,
// 256-thread workgroups, 64 per CU; TFLOP/s (an FMA = 2 flops) and ms per
// launch over `reps` timed launches after one untimed.
int pq3d_probe_fma64(int iters, int reps, float* ms_per, double* tflops) {
    clear_error();
    PQ_CHECK(iters > 0 && reps > 0 && ms_per && tflops, "pq3d_probe_fma64: bad arguments");
    int dev = 0;
    PQ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PQ_HIP(hipGetDeviceProperties(&p, dev));
    const unsigned blocks = (unsigned)p.multiProcessorCount * 64;
    double* out = nullptr;
    PQ_HIP(hipMalloc((void**)&out, blocks * sizeof(double)));
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t err = hipEventCreate(&a);
    if (err == hipSuccess) err = hipEventCreate(&b);
    float ms = 0;
    if (err == hipSuccess) {
        for (int r = -1; r < reps; ++r) {
            if (r == 0) (void)hipEventRecord(a, 0);
            hipLaunchKernelGGL(probe_fma64_kernel, dim3(blocks), dim3(256), 0, 0, iters, 1.0 + r, out);
        }
        err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(b, 0);
        if (err == hipSuccess) err = hipEventSynchronize(b);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(out);
    PQ_HIP(err);
    *ms_per = ms / reps;
    *tflops = 2.0 * 8.0 * iters * (double)blocks * 256 / (*ms_per * 1e-3) / 1e12;
    return 0;
}

}  // extern "C"
