// Shared host-side plumbing for the pq3d C-ABI library: error channel and
// device-memory helpers.  Every exported entry point returns an int status
// (0 = success) and records a message retrievable with pq3d_last_error().
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

namespace pq3d {

void set_error(const char* fmt, ...);
void clear_error();

#define PQ_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::pq3d::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                              __LINE__);                                                   \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

#define PQ_CHECK(cond, ...)              \
    do {                                 \
        if (!(cond)) {                   \
            ::pq3d::set_error(__VA_ARGS__); \
            return -2;                   \
        }                                \
    } while (0)

// Grow-only device scratch owned by the calling thread (used by the drop-in
// entry points whose callers pass host buffers).
struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
    int device = -1;
    ~Scratch();
};
int scratch_get(Scratch& s, size_t bytes, void** out);

// Matrix storage types of the dense traction operators.
// MAT_NONE: no dense matrices (the traction rates come from an attached H-matrix).
enum MatType : int { MAT_F64 = 0, MAT_F32 = 1, MAT_BF16 = 2, MAT_NONE = 3 };
inline size_t mat_elem_bytes(int t) { return t == MAT_F64 ? 8 : (t == MAT_F32 ? 4 : (t == MAT_BF16 ? 2 : 0)); }

// Bit k of *mask set iff device matrix k (ld x n elements of type mat_type;
// ld * elem bytes a multiple of 16) holds only +0/-0.  Reads every matrix
// once; synchronous.
int scan_zero_matrices(void* const* mats, int nmat, int mat_type, long long ld, long long n, int* mask);

}  // namespace pq3d
