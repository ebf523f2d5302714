// Shared host-side plumbing for the pq3d C-ABI library: error channel and
// device-memory helpers.  Every exported entry point returns an int status
// (0 = success) and records a message retrievable with pq3d_last_error().
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <tuple>
#include <utility>

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

// Enqueues kernels on a stream, or, when graph is set, appends them to that
// graph as a chain of kernel nodes (each depending on the one before).  Graphs
// are built node by node rather than by stream capture: a capture can be
// invalidated by synchronous HIP calls that other host threads make while it
// is open (e.g. a mesh loading on one thread while another starts a run).
struct Launch {
    hipStream_t stream = 0;
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;  // tail of the chain
    hipError_t err = hipSuccess;    // first failure, if any

    explicit Launch(hipStream_t s = 0, hipGraph_t g = nullptr) : stream(s), graph(g) {}

    template <typename... P, typename... A>
    void operator()(void (*kernel)(P...), dim3 grid, dim3 block, A&&... args) {
        static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
        std::tuple<P...> vals(std::forward<A>(args)...);  // each argument as the kernel's parameter type
        put(reinterpret_cast<const void*>(kernel), grid, block, vals, std::index_sequence_for<P...>{});
    }

  private:
    template <typename Tup, size_t... I>
    void put(const void* f, dim3 grid, dim3 block, Tup& vals, std::index_sequence<I...>) {
        void* argv[sizeof...(I) + 1] = {static_cast<void*>(&std::get<I>(vals))..., nullptr};
        hipError_t e;
        if (!graph) {
            e = hipLaunchKernel(f, grid, block, argv, 0, stream);
        } else {
            hipKernelNodeParams p{};
            p.func = const_cast<void*>(f);
            p.gridDim = grid;
            p.blockDim = block;
            p.sharedMemBytes = 0;
            p.kernelParams = argv;  // copied into the node
            p.extra = nullptr;
            hipGraphNode_t node = nullptr;
            e = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &p);
            if (e == hipSuccess) last = node;
        }
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
};

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
