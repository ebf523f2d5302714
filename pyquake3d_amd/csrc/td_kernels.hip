// HIP kernels for triangular-dislocation stress Green's functions (FP64) and
// the C-ABI entry points built on them:
//
//   * the reference's ctypes signatures TDstressFS_C, TDstressEachSourceAtReceiver_C
//     and TDstressFS_C1 (src/TDstressFS_C.cpp:2095-2152, :2155-2200, :570-609);
//   * pq3d_gf_assemble: the six dense traction Green's functions
//     A1s, A2s, Bs, A1d, A2d, Bd (QDsim_gpu.py:1745-1878 worker/get_coreAB_mulproess
//     with GetTtstress :1400-1417), optionally over row/column index subsets
//     (the dense leaves and ACA rows/columns of the H-matrix).
//
// Work decomposition: one source-receiver pair per lane.  A workgroup owns 64
// receivers (lanes) and walks a chunk of sources; per-source constants come
// from a precomputed table through wave-uniform (scalar) loads, and the
// outputs are stored source-major (column-major, element (r, s) at s*ld + r)
// so that each wave's stores are contiguous.
#include "common.hpp"
#include "td_field.hpp"

#include <mutex>

extern "C" const char* pq3d_last_error(void);

namespace pq3d {

struct Burgers2 {
    double b[2][3];
};
struct Burgers1 {
    double b[1][3];
};
struct Tri {
    double P1[3], P2[3], P3[3];
};

__device__ __forceinline__ void load_tri(const double* P1, const double* P2, const double* P3, long long s,
                                         double (*V)[3]) {
    for (int k = 0; k < 3; ++k) {
        V[0][k] = P1[3 * s + k];
        V[1][k] = P2[3 * s + k];
        V[2][k] = P3[3 * s + k];
    }
}

// Table entry j: the source's geometry and Burgers vectors, and the sides it
// shares with entries j - 1 and j + 1 (the assembly kernel evaluates a shared
// side's free-surface correction once for both entries).
template <int NB, class B>
__global__ void src_setup_kernel(const double* __restrict__ P1, const double* __restrict__ P2,
                                 const double* __restrict__ P3, const long long* __restrict__ cols, long long n,
                                 B burg, SrcGeom<NB>* __restrict__ out) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    long long s = cols ? cols[j] : j;
    SrcGeom<NB> g;
    src_setup<NB>(P1 + 3 * s, P2 + 3 * s, P3 + 3 * s, burg.b, g);
    double V[3][3], W[3][3];
    load_tri(P1, P2, P3, s, V);
    if (j > 0) {
        load_tri(P1, P2, P3, cols ? cols[j - 1] : j - 1, W);
        g.prev_side = shared_side(V, W, &g.prev_rev);
    }
    if (j + 1 < n) {
        int rev;
        load_tri(P1, P2, P3, cols ? cols[j + 1] : j + 1, W);
        g.next_side = shared_side(V, W, &rev);
    }
    if (g.prev_side >= 0 && g.prev_side == g.next_side) g.prev_side = -1;  // an edge of three triangles
    out[j] = g;
}

__global__ void src_setup_one_kernel(Tri t, Burgers1 burg, SrcGeom<1>* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        SrcGeom<1> g;
        src_setup<1>(t.P1, t.P2, t.P3, burg.b, g);
        *out = g;
    }
}

template <class T>
struct Mat6 {
    T* m[6];
};

template <class T> __device__ __forceinline__ T cvt_out(double x);
template <> __device__ __forceinline__ double cvt_out<double>(double x) { return x; }
template <> __device__ __forceinline__ float cvt_out<float>(double x) { return (float)x; }

// Dense traction Green's functions for receivers rows[0:nr] x sources (setup
// table entries 0:nc).  Output k in {A1s, A2s, Bs, A1d, A2d, Bd}, element
// (i, j) at out.m[k][j * ld + i].  One wave per workgroup (lanes = receivers)
// walks entries j0 .. j1 - 1 in table order.  A side shared by consecutive
// entries has its free-surface correction evaluated once: the first entry
// takes the responses to the three unit Burgers vectors of the side frame
// (kept per lane in LDS) and weights them by its own Burgers vectors; the
// next entry weights the same responses by its Burgers vectors, negated for a
// side it runs the other way (the correction of side BA is minus that of side
// AB, with the frame axes ey1, ey2 reversed and ey3 kept).  That replaces two
// of the six angular-dislocation end evaluations of most pairs.
template <class T>
__global__ __launch_bounds__(64) void gf_assemble_kernel(const SrcGeom<2>* __restrict__ geo,
                                                         const double* __restrict__ xg,
                                                         const double* __restrict__ ev,
                                                         const long long* __restrict__ rows, long long nr,
                                                         long long nc, int src_per_block, double mu,
                                                         double lam, double nu, int halfspace, Mat6<T> out,
                                                         long long ld) {
    __shared__ double carry[2][27][64];  // [entry parity][basis k * 9 + gradient component][lane]
    const int lane = threadIdx.x;
    const long long i = (long long)blockIdx.x * 64 + lane;
    const bool active = i < nr;
    const long long r = active ? (rows ? rows[i] : i) : 0;
    const double X = xg[3 * r + 0], Y = xg[3 * r + 1], Z = xg[3 * r + 2];
    double e[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = ev[9 * r + k];
    const long long j0 = (long long)blockIdx.y * src_per_block;
    const long long j1 = j0 + src_per_block < nc ? j0 + src_per_block : nc;
    for (long long j = j0; j < j1; ++j) {
        const int cur = (int)((j - j0) & 1);
        double G[2][9];
        td_pair_grad_shared<double>(
            geo[j], X, Y, Z, nu, halfspace != 0, j > j0,
            [&](int k, int c, double v) { carry[cur][9 * k + c][lane] = v; },
            [&](int k, int c) { return carry[cur ^ 1][9 * k + c][lane]; }, G);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            double eps[6], sig[6], t[3];
            grad_to_strain_stress<double>(G[b], mu, lam, eps, sig);
            stress_to_traction<double>(sig, e, t);
            if (active) {
                const size_t o = (size_t)j * (size_t)ld + (size_t)i;
#pragma unroll
                for (int k = 0; k < 3; ++k) out.m[b * 3 + k][o] = cvt_out<T>(t[k]);
            }
        }
    }
}

// n receivers x one source (lanes = receivers): stress and strain, n x 6
__global__ __launch_bounds__(256) void td_stress_rcv_kernel(const SrcGeom<1>* __restrict__ geo,
                                                            const double* __restrict__ X,
                                                            const double* __restrict__ Y,
                                                            const double* __restrict__ Z, long long n,
                                                            double mu, double lam, double nu, int halfspace,
                                                            double* __restrict__ stress,
                                                            double* __restrict__ strain) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double G[1][9];
    td_pair_grad<double, 1, true>(*geo, X[i], Y[i], Z[i], nu, halfspace != 0, G);
    double eps[6], sig[6];
    grad_to_strain_stress<double>(G[0], mu, lam, eps, sig);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        stress[6 * i + k] = sig[k];
        strain[6 * i + k] = eps[k];
    }
}

// one receiver x n sources (lanes = sources): stress and strain, n x 6
__global__ __launch_bounds__(256) void td_stress_src_kernel(const SrcGeom<1>* __restrict__ geo, double X,
                                                            double Y, double Z, long long n, double mu,
                                                            double lam, double nu, int halfspace,
                                                            double* __restrict__ stress,
                                                            double* __restrict__ strain) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double G[1][9];
    td_pair_grad<double, 1, true>(geo[i], X, Y, Z, nu, halfspace != 0, G);
    double eps[6], sig[6];
    grad_to_strain_stress<double>(G[0], mu, lam, eps, sig);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        stress[6 * i + k] = sig[k];
        strain[6 * i + k] = eps[k];
    }
}

static inline double poisson(double mu, double lam) { return 1.0 / (1.0 + lam / mu) / 2.0; }

static thread_local Scratch g_scratch_dropin;

// n receivers x 1 source through the GPU; host buffers in and out.
static int stress_rcv_host(const double* X, const double* Y, const double* Z, size_t n, const double* P1,
                           const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu,
                           double lam, bool hs, double* stress, double* strain) {
    if (n == 0) return 0;
    const size_t geo_bytes = (sizeof(SrcGeom<1>) + 255) / 256 * 256;
    const size_t bytes = geo_bytes + 3 * n * sizeof(double) + 12 * n * sizeof(double);
    void* base = nullptr;
    if (scratch_get(g_scratch_dropin, bytes, &base)) return -1;
    SrcGeom<1>* d_geo = (SrcGeom<1>*)base;
    double* d_xyz = (double*)((char*)base + geo_bytes);
    double* d_st = d_xyz + 3 * n;
    double* d_sn = d_st + 6 * n;
    PQ_HIP(hipMemcpy(d_xyz, X, n * sizeof(double), hipMemcpyHostToDevice));
    PQ_HIP(hipMemcpy(d_xyz + n, Y, n * sizeof(double), hipMemcpyHostToDevice));
    PQ_HIP(hipMemcpy(d_xyz + 2 * n, Z, n * sizeof(double), hipMemcpyHostToDevice));
    Tri t;
    for (int k = 0; k < 3; ++k) {
        t.P1[k] = P1[k];
        t.P2[k] = P2[k];
        t.P3[k] = P3[k];
    }
    Burgers1 b{{{Ts, Ss, Ds}}};
    hipLaunchKernelGGL(src_setup_one_kernel, dim3(1), dim3(64), 0, 0, t, b, d_geo);
    PQ_HIP(hipGetLastError());
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(td_stress_rcv_kernel, dim3(blocks), dim3(256), 0, 0, d_geo, d_xyz, d_xyz + n,
                       d_xyz + 2 * n, (long long)n, mu, lam, poisson(mu, lam), hs ? 1 : 0, d_st, d_sn);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipMemcpy(stress, d_st, 6 * n * sizeof(double), hipMemcpyDeviceToHost));
    PQ_HIP(hipMemcpy(strain, d_sn, 6 * n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

static void fail_loudly(const char* where, double* a, size_t na, double* b, size_t nb) {
    fprintf(stderr, "pq3d: %s failed: %s\n", where, pq3d_last_error());
    fflush(stderr);
    for (size_t i = 0; i < na; ++i) a[i] = NAN;
    for (size_t i = 0; i < nb; ++i) b[i] = NAN;
}

// ---------------------------------------------------------------------------
// pq3d_gf_assemble and its variant that allocates nothing share the staging
// layout and the launches; only where the staging memory comes from differs.
// ---------------------------------------------------------------------------
struct GfAssembleArgs {
    const double *P1, *P2, *P3;
    long long nsrc;
    const double *xg, *ev;
    long long nrcv;
    double mu, lam;
    int halfspace, out_type;
    void* const* d_out;
    long long ld;
    const long long* rows;
    long long nr;
    const long long* cols;
    long long nc;
    int src_per_block;
};

static int gf_assemble_check(GfAssembleArgs& a) {
    PQ_CHECK(a.out_type == MAT_F64 || a.out_type == MAT_F32, "pq3d_gf_assemble: out_type must be f64 or f32");
    if (!a.rows) a.nr = a.nrcv;
    if (!a.cols) a.nc = a.nsrc;
    PQ_CHECK(a.nr >= 0 && a.nc >= 0 && a.ld >= a.nr, "pq3d_gf_assemble: bad sizes (nr=%lld nc=%lld ld=%lld)",
             a.nr, a.nc, a.ld);
    for (long long i = 0; a.rows && i < a.nr; ++i)
        PQ_CHECK(a.rows[i] >= 0 && a.rows[i] < a.nrcv, "pq3d_gf_assemble: row index %lld out of range", a.rows[i]);
    for (long long j = 0; a.cols && j < a.nc; ++j)
        PQ_CHECK(a.cols[j] >= 0 && a.cols[j] < a.nsrc, "pq3d_gf_assemble: column index %lld out of range",
                 a.cols[j]);
    if (a.src_per_block <= 0) {
        // the longest source runs (a run's first source shares no side with a
        // predecessor) that still give the grid 8192 workgroups: eight
        // one-wave workgroups per SIMD of the 256-CU part
        const long long gx = (a.nr + 63) / 64;
        a.src_per_block = 4;
        for (int spb = 64; spb > 4; spb /= 2)
            if (gx * ((a.nc + spb - 1) / spb) >= 8192) {
                a.src_per_block = spb;
                break;
            }
    }
    // the grid's second dimension holds at most 65,535 runs
    if ((a.nc + a.src_per_block - 1) / a.src_per_block > 65535) a.src_per_block = (int)((a.nc + 65534) / 65535);
    return 0;
}

// staging: source-geometry table (nc entries), vertices (nsrc x 9), receiver
// positions and frames (nrcv x 12), index lists
static size_t gf_geo_bytes(const GfAssembleArgs& a) {
    return ((size_t)a.nc * sizeof(SrcGeom<2>) + 255) / 256 * 256;
}
static size_t gf_staging_bytes(const GfAssembleArgs& a) {
    return gf_geo_bytes(a) + (size_t)a.nsrc * 9 * sizeof(double) + (size_t)a.nrcv * 12 * sizeof(double) +
           (size_t)((a.rows ? a.nr : 0) + (a.cols ? a.nc : 0)) * sizeof(long long) + 64;
}

static int gf_assemble_staged(const GfAssembleArgs& a, char* base) {
    const long long nsrc = a.nsrc, nrcv = a.nrcv, nr = a.nr, nc = a.nc;
    SrcGeom<2>* d_geo = (SrcGeom<2>*)base;
    double* d_P = (double*)(base + gf_geo_bytes(a));
    double* d_xg = d_P + 9 * nsrc;
    double* d_ev = d_xg + 3 * nrcv;
    long long* d_rows = (long long*)(d_ev + 9 * nrcv);
    long long* d_cols = d_rows + (a.rows ? nr : 0);
    hipError_t e = hipMemcpy(d_P, a.P1, nsrc * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_P + 3 * nsrc, a.P2, nsrc * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_P + 6 * nsrc, a.P3, nsrc * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_xg, a.xg, nrcv * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ev, a.ev, nrcv * 9 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && a.rows) e = hipMemcpy(d_rows, a.rows, nr * sizeof(long long), hipMemcpyHostToDevice);
    if (e == hipSuccess && a.cols) e = hipMemcpy(d_cols, a.cols, nc * sizeof(long long), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        Burgers2 b{{{0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}};  // strike (Ss = 1), dip (Ds = 1)
        hipLaunchKernelGGL((src_setup_kernel<2, Burgers2>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, 0,
                           d_P, d_P + 3 * nsrc, d_P + 6 * nsrc, a.cols ? d_cols : nullptr, nc, b, d_geo);
        const int spb = a.src_per_block;
        dim3 grid((unsigned)((nr + 63) / 64), (unsigned)((nc + spb - 1) / spb));
        const double nu = poisson(a.mu, a.lam);
        if (a.out_type == MAT_F64) {
            Mat6<double> o;
            for (int k = 0; k < 6; ++k) o.m[k] = (double*)a.d_out[k];
            hipLaunchKernelGGL(gf_assemble_kernel<double>, grid, dim3(64), 0, 0, d_geo, d_xg, d_ev,
                               a.rows ? d_rows : nullptr, nr, nc, spb, a.mu, a.lam, nu, a.halfspace, o, a.ld);
        } else {
            Mat6<float> o;
            for (int k = 0; k < 6; ++k) o.m[k] = (float*)a.d_out[k];
            hipLaunchKernelGGL(gf_assemble_kernel<float>, grid, dim3(64), 0, 0, d_geo, d_xg, d_ev,
                               a.rows ? d_rows : nullptr, nr, nc, spb, a.mu, a.lam, nu, a.halfspace, o, a.ld);
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(0);  // the null stream only: other streams keep running
    }
    if (e != hipSuccess) {
        set_error("pq3d_gf_assemble: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

// The fixed device scratch of pq3d_gf_assemble_noalloc.  Its name is a host
// stand-in, not a device address: host code uses only the per-device address
// from hipGetSymbolAddress (each GPU holds its own copy of the array).
constexpr size_t GF_NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[GF_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {nullptr};

// the current device's address of `scratch`, looked up on the first call on
// that device
static int noalloc_scratch(void** out) {
    int dev = 0;
    PQ_HIP(hipGetDevice(&dev));
    PQ_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "noalloc scratch: device %d beyond %d", dev,
             NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        PQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PQ_CHECK(p != nullptr, "noalloc scratch: no device address on device %d", dev);
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

// Drop-in for TDstressFS_C (src/TDstressFS_C.cpp:2095-2152): n receivers x one
// TD, stress[i*6+k] and strain[i*6+k], k over [xx, yy, zz, xy, xz, yz].
void TDstressFS_C(const double* X, const double* Y, const double* Z, size_t n, const double* P1,
                  const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu, double lambda_,
                  bool jud_halfspace, double* stress, double* strain) {
    clear_error();
    if (jud_halfspace) {
        bool above = P1[2] > 0 || P2[2] > 0 || P3[2] > 0;
        for (size_t i = 0; i < n && !above; ++i) above = Z[i] > 0;
        if (above) {
            // same diagnostic as the reference (TDstressFS_C.cpp:1949-1953)
            printf("Half-space solution: Z coordinates must be negative!\n");
            fflush(stdout);
        }
    }
    if (stress_rcv_host(X, Y, Z, n, P1, P2, P3, Ss, Ds, Ts, mu, lambda_, jud_halfspace, stress, strain))
        fail_loudly("TDstressFS_C", stress, 6 * n, strain, 6 * n);
}

// Drop-in for TDstressEachSourceAtReceiver_C (src/TDstressFS_C.cpp:2155-2200):
// one receiver x n_sources TDs (P*_flat are n_sources x 3), outputs n_sources x 6.
void TDstressEachSourceAtReceiver_C(double x, double y, double z, const double* P1_flat, const double* P2_flat,
                                    const double* P3_flat, size_t n_sources, double Ss, double Ds, double Ts,
                                    double mu, double lambda_, bool jud_halfspace, double* stress_out,
                                    double* strain_out) {
    clear_error();
    const size_t n = n_sources;
    if (n == 0) return;
    if (jud_halfspace) {
        bool above = z > 0;
        for (size_t i = 0; i < n && !above; ++i)
            above = P1_flat[3 * i + 2] > 0 || P2_flat[3 * i + 2] > 0 || P3_flat[3 * i + 2] > 0;
        if (above) {
            printf("Half-space solution: Z coordinates must be negative!\n");
            fflush(stdout);
        }
    }
    const size_t geo_bytes = (n * sizeof(SrcGeom<1>) + 255) / 256 * 256;
    const size_t bytes = geo_bytes + 9 * n * sizeof(double) + 12 * n * sizeof(double);
    void* base = nullptr;
    int rc = scratch_get(g_scratch_dropin, bytes, &base);
    if (!rc) {
        SrcGeom<1>* d_geo = (SrcGeom<1>*)base;
        double* d_P = (double*)((char*)base + geo_bytes);
        double* d_st = d_P + 9 * n;
        double* d_sn = d_st + 6 * n;
        hipError_t e = hipMemcpy(d_P, P1_flat, 3 * n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_P + 3 * n, P2_flat, 3 * n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_P + 6 * n, P3_flat, 3 * n * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            Burgers1 b{{{Ts, Ss, Ds}}};
            const unsigned blocks = (unsigned)((n + 255) / 256);
            hipLaunchKernelGGL((src_setup_kernel<1, Burgers1>), dim3(blocks), dim3(256), 0, 0, d_P, d_P + 3 * n,
                               d_P + 6 * n, (const long long*)nullptr, (long long)n, b, d_geo);
            hipLaunchKernelGGL(td_stress_src_kernel, dim3(blocks), dim3(256), 0, 0, d_geo, x, y, z,
                               (long long)n, mu, lambda_, poisson(mu, lambda_), jud_halfspace ? 1 : 0, d_st, d_sn);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(stress_out, d_st, 6 * n * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(strain_out, d_sn, 6 * n * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error("TDstressEachSourceAtReceiver_C: %s", hipGetErrorString(e));
            rc = -1;
        }
    }
    if (rc) fail_loudly("TDstressEachSourceAtReceiver_C", stress_out, 6 * n, strain_out, 6 * n);
}

// Drop-in for TDstressFS_C1 (src/TDstressFS_C.cpp:570-609): full space only,
// twelve separate n-vectors.
void TDstressFS_C1(const double* X, const double* Y, const double* Z, size_t n, const double* P1, const double* P2,
                   const double* P3, double Ss, double Ds, double Ts, double mu, double lambda_, double* Sxx,
                   double* Syy, double* Szz, double* Sxy, double* Sxz, double* Syz, double* Exx, double* Eyy,
                   double* Ezz, double* Exy, double* Exz, double* Eyz) {
    clear_error();
    if (n == 0) return;
    double* st = (double*)malloc(12 * n * sizeof(double));
    if (!st) {
        fprintf(stderr, "pq3d: TDstressFS_C1: host allocation failed\n");
        return;
    }
    double* sn = st + 6 * n;
    if (stress_rcv_host(X, Y, Z, n, P1, P2, P3, Ss, Ds, Ts, mu, lambda_, false, st, sn))
        fail_loudly("TDstressFS_C1", st, 6 * n, sn, 6 * n);
    double* outs[12] = {Sxx, Syy, Szz, Sxy, Sxz, Syz, Exx, Eyy, Ezz, Exy, Exz, Eyz};
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 6; ++k) {
            outs[k][i] = st[6 * i + k];
            outs[6 + k][i] = sn[6 * i + k];
        }
    free(st);
}

// Dense traction Green's functions on the device.
//   P1, P2, P3 : nsrc x 3 source TD vertices (host)
//   xg, ev     : nrcv x 3 receiver positions and nrcv x 9 frames [e1, e2, n] (host)
//   rows/cols  : optional host index lists (nr / nc entries); null = all
//   d_out      : six device buffers (A1s, A2s, Bs, A1d, A2d, Bd); element (i, j)
//                of the nr x nc block at j*ld + i; out_type MAT_F64 or MAT_F32
int pq3d_gf_assemble(const double* P1, const double* P2, const double* P3, long long nsrc, const double* xg,
                     const double* ev, long long nrcv, double mu, double lam, int halfspace, int out_type,
                     void* const* d_out, long long ld, const long long* rows, long long nr, const long long* cols,
                     long long nc, int src_per_block) {
    clear_error();
    GfAssembleArgs a{P1, P2, P3, nsrc, xg, ev, nrcv, mu, lam, halfspace, out_type, d_out, ld, rows, nr, cols, nc,
                     src_per_block};
    if (int rc = gf_assemble_check(a)) return rc;
    if (a.nr == 0 || a.nc == 0) return 0;
    char* base = nullptr;
    PQ_HIP(hipMalloc((void**)&base, gf_staging_bytes(a)));
    const int rc = gf_assemble_staged(a, base);
    (void)hipFree(base);
    return rc;
}

// The same assembly with its staging (source-geometry table, copied geometry
// and index lists) in the library's fixed device scratch instead of a fresh
// allocation: nothing is allocated or freed.  Rejects inputs whose staging
// exceeds GF_NOALLOC_SCRATCH_BYTES.  Calls that overlap on one GPU race on
// the scratch; every call overwrites what it reads.
int pq3d_gf_assemble_noalloc(const double* P1, const double* P2, const double* P3, long long nsrc,
                             const double* xg, const double* ev, long long nrcv, double mu, double lam,
                             int halfspace, int out_type, void* const* d_out, long long ld, const long long* rows,
                             long long nr, const long long* cols, long long nc, int src_per_block) {
    clear_error();
    GfAssembleArgs a{P1, P2, P3, nsrc, xg, ev, nrcv, mu, lam, halfspace, out_type, d_out, ld, rows, nr, cols, nc,
                     src_per_block};
    if (int rc = gf_assemble_check(a)) return rc;
    if (a.nr == 0 || a.nc == 0) return 0;
    const size_t need = gf_staging_bytes(a);
    PQ_CHECK(need <= GF_NOALLOC_SCRATCH_BYTES,
             "pq3d_gf_assemble_noalloc: staging needs %zu bytes, the scratch holds %zu (%lld sources x %d B "
             "+ %lld vertices x 72 B + %lld receivers x 96 B + index lists); assemble in column chunks",
             need, GF_NOALLOC_SCRATCH_BYTES, a.nc, (int)sizeof(SrcGeom<2>), a.nsrc, a.nrcv);
    char* base = nullptr;
    if (int rc = noalloc_scratch((void**)&base)) return rc;
    return gf_assemble_staged(a, base);
}

}  // extern "C"
