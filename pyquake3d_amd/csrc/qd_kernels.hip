// Dense quasi-dynamic stepping on the device: traction matvec plus the fused
// rate-and-state RHS / Dormand-Prince attempt loop, with accept/reject decided
// on the device.
//
// Per step (the reference's simu_forwardGPU, QDsim_gpu.py:942-1027):
//   1. matvec_kernel: with u = V - Vpl (cos rake0, sin rake0),
//        AdotV1 = A1s u1 + A1d u2,  AdotV2 = A2s u1 + A2d u2,
//        dsigma/dt = Bs u1 + Bd u2 (+ normal loading, added in step 2).
//      Matrices are source-major (column j contiguous), so a workgroup owns a
//      row tile and a column chunk (split-K); each lane streams 16-byte
//      vectors of every matrix, accumulates in FP64, and writes one partial
//      sum per chunk.  Chunks are summed in a fixed order (deterministic).
//   2. rk_step_kernel: a persistent kernel whose lanes own one element each
//      for the whole step.  Each attempt evaluates the 7 Dormand-Prince stages
//      of the aging-law RHS with radiation damping in registers, reduces the
//      two relative-error maxima with wavefront shuffles and one atomic per
//      workgroup, meets the other workgroups at a grid barrier (bounded spin),
//      and every workgroup then applies the same accept/reject rule.  The
//      accepted state, slip rates, slip, rake and friction are written in the
//      same kernel.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace pq3d {

// ---------------------------------------------------------------------------
// Fields of a dense simulation (order is part of the C ABI)
// ---------------------------------------------------------------------------
enum Field : int {
    F_TNO = 0, F_TT1O, F_TT2O, F_STATE, F_SLIPV1, F_SLIPV2, F_SLIPV, F_SLIP1, F_SLIP2, F_SLIP, F_TT, F_RAKE,
    F_RAKE0, F_FRIC, F_A, F_B, F_DC, F_F0, F_SHEAR_LOADING, F_NORMAL_LOADING, F_P, F_DPDT0,
    F_ADOTV1, F_ADOTV2, F_DSIGMADT, F_COSR0, F_SINR0, NFIELDS
};

struct Ctrl {                    // device-resident step control
    double dttry;                // step size to try next
    double time;                 // simulated time
    double last_h, last_dtnext;
    double err1, err2;           // last attempt's relative error maxima (+1e-10)
    int status;                  // 0 ok, 1 dt too small / >20 rejections, 2 barrier timeout
    int nrjct;                   // rejections in the last step
    long long steps;             // accepted steps
    long long total_rejections;
    unsigned bar_count, bar_gen; // grid barrier
    unsigned long long slot[2][24][2];  // per step parity, per attempt: error maxima (bits)
    unsigned long long vmax[2];         // per step parity: max slip rate (bits)
    unsigned long long att_max[2];      // standalone attempt (pq3d_dense_dp45): error maxima (bits)
    double att_out[4];                  // standalone attempt: condition1, condition2, hnew1, hnew2
};

struct SimArgs {
    double* f[NFIELDS];
    const double* partial;       // nchunk x 3 x n
    Ctrl* ctrl;
    double* records;             // optional per-step log, 4 per step
    long long rec_index;         // slot in records for this step (or -1)
    long long n;
    int nchunk;
    int nblocks;
    double V0, Vpl, eta;         // eta = mu / (2 Cs)
    int fix_tn;
    int err_mode;                // 1: GPU semantics (normal-traction error dropped), 0: CPU
    long long spin_limit;
};

// ---------------------------------------------------------------------------
// Matvec: partial sums of the three traction rates over a column chunk
// ---------------------------------------------------------------------------
template <class T> struct Vec16;
template <> struct Vec16<float> { using type = float4; static constexpr int n = 4; };
template <> struct Vec16<double> { using type = double2; static constexpr int n = 2; };
struct bf16x8 { unsigned short v[8]; };
template <> struct Vec16<unsigned short> { using type = uint4; static constexpr int n = 8; };

__device__ __forceinline__ void unpack(const float4& a, double* o) { o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; }
__device__ __forceinline__ void unpack(const double2& a, double* o) { o[0] = a.x; o[1] = a.y; }
__device__ __forceinline__ double bf16_to_f64(unsigned int bits16) { return (double)__uint_as_float(bits16 << 16); }
__device__ __forceinline__ void unpack(const uint4& a, double* o) {
    unsigned int w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o[2 * k] = bf16_to_f64(w[k] & 0xffffu);
        o[2 * k + 1] = bf16_to_f64(w[k] >> 16);
    }
}

template <class T>
struct MatPtrs {
    const T* m[6];
};

// grid: (row tiles of 256*VN rows, column chunks of kc columns); block 256.
// n must be a multiple of VN (rows padded by the caller: ld >= n, ld % VN == 0).
template <class T, bool FIXTN>
__global__ __launch_bounds__(256) void matvec_kernel(MatPtrs<T> M, long long ld, long long n, int kc,
                                                     const double* __restrict__ V1, const double* __restrict__ V2,
                                                     const double* __restrict__ c0, const double* __restrict__ s0,
                                                     double Vpl, double* __restrict__ partial) {
    using VT = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    const long long r0 = ((long long)blockIdx.x * 256 + threadIdx.x) * VN;
    const long long j0 = (long long)blockIdx.y * kc;
    const long long j1 = j0 + kc < n ? j0 + kc : n;
    double acc[3][VN];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int v = 0; v < VN; ++v) acc[q][v] = 0.0;
    const bool active = r0 < n;
    if (active) {
        long long j = j0;
        for (; j + 3 < j1; j += 4) {  // four columns' loads in flight; the sums keep column order
            VT a[4][6];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    if (FIXTN && (k == 2 || k == 5)) continue;
                    a[t][k] = *reinterpret_cast<const VT*>(M.m[k] + (size_t)(j + t) * ld + r0);
                }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const double u1 = __dsub_rn(V1[j + t], __dmul_rn(Vpl, c0[j + t]));
                const double u2 = __dsub_rn(V2[j + t], __dmul_rn(Vpl, s0[j + t]));
                double x[6][VN];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    if (FIXTN && (k == 2 || k == 5)) continue;
                    unpack(a[t][k], x[k]);
                }
#pragma unroll
                for (int v = 0; v < VN; ++v) {
                    acc[0][v] = fma(x[0][v], u1, acc[0][v]);
                    acc[0][v] = fma(x[3][v], u2, acc[0][v]);
                    acc[1][v] = fma(x[1][v], u1, acc[1][v]);
                    acc[1][v] = fma(x[4][v], u2, acc[1][v]);
                    if (!FIXTN) {
                        acc[2][v] = fma(x[2][v], u1, acc[2][v]);
                        acc[2][v] = fma(x[5][v], u2, acc[2][v]);
                    }
                }
            }
        }
        for (; j < j1; ++j) {
            const double u1 = __dsub_rn(V1[j], __dmul_rn(Vpl, c0[j]));
            const double u2 = __dsub_rn(V2[j], __dmul_rn(Vpl, s0[j]));
            double x[6][VN];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                if (FIXTN && (k == 2 || k == 5)) continue;
                unpack(*reinterpret_cast<const VT*>(M.m[k] + (size_t)j * ld + r0), x[k]);
            }
#pragma unroll
            for (int v = 0; v < VN; ++v) {
                acc[0][v] = fma(x[0][v], u1, acc[0][v]);
                acc[0][v] = fma(x[3][v], u2, acc[0][v]);
                acc[1][v] = fma(x[1][v], u1, acc[1][v]);
                acc[1][v] = fma(x[4][v], u2, acc[1][v]);
                if (!FIXTN) {
                    acc[2][v] = fma(x[2][v], u1, acc[2][v]);
                    acc[2][v] = fma(x[5][v], u2, acc[2][v]);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int v = 0; v < VN; ++v)
                if (r0 + v < n) partial[((size_t)blockIdx.y * 3 + q) * n + r0 + v] = acc[q][v];
    }
}

// ---------------------------------------------------------------------------
// Rate-and-state RHS and Dormand-Prince (QDsim_gpu.py:830-891, :1128-1262)
// ---------------------------------------------------------------------------
struct Elem {        // per-element constants for one step
    double a, b, dc, f0, P, dPdt, sl;
    double A1, A2, dsig;   // frozen AdotV1, AdotV2, dsigma/dt [Pa/s]
    double V;              // frozen slip-rate magnitude
};

__device__ __forceinline__ double clip700(double x) { return x < -700.0 ? -700.0 : (x > 700.0 ? 700.0 : x); }

// returns (dTno, dTt1, dTt2, dstate) in the state's units (MPa/s, 1/s)
__device__ __forceinline__ void rsf_rhs(const Elem& e, double V0, double eta, double Tno, double Tt1,
                                        double Tt2, double st, double* k) {
    const double tn = Tno * 1e6, t1 = Tt1 * 1e6, t2 = Tt2 * 1e6;
    const double sg = tn - e.P;
    const double as = e.a * sg;
    const double ex = exp(clip700(-st / e.a));
    const double c1 = clip700(t1 / as), c2 = clip700(t2 / as);
    const double ep1 = exp(c1), em1 = exp(-c1), ep2 = exp(c2), em2 = exp(-c2);
    const double ch1 = (ep1 + em1) / 2, ch2 = (ep2 + em2) / 2;
    const double sh1 = (ep1 - em1) / 2, sh2 = (ep2 - em2) / 2;
    const double q = 2 * V0 / as;
    const double dV1dtau = q * ex * ch1;
    const double dV2dtau = q * ex * ch2;
    const double dV1dsig = -2 * V0 * t1 / (e.a * (sg * sg)) * ex * ch1;
    const double dV2dsig = -2 * V0 * t2 / (e.a * (sg * sg)) * ex * ch2;
    const double dV1dst = -2 * V0 / e.a * ex * sh1;
    const double dV2dst = -2 * V0 / e.a * ex * sh2;
    const double dstdt = e.b / e.dc * (V0 * exp(clip700((e.f0 - st) / e.b)) - e.V);
    const double ds = e.dsig - e.dPdt;
    const double dt1 = (-e.A1 + e.sl - eta * (dV1dsig * ds + dV1dst * dstdt)) / (1.0 + eta * dV1dtau);
    const double dt2 = (-e.A2 + e.sl - eta * (dV2dsig * ds + dV2dst * dstdt)) / (1.0 + eta * dV2dtau);
    k[0] = e.dsig / 1e6;
    k[1] = dt1 / 1e6;
    k[2] = dt2 / 1e6;
    k[3] = dstdt;
}

__device__ __forceinline__ double py_max(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double py_min(double a, double b) { return b < a ? b : a; }

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Block-wide max of two u64 values, then one device-scope atomic each.
__device__ __forceinline__ void block_max2_atomic(unsigned long long x, unsigned long long y,
                                                  unsigned long long* dst, unsigned long long* lds) {
    x = wave_max_u64(x);
    y = wave_max_u64(y);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        lds[2 * w] = x;
        lds[2 * w + 1] = y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mx = 0, my = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
            mx = lds[2 * i] > mx ? lds[2 * i] : mx;
            my = lds[2 * i + 1] > my ? lds[2 * i + 1] : my;
        }
        if (dst) {
            atomicMax(dst, mx);
            atomicMax(dst + 1, my);
        }
    }
    __syncthreads();
}

// Central sense-counting grid barrier; returns false on timeout.
__device__ bool grid_barrier(Ctrl* c, unsigned nblocks, long long spin_limit) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        const unsigned g = __hip_atomic_load(&c->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned arrived = atomicAdd(&c->bar_count, 1u);
        if (arrived == nblocks - 1) {
            __hip_atomic_store(&c->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&c->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long spins = 0;
            while (__hip_atomic_load(&c->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > spin_limit) {
                    atomicExch(&c->status, 2);
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
    return ok != 0;
}

__device__ __forceinline__ unsigned long long fbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsf(unsigned long long u) { return __longlong_as_double((long long)u); }

// Dormand-Prince tableau (QDsim_gpu.py:1129-1159)
#define DP_B21 (.2)
#define DP_B31 (3. / 40)
#define DP_B32 (9. / 40.)
#define DP_B41 (44. / 45.)
#define DP_B42 (-56. / 15)
#define DP_B43 (32. / 9)
#define DP_B51 (19372. / 6561.)
#define DP_B52 (-25360 / 2187.)
#define DP_B53 (64448. / 6561.)
#define DP_B54 (-212. / 729.)
#define DP_B61 (9017. / 3168.)
#define DP_B62 (-355. / 33.)
#define DP_B63 (-46732. / 5247.)
#define DP_B64 (49. / 176.)
#define DP_B65 (-5103. / 18656.)
#define DP_B71 (35. / 384.)
#define DP_B73 (500. / 1113.)
#define DP_B74 (125. / 192.)
#define DP_B75 (-2187. / 6784.)
#define DP_B76 (11. / 84.)
#define DP_B81 (5179. / 57600.)
#define DP_B83 (7571. / 16695.)
#define DP_B84 (393. / 640.)
#define DP_B85 (-92097. / 339200.)
#define DP_B86 (187. / 2100.)
#define DP_B87 (1. / 40.)

// One Dormand-Prince attempt for one element: y5 (returned solution) and y4.
__device__ __forceinline__ void dp45(const Elem& e, double V0, double eta, const double* y, double h,
                                     double* y5, double* y4) {
    double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], k7[4], t[4];
    rsf_rhs(e, V0, eta, y[0], y[1], y[2], y[3], k1);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * DP_B21 * k1[i];
    rsf_rhs(e, V0, eta, t[0], t[1], t[2], t[3], k2);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * (DP_B31 * k1[i] + DP_B32 * k2[i]);
    rsf_rhs(e, V0, eta, t[0], t[1], t[2], t[3], k3);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * (DP_B41 * k1[i] + DP_B42 * k2[i] + DP_B43 * k3[i]);
    rsf_rhs(e, V0, eta, t[0], t[1], t[2], t[3], k4);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t[i] = y[i] + h * (DP_B51 * k1[i] + DP_B52 * k2[i] + DP_B53 * k3[i] + DP_B54 * k4[i]);
    rsf_rhs(e, V0, eta, t[0], t[1], t[2], t[3], k5);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t[i] = y[i] + h * (DP_B61 * k1[i] + DP_B62 * k2[i] + DP_B63 * k3[i] + DP_B64 * k4[i] + DP_B65 * k5[i]);
    rsf_rhs(e, V0, eta, t[0], t[1], t[2], t[3], k6);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        y5[i] = y[i] + h * (DP_B71 * k1[i] + DP_B73 * k3[i] + DP_B74 * k4[i] + DP_B75 * k5[i] + DP_B76 * k6[i]);
    rsf_rhs(e, V0, eta, y5[0], y5[1], y5[2], y5[3], k7);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        y4[i] = y[i] + h * (DP_B81 * k1[i] + DP_B83 * k3[i] + DP_B84 * k4[i] + DP_B85 * k5[i] + DP_B86 * k6[i] +
                            DP_B87 * k7[i]);
}

__global__ __launch_bounds__(256) void rk_step_kernel(SimArgs A) {
    __shared__ unsigned long long lds[16];
    Ctrl* C = A.ctrl;
    // A failed step (status 1) or a grid-barrier timeout (status 2) ends the
    // run: the later launches of the same run or graph leave the state, the
    // reduction slots and their record rows untouched (the host marks those
    // rows NaN).  Every workgroup reads the status the previous launch left.
    if (__hip_atomic_load(&C->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < A.n;
    const long long ii = act ? i : 0;
    double** F = A.f;
    const long long n = A.n;
    const int par = (int)(__hip_atomic_load(&C->steps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1);
    // clear the other parity's reduction slots (used by the next step); the
    // previous step's max slip rate is logged before its slot is cleared
    if (blockIdx.x == 0 && threadIdx.x < 48) C->slot[par ^ 1][threadIdx.x >> 1][threadIdx.x & 1] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (A.records && A.rec_index > 0) A.records[4 * (A.rec_index - 1) + 3] = bitsf(C->vmax[par ^ 1]);
        C->vmax[par ^ 1] = 0ull;
    }

    // frozen stiffness terms (QDsim_gpu.py:946-957)
    Elem e;
    double A1 = 0, A2 = 0, ds = 0;
    for (int c = 0; c < A.nchunk; ++c) {
        A1 += A.partial[((size_t)c * 3 + 0) * n + ii];
        A2 += A.partial[((size_t)c * 3 + 1) * n + ii];
        ds += A.partial[((size_t)c * 3 + 2) * n + ii];
    }
    e.a = F[F_A][ii];
    e.b = F[F_B][ii];
    e.dc = F[F_DC][ii];
    e.f0 = F[F_F0][ii];
    e.P = F[F_P][ii];
    e.dPdt = F[F_DPDT0][ii];
    e.sl = F[F_SHEAR_LOADING][ii];
    e.A1 = A1;
    e.A2 = A2;
    e.dsig = A.fix_tn ? F[F_NORMAL_LOADING][ii] : ds + F[F_NORMAL_LOADING][ii];
    e.V = F[F_SLIPV][ii];
    if (act && blockIdx.y == 0) {
        F[F_ADOTV1][ii] = e.A1;
        F[F_ADOTV2][ii] = e.A2;
        F[F_DSIGMADT][ii] = e.dsig;
    }
    double y[4] = {F[F_TNO][ii], F[F_TT1O][ii], F[F_TT2O][ii], F[F_STATE][ii]};
    double y5[4], y4[4];

    double h = C->dttry;
    int nrjct = 0;
    bool accepted = false;
    double dtnext = h;
    for (int att = 0; att < 22; ++att) {
        dp45(e, A.V0, A.eta, y, h, y5, y4);
        // relative error maxima (QDsim_gpu.py:1221-1238)
        double r1 = 0.0, r2 = 0.0;
        if (act) {
            r1 = fabs(fabs(y4[3] - y5[3]) / y4[3]);
            const double rt1 = fabs(fabs(y4[1] - y5[1]) / y4[1]);
            const double rt2 = fabs(fabs(y4[2] - y5[2]) / y4[2]);
            r2 = (rt2 > rt1 || rt2 != rt2) ? rt2 : rt1;
            if (A.err_mode == 0) {
                const double rn = fabs(fabs(y4[0] - y5[0]) / y4[0]);
                r2 = (rn > r2 || rn != rn) ? rn : r2;
            }
        }
        // NaN propagates: a quiet NaN's bit pattern exceeds every finite positive
        block_max2_atomic(fbits(r1), fbits(r2), &C->slot[par][att][0], lds);
        if (!grid_barrier(C, (unsigned)A.nblocks, A.spin_limit)) return;
        const double e1 = bitsf(__hip_atomic_load(&C->slot[par][att][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) + 1e-10;
        const double e2 = bitsf(__hip_atomic_load(&C->slot[par][att][1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) + 1e-10;
        const double c1 = e1 / 1e-4, c2 = e2 / 1e-4;
        const double hnew1 = h * 0.9 * pow(1e-4 / e1, 0.2);
        const double hnew2 = h * 0.9 * pow(1e-4 / e2, 0.2);
        if (py_max(c1, c2) < 1.0 && !(c1 != c1 || c2 != c2)) {
            dtnext = py_min(1.5 * h, py_min(hnew1, hnew2));
            accepted = true;
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                C->err1 = e1;
                C->err2 = e2;
            }
            break;
        }
        nrjct += 1;
        dtnext = py_min(hnew1, hnew2);
        h = py_max(0.5 * h, dtnext);
        if (h < 1.e-15 || nrjct > 20) break;
    }
    if (!accepted) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            C->status = 1;
            C->nrjct = nrjct;
            C->last_h = h;
        }
        return;
    }
    // a workgroup that timed out at a barrier left the step's error maxima
    // incomplete: no workgroup writes the post-step state then
    if (__hip_atomic_load(&C->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 2) return;
    // post-step update (QDsim_gpu.py:987-1022)
    double vnow = 0.0;
    if (act) {
        const double Tno = y5[0], Tt1 = y5[1], Tt2 = y5[2], st = y5[3];
        F[F_TNO][i] = Tno;
        F[F_TT1O][i] = Tt1;
        F[F_TT2O][i] = Tt2;
        F[F_STATE][i] = st;
        const double Tt = sqrt(Tt1 * Tt1 + Tt2 * Tt2);
        F[F_TT][i] = Tt;
        const double ex = 2.0 * A.V0 * exp(-st / e.a);
        const double v1 = ex * sinh(Tt1 / Tno / e.a);
        const double v2 = ex * sinh(Tt2 / Tno / e.a);
        F[F_SLIPV1][i] = v1;
        F[F_SLIPV2][i] = v2;
        double v = sqrt(v1 * v1 + v2 * v2);
        F[F_RAKE][i] = atan2(v2, v1);
        if (v < 1e-30) v = 1e-30;
        F[F_SLIPV][i] = v;
        vnow = v;
        const double s1 = F[F_SLIP1][i] + v1 * h;
        const double s2 = F[F_SLIP2][i] + v2 * h;
        F[F_SLIP1][i] = s1;
        F[F_SLIP2][i] = s2;
        F[F_SLIP][i] = sqrt(s1 * s1 + s2 * s2);
        F[F_FRIC][i] = Tt / (Tno - e.P * 1e-6);
    }
    block_max2_atomic(fbits(vnow), 0ull, nullptr, lds);
    if (threadIdx.x == 0) {
        // block_max2_atomic left the block max in lds[0..]; recompute cheaply
        unsigned long long mx = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mx = lds[2 * w] > mx ? lds[2 * w] : mx;
        atomicMax(&C->vmax[par], mx);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        C->time = C->time + h;
        C->last_h = h;
        C->last_dtnext = dtnext;
        C->dttry = dtnext;
        C->nrjct = nrjct;
        C->total_rejections += nrjct;
        C->steps += 1;
        if (A.records && A.rec_index >= 0) {
            A.records[4 * A.rec_index + 0] = h;
            A.records[4 * A.rec_index + 1] = dtnext;
            A.records[4 * A.rec_index + 2] = C->time;
        }
    }
}

// ---------------------------------------------------------------------------
// The reference's stepping pieces one call at a time: derivative_gpu
// (QDsim_gpu.py:830-891) at given state vectors, and one Dormand-Prince
// attempt RungeKutte_solve_Dormand_Prince_gpu(h) (QDsim_gpu.py:1128-1262)
// from the current state, both with the frozen traction rates held in the
// AdotV1, AdotV2 and dsigmadt fields.  Same arithmetic as rk_step_kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ Elem frozen_elem(double* const* F, long long i) {
    Elem e;
    e.a = F[F_A][i];
    e.b = F[F_B][i];
    e.dc = F[F_DC][i];
    e.f0 = F[F_F0][i];
    e.P = F[F_P][i];
    e.dPdt = F[F_DPDT0][i];
    e.sl = F[F_SHEAR_LOADING][i];
    e.A1 = F[F_ADOTV1][i];
    e.A2 = F[F_ADOTV2][i];
    e.dsig = F[F_DSIGMADT][i];
    e.V = F[F_SLIPV][i];
    return e;
}

// out: 4 x n = dstate/dt, dsigma/dt, dtau1/dt, dtau2/dt (MPa/s for the
// tractions), the reference's return order
__global__ __launch_bounds__(256) void rhs_eval_kernel(SimArgs A, const double* __restrict__ Tno,
                                                       const double* __restrict__ Tt1, const double* __restrict__ Tt2,
                                                       const double* __restrict__ st, double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const Elem e = frozen_elem(A.f, i);
    double k[4];
    rsf_rhs(e, A.V0, A.eta, Tno[i], Tt1[i], Tt2[i], st[i], k);
    out[i] = k[3];
    out[A.n + i] = k[0];
    out[2 * A.n + i] = k[1];
    out[3 * A.n + i] = k[2];
}

// One attempt of step size h: y5 into out (4 x n: Tno, Tt1o, Tt2o, state) and
// the two relative-error maxima into C->att_max (cleared by the caller).
__global__ __launch_bounds__(256) void dp45_attempt_kernel(SimArgs A, double h, int err_mode,
                                                           double* __restrict__ out) {
    __shared__ unsigned long long lds[16];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < A.n;
    const long long ii = act ? i : 0;
    const Elem e = frozen_elem(A.f, ii);
    double* const* F = A.f;
    const double y[4] = {F[F_TNO][ii], F[F_TT1O][ii], F[F_TT2O][ii], F[F_STATE][ii]};
    double y5[4], y4[4];
    dp45(e, A.V0, A.eta, y, h, y5, y4);
    double r1 = 0.0, r2 = 0.0;
    if (act) {
        r1 = fabs(fabs(y4[3] - y5[3]) / y4[3]);
        const double rt1 = fabs(fabs(y4[1] - y5[1]) / y4[1]);
        const double rt2 = fabs(fabs(y4[2] - y5[2]) / y4[2]);
        r2 = (rt2 > rt1 || rt2 != rt2) ? rt2 : rt1;
        if (err_mode == 0) {
            const double rn = fabs(fabs(y4[0] - y5[0]) / y4[0]);
            r2 = (rn > r2 || rn != rn) ? rn : r2;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) out[(size_t)q * A.n + i] = y5[q];
    }
    block_max2_atomic(fbits(r1), fbits(r2), A.ctrl->att_max, lds);
}

// The attempt's error measures, conditions and proposed step sizes, with the
// expressions of rk_step_kernel (QDsim_gpu.py:1221-1259); the error maxima
// also become the control block's err1 / err2 (Relerrormax1 / 2).
__global__ void dp45_attempt_finish_kernel(Ctrl* C, double h) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double e1 = bitsf(C->att_max[0]) + 1e-10;
    const double e2 = bitsf(C->att_max[1]) + 1e-10;
    C->err1 = e1;
    C->err2 = e2;
    C->att_out[0] = e1 / 1e-4;
    C->att_out[1] = e2 / 1e-4;
    C->att_out[2] = h * 0.9 * pow(1e-4 / e1, 0.2);
    C->att_out[3] = h * 0.9 * pow(1e-4 / e2, 0.2);
}

__global__ void clear_u64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ void rake_trig_kernel(const double* __restrict__ rake0, double* __restrict__ c0,
                                 double* __restrict__ s0, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        c0[i] = cos(rake0[i]);
        s0[i] = sin(rake0[i]);
    }
}

__global__ void pad_zero_kernel(double* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Host side: the dense simulation object.  All of a simulation's work runs on
// the null stream and its calls wait for that stream only, so ensembles on
// their own (non-blocking) streams keep running meanwhile.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Dilatancy: pore-pressure diffusion normal to the fault, one backward-Euler
// update per accepted step (QDsim_gpu.py:628-671).  Each element owns a
// profile of K nodes on a stretched grid (node K-1 held at P0); the K-1
// unknowns solve (I + dt P) x = b with P tridiagonal and common to all
// elements, so one workgroup-wide elimination of the matrix (LDS) serves
// every element's forward/backward sweep.  Profiles are node-major [K][n].
// ---------------------------------------------------------------------------
constexpr int DILA_MAX_NODES = 256;

struct DilaConst {
    double c_hyd, eps, hs, beta, P0;
    double qb0;     // 2 c_hyd exp(-(z0 - dz0/2)) / dz0: flux source at node 0
    double qend;    // c_hyd exp(-(z[K-2] - dz[K-2]/2)) / dz[K-2]^2: the fixed boundary
    double e_lo0, e_hi0, ez0, emz0, dz0;
    double V0;
    int K;
};

// dt_given >= 0: update with that step size (the reference's
// Calc_P_implicit(dt)); otherwise with the last accepted step, unless it failed
__global__ __launch_bounds__(256) void dila_kernel(double* __restrict__ F, long long n,
                                                   const Ctrl* __restrict__ ctrl, double* __restrict__ parr,
                                                   const double* __restrict__ tri, DilaConst c, double dt_given) {
    __shared__ double s_cp[DILA_MAX_NODES], s_inv[DILA_MAX_NODES];
    const int m = c.K - 1;
    const double dt = dt_given >= 0.0 ? dt_given : ctrl->last_h;
    const bool ok = dt_given >= 0.0 || ctrl->status == 0;
    if (threadIdx.x == 0 && ok) {
        double cp = 0.0;
        for (int i = 0; i < m; ++i) {
            const double lo = i > 0 ? dt * tri[i] : 0.0;
            const double up = i < m - 1 ? dt * tri[2 * m + i] : 0.0;
            const double den = (1.0 + dt * tri[m + i]) - lo * cp;
            s_inv[i] = 1.0 / den;
            cp = up / den;
            s_cp[i] = cp;
        }
    }
    __syncthreads();
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!ok || e >= n) return;
    const double theta = F[F_DC * n + e] / c.V0 * exp((F[F_STATE * n + e] - F[F_F0 * n + e]) / F[F_B * n + e]);
    const double dthetadt = 1.0 - F[F_SLIPV * n + e] * theta / F[F_DC * n + e];
    const double g = -c.eps * c.hs / (2.0 * c.beta * c.c_hyd * theta) * dthetadt;
    double d = 0.0;
    for (int i = 0; i < m; ++i) {
        double b = parr[(size_t)i * n + e];
        if (i == 0) b -= c.qb0 * g * dt;
        if (i == m - 1) b += dt * c.qend * c.P0;
        const double lo = i > 0 ? dt * tri[i] : 0.0;
        d = (b - lo * d) * s_inv[i];
        parr[(size_t)i * n + e] = d;
    }
    double x = d, x1 = d;
    for (int i = m - 2; i >= 0; --i) {
        x1 = x;
        x = parr[(size_t)i * n + e] - s_cp[i] * x;
        parr[(size_t)i * n + e] = x;
    }
    const double term1 = -c.e_lo0 * (x - x1 + 2 * c.dz0 * g * c.ez0);
    const double term2 = c.e_hi0 * (x1 - x);
    F[F_DPDT0 * n + e] = c.c_hyd * c.emz0 * (term1 + term2) / (c.dz0 * c.dz0);
    F[F_P * n + e] = x;
}

// H-matrix traction rates (hmat_mvm.hip): out = 3 x n element-order sums
void hmat_launch(void* hm, const double* V1, const double* V2, const double* c0, const double* s0, double Vpl,
                 int fix_tn, double* out, Launch& L);
long long hmat_n(void* hm);

struct DenseSim {
    void* hmat = nullptr;          // attached H-matrix (replaces the dense matrices)
    int dila_on = 0;               // pore-pressure update after every accepted step
    double* dila_parr = nullptr;   // K x n profiles (Pa)
    double* dila_tri = nullptr;    // 3 x (K-1): lower, diagonal, upper of the operator
    DilaConst dila;
    long long n = 0, ld = 0;
    int mat_type = MAT_F32;
    void* mats[6] = {nullptr};
    double* fields = nullptr;      // NFIELDS x n
    double* partial = nullptr;
    Ctrl* ctrl = nullptr;
    int kc = 256, nchunk = 0;
    int rows_per_block = 0;
    int nblocks = 0;
    double V0 = 1e-6, Vpl = 1e-9, eta = 0.0;
    int fix_tn = 0, err_mode = 1;
    int zero_mask = 0;             // bit k: matrix k holds only zeros (from the last scan)
    int skip_normal = 0;           // Bs and Bd are zero: the product skips them
    long long spin_limit = 20000000;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    double* d_records = nullptr;
    long long rec_cap = 0;
};

static double host_bits_to_double(unsigned long long u) {
    double d;
    memcpy(&d, &u, sizeof(d));
    return d;
}

static int vec_n(int mat_type) { return mat_type == MAT_F64 ? 2 : (mat_type == MAT_F32 ? 4 : 8); }

static SimArgs make_args(DenseSim* s, double* records, long long rec_index) {
    SimArgs a;
    for (int k = 0; k < NFIELDS; ++k) a.f[k] = s->fields + (size_t)k * s->n;
    a.partial = s->partial;
    a.ctrl = s->ctrl;
    a.records = records;
    a.rec_index = rec_index;
    a.n = s->n;
    a.nchunk = s->nchunk;
    a.nblocks = s->nblocks;
    a.V0 = s->V0;
    a.Vpl = s->Vpl;
    a.eta = s->eta;
    a.fix_tn = s->fix_tn;
    a.err_mode = s->err_mode;
    a.spin_limit = s->spin_limit;
    return a;
}

template <class T>
static void launch_matvec_t(DenseSim* s, Launch& L) {
    MatPtrs<T> M;
    for (int k = 0; k < 6; ++k) M.m[k] = (const T*)s->mats[k];
    const int VN = Vec16<T>::n;
    dim3 grid((unsigned)((s->n + 256 * VN - 1) / (256 * VN)), (unsigned)s->nchunk);
    double* F = s->fields;
    const long long n = s->n;
    // with Bs = Bd = 0 the normal partials are exactly +0 either way
    if (s->fix_tn || s->skip_normal)
        L(matvec_kernel<T, true>, grid, dim3(256), M, s->ld, n, s->kc, F + F_SLIPV1 * n, F + F_SLIPV2 * n,
          F + F_COSR0 * n, F + F_SINR0 * n, s->Vpl, s->partial);
    else
        L(matvec_kernel<T, false>, grid, dim3(256), M, s->ld, n, s->kc, F + F_SLIPV1 * n, F + F_SLIPV2 * n,
          F + F_COSR0 * n, F + F_SINR0 * n, s->Vpl, s->partial);
}

static void launch_matvec(DenseSim* s, Launch& L) {
    if (s->hmat) {
        double* F = s->fields;
        const long long n = s->n;
        hmat_launch(s->hmat, F + F_SLIPV1 * n, F + F_SLIPV2 * n, F + F_COSR0 * n, F + F_SINR0 * n, s->Vpl,
                    s->fix_tn, s->partial, L);
        return;
    }
    if (s->mat_type == MAT_F64) launch_matvec_t<double>(s, L);
    else if (s->mat_type == MAT_F32) launch_matvec_t<float>(s, L);
    else launch_matvec_t<unsigned short>(s, L);
}

static void launch_step(DenseSim* s, Launch& L, double* records, long long rec_index) {
    launch_matvec(s, L);
    SimArgs a = make_args(s, records, rec_index);
    L(rk_step_kernel, dim3((unsigned)s->nblocks), dim3(256), a);
    if (s->dila_on)
        L(dila_kernel, dim3((unsigned)((s->n + 255) / 256)), dim3(256), s->fields, s->n, (const Ctrl*)s->ctrl,
          s->dila_parr, (const double*)s->dila_tri, s->dila, -1.0);
}

static int configure(DenseSim* s) {
    int dev = 0;
    PQ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PQ_HIP(hipGetDeviceProperties(&p, dev));
    int per_cu = 0;
    PQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rk_step_kernel, 256, 0));
    const long long need = (s->n + 255) / 256;
    const long long cap = (long long)per_cu * p.multiProcessorCount;
    PQ_CHECK(need <= cap, "dense sim: n=%lld needs %lld co-resident workgroups, device holds %lld", s->n, need,
             cap);
    s->nblocks = (int)need;
    return 0;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

int pq3d_dense_create(void** out, long long n, int mat_type, int kc) {
    clear_error();
    PQ_CHECK(n > 0, "pq3d_dense_create: n must be positive");
    PQ_CHECK(mat_type == MAT_F64 || mat_type == MAT_F32 || mat_type == MAT_BF16 || mat_type == MAT_NONE,
             "bad matrix type %d", mat_type);
    DenseSim* s = new DenseSim();
    s->n = n;
    s->mat_type = mat_type;
    const int VN = mat_type == MAT_NONE ? 1 : vec_n(mat_type);
    s->ld = (n + VN - 1) / VN * VN;
    s->kc = mat_type == MAT_NONE ? (int)std::min<long long>(n, 1 << 30) : (kc > 0 ? kc : 256);
    s->nchunk = (int)((n + s->kc - 1) / s->kc);
    int rc = configure(s);
    if (rc) {
        delete s;
        return rc;
    }
    const size_t mbytes = (size_t)s->ld * (size_t)n * mat_elem_bytes(mat_type);
    for (int k = 0; k < 6 && mbytes; ++k) {
        hipError_t e = hipMalloc(&s->mats[k], mbytes);
        if (e != hipSuccess) {
            set_error("pq3d_dense_create: matrix allocation of %zu bytes failed: %s", mbytes, hipGetErrorString(e));
            for (int j = 0; j < k; ++j) (void)hipFree(s->mats[j]);
            delete s;
            return -1;
        }
        (void)hipMemset(s->mats[k], 0, mbytes);
    }
    PQ_HIP(hipMalloc((void**)&s->fields, (size_t)NFIELDS * n * sizeof(double)));
    PQ_HIP(hipMemset(s->fields, 0, (size_t)NFIELDS * n * sizeof(double)));
    PQ_HIP(hipMalloc((void**)&s->partial, (size_t)s->nchunk * 3 * n * sizeof(double)));
    PQ_HIP(hipMemset(s->partial, 0, (size_t)s->nchunk * 3 * n * sizeof(double)));
    PQ_HIP(hipMalloc((void**)&s->ctrl, sizeof(Ctrl)));
    PQ_HIP(hipMemset(s->ctrl, 0, sizeof(Ctrl)));
    *out = s;
    return 0;
}

int pq3d_dense_destroy(void* h) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    if (!s) return 0;
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    for (int k = 0; k < 6; ++k) (void)hipFree(s->mats[k]);
    (void)hipFree(s->fields);
    (void)hipFree(s->partial);
    (void)hipFree(s->ctrl);
    (void)hipFree(s->d_records);
    (void)hipFree(s->dila_parr);
    (void)hipFree(s->dila_tri);
    delete s;
    return 0;
}

// Dilatancy / pore-pressure diffusion (QDsim_gpu.py:336-349, :611-671).
// zp: K node coordinates of the stretched grid; parr: n x K initial profiles
// (Pa, row per element, the reference's Parr layout); c_hyd hydraulic
// diffusivity, eps dilatancy coefficient, hs shear-zone thickness, beta
// effective compressibility, P0 fixed far-field pressure (Pa), V0 reference
// slip rate.  From the next step on, every accepted step updates the
// profiles, P and dP/dt.
int pq3d_dense_set_dilatancy(void* h, int K, const double* zp, const double* parr, double c_hyd, double eps,
                             double hs, double beta, double P0, double V0) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && K >= 3 && K <= DILA_MAX_NODES + 1, "pq3d_dense_set_dilatancy: K must be in [3, %d]",
             DILA_MAX_NODES + 1);
    const int m = K - 1;
    std::vector<double> dz(m), tri(3 * (size_t)m, 0.0);
    for (int i = 0; i < m; ++i) dz[i] = zp[i + 1] - zp[i];
    for (int i = 0; i < m; ++i) {
        const double lo = std::exp(-(2 * zp[i] - dz[i] / 2)), hi = std::exp(-(2 * zp[i] + dz[i] / 2));
        const double w = c_hyd / (dz[i] * dz[i]);
        tri[m + i] = w * (lo + hi);                              // diagonal
        if (i == 0) {
            tri[2 * m + 0] = -w * (lo + hi);                     // row 0 couples node 1 with both weights
        } else {
            tri[i] = -w * lo;                                    // lower
            if (i < m - 1) tri[2 * m + i] = -w * hi;             // upper (node K-1 is the fixed boundary)
        }
    }
    DilaConst c;
    c.c_hyd = c_hyd;
    c.eps = eps;
    c.hs = hs;
    c.beta = beta;
    c.P0 = P0;
    c.dz0 = dz[0];
    c.qb0 = 2.0 * c_hyd * std::exp(-(zp[0] - dz[0] / 2)) / dz[0];
    c.qend = c_hyd * std::exp(-(zp[m - 1] - dz[m - 1] / 2)) / (dz[m - 1] * dz[m - 1]);
    c.e_lo0 = std::exp(-(zp[0] - dz[0] / 2));
    c.e_hi0 = std::exp(-(zp[0] + dz[0] / 2));
    c.ez0 = std::exp(zp[0]);
    c.emz0 = std::exp(-zp[0]);
    c.V0 = V0;
    c.K = K;
    const long long n = s->n;
    std::vector<double> nm((size_t)K * n);
    for (long long e = 0; e < n; ++e)
        for (int k = 0; k < K; ++k) nm[(size_t)k * n + e] = parr[(size_t)e * K + k];
    if (s->dila_parr) (void)hipFree(s->dila_parr);
    if (s->dila_tri) (void)hipFree(s->dila_tri);
    s->dila_parr = nullptr;
    s->dila_tri = nullptr;
    PQ_HIP(hipMalloc((void**)&s->dila_parr, nm.size() * sizeof(double)));
    PQ_HIP(hipMalloc((void**)&s->dila_tri, tri.size() * sizeof(double)));
    PQ_HIP(hipMemcpy(s->dila_parr, nm.data(), nm.size() * sizeof(double), hipMemcpyHostToDevice));
    PQ_HIP(hipMemcpy(s->dila_tri, tri.data(), tri.size() * sizeof(double), hipMemcpyHostToDevice));
    s->dila = c;
    s->dila_on = 1;
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return 0;
}

// One pore-pressure update with step size dt from the current state and slip
// rates (the reference's Calc_P_implicit(dt), QDsim_gpu.py:628-671).
int pq3d_dense_dilatancy_update(void* h, double dt) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && s->dila_on, "pq3d_dense_dilatancy_update: dilatancy is not enabled");
    PQ_CHECK(dt >= 0.0, "pq3d_dense_dilatancy_update: dt must be >= 0");
    hipLaunchKernelGGL(dila_kernel, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0, 0, s->fields, s->n,
                       (const Ctrl*)s->ctrl, s->dila_parr, (const double*)s->dila_tri, s->dila, dt);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipStreamSynchronize(0));
    return 0;
}

// Current pore-pressure profiles into parr (n x K, row per element).
int pq3d_dense_get_dilatancy(void* h, double* parr) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && s->dila_on, "pq3d_dense_get_dilatancy: dilatancy is not enabled");
    const int K = s->dila.K;
    const long long n = s->n;
    std::vector<double> nm((size_t)K * n);
    PQ_HIP(hipMemcpy(nm.data(), s->dila_parr, nm.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long long e = 0; e < n; ++e)
        for (int k = 0; k < K; ++k) parr[(size_t)e * K + k] = nm[(size_t)k * n + e];
    return 0;
}

// Take the traction rates from an H-matrix (pq3d_hmat_create) instead of the
// dense matrices; the H-matrix must outlive the simulation's use of it.
int pq3d_dense_attach_hmatrix(void* h, void* hm) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && hm, "pq3d_dense_attach_hmatrix: null argument");
    PQ_CHECK(hmat_n(hm) == s->n, "pq3d_dense_attach_hmatrix: H-matrix has n=%lld, simulation n=%lld", hmat_n(hm),
             s->n);
    s->hmat = hm;
    s->nchunk = 1;
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return 0;
}

// Device memory the simulation holds: matrices, fields, partial sums,
// control block, step records and pore-pressure profiles.
int pq3d_dense_device_bytes(void* h, double* bytes) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && bytes, "pq3d_dense_device_bytes: null argument");
    double b = 0;
    for (int k = 0; k < 6; ++k)
        if (s->mats[k]) b += (double)s->ld * s->n * mat_elem_bytes(s->mat_type);
    b += (double)NFIELDS * s->n * sizeof(double) + (double)s->nchunk * 3 * s->n * sizeof(double) + sizeof(Ctrl);
    b += (double)s->rec_cap * 4 * sizeof(double);
    if (s->dila_on) b += (double)s->dila.K * s->n * sizeof(double) + 3.0 * (s->dila.K - 1) * sizeof(double);
    *bytes = b;
    return 0;
}

int pq3d_dense_matrix_ptrs(void* h, void** ptrs, long long* ld) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    for (int k = 0; k < 6; ++k) ptrs[k] = s->mats[k];
    *ld = s->ld;
    return 0;
}

int pq3d_dense_set_field(void* h, int field, const double* host) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(field >= 0 && field < NFIELDS, "bad field %d", field);
    double* dst = s->fields + (size_t)field * s->n;
    PQ_HIP(hipMemcpy(dst, host, s->n * sizeof(double), hipMemcpyHostToDevice));
    if (field == F_RAKE0) {
        hipLaunchKernelGGL(rake_trig_kernel, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0, 0, dst,
                           s->fields + (size_t)F_COSR0 * s->n, s->fields + (size_t)F_SINR0 * s->n, s->n);
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipStreamSynchronize(0));
    }
    return 0;
}

int pq3d_dense_get_field(void* h, int field, double* host) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(field >= 0 && field < NFIELDS, "bad field %d", field);
    PQ_HIP(hipMemcpy(host, s->fields + (size_t)field * s->n, s->n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int pq3d_dense_field_ptr(void* h, int field, void** dptr) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(field >= 0 && field < NFIELDS, "bad field %d", field);
    *dptr = s->fields + (size_t)field * s->n;
    return 0;
}

// V0, Vpl, eta = mu/(2 Cs), fix_Tn, err_mode (1: GPU-path error estimate)
int pq3d_dense_set_params(void* h, double V0, double Vpl, double eta, int fix_tn, int err_mode) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    s->V0 = V0;
    s->Vpl = Vpl;
    s->eta = eta;
    s->fix_tn = fix_tn;
    s->err_mode = err_mode;
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return 0;
}

int pq3d_dense_set_time(void* h, double t) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_HIP(hipMemcpy(&s->ctrl->time, &t, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// Read the control block: time, next dt, last h, last dtnext, max slip rate of
// the last step, rejections of the last step, status, accepted steps.
int pq3d_dense_state(void* h, double* out6, int* nrjct, int* status, long long* steps) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    Ctrl c;
    PQ_HIP(hipMemcpy(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    out6[0] = c.time;
    out6[1] = c.dttry;
    out6[2] = c.last_h;
    out6[3] = c.last_dtnext;
    const int par = (int)((c.steps + 1) & 1);  // parity of the last accepted step
    out6[4] = host_bits_to_double(c.vmax[par]);
    out6[5] = c.err2;
    *nrjct = c.nrjct;
    *status = c.status;
    *steps = c.steps;
    return 0;
}

// Relative-error measures of the last accepted step, or of the last
// pq3d_dense_dp45 attempt if one came after it (Relerrormax1 / Relerrormax2,
// QDsim_gpu.py:1221-1238: maxima + 1e-10).
int pq3d_dense_errors(void* h, double* err1, double* err2) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && err1 && err2, "pq3d_dense_errors: null argument");
    Ctrl c;
    PQ_HIP(hipMemcpy(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    *err1 = c.err1;
    *err2 = c.err2;
    return 0;
}

// derivative_gpu(Tno, Tt1o, Tt2o, state) (QDsim_gpu.py:830-891): the RHS at
// the given device vectors (n each; MPa, MPa, MPa, dimensionless) with the
// frozen traction rates of the AdotV1 / AdotV2 / dsigmadt fields and the
// current slip rate; d_out (4 x n) receives dstate/dt, dsigma/dt, dtau1/dt,
// dtau2/dt (MPa/s).
int pq3d_dense_derivative(void* h, const double* d_Tno, const double* d_Tt1o, const double* d_Tt2o,
                          const double* d_state, double* d_out) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && d_Tno && d_Tt1o && d_Tt2o && d_state && d_out, "pq3d_dense_derivative: null argument");
    SimArgs a = make_args(s, nullptr, -1);
    hipLaunchKernelGGL(rhs_eval_kernel, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0, 0, a, d_Tno, d_Tt1o,
                       d_Tt2o, d_state, d_out);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipStreamSynchronize(0));
    return 0;
}

// RungeKutte_solve_Dormand_Prince_gpu(h) (QDsim_gpu.py:1128-1262): one
// attempt of step size h from the current state with the frozen traction
// rates; nothing is committed.  d_out (4 x n): the fifth-order solution
// Tno, Tt1o, Tt2o, state.  out4: condition1, condition2, hnew1, hnew2.
// err_mode 1 drops the normal-traction error (the GPU method), 0 keeps it
// (the CPU method, QDsim_gpu.py:1373).  err1 / err2 of the control block
// become this attempt's (pq3d_dense_errors).
int pq3d_dense_dp45(void* h, double step, int err_mode, double* d_out, double* out4) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && d_out && out4, "pq3d_dense_dp45: null argument");
    SimArgs a = make_args(s, nullptr, -1);
    hipLaunchKernelGGL(clear_u64_kernel, dim3(1), dim3(64), 0, 0, s->ctrl->att_max, 2);
    hipLaunchKernelGGL(dp45_attempt_kernel, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0, 0, a, step,
                       err_mode, d_out);
    hipLaunchKernelGGL(dp45_attempt_finish_kernel, dim3(1), dim3(64), 0, 0, s->ctrl, step);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipMemcpy(out4, s->ctrl->att_out, 4 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// Frozen traction rates from the current slip rates (matvec + reduction only):
// fills the AdotV1, AdotV2, dsigmadt fields without stepping.
__global__ void reduce_partials_kernel(const double* __restrict__ partial, int nchunk, long long n,
                                       const double* __restrict__ nl, int fix_tn, double* __restrict__ a1,
                                       double* __restrict__ a2, double* __restrict__ dsg) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = 0, y = 0, z = 0;
    for (int c = 0; c < nchunk; ++c) {
        x += partial[((size_t)c * 3 + 0) * n + i];
        y += partial[((size_t)c * 3 + 1) * n + i];
        z += partial[((size_t)c * 3 + 2) * n + i];
    }
    a1[i] = x;
    a2[i] = y;
    dsg[i] = fix_tn ? nl[i] : z + nl[i];
}

// The captured step graph holds the matvec variant, so a change drops it.
static void set_structure(DenseSim* s, int mask) {
    const int skip = (mask & 0x24) == 0x24;
    if (skip != s->skip_normal && s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    s->zero_mask = mask;
    s->skip_normal = skip;
}

// Structure of the dense matrices: scan them (reads all six once) and let the
// product skip Bs and Bd when both hold only zeros (a planar fault: the
// normal traction of in-plane slip vanishes by mirror symmetry).  The results
// are bit-identical to the full product.  mask_out (optional): bit k set iff
// matrix k is all zero.
int pq3d_dense_scan_structure(void* h, int* mask_out) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s, "pq3d_dense_scan_structure: null handle");
    int mask = 0;
    if (s->mat_type != MAT_NONE) {
        PQ_HIP(hipStreamSynchronize(0));
        int rc = scan_zero_matrices(s->mats, 6, s->mat_type, s->ld, s->n, &mask);
        if (rc) return rc;
    }
    set_structure(s, mask);
    if (mask_out) *mask_out = mask;
    return 0;
}

// Override the structure (mask as from the scan; 0 = stream all six matrices).
// The caller vouches that every matrix named in mask holds only zeros.
int pq3d_dense_set_structure(void* h, int mask) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && mask >= 0 && mask < 64, "pq3d_dense_set_structure: bad argument");
    set_structure(s, mask);
    return 0;
}

int pq3d_dense_get_structure(void* h, int* mask, int* skip_normal) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && mask && skip_normal, "pq3d_dense_get_structure: null argument");
    *mask = s->zero_mask;
    *skip_normal = s->skip_normal;
    return 0;
}

// The traction product as the step consumes it: split-K partials, then their
// fixed-order sum into the traction-rate fields.
static void launch_product(DenseSim* s, Launch& L) {
    launch_matvec(s, L);
    const long long n = s->n;
    double* F = s->fields;
    L(reduce_partials_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), s->partial, s->nchunk, n,
      F + F_NORMAL_LOADING * n, s->fix_tn, F + F_ADOTV1 * n, F + F_ADOTV2 * n, F + F_DSIGMADT * n);
}

int pq3d_dense_matvec(void* h) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s->mat_type != MAT_NONE || s->hmat, "dense sim: no matrices and no H-matrix attached");
    Launch L;
    launch_product(s, L);
    PQ_HIP(L.err);
    PQ_HIP(hipStreamSynchronize(0));
    return 0;
}

// Re-tile the dense traction product with kc source columns per chunk (the
// split-K width of matvec_kernel).  Waits for the device, drops the captured
// step graph and reallocates the partial sums.  Not for a simulation that
// runs on an attached H-matrix or has no matrices.
int pq3d_dense_set_kc(void* h, int kc) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s, "pq3d_dense_set_kc: null handle");
    PQ_CHECK(s->mat_type != MAT_NONE && !s->hmat, "pq3d_dense_set_kc: simulation has no dense matrices");
    PQ_CHECK(kc >= 1, "pq3d_dense_set_kc: kc must be positive (got %d)", kc);
    const int kcn = (int)std::min<long long>(kc, s->n);
    const int nchunk = (int)((s->n + kcn - 1) / kcn);
    PQ_HIP(hipStreamSynchronize(0));
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    double* p = nullptr;
    const size_t bytes = (size_t)nchunk * 3 * s->n * sizeof(double);
    PQ_HIP(hipMalloc((void**)&p, bytes));
    PQ_HIP(hipMemset(p, 0, bytes));
    (void)hipFree(s->partial);
    s->partial = p;
    s->kc = kcn;
    s->nchunk = nchunk;
    return 0;
}

int pq3d_dense_get_kc(void* h, int* kc, int* nchunk) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s && kc && nchunk, "pq3d_dense_get_kc: null argument");
    *kc = s->kc;
    *nchunk = s->nchunk;
    return 0;
}

// Time the product with its partial-sum reduction (what a split-K width
// costs end to end): `reps` launches between two events, ms per product.
// Writes the same traction rates pq3d_dense_matvec does.
int pq3d_dense_time_product(void* h, int reps, float* ms_per) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s->mat_type != MAT_NONE || s->hmat, "dense sim: no matrices and no H-matrix attached");
    PQ_CHECK(reps > 0 && ms_per, "pq3d_dense_time_product: reps must be positive");
    hipEvent_t a, b;
    PQ_HIP(hipEventCreate(&a));
    PQ_HIP(hipEventCreate(&b));
    Launch L;
    launch_product(s, L);
    PQ_HIP(hipEventRecord(a, 0));
    for (int r = 0; r < reps; ++r) launch_product(s, L);
    PQ_HIP(hipEventRecord(b, 0));
    PQ_HIP(L.err);
    PQ_HIP(hipEventSynchronize(b));
    float ms = 0;
    PQ_HIP(hipEventElapsedTime(&ms, a, b));
    *ms_per = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 0;
}

// Time only the matvec: `reps` launches between two events; returns ms per launch.
int pq3d_dense_time_matvec(void* h, int reps, float* ms_per) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s->mat_type != MAT_NONE || s->hmat, "dense sim: no matrices and no H-matrix attached");
    hipEvent_t a, b;
    PQ_HIP(hipEventCreate(&a));
    PQ_HIP(hipEventCreate(&b));
    Launch L;
    launch_matvec(s, L);
    PQ_HIP(hipEventRecord(a, 0));
    for (int r = 0; r < reps; ++r) launch_matvec(s, L);
    PQ_HIP(hipEventRecord(b, 0));
    PQ_HIP(L.err);
    PQ_HIP(hipEventSynchronize(b));
    float ms = 0;
    PQ_HIP(hipEventElapsedTime(&ms, a, b));
    *ms_per = ms / reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 0;
}

// One step (drop-in simu_forwardGPU): returns the accepted h and dtnext.
// status 1 reproduces the reference's 'dt is too small' exit condition.
int pq3d_dense_step(void* h, double dttry, double* h_out, double* dtnext_out, int* nrjct_out) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s->mat_type != MAT_NONE || s->hmat, "dense sim: no matrices and no H-matrix attached");
    PQ_HIP(hipMemcpy(&s->ctrl->dttry, &dttry, sizeof(double), hipMemcpyHostToDevice));
    Launch L;
    launch_step(s, L, nullptr, -1);
    PQ_HIP(L.err);
    PQ_HIP(hipStreamSynchronize(0));
    Ctrl c;
    PQ_HIP(hipMemcpy(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    *h_out = c.last_h;
    *dtnext_out = c.last_dtnext;
    *nrjct_out = c.nrjct;
    if (c.status == 1) {
        set_error("error: dt is too small");
        return 1;
    }
    PQ_CHECK(c.status == 0, "dense step: device status %d (grid barrier timeout)", c.status);
    return 0;
}

// nsteps device-resident steps, no host synchronisation between them; the
// step sequence is captured once in a HIP graph of `graph_steps` steps.
// records (host, optional): nsteps x 4 = h, dtnext, time, max slip rate.
int pq3d_dense_run(void* h, long long nsteps, int graph_steps, double* records, int* status) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s->mat_type != MAT_NONE || s->hmat, "dense sim: no matrices and no H-matrix attached");
    if (graph_steps <= 0) graph_steps = 1;
    if (nsteps <= 0) return 0;
    if (records && s->rec_cap < nsteps) {
        if (s->d_records) (void)hipFree(s->d_records);
        PQ_HIP(hipMalloc((void**)&s->d_records, (size_t)nsteps * 4 * sizeof(double)));
        s->rec_cap = nsteps;
        if (s->graph) {
            (void)hipGraphExecDestroy(s->graph);
            s->graph = nullptr;
        }
    }
    // records are indexed per launch; the graph path writes no records
    long long steps0 = 0;
    if (records) {
        Ctrl c0;
        PQ_HIP(hipMemcpy(&c0, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
        steps0 = c0.steps;
    }
    long long done = 0;
    hipStream_t st = 0;
    if (!records) {
        if (!s->graph || s->graph_steps != graph_steps) {
            if (s->graph) (void)hipGraphExecDestroy(s->graph);
            s->graph = nullptr;
            hipGraph_t g;
            PQ_HIP(hipGraphCreate(&g, 0));
            Launch G(st, g);
            for (int k = 0; k < graph_steps; ++k) launch_step(s, G, nullptr, -1);
            hipError_t e = G.err;
            if (e == hipSuccess) e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PQ_HIP(e);
            s->graph_steps = graph_steps;
        }
        for (; done + graph_steps <= nsteps; done += graph_steps) PQ_HIP(hipGraphLaunch(s->graph, st));
    }
    Launch L(st);
    for (; done < nsteps; ++done) launch_step(s, L, records ? s->d_records : nullptr, records ? done : -1);
    PQ_HIP(L.err);
    PQ_HIP(hipStreamSynchronize(0));
    Ctrl c;
    PQ_HIP(hipMemcpy(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    if (records) {
        PQ_HIP(hipMemcpy(records, s->d_records, (size_t)nsteps * 4 * sizeof(double), hipMemcpyDeviceToHost));
        const long long acc = c.steps - steps0;  // rows written by accepted steps
        if (c.status == 0) {
            const int par = (int)((c.steps + 1) & 1);
            records[4 * (nsteps - 1) + 3] = host_bits_to_double(c.vmax[par]);
        } else {
            // the failing launch logged the last accepted step's max slip
            // rate; its row and the later ones were never written
            for (long long r = acc > 0 ? acc : 0; r < nsteps; ++r)
                for (int k = 0; k < 4; ++k) records[4 * r + k] = std::nan("");
        }
    }
    *status = c.status;
    return 0;
}

// Clear a failure (status, the grid barrier and the reduction slots) so that
// stepping can resume; time, step size, counters and the last step's
// measures are kept.
int pq3d_dense_reset_ctrl(void* h) {
    clear_error();
    DenseSim* s = (DenseSim*)h;
    PQ_CHECK(s, "pq3d_dense_reset_ctrl: null handle");
    Ctrl c;
    PQ_HIP(hipMemcpy(&c, s->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    c.status = 0;
    c.nrjct = 0;
    c.bar_count = 0;
    c.bar_gen = 0;
    memset(c.slot, 0, sizeof(c.slot));
    memset(c.att_max, 0, sizeof(c.att_max));
    PQ_HIP(hipMemcpy(s->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice));
    return 0;
}

}  // extern "C"
