// Ensemble mode: B simulations ("members") with their own friction parameters
// and state on one shared mesh and one shared set of traction Green's
// functions.  There is no reference counterpart: the reference runs
// independent QDsim_gpu instances (main_gpu.py:106-187), each streaming the
// six matrices once per step.
//
// Per global step:
//   1. ens_prep_u_kernel:  u = V - Vpl (cos rake0, sin rake0) for every member,
//      as FP32 [source][member] (members contiguous).
//   2. ens_gemm_kernel:    Y_q = A_q^s U1 + A_q^d U2 for q in {1, 2, n}: one
//      tall-skinny GEMM (N x 2N by 2N x B) per traction component on the MFMA
//      units (v_mfma_f32_16x16x4_f32, FP32 in, FP32 accumulate).  A wave owns
//      64 receivers x 64 members x 3 components (192 accumulators); lane l
//      feeds the four 16-row MFMA subtiles from ONE 16-byte load of the
//      source-major matrices (row = 4*(l%16) + subtile) and the four 16-member
//      subtiles from one 16-byte load of u (member = 4*(l%16) + subtile).
//      Split-K partials per column chunk are summed in a fixed order.
//   3. ens_rk_kernel:      persistent and masked.  Every member takes exactly
//      one accepted Dormand-Prince step with its own h; members that accept
//      drop out of the attempt rounds while the others retry (retries reuse
//      the step's GEMM, as the reference freezes AdotV per step).  Each round:
//      attempts for all pending (member, element) items -> per-member error
//      maxima (wave shuffles + one atomic per wave) -> ONE grid barrier -> every
//      workgroup applies the reference's accept/reject rule to the same
//      reduced errors in its own LDS copy of the member control -> accepted
//      members' items write the post-step state.
#include "common.hpp"

#include <cmath>
#include <cstdint>

namespace pq3d {

typedef float v4f __attribute__((ext_vector_type(4)));

enum EField : int {
    E_TNO = 0, E_TT1O, E_TT2O, E_STATE, E_SLIPV1, E_SLIPV2, E_SLIPV, E_SLIP1, E_SLIP2, E_SLIP, E_TT, E_RAKE,
    E_RAKE0, E_FRIC, E_A, E_B, E_DC, E_F0, E_SHEAR_LOADING, E_NORMAL_LOADING, E_P, E_DPDT0,
    E_ADOTV1, E_ADOTV2, E_DSIGMADT, E_COSR0, E_SINR0, E_Y5_0, E_Y5_1, E_Y5_2, E_Y5_3, E_NFIELDS
};

struct MemberCtrl {
    double h;          // current attempt's step size
    double dttry;      // step size to try at the next global step
    double time;
    double dtnext;
    double vmax;
    int nrjct;
    int done;          // accepted in this global step
    int status;        // 0 ok, 1 dt too small / >20 rejections
    int active;        // masked out when 0
    long long steps;
};

struct EnsCtrl {
    unsigned bar_count, bar_gen;
    int status;        // 2: barrier timeout
    int pad;
    long long gsteps;
};

struct EnsArgs {
    double* f;                 // E_NFIELDS x B x n
    const float* partial;      // nchunk x 3 x B x npad
    MemberCtrl* mc;            // B
    unsigned long long* slot;  // 2 x 24 x B x 2 (parity, attempt, member, {e1, e2})
    EnsCtrl* ctrl;
    long long n, npad;
    int B, nchunk, nblocks;
    double V0, Vpl, eta;
    long long spin_limit;
};

// ---------------------------------------------------------------------------
__global__ void ens_prep_u_kernel(const double* __restrict__ V1, const double* __restrict__ V2,
                                  const double* __restrict__ c0, const double* __restrict__ s0, double Vpl,
                                  long long n, int B, const MemberCtrl* __restrict__ mc, float* __restrict__ U1,
                                  float* __restrict__ U2) {
    // thread per (source j, member m); writes U[j*B + m]
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * B) return;
    const long long j = t / B;
    const int m = (int)(t - j * B);
    const size_t idx = (size_t)m * n + j;
    float u1 = 0.f, u2 = 0.f;
    if (mc[m].active) {
        u1 = (float)__dsub_rn(V1[idx], __dmul_rn(Vpl, c0[idx]));
        u2 = (float)__dsub_rn(V2[idx], __dmul_rn(Vpl, s0[idx]));
    }
    U1[t] = u1;
    U2[t] = u2;
}

// grid: (ceil(n / 64), nchunk); block 64 = one wave owning 64 rows x 64 members.
// Requires B == 64 and ld % 4 == 0 (rows padded, zero-filled).
__global__ __launch_bounds__(64) void ens_gemm_kernel(const float* __restrict__ A1s, const float* __restrict__ A2s,
                                                       const float* __restrict__ Bs, const float* __restrict__ A1d,
                                                       const float* __restrict__ A2d, const float* __restrict__ Bd,
                                                       long long ld, long long n, int kc,
                                                       const float* __restrict__ U1, const float* __restrict__ U2,
                                                       float* __restrict__ partial, long long npad) {
    const int lane = threadIdx.x & 63;
    const long long r0 = (long long)blockIdx.x * 64;  // this wave's first row
    const long long k0 = (long long)blockIdx.y * kc;
    long long k1 = k0 + kc;
    if (k1 > n) k1 = n;
    const int lr = lane & 15;   // row / member group within a 16-lane block
    const int lk = lane >> 4;   // k within the 4-deep MFMA step
    const long long row = r0 + 4 * lr;
    const bool row_ok = row < ld;  // ld is a multiple of 4: the float4 stays inside the column
    const float* mats[6] = {A1s, A2s, Bs, A1d, A2d, Bd};
    v4f acc[3][4][4];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    for (long long k = k0; k < k1; k += 4) {
        const long long s = k + lk;
        const bool s_ok = s < k1;
        float4 a[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            a[m] = (s_ok && row_ok) ? *reinterpret_cast<const float4*>(mats[m] + (size_t)s * ld + row)
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const float4 u1 = s_ok ? *reinterpret_cast<const float4*>(U1 + (size_t)s * 64 + 4 * lr)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 u2 = s_ok ? *reinterpret_cast<const float4*>(U2 + (size_t)s * 64 + 4 * lr)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
        const float uu1[4] = {u1.x, u1.y, u1.z, u1.w};
        const float uu2[4] = {u2.x, u2.y, u2.z, u2.w};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float as[4] = {a[q].x, a[q].y, a[q].z, a[q].w};
            const float ad[4] = {a[q + 3].x, a[q + 3].y, a[q + 3].z, a[q + 3].w};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[i], uu1[j], acc[q][i][j], 0, 0, 0);
                    acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ad[i], uu2[j], acc[q][i][j], 0, 0, 0);
                }
        }
    }
    // C fragment: lane holds tile rows 4*(lane/16) + v and tile column lane%16.
    // Tile (i, j) row t  -> physical row r0 + 4*t + i; tile column c -> member 4*c + j.
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int member = 4 * lr + j;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long prow = r0 + 4 * (4 * lk + v);
                if (prow < npad) {
                    float4 o = make_float4(acc[q][0][j][v], acc[q][1][j][v], acc[q][2][j][v], acc[q][3][j][v]);
                    *reinterpret_cast<float4*>(partial + (((size_t)blockIdx.y * 3 + q) * 64 + member) * npad + prow) = o;
                }
            }
        }
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ double eclip(double x) { return x < -700.0 ? -700.0 : (x > 700.0 ? 700.0 : x); }

struct EElem {
    double a, b, dc, f0, P, dPdt, sl, A1, A2, dsig, V;
};

__device__ __forceinline__ void ens_rhs(const EElem& e, double V0, double eta, const double* y, double* k) {
    const double tn = y[0] * 1e6, t1 = y[1] * 1e6, t2 = y[2] * 1e6, st = y[3];
    const double sg = tn - e.P;
    const double as = e.a * sg;
    const double ex = exp(eclip(-st / e.a));
    const double c1 = eclip(t1 / as), c2 = eclip(t2 / as);
    const double ep1 = exp(c1), em1 = exp(-c1), ep2 = exp(c2), em2 = exp(-c2);
    const double ch1 = (ep1 + em1) / 2, ch2 = (ep2 + em2) / 2;
    const double sh1 = (ep1 - em1) / 2, sh2 = (ep2 - em2) / 2;
    const double q = 2 * V0 / as;
    const double dV1dtau = q * ex * ch1, dV2dtau = q * ex * ch2;
    const double dV1dsig = -2 * V0 * t1 / (e.a * (sg * sg)) * ex * ch1;
    const double dV2dsig = -2 * V0 * t2 / (e.a * (sg * sg)) * ex * ch2;
    const double dV1dst = -2 * V0 / e.a * ex * sh1;
    const double dV2dst = -2 * V0 / e.a * ex * sh2;
    const double dstdt = e.b / e.dc * (V0 * exp(eclip((e.f0 - st) / e.b)) - e.V);
    const double ds = e.dsig - e.dPdt;
    k[0] = e.dsig / 1e6;
    k[1] = (-e.A1 + e.sl - eta * (dV1dsig * ds + dV1dst * dstdt)) / (1.0 + eta * dV1dtau) / 1e6;
    k[2] = (-e.A2 + e.sl - eta * (dV2dsig * ds + dV2dst * dstdt)) / (1.0 + eta * dV2dtau) / 1e6;
    k[3] = dstdt;
}

// Dormand-Prince tableau as used by the reference (QDsim_gpu.py:1129-1159,
// including its sign of a63)
__device__ __forceinline__ void ens_dp45(const EElem& e, double V0, double eta, const double* y, double h,
                                         double* y5, double* y4) {
    double k1[4], k2[4], k3[4], k4[4], k5[4], k6[4], k7[4], t[4];
    ens_rhs(e, V0, eta, y, k1);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * (.2) * k1[i];
    ens_rhs(e, V0, eta, t, k2);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * ((3. / 40) * k1[i] + (9. / 40.) * k2[i]);
    ens_rhs(e, V0, eta, t, k3);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] + h * ((44. / 45.) * k1[i] + (-56. / 15) * k2[i] + (32. / 9) * k3[i]);
    ens_rhs(e, V0, eta, t, k4);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t[i] = y[i] + h * ((19372. / 6561.) * k1[i] + (-25360 / 2187.) * k2[i] + (64448. / 6561.) * k3[i] +
                           (-212. / 729.) * k4[i]);
    ens_rhs(e, V0, eta, t, k5);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        t[i] = y[i] + h * ((9017. / 3168.) * k1[i] + (-355. / 33.) * k2[i] + (-46732. / 5247.) * k3[i] +
                           (49. / 176.) * k4[i] + (-5103. / 18656.) * k5[i]);
    ens_rhs(e, V0, eta, t, k6);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        y5[i] = y[i] + h * ((35. / 384.) * k1[i] + (500. / 1113.) * k3[i] + (125. / 192.) * k4[i] +
                            (-2187. / 6784.) * k5[i] + (11. / 84.) * k6[i]);
    ens_rhs(e, V0, eta, y5, k7);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        y4[i] = y[i] + h * ((5179. / 57600.) * k1[i] + (7571. / 16695.) * k3[i] + (393. / 640.) * k4[i] +
                            (-92097. / 339200.) * k5[i] + (187. / 2100.) * k6[i] + (1. / 40.) * k7[i]);
}

__device__ __forceinline__ double epy_max(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double epy_min(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ unsigned long long ebits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double eflt(unsigned long long u) { return __longlong_as_double((long long)u); }

__device__ bool ens_barrier(EnsCtrl* c, unsigned nblocks, long long spin_limit) {
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

// Max of (x, y) over the lanes whose member equals the wave's first member;
// lanes of other members (a wave straddling two members) post their own atomics.
__device__ __forceinline__ void member_max_atomic(int m, bool valid, unsigned long long x, unsigned long long y,
                                                  unsigned long long* slot_base /* [B][2] */) {
    const int m0 = __shfl(m, 0, 64);
    const bool same = (m == m0) || !valid;
    const unsigned long long all_same = __ballot(same);
    if (all_same == ~0ull) {
        unsigned long long vx = valid ? x : 0ull, vy = valid ? y : 0ull;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            unsigned long long wx = __shfl_xor(vx, o, 64), wy = __shfl_xor(vy, o, 64);
            vx = wx > vx ? wx : vx;
            vy = wy > vy ? wy : vy;
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMax(slot_base + 2 * m0, vx);
            atomicMax(slot_base + 2 * m0 + 1, vy);
        }
    } else if (valid) {
        atomicMax(slot_base + 2 * m, x);
        atomicMax(slot_base + 2 * m + 1, y);
    }
}

__global__ __launch_bounds__(256) void ens_rk_kernel(EnsArgs A) {
    const long long n = A.n;
    const int B = A.B;
    const long long items = n * B;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double* f = A.f;
    auto F = [&](int field, long long idx) -> double& { return f[(size_t)field * items + idx]; };
    const int par = (int)(__hip_atomic_load(&A.ctrl->gsteps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1);
    unsigned long long* slots = A.slot + (size_t)par * 24 * B * 2;
    // Every workgroup keeps its own copy of the member control in LDS and
    // applies the same deterministic accept/reject rule to the same reduced
    // errors, so one grid barrier per attempt round suffices.
    // s_st: 0 pending, 1 inactive, 2 accepted in this round, 3 failed, 4 accepted earlier
    __shared__ double s_h[64], s_dtnext[64];
    __shared__ int s_st[64], s_nrjct[64];
    __shared__ int s_any;
    if (threadIdx.x < B) {
        const MemberCtrl c = A.mc[threadIdx.x];
        s_h[threadIdx.x] = c.dttry;
        s_dtnext[threadIdx.x] = c.dttry;
        s_st[threadIdx.x] = c.active ? 0 : 1;
        s_nrjct[threadIdx.x] = 0;
    }
    for (long long t = t0; t < 24 * B * 2; t += stride) A.slot[(size_t)(par ^ 1) * 24 * B * 2 + t] = 0ull;
    // frozen traction rates (sum of the GEMM's split-K partials, fixed order);
    // each item is produced and later consumed by the same thread
    for (long long t = t0; t < items; t += stride) {
        const int m = (int)(t / n);
        const long long e = t - (long long)m * n;
        double s0 = 0, s1 = 0, s2 = 0;
        for (int c = 0; c < A.nchunk; ++c) {
            s0 += A.partial[(((size_t)c * 3 + 0) * B + m) * A.npad + e];
            s1 += A.partial[(((size_t)c * 3 + 1) * B + m) * A.npad + e];
            s2 += A.partial[(((size_t)c * 3 + 2) * B + m) * A.npad + e];
        }
        F(E_ADOTV1, t) = s0;
        F(E_ADOTV2, t) = s1;
        F(E_DSIGMADT, t) = s2 + F(E_NORMAL_LOADING, t);
    }
    __syncthreads();

    for (int att = 0; att < 22; ++att) {
        for (long long base = (long long)blockIdx.x * blockDim.x; base < items; base += stride) {
            const long long t = base + threadIdx.x;
            const bool in = t < items;
            const long long tt = in ? t : items - 1;
            const int m = (int)(tt / n);
            const bool pend = in && s_st[m] == 0;
            unsigned long long r1b = 0ull, r2b = 0ull;
            if (pend) {
                EElem e;
                e.a = F(E_A, tt); e.b = F(E_B, tt); e.dc = F(E_DC, tt); e.f0 = F(E_F0, tt);
                e.P = F(E_P, tt); e.dPdt = F(E_DPDT0, tt); e.sl = F(E_SHEAR_LOADING, tt);
                e.A1 = F(E_ADOTV1, tt); e.A2 = F(E_ADOTV2, tt); e.dsig = F(E_DSIGMADT, tt);
                e.V = F(E_SLIPV, tt);
                const double y[4] = {F(E_TNO, tt), F(E_TT1O, tt), F(E_TT2O, tt), F(E_STATE, tt)};
                double y5[4], y4[4];
                ens_dp45(e, A.V0, A.eta, y, s_h[m], y5, y4);
                F(E_Y5_0, tt) = y5[0]; F(E_Y5_1, tt) = y5[1]; F(E_Y5_2, tt) = y5[2]; F(E_Y5_3, tt) = y5[3];
                const double r1 = fabs(fabs(y4[3] - y5[3]) / y4[3]);
                const double rt1 = fabs(fabs(y4[1] - y5[1]) / y4[1]);
                const double rt2 = fabs(fabs(y4[2] - y5[2]) / y4[2]);
                const double r2 = (rt2 > rt1 || rt2 != rt2) ? rt2 : rt1;
                r1b = ebits(r1);
                r2b = ebits(r2);
            }
            member_max_atomic(m, pend, r1b, r2b, slots + (size_t)att * B * 2);
        }
        if (!ens_barrier(A.ctrl, (unsigned)A.nblocks, A.spin_limit)) return;
        // per-member decision (QDsim_gpu.py:959-985), replicated in every workgroup
        if (threadIdx.x < B && s_st[threadIdx.x] == 0) {
            const int m = threadIdx.x;
            const double e1 = eflt(__hip_atomic_load(slots + ((size_t)att * B + m) * 2 + 0, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) + 1e-10;
            const double e2 = eflt(__hip_atomic_load(slots + ((size_t)att * B + m) * 2 + 1, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) + 1e-10;
            const double c1 = e1 / 1e-4, c2 = e2 / 1e-4;
            const double h = s_h[m];
            const double hn1 = h * 0.9 * pow(1e-4 / e1, 0.2);
            const double hn2 = h * 0.9 * pow(1e-4 / e2, 0.2);
            if (epy_max(c1, c2) < 1.0 && !(c1 != c1 || c2 != c2)) {
                s_dtnext[m] = epy_min(1.5 * h, epy_min(hn1, hn2));
                s_st[m] = 2;
            } else {
                s_nrjct[m] += 1;
                const double hn = epy_max(0.5 * h, epy_min(hn1, hn2));
                s_h[m] = hn;
                if (hn < 1.e-15 || s_nrjct[m] > 20) s_st[m] = 3;
            }
        }
        __syncthreads();
        // post-step update of members accepted in this round (QDsim_gpu.py:987-1022)
        for (long long t = t0; t < items; t += stride) {
            const int m = (int)(t / n);
            if (s_st[m] != 2) continue;
            const double h = s_h[m];
            const double Tno = F(E_Y5_0, t), Tt1 = F(E_Y5_1, t), Tt2 = F(E_Y5_2, t), st = F(E_Y5_3, t);
            const double a = F(E_A, t);
            F(E_TNO, t) = Tno; F(E_TT1O, t) = Tt1; F(E_TT2O, t) = Tt2; F(E_STATE, t) = st;
            const double Tt = sqrt(Tt1 * Tt1 + Tt2 * Tt2);
            F(E_TT, t) = Tt;
            const double ex = 2.0 * A.V0 * exp(-st / a);
            const double v1 = ex * sinh(Tt1 / Tno / a), v2 = ex * sinh(Tt2 / Tno / a);
            F(E_SLIPV1, t) = v1; F(E_SLIPV2, t) = v2;
            double v = sqrt(v1 * v1 + v2 * v2);
            F(E_RAKE, t) = atan2(v2, v1);
            if (v < 1e-30) v = 1e-30;
            F(E_SLIPV, t) = v;
            const double s1 = F(E_SLIP1, t) + v1 * h, s2 = F(E_SLIP2, t) + v2 * h;
            F(E_SLIP1, t) = s1; F(E_SLIP2, t) = s2; F(E_SLIP, t) = sqrt(s1 * s1 + s2 * s2);
            F(E_FRIC, t) = Tt / (Tno - F(E_P, t) * 1e-6);
        }
        __syncthreads();
        if (threadIdx.x < B && s_st[threadIdx.x] == 2) s_st[threadIdx.x] = 4;
        __syncthreads();
        if (threadIdx.x == 0) {
            int any = 0;
            for (int m = 0; m < B; ++m) any |= (s_st[m] == 0);
            s_any = any;
        }
        __syncthreads();
        if (!s_any) break;
    }
    if (blockIdx.x == 0 && threadIdx.x < B) {
        const int m = threadIdx.x;
        MemberCtrl& c = A.mc[m];
        c.nrjct = s_nrjct[m];
        if (s_st[m] == 4) {
            c.time = c.time + s_h[m];
            c.h = s_h[m];
            c.dtnext = s_dtnext[m];
            c.dttry = s_dtnext[m];
            c.steps += 1;
        } else if (s_st[m] == 3 || s_st[m] == 0) {
            c.h = s_h[m];
            c.status = 1;
            c.active = 0;
        }
    }
    if (t0 == 0) A.ctrl->gsteps += 1;
}

__global__ void ens_trig_kernel(const double* __restrict__ r0, double* __restrict__ c0, double* __restrict__ s0,
                                long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        c0[i] = cos(r0[i]);
        s0[i] = sin(r0[i]);
    }
}

struct Ensemble {
    long long n = 0, ld = 0, npad = 0;
    int B = 64, kc = 512, nchunk = 0, nblocks = 0;
    const float* mats[6] = {nullptr};
    float* own_mats[6] = {nullptr};
    double* f = nullptr;
    float* U = nullptr;        // 2 x n x B
    float* partial = nullptr;  // nchunk x 3 x B x npad
    MemberCtrl* mc = nullptr;
    unsigned long long* slot = nullptr;
    EnsCtrl* ctrl = nullptr;
    double V0 = 1e-6, Vpl = 1e-9, eta = 0.0;
    long long spin_limit = 20000000;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
};

static EnsArgs ens_args(Ensemble* e) {
    EnsArgs a;
    a.f = e->f;
    a.partial = e->partial;
    a.mc = e->mc;
    a.slot = e->slot;
    a.ctrl = e->ctrl;
    a.n = e->n;
    a.npad = e->npad;
    a.B = e->B;
    a.nchunk = e->nchunk;
    a.nblocks = e->nblocks;
    a.V0 = e->V0;
    a.Vpl = e->Vpl;
    a.eta = e->eta;
    a.spin_limit = e->spin_limit;
    return a;
}

static void ens_launch_gemm(Ensemble* e, hipStream_t st) {
    const long long items = e->n * e->B;
    double* f = e->f;
    float* U1 = e->U;
    float* U2 = e->U + (size_t)e->n * e->B;
    hipLaunchKernelGGL(ens_prep_u_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, st,
                       f + (size_t)E_SLIPV1 * items, f + (size_t)E_SLIPV2 * items, f + (size_t)E_COSR0 * items,
                       f + (size_t)E_SINR0 * items, e->Vpl, e->n, e->B, e->mc, U1, U2);
    dim3 grid((unsigned)((e->n + 63) / 64), (unsigned)e->nchunk);
    hipLaunchKernelGGL(ens_gemm_kernel, grid, dim3(64), 0, st, e->mats[0], e->mats[1], e->mats[2], e->mats[3],
                       e->mats[4], e->mats[5], e->ld, e->n, e->kc, U1, U2, e->partial, e->npad);
}

static void ens_launch_step(Ensemble* e, hipStream_t st) {
    ens_launch_gemm(e, st);
    hipLaunchKernelGGL(ens_rk_kernel, dim3((unsigned)e->nblocks), dim3(256), 0, st, ens_args(e));
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

// B members (B == 64) on an n-element mesh.  If d_mats is non-null the six
// FP32 source-major matrices (leading dimension ld, multiple of 4) are
// borrowed from the caller (e.g. a dense simulation), else allocated here.
int pq3d_ens_create(void** out, long long n, int B, void* const* d_mats, long long ld, int kc) {
    clear_error();
    PQ_CHECK(B == 64, "pq3d_ens_create: B must be 64 (got %d)", B);
    PQ_CHECK(n > 0, "pq3d_ens_create: n must be positive");
    Ensemble* e = new Ensemble();
    e->n = n;
    e->B = B;
    if (kc > 0) {
        e->kc = (kc + 3) / 4 * 4;
    } else {
        // choose the number of column chunks so that the one-wave GEMM tiles
        // (64 rows x chunk) fill whole waves of the device's concurrent slots,
        // with a small penalty on the split-K partial traffic
        int dev = 0, per_cu = 0;
        PQ_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        PQ_HIP(hipGetDeviceProperties(&p, dev));
        PQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ens_gemm_kernel, 64, 0));
        const double slots = (double)per_cu * p.multiProcessorCount;
        const long long tiles = (n + 63) / 64;
        double best = -1.0;
        int best_nc = 8;
        for (int nc = 4; nc <= 32; ++nc) {
            const double W = (double)tiles * nc;
            const double waves = std::ceil(W / slots);
            const double score = W / (waves * slots) - 0.002 * nc;
            if (score > best) {
                best = score;
                best_nc = nc;
            }
        }
        e->kc = (int)(((n + best_nc - 1) / best_nc + 3) / 4 * 4);
    }
    e->nchunk = (int)((n + e->kc - 1) / e->kc);
    e->npad = (n + 255) / 256 * 256;
    if (d_mats) {
        PQ_CHECK(ld >= n && ld % 4 == 0, "pq3d_ens_create: ld must be >= n and a multiple of 4");
        e->ld = ld;
        for (int k = 0; k < 6; ++k) e->mats[k] = (const float*)d_mats[k];
    } else {
        e->ld = (n + 3) / 4 * 4;
        for (int k = 0; k < 6; ++k) {
            PQ_HIP(hipMalloc((void**)&e->own_mats[k], (size_t)e->ld * n * sizeof(float)));
            PQ_HIP(hipMemset(e->own_mats[k], 0, (size_t)e->ld * n * sizeof(float)));
            e->mats[k] = e->own_mats[k];
        }
    }
    const long long items = n * B;
    PQ_HIP(hipMalloc((void**)&e->f, (size_t)E_NFIELDS * items * sizeof(double)));
    PQ_HIP(hipMemset(e->f, 0, (size_t)E_NFIELDS * items * sizeof(double)));
    PQ_HIP(hipMalloc((void**)&e->U, (size_t)2 * items * sizeof(float)));
    PQ_HIP(hipMalloc((void**)&e->partial, (size_t)e->nchunk * 3 * B * e->npad * sizeof(float)));
    PQ_HIP(hipMemset(e->partial, 0, (size_t)e->nchunk * 3 * B * e->npad * sizeof(float)));
    PQ_HIP(hipMalloc((void**)&e->mc, B * sizeof(MemberCtrl)));
    PQ_HIP(hipMemset(e->mc, 0, B * sizeof(MemberCtrl)));
    PQ_HIP(hipMalloc((void**)&e->slot, (size_t)2 * 24 * B * 2 * sizeof(unsigned long long)));
    PQ_HIP(hipMemset(e->slot, 0, (size_t)2 * 24 * B * 2 * sizeof(unsigned long long)));
    PQ_HIP(hipMalloc((void**)&e->ctrl, sizeof(EnsCtrl)));
    PQ_HIP(hipMemset(e->ctrl, 0, sizeof(EnsCtrl)));
    // persistent RK grid: every workgroup must be co-resident
    int dev = 0;
    PQ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    PQ_HIP(hipGetDeviceProperties(&p, dev));
    int per_cu = 0;
    PQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ens_rk_kernel, 256, 0));
    long long want = (items + 255) / 256;
    long long cap = (long long)per_cu * p.multiProcessorCount;
    e->nblocks = (int)(want < cap ? want : cap);
    PQ_CHECK(e->nblocks > 0, "pq3d_ens_create: no co-resident workgroups");
    // members start active
    MemberCtrl h[64];
    memset(h, 0, sizeof(h));
    for (int m = 0; m < B; ++m) {
        h[m].active = 1;
        h[m].dttry = 1e-3;
    }
    PQ_HIP(hipMemcpy(e->mc, h, B * sizeof(MemberCtrl), hipMemcpyHostToDevice));
    *out = e;
    return 0;
}

int pq3d_ens_destroy(void* h) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    if (!e) return 0;
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    for (int k = 0; k < 6; ++k) (void)hipFree(e->own_mats[k]);
    (void)hipFree(e->f);
    (void)hipFree(e->U);
    (void)hipFree(e->partial);
    (void)hipFree(e->mc);
    (void)hipFree(e->slot);
    (void)hipFree(e->ctrl);
    delete e;
    return 0;
}

int pq3d_ens_matrix_ptrs(void* h, void** ptrs, long long* ld) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    for (int k = 0; k < 6; ++k) ptrs[k] = (void*)e->mats[k];
    *ld = e->ld;
    return 0;
}

// Field values for all members: host array B x n (member-major).  Field
// indices are those of the dense simulation.
int pq3d_ens_set_field(void* h, int field, const double* host) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    PQ_CHECK(field >= 0 && field < E_Y5_0, "bad field %d", field);
    const long long items = e->n * e->B;
    double* dst = e->f + (size_t)field * items;
    PQ_HIP(hipMemcpy(dst, host, items * sizeof(double), hipMemcpyHostToDevice));
    if (field == E_RAKE0) {
        hipLaunchKernelGGL(ens_trig_kernel, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, 0, dst,
                           e->f + (size_t)E_COSR0 * items, e->f + (size_t)E_SINR0 * items, items);
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipDeviceSynchronize());
    }
    return 0;
}

int pq3d_ens_get_field(void* h, int field, double* host) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    PQ_CHECK(field >= 0 && field < E_NFIELDS, "bad field %d", field);
    const long long items = e->n * e->B;
    PQ_HIP(hipMemcpy(host, e->f + (size_t)field * items, items * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int pq3d_ens_set_params(void* h, double V0, double Vpl, double eta) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    e->V0 = V0;
    e->Vpl = Vpl;
    e->eta = eta;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return 0;
}

// Member control: per member 5 doubles (time, next dt, last h, last dtnext,
// unused) and 4 ints (rejections in the last step, status, active, steps).
int pq3d_ens_members(void* h, double* d5, long long* i4) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    MemberCtrl c[64];
    PQ_HIP(hipMemcpy(c, e->mc, e->B * sizeof(MemberCtrl), hipMemcpyDeviceToHost));
    for (int m = 0; m < e->B; ++m) {
        d5[5 * m + 0] = c[m].time;
        d5[5 * m + 1] = c[m].dttry;
        d5[5 * m + 2] = c[m].h;
        d5[5 * m + 3] = c[m].dtnext;
        d5[5 * m + 4] = 0.0;
        i4[4 * m + 0] = c[m].nrjct;
        i4[4 * m + 1] = c[m].status;
        i4[4 * m + 2] = c[m].active;
        i4[4 * m + 3] = c[m].steps;
    }
    return 0;
}

// Set each member's next step size and active flag (host arrays of B).
int pq3d_ens_set_members(void* h, const double* dttry, const int* active) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    MemberCtrl c[64];
    PQ_HIP(hipMemcpy(c, e->mc, e->B * sizeof(MemberCtrl), hipMemcpyDeviceToHost));
    for (int m = 0; m < e->B; ++m) {
        if (dttry) c[m].dttry = dttry[m];
        if (active) c[m].active = active[m];
    }
    PQ_HIP(hipMemcpy(e->mc, c, e->B * sizeof(MemberCtrl), hipMemcpyHostToDevice));
    return 0;
}

// Only the GEMM (u from the current slip rates -> per-member traction rates
// into AdotV1/AdotV2/dsigmadt without normal loading); for tests and timing.
int pq3d_ens_gemm(void* h, int reps, float* ms_per) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    hipEvent_t a, b;
    PQ_HIP(hipEventCreate(&a));
    PQ_HIP(hipEventCreate(&b));
    ens_launch_gemm(e, 0);
    PQ_HIP(hipEventRecord(a, 0));
    for (int r = 0; r < reps; ++r) ens_launch_gemm(e, 0);
    PQ_HIP(hipEventRecord(b, 0));
    PQ_HIP(hipEventSynchronize(b));
    float ms = 0;
    PQ_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms_per) *ms_per = reps > 0 ? ms / reps : 0.f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 0;
}

// Copy the summed GEMM result (3 x B x n, FP64 host) -- sums the partials on
// the host side of the call through a small device pass.
__global__ void ens_sum_partials_kernel(const float* __restrict__ partial, int nchunk, int B, long long n,
                                        long long npad, double* __restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * B * n) return;
    const long long q = t / (B * n);
    const long long r = t - q * B * n;
    const long long m = r / n, e = r - m * n;
    double s = 0;
    for (int c = 0; c < nchunk; ++c) s += partial[(((size_t)c * 3 + q) * B + m) * npad + e];
    out[t] = s;
}

int pq3d_ens_gemm_result(void* h, double* host_3Bn) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    const long long tot = 3LL * e->B * e->n;
    double* d = nullptr;
    PQ_HIP(hipMalloc((void**)&d, tot * sizeof(double)));
    hipLaunchKernelGGL(ens_sum_partials_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, e->partial,
                       e->nchunk, e->B, e->n, e->npad, d);
    hipError_t err = hipMemcpy(host_3Bn, d, tot * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    PQ_HIP(err);
    return 0;
}

// nsteps global steps (each member takes one accepted step per global step),
// no host synchronisation in between; HIP graph of graph_steps steps.
int pq3d_ens_run(void* h, long long nsteps, int graph_steps, int* status) {
    clear_error();
    Ensemble* e = (Ensemble*)h;
    if (graph_steps <= 0) graph_steps = 1;
    long long done = 0;
    if (nsteps >= graph_steps) {
        if (!e->graph || e->graph_steps != graph_steps) {
            if (e->graph) (void)hipGraphExecDestroy(e->graph);
            hipStream_t cs;
            PQ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            PQ_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < graph_steps; ++k) ens_launch_step(e, cs);
            hipGraph_t g;
            PQ_HIP(hipStreamEndCapture(cs, &g));
            PQ_HIP(hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            (void)hipStreamDestroy(cs);
            e->graph_steps = graph_steps;
        }
        for (; done + graph_steps <= nsteps; done += graph_steps) PQ_HIP(hipGraphLaunch(e->graph, 0));
    }
    for (; done < nsteps; ++done) ens_launch_step(e, 0);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipDeviceSynchronize());
    EnsCtrl c;
    PQ_HIP(hipMemcpy(&c, e->ctrl, sizeof(EnsCtrl), hipMemcpyDeviceToHost));
    *status = c.status;
    return 0;
}

}  // extern "C"
