// Device-resident H-matrix of the six traction matrices and its product
// (the reference's BlockTree.blocks_process_MVM, src/Hmatrix.py:1773-1866,
// called three times per step pair-wise at src/QDsim.py:1153-1160).
//
// Layout.  Rows and columns are in the cluster order (depth-first leaves of
// the cluster tree, so every block is a contiguous [r0, r0+m) x [c0, c0+n)
// range).  The rows are cut into segments of at most 64 at every block row
// boundary; a block contributes one item per segment it covers:
//   * dense leaf item: FP32 [matrix 6][column n][row msp], msp = ms rounded
//     up to 4 (zero rows), so a lane loads 4 rows of a column as one 16-byte
//     vector;
//   * low-rank item: the segment's rows of U, [column][row mpu] in the factor
//     type (FP32 or BF16), mpu = ms rounded up to one 16-byte vector of rows;
//     the columns grouped by output (A1s then A1d, A2s then A2d, Bs then Bd),
//     so each output's columns are one contiguous run.
// Low-rank V factors: per block, the rank rows of the three matrices applied
// to x1 (A1s, A2s, Bs) one after another, padded to a multiple of VEC rows,
// then those applied to x2 (A1d, A2d, Bd) likewise; stored as slabs of VEC
// rows, [slab][column][VEC] (VEC = 16 bytes of the factor type), so a lane
// loads one 16-byte vector per column.  t (= V x) follows the same row order.
//
// Product y = (A1s x1 + A1d x2, A2s x1 + A2d x2, Bs x1 + Bd x2), four launches:
//   1. gather: x in cluster order (FP32);
//   2. phase A + dense: waves t = V x over 1-4 slabs of a block, FP32 sums,
//      then one wave per dense item, partial rows = dense * x (these do not
//      need t, so they share the launch);
//   3. low-rank: per wave a few low-rank items of a segment, partial rows =
//      sum of U * t -- lanes (row vector, column group), FP32 sums reduced
//      across groups;
//   4. reduce: per row, the items of its segment summed in FP64 in a fixed
//      order (deterministic: results do not depend on scheduling) and
//      scattered to element order.
// With fix_tn the normal-traction matrices Bs, Bd are skipped (the step then
// uses the normal loading alone, QDsim_gpu.py:948-953).
#include "common.hpp"

#include <algorithm>
#include <cstdint>
#include <vector>

namespace pq3d {

// one phase-A wave: nslab consecutive slabs of a block, ls lanes per slab
struct AItem {
    long long v_off;  // first element of the first slab (layout [slab][column][VEC])
    int n, c0;        // columns, first cluster-order column
    int t_off;        // first of the nslab * VEC outputs in t
    int nslab, ls;    // slabs (1-4), lanes per slab (64 / 2^k >= 16)
    int src_mask;     // bit s: slab s multiplies x2 (else x1)
    int fix_mask;     // bit s: slab s holds only Bs / Bd rows (or padding): skipped with fix_tn
    int pad;
};

struct BItem {
    long long off[6];  // dense: off[0] = data [n][6][msp]; low-rank: off[o] = first U column of output o
    int t_off[6];      // low-rank: t offsets per matrix
    int k[6];          // low-rank ranks
    int c0, n;         // dense: first cluster-order column, column count
    int s0, ms;        // segment rows [s0, s0+ms)
    int slot;          // partial sums at slot: [3][ms]
    int type;          // 0 dense, 1 low-rank
};

// one phase-B wave: `count` consecutive items of one segment and one kind
// (a dense item alone, or up to kLowRankPerWave low-rank items), partial sums
// [3][ms] at slot
constexpr int kLowRankPerWave = 2;
struct BWave {
    int item0, count, slot, pad;
};

struct Seg {
    int s0, ms;  // rows [s0, s0+ms)
    int nw;      // the segment's phase-B waves; their partial slots are consecutive from slot0
    int slot0;
};

struct HMat {
    long long n = 0;
    int ftype = MAT_F32;
    int vec = 4;
    int* perm = nullptr;      // cluster position -> element id
    AItem* a_items = nullptr;
    long long na = 0;
    BItem* b_items = nullptr;
    long long nbi = 0;
    BWave* b_waves = nullptr;  // the dense waves, then the low-rank waves
    long long nbw = 0, nbd = 0;
    Seg* segs = nullptr;
    long long nseg = 0;
    float* ddata = nullptr;   // dense leaves
    void* udata = nullptr;    // U rows (ftype)
    void* vdata = nullptr;    // V slabs (ftype)
    float* xp = nullptr;      // 2 x n, cluster order
    float* t = nullptr;
    float* partial = nullptr;
    double* xin = nullptr;    // 2 x n, element order (host-call inputs)
    double* yout = nullptr;   // 3 x n, element order (host-call outputs)
    long long nt = 0, npart = 0;
    // bytes read per product: algorithmic (unpadded) and as stored (padded), all
    // six matrices / without Bs, Bd (fix_tn)
    double alg = 0, alg_fix = 0, stored = 0, stored_fix = 0, alg_dense = 0, alg_factors = 0;
    long long lowrank_blocks = 0;
};

__device__ __forceinline__ float bf16f(unsigned int b) { return __uint_as_float(b << 16); }

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void hmat_gather_kernel(const int* __restrict__ perm, long long n,
                                                          const double* __restrict__ V1,
                                                          const double* __restrict__ V2,
                                                          const double* __restrict__ c0,
                                                          const double* __restrict__ s0, double Vpl,
                                                          float* __restrict__ xp) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int e = perm[p];
    double u1 = V1[e], u2 = V2[e];
    if (c0) {
        u1 = __dsub_rn(u1, __dmul_rn(Vpl, c0[e]));
        u2 = __dsub_rn(u2, __dmul_rn(Vpl, s0[e]));
    }
    xp[p] = (float)u1;
    xp[n + p] = (float)u2;
}

template <class T> struct Slab;
template <> struct Slab<float> {
    static constexpr int VEC = 4;
    __device__ static void load(const float* base, long long j, float* o) {
        const float4 a = *reinterpret_cast<const float4*>(base + 4 * j);
        o[0] = a.x;
        o[1] = a.y;
        o[2] = a.z;
        o[3] = a.w;
    }
    __device__ static float one(const float* p, long long i) { return p[i]; }
    // the 16-byte vector j as loaded, and its element r (r a constant)
    using Raw = float4;
    __device__ static Raw raw(const float* base, long long j) { return *reinterpret_cast<const float4*>(base + 4 * j); }
    // the vector at element offset e (32-bit) from a wave-uniform base
    __device__ static Raw raw32(const float* base, unsigned e) { return *reinterpret_cast<const float4*>(base + e); }
    __device__ static float at(const Raw& a, int r) { return r == 0 ? a.x : r == 1 ? a.y : r == 2 ? a.z : a.w; }
};
template <> struct Slab<unsigned short> {
    static constexpr int VEC = 8;
    __device__ static void load(const unsigned short* base, long long j, float* o) {
        const uint4 a = *reinterpret_cast<const uint4*>(base + 8 * j);
        const unsigned int w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            o[2 * k] = bf16f(w[k] & 0xffffu);
            o[2 * k + 1] = bf16f(w[k] >> 16);
        }
    }
    __device__ static float one(const unsigned short* p, long long i) { return bf16f(p[i]); }
    using Raw = uint4;
    __device__ static Raw raw(const unsigned short* base, long long j) {
        return *reinterpret_cast<const uint4*>(base + 8 * j);
    }
    __device__ static Raw raw32(const unsigned short* base, unsigned e) {
        return *reinterpret_cast<const uint4*>(base + e);
    }
    __device__ static float at(const Raw& a, int r) {
        const unsigned int w = (r >> 1) == 0 ? a.x : (r >> 1) == 1 ? a.y : (r >> 1) == 2 ? a.z : a.w;
        return (r & 1) ? bf16f(w >> 16) : bf16f(w & 0xffffu);
    }
};

// phase A: lanes [s ls, (s+1) ls) handle slab s of the item, columns
// strided by ls; t[t_off + s VEC + e] = sum_j V[s][j][e] x[c0 + j], the ls
// lanes' sums combined in a fixed butterfly.
template <class T>
__device__ __forceinline__ void phase_a_wave(const AItem& it, const T* __restrict__ vdata,
                                             const float* __restrict__ xp, long long n, float* __restrict__ t,
                                             int fix_tn, int lane) {
    constexpr int VEC = Slab<T>::VEC;
    const int ls = it.ls;
    const int sl = lane / ls, cl = lane - sl * ls;
    const bool live = sl < it.nslab && !(fix_tn && ((it.fix_mask >> sl) & 1));
    float acc[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc[e] = 0.f;
    if (live) {
        // 32-bit element offsets from wave-uniform bases
        const T* base = vdata + it.v_off;
        const float* x = xp + it.c0;
        const unsigned vs = (unsigned)(sl * it.n * VEC);
        const unsigned xs0 = ((it.src_mask >> sl) & 1) ? (unsigned)n : 0u;
        for (int j0 = cl; j0 < it.n; j0 += 4 * ls) {  // up to four 16-byte loads in flight per lane
            // past the last column: reload column j0 and weight it by zero
            // (unconditional loads stay in flight together)
            typename Slab<T>::Raw v[4];
            float xs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool live = j0 + u * ls < it.n;
                const unsigned j = (unsigned)(live ? j0 + u * ls : j0);
                v[u] = Slab<T>::raw32(base, vs + j * VEC);
                const float xv = x[xs0 + j];
                xs[u] = live ? xv : 0.f;
            }
#pragma unroll
            for (int e = 0; e < VEC; ++e)
                acc[e] = fmaf(Slab<T>::at(v[0], e), xs[0],
                              fmaf(Slab<T>::at(v[1], e), xs[1],
                                   fmaf(Slab<T>::at(v[2], e), xs[2], fmaf(Slab<T>::at(v[3], e), xs[3], acc[e]))));
        }
    }
    for (int o = 1; o < ls; o <<= 1) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] += __shfl_xor(acc[e], o, 64);
    }
    float mine = 0.f;
#pragma unroll
    for (int e = 0; e < VEC; ++e)
        if (cl == e) mine = acc[e];
    if (live && cl < VEC) t[it.t_off + sl * VEC + cl] = mine;
}

__device__ __forceinline__ int pow2ceil(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Phase-B epilogue: lanes (q, g) hold RA rows q*R.. of column group g; the
// groups are summed across lanes in a fixed butterfly and group 0 writes the
// partial rows [3][ms] at P.
template <int RA>
__device__ __forceinline__ void b_store(float (&acc)[RA][3], int R, int qp, int q, int g, int ms, float* P) {
    for (int o = qp; o < 64; o <<= 1) {
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            acc[r][0] += __shfl_xor(acc[r][0], o, 64);
            acc[r][1] += __shfl_xor(acc[r][1], o, 64);
            acc[r][2] += __shfl_xor(acc[r][2], o, 64);
        }
    }
    if (g == 0) {
#pragma unroll
        for (int r = 0; r < RA; ++r) {
            const int row = q * R + r;
            if (r < R && row < ms) {
                P[row] = acc[r][0];
                P[ms + row] = acc[r][1];
                P[2 * ms + row] = acc[r][2];
            }
        }
    }
}

// One dense item: each lane owns 4 consecutive rows of the segment (one
// 16-byte vector of a column) and a column group g.
__device__ __forceinline__ void dense_wave(const BWave& wv, const BItem* __restrict__ items,
                                           const float* __restrict__ ddata, const float* __restrict__ xp,
                                           long long n, float* __restrict__ partial, int fix_tn, int lane) {
    const BItem it = items[wv.item0];
    const int ms = it.ms;
    float acc[4][3];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r][0] = acc[r][1] = acc[r][2] = 0.f;
    const int msp = (ms + 3) & ~3, nq = msp >> 2;
    const int qp = pow2ceil(nq);
    const int G = 64 / qp;
    const int q = lane & (qp - 1);
    const int g = lane / qp;
    if (q < nq) {
        const float* D = ddata + it.off[0] + 4 * q;
        const float* x1 = xp + it.c0;
        const float* x2 = xp + n + it.c0;
        const int nm = fix_tn ? 4 : 6;                 // Bs, Bd last (skipped with fix_tn)
        for (int j0 = g; j0 < it.n; j0 += 2 * G) {     // two columns' loads in flight
            float4 m[2][6];
            float a[2], b[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = j0 + u * G < it.n ? j0 + u * G : j0;
                const float* col = D + (long long)j * 6 * msp;  // [column][matrix][row]
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    if (c < nm) m[u][c] = *reinterpret_cast<const float4*>(col + c * msp);
                const bool live = j0 + u * G < it.n;
                a[u] = live ? x1[j] : 0.f;
                b[u] = live ? x2[j] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const float v0[4] = {m[u][0].x, m[u][0].y, m[u][0].z, m[u][0].w};
                const float v1[4] = {m[u][1].x, m[u][1].y, m[u][1].z, m[u][1].w};
                const float v3[4] = {m[u][2].x, m[u][2].y, m[u][2].z, m[u][2].w};
                const float v4[4] = {m[u][3].x, m[u][3].y, m[u][3].z, m[u][3].w};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    acc[r][0] = fmaf(v0[r], a[u], fmaf(v3[r], b[u], acc[r][0]));
                    acc[r][1] = fmaf(v1[r], a[u], fmaf(v4[r], b[u], acc[r][1]));
                }
                if (nm == 6) {
                    const float v2[4] = {m[u][4].x, m[u][4].y, m[u][4].z, m[u][4].w};
                    const float v5[4] = {m[u][5].x, m[u][5].y, m[u][5].z, m[u][5].w};
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[r][2] = fmaf(v2[r], a[u], fmaf(v5[r], b[u], acc[r][2]));
                }
            }
        }
    }
    b_store<4>(acc, 4, qp, q, g, ms, partial + wv.slot);
}

// Launch 2: waves [0, na) are phase-A items, waves [na, na + nd) dense items.
template <class T>
__global__ __launch_bounds__(256) void hmat_phase_ad_kernel(const AItem* __restrict__ a_items, long long na,
                                                            const T* __restrict__ vdata,
                                                            const BWave* __restrict__ dwaves, long long nd,
                                                            const BItem* __restrict__ items,
                                                            const float* __restrict__ ddata,
                                                            const float* __restrict__ xp, long long n,
                                                            float* __restrict__ t, float* __restrict__ partial,
                                                            int fix_tn) {
    // wave-uniform (readfirstlane): the item tables are then read with scalar loads
    const long long w = (long long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w < na) {
        const AItem it = a_items[w];
        phase_a_wave<T>(it, vdata, xp, n, t, fix_tn, lane);
    } else if (w - na < nd) {
        const BWave wv = dwaves[w - na];
        dense_wave(wv, items, ddata, xp, n, partial, fix_tn, lane);
    }
}

// Launch 3: one wave per BWave of up to kLowRankPerWave low-rank items of the
// same segment accumulated together.  Each lane owns VR consecutive rows of
// the segment (one 16-byte vector of a U column) and a column group g; output
// o's columns (those of matrix o, then of matrix o + 3) are taken two per
// lane at a time, the three outputs' loads issued together.
template <class T>
__global__ __launch_bounds__(256) void hmat_lowrank_kernel(const BWave* __restrict__ waves, long long nwaves,
                                                           const BItem* __restrict__ items,
                                                           const T* __restrict__ udata,
                                                           const float* __restrict__ t,
                                                           float* __restrict__ partial, int fix_tn) {
    constexpr int VR = Slab<T>::VEC;  // U rows per 16-byte vector
    // wave-uniform (readfirstlane): the item tables are then read with scalar loads
    const long long w = (long long)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nwaves) return;  // whole wave
    const BWave wv = waves[w];
    const int ms = items[wv.item0].ms;
    float acc[VR][3];
#pragma unroll
    for (int r = 0; r < VR; ++r) acc[r][0] = acc[r][1] = acc[r][2] = 0.f;
    const int mpu = (ms + VR - 1) / VR * VR, no = mpu / VR;
    const int qp = pow2ceil(no);
    const int G = 64 / qp;
    const int q = lane & (qp - 1);
    const int g = lane / qp;
    if (q < no) {
        for (int x = 0; x < wv.count; ++x) {
            const BItem it = items[wv.item0 + x];  // wave-uniform: held in scalar registers
            // per output: columns K, of which the first mid are matrix o's;
            // U element offsets from the item's first column, t indices
            // ta + l (l < mid) and tb + l
            int K[3], mid[3], uo[3], ta[3], tb[3];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                mid[o] = it.k[o];
                K[o] = (fix_tn && o == 2) ? 0 : it.k[o] + it.k[o + 3];
                uo[o] = (int)(it.off[o] - it.off[0]);
                ta[o] = it.t_off[o];
                tb[o] = it.t_off[o + 3] - it.k[o];
            }
            const T* Ub = udata + it.off[0];
            const unsigned qo = (unsigned)(q * VR);
            const int kmax = max(K[0], max(K[1], K[2]));
            for (int l0 = g; l0 < kmax; l0 += 2 * G) {
                typename Slab<T>::Raw u[3][2];
                float tl[3][2];
#pragma unroll
                for (int o = 0; o < 3; ++o)
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        // past the output's last column: reload the item's
                        // first column and weight it by zero (unconditional
                        // loads stay in flight together)
                        const int l = l0 + v * G;
                        const bool live = l < K[o];
                        const unsigned e = live ? (unsigned)(uo[o] + l * mpu) : 0u;
                        u[o][v] = Slab<T>::raw32(Ub, e + qo);
                        const int ti = l + (l < mid[o] ? ta[o] : tb[o]);
                        const float tv = t[(unsigned)(live ? ti : ta[0])];
                        tl[o][v] = live ? tv : 0.f;
                    }
#pragma unroll
                for (int o = 0; o < 3; ++o)
#pragma unroll
                    for (int r = 0; r < VR; ++r)
                        acc[r][o] = fmaf(Slab<T>::at(u[o][0], r), tl[o][0],
                                         fmaf(Slab<T>::at(u[o][1], r), tl[o][1], acc[r][o]));
            }
        }
    }
    b_store<VR>(acc, VR, qp, q, g, ms, partial + wv.slot);
}

// reduce: one workgroup of four waves per segment; lane i of wave v sums, in
// FP64, the partial rows of the segment's phase-B waves v, v + 4, ... for row
// i (each load reads one wave's partial row block contiguously), and the four
// wave sums are added in wave order (deterministic: results do not depend on
// scheduling); scattered to element order: out[q * n + perm[p]]
__global__ __launch_bounds__(256) void hmat_reduce_kernel(const Seg* __restrict__ segs,
                                                          const int* __restrict__ perm, long long n,
                                                          const float* __restrict__ partial,
                                                          double* __restrict__ out) {
    __shared__ double part[4][3][64];
    const Seg s = segs[blockIdx.x];
    const int i = threadIdx.x & 63, v = threadIdx.x >> 6;
    const bool ok = i < s.ms;
    double y0 = 0, y1 = 0, y2 = 0;
    if (ok) {
        const long long stride = 3LL * s.ms;
        const float* P = partial + s.slot0 + i + stride * v;
        int w = v;
        // two waves' rows in flight per step
        for (; w + 4 < s.nw; w += 8, P += 8 * stride) {
            const float a0 = P[0], a1 = P[s.ms], a2 = P[2 * s.ms];
            const float b0 = P[4 * stride], b1 = P[4 * stride + s.ms], b2 = P[4 * stride + 2 * s.ms];
            y0 += (double)a0;
            y1 += (double)a1;
            y2 += (double)a2;
            y0 += (double)b0;
            y1 += (double)b1;
            y2 += (double)b2;
        }
        if (w < s.nw) {
            y0 += (double)P[0];
            y1 += (double)P[s.ms];
            y2 += (double)P[2 * s.ms];
        }
    }
    part[v][0][i] = y0;
    part[v][1][i] = y1;
    part[v][2][i] = y2;
    __syncthreads();
    if (v == 0 && ok) {
        const int e = perm[s.s0 + i];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            out[c * n + e] = ((part[0][c][i] + part[1][c][i]) + part[2][c][i]) + part[3][c][i];
    }
}

static unsigned short f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <class T> static T conv(double x);
template <> float conv<float>(double x) { return (float)x; }
template <> unsigned short conv<unsigned short>(double x) { return f2bf((float)x); }

template <class T>
static void launch_mvm_t(HMat* h, const double* V1, const double* V2, const double* c0, const double* s0,
                         double Vpl, int fix_tn, double* out, Launch& L) {
    const long long n = h->n;
    L(hmat_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), h->perm, n, V1, V2, c0, s0, Vpl, h->xp);
    if (h->na + h->nbd)
        L(hmat_phase_ad_kernel<T>, dim3((unsigned)((h->na + h->nbd + 3) / 4)), dim3(256), h->a_items, h->na,
          (const T*)h->vdata, h->b_waves, h->nbd, h->b_items, h->ddata, h->xp, n, h->t, h->partial, fix_tn);
    if (h->nbw > h->nbd)
        L(hmat_lowrank_kernel<T>, dim3((unsigned)((h->nbw - h->nbd + 3) / 4)), dim3(256), h->b_waves + h->nbd,
          h->nbw - h->nbd, h->b_items, (const T*)h->udata, h->t, h->partial, fix_tn);
    if (h->nseg) L(hmat_reduce_kernel, dim3((unsigned)h->nseg), dim3(256), h->segs, h->perm, n, h->partial, out);
}

// Traction rates of one step into out (3 x n, element order): x1 = V1 - Vpl
// cos(rake0), x2 = V2 - Vpl sin(rake0) (c0/s0 null: x = V as given).
void hmat_launch(void* hm, const double* V1, const double* V2, const double* c0, const double* s0, double Vpl,
                 int fix_tn, double* out, Launch& L) {
    HMat* h = (HMat*)hm;
    if (h->ftype == MAT_BF16) launch_mvm_t<unsigned short>(h, V1, V2, c0, s0, Vpl, fix_tn, out, L);
    else launch_mvm_t<float>(h, V1, V2, c0, s0, Vpl, fix_tn, out, L);
}

long long hmat_n(void* hm) { return ((HMat*)hm)->n; }

template <class T>
static int pack(HMat* h, long long n, const int* perm, long long nb, const long long* row_off, const int* row_idx,
                const long long* col_off, const int* col_idx, const int* kind, const long long* ranks,
                const double* dense, const double* U, const double* V) {
    constexpr int VEC = 16 / sizeof(T);
    h->vec = VEC;
    std::vector<int> inv(n, -1);
    for (long long p = 0; p < n; ++p) {
        PQ_CHECK(perm[p] >= 0 && perm[p] < n && inv[perm[p]] < 0, "pq3d_hmat_create: perm is not a permutation");
        inv[perm[p]] = (int)p;
    }
    std::vector<int> r0(nb), c0(nb);
    std::vector<int> bnd;
    bnd.push_back(0);
    bnd.push_back((int)n);
    double cover = 0;
    for (long long b = 0; b < nb; ++b) {
        const long long m = row_off[b + 1] - row_off[b], nc = col_off[b + 1] - col_off[b];
        PQ_CHECK(m > 0 && nc > 0, "pq3d_hmat_create: empty block %lld", b);
        int lo = INT32_MAX, hi = -1;
        for (long long i = row_off[b]; i < row_off[b + 1]; ++i) {
            PQ_CHECK(row_idx[i] >= 0 && row_idx[i] < n, "pq3d_hmat_create: row id out of range");
            lo = std::min(lo, inv[row_idx[i]]);
            hi = std::max(hi, inv[row_idx[i]]);
        }
        PQ_CHECK(hi - lo + 1 == m, "pq3d_hmat_create: rows of block %lld are not a cluster range", b);
        r0[b] = lo;
        lo = INT32_MAX;
        hi = -1;
        for (long long j = col_off[b]; j < col_off[b + 1]; ++j) {
            PQ_CHECK(col_idx[j] >= 0 && col_idx[j] < n, "pq3d_hmat_create: column id out of range");
            lo = std::min(lo, inv[col_idx[j]]);
            hi = std::max(hi, inv[col_idx[j]]);
        }
        PQ_CHECK(hi - lo + 1 == nc, "pq3d_hmat_create: columns of block %lld are not a cluster range", b);
        c0[b] = lo;
        bnd.push_back(r0[b]);
        bnd.push_back(r0[b] + (int)m);
        cover += (double)m * (double)nc;
        for (int c = 0; c < 6; ++c)
            PQ_CHECK(kind[b] == 0 || (ranks[6 * b + c] >= 0 && ranks[6 * b + c] <= std::min(m, nc)),
                     "pq3d_hmat_create: bad rank in block %lld", b);
    }
    PQ_CHECK(cover == (double)n * (double)n, "pq3d_hmat_create: blocks cover %.0f entries, not n^2", cover);
    std::sort(bnd.begin(), bnd.end());
    bnd.erase(std::unique(bnd.begin(), bnd.end()), bnd.end());
    std::vector<Seg> segs;
    for (size_t k = 0; k + 1 < bnd.size(); ++k) {
        const int a = bnd[k], len = bnd[k + 1] - bnd[k];
        const int pieces = (len + 63) / 64;
        for (int q = 0; q < pieces; ++q) {
            const int s = a + (int)((long long)len * q / pieces), e = a + (int)((long long)len * (q + 1) / pieces);
            segs.push_back(Seg{s, e - s, 0, 0});
        }
    }
    std::vector<int> seg_start(segs.size());
    for (size_t s = 0; s < segs.size(); ++s) seg_start[s] = segs[s].s0;
    // items: per segment, the blocks covering it in block order
    std::vector<std::vector<int>> seg_blocks(segs.size());
    for (long long b = 0; b < nb; ++b) {
        const long long m = row_off[b + 1] - row_off[b];
        size_t s = std::lower_bound(seg_start.begin(), seg_start.end(), r0[b]) - seg_start.begin();
        for (; s < segs.size() && segs[s].s0 < r0[b] + m; ++s) seg_blocks[s].push_back((int)b);
    }
    // offsets of every block's inputs
    std::vector<long long> doff(nb, -1), uoff(nb * 6, 0), voff(nb * 6, 0);
    {
        long long d = 0, u = 0, v = 0;
        for (long long b = 0; b < nb; ++b) {
            const long long m = row_off[b + 1] - row_off[b], nc = col_off[b + 1] - col_off[b];
            if (kind[b] == 0) {
                doff[b] = d;
                d += 6 * m * nc;
            } else {
                for (int c = 0; c < 6; ++c) {
                    uoff[6 * b + c] = u;
                    voff[6 * b + c] = v;
                    u += ranks[6 * b + c] * m;
                    v += ranks[6 * b + c] * nc;
                }
            }
        }
    }
    // block-local index of each cluster-order offset
    std::vector<long long> rinv_off(nb + 1, 0), cinv_off(nb + 1, 0);
    for (long long b = 0; b < nb; ++b) {
        rinv_off[b + 1] = rinv_off[b] + (row_off[b + 1] - row_off[b]);
        cinv_off[b + 1] = cinv_off[b] + (col_off[b + 1] - col_off[b]);
    }
    std::vector<int> rinv(rinv_off[nb]), cinv(cinv_off[nb]);
    for (long long b = 0; b < nb; ++b) {
        for (long long i = row_off[b]; i < row_off[b + 1]; ++i)
            rinv[rinv_off[b] + inv[row_idx[i]] - r0[b]] = (int)(i - row_off[b]);
        for (long long j = col_off[b]; j < col_off[b + 1]; ++j)
            cinv[cinv_off[b] + inv[col_idx[j]] - c0[b]] = (int)(j - col_off[b]);
    }
    // t offsets and V slabs (phase A): per block the rank rows of A1s, A2s, Bs
    // (applied to x1) padded to a multiple of VEC, then those of A1d, A2d, Bd
    // (applied to x2) likewise; phase-A waves take up to 64 / ls consecutive
    // slabs of a block, ls lanes per slab (about four columns per lane)
    std::vector<int> toff(nb * 6, 0);
    std::vector<AItem> aitems;
    std::vector<T> vpack;
    long long tcount = 0;
    h->lowrank_blocks = 0;
    for (long long b = 0; b < nb; ++b) {
        if (kind[b] == 0) continue;
        ++h->lowrank_blocks;
        const long long nc = col_off[b + 1] - col_off[b];
        int ls = 16;
        while (ls < 64 && 4LL * ls < nc) ls <<= 1;
        const int S = 64 / ls;
        std::vector<int> slab_src, slab_fix;
        const long long base = (long long)vpack.size();
        for (int gi = 0; gi < 2; ++gi) {
            const long long* kr = ranks + 6 * b + 3 * gi;
            const long long rows = kr[0] + kr[1] + kr[2], nsl = (rows + VEC - 1) / VEC;
            const long long row0 = (long long)slab_src.size() * VEC;  // first t row of the group
            long long start = 0;
            for (int c = 3 * gi; c < 3 * gi + 3; ++c) {
                toff[6 * b + c] = (int)(tcount + row0 + start);
                start += ranks[6 * b + c];
            }
            const long long s0 = base + row0 * nc;
            vpack.resize((size_t)(s0 + nsl * VEC * nc), T(0));
            long long r = 0;
            for (int c = 3 * gi; c < 3 * gi + 3; ++c) {
                const double* Vb = V + voff[6 * b + c];
                for (long long l = 0; l < ranks[6 * b + c]; ++l, ++r) {
                    const long long sl = r / VEC, e = r % VEC;
                    for (long long jp = 0; jp < nc; ++jp) {
                        const int j = cinv[cinv_off[b] + jp];
                        vpack[(size_t)(s0 + (sl * nc + jp) * VEC + e)] = conv<T>(Vb[l * nc + j]);
                    }
                }
                const double ab = (double)ranks[6 * b + c] * nc * sizeof(T);
                h->alg += ab;
                h->alg_factors += ab;
                if (c != 2 && c != 5) h->alg_fix += ab;
            }
            const long long keep = kr[0] + kr[1];  // rows before Bs / Bd
            for (long long sl = 0; sl < nsl; ++sl) {
                slab_src.push_back(gi);
                slab_fix.push_back(sl * VEC >= keep ? 1 : 0);
            }
        }
        const long long nsl = (long long)slab_src.size();
        for (long long s0 = 0; s0 < nsl; s0 += S) {
            AItem a;
            memset(&a, 0, sizeof(a));
            a.v_off = base + s0 * VEC * nc;
            a.n = (int)nc;
            a.c0 = c0[b];
            a.t_off = (int)(tcount + s0 * VEC);
            a.nslab = (int)std::min<long long>(S, nsl - s0);
            a.ls = ls;
            for (int q = 0; q < a.nslab; ++q) {
                a.src_mask |= slab_src[s0 + q] << q;
                a.fix_mask |= slab_fix[s0 + q] << q;
            }
            aitems.push_back(a);
        }
        double sb = 0, sbf = 0;
        for (long long sl = 0; sl < nsl; ++sl) {
            sb += (double)VEC * nc * sizeof(T);
            if (!slab_fix[sl]) sbf += (double)VEC * nc * sizeof(T);
        }
        h->stored += sb;
        h->stored_fix += sbf;
        tcount += nsl * VEC;
    }
    // phase B items: per segment the dense leaves first, then the low-rank
    // blocks, in block order within each kind
    std::vector<BItem> bitems;
    std::vector<BWave> dwaves, lwaves;
    std::vector<float> dpack;
    std::vector<T> upack;
    long long slot = 0;
    for (size_t s = 0; s < segs.size(); ++s) {
        Seg& sg = segs[s];
        std::stable_partition(seg_blocks[s].begin(), seg_blocks[s].end(), [&](int b) { return kind[b] == 0; });
        const int first_item = (int)bitems.size();
        for (int b : seg_blocks[s]) {
            const long long m = row_off[b + 1] - row_off[b], nc = col_off[b + 1] - col_off[b];
            const int ms = sg.ms;
            const int lr0 = sg.s0 - r0[b];  // first block-local cluster row of the segment
            BItem it;
            memset(&it, 0, sizeof(it));
            it.s0 = sg.s0;
            it.ms = ms;
            it.slot = (int)slot;
            it.type = kind[b];
            if (kind[b] == 0) {
                // columns padded to a multiple of 4 rows (16-byte vectors)
                const int msp = (ms + 3) & ~3;
                it.c0 = c0[b];
                it.n = (int)nc;
                const long long base = (long long)dpack.size();
                it.off[0] = base;
                dpack.resize((size_t)(base + 6 * nc * msp), 0.0f);
                const double* Db = dense + doff[b];
                // [column][matrix][row]: a column's six matrices are one
                // contiguous run, A1s, A2s, A1d, A2d first and Bs, Bd last
                // (the four a fix_tn product reads)
                static const int pos[6] = {0, 1, 4, 2, 3, 5};
                for (int c = 0; c < 6; ++c)
                    for (long long jp = 0; jp < nc; ++jp) {
                        const int j = cinv[cinv_off[b] + jp];
                        for (int ip = 0; ip < ms; ++ip) {
                            const int i = rinv[rinv_off[b] + lr0 + ip];
                            dpack[(size_t)(base + (jp * 6 + pos[c]) * msp + ip)] =
                                (float)Db[(long long)c * m * nc + (long long)j * m + i];
                        }
                    }
                const double ab = 6.0 * nc * ms * 4, sb = 6.0 * nc * msp * 4;
                h->alg += ab;
                h->alg_fix += ab * 4 / 6;
                h->stored += sb;
                h->stored_fix += sb * 4 / 6;
                h->alg_dense += ab;
            } else {
                // U columns padded to a multiple of VEC rows (16-byte vectors);
                // per output o the columns of matrix o, then of matrix o + 3
                const int mpu = (ms + VEC - 1) / VEC * VEC;
                for (int c = 0; c < 6; ++c) {
                    it.k[c] = (int)ranks[6 * b + c];
                    it.t_off[c] = toff[6 * b + c];
                }
                for (int o = 0; o < 3; ++o) {
                    const long long base = (long long)upack.size();
                    it.off[o] = base;
                    const long long kk = ranks[6 * b + o] + ranks[6 * b + o + 3];
                    upack.resize((size_t)(base + kk * mpu), T(0));
                    long long col = 0;
                    for (int c = o; c < 6; c += 3) {
                        const long long k = ranks[6 * b + c];
                        const double* Ub = U + uoff[6 * b + c];
                        for (long long l = 0; l < k; ++l, ++col)
                            for (int ip = 0; ip < ms; ++ip) {
                                const int i = rinv[rinv_off[b] + lr0 + ip];
                                upack[(size_t)(base + col * mpu + ip)] = conv<T>(Ub[l * m + i]);
                            }
                        const double ab = (double)k * ms * sizeof(T), sb = (double)k * mpu * sizeof(T);
                        h->alg += ab;
                        h->stored += sb;
                        h->alg_factors += ab;
                        if (c != 2 && c != 5) {
                            h->alg_fix += ab;
                            h->stored_fix += sb;
                        }
                    }
                }
            }
            bitems.push_back(it);
        }
        // waves: one per dense item, kLowRankPerWave low-rank items per wave
        // (many small waves hide memory latency better: 8 per wave measured
        // 15% slower than 1 on the 60k product)
        sg.slot0 = (int)slot;
        const int last_item = (int)bitems.size();
        for (int i = first_item; i < last_item;) {
            int cnt = 1;
            if (bitems[i].type == 1)
                while (cnt < kLowRankPerWave && i + cnt < last_item && bitems[i + cnt].type == 1) ++cnt;
            (bitems[i].type == 0 ? dwaves : lwaves).push_back(BWave{i, cnt, (int)slot, 0});
            for (int q = 0; q < cnt; ++q) bitems[i + q].slot = (int)slot;
            slot += 3 * sg.ms;
            ++sg.nw;
            i += cnt;
        }
    }
    // dense waves first (launch 2), then the low-rank waves (launch 3)
    std::vector<BWave> bwaves(dwaves);
    bwaves.insert(bwaves.end(), lwaves.begin(), lwaves.end());
    PQ_CHECK(slot < INT32_MAX && tcount < INT32_MAX, "pq3d_hmat_create: H-matrix too large for 32-bit offsets");
    // upload
    h->n = n;
    h->na = (long long)aitems.size();
    h->nbi = (long long)bitems.size();
    h->nbw = (long long)bwaves.size();
    h->nbd = (long long)dwaves.size();
    h->nseg = (long long)segs.size();
    h->nt = tcount;
    h->npart = slot;
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        PQ_HIP(hipMalloc(dst, bytes ? bytes : 16));
        if (bytes) PQ_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if (up((void**)&h->perm, perm, n * sizeof(int))) return -1;
    if (up((void**)&h->a_items, aitems.data(), aitems.size() * sizeof(AItem))) return -1;
    if (up((void**)&h->b_items, bitems.data(), bitems.size() * sizeof(BItem))) return -1;
    if (up((void**)&h->b_waves, bwaves.data(), bwaves.size() * sizeof(BWave))) return -1;
    if (up((void**)&h->segs, segs.data(), segs.size() * sizeof(Seg))) return -1;
    if (up((void**)&h->ddata, dpack.data(), dpack.size() * sizeof(float))) return -1;
    if (up(&h->udata, upack.data(), upack.size() * sizeof(T))) return -1;
    if (up(&h->vdata, vpack.data(), vpack.size() * sizeof(T))) return -1;
    PQ_HIP(hipMalloc((void**)&h->xp, 2 * n * sizeof(float)));
    PQ_HIP(hipMalloc((void**)&h->t, (tcount + 1) * sizeof(float)));
    PQ_HIP(hipMalloc((void**)&h->partial, (slot + 1) * sizeof(float)));
    PQ_HIP(hipMalloc((void**)&h->xin, 2 * n * sizeof(double)));
    PQ_HIP(hipMalloc((void**)&h->yout, 3 * n * sizeof(double)));
    PQ_HIP(hipMemset(h->xin, 0, 2 * n * sizeof(double)));
    PQ_HIP(hipMemset(h->t, 0, (tcount + 1) * sizeof(float)));
    return 0;
}

static void hmat_free(HMat* h) {
    void* ptrs[] = {h->perm, h->a_items, h->b_items, h->b_waves, h->segs, h->ddata, h->udata,
                    h->vdata, h->xp, h->t, h->partial, h->xin, h->yout};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete h;
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

// Build the device H-matrix.  perm (n): cluster position -> element id.
// Block b: rows row_idx[row_off[b]:row_off[b+1]], columns col_idx[...] (element
// ids; each must be a contiguous range of the cluster order); kind[b] = 0
// dense (six m x n column-major FP64 matrices in `dense`, blocks back to
// back), 1 low-rank (ranks[6b+c]; U m x k column-major in `U`, V k x n
// row-major in `V`, matrices A1s..Bd, blocks back to back).  factor_type:
// PQ3D_MAT_F32 or PQ3D_MAT_BF16 for the low-rank factors (dense leaves FP32).
int pq3d_hmat_create(void** out, long long n, const int* perm, long long nb, const long long* row_off,
                     const int* row_idx, const long long* col_off, const int* col_idx, const int* kind,
                     const long long* ranks, const double* dense, const double* U, const double* V,
                     int factor_type) {
    clear_error();
    PQ_CHECK(n > 0 && nb > 0, "pq3d_hmat_create: empty H-matrix");
    PQ_CHECK(factor_type == MAT_F32 || factor_type == MAT_BF16, "pq3d_hmat_create: factor type must be f32 or bf16");
    HMat* h = new HMat();
    h->ftype = factor_type;
    int rc = factor_type == MAT_BF16
                 ? pack<unsigned short>(h, n, perm, nb, row_off, row_idx, col_off, col_idx, kind, ranks, dense, U, V)
                 : pack<float>(h, n, perm, nb, row_off, row_idx, col_off, col_idx, kind, ranks, dense, U, V);
    if (rc) {
        hmat_free(h);
        return rc;
    }
    *out = h;
    return 0;
}

int pq3d_hmat_destroy(void* hm) {
    clear_error();
    if (hm) hmat_free((HMat*)hm);
    return 0;
}

// info10: dense-leaf bytes, factor bytes (unpadded), algorithmic bytes per
// product, same with fix_tn, segments, phase-B waves, phase-A waves, low-rank
// blocks, stored bytes read per product (padded), same with fix_tn
int pq3d_hmat_info(void* hm, double* info10) {
    clear_error();
    HMat* h = (HMat*)hm;
    PQ_CHECK(h, "pq3d_hmat_info: null");
    info10[0] = h->alg_dense;
    info10[1] = h->alg_factors;
    info10[2] = h->alg;
    info10[3] = h->alg_fix;
    info10[4] = (double)h->nseg;
    info10[5] = (double)h->nbw;
    info10[6] = (double)h->na;
    info10[7] = (double)h->lowrank_blocks;
    info10[8] = h->stored;
    info10[9] = h->stored_fix;
    return 0;
}

// y (host, 3 x n) = (A1s x1 + A1d x2, A2s x1 + A2d x2, Bs x1 + Bd x2); x1, x2 host
int pq3d_hmat_mvm(void* hm, const double* x1, const double* x2, int fix_tn, double* y) {
    clear_error();
    HMat* h = (HMat*)hm;
    PQ_CHECK(h, "pq3d_hmat_mvm: null");
    const long long n = h->n;
    PQ_HIP(hipMemcpy(h->xin, x1, n * sizeof(double), hipMemcpyHostToDevice));
    PQ_HIP(hipMemcpy(h->xin + n, x2, n * sizeof(double), hipMemcpyHostToDevice));
    Launch L;
    hmat_launch(h, h->xin, h->xin + n, nullptr, nullptr, 0.0, fix_tn, h->yout, L);
    PQ_HIP(L.err);
    PQ_HIP(hipMemcpy(y, h->yout, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// ms per product over `reps` back-to-back products on the last inputs
int pq3d_hmat_time_mvm(void* hm, int reps, int fix_tn, float* ms_per) {
    clear_error();
    HMat* h = (HMat*)hm;
    PQ_CHECK(h && reps > 0, "pq3d_hmat_time_mvm: bad arguments");
    const long long n = h->n;
    hipEvent_t a, b;
    PQ_HIP(hipEventCreate(&a));
    PQ_HIP(hipEventCreate(&b));
    Launch L;
    hmat_launch(h, h->xin, h->xin + n, nullptr, nullptr, 0.0, fix_tn, h->yout, L);
    PQ_HIP(hipEventRecord(a, 0));
    for (int r = 0; r < reps; ++r) hmat_launch(h, h->xin, h->xin + n, nullptr, nullptr, 0.0, fix_tn, h->yout, L);
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

}  // extern "C"
