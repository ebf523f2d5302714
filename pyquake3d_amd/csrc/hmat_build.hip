// H-matrix construction on the GPU: batched evaluation of traction Green's
// function sub-blocks (rows, columns, dense leaves) and adaptive cross
// approximation (ACA) of the admissible blocks.
//
// Reference interface replaced: the ACA of src/Hmatrix.py:661-1289
// (ACA_plus_full_traceS/D) with its row/column evaluators :329-658, and the
// dense leaves of ACA_worker :1292-1344.  The reference evaluates one row or
// column per ctypes call and pivots one block at a time; here every sweep
// evaluates the pending pivot rows/columns of ALL blocks in one launch, and
// the pivot logic of all blocks advances in lock step: on host threads
// (pq3d_aca_run, evaluation by ``gf_segments_kernel``), or on the device, one
// wave per (block, matrix), with the recompression there too
// (pq3d_aca_run_device, the ``aca_dev_*`` kernels).
//
// ACA per (block, matrix): partial pivoting with a reference row and column
// (ACA+): the next pivot is led by whichever of the reference row/column
// residuals holds the larger entry; a term u v^T is added with v = R[i*,:] /
// R[i*,j*], u = R[:,j*].  Stop when ||u|| ||v|| <= eps ||S_k||_F (running
// Frobenius norm of the approximation) and the reference residuals are below
// eps times the approximation's RMS entry, or when no unused row/column is
// left.  Entries below 1e-12 (pivots) or terms below 1e-9 (Frobenius) of the
// block's largest entry over all six matrices are treated as zero, so a
// matrix that vanishes by symmetry gets rank 0 instead of a fit to rounding
// noise.
//
// A negative eps selects the reference's absolute rule instead
// (Hmatrix.py:797,832-838, called with eps = 0.01 at :1298-1307): stop right
// after a term whose pivot is below 1e-4 or whose step ||u|| ||v|| is below
// |eps| (both in the matrices' units, Pa per metre of slip); no reference-
// residual check.  The reference pivots on A1s (A1d) and stops all three
// matrices of a slip direction by its criteria; here each matrix is its own
// instance and stops by its own.
#include "common.hpp"
#include "td_field.hpp"

#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdint>
#include <thread>
#include <unordered_map>
#include <vector>

extern "C" const char* pq3d_last_error(void);

namespace pq3d {

constexpr double ACA_REF_PIVOT = 1e-4;  // pivot floor of the reference's absolute stopping rule

__global__ void gfsrc_setup_kernel(const double* __restrict__ P, long long n, SrcGeom<2>* __restrict__ out) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double b[2][3] = {{0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};  // strike (Ss = 1), dip (Ds = 1)
    SrcGeom<2> g;
    src_setup<2>(P + 3 * j, P + 3 * (n + j), P + 3 * (2 * n + j), b, g);
    out[j] = g;
}

// one slip direction per table: strike for sources [0, n), dip for [n, 2n)
__global__ void gfsrc_setup1_kernel(const double* __restrict__ P, long long n, SrcGeom<1>* __restrict__ out) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 2 * n) return;
    const long long t = j < n ? j : j - n;
    const double b[2][1][3] = {{{0.0, 1.0, 0.0}}, {{0.0, 0.0, 1.0}}};
    SrcGeom<1> g;
    src_setup<1>(P + 3 * t, P + 3 * (n + t), P + 3 * (2 * n + t), b[j < n ? 0 : 1], g);
    out[j] = g;
}

struct GFSource {
    long long nsrc = 0, nrcv = 0;
    SrcGeom<2>* geo = nullptr;
    SrcGeom<1>* geo1 = nullptr;  // strike-only then dip-only tables (2 nsrc)
    double* xg = nullptr;
    double* ev = nullptr;
    double mu = 0, lam = 0, nu = 0;
    int halfspace = 1;
    // per-call buffers (grow only)
    void* seg = nullptr;
    size_t seg_cap = 0;
    double* out = nullptr;
    size_t out_cap = 0;
    int* idx = nullptr;
    size_t idx_cap = 0;
};

// One sub-block request: rows ridx[rcv_off : rcv_off+m] x columns
// sidx[src_off : src_off+n]; the pairs are numbered p = j*m + i from
// pair_off; matrices selected by `mask` (bit c for A1s, A2s, Bs, A1d, A2d, Bd)
// are written, in that order, each m*n values [column][row], from out_off.
struct SegDesc {
    long long pair_off;
    long long out_off;
    int rcv_off, m, src_off, n;
    int mask, pad;
};

static int grow(void** p, size_t* cap, size_t bytes) {
    if (bytes <= *cap) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t want = bytes + bytes / 4 + 4096;
    PQ_HIP(hipMalloc(p, want));
    *cap = want;
    return 0;
}

__global__ __launch_bounds__(256) void gf_segments_kernel(const SrcGeom<2>* __restrict__ geo,
                                                          const double* __restrict__ xg,
                                                          const double* __restrict__ ev,
                                                          const int* __restrict__ ridx,
                                                          const int* __restrict__ sidx,
                                                          const SegDesc* __restrict__ seg, int nseg,
                                                          long long total, double mu, double lam, double nu,
                                                          int halfspace, double* __restrict__ out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    int lo = 0, hi = nseg - 1;  // last segment with pair_off <= q (it has pairs)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].pair_off <= q) lo = mid;
        else hi = mid - 1;
    }
    const SegDesc d = seg[lo];
    const long long p = q - d.pair_off;
    const int i = (int)(p % d.m), j = (int)(p / d.m);
    const int r = ridx[d.rcv_off + i], s = sidx[d.src_off + j];
    const double X = xg[3 * r + 0], Y = xg[3 * r + 1], Z = xg[3 * r + 2];
    double e[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = ev[9 * (long long)r + k];
    double G[2][9];
    td_pair_grad<double, 2, false>(geo[s], X, Y, Z, nu, halfspace != 0, G);
    const long long mn = (long long)d.m * d.n;
    int slot = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        double eps[6], sig[6], t[3];
        grad_to_strain_stress<double>(G[b], mu, lam, eps, sig);
        stress_to_traction<double>(sig, e, t);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if ((d.mask >> (b * 3 + k)) & 1) {
                out[d.out_off + slot * mn + p] = t[k];
                ++slot;
            }
    }
}

static inline int popc(int x) { return __builtin_popcount((unsigned)x); }

static int d2h_staged(void* dst, const void* src, size_t bytes);

// Evaluate `segs` (descriptors with pair_off/out_off filled) whose index
// lists are already on the device; results copied into host `out`.
static int run_segments(GFSource* g, const int* d_ridx, const int* d_sidx, std::vector<SegDesc>& segs,
                        long long total_pairs, long long total_out, double* host_out, double* gpu_ms) {
    if (segs.empty() || total_pairs == 0) return 0;
    if (grow(&g->seg, &g->seg_cap, segs.size() * sizeof(SegDesc))) return -1;
    if (grow((void**)&g->out, &g->out_cap, (size_t)total_out * sizeof(double))) return -1;
    PQ_HIP(hipMemcpy(g->seg, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice));
    hipEvent_t a, b;
    PQ_HIP(hipEventCreate(&a));
    PQ_HIP(hipEventCreate(&b));
    PQ_HIP(hipEventRecord(a, 0));
    const long long nblk = (total_pairs + 255) / 256;
    PQ_CHECK(nblk < (1LL << 31), "segment evaluation too large (%lld pairs)", total_pairs);
    hipLaunchKernelGGL(gf_segments_kernel, dim3((unsigned)nblk), dim3(256), 0, 0, g->geo, g->xg, g->ev, d_ridx,
                       d_sidx, (const SegDesc*)g->seg, (int)segs.size(), total_pairs, g->mu, g->lam, g->nu,
                       g->halfspace, g->out);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipEventRecord(b, 0));
    if (d2h_staged(host_out, g->out, (size_t)total_out * sizeof(double))) return -1;
    float ms = 0;
    PQ_HIP(hipEventElapsedTime(&ms, a, b));
    if (gpu_ms) *gpu_ms += ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 0;
}

// ---------------------------------------------------------------------------
// ACA state of one (block, matrix)
// ---------------------------------------------------------------------------
enum AcaState : int { ST_REFS = 0, ST_LEAD = 1, ST_FOLLOW = 2, ST_DONE = 3 };

struct Req {
    int b;
    int type;  // 0 row, 1 column
    int idx;
    int mask;
};

struct Inst {
    int b = 0, c = 0, m = 0, n = 0, kmax = 0;
    int k = 0;
    std::vector<double> U, V;               // U: column l at l*m; V: row l at l*n
    std::vector<unsigned char> ur, uc;      // rows / columns used as pivots
    int iref = -1, jref = -1;
    std::vector<double> Rr, Rc;             // residual of the reference row / column
    std::vector<double> r, col;             // residual of the pivot row / column
    double norm2 = 0.0;
    double scale = -1.0;                    // largest |entry| of the block over the six matrices
    int state = ST_REFS;
    bool need_rr = true, need_rc = true;
    bool lead_row = false;
    int istar = -1, jstar = -1;
    int tries = 0;
    bool pending_conv = false;
    bool converged = false;
    uint64_t rng = 0;
    // requests issued this sweep: type, idx, request number
    int nq = 0;
    int qt[2], qi[2], qn[2];
};

static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int pick_unused(Inst& s, const std::vector<unsigned char>& used, int len) {
    // random start, then the first unused index from there
    int start = (int)(splitmix(s.rng) % (uint64_t)len);
    for (int t = 0; t < len; ++t) {
        int x = (start + t) % len;
        if (!used[x]) return x;
    }
    return -1;
}

static void push_req(Inst& s, int type, int idx) {
    s.qt[s.nq] = type;
    s.qi[s.nq] = idx;
    s.qn[s.nq] = -1;
    ++s.nq;
}

struct Results {
    const std::vector<Req>* reqs;
    const std::vector<SegDesc>* segs;
    const double* out;
};

// raw value vector of matrix c for request q (length m or n)
static const double* raw_of(const Results& R, int q, int c) {
    const Req& rq = (*R.reqs)[q];
    const SegDesc& d = (*R.segs)[q];
    const int slot = popc(rq.mask & ((1 << c) - 1));
    return R.out + d.out_off + (long long)slot * d.m * d.n;
}

static void residual_row(const Inst& s, const double* raw, int i, std::vector<double>& out) {
    out.assign(raw, raw + s.n);
    for (int l = 0; l < s.k; ++l) {
        const double a = s.U[(size_t)l * s.m + i];
        if (a == 0.0) continue;
        const double* v = s.V.data() + (size_t)l * s.n;
        for (int j = 0; j < s.n; ++j) out[j] -= a * v[j];
    }
}

static void residual_col(const Inst& s, const double* raw, int j, std::vector<double>& out) {
    out.assign(raw, raw + s.m);
    for (int l = 0; l < s.k; ++l) {
        const double a = s.V[(size_t)l * s.n + j];
        if (a == 0.0) continue;
        const double* u = s.U.data() + (size_t)l * s.m;
        for (int i = 0; i < s.m; ++i) out[i] -= a * u[i];
    }
}

static int argmax_unused(const std::vector<double>& x, const std::vector<unsigned char>& used, double* best) {
    int arg = -1;
    double b = -1.0;
    for (size_t t = 0; t < x.size(); ++t) {
        if (used[t]) continue;
        const double a = std::fabs(x[t]);
        if (a > b || (std::isnan(a) && b >= 0)) {
            b = a;
            arg = (int)t;
        }
    }
    *best = b < 0 ? 0.0 : b;
    return arg;
}

static double maxabs(const std::vector<double>& x) {
    double b = 0.0;
    for (double v : x) b = std::max(b, std::fabs(v));
    return b;
}

static void start_refs(Inst& s) {
    if (s.need_rr) {
        s.iref = pick_unused(s, s.ur, s.m);
        if (s.iref < 0) {
            s.need_rr = false;
            s.Rr.assign(s.n, 0.0);
        } else {
            push_req(s, 0, s.iref);
        }
    }
    if (s.need_rc) {
        s.jref = pick_unused(s, s.uc, s.n);
        if (s.jref < 0) {
            s.need_rc = false;
            s.Rc.assign(s.m, 0.0);
        } else {
            push_req(s, 1, s.jref);
        }
    }
    s.state = s.nq ? ST_REFS : ST_LEAD;
}

// choose the next pivot (or finish); called with valid reference residuals
static void decide(Inst& s, double eps) {
    const double tiny = 1e-12 * s.scale;
    const double rms = std::sqrt(s.norm2 / ((double)s.m * (double)s.n));
    double ar = 0, ac = 0;
    const int jc = argmax_unused(s.Rr, s.uc, &ar);
    const int ic = argmax_unused(s.Rc, s.ur, &ac);
    if (s.pending_conv) {
        s.pending_conv = false;
        if (ar <= eps * rms + tiny && ac <= eps * rms + tiny) {
            s.converged = true;
            s.state = ST_DONE;
            return;
        }
    }
    if (s.k >= s.kmax || (jc < 0 && ic < 0)) {
        s.converged = (jc < 0 && ic < 0) || s.k >= std::min(s.m, s.n);
        s.state = ST_DONE;
        return;
    }
    if (!(ar > tiny) && !(ac > tiny)) {
        // nothing left in the sampled residual: resample the references a few times
        if (++s.tries > 3) {
            s.converged = true;
            s.state = ST_DONE;
            return;
        }
        if (s.iref >= 0) s.ur[s.iref] = 1;
        if (s.jref >= 0) s.uc[s.jref] = 1;
        s.need_rr = s.need_rc = true;
        start_refs(s);
        if (s.state == ST_LEAD) decide(s, eps);
        return;
    }
    if (ac > ar) {
        s.lead_row = true;
        s.istar = ic;
        push_req(s, 0, ic);
    } else {
        s.lead_row = false;
        s.jstar = jc;
        push_req(s, 1, jc);
    }
    s.state = ST_LEAD;
}

static void add_term(Inst& s, double eps) {
    const double delta = s.r[s.jstar];
    const double tiny = 1e-12 * s.scale;
    s.ur[s.istar] = 1;
    s.uc[s.jstar] = 1;
    if (!(std::fabs(delta) > tiny)) {
        decide(s, eps);
        return;
    }
    // u = residual column, v = residual row / pivot
    const int m = s.m, n = s.n, k = s.k;
    s.U.resize((size_t)(k + 1) * m);
    s.V.resize((size_t)(k + 1) * n);
    double* u = s.U.data() + (size_t)k * m;
    double* v = s.V.data() + (size_t)k * n;
    for (int i = 0; i < m; ++i) u[i] = s.col[i];
    const double inv = 1.0 / delta;
    for (int j = 0; j < n; ++j) v[j] = s.r[j] * inv;
    double uu = 0, vv = 0;
    for (int i = 0; i < m; ++i) uu += u[i] * u[i];
    for (int j = 0; j < n; ++j) vv += v[j] * v[j];
    double cross = 0;
    for (int l = 0; l < k; ++l) {
        const double* ul = s.U.data() + (size_t)l * m;
        const double* vl = s.V.data() + (size_t)l * n;
        double a = 0, b = 0;
        for (int i = 0; i < m; ++i) a += ul[i] * u[i];
        for (int j = 0; j < n; ++j) b += vl[j] * v[j];
        cross += a * b;
    }
    s.norm2 = std::max(0.0, s.norm2 + 2.0 * cross + uu * vv);
    s.k = k + 1;
    // reference residuals
    if (s.iref >= 0) {
        const double a = u[s.iref];
        for (int j = 0; j < n; ++j) s.Rr[j] -= a * v[j];
        if (s.ur[s.iref]) s.need_rr = true;
    }
    if (s.jref >= 0) {
        const double a = v[s.jref];
        for (int i = 0; i < m; ++i) s.Rc[i] -= a * u[i];
        if (s.uc[s.jref]) s.need_rc = true;
    }
    const double step = std::sqrt(uu * vv);
    if (eps < 0) {
        // the reference's absolute rule (Hmatrix.py:832-838): stop after a
        // term whose pivot is below 1e-4 or whose step ||u|| ||v|| is below |eps|
        if (std::fabs(delta) < ACA_REF_PIVOT || step < -eps) {
            s.converged = true;
            s.state = ST_DONE;
            return;
        }
        s.pending_conv = false;
    } else {
        const double floor = 1e-9 * s.scale * std::sqrt((double)m * (double)n);
        s.pending_conv = step <= eps * std::sqrt(s.norm2) || step <= floor;
    }
    s.tries = 0;
    if (s.need_rr || s.need_rc) {
        start_refs(s);
        if (s.state == ST_LEAD) decide(s, eps);
    } else {
        decide(s, eps);
    }
}

// consume this sweep's results for one instance and issue its next requests
static void consume(Inst& s, const Results& R, double eps) {
    const int nq = s.nq;
    int qt[2], qi[2], qn[2];
    for (int t = 0; t < nq; ++t) {
        qt[t] = s.qt[t];
        qi[t] = s.qi[t];
        qn[t] = s.qn[t];
    }
    s.nq = 0;
    if (s.scale < 0) {
        // first results: reference row and column with all six matrices
        double sc = 0;
        for (int t = 0; t < nq; ++t) {
            const SegDesc& d = (*R.segs)[qn[t]];
            const int cnt = popc((*R.reqs)[qn[t]].mask);
            const double* p = R.out + d.out_off;
            const long long len = (long long)cnt * d.m * d.n;
            for (long long x = 0; x < len; ++x) sc = std::max(sc, std::fabs(p[x]));
        }
        s.scale = sc;
    }
    switch (s.state) {
        case ST_REFS: {
            for (int t = 0; t < nq; ++t) {
                const double* raw = raw_of(R, qn[t], s.c);
                if (qt[t] == 0) {
                    residual_row(s, raw, qi[t], s.Rr);
                    s.need_rr = false;
                } else {
                    residual_col(s, raw, qi[t], s.Rc);
                    s.need_rc = false;
                }
            }
            decide(s, eps);
            break;
        }
        case ST_LEAD: {
            const double* raw = raw_of(R, qn[0], s.c);
            double best = 0;
            if (s.lead_row) {
                residual_row(s, raw, s.istar, s.r);
                s.jstar = argmax_unused(s.r, s.uc, &best);
                if (s.jstar < 0 || !(best > 1e-12 * s.scale)) {
                    s.ur[s.istar] = 1;
                    decide(s, eps);
                    break;
                }
                push_req(s, 1, s.jstar);
            } else {
                residual_col(s, raw, s.jstar, s.col);
                s.istar = argmax_unused(s.col, s.ur, &best);
                if (s.istar < 0 || !(best > 1e-12 * s.scale)) {
                    s.uc[s.jstar] = 1;
                    decide(s, eps);
                    break;
                }
                push_req(s, 0, s.istar);
            }
            s.state = ST_FOLLOW;
            break;
        }
        case ST_FOLLOW: {
            const double* raw = raw_of(R, qn[0], s.c);
            if (s.lead_row) residual_col(s, raw, s.jstar, s.col);
            else residual_row(s, raw, s.istar, s.r);
            add_term(s, eps);
            break;
        }
        default:
            break;
    }
}

template <class F>
static void parallel_for(long long n, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < 64 || nt == 1) {
        for (long long i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (long long i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}

// Device -> pageable host copy through two pinned 64 MB buffers: the PCIe
// transfer of one chunk overlaps the host threads' copy (and first-touch page
// faults) of the previous one.  Copies of one chunk or less go direct.
static int d2h_staged(void* dst, const void* src, size_t bytes) {
    const size_t CH = (size_t)64 << 20;
    if (bytes <= CH) {
        PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return 0;
    }
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipHostMalloc(&pin[0], CH, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&pin[1], CH, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    const size_t nch = (bytes + CH - 1) / CH;
    auto len = [&](size_t c) { return std::min(CH, bytes - c * CH); };
    auto host_copy = [&](size_t c) {
        char* d = (char*)dst + c * CH;
        const char* sp = (const char*)pin[c & 1];
        const size_t L = len(c), parts = 64, per = (L + parts - 1) / parts;
        parallel_for((long long)parts, [&](long long q) {
            const size_t o = (size_t)q * per;
            if (o < L) memcpy(d + o, sp + o, std::min(per, L - o));
        });
    };
    for (size_t c = 0; c < nch && e == hipSuccess; ++c) {
        e = hipMemcpyAsync(pin[c & 1], (const char*)src + c * CH, len(c), hipMemcpyDeviceToHost, 0);
        if (e == hipSuccess) e = hipEventRecord(ev[c & 1], 0);
        if (e == hipSuccess && c >= 1) {
            e = hipEventSynchronize(ev[(c - 1) & 1]);
            if (e == hipSuccess) host_copy(c - 1);
        }
    }
    if (e == hipSuccess) {
        e = hipEventSynchronize(ev[(nch - 1) & 1]);
        if (e == hipSuccess) host_copy(nch - 1);
    }
    for (int k = 0; k < 2; ++k) {
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        if (pin[k]) (void)hipHostFree(pin[k]);
    }
    if (e != hipSuccess) {
        set_error("device to host copy: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Recompression of U V (m x k, k x n): QR of both factors by Gram-Schmidt
// with reorthogonalisation, one-sided Jacobi SVD of the k x k core, smallest
// rank whose discarded singular values have Frobenius norm <= eps ||U V||_F;
// the singular values are split evenly between the factors.
// ---------------------------------------------------------------------------
// A: rows x k, column j at j*rows -> Q (same layout), R (k x k, R[i*k + j])
static void qr_mgs2(int rows, int k, const double* A, std::vector<double>& Q, std::vector<double>& R) {
    Q.assign(A, A + (size_t)rows * k);
    R.assign((size_t)k * k, 0.0);
    for (int j = 0; j < k; ++j) {
        double* v = Q.data() + (size_t)j * rows;
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < j; ++i) {
                const double* q = Q.data() + (size_t)i * rows;
                double d = 0;
                for (int t = 0; t < rows; ++t) d += q[t] * v[t];
                R[(size_t)i * k + j] += d;
                for (int t = 0; t < rows; ++t) v[t] -= d * q[t];
            }
        double nv = 0;
        for (int t = 0; t < rows; ++t) nv += v[t] * v[t];
        nv = std::sqrt(nv);
        R[(size_t)j * k + j] = nv;
        if (nv > 0)
            for (int t = 0; t < rows; ++t) v[t] /= nv;
    }
}

// C (k x k, C[i*k + j]) = W diag(sig) Z^T; W, Z column j at j*k; sig descending
static void jacobi_svd(int k, const std::vector<double>& C, std::vector<double>& W, std::vector<double>& sig,
                       std::vector<double>& Z) {
    std::vector<double> A((size_t)k * k);  // columns of C
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) A[(size_t)j * k + i] = C[(size_t)i * k + j];
    std::vector<double> V((size_t)k * k, 0.0);
    for (int j = 0; j < k; ++j) V[(size_t)j * k + j] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                double* ap = A.data() + (size_t)p * k;
                double* aq = A.data() + (size_t)q * k;
                double al = 0, be = 0, ga = 0;
                for (int t = 0; t < k; ++t) {
                    al += ap[t] * ap[t];
                    be += aq[t] * aq[t];
                    ga += ap[t] * aq[t];
                }
                if (ga == 0.0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int x = 0; x < k; ++x) {
                    const double u = ap[x], w = aq[x];
                    ap[x] = c * u - s * w;
                    aq[x] = s * u + c * w;
                }
                double* vp = V.data() + (size_t)p * k;
                double* vq = V.data() + (size_t)q * k;
                for (int x = 0; x < k; ++x) {
                    const double u = vp[x], w = vq[x];
                    vp[x] = c * u - s * w;
                    vq[x] = s * u + c * w;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> nrm(k);
    std::vector<int> ord(k);
    for (int j = 0; j < k; ++j) {
        double a = 0;
        for (int t = 0; t < k; ++t) a += A[(size_t)j * k + t] * A[(size_t)j * k + t];
        nrm[j] = std::sqrt(a);
        ord[j] = j;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return nrm[a] > nrm[b]; });
    W.assign((size_t)k * k, 0.0);
    Z.assign((size_t)k * k, 0.0);
    sig.assign(k, 0.0);
    for (int jj = 0; jj < k; ++jj) {
        const int j = ord[jj];
        sig[jj] = nrm[j];
        for (int t = 0; t < k; ++t) {
            W[(size_t)jj * k + t] = nrm[j] > 0 ? A[(size_t)j * k + t] / nrm[j] : 0.0;
            Z[(size_t)jj * k + t] = V[(size_t)j * k + t];
        }
    }
}

// U: m x k (column l at l*m), V: k x n (row l at l*n) -> rank r, factors in
// place (first r columns / rows)
static int recompress_factors(int m, int n, int k, std::vector<double>& U, std::vector<double>& V, double eps) {
    if (k == 0) return 0;
    std::vector<double> Qu, Ru, Qv, Rv;
    qr_mgs2(m, k, U.data(), Qu, Ru);
    qr_mgs2(n, k, V.data(), Qv, Rv);  // V's rows are the columns of V^T
    std::vector<double> C((size_t)k * k, 0.0);  // Ru Rv^T
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) {
            double a = 0;
            for (int t = 0; t < k; ++t) a += Ru[(size_t)i * k + t] * Rv[(size_t)j * k + t];
            C[(size_t)i * k + j] = a;
        }
    std::vector<double> W, sig, Z;
    jacobi_svd(k, C, W, sig, Z);
    double tot = 0;
    for (double x : sig) tot += x * x;
    int r = 0;
    if (tot > 0) {
        double tail = tot;
        while (r < k && tail > eps * eps * tot) {
            tail -= sig[r] * sig[r];
            ++r;
        }
    }
    std::vector<double> U2((size_t)m * r, 0.0), V2((size_t)r * n, 0.0);
    for (int l = 0; l < r; ++l) {
        const double h = std::sqrt(sig[l]);
        double* u = U2.data() + (size_t)l * m;
        double* v = V2.data() + (size_t)l * n;
        for (int i = 0; i < k; ++i) {
            const double w = W[(size_t)l * k + i] * h;
            if (w != 0.0) {
                const double* q = Qu.data() + (size_t)i * m;
                for (int t = 0; t < m; ++t) u[t] += w * q[t];
            }
            const double z = Z[(size_t)l * k + i] * h;
            if (z != 0.0) {
                const double* q = Qv.data() + (size_t)i * n;
                for (int t = 0; t < n; ++t) v[t] += z * q[t];
            }
        }
    }
    U.swap(U2);
    V.swap(V2);
    return r;
}

struct AcaResult {
    long long nb = 0;
    std::vector<Inst> inst;
    double stats[8] = {0};
    double phase_s[4] = {0};  // request assembly, evaluation incl. transfers, ACA logic, recompression
    // device path: blocks redone by the host loop (an instance reached kcap),
    // seconds of set-up (instances, buffers, uploads), sweeps, factor readback
    double dev[4] = {0};
    // device path: final factors in the fetch layout; instances whose factors
    // live on the host (redone there) are marked in host_inst
    double* dU = nullptr;
    double* dV = nullptr;
    long long usz = 0, vsz = 0;
    std::vector<unsigned char> host_inst;
    ~AcaResult() {
        if (dU) (void)hipFree(dU);
        if (dV) (void)hipFree(dV);
    }
};

// Instances of nb blocks (six matrices each) with their first requests: the
// reference row and column of block b, drawn from the seed and bid[b] (the
// block's id in the caller's numbering; b when bid is null).
static void aca_init_instances(long long nb, const long long* row_off, const long long* col_off,
                               const long long* bid, int max_rank, uint64_t seed, std::vector<Inst>& inst) {
    inst.assign((size_t)nb * 6, Inst());
    for (long long b = 0; b < nb; ++b) {
        const int m = (int)(row_off[b + 1] - row_off[b]), n = (int)(col_off[b + 1] - col_off[b]);
        const long long id = bid ? bid[b] : b;
        uint64_t bs = seed ^ (0x51ED270B27A5ull * (uint64_t)(id + 1));
        const int i0 = (int)(splitmix(bs) % (uint64_t)m), j0 = (int)(splitmix(bs) % (uint64_t)n);
        for (int c = 0; c < 6; ++c) {
            Inst& s = inst[(size_t)b * 6 + c];
            s.b = (int)b;
            s.c = c;
            s.m = m;
            s.n = n;
            s.kmax = std::min(m, n);
            if (max_rank > 0) s.kmax = std::min(s.kmax, max_rank);
            s.ur.assign(m, 0);
            s.uc.assign(n, 0);
            s.Rr.assign(n, 0.0);
            s.Rc.assign(m, 0.0);
            s.rng = bs ^ (0x9E37ull * (uint64_t)(c + 1));
            // all six matrices of a block start from the same reference row/column
            s.iref = i0;
            s.jref = j0;
            push_req(s, 0, i0);
            push_req(s, 1, j0);
            s.state = ST_REFS;
        }
    }
}

struct AcaCounters {
    long long sweeps = 0;
    double pairs = 0, gpu_ms = 0;
    double phase_s[3] = {0, 0, 0};
};

// The sweeps of the host-side pivoting: every sweep evaluates the pending
// requests of all instances in one launch and advances each instance.
// d_idx: row ids (row_off[nb] of them) then column ids on the device.
static int aca_host_loop(GFSource* g, const int* d_idx, const long long* row_off, const long long* col_off,
                         long long nb, double eps, int max_rank, std::vector<Inst>& inst, AcaCounters& cnt) {
    const long long nr = row_off[nb];
    std::vector<Req> reqs;
    std::vector<SegDesc> segs;
    std::vector<double> out;
    std::unordered_map<uint64_t, int> key;
    const long long max_sweeps = 4LL * (long long)(max_rank > 0 ? max_rank : 4096) + 64;
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    while (true) {
        const auto t0 = clk::now();
        reqs.clear();
        key.clear();
        for (auto& s : inst) {
            if (s.state == ST_DONE) continue;
            for (int t = 0; t < s.nq; ++t) {
                const uint64_t kk = ((uint64_t)s.b << 33) | ((uint64_t)s.qt[t] << 32) | (uint32_t)s.qi[t];
                auto it = key.find(kk);
                int q;
                if (it == key.end()) {
                    q = (int)reqs.size();
                    key.emplace(kk, q);
                    reqs.push_back(Req{s.b, s.qt[t], s.qi[t], 0});
                } else {
                    q = it->second;
                }
                reqs[q].mask |= (s.scale < 0 ? 0x3f : (1 << s.c));
                s.qn[t] = q;
            }
        }
        if (reqs.empty()) return 0;
        if (++cnt.sweeps > max_sweeps) {
            set_error("pq3d_aca_run: no convergence after %lld sweeps", cnt.sweeps);
            return -3;
        }
        segs.resize(reqs.size());
        long long pairs = 0, outs = 0;
        for (size_t q = 0; q < reqs.size(); ++q) {
            const Req& r = reqs[q];
            const long long ro = row_off[r.b], co = col_off[r.b];
            const int m = (int)(row_off[r.b + 1] - ro), n = (int)(col_off[r.b + 1] - co);
            SegDesc d;
            if (r.type == 0) d = SegDesc{pairs, outs, (int)(ro + r.idx), 1, (int)co, n, r.mask, 0};
            else d = SegDesc{pairs, outs, (int)ro, m, (int)(co + r.idx), 1, r.mask, 0};
            segs[q] = d;
            pairs += (long long)d.m * d.n;
            outs += (long long)popc(r.mask) * d.m * d.n;
        }
        if ((long long)out.size() < outs) out.resize((size_t)outs);
        const auto t1 = clk::now();
        int rc = run_segments(g, d_idx, d_idx + nr, segs, pairs, outs, out.data(), &cnt.gpu_ms);
        if (rc) return rc;
        const auto t2 = clk::now();
        cnt.pairs += (double)pairs;
        Results R{&reqs, &segs, out.data()};
        parallel_for((long long)inst.size(), [&](long long x) {
            Inst& s = inst[(size_t)x];
            if (s.state != ST_DONE && s.nq) consume(s, R, eps);
        });
        const auto t3 = clk::now();
        cnt.phase_s[0] += secs(t0, t1);
        cnt.phase_s[1] += secs(t1, t2);
        cnt.phase_s[2] += secs(t2, t3);
    }
}

// ---------------------------------------------------------------------------
// ACA with the pivoting on the device.  The same algorithm as consume /
// decide / add_term above, one wave per (block, matrix).  Each sweep:
// aca_dev_logic_kernel consumes the instance's last evaluation and issues its
// next requests; aca_dev_scan_kernel numbers the requested pairs; and
// aca_dev_eval_kernel evaluates them, one pair per lane, straight into the
// requesting instance's buffers.  The host reads one count per sweep.  The
// factors of an instance hold at most kcap terms; an instance that would
// need more is marked and redone by the host loop.
// ---------------------------------------------------------------------------
struct DAca {
    int b, c, m, n;
    int kmax, kcap, k, state;
    int iref, jref, istar, jstar;
    int tries, nq, need_rr, need_rc;
    int lead_row, pending_conv, converged, capped;
    int first, pad0;
    int qt[2], qi[2];
    double norm2, scale;
    unsigned long long rng, scale_bits;
    long long offU, offV, offM, offN, offFm, offFn;
};

__device__ __forceinline__ unsigned long long dsplitmix(unsigned long long& s) {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// lane 0's value in every lane (a scalar register)
__device__ __forceinline__ double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum over the wave, the same bits in every lane.  The butterfly alone is not
// enough: the compiler may fuse a caller's product into its first add
// (fma(u_i, u_i, u_j^2) in lane i, fma(u_j, u_j, u_i^2) in lane j), and a
// branch on such a sum would split the wave; so lane 0's sum is broadcast.
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return wave_uniform(v);
}

__device__ __forceinline__ int wmin_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return __builtin_amdgcn_readfirstlane(v);  // same in every lane; now known to be
}

// argmax over unused entries with the host loop's semantics: the first
// maximum of |x| in index order; a NaN after the first valid entry wins.
__device__ __forceinline__ int dargmax_unused(const double* x, const unsigned char* used, int len, double* best) {
    const int lane = threadIdx.x & 63;
    double bv = -1.0;
    int bi = INT_MAX, fv = INT_MAX;
    bool has_nan = false;
    for (int t = lane; t < len; t += 64) {
        if (used[t]) continue;
        const double a = fabs(x[t]);
        if (a != a) {
            has_nan = true;
            continue;
        }
        if (t < fv) fv = t;
        if (a > bv) {
            bv = a;
            bi = t;
        }
    }
    // lane-wise first maxima -> wave maximum, smallest index on ties
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    fv = wmin_int(fv);
    if (__any(has_nan) && fv != INT_MAX) {
        int ni = INT_MAX;
        for (int t = lane; t < len; t += 64)
            if (!used[t] && t > fv && x[t] != x[t] && t < ni) ni = t;
        ni = wmin_int(ni);
        if (ni != INT_MAX) {
            *best = __longlong_as_double(0x7ff8000000000000ll);
            return ni;
        }
    }
    bi = __builtin_amdgcn_readfirstlane(bi);  // the reductions leave every lane equal
    bv = wave_uniform(bv);
    if (bi == INT_MAX) {
        *best = 0.0;
        return -1;
    }
    *best = bv;
    return bi;
}

__device__ __forceinline__ int dpick_unused(DAca& s, const unsigned char* used, int len) {
    const int lane = threadIdx.x & 63;
    const int start = (int)(dsplitmix(s.rng) % (unsigned long long)len);
    for (int base = 0; base < len; base += 64) {
        const int t = base + lane;
        const bool ok = t < len && !used[(start + t) % len];
        const unsigned long long bal = __ballot(ok);
        if (bal) return __builtin_amdgcn_readfirstlane((start + base + __ffsll((long long)bal) - 1) % len);
    }
    return -1;
}

struct DView {  // the instance's buffers
    double *U, *V, *Rc, *col, *rawC, *Rr, *r, *rawR;
    unsigned char *ur, *uc;
};

__device__ __forceinline__ DView dview(const DAca& s, double* dbuf, unsigned char* fbuf) {
    DView v;
    v.U = dbuf + s.offU;
    v.V = dbuf + s.offV;
    v.Rc = dbuf + s.offM;
    v.col = v.Rc + s.m;
    v.rawC = v.Rc + 2 * (long long)s.m;
    v.Rr = dbuf + s.offN;
    v.r = v.Rr + s.n;
    v.rawR = v.Rr + 2 * (long long)s.n;
    v.ur = fbuf + s.offFm;
    v.uc = fbuf + s.offFn;
    return v;
}

__device__ __forceinline__ void dset_flag(unsigned char* f, int i) {
    if ((threadIdx.x & 63) == 0) f[i] = 1;
    __threadfence_block();
}

// lane l's value in every lane (l uniform)
__device__ __forceinline__ double rd_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// out[t] = raw[t] - sum over l < k with c_l != 0 of c_l B[l ldb + t], t < len,
// the terms in order of l; c_l = C[l ldc] (k <= 64, one per lane).  Four
// entries per lane per pass and four terms at a time keep sixteen loads in
// flight (one term at a time, a 15k-row residual waited out one memory
// latency per term and pass).
__device__ __forceinline__ void dresidual(int k, const double* C, long long ldc, const double* B, long long ldb,
                          const double* raw, int len, double* out) {
    const int lane = threadIdx.x & 63;
    const double coef = lane < k ? C[(long long)lane * ldc] : 0.0;
    for (int base = 0; base < len; base += 256) {
        double v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int t = base + lane + 64 * c;
            v[c] = t < len ? raw[t] : 0.0;
        }
        // four terms' loads in flight (16 per lane), then their updates in
        // term order; a zero coefficient leaves v as it is (as skipping it would)
        int l = 0;
        for (; l + 4 <= k; l += 4) {
            double a4[4], bb[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a4[u] = rd_lane(coef, l + u);
                const double* b = B + (long long)(l + u) * ldb;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int t = base + lane + 64 * c;
                    bb[u][c] = t < len ? b[t] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int t = base + lane + 64 * c;
                    if (t < len && a4[u] != 0.0) v[c] -= a4[u] * bb[u][c];
                }
        }
        for (; l < k; ++l) {
            const double a = rd_lane(coef, l);
            if (a == 0.0) continue;
            const double* b = B + (long long)l * ldb;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int t = base + lane + 64 * c;
                if (t < len) v[c] -= a * b[t];
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int t = base + lane + 64 * c;
            if (t < len) out[t] = v[c];
        }
    }
    __threadfence_block();
}

// residual of row i: raw - sum_l U[l m + i] V[l, :]
__device__ __forceinline__ void dresidual_row(const DAca& s, const DView& w, const double* raw, int i, double* out) {
    dresidual(s.k, w.U + i, s.m, w.V, s.n, raw, s.n, out);
}

// residual of column j: raw - sum_l V[l n + j] U[l, :]
__device__ __forceinline__ void dresidual_col(const DAca& s, const DView& w, const double* raw, int j, double* out) {
    dresidual(s.k, w.V + j, s.n, w.U, s.m, raw, s.m, out);
}

__device__ __forceinline__ void dpush(DAca& s, int type, int idx) {
    if (s.nq == 0) {  // constant indices keep the instance in registers
        s.qt[0] = type;
        s.qi[0] = idx;
    } else {
        s.qt[1] = type;
        s.qi[1] = idx;
    }
    ++s.nq;
}

__device__ __forceinline__ void dstart_refs(DAca& s, const DView& w) {
    if (s.need_rr) {
        s.iref = dpick_unused(s, w.ur, s.m);
        if (s.iref < 0) {
            s.need_rr = 0;
            for (int j = threadIdx.x & 63; j < s.n; j += 64) w.Rr[j] = 0.0;
            __threadfence_block();
        } else {
            dpush(s, 0, s.iref);
        }
    }
    if (s.need_rc) {
        s.jref = dpick_unused(s, w.uc, s.n);
        if (s.jref < 0) {
            s.need_rc = 0;
            for (int i = threadIdx.x & 63; i < s.m; i += 64) w.Rc[i] = 0.0;
            __threadfence_block();
        } else {
            dpush(s, 1, s.jref);
        }
    }
    s.state = s.nq ? ST_REFS : ST_LEAD;
}

__device__ __forceinline__ void ddecide(DAca& s, const DView& w, double eps) {
    for (int guard = 0; guard < 16; ++guard) {
        const double tiny = 1e-12 * s.scale;
        const double rms = sqrt(s.norm2 / ((double)s.m * (double)s.n));
        double ar = 0, ac = 0;
        const int jc = dargmax_unused(w.Rr, w.uc, s.n, &ar);
        const int ic = dargmax_unused(w.Rc, w.ur, s.m, &ac);
        if (s.pending_conv) {
            s.pending_conv = 0;
            if (ar <= eps * rms + tiny && ac <= eps * rms + tiny) {
                s.converged = 1;
                s.state = ST_DONE;
                return;
            }
        }
        if (s.k >= s.kmax || (jc < 0 && ic < 0)) {
            s.converged = (jc < 0 && ic < 0) || s.k >= min(s.m, s.n);
            s.state = ST_DONE;
            return;
        }
        if (s.k >= s.kcap) {  // the host loop takes over
            s.capped = 1;
            s.state = ST_DONE;
            return;
        }
        if (!(ar > tiny) && !(ac > tiny)) {
            if (++s.tries > 3) {
                s.converged = 1;
                s.state = ST_DONE;
                return;
            }
            if (s.iref >= 0) dset_flag(w.ur, s.iref);
            if (s.jref >= 0) dset_flag(w.uc, s.jref);
            s.need_rr = s.need_rc = 1;
            dstart_refs(s, w);
            if (s.state == ST_LEAD) continue;
            return;
        }
        if (ac > ar) {
            s.lead_row = 1;
            s.istar = ic;
            dpush(s, 0, ic);
        } else {
            s.lead_row = 0;
            s.jstar = jc;
            dpush(s, 1, jc);
        }
        s.state = ST_LEAD;
        return;
    }
    s.state = ST_DONE;
}

__device__ __forceinline__ void dadd_term(DAca& s, const DView& w, double eps) {
    const int lane = threadIdx.x & 63;
    const double delta = w.r[s.jstar];
    const double tiny = 1e-12 * s.scale;
    dset_flag(w.ur, s.istar);
    dset_flag(w.uc, s.jstar);
    if (!(fabs(delta) > tiny)) {
        ddecide(s, w, eps);
        return;
    }
    const int m = s.m, n = s.n, k = s.k;
    double* u = w.U + (long long)k * m;
    double* v = w.V + (long long)k * n;
    const double inv = 1.0 / delta;
    double uu = 0, vv = 0;
    for (int i = lane; i < m; i += 64) {
        const double x = w.col[i];
        u[i] = x;
        uu += x * x;
    }
    for (int j = lane; j < n; j += 64) {
        const double x = w.r[j] * inv;
        v[j] = x;
        vv += x * x;
    }
    __threadfence_block();
    uu = wsum(uu);
    vv = wsum(vv);
    // cross = sum_l (U_l . u)(V_l . v), four terms' dot products per pass
    double cross = 0;
    int l = 0;
    for (; l + 4 <= k; l += 4) {
        double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
        for (int i = lane; i < m; i += 64) {
            const double x = u[i];
#pragma unroll
            for (int c = 0; c < 4; ++c) a[c] += w.U[(long long)(l + c) * m + i] * x;
        }
        for (int j = lane; j < n; j += 64) {
            const double x = v[j];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] += w.V[(long long)(l + c) * n + j] * x;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) cross += wsum(a[c]) * wsum(b[c]);
    }
    for (; l < k; ++l) {
        const double* ul = w.U + (long long)l * m;
        const double* vl = w.V + (long long)l * n;
        double a = 0, b = 0;
        for (int i = lane; i < m; i += 64) a += ul[i] * u[i];
        for (int j = lane; j < n; j += 64) b += vl[j] * v[j];
        cross += wsum(a) * wsum(b);
    }
    s.norm2 = fmax(0.0, s.norm2 + 2.0 * cross + uu * vv);
    s.k = k + 1;
    if (s.iref >= 0) {
        const double a = u[s.iref];
        for (int j = lane; j < n; j += 64) w.Rr[j] -= a * v[j];
        if (w.ur[s.iref]) s.need_rr = 1;
    }
    if (s.jref >= 0) {
        const double a = v[s.jref];
        for (int i = lane; i < m; i += 64) w.Rc[i] -= a * u[i];
        if (w.uc[s.jref]) s.need_rc = 1;
    }
    __threadfence_block();
    const double step = sqrt(uu * vv);
    if (eps < 0) {
        // the reference's absolute rule, as in add_term
        if (fabs(delta) < ACA_REF_PIVOT || step < -eps) {
            s.converged = 1;
            s.state = ST_DONE;
            return;
        }
        s.pending_conv = 0;
    } else {
        const double floor_ = 1e-9 * s.scale * sqrt((double)m * (double)n);
        s.pending_conv = step <= eps * sqrt(s.norm2) || step <= floor_;
    }
    s.tries = 0;
    if (s.need_rr || s.need_rc) {
        dstart_refs(s, w);
        if (s.state == ST_LEAD) ddecide(s, w, eps);
    } else {
        ddecide(s, w, eps);
    }
}

// one wave per instance: consume the last evaluation, issue the next requests
__global__ __launch_bounds__(256) void aca_dev_logic_kernel(DAca* __restrict__ inst, long long ni,
                                                            double* __restrict__ dbuf,
                                                            unsigned char* __restrict__ fbuf, double eps,
                                                            long long* __restrict__ cnt) {
    // the wave's instance, held in scalar registers
    const long long x = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (x >= ni) return;  // whole wave
    DAca s = inst[x];
    if (s.state == ST_DONE || s.nq == 0) {
        if ((threadIdx.x & 63) == 0) cnt[x] = 0;
        return;
    }
    const DView w = dview(s, dbuf, fbuf);
    const int nq = s.nq;
    const int qt0 = s.qt[0], qi0 = s.qi[0], qt1 = s.qt[1], qi1 = s.qi[1];
    s.nq = 0;
    if (s.first) {
        s.scale = __longlong_as_double((long long)s.scale_bits);
        s.first = 0;
    }
    double best = 0;
    switch (s.state) {
        case ST_REFS:
            for (int t = 0; t < nq; ++t) {
                const int qt = t ? qt1 : qt0, qi = t ? qi1 : qi0;
                if (qt == 0) {
                    dresidual_row(s, w, w.rawR, qi, w.Rr);
                    s.need_rr = 0;
                } else {
                    dresidual_col(s, w, w.rawC, qi, w.Rc);
                    s.need_rc = 0;
                }
            }
            ddecide(s, w, eps);
            break;
        case ST_LEAD:
            if (s.lead_row) {
                dresidual_row(s, w, w.rawR, s.istar, w.r);
                s.jstar = dargmax_unused(w.r, w.uc, s.n, &best);
                if (s.jstar < 0 || !(best > 1e-12 * s.scale)) {
                    dset_flag(w.ur, s.istar);
                    ddecide(s, w, eps);
                    break;
                }
                dpush(s, 1, s.jstar);
            } else {
                dresidual_col(s, w, w.rawC, s.jstar, w.col);
                s.istar = dargmax_unused(w.col, w.ur, s.m, &best);
                if (s.istar < 0 || !(best > 1e-12 * s.scale)) {
                    dset_flag(w.uc, s.jstar);
                    ddecide(s, w, eps);
                    break;
                }
                dpush(s, 0, s.istar);
            }
            s.state = ST_FOLLOW;
            break;
        case ST_FOLLOW:
            if (s.lead_row) dresidual_col(s, w, w.rawC, s.jstar, w.col);
            else dresidual_row(s, w, w.rawR, s.istar, w.r);
            dadd_term(s, w, eps);
            break;
        default:
            break;
    }
    if (s.state == ST_DONE) s.nq = 0;
    long long c = 0;
    if (s.nq > 0) c += s.qt[0] == 0 ? s.n : s.m;
    if (s.nq > 1) c += s.qt[1] == 0 ? s.n : s.m;
    if ((threadIdx.x & 63) == 0) {
        inst[x] = s;
        cnt[x] = c;
    }
}

// exclusive prefix sum of cnt[0, n) into off[0, n], off[n] = total; one workgroup
__global__ __launch_bounds__(1024) void aca_dev_scan_kernel(const long long* __restrict__ cnt, long long n,
                                                            long long* __restrict__ off) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const long long per = (n + 1023) / 1024;
    const long long lo = t * per, hi = lo + per < n ? lo + per : n;
    long long sum = 0;
    for (long long i = lo; i < hi; ++i) sum += cnt[i];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan of the partial sums
        const long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = t ? part[t - 1] : 0;
    for (long long i = lo; i < hi; ++i) {
        off[i] = run;
        run += cnt[i];
    }
    if (t == 1023) off[n] = part[1023];
}

// one pair per lane: the requested values of matrix c go to the instance's
// raw row/column buffer.  NB = 2 (an instance's first evaluation): both slip
// directions, and the largest finite |entry| over all six matrices (the
// block's scale); NB = 1: only the instance's own slip direction, from the
// strike-only or dip-only table geo1.  The pair's instance is found with two
// searches per workgroup and one over the few instances the workgroup spans.
template <int NB>
__global__ __launch_bounds__(256) void aca_dev_eval_kernel(const SrcGeom<2>* __restrict__ geo,
                                                           const SrcGeom<1>* __restrict__ geo1, long long nsrc,
                                                           const double* __restrict__ xg,
                                                           const double* __restrict__ ev,
                                                           const int* __restrict__ ridx, const int* __restrict__ sidx,
                                                           const long long* __restrict__ row_off,
                                                           const long long* __restrict__ col_off,
                                                           DAca* __restrict__ inst, const long long* __restrict__ off,
                                                           long long ni, double mu, double lam, double nu,
                                                           int halfspace, double* __restrict__ dbuf) {
    __shared__ long long s_lo, s_cnt;
    __shared__ long long s_off[257];
    const long long total = off[ni];
    const long long q0 = (long long)blockIdx.x * 256, q = q0 + threadIdx.x;  // q0 < total (grid fits)
    auto last_le = [&](long long qq, long long lo, long long hi) {  // last x in [lo, hi] with off[x] <= qq
        while (lo < hi) {
            const long long mid = (lo + hi + 1) >> 1;
            if (off[mid] <= qq) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    if (threadIdx.x == 0) {
        const long long a = last_le(q0, 0, ni - 1);
        const long long b = last_le(q0 + 255 < total ? q0 + 255 : total - 1, a, ni - 1);
        s_lo = a;
        s_cnt = b - a + 1;
    }
    __syncthreads();
    const long long lo0 = s_lo, cnt = s_cnt;
    if (cnt <= 256) {
        if (threadIdx.x < cnt) s_off[threadIdx.x] = off[lo0 + threadIdx.x];
        __syncthreads();
    }
    if (q >= total) return;
    long long x;
    if (cnt <= 256) {
        int lo = 0, hi = (int)cnt - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        x = lo0 + lo;
    } else {
        x = last_le(q, lo0, lo0 + cnt - 1);
    }
    const DAca& s = inst[x];
    long long p = q - off[x];
    const long long len0 = s.qt[0] == 0 ? s.n : s.m;
    const int t = p < len0 ? 0 : 1;
    if (t) p -= len0;
    const int type = s.qt[t], idx = s.qi[t];
    const int i = type == 0 ? idx : (int)p, j = type == 0 ? (int)p : idx;
    const int r = ridx[row_off[s.b] + i], src = sidx[col_off[s.b] + j];
    const double X = xg[3 * (long long)r + 0], Y = xg[3 * (long long)r + 1], Z = xg[3 * (long long)r + 2];
    double e[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = ev[9 * (long long)r + k];
    double v;
    if constexpr (NB == 2) {
        double G[2][9];
        td_pair_grad<double, 2, false>(geo[src], X, Y, Z, nu, halfspace != 0, G);
        double vals[6];
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            double eps[6], sig[6], tr[3];
            grad_to_strain_stress<double>(G[bb], mu, lam, eps, sig);
            stress_to_traction<double>(sig, e, tr);
#pragma unroll
            for (int k = 0; k < 3; ++k) vals[bb * 3 + k] = tr[k];
        }
        v = vals[0];
#pragma unroll
        for (int k = 1; k < 6; ++k)
            if (k == s.c) v = vals[k];
        if (s.first) {
            double mx = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double a = fabs(vals[k]);
                if (a == a && a > mx) mx = a;
            }
            atomicMax(&inst[x].scale_bits, (unsigned long long)__double_as_longlong(mx));
        }
    } else {
        double G[1][9];
        td_pair_grad<double, 1, false>(geo1[(s.c >= 3 ? nsrc : 0) + src], X, Y, Z, nu, halfspace != 0, G);
        double eps[6], sig[6], tr[3];
        grad_to_strain_stress<double>(G[0], mu, lam, eps, sig);
        stress_to_traction<double>(sig, e, tr);
        const int cc = s.c >= 3 ? s.c - 3 : s.c;
        v = cc == 0 ? tr[0] : cc == 1 ? tr[1] : tr[2];
    }
    if (type == 0) dbuf[s.offN + 2 * (long long)s.n + p] = v;   // rawR
    else dbuf[s.offM + 2 * (long long)s.m + p] = v;             // rawC
}

// ---- recompression on the device: the algorithm of recompress_factors
// (MGS2 QR of U and V^T, core Ru Rv^T, one-sided cyclic Jacobi SVD, Frobenius
// truncation), one wave per (block, matrix); sums run across the wave.

// MGS2 QR (the algorithm of qr_mgs2): Q (rows x k, column j at j*rows) ->
// orthonormal columns in place, R (k x k, R[i*k + j]) upper triangular.
// One wave, the column being orthogonalized held in registers (NV values per
// lane, rows <= 64 NV) and the next basis column loaded while the
// current one is applied: one pass over Q's columns per step instead of four.
template <int NV>
__device__ void dqr_mgs2_regs(int rows, int k, double* Q, double* R) {
    const int lane = threadIdx.x & 63;
    for (int e = lane; e < k * k; e += 64) R[e] = 0.0;
    __threadfence_block();
    for (int j = 0; j < k; ++j) {
        double* col = Q + (long long)j * rows;
        double v[NV];
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            const int t = lane + 64 * s;
            v[s] = t < rows ? col[t] : 0.0;
        }
        for (int pass = 0; pass < 2 && j > 0; ++pass) {
            double q[NV], qn[NV];
#pragma unroll
            for (int s = 0; s < NV; ++s) {
                const int t = lane + 64 * s;
                q[s] = t < rows ? Q[t] : 0.0;
                qn[s] = 0.0;
            }
            for (int i = 0; i < j; ++i) {
                if (i + 1 < j) {
                    const double* nx = Q + (long long)(i + 1) * rows;
#pragma unroll
                    for (int s = 0; s < NV; ++s) {
                        const int t = lane + 64 * s;
                        qn[s] = t < rows ? nx[t] : 0.0;
                    }
                }
                double d = 0;
#pragma unroll
                for (int s = 0; s < NV; ++s) d += q[s] * v[s];
                d = wsum(d);
                if (lane == 0) R[i * k + j] += d;
#pragma unroll
                for (int s = 0; s < NV; ++s) {
                    v[s] -= d * q[s];
                    q[s] = qn[s];
                }
            }
        }
        double nv = 0;
#pragma unroll
        for (int s = 0; s < NV; ++s) nv += v[s] * v[s];
        nv = sqrt(wsum(nv));
        if (lane == 0) R[j * k + j] = nv;
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            const int t = lane + 64 * s;
            if (nv > 0) v[s] /= nv;
            if (t < rows) col[t] = v[s];
        }
    }
    __threadfence_block();
}

// The same MGS2 on an NT-thread workgroup for long columns (rows <= NT NV):
// column in registers across the block, sums over the block through LDS in a
// fixed order (the same bits in every thread).  PF: load the next basis
// column while the current one is applied (costs NV registers).
template <int NV, int NT, bool PF>
__device__ void dqr_mgs2_block(int rows, int k, double* Q, double* R) {
    constexpr int NW = NT / 64;
    __shared__ double red[2][NW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int flip = 0;
    auto bsum = [&](double x) {
        x = wsum(x);
        if (lane == 0) red[flip][wv] = x;
        __syncthreads();
        double t = red[flip][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) t += red[flip][w];
        flip ^= 1;  // the other slot next time: no second barrier needed
        return t;
    };
    for (int e = tid; e < k * k; e += NT) R[e] = 0.0;
    __threadfence_block();
    __syncthreads();
    for (int j = 0; j < k; ++j) {
        double* col = Q + (long long)j * rows;
        double v[NV];
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            const int t = tid + NT * s;
            v[s] = t < rows ? col[t] : 0.0;
        }
        for (int pass = 0; pass < 2 && j > 0; ++pass) {
            double q[NV], qn[NV];
#pragma unroll
            for (int s = 0; s < NV; ++s) {
                const int t = tid + NT * s;
                q[s] = t < rows ? Q[t] : 0.0;
                qn[s] = 0.0;
            }
            for (int i = 0; i < j; ++i) {
                if (PF && i + 1 < j) {
                    const double* nx = Q + (long long)(i + 1) * rows;
#pragma unroll
                    for (int s = 0; s < NV; ++s) {
                        const int t = tid + NT * s;
                        qn[s] = t < rows ? nx[t] : 0.0;
                    }
                }
                double d = 0;
#pragma unroll
                for (int s = 0; s < NV; ++s) d += q[s] * v[s];
                d = bsum(d);
                if (tid == 0) R[i * k + j] += d;
#pragma unroll
                for (int s = 0; s < NV; ++s) v[s] -= d * q[s];
                if (PF) {
#pragma unroll
                    for (int s = 0; s < NV; ++s) q[s] = qn[s];
                } else if (i + 1 < j) {
                    const double* nx = Q + (long long)(i + 1) * rows;
#pragma unroll
                    for (int s = 0; s < NV; ++s) {
                        const int t = tid + NT * s;
                        q[s] = t < rows ? nx[t] : 0.0;
                    }
                }
            }
        }
        double nv = 0;
#pragma unroll
        for (int s = 0; s < NV; ++s) nv += v[s] * v[s];
        nv = sqrt(bsum(nv));
        if (tid == 0) R[j * k + j] = nv;
#pragma unroll
        for (int s = 0; s < NV; ++s) {
            const int t = tid + NT * s;
            if (nv > 0) v[s] /= nv;
            if (t < rows) col[t] = v[s];
        }
        __threadfence_block();
        __syncthreads();  // column j complete before other threads read it
    }
}

// The same MGS2 on an NT-thread workgroup with the columns in memory (any
// rows): each step reads the basis and current columns twice, spread over
// NT / 64 waves.
template <int NT>
__device__ void dqr_mgs2_mem_block(int rows, int k, double* Q, double* R) {
    constexpr int NW = NT / 64;
    __shared__ double red[2][NW];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int flip = 0;
    auto bsum = [&](double x) {
        x = wsum(x);
        if (lane == 0) red[flip][wv] = x;
        __syncthreads();
        double t = red[flip][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) t += red[flip][w];
        flip ^= 1;
        return t;
    };
    for (int e = tid; e < k * k; e += NT) R[e] = 0.0;
    __threadfence_block();
    __syncthreads();
    for (int j = 0; j < k; ++j) {
        double* v = Q + (long long)j * rows;
        for (int pass = 0; pass < 2; ++pass)
            for (int i = 0; i < j; ++i) {
                const double* q = Q + (long long)i * rows;
                double d = 0;
                for (int t = tid; t < rows; t += NT) d += q[t] * v[t];
                d = bsum(d);
                if (tid == 0) R[i * k + j] += d;
                for (int t = tid; t < rows; t += NT) v[t] -= d * q[t];
            }
        double nv = 0;
        for (int t = tid; t < rows; t += NT) nv += v[t] * v[t];
        nv = sqrt(bsum(nv));
        if (tid == 0) R[j * k + j] = nv;
        if (nv > 0)
            for (int t = tid; t < rows; t += NT) v[t] /= nv;
        __threadfence_block();
        __syncthreads();
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void aca_dev_qr_mem_block_kernel(const DAca* __restrict__ inst,
                                                                  const int* __restrict__ jobs,
                                                                  double* __restrict__ dbuf, double* __restrict__ ws,
                                                                  const long long* __restrict__ wsoff) {
    const int job = jobs[blockIdx.x], x = job >> 1, which = job & 1;
    const DAca& s = inst[x];
    const int k = s.k, rows = which ? s.n : s.m;
    dqr_mgs2_mem_block<NT>(rows, k, dbuf + (which ? s.offV : s.offU), ws + wsoff[x] + (which ? (long long)k * k : 0));
}

// QR jobs on workgroups: one job (as below) per NT-thread block
template <int NV, int NT, bool PF>
__global__ __launch_bounds__(NT) void aca_dev_qr_block_kernel(const DAca* __restrict__ inst,
                                                              const int* __restrict__ jobs,
                                                              double* __restrict__ dbuf, double* __restrict__ ws,
                                                              const long long* __restrict__ wsoff) {
    const int job = jobs[blockIdx.x], x = job >> 1, which = job & 1;
    const DAca& s = inst[x];
    const int k = s.k, rows = which ? s.n : s.m;
    dqr_mgs2_block<NV, NT, PF>(rows, k, dbuf + (which ? s.offV : s.offU),
                           ws + wsoff[x] + (which ? (long long)k * k : 0));
}

// QR jobs: job = 2 x + w, w = 0 for U (m rows, R at wsoff[x]), 1 for V^T (n
// rows, R at wsoff[x] + k^2); one wave per job
template <int NV>
__global__ __launch_bounds__(256) void aca_dev_qr_kernel(const DAca* __restrict__ inst,
                                                         const int* __restrict__ jobs, int njobs,
                                                         double* __restrict__ dbuf, double* __restrict__ ws,
                                                         const long long* __restrict__ wsoff) {
    const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (w >= njobs) return;  // whole wave
    const int job = jobs[w], x = job >> 1, which = job & 1;
    const DAca& s = inst[x];
    const int k = s.k, rows = which ? s.n : s.m;
    double* Q = dbuf + (which ? s.offV : s.offU);
    double* R = ws + wsoff[x] + (which ? (long long)k * k : 0);
    dqr_mgs2_regs<NV>(rows, k, Q, R);
}

// U (offU, m x k) and V (offV, k x n) of instance x have become Qu and Qv^T in
// place (aca_dev_qr_kernel) with Ru, Rv at ws + wsoff[x]; ws + wsoff[x] (4 k^2) receives the scaled singular vectors
// Wh = W_r sqrt(sig) (column l at l*k) and Zh = Z_r sqrt(sig) (at k^2);
// rank[x] = r.  k <= 64 (one column per lane in the SVD).
__global__ __launch_bounds__(256) void aca_dev_recompress_kernel(const DAca* __restrict__ inst, long long ni,
                                                                 double* __restrict__ ws,
                                                                 const long long* __restrict__ wsoff, double eps,
                                                                 int* __restrict__ rank) {
    const long long x = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (x >= ni) return;  // whole wave
    const int lane = threadIdx.x & 63;
    const DAca& s = inst[x];
    const int k = s.capped ? 0 : s.k;
    if (k == 0) {
        if (lane == 0) rank[x] = 0;
        return;
    }
    double* Ru = ws + wsoff[x];
    double* Rv = Ru + k * k;
    double* A = Rv + k * k;   // columns of C = Ru Rv^T, rotated in place
    double* Vj = A + k * k;   // accumulated rotations
    for (int e = lane; e < k * k; e += 64) {
        const int i = e / k, j = e % k;
        double a = 0;
        for (int t = 0; t < k; ++t) a += Ru[i * k + t] * Rv[j * k + t];
        A[j * k + i] = a;
        Vj[e] = (i == j) ? 1.0 : 0.0;
    }
    __threadfence_block();
    // one-sided Jacobi, cyclic order; lane t owns row t of A and Vj
    const bool own = lane < k;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                const double u = own ? A[p * k + lane] : 0.0, w = own ? A[q * k + lane] : 0.0;
                const double al = wsum(u * u), be = wsum(w * w), ga = wsum(u * w);
                if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                if (own) {
                    A[p * k + lane] = c * u - sn * w;
                    A[q * k + lane] = sn * u + c * w;
                    const double vu = Vj[p * k + lane], vw = Vj[q * k + lane];
                    Vj[p * k + lane] = c * vu - sn * vw;
                    Vj[q * k + lane] = sn * vu + c * vw;
                }
            }
        if (!rotated) break;
    }
    __threadfence_block();
    // column norms; lane j holds column j's, its stable descending position
    double nrm = 0;
    if (own) {
        double a = 0;
        for (int t = 0; t < k; ++t) a += A[lane * k + t] * A[lane * k + t];
        nrm = sqrt(a);
    }
    int pos = 0;
    for (int i = 0; i < k; ++i) {
        const double o = __shfl(nrm, i, 64);
        pos += (o > nrm) || (o == nrm && i < lane);
    }
    // truncation in descending order, as the host loop sums
    int r = 0;
    {
        double tot = 0;
        for (int l = 0; l < k; ++l) {
            const int src = __ffsll((long long)__ballot(own && pos == l)) - 1;
            const double sg = __shfl(nrm, src, 64);
            tot += sg * sg;
        }
        if (tot > 0) {
            double tail = tot;
            while (r < k && tail > eps * eps * tot) {
                const int src = __ffsll((long long)__ballot(own && pos == r)) - 1;
                const double sg = __shfl(nrm, src, 64);
                tail -= sg * sg;
                ++r;
            }
        }
    }
    // Wh, Zh over Ru, Rv (no longer needed)
    double* Wh = Ru;
    double* Zh = Rv;
    for (int l = 0; l < r; ++l) {
        const int src = __ffsll((long long)__ballot(own && pos == l)) - 1;
        if (src < 0) {  // positions are a permutation unless a norm is NaN
            r = l;
            break;
        }
        const double sg = __shfl(nrm, src, 64);
        const double h = sqrt(sg);
        if (own) {
            Wh[l * k + lane] = (sg > 0 ? A[src * k + lane] / sg : 0.0) * h;
            Zh[l * k + lane] = Vj[src * k + lane] * h;
        }
    }
    if (lane == 0) rank[x] = r;
}

// out[l len + t] = sum over i < k with c[l k + i] != 0 of c[l k + i] Q[i len + t]
// (the terms in order of i), l < r, t < len.  Eight output terms per pass: each
// Q entry is read once per eight terms (once per term, the 60k build re-read
// about 130 GB), four rows of Q in flight at a time.
__device__ __forceinline__ void form_terms(const double* __restrict__ c, const double* __restrict__ Q, int k,
                                           int len, int r, double* __restrict__ out) {
    constexpr int G = 8;
    for (int l0 = 0; l0 < r; l0 += G) {
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            double u[G];
#pragma unroll
            for (int j = 0; j < G; ++j) u[j] = 0.0;
            int i = 0;
            for (; i + 4 <= k; i += 4) {
                double q[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) q[a] = Q[(long long)(i + a) * len + t];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int j = 0; j < G; ++j)
                        if (l0 + j < r) {
                            const double w = c[(l0 + j) * k + i + a];
                            if (w != 0.0) u[j] += w * q[a];
                        }
            }
            for (; i < k; ++i) {
                const double q = Q[(long long)i * len + t];
#pragma unroll
                for (int j = 0; j < G; ++j)
                    if (l0 + j < r) {
                        const double w = c[(l0 + j) * k + i];
                        if (w != 0.0) u[j] += w * q;
                    }
            }
#pragma unroll
            for (int j = 0; j < G; ++j)
                if (l0 + j < r) out[(long long)(l0 + j) * len + t] = u[j];
        }
    }
}

// final factors in the fetch layout: U2 (m x r, column-major) at dU + uoff[x],
// V2 (r x n, row-major) at dV + voff[x]; U2 = Qu Wh, V2 = Zh^T Qv^T, or with
// ident the first k terms as they are (no recompression)
__global__ __launch_bounds__(256) void aca_dev_form_kernel(const DAca* __restrict__ inst,
                                                           const double* __restrict__ dbuf,
                                                           const double* __restrict__ ws,
                                                           const long long* __restrict__ wsoff,
                                                           const int* __restrict__ rank,
                                                           const long long* __restrict__ uoff,
                                                           const long long* __restrict__ voff, int ident,
                                                           double* __restrict__ dU, double* __restrict__ dV) {
    const long long x = blockIdx.x;
    const DAca& s = inst[x];
    if (s.capped) return;
    const int k = s.k, m = s.m, n = s.n;
    const int r = ident ? k : rank[x];
    if (r == 0) return;
    const double* Qu = dbuf + s.offU;
    const double* Qv = dbuf + s.offV;
    double* U2 = dU + uoff[x];
    double* V2 = dV + voff[x];
    if (ident) {
        for (long long t = threadIdx.x; t < (long long)k * m; t += blockDim.x) U2[t] = Qu[t];
        for (long long t = threadIdx.x; t < (long long)k * n; t += blockDim.x) V2[t] = Qv[t];
        return;
    }
    const double* Wh = ws + wsoff[x];
    const double* Zh = Wh + k * k;
    form_terms(Wh, Qu, k, m, r, U2);
    form_terms(Zh, Qv, k, n, r, V2);
}

}  // namespace pq3d

using namespace pq3d;

extern "C" {

// Device-resident geometry for batched Green's-function evaluation.
//   P1, P2, P3 : nsrc x 3 source TD vertices; xg: nrcv x 3; ev: nrcv x 9 frames [e1, e2, n]
int pq3d_gfsrc_create(void** out, const double* P1, const double* P2, const double* P3, long long nsrc,
                      const double* xg, const double* ev, long long nrcv, double mu, double lam, int halfspace) {
    clear_error();
    PQ_CHECK(nsrc > 0 && nrcv > 0, "pq3d_gfsrc_create: empty geometry");
    GFSource* g = new GFSource();
    g->nsrc = nsrc;
    g->nrcv = nrcv;
    g->mu = mu;
    g->lam = lam;
    g->nu = 1.0 / (1.0 + lam / mu) / 2.0;
    g->halfspace = halfspace;
    double* dP = nullptr;
    hipError_t e = hipMalloc((void**)&g->geo, (size_t)nsrc * sizeof(SrcGeom<2>));
    if (e == hipSuccess) e = hipMalloc((void**)&g->geo1, (size_t)nsrc * 2 * sizeof(SrcGeom<1>));
    if (e == hipSuccess) e = hipMalloc((void**)&g->xg, (size_t)nrcv * 12 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&dP, (size_t)nsrc * 9 * sizeof(double));
    if (e == hipSuccess) {
        g->ev = g->xg + 3 * nrcv;
        e = hipMemcpy(g->xg, xg, nrcv * 3 * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(g->ev, ev, nrcv * 9 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dP, P1, nsrc * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dP + 3 * nsrc, P2, nsrc * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dP + 6 * nsrc, P3, nsrc * 3 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gfsrc_setup_kernel, dim3((unsigned)((nsrc + 255) / 256)), dim3(256), 0, 0, dP, nsrc,
                           g->geo);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gfsrc_setup1_kernel, dim3((unsigned)((2 * nsrc + 255) / 256)), dim3(256), 0, 0, dP, nsrc,
                           g->geo1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (dP) (void)hipFree(dP);
    if (e != hipSuccess) {
        set_error("pq3d_gfsrc_create: %s", hipGetErrorString(e));
        (void)hipFree(g->geo);
        (void)hipFree(g->geo1);
        (void)hipFree(g->xg);
        delete g;
        return -1;
    }
    *out = g;
    return 0;
}

int pq3d_gfsrc_destroy(void* h) {
    clear_error();
    GFSource* g = (GFSource*)h;
    if (!g) return 0;
    (void)hipFree(g->geo);
    (void)hipFree(g->geo1);
    (void)hipFree(g->xg);
    (void)hipFree(g->seg);
    (void)hipFree(g->out);
    (void)hipFree(g->idx);
    delete g;
    return 0;
}

// Sub-blocks of the six traction matrices.  Segment k (seg[4k..4k+3] =
// rcv_off, m, src_off, n) covers receivers rcv_idx[rcv_off : rcv_off+m] x
// sources src_idx[src_off : src_off+n]; its output is 6 x n x m FP64 values
// ([matrix A1s..Bd][column][row]) in `out` (host), segments back to back.
int pq3d_gf_segments(void* h, const int* rcv_idx, long long nr_idx, const int* src_idx, long long ns_idx,
                     long long nseg, const long long* seg, double* out) {
    clear_error();
    GFSource* g = (GFSource*)h;
    PQ_CHECK(g, "pq3d_gf_segments: null source");
    for (long long i = 0; i < nr_idx; ++i)
        PQ_CHECK(rcv_idx[i] >= 0 && rcv_idx[i] < g->nrcv, "pq3d_gf_segments: receiver %d out of range", rcv_idx[i]);
    for (long long i = 0; i < ns_idx; ++i)
        PQ_CHECK(src_idx[i] >= 0 && src_idx[i] < g->nsrc, "pq3d_gf_segments: source %d out of range", src_idx[i]);
    std::vector<SegDesc> d;
    d.reserve((size_t)nseg);
    long long pairs = 0, outs = 0;
    for (long long k = 0; k < nseg; ++k) {
        const long long ro = seg[4 * k], m = seg[4 * k + 1], so = seg[4 * k + 2], n = seg[4 * k + 3];
        PQ_CHECK(m >= 0 && n >= 0 && ro >= 0 && so >= 0 && ro + m <= nr_idx && so + n <= ns_idx,
                 "pq3d_gf_segments: segment %lld out of its index lists", k);
        SegDesc s{pairs, outs, (int)ro, (int)m, (int)so, (int)n, 0x3f, 0};
        d.push_back(s);
        pairs += m * n;
        outs += 6 * m * n;
    }
    if (pairs == 0) return 0;
    // drop empty segments (the kernel's search assumes every listed one has pairs)
    std::vector<SegDesc> nz;
    for (auto& s : d)
        if ((long long)s.m * s.n > 0) nz.push_back(s);
    if (grow((void**)&g->idx, &g->idx_cap, (size_t)(nr_idx + ns_idx) * sizeof(int))) return -1;
    PQ_HIP(hipMemcpy(g->idx, rcv_idx, nr_idx * sizeof(int), hipMemcpyHostToDevice));
    PQ_HIP(hipMemcpy(g->idx + nr_idx, src_idx, ns_idx * sizeof(int), hipMemcpyHostToDevice));
    return run_segments(g, g->idx, g->idx + nr_idx, nz, pairs, outs, out, nullptr);
}

// ACA of all six matrices of nb blocks.  Block b: rows row_idx[row_off[b] :
// row_off[b+1]], columns col_idx[col_off[b] : col_off[b+1]] (element ids).
// eps: relative tolerance; recompress_eps > 0: recompress every factor pair
// to that tolerance afterwards; max_rank <= 0: min(m, n).  ranks (nb x 6) receive
// the ranks; *result holds the factors until pq3d_aca_free.
static int aca_check(GFSource* g, long long nb, const long long* row_off, const int* row_idx,
                     const long long* col_off, const int* col_idx, double eps) {
    PQ_CHECK(g, "pq3d_aca_run: null source");
    PQ_CHECK((eps > 0 && eps < 1) || eps < 0, "pq3d_aca_run: eps must be in (0, 1), or negative for the "
             "reference's absolute stopping rule");
    const long long nr = row_off[nb], nc = col_off[nb];
    for (long long b = 0; b < nb; ++b)
        PQ_CHECK(row_off[b + 1] > row_off[b] && col_off[b + 1] > col_off[b], "pq3d_aca_run: empty block %lld", b);
    for (long long i = 0; i < nr; ++i)
        PQ_CHECK(row_idx[i] >= 0 && row_idx[i] < g->nrcv, "pq3d_aca_run: row id out of range");
    for (long long i = 0; i < nc; ++i)
        PQ_CHECK(col_idx[i] >= 0 && col_idx[i] < g->nsrc, "pq3d_aca_run: column id out of range");
    return 0;
}

static int upload_ids(const long long* row_off, const int* row_idx, const long long* col_off, const int* col_idx,
                      long long nb, int** d_idx) {
    const long long nr = row_off[nb], nc = col_off[nb];
    *d_idx = nullptr;
    PQ_HIP(hipMalloc((void**)d_idx, (size_t)(nr + nc + 1) * sizeof(int)));
    hipError_t e = hipMemcpy(*d_idx, row_idx, nr * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(*d_idx + nr, col_idx, nc * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*d_idx);
        *d_idx = nullptr;
        set_error("pq3d_aca_run: %s", hipGetErrorString(e));
        return -1;
    }
    return 0;
}

// recompression, ranks and statistics shared by both entry points
static void aca_finish(AcaResult* res, double recompress_eps, long long* ranks, const AcaCounters& cnt,
                       std::chrono::steady_clock::time_point t_start) {
    using clk = std::chrono::steady_clock;
    const auto t_rc = clk::now();
    if (recompress_eps > 0)
        parallel_for((long long)res->inst.size(), [&](long long x) {
            Inst& s = res->inst[(size_t)x];
            s.k = recompress_factors(s.m, s.n, s.k, s.U, s.V, recompress_eps);
        });
    for (int k = 0; k < 3; ++k) res->phase_s[k] = cnt.phase_s[k];
    res->phase_s[3] = std::chrono::duration<double>(clk::now() - t_rc).count();
    int maxk = 0;
    long long unconv = 0;
    for (auto& s : res->inst) {
        ranks[(size_t)s.b * 6 + s.c] = s.k;
        maxk = std::max(maxk, s.k);
        if (!s.converged) ++unconv;
    }
    const double total_s = std::chrono::duration<double>(clk::now() - t_start).count();
    res->stats[0] = (double)cnt.sweeps;
    res->stats[1] = (double)cnt.sweeps;
    res->stats[2] = cnt.pairs;
    res->stats[3] = cnt.gpu_ms * 1e-3;
    res->stats[4] = total_s - cnt.gpu_ms * 1e-3;
    res->stats[5] = maxk;
    res->stats[6] = (double)unconv;
    res->stats[7] = total_s;
}

int pq3d_aca_run(void* h, long long nb, const long long* row_off, const int* row_idx, const long long* col_off,
                 const int* col_idx, double eps, double recompress_eps, int max_rank, unsigned long long seed,
                 long long* ranks, void** result) {
    clear_error();
    GFSource* g = (GFSource*)h;
    if (int rc = aca_check(g, nb, row_off, row_idx, col_off, col_idx, eps)) return rc;
    const auto t_start = std::chrono::steady_clock::now();
    int* d_idx = nullptr;
    if (upload_ids(row_off, row_idx, col_off, col_idx, nb, &d_idx)) return -1;
    AcaResult* res = new AcaResult();
    res->nb = nb;
    aca_init_instances(nb, row_off, col_off, nullptr, max_rank, seed, res->inst);
    AcaCounters cnt;
    const int rc = aca_host_loop(g, d_idx, row_off, col_off, nb, eps, max_rank, res->inst, cnt);
    (void)hipFree(d_idx);
    if (rc) {
        delete res;
        return rc;
    }
    aca_finish(res, recompress_eps, ranks, cnt, t_start);
    *result = res;
    return 0;
}

// The same ACA with the pivoting on the device (one wave per block and
// matrix; no per-sweep transfer of values).  Factors of at most kcap terms
// live on the device (kcap <= 0: 64, the most); an instance that needs more is redone
// by the host loop from its own seed.  Results, ranks and statistics as
// pq3d_aca_run (pq3d_aca_fetch/stats/phases/free apply).
int pq3d_aca_run_device(void* h, long long nb, const long long* row_off, const int* row_idx,
                        const long long* col_off, const int* col_idx, double eps, double recompress_eps,
                        int max_rank, unsigned long long seed, int kcap, long long* ranks, void** result) {
    clear_error();
    GFSource* g = (GFSource*)h;
    if (int rc = aca_check(g, nb, row_off, row_idx, col_off, col_idx, eps)) return rc;
    if (kcap <= 0) kcap = 64;
    PQ_CHECK(kcap <= 64, "pq3d_aca_run_device: kcap %d above 64 (one SVD column per lane)", kcap);
    if (nb == 0)  // nothing to launch: the host loop's empty result
        return pq3d_aca_run(h, nb, row_off, row_idx, col_off, col_idx, eps, recompress_eps, max_rank, seed, ranks,
                            result);
    using clk = std::chrono::steady_clock;
    const auto t_start = clk::now();
    AcaResult* res = new AcaResult();
    res->nb = nb;
    aca_init_instances(nb, row_off, col_off, nullptr, max_rank, seed, res->inst);
    const long long ni = (long long)res->inst.size();
    // device instances and their buffer offsets
    std::vector<DAca> dh((size_t)ni);
    long long dtot = 0, ftot = 0;
    for (long long x = 0; x < ni; ++x) {
        const Inst& s = res->inst[(size_t)x];
        DAca d;
        memset(&d, 0, sizeof(d));
        d.b = s.b;
        d.c = s.c;
        d.m = s.m;
        d.n = s.n;
        d.kmax = s.kmax;
        d.kcap = std::min(s.kmax, kcap);
        d.state = ST_REFS;
        d.iref = s.iref;
        d.jref = s.jref;
        d.istar = d.jstar = -1;
        d.nq = s.nq;
        d.need_rr = d.need_rc = 1;
        d.first = 1;
        for (int t = 0; t < 2; ++t) {
            d.qt[t] = s.qt[t];
            d.qi[t] = s.qi[t];
        }
        d.rng = s.rng;
        d.offU = dtot;
        dtot += (long long)d.kcap * s.m;
        d.offV = dtot;
        dtot += (long long)d.kcap * s.n;
        d.offM = dtot;
        dtot += 3LL * s.m;
        d.offN = dtot;
        dtot += 3LL * s.n;
        d.offFm = ftot;
        ftot += s.m;
        d.offFn = ftot;
        ftot += s.n;
        dh[(size_t)x] = d;
    }
    std::vector<long long> cnt0((size_t)ni);
    for (long long x = 0; x < ni; ++x) cnt0[(size_t)x] = (long long)dh[(size_t)x].m + dh[(size_t)x].n;
    int* d_idx = nullptr;
    DAca* d_inst = nullptr;
    double* d_buf = nullptr;
    unsigned char* d_flag = nullptr;
    long long *d_cnt = nullptr, *d_off = nullptr, *d_offs = nullptr;
    AcaCounters cnt;
    int rc = upload_ids(row_off, row_idx, col_off, col_idx, nb, &d_idx);
    hipError_t e = hipSuccess;
// each step runs only while every earlier one succeeded
#define ACA_TRY(x)                        \
    do {                                  \
        if (e == hipSuccess) e = (x);     \
    } while (0)
    if (!rc) {
        ACA_TRY(hipMalloc((void**)&d_inst, (size_t)ni * sizeof(DAca)));
        ACA_TRY(hipMalloc((void**)&d_buf, (size_t)std::max(dtot, 1LL) * sizeof(double)));
        ACA_TRY(hipMalloc((void**)&d_flag, (size_t)std::max(ftot, 1LL)));
        ACA_TRY(hipMalloc((void**)&d_cnt, (size_t)ni * sizeof(long long)));
        ACA_TRY(hipMalloc((void**)&d_off, (size_t)(ni + 1) * sizeof(long long)));
        ACA_TRY(hipMalloc((void**)&d_offs, (size_t)(2 * nb + 2) * sizeof(long long)));
        ACA_TRY(hipMemcpy(d_inst, dh.data(), (size_t)ni * sizeof(DAca), hipMemcpyHostToDevice));
        ACA_TRY(hipMemset(d_buf, 0, (size_t)std::max(dtot, 1LL) * sizeof(double)));
        ACA_TRY(hipMemset(d_flag, 0, (size_t)std::max(ftot, 1LL)));
        ACA_TRY(hipMemcpy(d_cnt, cnt0.data(), (size_t)ni * sizeof(long long), hipMemcpyHostToDevice));
        ACA_TRY(hipMemcpy(d_offs, row_off, (size_t)(nb + 1) * sizeof(long long), hipMemcpyHostToDevice));
        ACA_TRY(hipMemcpy(d_offs + nb + 1, col_off, (size_t)(nb + 1) * sizeof(long long), hipMemcpyHostToDevice));
    }
    const long long nr = row_off[nb];
    const long long max_sweeps = 4LL * (long long)(max_rank > 0 ? max_rank : 4096) + 64;
    hipEvent_t ea = nullptr, eb = nullptr;
    if (!rc && e == hipSuccess) {
        ACA_TRY(hipEventCreate(&ea));
        ACA_TRY(hipEventCreate(&eb));
    }
    const auto t_swept0 = clk::now();
    res->dev[1] = std::chrono::duration<double>(t_swept0 - t_start).count();
    bool first_sweep = true;
    while (!rc && e == hipSuccess) {
        const auto t0 = clk::now();
        if (!first_sweep)
            hipLaunchKernelGGL(aca_dev_logic_kernel, dim3((unsigned)((ni + 3) / 4)), dim3(256), 0, 0, d_inst, ni,
                               d_buf, d_flag, eps, d_cnt);
        first_sweep = false;
        hipLaunchKernelGGL(aca_dev_scan_kernel, dim3(1), dim3(1024), 0, 0, d_cnt, ni, d_off);
        long long total = 0;
        ACA_TRY(hipGetLastError());
        ACA_TRY(hipMemcpy(&total, d_off + ni, sizeof(long long), hipMemcpyDeviceToHost));
        if (e != hipSuccess) break;
        const auto t1 = clk::now();
        if (total == 0) break;
        if (++cnt.sweeps > max_sweeps) {
            set_error("pq3d_aca_run_device: no convergence after %lld sweeps", cnt.sweeps);
            rc = -3;
            break;
        }
        const long long nblk = (total + 255) / 256;
        if (nblk >= (1LL << 31)) {
            set_error("pq3d_aca_run_device: sweep too large (%lld pairs)", total);
            rc = -2;
            break;
        }
        ACA_TRY(hipEventRecord(ea, 0));
        if (cnt.sweeps == 1)  // every instance's first evaluation: both slip directions, the scale
            hipLaunchKernelGGL(aca_dev_eval_kernel<2>, dim3((unsigned)nblk), dim3(256), 0, 0, g->geo, g->geo1,
                               g->nsrc, g->xg, g->ev, d_idx, d_idx + nr, d_offs, d_offs + nb + 1, d_inst, d_off, ni,
                               g->mu, g->lam, g->nu, g->halfspace, d_buf);
        else
            hipLaunchKernelGGL(aca_dev_eval_kernel<1>, dim3((unsigned)nblk), dim3(256), 0, 0, g->geo, g->geo1,
                               g->nsrc, g->xg, g->ev, d_idx, d_idx + nr, d_offs, d_offs + nb + 1, d_inst, d_off, ni,
                               g->mu, g->lam, g->nu, g->halfspace, d_buf);
        ACA_TRY(hipGetLastError());
        ACA_TRY(hipEventRecord(eb, 0));
        ACA_TRY(hipEventSynchronize(eb));
        float ms = 0;
        ACA_TRY(hipEventElapsedTime(&ms, ea, eb));
        cnt.gpu_ms += ms;
        cnt.pairs += (double)total;
        const auto t2 = clk::now();
        cnt.phase_s[2] += std::chrono::duration<double>(t1 - t0).count();  // logic + scan (device)
        cnt.phase_s[1] += std::chrono::duration<double>(t2 - t1).count();  // evaluation
    }
    const auto t_swept1 = clk::now();
    res->dev[2] = std::chrono::duration<double>(t_swept1 - t_swept0).count();
    std::vector<long long> redo;
    if (!rc && e == hipSuccess) {
        ACA_TRY(hipMemcpy(dh.data(), d_inst, (size_t)ni * sizeof(DAca), hipMemcpyDeviceToHost));
        if (e == hipSuccess)
            for (long long x = 0; x < ni; ++x)
                if (dh[(size_t)x].capped && (redo.empty() || redo.back() != dh[(size_t)x].b))
                    redo.push_back(dh[(size_t)x].b);
    }
    // instances that reached kcap: the host loop redoes their blocks and
    // recompresses them on the host
    double host_rc_s = 0;
    if (!rc && e == hipSuccess && !redo.empty()) {
        std::vector<long long> sro(1, 0), sco(1, 0);
        std::vector<int> sri, sci;
        for (long long b : redo) {
            for (long long i = row_off[b]; i < row_off[b + 1]; ++i) sri.push_back(row_idx[i]);
            for (long long i = col_off[b]; i < col_off[b + 1]; ++i) sci.push_back(col_idx[i]);
            sro.push_back((long long)sri.size());
            sco.push_back((long long)sci.size());
        }
        const long long nbs = (long long)redo.size();
        std::vector<Inst> sub;
        aca_init_instances(nbs, sro.data(), sco.data(), redo.data(), max_rank, seed, sub);
        int* d_sidx = nullptr;
        AcaCounters hc;  // the host loop's own sweep limit
        rc = upload_ids(sro.data(), sri.data(), sco.data(), sci.data(), nbs, &d_sidx);
        if (!rc) rc = aca_host_loop(g, d_sidx, sro.data(), sco.data(), nbs, eps, max_rank, sub, hc);
        if (d_sidx) (void)hipFree(d_sidx);
        cnt.sweeps += hc.sweeps;
        cnt.pairs += hc.pairs;
        cnt.gpu_ms += hc.gpu_ms;
        for (int k = 0; k < 3; ++k) cnt.phase_s[k] += hc.phase_s[k];
        if (!rc) {
            std::vector<long long> mine;
            for (long long q = 0; q < nbs; ++q)
                for (int c = 0; c < 6; ++c) {
                    const long long x = redo[(size_t)q] * 6 + c;
                    if (!dh[(size_t)x].capped) continue;
                    Inst& s = res->inst[(size_t)x];
                    Inst& t = sub[(size_t)q * 6 + c];
                    s.k = t.k;
                    s.converged = t.converged;
                    s.state = ST_DONE;
                    s.U.swap(t.U);
                    s.V.swap(t.V);
                    mine.push_back(x);
                }
            const auto t_h = clk::now();
            if (recompress_eps > 0)
                parallel_for((long long)mine.size(), [&](long long y) {
                    Inst& s = res->inst[(size_t)mine[(size_t)y]];
                    s.k = recompress_factors(s.m, s.n, s.k, s.U, s.V, recompress_eps);
                });
            host_rc_s = std::chrono::duration<double>(clk::now() - t_h).count();
        }
        res->dev[0] = (double)nbs;
    }
    // recompression of the device instances on the device, then every device
    // instance's final factors in the fetch layout (block-major, then matrix)
    const auto t_rc0 = clk::now();
    double* d_ws = nullptr;
    long long *d_wsoff = nullptr, *d_uoff = nullptr, *d_voff = nullptr;
    int *d_rank = nullptr, *d_jobs = nullptr;
    if (!rc && e == hipSuccess) {
        const bool recompress = recompress_eps > 0;
        std::vector<long long> wsoff((size_t)ni, 0);
        long long wtot = 0;
        for (long long x = 0; x < ni; ++x) {
            const long long k = dh[(size_t)x].capped ? 0 : dh[(size_t)x].k;
            wsoff[(size_t)x] = wtot;
            wtot += 4 * k * k;
        }
        std::vector<int> rk((size_t)ni, 0);
        if (recompress) {
            ACA_TRY(hipMalloc((void**)&d_ws, (size_t)std::max(wtot, 1LL) * sizeof(double)));
            ACA_TRY(hipMalloc((void**)&d_wsoff, (size_t)ni * sizeof(long long)));
            ACA_TRY(hipMalloc((void**)&d_rank, (size_t)ni * sizeof(int)));
            ACA_TRY(hipMemcpy(d_wsoff, wsoff.data(), (size_t)ni * sizeof(long long), hipMemcpyHostToDevice));
            // QR jobs grouped by rows: one wave with 64 NV rows in registers
            // (NV 1..16), one workgroup with 256 NV rows (NV 8, 16, 32), one
            // 1024-thread workgroup with the columns in memory beyond that
            const int NB = 9;
            std::vector<int> jobs[NB];
            for (long long x = 0; x < ni; ++x) {
                const DAca& d = dh[(size_t)x];
                if (d.capped || d.k == 0) continue;
                for (int w = 0; w < 2; ++w) {
                    const int rows = w ? d.n : d.m;
                    const int bkt = rows <= 64 ? 0 : rows <= 128 ? 1 : rows <= 256 ? 2 : rows <= 512 ? 3
                                  : rows <= 1024 ? 4 : rows <= 2048 ? 5 : rows <= 4096 ? 6 : rows <= 8192 ? 7 : 8;
                    jobs[bkt].push_back((int)(2 * x + w));
                }
            }
            std::vector<int> all;
            size_t jo[NB + 1] = {0};
            for (int b = 0; b < NB; ++b) {
                all.insert(all.end(), jobs[b].begin(), jobs[b].end());
                jo[b + 1] = all.size();
            }
            ACA_TRY(hipMalloc((void**)&d_jobs, std::max(all.size(), (size_t)1) * sizeof(int)));
            ACA_TRY(hipMemcpy(d_jobs, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
            for (int b = 0; b < NB && e == hipSuccess; ++b) {
                const int nj = (int)(jo[b + 1] - jo[b]);
                if (!nj) continue;
                const dim3 grid((unsigned)((nj + 3) / 4)), gridb((unsigned)nj), blk(256), blk1024(1024);
                const int* jp = d_jobs + jo[b];
                switch (b) {
                    case 0: hipLaunchKernelGGL(aca_dev_qr_kernel<1>, grid, blk, 0, 0, d_inst, jp, nj, d_buf, d_ws, d_wsoff); break;
                    case 1: hipLaunchKernelGGL(aca_dev_qr_kernel<2>, grid, blk, 0, 0, d_inst, jp, nj, d_buf, d_ws, d_wsoff); break;
                    case 2: hipLaunchKernelGGL(aca_dev_qr_kernel<4>, grid, blk, 0, 0, d_inst, jp, nj, d_buf, d_ws, d_wsoff); break;
                    case 3: hipLaunchKernelGGL(aca_dev_qr_kernel<8>, grid, blk, 0, 0, d_inst, jp, nj, d_buf, d_ws, d_wsoff); break;
                    case 4: hipLaunchKernelGGL(aca_dev_qr_kernel<16>, grid, blk, 0, 0, d_inst, jp, nj, d_buf, d_ws, d_wsoff); break;
                    case 5: hipLaunchKernelGGL((aca_dev_qr_block_kernel<8, 256, true>), gridb, blk, 0, 0, d_inst, jp, d_buf, d_ws, d_wsoff); break;
                    case 6: hipLaunchKernelGGL((aca_dev_qr_block_kernel<16, 256, true>), gridb, blk, 0, 0, d_inst, jp, d_buf, d_ws, d_wsoff); break;
                    case 7: hipLaunchKernelGGL((aca_dev_qr_block_kernel<32, 256, true>), gridb, blk, 0, 0, d_inst, jp, d_buf, d_ws, d_wsoff); break;
                    default: hipLaunchKernelGGL(aca_dev_qr_mem_block_kernel<1024>, gridb, blk1024, 0, 0, d_inst, jp, d_buf, d_ws, d_wsoff); break;
                }
                ACA_TRY(hipGetLastError());
            }
            if (e == hipSuccess) {
                hipLaunchKernelGGL(aca_dev_recompress_kernel, dim3((unsigned)((ni + 3) / 4)), dim3(256), 0, 0,
                                   d_inst, ni, d_ws, d_wsoff, recompress_eps, d_rank);
                ACA_TRY(hipGetLastError());
            }
            ACA_TRY(hipMemcpy(rk.data(), d_rank, (size_t)ni * sizeof(int), hipMemcpyDeviceToHost));
        } else {
            for (long long x = 0; x < ni; ++x) rk[(size_t)x] = dh[(size_t)x].capped ? 0 : dh[(size_t)x].k;
        }
        res->host_inst.assign((size_t)ni, 0);
        std::vector<long long> uoff((size_t)ni), voff((size_t)ni);
        long long uo = 0, vo = 0;
        for (long long x = 0; x < ni && e == hipSuccess; ++x) {
            const DAca& d = dh[(size_t)x];
            Inst& s = res->inst[(size_t)x];
            if (d.capped) {
                res->host_inst[(size_t)x] = 1;
            } else {
                s.k = rk[(size_t)x];
                s.converged = d.converged != 0;
                s.state = ST_DONE;
            }
            uoff[(size_t)x] = uo;
            voff[(size_t)x] = vo;
            uo += (long long)s.k * s.m;
            vo += (long long)s.k * s.n;
        }
        res->usz = uo;
        res->vsz = vo;
        ACA_TRY(hipMalloc((void**)&res->dU, (size_t)std::max(uo, 1LL) * sizeof(double)));
        ACA_TRY(hipMalloc((void**)&res->dV, (size_t)std::max(vo, 1LL) * sizeof(double)));
        ACA_TRY(hipMalloc((void**)&d_uoff, (size_t)ni * sizeof(long long)));
        ACA_TRY(hipMalloc((void**)&d_voff, (size_t)ni * sizeof(long long)));
        ACA_TRY(hipMemcpy(d_uoff, uoff.data(), (size_t)ni * sizeof(long long), hipMemcpyHostToDevice));
        ACA_TRY(hipMemcpy(d_voff, voff.data(), (size_t)ni * sizeof(long long), hipMemcpyHostToDevice));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(aca_dev_form_kernel, dim3((unsigned)ni), dim3(256), 0, 0, d_inst, d_buf, d_ws,
                               d_wsoff, d_rank, d_uoff, d_voff, recompress ? 0 : 1, res->dU, res->dV);
            ACA_TRY(hipGetLastError());
        }
        ACA_TRY(hipDeviceSynchronize());
    }
    const double dev_rc_s = std::chrono::duration<double>(clk::now() - t_rc0).count();
    if (eb) (void)hipEventDestroy(eb);
    if (ea) (void)hipEventDestroy(ea);
    void* bufs[] = {d_idx, d_inst, d_buf, d_flag, d_cnt, d_off, d_offs, d_ws, d_wsoff, d_rank, d_uoff, d_voff, d_jobs};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
#undef ACA_TRY
    if (!rc && e != hipSuccess) {
        set_error("pq3d_aca_run_device: %s", hipGetErrorString(e));
        rc = -1;
    }
    if (rc) {
        delete res;
        return rc;
    }
    aca_finish(res, 0.0, ranks, cnt, t_start);  // recompressed above
    res->phase_s[3] = host_rc_s + dev_rc_s;
    *result = res;
    return 0;
}

// stats: sweeps, launches, pairs evaluated, GPU seconds, host seconds, max
// rank, unconverged (block, matrix) count, total seconds
int pq3d_aca_stats(void* r, double* stats8) {
    clear_error();
    AcaResult* res = (AcaResult*)r;
    PQ_CHECK(res, "pq3d_aca_stats: null result");
    for (int k = 0; k < 8; ++k) stats8[k] = res->stats[k];
    return 0;
}

// Wall seconds of the build's phases: request assembly, evaluation (kernel and
// transfers), ACA logic on the host threads, recompression.
int pq3d_aca_phases(void* r, double* secs4) {
    clear_error();
    AcaResult* res = (AcaResult*)r;
    PQ_CHECK(res && secs4, "pq3d_aca_phases: null argument");
    for (int k = 0; k < 4; ++k) secs4[k] = res->phase_s[k];
    return 0;
}

// pq3d_aca_run_device result: blocks the host loop redid because one of their
// matrices needed more than kcap terms, then wall seconds of set-up, sweeps and
// factor readback by pq3d_aca_fetch (all 0 for pq3d_aca_run).
int pq3d_aca_device_info(void* r, double* info4) {
    clear_error();
    AcaResult* res = (AcaResult*)r;
    PQ_CHECK(res && info4, "pq3d_aca_device_info: null argument");
    for (int k = 0; k < 4; ++k) info4[k] = res->dev[k];
    return 0;
}

// Factors, block-major then matrix A1s..Bd: U (m x k, column-major) into U,
// V (k x n, row-major) into V, back to back.
int pq3d_aca_fetch(void* r, double* U, double* V) {
    clear_error();
    AcaResult* res = (AcaResult*)r;
    PQ_CHECK(res, "pq3d_aca_fetch: null result");
    size_t uo = 0, vo = 0;
    if (res->dU) {  // device result: one copy each, then the instances redone on the host
        const auto t0 = std::chrono::steady_clock::now();
        if (res->usz && d2h_staged(U, res->dU, (size_t)res->usz * sizeof(double))) return -1;
        if (res->vsz && d2h_staged(V, res->dV, (size_t)res->vsz * sizeof(double))) return -1;
        for (size_t x = 0; x < res->inst.size(); ++x) {
            const Inst& s = res->inst[x];
            const size_t us = (size_t)s.k * s.m, vs = (size_t)s.k * s.n;
            if (res->host_inst[x]) {
                if (us) memcpy(U + uo, s.U.data(), us * sizeof(double));
                if (vs) memcpy(V + vo, s.V.data(), vs * sizeof(double));
            }
            uo += us;
            vo += vs;
        }
        res->dev[3] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    }
    for (auto& s : res->inst) {
        const size_t us = (size_t)s.k * s.m, vs = (size_t)s.k * s.n;
        if (us) memcpy(U + uo, s.U.data(), us * sizeof(double));
        if (vs) memcpy(V + vo, s.V.data(), vs * sizeof(double));
        uo += us;
        vo += vs;
    }
    return 0;
}

// Host-only: recompress U (m x k, column-major) V (k x n, row-major) to
// relative Frobenius tolerance eps; *r receives the new rank and Uout (m x r,
// column-major), Vout (r x n, row-major) the factors (capacity m*k, k*n).
int pq3d_lowrank_recompress(long long m, long long n, long long k, const double* U, const double* V, double eps,
                            long long* r, double* Uout, double* Vout) {
    clear_error();
    PQ_CHECK(m > 0 && n > 0 && k >= 0 && eps > 0, "pq3d_lowrank_recompress: bad arguments");
    std::vector<double> u(U, U + m * k), v(V, V + k * n);
    const int rr = recompress_factors((int)m, (int)n, (int)k, u, v, eps);
    if (rr) {
        memcpy(Uout, u.data(), (size_t)m * rr * sizeof(double));
        memcpy(Vout, v.data(), (size_t)rr * n * sizeof(double));
    }
    *r = rr;
    return 0;
}

int pq3d_aca_free(void* r) {
    clear_error();
    delete (AcaResult*)r;
    return 0;
}

}  // extern "C"
