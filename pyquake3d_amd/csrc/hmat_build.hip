// H-matrix construction on the GPU: batched evaluation of traction Green's
// function sub-blocks (rows, columns, dense leaves) and adaptive cross
// approximation (ACA) of the admissible blocks.
//
// Reference interface replaced: the ACA of src/Hmatrix.py:661-1289
// (ACA_plus_full_traceS/D) with its row/column evaluators :329-658, and the
// dense leaves of ACA_worker :1292-1344.  The reference evaluates one row or
// column per ctypes call and pivots one block at a time; here every sweep
// evaluates the pending pivot rows/columns of ALL blocks in one launch
// (``gf_segments_kernel``), and the pivot logic of all blocks advances in
// lock step on host threads.
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
#include "common.hpp"
#include "td_field.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <thread>
#include <unordered_map>
#include <vector>

extern "C" const char* pq3d_last_error(void);

namespace pq3d {

__global__ void gfsrc_setup_kernel(const double* __restrict__ P, long long n, SrcGeom<2>* __restrict__ out) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double b[2][3] = {{0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};  // strike (Ss = 1), dip (Ds = 1)
    SrcGeom<2> g;
    src_setup<2>(P + 3 * j, P + 3 * (n + j), P + 3 * (2 * n + j), b, g);
    out[j] = g;
}

struct GFSource {
    long long nsrc = 0, nrcv = 0;
    SrcGeom<2>* geo = nullptr;
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
    PQ_HIP(hipMemcpy(host_out, g->out, (size_t)total_out * sizeof(double), hipMemcpyDeviceToHost));
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
    const double floor = 1e-9 * s.scale * std::sqrt((double)m * (double)n);
    s.pending_conv = step <= eps * std::sqrt(s.norm2) || step <= floor;
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
};

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
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (dP) (void)hipFree(dP);
    if (e != hipSuccess) {
        set_error("pq3d_gfsrc_create: %s", hipGetErrorString(e));
        (void)hipFree(g->geo);
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
int pq3d_aca_run(void* h, long long nb, const long long* row_off, const int* row_idx, const long long* col_off,
                 const int* col_idx, double eps, double recompress_eps, int max_rank, unsigned long long seed,
                 long long* ranks, void** result) {
    clear_error();
    GFSource* g = (GFSource*)h;
    PQ_CHECK(g, "pq3d_aca_run: null source");
    PQ_CHECK(eps > 0 && eps < 1, "pq3d_aca_run: eps must be in (0, 1)");
    const long long nr = row_off[nb], nc = col_off[nb];
    for (long long b = 0; b < nb; ++b)
        PQ_CHECK(row_off[b + 1] > row_off[b] && col_off[b + 1] > col_off[b], "pq3d_aca_run: empty block %lld", b);
    for (long long i = 0; i < nr; ++i)
        PQ_CHECK(row_idx[i] >= 0 && row_idx[i] < g->nrcv, "pq3d_aca_run: row id out of range");
    for (long long i = 0; i < nc; ++i)
        PQ_CHECK(col_idx[i] >= 0 && col_idx[i] < g->nsrc, "pq3d_aca_run: column id out of range");
    auto t_start = std::chrono::steady_clock::now();
    AcaResult* res = new AcaResult();
    res->nb = nb;
    int* d_idx = nullptr;
    hipError_t e = hipMalloc((void**)&d_idx, (size_t)(nr + nc + 1) * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(d_idx, row_idx, nr * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_idx + nr, col_idx, nc * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d_idx);
        delete res;
        set_error("pq3d_aca_run: %s", hipGetErrorString(e));
        return -1;
    }
    res->inst.resize((size_t)nb * 6);
    for (long long b = 0; b < nb; ++b) {
        const int m = (int)(row_off[b + 1] - row_off[b]), n = (int)(col_off[b + 1] - col_off[b]);
        uint64_t bs = seed ^ (0x51ED270B27A5ull * (uint64_t)(b + 1));
        const int i0 = (int)(splitmix(bs) % (uint64_t)m), j0 = (int)(splitmix(bs) % (uint64_t)n);
        for (int c = 0; c < 6; ++c) {
            Inst& s = res->inst[(size_t)b * 6 + c];
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
    std::vector<Req> reqs;
    std::vector<SegDesc> segs;
    std::vector<double> out;
    std::unordered_map<uint64_t, int> key;
    double gpu_ms = 0, pairs_total = 0;
    long long sweeps = 0;
    int rc = 0;
    const long long max_sweeps = 4LL * (long long)(max_rank > 0 ? max_rank : 4096) + 64;
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    while (true) {
        const auto t0 = clk::now();
        reqs.clear();
        key.clear();
        for (auto& s : res->inst) {
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
        if (reqs.empty()) break;
        if (++sweeps > max_sweeps) {
            set_error("pq3d_aca_run: no convergence after %lld sweeps", sweeps);
            rc = -3;
            break;
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
        rc = run_segments(g, d_idx, d_idx + nr, segs, pairs, outs, out.data(), &gpu_ms);
        if (rc) break;
        const auto t2 = clk::now();
        pairs_total += (double)pairs;
        Results R{&reqs, &segs, out.data()};
        parallel_for((long long)res->inst.size(), [&](long long x) {
            Inst& s = res->inst[(size_t)x];
            if (s.state != ST_DONE && s.nq) consume(s, R, eps);
        });
        const auto t3 = clk::now();
        res->phase_s[0] += secs(t0, t1);
        res->phase_s[1] += secs(t1, t2);
        res->phase_s[2] += secs(t2, t3);
    }
    (void)hipFree(d_idx);
    if (rc) {
        delete res;
        return rc;
    }
    const auto t_rc = clk::now();
    if (recompress_eps > 0)
        parallel_for((long long)res->inst.size(), [&](long long x) {
            Inst& s = res->inst[(size_t)x];
            s.k = recompress_factors(s.m, s.n, s.k, s.U, s.V, recompress_eps);
        });
    res->phase_s[3] = secs(t_rc, clk::now());
    int maxk = 0;
    long long unconv = 0;
    for (auto& s : res->inst) {
        ranks[(size_t)s.b * 6 + s.c] = s.k;
        maxk = std::max(maxk, s.k);
        if (!s.converged) ++unconv;
    }
    const double total_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    res->stats[0] = (double)sweeps;
    res->stats[1] = (double)sweeps;
    res->stats[2] = pairs_total;
    res->stats[3] = gpu_ms * 1e-3;
    res->stats[4] = total_s - gpu_ms * 1e-3;
    res->stats[5] = maxk;
    res->stats[6] = (double)unconv;
    res->stats[7] = total_s;
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

// Factors, block-major then matrix A1s..Bd: U (m x k, column-major) into U,
// V (k x n, row-major) into V, back to back.
int pq3d_aca_fetch(void* r, double* U, double* V) {
    clear_error();
    AcaResult* res = (AcaResult*)r;
    PQ_CHECK(res, "pq3d_aca_fetch: null result");
    size_t uo = 0, vo = 0;
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
