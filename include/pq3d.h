/*
 * pq3d.h -- C ABI of libpq3d.so, the MI355X (gfx950) implementation of the
 * PyQuake3D per-step hot path.
 *
 * Plain pointers and sizes only.  Unless stated otherwise, every pq3d_*
 * function returns 0 on success and a negative status on failure, with a
 * message in pq3d_last_error() (thread-local).  Pointers named d_* are device
 * pointers (from pq3d_malloc); all other array pointers are host memory.
 *
 * Part 1 keeps the reference's ctypes signatures unchanged so that the
 * reference's own binding (src/Hmatrix.py:28-58) loads this library in place
 * of TDstressFS_C.so.
 */
#ifndef PQ3D_H
#define PQ3D_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * 1. Drop-in triangular-dislocation stress entry points
 * ---------------------------------------------------------------------- */

/* Replaces TDstressFS_C (src/TDstressFS_C.cpp:2095-2152), bound at
 * src/Hmatrix.py:32-43.  n receivers (X, Y, Z) x one TD (P1, P2, P3: 3 each);
 * unit-less Burgers components Ss (strike), Ds (dip), Ts (tensile); Lame mu,
 * lambda_; half space if jud_halfspace.  Writes stress[i*6+k], strain[i*6+k],
 * k over [xx, yy, zz, xy, xz, yz].  Computed on the GPU.  No status return
 * (as the reference); on failure a message goes to stderr and the outputs
 * are NaN. */
void TDstressFS_C(const double* X, const double* Y, const double* Z, size_t n, const double* P1,
                  const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu,
                  double lambda_, bool jud_halfspace, double* stress, double* strain);

/* Replaces TDstressEachSourceAtReceiver_C (src/TDstressFS_C.cpp:2155-2200),
 * bound at src/Hmatrix.py:45-56.  One receiver x n_sources TDs; P*_flat are
 * n_sources x 3 (C order); outputs n_sources x 6. */
void TDstressEachSourceAtReceiver_C(double x, double y, double z, const double* P1_flat,
                                    const double* P2_flat, const double* P3_flat, size_t n_sources,
                                    double Ss, double Ds, double Ts, double mu, double lambda_,
                                    bool jud_halfspace, double* stress_out, double* strain_out);

/* Replaces TDstressFS_C1 (src/TDstressFS_C.cpp:570-609): full space, twelve
 * separate n-vectors. */
void TDstressFS_C1(const double* X, const double* Y, const double* Z, size_t n, const double* P1,
                   const double* P2, const double* P3, double Ss, double Ds, double Ts, double mu,
                   double lambda_, double* Sxx, double* Syy, double* Szz, double* Sxy, double* Sxz,
                   double* Syz, double* Exx, double* Eyy, double* Ezz, double* Exy, double* Exz,
                   double* Eyz);

/* ------------------------------------------------------------------------
 * 2. Runtime: errors, devices, memory, events
 * ---------------------------------------------------------------------- */
const char* pq3d_last_error(void);
int pq3d_device_count(int* n);
int pq3d_device_info(int dev, char* name, int name_len, int* cu_count, double* hbm_bytes,
                     int* arch_major, int* arch_minor);
int pq3d_malloc(void** d_ptr, size_t bytes);
int pq3d_free(void* d_ptr);
int pq3d_memcpy_h2d(void* d_dst, const void* src, size_t bytes);
int pq3d_memcpy_d2h(void* dst, const void* d_src, size_t bytes);
int pq3d_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);
int pq3d_memset(void* d_dst, int value, size_t bytes);
int pq3d_synchronize(void);
int pq3d_event_create(void** ev);
int pq3d_event_destroy(void* ev);
int pq3d_event_record(void* ev);
int pq3d_event_elapsed_ms(void* start, void* stop, float* ms);

/* Measurement probes (no counterpart in the reference): what this device
 * sustains for the resources the hot path is bound by, quoted beside the
 * published peaks.  pq3d_probe_stream: read-only stream over a buffer of
 * `bytes` it allocates (a contiguous share per workgroup, 16-byte loads, eight
 * in flight per lane, 32 workgroups per CU), GB/s and ms of the fastest of
 * `reps` passes.  pq3d_probe_fma64:
 * independent FP64 FMA chains at full occupancy, TFLOP/s (an FMA = 2 flops)
 * and ms per launch. */
int pq3d_probe_stream(long long bytes, int reps, float* ms_per, double* gbs);
int pq3d_probe_fma64(int iters, int reps, float* ms_per, double* tflops);

/* Matrix storage types */
#define PQ3D_MAT_F64 0
#define PQ3D_MAT_F32 1
#define PQ3D_MAT_BF16 2

/* ------------------------------------------------------------------------
 * 3. Dense traction Green's functions
 * ---------------------------------------------------------------------- */

/* Six traction Green's functions A1s, A2s, Bs, A1d, A2d, Bd: receiver
 * centroid (row) x source TD (column) with unit strike (Ss=1) or dip (Ds=1)
 * slip, projected on the receiver frame [e1, e2, n] -- replaces
 * QDsim_gpu.get_coreAB_mulproess / worker / GetTtstress
 * (src/QDsim_gpu.py:1745-1878, :1400-1417) and, with index lists, the dense
 * block evaluators of src/Hmatrix.py:477-591.
 *   P1, P2, P3 : nsrc x 3 source vertices; xg: nrcv x 3; ev: nrcv x 9
 *   rows, cols : optional index lists (nr, nc entries), NULL = all
 *   d_out[6]   : device buffers; element (i, j) of the nr x nc block at
 *                j*ld + i (source-major), out_type PQ3D_MAT_F64 or _F32
 *   src_per_block: consecutive sources (columns) walked per one-wave
 *                workgroup (<= 0: the longest run up to 64 that still gives
 *                8,192 workgroups).  Within a run, the free-surface
 *                correction of a side shared by consecutive sources is
 *                evaluated once for both; entries agree with the
 *                every-side evaluation to rounding (a few 1e-11 relative). */
int pq3d_gf_assemble(const double* P1, const double* P2, const double* P3, long long nsrc,
                     const double* xg, const double* ev, long long nrcv, double mu, double lam,
                     int halfspace, int out_type, void* const* d_out, long long ld,
                     const long long* rows, long long nr, const long long* cols, long long nc,
                     int src_per_block);

/* No counterpart in the reference.  pq3d_gf_assemble for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc): the same arguments and results, with the staging (source
 * geometry table, copied vertices, receivers and index lists) in a fixed
 * 64 MiB device scratch that each GPU holds until the process ends.  Rejects
 * (status -2) an input whose staging exceeds it: about 1,184 B per assembled
 * source column + 72 B per source + 96 B per receiver + 8 B per index,
 * i.e. a whole mesh of up to ~50k elements per call, larger ones in column
 * chunks.  Calls that overlap on one GPU race on the scratch. */
int pq3d_gf_assemble_noalloc(const double* P1, const double* P2, const double* P3, long long nsrc,
                             const double* xg, const double* ev, long long nrcv, double mu, double lam,
                             int halfspace, int out_type, void* const* d_out, long long ld,
                             const long long* rows, long long nr, const long long* cols, long long nc,
                             int src_per_block);

/* ------------------------------------------------------------------------
 * 4. Dense quasi-dynamic simulation (device-resident state)
 *
 * Replaces QDsim_gpu.QDsim's GPU state and stepping
 * (src/QDsim_gpu.py:893-1027): initGPUvariable -> create/set_field,
 * simu_forwardGPU -> pq3d_dense_step, *_gpu.get() -> pq3d_dense_get_field.
 * Fields (index order): Tno Tt1o Tt2o state slipv1 slipv2 slipv slip1 slip2
 * slip Tt rake rake0 fric a b dc f0 shear_loading normal_loading P dPdt0
 * AdotV1 AdotV2 dsigmadt (units as in the reference: MPa, m/s, m, rad).
 * ---------------------------------------------------------------------- */
int pq3d_dense_create(void** sim, long long n, int mat_type, int kc);
int pq3d_dense_destroy(void* sim);
/* six device matrices A1s A2s Bs A1d A2d Bd, source-major with leading dim *ld */
int pq3d_dense_matrix_ptrs(void* sim, void** d_ptrs6, long long* ld);
/* device memory held by the simulation (no reference counterpart; used for
 * HBM budgets) */
int pq3d_dense_device_bytes(void* sim, double* bytes);
int pq3d_dense_set_field(void* sim, int field, const double* host_n);
int pq3d_dense_get_field(void* sim, int field, double* host_n);
int pq3d_dense_field_ptr(void* sim, int field, void** d_ptr);
/* eta = mu / (2 Cs); err_mode 1 = GPU-path error estimate (QDsim_gpu.py:1238),
 * 0 = CPU-path estimate (QDsim_gpu.py:1373) */
int pq3d_dense_set_params(void* sim, double V0, double Vpl, double eta, int fix_tn, int err_mode);
int pq3d_dense_set_time(void* sim, double t);
/* out6 = time, next dt, last h, last dtnext, max slip rate, last error2 */
int pq3d_dense_state(void* sim, double* out6, int* nrjct, int* status, long long* steps);
/* Relerrormax1 / Relerrormax2 (src/QDsim_gpu.py:1221-1238, read by the
 * reference driver's CPU branch, src/main_gpu.py:214): the relative-error
 * maxima (+1e-10) of the last accepted step, or of the last pq3d_dense_dp45
 * attempt if one came after it */
int pq3d_dense_errors(void* sim, double* err1, double* err2);
/* Replaces derivative_gpu(Tno, Tt1o, Tt2o, state) (src/QDsim_gpu.py:830-891):
 * the rate-and-state RHS at the given device vectors (n each: MPa, MPa, MPa,
 * dimensionless) with the frozen traction rates of the AdotV1 / AdotV2 /
 * dsigmadt fields and the current slip rate; d_out4 (device, 4 x n) receives
 * the reference's 4-tuple dstate/dt, dsigma/dt, dtau1/dt, dtau2/dt (MPa/s) */
int pq3d_dense_derivative(void* sim, const double* d_Tno, const double* d_Tt1o, const double* d_Tt2o,
                          const double* d_state, double* d_out4);
/* Replaces RungeKutte_solve_Dormand_Prince_gpu(h) (src/QDsim_gpu.py:1128-1262)
 * and, with err_mode 0, RungeKutte_solve_Dormand_Prince(h) (:1266-1397): one
 * Dormand-Prince attempt of step h from the current state with the frozen
 * traction rates, nothing committed.  d_out4 (device, 4 x n): the solution
 * Tno, Tt1o, Tt2o, state; out4 (host): condition1, condition2, hnew1, hnew2
 * (the rest of the reference's 8-tuple).  err_mode 1 drops the
 * normal-traction error as the GPU method does (:1238), 0 keeps it (:1373). */
int pq3d_dense_dp45(void* sim, double h, int err_mode, double* d_out4, double* out4);
/* traction rates from the current slip rates into AdotV1/AdotV2/dsigmadt */
int pq3d_dense_matvec(void* sim);
int pq3d_dense_time_matvec(void* sim, int reps, float* ms_per_launch);
/* split-K tile width of the dense product (source columns per tile; no
 * reference counterpart: the reference's product is one CuPy matmul).
 * set_kc waits for the device and reallocates the partial sums; time_product
 * times product + partial-sum pass, as the autotuner compares widths */
int pq3d_dense_set_kc(void* sim, int kc);
int pq3d_dense_get_kc(void* sim, int* kc, int* nchunk);
int pq3d_dense_time_product(void* sim, int reps, float* ms_per_product);
/* zero-matrix structure (no reference counterpart): scan the six matrices
 * (mask bit k = matrix k holds only +0/-0) and skip Bs, Bd in the product when
 * both are zero (planar fault); results are bit-identical.  Rescan after the
 * matrices change; set_structure(0) streams all six */
int pq3d_dense_scan_structure(void* sim, int* mask);
int pq3d_dense_set_structure(void* sim, int mask);
int pq3d_dense_get_structure(void* sim, int* mask, int* skip_normal);
/* one accepted step; returns 1 with 'error: dt is too small' in
 * pq3d_last_error() where the reference exits (QDsim_gpu.py:983-985) */
int pq3d_dense_step(void* sim, double dttry, double* h, double* dtnext, int* nrjct);
/* nsteps steps without host synchronisation (HIP graph of graph_steps steps);
 * records (host, optional): nsteps x 4 = h, dtnext, time, max slip rate */
int pq3d_dense_run(void* sim, long long nsteps, int graph_steps, double* records, int* status);
/* after a failed step (status 1: the reference's 'dt is too small' exit,
 * QDsim_gpu.py:983-985) every later launch of a run or graph leaves the state
 * unchanged and its records rows are NaN; reset_ctrl clears the failure (status,
 * grid barrier, reduction slots) keeping time, step size and counters */
int pq3d_dense_reset_ctrl(void* sim);
/* Dilatancy / pore-pressure diffusion (QDsim_gpu.py:336-349, :611-671):
 * zp K grid nodes, parr n x K initial profiles in Pa (row per element),
 * c_hyd, eps (dilatancy coefficient), hs, beta (effective compressibility),
 * P0 fixed far-field pressure (Pa), V0; every accepted step then updates the
 * profiles, P and dPdt0 on the device (K <= 257) */
int pq3d_dense_set_dilatancy(void* sim, int K, const double* zp, const double* parr, double c_hyd, double eps,
                             double hs, double beta, double P0, double V0);
int pq3d_dense_get_dilatancy(void* sim, double* parr);
/* one pore-pressure update with step dt (Calc_P_implicit(dt), QDsim_gpu.py:628) */
int pq3d_dense_dilatancy_update(void* sim, double dt);

/* ------------------------------------------------------------------------
 * 5. Ensemble: B = 64 simulations on one mesh (no reference counterpart;
 *    the reference runs independent QDsim_gpu instances)
 *
 * One global step = one tall-skinny MFMA GEMM streaming the six FP32 matrices
 * once for all members, then up to four Dormand-Prince attempt rounds per
 * active member with its own step size (a member still rejecting continues
 * at the next global step with the same frozen traction rates).  All work of
 * an ensemble runs on its own stream.  Field arrays are B x n (member-major),
 * field indices as in section 4.
 * ---------------------------------------------------------------------- */
int pq3d_ens_create(void** ens, long long n, int B, void* const* d_mats_or_null, long long ld, int kc);
int pq3d_ens_destroy(void* ens);
int pq3d_ens_matrix_ptrs(void* ens, void** d_ptrs6, long long* ld);
/* device memory held by the ensemble (fields, operands, partial sums, the
 * split GEMM's planes and panel copies; borrowed matrices not counted) */
int pq3d_ens_device_bytes(void* ens, double* bytes);
int pq3d_ens_set_field(void* ens, int field, const double* host_Bn);
int pq3d_ens_get_field(void* ens, int field, double* host_Bn);
int pq3d_ens_set_member_field(void* ens, int member, int field, const double* host_n);
int pq3d_ens_get_member_field(void* ens, int member, int field, double* host_n);
/* V0, Vpl, eta = mu / (2 Cs); fix_tn != 0 is the reference's Fix_Tn
 * (src/QDsim_gpu.py:950-953): dsigma/dt is the normal loading alone and the
 * GEMM leaves out Bs and Bd */
int pq3d_ens_set_params(void* ens, double V0, double Vpl, double eta, int fix_tn);
/* per member: d6 = time, next trial step, last accepted step, last dtnext,
 * t_end, 0; i6 = rejections since the last accepted step, status, active,
 * finished, accepted steps, global step at which it finished */
int pq3d_ens_members(void* ens, double* d6, long long* i6);
/* (re)start a slot: time 0, trial step h0, finish when time >= t_end (0: never) */
int pq3d_ens_reset_member(void* ens, int member, double h0, double t_end, int active);
/* one operand preparation + GEMM of the members' current slip rates, then
 * `reps` further launches of the GEMM kernel alone between two events (ms per
 * GEMM launch; reps = 0 only computes) */
int pq3d_ens_gemm(void* ens, int reps, float* ms_per_launch);
int pq3d_ens_gemm_result(void* ens, double* host_3Bn);
/* split-K tile width of the GEMM (64 receivers x 64 members x kc sources per
 * wave; kc rounded up to a multiple of 4, of 32 in precision mode 1);
 * time_product = operand prep + GEMM + one pass over the partial sums */
int pq3d_ens_set_kc(void* ens, int kc);
int pq3d_ens_get_kc(void* ens, int* kc, int* nchunk);
int pq3d_ens_time_product(void* ens, int reps, float* ms_per_product);
/* GEMM arithmetic (no reference counterpart): mode 0 FP32 MFMA; mode 1 each
 * FP32 operand split exactly into three BF16 parts and the six leading
 * products accumulated in FP32 on the BF16 MFMA units (dropped terms 2^-24
 * of each product and below): a shear pass over A1s, A2s, A1d, A2d and,
 * unless Bs = Bd = 0 or fix_tn, a normal pass over Bs, Bd in the same
 * launch.  Mode 1 reads panel copies of the matrices in its operand order
 * (64 rows x 32 sources per 8 KB block), built here and whenever the
 * structure is scanned or set: after changing the matrices, rescan.
 * split_active says whether the split GEMM is the one launched */
int pq3d_ens_set_precision(void* ens, int mode);
int pq3d_ens_get_precision(void* ens, int* mode, int* split_active);
/* zero-matrix structure of the GEMM's six matrices, as for the dense product */
int pq3d_ens_scan_structure(void* ens, int* mask);
int pq3d_ens_set_structure(void* ens, int mask);
int pq3d_ens_get_structure(void* ens, int* mask, int* skip_normal);
/* enqueue nsteps global steps (HIP graph of graph_steps); async != 0 returns
 * without waiting */
int pq3d_ens_run(void* ens, long long nsteps, int graph_steps, int async);
int pq3d_ens_sync(void* ens);
int pq3d_ens_query(void* ens, int* done);

/* ------------------------------------------------------------------------
 * 6. H-matrix: batched Green's-function sub-blocks, ACA, device product
 *
 * Replaces the block/ACA layer of src/Hmatrix.py: the row/column/sub-block
 * evaluators calc_stressfunc_fromC* / calc_rows* / calc_cols* (:329-658),
 * ACA_plus_full_traceS/D (:661-1289), ACA_worker (:1292-1344) and
 * blocks_process_MVM(_gpu) (:1773-1994, caller src/QDsim.py:1153-1160).
 * ---------------------------------------------------------------------- */
#define PQ3D_MAT_NONE 3 /* pq3d_dense_create without matrices (H-matrix attached) */

/* device-resident geometry: nsrc source TDs (P* nsrc x 3), nrcv receivers
 * (xg nrcv x 3, ev nrcv x 9 frames [e1, e2, n]) */
int pq3d_gfsrc_create(void** src, const double* P1, const double* P2, const double* P3, long long nsrc,
                      const double* xg, const double* ev, long long nrcv, double mu, double lam, int halfspace);
int pq3d_gfsrc_destroy(void* src);
/* nseg sub-blocks; seg[4k..4k+3] = rcv_off, m, src_off, n into rcv_idx /
 * src_idx (element ids); out (host) receives per segment 6 x n x m FP64
 * ([A1s..Bd][column][row]), segments back to back */
int pq3d_gf_segments(void* src, const int* rcv_idx, long long nr_idx, const int* src_idx, long long ns_idx,
                     long long nseg, const long long* seg, double* out);
/* ACA of the six matrices of nb blocks (rows row_idx[row_off[b]:row_off[b+1]],
 * columns likewise), relative tolerance eps (a negative eps selects the
 * reference's absolute rule of src/Hmatrix.py:797,832-838: stop after a term
 * whose pivot is below 1e-4 or whose step ||u|| ||v|| is below |eps|, the
 * reference calling it with 0.01), then (recompress_eps > 0) QR +
 * SVD recompression to that tolerance; max_rank <= 0 for min(m, n);
 * ranks: nb x 6 out; *result holds the factors */
int pq3d_aca_run(void* src, long long nb, const long long* row_off, const int* row_idx, const long long* col_off,
                 const int* col_idx, double eps, double recompress_eps, int max_rank, unsigned long long seed,
                 long long* ranks, void** result);
/* the same ACA (same seeds, same stopping rules) with the pivot logic on the
 * device: one wave per (block, matrix), one count read back per sweep instead
 * of every evaluated row and column, and the recompression (same algorithm)
 * on the device too; the final factors stay in HBM until pq3d_aca_fetch
 * copies them out.  Factors of at most kcap terms live on the device
 * (kcap <= 0: 64, the most); an instance needing more is redone and
 * recompressed by the host loop from its own seed.  Sums run in another order
 * than the host loop's, so pivots may differ where residual entries nearly
 * tie; the stopping rules and their tolerances are the same.  Results as
 * pq3d_aca_run. */
int pq3d_aca_run_device(void* src, long long nb, const long long* row_off, const int* row_idx,
                        const long long* col_off, const int* col_idx, double eps, double recompress_eps,
                        int max_rank, unsigned long long seed, int kcap, long long* ranks, void** result);
/* info4 of a pq3d_aca_run_device result: blocks redone by the host loop, wall
 * seconds of set-up, sweeps, factor readback in pq3d_aca_fetch (zeros for
 * pq3d_aca_run) */
int pq3d_aca_device_info(void* result, double* info4);
/* host only: U (m x k column-major) V (k x n row-major) recompressed to
 * relative Frobenius tolerance eps; rank in *r, factors in Uout / Vout
 * (the reference's SVD_recompress, src/Hmatrix.py:116-141) */
int pq3d_lowrank_recompress(long long m, long long n, long long k, const double* U, const double* V, double eps,
                            long long* r, double* Uout, double* Vout);
/* stats8: sweeps, launches, pairs evaluated, GPU s, host s, max rank, unconverged, total s */
int pq3d_aca_stats(void* result, double* stats8);
/* wall seconds: request assembly, evaluation incl. transfers, host ACA logic,
 * recompression */
int pq3d_aca_phases(void* result, double* secs4);
/* U (m x k column-major) and V (k x n row-major) per block and matrix A1s..Bd, back to back */
int pq3d_aca_fetch(void* result, double* U, double* V);
int pq3d_aca_free(void* result);

/* device H-matrix from leaves (kind 0 dense: six m x n column-major FP64;
 * kind 1 low-rank: ranks[6b+c], U m x k column-major, V k x n row-major);
 * perm: cluster position -> element id; factor_type PQ3D_MAT_F32 or _BF16 */
int pq3d_hmat_create(void** hm, long long n, const int* perm, long long nb, const long long* row_off,
                     const int* row_idx, const long long* col_off, const int* col_idx, const int* kind,
                     const long long* ranks, const double* dense, const double* U, const double* V,
                     int factor_type);
int pq3d_hmat_destroy(void* hm);
/* info10: dense-leaf bytes, factor bytes, algorithmic bytes per product,
 * same with fix_tn, segments, row-item waves, column-slab waves, low-rank blocks,
 * stored (padded) bytes read per product, same with fix_tn */
int pq3d_hmat_info(void* hm, double* info10);
/* y (3 x n) = (A1s x1 + A1d x2, A2s x1 + A2d x2, Bs x1 + Bd x2) (host arrays) */
int pq3d_hmat_mvm(void* hm, const double* x1, const double* x2, int fix_tn, double* y);
int pq3d_hmat_time_mvm(void* hm, int reps, int fix_tn, float* ms_per_product);
/* the dense simulation's traction rates come from hm from now on */
int pq3d_dense_attach_hmatrix(void* sim, void* hm);

#ifdef __cplusplus
}
#endif

#endif /* PQ3D_H */
