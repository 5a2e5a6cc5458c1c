/*
 * CPU oracle for the Muskingum-Cunge (MC) routing path -- TEST INFRASTRUCTURE.
 *
 * This is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The shipped routing path is the
 * HIP library under noaa-gpu-kernels_amd/; it has no CPU fallback into this code.
 *
 * Every function restates the reference's FP64 host algorithm in plain C with
 * the same operation order (compile with -O2 -ffp-contract=off so no FMA
 * contraction changes rounding).  Citations are to the reference tree
 * (consigcody94/noaa-gpu-kernels):
 *
 *   orc_mc_solve            water-prediction/t-route/wavefront/wavefront_mc.cu:215-337 (mc_solve_d)
 *                           geometry helper :200-213 (hgeo_inline_d)
 *   orc_wavefront_f32in     wavefront_mc.cu:495-564 (cpu_wavefront_fp64)
 *   orc_wavefront_f64       same algorithm, every array double (FP64 end to end)
 *   orc_reach_parallel      water-prediction/t-route/troute_mc_final.cu:167-252 (cpu_mc)
 *   orc_compute_stats       wavefront_mc.cu:566-572,822-858 (pct / accuracy block)
 *
 * Converged mode (no reference counterpart; the reference stops its secant at a
 * loose 1 % / 1 cm tolerance):
 *   orc_mc_solve_converged  leftmost downcrossing of the same MC residual
 *                           (certified scan), ended in its canonical cell
 *   orc_canonical_root      that cell's root next to an estimate
 *   orc_wavefront_converged DAG + time loop of wavefront_mc.cu:495-564 with the
 *                           converged solve
 *   orc_normal_depth        depth h with Manning flow Qmn(h) == Q (training
 *                           targets for the learned initial guess)
 *
 * Linearized MC (fixed coefficients, no secant):
 *   orc_linear_coeffs       wavefront/linear_mc.cu:57-100 (compute_mc_coeffs),
 *                           float arithmetic; each powf is taken as the float
 *                           rounding of the double pow (the reference's device
 *                           powf is not correctly rounded, so no CPU code can
 *                           reproduce its coefficients bit for bit)
 *   orc_linear_mc_fp64      wavefront/linear_mc.cu:238-289 (cpu_linear_mc_fp64)
 *
 * Matrix form of the linearized MC (FP32):
 *   orc_matrix_coeffs       wavefront/linear_mc_matrix.cu:58-84 (compute_mc_coeffs_cpu)
 *   orc_matrix_build        linear_mc_matrix.cu:118-229 (build_propagation_matrices)
 *   orc_matrix_route        the step loop of linear_mc_matrix.cu:340-380 /
 *                           linear_mc_matrix_graph.cu:179-205, 355-392, with a
 *                           stated summation order
 *
 * Diffusive wave (one implicit step per reach, FP32):
 *   orc_diffusive_tridiag   water-prediction/owp_extended_kernels.cu:35-94
 *                           (cpu_diffusive_tridiag)
 *
 * Two builds of the same source: liborc.so calls the host libm's pow, as the
 * reference does (bit-for-bit the reference); liborc_cr.so (-DORC_CR_POW)
 * uses a correctly rounded pow from quad precision, the arithmetic the GPU
 * library implements with its own correctly rounded pow.
 *
 * Parity pinning (how each function is checked, see tests/test_oracle.py):
 *   - orc_mc_solve, orc_wavefront_f32in, orc_reach_parallel: bit-for-bit against
 *     the reference's own host routines compiled from the reference sources into
 *     oracle/_ref (when that build exists) and against committed golden vectors
 *     written from it (tests/golden/), plus invariants (C1+C2+C3 == 1, steady
 *     state, guards, hand-traced tiny networks).
 *   - orc_wavefront_f64: identical to orc_wavefront_f32in on float-representable
 *     inputs (checked), so it inherits that pin.
 *   - orc_linear_mc_fp64: bit-for-bit against the reference's own
 *     cpu_linear_mc_fp64 compiled into oracle/_ref, plus the steady state.
 *   - orc_linear_coeffs: compute_mc_coeffs is a __host__ __device__ function
 *     that cannot be compiled here without CUDA headers; pinned by what its
 *     arithmetic fixes: X = 0.5 for valid channels, hence C1 == 1 and
 *     C3 == -C2 exactly, C1 + C2 + C3 == 1, C4/ql == dt / D, and the
 *     celerity against a double-precision closed form.
 *   - orc_diffusive_tridiag: bit-for-bit against the reference's own
 *     cpu_diffusive_tridiag compiled into oracle/_ref, on inputs from the
 *     reference's own gen_diff_data; and its output solves the tridiagonal
 *     system it eliminates (residual in double).
 *   - orc_matrix_coeffs / orc_matrix_build: bit-for-bit against the
 *     reference's own compute_mc_coeffs_cpu and build_propagation_matrices
 *     compiled into oracle/_ref (coefficients: powf of the host libm there,
 *     the double pow rounded to float here, equal but for rare double
 *     rounding).  orc_matrix_route: against the FP64 level-scheduled linear MC
 *     of the same coefficients (with tol = 0 the matrix form is the same
 *     recurrence), and its summation orders against one another.
 *   - orc_mc_solve_converged / orc_normal_depth: residual |F(h*)| at the root,
 *     brute-force dense scan on small cases, the sign change bracketing the root,
 *     and the steady-state closed form.  No reference counterpart exists, so
 *     these are pinned by mathematics only.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Channel parameters of one segment, in the reference's field order
 * (troute_mc_final.cu:26-28 double, wavefront_mc.cu:41-44 float). */
typedef struct { double dx, bw, tw, twcc, n_ch, ncc, cs, s0; } orc_chparams_f64;
typedef struct { float  dx, bw, tw, twcc, n_ch, ncc, cs, s0; } orc_chparams_f32;

/* One segment solve; returns qdc, writes depth and velocity.  *iters (may be
 * NULL) receives the number of secant iterations performed over all attempts. */
double orc_mc_solve(double dt, double qup, double quc, double qdp, double ql,
                    const orc_chparams_f64 *p, double depthp,
                    double *depthc, double *velc, int *iters);

/* Literal twin of cpu_wavefront_fp64: float params/forcings, double outputs
 * (final-step qdc per segment and final-step last-segment qdc per reach).
 * iter_sum (may be NULL) accumulates secant iterations over all solves. */
void orc_wavefront_f32in(const orc_chparams_f32 *params,
                         const int *reach_seg_start, const int *reach_seg_len,
                         const int *reach_up_ptr, const int *reach_up_idx,
                         const int *level_ptr, const int *level_reach,
                         int n_reaches, int n_levels, int n_segments,
                         const float *qlat_ts, const float *qdp0_seg,
                         const float *dp0_seg, const float *reach_last0,
                         double *qdc_seg_out_d, double *reach_last_out_d,
                         int n_timesteps, double dt, long long *iter_sum);

/* Same algorithm, every array double.  depth_out / vel_out (may be NULL) get
 * the final-step per-segment depth and velocity.  reach_hist (may be NULL)
 * gets reach_last after every step: reach_hist[t*n_reaches + r]. */
void orc_wavefront_f64(const orc_chparams_f64 *params,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const double *qlat_ts, const double *qdp0_seg,
                       const double *dp0_seg, const double *reach_last0,
                       double *qdc_seg_out, double *reach_last_out,
                       double *depth_out, double *vel_out, double *reach_hist,
                       int n_timesteps, double dt, int mode, long long *iter_sum);

/* orc_wavefront_f64 on n_threads host threads (OpenMP over the reaches of one
 * level); bit for bit the one-thread result. */
void orc_wavefront_f64_mt(const orc_chparams_f64 *params,
                          const int *reach_seg_start, const int *reach_seg_len,
                          const int *reach_up_ptr, const int *reach_up_idx,
                          const int *level_ptr, const int *level_reach,
                          int n_reaches, int n_levels, int n_segments,
                          const double *qlat_ts, const double *qdp0_seg,
                          const double *dp0_seg, const double *reach_last0,
                          double *qdc_seg_out, double *reach_last_out,
                          double *depth_out, double *vel_out, double *reach_hist,
                          int n_timesteps, double dt, int mode, long long *iter_sum, int n_threads);

/* FP32 REFERENCE rule: the reference's float solve (wavefront_mc.cu:66-193)
 * and its float DAG loop cpu_wavefront (:413-491), float in, float out.
 * depth_out / reach_hist may be NULL; n_threads > 1 shares the reaches of a
 * level over threads (bit for bit the one-thread result). */
float orc_mc_solve_f32(float dt, float qup, float quc, float qdp, float ql,
                       const orc_chparams_f32 *p, float depthp,
                       float *depthc, float *velc, int *iters);
void orc_wavefront_f32(const orc_chparams_f32 *params,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const float *qlat_ts, const float *qdp0_seg,
                       const float *dp0_seg, const float *reach_last0,
                       float *qdc_seg_out, float *reach_last_out,
                       float *depth_out, float *reach_hist,
                       int n_timesteps, float dt, long long *iter_sum, int n_threads);

/* cpu_mc: independent reaches, one step, constant upstream inflow uq into
 * segment 0 of every reach. */
void orc_reach_parallel(const orc_chparams_f64 *p, const double *ql,
                        const int *rs, const int *rl,
                        const double *qu, const double *qd, const double *dp,
                        double *qo, double *vo, double *dpo,
                        double uq, double dt, int nr, long long *iter_sum);

/* Converged-mode residual F(h) = Qmc(h) - Qmn(h) (see mc_oracle.c). */
double orc_mc_residual(double dt, double qup, double quc, double qdp, double ql,
                       const orc_chparams_f64 *p, double h, double *qmc_out);

double orc_mc_solve_converged(double dt, double qup, double quc, double qdp, double ql,
                              const orc_chparams_f64 *p, double depthp,
                              double *depthc, double *velc);

/* Enclosure [*lo, *hi] of the converged residual F over depths [a, b] (both in
 * the main channel, or both at or above bankfull on a floodplain); the
 * certified scan of orc_mc_solve_converged is built on it. */
void orc_residual_enclosure(double dt, double qup, double quc, double qdp, double ql,
                            const orc_chparams_f64 *p, double a, double b, double *lo, double *hi);

/* Enclosure [*lo, *hi] of dF/dh over [a, b] (same regimes); returns 0 where
 * none is formed (the celerity may vanish on [a, b]). */
int orc_residual_slope_enclosure(double dt, double qup, double quc, double qdp, double ql,
                                 const orc_chparams_f64 *p, double a, double b, double *lo, double *hi);

/* Enclosures evaluated by certified scans so far (diagnostics). */
long long orc_cert_steps(void);

/* Canonical root (see mc_oracle.c): the root of the canonical cell next to the
 * estimate r, NaN when there is none (then orc_mc_solve_converged keeps its
 * estimate and counts it in orc_noncanon_count). */
double orc_canonical_root(double dt, double qup, double quc, double qdp, double ql,
                          const orc_chparams_f64 *p, double r, double *qout);
/* Fc(h), the residual form whose signs decide the cell, and its outflow. */
double orc_canon_residual(double dt, double qup, double quc, double qdp, double ql,
                          const orc_chparams_f64 *p, double h, double *qout);
/* R^(2/3), R^(5/3) of Fc (within two ulps, the same bits on every machine). */
void orc_pow2353_fast(double x, double *p23, double *p53);
long long orc_noncanon_count(void);
void orc_noncanon_reset(void);

/* orc_wavefront_f32in on n_threads host threads (OpenMP over the reaches of
 * one level); bit for bit the one-thread result. */
void orc_wavefront_f32in_mt(const orc_chparams_f32 *params,
                            const int *reach_seg_start, const int *reach_seg_len,
                            const int *reach_up_ptr, const int *reach_up_idx,
                            const int *level_ptr, const int *level_reach,
                            int n_reaches, int n_levels, int n_segments,
                            const float *qlat_ts, const float *qdp0_seg,
                            const float *dp0_seg, const float *reach_last0,
                            double *qdc_seg_out_d, double *reach_last_out_d,
                            int n_timesteps, double dt, long long *iter_sum, int n_threads);

/* Linearized MC coefficients at half bankfull depth, float (n segments). */
void orc_linear_coeffs(const orc_chparams_f32 *params, int n, float dt,
                       float *C1, float *C2, float *C3, float *C4oq);

/* FP64 linearized MC over T steps from float coefficients, forcing and
 * initial flow; Q_final = flow of every segment after the last step. */
void orc_linear_mc_fp64(const float *C1, const float *C2, const float *C3, const float *C4oq,
                        const int *reach_seg_start, const int *reach_seg_len,
                        const int *reach_up_ptr, const int *reach_up_idx,
                        const int *level_ptr, const int *level_reach,
                        int n_reaches, int n_levels, int n_segments,
                        const float *qlat_ts, const float *Q0, double *Q_final,
                        int n_timesteps);

/* Matrix form of the linearized MC (wavefront/linear_mc_matrix.cu): X = 0.25
 * coefficients; P_Q and P_q built row by row with the reference's merge,
 * prune and insert steps (returns -1 when an upstream index is not below its
 * reach; arrays malloc'd, free with orc_free; nnz[0] = P_Q's, nnz[1] = P_q's);
 * T float steps from Q (in place) over each row's union of entries with G
 * lanes' fixed summation order (G = 0: each row's G from its union length, as
 * the library's default). */
void orc_matrix_coeffs(const orc_chparams_f32 *params, int n, float dt,
                       float *C1, float *C2, float *C3, float *C4oq);
int orc_matrix_build(int n, const int *up_ptr, const int *up_idx,
                     const float *C1, const float *C2, const float *C3, const float *C4oq, float tol,
                     int **rpQ, int **colQ, float **valQ, int **rpq, int **colq, float **valq, int *nnz);
void orc_free(void *p);
void orc_matrix_route(int n, const int *rpQ, const int *colQ, const float *valQ,
                      const int *rpq, const int *colq, const float *valq,
                      const float *qlat, float *Q, int T, int G);

/* Diffusive-wave step: reach j has rp[j] = {ncomp, theta, dtini} and node i
 * at j * ORC_DIFF_MAX_NODES + i of dx, diffusivity, celerity (unused by the
 * step), qp_prev and qp_out; 1 <= ncomp <= ORC_DIFF_MAX_NODES. */
#define ORC_DIFF_MAX_NODES 500
typedef struct { int ncomp; float theta, dtini; } orc_diff_reach;
void orc_diffusive_tridiag(const orc_diff_reach *rp, const float *dx, const float *diffusivity,
                           const float *celerity, const float *qp_prev, float *qp_out, int nreach);

/* Manning normal depth: h >= 0 with Qmn(h) == Q (Q > 0). */
double orc_normal_depth(double Q, const orc_chparams_f64 *p);
double orc_manning_flow(double h, const orc_chparams_f64 *p);

/* The pow the build uses: libm pow (liborc.so) or correctly rounded pow
 * (liborc_cr.so, built with -DORC_CR_POW). */
double orc_pow_fn(double x, double y);
/* The powf of the float rule: libm powf (liborc.so) or correctly rounded
 * powf (liborc_cr.so). */
float orc_powf_fn(float x, float y);

/* Error statistics of wavefront_mc.cu:822-858: relative error over entries
 * with |truth| > 1e-3.  out[0..8] = max_abs, p50, p90, p99, p99.9, max_rel,
 * frac within 1 %, frac within 10 %, nan count. */
void orc_compute_stats(const double *pred, const double *truth, int n, double *out);

#define ORC_MODE_REFERENCE 0
#define ORC_MODE_CONVERGED 1

#ifdef __cplusplus
}
#endif
#endif
