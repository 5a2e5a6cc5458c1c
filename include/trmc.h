/*
 * trmc -- Muskingum-Cunge river routing on AMD Instinct MI355X (gfx950).
 *
 * C-ABI of the routing library (noaa-gpu-kernels_amd/libtrmc.so).  Plain
 * pointers and sizes only; every array is caller-owned host memory unless the
 * name says _dev.  Every entry point returns an int status (TRMC_OK == 0,
 * negative on error) and never exits the process; trmc_last_error() gives the
 * message of the calling thread's last failure.
 *
 * Reference interfaces replaced (consigcody94/noaa-gpu-kernels):
 *   trmc_wavefront_f32in_f64  <- cpu_wavefront_fp64,
 *                                water-prediction/t-route/wavefront/wavefront_mc.cu:495-507
 *                                (and the GPU time/level loop of :737-762 with
 *                                kernel_level_wavefront :342-408)
 *   trmc_wavefront_f64        <- same call, every array double (FP64 end to end)
 *   trmc_wavefront_f32        <- cpu_wavefront, wavefront_mc.cu:413-427 (the FP32 rule:
 *                                mc_solve :66-193 in float, float in and out)
 *   trmc_reach_parallel_f64   <- cpu_mc, water-prediction/t-route/troute_mc_final.cu:167-171
 *                                and kernel_mc_gpu :142-148
 *   trmc_mc_solve_f64         <- mc_solve_d, wavefront_mc.cu:215-220 (batched)
 *   trmc_load_topo / trmc_load_params / trmc_load_forcings
 *                             <- load_topo wavefront_mc.cu:599-632 and the
 *                                params/forcings reads :654-680 (format of
 *                                wavefront/gen_network.py:10-37)
 *   trmc_compute_stats        <- compute_stats lambda, wavefront_mc.cu:822-858
 * Entry points with no reference counterpart are marked [new].
 */
#ifndef TRMC_H
#define TRMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRMC_OK                 0
#define TRMC_ERR_ARG           -1  /* bad argument (null pointer, negative size) */
#define TRMC_ERR_TOPOLOGY      -2  /* level/CSR structure is not a valid DAG schedule */
#define TRMC_ERR_HIP           -3  /* HIP runtime call failed */
#define TRMC_ERR_NO_DEVICE     -4  /* no usable gfx950 device */
#define TRMC_ERR_TIMEOUT       -5  /* persistent-kernel barrier wait exceeded its limit */
#define TRMC_ERR_NOMEM         -6  /* host or device allocation failed */
#define TRMC_ERR_IO            -7  /* file missing, short or inconsistent */
#define TRMC_ERR_STATE         -8  /* call order violated (e.g. run before params) */
#define TRMC_ERR_NOT_CONVERGED -9  /* converged/learned solve hit its iteration cap */

/* Per-segment channel parameters in the reference's field order:
 * FP32 as wavefront_mc.cu:41-44, FP64 as troute_mc_final.cu:26-28. */
typedef struct { float  dx, bw, tw, twcc, n_ch, ncc, cs, s0; } trmc_chparams_f32;
typedef struct { double dx, bw, tw, twcc, n_ch, ncc, cs, s0; } trmc_chparams_f64;

/* Secant stopping rule.
 *  REFERENCE: the reference's rule, 1 % relative or 1 cm absolute depth change,
 *             up to 5 widening attempts (wavefront_mc.cu:242-318).  Bit-for-bit
 *             drop-in semantics.
 *  CONVERGED: [new] the fixed point of the reference iteration (X0 = 0.5),
 *             F(h) = Qmc(h) - Qmn(h), solved from the reference's initial
 *             bracket (0.67 d, 1.33 d + 0.01).  The depth is the leftmost
 *             downcrossing of F on [1 cm, inf) (1 cm when F is never
 *             positive), ended in its canonical cell: the neighbouring depths
 *             g < g+ whose significands have their low 28 bits zero with
 *             F(g) > 0 >= F(g+), and the secant point of that cell (cut at
 *             bankfull when the floodplain's jump lies inside).  F has the same
 *             bits on every machine (IEEE operations only), so any start that
 *             converges returns the same depth and outflow bits, the CPU
 *             oracle's included.
 *  LEARNED:   [new] as CONVERGED, same root and bits, but started from F at the
 *             previous depth d and a guess along the on-device MLP's rating
 *             curve from Qmn(d) to the routed flow Qmc(d) (trmc_mlp_*); a
 *             segment not converged within learned_budget evaluations
 *             restarts from the reference bracket.
 *  LINEAR:    the linearized MC of wavefront/linear_mc.cu:130-182 with fixed
 *             per-segment coefficients (trmc_network_set_linear_coeffs):
 *             Q = C1 Qup_new + C2 Qup_old + C3 Q_old + C4/ql ql, no secant;
 *             DATAFLOW schedule only.  Depth and velocity are not updated.
 *  REFERENCE_F32: the reference's FP32 solve (wavefront_mc.cu:66-193): the
 *             REFERENCE rule in float arithmetic, parameters rounded to float,
 *             float upstream sums (:444-451), every powf correctly rounded;
 *             flows and depths are float values (held exactly in the double
 *             state).  Interleaved DATAFLOW schedule only. */
#define TRMC_MODE_REFERENCE 0
#define TRMC_MODE_CONVERGED 1
#define TRMC_MODE_LEARNED   2
#define TRMC_MODE_LINEAR    3
#define TRMC_MODE_REFERENCE_F32 4

/* Schedules of the level-ordered DAG sweep.
 *  AUTO:    pick from the autotuner's table (or SKEWED).
 *  SKEWED:  one persistent kernel over super-levels k = level(unit) + t; a
 *           device-scope barrier between super-levels (L + T - 1 barriers).
 *  LEVEL:   one launch per (timestep, level), as the reference's
 *           kernel_level_wavefront loop; the per-step launch sequence is
 *           captured once in a hipGraph and replayed T times.
 *  DATAFLOW: one persistent kernel without grid barriers: each (unit, step)
 *           starts as soon as its upstream outflows for steps t and t-1 are
 *           in device memory (a full outflow history, (T+1) x units x
 *           members doubles).  Confluences of at most 4 upstream reaches.
 * AUTO picks DATAFLOW when the network allows it, else SKEWED. */
#define TRMC_SCHED_AUTO     0
#define TRMC_SCHED_SKEWED   1
#define TRMC_SCHED_LEVEL    2
#define TRMC_SCHED_DATAFLOW 3

typedef struct {
    int    mode;            /* TRMC_MODE_* (default REFERENCE) */
    int    schedule;        /* TRMC_SCHED_* (default AUTO) */
    int    unit_segments;   /* max segments per work unit; 0 = auto (REFERENCE 4,
                               CONVERGED and LEARNED 2, LINEAR 8) */
    int    block_size;      /* threads per workgroup; 0 = auto (64) */
    int    n_members;       /* ensemble members sharing params (>= 1) */
    int    record_history;  /* keep every step's last-segment outflow per reach */
    double tol;             /* CONVERGED/LEARNED: relative tolerance of the root estimate the
                               canonical cell is found from (default 1e-10; the cell fixes the
                               result, the tolerance only how soon the search hands over) */
    int    max_iter;        /* CONVERGED/LEARNED iteration cap (default 60) */
    int    learned_budget;  /* LEARNED: evaluations before falling back (default 12) */
    int    grid_blocks;     /* persistent grid size; 0 = fill the device */
    int    dataflow_variant;/* DATAFLOW: 0 = secant iterations interleaved across lanes,
                               1 = one whole item per lane pass */
    int    device;          /* HIP device ordinal */
    int    trace;           /* SKEWED: record per-super-level time and max iterations */
    int    batch_iters;     /* DATAFLOW interleaved: secant iterations between scheduling
                               phases (0 = default 8) */
    int    guess_model;     /* LEARNED: 0 = the network's rating curve (default),
                               1 = the wide-channel power law h ~ Q^(3/5) alone
                               (ablation: the same anchored guess without the
                               network) */
} trmc_options;

typedef struct {
    double kernel_ms;        /* device time of the routing launches (HIP events) */
    double wall_ms;          /* host wall time of trmc_network_run */
    long long seg_steps;     /* segments x members x timesteps routed */
    long long secant_iters;  /* total secant iterations over all solves */
    long long fallbacks;     /* LEARNED: solves that restarted from the bracket */
    long long not_converged; /* CONVERGED/LEARNED: solves that hit max_iter */
    int    launches;         /* kernel launches (or graph replays) issued */
    int    barriers;         /* device-scope barriers crossed (SKEWED) */
    int    super_levels;     /* L + T - 1 for SKEWED */
    int    grid_blocks;      /* persistent grid actually used */
    int    block_size;
    double barrier_wait_frac;/* fraction of block-time spent waiting in barriers */
    long long mlp_calls;     /* LEARNED: wave-wide network forwards (64 samples each) */
    double mlp_frac;         /* LEARNED: fraction of wave-time spent in them */
    long long certified_solves; /* CONVERGED/LEARNED: solves with F(mindepth) <= 0 done by the
                                   certified scan (on the server waves of the DATAFLOW kernel) */
    long long noncanonical;  /* CONVERGED/LEARNED: solves whose root estimate was kept because
                                no canonical cell lies next to it (0 in every measured run) */
} trmc_run_stats;

/* ---------------------------------------------------------------- basics */
const char *trmc_last_error(void);
int  trmc_version(void);                       /* major*10000 + minor*100 + patch */
int  trmc_device_count(int *count);            /* gfx950 devices visible */
void trmc_default_options(trmc_options *opt);
/* Waits for all work the library queued on `device` (hipDeviceSynchronize). */
int  trmc_device_synchronize(int device);

/* ------------------------------------------------------ drop-in routines */

/* cpu_wavefront_fp64 (wavefront_mc.cu:495-507), same argument list: float
 * params/forcings widened to double, FP64 routing on the GPU, final-step
 * per-segment qdc and per-reach last-segment qdc as double.  qlat_ts is
 * [n_timesteps][n_segments].  Uses the REFERENCE stopping rule and the
 * default schedule. */
int trmc_wavefront_f32in_f64(const trmc_chparams_f32 *params,
                             const int *reach_seg_start, const int *reach_seg_len,
                             const int *reach_up_ptr, const int *reach_up_idx,
                             const int *level_ptr, const int *level_reach,
                             int n_reaches, int n_levels, int n_segments,
                             const float *qlat_ts, const float *qdp0_seg,
                             const float *dp0_seg, const float *reach_last0,
                             double *qdc_seg_out_d, double *reach_last_out_d,
                             int n_timesteps, double dt);

/* Same call with every array double (FP64 end to end).  opt may be NULL. */
/* cpu_wavefront (wavefront_mc.cu:413-427), same argument list: the FP32
 * rule end to end (TRMC_MODE_REFERENCE_F32 on the DATAFLOW schedule), final-
 * step per-segment qdc and per-reach last-segment qdc as float.  Results are
 * the reference routine's with every powf correctly rounded (the reference's
 * host libm powf is not, in about 1 of 1,600 calls). */
int trmc_wavefront_f32(const trmc_chparams_f32 *params,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const float *qlat_ts, const float *qdp0_seg,
                       const float *dp0_seg, const float *reach_last0,
                       float *qdc_seg_out, float *reach_last_out,
                       int n_timesteps, float dt);

int trmc_wavefront_f64(const trmc_chparams_f64 *params,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const double *qlat_ts, const double *qdp0_seg,
                       const double *dp0_seg, const double *reach_last0,
                       double *qdc_seg_out, double *reach_last_out,
                       int n_timesteps, double dt, const trmc_options *opt);

/* cpu_mc / kernel_mc_gpu (troute_mc_final.cu:167-171, :142-148): independent
 * reaches, one step, upstream inflow uq into segment 0 of every reach. */
int trmc_reach_parallel_f64(const trmc_chparams_f64 *p, const double *ql,
                            const int *rs, const int *rl,
                            const double *qu, const double *qd, const double *dp,
                            double *qo, double *vo, double *dpo,
                            double uq, double dt, int nr);

/* cpu_linear_mc_fp64 (wavefront/linear_mc.cu:238-289), called from main at
 * :561-567: the linearized MC over n_timesteps with float coefficients C1, C2,
 * C3, C4/ql per segment, float lateral inflow [T][N] and initial flow Q0 [N];
 * Q_final [N] = every segment's flow after the last step, FP64 on the GPU.
 * Same argument list; returns a status.  Every reach needs at least one
 * segment (the reference writes a reach's first segment unconditionally). */
int trmc_linear_mc_f64(const float *C1, const float *C2, const float *C3, const float *C4oq,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const float *qlat_ts, const float *Q0, double *Q_final,
                       int n_timesteps);

/* compute_mc_coeffs / kernel_precompute_coeffs (wavefront/linear_mc.cu:57-113):
 * the linearized MC coefficients at half bankfull depth, float, on the GPU
 * (each power is the float rounding of the double pow). */
int trmc_linear_coeffs_f32(const trmc_chparams_f32 *params, int n, float dt,
                           float *C1, float *C2, float *C3, float *C4oq);

/* Matrix form of the linearized MC (wavefront/linear_mc_matrix.cu,
 * linear_mc_matrix_graph.cu), one segment per reach, reaches numbered so that
 * every upstream reach precedes its downstream one.  Each step is
 *   Q_new = P_Q Q_old + P_q qlat_t
 * with P_Q = (I - C1 N)^-1 (C2 N + C3 I) and P_q = (I - C1 N)^-1 diag(C4/ql),
 * N the upstream adjacency; entries below a tolerance are dropped. */
typedef struct {
    int    n, nnz;
    int   *rowptr;   /* [n + 1] */
    int   *col;      /* [nnz], ascending within a row */
    float *val;      /* [nnz] */
} trmc_csr;

/* compute_mc_coeffs_cpu (linear_mc_matrix.cu:58-84): the linearized MC
 * coefficients at half bankfull depth with the weighting X = 0.25, float, on
 * the GPU (each power is the float rounding of the double pow). */
int trmc_matrix_coeffs_f32(const trmc_chparams_f32 *params, int n, float dt,
                           float *C1, float *C2, float *C3, float *C4oq);

/* The secant-fitted coefficients of wavefront/linear_mc_matrix_secant.cu
 * (mc_secant_solve_cpu :57-195, called from main :358-418): per reach, the
 * float MC secant on depth at a steady reference flow (upstream reference
 * flows + own first-step lateral inflow + the mean first-step inflow over all
 * segments, in level order), its last iterate's C1, C2, C3 and C4/ql; the
 * half-bankfull X = 0.25 coefficients for invalid channels.  One segment per
 * reach: params [n_reaches], qlat0 = the first step's lateral inflow
 * [n_segments >= n_reaches].  Host threads; float operations in the
 * reference's order. */
int trmc_matrix_secant_coeffs_f32(const trmc_chparams_f32 *params, int n_reaches,
                                  const int *reach_up_ptr, const int *reach_up_idx,
                                  const int *level_ptr, const int *level_reach, int n_levels,
                                  const float *qlat0, int n_segments, float dt,
                                  float *C1, float *C2, float *C3, float *C4oq);

/* build_propagation_matrices (linear_mc_matrix.cu:118-229): P_Q and P_q row
 * by row in reach order, each row C1 x the sum of its upstream rows (pruned
 * below tol), plus C2 at each upstream column and C3 (P_Q) or C4/ql (P_q) on
 * the diagonal, pruned again; the reference's float operations in its order,
 * on the host.  An upstream index not below its reach is TRMC_ERR_TOPOLOGY.
 * Fills *P_Q and *P_q (free with trmc_free_csr). */
int trmc_matrix_build(int n_reaches, const int *reach_up_ptr, const int *reach_up_idx,
                      const float *C1, const float *C2, const float *C3, const float *C4oq,
                      float tol, trmc_csr *P_Q, trmc_csr *P_q);
void trmc_free_csr(trmc_csr *m);

/* The timestep loop of linear_mc_matrix.cu:340-380 and
 * linear_mc_matrix_graph.cu:355-392: n_timesteps steps from Q0 [n] with
 * lateral inflow qlat [T][n], both products and the sum fused in one FP32
 * kernel per step, the steps replayed as one HIP graph; Q_final [n] = the flow
 * after the last step.  Columns must ascend within each row.  Summation order
 * (fixed, so results do not depend on scheduling): each row is the union of
 * its two CSR rows by column (a matrix without an entry there contributes
 * 0 * value); G lanes share a row, lane l taking union entries l, l + G, ...,
 * each as s = s + vQ * Q[c] then s = s + vq * ql[c]; the lanes' sums are
 * added pairwise by xor distance G/2, ..., 1.  lanes_per_row = G for every
 * row (1, 2, 4, ..., 64), or 0: each row's G from its union length U, the
 * smallest power of two with 2 G >= U, at most 64. */
int trmc_matrix_route_f32(const trmc_csr *P_Q, const trmc_csr *P_q, const float *qlat,
                          const float *Q0, int n_timesteps, int lanes_per_row, float *Q_final);
/* [new] The same with reps timed replays of the whole graph after one
 * untimed one, inputs resident; *avg_pass_ms = mean HIP-event time of one
 * replay (all n_timesteps steps). */
int trmc_matrix_route_timed_f32(const trmc_csr *P_Q, const trmc_csr *P_q, const float *qlat,
                                const float *Q0, int n_timesteps, int lanes_per_row, float *Q_final,
                                int reps, double *avg_pass_ms);

/* Diffusive-wave step (owp_extended_kernels.cu:23-135): one implicit
 * (theta-weighted) diffusion step per reach, solved by forward elimination and
 * back substitution, FP32.  Node i of reach j is element
 * j * TRMC_DIFF_MAX_NODES + i of every node array. */
#define TRMC_DIFF_MAX_NODES 500
typedef struct { int ncomp; float theta; float dtini; } trmc_diff_reach;   /* DiffReachParams, :28-32 */

/* cpu_diffusive_tridiag / kernel_diffusive_tridiag (owp_extended_kernels.cu:35-94,
 * :97-158), same argument list; 1 <= ncomp <= TRMC_DIFF_MAX_NODES per reach.
 * celerity is accepted and unused (the reference computes a Courant number
 * from it that it never uses).  Entries beyond a reach's ncomp are untouched. */
int trmc_diffusive_tridiag_f32(const trmc_diff_reach *rp, const float *dx, const float *diffusivity,
                               const float *celerity, const float *qp_prev, float *qp_out, int nreach);

/* [new] The same step with the inputs resident: one launch, then reps timed
 * launches (HIP events); *avg_kernel_ms = their average. */
int trmc_diffusive_tridiag_timed_f32(const trmc_diff_reach *rp, const float *dx, const float *diffusivity,
                                     const float *celerity, const float *qp_prev, float *qp_out, int nreach,
                                     int reps, double *avg_kernel_ms);

/* [new] No counterpart in the reference: for callers that cannot allocate
 * device memory (serving loops, harnesses that forbid hipMalloc).  The step of
 * trmc_diffusive_tridiag_f32 with device-resident inputs and outputs (device
 * pointers, same layout; stream = a hipStream_t, NULL for the default stream).
 * Its scratch (the elimination factors, a count of invalid reach records)
 * lives in a static device array of 32.8 MB per GPU, held until the process
 * ends, so the call allocates nothing.  At most TRMC_DIFF_NOALLOC_MAX_REACHES
 * reaches; more is TRMC_ERR_ARG before any device work.  Synchronous: returns
 * after the step; a reach record with ncomp outside 1..TRMC_DIFF_MAX_NODES
 * leaves that reach's nodes untouched and the call returns TRMC_ERR_ARG.
 * Calls that overlap on one GPU share its scratch and race. */
#define TRMC_DIFF_NOALLOC_MAX_REACHES 16384
int trmc_diffusive_tridiag_f32_noalloc(const trmc_diff_reach *d_rp, const float *d_dx, const float *d_diffusivity,
                                       const float *d_celerity, const float *d_qp_prev, float *d_qp_out,
                                       int nreach, void *stream);

/* Forcing ingest (data_parsers/parse_qlat_fast_cpp.cpp:76-224): lateral inflow
 * from a directory of per-nexus CSVs "nex-<id>_output.csv" (rows "timestep,
 * time, value"), pivoted to qlat [n_timesteps][n_reaches] float with reach r
 * taking the file of nexus topo_to_wbid[r]; host threads (n_threads <= 0: all
 * cores).  Unnamed reaches and timesteps stay 0; a reach named by two files
 * takes the later one in name order.  counts (may be NULL) = {files read,
 * files skipped: no number, unknown id or unreadable}. */
int trmc_ingest_nex_csv(const char *csv_dir, const int64_t *topo_to_wbid, int n_reaches,
                        int n_timesteps, int n_threads, float *qlat, int64_t *counts);

/* The program's whole path: topo_to_wbid from a .npy (int64) file, then
 * forcings.bin = int32 T, float qlat [T][R], qup0 [R] = 0, qdp0 [R] = 0,
 * dp0 [R] = 0.5. */
int trmc_ingest_nex_csv_to_forcings(const char *csv_dir, const char *wbid_npy_path,
                                    const char *out_path, int n_timesteps, int n_threads,
                                    int64_t *counts);

/* The same pivot under the parsing rules of the reference's Rust twin
 * (rust_preprocessor/src/main.rs:28-91, 116-273) or of the C++ program:
 *  TRMC_INGEST_CPP   parse_qlat_fast_cpp.cpp (atoi / atof, the value rounded
 *                    from double to float);
 *  TRMC_INGEST_RUST  names "nex-<i64>_output.csv"; a row's timestep is its
 *                    leading digits, its value the text after the last comma
 *                    in Rust's float grammar, correctly rounded to float; a
 *                    row whose value does not read is skipped.
 * A reach named by two files takes the later one in name order (the Rust
 * program's order is the directory's). */
#define TRMC_INGEST_CPP  0
#define TRMC_INGEST_RUST 1
int trmc_ingest_nex_csv_rules(const char *csv_dir, const int64_t *topo_to_wbid, int n_reaches,
                              int n_timesteps, int n_threads, int rules, float *qlat, int64_t *counts);
int trmc_ingest_nex_csv_to_forcings_rules(const char *csv_dir, const char *wbid_npy_path,
                                          const char *out_path, int n_timesteps, int n_threads,
                                          int rules, int64_t *counts);

/* mc_solve_d (wavefront_mc.cu:215-220) over n independent segment inputs on
 * the GPU.  in = [n][5] {qup, quc, qdp, ql, depthp}; out = [n][3]
 * {qdc, depth, velocity}; iters (may be NULL) = [n] secant iterations.
 * mode selects the stopping rule (REFERENCE or CONVERGED). */
int trmc_mc_solve_f64(const trmc_chparams_f64 *params, const double *in,
                      double *out, int *iters, int n, double dt, int mode);

/* ------------------------------------------------- host-only schedule plan */
/* [new] The schedule the library prepares for a topology, without a device:
 * reaches cut into units of <= unit_segments segments, ALAP unit levels,
 * level-sorted device order.  For inspection and tests; trmc_network_create
 * builds the same plan internally. */
typedef struct trmc_plan trmc_plan;
int  trmc_plan_build(const int *reach_seg_start, const int *reach_seg_len,
                     const int *reach_up_ptr, const int *reach_up_idx,
                     const int *level_ptr, const int *level_reach,
                     int n_reaches, int n_levels, int n_segments, int unit_segments,
                     trmc_plan **out);
void trmc_plan_destroy(trmc_plan *plan);
/* info[0..7] = n_units, n_unit_levels, ring depth, super-levels for
 * n_timesteps, widest super-level in items (x n_members), device segments,
 * max level span of an edge, upstream links. */
int  trmc_plan_info(const trmc_plan *plan, int n_timesteps, int n_members,
                    int64_t *info, int n_info);
/* Copies plan array `which` into out[n] and returns its length:
 * 0 level_ptr, 1 unit_level, 2 unit_seg, 3 up_ptr, 4 up_idx, 5 unit_flags,
 * 6 unit_reach, 7 unit_init, 8 seg_of_dev, 9 reach_last_unit,
 * 10 unit_rec (8 ints per unit: level, first and end segment,
 *    flags | n_up << 8 with flags bit 0 internal boundary and bit 1 no slack
 *    (ASAP level == ALAP level), up to 4 upstream units; empty when a unit
 *    has more than 4 upstream units). */
int  trmc_plan_export(const trmc_plan *plan, int which, int *out, int n);

/* ----------------------------------------------- resident network handle */
/* [new] Owns device buffers across calls: topology, parameters, forcing and
 * state stay in HBM; trmc_network_run routes with inputs already resident. */
typedef struct trmc_network trmc_network;

/* Topology in the reference's CSR/level layout (wavefront_mc.cu:577-632). */
int trmc_network_create(const int *reach_seg_start, const int *reach_seg_len,
                        const int *reach_up_ptr, const int *reach_up_idx,
                        const int *level_ptr, const int *level_reach,
                        int n_reaches, int n_levels, int n_segments,
                        const trmc_options *opt, trmc_network **out);
void trmc_network_destroy(trmc_network *net);

int trmc_network_set_params_f64(trmc_network *net, const trmc_chparams_f64 *params);
int trmc_network_set_params_f32(trmc_network *net, const trmc_chparams_f32 *params);

/* Initial state per member: qdp0/dp0 are [n_members][n_segments], reach_last0
 * is [n_members][n_reaches] (NULL -> qdp0 of each reach's last segment,
 * as wavefront_mc.cu:676-680). */
int trmc_network_set_state_f64(trmc_network *net, const double *qdp0_seg,
                               const double *dp0_seg, const double *reach_last0);

/* Lateral inflow [n_members][n_timesteps][n_segments] (double or float). */
int trmc_network_set_forcing_f64(trmc_network *net, const double *qlat, int n_timesteps);
int trmc_network_set_forcing_f32(trmc_network *net, const float *qlat, int n_timesteps);

/* [new] Synthetic lateral inflow generated on the device, in place of
 * set_forcing: exactly the values trmc_generate_forcing(n_segments,
 * n_timesteps, n_members, seed) returns, without the host array or the copy
 * (a 1M-segment x 16-member x 288-step forcing is 37 GB). */
int trmc_network_generate_forcing(trmc_network *net, int n_timesteps, uint64_t seed);

/* [new] Weights of the LEARNED depth guess (TRMC_MLP_NPARAM floats, layout of
 * trmc_mlp_train), copied to the device; required before a LEARNED run. */
int trmc_network_set_mlp(trmc_network *net, const float *weights);

/* [new] Coefficients of the LINEAR mode, one float per segment each, in the
 * caller's segment order (as trmc_linear_coeffs_f32 returns them); required
 * before a LINEAR run. */
int trmc_network_set_linear_coeffs(trmc_network *net, const float *C1, const float *C2,
                                   const float *C3, const float *C4oq);

/* Route n_timesteps steps (<= the forcing's) from the current state; the final
 * state becomes the new current state (calls chain: run(T1) then run(T2) ==
 * run(T1+T2) bit for bit).  stats may be NULL. */
int trmc_network_run(trmc_network *net, int n_timesteps, double dt, trmc_run_stats *stats);

/* Restore the state given to the last trmc_network_set_state_f64 (kept in
 * device memory; device-to-device copies only) and rewind the forcing to its
 * first step.  Queued on the network's stream. */
int trmc_network_reset(trmc_network *net);

/* Current state per member in the caller's (reference) segment order.
 * Any output pointer may be NULL.  qdc/depth/vel are [n_members][n_segments],
 * reach_last [n_members][n_reaches]. */
int trmc_network_get_state_f64(const trmc_network *net, double *qdc_seg,
                               double *depth_seg, double *vel_seg, double *reach_last);

/* With record_history: last-segment outflow of every reach after every step
 * of the last run, [n_members][n_timesteps][n_reaches]. */
int trmc_network_get_history_f64(const trmc_network *net, double *reach_hist);

/* With trace: for each super-level of the last SKEWED run, the wall time from
 * the previous barrier's release to this one's (microseconds, block 0's view)
 * and the largest secant iteration count of any item in it.  Returns the
 * number of super-levels written (at most n). */
int trmc_network_get_trace(const trmc_network *net, double *us, int *max_iters, int n);

/* Shape facts of the prepared schedule, as trmc_plan_info with the
 * network's member count. */
int trmc_network_info(const trmc_network *net, int n_timesteps, int64_t *info, int n_info);

/* [new] Streaming device copy of `bytes` (read + write), reps times after
 * one warm-up; *gbps = 2 bytes reps / time.  The reference for achievable
 * HBM bandwidth and for scaling memory-traffic counters. */
int trmc_stream_copy(int device, size_t bytes, int reps, double *gbps);

/* [new] Lateral inflow read from host memory while the routing runs (the
 * reference copies each step's inflow to the device inside its time loop,
 * wavefront_mc.cu:739, before that step's kernels).  The forcing stays in
 * page-locked host memory (trmc_host_alloc), mapped into the device's address
 * space, and the routing kernels read each segment-step's inflow from there
 * over the host link as they route it: the transfer overlaps the routing,
 * with no copy and no second stream.  qlat_dev_order holds the steps in the
 * network's device order, [n_timesteps][n_dev_segments][n_members] doubles,
 * as trmc_network_order_forcing_f64 writes them from qlat
 * [n_members][n_timesteps][n_segments] (the layout of
 * trmc_network_set_forcing_f64).  The network reads the buffer in place: it
 * must stay allocated and unchanged until another forcing is set or the
 * network is destroyed.  Results are the same bits as with the forcing in
 * HBM.  The REFERENCE, LINEAR and REFERENCE_F32 rules only: a CONVERGED or
 * LEARNED run with forcing in host memory returns TRMC_ERR_ARG. */
int trmc_network_order_forcing_f64(const trmc_network *net, const double *qlat, int n_timesteps, double *out);
int trmc_network_set_forcing_mapped(trmc_network *net, const double *qlat_dev_order, int n_timesteps);
int trmc_host_alloc(size_t bytes, void **ptr);
void trmc_host_free(void *ptr);

/* [new] Autotuner for the schedule of a network (or a union of networks):
 * times candidate schedules on the first n_probe steps of the forcing --
 * qlat [n_members][n_probe][n_segments], or generated on the device from
 * forcing_seed when qlat is NULL -- from the given state, and returns the
 * options with the fastest kernel time (base's other fields kept) and that
 * time per step.  Candidates, each at unit sizes 2, 4 and 8 segments:
 *   tile      DATAFLOW, interleaved (dataflow_variant 0), solve batch 6, 8,
 *             12, workgroup 64 or 128;
 *   whole     DATAFLOW, one whole item per lane pass (variant 1), 64 or 128;
 *   wavefront SKEWED (super-level wavefront between device-scope barriers),
 *             workgroup 64 or 256.
 * A base->schedule other than TRMC_SCHED_AUTO restricts the search to that
 * schedule; LINEAR and REFERENCE_F32 search the tile schedule only.
 * qdp0/dp0 are [n_members][n_segments] (base->n_members); mlp_weights is
 * needed for LEARNED only. */
int trmc_autotune(const int *reach_seg_start, const int *reach_seg_len, const int *reach_up_ptr,
                  const int *reach_up_idx, const int *level_ptr, const int *level_reach, int n_reaches,
                  int n_levels, int n_segments, const trmc_chparams_f64 *params, const double *qdp0_seg,
                  const double *dp0_seg, const double *reach_last0, const double *qlat, uint64_t forcing_seed,
                  int n_probe, double dt, const float *mlp_weights, const trmc_options *base, trmc_options *best,
                  double *best_ms_per_step);

/* ------------------------------------------------------ learned guess */
/* [new] 5->32->32->1 tanh MLP proposing the initial secant depth from
 * (log Q, Manning n, bottom width, side slope, log S0).  FP32 weights;
 * trained on the GPU by the library's own forward/backward/Adam kernels on
 * depth pairs from the FP64 Manning normal-depth solver; evaluated with MFMA
 * inside the routing kernel.  Output y = log(h_normal / h_wide) with
 * h_wide = (Q n / (bw sqrt S0))^(3/5). */
#define TRMC_MLP_NPARAM 1281

/* Generate n training pairs on the device (kept for trmc_mlp_train):
 * normalised features [n][5], target y [n], and optionally the raw draws
 * [n][9] = Q, dx, bw, tw, twcc, n_ch, ncc, cs, s0 (host arrays, each may be
 * NULL). */
int trmc_mlp_make_dataset(int n, uint64_t seed, float *features, float *target, double *raw);

/* Glorot-uniform initial weights (zero biases) from a seed. */
void trmc_mlp_init_weights(uint64_t seed, float *weights);

/* Adam training over the last generated set: weights in/out, batch a
 * multiple of 64, learning rate lr (decayed x0.3 after 60 % and 85 % of the
 * epochs), loss_curve[epochs] (may be NULL) = mean 0.5 (y - target)^2. */
int trmc_mlp_train(float *weights, int epochs, int batch, float lr, uint64_t seed, float *loss_curve);

/* Forward and backward on the GPU for n samples (n multiple of 64):
 * pred [n], grad [TRMC_MLP_NPARAM] = d(mean 0.5 (pred - target)^2)/dW. */
int trmc_mlp_forward_backward(const float *weights, const float *features, const float *target, int n,
                              float *pred, float *grad);

/* Inference for n samples with the routing kernel's MFMA wave forward
 * (use_mfma = 1), the scalar forward (0), or the routing kernel's packed form
 * (2: sample pairs (2i, 2i+1) of the lanes i with i % 3 != 1, up to 32 pairs
 * per wave forward; out is NaN for the other lanes' pairs; n even). */
int trmc_mlp_predict(const float *weights, const float *features, int n, float *out, int use_mfma);

/* ------------------------------------------------- formats and generator */
/* Binary formats of wavefront/gen_network.py:10-37 (little-endian int32 /
 * float32).  Arrays are malloc'd by the library; free with trmc_free. */
typedef struct {
    int n_reaches, n_segments, n_levels, max_upstream;
    int *reach_seg_start, *reach_seg_len, *reach_level, *reach_n_up;
    int *reach_up_ptr, *reach_up_idx, *level_ptr, *level_reach;
} trmc_topo;

int  trmc_load_topo(const char *path, trmc_topo *out);
/* [new] Disjoint union of n networks for routing them in one launch:
 * network i's segments and reaches are numbered after network i-1's, level L
 * of the union lists level L of every network.  seg_offset / reach_offset
 * [n + 1] (may be NULL) receive where each network starts.  Free with
 * trmc_free_topo. */
int  trmc_topo_concat(const trmc_topo *nets, int n, trmc_topo *out, int *seg_offset, int *reach_offset);
int  trmc_save_topo(const char *path, const trmc_topo *t);
void trmc_free_topo(trmc_topo *t);
int  trmc_load_params(const char *path, int n_segments, trmc_chparams_f32 *out);
int  trmc_save_params(const char *path, int n_segments, const trmc_chparams_f32 *p);
/* forcings.bin: int32 T; float32 qlat[T*N]; qup0[R]; qdp0[N]; dp0[N].
 * Pass NULL outputs to query *n_timesteps first. */
int  trmc_load_forcings(const char *path, int n_reaches, int n_segments, int *n_timesteps,
                        float *qlat, float *qup0, float *qdp0, float *dp0);
int  trmc_save_forcings(const char *path, int n_reaches, int n_segments, int n_timesteps,
                        const float *qlat, const float *qup0, const float *qdp0,
                        const float *dp0);
void trmc_free(void *p);

/* Hydrofabric network builder (data_parsers/parse_gpkg.py:43-270): the
 * flowpaths table's rows whose id parses, wb_id[n_rows] and to_id[n_rows]
 * (waterbody "wb-<k>" draining through nexus "nex-<m>", which feeds waterbody m;
 * to_id < 0 for a NULL, unparsable or terminal toid), and the
 * flowpath-attributes rows whose link parses, link[n_attr] with vals[n_attr][8]
 * = {Length_m, n, nCC, BtmWdth, TopWdth, TopWdthCC, ChSlp, So} (NaN = NULL).
 * The first row of a repeated id is kept; reaches are ordered by Kahn's sort
 * with a stack of ready reaches; one segment per reach; params start at
 * {1000, 20, 25, 60, 0.06, 0.12, 2, 0.001} and take each attribute row in
 * table order under the reference's clamps.  Fills *topo (free with
 * trmc_free_topo) and returns *params [n] and *topo_to_wbid [n] allocated by
 * the library (free with trmc_free).  A cycle is TRMC_ERR_TOPOLOGY. */
int trmc_build_flowpath_network(const int64_t *wb_id, const int64_t *to_id, int n_rows,
                                const int64_t *link, const double *vals, int n_attr,
                                trmc_topo *topo, trmc_chparams_f32 **params,
                                int64_t **topo_to_wbid);

/* [new] Seeded synthetic networks for the benchmark workloads.
 *  kind 1: binary-confluence tree, reaches of U{1..40} segments.
 *  kind 2: Shreve-style random tree: uniform random binary topology grown by
 *          random link attachment (n_reaches/2 headwaters, one single-upstream
 *          reach when n_reaches is even), reach lengths geometric with mean
 *          n_segments/n_reaches, clipped to [1, 40], total exactly n_segments.
 *  kind 3: balanced random merge tree with chains, the construction of the
 *          reference's wavefront/gen_network.py:62-246 (n_reaches/6
 *          headwaters merged two or three at a time, Poisson(2) chain reaches
 *          on every link, max(1, Poisson(12)) segments per reach; tens of
 *          levels, confluences of up to 3); n_segments is ignored (> 0).
 * n_reaches is a target (kinds 2, 3) or ignored (kind 1).  Parameters are
 * float-representable doubles so f32 and f64 entry points see equal values. */
int trmc_generate_network(int kind, int n_segments, int n_reaches, uint64_t seed,
                          trmc_topo *out);
int trmc_generate_params(int n_segments, uint64_t seed, trmc_chparams_f64 *out);
/* qlat [n_members][n_timesteps][n_segments], lognormal with mean 0.1 m3/s
 * (sigma_ln 1), member m drawn from its own stream. */
int trmc_generate_forcing(int n_segments, int n_timesteps, int n_members, uint64_t seed,
                          double *qlat);
/* [new] member `member` alone: qlat [n_timesteps][n_segments], the values of
 * member `member` of trmc_generate_forcing (and of the device generator). */
int trmc_generate_forcing_member(int n_segments, int n_timesteps, int member, uint64_t seed,
                                 double *qlat);
/* Initial state: qdp0 = q0, dp0 = wide-channel normal depth at q0. */
int trmc_generate_state(int n_segments, const trmc_chparams_f64 *params, double q0,
                        double *qdp0, double *dp0);

/* Error statistics of wavefront_mc.cu:822-858: out[0..8] = max_abs, p50, p90,
 * p99, p99.9, max_rel (relative error over |truth| > 1e-3), fraction < 1 %,
 * fraction < 10 %, NaN/Inf count. */
int trmc_compute_stats(const double *pred, const double *truth, int n, double *out);

#ifdef __cplusplus
}
#endif
#endif /* TRMC_H */
