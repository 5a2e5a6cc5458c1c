/*
 * CPU oracle for Muskingum-Cunge routing -- TEST INFRASTRUCTURE ONLY.
 * (Checker for tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.)
 *
 * Plain-C restatement of the reference's FP64 host routines.  The arithmetic
 * keeps the reference's operation order term for term so that, built with
 * `gcc -O2 -ffp-contract=off`, the results are bit-identical to the reference
 * routine compiled the same way.  File:line citations are into the reference
 * tree (water-prediction/t-route/...).
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* The pow of the MC arithmetic.  Default: the host libm's pow, as the
 * reference calls it (bit-for-bit the reference's results).  With
 * ORC_CR_POW: the correctly rounded pow, computed in quad precision and
 * rounded once (libquadmath) -- an implementation independent of the GPU
 * library's, used to check that the GPU is the same algorithm to the bit. */
#ifdef ORC_CR_POW
#include <quadmath.h>
/* Ziv's two-step rounding: the x87 extended pow (64-bit significand) decides
 * the rounding unless its value lies within 2^-60 relative of a double
 * rounding boundary; then the quad-precision pow (113-bit) decides. */
static double orc_pow(double x, double y)
{
    long double r = powl((long double)x, (long double)y);
    double d = (double)r;
    if (isfinite(d) && d > 0.0) {
        long double up = (long double)nextafter(d, INFINITY) - (long double)d;
        long double dn = (long double)d - (long double)nextafter(d, 0.0);
        long double dist = r >= (long double)d ? up / 2 - (r - (long double)d)
                                               : dn / 2 - ((long double)d - r);
        if (dist > fabsl(r) * 0x1p-60L) return d;
    }
    return (double)powq((__float128)x, (__float128)y);
}
#else
static double orc_pow(double x, double y) { return pow(x, y); }
#endif

double orc_pow_fn(double x, double y) { return orc_pow(x, y); }

/* Per-segment constants derived once, wavefront_mc.cu:221-225. */
typedef struct {
    double dx, bw, tw, twcc, n_ch, ncc, s0;
    double z;    /* 1/cs, or 1 when cs == 0 */
    double bfd;  /* bankfull depth */
} chan_t;

/* Hydraulic geometry of the trapezoid + compound channel at depth h,
 * wavefront_mc.cu:200-213. */
typedef struct { double twl, R, A, AC, WP, WPC; } geo_t;

/* Muskingum routing coefficients, wavefront_mc.cu:262-266 / :288-292. */
typedef struct { double c1, c2, c3, c4; } mcoef_t;

static void chan_init(const orc_chparams_f64 *p, chan_t *c)
{
    c->dx = p->dx; c->bw = p->bw; c->tw = p->tw; c->twcc = p->twcc;
    c->n_ch = p->n_ch; c->ncc = p->ncc; c->s0 = p->s0;
    c->z = (p->cs == 0.0) ? 1.0 : 1.0 / p->cs;
    if (c->bw > c->tw)       c->bfd = c->bw / 0.00001;
    else if (c->bw == c->tw) c->bfd = c->bw / (2.0 * c->z);
    else                     c->bfd = (c->tw - c->bw) / (2.0 * c->z);
}

/* wavefront_mc.cu:227-229 */
static int chan_invalid(const chan_t *c)
{
    return c->n_ch <= 0.0 || c->s0 <= 0.0 || c->z <= 0.0 || c->bw <= 0.0;
}

static void geometry(double h, const chan_t *c, geo_t *g)
{
    g->twl = c->bw + 2.0 * c->z * h;
    double over = fmax(h - c->bfd, 0.0);   /* depth above bankfull */
    double in = fmin(c->bfd, h);           /* depth inside the main channel */
    if (over > 0.0 && c->twcc <= 0.0) { over = 0.0; in = h; }
    g->A = (c->bw + in * c->z) * in;
    g->WP = c->bw + 2.0 * in * sqrt(1.0 + c->z * c->z);
    g->AC = c->twcc * over;
    g->WPC = (over > 0.0) ? (c->twcc + 2.0 * over) : 0.0;
    g->R = (g->WP + g->WPC > 0.0) ? (g->A + g->AC) / (g->WP + g->WPC) : 0.0;
}

/* Kinematic wave celerity, wavefront_mc.cu:249-254 (h_0) and :276-281 (h). */
static double celerity(double h, const geo_t *g, const chan_t *c)
{
    if (!(h > 0.0)) return 0.0;
    double r23 = orc_pow(g->R, 2.0 / 3.0);
    double r53 = orc_pow(g->R, 5.0 / 3.0);
    return fmax(0.0, (sqrt(c->s0) / c->n_ch) *
                ((5.0 / 3.0) * r23 - (2.0 / 3.0) * r53 *
                 (2.0 * sqrt(1.0 + c->z * c->z) / (c->bw + 2.0 * h * c->z))));
}

/* Storage constant Km, wavefront_mc.cu:255 / :282. */
static double storage_km(double ck, double dt, const chan_t *c)
{
    return (ck > 0.0) ? fmax(dt, c->dx / ck) : dt;
}

/* Top width used by the weighting factor, wavefront_mc.cu:256 / :283. */
static double weight_width(double h, const geo_t *g, const chan_t *c)
{
    return (h > c->bfd && c->twcc > 0.0) ? c->twcc : g->twl;
}

/* Muskingum weighting X clamped to [lo, 0.5] from a flow q,
 * wavefront_mc.cu:257-259 (lo = 0, q = previous Qj_0) and :284-287
 * (lo = 0.25, q = routed flow with the h_0 coefficients). */
static double weighting(double ck, double twu, double q, double lo, const chan_t *c)
{
    if (ck > 0.0 && twu * c->s0 * ck * c->dx > 0.0)
        return fmin(0.5, fmax(lo, 0.5 * (1.0 - q / (2.0 * twu * c->s0 * ck * c->dx))));
    return 0.5;
}

static void muskingum(double km, double x, double dt, double ql, mcoef_t *m)
{
    double D = km * (1.0 - x) + dt / 2.0;
    if (D == 0.0) D = 1.0;
    m->c1 = (km * x + dt / 2.0) / D;
    m->c2 = (dt / 2.0 - km * x) / D;
    m->c3 = (km * (1.0 - x) - dt / 2.0) / D;
    m->c4 = (ql * dt) / D;
}

/* Manning flow of the compound section, wavefront_mc.cu:267-270 / :298-302. */
static double manning(const geo_t *g, const chan_t *c)
{
    if (!(g->WP + g->WPC > 0.0)) return 0.0;
    return (1.0 / (((g->WP * c->n_ch) + (g->WPC * c->ncc)) / (g->WP + g->WPC)))
           * (g->A + g->AC) * orc_pow(g->R, 2.0 / 3.0) * sqrt(c->s0);
}

/* Outflow from the final coefficients, wavefront_mc.cu:321-328. */
static double outflow(const mcoef_t *m, double qup, double quc, double qdp)
{
    double routed = m->c1 * qup + m->c2 * quc + m->c3 * qdp;
    double qmc = routed + m->c4;
    if (qmc < 0.0) {
        if (m->c4 < 0.0 && fabs(m->c4) > routed) return 0.0;
        return fmax(m->c1 * qup + m->c2 * quc + m->c4, m->c1 * qup + m->c3 * qdp + m->c4);
    }
    return qmc;
}

/* Trapezoid-only velocity at the final depth, wavefront_mc.cu:330-336. */
static double velocity(double h, const chan_t *c)
{
    double twl = c->bw + 2.0 * c->z * h;
    double Rv = (h * (c->bw + twl) / 2.0) /
                (c->bw + 2.0 * sqrt(((twl - c->bw) / 2.0) * ((twl - c->bw) / 2.0) + h * h));
    if (Rv < 0.0) Rv = 0.0;
    return (1.0 / c->n_ch) * orc_pow(Rv, 2.0 / 3.0) * sqrt(c->s0);
}

double orc_mc_solve(double dt, double qup, double quc, double qdp, double ql,
                    const orc_chparams_f64 *p, double depthp,
                    double *depthc, double *velc, int *iters)
{
    chan_t c;
    chan_init(p, &c);
    if (iters) *iters = 0;
    /* guards, wavefront_mc.cu:227-232 */
    if (chan_invalid(&c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0)) {
        *depthc = 0.0; *velc = 0.0;
        return 0.0;
    }
    const double mindepth = 0.01;
    double d0 = fmax(depthp, 0.0);
    double h = d0 * 1.33 + mindepth;   /* upper secant point, :235 */
    double h_0 = d0 * 0.67;            /* lower secant point, :236 */
    mcoef_t C = {0.0, 0.0, 0.0, 0.0};
    double X = 0.25, Qj_0 = 0.0, Qj = 0.0;
    int maxiter = 100, total = 0;
    (void)X;

    /* attempt loop with bracket widening, :242 and :315-318 */
    for (int attempt = 0; attempt < 5; ++attempt) {
        double aerror = 0.01, rerror = 1.0;
        int iter = 0;
        while (rerror > 0.01 && aerror >= mindepth && iter <= maxiter) {
            geo_t g0, g1;
            mcoef_t m0;
            geometry(h_0, &c, &g0);
            double ck0 = celerity(h_0, &g0, &c);
            double km0 = storage_km(ck0, dt, &c);
            double x0 = weighting(ck0, weight_width(h_0, &g0, &c), Qj_0, 0.0, &c);
            muskingum(km0, x0, dt, ql, &m0);
            double q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
            Qj_0 = q0 - manning(&g0, &c);

            geometry(h, &c, &g1);
            double ck1 = celerity(h, &g1, &c);
            double km1 = storage_km(ck1, dt, &c);
            X = weighting(ck1, weight_width(h, &g1, &c), q0, 0.25, &c);
            muskingum(km1, X, dt, ql, &C);
            double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
            if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
            Qj = (routed + C.c4) - manning(&g1, &c);

            /* secant update, :306-313 */
            double h_1 = (Qj_0 - Qj != 0.0) ? h - (Qj * (h_0 - h)) / (Qj_0 - Qj) : h;
            if (h_1 < 0.0) h_1 = h;
            if (h > 0.0) { rerror = fabs((h_1 - h) / h); aerror = fabs(h_1 - h); }
            else { rerror = 0.0; aerror = 0.9; }
            h_0 = fmax(0.0, h);
            h = fmax(0.0, h_1);
            ++iter;
            if (h < mindepth) break;
        }
        total += iter;
        if (iter < maxiter) break;
        h *= 1.33;
        h_0 *= 0.67;
        maxiter += 25;
    }
    if (iters) *iters = total;
    *velc = velocity(h, &c);
    *depthc = h;
    return outflow(&C, qup, quc, qdp);
}

/* ------------------------------------------------------------------------ */
/* DAG routing over T steps, wavefront_mc.cu:495-564.                       */

static void route_dag(const orc_chparams_f64 *params,
                      const int *reach_seg_start, const int *reach_seg_len,
                      const int *reach_up_ptr, const int *reach_up_idx,
                      const int *level_ptr, const int *level_reach,
                      int n_reaches, int n_levels, int n_segments,
                      const float *qlat_f, const double *qlat_d,
                      double *q_prev, double *d_prev, double *rl_prev,
                      double *v_last, double *reach_hist,
                      int n_timesteps, double dt, int mode, long long *iter_sum, int n_threads)
{
    double *q_cur = (double *)calloc((size_t)n_segments + 1, sizeof(double));
    double *d_cur = (double *)calloc((size_t)n_segments + 1, sizeof(double));
    double *rl_cur = (double *)calloc((size_t)n_reaches + 1, sizeof(double));
    long long its = 0;

    for (int t = 0; t < n_timesteps; ++t) {
        size_t row = (size_t)t * (size_t)n_segments;
        for (int L = 0; L < n_levels; ++L) {
            /* the reaches of one level are independent: with n_threads > 1
             * they are shared out over threads, each reach computed exactly
             * as on one thread */
#pragma omp parallel for schedule(dynamic, 32) num_threads(n_threads) if (n_threads > 1) reduction(+ : its)
            for (int ii = level_ptr[L]; ii < level_ptr[L + 1]; ++ii) {
                int r = level_reach[ii];
                /* fixed-order upstream sums in CSR order, :523-529 */
                double qup_r = 0.0, quc_r = 0.0;
                for (int k = reach_up_ptr[r]; k < reach_up_ptr[r + 1]; ++k) {
                    int u = reach_up_idx[k];
                    qup_r += rl_prev[u];
                    quc_r += rl_cur[u];
                }
                double quc_next = 0.0, last = 0.0;
                int s0 = reach_seg_start[r];
                for (int i = 0; i < reach_seg_len[r]; ++i) {
                    int idx = s0 + i;
                    double qup = (i == 0) ? qup_r : q_prev[idx - 1];
                    double quc = (i == 0) ? quc_r : quc_next;
                    double ql = qlat_f ? (double)qlat_f[row + idx] : qlat_d[row + idx];
                    double dep, vel, qdc;
                    int it = 0;
                    if (mode == ORC_MODE_CONVERGED)
                        qdc = orc_mc_solve_converged(dt, qup, quc, q_prev[idx], ql,
                                                     &params[idx], d_prev[idx], &dep, &vel);
                    else
                        qdc = orc_mc_solve(dt, qup, quc, q_prev[idx], ql, &params[idx],
                                           d_prev[idx], &dep, &vel, &it);
                    its += it;
                    q_cur[idx] = qdc;
                    d_cur[idx] = dep;
                    if (v_last && t == n_timesteps - 1) v_last[idx] = vel;
                    quc_next = qdc;
                    last = qdc;
                }
                rl_cur[r] = last;
            }
        }
        /* roll state, :554-559 */
        memcpy(q_prev, q_cur, (size_t)n_segments * sizeof(double));
        memcpy(d_prev, d_cur, (size_t)n_segments * sizeof(double));
        memcpy(rl_prev, rl_cur, (size_t)n_reaches * sizeof(double));
        if (reach_hist) memcpy(reach_hist + (size_t)t * n_reaches, rl_cur, (size_t)n_reaches * sizeof(double));
        memset(rl_cur, 0, (size_t)n_reaches * sizeof(double));
        memset(q_cur, 0, (size_t)n_segments * sizeof(double));
        memset(d_cur, 0, (size_t)n_segments * sizeof(double));
    }
    if (iter_sum) *iter_sum += its;
    free(q_cur); free(d_cur); free(rl_cur);
}

static void wavefront_f32in(const orc_chparams_f32 *params,
                            const int *reach_seg_start, const int *reach_seg_len,
                            const int *reach_up_ptr, const int *reach_up_idx,
                            const int *level_ptr, const int *level_reach,
                            int n_reaches, int n_levels, int n_segments,
                            const float *qlat_ts, const float *qdp0_seg,
                            const float *dp0_seg, const float *reach_last0,
                            double *qdc_seg_out_d, double *reach_last_out_d,
                            int n_timesteps, double dt, long long *iter_sum, int n_threads)
{
    /* float params are widened field by field, :536-538 */
    orc_chparams_f64 *pd = (orc_chparams_f64 *)malloc(((size_t)n_segments + 1) * sizeof(*pd));
    double *q = (double *)malloc(((size_t)n_segments + 1) * sizeof(double));
    double *d = (double *)malloc(((size_t)n_segments + 1) * sizeof(double));
    double *rl = (double *)malloc(((size_t)n_reaches + 1) * sizeof(double));
    for (int i = 0; i < n_segments; ++i) {
        const orc_chparams_f32 *s = &params[i];
        orc_chparams_f64 w = {s->dx, s->bw, s->tw, s->twcc, s->n_ch, s->ncc, s->cs, s->s0};
        pd[i] = w;
        q[i] = qdp0_seg[i];
        d[i] = dp0_seg[i];
    }
    for (int r = 0; r < n_reaches; ++r) rl[r] = reach_last0[r];
    route_dag(pd, reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx,
              level_ptr, level_reach, n_reaches, n_levels, n_segments,
              qlat_ts, NULL, q, d, rl, NULL, NULL, n_timesteps, dt,
              ORC_MODE_REFERENCE, iter_sum, n_threads);
    memcpy(qdc_seg_out_d, q, (size_t)n_segments * sizeof(double));
    memcpy(reach_last_out_d, rl, (size_t)n_reaches * sizeof(double));
    free(pd); free(q); free(d); free(rl);
}

void orc_wavefront_f32in(const orc_chparams_f32 *params,
                         const int *reach_seg_start, const int *reach_seg_len,
                         const int *reach_up_ptr, const int *reach_up_idx,
                         const int *level_ptr, const int *level_reach,
                         int n_reaches, int n_levels, int n_segments,
                         const float *qlat_ts, const float *qdp0_seg,
                         const float *dp0_seg, const float *reach_last0,
                         double *qdc_seg_out_d, double *reach_last_out_d,
                         int n_timesteps, double dt, long long *iter_sum)
{
    wavefront_f32in(params, reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr,
                    level_reach, n_reaches, n_levels, n_segments, qlat_ts, qdp0_seg, dp0_seg, reach_last0,
                    qdc_seg_out_d, reach_last_out_d, n_timesteps, dt, iter_sum, 1);
}

/* The same on n_threads host threads, level by level (a labelled all-core
 * CPU baseline; results equal the one-thread routine's bit for bit). */
void orc_wavefront_f32in_mt(const orc_chparams_f32 *params,
                            const int *reach_seg_start, const int *reach_seg_len,
                            const int *reach_up_ptr, const int *reach_up_idx,
                            const int *level_ptr, const int *level_reach,
                            int n_reaches, int n_levels, int n_segments,
                            const float *qlat_ts, const float *qdp0_seg,
                            const float *dp0_seg, const float *reach_last0,
                            double *qdc_seg_out_d, double *reach_last_out_d,
                            int n_timesteps, double dt, long long *iter_sum, int n_threads)
{
    wavefront_f32in(params, reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr,
                    level_reach, n_reaches, n_levels, n_segments, qlat_ts, qdp0_seg, dp0_seg, reach_last0,
                    qdc_seg_out_d, reach_last_out_d, n_timesteps, dt, iter_sum, n_threads < 1 ? 1 : n_threads);
}

static void wavefront_f64(const orc_chparams_f64 *params,
                          const int *reach_seg_start, const int *reach_seg_len,
                          const int *reach_up_ptr, const int *reach_up_idx,
                          const int *level_ptr, const int *level_reach,
                          int n_reaches, int n_levels, int n_segments,
                          const double *qlat_ts, const double *qdp0_seg,
                          const double *dp0_seg, const double *reach_last0,
                          double *qdc_seg_out, double *reach_last_out,
                          double *depth_out, double *vel_out, double *reach_hist,
                          int n_timesteps, double dt, int mode, long long *iter_sum, int n_threads)
{
    double *q = (double *)malloc(((size_t)n_segments + 1) * sizeof(double));
    double *d = (double *)malloc(((size_t)n_segments + 1) * sizeof(double));
    double *rl = (double *)malloc(((size_t)n_reaches + 1) * sizeof(double));
    memcpy(q, qdp0_seg, (size_t)n_segments * sizeof(double));
    memcpy(d, dp0_seg, (size_t)n_segments * sizeof(double));
    memcpy(rl, reach_last0, (size_t)n_reaches * sizeof(double));
    if (vel_out) memset(vel_out, 0, (size_t)n_segments * sizeof(double));
    route_dag(params, reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx,
              level_ptr, level_reach, n_reaches, n_levels, n_segments,
              NULL, qlat_ts, q, d, rl, vel_out, reach_hist, n_timesteps, dt, mode, iter_sum, n_threads);
    memcpy(qdc_seg_out, q, (size_t)n_segments * sizeof(double));
    memcpy(reach_last_out, rl, (size_t)n_reaches * sizeof(double));
    if (depth_out) memcpy(depth_out, d, (size_t)n_segments * sizeof(double));
    free(q); free(d); free(rl);
}

void orc_wavefront_f64(const orc_chparams_f64 *params,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const double *qlat_ts, const double *qdp0_seg,
                       const double *dp0_seg, const double *reach_last0,
                       double *qdc_seg_out, double *reach_last_out,
                       double *depth_out, double *vel_out, double *reach_hist,
                       int n_timesteps, double dt, int mode, long long *iter_sum)
{
    wavefront_f64(params, reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr, level_reach,
                  n_reaches, n_levels, n_segments, qlat_ts, qdp0_seg, dp0_seg, reach_last0, qdc_seg_out,
                  reach_last_out, depth_out, vel_out, reach_hist, n_timesteps, dt, mode, iter_sum, 1);
}

void orc_wavefront_f64_mt(const orc_chparams_f64 *params,
                          const int *reach_seg_start, const int *reach_seg_len,
                          const int *reach_up_ptr, const int *reach_up_idx,
                          const int *level_ptr, const int *level_reach,
                          int n_reaches, int n_levels, int n_segments,
                          const double *qlat_ts, const double *qdp0_seg,
                          const double *dp0_seg, const double *reach_last0,
                          double *qdc_seg_out, double *reach_last_out,
                          double *depth_out, double *vel_out, double *reach_hist,
                          int n_timesteps, double dt, int mode, long long *iter_sum, int n_threads)
{
    wavefront_f64(params, reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr, level_reach,
                  n_reaches, n_levels, n_segments, qlat_ts, qdp0_seg, dp0_seg, reach_last0, qdc_seg_out,
                  reach_last_out, depth_out, vel_out, reach_hist, n_timesteps, dt, mode, iter_sum, n_threads);
}

/* ------------------------------------------------------------------------ */
/* FP32 REFERENCE rule: the reference's float solve mc_solve                */
/* (wavefront_mc.cu:51-193, hgeo_inline :51-64) and its single-thread DAG   */
/* loop cpu_wavefront (:413-491), every operation in float as there (x86-64 */
/* SSE float: no excess precision; -ffp-contract=off: no FMA).              */

/* powf of the float arithmetic: the host libm's powf (liborc.so, as the
 * reference calls it), or the correctly rounded powf (liborc_cr.so): the
 * double pow (relative error below 2^-52) decides the float rounding unless
 * it lies within 2^-46 relative of a float rounding midpoint (float midpoints
 * are doubles), then the quad pow does. */
#ifdef ORC_CR_POW
static float orc_powf(float x, float y)
{
    double r = pow((double)x, (double)y);
    float f = (float)r;
    if (isfinite(f) && f > 0.0f) {
        double up = ((double)nextafterf(f, INFINITY) - (double)f) / 2;
        double dn = ((double)f - (double)nextafterf(f, 0.0f)) / 2;
        double dist = r >= (double)f ? up - (r - (double)f) : dn - ((double)f - r);
        if (dist > fabs(r) * 0x1p-46) return f;
    }
    return (float)powq((__float128)x, (__float128)y);
}
#else
static float orc_powf(float x, float y) { return powf(x, y); }
#endif

float orc_powf_fn(float x, float y) { return orc_powf(x, y); }

/* float channel: the parameters and z, bankfull depth as :73-77 computes them */
typedef struct { float dx, bw, tw, twcc, n_ch, ncc, s0, z, bfd; } chan32_t;
typedef struct { float twl, R, A, AC, WP, WPC; } geo32_t;
typedef struct { float c1, c2, c3, c4; } coef32_t;

static void chan32_init(const orc_chparams_f32 *p, chan32_t *c)
{
    c->dx = p->dx; c->bw = p->bw; c->tw = p->tw; c->twcc = p->twcc;
    c->n_ch = p->n_ch; c->ncc = p->ncc; c->s0 = p->s0;
    c->z = (p->cs == 0.0f) ? 1.0f : 1.0f / p->cs;
    if (p->bw > p->tw) c->bfd = p->bw / 0.00001f;
    else if (p->bw == p->tw) c->bfd = p->bw / (2.0f * c->z);
    else c->bfd = (p->tw - p->bw) / (2.0f * c->z);
}

static void geometry32(float h, const chan32_t *c, geo32_t *g)
{
    g->twl = c->bw + 2.0f * c->z * h;
    float over = fmaxf(h - c->bfd, 0.0f);
    float in = fminf(c->bfd, h);
    if (over > 0.0f && c->twcc <= 0.0f) { over = 0.0f; in = h; }
    g->A = (c->bw + in * c->z) * in;
    g->WP = c->bw + 2.0f * in * sqrtf(1.0f + c->z * c->z);
    g->AC = c->twcc * over;
    g->WPC = (over > 0.0f) ? (c->twcc + 2.0f * over) : 0.0f;
    g->R = (g->WP + g->WPC > 0.0f) ? (g->A + g->AC) / (g->WP + g->WPC) : 0.0f;
}

/* one secant point's depth functions (:102-124 / :129-156) */
static float celerity32(float h, const geo32_t *g, const chan32_t *c)
{
    if (!(h > 0.0f)) return 0.0f;
    const float r23 = orc_powf(g->R, 2.0f / 3.0f);
    const float r53 = orc_powf(g->R, 5.0f / 3.0f);
    return fmaxf(0.0f, (sqrtf(c->s0) / c->n_ch) *
                       ((5.0f / 3.0f) * r23 - (2.0f / 3.0f) * r53 *
                        (2.0f * sqrtf(1.0f + c->z * c->z) / (c->bw + 2.0f * h * c->z))));
}

static float weighting32(float ck, float twu, float q, float lo, const chan32_t *c)
{
    if (ck > 0.0f && twu * c->s0 * ck * c->dx > 0.0f)
        return fminf(0.5f, fmaxf(lo, 0.5f * (1.0f - q / (2.0f * twu * c->s0 * ck * c->dx))));
    return 0.5f;
}

static void muskingum32(float km, float x, float dt, float ql, coef32_t *m)
{
    float D = km * (1.0f - x) + dt / 2.0f;
    if (D == 0.0f) D = 1.0f;
    m->c1 = (km * x + dt / 2.0f) / D;
    m->c2 = (dt / 2.0f - km * x) / D;
    m->c3 = (km * (1.0f - x) - dt / 2.0f) / D;
    m->c4 = (ql * dt) / D;
}

static float manning32(const geo32_t *g, const chan32_t *c)
{
    if (!(g->WP + g->WPC > 0.0f)) return 0.0f;
    return (1.0f / (((g->WP * c->n_ch) + (g->WPC * c->ncc)) / (g->WP + g->WPC))) * (g->A + g->AC) *
           orc_powf(g->R, 2.0f / 3.0f) * sqrtf(c->s0);
}

float orc_mc_solve_f32(float dt, float qup, float quc, float qdp, float ql,
                       const orc_chparams_f32 *p, float depthp,
                       float *depthc, float *velc, int *iters)
{
    chan32_t c;
    chan32_init(p, &c);
    if (iters) *iters = 0;
    if (c.n_ch <= 0.0f || c.s0 <= 0.0f || c.z <= 0.0f || c.bw <= 0.0f ||
        (ql <= 0.0f && qup <= 0.0f && quc <= 0.0f && qdp <= 0.0f)) {      /* :78-84 */
        *depthc = 0.0f; *velc = 0.0f;
        return 0.0f;
    }
    const float mindepth = 0.01f;
    float d0 = fmaxf(depthp, 0.0f);
    float h = d0 * 1.33f + mindepth;
    float h_0 = d0 * 0.67f;
    coef32_t C = {0.0f, 0.0f, 0.0f, 0.0f};
    float Qj_0 = 0.0f;
    int maxiter = 100, total = 0;
    for (int attempt = 0; attempt < 5; ++attempt) {                      /* :95-175 */
        float aerror = 0.01f, rerror = 1.0f;
        int iter = 0;
        while (rerror > 0.01f && aerror >= mindepth && iter <= maxiter) {
            geo32_t g0, g1;
            coef32_t m0;
            geometry32(h_0, &c, &g0);
            const float ck0 = celerity32(h_0, &g0, &c);
            const float km0 = (ck0 > 0.0f) ? fmaxf(dt, c.dx / ck0) : dt;
            const float twu0 = (h_0 > c.bfd && c.twcc > 0.0f) ? c.twcc : g0.twl;
            muskingum32(km0, weighting32(ck0, twu0, Qj_0, 0.0f, &c), dt, ql, &m0);
            const float q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
            Qj_0 = q0 - manning32(&g0, &c);

            geometry32(h, &c, &g1);
            const float ck1 = celerity32(h, &g1, &c);
            const float km1 = (ck1 > 0.0f) ? fmaxf(dt, c.dx / ck1) : dt;
            const float twu1 = (h > c.bfd && c.twcc > 0.0f) ? c.twcc : g1.twl;
            muskingum32(km1, weighting32(ck1, twu1, q0, 0.25f, &c), dt, ql, &C);
            const float routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
            if (C.c4 < 0.0f && fabsf(C.c4) > routed) C.c4 = -routed;
            const float Qj = (routed + C.c4) - manning32(&g1, &c);

            float h_1 = (Qj_0 - Qj != 0.0f) ? h - (Qj * (h_0 - h)) / (Qj_0 - Qj) : h;
            if (h_1 < 0.0f) h_1 = h;
            if (h > 0.0f) { rerror = fabsf((h_1 - h) / h); aerror = fabsf(h_1 - h); }
            else { rerror = 0.0f; aerror = 0.9f; }
            h_0 = fmaxf(0.0f, h);
            h = fmaxf(0.0f, h_1);
            ++iter;
            if (h < mindepth) break;
        }
        total += iter;
        if (iter < maxiter) break;
        h *= 1.33f;
        h_0 *= 0.67f;
        maxiter += 25;
    }
    if (iters) *iters = total;
    /* outflow rule :177-184 */
    const float routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
    const float Qmc = routed + C.c4;
    float qdc = Qmc;
    if (Qmc < 0.0f) {
        if (C.c4 < 0.0f && fabsf(C.c4) > routed) qdc = 0.0f;
        else qdc = fmaxf(C.c1 * qup + C.c2 * quc + C.c4, C.c1 * qup + C.c3 * qdp + C.c4);
    }
    /* trapezoid velocity :186-190 */
    const float twl = c.bw + 2.0f * c.z * h;
    float Rv = (h * (c.bw + twl) / 2.0f) /
               (c.bw + 2.0f * sqrtf(((twl - c.bw) / 2.0f) * ((twl - c.bw) / 2.0f) + h * h));
    if (Rv < 0.0f) Rv = 0.0f;
    *velc = (1.0f / c.n_ch) * orc_powf(Rv, 2.0f / 3.0f) * sqrtf(c.s0);
    *depthc = h;
    return qdc;
}

/* cpu_wavefront (:413-491): levels in order, reaches of a level in their
 * order, float upstream sums from 0 in CSR order, segments downstream.  With
 * n_threads > 1 the reaches of one level are shared out over threads (each
 * computed exactly as on one thread).  depth_out (may be NULL) gets the final
 * depths, reach_hist (may be NULL) reach_last after every step. */
void orc_wavefront_f32(const orc_chparams_f32 *params,
                       const int *reach_seg_start, const int *reach_seg_len,
                       const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach,
                       int n_reaches, int n_levels, int n_segments,
                       const float *qlat_ts, const float *qdp0_seg,
                       const float *dp0_seg, const float *reach_last0,
                       float *qdc_seg_out, float *reach_last_out,
                       float *depth_out, float *reach_hist,
                       int n_timesteps, float dt, long long *iter_sum, int n_threads)
{
    const size_t ns = (size_t)n_segments + 1, nr = (size_t)n_reaches + 1;
    float *q_prev = (float *)malloc(ns * sizeof(float)), *d_prev = (float *)malloc(ns * sizeof(float));
    float *q_cur = (float *)calloc(ns, sizeof(float)), *d_cur = (float *)calloc(ns, sizeof(float));
    float *rl_prev = (float *)malloc(nr * sizeof(float)), *rl_cur = (float *)calloc(nr, sizeof(float));
    memcpy(q_prev, qdp0_seg, (size_t)n_segments * sizeof(float));
    memcpy(d_prev, dp0_seg, (size_t)n_segments * sizeof(float));
    memcpy(rl_prev, reach_last0, (size_t)n_reaches * sizeof(float));
    if (n_threads < 1) n_threads = 1;
    long long its = 0;
    for (int t = 0; t < n_timesteps; ++t) {
        const float *qlat = qlat_ts + (size_t)t * (size_t)n_segments;
        for (int L = 0; L < n_levels; ++L) {
#pragma omp parallel for schedule(dynamic, 32) num_threads(n_threads) if (n_threads > 1) reduction(+ : its)
            for (int ii = level_ptr[L]; ii < level_ptr[L + 1]; ++ii) {
                const int r = level_reach[ii];
                float qup_r = 0.0f, quc_r = 0.0f;
                for (int k = reach_up_ptr[r]; k < reach_up_ptr[r + 1]; ++k) {
                    qup_r += rl_prev[reach_up_idx[k]];
                    quc_r += rl_cur[reach_up_idx[k]];
                }
                float quc_next = 0.0f, last = 0.0f;
                for (int i = 0; i < reach_seg_len[r]; ++i) {
                    const int idx = reach_seg_start[r] + i;
                    float dep, vel;
                    int it = 0;
                    const float qdc = orc_mc_solve_f32(dt, i == 0 ? qup_r : q_prev[idx - 1],
                                                       i == 0 ? quc_r : quc_next, q_prev[idx], qlat[idx],
                                                       &params[idx], d_prev[idx], &dep, &vel, &it);
                    its += it;
                    q_cur[idx] = qdc;
                    d_cur[idx] = dep;
                    quc_next = qdc;
                    last = qdc;
                }
                rl_cur[r] = last;
            }
        }
        memcpy(q_prev, q_cur, (size_t)n_segments * sizeof(float));
        memcpy(d_prev, d_cur, (size_t)n_segments * sizeof(float));
        memcpy(rl_prev, rl_cur, (size_t)n_reaches * sizeof(float));
        if (reach_hist) memcpy(reach_hist + (size_t)t * n_reaches, rl_cur, (size_t)n_reaches * sizeof(float));
        memset(rl_cur, 0, (size_t)n_reaches * sizeof(float));
    }
    memcpy(qdc_seg_out, q_prev, (size_t)n_segments * sizeof(float));
    memcpy(reach_last_out, rl_prev, (size_t)n_reaches * sizeof(float));
    if (depth_out) memcpy(depth_out, d_prev, (size_t)n_segments * sizeof(float));
    if (iter_sum) *iter_sum += its;
    free(q_prev); free(d_prev); free(q_cur); free(d_cur); free(rl_prev); free(rl_cur);
}

/* troute_mc_final.cu:167-252.  The inlined solve there is the same
 * arithmetic as mc_solve_d; the invalid/zero-flow guard zeroes the outputs
 * and passes quc = 0 to the next segment (:187-190), which is what a zero
 * return from the solve does as well. */
void orc_reach_parallel(const orc_chparams_f64 *p, const double *ql,
                        const int *rs, const int *rl,
                        const double *qu, const double *qd, const double *dp,
                        double *qo, double *vo, double *dpo,
                        double uq, double dt, int nr, long long *iter_sum)
{
    long long its = 0;
    for (int r = 0; r < nr; ++r) {
        double quc = uq;
        for (int i = 0; i < rl[r]; ++i) {
            int idx = rs[r] + i;
            double dep, vel;
            int it = 0;
            double qdc = orc_mc_solve(dt, qu[idx], quc, qd[idx], ql[idx], &p[idx],
                                      dp[idx], &dep, &vel, &it);
            its += it;
            qo[idx] = qdc; vo[idx] = vel; dpo[idx] = dep;
            quc = qdc;
        }
    }
    if (iter_sum) *iter_sum += its;
}

/* ------------------------------------------------------------------------ */
/* Converged mode.  No reference counterpart.
 *
 * The reference's secant (mc_solve_d) stops at 1 % / 1 cm, so its answer is
 * an iterate, not a root.  At a fixed point of that iteration h_0 == h and
 * Qj_0 == 0, so X0 == 0.5 and the residual reduces to the single-valued
 *
 *   F(h) = Qmc(h) - Qmn(h),
 *   Qmc(h) = C1 qup + C2 quc + C3 qdp + C4 with C(Km(h), X(h)) and the C4
 *            clamp of wavefront_mc.cu:293-296,
 *   X(h)   = weighting with lo = 0.25 from the routed flow of the X0 = 0.5
 *            coefficients at the same h.
 *
 * The converged depth h* is the leftmost downcrossing of F on [mindepth, inf):
 * the smallest h where F passes from > 0 (just below) to <= 0 (at h); h* =
 * mindepth when F is positive nowhere.  F need not be monotone: F(mindepth)
 * <= 0 on a sharply rising hydrograph (the routed flow at small depth is
 * negative), with a positive band above it, and the floodplain's jump in
 * wetted perimeter at bankfull can make F positive again above an in-bank
 * root.  A search over sample points can step over a narrow band, so h* is
 * found by a CERTIFIED scan: every interval the scan passes over is shown to
 * keep the sign of F by an enclosure of F on the whole interval
 * (f_enclosure below), not by samples.  Only features narrower than
 * CERT_RES relative depth (where the scan falls back to point values) are
 * below its resolution.  Outputs then follow wavefront_mc.cu:321-336 at h*. */

static double residual_at(double h, double dt, double qup, double quc, double qdp,
                          double ql, const chan_t *c, mcoef_t *Cout)
{
    geo_t g;
    mcoef_t m0, C;
    geometry(h, c, &g);
    double ck = celerity(h, &g, c);
    double km = storage_km(ck, dt, c);
    double twu = weight_width(h, &g, c);
    muskingum(km, 0.5, dt, ql, &m0);
    double q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
    double X = weighting(ck, twu, q0, 0.25, c);
    muskingum(km, X, dt, ql, &C);
    double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
    if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
    if (Cout) *Cout = C;
    return (routed + C.c4) - manning(&g, c);
}

double orc_mc_residual(double dt, double qup, double quc, double qdp, double ql,
                       const orc_chparams_f64 *p, double h, double *qmc_out)
{
    chan_t c;
    mcoef_t C;
    chan_init(p, &c);
    double f = residual_at(h, dt, qup, quc, qdp, ql, &c, &C);
    if (qmc_out) *qmc_out = C.c1 * qup + C.c2 * quc + C.c3 * qdp + C.c4;
    return f;
}

/* Leftmost downcrossing of f on [lo, inf): the smallest h where f changes
 * from > 0 (just below) to <= 0 (at h).  Geometric scan (x1.25 from lo,
 * visiting the breakpoint bp = bankfull) for the first f > 0, then on for the
 * first f <= 0, then bisection to a 2-ulp bracket.  Returns lo when f is
 * never positive on the scan. */
typedef double (*scalar_fn)(double h, const void *ctx);

static double leftmost_downcrossing(scalar_fn f, const void *ctx, double lo, double bp)
{
    double x = lo, fx = f(x, ctx);
    int used_bp = !(bp > lo);
    int k = 0;
    /* find the first point with f > 0 */
    while (!(fx > 0.0)) {
        if (++k > 400) return lo;
        double nx = x * 1.25;
        if (!used_bp && nx >= bp) { nx = bp; used_bp = 1; }
        if (nx > 1.0e4) return lo;
        x = nx;
        fx = f(x, ctx);
    }
    /* then the first point after it with f <= 0 */
    double a = x, b = x;
    for (;;) {
        if (++k > 800) return b;
        double nb = b * 1.25;
        if (!used_bp && nb >= bp) { nb = bp; used_bp = 1; }
        double fnb = f(nb, ctx);
        if (!(fnb > 0.0)) { b = nb; break; }
        a = nb;
        b = nb;
    }
    for (int i = 0; i < 200; ++i) {
        double m = a + (b - a) / 2.0;
        if (!(m > a && m < b)) break;
        if (f(m, ctx) > 0.0) a = m; else b = m;
    }
    return a + (b - a) / 2.0;
}

/* ---- enclosure of F on an interval --------------------------------------
 * F depends on h through the geometry and through the storage constant K and
 * weighting X.  For fixed X the routed flow R = C1 qup + C2 quc + C3 qdp + C4
 * is a ratio of functions linear in K (denominator K (1 - X) + dt/2 > 0),
 * and for fixed K a ratio of functions linear in X, so R is monotone in each
 * of K and X separately: over a box of (K, X) its extremes are at corners
 * (K -> inf included as the limit C1 = X/(1-X) = -C2, C3 = 1, C4 = 0).  The
 * C4 clamp turns R into max(R, 0) when ql < 0, which keeps that property.
 *
 * On [a, b] inside one regime (in the main channel, or on the floodplain
 * above bankfull) every depth function used is monotone or a difference of
 * monotone terms:
 *   main channel: area A, hydraulic radius R_h (dR_h/dh > 0 for a trapezoid),
 *                 top width, Manning flow Qmn all rise with h;
 *   floodplain:   A = A_b + twcc o and wetted perimeter P = P_b + twcc + 2 o
 *                 are linear in o = h - bfd, so R_h = A / P and the blended
 *                 roughness (P_b n + (twcc + 2 o) ncc) / P are monotone in o,
 *                 extremes at the ends (the end at bankfull is the limit from
 *                 above, o -> 0+); Qmn >= A(a) min R_h^(2/3) sqrt(s0) / max n;
 *   celerity:     5/3 R_h^(2/3) - 2/3 R_h^(5/3) w, w = 2 sqrt(1+z^2)/(bw+2hz)
 *                 falling in h: bounded term by term;
 *   K = max(dt, dx/ck) from the celerity range (K spans [dt, inf] when the
 *                 celerity can be 0); q0 = R(K, 0.5) from the K range; X =
 *                 clamp(0.5 (1 - q0 / (2 twu s0 ck dx)), 0.25, 0.5) from the
 *                 ranges of q0 and of the denominator.
 * The enclosure is [min R - max Qmn, max R - min Qmn], widened by 1e-12 of
 * the flows for rounding.  It is loose on wide intervals and tight as they
 * shrink, which is all the scan below needs. */
typedef struct { double dt, qup, quc, qdp, ql; const chan_t *c; } route_ctx;

static double routed_kx(const route_ctx *x, double K, double X, int clamp)
{
    double c1, c2, c3, c4;
    if (isinf(K)) {
        c1 = X / (1.0 - X); c2 = -c1; c3 = 1.0; c4 = 0.0;
    } else {
        double D = K * (1.0 - X) + x->dt / 2.0;
        c1 = (K * X + x->dt / 2.0) / D;
        c2 = (x->dt / 2.0 - K * X) / D;
        c3 = (K * (1.0 - X) - x->dt / 2.0) / D;
        c4 = x->ql * x->dt / D;
    }
    double v = c1 * x->qup + c2 * x->quc + c3 * x->qdp + c4;
    if (clamp && c4 < 0.0 && v < 0.0) v = 0.0;
    return v;
}

typedef struct { double A, R, n_eff, twu; } hbox_t;   /* depth functions at one end */

static hbox_t depth_end(double h, int flood, const chan_t *c)
{
    hbox_t e;
    double sq = sqrt(1.0 + c->z * c->z);
    if (!flood) {
        e.A = (c->bw + h * c->z) * h;
        e.R = e.A / (c->bw + 2.0 * h * sq);
        e.n_eff = c->n_ch;
        e.twu = c->bw + 2.0 * c->z * h;
    } else {
        double o = fmax(h - c->bfd, 0.0);
        double Ab = (c->bw + c->bfd * c->z) * c->bfd, Pb = c->bw + 2.0 * c->bfd * sq;
        double P = Pb + c->twcc + 2.0 * o;
        e.A = Ab + c->twcc * o;
        e.R = e.A / P;
        e.n_eff = (Pb * c->n_ch + (c->twcc + 2.0 * o) * c->ncc) / P;
        e.twu = c->twcc;
    }
    return e;
}

/* Enclosure [*lo, *hi] of F on [a, b]; flood = the floodplain regime
 * (a >= bfd, twcc > 0), else the main channel (b <= bfd or no floodplain). */
static void f_enclosure(const route_ctx *x, double a, double b, int flood, double *lo, double *hi)
{
    const chan_t *c = x->c;
    const double dt = x->dt;
    hbox_t ea = depth_end(a, flood, c), eb = depth_end(b, flood, c);
    double Rlo = fmin(ea.R, eb.R), Rhi = fmax(ea.R, eb.R);
    double nlo = fmin(ea.n_eff, eb.n_eff), nhi = fmax(ea.n_eff, eb.n_eff);
    double sqs0 = sqrt(c->s0), ks = sqs0 / c->n_ch, sq = sqrt(1.0 + c->z * c->z);
    /* Manning flow */
    double qmn_lo = ea.A * orc_pow(Rlo, 2.0 / 3.0) * sqs0 / nhi;
    double qmn_hi = eb.A * orc_pow(Rhi, 2.0 / 3.0) * sqs0 / nlo;
    /* celerity */
    double wa = 2.0 * sq / (c->bw + 2.0 * a * c->z), wb = 2.0 * sq / (c->bw + 2.0 * b * c->z);
    double ck_lo = fmax(0.0, ks * ((5.0 / 3.0) * orc_pow(Rlo, 2.0 / 3.0) - (2.0 / 3.0) * orc_pow(Rhi, 5.0 / 3.0) * wa));
    double ck_hi = fmax(0.0, ks * ((5.0 / 3.0) * orc_pow(Rhi, 2.0 / 3.0) - (2.0 / 3.0) * orc_pow(Rlo, 5.0 / 3.0) * wb));
    /* storage constant and weighting */
    double K_lo, K_hi, X_lo, X_hi;
    if (!(ck_hi > 0.0)) {                   /* ck == 0: K = dt, X = 0.5 */
        K_lo = K_hi = dt; X_lo = X_hi = 0.5;
    } else if (!(ck_lo > 0.0)) {            /* ck may be 0 or tiny */
        K_lo = dt; K_hi = INFINITY; X_lo = 0.25; X_hi = 0.5;
    } else {
        K_lo = fmax(dt, c->dx / ck_hi);
        K_hi = fmax(dt, c->dx / ck_lo);
        double q0a = routed_kx(x, K_lo, 0.5, 0), q0b = routed_kx(x, K_hi, 0.5, 0);
        double q0lo = fmin(q0a, q0b), q0hi = fmax(q0a, q0b);
        double twlo = fmin(ea.twu, eb.twu), twhi = fmax(ea.twu, eb.twu);
        double dlo = twlo * c->s0 * ck_lo * c->dx, dhi = twhi * c->s0 * ck_hi * c->dx;
        if (!(dlo > 0.0)) {
            X_lo = 0.25; X_hi = 0.5;
        } else {
            double rlo = fmin(q0lo / (2.0 * dlo), q0lo / (2.0 * dhi));
            double rhi = fmax(q0hi / (2.0 * dlo), q0hi / (2.0 * dhi));
            X_lo = fmin(0.5, fmax(0.25, 0.5 * (1.0 - rhi)));
            X_hi = fmin(0.5, fmax(0.25, 0.5 * (1.0 - rlo)));
        }
    }
    double r1 = routed_kx(x, K_lo, X_lo, 1), r2 = routed_kx(x, K_lo, X_hi, 1);
    double r3 = routed_kx(x, K_hi, X_lo, 1), r4 = routed_kx(x, K_hi, X_hi, 1);
    double Rt_lo = fmin(fmin(r1, r2), fmin(r3, r4)), Rt_hi = fmax(fmax(r1, r2), fmax(r3, r4));
    double slack = 1e-12 * (fabs(x->qup) + fabs(x->quc) + fabs(x->qdp) + 2.0 * fabs(x->ql) + qmn_hi);
    *lo = Rt_lo - qmn_hi - slack;
    *hi = Rt_hi - qmn_lo + slack;
}

/* ---- mean-value enclosure --------------------------------------------------
 * The direct enclosure bounds the routed flow and Qmn separately, so its
 * width grows like (|R'| + |Qmn'|) (b - a) while F changes like |R' - Qmn'|:
 * where the two nearly cancel (a narrow band of F) it is far too wide.  The
 * mean-value form F(h) in F(m) + F'([a, b]) [a - m, b - m], m the midpoint,
 * is tight to second order.  F' = R_K K' + R_X X' - Qmn' by the chain rule,
 * every factor bounded by interval arithmetic over [a, b]:
 *   R_K = (dt/2) (X d + (1 - X)(qdp - S)) / D^2,  R_X = K (K d + dt (qup - qdp + ql)) / D^2,
 *     d = qup - quc, S = qup + quc - qdp + 2 ql, D = K (1 - X) + dt/2;
 *   K' = -dx ck' / ck^2 where dx / ck > dt (0 where K = dt),
 *   ck' = ks (g_rho rho' + g_w w'), g_rho = 10/9 rho^(-1/3) (1 - rho w), g_w = -2/3 rho^(5/3),
 *   X' = -(q0' den - q0 den') / (4 den^2) where the weighting is not clamped,
 *   Qmn' = sqrt(s0) (A' rho^(2/3) / n + A 2/3 rho^(-1/3) rho' / n - A rho^(2/3) n' / n^2);
 * at a kink (a clamp that may switch inside [a, b]) the range takes the hull
 * with 0, which keeps the mean-value bound (F is continuous within a regime).
 * Returns 0 where the celerity may vanish (K unbounded); the direct enclosure
 * alone is used there. */
typedef struct { double lo, hi; } iv_t;

static iv_t iv(double a, double b) { iv_t r = {fmin(a, b), fmax(a, b)}; return r; }
static iv_t iv_add(iv_t a, iv_t b) { return iv(a.lo + b.lo, a.hi + b.hi); }
static iv_t iv_sub(iv_t a, iv_t b) { return iv(a.lo - b.hi, a.hi - b.lo); }
static iv_t iv_mul(iv_t a, iv_t b)
{
    double p1 = a.lo * b.lo, p2 = a.lo * b.hi, p3 = a.hi * b.lo, p4 = a.hi * b.hi;
    iv_t r = {fmin(fmin(p1, p2), fmin(p3, p4)), fmax(fmax(p1, p2), fmax(p3, p4))};
    return r;
}
static iv_t iv_scale(iv_t a, double k) { return iv(a.lo * k, a.hi * k); }
static iv_t iv_inv(iv_t a) { return iv(1.0 / a.hi, 1.0 / a.lo); }   /* 0 not in a */
static iv_t iv_hull0(iv_t a) { return iv(fmin(a.lo, 0.0), fmax(a.hi, 0.0)); }
static iv_t iv_pow(iv_t a, double y) { return iv(orc_pow(a.lo, y), orc_pow(a.hi, y)); }   /* a > 0 */

static int fprime_enclosure(const route_ctx *x, double a, double b, int flood, iv_t *out)
{
    const chan_t *c = x->c;
    const double dt = x->dt, hdt = 0.5 * dt;
    const double sq = sqrt(1.0 + c->z * c->z), sqs0 = sqrt(c->s0), ks = sqs0 / c->n_ch;
    iv_t A, Ad, rho, rhod, n_eff, n_effd, twu, twud;
    if (!flood) {
        A = iv((c->bw + a * c->z) * a, (c->bw + b * c->z) * b);
        Ad = iv(c->bw + 2.0 * c->z * a, c->bw + 2.0 * c->z * b);
        double Pa = c->bw + 2.0 * a * sq, Pb_ = c->bw + 2.0 * b * sq;
        rho = iv(((c->bw + a * c->z) * a) / Pa, ((c->bw + b * c->z) * b) / Pb_);
        /* rho' = (bw^2 + 2 z h bw + 2 z h^2 sqrt(1+z^2)) / P^2, numerator rising */
        double na = c->bw * c->bw + 2.0 * c->z * a * c->bw + 2.0 * c->z * a * a * sq;
        double nb = c->bw * c->bw + 2.0 * c->z * b * c->bw + 2.0 * c->z * b * b * sq;
        rhod = iv(na / (Pb_ * Pb_), nb / (Pa * Pa));
        n_eff = iv(c->n_ch, c->n_ch);
        n_effd = iv(0.0, 0.0);
        twu = iv(c->bw + 2.0 * c->z * a, c->bw + 2.0 * c->z * b);
        twud = iv(2.0 * c->z, 2.0 * c->z);
    } else {
        double oa = fmax(a - c->bfd, 0.0), ob = fmax(b - c->bfd, 0.0);
        double Ab = (c->bw + c->bfd * c->z) * c->bfd, Pb0 = c->bw + 2.0 * c->bfd * sq;
        double Pa = Pb0 + c->twcc + 2.0 * oa, Pb_ = Pb0 + c->twcc + 2.0 * ob;
        A = iv(Ab + c->twcc * oa, Ab + c->twcc * ob);
        Ad = iv(c->twcc, c->twcc);
        rho = iv((Ab + c->twcc * oa) / Pa, (Ab + c->twcc * ob) / Pb_);
        double c0 = c->twcc * (Pb0 + c->twcc) - 2.0 * Ab;          /* rho' = c0 / P^2 */
        rhod = iv(c0 / (Pa * Pa), c0 / (Pb_ * Pb_));
        n_eff = iv((Pb0 * c->n_ch + (c->twcc + 2.0 * oa) * c->ncc) / Pa,
                   (Pb0 * c->n_ch + (c->twcc + 2.0 * ob) * c->ncc) / Pb_);
        double c1 = 2.0 * Pb0 * (c->ncc - c->n_ch);                  /* n' = c1 / P^2 */
        n_effd = iv(c1 / (Pa * Pa), c1 / (Pb_ * Pb_));
        twu = iv(c->twcc, c->twcc);
        twud = iv(0.0, 0.0);
    }
    iv_t r23 = iv_pow(rho, 2.0 / 3.0), r53 = iv_pow(rho, 5.0 / 3.0), rm13 = iv_pow(rho, -1.0 / 3.0);
    iv_t inv_n = iv_inv(n_eff);
    /* Qmn' */
    iv_t q1 = iv_mul(iv_mul(Ad, r23), inv_n);
    iv_t q2 = iv_mul(iv_mul(iv_mul(A, rm13), rhod), iv_scale(inv_n, 2.0 / 3.0));
    iv_t q3 = iv_mul(iv_mul(iv_mul(A, r23), n_effd), iv_mul(inv_n, inv_n));
    iv_t qmnd = iv_scale(iv_sub(iv_add(q1, q2), q3), sqs0);
    /* celerity and its derivative */
    double wa = 2.0 * sq / (c->bw + 2.0 * a * c->z), wb = 2.0 * sq / (c->bw + 2.0 * b * c->z);
    iv_t wv = iv(wb, wa);
    double ea = c->bw + 2.0 * a * c->z, eb = c->bw + 2.0 * b * c->z;
    iv_t wd = iv(-4.0 * sq * c->z / (ea * ea), -4.0 * sq * c->z / (eb * eb));
    iv_t g = iv_sub(iv_scale(r23, 5.0 / 3.0), iv_mul(iv_scale(r53, 2.0 / 3.0), wv));
    if (!(g.lo > 0.0)) return 0;
    iv_t ck = iv_scale(g, ks);
    iv_t g_rho = iv_mul(iv_scale(rm13, 10.0 / 9.0), iv_sub(iv(1.0, 1.0), iv_mul(rho, wv)));
    iv_t ckd = iv_scale(iv_add(iv_mul(g_rho, rhod), iv_mul(iv_scale(r53, -2.0 / 3.0), wd)), ks);
    /* storage constant */
    iv_t K, Kd;
    iv_t kd_branch = iv_scale(iv_mul(ckd, iv_inv(iv_mul(ck, ck))), -c->dx);
    if (c->dx / ck.hi >= dt) { K = iv(c->dx / ck.hi, c->dx / ck.lo); Kd = kd_branch; }
    else if (c->dx / ck.lo <= dt) { K = iv(dt, dt); Kd = iv(0.0, 0.0); }
    else { K = iv(dt, c->dx / ck.lo); Kd = iv_hull0(kd_branch); }
    const double d = x->qup - x->quc, S = x->qup + x->quc - x->qdp + 2.0 * x->ql;
    /* routed flow partials over a (K, X) box */
#define RK(Kv, Xv) iv_mul(iv_scale(iv_add(iv_scale(Xv, d), iv_scale(iv_sub(iv(1.0, 1.0), Xv), x->qdp - S)), hdt), \
                          iv_inv(iv_mul(iv_add(iv_mul(Kv, iv_sub(iv(1.0, 1.0), Xv)), iv(hdt, hdt)), \
                                        iv_add(iv_mul(Kv, iv_sub(iv(1.0, 1.0), Xv)), iv(hdt, hdt)))))
    /* weighting */
    iv_t half = iv(0.5, 0.5);
    double q0a = routed_kx(x, K.lo, 0.5, 0), q0b = routed_kx(x, K.hi, 0.5, 0);
    iv_t q0 = iv(q0a, q0b);
    iv_t q0d = iv_mul(RK(K, half), Kd);
    iv_t den = iv_scale(iv_mul(twu, ck), c->s0 * c->dx);
    iv_t dend = iv_scale(iv_add(iv_mul(twud, ck), iv_mul(twu, ckd)), c->s0 * c->dx);
    iv_t u = iv_sub(half, iv_mul(q0, iv_inv(iv_scale(den, 4.0))));
    iv_t ud = iv_scale(iv_mul(iv_sub(iv_mul(q0d, den), iv_mul(q0, dend)), iv_inv(iv_mul(den, den))), -0.25);
    iv_t X = iv(fmin(0.5, fmax(0.25, u.lo)), fmin(0.5, fmax(0.25, u.hi)));
    iv_t Xd;
    if (u.lo > 0.25 && u.hi < 0.5) Xd = ud;
    else if (u.lo >= 0.5 || u.hi <= 0.25) Xd = iv(0.0, 0.0);
    else Xd = iv_hull0(ud);
    iv_t Dx = iv_add(iv_mul(K, iv_sub(iv(1.0, 1.0), X)), iv(hdt, hdt));
    iv_t RX = iv_mul(iv_mul(K, iv_add(iv_scale(K, d), iv(dt * (x->qup - x->qdp + x->ql), dt * (x->qup - x->qdp + x->ql)))),
                     iv_inv(iv_mul(Dx, Dx)));
    iv_t Rd = iv_add(iv_mul(RK(K, X), Kd), iv_mul(RX, Xd));
#undef RK
    if (x->ql < 0.0) Rd = iv_hull0(Rd);        /* the C4 clamp may switch inside */
    iv_t Fd = iv_sub(Rd, qmnd);
    double mag = fmax(fabs(Fd.lo), fabs(Fd.hi));
    Fd.lo -= 1e-9 * mag + 1e-300;
    Fd.hi += 1e-9 * mag + 1e-300;
    *out = Fd;
    return isfinite(Fd.lo) && isfinite(Fd.hi);
}

/* The direct enclosure intersected with the mean-value one. */
static void f_enclosure_mv(const route_ctx *x, double a, double b, int flood, double *lo, double *hi)
{
    f_enclosure(x, a, b, flood, lo, hi);
    iv_t Fd;
    if (fprime_enclosure(x, a, b, flood, &Fd)) {
        double m = a + 0.5 * (b - a);
        double fm = residual_at(m, x->dt, x->qup, x->quc, x->qdp, x->ql, x->c, NULL);
        double r = 0.5 * (b - a) * fmax(fabs(Fd.lo), fabs(Fd.hi));
        double slack = 1e-12 * (fabs(x->qup) + fabs(x->quc) + fabs(x->qdp) + 2.0 * fabs(x->ql) + fabs(fm));
        *lo = fmax(*lo, fm - r - slack);
        *hi = fmin(*hi, fm + r + slack);
    }
}

/* The largest routed flow any depth can give: the four corners of
 * K in [dt, inf], X in [0.25, 0.5]. */
static double routed_max(const route_ctx *x)
{
    double r1 = routed_kx(x, x->dt, 0.25, 1), r2 = routed_kx(x, x->dt, 0.5, 1);
    double r3 = routed_kx(x, INFINITY, 0.25, 1), r4 = routed_kx(x, INFINITY, 0.5, 1);
    return fmax(fmax(r1, r2), fmax(r3, r4));
}

/* A lower bound of Qmn on [h, inf) (Qmn rises in the main channel; on the
 * floodplain it is at least A(h) min(R_h(h), twcc/2)^(2/3) sqrt(s0) / max n,
 * R_h tending to twcc/2). */
static double qmn_floor_above(double h, const chan_t *c)
{
    int flood = c->twcc > 0.0 && c->bfd > 0.01 && c->bfd < 1.0e4;
    double sqs0 = sqrt(c->s0);
    double nmax = fmax(c->n_ch, c->ncc);
    hbox_t fb = depth_end(fmax(h, c->bfd), 1, c);
    double flood_floor = fb.A * orc_pow(fmin(fb.R, 0.5 * c->twcc), 2.0 / 3.0) * sqs0 / nmax;
    if (!flood) {
        hbox_t e = depth_end(h, 0, c);
        return e.A * orc_pow(e.R, 2.0 / 3.0) * sqs0 / c->n_ch;
    }
    if (h >= c->bfd) return flood_floor;
    hbox_t e = depth_end(h, 0, c);
    return fmin(e.A * orc_pow(e.R, 2.0 / 3.0) * sqs0 / c->n_ch, flood_floor);
}

static double route_residual_fn(double h, const void *vctx)
{
    const route_ctx *x = (const route_ctx *)vctx;
    return residual_at(h, x->dt, x->qup, x->quc, x->qdp, x->ql, x->c, NULL);
}

#define CERT_RES 1e-6         /* relative depth resolution of the scan */
#define CERT_HMAX 1.0e4       /* depths searched: [mindepth, CERT_HMAX] */

static long long g_cert_steps;  /* enclosures evaluated (diagnostics, racy under threads) */

typedef struct {
    const route_ctx *x;
    int has_flood;
    double bfd;
    long long steps;
} scan_t;

/* Bounds of F on [a, b] from F at its ends and F' in [sl, sh]: for t = h - a
 * in [0, w], F(h) <= min(F(a) + sh t, F(b) - sl (w - t)) and F(h) >=
 * max(F(a) + sl t, F(b) - sh (w - t)); a minimum of two lines is concave, a
 * maximum convex, so their extremes lie at t = 0, t = w or where the lines
 * cross.  have_fa = 0 drops the left line (an interval that starts at
 * bankfull on the floodplain, where F(a) is the main channel's value). */
static void slope_bounds(double w, double fa, int have_fa, double fb, double sl, double sh,
                         double *lo, double *hi)
{
    double u0 = fb - sl * w, uw = fb, l0 = fb - sh * w, lw = fb;
    if (have_fa) {
        u0 = fmin(u0, fa); uw = fmin(uw, fa + sh * w);
        l0 = fmax(l0, fa); lw = fmax(lw, fa + sl * w);
    }
    double up = fmax(u0, uw), dn = fmin(l0, lw);
    if (have_fa && sh > sl) {
        double t = (fb - sl * w - fa) / (sh - sl);            /* fa + sh t == fb - sl (w - t) */
        if (t > 0.0 && t < w) up = fmax(up, fa + sh * t);
        double t2 = (fb - sh * w - fa) / (sl - sh);           /* fa + sl t == fb - sh (w - t) */
        if (t2 > 0.0 && t2 < w) dn = fmin(dn, fa + sl * t2);
    }
    *lo = dn;
    *hi = up;
}

/* Does F keep its sign (pos: F > 0, else F <= 0) on [a, b] (not crossing
 * bankfull)?  fa = F(a) (unused when !have_fa), fb = F(b), both of that sign.
 * The direct enclosure intersected with the slope bounds; *sl, *sh receive
 * the slope enclosure (+-inf when none is formed). */
static int sign_holds(scan_t *S, double a, double b, int flood, double fa, int have_fa, double fb, int pos)
{
    const route_ctx *x = S->x;
    double elo, ehi;
    ++S->steps;
    f_enclosure(x, a, b, flood, &elo, &ehi);
    if (pos ? elo > 0.0 : !(ehi > 0.0)) return 1;
    iv_t Fd;
    if (!fprime_enclosure(x, a, b, flood, &Fd)) return 0;
    double lo, hi;
    slope_bounds(b - a, fa, have_fa, fb, Fd.lo, Fd.hi, &lo, &hi);
    double slack = 1e-12 * (fabs(x->qup) + fabs(x->quc) + fabs(x->qdp) + 2.0 * fabs(x->ql) + fabs(fb));
    return pos ? lo - slack > 0.0 : !(hi + slack > 0.0);
}

/* The next step after [a, b] (width w) held its sign, F(a) = fa, F(b) = fb:
 * where F moves toward 0, 0.95 of the distance at which the average slope
 * over [a, b] would reach it (a Newton step that stops short); otherwise,
 * or beyond that, four times w. */
static double next_step(double w, double fa, double fb)
{
    double toward = fabs(fb) < fabs(fa) ? (fabs(fa) - fabs(fb)) / w : 0.0;
    double wn = (toward > 0.0) ? 0.95 * fabs(fb) / toward : 4.0 * w;
    return fmin(4.0 * w, wn);
}

/* Root of F in [l, r] with F(l) > 0 >= F(r), to two ulps: regula falsi with
 * the Illinois modification, bisecting whenever a step fails to halve the
 * bracket. */
static double bracket_root(const route_ctx *x, double l, double fl, double r, double fr)
{
    int side = 0;
    double width = r - l;
    for (int i = 0; i < 400; ++i) {
        double m = (fr - fl != 0.0) ? (l * fr - r * fl) / (fr - fl) : l + (r - l) / 2.0;
        if ((i % 3) == 2 && r - l > 0.5 * width) m = l + (r - l) / 2.0;
        if ((i % 3) == 2) width = r - l;
        if (!(m > l && m < r)) m = l + (r - l) / 2.0;
        if (!(m > l && m < r)) break;
        double fm = route_residual_fn(m, x);
        if (fm > 0.0) {
            l = m; fl = fm;
            if (side == 1) fr *= 0.5;
            side = 1;
        } else {
            r = m; fr = fm;
            if (side == -1) fl *= 0.5;
            side = -1;
        }
    }
    return l + (r - l) / 2.0;
}

/* Generic certified scan from a, where F's sign is pos and its value fa.  It
 * keeps a point a whose sign is known, evaluates F at b = a + w (never
 * crossing bankfull), and:
 *   - F(b) of the other sign: a sign change lies in (a, b]; at the resolution
 *     (w <= CERT_RES a) b is taken as the first point of the other sign,
 *     else b moves back toward the secant estimate of the crossing;
 *   - F(b) of the same sign and the sign shown to hold on all of [a, b]:
 *     a moves to b, and the next step is 0.9 of the distance the slope bound
 *     lets F travel toward 0 (at most 4 w);
 *   - otherwise the step halves, and at the resolution a moves to b.
 * Looking for F > 0 it stops when Qmn on [a, inf) exceeds every routed flow
 * (F < 0 beyond: no root, the answer is lo); looking for F <= 0, the first
 * such point b closes a bracket [a, b] solved by bracket_root. */
static double scan_generic(scan_t *S, double a, double fa, int pos, double lo)
{
    const route_ctx *x = S->x;
    const double rmax = routed_max(x);
    double w = a;
    for (;;) {
        if (!pos && qmn_floor_above(a, x->c) > rmax + 1e-12 * (fabs(rmax) + 1.0)) return lo;
        if (a >= CERT_HMAX) return pos ? CERT_HMAX : lo;
        double b = fmin(a + w, CERT_HMAX);
        if (S->has_flood && a < S->bfd && b > S->bfd) b = S->bfd;
        const int flood = S->has_flood && a >= S->bfd;
        const int have_fa = !(flood && a <= S->bfd);
        double fb = route_residual_fn(b, x);
        if (pos ? !(fb > 0.0) : fb > 0.0) {          /* a sign change in (a, b] */
            if (b - a <= CERT_RES * a) {
                if (pos) return bracket_root(x, a, fa, b, fb);
                pos = 1; a = b; fa = fb;
                continue;
            }
            double t = (have_fa && fa != fb) ? (b - a) * fa / (fa - fb) : 0.5 * (b - a);
            w = fmax(0.5 * CERT_RES * a, fmin(0.5 * (b - a), 0.98 * t));
            continue;
        }
        if (sign_holds(S, a, b, flood, fa, have_fa, fb, pos)) {
            w = fmax(CERT_RES * b, have_fa ? next_step(b - a, fa, fb) : 4.0 * (b - a));
            a = b; fa = fb;
            continue;
        }
        if (b - a > CERT_RES * a) { w = 0.5 * (b - a); continue; }
        a = b; fa = fb;                                /* at the resolution: move on */
        w = 2.0 * w;
    }
}

/* Certified leftmost downcrossing of F on [lo, CERT_HMAX].
 * F(lo) <= 0: the generic scan (first a positive point, then the first
 * non-positive one).  F(lo) > 0, the usual case: a candidate root first (F
 * at lo 2^k, visiting bankfull, up to the first F <= 0; then bracket_root),
 * then a certificate that F > 0 on [lo, candidate) with steps that are a
 * fraction of the remaining distance; if a step cannot be certified even at
 * the resolution, F has another sign change first and the generic scan takes
 * over from the last certified point. */
static double certified_leftmost_downcrossing(const route_ctx *x, double lo)
{
    const chan_t *c = x->c;
    scan_t S = {x, c->twcc > 0.0 && c->bfd > lo && c->bfd < CERT_HMAX, c->bfd, 0};
    double root;
    double flo = route_residual_fn(lo, x);
    if (!(flo > 0.0)) {
        root = scan_generic(&S, lo, flo, 0, lo);
        g_cert_steps += S.steps;
        return root;
    }
    /* candidate */
    double p = lo, fp = flo, q = lo, fq = flo;
    int used_bfd = !S.has_flood;
    while (fq > 0.0 && q < CERT_HMAX) {
        p = q; fp = fq;
        q = fmin(2.0 * q, CERT_HMAX);
        if (!used_bfd && q >= S.bfd) { q = S.bfd; used_bfd = 1; }
        fq = route_residual_fn(q, x);
    }
    if (fq > 0.0) { g_cert_steps += S.steps; return CERT_HMAX; }
    double r = bracket_root(x, p, fp, q, fq);
    /* certificate on [lo, r) */
    double a = lo, fa = flo, beta = 0.5;
    for (;;) {
        double d = r - a;
        if (d <= CERT_RES * a) break;
        double b = a + beta * d;
        if (S.has_flood && a < S.bfd && b > S.bfd) b = S.bfd;
        const int flood = S.has_flood && a >= S.bfd;
        double fb = route_residual_fn(b, x);
        if (!(fb > 0.0)) {                            /* another sign change first */
            r = scan_generic(&S, a, fa, 1, lo);
            break;
        }
        const int have_fa = !(flood && a <= S.bfd);
        if (sign_holds(&S, a, b, flood, fa, have_fa, fb, 1)) {
            double wn = have_fa ? next_step(b - a, fa, fb) : 4.0 * (b - a);
            a = b;
            fa = fb;
            beta = fmin(0.95, fmax(0.05, wn / (r - a)));
        } else {
            beta *= 0.5;
            if (beta * d <= CERT_RES * a) {
                r = scan_generic(&S, a, fa, 1, lo);
                break;
            }
        }
    }
    g_cert_steps += S.steps;
    return r;
}

long long orc_cert_steps(void) { return g_cert_steps; }

void orc_residual_enclosure(double dt, double qup, double quc, double qdp, double ql,
                            const orc_chparams_f64 *p, double a, double b, double *lo, double *hi)
{
    chan_t c;
    chan_init(p, &c);
    route_ctx x = {dt, qup, quc, qdp, ql, &c};
    int flood = c.twcc > 0.0 && c.bfd > 0.01 && c.bfd < CERT_HMAX && a >= c.bfd;
    f_enclosure_mv(&x, a, b, flood, lo, hi);
}

int orc_residual_slope_enclosure(double dt, double qup, double quc, double qdp, double ql,
                                 const orc_chparams_f64 *p, double a, double b, double *lo, double *hi)
{
    chan_t c;
    chan_init(p, &c);
    route_ctx x = {dt, qup, quc, qdp, ql, &c};
    int flood = c.twcc > 0.0 && c.bfd > 0.01 && c.bfd < CERT_HMAX && a >= c.bfd;
    iv_t Fd;
    if (!fprime_enclosure(&x, a, b, flood, &Fd)) return 0;
    *lo = Fd.lo;
    *hi = Fd.hi;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Canonical root of the converged rules.  No reference counterpart.
 *
 * Two solvers of one F that start from different points stop at different
 * depths inside the band where the rounding noise of F hides its sign
 * (about 1e-14 relative); the routing carries such differences on.  Both
 * converged rules therefore end a solve whose root is a downcrossing in its
 * canonical cell, decided by the signs of one fixed restatement Fc of F:
 *   grid   the depths whose binary significand has its low 28 bits zero
 *          (spacing 2^-24 of the depth's binade);
 *   cell   neighbouring grid points g < g+ with Fc(g) > 0 >= Fc(g+), found
 *          from the grid point at or below the root estimate, one point at a
 *          time (at most CANON_MOVES moves; F > 0 at both ends moves up,
 *          F <= 0 at the lower end moves down);
 *   root   t = Fc(g) / (Fc(g) - Fc(g+)), h* = g + (g+ - g) t, outflow
 *          Q(g) + (Q(g+) - Q(g)) t with Q the outflow rule of Fc's
 *          coefficients; a cell holding the floodplain's jump at bankfull is
 *          first cut there (Fc(bankfull) <= 0 keeps [g, bankfull]; else the
 *          next double above bankfull is the root if Fc <= 0 there, or the
 *          cell's new lower end).
 * Fc is the device's form of the residual (two divisions, the powers
 * R^(2/3), R^(5/3) from an integer estimate of R^(-1/3) and four Newton
 * steps): every operation IEEE, so its bits are the same on every machine.  The certified scan above (on residual_at, the
 * reference's form) only supplies the estimate. */
#define CANON_BITS 28
#define CANON_MOVES 4

static double canon_floor(double r)
{
    uint64_t b;
    memcpy(&b, &r, 8);
    b &= ~((UINT64_C(1) << CANON_BITS) - 1);
    memcpy(&r, &b, 8);
    return r;
}

static double canon_step(double g, int up)
{
    uint64_t b;
    memcpy(&b, &g, 8);
    b = up ? b + (UINT64_C(1) << CANON_BITS) : b - (UINT64_C(1) << CANON_BITS);
    memcpy(&g, &b, 8);
    return g;
}

/* the next double above x > 0 */
static double canon_step_ulp(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    b += 1;
    memcpy(&x, &b, 8);
    return x;
}

/* R^(2/3), R^(5/3) within two ulps: hi(s) = 0x553f1348 - hi(x) / 3 gives
 * x^(-1/3) to 4.1e-2, four Newton steps s += (s/3)(1 - x s^3) to 1.6e-18;
 * outside [2^-1000, 2^1000] the correctly rounded pow. */
static void pow2353_fast(double x, double *p23, double *p53)
{
    if (!(x >= 0x1p-1000 && x <= 0x1p1000)) {
        *p23 = orc_pow(x, 2.0 / 3.0);
        *p53 = orc_pow(x, 5.0 / 3.0);
        return;
    }
    uint64_t bx;
    memcpy(&bx, &x, 8);
    uint32_t hs = 0x553f1348u - (uint32_t)(bx >> 32) / 3u;
    uint64_t bs = (uint64_t)hs << 32;
    double s;
    memcpy(&s, &bs, 8);
    for (int i = 0; i < 4; ++i) {
        double e = fma(-x * s, s * s, 1.0);
        s = fma(s * (1.0 / 3.0), e, s);
    }
    *p23 = x * s;
    *p53 = *p23 * x;
}

/* Fc(h) and the outflow of its coefficients: the reference's formulas
 * (wavefront_mc.cu:200-213, :249-328) in the device's arrangement -- the
 * hydraulic radius and the Manning factor share 1 / (P (P_ch n + P_fp ncc)),
 * the celerity is carried as cn = ck twl and the storage constant as
 * K cn = max(dt cn, dx twl), and the weighting X = 0.5 - u / v as a ratio
 * (clamped by comparing u and v), so C1..C4 take the one division
 * 1 / (cn xd (K (1 - X) + dt/2)). */
static double canon_residual(double h, double dt, double qup, double quc, double qdp, double ql,
                             const chan_t *c, double *qout)
{
    const double sq = sqrt(1.0 + c->z * c->z), sqs0 = sqrt(c->s0), ks = sqrt(c->s0) / c->n_ch;
    const double twl = c->bw + 2.0 * c->z * h;
    double over = fmax(h - c->bfd, 0.0);
    double in = fmin(c->bfd, h);
    if (over > 0.0 && c->twcc <= 0.0) { over = 0.0; in = h; }
    const double asum = (c->bw + in * c->z) * in + c->twcc * over;
    const double wpc = c->bw + 2.0 * in * sq;
    const double wpf = (over > 0.0) ? (c->twcc + 2.0 * over) : 0.0;
    const double wp = wpc + wpf;
    const double nsum = wpc * c->n_ch + wpf * c->ncc;
    double R = 0.0, mq = 0.0;
    if (wp > 0.0) {
        const double inv = 1.0 / (wp * nsum);
        R = asum * nsum * inv;
        mq = wp * wp * inv;
    }
    double r23, r53;
    pow2353_fast(R, &r23, &r53);
    const double qmn = asum * r23 * sqs0 * mq;
    double cn = 0.0;
    if (h > 0.0) cn = fmax(0.0, ks * ((5.0 / 3.0) * r23 * twl - (4.0 / 3.0) * r53 * sq));
    const double hdt = 0.5 * dt;
    const double cs = (cn > 0.0) ? cn : 1.0;
    const double P = (cn > 0.0) ? fmax(dt * cn, c->dx * twl) : dt;
    const double b = hdt * cs;
    /* X = xn / xd, 1 - X = omx / xd; X = 0.5 - u / v clamped by comparing u, v */
    double xn = 0.5, omx = 0.5, xd = 1.0;
    if (cn > 0.0) {
        const double a = 0.5 * P;
        const double n0 = (a + b) * qup + (b - a) * quc + (a - b) * qdp + ql * dt * cn;
        const double twu = (h > c->bfd && c->twcc > 0.0) ? c->twcc : twl;
        const double sdx = twu * c->s0 * c->dx;
        if (sdx > 0.0) {
            const double u = n0 * twl, v = (a + b) * 4.0 * sdx * cn;
            if (u >= 0.25 * v) { xn = 0.25; omx = 0.75; }
            else if (u > 0.0) { xn = 0.5 * v - u; omx = 0.5 * v + u; xd = v; }
        }
    }
    double D = P * omx + b * xd;
    if (D == 0.0) D = 1.0;
    const double iD = 1.0 / D;
    mcoef_t C;
    C.c1 = (P * xn + b * xd) * iD;
    C.c2 = (b * xd - P * xn) * iD;
    C.c3 = (P * omx - b * xd) * iD;
    C.c4 = (ql * dt * cs * xd) * iD;
    const double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
    if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
    if (qout) *qout = outflow(&C, qup, quc, qdp);
    return (routed + C.c4) - qmn;
}

static long long g_noncanon = 0;
long long orc_noncanon_count(void) { return g_noncanon; }
void orc_noncanon_reset(void) { g_noncanon = 0; }

/* The canonical root next to the estimate r: 1 with *h, *q set, 0 when r is
 * kept (below mindepth's cell, at the search's end, or no cell within the
 * moves). */
static int canonical_root(const route_ctx *x, double r, double *h, double *q)
{
    double a = canon_floor(r), b = canon_step(a, 1);
    if (!(a >= 0.01 && b <= CERT_HMAX)) return 0;
    double qa, qb;
    double fa = canon_residual(a, x->dt, x->qup, x->quc, x->qdp, x->ql, x->c, &qa);
    double fb = canon_residual(b, x->dt, x->qup, x->quc, x->qdp, x->ql, x->c, &qb);
    for (int moves = 0;; ++moves) {
        if (fa > 0.0 && !(fb > 0.0)) {
            const chan_t *c = x->c;
            if (c->twcc > 0.0 && a <= c->bfd && c->bfd < b) {
                /* the floodplain's jump inside the cell: Fc is continuous on
                 * [a, bankfull] and on (bankfull, b] */
                double fm = fa, qm = qa;
                if (a < c->bfd) fm = canon_residual(c->bfd, x->dt, x->qup, x->quc, x->qdp, x->ql, c, &qm);
                if (!(fm > 0.0)) {
                    b = c->bfd; fb = fm; qb = qm;
                } else {
                    double up = canon_step_ulp(c->bfd), qp;
                    double fp = canon_residual(up, x->dt, x->qup, x->quc, x->qdp, x->ql, c, &qp);
                    if (!(fp > 0.0)) { *h = up; *q = qp; return 1; }
                    a = up; fa = fp; qa = qp;
                }
            }
            double t = fa / (fa - fb);
            *h = a + (b - a) * t;
            *q = qa + (qb - qa) * t;
            return 1;
        }
        if (moves == CANON_MOVES) return 0;
        if (fa > 0.0) {
            a = b; fa = fb; qa = qb;
            b = canon_step(b, 1);
            if (!(b <= CERT_HMAX)) return 0;
            fb = canon_residual(b, x->dt, x->qup, x->quc, x->qdp, x->ql, x->c, &qb);
        } else {
            b = a; fb = fa; qb = qa;
            a = canon_step(a, 0);
            if (!(a >= 0.01)) return 0;
            fa = canon_residual(a, x->dt, x->qup, x->quc, x->qdp, x->ql, x->c, &qa);
        }
    }
}

double orc_canonical_root(double dt, double qup, double quc, double qdp, double ql,
                          const orc_chparams_f64 *p, double r, double *qout)
{
    chan_t c;
    chan_init(p, &c);
    route_ctx x = {dt, qup, quc, qdp, ql, &c};
    double h, q;
    if (!canonical_root(&x, r, &h, &q)) return NAN;
    if (qout) *qout = q;
    return h;
}

double orc_canon_residual(double dt, double qup, double quc, double qdp, double ql,
                          const orc_chparams_f64 *p, double h, double *qout)
{
    chan_t c;
    chan_init(p, &c);
    return canon_residual(h, dt, qup, quc, qdp, ql, &c, qout);
}

void orc_pow2353_fast(double x, double *p23, double *p53) { pow2353_fast(x, p23, p53); }

/* The converged depth: the certified leftmost downcrossing of F, ended in
 * its canonical cell; mindepth (F positive nowhere) with the outflow of Fc's
 * coefficients there. */
double orc_mc_solve_converged(double dt, double qup, double quc, double qdp, double ql,
                              const orc_chparams_f64 *p, double depthp,
                              double *depthc, double *velc)
{
    chan_t c;
    mcoef_t C;
    (void)depthp;
    chan_init(p, &c);
    if (chan_invalid(&c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0)) {
        *depthc = 0.0; *velc = 0.0;
        return 0.0;
    }
    route_ctx x = {dt, qup, quc, qdp, ql, &c};
    double h = certified_leftmost_downcrossing(&x, 0.01), q;
    if (h == 0.01) {
        canon_residual(h, dt, qup, quc, qdp, ql, &c, &q);
    } else if (!canonical_root(&x, h, &h, &q)) {
        __atomic_fetch_add(&g_noncanon, 1, __ATOMIC_RELAXED);
        residual_at(h, dt, qup, quc, qdp, ql, &c, &C);
        q = outflow(&C, qup, quc, qdp);
    }
    *velc = velocity(h, &c);
    *depthc = h;
    return q;
}

double orc_manning_flow(double h, const orc_chparams_f64 *p)
{
    chan_t c;
    geo_t g;
    chan_init(p, &c);
    geometry(h, &c, &g);
    return manning(&g, &c);
}

typedef struct { double Q; const chan_t *c; } nd_ctx;

static double normal_residual_fn(double h, const void *vctx)
{
    const nd_ctx *x = (const nd_ctx *)vctx;
    geo_t g;
    geometry(h, x->c, &g);
    return x->Q - manning(&g, x->c);
}

double orc_normal_depth(double Q, const orc_chparams_f64 *p)
{
    chan_t c;
    chan_init(p, &c);
    if (!(Q > 0.0) || chan_invalid(&c)) return 0.0;
    nd_ctx x = {Q, &c};
    double bp = (c.twcc > 0.0) ? c.bfd : -1.0;
    return leftmost_downcrossing(normal_residual_fn, &x, 1e-6, bp);
}

/* ------------------------------------------------------------------------ */
/* Accuracy statistics, wavefront_mc.cu:566-572 (pct) and :822-858.          */

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static double pct_sorted(const double *v, size_t n, double p)
{
    if (n == 0) return 0.0;
    size_t k = (size_t)((p / 100.0) * (double)(n - 1));
    return v[k];
}

void orc_compute_stats(const double *pred, const double *truth, int n, double *out)
{
    double *rel = (double *)malloc(((size_t)n + 1) * sizeof(double));
    size_t m = 0, w1 = 0, w10 = 0;
    int nans = 0;
    double max_abs = 0.0;
    for (int i = 0; i < n; ++i) {
        double g = pred[i], t = truth[i];
        if (isnan(g) || isinf(g)) { ++nans; continue; }
        double ae = fabs(g - t);
        max_abs = fmax(max_abs, ae);
        if (fabs(t) > 1e-3) {
            double re = ae / fabs(t);
            rel[m++] = re;
            if (re < 0.01) ++w1;
            if (re < 0.1) ++w10;
        }
    }
    qsort(rel, m, sizeof(double), cmp_double);
    out[0] = max_abs;
    out[1] = pct_sorted(rel, m, 50.0);
    out[2] = pct_sorted(rel, m, 90.0);
    out[3] = pct_sorted(rel, m, 99.0);
    out[4] = pct_sorted(rel, m, 99.9);
    out[5] = m ? rel[m - 1] : 0.0;
    out[6] = m ? (double)w1 / (double)m : 1.0;
    out[7] = m ? (double)w10 / (double)m : 1.0;
    out[8] = (double)nans;
    free(rel);
}


/* ------------------------------------------------------------ linearized MC */

/* x^y in float as the float rounding of the double pow. */
static float powf_from_double(float x, double y)
{
    return (float)pow((double)x, y);
}

/* wavefront/linear_mc.cu:57-100: coefficients at the reference depth
 * h = max(0.5 bankfull, 0.1), weighting X from the Cunge formula with Qj = 0.
 * The reference evaluates 2.0f / 3.0f and 5.0f / 3.0f in float. */
static void linear_coeffs(const orc_chparams_f32 *params, int n, float dt, int fixed_x,
                          float *C1, float *C2, float *C3, float *C4oq)
{
    for (int i = 0; i < n; ++i) {
        const orc_chparams_f32 *p = &params[i];
        float z = (p->cs == 0.0f) ? 1.0f : 1.0f / p->cs;
        float bfd;
        if (p->bw > p->tw) bfd = p->bw / 0.00001f;
        else if (p->bw == p->tw) bfd = p->bw / (2.0f * z);
        else bfd = (p->tw - p->bw) / (2.0f * z);
        float h = fmaxf(0.5f * bfd, 0.1f);
        float twl = p->bw + 2.0f * z * h;
        float A = (p->bw + h * z) * h;
        float WP = p->bw + 2.0f * h * sqrtf(1.0f + z * z);
        float R = (WP > 0.0f) ? A / WP : 0.0f;
        float r23 = powf_from_double(R, (double)(2.0f / 3.0f));
        float r53 = powf_from_double(R, (double)(5.0f / 3.0f));
        float Ck = fmaxf(1e-6f, (sqrtf(p->s0) / p->n_ch) *
                                    ((5.0f / 3.0f) * r23 - (2.0f / 3.0f) * r53 *
                                     (2.0f * sqrtf(1.0f + z * z) / (p->bw + 2.0f * h * z))));
        float Km = fmaxf(dt, p->dx / Ck);
        float tw_use = (h > bfd && p->twcc > 0.0f) ? p->twcc : twl;
        float X = 0.25f;
        if (!fixed_x && Ck > 0.0f && tw_use * p->s0 * Ck * p->dx > 0.0f)
            X = fminf(0.5f, fmaxf(0.0f, 0.5f * (1.0f - 0.0f / (2.0f * tw_use * p->s0 * Ck * p->dx))));
        float D = Km * (1.0f - X) + dt / 2.0f;
        if (D == 0.0f) D = 1.0f;
        C1[i] = (Km * X + dt / 2.0f) / D;
        C2[i] = (dt / 2.0f - Km * X) / D;
        C3[i] = (Km * (1.0f - X) - dt / 2.0f) / D;
        C4oq[i] = dt / D;
    }
}

void orc_linear_coeffs(const orc_chparams_f32 *params, int n, float dt,
                       float *C1, float *C2, float *C3, float *C4oq)
{
    linear_coeffs(params, n, dt, 0, C1, C2, C3, C4oq);
}

/* wavefront/linear_mc_matrix.cu:58-84 (compute_mc_coeffs_cpu): the same
 * arithmetic with the weighting fixed at X = 0.25. */
void orc_matrix_coeffs(const orc_chparams_f32 *params, int n, float dt,
                       float *C1, float *C2, float *C3, float *C4oq)
{
    linear_coeffs(params, n, dt, 1, C1, C2, C3, C4oq);
}

/* ---- matrix form: propagation matrices (linear_mc_matrix.cu:118-229) ----
 * Rows as sorted (column, value) lists.  Row r: the upstream rows merged one
 * after another into an initially empty list, each scaled by C1[r] (a column
 * new to the list takes the product, a present one adds it), pruned (drop
 * |v| < tol), then C2[r] added at each upstream column and C3[r] (P_Q) or
 * C4/ql[r] (P_q) at column r (inserted when absent), pruned again. */
typedef struct { int n, cap; int *col; float *val; } orc_row;

static void row_reserve(orc_row *r, int need)
{
    if (need <= r->cap) return;
    int cap = r->cap ? r->cap : 4;
    while (cap < need) cap *= 2;
    r->col = (int *)realloc(r->col, sizeof(int) * (size_t)cap);
    r->val = (float *)realloc(r->val, sizeof(float) * (size_t)cap);
    r->cap = cap;
}

/* dst = merge(dst, scale * src), both sorted by column */
static void row_merge_scaled(orc_row *dst, const orc_row *src, float scale, orc_row *tmp)
{
    tmp->n = 0;
    row_reserve(tmp, dst->n + src->n);
    int i = 0, j = 0;
    while (i < dst->n && j < src->n) {
        if (dst->col[i] < src->col[j]) {
            tmp->col[tmp->n] = dst->col[i]; tmp->val[tmp->n++] = dst->val[i++];
        } else if (dst->col[i] > src->col[j]) {
            tmp->col[tmp->n] = src->col[j]; tmp->val[tmp->n++] = src->val[j] * scale; j++;
        } else {
            tmp->col[tmp->n] = dst->col[i]; tmp->val[tmp->n++] = dst->val[i] + src->val[j] * scale;
            i++; j++;
        }
    }
    while (i < dst->n) { tmp->col[tmp->n] = dst->col[i]; tmp->val[tmp->n++] = dst->val[i++]; }
    while (j < src->n) { tmp->col[tmp->n] = src->col[j]; tmp->val[tmp->n++] = src->val[j] * scale; j++; }
    orc_row t = *dst; *dst = *tmp; *tmp = t;
}

static void row_add_col(orc_row *r, int col, float v)
{
    int lo = 0, hi = r->n;
    while (lo < hi) { int m = (lo + hi) / 2; if (r->col[m] < col) lo = m + 1; else hi = m; }
    if (lo < r->n && r->col[lo] == col) { r->val[lo] += v; return; }
    row_reserve(r, r->n + 1);
    memmove(r->col + lo + 1, r->col + lo, sizeof(int) * (size_t)(r->n - lo));
    memmove(r->val + lo + 1, r->val + lo, sizeof(float) * (size_t)(r->n - lo));
    r->col[lo] = col; r->val[lo] = v; r->n++;
}

static void row_prune(orc_row *r, float tol)
{
    int k = 0;
    for (int i = 0; i < r->n; i++)
        if (!(fabsf(r->val[i]) < tol)) { r->col[k] = r->col[i]; r->val[k++] = r->val[i]; }
    r->n = k;
}

static int rows_to_csr(orc_row *rows, int n, int **rowptr, int **col, float **val)
{
    size_t nnz = 0;
    for (int r = 0; r < n; r++) nnz += (size_t)rows[r].n;
    *rowptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    *col = (int *)malloc(sizeof(int) * (nnz ? nnz : 1));
    *val = (float *)malloc(sizeof(float) * (nnz ? nnz : 1));
    size_t p = 0;
    (*rowptr)[0] = 0;
    for (int r = 0; r < n; r++) {
        memcpy(*col + p, rows[r].col, sizeof(int) * (size_t)rows[r].n);
        memcpy(*val + p, rows[r].val, sizeof(float) * (size_t)rows[r].n);
        p += (size_t)rows[r].n;
        (*rowptr)[r + 1] = (int)p;
        free(rows[r].col); free(rows[r].val);
    }
    free(rows);
    return (int)nnz;
}

int orc_matrix_build(int n, const int *up_ptr, const int *up_idx,
                     const float *C1, const float *C2, const float *C3, const float *C4oq, float tol,
                     int **rpQ, int **colQ, float **valQ, int **rpq, int **colq, float **valq, int *nnz)
{
    orc_row *Q = (orc_row *)calloc((size_t)(n ? n : 1), sizeof(orc_row));
    orc_row *q = (orc_row *)calloc((size_t)(n ? n : 1), sizeof(orc_row));
    orc_row tmp = {0, 0, NULL, NULL};
    for (int r = 0; r < n; r++) {
        for (int k = up_ptr[r]; k < up_ptr[r + 1]; k++) {
            const int u = up_idx[k];
            if (u < 0 || u >= r) return -1;   /* not an earlier reach */
            row_merge_scaled(&Q[r], &Q[u], C1[r], &tmp);
            row_merge_scaled(&q[r], &q[u], C1[r], &tmp);
        }
        row_prune(&Q[r], tol);
        row_prune(&q[r], tol);
        for (int k = up_ptr[r]; k < up_ptr[r + 1]; k++) row_add_col(&Q[r], up_idx[k], C2[r]);
        row_add_col(&Q[r], r, C3[r]);
        row_add_col(&q[r], r, C4oq[r]);
        row_prune(&Q[r], tol);
        row_prune(&q[r], tol);
    }
    free(tmp.col); free(tmp.val);
    nnz[0] = rows_to_csr(Q, n, rpQ, colQ, valQ);
    nnz[1] = rows_to_csr(q, n, rpq, colq, valq);
    return 0;
}

void orc_free(void *p) { free(p); }

/* Steps of the matrix form, y = P_Q x + P_q ql_t, in float with a fixed
 * summation order over each row's union of its two CSR rows (columns
 * ascending; a matrix without an entry at a column contributes 0 * value):
 * with G lanes per row, partial l (0 <= l < G) takes union entries l, l + G,
 * ... and adds, per entry, s = s + vQ * x[c] then s = s + vq * ql[c] (no
 * fused multiply-add); the partials are combined as p[l] = p[l] + p[l ^ o]
 * for o = G/2, ..., 1 and the row's value is p[0].  G = 1 is the sequential
 * row sum; G = 0 takes each row's G from its union length U: the smallest
 * power of two with 2 G >= U, at most 64.  Q [n] holds Q0 on entry and the
 * flow after T steps on return. */
void orc_matrix_route(int n, const int *rpQ, const int *colQ, const float *valQ,
                      const int *rpq, const int *colq, const float *valq,
                      const float *qlat, float *Q, int T, int Gfix)
{
    float *y = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
    int ucap = 64;
    int *uc = (int *)malloc(sizeof(int) * (size_t)ucap);
    float *uq = (float *)malloc(sizeof(float) * (size_t)ucap), *uv = (float *)malloc(sizeof(float) * (size_t)ucap);
    float p[64], p2[64];
    for (int t = 0; t < T; t++) {
        const float *ql = qlat + (size_t)t * n;
        for (int r = 0; r < n; r++) {
            /* union of the row's entries, ascending columns */
            int a = rpQ[r], b = rpq[r], U = 0;
            const int need = (rpQ[r + 1] - a) + (rpq[r + 1] - b);
            if (need > ucap) {
                ucap = need;
                uc = (int *)realloc(uc, sizeof(int) * (size_t)ucap);
                uq = (float *)realloc(uq, sizeof(float) * (size_t)ucap);
                uv = (float *)realloc(uv, sizeof(float) * (size_t)ucap);
            }
            while (a < rpQ[r + 1] || b < rpq[r + 1]) {
                const int ca = a < rpQ[r + 1] ? colQ[a] : 0x7fffffff, cb = b < rpq[r + 1] ? colq[b] : 0x7fffffff;
                const int c = ca < cb ? ca : cb;
                uc[U] = c;
                uq[U] = (ca == c) ? valQ[a++] : 0.0f;
                uv[U] = (cb == c) ? valq[b++] : 0.0f;
                U++;
            }
            int G = Gfix;
            if (G == 0)
                for (G = 1; G < 64 && G * 2 < U; G *= 2) {}
            for (int l = 0; l < G; l++) {
                float s = 0.0f;
                for (int k = l; k < U; k += G) {
                    s = s + uq[k] * Q[uc[k]];
                    s = s + uv[k] * ql[uc[k]];
                }
                p[l] = s;
            }
            for (int o = G / 2; o > 0; o >>= 1) {
                for (int l = 0; l < G; l++) p2[l] = p[l] + p[l ^ o];
                memcpy(p, p2, sizeof(float) * (size_t)G);
            }
            y[r] = p[0];
        }
        memcpy(Q, y, sizeof(float) * (size_t)n);
    }
    free(uc); free(uq); free(uv);
    free(y);
}

/* wavefront/linear_mc.cu:238-289: level order, reaches in level_reach
 * order; a reach's first segment takes the sums (from 0.0, CSR order) of its
 * upstream reaches' last-segment flows, new and old; later segments take
 * the previous segment's.  Terms in the reference's order:
 * C1 new + C2 old + C3 Q_old + C4/ql qlat. */
void orc_linear_mc_fp64(const float *C1f, const float *C2f, const float *C3f, const float *C4f,
                        const int *reach_seg_start, const int *reach_seg_len,
                        const int *reach_up_ptr, const int *reach_up_idx,
                        const int *level_ptr, const int *level_reach,
                        int n_reaches, int n_levels, int n_segments,
                        const float *qlat_ts, const float *Q0, double *Q_final,
                        int n_timesteps)
{
    (void)n_reaches;
    double *Qp = (double *)malloc(sizeof(double) * (size_t)(n_segments > 0 ? n_segments : 1));
    double *Qn = (double *)calloc((size_t)(n_segments > 0 ? n_segments : 1), sizeof(double));
    for (int i = 0; i < n_segments; ++i) Qp[i] = Q0[i];
    for (int t = 0; t < n_timesteps; ++t) {
        const float *qlat = qlat_ts + (size_t)t * n_segments;
        for (int L = 0; L < n_levels; ++L) {
            for (int kk = level_ptr[L]; kk < level_ptr[L + 1]; ++kk) {
                int r = level_reach[kk];
                int s0 = reach_seg_start[r], sl = reach_seg_len[r];
                double sum_new = 0.0, sum_old = 0.0;
                for (int k = reach_up_ptr[r]; k < reach_up_ptr[r + 1]; ++k) {
                    int u = reach_up_idx[k];
                    int last_u = reach_seg_start[u] + reach_seg_len[u] - 1;
                    sum_new += Qn[last_u];
                    sum_old += Qp[last_u];
                }
                if (sl <= 0) continue;
                int idx = s0;
                Qn[idx] = (double)C1f[idx] * sum_new + (double)C2f[idx] * sum_old +
                          (double)C3f[idx] * Qp[idx] + (double)C4f[idx] * (double)qlat[idx];
                double prev_new = Qn[idx], prev_old = Qp[idx];
                for (int i = 1; i < sl; ++i) {
                    idx = s0 + i;
                    Qn[idx] = (double)C1f[idx] * prev_new + (double)C2f[idx] * prev_old +
                              (double)C3f[idx] * Qp[idx] + (double)C4f[idx] * (double)qlat[idx];
                    prev_new = Qn[idx];
                    prev_old = Qp[idx];
                }
            }
        }
        double *tmp = Qp; Qp = Qn; Qn = tmp;
        memset(Qn, 0, sizeof(double) * (size_t)n_segments);
    }
    for (int i = 0; i < n_segments; ++i) Q_final[i] = Qp[i];
    free(Qp);
    free(Qn);
}


/* ------------------------------------------------------------ diffusive wave */

/* owp_extended_kernels.cu:35-94: per reach, forward elimination of the
 * implicit (theta-weighted) diffusion system with the upstream boundary
 * eei[0] = 1, ffi[0] = 0, then back substitution from the downstream ghost
 * node qp_prev[nc - 1].  Float arithmetic in the reference's order; the
 * reference's alpha == 1 is kept as a variable so the expressions match. */
void orc_diffusive_tridiag(const orc_diff_reach *rp, const float *dx, const float *diffusivity,
                           const float *celerity, const float *qp_prev, float *qp_out, int nreach)
{
    float eei[ORC_DIFF_MAX_NODES], ffi[ORC_DIFF_MAX_NODES];
    (void)celerity;   /* the reference's Courant clamp is computed and not used */
    for (int j = 0; j < nreach; j++) {
        const int nc = rp[j].ncomp;
        const float theta = rp[j].theta, dtini = rp[j].dtini;
        const size_t base = (size_t)j * ORC_DIFF_MAX_NODES;
        eei[0] = 1.0f;
        ffi[0] = 0.0f;
        for (int i = 1; i < nc; i++) {
            float dxi = dx[base + i - 1];
            if (dxi < 1.0f) dxi = 1.0f;
            const float alpha = 1.0f;
            float diff_i = diffusivity[base + i];
            if (diff_i < 0.01f) diff_i = 0.01f;
            const float ppi = -theta * diff_i * dtini / (dxi * dxi) * 2.0f / (alpha * (alpha + 1.0f)) * alpha;
            const float qqi = 1.0f - ppi * (alpha + 1.0f) / alpha;
            const float rri = ppi / alpha;
            const float next = (i < nc - 1) ? qp_prev[base + i + 1] : qp_prev[base + i];
            const float ssi = qp_prev[base + i] + dtini * diff_i * (1.0f - theta) *
                                                      (qp_prev[base + i - 1] - 2.0f * qp_prev[base + i] + next) /
                                                      (dxi * dxi);
            float denom = ppi * eei[i - 1] + qqi;
            if (fabsf(denom) < 1e-20f) denom = 1e-20f;
            eei[i] = -rri / denom;
            ffi[i] = (ssi - ppi * ffi[i - 1]) / denom;
        }
        qp_out[base + nc - 1] = eei[nc - 1] * qp_prev[base + nc - 1] + ffi[nc - 1];
        for (int i = nc - 2; i >= 0; i--) qp_out[base + i] = eei[i] * qp_out[base + i + 1] + ffi[i];
    }
}
