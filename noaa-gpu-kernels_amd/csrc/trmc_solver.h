// Device-side Muskingum-Cunge segment solve for gfx950, FP64.
//
// REFERENCE rule: the reference's secant (wavefront_mc.cu:215-337, identical
// arithmetic in troute_mc_final.cu:42-139) with its operation order kept term
// for term (-ffp-contract=off), and every pow replaced by the correctly
// rounded pow23_53 below.  Derived per-segment constants are computed once per
// segment by the planner with the same IEEE operations the reference performs
// inline (1/cs, sqrt(s0)/n, sqrt(1+z*z), bankfull depth), so they are
// bit-identical to the inline ones.
//
// CONVERGED rule: the fixed point of the same iteration (h_0 == h, Qj_0 == 0,
// hence X0 == 0.5) gives the single-valued residual
//     F(h) = Qmc(h) - Qmn(h)
// whose smallest root on [mindepth, inf) is found by a bracket-safeguarded
// secant to a relative depth tolerance.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "trmc_device.h"

// Device functions are also host-callable so a CPU build of this header can
// exercise the same solver logic in tests.
#define TRMC_HD __host__ __device__ __forceinline__

namespace trmc {

struct Seg {
    double dx, bw, twcc, n_ch, ncc, s0, z, bfd, ks, sq, sqs0;
};

TRMC_HD Seg load_seg(const double *__restrict__ sc, size_t stride, int s)
{
    Seg c;
    c.dx = sc[SC_DX * stride + s];
    c.bw = sc[SC_BW * stride + s];
    c.twcc = sc[SC_TWCC * stride + s];
    c.n_ch = sc[SC_N * stride + s];
    c.ncc = sc[SC_NCC * stride + s];
    c.s0 = sc[SC_S0 * stride + s];
    c.z = sc[SC_Z * stride + s];
    c.bfd = sc[SC_BFD * stride + s];
    c.ks = sc[SC_KS * stride + s];
    c.sq = sc[SC_SQ * stride + s];
    c.sqs0 = sc[SC_SQS0 * stride + s];
    return c;
}

// Correctly rounded x^y for the two exponents the MC arithmetic uses,
// y = double(2/3) and y = double(5/3).  The reference calls the host libm's
// pow (wavefront_mc.cu:251-252, :270, :334); matching it to the last ulp is
// what keeps the loose-tolerance secant on the same path, so both powers are
// computed to ~2^-100 relative before one final rounding:
//   * x = m 2^(3k) with m in [1, 8);
//   * c = cbrt(m) refined by one Newton step carried in double-double
//     (error ~ eps(cbrt)^2 ~ 2^-102);
//   * m^(2/3) = c^2 and m^(5/3) = m c^2, both as double-double;
//   * the exponents are not 2/3 and 5/3 exactly: 3 double(2/3) = 2 - 2^-53 and
//     3 double(5/3) = 5 + 2^-52, so x^double(2/3) = x^(2/3) exp(-L) and
//     x^double(5/3) = x^(5/3) exp(2L) with L = ln(x) 2^-53 / 3 (|L| < 2^-44),
//     applied to second order;
//   * one rounding of the double-double sum, then an exact 2^(2k) / 2^(5k)
//     scale.
// Results are exact roundings for every x whose powers are normal numbers,
// except when the true value lies within ~2^-100 relative of a rounding
// midpoint.  One cube root and one log serve both powers.
struct Pow2353 { double p23, p53; };

TRMC_HD Pow2353 pow23_53(double x)
{
    Pow2353 r;
    if (!(x > 0.0) || x == HUGE_VAL) {          // 0, negative, NaN, +inf
        double v = (x == 0.0) ? 0.0 : (x == HUGE_VAL ? HUGE_VAL : NAN);
        r.p23 = v;
        r.p53 = v;
        return r;
    }
    int e = 0;
    const double f = frexp(x, &e);                // x = f 2^e, f in [0.5, 1)
    const int e1 = e - 1;
    const int k = (e1 >= 0) ? e1 / 3 : -((2 - e1) / 3);   // floor(e1 / 3)
    const double m = ldexp(2.0 * f, e1 - 3 * k);  // [1, 8)
    const double c0 = cbrt(m);
    const double s_hi = c0 * c0, s_lo = fma(c0, c0, -s_hi);
    const double t_hi = s_hi * c0, t_lo = fma(s_hi, c0, -t_hi) + s_lo * c0;
    const double rho = (t_hi - m) + t_lo;         // c0^3 - m
    const double corr = rho / (3.0 * s_hi);
    const double c_hi = c0 - corr;
    const double c_lo = (c0 - c_hi) - corr;
    const double p_hi = c_hi * c_hi;
    const double p_lo = fma(c_hi, c_hi, -p_hi) + 2.0 * c_hi * c_lo;
    const double q_hi = m * p_hi;
    const double q_lo = fma(m, p_hi, -q_hi) + m * p_lo;
    const double L = ldexp(log(x) / 3.0, -53);
    const double e23 = -L + 0.5 * L * L;          // exp(-L) - 1
    const double e53 = 2.0 * L + 2.0 * L * L;     // exp(2L) - 1
    r.p23 = ldexp(p_hi + (p_lo + p_hi * e23), 2 * k);
    r.p53 = ldexp(q_hi + (q_lo + q_hi * e53), 5 * k);
    return r;
}

struct Geo { double twl, R, A, AC, WP, WPC; };
struct Coef { double c1, c2, c3, c4; };

// compound trapezoid geometry, wavefront_mc.cu:200-213
TRMC_HD Geo geometry(double h, const Seg &c)
{
    Geo g;
    g.twl = c.bw + 2.0 * c.z * h;
    double over = fmax(h - c.bfd, 0.0);
    double in = fmin(c.bfd, h);
    if (over > 0.0 && c.twcc <= 0.0) { over = 0.0; in = h; }
    g.A = (c.bw + in * c.z) * in;
    g.WP = c.bw + 2.0 * in * c.sq;
    g.AC = c.twcc * over;
    g.WPC = (over > 0.0) ? (c.twcc + 2.0 * over) : 0.0;
    g.R = (g.WP + g.WPC > 0.0) ? (g.A + g.AC) / (g.WP + g.WPC) : 0.0;
    return g;
}

// kinematic celerity from R^(2/3), R^(5/3), wavefront_mc.cu:249-254
TRMC_HD double celerity(double h, double r23, double r53, const Seg &c)
{
    if (!(h > 0.0)) return 0.0;
    return fmax(0.0, c.ks * ((5.0 / 3.0) * r23 - (2.0 / 3.0) * r53 *
                             (2.0 * c.sq / (c.bw + 2.0 * h * c.z))));
}

TRMC_HD double storage_km(double ck, double dt, const Seg &c)
{
    return (ck > 0.0) ? fmax(dt, c.dx / ck) : dt;
}

TRMC_HD double weight_width(double h, const Geo &g, const Seg &c)
{
    return (h > c.bfd && c.twcc > 0.0) ? c.twcc : g.twl;
}

TRMC_HD double weighting(double ck, double twu, double q, double lo, const Seg &c)
{
    if (ck > 0.0 && twu * c.s0 * ck * c.dx > 0.0)
        return fmin(0.5, fmax(lo, 0.5 * (1.0 - q / (2.0 * twu * c.s0 * ck * c.dx))));
    return 0.5;
}

TRMC_HD Coef muskingum(double km, double x, double dt, double ql)
{
    Coef m;
    double D = km * (1.0 - x) + dt / 2.0;
    if (D == 0.0) D = 1.0;
    m.c1 = (km * x + dt / 2.0) / D;
    m.c2 = (dt / 2.0 - km * x) / D;
    m.c3 = (km * (1.0 - x) - dt / 2.0) / D;
    m.c4 = (ql * dt) / D;
    return m;
}

// Manning flow with the R^(2/3) already at hand, wavefront_mc.cu:267-270
TRMC_HD double manning(const Geo &g, double r23, const Seg &c)
{
    if (!(g.WP + g.WPC > 0.0)) return 0.0;
    return (1.0 / (((g.WP * c.n_ch) + (g.WPC * c.ncc)) / (g.WP + g.WPC))) * (g.A + g.AC) * r23 * c.sqs0;
}

// wavefront_mc.cu:321-328
TRMC_HD double outflow(const Coef &m, double qup, double quc, double qdp)
{
    double routed = m.c1 * qup + m.c2 * quc + m.c3 * qdp;
    double qmc = routed + m.c4;
    if (qmc < 0.0) {
        if (m.c4 < 0.0 && fabs(m.c4) > routed) return 0.0;
        return fmax(m.c1 * qup + m.c2 * quc + m.c4, m.c1 * qup + m.c3 * qdp + m.c4);
    }
    return qmc;
}

// wavefront_mc.cu:330-336
TRMC_HD double velocity(double h, const Seg &c)
{
    double twl = c.bw + 2.0 * c.z * h;
    double Rv = (h * (c.bw + twl) / 2.0) /
                (c.bw + 2.0 * sqrt(((twl - c.bw) / 2.0) * ((twl - c.bw) / 2.0) + h * h));
    if (Rv < 0.0) Rv = 0.0;
    return (1.0 / c.n_ch) * pow23_53(Rv).p23 * c.sqs0;
}

TRMC_HD bool seg_invalid(const Seg &c)
{
    return c.n_ch <= 0.0 || c.s0 <= 0.0 || c.z <= 0.0 || c.bw <= 0.0;
}

struct SolveOut {
    double q, depth, vel;
    int iters;
    int flags;   // bit0: hit the iteration cap (CONVERGED), bit1: fell back (LEARNED)
};

// Everything of one secant point that depends only on the depth: geometry,
// R^(2/3), R^(5/3), celerity, storage constant, weighting width, Manning flow.
struct Point {
    double h, r23, ck, km, twu, qmn;
};

TRMC_HD Point eval_point(double h, double dt, const Seg &c)
{
    Point p;
    p.h = h;
    const Geo g = geometry(h, c);
    const Pow2353 w = pow23_53(g.R);
    p.r23 = w.p23;
    p.ck = celerity(h, w.p23, w.p53, c);
    p.km = storage_km(p.ck, dt, c);
    p.twu = weight_width(h, g, c);
    p.qmn = manning(g, w.p23, c);
    return p;
}

// REFERENCE stopping rule, wavefront_mc.cu:215-337.  Each iteration evaluates
// the secant points h_0 and h; the next iteration's h_0 is max(0, h) == h, so
// the depth-only part of that point is carried over instead of recomputed
// (same inputs, same operations: the values are identical).
TRMC_HD SolveOut solve_reference(double dt, double qup, double quc, double qdp, double ql,
                                 const Seg &c, double depthp)
{
    SolveOut o;
    o.iters = 0;
    o.flags = 0;
    if (seg_invalid(c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0)) {
        o.q = 0.0; o.depth = 0.0; o.vel = 0.0;
        return o;
    }
    const double mindepth = 0.01;
    double d0 = fmax(depthp, 0.0);
    double h = d0 * 1.33 + mindepth;
    double h_0 = d0 * 0.67;
    Coef C = {0.0, 0.0, 0.0, 0.0};
    double Qj_0 = 0.0;
    int maxiter = 100, total = 0;
    Point p0;
    bool have0 = false;
    for (int attempt = 0; attempt < 5; ++attempt) {
        double aerror = 0.01, rerror = 1.0;
        int iter = 0;
        while (rerror > 0.01 && aerror >= mindepth && iter <= maxiter) {
            if (!have0 || !(p0.h == h_0)) p0 = eval_point(h_0, dt, c);
            const double x0 = weighting(p0.ck, p0.twu, Qj_0, 0.0, c);
            const Coef m0 = muskingum(p0.km, x0, dt, ql);
            const double q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
            Qj_0 = q0 - p0.qmn;

            const Point p1 = eval_point(h, dt, c);
            const double X = weighting(p1.ck, p1.twu, q0, 0.25, c);
            C = muskingum(p1.km, X, dt, ql);
            const double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
            if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
            const double Qj = (routed + C.c4) - p1.qmn;

            double h_1 = (Qj_0 - Qj != 0.0) ? h - (Qj * (h_0 - h)) / (Qj_0 - Qj) : h;
            if (h_1 < 0.0) h_1 = h;
            if (h > 0.0) { rerror = fabs((h_1 - h) / h); aerror = fabs(h_1 - h); }
            else { rerror = 0.0; aerror = 0.9; }
            h_0 = fmax(0.0, h);
            h = fmax(0.0, h_1);
            p0 = p1;
            have0 = true;
            ++iter;
            if (h < mindepth) break;
        }
        total += iter;
        if (iter < maxiter) break;
        h *= 1.33;
        h_0 *= 0.67;
        maxiter += 25;
    }
    o.iters = total;
    o.vel = velocity(h, c);
    o.depth = h;
    o.q = outflow(C, qup, quc, qdp);
    return o;
}

// The REFERENCE solve as a resumable state machine: start() does the guards
// and the initial bracket, step() one secant iteration plus the loop/attempt
// bookkeeping of wavefront_mc.cu:246-318, finish() the outputs.  Driving
// start(); while (!step()); finish() performs exactly the operations of
// solve_reference in the same order, so the results are identical; a kernel
// can interleave the iterations of many solves.
struct SecantRef {
    Point p0;
    Coef C;
    double qup, quc, qdp, ql, dt;
    double h, h_0, Qj_0, rerror, aerror;
    int iter, maxiter, attempt, total;
    bool have0;
    bool trivial;     // guard hit: outputs are zero

    // returns true when no iteration is needed (guard)
    TRMC_HD bool start(double dt_, double qup_, double quc_, double qdp_, double ql_, const Seg &c,
                       double depthp)
    {
        dt = dt_; qup = qup_; quc = quc_; qdp = qdp_; ql = ql_;
        total = 0;
        trivial = seg_invalid(c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0);
        if (trivial) return true;
        const double d0 = fmax(depthp, 0.0);
        h = d0 * 1.33 + 0.01;
        h_0 = d0 * 0.67;
        C.c1 = 0.0; C.c2 = 0.0; C.c3 = 0.0; C.c4 = 0.0;
        Qj_0 = 0.0;
        maxiter = 100;
        attempt = 0;
        iter = 0;
        aerror = 0.01;
        rerror = 1.0;
        have0 = false;
        return false;
    }

    // one iteration; returns true when the solve has ended
    TRMC_HD bool step(const Seg &c)
    {
        const double mindepth = 0.01;
        if (!have0 || !(p0.h == h_0)) p0 = eval_point(h_0, dt, c);
        const double x0 = weighting(p0.ck, p0.twu, Qj_0, 0.0, c);
        const Coef m0 = muskingum(p0.km, x0, dt, ql);
        const double q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
        Qj_0 = q0 - p0.qmn;
        const Point p1 = eval_point(h, dt, c);
        const double X = weighting(p1.ck, p1.twu, q0, 0.25, c);
        C = muskingum(p1.km, X, dt, ql);
        const double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
        if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
        const double Qj = (routed + C.c4) - p1.qmn;
        double h_1 = (Qj_0 - Qj != 0.0) ? h - (Qj * (h_0 - h)) / (Qj_0 - Qj) : h;
        if (h_1 < 0.0) h_1 = h;
        if (h > 0.0) { rerror = fabs((h_1 - h) / h); aerror = fabs(h_1 - h); }
        else { rerror = 0.0; aerror = 0.9; }
        h_0 = fmax(0.0, h);
        h = fmax(0.0, h_1);
        p0 = p1;
        have0 = true;
        ++iter;
        const bool more = !(h < mindepth) && rerror > 0.01 && aerror >= mindepth && iter <= maxiter;
        if (more) return false;
        total += iter;                       // attempt over
        if (iter < maxiter) return true;
        h *= 1.33;
        h_0 *= 0.67;
        maxiter += 25;
        if (++attempt == 5) return true;
        iter = 0;                            // the next attempt always iterates at least once
        aerror = 0.01;
        rerror = 1.0;
        return false;
    }

    TRMC_HD SolveOut finish(const Seg &c) const
    {
        SolveOut o;
        o.flags = 0;
        o.iters = total;
        if (trivial) { o.q = 0.0; o.depth = 0.0; o.vel = 0.0; return o; }
        o.vel = velocity(h, c);
        o.depth = h;
        o.q = outflow(C, qup, quc, qdp);
        return o;
    }
};

// ------------------------------------------------------------- CONVERGED
// F(h) with the X0 = 0.5 coefficients feeding the weighting X(h); also returns
// the final coefficients at h for the outflow rule.
TRMC_HD double residual(double h, double dt, double qup, double quc, double qdp,
                                           double ql, const Seg &c, Coef &C)
{
    Geo g = geometry(h, c);
    const Pow2353 w = pow23_53(g.R);
    const double r23 = w.p23, r53 = w.p53;
    double ck = celerity(h, r23, r53, c);
    double km = storage_km(ck, dt, c);
    Coef m0 = muskingum(km, 0.5, dt, ql);
    double q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
    double X = weighting(ck, weight_width(h, g, c), q0, 0.25, c);
    C = muskingum(km, X, dt, ql);
    double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
    if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
    return (routed + C.c4) - manning(g, r23, c);
}

struct RootState {
    double lo_pos;   // largest depth seen with F > 0 (below the root), or -1
    double hi_neg;   // smallest depth seen with F <= 0 (above the root), or +inf
};

// Bracket-safeguarded secant for the smallest root of F on [mindepth, hmax]
// starting from (ha, hb).  Returns the last evaluated depth (its coefficients
// in C) and counts evaluations in *evals.
TRMC_HD double secant_root(double ha, double hb, double hmax, double dt,
                                              double qup, double quc, double qdp, double ql,
                                              const Seg &c, double tol, int max_eval,
                                              Coef &C, int *evals, bool *capped)
{
    const double lo = 0.01;
    RootState b = {-1.0, hmax};
    ha = fmin(fmax(ha, lo), hmax);
    hb = fmin(fmax(hb, lo), hmax);
    if (hb == ha) hb = (ha < hmax) ? fmin(hmax, ha * 1.01 + 1e-6) : ha * 0.99;
    Coef Ca;
    double fa = residual(ha, dt, qup, quc, qdp, ql, c, Ca);
    int n = 1;
    if (fa > 0.0) b.lo_pos = ha; else b.hi_neg = fmin(b.hi_neg, ha);
    double fb = residual(hb, dt, qup, quc, qdp, ql, c, C);
    ++n;
    if (fb > 0.0) b.lo_pos = fmax(b.lo_pos, hb); else b.hi_neg = fmin(b.hi_neg, hb);
    double h = hb;
    *capped = false;
    for (;;) {
        // F(mindepth) <= 0: the root is clamped at mindepth
        if (b.hi_neg == lo) { h = lo; if (ha != lo && hb != lo) { residual(lo, dt, qup, quc, qdp, ql, c, C); ++n; } break; }
        double den = fa - fb;
        double hn = (den != 0.0) ? hb - fb * (ha - hb) / den : 0.5 * (ha + hb);
        double left = fmax(b.lo_pos, lo);
        bool bracketed = b.lo_pos > 0.0 && b.hi_neg < hmax;
        if (!(hn == hn) || hn <= left || hn >= b.hi_neg) {
            if (bracketed) hn = 0.5 * (left + b.hi_neg);
            else if (b.lo_pos < 0.0) hn = lo;                 // all F <= 0: try mindepth
            else hn = fmin(hmax, fmax(hb, ha) * 2.0);         // all F > 0: expand upward
        }
        if (n >= max_eval) { *capped = true; break; }
        double fn = residual(hn, dt, qup, quc, qdp, ql, c, C);
        ++n;
        h = hn;
        if (fn > 0.0) b.lo_pos = fmax(b.lo_pos, hn); else b.hi_neg = fmin(b.hi_neg, hn);
        double step = fabs(hn - hb);
        ha = hb; fa = fb; hb = hn; fb = fn;
        if (fn == 0.0 || step <= tol * hn) break;
        if (b.lo_pos > 0.0 && b.hi_neg < hmax && (b.hi_neg - b.lo_pos) <= tol * b.hi_neg) break;
    }
    *evals = n;
    return h;
}

TRMC_HD SolveOut solve_converged(double dt, double qup, double quc, double qdp,
                                                    double ql, const Seg &c, double ha, double hb,
                                                    double tol, int max_eval)
{
    SolveOut o;
    o.iters = 0;
    o.flags = 0;
    if (seg_invalid(c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0)) {
        o.q = 0.0; o.depth = 0.0; o.vel = 0.0;
        return o;
    }
    const double hmax = 1.0e4;
    Coef C;
    int evals = 0;
    bool capped = false;
    double h = secant_root(ha, hb, hmax, dt, qup, quc, qdp, ql, c, tol, max_eval, C, &evals, &capped);
    // the smallest root lies in the main channel when F(bankfull) <= 0
    if (h > c.bfd && c.twcc > 0.0 && c.bfd > 0.01) {
        Coef Cb;
        double fbf = residual(c.bfd, dt, qup, quc, qdp, ql, c, Cb);
        ++evals;
        if (!(fbf > 0.0)) {
            int e2 = 0;
            bool cap2 = false;
            h = secant_root(0.01, c.bfd, c.bfd, dt, qup, quc, qdp, ql, c, tol, max_eval, C, &e2, &cap2);
            evals += e2;
            capped = capped || cap2;
        }
    }
    o.iters = evals;
    o.flags = capped ? 1 : 0;
    o.depth = h;
    o.vel = velocity(h, c);
    o.q = outflow(C, qup, quc, qdp);
    return o;
}

}  // namespace trmc
