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
// whose leftmost downcrossing on [mindepth, inf) (the smallest h where F goes
// from > 0 to <= 0; mindepth when F is never positive) is found by a
// bracket-safeguarded secant to a relative depth tolerance.
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

// x^(2/3) and x^(5/3) to within a few ulps (no correctly rounded finish): the
// CONVERGED/LEARNED residual and velocity, whose root is wanted to 1e-9, not
// bit for bit.  On the device, for x in [2^-100, 2^100]: s = x^(-1/3) from
// the FP32 hardware log2/exp2 (relative error < 3e-6), then two Newton steps
// s <- s + (s/3)(1 - x s^3), each of which squares the relative error
// (e -> -2 e^2: 3e-6 -> 2e-11 -> 1e-21); x^(2/3) = x s, x^(5/3) = x^(2/3) x.
// Elsewhere (and in host builds) through cbrt.
TRMC_HD Pow2353 pow23_53_faithful(double x)
{
    Pow2353 r;
    if (!(x > 0.0) || x == HUGE_VAL) {
        double v = (x == 0.0) ? 0.0 : (x == HUGE_VAL ? HUGE_VAL : NAN);
        r.p23 = v;
        r.p53 = v;
        return r;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    if (x >= 0x1p-100 && x <= 0x1p100) {
        double s = (double)__builtin_amdgcn_exp2f(-(1.0f / 3.0f) * __builtin_amdgcn_logf((float)x));
        double e = fma(-x * s, s * s, 1.0);
        s = fma(s * (1.0 / 3.0), e, s);
        e = fma(-x * s, s * s, 1.0);
        s = fma(s * (1.0 / 3.0), e, s);
        r.p23 = x * s;
        r.p53 = r.p23 * x;
        return r;
    }
#endif
    const double cr = cbrt(x);
    r.p23 = cr * cr;
    r.p53 = r.p23 * x;
    return r;
}

// 1/x for the CONVERGED/LEARNED residual, within about an ulp.  On the device,
// for x in [2^-1000, 2^1000]: the hardware reciprocal estimate refined by two
// Newton steps r <- r + r (1 - x r) (each squares the relative error); other
// x (and host builds) by division.
TRMC_HD double recip(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double ax = fabs(x);
    if (ax >= 0x1p-1000 && ax <= 0x1p1000) {
        double r = __builtin_amdgcn_rcp(x);
        double e = fma(-x, r, 1.0);
        r = fma(r, e, r);
        e = fma(-x, r, 1.0);
        return fma(r, e, r);
    }
#endif
    return 1.0 / x;
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

// wavefront_mc.cu:330-336; FAITHFUL (the CONVERGED/LEARNED output) takes
// R^(2/3) within a few ulps instead of correctly rounded
template <bool FAITHFUL = false>
TRMC_HD double velocity(double h, const Seg &c)
{
    double twl = c.bw + 2.0 * c.z * h;
    double Rv = (h * (c.bw + twl) / 2.0) /
                (c.bw + 2.0 * sqrt(((twl - c.bw) / 2.0) * ((twl - c.bw) / 2.0) + h * h));
    if (Rv < 0.0) Rv = 0.0;
    return (1.0 / c.n_ch) * (FAITHFUL ? pow23_53_faithful(Rv) : pow23_53(Rv)).p23 * c.sqs0;
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
// the final coefficients at h for the outflow rule and the Manning flow Qmn(h).
// The reference's formulas (wavefront_mc.cu:249-328) with each shared
// denominator inverted once (recip) and the powers from pow23_53_faithful:
// F moves by a few ulps of its terms, its root by far less than the 1e-9
// the CONVERGED/LEARNED outputs are held to.
TRMC_HD double residual(double h, double dt, double qup, double quc, double qdp,
                                           double ql, const Seg &c, Coef &C, double &qmn)
{
    const Geo g = geometry(h, c);
    const Pow2353 w = pow23_53_faithful(g.R);
    const double r23 = w.p23, r53 = w.p53;
    // celerity (wavefront_mc.cu:249-254) and storage constant
    double ck = 0.0;
    if (h > 0.0)
        ck = fmax(0.0, c.ks * ((5.0 / 3.0) * r23 - (2.0 / 3.0) * r53 * (2.0 * c.sq * recip(c.bw + 2.0 * h * c.z))));
    const double km = (ck > 0.0) ? fmax(dt, c.dx * recip(ck)) : dt;
    const double hdt = 0.5 * dt;
    // X0 = 0.5 routing, then the weighting X from it (lower limit 0.25)
    const double k5 = 0.5 * km;
    const double q0 = ((k5 + hdt) * qup + (hdt - k5) * quc + (k5 - hdt) * qdp + ql * dt) * recip(k5 + hdt);
    const double twu = weight_width(h, g, c);
    const double den = twu * c.s0 * ck * c.dx;
    double X = 0.5;
    if (ck > 0.0 && den > 0.0) X = fmin(0.5, fmax(0.25, 0.5 * (1.0 - q0 * recip(2.0 * den))));
    double D = km * (1.0 - X) + hdt;
    if (D == 0.0) D = 1.0;
    const double iD = recip(D);
    C.c1 = (km * X + hdt) * iD;
    C.c2 = (hdt - km * X) * iD;
    C.c3 = (km * (1.0 - X) - hdt) * iD;
    C.c4 = (ql * dt) * iD;
    const double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
    if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
    // Manning flow (wavefront_mc.cu:267-270)
    const double wp = g.WP + g.WPC;
    qmn = (wp > 0.0) ? (g.A + g.AC) * r23 * c.sqs0 * wp * recip(g.WP * c.n_ch + g.WPC * c.ncc) : 0.0;
    return (routed + C.c4) - qmn;
}

// Resumable CONVERGED / LEARNED solve.  Every step() is one evaluation of
// F, so interleaved lanes advance at a uniform cost per step.
//   root:         the leftmost downcrossing of F on [mindepth, inf) -- the
//                 smallest h where F passes from > 0 to <= 0 -- or mindepth
//                 when F is never positive (orc_mc_solve_converged).
//   start points: CONVERGED -- the reference's bracket (0.67 d, 1.33 d + 0.01);
//                 LEARNED   -- F at the previous depth d first; then a guess
//                 along the MLP's rating curve from d (the Manning flow
//                 Qmn(d)) to the flow Qmc(d) (need_guess / guess_* / set_guess).
//   iteration:    secant on the last two points, kept inside the downcrossing
//                 bracket [lo_pos, hi_neg] seen so far (bisection when the
//                 step leaves it, expansion while no F <= 0 lies above a
//                 positive point, mindepth while every F <= 0).
//   stop:         |step| <= tol h, F zero to within its rounding, or
//                 bracket width <= tol h.
//   F(mindepth) <= 0: the oracle's geometric scan (x1.25, visiting bankfull)
//                 for a positive point, then for the next non-positive one;
//                 none positive leaves the root at mindepth.
//   bankfull:     a root above bankfull is kept only if F(bankfull) > 0 or
//                 F(mindepth) <= 0; otherwise [mindepth, bankfull] is solved.
//   fallback:     LEARNED solves that have not stopped after `budget`
//                 evaluations restart from the CONVERGED start points.
// Launch-uniform settings of a SecantConv solve (kept out of the per-lane
// state so they stay in scalar registers).
struct ConvCfg {
    double dt, tol;
    int max_eval, budget;
};

struct SecantConv {
    double qup, quc, qdp, ql, dp;
    double ha, fa, hb, fb, lo_pos, hi_neg;
    double qa, qb;           // outflow at ha and hb
    double h_cur, q_cur;     // last evaluated depth and the outflow its coefficients give
    double qmn_cur;          // Manning flow Qmn(h_cur)
    double root_h, root_q;
    double s_prev, s_prev2;  // the last two step sizes taken while bracketed
    int evals, stage;
    bool learned, need_guess, fell_back, capped, trivial, inbank, have_root, scan_bp, neg_seen;

    // upper end of the search: bankfull while the in-bank interval is solved
    TRMC_HD double hmax(const Seg &c) const { return inbank ? c.bfd : 1.0e4; }

    // stage: 0 eval ha, 1 eval hb, 2 secant, 3 bankfull check, 4 settle at
    // mindepth, 5 done, 6 scan upward from mindepth, 7 re-evaluate hi_neg
    TRMC_HD bool start(double qup_, double quc_, double qdp_, double ql_, const Seg &c, double depthp,
                       bool learned_)
    {
        qup = qup_; quc = quc_; qdp = qdp_; ql = ql_;
        evals = 0;
        learned = learned_;
        need_guess = false;
        fell_back = false;
        capped = false;
        inbank = false;
        have_root = false;
        trivial = seg_invalid(c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0);
        if (trivial) return true;
        dp = fmax(depthp, 0.0);
        begin_plain_or_learned(c);
        return false;
    }

    TRMC_HD void begin_plain_or_learned(const Seg &c)
    {
        const double hm = hmax(c);
        lo_pos = -1.0;
        hi_neg = hm;
        neg_seen = false;
        s_prev = s_prev2 = 1.0e30;
        if (learned) {
            ha = fmin(fmax(dp, 0.01), hm);
            hb = ha;                          // set by set_guess
        } else {
            ha = fmin(fmax(dp * 0.67, 0.01), hm);
            hb = fmin(fmax(dp * 1.33 + 0.01, 0.01), hm);
            if (hb == ha) hb = (ha < hm) ? fmin(hm, ha * 1.01 + 1e-6) : ha * 0.99;
        }
        stage = 0;
    }

    // What a LEARNED guess is made from, valid while need_guess: the anchor
    // depth d (evaluated first), the flow the routing gives there (Qmc(d),
    // the outflow rule) and the Manning flow Qmn(d) of depth d.
    TRMC_HD double guess_anchor() const { return ha; }
    TRMC_HD double guess_flow_new() const { return q_cur; }
    TRMC_HD double guess_flow_old() const { return qmn_cur; }

    // the guess for stage 1 (LEARNED)
    TRMC_HD void set_guess(double h)
    {
        if (!(h == h) || h <= 0.0) h = ha * 1.5;
        const double hm = 1.0e4;              // a guess is only asked for before any bankfull check
        hb = fmin(fmax(h, 0.01), hm);
        if (hb == ha) hb = (ha < hm) ? fmin(hm, ha * 1.01 + 1e-6) : ha * 0.99;
        need_guess = false;
    }

    // Keeps [lo_pos, hi_neg] around a downcrossing (F(lo_pos) > 0 >= F(hi_neg),
    // lo_pos < hi_neg; lo_pos < 0 while no positive point lies below hi_neg,
    // neg_seen once hi_neg is an evaluated point rather than hmax).
    // F need not be monotone: a positive point above a bracket is ignored, a
    // non-positive point below lo_pos (an upcrossing lies between) too.  Once
    // F(mindepth) <= 0 is the lowest point (hi_neg == 0.01), the first
    // positive point above it opens an upward search.
    TRMC_HD double eval(double h, const Seg &c, const ConvCfg &k)
    {
        ++evals;
        h_cur = h;
        Coef C;
        double qmn;
        const double f = residual(h, k.dt, qup, quc, qdp, ql, c, C, qmn);
        q_cur = outflow(C, qup, quc, qdp);
        qmn_cur = qmn;
        if (f > 0.0) {
            if (h < hi_neg) lo_pos = fmax(lo_pos, h);
            else if (lo_pos < 0.0 && hi_neg == 0.01) { lo_pos = h; hi_neg = hmax(c); neg_seen = false; }
        } else if (h <= hi_neg && (h > lo_pos || lo_pos < 0.0)) {
            hi_neg = h;
            neg_seen = true;
        }
        return f;
    }

    // F at the last point is zero to within its own rounding: the routed sum
    // C1 qup + C2 quc + C3 qdp + C4 and Qmn carry errors of a few ulps of
    // their terms, so no point can be told apart from a root beyond that.
    TRMC_HD bool f_at_noise(double f) const
    {
        return fabs(f) <= 1.6e-14 * (fabs(qup) + fabs(quc) + fabs(qdp) + qmn_cur);
    }

    // A lower bound of Qmn on [h, inf).  In bank Qmn rises with h (area and
    // hydraulic radius both do), so below bankfull it is Qmn(h) for the rest
    // of the bank.  Above bankfull, with o = h - bfd: the area A_b + twcc o
    // rises, R = (A_b + twcc o) / (WP_b + twcc + 2 o) is monotone in o (a
    // ratio of linear functions), so R >= min(R(o), twcc / 2), and the
    // blended roughness is at most max(n, ncc).
    TRMC_HD double qmn_never_below(double h, const Seg &c) const
    {
        const bool overbank = c.twcc > 0.0 && c.bfd > 0.01;
        const double nmax = fmax(c.n_ch, c.ncc);
        auto bound_above = [&](double o) {
            const double ab = (c.bw + c.bfd * c.z) * c.bfd, wpb = c.bw + 2.0 * c.bfd * c.sq;
            const double a = ab + c.twcc * o, r = a / (wpb + c.twcc + 2.0 * o);
            return a * pow23_53_faithful(fmin(r, 0.5 * c.twcc)).p23 * c.sqs0 / nmax;
        };
        if (!overbank) return qmn_cur;              // in bank for every depth
        if (h <= c.bfd) return fmin(qmn_cur, bound_above(0.0));
        return bound_above(h - c.bfd);
    }

    TRMC_HD void accept_at(double h, double q, const Seg &c)
    {
        root_h = h;
        root_q = q;
        have_root = true;
        stage = (!inbank && root_h > c.bfd && c.twcc > 0.0 && c.bfd > 0.01) ? 3 : 5;
    }

    TRMC_HD void accept_root(const Seg &c) { accept_at(h_cur, q_cur, c); }

    // returns true when the solve is finished (MLP-guess requests leave it running)
    TRMC_HD bool step(const Seg &c, const ConvCfg &k)
    {
        if (stage == 5) return true;
        if (evals >= k.max_eval) {
            capped = true;
            if (!have_root) { root_h = h_cur; root_q = q_cur; have_root = true; }
            stage = 5;
            return true;
        }
        // a guess that has not led to a bracket (or to the mindepth scan)
        // within the budget: restart from the reference bracket
        if (learned && !need_guess && stage <= 2 && evals >= k.budget && !(lo_pos > 0.0 && neg_seen) &&
            hi_neg != 0.01) {
            learned = false;
            fell_back = true;
            begin_plain_or_learned(c);
        }
        switch (stage) {
        case 0:
            fa = eval(ha, c, k);
            qa = q_cur;
            if (learned) need_guess = true;
            stage = 1;
            return false;
        case 1:
            if (need_guess) return false;     // waiting for the wave's MLP call
            fb = eval(hb, c, k);
            qb = q_cur;
            stage = 2;
            if (f_at_noise(fb) && (lo_pos > 0.0 || neg_seen)) accept_root(c);
            return stage == 5;
        case 2: {
            if (lo_pos < 0.0 && hi_neg == 0.01) {   // F(mindepth) <= 0, nothing positive above yet
                hb = 0.01;                    // scan position; fb, qb stand in for F, outflow there
                fb = -1.0;                    // (used only once the scan has replaced both points)
                qb = q_cur;
                scan_bp = !(c.twcc > 0.0 && c.bfd > 0.01);
                stage = 6;
                return false;
            }
            const double hm = hmax(c);
            const double den = fa - fb;
            double hn = (den != 0.0) ? hb - fb * (ha - hb) / den : 0.5 * (ha + hb);
            // Converged without another evaluation: the secant step points
            // to a downcrossing next to the point just evaluated (toward lower
            // h when F <= 0 there, higher when F > 0), inside the bracket, and
            // the secant error estimate |hn - hb| |hn - ha| / h^2 is within
            // tol.  The root is the prediction hn; its outflow is the secant
            // interpolation of the outflows at ha and hb (an error of the
            // same second order).
            if (h_cur == hb && hn == hn && (fb > 0.0 ? hn >= hb : hn <= hb) && (lo_pos > 0.0 || neg_seen) &&
                !(hn <= lo_pos) && !(neg_seen && hn > hi_neg) && fabs(hn - hb) <= 1e-6 * hb &&
                fabs(hn - hb) * fabs(hn - ha) <= k.tol * hb * hb) {
                const double q = (ha != hb) ? qb + (qa - qb) * ((hn - hb) / (ha - hb)) : qb;
                accept_at(hn, q, c);
                return stage == 5;
            }
            const double left = fmax(lo_pos, 0.01);
            const bool bracketed = lo_pos > 0.0 && neg_seen;
            // Once bracketed, every step must be under half the step before
            // the last one, else it bisects the bracket (geometrically while
            // it spans more than a factor 4), after Brent: a secant pinned to
            // one end of a strongly curved F then still converges.
            // The same rule without a bracket: a one-sided secant that creeps
            // (every F <= 0, or every F > 0) jumps to mindepth or expands.
            const bool bisect = hn == hn && fabs(hn - hb) > 0.5 * s_prev2;
            if (bisect || !(hn == hn) || hn <= left || hn >= hi_neg) {
                if (bracketed) hn = (hi_neg > 4.0 * left) ? sqrt(left * hi_neg) : 0.5 * (left + hi_neg);
                else if (lo_pos < 0.0) hn = 0.01;
                else hn = fmin(hm, fmax(ha, hb) * 2.0);
            }
            s_prev2 = s_prev;
            s_prev = fabs(hn - hb);
            const double fn = eval(hn, c, k);
            const double stepsz = fabs(hn - hb);
            ha = hb; fa = fb; hb = hn; fb = fn;
            qa = qb; qb = q_cur;
            // a converged step from either side of the bracket: points with
            // F <= 0 only approach from above (hn < hi_neg), so this is a
            // downcrossing too
            if ((lo_pos > 0.0 || neg_seen) &&
                (f_at_noise(fn) || stepsz <= k.tol * hn ||
                 (lo_pos > 0.0 && neg_seen && (hi_neg - lo_pos) <= k.tol * hi_neg))) {
                if (!(h_cur == hi_neg || h_cur == lo_pos)) {
                    stage = 7;                // last point outside the bracket: take its right end
                    return false;
                }
                accept_root(c);
            }
            return stage == 5;
        }
        case 3: {                             // a downcrossing below bankfull?
            const double rh = root_h, rq = root_q;
            const double fbf = eval(c.bfd, c, k);
            const double qbf = q_cur;
            inbank = true;
            if (fbf > 0.0) {                  // no upcrossing between: keep the root
                root_h = rh;
                root_q = rq;
                stage = 5;
                return true;
            }
            // solve [mindepth, bankfull] unless F(mindepth) <= 0
            lo_pos = -1.0;
            hi_neg = c.bfd;
            neg_seen = true;
            s_prev = s_prev2 = 1.0e30;
            ha = 0.01;
            hb = c.bfd;
            fb = fbf;
            qb = qbf;
            fa = eval(ha, c, k);
            qa = q_cur;
            if (!(fa > 0.0)) {
                root_h = rh;
                root_q = rq;
                stage = 5;
                return true;
            }
            stage = 2;
            return false;
        }
        case 4:
            eval(0.01, c, k);
            root_h = 0.01;
            root_q = q_cur;
            have_root = true;
            stage = 5;
            return true;
        case 6: {
            // The oracle's scan (leftmost_downcrossing in mc_oracle.c): x1.25
            // from mindepth, visiting bankfull, first to a point with F > 0,
            // then on to the first point with F <= 0; the secant then solves
            // that bracket.  The search for F > 0 stops early once Qmn on
            // [h, inf) provably exceeds every flow the routing can produce,
            // |qup| + |quc| + |qdp| + 2 |ql| (|C1|, |C2|, |C3| <= 1, C4 <= 2 ql
            // since D >= dt / 2): F < 0 there, so the oracle finds nothing
            // above either.
            double hs = hb * 1.25;
            if (!scan_bp && hs >= c.bfd) { hs = c.bfd; scan_bp = true; }
            if (hs > hmax(c)) { stage = 4; return false; }
            const double fs = eval(hs, c, k);
            ha = hb; fa = fb;
            hb = hs; fb = fs;
            qa = qb; qb = q_cur;
            if (lo_pos > 0.0) {                 // past a positive point: F <= 0 closes the bracket
                if (!(fs > 0.0)) stage = 2;
            } else if (!(fs > 0.0) && qmn_never_below(hs, c) > fabs(qup) + fabs(quc) + fabs(qdp) + 2.0 * fabs(ql)) {
                stage = 4;
            }
            return false;
        }
        case 7:
            eval(hi_neg, c, k);
            accept_root(c);
            return stage == 5;
        }
        return true;
    }

    TRMC_HD SolveOut finish(const Seg &c) const
    {
        SolveOut o;
        o.iters = evals;
        o.flags = (capped ? 1 : 0) | (fell_back ? 2 : 0);
        if (trivial) { o.q = 0.0; o.depth = 0.0; o.vel = 0.0; return o; }
        o.depth = root_h;
        o.vel = velocity<true>(root_h, c);
        o.q = root_q;
        return o;
    }
};

// Whole CONVERGED solve from the reference bracket (no MLP).
TRMC_HD SolveOut solve_converged(double dt, double qup, double quc, double qdp, double ql, const Seg &c,
                                 double depthp, double tol, int max_eval)
{
    const ConvCfg k = {dt, tol, max_eval, 0};
    SecantConv s;
    if (!s.start(qup, quc, qdp, ql, c, depthp, false))
        while (!s.step(c, k)) {
        }
    return s.finish(c);
}

// ------------------------------------------------------------- normal depth
// Leftmost downcrossing of f on [lo, inf): geometric scan (x1.25 from lo,
// visiting the breakpoint bp) for the first f > 0, then on for the first
// f <= 0, then bisection to a two-ulp bracket; lo when f is never positive
// below 1e4 -- the scan the CPU oracle uses (leftmost_downcrossing in mc_oracle.c).
template <class F>
TRMC_HD double leftmost_downcrossing(const F &f, double lo, double bp)
{
    double x = lo;
    bool used_bp = !(bp > lo);
    int k = 0;
    while (!(f(x) > 0.0)) {
        if (++k > 400) return lo;
        double nx = x * 1.25;
        if (!used_bp && nx >= bp) { nx = bp; used_bp = true; }
        if (nx > 1.0e4) return lo;
        x = nx;
    }
    double a = x, b = x;
    for (;;) {
        if (++k > 800) return b;
        double nb = b * 1.25;
        if (!used_bp && nb >= bp) { nb = bp; used_bp = true; }
        if (!(f(nb) > 0.0)) { b = nb; break; }
        a = nb;
        b = nb;
    }
    for (int i = 0; i < 200; ++i) {
        const double m = a + (b - a) / 2.0;
        if (!(m > a && m < b)) break;
        if (f(m) > 0.0) a = m; else b = m;
    }
    return a + (b - a) / 2.0;
}

// Manning normal depth: the leftmost downcrossing of Q - Qmn(h) from 1e-6 m
// (orc_normal_depth).
TRMC_HD double manning_at(double h, const Seg &c)
{
    const Geo g = geometry(h, c);
    return manning(g, pow23_53(g.R).p23, c);
}

TRMC_HD double normal_depth(double Q, const Seg &c)
{
    if (!(Q > 0.0) || seg_invalid(c)) return 0.0;
    const double bp = (c.twcc > 0.0) ? c.bfd : -1.0;
    return leftmost_downcrossing([&](double h) { return Q - manning_at(h, c); }, 1e-6, bp);
}

}  // namespace trmc
