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

// Per-segment constants of one precision: the parameters and the derived
// z = 1/cs, bankfull depth, ks = sqrt(s0)/n, sq = sqrt(1+z^2), sqs0 = sqrt(s0).
// Seg (double) serves the FP64 rules; Seg32 (float, computed with float
// operations as the reference's float solve computes them inline) the FP32
// REFERENCE rule.
template <class T> struct SegT {
    T dx, bw, twcc, n_ch, ncc, s0, z, bfd, ks, sq, sqs0;
};
typedef SegT<double> Seg;
typedef SegT<float> Seg32;

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

TRMC_HD Seg32 load_seg32(const float *__restrict__ sc, size_t stride, int s)
{
    Seg32 c;
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

// fmax / fmin / fabs of one precision (an unqualified call on a float could
// otherwise reach the double C function on the host and widen the expression)
TRMC_HD double tmax(double a, double b) { return fmax(a, b); }
TRMC_HD float tmax(float a, float b) { return fmaxf(a, b); }
TRMC_HD double tmin(double a, double b) { return fmin(a, b); }
TRMC_HD float tmin(float a, float b) { return fminf(a, b); }
TRMC_HD double tabs(double a) { return fabs(a); }
TRMC_HD float tabs(float a) { return fabsf(a); }
TRMC_HD double tsqrt(double a) { return sqrt(a); }
TRMC_HD float tsqrt(float a) { return sqrtf(a); }

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

// The double-double core of the correctly rounded powers: for x > 0 finite,
// x = m 2^(3k) with m in [1, 8), and m^(2/3) = p_hi + p_lo, m^(5/3) = q_hi +
// q_lo to ~2^-100 relative.
struct Cbrt23 { int k; double p_hi, p_lo, q_hi, q_lo; };

TRMC_HD Cbrt23 cbrt23_core(double x)
{
    Cbrt23 r;
    int e = 0;
    const double f = frexp(x, &e);                // x = f 2^e, f in [0.5, 1)
    const int e1 = e - 1;
    r.k = (e1 >= 0) ? e1 / 3 : -((2 - e1) / 3);   // floor(e1 / 3)
    const double m = ldexp(2.0 * f, e1 - 3 * r.k);  // [1, 8)
    const double c0 = cbrt(m);
    const double s_hi = c0 * c0, s_lo = fma(c0, c0, -s_hi);
    const double t_hi = s_hi * c0, t_lo = fma(s_hi, c0, -t_hi) + s_lo * c0;
    const double rho = (t_hi - m) + t_lo;         // c0^3 - m
    const double corr = rho / (3.0 * s_hi);
    const double c_hi = c0 - corr;
    const double c_lo = (c0 - c_hi) - corr;
    r.p_hi = c_hi * c_hi;
    r.p_lo = fma(c_hi, c_hi, -r.p_hi) + 2.0 * c_hi * c_lo;
    r.q_hi = m * r.p_hi;
    r.q_lo = fma(m, r.p_hi, -r.q_hi) + m * r.p_lo;
    return r;
}

TRMC_HD Pow2353 pow23_53(double x)
{
    Pow2353 r;
    if (!(x > 0.0) || x == HUGE_VAL) {          // 0, negative, NaN, +inf
        double v = (x == 0.0) ? 0.0 : (x == HUGE_VAL ? HUGE_VAL : NAN);
        r.p23 = v;
        r.p53 = v;
        return r;
    }
    const Cbrt23 c = cbrt23_core(x);
    const double L = ldexp(log(x) / 3.0, -53);
    const double e23 = -L + 0.5 * L * L;          // exp(-L) - 1
    const double e53 = 2.0 * L + 2.0 * L * L;     // exp(2L) - 1
    r.p23 = ldexp(c.p_hi + (c.p_lo + c.p_hi * e23), 2 * c.k);
    r.p53 = ldexp(c.q_hi + (c.q_lo + c.q_hi * e53), 5 * c.k);
    return r;
}

// The float a double-double hi + lo rounds to (round to nearest even): the
// rounding of hi, unless hi lies exactly halfway between two floats, where
// the sign of lo decides.
TRMC_HD float round_dd_to_float(double hi, double lo)
{
    float f = (float)hi;
    const double fd = (double)f;
    const double diff = hi - fd;                  // exact
    if (diff != 0.0 && lo != 0.0 && (lo > 0.0) == (diff > 0.0)) {
        const float nb = nextafterf(f, diff > 0.0 ? INFINITY : -INFINITY);
        if (diff == ((double)nb - fd) * 0.5) f = nb;
    }
    return f;
}

// Correctly rounded float x^y for the two float exponents of the FP32 rule,
// y = 2.0f/3.0f = 2/3 + 2^-24/3 and y = 5.0f/3.0f = 5/3 - 2^-23/3 (the
// reference's float solve calls powf with them, wavefront_mc.cu:105-106,
// :124, :132-133, :156, :190): the double-double m^(2/3), m^(5/3) above, the
// exponent offsets as x^(2/3) exp(u) and x^(5/3) exp(-2u) with
// u = ln(x) 2^-24 / 3 (|u| < 2^-18) to third order, and one rounding of the
// double-double result (relative error ~2^-68) to float.
struct Powf2353 { float p23, p53; };

static_assert(2.0f / 3.0f == (float)(11184811.0 / 16777216.0), "float(2/3) = 2/3 + 2^-24/3");
static_assert(5.0f / 3.0f == (float)(13981013.0 / 8388608.0), "float(5/3) = 5/3 - 2^-23/3");

TRMC_HD Powf2353 powf23_53(float xf)
{
    Powf2353 r;
    if (!(xf > 0.0f) || xf == INFINITY) {
        const float v = (xf == 0.0f) ? 0.0f : (xf == INFINITY ? INFINITY : NAN);
        r.p23 = v;
        r.p53 = v;
        return r;
    }
    const double x = (double)xf;
    const Cbrt23 c = cbrt23_core(x);
    const double u = ldexp(log(x) / 3.0, -24);
    const double e23 = u + u * u * (0.5 + u * (1.0 / 6.0));          // exp(u) - 1
    const double e53 = -2.0 * u + u * u * (2.0 - u * (4.0 / 3.0));    // exp(-2u) - 1
    double t = c.p_lo + c.p_hi * e23;
    double hi = c.p_hi + t;
    double lo = (c.p_hi - hi) + t;
    r.p23 = round_dd_to_float(ldexp(hi, 2 * c.k), ldexp(lo, 2 * c.k));
    t = c.q_lo + c.q_hi * e53;
    hi = c.q_hi + t;
    lo = (c.q_hi - hi) + t;
    r.p53 = round_dd_to_float(ldexp(hi, 5 * c.k), ldexp(lo, 5 * c.k));
    return r;
}

// the correctly rounded power pair of one precision
TRMC_HD Pow2353 pow_pair(double x) { return pow23_53(x); }
TRMC_HD Powf2353 pow_pair(float x) { return powf23_53(x); }

// x^(2/3) and x^(5/3) to within two ulps (no correctly rounded finish): the
// CONVERGED/LEARNED residual and its enclosures.  Every step is an IEEE
// double or an integer operation, so a host build (and the CPU oracle's
// restatement, which the converged rules' canonical root depends on) gives
// the device's bits.  For x in [2^-1000, 2^1000]:
//   s = x^(-1/3) from the high word of x, hi(s) = 0x553f1348 - hi(x) / 3
//   (relative error below 4.1e-2 for every normal x), then four Newton steps
//   s <- s + (s/3)(1 - x s^3), each taking the relative error e to -2 e^2
//   (4.1e-2 -> 3.3e-3 -> 2.1e-5 -> 8.8e-10 -> 1.6e-18); x^(2/3) = x s,
//   x^(5/3) = x^(2/3) x.
// Other x through the correctly rounded pow23_53 (as reproducible).
TRMC_HD Pow2353 pow23_53_faithful(double x)
{
    Pow2353 r;
    if (!(x >= 0x1p-1000 && x <= 0x1p1000)) return pow23_53(x);
    const unsigned long long bx = __builtin_bit_cast(unsigned long long, x);
    const unsigned int hs = 0x553f1348u - (unsigned int)(bx >> 32) / 3u;
    double s = __builtin_bit_cast(double, (unsigned long long)hs << 32);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double e = fma(-x * s, s * s, 1.0);
        s = fma(s * (1.0 / 3.0), e, s);
    }
    r.p23 = x * s;
    r.p53 = r.p23 * x;
    return r;
}

// 1/x for the CONVERGED/LEARNED residual: IEEE division (correctly rounded
// on the device and the host alike, so the residual's bits do not depend on
// where it runs).
TRMC_HD double recip(double x) { return 1.0 / x; }

// The REFERENCE arithmetic below is written once for both precisions: T =
// double is the FP64 rule (mc_solve_d, wavefront_mc.cu:200-337), T = float
// the FP32 rule (mc_solve, :51-193), whose operations are the same ones in
// float.  Every constant is T(...) of a literal that rounds to the same float
// as the reference's f-suffixed one (checked below).
static_assert((float)1.33 == 1.33f && (float)0.67 == 0.67f && (float)0.01 == 0.01f && (float)0.9 == 0.9f &&
                  (float)0.25 == 0.25f && (float)0.5 == 0.5f,
              "float constants of the FP32 rule");

template <class T> struct GeoT { T twl, R, A, AC, WP, WPC; };
template <class T> struct CoefT { T c1, c2, c3, c4; };
typedef GeoT<double> Geo;
typedef CoefT<double> Coef;

// compound trapezoid geometry, wavefront_mc.cu:200-213 (float: :51-64)
template <class T>
TRMC_HD GeoT<T> geometry(T h, const SegT<T> &c)
{
    GeoT<T> g;
    g.twl = c.bw + T(2) * c.z * h;
    T over = tmax(h - c.bfd, T(0));
    T in = tmin(c.bfd, h);
    if (over > T(0) && c.twcc <= T(0)) { over = T(0); in = h; }
    g.A = (c.bw + in * c.z) * in;
    g.WP = c.bw + T(2) * in * c.sq;
    g.AC = c.twcc * over;
    g.WPC = (over > T(0)) ? (c.twcc + T(2) * over) : T(0);
    g.R = (g.WP + g.WPC > T(0)) ? (g.A + g.AC) / (g.WP + g.WPC) : T(0);
    return g;
}

// kinematic celerity from R^(2/3), R^(5/3), wavefront_mc.cu:249-254 (float: :104-110)
template <class T>
TRMC_HD T celerity(T h, T r23, T r53, const SegT<T> &c)
{
    if (!(h > T(0))) return T(0);
    return tmax(T(0), c.ks * ((T(5) / T(3)) * r23 - (T(2) / T(3)) * r53 *
                              (T(2) * c.sq / (c.bw + T(2) * h * c.z))));
}

template <class T>
TRMC_HD T storage_km(T ck, T dt, const SegT<T> &c)
{
    return (ck > T(0)) ? tmax(dt, c.dx / ck) : dt;
}

template <class T>
TRMC_HD T weight_width(T h, const GeoT<T> &g, const SegT<T> &c)
{
    return (h > c.bfd && c.twcc > T(0)) ? c.twcc : g.twl;
}

template <class T>
TRMC_HD T weighting(T ck, T twu, T q, T lo, const SegT<T> &c)
{
    if (ck > T(0) && twu * c.s0 * ck * c.dx > T(0))
        return tmin(T(0.5), tmax(lo, T(0.5) * (T(1) - q / (T(2) * twu * c.s0 * ck * c.dx))));
    return T(0.5);
}

template <class T>
TRMC_HD CoefT<T> muskingum(T km, T x, T dt, T ql)
{
    CoefT<T> m;
    T D = km * (T(1) - x) + dt / T(2);
    if (D == T(0)) D = T(1);
    m.c1 = (km * x + dt / T(2)) / D;
    m.c2 = (dt / T(2) - km * x) / D;
    m.c3 = (km * (T(1) - x) - dt / T(2)) / D;
    m.c4 = (ql * dt) / D;
    return m;
}

// Manning flow with the R^(2/3) already at hand, wavefront_mc.cu:267-270
template <class T>
TRMC_HD T manning(const GeoT<T> &g, T r23, const SegT<T> &c)
{
    if (!(g.WP + g.WPC > T(0))) return T(0);
    return (T(1) / (((g.WP * c.n_ch) + (g.WPC * c.ncc)) / (g.WP + g.WPC))) * (g.A + g.AC) * r23 * c.sqs0;
}

// wavefront_mc.cu:321-328 (float: :177-184)
template <class T>
TRMC_HD T outflow(const CoefT<T> &m, T qup, T quc, T qdp)
{
    T routed = m.c1 * qup + m.c2 * quc + m.c3 * qdp;
    T qmc = routed + m.c4;
    if (qmc < T(0)) {
        if (m.c4 < T(0) && tabs(m.c4) > routed) return T(0);
        return tmax(m.c1 * qup + m.c2 * quc + m.c4, m.c1 * qup + m.c3 * qdp + m.c4);
    }
    return qmc;
}

// wavefront_mc.cu:330-336 (float: :186-190); FAITHFUL (the CONVERGED/LEARNED output,
// double only) takes R^(2/3) within a few ulps instead of correctly rounded
template <bool FAITHFUL = false, class T>
TRMC_HD T velocity(T h, const SegT<T> &c)
{
    T twl = c.bw + T(2) * c.z * h;
    T Rv = (h * (c.bw + twl) / T(2)) /
           (c.bw + T(2) * tsqrt(((twl - c.bw) / T(2)) * ((twl - c.bw) / T(2)) + h * h));
    if (Rv < T(0)) Rv = T(0);
    T r23;
    if constexpr (FAITHFUL) r23 = pow23_53_faithful(Rv).p23;
    else r23 = pow_pair(Rv).p23;
    return (T(1) / c.n_ch) * r23 * c.sqs0;
}

template <class T>
TRMC_HD bool seg_invalid(const SegT<T> &c)
{
    return c.n_ch <= T(0) || c.s0 <= T(0) || c.z <= T(0) || c.bw <= T(0);
}

struct SolveOut {
    double q, depth, vel;
    int iters;
    int flags;   // bit0: hit the iteration cap (CONVERGED), bit1: fell back (LEARNED)
};

// Everything of one secant point that depends only on the depth: geometry,
// R^(2/3), R^(5/3), celerity, storage constant, weighting width, Manning flow.
template <class T> struct PointT {
    T h, r23, ck, km, twu, qmn;
};
typedef PointT<double> Point;

template <class T>
TRMC_HD PointT<T> eval_point(T h, T dt, const SegT<T> &c)
{
    PointT<T> p;
    p.h = h;
    const GeoT<T> g = geometry(h, c);
    const auto w = pow_pair(g.R);
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
// with_vel false leaves o.vel at 0 (a caller that stores no velocity for this
// step skips its power and division)
TRMC_HD SolveOut solve_reference(double dt, double qup, double quc, double qdp, double ql,
                                 const Seg &c, double depthp, bool with_vel = true)
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
    o.vel = with_vel ? velocity(h, c) : 0.0;
    o.depth = h;
    o.q = outflow(C, qup, quc, qdp);
    return o;
}

// The REFERENCE solve as a resumable state machine: start() does the guards
// and the initial bracket, step() one secant iteration plus the loop/attempt
// bookkeeping of wavefront_mc.cu:246-318, finish() the outputs.  Driving
// start(); while (!step()); finish() performs exactly the operations of
// solve_reference in the same order, so the results are identical; a kernel
// can interleave the iterations of many solves.  T = float is the FP32
// rule's solve (wavefront_mc.cu:66-193): the same steps in float.
template <class T>
struct SecantRefT {
    PointT<T> p0;
    CoefT<T> C;
    T qup, quc, qdp, ql, dt;
    T h, h_0, Qj_0, rerror, aerror;
    int iter, maxiter, attempt, total;
    bool have0;
    bool half0;       // p0 was evaluated by the step before: this step evaluates h
    bool trivial;     // guard hit: outputs are zero

    // returns true when no iteration is needed (guard)
    TRMC_HD bool start(T dt_, T qup_, T quc_, T qdp_, T ql_, const SegT<T> &c, T depthp)
    {
        dt = dt_; qup = qup_; quc = quc_; qdp = qdp_; ql = ql_;
        total = 0;
        trivial = seg_invalid(c) || (ql <= T(0) && qup <= T(0) && quc <= T(0) && qdp <= T(0));
        if (trivial) return true;
        const T d0 = tmax(depthp, T(0));
        h = d0 * T(1.33) + T(0.01);
        h_0 = d0 * T(0.67);
        C.c1 = T(0); C.c2 = T(0); C.c3 = T(0); C.c4 = T(0);
        Qj_0 = T(0);
        maxiter = 100;
        attempt = 0;
        iter = 0;
        aerror = T(0.01);
        rerror = T(1);
        have0 = false;
        half0 = false;
        return false;
    }

    // One iteration; returns true when the solve has ended.  kOnePoint: one
    // point per call, the iteration's h_0 point (when it is not carried over:
    // the first iteration of an attempt) in a call of its own, so that a wave
    // whose lanes are in different iterations pays for one point per call;
    // else both points in one call.  The values are the same either way.
    template <bool kOnePoint = false>
    TRMC_HD bool step(const SegT<T> &c)
    {
        const T mindepth = T(0.01);
        const bool point0 = !half0 && (!have0 || !(p0.h == h_0));
        PointT<T> pe;
        if constexpr (kOnePoint) {
            pe = eval_point(point0 ? h_0 : h, dt, c);
            if (point0) {
                p0 = pe;
                half0 = true;
                return false;
            }
        } else {
            if (point0) p0 = eval_point(h_0, dt, c);
            pe = eval_point(h, dt, c);
        }
        half0 = false;
        const T x0 = weighting(p0.ck, p0.twu, Qj_0, T(0), c);
        const CoefT<T> m0 = muskingum(p0.km, x0, dt, ql);
        const T q0 = m0.c1 * qup + m0.c2 * quc + m0.c3 * qdp + m0.c4;
        Qj_0 = q0 - p0.qmn;
        const PointT<T> &p1 = pe;
        const T X = weighting(p1.ck, p1.twu, q0, T(0.25), c);
        C = muskingum(p1.km, X, dt, ql);
        const T routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
        if (C.c4 < T(0) && tabs(C.c4) > routed) C.c4 = -routed;
        const T Qj = (routed + C.c4) - p1.qmn;
        T h_1 = (Qj_0 - Qj != T(0)) ? h - (Qj * (h_0 - h)) / (Qj_0 - Qj) : h;
        if (h_1 < T(0)) h_1 = h;
        if (h > T(0)) { rerror = tabs((h_1 - h) / h); aerror = tabs(h_1 - h); }
        else { rerror = T(0); aerror = T(0.9); }
        h_0 = tmax(T(0), h);
        h = tmax(T(0), h_1);
        p0 = p1;
        have0 = true;
        ++iter;
        const bool more = !(h < mindepth) && rerror > T(0.01) && aerror >= mindepth && iter <= maxiter;
        if (more) return false;
        total += iter;                       // attempt over
        if (iter < maxiter) return true;
        h *= T(1.33);
        h_0 *= T(0.67);
        maxiter += 25;
        if (++attempt == 5) return true;
        iter = 0;                            // the next attempt always iterates at least once
        aerror = T(0.01);
        rerror = T(1);
        return false;
    }

    // with_vel false: o.vel is left at 0 (not wanted for this step)
    TRMC_HD SolveOut finish(const SegT<T> &c, bool with_vel = true) const
    {
        SolveOut o;
        o.flags = 0;
        o.iters = total;
        if (trivial) { o.q = 0.0; o.depth = 0.0; o.vel = 0.0; return o; }
        o.vel = with_vel ? (double)velocity(h, c) : 0.0;
        o.depth = (double)h;
        o.q = (double)outflow(C, qup, quc, qdp);
        return o;
    }
};
typedef SecantRefT<double> SecantRef;
typedef SecantRefT<float> SecantRef32;

// ------------------------------------------------------------- CONVERGED
// F(h) with the X0 = 0.5 coefficients feeding the weighting X(h); also returns
// the final coefficients at h for the outflow rule and the Manning flow Qmn(h).
// The reference's formulas (wavefront_mc.cu:200-213, :249-328) rearranged so
// that one evaluation divides twice, each on the dependency chain once:
//   * the hydraulic radius A/P and the Manning factor P/(P_ch n + P_fp ncc)
//     share the reciprocal of P (P_ch n + P_fp ncc);
//   * the celerity is carried times the top width, cn = ck twl =
//     ks ((5/3) R^(2/3) twl - (4/3) R^(5/3) sqrt(1+z^2)), and the storage
//     constant as K cn = max(dt cn, dx twl); X0 = 0.5 routing and the
//     weighting X = 0.5 - q0 / (4 twu s0 ck dx) are kept as a ratio xn / xd
//     (its clamp to [0.25, 0.5] decided by comparing numerator and
//     denominator), so that C1..C4, ratios with the common factor cn xd, need
//     the one division 1 / (cn xd (K (1 - X) + dt/2));
//   * R^(2/3), R^(5/3) from pow23_53_faithful.
// F moves by a few ulps of its terms against the reference's order, its root
// by far less than the 1e-9 the CONVERGED/LEARNED outputs are held to.  Every
// operation is IEEE (no hardware estimate), so this F has the same bits on the
// device, in a host build and in the CPU oracle's restatement: the canonical
// cell of the root (canon_* below) is decided by its signs.
TRMC_HD double residual(double h, double dt, double qup, double quc, double qdp,
                        double ql, const Seg &c, Coef &C, double &qmn)
{
    // compound section at h
    const double twl = c.bw + 2.0 * c.z * h;
    double over = fmax(h - c.bfd, 0.0);
    double in = fmin(c.bfd, h);
    if (over > 0.0 && c.twcc <= 0.0) { over = 0.0; in = h; }
    const double asum = (c.bw + in * c.z) * in + c.twcc * over;
    const double wpc = c.bw + 2.0 * in * c.sq;
    const double wpf = (over > 0.0) ? (c.twcc + 2.0 * over) : 0.0;
    const double wp = wpc + wpf;
    const double nsum = wpc * c.n_ch + wpf * c.ncc;
    double R = 0.0, mq = 0.0;                    // A / P and P / (P_ch n + P_fp ncc)
    if (wp > 0.0) {
        const double inv = 1.0 / (wp * nsum);
        R = asum * nsum * inv;
        mq = wp * wp * inv;
    }
    const Pow2353 w = pow23_53_faithful(R);
    const double r23 = w.p23, r53 = w.p53;
    qmn = asum * r23 * c.sqs0 * mq;
    // celerity x top width, storage constant x that
    double cn = 0.0;
    if (h > 0.0) cn = fmax(0.0, c.ks * ((5.0 / 3.0) * r23 * twl - (4.0 / 3.0) * r53 * c.sq));
    const double hdt = 0.5 * dt;
    const double cs = (cn > 0.0) ? cn : 1.0;     // ck = 0: K = dt, X = 0.5
    const double P = (cn > 0.0) ? fmax(dt * cn, c.dx * twl) : dt;
    const double b = hdt * cs;
    // X = xn / xd and 1 - X = omx / xd: X = 0.5 - u / v from the X0 = 0.5
    // routed flow (u = q0 (K/2 + dt/2) cn twl, v = (K/2 + dt/2) cn 4 twu s0 dx cn),
    // clamped to [0.25, 0.5] by comparing u with v
    double xn = 0.5, omx = 0.5, xd = 1.0;
    if (cn > 0.0) {
        const double a = 0.5 * P;
        const double n0 = (a + b) * qup + (b - a) * quc + (a - b) * qdp + ql * dt * cn;
        const double twu = (h > c.bfd && c.twcc > 0.0) ? c.twcc : twl;
        const double sdx = twu * c.s0 * c.dx;
        if (sdx > 0.0) {
            const double u = n0 * twl, v = (a + b) * 4.0 * sdx * cn;
            if (u >= 0.25 * v) { xn = 0.25; omx = 0.75; }
            else if (u > 0.0) { xn = 0.5 * v - u; omx = 0.5 * v + u; xd = v; }
        }
    }
    double D = P * omx + b * xd;
    if (D == 0.0) D = 1.0;
    const double iD = 1.0 / D;
    C.c1 = (P * xn + b * xd) * iD;
    C.c2 = (b * xd - P * xn) * iD;
    C.c3 = (P * omx - b * xd) * iD;
    C.c4 = (ql * dt * cs * xd) * iD;
    const double routed = C.c1 * qup + C.c2 * quc + C.c3 * qdp;
    if (C.c4 < 0.0 && fabs(C.c4) > routed) C.c4 = -routed;
    return (routed + C.c4) - qmn;
}

// ---------------------------------------------- certified sign of F
// The solves where F(mindepth) <= 0 (a sharply rising inflow makes the routed
// flow at small depth negative) may hold a positive band of F narrower than
// any sample spacing; there the leftmost downcrossing is found by a scan that
// proves the sign of F on whole intervals.  Two facts make a cheap enclosure:
//   * for fixed X the routed flow R = C1 qup + C2 quc + C3 qdp + C4 is a ratio
//     of functions linear in K (positive denominator K (1 - X) + dt/2), and
//     for fixed K a ratio linear in X: R is monotone in K and in X separately,
//     so over a box of (K, X) it is extreme at a corner (K = inf as the limit
//     C1 = -C2 = X / (1 - X), C3 = 1, C4 = 0; the C4 clamp, max(R, 0) when
//     ql < 0, keeps this);
//   * within the main channel (h <= bankfull) area, hydraulic radius, top
//     width and Manning flow rise with h; on the floodplain area and wetted
//     perimeter are linear in o = h - bankfull, so the hydraulic radius and the
//     blended roughness are monotone in o, extreme at the ends (the bankfull end
//     as the limit o -> 0+).
// The celerity is bounded term by term, K = max(dt, dx / ck) and the weighting
// X = clamp(0.5 (1 - q0 / (2 twu s0 ck dx)), 0.25, 0.5) from those ranges.

TRMC_HD double routed_corner(double K, bool k_inf, double X, double dt, double qup, double quc, double qdp,
                             double ql, bool clamp)
{
    double c1, c2, c3, c4;
    if (k_inf) {
        c1 = X / (1.0 - X); c2 = -c1; c3 = 1.0; c4 = 0.0;
    } else {
        const double D = K * (1.0 - X) + 0.5 * dt;
        c1 = (K * X + 0.5 * dt) / D;
        c2 = (0.5 * dt - K * X) / D;
        c3 = (K * (1.0 - X) - 0.5 * dt) / D;
        c4 = ql * dt / D;
    }
    double v = c1 * qup + c2 * quc + c3 * qdp + c4;
    if (clamp && c4 < 0.0 && v < 0.0) v = 0.0;
    return v;
}

struct DepthEnd { double A, R, n_eff, twu; };

TRMC_HD DepthEnd depth_end(double h, bool flood, const Seg &c)
{
    DepthEnd e;
    if (!flood) {
        e.A = (c.bw + h * c.z) * h;
        e.R = e.A / (c.bw + 2.0 * h * c.sq);
        e.n_eff = c.n_ch;
        e.twu = c.bw + 2.0 * c.z * h;
    } else {
        const double o = fmax(h - c.bfd, 0.0);
        const double Ab = (c.bw + c.bfd * c.z) * c.bfd, Pb = c.bw + 2.0 * c.bfd * c.sq;
        const double P = Pb + c.twcc + 2.0 * o;
        e.A = Ab + c.twcc * o;
        e.R = e.A / P;
        e.n_eff = (Pb * c.n_ch + (c.twcc + 2.0 * o) * c.ncc) / P;
        e.twu = c.twcc;
    }
    return e;
}

// Enclosure [lo, hi] of F on [a, b], both in the main channel (flood false)
// or both at or above bankfull on a floodplain (flood true), widened by 1e-12
// of the flows for rounding.
struct Enclosure { double lo, hi; };

TRMC_HD Enclosure f_enclosure(double a, double b, bool flood, double dt, double qup, double quc, double qdp,
                              double ql, const Seg &c)
{
    const DepthEnd ea = depth_end(a, flood, c), eb = depth_end(b, flood, c);
    const double Rlo = fmin(ea.R, eb.R), Rhi = fmax(ea.R, eb.R);
    const double nlo = fmin(ea.n_eff, eb.n_eff), nhi = fmax(ea.n_eff, eb.n_eff);
    const Pow2353 plo = pow23_53_faithful(Rlo), phi = pow23_53_faithful(Rhi);
    const double qmn_lo = ea.A * plo.p23 * c.sqs0 / nhi;
    const double qmn_hi = eb.A * phi.p23 * c.sqs0 / nlo;
    const double wa = 2.0 * c.sq / (c.bw + 2.0 * a * c.z), wb = 2.0 * c.sq / (c.bw + 2.0 * b * c.z);
    const double ck_lo = fmax(0.0, c.ks * ((5.0 / 3.0) * plo.p23 - (2.0 / 3.0) * phi.p53 * wa));
    const double ck_hi = fmax(0.0, c.ks * ((5.0 / 3.0) * phi.p23 - (2.0 / 3.0) * plo.p53 * wb));
    double K_lo = dt, K_hi = dt, X_lo = 0.5, X_hi = 0.5;
    bool k_inf = false;
    if (ck_hi > 0.0) {
        if (!(ck_lo > 0.0)) {                 // the celerity may vanish: K anywhere in [dt, inf]
            k_inf = true;
            X_lo = 0.25;
        } else {
            K_lo = fmax(dt, c.dx / ck_hi);
            K_hi = fmax(dt, c.dx / ck_lo);
            const double q0a = routed_corner(K_lo, false, 0.5, dt, qup, quc, qdp, ql, false);
            const double q0b = routed_corner(K_hi, false, 0.5, dt, qup, quc, qdp, ql, false);
            const double q0lo = fmin(q0a, q0b), q0hi = fmax(q0a, q0b);
            const double dlo = fmin(ea.twu, eb.twu) * c.s0 * ck_lo * c.dx;
            const double dhi = fmax(ea.twu, eb.twu) * c.s0 * ck_hi * c.dx;
            if (!(dlo > 0.0)) {
                X_lo = 0.25;
            } else {
                const double rlo = fmin(q0lo / (2.0 * dlo), q0lo / (2.0 * dhi));
                const double rhi = fmax(q0hi / (2.0 * dlo), q0hi / (2.0 * dhi));
                X_lo = fmin(0.5, fmax(0.25, 0.5 * (1.0 - rhi)));
                X_hi = fmin(0.5, fmax(0.25, 0.5 * (1.0 - rlo)));
            }
        }
    }
    const double r1 = routed_corner(K_lo, false, X_lo, dt, qup, quc, qdp, ql, true);
    const double r2 = routed_corner(K_lo, false, X_hi, dt, qup, quc, qdp, ql, true);
    const double r3 = routed_corner(K_hi, k_inf, X_lo, dt, qup, quc, qdp, ql, true);
    const double r4 = routed_corner(K_hi, k_inf, X_hi, dt, qup, quc, qdp, ql, true);
    const double slack = 1e-12 * (fabs(qup) + fabs(quc) + fabs(qdp) + 2.0 * fabs(ql) + qmn_hi);
    Enclosure e;
    e.lo = fmin(fmin(r1, r2), fmin(r3, r4)) - qmn_hi - slack;
    e.hi = fmax(fmax(r1, r2), fmax(r3, r4)) - qmn_lo + slack;
    return e;
}

// The largest routed flow any depth gives: the corners of K in [dt, inf],
// X in [0.25, 0.5].
TRMC_HD double routed_max(double dt, double qup, double quc, double qdp, double ql)
{
    const double r1 = routed_corner(dt, false, 0.25, dt, qup, quc, qdp, ql, true);
    const double r2 = routed_corner(dt, false, 0.5, dt, qup, quc, qdp, ql, true);
    const double r3 = routed_corner(dt, true, 0.25, dt, qup, quc, qdp, ql, true);
    const double r4 = routed_corner(dt, true, 0.5, dt, qup, quc, qdp, ql, true);
    return fmax(fmax(r1, r2), fmax(r3, r4));
}

// A lower bound of Qmn on [h, inf): Qmn rises in the main channel; on the
// floodplain it is at least A(h) min(R_h(h), twcc / 2)^(2/3) sqrt(s0) / max(n, ncc)
// (the hydraulic radius tends to twcc / 2 there).
TRMC_HD double qmn_floor(double h, const Seg &c)
{
    const bool flood = c.twcc > 0.0 && c.bfd > 0.01 && c.bfd < 1.0e4;
    const DepthEnd e = depth_end(h, false, c);
    const double in_bank = e.A * pow23_53_faithful(e.R).p23 * c.sqs0 / c.n_ch;
    if (!flood) return in_bank;
    const DepthEnd f = depth_end(fmax(h, c.bfd), true, c);
    const double fl = f.A * pow23_53_faithful(fmin(f.R, 0.5 * c.twcc)).p23 * c.sqs0 / fmax(c.n_ch, c.ncc);
    return (h >= c.bfd) ? fl : fmin(in_bank, fl);
}

// F(mindepth) > 0 for certain: the smallest routed flow over every K >= dt and
// X in [0.25, 0.5] (four corners) exceeds an upper bound of Qmn(0.01) (the
// main channel's, R_h < h).  Otherwise F(mindepth) is evaluated first.
TRMC_HD bool f_mindepth_positive(double qup, double quc, double qdp, double ql, const Seg &c)
{
    if (!(c.bfd >= 0.01)) return false;
    // (products by the constants' roundings differ from the corners by an ulp
    // or so; the margin below covers that)
    const double r = fmin(fmin(0.6 * qup + 0.2 * quc + 0.2 * qdp + 0.8 * ql, qup + ql),
                          fmin(qdp + (qup - quc) * (1.0 / 3.0), qdp + (qup - quc)));
    const double qmn_ub = (c.bw + 0.01 * c.z) * (0.01 * 0.046415888336127789) * c.ks;   // ks = sqrt(s0) / n
    return r > qmn_ub * (1.0 + 1e-9) + 1e-13 * (fabs(qup) + fabs(quc) + fabs(qdp) + fabs(ql));
}

// Mean-value form.  The enclosure above bounds the routed flow and Qmn
// separately, so its width grows like (|R'| + |Qmn'|) (b - a) while F changes
// like |R' - Qmn'|: where the two nearly cancel (a narrow band of F) it is far
// too wide.  An enclosure of F' over [a, b] by the chain rule,
//   F' = R_K K' + R_X X' - Qmn',
//   R_K = (dt/2)(X d + (1 - X)(qdp - S)) / D^2,  R_X = K (K d + dt (qup - qdp + ql)) / D^2,
//   d = qup - quc, S = qup + quc - qdp + 2 ql, D = K (1 - X) + dt/2,
//   K' = -dx ck' / ck^2 (0 where K = dt),  ck' = ks (g_rho rho' + g_w w'),
//   g_rho = 10/9 rho^(-1/3) (1 - rho w),  g_w = -2/3 rho^(5/3),
//   X' = -(q0' den - q0 den') / (4 den^2) where the weighting is not clamped,
//   Qmn' = sqrt(s0) (A' rho^(2/3) / n + A 2/3 rho^(-1/3) rho' / n - A rho^(2/3) n' / n^2),
// every factor in interval arithmetic (a clamp that may switch inside [a, b]
// takes the hull with 0), with F at both ends gives bounds that are tight to
// second order.  None is formed where the celerity may vanish.
struct Iv { double lo, hi; };

TRMC_HD Iv iv_of(double a, double b) { return Iv{fmin(a, b), fmax(a, b)}; }
TRMC_HD Iv iv_add(Iv a, Iv b) { return Iv{a.lo + b.lo, a.hi + b.hi}; }
TRMC_HD Iv iv_sub(Iv a, Iv b) { return Iv{a.lo - b.hi, a.hi - b.lo}; }
TRMC_HD Iv iv_mul(Iv a, Iv b)
{
    const double p1 = a.lo * b.lo, p2 = a.lo * b.hi, p3 = a.hi * b.lo, p4 = a.hi * b.hi;
    return Iv{fmin(fmin(p1, p2), fmin(p3, p4)), fmax(fmax(p1, p2), fmax(p3, p4))};
}
TRMC_HD Iv iv_scale(Iv a, double k) { return iv_of(a.lo * k, a.hi * k); }
TRMC_HD Iv iv_inv(Iv a) { return Iv{1.0 / a.hi, 1.0 / a.lo}; }     // 0 not in a
TRMC_HD Iv iv_hull0(Iv a) { return Iv{fmin(a.lo, 0.0), fmax(a.hi, 0.0)}; }

// R_K over a (K, X) box (numerator linear in X; D > 0)
TRMC_HD Iv routed_dk(Iv K, Iv X, double hdt, double d, double qdp_minus_S)
{
    const Iv one_x = iv_sub(Iv{1.0, 1.0}, X);
    const Iv D = iv_add(iv_mul(K, one_x), Iv{hdt, hdt});
    const Iv num = iv_scale(iv_add(iv_scale(X, d), iv_scale(one_x, qdp_minus_S)), hdt);
    return iv_mul(num, iv_inv(iv_mul(D, D)));
}

TRMC_HD bool fprime_enclosure(double a, double b, bool flood, double dt, double qup, double quc, double qdp,
                              double ql, const Seg &c, Iv &out)
{
    const double hdt = 0.5 * dt;
    Iv A, Ad, rho, rhod, n_eff, n_effd, twu, twud;
    if (!flood) {
        const double Pa = c.bw + 2.0 * a * c.sq, Pb = c.bw + 2.0 * b * c.sq;
        const double Aa = (c.bw + a * c.z) * a, Ab = (c.bw + b * c.z) * b;
        A = Iv{Aa, Ab};
        Ad = Iv{c.bw + 2.0 * c.z * a, c.bw + 2.0 * c.z * b};
        rho = Iv{Aa / Pa, Ab / Pb};
        // rho' = (bw^2 + 2 z h bw + 2 z h^2 sqrt(1+z^2)) / P^2, numerator rising
        const double na = c.bw * c.bw + 2.0 * c.z * a * c.bw + 2.0 * c.z * a * a * c.sq;
        const double nb = c.bw * c.bw + 2.0 * c.z * b * c.bw + 2.0 * c.z * b * b * c.sq;
        rhod = Iv{na / (Pb * Pb), nb / (Pa * Pa)};
        n_eff = Iv{c.n_ch, c.n_ch};
        n_effd = Iv{0.0, 0.0};
        twu = Iv{c.bw + 2.0 * c.z * a, c.bw + 2.0 * c.z * b};
        twud = Iv{2.0 * c.z, 2.0 * c.z};
    } else {
        const double oa = fmax(a - c.bfd, 0.0), ob = fmax(b - c.bfd, 0.0);
        const double Abf = (c.bw + c.bfd * c.z) * c.bfd, P0 = c.bw + 2.0 * c.bfd * c.sq;
        const double Pa = P0 + c.twcc + 2.0 * oa, Pb = P0 + c.twcc + 2.0 * ob;
        A = Iv{Abf + c.twcc * oa, Abf + c.twcc * ob};
        Ad = Iv{c.twcc, c.twcc};
        rho = iv_of((Abf + c.twcc * oa) / Pa, (Abf + c.twcc * ob) / Pb);
        const double c0 = c.twcc * (P0 + c.twcc) - 2.0 * Abf;   // rho' = c0 / P^2
        rhod = iv_of(c0 / (Pa * Pa), c0 / (Pb * Pb));
        n_eff = iv_of((P0 * c.n_ch + (c.twcc + 2.0 * oa) * c.ncc) / Pa, (P0 * c.n_ch + (c.twcc + 2.0 * ob) * c.ncc) / Pb);
        const double c1 = 2.0 * P0 * (c.ncc - c.n_ch);           // n' = c1 / P^2
        n_effd = iv_of(c1 / (Pa * Pa), c1 / (Pb * Pb));
        twu = Iv{c.twcc, c.twcc};
        twud = Iv{0.0, 0.0};
    }
    const Pow2353 pl = pow23_53_faithful(rho.lo), ph = pow23_53_faithful(rho.hi);
    const Iv r23 = Iv{pl.p23, ph.p23}, r53 = Iv{pl.p53, ph.p53};
    const Iv rm13o = Iv{ph.p23 / rho.hi, pl.p23 / rho.lo};      // rho^(-1/3) = rho^(2/3) / rho, falling
    const Iv inv_n = iv_inv(n_eff);
    const Iv q1 = iv_mul(iv_mul(Ad, r23), inv_n);
    const Iv q2 = iv_mul(iv_mul(iv_mul(A, rm13o), rhod), iv_scale(inv_n, 2.0 / 3.0));
    const Iv q3 = iv_mul(iv_mul(iv_mul(A, r23), n_effd), iv_mul(inv_n, inv_n));
    const Iv qmnd = iv_scale(iv_sub(iv_add(q1, q2), q3), c.sqs0);
    const double ea = c.bw + 2.0 * a * c.z, eb = c.bw + 2.0 * b * c.z;
    const Iv wv = Iv{2.0 * c.sq / eb, 2.0 * c.sq / ea};
    const Iv wd = Iv{-4.0 * c.sq * c.z / (ea * ea), -4.0 * c.sq * c.z / (eb * eb)};
    const Iv g = iv_sub(iv_scale(r23, 5.0 / 3.0), iv_mul(iv_scale(r53, 2.0 / 3.0), wv));
    if (!(g.lo > 0.0)) return false;
    const Iv ck = iv_scale(g, c.ks);
    const Iv g_rho = iv_mul(iv_scale(rm13o, 10.0 / 9.0), iv_sub(Iv{1.0, 1.0}, iv_mul(rho, wv)));
    const Iv ckd = iv_scale(iv_add(iv_mul(g_rho, rhod), iv_mul(iv_scale(r53, -2.0 / 3.0), wd)), c.ks);
    Iv K, Kd;
    const Iv kd_branch = iv_scale(iv_mul(ckd, iv_inv(iv_mul(ck, ck))), -c.dx);
    if (c.dx / ck.hi >= dt) { K = Iv{c.dx / ck.hi, c.dx / ck.lo}; Kd = kd_branch; }
    else if (c.dx / ck.lo <= dt) { K = Iv{dt, dt}; Kd = Iv{0.0, 0.0}; }
    else { K = Iv{dt, c.dx / ck.lo}; Kd = iv_hull0(kd_branch); }
    const double d = qup - quc, qdp_minus_S = qdp - (qup + quc - qdp + 2.0 * ql);
    const Iv half = Iv{0.5, 0.5};
    const Iv q0 = iv_of(routed_corner(K.lo, false, 0.5, dt, qup, quc, qdp, ql, false),
                        routed_corner(K.hi, false, 0.5, dt, qup, quc, qdp, ql, false));
    const Iv q0d = iv_mul(routed_dk(K, half, hdt, d, qdp_minus_S), Kd);
    const Iv den = iv_scale(iv_mul(twu, ck), c.s0 * c.dx);
    const Iv dend = iv_scale(iv_add(iv_mul(twud, ck), iv_mul(twu, ckd)), c.s0 * c.dx);
    const Iv u = iv_sub(half, iv_mul(q0, iv_inv(iv_scale(den, 4.0))));
    const Iv ud = iv_scale(iv_mul(iv_sub(iv_mul(q0d, den), iv_mul(q0, dend)), iv_inv(iv_mul(den, den))), -0.25);
    const Iv X = Iv{fmin(0.5, fmax(0.25, u.lo)), fmin(0.5, fmax(0.25, u.hi))};
    Iv Xd;
    if (u.lo > 0.25 && u.hi < 0.5) Xd = ud;
    else if (u.lo >= 0.5 || u.hi <= 0.25) Xd = Iv{0.0, 0.0};
    else Xd = iv_hull0(ud);
    const Iv Dx = iv_add(iv_mul(K, iv_sub(Iv{1.0, 1.0}, X)), Iv{hdt, hdt});
    const double e = dt * (qup - qdp + ql);
    const Iv RX = iv_mul(iv_mul(K, iv_add(iv_scale(K, d), Iv{e, e})), iv_inv(iv_mul(Dx, Dx)));
    Iv Rd = iv_add(iv_mul(routed_dk(K, X, hdt, d, qdp_minus_S), Kd), iv_mul(RX, Xd));
    if (ql < 0.0) Rd = iv_hull0(Rd);                          // the C4 clamp may switch inside
    Iv Fd = iv_sub(Rd, qmnd);
    const double mag = fmax(fabs(Fd.lo), fabs(Fd.hi));
    Fd.lo -= 1e-9 * mag;
    Fd.hi += 1e-9 * mag;
    out = Fd;
    return fabs(Fd.lo) < HUGE_VAL && fabs(Fd.hi) < HUGE_VAL;     // false for inf and NaN
}

// Bounds of F on [a, b] (width w) from F(a) = fa (when have_fa), F(b) = fb
// and F' in [sl, sh]: with t = h - a, F(h) <= min(fa + sh t, fb - sl (w - t))
// and F(h) >= max(fa + sl t, fb - sh (w - t)); a minimum of two lines is
// concave, a maximum convex, so the extremes lie at t = 0, w or the crossing.
TRMC_HD Enclosure slope_bounds(double w, double fa, bool have_fa, double fb, double sl, double sh)
{
    double u0 = fb - sl * w, uw = fb, l0 = fb - sh * w, lw = fb;
    if (have_fa) {
        u0 = fmin(u0, fa); uw = fmin(uw, fa + sh * w);
        l0 = fmax(l0, fa); lw = fmax(lw, fa + sl * w);
    }
    Enclosure e;
    e.hi = fmax(u0, uw);
    e.lo = fmin(l0, lw);
    if (have_fa && sh > sl) {
        const double t = (fb - sl * w - fa) / (sh - sl);
        if (t > 0.0 && t < w) e.hi = fmax(e.hi, fa + sh * t);
        const double t2 = (fb - sh * w - fa) / (sl - sh);
        if (t2 > 0.0 && t2 < w) e.lo = fmin(e.lo, fa + sl * t2);
    }
    return e;
}

// F keeps its sign (pos: F > 0, else F <= 0) on [a, b], shown by the slope
// bounds from F(a) = fa (when have_fa) and F(b) = fb.
TRMC_HD bool slope_sign_holds(double a, double b, bool flood, double fa, bool have_fa, double fb, bool pos,
                              double dt, double qup, double quc, double qdp, double ql, const Seg &c)
{
    Iv Fd;
    if (!fprime_enclosure(a, b, flood, dt, qup, quc, qdp, ql, c, Fd)) return false;
    const Enclosure sb = slope_bounds(b - a, fa, have_fa, fb, Fd.lo, Fd.hi);
    const double slack = 1e-12 * (fabs(qup) + fabs(quc) + fabs(qdp) + 2.0 * fabs(ql) + fabs(fb));
    return pos ? sb.lo - slack > 0.0 : !(sb.hi + slack > 0.0);
}

// The scan's next step after [a, b] (width w) held its sign: where F moves
// toward 0, 0.95 of the distance at which the average slope over [a, b]
// would reach it (a Newton step that stops short); else four times w.
TRMC_HD double scan_next_step(double w, double fa, double fb)
{
    const double toward = fabs(fb) < fabs(fa) ? (fabs(fa) - fabs(fb)) / w : 0.0;
    return (toward > 0.0) ? fmin(4.0 * w, 0.95 * fabs(fb) / toward) : 4.0 * w;
}

// relative depth resolution of the certified scan, and its step limit
constexpr double kScanRes = 1e-6;
constexpr int kScanCap = 400;

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
//   F(mindepth) <= 0: possible only when the routed flow can fall to Qmn at
//                 1 cm (f_mindepth_positive false, under 1 % of solves); then
//                 F(mindepth) is evaluated first, and if it is <= 0 a certified
//                 scan (f_enclosure on whole intervals) finds the first
//                 positive point and the first non-positive one after it,
//                 whose bracket the secant solves; none positive leaves the
//                 root at mindepth.
//   bankfull:     a root above bankfull is kept only if F(bankfull) > 0 or
//                 F(mindepth) <= 0; otherwise [mindepth, bankfull] is solved.
//   fallback:     LEARNED solves that have not stopped after `budget`
//                 evaluations restart from the CONVERGED start points.
// ---------------------------------------------- canonical root
// Two solves of one F that start from different points stop at different
// depths inside the band where the rounding noise of F hides its sign (about
// 1e-14 relative), and the routing carries such differences on from step to
// step.  So the converged rules end every solve whose root is a downcrossing
// of F in its canonical cell.  The grid is the depths whose binary
// significand has its low kCanonBits bits zero (spacing 2^-24 of the depth's
// binade, 6e-8 relative at most); the cell is the pair of neighbouring grid
// points g < g+ with F(g) > 0 >= F(g+) next to the solver's root estimate
// (found from the grid point at or below it, moving one point at a time, at
// most kCanonMoves moves).  F at grid points has fixed bits, so every solver
// whose estimate lies in the noise band picks the same cell (when the band
// straddles a grid point the sign there decides for all of them) and returns
// the same point of it:
//   t = F(g) / (F(g) - F(g+)),  h* = g + (g+ - g) t,  Q* = Q(g) + (Q(g+) - Q(g)) t,
// within about (F''/F') (g+ - g)^2 / 8 of the root where F is smooth.
// Estimates below mindepth's cell or at the search's end (1e4) are kept as
// they are, and so is one whose cell is not found within the moves.
constexpr int kCanonBits = 28;
constexpr int kCanonMoves = 4;
constexpr double kSnapRel = 1e-4;   // secant steps below this (relative) land on grid points

TRMC_HD double canon_floor(double r)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, r);
    return __builtin_bit_cast(double, b & ~((1ull << kCanonBits) - 1ull));
}

// the next grid point above / below a grid point g > 0 (the step crosses
// binades: the significand's carry or borrow moves the exponent)
TRMC_HD double canon_next(double g)
{
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, g) + (1ull << kCanonBits));
}

TRMC_HD double canon_prev(double g)
{
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, g) - (1ull << kCanonBits));
}

// the next double above x > 0
TRMC_HD double canon_above(double x)
{
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) + 1ull);
}

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
    double f_cur;            // F(h_cur)
    double qmn_cur;          // Manning flow Qmn(h_cur)
    double root_h, root_q;
    double s_prev, s_prev2;  // the last two step sizes taken while bracketed
    int evals, stage, scan_n;
    int cmoves;              // canonical cell: moves made from the first cell
    bool noncanon;           // the root estimate was kept (no canonical cell)
    bool learned, need_guess, fell_back, capped, trivial, inbank, have_root, neg_seen;
    bool spos;               // certified scan: F > 0 at its position
    bool certified;          // the bracket came from the certified scan (no bankfull check)

    // upper end of the search: bankfull while the in-bank interval is solved
    TRMC_HD double hmax(const Seg &c) const { return inbank ? c.bfd : 1.0e4; }

    // stage: 0 eval ha, 1 eval hb, 2 secant, 3 bankfull check, 4 settle at
    // mindepth, 5 done, 7 re-evaluate hi_neg, 8 F(mindepth) first,
    // 9 certified scan, 10 F at the scan bracket's lower end, 11 waiting
    // for a deferred certified solve, 13-17 the canonical cell
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
        noncanon = false;
        trivial = seg_invalid(c) || (ql <= 0.0 && qup <= 0.0 && quc <= 0.0 && qdp <= 0.0);
        if (trivial) return true;
        dp = fmax(depthp, 0.0);
        scan_n = 0;
        certified = false;
        begin_plain_or_learned(c);
        if (!f_mindepth_positive(qup, quc, qdp, ql, c)) stage = 8;
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
        f_cur = f;
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

    // The certified scan keeps its position a (where the sign of F is known)
    // in ha, F(a) in fa, the outflow there in qa and its step in hb: the
    // secant's points are unused until the scan hands it a bracket.  Entered
    // right after F(a) was evaluated.
    TRMC_HD void begin_scan(double a, bool pos)
    {
        ha = a;
        fa = -1.0;
        qa = q_cur;
        hb = a;
        spos = pos;
        stage = 9;
    }

    TRMC_HD void accept_at(double h, double q, const Seg &c)
    {
        root_h = h;
        root_q = q;
        have_root = true;
        stage = (!inbank && !certified && root_h > c.bfd && c.twcc > 0.0 && c.bfd > 0.01) ? 3 : canon_begin();
    }

    // The canonical cell around the root estimate root_h: its ends are kept
    // in ha, hb (F in fa, fb, outflows in qa, qb; the secant's points are no
    // longer needed).  Returns the stage to go on with.
    TRMC_HD int canon_begin()
    {
        const double g = canon_floor(root_h);
        if (!(g >= 0.01 && canon_next(g) <= 1.0e4)) {
            noncanon = true;
            return 5;
        }
        ha = g;
        hb = canon_next(g);
        cmoves = 0;
        // the last evaluated point is an end of the cell (the secant's points
        // near the root lie on the grid): F there is known
        if (h_cur == ha) { fa = f_cur; qa = q_cur; return 14; }
        if (h_cur == hb) { fb = f_cur; qb = q_cur; return 15; }
        return 13;
    }

    // The secant point of the cell [ha, hb] (F continuous on it).
    TRMC_HD void canon_finish()
    {
        const double t = fa / (fa - fb);
        root_h = ha + (hb - ha) * t;
        root_q = qa + (qb - qa) * t;
        stage = 5;
    }

    // Both ends evaluated: the cell's point, or a move toward the cell.  A
    // cell holding the floodplain's jump (F continuous on [g, bankfull] and on
    // (bankfull, g+]) is split there: F(bankfull) <= 0 leaves [g, bankfull];
    // else F at the next double above bankfull, where F <= 0 makes that double
    // the downcrossing and F > 0 leaves [that double, g+].
    TRMC_HD void canon_decide(const Seg &c)
    {
        if (fa > 0.0 && !(fb > 0.0)) {
            if (c.twcc > 0.0 && ha <= c.bfd && c.bfd < hb) stage = (ha == c.bfd) ? 17 : 16;
            else canon_finish();
        } else if (++cmoves > kCanonMoves) {
            noncanon = true;                     // keep the estimate
            stage = 5;
        } else if (fa > 0.0) {                   // F > 0 at both ends: the cell lies above
            ha = hb; fa = fb; qa = qb;
            hb = canon_next(hb);
            stage = 14;
            if (!(hb <= 1.0e4)) { noncanon = true; stage = 5; }
        } else {                                 // F <= 0 at the lower end: the cell lies below
            hb = ha; fb = fa; qb = qa;
            ha = canon_prev(ha);
            stage = 15;
            if (!(ha >= 0.01)) { noncanon = true; stage = 5; }
        }
    }

    TRMC_HD void accept_root(const Seg &c) { accept_at(h_cur, q_cur, c); }

    // A certified solve done elsewhere (kDefer): its root, the outflow there
    // and the evaluations it took.
    TRMC_HD void take_scan_result(double h, double q, int n_evals)
    {
        root_h = h;
        root_q = q;
        have_root = true;
        evals += n_evals;
        stage = 5;
    }

    TRMC_HD bool awaiting_scan() const { return stage == 11; }

    // Certified scan (mirrors the oracle's) from a = ha, where F has the sign
    // spos and the value fa, with step hb: F(b) at b = a + hb (not crossing
    // bankfull); the other sign there means a sign change in (a, b]: at the
    // resolution kScanRes a, b is the first point of that sign, else b moves
    // back toward the secant estimate of the crossing; the same sign and the
    // sign shown on all of [a, b] (sign_holds) move a to b with a Newton-like
    // next step; otherwise the step halves, and at the resolution a moves on.
    // Looking for F > 0 the scan ends at mindepth once Qmn on [a, inf)
    // exceeds every routed flow; the first F <= 0 after a positive point
    // closes the bracket the secant then solves.  (Stage 9; host builds and
    // the serial form only: the routing kernel defers these solves.)
    TRMC_HD bool scan_step(const Seg &c, const ConvCfg &k)
    {
        if (++scan_n > kScanCap) {
            capped = true;
            if (spos) hi_neg = ha;
            stage = spos ? 7 : 4;
            return false;
        }
        const double a = ha, w = hb;
        if (!spos) {
            const double rmax = routed_max(k.dt, qup, quc, qdp, ql);
            if (a >= 1.0e4 || qmn_floor(a, c) > rmax + 1e-12 * (fabs(rmax) + 1.0)) { stage = 4; return false; }
        } else if (a >= 1.0e4) {
            hi_neg = 1.0e4;
            stage = 7;
            return false;
        }
        const bool has_flood = c.twcc > 0.0 && c.bfd > 0.01 && c.bfd < 1.0e4;
        double b = fmin(a + w, 1.0e4);
        if (has_flood && a < c.bfd && b > c.bfd) b = c.bfd;
        const bool flood = has_flood && a >= c.bfd;
        // fa is NaN where a was reached by the enclosure alone (F(a) not evaluated)
        const bool have_fa = !(flood && a <= c.bfd) && fa == fa;
        {                                             // the direct enclosure needs no F(b)
            const Enclosure en = f_enclosure(a, b, flood, k.dt, qup, quc, qdp, ql, c);
            if (spos ? en.lo > 0.0 : !(en.hi > 0.0)) {
                ha = b;
                fa = NAN;
                hb = 2.0 * (b - a);
                return false;
            }
        }
        const double fbv = eval(b, c, k);
        if (spos ? !(fbv > 0.0) : fbv > 0.0) {          // a sign change in (a, b]
            if (b - a <= kScanRes * a) {
                if (spos) {                              // the bracket [a, b]
                    lo_pos = a;
                    hi_neg = b;
                    neg_seen = true;
                    certified = true;
                    hb = b; fb = fbv; qb = q_cur;
                    s_prev = s_prev2 = 1.0e30;
                    stage = have_fa ? 2 : 10;            // the secant needs F(a) as well
                    return false;
                }
                spos = true;
                ha = b; fa = fbv; qa = q_cur;
                return false;
            }
            const double t = have_fa && fa != fbv ? (b - a) * fa / (fa - fbv) : 0.5 * (b - a);
            hb = fmax(0.5 * kScanRes * a, fmin(0.5 * (b - a), 0.98 * t));
            return false;
        }
        if (slope_sign_holds(a, b, flood, fa, have_fa, fbv, spos, k.dt, qup, quc, qdp, ql, c)) {
            hb = fmax(kScanRes * b, have_fa ? scan_next_step(b - a, fa, fbv) : 2.0 * (b - a));
            ha = b; fa = fbv; qa = q_cur;
            return false;
        }
        if (b - a > kScanRes * a) { hb = 0.5 * (b - a); return false; }
        ha = b; fa = fbv; qa = q_cur;                    // at the resolution: move on
        hb = 2.0 * w;
        return false;
    }

    // One step of the solve: at most one evaluation of F.  Each stage first
    // names the depth it needs F at (or ends the step without one), then a
    // single evaluation serves every stage, then the stage reads the value.
    // On a SIMT machine the lanes of a wave are in different stages; with
    // one evaluation site a wave pays for one residual per step, whatever
    // mix of stages its lanes are in.  Returns true when the solve is
    // finished (a LEARNED solve waiting for its guess keeps running).
    // kDefer: a solve whose F(mindepth) <= 0 does not run the certified scan
    // here but waits in stage 11 for take_scan_result (the routing kernel
    // hands it to separate waves, which keeps the scan's registers out of its
    // hot loop).
    // Stages: 0 F at ha, 1 F at hb, 2 secant, 3 bankfull check (F at
    // bankfull), 12 its second point (F at mindepth), 4 settle at mindepth,
    // 5 done, 7 F at hi_neg, 8 F(mindepth) first, 9 certified scan, 10 F at
    // the scan bracket's lower end, 11 waiting for a deferred certified solve.
    template <bool kDefer = false>
    TRMC_HD bool step(const Seg &c, const ConvCfg &k)
    {
        if (stage == 5) return true;
        if (stage == 11) return false;           // waiting for a deferred certified solve
        if constexpr (!kDefer) {
            if (stage == 9) return scan_step(c, k);
        }
        // the cap counts evaluations outside the certified scan (which has
        // its own); stage 12 is the second half of the bankfull check
        if (stage != 9 && stage != 12 && stage < 13 && evals - scan_n >= k.max_eval) {
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
        // ---------------------------------------- (A) where F is needed
        // The stages a solve spends most steps in (0, 1, 2) come first; the
        // rest are grouped behind one test, which a wave whose lanes are all
        // in the common stages passes over at the cost of a branch.
        double h;
        double step_len = 0.0;                   // stage 2: the step to the new point
        if (stage <= 2) {
            if (stage == 2) {
                if (lo_pos < 0.0 && hi_neg == 0.01) {   // F(mindepth) <= 0, nothing positive above yet
                    stage = 8;                   // F(mindepth) again, then the certified scan
                    return false;
                }
                const double hm = hmax(c);
                const double den = fa - fb;
                double hn = (den != 0.0) ? hb - fb * (ha - hb) / den : 0.5 * (ha + hb);
                // Converged without another evaluation: the secant step points
                // to a downcrossing next to the point just evaluated (toward
                // lower h when F <= 0 there, higher when F > 0), inside the
                // bracket, and the secant error estimate |hn - hb| |hn - ha| / h^2
                // is within tol.  The root is the prediction hn; its outflow is
                // the secant interpolation of the outflows at ha and hb (an
                // error of the same second order).
                // F jumps at bankfull on a floodplain, where the error estimate
                // means nothing: a bracket across it (and the certified scan's,
                // which may end there) converges by steps instead.
                const bool smooth = !certified && !(c.twcc > 0.0 && fmin(ha, hb) <= c.bfd && fmax(ha, hb) > c.bfd);
                if (smooth && h_cur == hb && hn == hn && (fb > 0.0 ? hn >= hb : hn <= hb) &&
                    (lo_pos > 0.0 || neg_seen) && !(hn <= lo_pos) && !(neg_seen && hn > hi_neg) &&
                    fabs(hn - hb) <= 1e-6 * hb && fabs(hn - hb) * fabs(hn - ha) <= k.tol * hb * hb) {
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
                // Near the root the new point is the nearer grid point (inside
                // the bracket): it then serves as an end of the root's
                // canonical cell, whose F need not be evaluated again.
                if (fabs(hn - hb) <= kSnapRel * hb) {
                    const double g = canon_floor(hn), gn = canon_next(g);
                    const double hs = (hn - g <= gn - hn) ? g : gn;
                    if (hs > left && hs < hi_neg && hs != hb && hs != ha) hn = hs;
                }
                s_prev2 = s_prev;
                s_prev = fabs(hn - hb);
                step_len = fabs(hn - hb);
                h = hn;
            } else {
                if (stage == 1 && need_guess) return false;   // waiting for the wave's guess
                h = stage == 0 ? ha : hb;
            }
        } else {
            switch (stage) {
            case 3: h = c.bfd; break;            // a downcrossing below bankfull?
            case 12:                             // solve [mindepth, bankfull] from F(mindepth)
                lo_pos = -1.0;
                hi_neg = c.bfd;
                neg_seen = true;
                s_prev = s_prev2 = 1.0e30;
                ha = 0.01;
                hb = c.bfd;
                h = 0.01;
                break;
            case 4: h = 0.01; break;
            case 7: h = hi_neg; break;
            case 8: h = 0.01; break;             // F(mindepth) first: it may be <= 0
            case 10: h = lo_pos; break;          // F at the scan bracket's lower end
            case 13: h = ha; break;              // the canonical cell's lower end, then its upper end
            case 14: h = hb; break;              // its upper end
            case 15: h = ha; break;              // a new lower end after a move down
            case 16: h = c.bfd; break;           // the floodplain's jump inside the cell
            case 17: h = canon_above(c.bfd); break;
            default: return true;
            }
        }
        // ---------------------------------------- (B) the one evaluation
        const double f = eval(h, c, k);
        // ---------------------------------------- (C) what it means
        if (stage <= 2) {
            if (stage == 2) {
                ha = hb; fa = fb; hb = h; fb = f;
                qa = qb; qb = q_cur;
                // a converged step from either side of the bracket: points with
                // F <= 0 only approach from above (hn < hi_neg), so this is a
                // downcrossing too
                if ((lo_pos > 0.0 || neg_seen) &&
                    (f_at_noise(f) || step_len <= k.tol * h ||
                     (lo_pos > 0.0 && neg_seen && (hi_neg - lo_pos) <= k.tol * hi_neg))) {
                    if (!(h_cur == hi_neg || h_cur == lo_pos)) {
                        stage = 7;               // last point outside the bracket: take its right end
                        return false;
                    }
                    accept_root(c);
                }
                return stage == 5;
            }
            if (stage == 0) {
                fa = f;
                qa = q_cur;
                if (learned) need_guess = true;
                stage = 1;
                return false;
            }
            fb = f;                              // stage 1
            qb = q_cur;
            stage = 2;
            if (f_at_noise(fb) && (lo_pos > 0.0 || neg_seen)) accept_root(c);
            return stage == 5;
        }
        switch (stage) {
        case 3:
            inbank = true;
            if (f > 0.0) {                       // no upcrossing between: keep the root
                stage = canon_begin();
                return stage == 5;
            }
            fb = f;                              // F(bankfull) <= 0: the bracket's right end
            qb = q_cur;
            stage = 12;
            return false;
        case 12:
            fa = f;
            qa = q_cur;
            if (!(fa > 0.0)) {                   // F(mindepth) <= 0: keep the root
                stage = canon_begin();
                return stage == 5;
            }
            stage = 2;
            return false;
        case 4:
            root_h = 0.01;
            root_q = q_cur;
            have_root = true;
            stage = 5;
            return true;
        case 7:
            accept_root(c);
            return stage == 5;
        case 8:
            if (f > 0.0) {
                stage = 0;
            } else if constexpr (kDefer) {
                stage = 11;
            } else {
                begin_scan(0.01, false);
                fa = f;
            }
            return false;
        case 10:
            if (!(f > 0.0)) { accept_root(c); return stage == 5; }
            ha = lo_pos; fa = f; qa = q_cur;
            stage = 2;
            return false;
        case 13:
            fa = f; qa = q_cur;
            stage = 14;
            return false;
        case 14:
            fb = f; qb = q_cur;
            canon_decide(c);
            return stage == 5;
        case 15:
            fa = f; qa = q_cur;
            canon_decide(c);
            return stage == 5;
        case 16:                                 // F at bankfull inside the cell
            if (!(f > 0.0)) {
                hb = h; fb = f; qb = q_cur;
                canon_finish();
            } else {
                stage = 17;
            }
            return stage == 5;
        case 17:                                 // F just above bankfull
            if (!(f > 0.0)) {
                root_h = h; root_q = q_cur;
                stage = 5;
            } else {
                ha = h; fa = f; qa = q_cur;
                canon_finish();
            }
            return stage == 5;
        }
        return true;
    }

    // with_vel false: o.vel is left at 0 (not wanted for this step)
    TRMC_HD SolveOut finish(const Seg &c, bool with_vel = true) const
    {
        SolveOut o;
        o.iters = evals + scan_n;
        o.flags = (capped ? 1 : 0) | (fell_back ? 2 : 0) | ((noncanon && !trivial) ? 4 : 0);
        if (trivial) { o.q = 0.0; o.depth = 0.0; o.vel = 0.0; return o; }
        o.depth = root_h;
        o.vel = with_vel ? velocity(root_h, c) : 0.0;
        o.q = root_q;
        return o;
    }
};

// ---------------------------------------------- wave-parallel certified scan
// One certified solve by the 64 lanes of a wave together (the server waves of
// the dataflow kernel).  Each round splits the span ahead of the scan
// position a, [a, b], into 64 pieces (geometric when b > 1.5 a, else even);
// lane i evaluates F at its piece's right end and proves with the serial
// scan's enclosures (the direct one, then the mean-value one with F at both
// ends) that F has no point of the searched sign on its piece.  The first
// piece that is not proven is where the scan goes on:
//   * a sign change at its right end narrower than kScanRes relative ends the
//     search for that sign (the first positive point; after it the bracket's
//     right end); wider, the next round splits that piece;
//   * no sign change: at the resolution the piece is passed, as the serial
//     scan passes it; wider, the next round splits it.
// Every piece before the first unproven one is proven, so the scan returns
// the serial scan's leftmost downcrossing to the same resolution (not the
// same points).  Pieces end at bankfull, where F jumps; above it the
// floodplain enclosures apply.  Looking for F > 0, the search ends at
// mindepth once Qmn on [a, inf) exceeds every routed flow.
struct ScanWave {
    double a, fa, qa;   // position, F(a) (NaN: not known), outflow at a
    double b;           // end of the next round's span
    bool spos;          // searching F <= 0 after a positive point
    int rounds;
    int state;          // 0 scanning, 1 root at mindepth, 2 bracket [lo, hi], 3 root at 1e4,
                        // 4 root at a (the round cap while looking for F <= 0)
    double lo, flo, qlo, hi, fhi, qhi;
};

struct ScanPiece {
    double f, q;        // F and the outflow at the piece's right end
    bool proven;        // no point of the searched sign on the piece
    bool change;        // the searched sign at the right end
};

// the scan from mindepth, where F(0.01) = f001 <= 0 with outflow q001
TRMC_HD void scan_wave_begin(ScanWave &w, double f001, double q001)
{
    w.a = 0.01;
    w.fa = f001;
    w.qa = q001;
    w.b = 1.0e4;
    w.spos = false;
    w.rounds = 0;
    w.state = 0;
}

// the round's span end and spacing; false when the scan has ended here
TRMC_HD bool scan_wave_round(ScanWave &w, const Seg &c, double rmax, double &b_end, bool &geo)
{
    if (!w.spos) {
        if (w.a >= 1.0e4 || qmn_floor(w.a, c) > rmax + 1e-12 * (fabs(rmax) + 1.0)) { w.state = 1; return false; }
    } else if (w.a >= 1.0e4) {
        w.state = 3;
        return false;
    }
    b_end = fmin(w.b, 1.0e4);
    const bool has_flood = c.twcc > 0.0 && c.bfd > 0.01 && c.bfd < 1.0e4;
    if (has_flood && w.a < c.bfd && b_end > c.bfd) b_end = c.bfd;
    geo = b_end > 1.5 * w.a;
    return true;
}

// the left end of piece i (i = 64: the span's end), strictly increasing in i
TRMC_HD double scan_piece_x(const ScanWave &w, double b_end, bool geo, int i)
{
    if (i <= 0) return w.a;
    if (i >= 64) return b_end;
    const double t = (double)i * (1.0 / 64.0);
    return geo ? w.a * exp2(t * log2(b_end / w.a)) : w.a + (b_end - w.a) * t;
}

// F at the right end x1 of piece [x0, x1] (every lane)
TRMC_HD ScanPiece scan_piece_eval(double x1, bool spos, double dt, double qup, double quc, double qdp, double ql,
                                  const Seg &c)
{
    ScanPiece p;
    Coef C;
    double qmn;
    p.f = residual(x1, dt, qup, quc, qdp, ql, c, C, qmn);
    p.q = outflow(C, qup, quc, qdp);
    p.change = spos ? !(p.f > 0.0) : p.f > 0.0;
    p.proven = false;
    return p;
}

// the proof of piece [x0, x1] with F(x0) = f0 (NaN: not known)
TRMC_HD bool scan_piece_proof(double x0, double x1, double f0, const ScanPiece &p, bool spos, double dt, double qup,
                              double quc, double qdp, double ql, const Seg &c, double rmax)
{
    if (p.change) return false;
    if (!spos && qmn_floor(x0, c) > rmax + 1e-12 * (fabs(rmax) + 1.0)) return true;   // F < 0 on [x0, inf)
    const bool has_flood = c.twcc > 0.0 && c.bfd > 0.01 && c.bfd < 1.0e4;
    const bool flood = has_flood && x0 >= c.bfd;
    const Enclosure en = f_enclosure(x0, x1, flood, dt, qup, quc, qdp, ql, c);
    if (spos ? en.lo > 0.0 : !(en.hi > 0.0)) return true;
    const bool have_f0 = !(flood && x0 <= c.bfd) && f0 == f0;
    return slope_sign_holds(x0, x1, flood, f0, have_f0, p.f, spos, dt, qup, quc, qdp, ql, c);
}

// The round's outcome from the first unproven piece `first` (64: every piece
// proven): its ends x_first < x_next, F and outflow at x_first (f_left,
// q_left: F(a) for the first piece), its own right-end values, and the
// values at the span's end b_end (piece 63's right end).
TRMC_HD void scan_wave_combine(ScanWave &w, int first, double x_first, double x_next, double f_left, double q_left,
                               const ScanPiece &pf, double b_end, double f_end, double q_end)
{
    ++w.rounds;
    if (first >= 64) {                                   // all proven: on past the span, twice as far
        const double a0 = w.a;
        w.a = b_end;
        w.fa = f_end;
        w.qa = q_end;
        w.b = b_end + 2.0 * (b_end - a0);
        return;
    }
    const double width = x_next - x_first;
    w.a = x_first;
    w.fa = f_left;
    w.qa = q_left;
    if (width > kScanRes * x_first) {                    // split this piece next
        w.b = x_next;
        return;
    }
    if (pf.change && w.spos) {                           // the bracket [x_first, x_next]
        w.state = 2;
        w.lo = x_first; w.flo = f_left; w.qlo = q_left;
        w.hi = x_next; w.fhi = pf.f; w.qhi = pf.q;
        return;
    }
    if (pf.change) w.spos = true;                        // the first positive point
    w.a = x_next;                                        // (or an unproven piece at the resolution: passed)
    w.fa = pf.f;
    w.qa = pf.q;
    w.b = pf.change ? 1.0e4 : x_next + 256.0 * width;
}

// After the scan: the root and its outflow, with the secant on the bracket
// (state 2) as SecantConv solves it; returns the evaluations it took.
TRMC_HD int scan_wave_finish(const ScanWave &w, double qup, double quc, double qdp, double ql, const Seg &c,
                             const ConvCfg &k, double &h, double &q)
{
    SecantConv s;
    s.qup = qup; s.quc = quc; s.qdp = qdp; s.ql = ql;
    s.evals = 0;
    s.scan_n = 0;
    s.learned = false; s.need_guess = false; s.fell_back = false; s.capped = false;
    s.trivial = false; s.have_root = false;
    s.inbank = false;
    if (w.state != 2) {                                  // mindepth, the search's end, or the cap
        const double at = (w.state == 1) ? 0.01 : (w.state == 3) ? 1.0e4 : w.a;
        s.eval(at, c, k);
        h = at;
        q = s.q_cur;
        return s.evals;
    }
    s.certified = true;
    s.lo_pos = w.lo; s.hi_neg = w.hi; s.neg_seen = true;
    s.ha = w.lo; s.fa = w.flo; s.qa = w.qlo;
    s.hb = w.hi; s.fb = w.fhi; s.qb = w.qhi;
    s.h_cur = w.hi; s.q_cur = w.qhi; s.f_cur = w.fhi;
    s.s_prev = s.s_prev2 = 1.0e30;
    s.stage = 2;
    while (!s.step(c, k)) {
    }
    h = s.root_h;
    q = s.root_q;
    return s.evals;
}

// The wave-parallel scan with its 64 lanes run one after another (host
// builds and tests; the device runs the lanes side by side).
TRMC_HD int certified_solve_lanes(double qup, double quc, double qdp, double ql, const Seg &c, const ConvCfg &k,
                                  double &h, double &q)
{
    Coef C0;
    double qmn0;
    const double f001 = residual(0.01, k.dt, qup, quc, qdp, ql, c, C0, qmn0);
    const double q001 = outflow(C0, qup, quc, qdp);
    if (f001 > 0.0) return -1;                           // not a certified solve
    const double rmax = routed_max(k.dt, qup, quc, qdp, ql);
    ScanWave w;
    scan_wave_begin(w, f001, q001);
    int evals = 1;
    double b_end = 0.0;
    bool geo = false;
    while (w.state == 0 && w.rounds < kScanCap && scan_wave_round(w, c, rmax, b_end, geo)) {
        ScanPiece pc[64];
        double x[65];
        for (int i = 0; i <= 64; ++i) x[i] = scan_piece_x(w, b_end, geo, i);
        for (int i = 0; i < 64; ++i) pc[i] = scan_piece_eval(x[i + 1], w.spos, k.dt, qup, quc, qdp, ql, c);
        int first = 64;
        for (int i = 0; i < 64 && first == 64; ++i) {
            const double f0 = (i == 0) ? w.fa : pc[i - 1].f;
            if (!scan_piece_proof(x[i], x[i + 1], f0, pc[i], w.spos, k.dt, qup, quc, qdp, ql, c, rmax)) first = i;
        }
        evals += 1;
        const double fl = (first == 0 || first == 64) ? w.fa : pc[first - 1].f;
        const double ql_ = (first == 0 || first == 64) ? w.qa : pc[first - 1].q;
        const int f = first < 64 ? first : 63;
        scan_wave_combine(w, first, x[f], x[f + 1], fl, ql_, pc[f], b_end, pc[63].f, pc[63].q);
    }
    if (w.state == 0) w.state = w.spos ? 4 : 1;          // the round cap: as the serial scan's cap
    return evals + scan_wave_finish(w, qup, quc, qdp, ql, c, k, h, q);
}

// Whole CONVERGED solve from the reference bracket (no MLP).
TRMC_HD SolveOut solve_converged(double dt, double qup, double quc, double qdp, double ql, const Seg &c,
                                 double depthp, double tol, int max_eval, bool with_vel = true)
{
    const ConvCfg k = {dt, tol, max_eval, 0};
    SecantConv s;
    if (!s.start(qup, quc, qdp, ql, c, depthp, false))
        while (!s.step(c, k)) {
        }
    return s.finish(c, with_vel);
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
