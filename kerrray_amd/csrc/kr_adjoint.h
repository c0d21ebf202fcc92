// Discrete adjoint of the FP64 Dormand-Prince 5(4) photon path.
//
// The map differentiated is the one the adaptive integrator actually
// applied: the sequence of accepted steps y_{k+1} = Phi(y_k; h_k, a) with
// every step size h_k frozen at the value it had (rejected attempts leave no
// trace in it), started from the Bardeen image-plane state
// y_0(alpha, beta, a, theta_o) (benchmarks/rayset.py:158-165,
// geodesics/initial_conditions.py:50-112).  For the escape direction
// (theta_N, phi_N) of an escaping photon it returns the gradient of
// g_theta theta_N + g_phi phi_N with respect to (alpha, beta, spin, theta_o).
//
// Reverse sweep of one step (explicit Runge-Kutta, stages Y_i =
// y + h sum_{j<i} a_ij k_j, k_i = f(Y_i), y_new = y + h sum_i b_i k_i):
//   kbar_i = h b_i ybar_new,  ybar = ybar_new,
//   for i = 5 .. 0:  Ybar_i = J_f(Y_i)^T kbar_i;  ybar += Ybar_i;
//                    kbar_j += h a_ij Ybar_i (j < i);
//                    (p_t, p_phi, a)bar += (df/d(p_t, p_phi, a))^T kbar_i.
// J_f^T v is formed by one reverse sweep through the inverse metric (the same
// 15 terms as numba_kernel.py:111-154, with the spin a variable; metric_vjp)
// for the r, theta and a columns, and from the closed-form momentum
// derivatives of the Hamiltonian right-hand side (numba_kernel.py:158-197).
//
// Step Jacobians instead of a reverse sweep: t and phi enter no right-hand
// side, so a step's Jacobian with respect to (y, p_t, p_phi, a) is the
// identity plus an increment D with seven nonzero columns (r, theta, p_r,
// p_theta, p_t, p_phi, a).  D follows from the stage states alone, by
// forward-mode tangents through the stages:
//   Ydot_i = e + h sum_{j<i} a_ij kdot_j,  kdot_i = J_f(Y_i) Ydot_i (+ df/dp),
//   D = h sum_i b_i kdot_i,
// independently for every step, after which the cotangent recurrence is
//   ybar_k = ybar_{k+1} + D_k^T ybar_{k+1}  (active columns; t and phi keep
//   theirs),  (p_t, p_phi, a)bar += D_k^T ybar_{k+1}  (parameter columns).
//
// HIP-free: the kernels include it, and so does the host build of the
// kernel source used by the tests.
#pragma once

#include "kr_device.h"

namespace kr {
namespace ad {

// The derivative arithmetic (dual numbers, vector-Jacobian products, step
// Jacobians) is not bound to the reference's roundings, only the primal steps
// are: it may fuse a * b + c into one FMA.  The primal functions this file
// calls (rhs, dp5_step) and the escape tail on plain doubles lie outside
// these regions and keep the build's separate roundings.
#pragma clang fp contract(fast)

// Forward-mode dual number with N tangent directions.
template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
KR_HD Dual<N> seed(double v, int k) {
    Dual<N> x;
    x.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) x.d[i] = (i == k) ? 1.0 : 0.0;
    return x;
}

template <int N>
KR_HD Dual<N> constant(double v) {
    return seed<N>(v, -1);
}

template <int N>
KR_HD Dual<N> operator+(const Dual<N>& x, const Dual<N>& y) {
    Dual<N> z;
    z.v = x.v + y.v;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = x.d[i] + y.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> operator-(const Dual<N>& x, const Dual<N>& y) {
    Dual<N> z;
    z.v = x.v - y.v;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = x.d[i] - y.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> operator-(const Dual<N>& x) {
    Dual<N> z;
    z.v = -x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = -x.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> operator*(const Dual<N>& x, const Dual<N>& y) {
    Dual<N> z;
    z.v = x.v * y.v;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = x.d[i] * y.v + x.v * y.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> operator*(double c, const Dual<N>& x) {
    Dual<N> z;
    z.v = c * x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = c * x.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> operator+(double c, const Dual<N>& x) {
    Dual<N> z = x;
    z.v = c + x.v;
    return z;
}

template <int N>
KR_HD Dual<N> operator-(const Dual<N>& x, double c) {
    Dual<N> z = x;
    z.v = x.v - c;
    return z;
}

template <int N>
KR_HD Dual<N> operator/(const Dual<N>& x, const Dual<N>& y) {
    Dual<N> z;
    const double inv = 1.0 / y.v;
    z.v = x.v * inv;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = (x.d[i] - z.v * y.d[i]) * inv;
    return z;
}

template <int N>
KR_HD Dual<N> operator/(double c, const Dual<N>& y) {
    return constant<N>(c) / y;
}

template <int N>
KR_HD Dual<N> dsin(const Dual<N>& x) {
    Dual<N> z;
    double s, c;
    sin_cos(x.v, s, c);
    z.v = s;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = c * x.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> dcos(const Dual<N>& x) {
    Dual<N> z;
    double s, c;
    sin_cos(x.v, s, c);
    z.v = c;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = -s * x.d[i];
    return z;
}

template <int N>
KR_HD Dual<N> dsqrt(const Dual<N>& x) {
    Dual<N> z;
    z.v = sqrt(x.v);
    const double k = z.v > 0.0 ? 0.5 / z.v : 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) z.d[i] = k * x.d[i];
    return z;
}

#pragma clang fp contract(off)

KR_HD double dsin(double x) { return sin(x); }
KR_HD double dcos(double x) { return cos(x); }
KR_HD double dsqrt(double x) { return sqrt(x); }
KR_HD void dsincos(double x, double& s, double& c) { sin_cos(x, s, c); }

#pragma clang fp contract(fast)

template <int N>
KR_HD void dsincos(const Dual<N>& x, Dual<N>& s, Dual<N>& c) {
    double sv, cv;
    sin_cos(x.v, sv, cv);
    s.v = sv;
    c.v = cv;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        s.d[i] = cv * x.d[i];
        c.d[i] = -sv * x.d[i];
    }
}

// The inverse Kerr metric and its r and theta derivatives (the 15 terms of
// numba_kernel.py:111-154, in the same order) as dual numbers in (r, theta,
// a).  The same quantities as the reference's expressions, arranged for
// derivatives: three reciprocals (Sigma, Delta, sin^2) serve every quotient,
// and sin 2 theta = 2 sin cos.
template <int N>
KR_HD void metric_terms(const Dual<N>& r, const Dual<N>& th, const Dual<N>& a, double m, Dual<N> (&g)[15]) {
    using S = Dual<N>;
    S s, c;
    dsincos(th, s, c);
    const S s2 = s * s;
    const S sin2th = 2.0 * (s * c);
    const S r2 = r * r;
    const S a2 = a * a;
    const S sig = r2 + a2 * (c * c);
    const S dlt = r2 - (2.0 * m) * r + a2;
    const S r2a2 = r2 + a2;
    const S big_a = r2a2 * r2a2 - a2 * dlt * s2;
    const S sd = sig * dlt;
    const S ma2 = (2.0 * m) * a;
    const S i_sig = 1.0 / sig;
    const S i_dlt = 1.0 / dlt;
    const S i_s2 = 1.0 / s2;
    const S i_sd = i_sig * i_dlt;
    const S i_sd2 = i_sd * i_sd;
    const S i_sig2 = i_sig * i_sig;
    g[0] = -(big_a * i_sd);
    g[1] = -((ma2 * r) * i_sd);
    g[2] = dlt * i_sig;
    g[3] = i_sig;
    g[4] = (dlt - a2 * s2) * (i_sd * i_s2);
    const S sig_r = 2.0 * r;
    const S sig_th = -(a2 * sin2th);
    const S dlt_r = 2.0 * (r - m);
    const S a_r = 4.0 * (r * r2a2) - 2.0 * (a2 * (r - m) * s2);
    const S a_th = -(a2 * dlt * sin2th);
    const S sd_r = sig_r * dlt + sig * dlt_r;
    const S sd_th = sig_th * dlt;
    g[5] = -((a_r * sd - big_a * sd_r) * i_sd2);
    g[6] = -((ma2 * (sd - r * sd_r)) * i_sd2);
    g[7] = (dlt_r * sig - dlt * sig_r) * i_sig2;
    g[8] = -(sig_r * i_sig2);
    g[9] = -(sig_r * (i_sig2 * i_s2)) + (a2 * sd_r) * i_sd2;
    g[10] = -((a_th * sd - big_a * sd_th) * i_sd2);
    g[11] = (ma2 * r * sd_th) * i_sd2;
    g[12] = -((dlt * sig_th) * i_sig2);
    g[13] = -(sig_th * i_sig2);
    g[14] = -((sig_th * s2 + sig * sin2th) * (i_sig2 * (i_s2 * i_s2))) + (a2 * sd_th) * i_sd2;
}

// The 15 metric terms g of metric_terms at (r, theta, a) on plain doubles,
// and the cotangent of (r, theta, a) for the cotangent gb of g: one reverse
// sweep through the same expression graph (each intermediate's adjoint is
// the sum, over its uses, of the use's adjoint times the use's partial
// derivative), instead of three tangent directions carried through every
// operation.  The forward values are kept for the sweep: about 40 doubles,
// against four per value with Dual<3>.
KR_HD void metric_vjp(double r, double th, double a, double m, const double (&gb)[15], double (&g)[15],
                      double& r_b, double& th_b, double& a_b) {
    // Forward (metric_terms' operations, values only).
    double s, c;
    sin_cos(th, s, c);
    const double s2 = s * s;
    const double sin2th = 2.0 * (s * c);
    const double r2 = r * r;
    const double a2 = a * a;
    const double cc = c * c;
    const double sig = r2 + a2 * cc;
    const double dlt = r2 - (2.0 * m) * r + a2;
    const double r2a2 = r2 + a2;
    const double a2dlt = a2 * dlt;
    const double big_a = r2a2 * r2a2 - a2dlt * s2;
    const double sd = sig * dlt;
    const double ma2 = (2.0 * m) * a;
    const double i_sig = 1.0 / sig;
    const double i_dlt = 1.0 / dlt;
    const double i_s2 = 1.0 / s2;
    const double i_sd = i_sig * i_dlt;
    const double i_sd2 = i_sd * i_sd;
    const double i_sig2 = i_sig * i_sig;
    const double ma2r = ma2 * r;
    const double u4 = dlt - a2 * s2, w4 = i_sd * i_s2;
    const double sig_r = 2.0 * r;
    const double sig_th = -(a2 * sin2th);
    const double dlt_r = 2.0 * (r - m);
    const double rm = r - m;
    const double p_ar = a2 * rm;
    const double a_r = 4.0 * (r * r2a2) - 2.0 * (p_ar * s2);
    const double a_th = -(a2dlt * sin2th);
    const double sd_r = sig_r * dlt + sig * dlt_r;
    const double sd_th = sig_th * dlt;
    const double u5 = a_r * sd - big_a * sd_r;
    const double u6 = sd - r * sd_r;
    const double v6 = ma2 * u6;
    const double u7 = dlt_r * sig - dlt * sig_r;
    const double w9 = i_sig2 * i_s2;
    const double z9 = a2 * sd_r;
    const double u10 = a_th * sd - big_a * sd_th;
    const double u11 = ma2r * sd_th;
    const double u12 = dlt * sig_th;
    const double u14 = sig_th * s2 + sig * sin2th;
    const double q14 = i_s2 * i_s2;
    const double w14 = i_sig2 * q14;
    const double z14 = a2 * sd_th;
    g[0] = -(big_a * i_sd);
    g[1] = -(ma2r * i_sd);
    g[2] = dlt * i_sig;
    g[3] = i_sig;
    g[4] = u4 * w4;
    g[5] = -(u5 * i_sd2);
    g[6] = -(v6 * i_sd2);
    g[7] = u7 * i_sig2;
    g[8] = -(sig_r * i_sig2);
    g[9] = -(sig_r * w9) + z9 * i_sd2;
    g[10] = -(u10 * i_sd2);
    g[11] = u11 * i_sd2;
    g[12] = -(u12 * i_sig2);
    g[13] = -(sig_th * i_sig2);
    g[14] = -(u14 * w14) + z14 * i_sd2;
    // Reverse: the uses of each intermediate, last definitions first.
    double big_a_b = -gb[0] * i_sd;
    double i_sd_b = -gb[0] * big_a - gb[1] * ma2r;
    double ma2r_b = -gb[1] * i_sd;
    double dlt_b = gb[2] * i_sig;
    double i_sig_b = gb[2] * dlt + gb[3];
    const double u4_b = gb[4] * w4, w4_b = gb[4] * u4;
    dlt_b += u4_b;
    double a2_b = -u4_b * s2;
    double s2_b = -u4_b * a2;
    i_sd_b += w4_b * i_s2;
    double i_s2_b = w4_b * i_sd;
    const double u5_b = -gb[5] * i_sd2;
    double i_sd2_b = -gb[5] * u5 - gb[6] * v6 + gb[9] * z9 - gb[10] * u10 + gb[11] * u11 + gb[14] * z14;
    double a_r_b = u5_b * sd;
    double sd_b = u5_b * a_r;
    big_a_b -= u5_b * sd_r;
    double sd_r_b = -u5_b * big_a;
    const double v6_b = -gb[6] * i_sd2;
    double ma2_b = v6_b * u6;
    const double u6_b = v6_b * ma2;
    sd_b += u6_b;
    double r_bb = -u6_b * sd_r;
    sd_r_b -= u6_b * r;
    const double u7_b = gb[7] * i_sig2;
    double i_sig2_b = gb[7] * u7 - gb[8] * sig_r - gb[12] * u12 - gb[13] * sig_th;
    double dlt_r_b = u7_b * sig;
    double sig_b = u7_b * dlt_r;
    dlt_b -= u7_b * sig_r;
    double sig_r_b = -u7_b * dlt - gb[8] * i_sig2 - gb[9] * w9;
    const double w9_b = -gb[9] * sig_r;
    i_sig2_b += w9_b * i_s2;
    i_s2_b += w9_b * i_sig2;
    const double z9_b = gb[9] * i_sd2;
    a2_b += z9_b * sd_r;
    sd_r_b += z9_b * a2;
    const double u10_b = -gb[10] * i_sd2;
    double a_th_b = u10_b * sd;
    sd_b += u10_b * a_th;
    big_a_b -= u10_b * sd_th;
    double sd_th_b = -u10_b * big_a;
    const double u11_b = gb[11] * i_sd2;
    ma2r_b += u11_b * sd_th;
    sd_th_b += u11_b * ma2r;
    const double u12_b = -gb[12] * i_sig2;
    dlt_b += u12_b * sig_th;
    double sig_th_b = u12_b * dlt - gb[13] * i_sig2;
    const double u14_b = -gb[14] * w14, w14_b = -gb[14] * u14;
    sig_th_b += u14_b * s2;
    s2_b += u14_b * sig_th;
    sig_b += u14_b * sin2th;
    double sin2th_b = u14_b * sig;
    i_sig2_b += w14_b * q14;
    i_s2_b += 2.0 * (w14_b * i_sig2) * i_s2;
    const double z14_b = gb[14] * i_sd2;
    a2_b += z14_b * sd_th;
    sd_th_b += z14_b * a2;
    // sd_th = sig_th dlt; sd_r = sig_r dlt + sig dlt_r
    sig_th_b += sd_th_b * dlt;
    dlt_b += sd_th_b * sig_th + sd_r_b * sig_r;
    sig_r_b += sd_r_b * dlt;
    sig_b += sd_r_b * dlt_r;
    dlt_r_b += sd_r_b * sig;
    // a_th = -(a2 dlt) sin2th; a_r = 4 r r2a2 - 2 (a2 (r - m)) s2
    const double a2dlt_b = -a_th_b * sin2th - big_a_b * s2;
    sin2th_b -= a_th_b * a2dlt;
    r_bb += 4.0 * a_r_b * r2a2;
    double r2a2_b = 4.0 * a_r_b * r + 2.0 * big_a_b * r2a2;
    const double p_ar_b = -2.0 * a_r_b * s2;
    s2_b -= 2.0 * a_r_b * p_ar + big_a_b * a2dlt;
    a2_b += p_ar_b * rm;
    r_bb += p_ar_b * a2 + 2.0 * dlt_r_b + 2.0 * sig_r_b;
    // sig_th = -(a2 sin2th)
    a2_b -= sig_th_b * sin2th;
    sin2th_b -= sig_th_b * a2;
    // i_sig2, i_sd2, i_sd, ma2 r
    i_sig_b += 2.0 * i_sig2_b * i_sig;
    i_sd_b += 2.0 * i_sd2_b * i_sd;
    i_sig_b += i_sd_b * i_dlt;
    const double i_dlt_b = i_sd_b * i_sig;
    ma2_b += ma2r_b * r;
    r_bb += ma2r_b * ma2;
    // the reciprocals
    s2_b -= i_s2_b * i_s2 * i_s2;
    dlt_b -= i_dlt_b * i_dlt * i_dlt;
    sig_b -= i_sig_b * i_sig * i_sig;
    // sd = sig dlt; a2 dlt; r2a2 = r2 + a2; dlt = r2 - 2 m r + a2; sig = r2 + a2 c^2
    sig_b += sd_b * dlt;
    dlt_b += sd_b * sig + a2dlt_b * a2;
    a2_b += a2dlt_b * dlt + r2a2_b + dlt_b + sig_b * cc;
    const double r2_b = r2a2_b + dlt_b + sig_b;
    r_bb -= (2.0 * m) * dlt_b;
    double c_b = 2.0 * (sig_b * a2) * c;
    // a2 = a a; r2 = r r; sin2th = 2 s c; s2 = s s; (s, c) = (sin th, cos th)
    a_b = (2.0 * m) * ma2_b + 2.0 * a * a2_b;
    r_b = r_bb + 2.0 * r * r2_b;
    double s_b = 2.0 * sin2th_b * c + 2.0 * s2_b * s;
    c_b += 2.0 * sin2th_b * s;
    th_b = s_b * c - c_b * s;
}

// v^T df, for the right-hand side f(y; p_t, p_phi, a) of the six integrated
// components (t, r, theta, phi, p_r, p_theta) at the state Y.  Accumulates
// into ybar (six components), ptbar, pphbar and abar.  f is linear in the
// metric terms g for given momenta, so v^T df/d(r, theta, a) is
// gbar^T dg/d(r, theta, a) with gbar = (df/dg)^T v, which metric_vjp takes
// in one reverse sweep.
KR_HD void rhs_vjp(double m, double a, const double (&Y)[NA], double pt, double pph, const double (&v)[NA],
                   double (&ybar)[NA], double& ptbar, double& pphbar, double& abar) {
    const double pr = Y[4], pth = Y[5];
    const double pt2 = pt * pt, ptpph = pt * pph, pr2 = pr * pr, pth2 = pth * pth, pph2 = pph * pph;
    // The polar-axis guard of the primal right-hand side zeroes g[4], g[9]
    // and g[14] (values and derivatives) when p_phi = 0.
    const double keep = pph == 0.0 ? 0.0 : 1.0;
    const double gb[15] = {v[0] * pt,
                           v[0] * pph + v[3] * pt,
                           v[1] * pr,
                           v[2] * pth,
                           keep * (v[3] * pph),
                           -0.5 * v[4] * pt2,
                           -v[4] * ptpph,
                           -0.5 * v[4] * pr2,
                           -0.5 * v[4] * pth2,
                           keep * (-0.5 * v[4] * pph2),
                           -0.5 * v[5] * pt2,
                           -v[5] * ptpph,
                           -0.5 * v[5] * pr2,
                           -0.5 * v[5] * pth2,
                           keep * (-0.5 * v[5] * pph2)};
    double g[15], r_b, th_b, a_b;
    metric_vjp(Y[1], Y[2], a, m, gb, g, r_b, th_b, a_b);
    if (pph == 0.0) {
        g[4] = 0.0;
        g[9] = 0.0;
        g[14] = 0.0;
    }
    ybar[1] += r_b;
    ybar[2] += th_b;
    abar += a_b;
    ybar[4] += v[1] * g[2] - v[4] * g[7] * pr - v[5] * g[12] * pr;
    ybar[5] += v[2] * g[3] - v[4] * g[8] * pth - v[5] * g[13] * pth;
    ptbar += v[0] * g[0] + v[3] * g[1] - v[4] * (g[5] * pt + g[6] * pph) - v[5] * (g[10] * pt + g[11] * pph);
    pphbar += v[0] * g[1] + v[3] * g[4] - v[4] * (g[6] * pt + g[9] * pph) - v[5] * (g[11] * pt + g[14] * pph);
}

// One accepted step of the primal, recomputed: y <- Phi(y; h).  Uses the
// primal's own FP64 stage code, so a recomputed trajectory reproduces the
// recorded one bit for bit.
KR_HD void step_forward(const Params& P, double (&y)[NA], double pt, double pph, double h) {
    double k0[NA], y_new[NA], err[NA], k6[NA], s, c;
    rhs(P, y, pt, pph, k0, s, c);
    dp5_step(P, y, pt, pph, h, k0, y_new, err, k6, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i) y[i] = y_new[i];
}

// Reverse sweep of one accepted step at the state y with step h: ybar (the
// cotangent of y_new on entry) becomes the cotangent of y.
KR_HD void step_adjoint(const Params& P, double m, double a, const double (&y)[NA], double pt, double pph, double h,
                        double (&ybar)[NA], double& ptbar, double& pphbar, double& abar) {
    using namespace dp;
    // Stage slopes k_0 .. k_5 of the step (k_6 does not enter y_new).
    double k[6][NA], ys[NA], s, c;
    rhs(P, y, pt, pph, k[0], s, c);
    const double A[6][5] = {{0, 0, 0, 0, 0},         {a10, 0, 0, 0, 0},     {a20, a21, 0, 0, 0},
                            {a30, a31, a32, 0, 0},   {a40, a41, a42, a43, 0}, {a50, a51, a52, a53, a54}};
    const double b[6] = {a60, 0.0, a62, a63, a64, a65};
#pragma unroll
    for (int st = 1; st < 6; ++st) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < st; ++j) acc += A[st][j] * k[j][i];
            ys[i] = y[i] + h * acc;
        }
        rhs(P, ys, pt, pph, k[st], s, c);
    }
    double kbar[6][NA];
#pragma unroll
    for (int st = 0; st < 6; ++st)
#pragma unroll
        for (int i = 0; i < NA; ++i) kbar[st][i] = h * b[st] * ybar[i];
#pragma unroll
    for (int st = 5; st >= 0; --st) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < st; ++j) acc += A[st][j] * k[j][i];
            ys[i] = y[i] + h * acc;
        }
        double Ybar[NA] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        rhs_vjp(m, a, ys, pt, pph, kbar[st], Ybar, ptbar, pphbar, abar);
#pragma unroll
        for (int i = 0; i < NA; ++i) ybar[i] += Ybar[i];
#pragma unroll
        for (int j = 0; j < st; ++j)
#pragma unroll
            for (int i = 0; i < NA; ++i) kbar[j][i] += h * A[st][j] * Ybar[i];
    }
}

// ---------------------------------------------------------------------------
// Step Jacobians.  Columns of D: the active state components r, theta, p_r,
// p_theta (integrated-state indices 1, 2, 4, 5), then p_t, p_phi, a.  D[i][c]
// is d(y_new[i] - y[i]) / d x_c.
constexpr int JC = 7;
constexpr int JW = NA * JC;  // entries of one step's D, row-major
constexpr int ACTIVE[4] = {1, 2, 4, 5};

// The right-hand side's partial derivatives at one state: every row depends
// on r, theta and a; the momenta enter only the rows listed.
struct RhsJac {
    double r[NA], th[NA], a[NA];
    double pr1, pr4, pr5;           // d f_1, f_4, f_5 / d p_r
    double pth2, pth4, pth5;        // d f_2, f_4, f_5 / d p_theta
    double pt0, pt3, pt4, pt5;      // d f_0, f_3, f_4, f_5 / d p_t
    double pph0, pph3, pph4, pph5;  // d f_0, f_3, f_4, f_5 / d p_phi
};

// J_f at the state Y (the transposes of rhs_vjp's products, entry by entry).
KR_HD void rhs_jacobian(double m, double a, const double (&Y)[NA], double pt, double pph, RhsJac& J) {
    Dual<3> g[15];
    metric_terms(seed<3>(Y[1], 0), seed<3>(Y[2], 1), seed<3>(a, 2), m, g);
    if (pph == 0.0) {  // polar-axis guard of the primal right-hand side
        g[4] = constant<3>(0.0);
        g[9] = constant<3>(0.0);
        g[14] = constant<3>(0.0);
    }
    const double pr = Y[4], pth = Y[5];
    const double pt2 = pt * pt, ptpph = pt * pph, pr2 = pr * pr, pth2 = pth * pth, pph2 = pph * pph;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double* o = k == 0 ? J.r : (k == 1 ? J.th : J.a);
        o[0] = g[0].d[k] * pt + g[1].d[k] * pph;
        o[1] = g[2].d[k] * pr;
        o[2] = g[3].d[k] * pth;
        o[3] = g[1].d[k] * pt + g[4].d[k] * pph;
        o[4] = -0.5 * (g[5].d[k] * pt2 + 2.0 * g[6].d[k] * ptpph + g[7].d[k] * pr2 + g[8].d[k] * pth2 +
                       g[9].d[k] * pph2);
        o[5] = -0.5 * (g[10].d[k] * pt2 + 2.0 * g[11].d[k] * ptpph + g[12].d[k] * pr2 + g[13].d[k] * pth2 +
                       g[14].d[k] * pph2);
    }
    J.pr1 = g[2].v;
    J.pr4 = -(g[7].v * pr);
    J.pr5 = -(g[12].v * pr);
    J.pth2 = g[3].v;
    J.pth4 = -(g[8].v * pth);
    J.pth5 = -(g[13].v * pth);
    J.pt0 = g[0].v;
    J.pt3 = g[1].v;
    J.pt4 = -(g[5].v * pt + g[6].v * pph);
    J.pt5 = -(g[10].v * pt + g[11].v * pph);
    J.pph0 = g[1].v;
    J.pph3 = g[4].v;
    J.pph4 = -(g[6].v * pt + g[9].v * pph);
    J.pph5 = -(g[11].v * pt + g[14].v * pph);
}

// Columns [C0, C0 + NC) of one step's increment D from its six stage states:
// stage(s, Y) loads stage s's state into Y; out(i, c, v) receives D[i][c].
template <int C0, int NC, class StageIn, class Out>
KR_HD void step_increment(double m, double a, double pt, double pph, double h, const StageIn& stage, const Out& out) {
    using namespace dp;
    const double A[6][5] = {{0, 0, 0, 0, 0},         {a10, 0, 0, 0, 0},     {a20, a21, 0, 0, 0},
                            {a30, a31, a32, 0, 0},   {a40, a41, a42, a43, 0}, {a50, a51, a52, a53, a54}};
    const double b[6] = {a60, 0.0, a62, a63, a64, a65};
    double kact[5][4][NC];  // active rows of the stage tangents 0..4
    double acc[NA][NC];     // sum_i b_i kdot_i
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[i][c] = 0.0;
    double dep = 0.0;  // a result of the previous stage
#pragma unroll
    for (int st = 0; st < 6; ++st) {
        double Y[NA];
        stage(st, Y);
        // One stage's metric derivatives at a time: the stages are
        // independent until their tangents combine, and interleaved they
        // would hold several stages' registers at once.
        KR_AFTER(dep, Y[0], Y[1], Y[2], Y[3], Y[4], Y[5]);
        RhsJac J;
        rhs_jacobian(m, a, Y, pt, pph, J);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int col = C0 + c;
            // Tangent of the stage state (active rows) along column col.
            double d[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double sum = 0.0;
#pragma unroll
                for (int j = 0; j < st; ++j) sum += A[st][j] * kact[j][r][c];
                const double base = (col == r) ? 1.0 : 0.0;
                d[r] = st == 0 ? base : base + h * sum;
            }
            double k[NA];
#pragma unroll
            for (int i = 0; i < NA; ++i) k[i] = J.r[i] * d[0] + J.th[i] * d[1];
            k[1] += J.pr1 * d[2];
            k[4] += J.pr4 * d[2] + J.pth4 * d[3];
            k[5] += J.pr5 * d[2] + J.pth5 * d[3];
            k[2] += J.pth2 * d[3];
            if (col == 4) {
                k[0] += J.pt0;
                k[3] += J.pt3;
                k[4] += J.pt4;
                k[5] += J.pt5;
            } else if (col == 5) {
                k[0] += J.pph0;
                k[3] += J.pph3;
                k[4] += J.pph4;
                k[5] += J.pph5;
            } else if (col == 6) {
#pragma unroll
                for (int i = 0; i < NA; ++i) k[i] += J.a[i];
            }
            if (st < 5) {
#pragma unroll
                for (int r = 0; r < 4; ++r) kact[st][r][c] = k[ACTIVE[r]];
            }
            if (st != 1) {
#pragma unroll
                for (int i = 0; i < NA; ++i) acc[i][c] += b[st] * k[i];
            }
            if (c == 0) dep = k[0] + k[5];
        }
    }
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int c = 0; c < NC; ++c) out(i, C0 + c, h * acc[i][c]);
}

// The exact FP64 Dormand-Prince attempt (the primal's own dp5_step, so the
// trajectory is the traced one bit for bit) handing its stage states to sink.
template <class Sink>
struct StageStepper {
    Sink sink;
    KR_HD bool operator()(const Params& P, const RayState<double>& R, double h_eff, double (&y_new)[NA],
                          double (&err)[NA], double (&k6)[NA], double& s_new, double& c_new) const {
        dp5_step(P, R.y, R.p_t, R.p_ph, h_eff, R.f, y_new, err, k6, s_new, c_new, sink);
        return true;
    }
};

// One step of the cotangent recurrence with the step's increment D (entry
// e = i * JC + c from d(e)).
template <class In>
KR_HD void increment_vjp(const In& d, double (&ybar)[NA], double& ptbar, double& pphbar, double& abar) {
    double v[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) v[i] = ybar[i];
    double s[JC];
#pragma unroll
    for (int c = 0; c < JC; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i) acc += d(i * JC + c) * v[i];
        s[c] = acc;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) ybar[ACTIVE[r]] += s[r];
    ptbar += s[4];
    pphbar += s[5];
    abar += s[6];
}

// Bardeen image-plane state (benchmarks/rayset.py:158-165 with E = 1,
// inward) as a function of (alpha, beta, a, theta_o).  p_theta is written
// as -beta: with Theta(theta_o) = beta^2 and sign_theta = -1 for beta >= 0,
// +1 otherwise, sign_theta sqrt(Theta) = -beta exactly.
template <int N>
KR_HD void bardeen_state_of(const Dual<N>& alpha, const Dual<N>& beta, const Dual<N>& a, const Dual<N>& th_o,
                            double m, double r_o, Dual<N> (&y)[8]) {
    using S = Dual<N>;
    S sin_o, cos_o;
    dsincos(th_o, sin_o, cos_o);
    const S xi = -(alpha * sin_o);
    const S cot = cos_o / sin_o;
    const S a2 = a * a;
    const S eta = beta * beta - a2 * (cos_o * cos_o) + xi * xi * (cot * cot);
    const S term1 = (r_o * r_o) + a2 - a * xi;
    const S d = xi - a;
    const S dlt = (r_o * r_o - 2.0 * m * r_o) + a2;
    const S big_r = term1 * term1 - dlt * (eta + d * d);
    y[0] = constant<N>(0.0);
    y[1] = constant<N>(r_o);
    y[2] = th_o;
    y[3] = constant<N>(0.0);
    y[4] = constant<N>(-1.0);
    y[5] = -(dsqrt(big_r) / dlt);
    y[6] = -beta;
    y[7] = xi;
}


#pragma clang fp contract(off)

// ---------------------------------------------------------------------------
// Escape direction at infinity.  An escaping photon ends at some r_N >=
// r_esc, where theta and phi are still changing (d theta / dr ~ p_theta /
// r^2), so the final state's angles depend on where the last step landed.
// The escape direction is therefore taken at r = infinity: from the final
// state, (theta, phi, p_theta) are integrated in u = 1/r from u_N = 1/r_N to
// 0 with the geodesic equations divided by du/dlambda, written in terms of
//   Sigma^ = u^2 Sigma = 1 + a^2 u^2 cos^2,  Delta^ = u^2 Delta = 1 - 2 M u + a^2 u^2,
//   A^ = u^4 A = (1 + a^2 u^2)^2 - a^2 u^2 Delta^ sin^2,
// in which every term is finite at u = 0:
//   d theta/du = -p_theta / (Delta^ p_r),
//   d phi/du   = (2 M a u p_t - (Delta^ - a^2 u^2 sin^2) p_phi / sin^2) / (Delta^2 p_r),
//   d p_theta/du = Sigma^ (T1 p_t^2 + 2 T2 p_t p_phi + T3 p_r^2 + T4 p_theta^2 + T5 p_phi^2) / (2 Delta^ p_r),
// with T_i = d_theta g^{..} / u^2 and p_r > 0 from the null condition
//   p_r^2 = (A^ p_t^2 + 4 M a u^3 p_t p_phi - u^2 Delta^ p_theta^2
//            - u^2 (Delta^ - a^2 u^2 sin^2) p_phi^2 / sin^2) / Delta^2.
// Classical RK4 with TAIL_STEPS steps in u.
constexpr int TAIL_STEPS = 16;

template <class S>
KR_HD void tail_rhs(const S& u, const S& th, const S& pth, const S& pt, const S& pph, const S& a, double m,
                    S& dth, S& dph, S& dpth) {
    S s, c;
    dsincos(th, s, c);
    const S s2 = s * s, sc = s * c;
    const S u2 = u * u, a2 = a * a;
    const S a2u2 = a2 * u2;
    const S sg = 1.0 + a2u2 * (c * c);
    const S dl = 1.0 + (a2u2 - (2.0 * m) * u);
    const S one_a2u2 = 1.0 + a2u2;
    const S ah = one_a2u2 * one_a2u2 - a2u2 * dl * s2;
    const S dphi_core = dl - a2u2 * s2;
    const S mau3 = (4.0 * m) * (a * u2 * u);
    const S pr2 = (ah * pt * pt + mau3 * pt * pph - u2 * dl * pth * pth - u2 * dphi_core * pph * pph / s2) / (dl * dl);
    const S pr = dsqrt(pr2);
    dth = -(pth / (dl * pr));
    dph = ((2.0 * m) * (a * u * pt) - dphi_core * pph / s2) / (dl * dl * pr);
    const S sg2 = sg * sg;
    const S t1 = 2.0 * (a2 * sc * (dl * sg - ah)) / (sg2 * dl);
    const S t2 = -((4.0 * m) * (a2 * a * u2 * u * sc)) / (sg2 * dl);
    const S t3 = 2.0 * (a2 * sc * dl) / sg2;
    const S t4 = 2.0 * (a2u2 * sc) / sg2;
    const S t5 = 2.0 * (sc * (a2u2 * s2 - sg)) / (sg2 * (s2 * s2)) - 2.0 * (a2u2 * a2u2 * sc) / (sg2 * dl);
    dpth = 0.5 * (sg * (t1 * pt * pt + 2.0 * (t2 * pt * pph) + t3 * pr2 + t4 * pth * pth + t5 * pph * pph)) / (dl * pr);
}

// (theta, phi) at r = infinity from the state (r, theta, phi, p_theta) with
// constants p_t, p_phi and spin parameter a.
template <class S>
KR_HD void escape_tail(const S& r, const S& th0, const S& ph0, const S& pth0, const S& pt, const S& pph, const S& a,
                       double m, S& th_inf, S& ph_inf) {
    S u = 1.0 / r;
    const S h = (-1.0 / static_cast<double>(TAIL_STEPS)) * u;
    S th = th0, ph = ph0, pth = pth0;
    for (int k = 0; k < TAIL_STEPS; ++k) {
        S k1t, k1p, k1q, k2t, k2p, k2q, k3t, k3p, k3q, k4t, k4p, k4q;
        tail_rhs(u, th, pth, pt, pph, a, m, k1t, k1p, k1q);
        const S hh = 0.5 * h;
        tail_rhs(u + hh, th + hh * k1t, pth + hh * k1q, pt, pph, a, m, k2t, k2p, k2q);
        tail_rhs(u + hh, th + hh * k2t, pth + hh * k2q, pt, pph, a, m, k3t, k3p, k3q);
        const S u_next = (k == TAIL_STEPS - 1) ? 0.0 * u : u + h;
        tail_rhs(u_next, th + h * k3t, pth + h * k3q, pt, pph, a, m, k4t, k4p, k4q);
        const S w = (1.0 / 6.0) * h;
        th = th + w * (k1t + 2.0 * k2t + 2.0 * k3t + k4t);
        ph = ph + w * (k1p + 2.0 * k2p + 2.0 * k3p + k4p);
        pth = pth + w * (k1q + 2.0 * k2q + 2.0 * k3q + k4q);
        u = u_next;
    }
    th_inf = th;
    ph_inf = ph;
}

// Jacobian of the escape direction at infinity (row 0: theta_inf, row 1:
// phi_inf) with respect to the final state (the NA integrated components:
// t and p_r do not enter, phi enters additively) and p_t, p_phi, a
// (columns NA .. NA + 2).
KR_HD void escape_tail_jacobian(double m, double a, const double (&y)[NA], double pt, double pph,
                                double (&J)[2][NA + 3]) {
    using D = Dual<5>;
    const D r = seed<5>(y[1], 0), th = seed<5>(y[2], 1), pth = seed<5>(y[5], 2), dpt = seed<5>(pt, 3),
            dpph = seed<5>(pph, 4), da = constant<5>(a);
    D out[2];
    escape_tail(r, th, constant<5>(y[3]), pth, dpt, dpph, da, m, out[0], out[1]);
    // d/da through the tail, by a separate one-direction pass.
    using E = Dual<1>;
    E out_a[2];
    escape_tail(constant<1>(y[1]), constant<1>(y[2]), constant<1>(y[3]), constant<1>(y[5]), constant<1>(pt),
                constant<1>(pph), seed<1>(a, 0), m, out_a[0], out_a[1]);
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        J[o][0] = 0.0;
        J[o][1] = out[o].d[0];
        J[o][2] = out[o].d[1];
        J[o][3] = o == 1 ? 1.0 : 0.0;
        J[o][4] = 0.0;
        J[o][5] = out[o].d[2];
        J[o][6] = out[o].d[3];
        J[o][7] = out[o].d[4];
        J[o][8] = out_a[o].d[0];
    }
}

// Cotangent of the final state for the cotangent (g_theta, g_phi) of the
// escape direction at infinity: fills ybar (r, theta, phi, p_theta; p_r and
// t do not enter) and the p_t, p_phi and a cotangents.
KR_HD void escape_tail_vjp(double m, double a, const double (&y)[NA], double pt, double pph, double g_theta,
                           double g_phi, double (&ybar)[NA], double& ptbar, double& pphbar, double& abar) {
    double J[2][NA + 3];
    escape_tail_jacobian(m, a, y, pt, pph, J);
    double v[NA + 3];
#pragma unroll
    for (int k = 0; k < NA + 3; ++k) v[k] = g_theta * J[0][k] + g_phi * J[1][k];
#pragma unroll
    for (int c = 0; c < NA; ++c) ybar[c] = v[c];
    ptbar = v[NA];
    pphbar = v[NA + 1];
    abar = v[NA + 2];
}

}  // namespace ad
}  // namespace kr
