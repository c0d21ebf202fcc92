// Device-side building blocks of the per-photon null-geodesic integrator.
//
// One lane integrates one photon; its whole state lives in VGPRs.  Every
// function reproduces the reference's scalar Numba kernel
// (src/kerrray/raytracing/backends/numba_kernel.py) in the same
// floating-point operation order; the library is compiled with
// -ffp-contract=off so no multiply-add is fused.  The working type T is
// double (the reference's float64) or float (its float32 mode, where the
// metric is still evaluated in double and cast, numba_kernel.py:165-167).
//
// Two exact simplifications relative to the reference's eight-component
// loops (both leave every output bit unchanged):
//  * dp_t/dlambda and dp_phi/dlambda are identically zero
//    (numba_kernel.py:185,197), so every stage value of p_t and p_phi is
//    y + h * (+0) = y + 0.0 (the first tableau coefficient of every stage row
//    is positive, so the zero sum is +0); only the six components
//    t, r, theta, phi, p_r, p_theta are integrated and p_t, p_phi are
//    carried as y + 0.0.
//  * the error-norm terms of p_t and p_phi are +0, so NumPy's pairwise
//    sum ((q0+q1)+(q2+q3))+((q4+q5)+(q6+q7)) equals ((q0+q1)+(q2+q3))+(q5+q6)
//    bit for bit.
// sin and cos of theta computed inside the metric of the last stage (at the
// accepted state) are reused by the classifier and the Carter diagnostic,
// which evaluate exactly the same functions of exactly the same argument.
#pragma once

#include <stdint.h>

// KR_HOST_BUILD compiles these functions as plain host C++ (the test suite
// builds them with g++ against glibc to check the kernel logic bit for bit
// against the oracle on the CPU); otherwise they are HIP device functions.
#ifdef KR_HOST_BUILD
#include <math.h>

#include <cmath>
#define KR_HD inline
using std::cos;
using std::fabs;
using std::isfinite;
using std::pow;
using std::sin;
using std::sqrt;
#else
#include <hip/hip_runtime.h>
#define KR_HD __host__ __device__ __forceinline__
#endif

// KR_AFTER(dep, x0..x5): in device code, makes the six values x0..x5 depend
// on dep (an empty asm statement that claims to read dep and rewrite them;
// no instruction is emitted), so no work on x can start before dep is
// computed.  Serialises independent computations that the compiler would
// otherwise interleave, and the registers they hold.  No effect on the host.
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
#define KR_AFTER(dep, x0, x1, x2, x3, x4, x5) ((void)0)
#else
#define KR_AFTER(dep, x0, x1, x2, x3, x4, x5) \
    asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5) : "v"(dep))
#endif

namespace kr {

// Flag bits a kernel ORs into the caller's flag word (the KR_FLAG_* values of
// the C ABI), and the camera kinds of an image (KR_CAMERA_*).
enum : int32_t { FLAG_RADICAND = 1, FLAG_NOT_NULL = 2, FLAG_SKY = 4, FLAG_DISK = 8, FLAG_EDGE = 16, FLAG_REPAIR = 32 };
enum : int64_t { CAMERA_BARDEEN = 0, CAMERA_ZAMO = 1 };

// Termination codes (src/kerrray/photons/classification.py:60-69).
enum : int { RUNNING = 0, ESCAPED = 1, CAPTURED = 2, MAX_AFFINE = 3, FAILURE = 4, OUT_OF_DOMAIN = 5, DISK_HIT = 6 };

// Indices of the reference packing (numba_kernel.py:89-107).
enum { G_MASS, G_A, G_A2, G_NEG_2MA, G_POS_2MA, G_TWO_M, G_TWO_A2 };
enum {
    C_ZERO, C_ONE, C_TWO, C_NEG_HALF, C_EIGHT, C_SAFETY, C_FMIN, C_FMAX, C_EXP, C_RTOL,
    C_ATOL, C_HMIN, C_RCAP, C_RESC, C_THLO, C_THHI, C_AXIS, C_H0, C_RIN, C_ROUT, C_HALFPI
};
enum { P_LAMMAX, P_HALFPI, P_HMAX };
enum { I_MAX_STEPS, I_GUARD, I_ADAPTIVE, I_DISK };

// Constants the reference's host always packs with the same values (the
// module constants behind numba_backend.py:100-115: step controller and
// classification tolerances).  The C ABI rejects any other value, so the
// kernels use them as literals rather than holding them in scalar registers.
namespace fixed {
constexpr double safety = 0.9, factor_min = 0.2, factor_max = 5.0, exponent = -1.0 / 5.0;
constexpr double theta_lo = -1e-9, theta_hi = 3.141592653589793 + 1e-9, axis_sin = 1e-12;
constexpr double half_pi = 0.5 * 3.141592653589793;
}  // namespace fixed

// Dormand-Prince 5(4) tableau (Dormand & Prince 1980; HNW 1993 Table 5.2),
// the values of src/kerrray/geodesics/tableaus.py:63-86 evaluated in double.
// The host entry points reject any other tableau, so these literals are what
// the reference passes in (numba_backend.py:118-128).
namespace dp {
constexpr double a10 = 1.0 / 5.0;
constexpr double a20 = 3.0 / 40.0, a21 = 9.0 / 40.0;
constexpr double a30 = 44.0 / 45.0, a31 = -56.0 / 15.0, a32 = 32.0 / 9.0;
constexpr double a40 = 19372.0 / 6561.0, a41 = -25360.0 / 2187.0, a42 = 64448.0 / 6561.0, a43 = -212.0 / 729.0;
constexpr double a50 = 9017.0 / 3168.0, a51 = -355.0 / 33.0, a52 = 46732.0 / 5247.0, a53 = 49.0 / 176.0,
                 a54 = -5103.0 / 18656.0;
constexpr double a60 = 35.0 / 384.0, a62 = 500.0 / 1113.0, a63 = 125.0 / 192.0, a64 = -2187.0 / 6784.0,
                 a65 = 11.0 / 84.0;
constexpr double e0 = 5179.0 / 57600.0 - 35.0 / 384.0, e2 = 7571.0 / 16695.0 - 500.0 / 1113.0,
                 e3 = 393.0 / 640.0 - 125.0 / 192.0, e4 = -92097.0 / 339200.0 - -2187.0 / 6784.0,
                 e5 = 187.0 / 2100.0 - 11.0 / 84.0, e6 = 1.0 / 40.0 - 0.0;
}  // namespace dp
namespace rk4 {
constexpr double c1 = 0.5, c2 = 0.5, c3 = 1.0;
constexpr double b0 = 1.0 / 6.0, b1 = 1.0 / 3.0, b2 = 1.0 / 3.0, b3 = 1.0 / 6.0;
}  // namespace rk4

// Everything one launch needs, passed by value in the kernel arguments
// (lands in the scalar cache; every lane reads the same words).
struct Params {
    double geom[7];
    double consts[21];  // float64 values; cast to the working type where used
    double params[3];
    int64_t ipar[4];
};

// The six integrated components of y = [t, r, th, ph, p_t, p_r, p_th, p_ph]
// (state.py:35) are stored as v[0..5] = t, r, th, ph, p_r, p_th.
constexpr int NA = 6;
constexpr int full_index(int i) { return i < 4 ? i : i + 1; }

template <typename T>
struct Photon {
    T v[NA];   // t, r, th, ph, p_r, p_th
    T p_t;     // constant of motion (-E)
    T p_ph;    // constant of motion (L_z)
};

template <typename T> KR_HD T pymax(T a, T b) { return b > a ? b : a; }
template <typename T> KR_HD T pymin(T a, T b) { return b < a ? b : a; }

KR_HD double wsqrt(double x) { return sqrt(x); }
KR_HD float wsqrt(float x) { return sqrtf(x); }
KR_HD double wpow(double x, double y) { return pow(x, y); }
KR_HD float wpow(float x, float y) { return powf(x, y); }
KR_HD double wabs(double x) { return fabs(x); }
KR_HD float wabs(float x) { return fabsf(x); }
KR_HD bool wfinite(double x) { return isfinite(x); }
KR_HD bool wfinite(float x) { return isfinite(x); }

// a / b rounded to nearest, from y = 1 / b rounded to nearest, by
// Markstein's theorem (Markstein 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, section 4.7): if y = RN(1/b) and q is within
// one ulp of a/b, then r = a - b q is exact under a fused multiply-add and
// RN(q + r y) = RN(a/b).  q0 = RN(a y) can be up to 1.5 ulp off; one
// correction brings it within 1/2 ulp plus a term of order 2^-52 ulp, and the
// second correction, under the theorem, gives exactly the IEEE quotient.  The
// sign is taken from q0 = a y, which always carries the quotient's sign (this
// also gives a zero numerator the signed zero the division gives).  Valid away
// from overflow and underflow, which the metric's operands here never reach; a
// zero or non-finite b gives a non-finite result, as the division does.
KR_HD double div_by(double a, double b, double y) {
    const double q0 = a * y;
    const double q1 = fma(fma(-b, q0, a), y, q0);
    const double q2 = fma(fma(-b, q1, a), y, q1);
    return copysign(q2, q0);
}

// sin(x) and cos(x).  On the device one sincos call: the device library
// derives both from one argument reduction and one polynomial pair, so the
// results are bit for bit those of separate sin and cos calls, at about half
// the instructions (separate calls are not merged by the compiler).  The host
// build keeps the C library's separate calls.
KR_HD void sin_cos(double x, double& s, double& c) {
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    s = sin(x);
    c = cos(x);
#else
    sincos(x, &s, &c);
#endif
}

// The metric's trigonometry: s = sin(th), c = cos(th) and s2 = sin(2 th),
// one lane per photon.
struct TrigSolo {
    KR_HD void operator()(double th, double& s, double& c, double& s2) const {
        sin_cos(th, s, c);
        s2 = sin(2.0 * th);
    }
};

#ifndef KR_HOST_BUILD
// x of the other lane of this lane's pair (lanes 2k and 2k + 1): a DPP
// quad_perm [1, 0, 3, 2] move of each 32-bit half.  Both lanes must be
// active.
__device__ __forceinline__ double pair_swap(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
#else
    return x;
#endif
}

// The same three values with both lanes of a pair carrying one photon (the
// same state in both, so the same control flow): the even lane evaluates
// sincos(th), the odd one sincos(2 th) (2 th is exact), and each takes what
// it lacks from the other.  The sine of a sincos call is bit for bit the
// sin call's, so every value is the one TrigSolo gives, for one library call
// per lane instead of two: the calls are long, and a wavefront whose photons
// set the end of a launch runs alone on its SIMD, where each instruction it
// issues is time on the critical path.
struct TrigPair {
    bool odd;
    __device__ void operator()(double th, double& s, double& c, double& s2) const {
        double sa, ca;
        sincos(odd ? 2.0 * th : th, &sa, &ca);
        const double so = pair_swap(sa), co = pair_swap(ca);
        s = odd ? so : sa;
        c = odd ? co : ca;
        s2 = odd ? sa : so;
    }
};
#endif

// numba_kernel.py:111-154 (= geometry/metric.py:166-195,265-313). Also
// returns sin(th), cos(th) for reuse by the classifier and diagnostics.
template <class Trig = TrigSolo>
KR_HD void inverse_metric_terms(const Params& P, double r, double th, double (&out)[15],
                                                     double& s_out, double& c_out, const Trig& trig = Trig{}) {
    const double m = P.geom[G_MASS];
    const double a2 = P.geom[G_A2];
    double s, c, sin2th;
    trig(th, s, c, sin2th);
    const double s2 = s * s;
    const double r2 = r * r;
    const double sig = r2 + a2 * (c * c);
    const double dlt = r2 - P.geom[G_TWO_M] * r + a2;
    const double r2a2 = r2 + a2;
    const double big_a = r2a2 * r2a2 - a2 * dlt * s2;
    const double sd = sig * dlt;
    // The seventeen quotients of numba_kernel.py:122-154 share four
    // denominators; each of those is inverted once (correctly rounded) and
    // div_by gives every quotient bit for bit.
    const double i_sd = 1.0 / sd;
    const double i_sig = 1.0 / sig;
    out[0] = div_by(-big_a, sd, i_sd);
    out[1] = div_by(P.geom[G_NEG_2MA] * r, sd, i_sd);
    out[2] = div_by(dlt, sig, i_sig);
    out[3] = i_sig;
    out[4] = (dlt - a2 * s2) / (sd * s2);
    const double sig_r = 2.0 * r;
    const double sig_th = (-a2) * sin2th;
    const double dlt_r = 2.0 * (r - m);
    const double a_r = 4.0 * r * r2a2 - P.geom[G_TWO_A2] * (r - m) * s2;
    const double a_th = (-a2) * dlt * sin2th;
    const double sd_r = sig_r * dlt + sig * dlt_r;
    const double sd_th = sig_th * dlt;
    const double sd2 = sd * sd;
    const double sig2 = sig * sig;
    const double i_sd2 = 1.0 / sd2;
    const double i_sig2 = 1.0 / sig2;
    out[5] = div_by(-(a_r * sd - big_a * sd_r), sd2, i_sd2);
    out[6] = div_by(P.geom[G_NEG_2MA] * (sd - r * sd_r), sd2, i_sd2);
    out[7] = div_by(dlt_r * sig - dlt * sig_r, sig2, i_sig2);
    out[8] = div_by(-sig_r, sig2, i_sig2);
    out[9] = -sig_r / (sig2 * s2) + div_by(a2 * sd_r, sd2, i_sd2);
    out[10] = div_by(-(a_th * sd - big_a * sd_th), sd2, i_sd2);
    out[11] = div_by(P.geom[G_POS_2MA] * r * sd_th, sd2, i_sd2);
    out[12] = div_by(-dlt * sig_th, sig2, i_sig2);
    out[13] = div_by(-sig_th, sig2, i_sig2);
    out[14] = -(sig_th * s2 + sig * sin2th) / (sig2 * (s2 * s2)) + div_by(a2 * sd_th, sd2, i_sd2);
    s_out = s;
    c_out = c;
}

// numba_kernel.py:158-197: f = dy/dlambda for the six integrated components
// (f[p_t] = f[p_phi] = 0 identically).
template <typename T, class Trig = TrigSolo>
KR_HD void rhs(const Params& P, const T (&v)[NA], T p_t, T p_ph, T (&f)[NA], double& s,
                                    double& c, const Trig& trig = Trig{}) {
    double g64[15];
    inverse_metric_terms(P, static_cast<double>(v[1]), static_cast<double>(v[2]), g64, s, c, trig);
    T g[15];
#pragma unroll
    for (int i = 0; i < 15; ++i) g[i] = static_cast<T>(g64[i]);
    const T zero = T(0);
    const T two = T(2);
    const T p_r = v[4], p_th = v[5];
    T gphph = g[4], r_gphph = g[9], t_gphph = g[14];
    if (p_ph == zero) {  // polar-axis guard (numba_kernel.py:177-180)
        gphph = zero;
        r_gphph = zero;
        t_gphph = zero;
    }
    f[0] = g[0] * p_t + g[1] * p_ph;
    f[1] = g[2] * p_r;
    f[2] = g[3] * p_th;
    f[3] = g[1] * p_t + gphph * p_ph;
    const T pt2 = p_t * p_t, ptpph = p_t * p_ph, pr2 = p_r * p_r, pth2 = p_th * p_th, pph2 = p_ph * p_ph;
    f[4] = T(-0.5) * (g[5] * pt2 + two * g[6] * ptpph + g[7] * pr2 + g[8] * pth2 + r_gphph * pph2);
    f[5] = T(-0.5) * (g[10] * pt2 + two * g[11] * ptpph + g[12] * pr2 + g[13] * pth2 + t_gphph * pph2);
}

template <typename T>
KR_HD bool all_finite(const T (&a)[NA], const T (&b)[NA]) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NA; ++i) ok = ok && wfinite(a[i]) && wfinite(b[i]);
    return ok;
}

// Receives the stage states of a Dormand-Prince attempt (none by default).
struct NoStages {
    template <typename T>
    KR_HD void operator()(int, const T (&)[NA]) const {}
};

// numba_kernel.py:201-238.  k0 = f(y) on entry (FSAL); on exit y_new, err and
// k6 = f(y_new), plus sin/cos of theta at y_new.  `stages(s, ys)` sees the
// state at which stage s's slope is evaluated (s = 0..5; stage 0's is y).
template <typename T, class Stages = NoStages, class Trig = TrigSolo>
KR_HD void dp5_step(const Params& P, const T (&y)[NA], T p_t, T p_ph, T h, const T (&k0)[NA],
                                         T (&y_new)[NA], T (&err)[NA], T (&k6)[NA], double& s_new, double& c_new,
                                         const Stages& stages = Stages{}, const Trig& trig = Trig{}) {
    using namespace dp;
    T k1[NA], k2[NA], k3[NA], k4[NA], k5[NA], ys[NA];
    double s, c;
    const T pt = p_t + T(0), pph = p_ph + T(0);
    stages(0, y);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(a10) * k0[i]);
    stages(1, ys);
    rhs(P, ys, pt, pph, k1, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(a20) * k0[i] + T(a21) * k1[i]);
    stages(2, ys);
    rhs(P, ys, pt, pph, k2, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(a30) * k0[i] + T(a31) * k1[i] + T(a32) * k2[i]);
    stages(3, ys);
    rhs(P, ys, pt, pph, k3, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y[i] + h * (T(a40) * k0[i] + T(a41) * k1[i] + T(a42) * k2[i] + T(a43) * k3[i]);
    stages(4, ys);
    rhs(P, ys, pt, pph, k4, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y[i] + h * (T(a50) * k0[i] + T(a51) * k1[i] + T(a52) * k2[i] + T(a53) * k3[i] + T(a54) * k4[i]);
    stages(5, ys);
    rhs(P, ys, pt, pph, k5, s, c, trig);
    T esum[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        y_new[i] = y[i] + h * (T(a60) * k0[i] + T(a62) * k2[i] + T(a63) * k3[i] + T(a64) * k4[i] + T(a65) * k5[i]);
        esum[i] = T(e0) * k0[i] + T(e2) * k2[i] + T(e3) * k3[i] + T(e4) * k4[i] + T(e5) * k5[i];
    }
    rhs(P, y_new, pt, pph, k6, s_new, c_new, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) err[i] = h * (esum[i] + T(e6) * k6[i]);
}

// Where dp5_step_stored keeps the stage slopes k1..k4 of one lane: slope s,
// component c at base[(s * NA + c) * stride].  In the persistent kernels
// base points into LDS at the lane's own column (stride = lanes per block),
// which takes 24 values per lane out of the register file for the whole
// attempt.
template <typename T>
struct SlopeStore {
    T* base;
    int stride;
    KR_HD T& at(int s, int c) const { return base[(s * NA + c) * stride]; }
};

// dp5_step with k1..k4 held in a SlopeStore.  Every value and every
// operation is the one dp5_step performs, in the same order, so the results
// are identical bit for bit.  With PARK the partial error sums wait in the
// store (k1's slot, no longer read) while the last right-hand side is
// evaluated: for a store in LDS that takes six more values per lane out of
// the register file across the deepest point of the attempt.
template <typename T, bool PARK = false, class Trig = TrigSolo>
KR_HD void dp5_step_stored(const Params& P, const T (&y)[NA], T p_t, T p_ph, T h, const T (&k0)[NA],
                           const SlopeStore<T>& K, T (&y_new)[NA], T (&err)[NA], T (&k6)[NA], double& s_new,
                           double& c_new, const Trig& trig = Trig{}) {
    using namespace dp;
    T k[NA], ys[NA];
    double s, c;
    const T pt = p_t + T(0), pph = p_ph + T(0);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(a10) * k0[i]);
    rhs(P, ys, pt, pph, k, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) K.at(0, i) = k[i];
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(a20) * k0[i] + T(a21) * K.at(0, i));
    rhs(P, ys, pt, pph, k, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) K.at(1, i) = k[i];
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(a30) * k0[i] + T(a31) * K.at(0, i) + T(a32) * K.at(1, i));
    rhs(P, ys, pt, pph, k, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) K.at(2, i) = k[i];
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y[i] + h * (T(a40) * k0[i] + T(a41) * K.at(0, i) + T(a42) * K.at(1, i) + T(a43) * K.at(2, i));
    rhs(P, ys, pt, pph, k, s, c, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) K.at(3, i) = k[i];
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y[i] + h * (T(a50) * k0[i] + T(a51) * K.at(0, i) + T(a52) * K.at(1, i) + T(a53) * K.at(2, i) +
                            T(a54) * K.at(3, i));
    rhs(P, ys, pt, pph, k, s, c, trig);  // k5
    T esum[PARK ? 1 : NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        y_new[i] = y[i] + h * (T(a60) * k0[i] + T(a62) * K.at(1, i) + T(a63) * K.at(2, i) + T(a64) * K.at(3, i) +
                               T(a65) * k[i]);
        T& e = PARK ? K.at(0, i) : esum[PARK ? 0 : i];
        e = T(e0) * k0[i] + T(e2) * K.at(1, i) + T(e3) * K.at(2, i) + T(e4) * K.at(3, i) + T(e5) * k[i];
    }
    rhs(P, y_new, pt, pph, k6, s_new, c_new, trig);
#pragma unroll
    for (int i = 0; i < NA; ++i) err[i] = h * ((PARK ? K.at(0, i) : esum[PARK ? 0 : i]) + T(e6) * k6[i]);
}

// numba_kernel.py:242-257.
template <typename T>
KR_HD void rk4_step(const Params& P, const T (&y)[NA], T p_t, T p_ph, T h, const T (&k0)[NA],
                                         T (&y_new)[NA], T (&k6)[NA], double& s_new, double& c_new) {
    using namespace rk4;
    T k1[NA], k2[NA], k3[NA], ys[NA];
    double s, c;
    const T pt = p_t + T(0), pph = p_ph + T(0);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(c1) * k0[i]);
    rhs(P, ys, pt, pph, k1, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(c2) * k1[i]);
    rhs(P, ys, pt, pph, k2, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y[i] + h * (T(c3) * k2[i]);
    rhs(P, ys, pt, pph, k3, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i) y_new[i] = y[i] + h * (T(b0) * k0[i] + T(b1) * k1[i] + T(b2) * k2[i] + T(b3) * k3[i]);
    rhs(P, y_new, pt, pph, k6, s_new, c_new);
}

// numba_kernel.py:261-268 (scale uses max(|y|, |y_new|); p_t/p_phi terms are +0).
template <typename T>
KR_HD T error_norm(T rtol, T atol, const T (&err)[NA], const T (&y)[NA], const T (&y_new)[NA]) {
    T q[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const T scale = atol + rtol * pymax(wabs(y[i]), wabs(y_new[i]));
        const T v = err[i] / scale;
        q[i] = v * v;
    }
    // q index: 0 t, 1 r, 2 th, 3 ph, 4 p_r (full 5), 5 p_th (full 6)
    const T total = ((q[0] + q[1]) + (q[2] + q[3])) + (q[4] + q[5]);
    return wsqrt(total / T(8));
}

// 1/x and sqrt(x) for the bookkeeping of FP32-bulk attempts (error norm,
// drift diagnostics), which need not round as the reference does: the
// hardware reciprocal estimate (about 2^-23 relative) with one Newton step
// (about 2^-46), and the hardware square root (about 2^-23).  Zero and
// non-finite operands give non-finite results, as the divisions do.  The
// host build divides and calls sqrt.
KR_HD double rcp_bulk(double x) {
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    return 1.0 / x;
#else
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
#endif
}

KR_HD double sqrt_bulk(double x) {
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    return sqrt(x);
#else
    return __builtin_amdgcn_sqrt(x);
#endif
}

// error_norm for an FP32-bulk attempt: the same quantity, with the six
// quotients and the root from rcp_bulk and sqrt_bulk.  It moves en by about
// 1e-7 relative, which changes an acceptance only for en within that of 1;
// the bulk's rows are not the reference's bits in any case, and its mask is
// checked and repaired in FP64 after the trace.
KR_HD double error_norm_bulk(double rtol, double atol, const double (&err)[NA], const double (&y)[NA],
                             const double (&y_new)[NA]) {
    double q[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const double scale = atol + rtol * pymax(fabs(y[i]), fabs(y_new[i]));
        const double v = err[i] * rcp_bulk(scale);
        q[i] = v * v;
    }
    const double total = ((q[0] + q[1]) + (q[2] + q[3])) + (q[4] + q[5]);
    return sqrt_bulk(total * 0.125);
}

// x^(-1/5) for 1e-5 <= x <= 1e5 in FP64 to within a few ulp (not the library
// pow's rounding): an FP32 estimate (hardware log2 and exp2 on the device)
// refined by two Newton corrections y <- y + (y / 5)(1 - x y^5), each of
// which squares the relative error (1e-7, 1e-14, then rounding).
KR_HD double pow_m_fifth(double x) {
    const float xf = static_cast<float>(x);
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    double y = static_cast<double>(exp2f(-0.2f * log2f(xf)));
#else
    double y = static_cast<double>(__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf(xf)));
#endif
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double y2 = y * y;
        const double t = x * (y2 * y2 * y);
        y = fma(y * 0.2, 1.0 - t, y);
    }
    return y;
}

// log2 in FP32: the hardware estimate on the device (v_log_f32), the C
// library's on the host.
KR_HD float log2_fast(float x) {
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    return log2f(x);
#else
    return __builtin_amdgcn_logf(x);
#endif
}

// x^(-1/5) as y (1 + c): y = pow_m_fifth(x) (within 2 ulp) and
// c = r/5 + 3 r^2/25 with r = 1 - x y^5 (x y^5 = 1 - r, so x^(-1/5) =
// y (1 - r)^(-1/5)).  r is formed to about 2^-105 with the rounding errors of
// the products (fused multiply-adds): 1 - x y^5 is exact in double (x y^5 is
// within 1e-15 of 1), the error terms are added after.
KR_HD void m_fifth_residual(double x, double& y, double& c) {
    y = pow_m_fifth(x);
    const double y2 = y * y, y2e = fma(y, y, -y2);
    const double y4 = y2 * y2, y4e = fma(y2, y2, -y4) + 2.0 * y2 * y2e;
    const double y5 = y4 * y, y5e = fma(y4, y, -y5) + y4e * y;
    const double p = x * y5, pe = fma(x, y5, -p) + x * y5e;
    const double r = (1.0 - p) - pe;
    c = r * 0.2 + 0.12 * (r * r);
}

// pow(x, -0.2) correctly rounded, for 1e-5 <= x <= 1e5 (Ziv's method: an
// approximation with a proven error bound, rounded when the bound cannot
// straddle a rounding boundary, else recomputed more accurately).  The double
// nearest -1/5 is -(2^54 + 1) / (5 2^54), so x^-0.2 = x^(-1/5) x^(-2^-54 / 5)
// = y (1 + c - ln(x) 2^-54 / 5) to far below an ulp.  The fast path takes
// ln x from the FP32 log2 (absolute error < 4e-6): total error under
// 2^-75 y, tested against 2^-72 y.  Otherwise ln x comes from the double log:
// error under 2^-102 y, tested against 2^-99 y.  Both tests fail together
// with probability below 1e-16 per call; the rounded sum is then returned.
KR_HD double pow_m02_cr(double x) {
    double y, c;
    m_fifth_residual(x, y, c);
    constexpr double k = 0x1p-54 / 5.0;  // 2^-54 / 5, the exponent's excess over -1/5
    constexpr double ln2 = 0.6931471805599453;
    const double ln_fast = static_cast<double>(log2_fast(static_cast<float>(x))) * ln2;
    double lo = y * (c - k * ln_fast);
    double e = 0x1p-72 * y;
    if (y + (lo + e) == y + (lo - e)) return y + lo;
    lo = y * (c - k * log(x));
    e = 0x1p-99 * y;
    return y + lo;
}

// powf(x, -0.2f) correctly rounded for 1e-5 <= x <= 1e5, in double.  -0.2f is
// -(2^26 + 1) / (5 2^26), so x^-0.2f = x^(-1/5) exp(-u) with u = ln(x) 2^-26 / 5
// (|u| < 4e-8): z = y (1 + c - u + u^2 / 2) to about 1e-14 relative with the
// FP32 log2, to about 3e-16 with the double log.  z is rounded to float when
// the error bound cannot straddle a float rounding boundary.
KR_HD float powf_m02_cr(float xf) {
    const double x = static_cast<double>(xf);
    double y, c;
    m_fifth_residual(x, y, c);
    constexpr double k = 0x1p-26 / 5.0;
    constexpr double ln2 = 0.6931471805599453;
    double u = k * (static_cast<double>(log2_fast(xf)) * ln2);
    double z = y + y * (c - u + 0.5 * (u * u));
    double e = 1e-13 * z;
    if (static_cast<float>(z + e) == static_cast<float>(z - e)) return static_cast<float>(z);
    u = k * log(x);
    z = y + y * (c - u + 0.5 * (u * u));
    return static_cast<float>(z);
}

// en ** expo of the step controller (numba_kernel.py:277), expo being the
// float64 -0.2 or its float32 cast.  The host build calls the C library's
// pow / powf, as the reference's kernel does.  The device calls the
// correctly rounded routines above: they equal the C library's pow(x, -0.2)
// on 99.92 % of 1e7 arguments and powf(x, -0.2f) on 99.94 % of every float
// in [1e-5, 1e5] (where the library misrounds, they give the correctly
// rounded value), against about 83 % for the device library's pow.  Outside
// [1e-5, 1e5] only the clamps of step_factor see the value (0.9 en^-0.2 is
// above 5 below 1e-5 and below 0.2 above 1e5), so any value beyond them
// gives the same factor.
template <typename T>
KR_HD T controller_power(T en, T expo) {
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    return wpow(en, expo);
#else
    (void)expo;
    if (en < T(1e-5)) return T(1e30);
    if (en > T(1e5)) return T(0);
    if constexpr (sizeof(T) == sizeof(double))
        return pow_m02_cr(en);
    else
        return powf_m02_cr(en);
#endif
}

// numba_kernel.py:272-281.
template <typename T>
KR_HD T step_factor(T safety, T fmin, T fmax, T expo, T en, bool rejected_before) {
    T factor;
    if (en == T(0)) {
        factor = fmax;
    } else {
        const T raw = safety * controller_power(en, expo);
        factor = pymin(fmax, pymax(fmin, raw));
    }
    if (rejected_before) factor = pymin(factor, T(1));
    return factor;
}

// The step factor of an FP32-bulk attempt of the precision form: the
// controller of numba_kernel.py:272-281 in FP64 (same constants and clamps)
// with en^(-1/5) from pow_m_fifth instead of the library pow, which is
// a fifth of such an attempt's instructions.  Outside [1e-5, 1e5] the clamps
// decide whatever the power: 0.9 en^(-1/5) exceeds 5 below and is under 0.2
// above.
template <typename T>
KR_HD T bulk_step_factor(T en, bool rejected_before) {
    T factor;
    if (en < T(1e-5))
        factor = static_cast<T>(fixed::factor_max);
    else if (en > T(1e5))
        factor = static_cast<T>(fixed::factor_min);
    else
        factor = pymin(static_cast<T>(fixed::factor_max),
                       pymax(static_cast<T>(fixed::factor_min),
                             static_cast<T>(fixed::safety) * static_cast<T>(pow_m_fifth(static_cast<double>(en)))));
    if (rejected_before) factor = pymin(factor, T(1));
    return factor;
}

// numba_kernel.py:293-311.  sin_th is sin(y[th]) in the working type.
template <typename T>
KR_HD int classify(const Params& P, const T (&v)[NA], T p_ph, const T (&f)[NA], double lam,
                                        int64_t n_steps, T h, bool check_h, T sin_th) {
    if (!all_finite(v, f)) return FAILURE;
    if (check_h && h < static_cast<T>(P.consts[C_HMIN])) return FAILURE;
    const T r = v[1];
    const T th = v[2];
    if (r < T(0) || th < static_cast<T>(fixed::theta_lo) || th > static_cast<T>(fixed::theta_hi))
        return OUT_OF_DOMAIN;
    if (p_ph != T(0) && wabs(sin_th) < static_cast<T>(fixed::axis_sin)) return OUT_OF_DOMAIN;
    if (r <= static_cast<T>(P.consts[C_RCAP])) return CAPTURED;
    if (r >= static_cast<T>(P.consts[C_RESC]) && f[1] > T(0)) return ESCAPED;
    if (lam >= P.params[P_LAMMAX] || n_steps >= P.ipar[I_MAX_STEPS]) return MAX_AFFINE;
    return RUNNING;
}

template <typename T> KR_HD T working_sin(double s_double, T th);
template <> KR_HD double working_sin<double>(double s_double, double) { return s_double; }
template <> KR_HD float working_sin<float>(double, float th) { return sinf(th); }

// numba_kernel.py:315-321.
template <typename T>
KR_HD double null_error(const T (&v)[NA], T p_t, T p_ph, const T (&f)[NA]) {
    double total = static_cast<double>(p_t) * static_cast<double>(f[0]);
    total = total + static_cast<double>(v[4]) * static_cast<double>(f[1]);
    total = total + static_cast<double>(v[5]) * static_cast<double>(f[2]);
    total = total + static_cast<double>(p_ph) * static_cast<double>(f[3]);
    return fabs(0.5 * total);
}

// numba_kernel.py:325-336 with s = sin(th), c = cos(th) of double(th).
// BULK: the quotient from rcp_bulk (FP32-bulk attempts).
template <bool BULK = false, typename T>
KR_HD double carter(double a2, const T (&v)[NA], T p_t, T p_ph, double s, double c) {
    const double p_th = static_cast<double>(v[5]);
    const double lz = static_cast<double>(p_ph);
    const double e = -static_cast<double>(p_t);
    double lz2_over_s2 = 0.0;
    if (lz != 0.0) lz2_over_s2 = BULK ? (lz * lz) * rcp_bulk(s * s) : (lz * lz) / (s * s);
    return p_th * p_th + (c * c) * (lz2_over_s2 - a2 * e * e);
}


// The per-photon integrator of numba_kernel.py:340-463 as a small state
// machine: ray_init (:344-375), one loop iteration per ray_attempt
// (:376-455), ray_result (:456-463).  integrate_ray runs them back to back;
// the persistent kernel interleaves photons between attempts.  Either way
// a photon sees exactly the same arithmetic.
template <typename T>
struct RayState {
    T y[NA];
    T p_t, p_ph;
    T f[NA];
    double s_cur, c_cur;  // sin, cos of the current theta (from the last right-hand side)
    double e0, lz0, fe, fl, fq, fe2, q0;
    double m_null, m_e, m_lz, m_q, lam;
    T h;
    int64_t n_steps, n_rej, attempts;
    bool rejected_before;
    bool ev_hit;
    int state;
    T ev[8];
};

// numba_kernel.py:344-375.  (Y0 arrives in float64 and is cast to T,
// numba_backend.py:167; the drift baselines e0, lz0, q0 and the floors come
// from the float64 Y0, numba_backend.py:174-178.)
template <typename T, bool ADAPTIVE, class Trig = TrigSolo>
KR_HD void ray_init(const Params& P, const double (&y64)[8], RayState<T>& R, const Trig& trig = Trig{}) {
    const double a2 = P.geom[G_A2];
#pragma unroll
    for (int i = 0; i < NA; ++i) R.y[i] = static_cast<T>(y64[full_index(i)]);
    R.p_t = static_cast<T>(y64[4]);
    R.p_ph = static_cast<T>(y64[7]);
    R.e0 = -y64[4];
    R.lz0 = y64[7];
    double fe = fabs(R.e0);
    fe = fe > 0.0 ? fe : 1.0;
    const double mass = P.geom[G_MASS];
    R.fe = fe;
    R.fl = fe * mass;
    R.fq = fe * fe * mass * mass;
    R.fe2 = fe * fe;
    rhs(P, R.y, R.p_t, R.p_ph, R.f, R.s_cur, R.c_cur, trig);
    {
        // carter_constant(st, Y0) on the float64 Y0 (same expression as numba_kernel.py:325-336).
        const double th0 = y64[2];
        double s0 = R.s_cur, c0 = R.c_cur;
        if (static_cast<double>(static_cast<T>(th0)) != th0) {
            s0 = sin(th0);
            c0 = cos(th0);
        }
        const double lz = y64[7], e = -y64[4];
        double lz2_over_s2 = 0.0;
        if (lz != 0.0) lz2_over_s2 = (lz * lz) / (s0 * s0);
        R.q0 = y64[6] * y64[6] + (c0 * c0) * (lz2_over_s2 - a2 * e * e);
    }
    R.m_null = null_error(R.y, R.p_t, R.p_ph, R.f) / R.fe2;
    R.m_e = 0.0;
    R.m_lz = 0.0;
    R.m_q = 0.0;
    R.lam = 0.0;
    R.h = static_cast<T>(P.consts[C_H0]);
    if (ADAPTIVE && static_cast<double>(R.h) > P.params[P_HMAX]) R.h = static_cast<T>(P.params[P_HMAX]);
    R.n_steps = 0;
    R.n_rej = 0;
    R.attempts = 0;
    R.rejected_before = false;
    R.ev_hit = false;
    R.state = classify(P, R.y, R.p_ph, R.f, R.lam, R.n_steps, R.h, false, working_sin<T>(R.s_cur, R.y[2]));
}

// The exact step of the reference: Dormand-Prince 5(4) or RK4 in the
// working type.
template <typename T, bool ADAPTIVE, class Trig = TrigSolo>
struct ExactStepper {
    Trig trig = Trig{};
    KR_HD bool operator()(const Params& P, const RayState<T>& R, T h_eff, T (&y_new)[NA], T (&err)[NA],
                          T (&k6)[NA], double& s_new, double& c_new) const {
        if (ADAPTIVE)
            dp5_step(P, R.y, R.p_t, R.p_ph, h_eff, R.f, y_new, err, k6, s_new, c_new, NoStages{}, trig);
        else
            rk4_step(P, R.y, R.p_t, R.p_ph, h_eff, R.f, y_new, k6, s_new, c_new);
        return true;
    }
};

// The exact attempt with the stage slopes in a SlopeStore (RK4 attempts
// are short and keep theirs in registers); PARK as in dp5_step_stored.
template <typename T, bool ADAPTIVE, bool PARK = false, class Trig = TrigSolo>
struct StoredStepper {
    SlopeStore<T> K;
    Trig trig = Trig{};
    KR_HD bool operator()(const Params& P, const RayState<T>& R, T h_eff, T (&y_new)[NA], T (&err)[NA],
                          T (&k6)[NA], double& s_new, double& c_new) const {
        if (ADAPTIVE)
            dp5_step_stored<T, PARK, Trig>(P, R.y, R.p_t, R.p_ph, h_eff, R.f, K, y_new, err, k6, s_new, c_new, trig);
        else
            rk4_step(P, R.y, R.p_t, R.p_ph, h_eff, R.f, y_new, k6, s_new, c_new);
        return true;
    }
};

// One iteration of the while loop of numba_kernel.py:376-455 (a `break`
// there is a return here with R.state set; `continue` is a return with
// R.state still RUNNING).  `step` computes y_new, the error estimate and
// f(y_new); everything else (error norm, acceptance, step control,
// bookkeeping, classification) is the reference's.
template <typename T, bool ADAPTIVE, bool DISK, class Stepper>
KR_HD void ray_attempt_with(const Params& P, RayState<T>& R, const Stepper& step) {
    const double lambda_max = P.params[P_LAMMAX];
    const double h_cap = P.params[P_HMAX];
    R.attempts += 1;
    if (R.attempts > P.ipar[I_GUARD]) {
        R.state = FAILURE;
        return;
    }
    const double remaining = lambda_max - R.lam;
    const bool last = static_cast<double>(R.h) >= remaining;
    T h_eff = R.h;
    if (last) h_eff = static_cast<T>(remaining);
    T y_new[NA], k6[NA], err[NA];
    double s_new, c_new;
    // The stepper reports whether the attempt was the exact one (the
    // reference's arithmetic throughout, controller included).
    const bool exact = step(P, R, h_eff, y_new, err, k6, s_new, c_new);
    bool finite = all_finite(y_new, k6);
    bool accept = finite;
    T h_new = R.h;
    if (ADAPTIVE && finite) {
        T en;
        if constexpr (sizeof(T) == sizeof(double)) {
            en = exact ? error_norm(P.consts[C_RTOL], P.consts[C_ATOL], err, R.y, y_new)
                       : error_norm_bulk(P.consts[C_RTOL], P.consts[C_ATOL], err, R.y, y_new);
        } else {
            en = error_norm(static_cast<T>(P.consts[C_RTOL]), static_cast<T>(P.consts[C_ATOL]), err, R.y, y_new);
        }
        if (!wfinite(en)) {
            finite = false;
            accept = false;
        } else {
            accept = en <= T(1);
            const T factor = exact ? step_factor(static_cast<T>(fixed::safety), static_cast<T>(fixed::factor_min),
                                                 static_cast<T>(fixed::factor_max),
                                                 static_cast<T>(fixed::exponent), en, R.rejected_before)
                                   : bulk_step_factor(en, R.rejected_before);
            if (!(accept && last)) {
                h_new = h_eff * factor;
                if (static_cast<double>(h_new) > h_cap) h_new = static_cast<T>(h_cap);
            }
        }
    }
    if (!finite) {
        R.state = FAILURE;
        return;
    }
    const bool rejected = !accept;
    if (rejected) R.n_rej += 1;
    R.rejected_before = rejected;
    R.h = h_new;
    if (rejected) {
        if (h_new < static_cast<T>(P.consts[C_HMIN])) R.state = FAILURE;
        return;
    }
    T y_prev[NA];
    const T p_t_prev = R.p_t, p_ph_prev = R.p_ph;
    if (DISK) {
#pragma unroll
        for (int i = 0; i < NA; ++i) y_prev[i] = R.y[i];
    }
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        R.y[i] = y_new[i];
        R.f[i] = k6[i];
    }
    R.p_t = R.p_t + T(0);
    R.p_ph = R.p_ph + T(0);
    R.s_cur = s_new;
    R.c_cur = c_new;
    const double lam_old = R.lam;
    if (last)
        R.lam = lambda_max;
    else
        R.lam = R.lam + static_cast<double>(h_eff);
    R.n_steps += 1;
    if (sizeof(T) != sizeof(double)) {
        // In float64 E and L_z are bit-constant and both drifts stay 0.
        const double e = -static_cast<double>(R.p_t);
        const double lz = static_cast<double>(R.p_ph);
        R.m_e = pymax(R.m_e, fabs(e - R.e0) / pymax(fabs(R.e0), R.fe));
        R.m_lz = pymax(R.m_lz, fabs(lz - R.lz0) / pymax(fabs(R.lz0), R.fl));
    }
    if (exact) {
        R.m_q = pymax(R.m_q, fabs(carter(P.geom[G_A2], R.y, R.p_t, R.p_ph, R.s_cur, R.c_cur) - R.q0) /
                                 pymax(fabs(R.q0), R.fq));
        R.m_null = pymax(R.m_null, null_error(R.y, R.p_t, R.p_ph, R.f) / R.fe2);
    } else {
        R.m_q = pymax(R.m_q, fabs(carter<true>(P.geom[G_A2], R.y, R.p_t, R.p_ph, R.s_cur, R.c_cur) - R.q0) *
                                 rcp_bulk(pymax(fabs(R.q0), R.fq)));
        R.m_null = pymax(R.m_null, null_error(R.y, R.p_t, R.p_ph, R.f) * rcp_bulk(R.fe2));
    }
    int code = classify(P, R.y, R.p_ph, R.f, R.lam, R.n_steps, R.h, true, working_sin<T>(R.s_cur, R.y[2]));
    if (DISK) {
        // numba_kernel.py:439-454
        const double half_pi = fixed::half_pi;
        const double th_old = static_cast<double>(y_prev[2]);
        const double th_new = static_cast<double>(R.y[2]);
        if ((th_old - half_pi) * (th_new - half_pi) < 0.0) {
            const double sfrac = (half_pi - th_old) / (th_new - th_old);
            T yc[8];
            const T full_new[8] = {R.y[0], R.y[1], R.y[2], R.y[3], R.p_t, R.y[4], R.y[5], R.p_ph};
            const T full_old[8] = {y_prev[0], y_prev[1], y_prev[2], y_prev[3], p_t_prev, y_prev[4], y_prev[5], p_ph_prev};
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const T cast = static_cast<T>(sfrac * static_cast<double>(full_new[i] - full_old[i]));
                yc[i] = full_old[i] + cast;
            }
            yc[2] = static_cast<T>(fixed::half_pi);
            if (yc[1] >= static_cast<T>(P.consts[C_RIN]) && yc[1] <= static_cast<T>(P.consts[C_ROUT])) {
#pragma unroll
                for (int i = 0; i < 8; ++i) R.ev[i] = yc[i];
#pragma unroll
                for (int i = 0; i < NA; ++i) R.y[i] = yc[full_index(i)];
                R.p_t = yc[4];
                R.p_ph = yc[7];
                R.ev_hit = true;
                R.lam = lam_old + sfrac * static_cast<double>(h_eff);
                code = DISK_HIT;
            }
        }
    }
    R.state = code;
}

template <typename T, bool ADAPTIVE, bool DISK>
KR_HD void ray_attempt(const Params& P, RayState<T>& R) {
    ray_attempt_with<T, ADAPTIVE, DISK>(P, R, ExactStepper<T, ADAPTIVE>{});
}

template <typename T>
struct RayResult {
    T y[8];
    double lam;
    int64_t n_steps;
    int64_t n_rej;
    int state;
    double m_null, m_e, m_lz, m_q;
    bool ev_hit;
    T ev[8];
};

template <typename T, bool DISK>
KR_HD void ray_result(const RayState<T>& R, RayResult<T>& out) {
#pragma unroll
    for (int i = 0; i < NA; ++i) out.y[full_index(i)] = R.y[i];
    out.y[4] = R.p_t;
    out.y[7] = R.p_ph;
    out.lam = R.lam;
    out.n_steps = R.n_steps;
    out.n_rej = R.n_rej;
    out.state = R.state;
    out.m_null = R.m_null;
    out.m_e = R.m_e;
    out.m_lz = R.m_lz;
    out.m_q = R.m_q;
    out.ev_hit = R.ev_hit;
    if (DISK) {
#pragma unroll
        for (int i = 0; i < 8; ++i) out.ev[i] = R.ev_hit ? R.ev[i] : T(0);
    }
}

// Output rows of a batch (the BatchResult arrays, kr_outputs without the
// disk event) and the store of one finished photon into row `idx`.
template <typename T>
struct RowsOut {
    T* Y;
    double* lam;
    int64_t* n_steps;
    int64_t* n_rej;
    int64_t* state;
    double* m_null;
    double* m_e;
    double* m_lz;
    double* m_q;
};

template <typename T>
KR_HD void store_row(const RowsOut<T>& o, int64_t idx, const RayState<T>& R) {
    RayResult<T> r;
    ray_result<T, false>(R, r);
#pragma unroll
    for (int c = 0; c < 8; ++c) o.Y[idx * 8 + c] = r.y[c];
    o.lam[idx] = r.lam;
    o.n_steps[idx] = r.n_steps;
    o.n_rej[idx] = r.n_rej;
    o.state[idx] = r.state;
    o.m_null[idx] = r.m_null;
    o.m_e[idx] = r.m_e;
    o.m_lz[idx] = r.m_lz;
    o.m_q[idx] = r.m_q;
}

// The disk event of a finished photon into row idx of the event outputs:
// the crossing state where the photon hit the disk (numba_kernel.py:450-452),
// zeros elsewhere (the zeroed arrays of numba_backend.py:180-181).
template <typename T>
KR_HD void store_event(T* ev_Y, uint8_t* ev_hit, int64_t idx, const RayState<T>& R) {
#pragma unroll
    for (int c = 0; c < 8; ++c) ev_Y[idx * 8 + c] = R.ev_hit ? R.ev[c] : T(0);
    ev_hit[idx] = R.ev_hit ? 1 : 0;
}

// numba_kernel.py:340-463 for one photon whose float64 initial state is y64.
template <typename T, bool ADAPTIVE, bool DISK>
KR_HD void integrate_ray(const Params& P, const double (&y64)[8], RayResult<T>& out) {
    RayState<T> R;
    ray_init<T, ADAPTIVE>(P, y64, R);
    while (R.state == RUNNING) ray_attempt<T, ADAPTIVE, DISK>(P, R);
    ray_result<T, DISK>(R, out);
}

// Per-image scalars of the ZAMO camera (camera.py:225-381): the observer's
// tetrad rows e_(a)^mu, the covariant metric g_mu_nu and the contravariant
// g^mu_nu at (r_o, theta_o), computed on the host as geometry/metric.py:147-195
// and camera.py:256-262 compute them.
struct ZamoImage {
    double r_o, theta_o, phi_o;
    double r_o2;                              // r_o**2
    double e00, e03, e11, e22, e33;           // 1/lapse, omega/lapse, 1/sqrt(g_rr), 1/sqrt(g_thth), 1/sqrt(g_phph)
    double g_tt, g_tph, g_rr, g_thth, g_phph;
    double gi_tt, gi_tph, gi_rr, gi_thth, gi_phph;
    double sign;                              // -1: reversed for backward tracing, +1: the physical photon
};

// Per-image scalars of the Bardeen initial conditions, computed on the host
// with the reference's own Python float expressions (rayset.py:158-165,
// initial_conditions.py:50-112, horizons via delta()), and the camera whose
// pixels an image traces:
struct BardeenImage {
    double r_o, theta_o;     // observer radius and polar angle
    double sin_o;            // math.sin(theta_o)
    double a;                // spin * mass
    double a2c2;             // a**2 * cos_o**2
    double cot2;             // (cos_o / sin_o)**2
    double r2a2;             // E * (r*r + a*a) with E = 1
    double dlt;              // r**2 - 2.0*m*r + a**2 at r_o
    double c2, s2;           // math.cos(theta)**2, math.sin(theta)**2 at theta_o
    double aaEE;             // a * a * E * E
    double fov;              // half-width of the image plane (M)
    int64_t n;               // resolution (n x n pixels)
    int64_t camera;          // KR_CAMERA_BARDEEN or KR_CAMERA_ZAMO
    ZamoImage Z;             // KR_CAMERA_ZAMO only
};

// _off_equatorial_state + photon_from_constants for one image-plane point.
// Python's x**2 on a float is glibc pow(x, 2.0), which returns x*x exactly.
// Returns false when a radicand is negative beyond round-off
// (initial_conditions.py:71-74).
KR_HD bool bardeen_state(const BardeenImage& B, double alpha, double beta, double (&y)[8]) {
    const double xi = -alpha * B.sin_o;
    const double eta = beta * beta - B.a2c2 + xi * xi * B.cot2;
    const double term1 = B.r2a2 - B.a * xi;
    const double d = xi - B.a * 1.0;
    const double term2 = B.dlt * (eta + d * d);
    const double big_r = term1 * term1 - term2;
    const double scale_r = term1 * term1 + fabs(term2);
    const double lz_term = (xi == 0.0) ? 0.0 : xi * xi / B.s2;
    const double big_theta = eta - B.c2 * (lz_term - B.aaEE);
    const double scale_theta = fabs(eta) + B.c2 * (lz_term + B.aaEE);
    const bool ok = !(big_r < -1e-12 * pymax(scale_r, 1.0)) && !(big_theta < -1e-12 * pymax(scale_theta, 1.0));
    const double sth = beta >= 0.0 ? -1.0 : 1.0;
    y[0] = 0.0;
    y[1] = B.r_o;
    y[2] = B.theta_o;
    y[3] = 0.0;
    y[4] = -1.0;
    y[5] = -1.0 * sqrt(pymax(big_r, 0.0)) / B.dlt;
    y[6] = sth * sqrt(pymax(big_theta, 0.0));
    y[7] = xi;
    return ok;
}

// Launch constants of a critical-impact search: photons from (r0, theta)
// with E = 1, Q = 0 (kerr_bisection.py:117-118 -> photon_from_constants).
struct ImpactLaunch {
    double m, a;        // mass, spin * mass
    double r0, theta;
    double c2, s2;      // math.cos(theta) ** 2, math.sin(theta) ** 2
};

// photon_from_constants(st, r0, theta, E = 1, L_z, Q = 0, sign_r = -1,
// sign_theta = +1) with carter_potentials (initial_conditions.py:50-112),
// operation for operation (Python's x ** 2 on a float is x * x exactly).
// Returns false when a radicand is negative beyond round-off.
KR_HD bool impact_state(const ImpactLaunch& L, double lz, double (&y)[8]) {
    const double e = 1.0, q = 0.0, r = L.r0, a = L.a;
    const double dlt = r * r - 2.0 * L.m * r + a * a;
    const double term1 = e * (r * r + a * a) - a * lz;
    const double d = lz - a * e;
    const double term2 = dlt * (q + d * d);
    const double big_r = term1 * term1 - term2;
    const double scale_r = term1 * term1 + fabs(term2);
    const double lz_term = lz == 0.0 ? 0.0 : lz * lz / L.s2;
    const double big_theta = q - L.c2 * (lz_term - a * a * e * e);
    const double scale_theta = fabs(q) + L.c2 * (lz_term + a * a * e * e);
    const bool ok = !(big_r < -1e-12 * pymax(scale_r, 1.0)) && !(big_theta < -1e-12 * pymax(scale_theta, 1.0));
    y[0] = 0.0;
    y[1] = r;
    y[2] = L.theta;
    y[3] = 0.0;
    y[4] = -e;
    y[5] = -1.0 * sqrt(pymax(big_r, 0.0)) / dlt;
    y[6] = 1.0 * sqrt(pymax(big_theta, 0.0));
    y[7] = lz;
    return ok;
}

// Camera.pixel_coordinates (camera.py:206-222): fov (2 j + 1 - n) / n.
KR_HD double pixel_centre(double fov, int64_t j, int64_t n) {
    const double nd = static_cast<double>(n);
    return (fov * ((2.0 * static_cast<double>(j) + 1.0) - nd)) / nd;
}

// The ZAMO photon of initial_states (camera.py:298-381) for the image point
// (alpha, beta), followed by _null_pt (camera.py:265-295).  Returns 0, or
// KR_FLAG_SKY when alpha^2 + beta^2 >= r_o^2 (no such direction on the sky)
// and KR_FLAG_NOT_NULL when the tetrad projection is not null to 1e-10 or no
// null p_t exists.
//
// The tetrad and metric products are the reference's 4 x 4 matrix products,
// p^mu = sum_a p^(a) e_(a)^mu and p_mu = sum_nu p^nu g_nu_mu.  NumPy hands
// them to BLAS, whose double kernels accumulate the terms in index order with
// fused multiply-adds from zero.  Products with the zero entries of the
// tetrad and the metric leave the sum unchanged, so each component is one
// product, or one fused multiply-add onto one product; the expressions below
// are those, term for term.
KR_HD int32_t zamo_state(const ZamoImage& Z, double alpha, double beta, double (&y)[8]) {
    int32_t bad = 0;
    const double rho2 = (alpha * alpha + beta * beta) / Z.r_o2;
    if (!(rho2 < 1.0)) bad |= FLAG_SKY;
    const double l_r = sqrt(1.0 - rho2);
    const double l_th = beta / Z.r_o;
    const double l_ph = -alpha / Z.r_o;
    // p^mu; p^(t) = 1
    const double u_t = Z.e00;
    const double u_r = l_r * Z.e11;
    const double u_th = l_th * Z.e22;
    const double u_ph = fma(l_ph, Z.e33, Z.e03);
    // p_mu
    double p_t = fma(u_ph, Z.g_tph, u_t * Z.g_tt);
    double p_r = u_r * Z.g_rr;
    double p_th = u_th * Z.g_thth;
    double p_ph = fma(u_ph, Z.g_phph, u_t * Z.g_tph);
    const double residual = Z.gi_tt * (p_t * p_t) + 2.0 * Z.gi_tph * p_t * p_ph + Z.gi_rr * (p_r * p_r) +
                            Z.gi_thth * (p_th * p_th) + Z.gi_phph * (p_ph * p_ph);
    if (!(fabs(residual) < 1e-10)) bad |= FLAG_NOT_NULL;
    p_t *= Z.sign;
    p_r *= Z.sign;
    p_th *= Z.sign;
    p_ph *= Z.sign;
    // g^{mu nu} p_mu p_nu = 0 solved for p_t: the root nearest the projected one.
    const double qa = Z.gi_tt;
    const double qb = 2.0 * Z.gi_tph * p_ph;
    const double qc = Z.gi_rr * (p_r * p_r) + Z.gi_thth * (p_th * p_th) + Z.gi_phph * (p_ph * p_ph);
    const double disc = qb * qb - 4.0 * qa * qc;
    if (disc < -1e-10 * (qb * qb + fabs(4.0 * qa * qc))) bad |= FLAG_NOT_NULL;
    const double q = -0.5 * (qb + copysign(sqrt(disc >= 0.0 ? disc : 0.0), qb));
    const double root_1 = q / qa;
    const double root_2 = q != 0.0 ? qc / q : root_1;
    y[0] = 0.0;
    y[1] = Z.r_o;
    y[2] = Z.theta_o;
    y[3] = Z.phi_o;
    y[4] = fabs(root_1 - p_t) <= fabs(root_2 - p_t) ? root_1 : root_2;
    y[5] = p_r;
    y[6] = p_th;
    y[7] = p_ph;
    return bad;
}

// Per-image scalars of the disk shading (physics/accretion.py:130-205,
// physics/redshift.py:63-151, photons/orbits.py:281-292), computed on the
// host.  The emitter is the Keplerian gas at the crossing radius r,
// Omega = omega_num / (r^1.5 + omega_c), in the equatorial metric at
// theta = 0.5 pi (s2, c2 = sin^2, cos^2 of that angle).
struct DiskShade {
    double m, a, a2;          // mass, spin * mass, a ** 2
    double s2, c2;            // np.sin(0.5 pi) ** 2, np.cos(0.5 pi) ** 2
    double omega_num;         // s sign(a) sqrt(M)  (s = +1 prograde, -1 retrograde)
    double omega_c;           // s |a| sqrt(M)
    double r_in, r_out;       // emitting annulus
    double neg_p;             // -emissivity_index
    double n_pow;             // 3 or 4: I_obs = g^n I_em
};

// g, r_hit and I_obs of one disk crossing y (camera.py convention: reversed
// momentum; g is invariant under the reversal).  Returns 0, or FLAG_DISK when
// the reference raises: a non-finite or non-equatorial state, no timelike
// circular orbit at r, or a non-positive frequency ratio.
KR_HD int32_t disk_shade(const DiskShade& D, const double (&y)[8], double& intensity, double& g, double& r_hit) {
    bool finite = true;
    for (int c = 0; c < 8; ++c) finite = finite && isfinite(y[c]);
    const double r = y[1];
    r_hit = r;
    intensity = 0.0;
    g = 0.0;
    if (!finite || fabs(y[2] - 0.5 * 3.141592653589793) > 1e-6) return FLAG_DISK;
    // keplerian_angular_velocity and metric_components at theta = pi/2
    const double omega = D.omega_num / (pow(r, 1.5) + D.omega_c);
    const double sig = r * r + D.a2 * D.c2;
    const double g_tt = -(1.0 - 2.0 * D.m * r / sig);
    const double g_tph = -2.0 * D.m * D.a * r * D.s2 / sig;
    const double g_phph = (r * r + D.a2 + 2.0 * D.m * D.a2 * r * D.s2 / sig) * D.s2;
    const double radicand = -(g_tt + 2.0 * omega * g_tph + omega * omega * g_phph);
    if (!isfinite(radicand) || radicand <= 0.0) return FLAG_DISK;
    const double u_t = 1.0 / sqrt(radicand);
    const double u_ph = omega * u_t;
    // -p_mu u^mu over the four components; the r and theta terms are zeros.
    const double p_dot_u = y[4] * u_t + y[5] * 0.0 + y[6] * 0.0 + y[7] * u_ph;
    g = y[4] / p_dot_u;
    if (!isfinite(g) || g <= 0.0) return FLAG_DISK;
    const double emitted = (r >= D.r_in && r <= D.r_out) ? pow(r / D.r_in, D.neg_p) : 0.0;
    intensity = emitted * pow(g, D.n_pow);
    return 0;
}

// Bilinear value of a 0/1 mask (rows x cols, row-major bytes) at fractional
// (row v, column u), with a ring of zeros around it: boundary.py:95-112 on one
// point, operation for operation.
KR_HD double mask_bilinear(const uint8_t* mask, int64_t rows, int64_t cols, double u, double v) {
    const double up = u + 1.0, vp = v + 1.0;
    const bool inside = up >= 0.0 && up <= static_cast<double>(cols) && vp >= 0.0 && vp <= static_cast<double>(rows);
    if (!inside) return 0.0;
    int64_t j0 = static_cast<int64_t>(floor(up)), i0 = static_cast<int64_t>(floor(vp));
    j0 = j0 < 0 ? 0 : (j0 > cols ? cols : j0);
    i0 = i0 < 0 ? 0 : (i0 > rows ? rows : i0);
    const double fu = pymin(pymax(up - static_cast<double>(j0), 0.0), 1.0);
    const double fv = pymin(pymax(vp - static_cast<double>(i0), 0.0), 1.0);
    const int64_t j1 = j0 + 1 < cols + 1 ? j0 + 1 : cols + 1;
    const int64_t i1 = i0 + 1 < rows + 1 ? i0 + 1 : rows + 1;
    auto at = [&](int64_t i, int64_t j) -> double {
        return (i >= 1 && i <= rows && j >= 1 && j <= cols && mask[(i - 1) * cols + (j - 1)]) ? 1.0 : 0.0;
    };
    return at(i0, j0) * (1.0 - fu) * (1.0 - fv) + at(i0, j1) * fu * (1.0 - fv) + at(i1, j0) * (1.0 - fu) * fv +
           at(i1, j1) * fu * fv;
}

// Sub-pixel edge radius of a mask along one direction (cos_k, sin_k) from
// (ca, cb): samples s_m = ds m, m = 0 .. n_samples - 1, the first one whose
// bilinear value is below 0.5, and the linear crossing between it and the
// sample before (boundary.py:156-170).  Grid: first pixel centre (a0, b0),
// pixel size d, alpha along columns, beta down the rows.  Returns false when
// the centre itself is below 0.5 or no sample is.
KR_HD bool mask_edge_radius(const uint8_t* mask, int64_t rows, int64_t cols, double a0, double b0, double d,
                            double ca, double cb, double cos_k, double sin_k, int64_t n_samples, double& radius) {
    const double ds = 0.125 * d;
    double f_prev = 0.0, s_prev = 0.0;
    for (int64_t m = 0; m < n_samples; ++m) {
        const double s = ds * static_cast<double>(m);
        const double x = ca + cos_k * s;
        const double y = cb + sin_k * s;
        const double f = mask_bilinear(mask, rows, cols, (x - a0) / d, (b0 - y) / d);
        if (f < 0.5) {
            if (m == 0) return false;
            radius = s_prev + ds * (f_prev - 0.5) / (f_prev - f);
            return true;
        }
        f_prev = f;
        s_prev = s;
    }
    return false;
}

// The initial state of the image point (alpha, beta) for the image's camera:
// 0, or the KR_FLAG_* bits of what failed.
KR_HD int32_t image_state(const BardeenImage& B, double alpha, double beta, double (&y)[8]) {
    if (B.camera == CAMERA_ZAMO) return zamo_state(B.Z, alpha, beta, y);
    return bardeen_state(B, alpha, beta, y) ? 0 : FLAG_RADICAND;
}

}  // namespace kr
