// FP32 bulk integration with FP64 where stability needs it.
//
// The photon state, the Dormand-Prince update y_new = y + h sum b_i k_i, the
// embedded error estimate, the error norm, acceptance and step-size control
// stay in FP64 (ray_attempt_with of kr_device.h); what runs in FP32 is the
// expensive part, the six right-hand-side evaluations of an attempt
// (trigonometry, seventeen quotients per evaluation) and the stage inputs
// fed to them.  An attempt switches to the exact FP64 step of the reference
// when the current state is
//   * near the horizon, r < r_+ + horizon_margin (where p_r ~ 1/Delta
//     grows without bound; Delta itself is evaluated in factored form, see
//     Geom32),
//   * near the polar axis, |sin theta| < axis_sin,
//   * near a radial turning point in the strong field,
//     r < turn_rmax and |dr/dlambda| < turn_drdl.
// Photons that then end within a few pixels of the capture boundary, or
// that fail, are re-integrated by the FP64 kernel (kr_mixed.hip).
#pragma once

#include "kr_device.h"

namespace kr {

struct MixedCfg {
    double r_switch;   // r_+ + horizon margin
    double axis_sin;
    double turn_rmax;
    double turn_drdl;
};

// Single-precision copies of the geometry constants, plus the horizons.
// In the FP32 stages the radius is carried as x = r - r_+ (formed in FP64)
// and Delta = r^2 - 2Mr + a^2 is evaluated as x (x + r_+ - r_-): the
// product keeps full relative precision down to the capture radius, where
// the expanded form would lose every digit to cancellation.
struct Geom32 {
    float m, a2, neg_2ma, pos_2ma, two_a2;
    float r_plus, r_gap;  // r_+, r_+ - r_-
    double r_plus64;
};

KR_HD Geom32 geom32(const Params& P) {
    const double m = P.geom[G_MASS];
    const double root = sqrt(m * m - P.geom[G_A2]);
    return Geom32{static_cast<float>(m),
                  static_cast<float>(P.geom[G_A2]),
                  static_cast<float>(P.geom[G_NEG_2MA]),
                  static_cast<float>(P.geom[G_POS_2MA]),
                  static_cast<float>(P.geom[G_TWO_A2]),
                  static_cast<float>(m + root),
                  static_cast<float>(2.0 * root),
                  m + root};
}

// Reciprocal to ~0.5 ulp: hardware estimate plus one Newton step.
KR_HD float frcp(float x) {
#ifdef KR_HOST_BUILD
    return 1.0f / x;
#else
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
#endif
}

// The inverse Kerr metric, its r and theta derivatives and the Hamiltonian
// right-hand side (same formulas as inverse_metric_terms / rhs), in FP32,
// with the seventeen quotients rewritten over three reciprocals
// (1/Sigma, 1/Delta, 1/sin^2) and sin(2 theta) = 2 sin cos.
KR_HD void rhs32(const Geom32& G, const float (&v)[NA], float p_t, float p_ph, float (&f)[NA], float& s_out,
                 float& c_out) {
    // Nothing here has to reproduce the reference's roundings, so products
    // and sums may fuse into FMAs (about a third fewer instructions).
#pragma clang fp contract(fast)
    const float x = v[1], th = v[2];  // x = r - r_+
    const float r = G.r_plus + x;
    float s, c;
#ifdef KR_HOST_BUILD
    s = sinf(th);
    c = cosf(th);
#else
    sincosf(th, &s, &c);
#endif
    const float a2 = G.a2;
    const float s2 = s * s, c2 = c * c;
    const float sin2th = 2.0f * s * c;
    const float r2 = r * r;
    const float sig = r2 + a2 * c2;
    const float dlt = x * (x + G.r_gap);
    const float r2a2 = r2 + a2;
    const float big_a = r2a2 * r2a2 - a2 * dlt * s2;
    const float isig = frcp(sig), idlt = frcp(dlt), is2 = frcp(s2);
    const float isd = isig * idlt;
    const float isd2 = isd * isd, isig2 = isig * isig;
    const float g0 = -big_a * isd;
    const float g1 = G.neg_2ma * r * isd;
    const float g2 = dlt * isig;
    const float g3 = isig;
    const float g4 = (dlt - a2 * s2) * isd * is2;
    const float sd = sig * dlt;
    const float sig_r = 2.0f * r;
    const float sig_th = -a2 * sin2th;
    const float dlt_r = 2.0f * (r - G.m);
    const float a_r = 4.0f * r * r2a2 - G.two_a2 * (r - G.m) * s2;
    const float a_th = -a2 * dlt * sin2th;
    const float sd_r = sig_r * dlt + sig * dlt_r;
    const float sd_th = sig_th * dlt;
    const float r5 = -(a_r * sd - big_a * sd_r) * isd2;
    const float r6 = G.neg_2ma * (sd - r * sd_r) * isd2;
    const float r7 = (dlt_r * sig - dlt * sig_r) * isig2;
    const float r8 = -sig_r * isig2;
    float r9 = -sig_r * isig2 * is2 + a2 * sd_r * isd2;
    const float t10 = -(a_th * sd - big_a * sd_th) * isd2;
    const float t11 = G.pos_2ma * r * sd_th * isd2;
    const float t12 = -dlt * sig_th * isig2;
    const float t13 = -sig_th * isig2;
    float t14 = -(sig_th * s2 + sig * sin2th) * isig2 * is2 * is2 + a2 * sd_th * isd2;
    float gphph = g4;
    if (p_ph == 0.0f) {
        gphph = 0.0f;
        r9 = 0.0f;
        t14 = 0.0f;
    }
    const float p_r = v[4], p_th = v[5];
    f[0] = g0 * p_t + g1 * p_ph;
    f[1] = g2 * p_r;
    f[2] = g3 * p_th;
    f[3] = g1 * p_t + gphph * p_ph;
    const float pt2 = p_t * p_t, ptpph = p_t * p_ph, pr2 = p_r * p_r, pth2 = p_th * p_th, pph2 = p_ph * p_ph;
    f[4] = -0.5f * (r5 * pt2 + 2.0f * r6 * ptpph + r7 * pr2 + r8 * pth2 + r9 * pph2);
    f[5] = -0.5f * (t10 * pt2 + 2.0f * t11 * ptpph + t12 * pr2 + t13 * pth2 + t14 * pph2);
    s_out = s;
    c_out = c;
}

// Dormand-Prince 5(4) with FP32 stage evaluations and an FP64 update and
// error estimate (see the file comment).
KR_HD void dp5_step_mixed(const Geom32& G, const double (&y)[NA], double p_t, double p_ph, double h,
                          const double (&k0)[NA], double (&y_new)[NA], double (&err)[NA], double (&k6)[NA],
                          double& s_new, double& c_new) {
#pragma clang fp contract(fast)
    using namespace dp;
    float y32[NA], k0f[NA], k1[NA], k2[NA], k3[NA], k4[NA], k5[NA], k6f[NA], ys[NA];
    const float pt = static_cast<float>(p_t), pph = static_cast<float>(p_ph);
    const float h32 = static_cast<float>(h);
    float s, c;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        y32[i] = static_cast<float>(i == 1 ? y[i] - G.r_plus64 : y[i]);
        k0f[i] = static_cast<float>(k0[i]);
    }
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y32[i] + h32 * (float(a10) * k0f[i]);
    rhs32(G, ys, pt, pph, k1, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i) ys[i] = y32[i] + h32 * (float(a20) * k0f[i] + float(a21) * k1[i]);
    rhs32(G, ys, pt, pph, k2, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y32[i] + h32 * (float(a30) * k0f[i] + float(a31) * k1[i] + float(a32) * k2[i]);
    rhs32(G, ys, pt, pph, k3, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y32[i] + h32 * (float(a40) * k0f[i] + float(a41) * k1[i] + float(a42) * k2[i] + float(a43) * k3[i]);
    rhs32(G, ys, pt, pph, k4, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i)
        ys[i] = y32[i] + h32 * (float(a50) * k0f[i] + float(a51) * k1[i] + float(a52) * k2[i] +
                                float(a53) * k3[i] + float(a54) * k4[i]);
    rhs32(G, ys, pt, pph, k5, s, c);
    double esum[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        const double d2 = k2[i], d3 = k3[i], d4 = k4[i], d5 = k5[i];
        y_new[i] = y[i] + h * (a60 * k0[i] + a62 * d2 + a63 * d3 + a64 * d4 + a65 * d5);
        esum[i] = e0 * k0[i] + e2 * d2 + e3 * d3 + e4 * d4 + e5 * d5;
        ys[i] = static_cast<float>(i == 1 ? y_new[i] - G.r_plus64 : y_new[i]);
    }
    rhs32(G, ys, pt, pph, k6f, s, c);
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        k6[i] = k6f[i];
        err[i] = h * (esum[i] + e6 * k6[i]);
    }
    s_new = s;
    c_new = c;
}

KR_HD bool needs_fp64(const MixedCfg& M, const RayState<double>& R) {
    const double r = R.y[1];
    if (r < M.r_switch) return true;
    if (fabs(R.s_cur) < M.axis_sin) return true;
    if (r < M.turn_rmax && fabs(R.f[1]) < M.turn_drdl) return true;
    return false;
}

// Chooses the FP32-bulk or the exact FP64 step for each attempt (the FP64
// one with its stage slopes in K; PARK as in dp5_step_stored).
template <bool PARK = false>
struct MixedStepper {
    Geom32 G;
    MixedCfg M;
    int64_t* fp64_attempts;  // per-lane counters (registers), may be null
    SlopeStore<double> K;
    KR_HD bool operator()(const Params& P, const RayState<double>& R, double h_eff, double (&y_new)[NA],
                          double (&err)[NA], double (&k6)[NA], double& s_new, double& c_new) const {
        if (needs_fp64(M, R)) {
            if (fp64_attempts) *fp64_attempts += 1;
            dp5_step_stored<double, PARK>(P, R.y, R.p_t, R.p_ph, h_eff, R.f, K, y_new, err, k6, s_new, c_new);
            return true;
        }
        dp5_step_mixed(G, R.y, R.p_t + 0.0, R.p_ph + 0.0, h_eff, R.f, y_new, err, k6, s_new, c_new);
        return false;
    }
};

// Pixels are classed in NBINS bins by their distance d to the analytic edge,
// which is what sets a photon's length (near the edge it orbits before it
// leaves or falls in).  Bin 0: d <= band, or |xi| = |alpha sin theta_o| <
// axis_lz (photons with L_z near 0 pass close to the poles); bin b >= 1:
// band 2^(b-1) < d <= band 2^b, the last bin open-ended.  Bin 0 is traced in
// FP64; the bin also sets a wavefront's issue priority (band_priority).
constexpr int NBINS = 8;

struct PlanDev {
    double ca, cb;          // centre of the polar radii table
    const double* radii;    // n_angles radii at angles 2 pi k / n_angles
    int32_t n_angles;
    double band;            // band half-width in units of M
    double axis_lz;         // |xi| below which a pixel belongs to bin 0
};

struct PixelPlan {
    int bin;
    bool inside;
};

KR_HD PixelPlan plan_pixel(const PlanDev& pl, const BardeenImage& B, int64_t idx) {
    const int64_t i = idx / B.n, j = idx - (idx / B.n) * B.n;
    const double alpha = pixel_centre(B.fov, j, B.n);
    const double da = alpha - pl.ca;
    const double db = -pixel_centre(B.fov, i, B.n) - pl.cb;
    const double rho = sqrt(da * da + db * db);
    double phi = atan2(db, da);
    if (phi < 0.0) phi += 2.0 * M_PI;
    const double u = phi * (pl.n_angles / (2.0 * M_PI));
    int k0 = static_cast<int>(floor(u));
    const double w = u - k0;
    k0 = ((k0 % pl.n_angles) + pl.n_angles) % pl.n_angles;
    const int k1 = (k0 + 1) % pl.n_angles;
    const double rc = (1.0 - w) * pl.radii[k0] + w * pl.radii[k1];
    const double d = fabs(rho - rc);
    int bin = 0;
    if (!(d <= pl.band || fabs(alpha * B.sin_o) < pl.axis_lz)) {
        bin = 1;
        for (double edge = 2.0 * pl.band; bin < NBINS - 1 && d > edge; edge *= 2.0) ++bin;
    }
    return PixelPlan{bin, rho < rc};
}


}  // namespace kr
