// Host build of the device kernel source (kerrray_amd/csrc/kr_device.h),
// test infrastructure only: compiled with g++ against glibc, it runs exactly
// the per-photon code the GPU runs, so the test suite can show on the CPU
// that the kernel logic reproduces the oracle bit for bit and that the only
// GPU-vs-oracle differences come from the device math library.
#define KR_HOST_BUILD 1
#include <omp.h>

#include "kerrray_hip.h"
#include "kr_device.h"
#include "kr_image.h"
#include "kr_repair.h"
#include <utility>
#include <vector>
#include <cmath>
#include <cstring>

namespace {

template <typename T, bool ADAPTIVE, bool DISK>
void store(const kr::RayResult<T>& r, const kr_outputs* o, int64_t idx) {
    T* Y = static_cast<T*>(o->Y);
    for (int c = 0; c < 8; ++c) Y[idx * 8 + c] = r.y[c];
    o->lam[idx] = r.lam;
    o->n_steps[idx] = r.n_steps;
    o->n_rejected[idx] = r.n_rej;
    o->state[idx] = r.state;
    o->max_null_error[idx] = r.m_null;
    o->max_energy_drift[idx] = r.m_e;
    o->max_lz_drift[idx] = r.m_lz;
    o->max_carter_drift[idx] = r.m_q;
    if (DISK) {
        T* E = static_cast<T*>(o->event_Y);
        for (int c = 0; c < 8; ++c) E[idx * 8 + c] = r.ev[c];
        o->event_hit[idx] = r.ev_hit ? 1 : 0;
    }
}

template <typename T, bool ADAPTIVE, bool DISK>
void run_y0(const kr::Params& P, const double* Y0, int64_t n, const kr_outputs* o, int nt) {
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        double y0[8];
        for (int c = 0; c < 8; ++c) y0[c] = Y0[i * 8 + c];
        kr::RayResult<T> r;
        kr::integrate_ray<T, ADAPTIVE, DISK>(P, y0, r);
        store<T, ADAPTIVE, DISK>(r, o, i);
    }
}

template <typename T, bool ADAPTIVE>
void run_image(const kr::Params& P, const kr::BardeenImage& B, const kr_outputs* o, int nt, int* bad) {
    const int64_t n = B.n;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
    for (int64_t idx = 0; idx < n * n; ++idx) {
        const int64_t i = idx / n, j = idx % n;
        double y0[8];
        if (const int f = kr::image_state(B, kr::pixel_centre(B.fov, j, n), -kr::pixel_centre(B.fov, i, n), y0)) {
#pragma omp atomic
            *bad |= f;
        }
        kr::RayResult<T> r;
        kr::integrate_ray<T, ADAPTIVE, false>(P, y0, r);
        store<T, ADAPTIVE, false>(r, o, idx);
    }
}

kr::Params params_of(const kr_params* p) {
    kr::Params P;
    for (int i = 0; i < 7; ++i) P.geom[i] = p->geom[i];
    for (int i = 0; i < 21; ++i) P.consts[i] = p->consts[i];
    for (int i = 0; i < 3; ++i) P.params[i] = p->params[i];
    for (int i = 0; i < 4; ++i) P.ipar[i] = p->ipar[i];
    return P;
}

}  // namespace

extern "C" int krh_integrate(const kr_params* p, const double* Y0, int64_t n, const kr_outputs* o, int nt) {
    const kr::Params P = params_of(p);
    const bool ad = P.ipar[kr::I_ADAPTIVE] != 0, disk = P.ipar[kr::I_DISK] != 0;
    if (nt <= 0) nt = omp_get_max_threads();
    if (p->dtype == KR_FLOAT64) {
        if (ad) disk ? run_y0<double, true, true>(P, Y0, n, o, nt) : run_y0<double, true, false>(P, Y0, n, o, nt);
        else disk ? run_y0<double, false, true>(P, Y0, n, o, nt) : run_y0<double, false, false>(P, Y0, n, o, nt);
    } else {
        if (ad) disk ? run_y0<float, true, true>(P, Y0, n, o, nt) : run_y0<float, true, false>(P, Y0, n, o, nt);
        else disk ? run_y0<float, false, true>(P, Y0, n, o, nt) : run_y0<float, false, false>(P, Y0, n, o, nt);
    }
    return 0;
}

extern "C" int krh_trace_image(const kr_params* p, const kr_image* img, const kr_outputs* o, int nt) {
    kr::BardeenImage B;
    if (kr::bardeen_image_scalars(img, true, &B)) return 1;
    const kr::Params P = params_of(p);
    if (nt <= 0) nt = omp_get_max_threads();
    int bad = 0;
    const bool ad = P.ipar[kr::I_ADAPTIVE] != 0;
    if (p->dtype == KR_FLOAT64) ad ? run_image<double, true>(P, B, o, nt, &bad) : run_image<double, false>(P, B, o, nt, &bad);
    else ad ? run_image<float, true>(P, B, o, nt, &bad) : run_image<float, false>(P, B, o, nt, &bad);
    return bad ? 3 : 0;
}

// Initial states of every pixel of an image (the device's image_states_kernel
// on the CPU); returns the KR_FLAG_* bits, or -1 for an invalid image.
extern "C" int krh_image_states(const kr_image* img, int reverse, double* Y0) {
    kr::BardeenImage B;
    if (kr::bardeen_image_scalars(img, true, &B)) return -1;
    if (B.camera == kr::CAMERA_ZAMO) B.Z.sign = reverse ? -1.0 : 1.0;
    int bad = 0;
    for (int64_t idx = 0; idx < B.n * B.n; ++idx) {
        const int64_t i = idx / B.n, j = idx % B.n;
        double y[8];
        bad |= kr::image_state(B, kr::pixel_centre(B.fov, j, B.n), -kr::pixel_centre(B.fov, i, B.n), y);
        for (int c = 0; c < 8; ++c) Y0[idx * 8 + c] = y[c];
    }
    return bad;
}

#include "kr_mixed.h"

// Every photon of an image with the stage slopes k1..k4 kept in a
// SlopeStore (an interleaved buffer, as in LDS) instead of registers.
extern "C" int krh_trace_image_stored(const kr_params* p, const kr_image* img, const kr_outputs* o, int nt) {
    kr::BardeenImage B;
    if (kr::bardeen_image_scalars(img, true, &B)) return 1;
    const kr::Params P = params_of(p);
    if (nt <= 0) nt = omp_get_max_threads();
    const int64_t n = B.n;
    constexpr int LANES = 16;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
    for (int64_t idx = 0; idx < n * n; ++idx) {
        const int64_t i = idx / n, j = idx % n;
        double shared[4 * kr::NA * LANES];  // this photon uses column idx % LANES
        const kr::SlopeStore<double> K{&shared[idx % LANES], LANES};
        double y0[8];
        kr::bardeen_state(B, kr::pixel_centre(B.fov, j, n), -kr::pixel_centre(B.fov, i, n), y0);
        kr::RayState<double> R;
        kr::ray_init<double, true>(P, y0, R);
        while (R.state == kr::RUNNING) kr::ray_attempt_with<double, true, false>(P, R, kr::StoredStepper<double, true, true>{K});
        kr::RayResult<double> r;
        kr::ray_result<double, false>(R, r);
        store<double, true, false>(r, o, idx);
    }
    return 0;
}

// Every photon of an image with the FP32-bulk stepper (no band, no repair):
// the CPU view of what the mixed attempt does to accuracy.
extern "C" int krh_trace_image_mixed(const kr_params* p, const kr_image* img, double horizon_margin,
                                     double axis_sin, double turn_rmax, double turn_drdl, const kr_outputs* o,
                                     int64_t* fp64_attempts, int nt) {
    kr::BardeenImage B;
    if (kr::bardeen_image_scalars(img, true, &B)) return 1;
    const kr::Params P = params_of(p);
    if (nt <= 0) nt = omp_get_max_threads();
    const double r_plus = P.geom[kr::G_MASS] + std::sqrt(P.geom[kr::G_MASS] * P.geom[kr::G_MASS] - P.geom[kr::G_A2]);
    const kr::MixedCfg M{r_plus + horizon_margin, axis_sin, turn_rmax, turn_drdl};
    const kr::Geom32 G = kr::geom32(P);
    const int64_t n = B.n;
    int64_t total_fp64 = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16) reduction(+ : total_fp64)
    for (int64_t idx = 0; idx < n * n; ++idx) {
        const int64_t i = idx / n, j = idx % n;
        double y0[8];
        kr::bardeen_state(B, kr::pixel_centre(B.fov, j, n), -kr::pixel_centre(B.fov, i, n), y0);
        kr::RayState<double> R;
        kr::ray_init<double, true>(P, y0, R);
        int64_t f64 = 0;
        double slopes[4 * kr::NA];
        const kr::MixedStepper<> step{G, M, &f64, kr::SlopeStore<double>{slopes, 1}};
        while (R.state == kr::RUNNING) kr::ray_attempt_with<double, true, false>(P, R, step);
        kr::RayResult<double> r;
        kr::ray_result<double, false>(R, r);
        store<double, true, false>(r, o, idx);
        total_fp64 += f64;
    }
    *fp64_attempts = total_fp64;
    return 0;
}

#include <vector>

#include "kr_adjoint.h"

// The discrete adjoint on the CPU, with the kernels' algorithm: trace,
// record accepted steps and checkpoints every `segment` steps, then per
// photon sweep the segments backwards, recomputing each from its checkpoint;
// segment 0: record every accepted step's stage states and chain the step
// Jacobians built from them.
// theta/phi = escape direction at infinity (NaN unless escaped);
// grad[4 i + j] = g_theta[i] d theta_inf/dx_j + g_phi[i] d phi_inf/dx_j,
// x = (alpha, beta, spin, theta_o); 0 for photons that did not escape.
// Copies the stage states of one attempt into a 6 x NA buffer.
struct HostStages {
    double* buf;
    void operator()(int s, const double (&ys)[kr::NA]) const {
        for (int c = 0; c < kr::NA; ++c) buf[s * kr::NA + c] = ys[c];
    }
};

extern "C" int krh_escape_vjp(const kr_params* p, const kr_image* img, const double* alpha, const double* beta,
                              int64_t n, const double* g_theta, const double* g_phi, int segment, int64_t* state,
                              double* theta, double* phi, double* grad, int nt) {
    kr::BardeenImage B;
    if (kr::bardeen_image_scalars(img, false, &B)) return 1;
    const kr::Params P = params_of(p);
    if (nt <= 0) nt = omp_get_max_threads();
    const double mass = img->mass, a = img->spin * img->mass;
    int bad = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4) reduction(| : bad)
    for (int64_t i = 0; i < n; ++i) {
        double y0[8];
        kr::bardeen_state(B, alpha[i], beta[i], y0);
        kr::RayState<double> R;
        kr::ray_init<double, true>(P, y0, R);
        std::vector<double> hs;
        std::vector<double> ck;
        for (int c = 0; c < kr::NA; ++c) ck.push_back(R.y[c]);
        while (R.state == kr::RUNNING) {
            const double remaining = P.params[kr::P_LAMMAX] - R.lam;
            const double h_eff = (R.h >= remaining) ? remaining : R.h;
            const int64_t k = R.n_steps;
            kr::ray_attempt<double, true, false>(P, R);
            if (R.n_steps != k) {
                hs.push_back(h_eff);
                if (segment > 0 && R.n_steps % segment == 0 && R.state == kr::RUNNING)
                    for (int c = 0; c < kr::NA; ++c) ck.push_back(R.y[c]);
            }
        }
        kr::RayResult<double> r;
        kr::ray_result<double, false>(R, r);
        state[i] = r.state;
        theta[i] = NAN;
        phi[i] = NAN;
        for (int j = 0; j < 4; ++j) grad[i * 4 + j] = 0.0;
        if (r.state != kr::ESCAPED) continue;
        const double pt = y0[4], pph = y0[7];
        kr::ad::escape_tail(R.y[1], R.y[2], R.y[3], R.y[5], pt, pph, a, mass, theta[i], phi[i]);
        const int64_t N = static_cast<int64_t>(hs.size());
        double ybar[kr::NA], ptbar, pphbar, abar;
        kr::ad::escape_tail_vjp(mass, a, R.y, pt, pph, g_theta[i], g_phi[i], ybar, ptbar, pphbar, abar);
        if (segment == 0) {
            // Step Jacobians from the recorded stage states, then the
            // transposed chain.
            std::vector<double> stg;
            kr::RayState<double> Q;
            kr::ray_init<double, true>(P, y0, Q);
            double cur[6 * kr::NA];
            const kr::ad::StageStepper<HostStages> rec{HostStages{cur}};
            while (Q.state == kr::RUNNING) {
                const int64_t k = Q.n_steps;
                kr::ray_attempt_with<double, true, false>(P, Q, rec);
                if (Q.n_steps != k) stg.insert(stg.end(), cur, cur + 6 * kr::NA);
            }
            for (int c = 0; c < kr::NA; ++c)
                if (!(Q.y[c] == R.y[c])) bad |= 2;
            for (int64_t k = N - 1; k >= 0; --k) {
                double D[kr::ad::JW];
                const double* sk = &stg[static_cast<size_t>(k) * 6 * kr::NA];
                auto stage = [&](int s, double (&Y)[kr::NA]) {
                    for (int c = 0; c < kr::NA; ++c) Y[c] = sk[s * kr::NA + c];
                };
                auto out = [&](int r, int c, double v) { D[r * kr::ad::JC + c] = v; };
                kr::ad::step_increment<0, kr::ad::JC>(mass, a, pt, pph, hs[k], stage, out);
                kr::ad::increment_vjp([&](int e) { return D[e]; }, ybar, ptbar, pphbar, abar);
            }
        }
        const int64_t nseg = segment == 0 ? 0 : (N + segment - 1) / segment;
        std::vector<double> seg(static_cast<size_t>(segment) * kr::NA);
        for (int64_t s = nseg - 1; s >= 0; --s) {
            double y[kr::NA];
            for (int c = 0; c < kr::NA; ++c) y[c] = ck[s * kr::NA + c];
            const int64_t kb = s * segment, ke = std::min<int64_t>(N, kb + segment);
            for (int64_t k = kb; k < ke; ++k) {
                for (int c = 0; c < kr::NA; ++c) seg[(k - kb) * kr::NA + c] = y[c];
                kr::ad::step_forward(P, y, pt, pph, hs[k]);
            }
            if (s == nseg - 1)
                for (int c = 0; c < kr::NA; ++c)
                    if (!(y[c] == r.y[kr::full_index(c)])) bad |= 2;
            for (int64_t k = ke - 1; k >= kb; --k) {
                for (int c = 0; c < kr::NA; ++c) y[c] = seg[(k - kb) * kr::NA + c];
                kr::ad::step_adjoint(P, mass, a, y, pt, pph, hs[k], ybar, ptbar, pphbar, abar);
            }
        }
        kr::ad::Dual<4> yd[8];
        kr::ad::bardeen_state_of<4>(kr::ad::seed<4>(alpha[i], 0), kr::ad::seed<4>(beta[i], 1),
                                    kr::ad::seed<4>(a, 2), kr::ad::seed<4>(B.theta_o, 3), mass, B.r_o, yd);
        for (int j = 0; j < 4; ++j) {
            double g = ptbar * yd[4].d[j] + pphbar * yd[7].d[j];
            for (int c = 0; c < kr::NA; ++c) g += ybar[c] * yd[kr::full_index(c)].d[j];
            if (j == 2) g = (g + abar) * mass;
            grad[i * 4 + j] = g;
        }
    }
    return bad;
}

// Building blocks for unit tests: v^T df at one state, and one step's
// forward map and reverse sweep (out = ybar[6], ptbar, pphbar, abar).
extern "C" void krh_rhs_vjp(double m, double a, const double* Y, double pt, double pph, const double* v, double* out) {
    double y[kr::NA], vv[kr::NA], yb[kr::NA] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < kr::NA; ++c) {
        y[c] = Y[c];
        vv[c] = v[c];
    }
    double ptb = 0.0, pphb = 0.0, ab = 0.0;
    kr::ad::rhs_vjp(m, a, y, pt, pph, vv, yb, ptb, pphb, ab);
    for (int c = 0; c < kr::NA; ++c) out[c] = yb[c];
    out[6] = ptb;
    out[7] = pphb;
    out[8] = ab;
}

extern "C" void krh_step_forward(const kr_params* p, const double* Y, double pt, double pph, double h, double* out) {
    const kr::Params P = params_of(p);
    double y[kr::NA];
    for (int c = 0; c < kr::NA; ++c) y[c] = Y[c];
    kr::ad::step_forward(P, y, pt, pph, h);
    for (int c = 0; c < kr::NA; ++c) out[c] = y[c];
}

extern "C" void krh_step_adjoint(const kr_params* p, double m, double a, const double* Y, double pt, double pph,
                                 double h, const double* ybar_new, double* out) {
    const kr::Params P = params_of(p);
    double y[kr::NA], yb[kr::NA];
    for (int c = 0; c < kr::NA; ++c) {
        y[c] = Y[c];
        yb[c] = ybar_new[c];
    }
    double ptb = 0.0, pphb = 0.0, ab = 0.0;
    kr::ad::step_adjoint(P, m, a, y, pt, pph, h, yb, ptb, pphb, ab);
    for (int c = 0; c < kr::NA; ++c) out[c] = yb[c];
    out[6] = ptb;
    out[7] = pphb;
    out[8] = ab;
}

// One step's increment D (NA x 7, row-major; columns r, theta, p_r, p_theta,
// p_t, p_phi, a) from the state Y, through the step's own stage states.
extern "C" void krh_step_increment(const kr_params* p, double m, double a, const double* Y, double pt, double pph,
                                   double h, double* out) {
    const kr::Params P = params_of(p);
    double y[kr::NA], k0[kr::NA], y_new[kr::NA], err[kr::NA], k6[kr::NA], s, c;
    for (int i = 0; i < kr::NA; ++i) y[i] = Y[i];
    kr::rhs(P, y, pt, pph, k0, s, c);
    double stg[6 * kr::NA];
    kr::dp5_step(P, y, pt, pph, h, k0, y_new, err, k6, s, c, HostStages{stg});
    auto stage = [&](int st, double (&Ys)[kr::NA]) {
        for (int i = 0; i < kr::NA; ++i) Ys[i] = stg[st * kr::NA + i];
    };
    auto put = [&](int r, int col, double v) { out[r * kr::ad::JC + col] = v; };
    kr::ad::step_increment<0, 4>(m, a, pt, pph, h, stage, put);
    kr::ad::step_increment<4, 3>(m, a, pt, pph, h, stage, put);
}

// The FP32-bulk controller's pieces on n inputs: out[3 i] = pow_m_fifth(x_i)
// (x_i in [1e-5, 1e5]), out[3 i + 1] = the exact step factor of en = x_i,
// out[3 i + 2] = the bulk step factor (neither rejected before).
extern "C" void krh_bulk_controller(const double* x, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        const double en = x[i];
        out[3 * i] = (en >= 1e-5 && en <= 1e5) ? kr::pow_m_fifth(en) : 0.0;
        out[3 * i + 1] = kr::step_factor(kr::fixed::safety, kr::fixed::factor_min, kr::fixed::factor_max,
                                         kr::fixed::exponent, en, false);
        out[3 * i + 2] = kr::bulk_step_factor(en, false);
    }
}

// div_by (kr_device.h) against the IEEE quotient on n operand pairs:
// returns the number of results whose bits differ.
extern "C" int64_t krh_div_by_mismatches(const double* a, const double* b, int64_t n) {
    int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (int64_t i = 0; i < n; ++i) {
        const double q = a[i] / b[i];
        const double m = kr::div_by(a[i], b[i], 1.0 / b[i]);
        bad += std::memcmp(&q, &m, sizeof(double)) != 0;
    }
    return bad;
}

// Disk shading of n crossing rows by the kernel source (disk_shade); returns
// the KR_FLAG_* bits, or -1 for an invalid disk.
extern "C" int krh_disk_shade(const kr_disk* disk, double mass, double spin, const double* Y, int64_t n,
                              double* intensity, double* g, double* r_hit) {
    kr::DiskShade D;
    if (kr::disk_shade_scalars(disk, mass, spin, &D)) return -1;
    int bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        double y[8];
        for (int c = 0; c < 8; ++c) y[c] = Y[i * 8 + c];
        bad |= kr::disk_shade(D, y, intensity[i], g[i], r_hit[i]);
    }
    return bad;
}

// Mask edge radii by the kernel source (mask_edge_radius), one direction at a
// time; returns the number of directions that failed.
extern "C" int krh_mask_edge(const uint8_t* mask, int64_t rows, int64_t cols, double a0, double b0, double d,
                             double ca, double cb, const double* cosk, const double* sink, int64_t n_angles,
                             double reach, double* radii) {
    const int64_t n_samples = static_cast<int64_t>(std::ceil(reach / (0.125 * d)) + 1.0);
    int bad = 0;
    for (int64_t k = 0; k < n_angles; ++k)
        if (!kr::mask_edge_radius(mask, rows, cols, a0, b0, d, ca, cb, cosk[k], sink[k], n_samples, radii[k])) ++bad;
    return bad;
}

// The verify/repair rounds of a planned trace (kr_mixed.hip: verify_kernel,
// repair_kernel) run serially with the same kr_repair.h functions, the FP64
// trace of a pixel replaced by a given class: state holds the traced classes
// (updated in place), state64 the FP64 class of every pixel, marks the mark
// bytes (MARK_FP64 for pixels traced in FP64).  listed[r] (max_rounds + 1
// entries) receives round r's list length, listed[max_rounds] the pushes
// past the last round.  Returns FLAG_REPAIR when the limit was hit, else 0.
extern "C" int krh_repair_rounds(int64_t n, int band_px, int64_t* state, const int64_t* state64, uint8_t* marks,
                                 int max_rounds, int64_t* listed) {
    std::vector<int64_t> cur, next;
    for (int64_t idx = 0; idx < n * n; ++idx)
        if (kr::verify_pixel(state, marks, n, band_px, idx)) cur.push_back(idx);
    for (int64_t idx : cur) marks[idx] = kr::MARK_QUEUED;
    int flags = 0;
    for (int round = 0; round < max_rounds; ++round) {
        listed[round] = static_cast<int64_t>(cur.size());
        int64_t pushed = 0;
        next.clear();
        for (int64_t idx : cur) {
            const bool was_captured = state[idx] == kr::CAPTURED;
            state[idx] = state64[idx];
            kr::mark_or(marks, idx, kr::MARK_FP64);
            if ((state[idx] == kr::CAPTURED) == was_captured) continue;
            kr::for_each_neighbour(n, band_px, idx, [&](int64_t pos) {
                if (kr::mark_or(marks, pos, kr::MARK_QUEUED) != 0) return;
                ++pushed;
                if (round + 1 < max_rounds)
                    next.push_back(pos);
                else
                    flags |= kr::FLAG_REPAIR;
            });
        }
        listed[round + 1] = pushed;
        std::swap(cur, next);
    }
    return flags;
}

// The device's initial state of a critical-impact probe (kr_device.h:
// impact_state) for n angular momenta lz; Y (n, 8).  Returns the number of
// states whose radicand check failed.
extern "C" int krh_impact_states(double m, double a, double r0, double theta, double c2, double s2, const double* lz,
                                 int64_t n, double* Y) {
    const kr::ImpactLaunch L{m, a, r0, theta, c2, s2};
    int bad = 0;
    for (int64_t k = 0; k < n; ++k) {
        double y[8];
        if (!kr::impact_state(L, lz[k], y)) ++bad;
        for (int c = 0; c < 8; ++c) Y[k * 8 + c] = y[c];
    }
    return bad;
}

// The step controller's power, as the device computes it (kr_device.h), on n
// arguments: pow_m02_cr (double) and powf_m02_cr (float).
extern "C" void krh_pow_m02(const double* x, int64_t n, double* out) {
    for (int64_t k = 0; k < n; ++k) out[k] = kr::pow_m02_cr(x[k]);
}
extern "C" void krh_powf_m02(const float* x, int64_t n, float* out) {
    for (int64_t k = 0; k < n; ++k) out[k] = kr::powf_m02_cr(x[k]);
}

// Agreement with the C library: over n arguments log-uniform in [1e-5, 1e5]
// (a splitmix64 sequence from seed), the number whose pow_m02_cr differs from
// pow(x, -0.2); the first `keep` of them go to bad_x.
extern "C" int64_t krh_pow_m02_disagreements(int64_t n, uint64_t seed, double* bad_x, int64_t keep) {
    int64_t bad = 0;
    uint64_t s = seed;
    for (int64_t k = 0; k < n; ++k) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = static_cast<double>(z >> 11) * 0x1p-53;
        const double x = std::exp(std::log(1e-5) + u * (std::log(1e5) - std::log(1e-5)));
        if (kr::pow_m02_cr(x) != std::pow(x, -0.2)) {
            if (bad < keep) bad_x[bad] = x;
            ++bad;
        }
    }
    return bad;
}

// The same for powf over every float in [lo, hi] (exhaustive).
extern "C" int64_t krh_powf_m02_disagreements(float lo, float hi, float* bad_x, int64_t keep, int64_t* tested) {
    int64_t bad = 0, count = 0;
    for (float x = lo; x <= hi; x = std::nextafter(x, 2.0f * hi)) {
        ++count;
        if (kr::powf_m02_cr(x) != std::pow(x, -0.2f)) {
            if (bad < keep) bad_x[bad] = x;
            ++bad;
        }
    }
    *tested = count;
    return bad;
}
