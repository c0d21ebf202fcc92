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
#include <cmath>

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
        if (!kr::bardeen_state(B, kr::pixel_centre(B.fov, j, n), -kr::pixel_centre(B.fov, i, n), y0)) *bad = 1;
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

#include "kr_mixed.h"

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
        const kr::MixedStepper step{G, M, &f64};
        while (R.state == kr::RUNNING) kr::ray_attempt_with<double, true, false>(P, R, step);
        kr::RayResult<double> r;
        kr::ray_result<double, false>(R, r);
        store<double, true, false>(r, o, idx);
        total_fp64 += f64;
    }
    *fp64_attempts = total_fp64;
    return 0;
}
