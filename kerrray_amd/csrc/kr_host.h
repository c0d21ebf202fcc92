// Host-side helpers shared by the C ABI translation units: error state,
// handle layout, argument validation and the per-image scalar precompute.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "kerrray_hip.h"
#include "kr_device.h"

struct kr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr;
    hipEvent_t ev1 = nullptr;
    void* buf = nullptr;
    size_t buf_bytes = 0;
};

#define KR_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t kr_e_ = (expr);                                                                     \
        if (kr_e_ != hipSuccess)                                                                       \
            return ::kr::host::fail(KR_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(kr_e_)); \
    } while (0)

namespace kr {
namespace host {

inline std::string& last_error() {
    static thread_local std::string msg;
    return msg;
}

inline void clear_error() { last_error().clear(); }

inline kr_status fail(kr_status st, const std::string& msg) {
    last_error() = msg;
    return st;
}

// Sequential sub-allocation of the handle's device buffer (256-byte aligned).
struct Carver {
    size_t total = 0;
    size_t take(int64_t bytes) {
        const size_t off = total;
        total += (static_cast<size_t>(bytes < 0 ? 0 : bytes) + 255) & ~static_cast<size_t>(255);
        return off;
    }
};

inline kr_status reserve(kr_context* h, size_t bytes) {
    if (bytes <= h->buf_bytes) return KR_OK;
    if (h->buf) {
        KR_HIP(hipStreamSynchronize(h->stream));
        KR_HIP(hipFree(h->buf));
        h->buf = nullptr;
        h->buf_bytes = 0;
    }
    const size_t rounded = (bytes + (1u << 20) - 1) & ~static_cast<size_t>((1u << 20) - 1);
    KR_HIP(hipMalloc(&h->buf, rounded));
    h->buf_bytes = rounded;
    return KR_OK;
}

inline bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// The tableau the kernels implement (tableaus.py:59-86, dense as in
// numba_backend.py:118-128).
inline bool tableau_ok(const kr_params* p) {
    using namespace kr::dp;
    const double A[49] = {0, 0, 0, 0, 0, 0, 0,
                          a10, 0, 0, 0, 0, 0, 0,
                          a20, a21, 0, 0, 0, 0, 0,
                          a30, a31, a32, 0, 0, 0, 0,
                          a40, a41, a42, a43, 0, 0, 0,
                          a50, a51, a52, a53, a54, 0, 0,
                          a60, 0.0, a62, a63, a64, a65, 0};
    const double E[7] = {e0, 0.0 - 0.0, e2, e3, e4, e5, e6};
    const double B[4] = {kr::rk4::b0, kr::rk4::b1, kr::rk4::b2, kr::rk4::b3};
    const double C[4] = {0.0, kr::rk4::c1, kr::rk4::c2, kr::rk4::c3};
    for (int i = 0; i < 49; ++i)
        if (!same_bits(A[i], p->tab_a[i])) return false;
    for (int i = 0; i < 7; ++i)
        if (!same_bits(E[i], p->tab_e[i])) return false;
    for (int i = 0; i < 4; ++i)
        if (!same_bits(B[i], p->rk4_b[i]) || !same_bits(C[i], p->rk4_c[i])) return false;
    return true;
}

// Checks a kr_params block and copies it into the kernel-argument layout.
inline kr_status validate(kr_context* h, const kr_params* p, int64_t n, kr::Params* out) {
    if (!h) return fail(KR_ERR_INVALID, "handle is NULL");
    if (!p) return fail(KR_ERR_INVALID, "params is NULL");
    if (n < 0) return fail(KR_ERR_INVALID, "negative photon count");
    if (p->dtype != KR_FLOAT64 && p->dtype != KR_FLOAT32) return fail(KR_ERR_INVALID, "dtype must be 0 or 1");
    const double fixed[5] = {0.0, 1.0, 2.0, -0.5, 8.0};
    for (int i = 0; i < 5; ++i)
        if (!same_bits(p->consts[i], fixed[i]))
            return fail(KR_ERR_INVALID, "consts[0:5] must be (0, 1, 2, -0.5, 8) as in numba_backend.py:100-108");
    if (!tableau_ok(p))
        return fail(KR_ERR_UNSUPPORTED, "only the Dormand-Prince 5(4) / classical RK4 tableaus are implemented");
    if (p->ipar[kr::I_ADAPTIVE] != 0 && p->ipar[kr::I_ADAPTIVE] != 1)
        return fail(KR_ERR_INVALID, "ipar[2] (adaptive) must be 0 or 1");
    if (p->ipar[kr::I_DISK] != 0 && p->ipar[kr::I_DISK] != 1) return fail(KR_ERR_INVALID, "ipar[3] (disk) must be 0 or 1");
    if (p->ipar[kr::I_MAX_STEPS] < 1 || p->ipar[kr::I_GUARD] < 1)
        return fail(KR_ERR_INVALID, "ipar[0] (max_steps) and ipar[1] (guard) must be >= 1");
    if (!(p->params[kr::P_LAMMAX] > 0.0) || std::isnan(p->params[kr::P_HMAX]))
        return fail(KR_ERR_INVALID, "params: lambda_max must be > 0 and h_max a number or +inf");
    std::memcpy(out->geom, p->geom, sizeof(out->geom));
    std::memcpy(out->consts, p->consts, sizeof(out->consts));
    std::memcpy(out->params, p->params, sizeof(out->params));
    std::memcpy(out->ipar, p->ipar, sizeof(out->ipar));
    return KR_OK;
}

// libm calls made through volatile pointers so the compiler cannot merge a
// sin/cos pair into sincos (glibc's sincos is not bit-identical to the
// separate calls the reference makes through Python's math module).
inline double libm_sin(double x) {
    static double (*volatile f)(double) = ::sin;
    return f(x);
}
inline double libm_cos(double x) {
    static double (*volatile f)(double) = ::cos;
    return f(x);
}
inline double libm_pow(double x, double y) {
    static double (*volatile f)(double, double) = ::pow;
    return f(x, y);
}

// Per-image scalars of the Bardeen initial conditions, evaluated with the
// Python float expressions of rayset.py:158-165 and
// initial_conditions.py:50-68 (x ** 2 is pow(x, 2.0); math.sin/math.cos).
inline kr_status bardeen_image(const kr_image* img, bool need_grid, kr::BardeenImage* B) {
    if (!img) return fail(KR_ERR_INVALID, "image is NULL");
    const double m = img->mass, spin = img->spin, r = img->r_o, th = img->theta_o;
    if (!(std::isfinite(m) && m > 0.0)) return fail(KR_ERR_INVALID, "mass must be finite and > 0");
    if (!(std::isfinite(spin) && std::fabs(spin) < 1.0)) return fail(KR_ERR_INVALID, "spin must satisfy |spin| < 1");
    const double a = spin * m;
    const double r_plus = m + std::sqrt(m * m - a * a);
    if (!(std::isfinite(r) && r > r_plus)) return fail(KR_ERR_INVALID, "r_o must lie outside the outer horizon");
    if (!(th > 0.0 && th < M_PI)) return fail(KR_ERR_INVALID, "theta_o must lie strictly between 0 and pi");
    if (need_grid) {
        if (img->resolution < 1 || img->resolution > (1 << 20))
            return fail(KR_ERR_INVALID, "resolution must be in [1, 2^20]");
        if (!(std::isfinite(img->fov) && img->fov > 0.0)) return fail(KR_ERR_INVALID, "fov must be finite and > 0");
    }
    const double sin_o = libm_sin(th), cos_o = libm_cos(th);
    B->r_o = r;
    B->theta_o = th;
    B->sin_o = sin_o;
    B->a = a;
    B->a2c2 = libm_pow(a, 2.0) * libm_pow(cos_o, 2.0);
    B->cot2 = libm_pow(cos_o / sin_o, 2.0);
    B->r2a2 = 1.0 * (r * r + a * a);
    B->dlt = r * r - 2.0 * m * r + libm_pow(a, 2.0);
    B->c2 = libm_pow(libm_cos(th), 2.0);
    B->s2 = libm_pow(libm_sin(th), 2.0);
    B->aaEE = a * a * 1.0 * 1.0;
    B->fov = img->fov;
    B->n = img->resolution;
    return KR_OK;
}

}  // namespace host
}  // namespace kr
