// Host-side precompute of the per-image scalars of the Bardeen initial
// conditions.  Plain C++ (no HIP), shared by the C ABI and the test suite's
// host build of the kernel source.
#pragma once

#include <cmath>

#include "kerrray_hip.h"
#include "kr_device.h"

namespace kr {

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

// Fills *B with the Python float expressions of rayset.py:158-165 and
// initial_conditions.py:50-68 (x ** 2 is pow(x, 2.0); math.sin/math.cos).
// Returns nullptr, or a message for an invalid image.
inline const char* bardeen_image_scalars(const kr_image* img, bool need_grid, BardeenImage* B) {
    if (!img) return "image is NULL";
    const double m = img->mass, spin = img->spin, r = img->r_o, th = img->theta_o;
    if (!(std::isfinite(m) && m > 0.0)) return "mass must be finite and > 0";
    if (!(std::isfinite(spin) && std::fabs(spin) < 1.0)) return "spin must satisfy |spin| < 1";
    const double a = spin * m;
    const double r_plus = m + std::sqrt(m * m - a * a);
    if (!(std::isfinite(r) && r > r_plus)) return "r_o must lie outside the outer horizon";
    if (!(th > 0.0 && th < M_PI)) return "theta_o must lie strictly between 0 and pi";
    if (need_grid) {
        if (img->resolution < 1 || img->resolution > (1 << 20)) return "resolution must be in [1, 2^20]";
        if (!(std::isfinite(img->fov) && img->fov > 0.0)) return "fov must be finite and > 0";
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
    return nullptr;
}

}  // namespace kr
