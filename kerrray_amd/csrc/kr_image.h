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

// The ZAMO frame at (r, th): metric_components and inverse_metric_components
// (geometry/metric.py:147-195) and zamo_tetrad (camera.py:225-262) as NumPy
// evaluates them on float64 scalars (x ** 2 of an array is x * x, a ** 2 of
// the Python float a is pow(a, 2.0) = a * a, np.sin/np.cos are glibc's).
// sign is -1 for the reversed (backward-traced) photon, +1 for the physical
// one.  Returns nullptr, or a message when the frame does not exist.
inline const char* zamo_frame_scalars(double m, double a, double r, double th, double ph, double sign,
                                      ZamoImage* Z) {
    if (!std::isfinite(ph)) return "phi_o must be finite";
    const double s = libm_sin(th), c = libm_cos(th);
    const double s2 = s * s;
    const double a2 = libm_pow(a, 2.0);
    const double sig = r * r + a2 * (c * c);
    const double dlt = r * r - 2.0 * m * r + a2;
    // metric_components
    const double g_tt = -(1.0 - 2.0 * m * r / sig);
    const double g_tph = -2.0 * m * a * r * s2 / sig;
    const double g_rr = sig / dlt;
    const double g_thth = sig + 0.0 * dlt;
    const double g_phph = (r * r + a2 + 2.0 * m * a2 * r * s2 / sig) * s2;
    // zamo_tetrad
    const double omega = -g_tph / g_phph;
    const double lapse2 = -(g_tt - g_tph * g_tph / g_phph);
    if (!(lapse2 > 0.0)) return "lapse^2 = Sigma Delta / A is not positive at the observer";
    const double lapse = std::sqrt(lapse2);
    // inverse_metric_components
    const double r2a2 = r * r + a2;
    const double big_a = r2a2 * r2a2 - a2 * dlt * s2;
    const double sd = sig * dlt;
    Z->r_o = r;
    Z->theta_o = th;
    Z->phi_o = ph;
    Z->r_o2 = libm_pow(r, 2.0);
    Z->e00 = 1.0 / lapse;
    Z->e03 = omega / lapse;
    Z->e11 = 1.0 / std::sqrt(g_rr);
    Z->e22 = 1.0 / std::sqrt(g_thth);
    Z->e33 = 1.0 / std::sqrt(g_phph);
    Z->g_tt = g_tt;
    Z->g_tph = g_tph;
    Z->g_rr = g_rr;
    Z->g_thth = g_thth;
    Z->g_phph = g_phph;
    Z->gi_tt = -big_a / sd;
    Z->gi_tph = -2.0 * m * a * r / sd;
    Z->gi_rr = dlt / sig;
    Z->gi_thth = 1.0 / sig + 0.0 * dlt;
    Z->gi_phph = (dlt - a2 * s2) / (sd * s2);
    Z->sign = sign;
    if (!(Z->gi_tt < 0.0)) return "null re-normalisation needs g^tt < 0 at the observer";
    return nullptr;
}

// Fills *B with the Python float expressions of rayset.py:158-165 and
// initial_conditions.py:50-68 (x ** 2 is pow(x, 2.0); math.sin/math.cos),
// and for a ZAMO camera its frame (reversed photons, as the tracer takes
// them).  Returns nullptr, or a message for an invalid image.
inline const char* bardeen_image_scalars(const kr_image* img, bool need_grid, BardeenImage* B) {
    if (!img) return "image is NULL";
    if (img->camera != KR_CAMERA_BARDEEN && img->camera != KR_CAMERA_ZAMO)
        return "camera must be KR_CAMERA_BARDEEN or KR_CAMERA_ZAMO";
    const double m = img->mass, spin = img->spin, r = img->r_o, th = img->theta_o;
    if (!(std::isfinite(m) && m > 0.0)) return "mass must be finite and > 0";
    if (!(std::isfinite(spin) && std::fabs(spin) < 1.0)) return "spin must satisfy |spin| < 1";
    const double a = spin * m;
    const double r_plus = m + std::sqrt(m * m - a * a);
    if (!(std::isfinite(r) && r > r_plus)) return "r_o must lie outside the outer horizon";
    if (!(th > 0.0 && th < M_PI)) return "theta_o must lie strictly between 0 and pi";
    if (need_grid) {
        // n * n rows at most 2^31: one lane per pixel fits a launch grid.
        if (img->resolution < 1 || img->resolution > 46340) return "resolution must be in [1, 46340]";
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
    B->camera = img->camera;
    B->Z = ZamoImage{};
    if (img->camera == KR_CAMERA_ZAMO) {
        if (const char* err = zamo_frame_scalars(m, a, r, th, img->phi_o, -1.0, &B->Z)) return err;
        if (need_grid) {
            // The corner pixels are the farthest from the centre
            // (camera.py:342-349 checks every pixel).
            const double x = pixel_centre(img->fov, 0, img->resolution);
            if (!((x * x + x * x) / B->Z.r_o2 < 1.0))
                return "fov is too large for r_o: pixels with alpha^2 + beta^2 >= r_o^2 are not directions on the sky";
        }
    }
    return nullptr;
}

// The disk shading scalars of DiskModel (physics/accretion.py:130-178) in
// spacetime (mass, spin): keplerian_angular_velocity's sign and offset
// (photons/orbits.py:281-292, Python int and float arithmetic) and the
// equatorial sin^2, cos^2 of metric_components(st, r, 0.5 pi).  Returns
// nullptr, or a message for an invalid disk.
inline const char* disk_shade_scalars(const kr_disk* disk, double mass, double spin, DiskShade* D) {
    if (!disk) return "disk is NULL";
    if (!(std::isfinite(mass) && mass > 0.0)) return "mass must be finite and > 0";
    if (!(std::isfinite(spin) && std::fabs(spin) < 1.0)) return "spin must satisfy |spin| < 1";
    if (!(std::isfinite(disk->r_in) && std::isfinite(disk->r_out) && 0.0 < disk->r_in && disk->r_in < disk->r_out))
        return "need finite 0 < r_in < r_out";
    if (!std::isfinite(disk->emissivity_index)) return "emissivity_index must be finite";
    if (disk->redshift_power != 3 && disk->redshift_power != 4) return "redshift_power must be 3 or 4";
    if (disk->prograde != 0 && disk->prograde != 1) return "prograde must be 0 or 1";
    const double a = spin * mass;
    const double th = 0.5 * M_PI;
    const double sq_m = std::sqrt(mass);
    const double s = disk->prograde ? 1.0 : -1.0;
    const double sign_a = a < 0.0 ? -1.0 : 1.0;
    D->m = mass;
    D->a = a;
    D->a2 = libm_pow(a, 2.0);
    D->s2 = libm_sin(th) * libm_sin(th);
    D->c2 = libm_cos(th) * libm_cos(th);
    D->omega_num = (s * sign_a) * sq_m;
    D->omega_c = s * std::fabs(a) * sq_m;
    D->r_in = disk->r_in;
    D->r_out = disk->r_out;
    D->neg_p = -disk->emissivity_index;
    D->n_pow = static_cast<double>(disk->redshift_power);
    return nullptr;
}

}  // namespace kr
