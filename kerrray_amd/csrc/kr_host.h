// Host-side helpers shared by the C ABI translation units: error state,
// handle layout, argument validation and the per-image scalar precompute.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "kerrray_hip.h"
#include "kr_device.h"
#include "kr_image.h"

struct kr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr;
    hipEvent_t ev1 = nullptr;
    void* buf = nullptr;
    size_t buf_bytes = 0;
    void* scratch = nullptr;  // device address of the planned trace's static scratch on this device
    void* ws = nullptr;       // workspace of the _dev row traces (grow-only; freed after a device sync)
    size_t ws_bytes = 0;
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

// Grow-only workspace of the _dev entry points, whose kernels run on the
// caller's streams: it is replaced only after the whole device is idle.
inline kr_status reserve_ws(kr_context* h, size_t bytes) {
    if (bytes <= h->ws_bytes) return KR_OK;
    if (h->ws) {
        KR_HIP(hipDeviceSynchronize());
        KR_HIP(hipFree(h->ws));
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    const size_t rounded = (bytes + (1u << 20) - 1) & ~static_cast<size_t>((1u << 20) - 1);
    KR_HIP(hipMalloc(&h->ws, rounded));
    h->ws_bytes = rounded;
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
    if (n > (int64_t(1) << 31)) return fail(KR_ERR_INVALID, "at most 2^31 photons per call");
    if (p->dtype != KR_FLOAT64 && p->dtype != KR_FLOAT32) return fail(KR_ERR_INVALID, "dtype must be 0 or 1");
    const double fixed[5] = {0.0, 1.0, 2.0, -0.5, 8.0};
    for (int i = 0; i < 5; ++i)
        if (!same_bits(p->consts[i], fixed[i]))
            return fail(KR_ERR_INVALID, "consts[0:5] must be (0, 1, 2, -0.5, 8) as in numba_backend.py:100-108");
    const int fixed_idx[8] = {kr::C_SAFETY, kr::C_FMIN, kr::C_FMAX, kr::C_EXP, kr::C_THLO, kr::C_THHI, kr::C_AXIS,
                              kr::C_HALFPI};
    const double fixed_val[8] = {kr::fixed::safety,   kr::fixed::factor_min, kr::fixed::factor_max,
                                 kr::fixed::exponent, kr::fixed::theta_lo,   kr::fixed::theta_hi,
                                 kr::fixed::axis_sin, kr::fixed::half_pi};
    for (int i = 0; i < 8; ++i)
        if (!same_bits(p->consts[fixed_idx[i]], fixed_val[i]))
            return fail(KR_ERR_UNSUPPORTED,
                        "consts[5:9], [14:17] and [20] must be the reference's controller and classification constants");
    if (!same_bits(p->params[kr::P_HALFPI], kr::fixed::half_pi))
        return fail(KR_ERR_UNSUPPORTED, "params[1] must be pi / 2");
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

// Every BatchResult array of an output block is set (and the disk event's
// two when the event is on).
inline kr_status check_outputs(const kr_outputs* o, bool disk) {
    if (!o) return fail(KR_ERR_INVALID, "outputs struct is NULL");
    if (!o->Y || !o->lam || !o->n_steps || !o->n_rejected || !o->state || !o->max_null_error ||
        !o->max_energy_drift || !o->max_lz_drift || !o->max_carter_drift)
        return fail(KR_ERR_INVALID, "an output array is NULL");
    if (disk && (!o->event_Y || !o->event_hit))
        return fail(KR_ERR_INVALID, "disk event enabled but event_Y/event_hit is NULL");
    return KR_OK;
}

// The error a device flag word stands for (KR_FLAG_* bits; 0 is KR_OK).
inline kr_status flag_status(int32_t flags) {
    static_assert(KR_FLAG_RADICAND == kr::FLAG_RADICAND && KR_FLAG_NOT_NULL == kr::FLAG_NOT_NULL &&
                      KR_FLAG_SKY == kr::FLAG_SKY && KR_FLAG_DISK == kr::FLAG_DISK && KR_FLAG_EDGE == kr::FLAG_EDGE &&
                      KR_FLAG_REPAIR == kr::FLAG_REPAIR && KR_CAMERA_ZAMO == kr::CAMERA_ZAMO,
                  "flag and camera values of the ABI and the kernels");
    if (flags & KR_FLAG_REPAIR)
        return fail(KR_ERR_REPAIR,
                    "the FP64 verify/repair of the mask did not reach a fixed point within its round limit");
    if (flags & KR_FLAG_RADICAND)
        return fail(KR_ERR_RADICAND, "a Carter radicand is negative beyond round-off at the observer");
    if (flags & KR_FLAG_SKY)
        return fail(KR_ERR_INVALID, "an image point has alpha^2 + beta^2 >= r_o^2: no direction on the observer's sky");
    if (flags & KR_FLAG_EDGE) return fail(KR_ERR_INVALID, "the centre lies outside the captured region");
    if (flags & KR_FLAG_DISK)
        return fail(KR_ERR_INVALID,
                    "a disk crossing cannot be shaded: not equatorial, no circular orbit at its radius, or g <= 0");
    if (flags & KR_FLAG_NOT_NULL)
        return fail(KR_ERR_NOT_NULL, "the ZAMO tetrad projection is not null to 1e-10, or no null p_t exists");
    return flags ? fail(KR_ERR_DEVICE, "unknown device flag") : KR_OK;
}

// Per-image scalars of the Bardeen initial conditions (kr_image.h).
inline kr_status bardeen_image(const kr_image* img, bool need_grid, kr::BardeenImage* B) {
    const char* err = kr::bardeen_image_scalars(img, need_grid, B);
    return err ? fail(KR_ERR_INVALID, err) : KR_OK;
}

// The trace of given rows (working dtype `dtype`) on the longest-first
// persistent kernels (kr_mixed.hip), enqueued on s with a workspace of
// rows_workspace_bytes(n).
size_t rows_workspace_bytes(int64_t n);
kr_status enqueue_rows_planned(kr_context* h, const kr::Params& P, int dtype, const double* d_Y0, int64_t n,
                               const kr_outputs& d_out, void* d_work, hipStream_t s);

// Finds the device address of the planned trace's static scratch array on
// the current device (once per device; kr_mixed.hip) for kr_create.
kr_status planned_scratch_setup(int device, void** address);

}  // namespace host
}  // namespace kr
