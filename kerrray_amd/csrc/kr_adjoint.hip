// Differentiation form: a tape of the FP64 adaptive trace of image-plane
// points and the discrete-adjoint sweep over it (kr_adjoint.h).
//
// kr_tape_create traces the points once (the plain FP64 kernel), keeps the
// escaped photons, and re-integrates those with a recording kernel that
// stores every accepted step size h_k and the state every `segment` accepted
// steps (checkpoints).  kr_tape_vjp runs one lane per recorded photon,
// backwards over its segments: from each checkpoint it recomputes the
// segment's states with the frozen steps (same FP64 step code, so the
// recomputed trajectory is the recorded one bit for bit), stores them in a
// per-lane segment buffer laid out [slot][component][photon] so that lanes
// touch consecutive addresses, and reverses the steps of the segment.
//
// HBM per recorded photon: 8 B per accepted step, 48 B per checkpoint,
// 48 * segment B of segment buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "kerrray_hip.h"
#include "kr_adjoint.h"
#include "kr_host.h"

namespace kr {
namespace {

struct TapeDev {
    const double* alpha;
    const double* beta;
    const int64_t* list;   // recorded photon -> point index
    const int64_t* h_off;  // (n_rec + 1) offsets into hs
    const int64_t* c_off;  // (n_rec + 1) offsets into ckpt (in states)
    double* hs;
    double* ckpt;  // 6 doubles per checkpoint
    int64_t n_rec;
    int segment;
};

constexpr int32_t TAPE_FLAG_REPLAY = 1 << 8;     // re-integration took a different number of steps
constexpr int32_t TAPE_FLAG_RECOMPUTE = 1 << 9;  // recomputed final state differs from the traced one

// Re-integrates recorded photon t exactly as trace_kernel does and stores its
// accepted step sizes and checkpoints.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void record_kernel(
    Params P, BardeenImage B, TapeDev T, double* theta_inf, double* phi_inf, int32_t* flags) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T.n_rec) return;
    const int64_t i = T.list[t];
    double y0[8];
    bardeen_state(B, T.alpha[i], T.beta[i], y0);
    RayState<double> R;
    ray_init<double, true>(P, y0, R);
    const int64_t hb = T.h_off[t], cap = T.h_off[t + 1] - hb;
    const int64_t cb = T.c_off[t], ccap = T.c_off[t + 1] - cb;
    const double lambda_max = P.params[P_LAMMAX];
    if (ccap > 0) {
#pragma unroll
        for (int c = 0; c < NA; ++c) T.ckpt[cb * NA + c] = R.y[c];
    }
    while (R.state == RUNNING) {
        const double remaining = lambda_max - R.lam;
        const double h_eff = (R.h >= remaining) ? remaining : R.h;
        const int64_t k = R.n_steps;
        ray_attempt<double, true, false>(P, R);
        if (R.n_steps != k) {
            if (k < cap) T.hs[hb + k] = h_eff;
            const int64_t n = R.n_steps;
            if (n % T.segment == 0) {
                const int64_t s = n / T.segment;
                if (s < ccap) {
#pragma unroll
                    for (int c = 0; c < NA; ++c) T.ckpt[(cb + s) * NA + c] = R.y[c];
                }
            }
        }
    }
    if (R.n_steps != cap || R.state != ESCAPED) atomicOr(flags, TAPE_FLAG_REPLAY);
    double th, ph;
    ad::escape_tail(R.y[1], R.y[2], R.y[3], R.y[5], static_cast<double>(R.p_t), static_cast<double>(R.p_ph),
                    P.geom[G_A], P.geom[G_MASS], th, ph);
    theta_inf[i] = th;
    phi_inf[i] = ph;
}

__global__ void fill_nan_kernel(double* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = __builtin_nan("");
}

// Reverse sweep for recorded photon t with cotangent (g_theta, g_phi) of its
// escape direction at infinity; writes d/d(alpha, beta, spin, theta_o) to
// grad[i].
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void adjoint_kernel(
    Params P, BardeenImage B, double mass, double spin, TapeDev T, const double* Yfinal, const double* g_theta,
    const double* g_phi, double* seg, double* grad, int32_t* flags) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T.n_rec) return;
    const int64_t i = T.list[t];
    const double alpha = T.alpha[i], beta = T.beta[i];
    double y0[8];
    bardeen_state(B, alpha, beta, y0);
    const double pt = y0[4], pph = y0[7];
    const double a = spin * mass;
    const int64_t hb = T.h_off[t], n = T.h_off[t + 1] - hb;
    const int64_t cb = T.c_off[t];
    const int S = T.segment;
    const int64_t nseg = (n + S - 1) / S;
    double ybar[NA], ptbar, pphbar, abar;
    {
        double yN[NA];
#pragma unroll
        for (int c = 0; c < NA; ++c) yN[c] = Yfinal[i * 8 + full_index(c)];
        ad::escape_tail_vjp(mass, a, yN, pt, pph, g_theta[i], g_phi[i], ybar, ptbar, pphbar, abar);
    }
    const int64_t stride = T.n_rec;
    for (int64_t s = nseg - 1; s >= 0; --s) {
        double y[NA];
#pragma unroll
        for (int c = 0; c < NA; ++c) y[c] = T.ckpt[(cb + s) * NA + c];
        const int64_t kb = s * S;
        const int64_t ke = (kb + S < n) ? kb + S : n;
        for (int64_t k = kb; k < ke; ++k) {
#pragma unroll
            for (int c = 0; c < NA; ++c) seg[((k - kb) * NA + c) * stride + t] = y[c];
            ad::step_forward(P, y, pt, pph, T.hs[hb + k]);
        }
        if (s == nseg - 1) {
            bool same = true;
#pragma unroll
            for (int c = 0; c < NA; ++c) same = same && (y[c] == Yfinal[i * 8 + full_index(c)]);
            if (!same) atomicOr(flags, TAPE_FLAG_RECOMPUTE);
        }
        for (int64_t k = ke - 1; k >= kb; --k) {
#pragma unroll
            for (int c = 0; c < NA; ++c) y[c] = seg[((k - kb) * NA + c) * stride + t];
            ad::step_adjoint(P, mass, a, y, pt, pph, T.hs[hb + k], ybar, ptbar, pphbar, abar);
        }
    }
    // Chain rule through the initial state y_0(alpha, beta, a, theta_o).
    ad::Dual<4> yd[8];
    ad::bardeen_state_of<4>(ad::seed<4>(alpha, 0), ad::seed<4>(beta, 1), ad::seed<4>(a, 2),
                            ad::seed<4>(B.theta_o, 3), mass, B.r_o, yd);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double g = ptbar * yd[4].d[j] + pphbar * yd[7].d[j];
#pragma unroll
        for (int c = 0; c < NA; ++c) g += ybar[c] * yd[full_index(c)].d[j];
        if (j == 2) g = (g + abar) * mass;  // d/d(spin) = M d/da
        grad[i * 4 + j] = g;
    }
}

// Escape direction at infinity of every escaped row of a result (NaN for
// the others).
__global__ void escape_direction_kernel(const double* Y, const int64_t* state, int64_t n, double a, double m,
                                        double* theta_inf, double* phi_inf) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double th = __builtin_nan(""), ph = __builtin_nan("");
    if (state[i] == ESCAPED) {
        const double* y = Y + i * 8;
        ad::escape_tail(y[1], y[2], y[3], y[6], y[4], y[7], a, m, th, ph);
    }
    theta_inf[i] = th;
    phi_inf[i] = ph;
}

__global__ void fill_f64_kernel(double* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void fill_flag_kernel(int32_t* p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0;
}

}  // namespace
}  // namespace kr

using namespace kr;

struct kr_tape_s {
    kr_context* h = nullptr;
    Params P{};
    BardeenImage B{};
    double mass = 0.0, spin = 0.0;
    int64_t n = 0, n_rec = 0, total_steps = 0, total_ckpt = 0;
    int segment = 0;
    double *alpha = nullptr, *beta = nullptr, *Y = nullptr, *hs = nullptr, *ckpt = nullptr, *seg = nullptr;
    int64_t *list = nullptr, *h_off = nullptr, *c_off = nullptr;
    int32_t* flags = nullptr;
    double record_ms = 0.0;

    TapeDev dev() const { return TapeDev{alpha, beta, list, h_off, c_off, hs, ckpt, n_rec, segment}; }
};

namespace {

void free_tape(kr_tape_s* t) {
    if (!t) return;
    double* d[] = {t->alpha, t->beta, t->Y, t->hs, t->ckpt, t->seg};
    for (double* p : d)
        if (p) (void)hipFree(p);
    int64_t* q[] = {t->list, t->h_off, t->c_off};
    for (int64_t* p : q)
        if (p) (void)hipFree(p);
    if (t->flags) (void)hipFree(t->flags);
    delete t;
}

template <typename T>
kr_status dev_alloc(T** p, int64_t count) {
    KR_HIP(hipMalloc(reinterpret_cast<void**>(p), static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
    return KR_OK;
}

}  // namespace

extern "C" kr_status kr_tape_create(kr_handle h, const kr_params* p, const kr_image* img, const double* d_alpha,
                                    const double* d_beta, int64_t n, int32_t segment, const kr_outputs* d_out,
                                    double* d_theta_inf, double* d_phi_inf, kr_tape* out) {
    host::clear_error();
    if (!out) return host::fail(KR_ERR_INVALID, "tape out-pointer is NULL");
    *out = nullptr;
    BardeenImage B;
    kr_status st = host::bardeen_image(img, false, &B);
    if (st != KR_OK) return st;
    if (B.camera != CAMERA_BARDEEN)
        return host::fail(KR_ERR_UNSUPPORTED, "the adjoint differentiates the Bardeen image-plane initial state only");
    Params P;
    if ((st = host::validate(h, p, n, &P)) != KR_OK) return st;
    if (p->dtype != KR_FLOAT64 || P.ipar[I_ADAPTIVE] != 1 || P.ipar[I_DISK] != 0)
        return host::fail(KR_ERR_UNSUPPORTED, "the adjoint is implemented for FP64 Dormand-Prince 5(4) without events");
    const double a = img->spin * img->mass;
    if (!(P.geom[G_MASS] == img->mass) || std::fabs(P.geom[G_A] - a) > 1e-15 * std::max(1.0, std::fabs(a)))
        return host::fail(KR_ERR_INVALID, "params geometry (mass, a) differs from the image's mass and spin");
    if (segment < 1 || segment > 4096) return host::fail(KR_ERR_INVALID, "segment must be in [1, 4096]");
    if (n > 0 && (!d_alpha || !d_beta)) return host::fail(KR_ERR_INVALID, "alpha/beta is NULL");
    if (n > 0 && (!d_theta_inf || !d_phi_inf)) return host::fail(KR_ERR_INVALID, "theta_inf/phi_inf is NULL");
    if (!d_out || !d_out->Y || !d_out->state || !d_out->n_steps || !d_out->lam || !d_out->n_rejected ||
        !d_out->max_null_error || !d_out->max_energy_drift || !d_out->max_lz_drift || !d_out->max_carter_drift)
        return host::fail(KR_ERR_INVALID, "an output array is NULL");
    KR_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;

    // 1. The plain FP64 trace of every point, into the caller's buffers.
    if ((st = kr_trace_points_dev(h, p, img, d_alpha, d_beta, n, d_out, nullptr, s)) != KR_OK) return st;
    std::vector<int64_t> state(static_cast<size_t>(n)), steps(static_cast<size_t>(n));
    if (n > 0) {
        KR_HIP(hipMemcpyAsync(state.data(), d_out->state, n * 8, hipMemcpyDeviceToHost, s));
        KR_HIP(hipMemcpyAsync(steps.data(), d_out->n_steps, n * 8, hipMemcpyDeviceToHost, s));
    }
    KR_HIP(hipStreamSynchronize(s));

    // 2. Escaped photons and their offsets.
    std::vector<int64_t> list, h_off(1, 0), c_off(1, 0);
    for (int64_t i = 0; i < n; ++i) {
        if (state[i] != ESCAPED) continue;
        list.push_back(i);
        h_off.push_back(h_off.back() + steps[i]);
        c_off.push_back(c_off.back() + (steps[i] + segment - 1) / segment);
    }
    kr_tape_s* T = new kr_tape_s();
    T->h = h;
    T->P = P;
    T->B = B;
    T->mass = img->mass;
    T->spin = img->spin;
    T->n = n;
    T->n_rec = static_cast<int64_t>(list.size());
    T->total_steps = h_off.back();
    T->total_ckpt = c_off.back();
    T->segment = segment;
    auto bail = [&](kr_status e) {
        free_tape(T);
        return e;
    };
    if ((st = dev_alloc(&T->alpha, n)) != KR_OK || (st = dev_alloc(&T->beta, n)) != KR_OK ||
        (st = dev_alloc(&T->Y, n * 8)) != KR_OK || (st = dev_alloc(&T->list, T->n_rec)) != KR_OK ||
        (st = dev_alloc(&T->h_off, T->n_rec + 1)) != KR_OK || (st = dev_alloc(&T->c_off, T->n_rec + 1)) != KR_OK ||
        (st = dev_alloc(&T->hs, T->total_steps)) != KR_OK || (st = dev_alloc(&T->ckpt, T->total_ckpt * NA)) != KR_OK ||
        (st = dev_alloc(&T->seg, T->n_rec * static_cast<int64_t>(segment) * NA)) != KR_OK ||
        (st = dev_alloc(&T->flags, 1)) != KR_OK)
        return bail(st);
    auto hipok = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        host::fail(KR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return false;
    };
    if (n > 0) {
        if (!hipok(hipMemcpyAsync(T->alpha, d_alpha, n * 8, hipMemcpyDeviceToDevice, s), "copy alpha") ||
            !hipok(hipMemcpyAsync(T->beta, d_beta, n * 8, hipMemcpyDeviceToDevice, s), "copy beta") ||
            !hipok(hipMemcpyAsync(T->Y, d_out->Y, n * 64, hipMemcpyDeviceToDevice, s), "copy Y"))
            return bail(KR_ERR_DEVICE);
    }
    if (T->n_rec > 0) {
        if (!hipok(hipMemcpyAsync(T->list, list.data(), T->n_rec * 8, hipMemcpyHostToDevice, s), "copy list") ||
            !hipok(hipMemcpyAsync(T->h_off, h_off.data(), (T->n_rec + 1) * 8, hipMemcpyHostToDevice, s), "copy h_off") ||
            !hipok(hipMemcpyAsync(T->c_off, c_off.data(), (T->n_rec + 1) * 8, hipMemcpyHostToDevice, s), "copy c_off"))
            return bail(KR_ERR_DEVICE);
    }
    hipLaunchKernelGGL(fill_flag_kernel, dim3(1), dim3(64), 0, s, T->flags);
    if (!hipok(hipGetLastError(), "fill flags")) return bail(KR_ERR_DEVICE);
    if (n > 0) {
        const dim3 g(static_cast<unsigned>((n + 255) / 256));
        hipLaunchKernelGGL(fill_nan_kernel, g, dim3(256), 0, s, d_theta_inf, n);
        hipLaunchKernelGGL(fill_nan_kernel, g, dim3(256), 0, s, d_phi_inf, n);
        if (!hipok(hipGetLastError(), "fill directions")) return bail(KR_ERR_DEVICE);
    }

    // 3. Record the escaped photons.
    if (!hipok(hipEventRecord(h->ev0, s), "event")) return bail(KR_ERR_DEVICE);
    if (T->n_rec > 0) {
        hipLaunchKernelGGL(record_kernel, dim3(static_cast<unsigned>((T->n_rec + 255) / 256)), dim3(256), 0, s, P, B,
                           T->dev(), d_theta_inf, d_phi_inf, T->flags);
        if (!hipok(hipGetLastError(), "record_kernel")) return bail(KR_ERR_DEVICE);
    }
    if (!hipok(hipEventRecord(h->ev1, s), "event") || !hipok(hipStreamSynchronize(s), "record"))
        return bail(KR_ERR_DEVICE);
    float ms = 0.0f;
    if (!hipok(hipEventElapsedTime(&ms, h->ev0, h->ev1), "elapsed")) return bail(KR_ERR_DEVICE);
    T->record_ms = ms;
    int32_t flags = 0;
    if (!hipok(hipMemcpy(&flags, T->flags, 4, hipMemcpyDeviceToHost), "read flags")) return bail(KR_ERR_DEVICE);
    if (flags & TAPE_FLAG_REPLAY)
        return bail(host::fail(KR_ERR_DEVICE, "re-integration of a recorded photon differs from its trace"));
    *out = T;
    return KR_OK;
}

extern "C" int64_t kr_tape_recorded(kr_tape t) { return t ? t->n_rec : -1; }

extern "C" int64_t kr_tape_steps(kr_tape t) { return t ? t->total_steps : -1; }

extern "C" double kr_tape_record_ms(kr_tape t) { return t ? t->record_ms : -1.0; }

extern "C" kr_status kr_tape_vjp(kr_tape t, const double* d_g_theta, const double* d_g_phi, double* d_grad,
                                 void* stream) {
    host::clear_error();
    if (!t) return host::fail(KR_ERR_INVALID, "tape is NULL");
    if (t->n > 0 && (!d_g_theta || !d_g_phi || !d_grad))
        return host::fail(KR_ERR_INVALID, "a cotangent or gradient array is NULL");
    KR_HIP(hipSetDevice(t->h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL is the default stream, as for the other _dev calls
    if (t->n > 0) {
        hipLaunchKernelGGL(fill_f64_kernel, dim3(static_cast<unsigned>((t->n * 4 + 255) / 256)), dim3(256), 0, s,
                           d_grad, t->n * 4);
        KR_HIP(hipGetLastError());
    }
    if (t->n_rec > 0) {
        hipLaunchKernelGGL(adjoint_kernel, dim3(static_cast<unsigned>((t->n_rec + 63) / 64)), dim3(64), 0, s, t->P,
                           t->B, t->mass, t->spin, t->dev(), t->Y, d_g_theta, d_g_phi, t->seg, d_grad, t->flags);
        KR_HIP(hipGetLastError());
    }
    return KR_OK;
}

extern "C" kr_status kr_tape_flags(kr_tape t, int32_t* flags) {
    host::clear_error();
    if (!t || !flags) return host::fail(KR_ERR_INVALID, "tape or flags is NULL");
    KR_HIP(hipSetDevice(t->h->device));
    KR_HIP(hipDeviceSynchronize());
    KR_HIP(hipMemcpy(flags, t->flags, 4, hipMemcpyDeviceToHost));
    return KR_OK;
}

extern "C" kr_status kr_tape_destroy(kr_tape t) {
    host::clear_error();
    if (!t) return KR_OK;
    (void)hipSetDevice(t->h->device);
    (void)hipStreamSynchronize(t->h->stream);
    free_tape(t);
    return KR_OK;
}

extern "C" kr_status kr_escape_directions_dev(kr_handle h, const kr_params* p, const double* d_Y,
                                              const int64_t* d_state, int64_t n, double* d_theta_inf,
                                              double* d_phi_inf, void* stream) {
    host::clear_error();
    Params P;
    kr_status st = host::validate(h, p, n, &P);
    if (st != KR_OK) return st;
    if (p->dtype != KR_FLOAT64) return host::fail(KR_ERR_UNSUPPORTED, "escape directions are computed from FP64 rows");
    if (n > 0 && (!d_Y || !d_state || !d_theta_inf || !d_phi_inf)) return host::fail(KR_ERR_INVALID, "an array is NULL");
    KR_HIP(hipSetDevice(h->device));
    if (n > 0) {
        hipLaunchKernelGGL(escape_direction_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), d_Y, d_state, n, P.geom[G_A], P.geom[G_MASS],
                           d_theta_inf, d_phi_inf);
        KR_HIP(hipGetLastError());
    }
    return KR_OK;
}
