// Plain per-photon kernels (one lane = one photon) and the C ABI around
// them: batches of given initial states, Bardeen images generated on the
// device, arbitrary image points, and the device-side edge bisection.
//
// Reference: src/kerrray/raytracing/backends/numba_kernel.py:466-474 (one ray
// per prange iteration) behind numba_backend.py:131-256.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kerrray_hip.h"
#include "kr_device.h"
#include "kr_host.h"

namespace kr {

template <typename T>
struct Outputs {
    T* Y;
    double* lam;
    int64_t* n_steps;
    int64_t* n_rej;
    int64_t* state;
    double* m_null;
    double* m_e;
    double* m_lz;
    double* m_q;
    T* ev_Y;
    uint8_t* ev_hit;
};

enum { IC_Y0 = 0, IC_IMAGE = 1, IC_POINTS = 2 };

struct ICArgs {
    const double* Y0;     // IC_Y0: (n, 8)
    BardeenImage B;       // IC_IMAGE / IC_POINTS
    const double* alpha;  // IC_POINTS
    const double* beta;
    int32_t* flags;       // may be null
};

template <typename T, bool DISK>
__device__ __forceinline__ void write_result(const Outputs<T>& o, int64_t idx, const RayResult<T>& r) {
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
    if (DISK) {
#pragma unroll
        for (int c = 0; c < 8; ++c) o.ev_Y[idx * 8 + c] = r.ev[c];
        o.ev_hit[idx] = r.ev_hit ? 1 : 0;
    }
}

// Image lanes: each 64-lane wavefront takes an 8 x 8 pixel tile so that a
// wave's photons are neighbours on the sky and have similar step counts.
__device__ __forceinline__ bool image_pixel(int64_t n, int64_t& i, int64_t& j) {
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t tiles_x = (n + 7) >> 3;
    const int64_t ty = wave / tiles_x, tx = wave - ty * tiles_x;
    i = ty * 8 + (lane >> 3);
    j = tx * 8 + (lane & 7);
    return i < n && j < n;
}

// Two wavefronts per SIMD: the register budget (256 VGPRs) the per-photon
// state and the seven Dormand-Prince stages fit in without spilling.
template <typename T, bool ADAPTIVE, bool DISK, int IC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void trace_kernel(Params P, ICArgs ic, Outputs<T> out, int64_t n) {
    int64_t idx;
    double y0[8];
    if (IC == IC_IMAGE) {
        int64_t i, j;
        if (!image_pixel(ic.B.n, i, j)) return;
        idx = i * ic.B.n + j;
        const double alpha = pixel_centre(ic.B.fov, j, ic.B.n);
        const double beta = -pixel_centre(ic.B.fov, i, ic.B.n);
        const int32_t bad = image_state(ic.B, alpha, beta, y0);
        if (bad && ic.flags) atomicOr(ic.flags, bad);
    } else {
        idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (idx >= n) return;
        if (IC == IC_Y0) {
#pragma unroll
            for (int c = 0; c < 8; ++c) y0[c] = ic.Y0[idx * 8 + c];
        } else {
            const int32_t bad = image_state(ic.B, ic.alpha[idx], ic.beta[idx], y0);
            if (bad && ic.flags) atomicOr(ic.flags, bad);
        }
    }
    RayResult<T> r;
    integrate_ray<T, ADAPTIVE, DISK>(P, y0, r);
    write_result<T, DISK>(out, idx, r);
}

// Given rows or points, FP64 DP5, no disk, at most PAIR_ROWS_MAX photons:
// such a launch fills few SIMDs and lasts as long as its longest photon, so
// each photon is carried by two lanes (lanes 2k and 2k + 1 hold the same
// state; TrigPair shares out the trigonometry of every right-hand side) and
// each wavefront has a SIMD of its own.  The rows are trace_kernel's bit for
// bit; the even lane writes them.
constexpr int64_t PAIR_ROWS_MAX = 16384;

template <int IC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void trace_pair_kernel(
    Params P, ICArgs ic, Outputs<double> out, int64_t n) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t idx = t >> 1;
    const bool odd = (t & 1) != 0;
    if (idx >= n) return;  // both lanes of a pair leave together
    double y0[8];
    if (IC == IC_Y0) {
#pragma unroll
        for (int c = 0; c < 8; ++c) y0[c] = ic.Y0[idx * 8 + c];
    } else {
        const int32_t bad = image_state(ic.B, ic.alpha[idx], ic.beta[idx], y0);
        if (bad && ic.flags && !odd) atomicOr(ic.flags, bad);
    }
    const TrigPair trig{odd};
    RayState<double> R;
    ray_init<double, true>(P, y0, R, trig);
    while (R.state == RUNNING) ray_attempt_with<double, true, false>(P, R, ExactStepper<double, true, TrigPair>{trig});
    RayResult<double> r;
    ray_result<double, false>(R, r);
    if (!odd) write_result<double, false>(out, idx, r);
}

// One bisection round, probe kernel: lane (k, j) traces probe j of direction k.
__device__ __forceinline__ double probe_radius(double lo, double hi, int j, int n_probe) {
    if (n_probe == 1) return 0.5 * (lo + hi);
    return lo + (hi - lo) * (static_cast<double>(j + 1) / static_cast<double>(n_probe + 1));
}

template <typename T, bool ADAPTIVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void bisect_probe_kernel(Params P, BardeenImage B, double ca, double cb,
                                                          const double* cosk, const double* sink, const double* lo,
                                                          const double* hi, int64_t n_angles, int n_probe,
                                                          uint8_t* escaped, unsigned long long* failures,
                                                          int32_t* flags) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_angles * n_probe) return;
    const int64_t k = t / n_probe;
    const int j = static_cast<int>(t - k * n_probe);
    const double mid = probe_radius(lo[k], hi[k], j, n_probe);
    const double alpha = ca + mid * cosk[k];
    const double beta = cb + mid * sink[k];
    double y0[8];
    if (const int32_t bad = image_state(B, alpha, beta, y0)) atomicOr(flags, bad);
    RayResult<T> r;
    integrate_ray<T, ADAPTIVE, false>(P, y0, r);
    escaped[t] = r.state == ESCAPED ? 1 : 0;
    if (r.state != ESCAPED && r.state != CAPTURED) atomicAdd(failures, 1ull);
}

// The bracket update of one direction from its round's probe results
// escaped[0, n_probe) (study_edge_bisection.py:66-67 for n_probe = 1): the
// sub-bracket around the first escaping probe.  Returns true when the probes
// were not ordered captured -> escaped.
__device__ __forceinline__ bool bisect_update(double& lo, double& hi, const uint8_t* escaped, int n_probe) {
    const double l = lo, h = hi;
    int first = -1;
    bool bad = false;
    for (int j = 0; j < n_probe; ++j) {
        const bool e = escaped[j] != 0;
        if (e && first < 0) first = j;
        if (!e && first >= 0) bad = true;
    }
    if (first < 0) {
        lo = probe_radius(l, h, n_probe - 1, n_probe);
    } else {
        hi = probe_radius(l, h, first, n_probe);
        if (first > 0) lo = probe_radius(l, h, first - 1, n_probe);
    }
    return bad;
}

// Bracket update kernel, one thread per direction.
__global__ void bisect_update_kernel(double* lo, double* hi, const uint8_t* escaped, int64_t n_angles, int n_probe,
                                     unsigned long long* nonmonotone) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n_angles) return;
    double l = lo[k], h = hi[k];
    if (bisect_update(l, h, escaped + k * n_probe, n_probe)) atomicAdd(nonmonotone, 1ull);
    lo[k] = l;
    hi[k] = h;
}

// Every round of the bisection in one launch (FP64 DP5).  A workgroup of
// BLOCK lanes carries `dirs` directions (n_probe lanes each, dirs * n_probe
// <= BLOCK) through all `iterations` rounds, its brackets in LDS between
// rounds, so a direction waits for its own slowest probe only, never for
// another workgroup's.  One wavefront per SIMD: the launch is used only when
// all its workgroups fit on the chip at once.  The probe radii and the update
// are those of the per-round kernels above: the brackets are the same bit for
// bit.
template <typename T, bool ADAPTIVE, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(1, 1))) void bisect_rounds_kernel(
    Params P, BardeenImage B, double ca, double cb, const double* cosk, const double* sink, double* lo, double* hi,
    int64_t n_angles, int n_probe, int dirs, int iterations, unsigned long long* counts, int32_t* flags) {
    // The directions' brackets and unit vectors stay in LDS, so that the
    // lanes carry nothing but their photon across the rounds.
    __shared__ uint8_t esc[BLOCK];
    __shared__ double lo_s[64], hi_s[64], c_s[64], s_s[64];  // dirs <= 64
    const int d = static_cast<int>(threadIdx.x) / n_probe;  // direction within the workgroup
    const int j = static_cast<int>(threadIdx.x) - d * n_probe;
    const int64_t k = static_cast<int64_t>(blockIdx.x) * dirs + d;
    const bool active = d < dirs && k < n_angles;
    if (active && j == 0) {
        lo_s[d] = lo[k];
        hi_s[d] = hi[k];
        c_s[d] = cosk[k];
        s_s[d] = sink[k];
    }
    __syncthreads();
    for (int it = 0; it < iterations; ++it) {
        if (active) {
            const double mid = probe_radius(lo_s[d], hi_s[d], j, n_probe);
            const double alpha = ca + mid * c_s[d];
            const double beta = cb + mid * s_s[d];
            double y0[8];
            if (const int32_t bad = image_state(B, alpha, beta, y0)) atomicOr(flags, bad);
            RayResult<T> r;
            integrate_ray<T, ADAPTIVE, false>(P, y0, r);
            esc[threadIdx.x] = r.state == ESCAPED ? 1 : 0;
            if (r.state != ESCAPED && r.state != CAPTURED) atomicAdd(&counts[0], 1ull);
        }
        __syncthreads();
        if (active && j == 0) {
            double l = lo_s[d], h = hi_s[d];
            if (bisect_update(l, h, esc + d * n_probe, n_probe)) atomicAdd(&counts[1], 1ull);
            lo_s[d] = l;
            hi_s[d] = h;
        }
        __syncthreads();
    }
    if (active && j == 0) {
        lo[k] = lo_s[d];
        hi[k] = hi_s[d];
    }
}

// Critical-impact search (kr_impact_search).  Searches live on the device;
// a round is one probe launch (lane = (search, probe)) and one update launch
// (one workgroup, thread = search).  After an error every later launch
// returns at once, leaving the failing round's probes and open flags intact.
struct ImpactSearchDev {
    double lo, hi;
    double lz_sign;
    int64_t iterations;
    int32_t open;
    int32_t pad;
};

struct ImpactCounters {
    unsigned long long rays, max_steps, crossings;
    int32_t error, error_round;
    int32_t round;  // rounds enqueued so far (0: the endpoint check)
    int32_t pad;
};

struct ImpactCfg {
    ImpactLaunch L;
    double tol, r_capture;
    int64_t max_iterations;
    int32_t n_search, n_probe, relabel;
};

template <typename T, bool ADAPTIVE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void impact_probe_kernel(
    Params P, ImpactCfg C, ImpactSearchDev* s, ImpactCounters* c, double* probe_b, int64_t* probe_state,
    int endpoints, int32_t* flags) {
    if (c->error) return;
    const int per = endpoints ? 2 : C.n_probe;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(C.n_search) * per) return;
    const int k = static_cast<int>(t / per), j = static_cast<int>(t - static_cast<int64_t>(k) * per);
    const ImpactSearchDev S = s[k];
    if (!endpoints && !S.open) return;
    // kerr_bisection.py:233: lo + width * (j + 1) / (n_probe + 1)
    const double b = endpoints ? (j == 0 ? S.lo : S.hi)
                               : S.lo + (S.hi - S.lo) * (static_cast<double>(j + 1) / static_cast<double>(per + 1));
    double y0[8];
    if (!impact_state(C.L, S.lz_sign * b, y0)) atomicOr(flags, FLAG_RADICAND);
    RayResult<T> r;
    integrate_ray<T, ADAPTIVE, false>(P, y0, r);
    int64_t code = r.state;
    const double r_end = static_cast<double>(r.y[1]);
    if (C.relabel && code == OUT_OF_DOMAIN && isfinite(r_end) && r_end <= C.r_capture) {  // kerr_bisection.py:121-126
        code = CAPTURED;
        atomicAdd(&c->crossings, 1ull);
    }
    probe_b[t] = b;
    probe_state[t] = code;
    atomicAdd(&c->rays, 1ull);
    atomicMax(&c->max_steps, static_cast<unsigned long long>(r.n_steps));
}

// One workgroup; thread k owns search k.  Checks the round, then narrows the
// brackets (kerr_bisection.py:143-156,240-256) only if no search failed.
__global__ __launch_bounds__(1024) void impact_update_kernel(ImpactCfg C, ImpactSearchDev* s, ImpactCounters* c,
                                                             const double* probe_b, const int64_t* probe_state,
                                                             int endpoints) {
    __shared__ int bad_state, bad_order;
    if (c->error) return;
    const int k = threadIdx.x;
    if (k == 0) bad_state = bad_order = 0;
    __syncthreads();
    const int per = endpoints ? 2 : C.n_probe;
    ImpactSearchDev S = k < C.n_search ? s[k] : ImpactSearchDev{};
    const bool active = k < C.n_search && (endpoints || S.open);
    double cap_max = 0.0, esc_min = 0.0;
    bool any_cap = false, any_esc = false;
    if (active) {
        for (int j = 0; j < per; ++j) {
            const int64_t code = probe_state[static_cast<int64_t>(k) * per + j];
            const double b = probe_b[static_cast<int64_t>(k) * per + j];
            if (endpoints) {
                if (code != (j == 0 ? CAPTURED : ESCAPED)) atomicOr(&bad_state, 1);
                continue;
            }
            if (code == CAPTURED) {
                cap_max = any_cap ? pymax(cap_max, b) : b;
                any_cap = true;
            } else if (code == ESCAPED) {
                esc_min = any_esc ? pymin(esc_min, b) : b;
                any_esc = true;
            } else {
                atomicOr(&bad_state, 1);
            }
        }
        if (any_cap && any_esc && cap_max > esc_min) atomicOr(&bad_order, 1);
    }
    __syncthreads();
    if (bad_state || bad_order) {
        if (k == 0) {
            c->error = endpoints ? 1 : (bad_state ? 2 : 3);
            c->error_round = c->round;
        }
        return;
    }
    if (active && !endpoints) {
        if (any_cap) S.lo = pymax(S.lo, cap_max);
        if (any_esc) S.hi = pymin(S.hi, esc_min);
        S.iterations += 1;
    }
    if (k < C.n_search) {
        S.open = (S.hi - S.lo > C.tol) && (S.iterations < C.max_iterations);
        s[k] = S;
    }
    if (k == 0) c->round += 1;
}

// Initial states of every pixel of an image, row-major (n*n, 8): one lane per
// pixel, each lane writing its own 64-byte row.
__global__ __launch_bounds__(256) void image_states_kernel(BardeenImage B, double* Y0, int32_t* flags) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= B.n * B.n) return;
    const int64_t i = idx / B.n, j = idx - i * B.n;
    double y[8];
    const int32_t bad = image_state(B, pixel_centre(B.fov, j, B.n), -pixel_centre(B.fov, i, B.n), y);
    if (bad && flags) atomicOr(flags, bad);
#pragma unroll
    for (int c = 0; c < 8; ++c) Y0[idx * 8 + c] = y[c];
}

// Initial states of given image points, one lane per point.
__global__ __launch_bounds__(256) void point_states_kernel(BardeenImage B, const double* alpha, const double* beta,
                                                           int64_t n, double* Y0, int32_t* flags) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    double y[8];
    const int32_t bad = image_state(B, alpha[idx], beta[idx], y);
    if (bad && flags) atomicOr(flags, bad);
#pragma unroll
    for (int c = 0; c < 8; ++c) Y0[idx * 8 + c] = y[c];
}

// Shading of disk crossings, one lane per row: g, r_hit and I_obs of every
// DISK_HIT row; 0 intensity and NaN elsewhere (renderer.py:190-205).
template <typename T>
__global__ __launch_bounds__(256) void disk_shade_kernel(DiskShade D, const T* ev_Y, const uint8_t* ev_hit, int64_t n,
                                                         double* intensity, double* g, double* r_hit, int32_t* flags) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const double nan = __builtin_nan("");
    double I = 0.0, gg = nan, rr = nan;
    if (ev_hit[idx]) {
        double y[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) y[c] = static_cast<double>(ev_Y[idx * 8 + c]);
        if (const int32_t bad = disk_shade(D, y, I, gg, rr)) {
            if (flags) atomicOr(flags, bad);
            I = 0.0;
            gg = nan;
        }
    }
    intensity[idx] = I;
    g[idx] = gg;
    r_hit[idx] = rr;
}

// Mask edge radii, one wavefront per direction: the lanes evaluate 64
// consecutive samples at a time and a ballot finds the first below 0.5, so a
// direction's march takes n_samples / 64 steps instead of n_samples.  Each
// sample's value and the crossing are mask_edge_radius's expressions, so the
// radii are the same bit for bit.  The mask (at most a few MB) is read
// through the caches.
__global__ __launch_bounds__(256) void mask_edge_kernel(const uint8_t* mask, int64_t rows, int64_t cols, double a0,
                                                        double b0, double d, double ca, double cb, const double* cosk,
                                                        const double* sink, int64_t n_angles, int64_t n_samples,
                                                        double* radii, int32_t* flags) {
    const int64_t k = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (k >= n_angles) return;  // whole wavefronts leave together
    const double ds = 0.125 * d;
    const double cos_k = cosk[k], sin_k = sink[k];
    double f_carry = 0.0;  // the value of the last sample of the previous step
    double r = __builtin_nan("");
    bool found = false;
    for (int64_t m0 = 0; m0 < n_samples; m0 += 64) {
        const int64_t m = m0 + lane;
        double f = 1.0;
        if (m < n_samples) {
            const double sm = ds * static_cast<double>(m);
            const double x = ca + cos_k * sm;
            const double y = cb + sin_k * sm;
            f = mask_bilinear(mask, rows, cols, (x - a0) / d, (b0 - y) / d);
        }
        const uint64_t below = __ballot(m < n_samples && f < 0.5);
        const double f_before = __shfl_up(f, 1);  // lane L gets lane L - 1's value
        if (below != 0ull) {
            const int first = __ffsll(static_cast<unsigned long long>(below)) - 1;
            if (lane == first) {
                if (m > 0) {
                    const double f_prev = lane == 0 ? f_carry : f_before;
                    const double s_prev = ds * static_cast<double>(m - 1);
                    r = s_prev + ds * (f_prev - 0.5) / (f_prev - f);
                    found = true;
                }
                radii[k] = r;
                if (!found && flags) atomicOr(flags, FLAG_EDGE);
            }
            return;
        }
        f_carry = __shfl(f, 63);
    }
    if (lane == 0) {
        radii[k] = r;
        if (flags) atomicOr(flags, FLAG_EDGE);
    }
}

__global__ void fill_u64_kernel(unsigned long long* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- launches

template <typename T>
Outputs<T> typed(const kr_outputs& o) {
    return Outputs<T>{static_cast<T*>(o.Y), o.lam, o.n_steps, o.n_rejected, o.state, o.max_null_error,
                      o.max_energy_drift, o.max_lz_drift, o.max_carter_drift, static_cast<T*>(o.event_Y),
                      o.event_hit};
}

template <typename T, bool ADAPTIVE, bool DISK, int IC>
hipError_t launch_trace_t(const Params& P, const ICArgs& ic, const kr_outputs& o, int64_t n, hipStream_t s) {
    int64_t threads;
    if (IC == IC_IMAGE) {
        const int64_t tiles = ((ic.B.n + 7) >> 3) * ((ic.B.n + 7) >> 3);
        threads = tiles * 64;
    } else {
        threads = n;
    }
    if (threads == 0) return hipSuccess;
    if constexpr (IC != IC_IMAGE && sizeof(T) == sizeof(double) && ADAPTIVE && !DISK) {
        if (n <= PAIR_ROWS_MAX) {
            hipLaunchKernelGGL((trace_pair_kernel<IC>), dim3(static_cast<unsigned>((2 * n + 255) / 256)), dim3(256),
                               0, s, P, ic, typed<T>(o), n);
            return hipGetLastError();
        }
    }
    const int64_t blocks = (threads + 255) / 256;
    hipLaunchKernelGGL((trace_kernel<T, ADAPTIVE, DISK, IC>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       P, ic, typed<T>(o), n);
    return hipGetLastError();
}

template <int IC>
hipError_t launch_trace(const Params& P, int dtype, const ICArgs& ic, const kr_outputs& o, int64_t n, hipStream_t s) {
    const bool adaptive = P.ipar[I_ADAPTIVE] != 0;
    const bool disk = P.ipar[I_DISK] != 0;
    if (dtype == KR_FLOAT64) {
        if (adaptive)
            return disk ? launch_trace_t<double, true, true, IC>(P, ic, o, n, s)
                        : launch_trace_t<double, true, false, IC>(P, ic, o, n, s);
        return disk ? launch_trace_t<double, false, true, IC>(P, ic, o, n, s)
                    : launch_trace_t<double, false, false, IC>(P, ic, o, n, s);
    }
    if (adaptive)
        return disk ? launch_trace_t<float, true, true, IC>(P, ic, o, n, s)
                    : launch_trace_t<float, true, false, IC>(P, ic, o, n, s);
    return disk ? launch_trace_t<float, false, true, IC>(P, ic, o, n, s)
                : launch_trace_t<float, false, false, IC>(P, ic, o, n, s);
}

hipError_t fill_i32(int32_t* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t fill_u64(unsigned long long* p, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace kr

// ------------------------------------------------------------------ C ABI

using namespace kr;

extern "C" int kr_abi_version(void) { return KR_ABI_VERSION; }

extern "C" int kr_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return -1;
    return n;
}

extern "C" const char* kr_last_error(void) { return host::last_error().c_str(); }

extern "C" kr_status kr_create(int device, kr_handle* out) {
    host::clear_error();
    if (!out) return host::fail(KR_ERR_INVALID, "kr_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return host::fail(KR_ERR_DEVICE, std::string("no HIP device available: ") + hipGetErrorString(e));
    if (device < 0 || device >= count) return host::fail(KR_ERR_INVALID, "kr_create: device index out of range");
    kr_context* h = new (std::nothrow) kr_context();
    if (!h) return host::fail(KR_ERR_DEVICE, "kr_create: out of host memory");
    h->device = device;
    KR_HIP(hipSetDevice(device));
    KR_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    KR_HIP(hipEventCreate(&h->ev0));
    KR_HIP(hipEventCreate(&h->ev1));
    if (kr_status st = host::planned_scratch_setup(device, &h->scratch); st != KR_OK) {
        hipEventDestroy(h->ev1);
        hipEventDestroy(h->ev0);
        hipStreamDestroy(h->stream);
        delete h;
        return st;
    }
    *out = h;
    return KR_OK;
}

extern "C" kr_status kr_destroy(kr_handle h) {
    host::clear_error();
    if (!h) return KR_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->buf) hipFree(h->buf);
    if (h->ws) {
        hipDeviceSynchronize();
        hipFree(h->ws);
    }
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return KR_OK;
}

namespace {

// Device layout of one batch in the handle's buffer.
struct BatchBuffers {
    double* Y0;
    kr_outputs d;
    int32_t* flags;
    char* extra;
};

// `extra` more bytes (for the caller's own arrays) start at b.extra.
kr_status carve(kr_handle h, int64_t n, int dtype, bool need_y0, bool disk, BatchBuffers& b, size_t extra = 0) {
    const size_t ts = dtype == KR_FLOAT64 ? 8 : 4;
    host::Carver c;
    const size_t oY0 = need_y0 ? c.take(n * 64) : 0;
    const size_t oY = c.take(n * 8 * ts);
    const size_t olam = c.take(n * 8), osteps = c.take(n * 8), orej = c.take(n * 8), ostate = c.take(n * 8);
    const size_t om0 = c.take(n * 8), om1 = c.take(n * 8), om2 = c.take(n * 8), om3 = c.take(n * 8);
    const size_t oev = disk ? c.take(n * 8 * ts) : 0;
    const size_t ohit = disk ? c.take(n) : 0;
    const size_t oflags = c.take(256);
    const size_t oextra = c.take(static_cast<int64_t>(extra));
    kr_status st = host::reserve(h, c.total);
    if (st != KR_OK) return st;
    char* base = static_cast<char*>(h->buf);
    b.Y0 = need_y0 ? reinterpret_cast<double*>(base + oY0) : nullptr;
    b.d.Y = base + oY;
    b.d.lam = reinterpret_cast<double*>(base + olam);
    b.d.n_steps = reinterpret_cast<int64_t*>(base + osteps);
    b.d.n_rejected = reinterpret_cast<int64_t*>(base + orej);
    b.d.state = reinterpret_cast<int64_t*>(base + ostate);
    b.d.max_null_error = reinterpret_cast<double*>(base + om0);
    b.d.max_energy_drift = reinterpret_cast<double*>(base + om1);
    b.d.max_lz_drift = reinterpret_cast<double*>(base + om2);
    b.d.max_carter_drift = reinterpret_cast<double*>(base + om3);
    b.d.event_Y = disk ? base + oev : nullptr;
    b.d.event_hit = disk ? reinterpret_cast<uint8_t*>(base + ohit) : nullptr;
    b.flags = reinterpret_cast<int32_t*>(base + oflags);
    b.extra = base + oextra;
    return KR_OK;
}

kr_status copy_back(kr_handle h, const BatchBuffers& b, const kr_outputs& out, int64_t n, int dtype, bool disk) {
    const size_t ts = dtype == KR_FLOAT64 ? 8 : 4;
    hipStream_t s = h->stream;
    KR_HIP(hipMemcpyAsync(out.Y, b.d.Y, n * 8 * ts, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.lam, b.d.lam, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.n_steps, b.d.n_steps, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.n_rejected, b.d.n_rejected, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.state, b.d.state, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.max_null_error, b.d.max_null_error, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.max_energy_drift, b.d.max_energy_drift, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.max_lz_drift, b.d.max_lz_drift, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out.max_carter_drift, b.d.max_carter_drift, n * 8, hipMemcpyDeviceToHost, s));
    if (disk) {
        if (out.event_Y) KR_HIP(hipMemcpyAsync(out.event_Y, b.d.event_Y, n * 8 * ts, hipMemcpyDeviceToHost, s));
        if (out.event_hit) KR_HIP(hipMemcpyAsync(out.event_hit, b.d.event_hit, n, hipMemcpyDeviceToHost, s));
    }
    KR_HIP(hipStreamSynchronize(s));
    return KR_OK;
}

}  // namespace

namespace {

// Dormand-Prince batches of at least this many rows run on the longest-first
// persistent kernels (kr_mixed.hip): FP64 from 2^15 rows, 2^17 with the disk
// event (whose photons end at the disk plane, which leaves less imbalance to
// remove), float32 from 2^17 (at 256^2 neither gains).  Smaller ones, and
// RK4 (whose photons take about as many steps each), run on the plain
// one-photon-per-lane kernel.  Both give the same rows and events bit for bit.
constexpr int64_t ROWS_PLANNED_MIN = int64_t(1) << 15;
constexpr int64_t ROWS_PLANNED_MIN_WIDE = int64_t(1) << 17;

bool rows_planned(const Params& P, int dtype, int64_t n) {
    if (P.ipar[I_ADAPTIVE] != 1) return false;
    const bool fp64 = dtype == KR_FLOAT64;
    return n >= (fp64 && !P.ipar[I_DISK] ? ROWS_PLANNED_MIN : ROWS_PLANNED_MIN_WIDE);
}

// The workspace of a planned disk render: the pixels' initial states (n, 8),
// then the row planner's workspace.
size_t round256(size_t bytes) { return (bytes + 255) & ~static_cast<size_t>(255); }
size_t disk_rows_bytes(int64_t n) { return round256(static_cast<size_t>(n) * 64) + host::rows_workspace_bytes(n); }

// The image trace of a disk render into `o`: planned rows on the image's
// initial states (in `ws`, disk_rows_bytes(n)) where rows_planned holds, the
// plain image kernel elsewhere; the same rows and events either way.
kr_status disk_trace(kr_handle h, const Params& P, int dtype, const BardeenImage& B, const kr_outputs& o,
                     int32_t* flags, char* ws, hipStream_t s) {
    const int64_t n = B.n * B.n;
    if (rows_planned(P, dtype, n)) {
        double* Y0 = reinterpret_cast<double*>(ws);
        hipLaunchKernelGGL(image_states_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, B, Y0,
                           flags);
        KR_HIP(hipGetLastError());
        return host::enqueue_rows_planned(h, P, dtype, Y0, n, o, ws + round256(static_cast<size_t>(n) * 64), s);
    }
    ICArgs ic{};
    ic.B = B;
    ic.flags = flags;
    KR_HIP(launch_trace<IC_IMAGE>(P, dtype, ic, o, n, s));
    return KR_OK;
}

}  // namespace

extern "C" kr_status kr_integrate_batch_dev(kr_handle h, const kr_params* p, const double* d_Y0, int64_t n,
                                            const kr_outputs* d_out, void* stream) {
    host::clear_error();
    Params P;
    kr_status st = host::validate(h, p, n, &P);
    if (st != KR_OK) return st;
    const bool disk = P.ipar[I_DISK] != 0;
    if ((st = host::check_outputs(d_out, disk)) != KR_OK) return st;
    if (n > 0 && !d_Y0) return host::fail(KR_ERR_INVALID, "Y0 is NULL");
    KR_HIP(hipSetDevice(h->device));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (rows_planned(P, p->dtype, n)) {
        if ((st = host::reserve_ws(h, host::rows_workspace_bytes(n))) != KR_OK) return st;
        return host::enqueue_rows_planned(h, P, p->dtype, d_Y0, n, *d_out, h->ws, s);
    }
    ICArgs ic{};
    ic.Y0 = d_Y0;
    KR_HIP(launch_trace<IC_Y0>(P, p->dtype, ic, *d_out, n, s));
    return KR_OK;
}

extern "C" kr_status kr_integrate_batch(kr_handle h, const kr_params* p, const double* Y0, int64_t n,
                                        const kr_outputs* out, double* kernel_ms) {
    host::clear_error();
    Params P;
    kr_status st = host::validate(h, p, n, &P);
    if (st != KR_OK) return st;
    const bool disk = P.ipar[I_DISK] != 0;
    if ((st = host::check_outputs(out, disk)) != KR_OK) return st;
    if (n > 0 && !Y0) return host::fail(KR_ERR_INVALID, "Y0 is NULL");
    KR_HIP(hipSetDevice(h->device));
    BatchBuffers b;
    const bool planned = rows_planned(P, p->dtype, n);
    if ((st = carve(h, n, p->dtype, true, disk, b, planned ? host::rows_workspace_bytes(n) : 0)) != KR_OK) return st;
    hipStream_t s = h->stream;
    if (n > 0) KR_HIP(hipMemcpyAsync(b.Y0, Y0, n * 64, hipMemcpyHostToDevice, s));
    ICArgs ic{};
    ic.Y0 = b.Y0;
    KR_HIP(hipEventRecord(h->ev0, s));
    if (planned) {
        if ((st = host::enqueue_rows_planned(h, P, p->dtype, b.Y0, n, b.d, b.extra, s)) != KR_OK) return st;
    } else {
        KR_HIP(launch_trace<IC_Y0>(P, p->dtype, ic, b.d, n, s));
    }
    KR_HIP(hipEventRecord(h->ev1, s));
    if ((st = copy_back(h, b, *out, n, p->dtype, disk)) != KR_OK) return st;
    if (kernel_ms) {
        float ms = 0.f;
        KR_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *kernel_ms = ms;
    }
    return KR_OK;
}

extern "C" kr_status kr_trace_image_dev(kr_handle h, const kr_params* p, const kr_image* img,
                                        const kr_outputs* d_out, int32_t* d_flags, void* stream) {
    host::clear_error();
    BardeenImage B;
    kr_status st = host::bardeen_image(img, true, &B);
    if (st != KR_OK) return st;
    Params P;
    if ((st = host::validate(h, p, B.n * B.n, &P)) != KR_OK) return st;
    if ((st = host::check_outputs(d_out, P.ipar[I_DISK] != 0)) != KR_OK) return st;
    KR_HIP(hipSetDevice(h->device));
    ICArgs ic{};
    ic.B = B;
    ic.flags = d_flags;
    KR_HIP(launch_trace<IC_IMAGE>(P, p->dtype, ic, *d_out, B.n * B.n, static_cast<hipStream_t>(stream)));
    return KR_OK;
}

namespace {

hipError_t launch_shade(const DiskShade& D, int dtype, const void* ev_Y, const uint8_t* ev_hit, int64_t n,
                        double* intensity, double* g, double* r_hit, int32_t* flags, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
    if (dtype == KR_FLOAT64)
        hipLaunchKernelGGL(disk_shade_kernel<double>, grid, block, 0, s, D, static_cast<const double*>(ev_Y), ev_hit,
                           n, intensity, g, r_hit, flags);
    else
        hipLaunchKernelGGL(disk_shade_kernel<float>, grid, block, 0, s, D, static_cast<const float*>(ev_Y), ev_hit, n,
                           intensity, g, r_hit, flags);
    return hipGetLastError();
}

// The disk's shading scalars, checked against the event constants the trace uses.
kr_status disk_setup(const kr_disk* disk, const kr_image* img, const Params& P, DiskShade* D) {
    if (const char* err = disk_shade_scalars(disk, img->mass, img->spin, D)) return host::fail(KR_ERR_INVALID, err);
    if (P.ipar[I_DISK] != 1 || !(P.consts[C_RIN] == disk->r_in) || !(P.consts[C_ROUT] == disk->r_out))
        return host::fail(KR_ERR_INVALID, "params must enable the disk event (ipar[3] = 1) with the disk's r_in, r_out");
    return KR_OK;
}

// kr_trace_image, and with `disk` kr_render_disk: host arrays in and out.
kr_status image_host(kr_handle h, const kr_params* p, const kr_image* img, const kr_disk* disk, const kr_outputs* out,
                     double* intensity, double* g, double* r_hit, double* kernel_ms) {
    host::clear_error();
    BardeenImage B;
    kr_status st = host::bardeen_image(img, true, &B);
    if (st != KR_OK) return st;
    const int64_t n = B.n * B.n;
    Params P;
    if ((st = host::validate(h, p, n, &P)) != KR_OK) return st;
    const bool events = P.ipar[I_DISK] != 0;
    if ((st = host::check_outputs(out, events)) != KR_OK) return st;
    DiskShade D{};
    if (disk) {
        if ((st = disk_setup(disk, img, P, &D)) != KR_OK) return st;
        if (!out->event_Y || !out->event_hit || !intensity || !g || !r_hit)
            return host::fail(KR_ERR_INVALID, "render_disk needs event_Y, event_hit, intensity, g and r_hit");
    }
    KR_HIP(hipSetDevice(h->device));
    BatchBuffers b;
    // intensity, g, r_hit follow the batch buffers; a render's trace
    // workspace follows them
    const size_t shade_bytes = disk ? round256(static_cast<size_t>(n) * 24) : 0;
    const size_t ws_bytes = disk && rows_planned(P, p->dtype, n) ? disk_rows_bytes(n) : 0;
    if ((st = carve(h, n, p->dtype, false, events, b, shade_bytes + ws_bytes)) != KR_OK) return st;
    double* shade = reinterpret_cast<double*>(b.extra);
    hipStream_t s = h->stream;
    KR_HIP(fill_i32(b.flags, 1, s));
    KR_HIP(hipEventRecord(h->ev0, s));
    if (disk) {
        if ((st = disk_trace(h, P, p->dtype, B, b.d, b.flags, b.extra + shade_bytes, s)) != KR_OK) return st;
    } else {
        ICArgs ic{};
        ic.B = B;
        ic.flags = b.flags;
        KR_HIP(launch_trace<IC_IMAGE>(P, p->dtype, ic, b.d, n, s));
    }
    if (disk)
        KR_HIP(launch_shade(D, p->dtype, b.d.event_Y, b.d.event_hit, n, shade, shade + n, shade + 2 * n, b.flags, s));
    KR_HIP(hipEventRecord(h->ev1, s));
    int32_t flags = 0;
    KR_HIP(hipMemcpyAsync(&flags, b.flags, 4, hipMemcpyDeviceToHost, s));
    if (disk) {
        KR_HIP(hipMemcpyAsync(intensity, shade, n * 8, hipMemcpyDeviceToHost, s));
        KR_HIP(hipMemcpyAsync(g, shade + n, n * 8, hipMemcpyDeviceToHost, s));
        KR_HIP(hipMemcpyAsync(r_hit, shade + 2 * n, n * 8, hipMemcpyDeviceToHost, s));
    }
    if ((st = copy_back(h, b, *out, n, p->dtype, events)) != KR_OK) return st;
    if (flags) return host::flag_status(flags);
    if (kernel_ms) {
        float ms = 0.f;
        KR_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *kernel_ms = ms;
    }
    return KR_OK;
}

}  // namespace

extern "C" kr_status kr_trace_image(kr_handle h, const kr_params* p, const kr_image* img, const kr_outputs* out,
                                    double* kernel_ms) {
    return image_host(h, p, img, nullptr, out, nullptr, nullptr, nullptr, kernel_ms);
}

extern "C" kr_status kr_render_disk(kr_handle h, const kr_params* p, const kr_image* img, const kr_disk* disk,
                                    const kr_outputs* out, double* intensity, double* g, double* r_hit,
                                    double* kernel_ms) {
    if (!disk) {
        host::clear_error();
        return host::fail(KR_ERR_INVALID, "disk is NULL");
    }
    return image_host(h, p, img, disk, out, intensity, g, r_hit, kernel_ms);
}

extern "C" kr_status kr_render_disk_dev(kr_handle h, const kr_params* p, const kr_image* img, const kr_disk* disk,
                                        const kr_outputs* d_out, double* d_intensity, double* d_g, double* d_r_hit,
                                        int32_t* d_flags, void* stream) {
    host::clear_error();
    BardeenImage B;
    kr_status st = host::bardeen_image(img, true, &B);
    if (st != KR_OK) return st;
    const int64_t n = B.n * B.n;
    Params P;
    if ((st = host::validate(h, p, n, &P)) != KR_OK) return st;
    if ((st = host::check_outputs(d_out, true)) != KR_OK) return st;
    DiskShade D{};
    if ((st = disk_setup(disk, img, P, &D)) != KR_OK) return st;
    if (!d_intensity || !d_g || !d_r_hit) return host::fail(KR_ERR_INVALID, "intensity, g or r_hit is NULL");
    KR_HIP(hipSetDevice(h->device));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (rows_planned(P, p->dtype, n) && (st = host::reserve_ws(h, disk_rows_bytes(n))) != KR_OK) return st;
    if ((st = disk_trace(h, P, p->dtype, B, *d_out, d_flags, static_cast<char*>(h->ws), s)) != KR_OK) return st;
    KR_HIP(launch_shade(D, p->dtype, d_out->event_Y, d_out->event_hit, n, d_intensity, d_g, d_r_hit, d_flags, s));
    return KR_OK;
}

extern "C" kr_status kr_mask_edge_dev(kr_handle h, const uint8_t* d_mask, int64_t rows, int64_t cols, double alpha0,
                                      double beta0, double d, double centre_alpha, double centre_beta,
                                      const double* d_cos, const double* d_sin, int64_t n_angles, double reach,
                                      double* d_radii, int32_t* d_flags, void* stream) {
    host::clear_error();
    if (!h) return host::fail(KR_ERR_INVALID, "handle is NULL");
    if (rows < 2 || cols < 2 || rows > 65536 || cols > 65536) return host::fail(KR_ERR_INVALID, "need 2 <= rows, cols <= 65536");
    if (!(d > 0.0) || !std::isfinite(d) || !std::isfinite(alpha0) || !std::isfinite(beta0))
        return host::fail(KR_ERR_INVALID, "the grid needs a finite first centre and pixel size d > 0");
    if (!std::isfinite(centre_alpha) || !std::isfinite(centre_beta) || !(reach > 0.0) || !std::isfinite(reach))
        return host::fail(KR_ERR_INVALID, "centre and reach must be finite, reach > 0");
    if (n_angles < 0 || n_angles > (int64_t(1) << 24)) return host::fail(KR_ERR_INVALID, "n_angles must be in [0, 2^24]");
    if (n_angles > 0 && (!d_mask || !d_cos || !d_sin || !d_radii)) return host::fail(KR_ERR_INVALID, "an array is NULL");
    const double ds = 0.125 * d;
    const double samples = std::ceil(reach / ds) + 1.0;
    if (!(samples <= 1e8)) return host::fail(KR_ERR_INVALID, "reach / (d / 8) must stay below 1e8 samples");
    if (n_angles == 0) return KR_OK;
    KR_HIP(hipSetDevice(h->device));
    hipLaunchKernelGGL(mask_edge_kernel, dim3(static_cast<unsigned>((n_angles + 3) / 4)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), d_mask, rows, cols, alpha0, beta0, d, centre_alpha,
                       centre_beta, d_cos, d_sin, n_angles, static_cast<int64_t>(samples), d_radii, d_flags);
    KR_HIP(hipGetLastError());
    return KR_OK;
}

extern "C" kr_status kr_disk_shade_dev(kr_handle h, const kr_disk* disk, double mass, double spin, int32_t dtype,
                                       const void* d_event_Y, const uint8_t* d_event_hit, int64_t n,
                                       double* d_intensity, double* d_g, double* d_r_hit, int32_t* d_flags,
                                       void* stream) {
    host::clear_error();
    if (!h) return host::fail(KR_ERR_INVALID, "handle is NULL");
    DiskShade D{};
    if (const char* err = disk_shade_scalars(disk, mass, spin, &D)) return host::fail(KR_ERR_INVALID, err);
    if (dtype != KR_FLOAT64 && dtype != KR_FLOAT32) return host::fail(KR_ERR_INVALID, "dtype must be float64 or float32");
    if (n < 0 || n > (int64_t(1) << 31)) return host::fail(KR_ERR_INVALID, "n must be in [0, 2^31]");
    if (n > 0 && (!d_event_Y || !d_event_hit || !d_intensity || !d_g || !d_r_hit))
        return host::fail(KR_ERR_INVALID, "an array is NULL");
    KR_HIP(hipSetDevice(h->device));
    KR_HIP(launch_shade(D, dtype, d_event_Y, d_event_hit, n, d_intensity, d_g, d_r_hit, d_flags,
                        static_cast<hipStream_t>(stream)));
    return KR_OK;
}

extern "C" kr_status kr_image_states_dev(kr_handle h, const kr_image* img, int32_t reverse, double* d_Y0,
                                         int32_t* d_flags, void* stream) {
    host::clear_error();
    if (!h) return host::fail(KR_ERR_INVALID, "handle is NULL");
    BardeenImage B;
    kr_status st = host::bardeen_image(img, true, &B);
    if (st != KR_OK) return st;
    if (reverse != 0 && reverse != 1) return host::fail(KR_ERR_INVALID, "reverse must be 0 or 1");
    if (B.camera == CAMERA_BARDEEN && reverse != 1)
        return host::fail(KR_ERR_INVALID, "Bardeen image photons are the inward rays: reverse must be 1");
    if (B.camera == CAMERA_ZAMO) B.Z.sign = reverse ? -1.0 : 1.0;
    const int64_t n = B.n * B.n;
    if (!d_Y0) return host::fail(KR_ERR_INVALID, "Y0 is NULL");
    KR_HIP(hipSetDevice(h->device));
    hipLaunchKernelGGL(image_states_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), B, d_Y0, d_flags);
    KR_HIP(hipGetLastError());
    return KR_OK;
}

extern "C" kr_status kr_point_states_dev(kr_handle h, const kr_image* img, int32_t reverse, const double* d_alpha,
                                         const double* d_beta, int64_t n, double* d_Y0, int32_t* d_flags,
                                         void* stream) {
    host::clear_error();
    if (!h) return host::fail(KR_ERR_INVALID, "handle is NULL");
    BardeenImage B;
    kr_status st = host::bardeen_image(img, false, &B);
    if (st != KR_OK) return st;
    if (reverse != 0 && reverse != 1) return host::fail(KR_ERR_INVALID, "reverse must be 0 or 1");
    if (B.camera == CAMERA_BARDEEN && reverse != 1)
        return host::fail(KR_ERR_INVALID, "Bardeen image photons are the inward rays: reverse must be 1");
    if (B.camera == CAMERA_ZAMO) B.Z.sign = reverse ? -1.0 : 1.0;
    if (n < 0 || n > (int64_t(1) << 31)) return host::fail(KR_ERR_INVALID, "n must be in [0, 2^31]");
    if (n == 0) return KR_OK;
    if (!d_alpha || !d_beta || !d_Y0) return host::fail(KR_ERR_INVALID, "alpha, beta or Y0 is NULL");
    KR_HIP(hipSetDevice(h->device));
    hipLaunchKernelGGL(point_states_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), B, d_alpha, d_beta, n, d_Y0, d_flags);
    KR_HIP(hipGetLastError());
    return KR_OK;
}

extern "C" kr_status kr_zamo_frame(const kr_image* img, double* tetrad, double* metric, double* inverse_metric) {
    host::clear_error();
    if (!tetrad || !metric || !inverse_metric) return host::fail(KR_ERR_INVALID, "an output array is NULL");
    kr_image zi = img ? *img : kr_image{};
    zi.camera = KR_CAMERA_ZAMO;
    BardeenImage B;
    kr_status st = host::bardeen_image(img ? &zi : nullptr, false, &B);
    if (st != KR_OK) return st;
    const ZamoImage& Z = B.Z;
    for (int k = 0; k < 16; ++k) tetrad[k] = metric[k] = inverse_metric[k] = 0.0;
    tetrad[0] = Z.e00;
    tetrad[3] = Z.e03;
    tetrad[5] = Z.e11;
    tetrad[10] = Z.e22;
    tetrad[15] = Z.e33;
    const double g[5] = {Z.g_tt, Z.g_tph, Z.g_rr, Z.g_thth, Z.g_phph};
    const double gi[5] = {Z.gi_tt, Z.gi_tph, Z.gi_rr, Z.gi_thth, Z.gi_phph};
    double* mats[2] = {metric, inverse_metric};
    const double* comps[2] = {g, gi};
    for (int m = 0; m < 2; ++m) {
        mats[m][0] = comps[m][0];
        mats[m][3] = mats[m][12] = comps[m][1];
        mats[m][5] = comps[m][2];
        mats[m][10] = comps[m][3];
        mats[m][15] = comps[m][4];
    }
    return KR_OK;
}

extern "C" kr_status kr_trace_points_dev(kr_handle h, const kr_params* p, const kr_image* img, const double* d_alpha,
                                         const double* d_beta, int64_t n, const kr_outputs* d_out, int32_t* d_flags,
                                         void* stream) {
    host::clear_error();
    BardeenImage B;
    kr_status st = host::bardeen_image(img, false, &B);
    if (st != KR_OK) return st;
    Params P;
    if ((st = host::validate(h, p, n, &P)) != KR_OK) return st;
    if (P.ipar[I_DISK] != 0) return host::fail(KR_ERR_UNSUPPORTED, "the point entry points have no disk event");
    if ((st = host::check_outputs(d_out, false)) != KR_OK) return st;
    if (n > 0 && (!d_alpha || !d_beta)) return host::fail(KR_ERR_INVALID, "alpha/beta is NULL");
    KR_HIP(hipSetDevice(h->device));
    ICArgs ic{};
    ic.B = B;
    ic.alpha = d_alpha;
    ic.beta = d_beta;
    ic.flags = d_flags;
    KR_HIP(launch_trace<IC_POINTS>(P, p->dtype, ic, *d_out, n, static_cast<hipStream_t>(stream)));
    return KR_OK;
}

namespace {

// Largest workgroup of the one-launch bisection: 256 lanes, one per SIMD of
// a CU at one wavefront per SIMD.
constexpr int BISECT_ROUNDS_MAX_PROBE = 256;

// Launches bisect_rounds_kernel on the smallest workgroup (64, 128 or 256
// lanes) that holds a direction's probes (at most 64 probes: as many
// directions as fit in one wavefront) and returns true, if every workgroup
// fits on the chip at one wavefront per SIMD (4 per CU); else launches
// nothing and returns false.  FP64 Dormand-Prince only.
bool launch_bisect_rounds(int cus, const Params& P, int dtype, const BardeenImage& B, double ca, double cb,
                          const double* cosk, const double* sink, double* lo, double* hi, int64_t n_angles,
                          int n_probe, int iterations, unsigned long long* counts, int32_t* flags, hipStream_t s) {
    if (dtype != KR_FLOAT64 || P.ipar[I_ADAPTIVE] == 0 || n_probe > BISECT_ROUNDS_MAX_PROBE) return false;
    const int block = n_probe <= 64 ? 64 : n_probe <= 128 ? 128 : 256;
    const int dirs = block / n_probe;
    const int64_t grid = (n_angles + dirs - 1) / dirs;
    if (grid > static_cast<int64_t>(cus) * (256 / block)) return false;
    const dim3 g(static_cast<unsigned>(grid)), b(static_cast<unsigned>(block));
    if (block == 64)
        hipLaunchKernelGGL((bisect_rounds_kernel<double, true, 64>), g, b, 0, s, P, B, ca, cb, cosk, sink, lo, hi,
                           n_angles, n_probe, dirs, iterations, counts, flags);
    else if (block == 128)
        hipLaunchKernelGGL((bisect_rounds_kernel<double, true, 128>), g, b, 0, s, P, B, ca, cb, cosk, sink, lo, hi,
                           n_angles, n_probe, dirs, iterations, counts, flags);
    else
        hipLaunchKernelGGL((bisect_rounds_kernel<double, true, 256>), g, b, 0, s, P, B, ca, cb, cosk, sink, lo, hi,
                           n_angles, n_probe, dirs, iterations, counts, flags);
    return true;
}

}  // namespace

extern "C" kr_status kr_bisect_edge(kr_handle h, const kr_params* p, const kr_image* img, double centre_alpha,
                                    double centre_beta, const double* cos_angles, const double* sin_angles,
                                    int64_t n_angles, double* lo, double* hi, int32_t iterations, int32_t n_probe,
                                    int64_t* failures, int64_t* nonmonotone, double* kernel_ms) {
    host::clear_error();
    BardeenImage B;
    kr_status st = host::bardeen_image(img, false, &B);
    if (st != KR_OK) return st;
    if (n_angles < 0 || iterations < 0 || n_probe < 1 || n_probe > 4096)
        return host::fail(KR_ERR_INVALID, "need n_angles >= 0, iterations >= 0 and 1 <= n_probe <= 4096");
    Params P;
    if ((st = host::validate(h, p, n_angles * n_probe, &P)) != KR_OK) return st;
    if (P.ipar[I_DISK] != 0) return host::fail(KR_ERR_UNSUPPORTED, "edge bisection has no disk event");
    if (n_angles > 0 && (!cos_angles || !sin_angles || !lo || !hi))
        return host::fail(KR_ERR_INVALID, "an angle or bracket array is NULL");
    KR_HIP(hipSetDevice(h->device));
    host::Carver c;
    const int64_t nt = n_angles * n_probe;
    const size_t ocos = c.take(n_angles * 8), osin = c.take(n_angles * 8), olo = c.take(n_angles * 8),
                 ohi = c.take(n_angles * 8), oesc = c.take(nt), ocnt = c.take(16), oflags = c.take(4);
    if ((st = host::reserve(h, c.total)) != KR_OK) return st;
    char* base = static_cast<char*>(h->buf);
    double* d_cos = reinterpret_cast<double*>(base + ocos);
    double* d_sin = reinterpret_cast<double*>(base + osin);
    double* d_lo = reinterpret_cast<double*>(base + olo);
    double* d_hi = reinterpret_cast<double*>(base + ohi);
    uint8_t* d_esc = reinterpret_cast<uint8_t*>(base + oesc);
    unsigned long long* d_cnt = reinterpret_cast<unsigned long long*>(base + ocnt);
    int32_t* d_flags = reinterpret_cast<int32_t*>(base + oflags);
    hipStream_t s = h->stream;
    if (n_angles > 0) {
        KR_HIP(hipMemcpyAsync(d_cos, cos_angles, n_angles * 8, hipMemcpyHostToDevice, s));
        KR_HIP(hipMemcpyAsync(d_sin, sin_angles, n_angles * 8, hipMemcpyHostToDevice, s));
        KR_HIP(hipMemcpyAsync(d_lo, lo, n_angles * 8, hipMemcpyHostToDevice, s));
        KR_HIP(hipMemcpyAsync(d_hi, hi, n_angles * 8, hipMemcpyHostToDevice, s));
    }
    KR_HIP(fill_u64(d_cnt, 2, s));
    KR_HIP(fill_i32(d_flags, 1, s));
    KR_HIP(hipEventRecord(h->ev0, s));
    const bool adaptive = P.ipar[I_ADAPTIVE] != 0;
    int cus = 0;
    KR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    bool fused = false;
    if (nt > 0 && iterations > 0) {
        fused = launch_bisect_rounds(cus, P, p->dtype, B, centre_alpha, centre_beta, d_cos, d_sin, d_lo, d_hi,
                                     n_angles, n_probe, iterations, d_cnt, d_flags, s);
        KR_HIP(hipGetLastError());
    }
    for (int it = 0; !fused && it < iterations && nt > 0; ++it) {
        const unsigned blocks = static_cast<unsigned>((nt + 255) / 256);
        if (p->dtype == KR_FLOAT64) {
            if (adaptive)
                hipLaunchKernelGGL((bisect_probe_kernel<double, true>), dim3(blocks), dim3(256), 0, s, P, B,
                                   centre_alpha, centre_beta, d_cos, d_sin, d_lo, d_hi, n_angles, n_probe, d_esc,
                                   d_cnt, d_flags);
            else
                hipLaunchKernelGGL((bisect_probe_kernel<double, false>), dim3(blocks), dim3(256), 0, s, P, B,
                                   centre_alpha, centre_beta, d_cos, d_sin, d_lo, d_hi, n_angles, n_probe, d_esc,
                                   d_cnt, d_flags);
        } else {
            if (adaptive)
                hipLaunchKernelGGL((bisect_probe_kernel<float, true>), dim3(blocks), dim3(256), 0, s, P, B,
                                   centre_alpha, centre_beta, d_cos, d_sin, d_lo, d_hi, n_angles, n_probe, d_esc,
                                   d_cnt, d_flags);
            else
                hipLaunchKernelGGL((bisect_probe_kernel<float, false>), dim3(blocks), dim3(256), 0, s, P, B,
                                   centre_alpha, centre_beta, d_cos, d_sin, d_lo, d_hi, n_angles, n_probe, d_esc,
                                   d_cnt, d_flags);
        }
        KR_HIP(hipGetLastError());
        hipLaunchKernelGGL(bisect_update_kernel, dim3(static_cast<unsigned>((n_angles + 255) / 256)), dim3(256), 0,
                           s, d_lo, d_hi, d_esc, n_angles, n_probe, d_cnt + 1);
        KR_HIP(hipGetLastError());
    }
    KR_HIP(hipEventRecord(h->ev1, s));
    unsigned long long cnt[2] = {0, 0};
    int32_t flags = 0;
    if (n_angles > 0) {
        KR_HIP(hipMemcpyAsync(lo, d_lo, n_angles * 8, hipMemcpyDeviceToHost, s));
        KR_HIP(hipMemcpyAsync(hi, d_hi, n_angles * 8, hipMemcpyDeviceToHost, s));
    }
    KR_HIP(hipMemcpyAsync(cnt, d_cnt, 16, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(&flags, d_flags, 4, hipMemcpyDeviceToHost, s));
    KR_HIP(hipStreamSynchronize(s));
    if (failures) *failures = static_cast<int64_t>(cnt[0]);
    if (nonmonotone) *nonmonotone = static_cast<int64_t>(cnt[1]);
    if (flags) return host::flag_status(flags);
    if (kernel_ms) {
        float ms = 0.f;
        KR_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *kernel_ms = ms;
    }
    return KR_OK;
}

namespace {

void launch_impact_probe(const Params& P, int dtype, const ImpactCfg& C, ImpactSearchDev* s, ImpactCounters* c,
                         double* pb, int64_t* ps, int endpoints, int32_t* flags, hipStream_t st) {
    const int64_t lanes = static_cast<int64_t>(C.n_search) * (endpoints ? 2 : C.n_probe);
    const dim3 grid(static_cast<unsigned>((lanes + 255) / 256)), block(256);
    const bool adaptive = P.ipar[I_ADAPTIVE] != 0;
    if (dtype == KR_FLOAT64) {
        if (adaptive)
            hipLaunchKernelGGL((impact_probe_kernel<double, true>), grid, block, 0, st, P, C, s, c, pb, ps, endpoints,
                               flags);
        else
            hipLaunchKernelGGL((impact_probe_kernel<double, false>), grid, block, 0, st, P, C, s, c, pb, ps, endpoints,
                               flags);
    } else {
        if (adaptive)
            hipLaunchKernelGGL((impact_probe_kernel<float, true>), grid, block, 0, st, P, C, s, c, pb, ps, endpoints,
                               flags);
        else
            hipLaunchKernelGGL((impact_probe_kernel<float, false>), grid, block, 0, st, P, C, s, c, pb, ps, endpoints,
                               flags);
    }
}

}  // namespace

extern "C" kr_status kr_impact_search(kr_handle h, const kr_params* p, const kr_impact_opts* o, int32_t n_search,
                                      const double* lz_sign, double* lo, double* hi, int64_t* iterations,
                                      double* probe_b, int64_t* probe_state, uint8_t* was_open, kr_impact_result* res,
                                      double* kernel_ms) {
    host::clear_error();
    if (!o || !res) return host::fail(KR_ERR_INVALID, "options or result is NULL");
    if (n_search < 1 || n_search > 1024) return host::fail(KR_ERR_INVALID, "n_search must be in [1, 1024]");
    if (o->n_probe < 1 || o->n_probe > 1024) return host::fail(KR_ERR_INVALID, "n_probe must be in [1, 1024]");
    if (!(o->tol > 0.0) || o->max_iterations < 1)
        return host::fail(KR_ERR_INVALID, "need tol > 0 and max_iterations >= 1");
    if (o->horizon_crossing_as_capture != 0 && o->horizon_crossing_as_capture != 1)
        return host::fail(KR_ERR_INVALID, "horizon_crossing_as_capture must be 0 or 1");
    if (!lz_sign || !lo || !hi || !iterations || !probe_b || !probe_state || !was_open)
        return host::fail(KR_ERR_INVALID, "an array is NULL");
    const int per = o->n_probe > 2 ? o->n_probe : 2;
    const int64_t lanes = static_cast<int64_t>(n_search) * per;
    Params P;
    kr_status st = host::validate(h, p, lanes, &P);
    if (st != KR_OK) return st;
    if (P.ipar[I_DISK] != 0) return host::fail(KR_ERR_UNSUPPORTED, "the impact search has no disk event");
    const double m = P.geom[G_MASS], a = P.geom[G_A];
    if (!(o->r0 > m + std::sqrt(m * m - P.geom[G_A2])) || !std::isfinite(o->r0))
        return host::fail(KR_ERR_INVALID, "r0 must lie outside the outer horizon");
    for (int k = 0; k < n_search; ++k) {
        if (lz_sign[k] != 1.0 && lz_sign[k] != -1.0) return host::fail(KR_ERR_INVALID, "lz_sign must be +1 or -1");
        if (!(lo[k] < hi[k])) return host::fail(KR_ERR_INVALID, "each bracket needs lo < hi");
    }
    ImpactCfg C{};
    C.L = ImpactLaunch{m, a, o->r0, o->theta, o->cos2_theta, o->sin2_theta};
    C.tol = o->tol;
    C.r_capture = o->capture_radius;
    C.max_iterations = o->max_iterations;
    C.n_search = n_search;
    C.n_probe = o->n_probe;
    C.relabel = o->horizon_crossing_as_capture;
    KR_HIP(hipSetDevice(h->device));
    host::Carver cv;
    const size_t os = cv.take(n_search * static_cast<int64_t>(sizeof(ImpactSearchDev))),
                 oc = cv.take(sizeof(ImpactCounters)), ob = cv.take(lanes * 8), ost = cv.take(lanes * 8),
                 of = cv.take(4);
    if ((st = host::reserve(h, cv.total)) != KR_OK) return st;
    char* base = static_cast<char*>(h->buf);
    ImpactSearchDev* d_s = reinterpret_cast<ImpactSearchDev*>(base + os);
    ImpactCounters* d_c = reinterpret_cast<ImpactCounters*>(base + oc);
    double* d_b = reinterpret_cast<double*>(base + ob);
    int64_t* d_state = reinterpret_cast<int64_t*>(base + ost);
    int32_t* d_flags = reinterpret_cast<int32_t*>(base + of);
    std::vector<ImpactSearchDev> S(n_search);
    double widest = 0.0;
    for (int k = 0; k < n_search; ++k) {
        S[k] = ImpactSearchDev{lo[k], hi[k], lz_sign[k], 0, 0, 0};
        widest = std::max(widest, hi[k] - lo[k]);
    }
    ImpactCounters zero{};
    hipStream_t s = h->stream;
    KR_HIP(hipMemcpyAsync(d_s, S.data(), S.size() * sizeof(ImpactSearchDev), hipMemcpyHostToDevice, s));
    KR_HIP(hipMemcpyAsync(d_c, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    KR_HIP(fill_i32(d_flags, 1, s));
    KR_HIP(hipEventRecord(h->ev0, s));
    // Rounds: the bracket shrinks (n_probe + 1)-fold per round, so the host
    // enqueues that many (plus a margin) without waiting, then checks.
    const double shrink = std::log(static_cast<double>(o->n_probe) + 1.0);
    const double want = widest > o->tol ? std::ceil(std::log(widest / o->tol) / shrink) + 2.0 : 1.0;
    int64_t batch = static_cast<int64_t>(std::min(want, 4096.0));
    int64_t enqueued = 0;
    ImpactCounters cnt{};
    for (int pass = 0;; ++pass) {
        if (pass == 0) {
            launch_impact_probe(P, p->dtype, C, d_s, d_c, d_b, d_state, 1, d_flags, s);
            KR_HIP(hipGetLastError());
            hipLaunchKernelGGL(impact_update_kernel, dim3(1), dim3(1024), 0, s, C, d_s, d_c, d_b, d_state, 1);
            KR_HIP(hipGetLastError());
        }
        const int64_t todo = std::min<int64_t>(batch, o->max_iterations - enqueued);
        for (int64_t r = 0; r < todo; ++r) {
            launch_impact_probe(P, p->dtype, C, d_s, d_c, d_b, d_state, 0, d_flags, s);
            KR_HIP(hipGetLastError());
            hipLaunchKernelGGL(impact_update_kernel, dim3(1), dim3(1024), 0, s, C, d_s, d_c, d_b, d_state, 0);
            KR_HIP(hipGetLastError());
        }
        enqueued += todo;
        KR_HIP(hipEventRecord(h->ev1, s));
        KR_HIP(hipMemcpyAsync(S.data(), d_s, S.size() * sizeof(ImpactSearchDev), hipMemcpyDeviceToHost, s));
        KR_HIP(hipMemcpyAsync(&cnt, d_c, sizeof(cnt), hipMemcpyDeviceToHost, s));
        KR_HIP(hipStreamSynchronize(s));
        bool open = false;
        for (const ImpactSearchDev& x : S) open = open || x.open;
        if (cnt.error || !open || enqueued >= o->max_iterations) break;
        batch = 4;  // the estimate fell short (rounding): a few more rounds at a time
    }
    int32_t flags = 0;
    KR_HIP(hipMemcpyAsync(&flags, d_flags, 4, hipMemcpyDeviceToHost, s));
    if (cnt.error) {
        KR_HIP(hipMemcpyAsync(probe_b, d_b, lanes * 8, hipMemcpyDeviceToHost, s));
        KR_HIP(hipMemcpyAsync(probe_state, d_state, lanes * 8, hipMemcpyDeviceToHost, s));
    }
    KR_HIP(hipStreamSynchronize(s));
    if (flags) return host::flag_status(flags);
    for (int k = 0; k < n_search; ++k) {
        lo[k] = S[k].lo;
        hi[k] = S[k].hi;
        iterations[k] = S[k].iterations;
        was_open[k] = S[k].open ? 1 : 0;
    }
    res->rays = static_cast<int64_t>(cnt.rays);
    res->max_steps = static_cast<int64_t>(cnt.max_steps);
    res->crossings = static_cast<int64_t>(cnt.crossings);
    res->error = cnt.error;
    res->error_round = cnt.error_round;
    if (kernel_ms) {
        float ms = 0.f;
        KR_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *kernel_ms = ms;
    }
    return KR_OK;
}
