// CPU oracle for the per-photon null-geodesic path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: it is loaded by the
// tests, by the smoke check and by the benchmark's CPU-baseline leg, and by
// nothing else. The product path (kerrray_amd/) never links or calls it.
//
// It restates, operation by operation and in the same floating-point order,
// the reference's scalar Numba kernel
//   src/kerrray/raytracing/backends/numba_kernel.py:111-474
// (inverse_metric_terms :111-154, geodesic_rhs_scalar :158-197,
//  dormand_prince_step :201-238, rk4_step :242-257, _error_norm :261-268,
//  _step_factor :272-281, _all_finite :285-289, _classify :293-311,
//  _null_error :315-321, _carter :325-336, _integrate_ray :340-463,
//  integrate_rays_kernel :466-474)
// and the Bardeen image-plane initial conditions
//   src/kerrray/raytracing/camera.py:206-222          (pixel centres)
//   src/kerrray/benchmarks/rayset.py:158-165          (_off_equatorial_state)
//   src/kerrray/geodesics/initial_conditions.py:50-112 (carter_potentials,
//                                                      photon_from_constants)
// Build: oracle/Makefile (g++ -O2 -fno-fast-math -ffp-contract=off -fopenmp),
// so that no contraction or reassociation changes a rounding.  Transcendentals
// are glibc's sin, cos, pow (the same libm the compiled Numba kernel calls).
//
// Working dtype: the reference specialises the kernel on float32/float64;
// here the template parameter T plays that role.  The metric is always
// evaluated in double and cast, as in numba_kernel.py:165-167.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>

#include <vector>

namespace {

enum { RUNNING = 0, ESCAPED = 1, CAPTURED = 2, MAX_AFFINE = 3, FAILURE = 4, OUT_OF_DOMAIN = 5, DISK_HIT = 6 };
enum { G_MASS, G_A, G_A2, G_NEG_2MA, G_POS_2MA, G_TWO_M, G_TWO_A2 };
enum {
    C_ZERO, C_ONE, C_TWO, C_NEG_HALF, C_EIGHT, C_SAFETY, C_FMIN, C_FMAX, C_EXP, C_RTOL,
    C_ATOL, C_HMIN, C_RCAP, C_RESC, C_THLO, C_THHI, C_AXIS, C_H0, C_RIN, C_ROUT, C_HALFPI,
    N_CONSTS
};
enum { P_LAMMAX, P_HALFPI, P_HMAX };
enum { I_MAX_STEPS, I_GUARD, I_ADAPTIVE, I_DISK };
enum { IDX_T, IDX_R, IDX_TH, IDX_PH, IDX_PT, IDX_PR, IDX_PTH, IDX_PPH };
constexpr int S = 8;

// Python's builtin max/min on two floats: the first argument unless the
// second compares strictly greater (smaller).
template <typename T> inline T pymax(T a, T b) { return b > a ? b : a; }
template <typename T> inline T pymin(T a, T b) { return b < a ? b : a; }

inline double wpow(double x, double y) { return std::pow(x, y); }
inline float wpow(float x, float y) { return powf(x, y); }
inline double wsqrt(double x) { return std::sqrt(x); }
inline float wsqrt(float x) { return sqrtf(x); }
inline double wsin(double x) { return std::sin(x); }
inline float wsin(float x) { return sinf(x); }

// numba_kernel.py:111-154
void inverse_metric_terms(const double* geom, double r, double th, double* out) {
    const double m = geom[G_MASS];
    const double a2 = geom[G_A2];
    const double s = std::sin(th);
    const double c = std::cos(th);
    const double s2 = s * s;
    const double sin2th = std::sin(2.0 * th);
    const double r2 = r * r;
    const double sig = r2 + a2 * (c * c);
    const double dlt = r2 - geom[G_TWO_M] * r + a2;
    const double r2a2 = r2 + a2;
    const double big_a = r2a2 * r2a2 - a2 * dlt * s2;
    const double sd = sig * dlt;
    out[0] = -big_a / sd;
    out[1] = geom[G_NEG_2MA] * r / sd;
    out[2] = dlt / sig;
    out[3] = 1.0 / sig;
    out[4] = (dlt - a2 * s2) / (sd * s2);
    const double sig_r = 2.0 * r;
    const double sig_th = (-a2) * sin2th;
    const double dlt_r = 2.0 * (r - m);
    const double a_r = 4.0 * r * r2a2 - geom[G_TWO_A2] * (r - m) * s2;
    const double a_th = (-a2) * dlt * sin2th;
    const double sd_r = sig_r * dlt + sig * dlt_r;
    const double sd_th = sig_th * dlt;
    const double sd2 = sd * sd;
    const double sig2 = sig * sig;
    out[5] = -(a_r * sd - big_a * sd_r) / sd2;
    out[6] = geom[G_NEG_2MA] * (sd - r * sd_r) / sd2;
    out[7] = (dlt_r * sig - dlt * sig_r) / sig2;
    out[8] = -sig_r / sig2;
    out[9] = -sig_r / (sig2 * s2) + a2 * sd_r / sd2;
    out[10] = -(a_th * sd - big_a * sd_th) / sd2;
    out[11] = geom[G_POS_2MA] * r * sd_th / sd2;
    out[12] = -dlt * sig_th / sig2;
    out[13] = -sig_th / sig2;
    out[14] = -(sig_th * s2 + sig * sin2th) / (sig2 * (s2 * s2)) + a2 * sd_th / sd2;
}

// numba_kernel.py:158-197
template <typename T>
void geodesic_rhs_scalar(const double* geom, const T* consts, const T* y, T* f) {
    double g64[15];
    T g[15];
    inverse_metric_terms(geom, static_cast<double>(y[IDX_R]), static_cast<double>(y[IDX_TH]), g64);
    for (int i = 0; i < 15; ++i) g[i] = static_cast<T>(g64[i]);
    const T zero = consts[C_ZERO];
    const T two = consts[C_TWO];
    const T p_t = y[IDX_PT], p_r = y[IDX_PR], p_th = y[IDX_PTH], p_ph = y[IDX_PPH];
    T gphph = g[4], r_gphph = g[9], t_gphph = g[14];
    if (p_ph == zero) {
        gphph = zero;
        r_gphph = zero;
        t_gphph = zero;
    }
    f[IDX_T] = g[0] * p_t + g[1] * p_ph;
    f[IDX_R] = g[2] * p_r;
    f[IDX_TH] = g[3] * p_th;
    f[IDX_PH] = g[1] * p_t + gphph * p_ph;
    f[IDX_PT] = zero;
    const T pt2 = p_t * p_t, ptpph = p_t * p_ph, pr2 = p_r * p_r, pth2 = p_th * p_th, pph2 = p_ph * p_ph;
    f[IDX_PR] = consts[C_NEG_HALF] * (g[5] * pt2 + two * g[6] * ptpph + g[7] * pr2 + g[8] * pth2 + r_gphph * pph2);
    f[IDX_PTH] = consts[C_NEG_HALF] * (g[10] * pt2 + two * g[11] * ptpph + g[12] * pr2 + g[13] * pth2 + t_gphph * pph2);
    f[IDX_PPH] = zero;
}

// numba_kernel.py:201-238 (k[0] = f(y) on entry; fills k[1..6], y_new, err)
template <typename T>
void dormand_prince_step(const double* geom, const T* consts, const T* A, const T* E, const T* y, T h,
                         T (*k)[S], T* ys, T* y_new, T* err) {
#define KA(i, j) A[(i) * 7 + (j)]
    for (int i = 0; i < S; ++i) ys[i] = y[i] + h * (KA(1, 0) * k[0][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[1]);
    for (int i = 0; i < S; ++i) ys[i] = y[i] + h * (KA(2, 0) * k[0][i] + KA(2, 1) * k[1][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[2]);
    for (int i = 0; i < S; ++i) ys[i] = y[i] + h * (KA(3, 0) * k[0][i] + KA(3, 1) * k[1][i] + KA(3, 2) * k[2][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[3]);
    for (int i = 0; i < S; ++i)
        ys[i] = y[i] + h * (KA(4, 0) * k[0][i] + KA(4, 1) * k[1][i] + KA(4, 2) * k[2][i] + KA(4, 3) * k[3][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[4]);
    for (int i = 0; i < S; ++i)
        ys[i] = y[i] + h * (KA(5, 0) * k[0][i] + KA(5, 1) * k[1][i] + KA(5, 2) * k[2][i] + KA(5, 3) * k[3][i] +
                            KA(5, 4) * k[4][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[5]);
    for (int i = 0; i < S; ++i)
        y_new[i] = y[i] + h * (KA(6, 0) * k[0][i] + KA(6, 2) * k[2][i] + KA(6, 3) * k[3][i] + KA(6, 4) * k[4][i] +
                               KA(6, 5) * k[5][i]);
    geodesic_rhs_scalar(geom, consts, y_new, k[6]);
    for (int i = 0; i < S; ++i)
        err[i] = h * (E[0] * k[0][i] + E[2] * k[2][i] + E[3] * k[3][i] + E[4] * k[4][i] + E[5] * k[5][i] +
                      E[6] * k[6][i]);
#undef KA
}

// numba_kernel.py:242-257
template <typename T>
void rk4_step(const double* geom, const T* consts, const T* b, const T* c, const T* y, T h, T (*k)[S], T* ys,
              T* y_new) {
    for (int i = 0; i < S; ++i) ys[i] = y[i] + h * (c[1] * k[0][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[1]);
    for (int i = 0; i < S; ++i) ys[i] = y[i] + h * (c[2] * k[1][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[2]);
    for (int i = 0; i < S; ++i) ys[i] = y[i] + h * (c[3] * k[2][i]);
    geodesic_rhs_scalar(geom, consts, ys, k[3]);
    for (int i = 0; i < S; ++i) y_new[i] = y[i] + h * (b[0] * k[0][i] + b[1] * k[1][i] + b[2] * k[2][i] + b[3] * k[3][i]);
    geodesic_rhs_scalar(geom, consts, y_new, k[6]);
}

// numba_kernel.py:261-268: HNW scaled RMS, NumPy's pairwise order over 8 terms.
template <typename T>
T error_norm(const T* consts, const T* err, const T* y, const T* y_new) {
    T q[S];
    for (int i = 0; i < S; ++i) {
        const T scale = consts[C_ATOL] + consts[C_RTOL] * pymax(std::fabs(y[i]), std::fabs(y_new[i]));
        const T v = err[i] / scale;
        q[i] = v * v;
    }
    const T total = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
    return wsqrt(total / consts[C_EIGHT]);
}

// numba_kernel.py:272-281
template <typename T>
T step_factor(const T* consts, T en, bool rejected_before) {
    T factor;
    if (en == consts[C_ZERO]) {
        factor = consts[C_FMAX];
    } else {
        const T raw = consts[C_SAFETY] * wpow(en, consts[C_EXP]);
        factor = pymin(consts[C_FMAX], pymax(consts[C_FMIN], raw));
    }
    if (rejected_before) factor = pymin(factor, consts[C_ONE]);
    return factor;
}

template <typename T>
bool all_finite(const T* y, const T* f) {
    for (int i = 0; i < S; ++i)
        if (!(std::isfinite(y[i]) && std::isfinite(f[i]))) return false;
    return true;
}

// numba_kernel.py:293-311
template <typename T>
int classify(const T* consts, const double* params, const int64_t* ipar, const T* y, const T* f, double lam,
             int64_t n_steps, T h, bool check_h) {
    if (!all_finite(y, f)) return FAILURE;
    if (check_h && h < consts[C_HMIN]) return FAILURE;
    const T r = y[IDX_R];
    const T th = y[IDX_TH];
    if (r < consts[C_ZERO] || th < consts[C_THLO] || th > consts[C_THHI]) return OUT_OF_DOMAIN;
    if (y[IDX_PPH] != consts[C_ZERO] && std::fabs(wsin(th)) < consts[C_AXIS]) return OUT_OF_DOMAIN;
    if (r <= consts[C_RCAP]) return CAPTURED;
    if (r >= consts[C_RESC] && f[IDX_R] > consts[C_ZERO]) return ESCAPED;
    if (lam >= params[P_LAMMAX] || n_steps >= ipar[I_MAX_STEPS]) return MAX_AFFINE;
    return RUNNING;
}

// numba_kernel.py:315-321
template <typename T>
double null_error(const T* y, const T* f) {
    double total = static_cast<double>(y[IDX_PT]) * static_cast<double>(f[IDX_T]);
    total = total + static_cast<double>(y[IDX_PR]) * static_cast<double>(f[IDX_R]);
    total = total + static_cast<double>(y[IDX_PTH]) * static_cast<double>(f[IDX_TH]);
    total = total + static_cast<double>(y[IDX_PPH]) * static_cast<double>(f[IDX_PH]);
    return std::fabs(0.5 * total);
}

// numba_kernel.py:325-336
template <typename T>
double carter(double a2, const T* y) {
    const double th = static_cast<double>(y[IDX_TH]);
    const double p_th = static_cast<double>(y[IDX_PTH]);
    const double lz = static_cast<double>(y[IDX_PPH]);
    const double e = -static_cast<double>(y[IDX_PT]);
    const double s = std::sin(th);
    const double c = std::cos(th);
    double lz2_over_s2 = 0.0;
    if (lz != 0.0) lz2_over_s2 = (lz * lz) / (s * s);
    return p_th * p_th + (c * c) * (lz2_over_s2 - a2 * e * e);
}

struct RayArgs {
    const double* geom;
    const double* params;
    const int64_t* ipar;
    // working-dtype copies are built per call (below)
    double* lam_out;
    int64_t* n_steps_out;
    int64_t* n_rej_out;
    int64_t* state_out;
    double* max_null;
    double* max_e;
    double* max_lz;
    double* max_q;
    const double* e0;
    const double* lz0;
    const double* q0;
    const double* fe;
    const double* fl;
    const double* fq;
    const double* fe2;
    uint8_t* event_hit;
    std::vector<double>* steps = nullptr;  // when set: accepted step sizes of ray n go to steps[n]
};

// numba_kernel.py:340-463
template <typename T>
void integrate_ray(int64_t n, const RayArgs& a, const T* consts, const T* tab_a, const T* tab_e, const T* rk4_b,
                   const T* rk4_c, T* Y, T* F, T* event_Y) {
    T* y = Y + n * S;
    T* f = F + n * S;
    T k[7][S], ys[S], y_new[S], y_prev[S], err[S];
    const double* geom = a.geom;
    const double a2 = geom[G_A2];
    const bool adaptive = a.ipar[I_ADAPTIVE] != 0;
    const bool disk = a.ipar[I_DISK] != 0;
    const double lambda_max = a.params[P_LAMMAX];
    const double half_pi = a.params[P_HALFPI];
    geodesic_rhs_scalar(geom, consts, y, f);
    double m_null = null_error(y, f) / a.fe2[n];
    double m_e = 0.0, m_lz = 0.0, m_q = 0.0;
    double lam = 0.0;
    T h = consts[C_H0];
    const double h_cap = a.params[P_HMAX];
    if (adaptive && static_cast<double>(h) > h_cap) h = static_cast<T>(h_cap);
    int64_t n_steps = 0, n_rej = 0, attempts = 0;
    bool rejected_before = false;
    int state = classify(consts, a.params, a.ipar, y, f, lam, n_steps, h, false);
    while (state == RUNNING) {
        attempts += 1;
        if (attempts > a.ipar[I_GUARD]) {
            state = FAILURE;
            break;
        }
        const double remaining = lambda_max - lam;
        const bool last = static_cast<double>(h) >= remaining;
        T h_eff = h;
        if (last) h_eff = static_cast<T>(remaining);
        for (int i = 0; i < S; ++i) {
            k[0][i] = f[i];
            y_prev[i] = y[i];
        }
        if (adaptive)
            dormand_prince_step(geom, consts, tab_a, tab_e, y, h_eff, k, ys, y_new, err);
        else
            rk4_step(geom, consts, rk4_b, rk4_c, y, h_eff, k, ys, y_new);
        bool finite = all_finite(y_new, k[6]);
        bool accept = finite;
        T h_new = h;
        if (adaptive && finite) {
            const T en = error_norm(consts, err, y, y_new);
            if (!std::isfinite(en)) {
                finite = false;
                accept = false;
            } else {
                accept = en <= consts[C_ONE];
                const T factor = step_factor(consts, en, rejected_before);
                if (!(accept && last)) {
                    h_new = h_eff * factor;
                    if (static_cast<double>(h_new) > h_cap) h_new = static_cast<T>(h_cap);
                }
            }
        }
        if (!finite) {
            state = FAILURE;
            break;
        }
        const bool rejected = !accept;
        if (rejected) n_rej += 1;
        rejected_before = rejected;
        h = h_new;
        if (rejected) {
            if (h_new < consts[C_HMIN]) {
                state = FAILURE;
                break;
            }
            continue;
        }
        for (int i = 0; i < S; ++i) {
            y[i] = y_new[i];
            f[i] = k[6][i];
        }
        if (a.steps) a.steps[n].push_back(static_cast<double>(h_eff));
        const double lam_old = lam;
        if (last)
            lam = lambda_max;
        else
            lam = lam + static_cast<double>(h_eff);
        n_steps += 1;
        const double e = -static_cast<double>(y[IDX_PT]);
        const double lz = static_cast<double>(y[IDX_PPH]);
        m_e = pymax(m_e, std::fabs(e - a.e0[n]) / pymax(std::fabs(a.e0[n]), a.fe[n]));
        m_lz = pymax(m_lz, std::fabs(lz - a.lz0[n]) / pymax(std::fabs(a.lz0[n]), a.fl[n]));
        m_q = pymax(m_q, std::fabs(carter(a2, y) - a.q0[n]) / pymax(std::fabs(a.q0[n]), a.fq[n]));
        m_null = pymax(m_null, null_error(y, f) / a.fe2[n]);
        int code = classify(consts, a.params, a.ipar, y, f, lam, n_steps, h, true);
        if (disk) {
            const double th_old = static_cast<double>(y_prev[IDX_TH]);
            const double th_new = static_cast<double>(y[IDX_TH]);
            if ((th_old - half_pi) * (th_new - half_pi) < 0.0) {
                const double s = (half_pi - th_old) / (th_new - th_old);
                for (int i = 0; i < S; ++i) {
                    const T cast = static_cast<T>(s * static_cast<double>(y[i] - y_prev[i]));
                    ys[i] = y_prev[i] + cast;
                }
                ys[IDX_TH] = consts[C_HALFPI];
                if (ys[IDX_R] >= consts[C_RIN] && ys[IDX_R] <= consts[C_ROUT]) {
                    for (int i = 0; i < S; ++i) {
                        event_Y[n * S + i] = ys[i];
                        y[i] = ys[i];
                    }
                    a.event_hit[n] = 1;
                    lam = lam_old + s * static_cast<double>(h_eff);
                    code = DISK_HIT;
                }
            }
        }
        state = code;
    }
    a.state_out[n] = state;
    a.n_steps_out[n] = n_steps;
    a.n_rej_out[n] = n_rej;
    a.lam_out[n] = lam;
    a.max_null[n] = m_null;
    a.max_e[n] = m_e;
    a.max_lz[n] = m_lz;
    a.max_q[n] = m_q;
}

template <typename T>
void run_all(int64_t n_rays, const RayArgs& a, const double* consts64, const double* tab_a64, const double* tab_e64,
             const double* rk4_b64, const double* rk4_c64, void* Y, void* F, void* event_Y, int n_threads,
             int dynamic_schedule) {
    // Working-dtype copies, as np.array(..., dtype=dtype) does in numba_backend.py:100-128.
    T consts[N_CONSTS], tab_a[49], tab_e[7], rk4_b[4], rk4_c[4];
    for (int i = 0; i < N_CONSTS; ++i) consts[i] = static_cast<T>(consts64[i]);
    for (int i = 0; i < 49; ++i) tab_a[i] = static_cast<T>(tab_a64[i]);
    for (int i = 0; i < 7; ++i) tab_e[i] = static_cast<T>(tab_e64[i]);
    for (int i = 0; i < 4; ++i) {
        rk4_b[i] = static_cast<T>(rk4_b64[i]);
        rk4_c[i] = static_cast<T>(rk4_c64[i]);
    }
    T* Yt = static_cast<T*>(Y);
    T* Ft = static_cast<T*>(F);
    T* Et = static_cast<T*>(event_Y);
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    if (dynamic_schedule) {
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 16)
        for (int64_t n = 0; n < n_rays; ++n) integrate_ray<T>(n, a, consts, tab_a, tab_e, rk4_b, rk4_c, Yt, Ft, Et);
    } else {
#pragma omp parallel for num_threads(n_threads) schedule(static)
        for (int64_t n = 0; n < n_rays; ++n) integrate_ray<T>(n, a, consts, tab_a, tab_e, rk4_b, rk4_c, Yt, Ft, Et);
    }
}

}  // namespace

extern "C" {

// integrate_rays_kernel (numba_kernel.py:466-474) with the same argument
// meaning.  dtype: 0 = float64, 1 = float32 (Y, F, event_Y in that dtype).
// Y (n, 8) holds the initial states on entry and the final ones on return.
int kro_integrate_rays(int dtype, const double* geom, const double* consts, const double* params, const int64_t* ipar,
                       const double* tab_a, const double* tab_e, const double* rk4_b, const double* rk4_c, void* Y,
                       void* F, double* lam, int64_t* n_steps, int64_t* n_rej, int64_t* state, double* max_null,
                       double* max_e, double* max_lz, double* max_q, const double* e0, const double* lz0,
                       const double* q0, const double* fe, const double* fl, const double* fq, const double* fe2,
                       void* event_Y, uint8_t* event_hit, int64_t n_rays, int n_threads, int dynamic_schedule) {
    RayArgs a{geom, params, ipar, lam, n_steps, n_rej, state, max_null, max_e, max_lz, max_q,
              e0, lz0, q0, fe, fl, fq, fe2, event_hit};
    if (dtype == 0)
        run_all<double>(n_rays, a, consts, tab_a, tab_e, rk4_b, rk4_c, Y, F, event_Y, n_threads, dynamic_schedule);
    else if (dtype == 1)
        run_all<float>(n_rays, a, consts, tab_a, tab_e, rk4_b, rk4_c, Y, F, event_Y, n_threads, dynamic_schedule);
    else
        return 1;
    return 0;
}

// Camera.pixel_coordinates (camera.py:206-222): centre_j = fov (2 j + 1 - n) / n,
// evaluated as NumPy does: ((fov * ((2.0 * j + 1.0) - n)) / n).
void kro_pixel_centres(double fov, int64_t n, double* centres) {
    const double nd = static_cast<double>(n);
    for (int64_t j = 0; j < n; ++j) centres[j] = (fov * ((2.0 * static_cast<double>(j) + 1.0) - nd)) / nd;
}

// Bardeen-constant photon at the observer (rayset.py:158-165 ->
// initial_conditions.py:83-112 with E = 1, sign_r = -1).  Python's float
// `x ** 2` is glibc pow(x, 2.0) and is written as such.  Returns 0, or 1 when
// a radicand is negative beyond RADICAND_TOLERANCE (initial_conditions.py:71-74).
int kro_bardeen_state(double mass, double a, double r0, double theta_o, double alpha, double beta, int sign_theta,
                      double* y) {
    const double sin_o = std::sin(theta_o), cos_o = std::cos(theta_o);
    const double xi = -alpha * sin_o;
    const double eta = beta * beta - std::pow(a, 2.0) * std::pow(cos_o, 2.0) + xi * xi * std::pow(cos_o / sin_o, 2.0);
    const double E = 1.0, Lz = xi, Q = eta;
    // carter_potentials (initial_conditions.py:50-68); delta() is NumPy: r**2 is r*r.
    const double dlt = r0 * r0 - 2.0 * mass * r0 + std::pow(a, 2.0);
    const double term1 = E * (r0 * r0 + a * a) - a * Lz;
    const double term2 = dlt * (Q + std::pow(Lz - a * E, 2.0));
    const double big_r = term1 * term1 - term2;
    const double scale_r = term1 * term1 + std::fabs(term2);
    const double c2 = std::pow(std::cos(theta_o), 2.0), s2 = std::pow(std::sin(theta_o), 2.0);
    const double lz_term = (Lz == 0.0) ? 0.0 : Lz * Lz / s2;
    const double big_theta = Q - c2 * (lz_term - a * a * E * E);
    const double scale_theta = std::fabs(Q) + c2 * (lz_term + a * a * E * E);
    const double tol = 1e-12;
    if (big_r < -tol * pymax(scale_r, 1.0)) return 1;
    if (big_theta < -tol * pymax(scale_theta, 1.0)) return 1;
    const double p_r = -1.0 * std::sqrt(pymax(big_r, 0.0)) / dlt;
    const double p_th = static_cast<double>(sign_theta) * std::sqrt(pymax(big_theta, 0.0));
    y[0] = 0.0;
    y[1] = r0;
    y[2] = theta_o;
    y[3] = 0.0;
    y[4] = -E;
    y[5] = p_r;
    y[6] = p_th;
    y[7] = Lz;
    return 0;
}

// Row-major image of Bardeen states (pixel (i, j) -> row i*n + j), with
// alpha = centre_j, beta = -centre_i and sign_theta = -1 for beta >= 0, +1
// otherwise (the backward ray leaves towards the image's own side).
int kro_bardeen_image(double mass, double a, double r0, double theta_o, double fov, int64_t n, double* Y) {
    double* centres = new double[n];
    kro_pixel_centres(fov, n, centres);
    int bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double beta = -centres[i];
        for (int64_t j = 0; j < n; ++j) {
            const double alpha = centres[j];
            bad |= kro_bardeen_state(mass, a, r0, theta_o, alpha, beta, beta >= 0.0 ? -1 : 1, Y + (i * n + j) * S);
        }
    }
    delete[] centres;
    return bad;
}

// Single-call building blocks, exposed for known-answer tests.
void kro_inverse_metric_terms(const double* geom, double r, double th, double* out) {
    inverse_metric_terms(geom, r, th, out);
}

void kro_rhs(const double* geom, const double* consts64, const double* y, double* f) {
    double consts[N_CONSTS];
    std::memcpy(consts, consts64, sizeof(consts));
    geodesic_rhs_scalar<double>(geom, consts, y, f);
}

double kro_error_norm(const double* consts64, const double* err, const double* y, const double* y_new) {
    return error_norm<double>(consts64, err, y, y_new);
}

double kro_step_factor(const double* consts64, double en, int rejected_before) {
    return step_factor<double>(consts64, en, rejected_before != 0);
}

void kro_dp5_step(const double* geom, const double* consts64, const double* tab_a, const double* tab_e,
                  const double* y, double h, const double* f0, double* y_new, double* err, double* f_new) {
    double k[7][S], ys[S];
    std::memcpy(k[0], f0, sizeof(double) * S);
    dormand_prince_step<double>(geom, consts64, tab_a, tab_e, y, h, k, ys, y_new, err);
    std::memcpy(f_new, k[6], sizeof(double) * S);
}

int kro_classify(const double* consts64, const double* params, const int64_t* ipar, const double* y,
                 const double* f, double lam, int64_t n_steps, double h, int check_h) {
    return classify<double>(consts64, params, ipar, y, f, lam, n_steps, h, check_h != 0);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Forward-mode variational integration (checker of the discrete adjoint).
//
// For an escaping photon the oracle first integrates exactly as above and
// keeps the accepted step sizes h_k.  It then integrates the variational
// equations along the same steps: the Dormand-Prince 5(4) update applied to
// the augmented state (y, dy/dx), which for an explicit Runge-Kutta method
// with the steps held fixed is the derivative of the discrete map
// y_{k+1} = Phi(y_k; h_k).  Here every quantity carries its derivatives with
// respect to x = (alpha, beta, spin, theta_o) (forward-mode dual numbers),
// from the Bardeen initial state (initial_conditions.py:50-112 with
// Theta evaluated in full) through the inverse metric with the spin a
// variable (numba_kernel.py:111-154) and the Hamiltonian right-hand side
// (numba_kernel.py:158-197) to the final (theta, phi).
namespace {

constexpr int NX = 4;

struct Dx {
    double v;
    double d[NX];
};

inline Dx dx_const(double v) {
    Dx z{v, {0.0, 0.0, 0.0, 0.0}};
    return z;
}
inline Dx dx_var(double v, int k) {
    Dx z = dx_const(v);
    z.d[k] = 1.0;
    return z;
}
inline Dx operator+(Dx x, const Dx& y) {
    x.v += y.v;
    for (int i = 0; i < NX; ++i) x.d[i] += y.d[i];
    return x;
}
inline Dx operator-(Dx x, const Dx& y) {
    x.v -= y.v;
    for (int i = 0; i < NX; ++i) x.d[i] -= y.d[i];
    return x;
}
inline Dx operator-(Dx x) {
    x.v = -x.v;
    for (int i = 0; i < NX; ++i) x.d[i] = -x.d[i];
    return x;
}
inline Dx operator*(const Dx& x, const Dx& y) {
    Dx z;
    z.v = x.v * y.v;
    for (int i = 0; i < NX; ++i) z.d[i] = x.d[i] * y.v + x.v * y.d[i];
    return z;
}
inline Dx operator*(double c, Dx x) {
    x.v *= c;
    for (int i = 0; i < NX; ++i) x.d[i] *= c;
    return x;
}
inline Dx operator/(const Dx& x, const Dx& y) {
    Dx z;
    z.v = x.v / y.v;
    for (int i = 0; i < NX; ++i) z.d[i] = (x.d[i] * y.v - x.v * y.d[i]) / (y.v * y.v);
    return z;
}
inline Dx dx_sin(const Dx& x) {
    Dx z;
    z.v = std::sin(x.v);
    const double c = std::cos(x.v);
    for (int i = 0; i < NX; ++i) z.d[i] = c * x.d[i];
    return z;
}
inline Dx dx_cos(const Dx& x) {
    Dx z;
    z.v = std::cos(x.v);
    const double s = std::sin(x.v);
    for (int i = 0; i < NX; ++i) z.d[i] = -s * x.d[i];
    return z;
}
inline Dx dx_sqrt(const Dx& x) {
    Dx z;
    z.v = std::sqrt(x.v);
    for (int i = 0; i < NX; ++i) z.d[i] = x.d[i] / (2.0 * z.v);
    return z;
}

// numba_kernel.py:111-154 with m fixed and a = spin * m carried as a dual.
void metric_dx(double m, const Dx& a, const Dx& r, const Dx& th, Dx* g) {
    const Dx two_m = dx_const(2.0 * m);
    const Dx a2 = a * a;
    const Dx s = dx_sin(th), c = dx_cos(th);
    const Dx s2 = s * s;
    const Dx sin2th = dx_sin(2.0 * th);
    const Dx r2 = r * r;
    const Dx sig = r2 + a2 * (c * c);
    const Dx dlt = r2 - two_m * r + a2;
    const Dx r2a2 = r2 + a2;
    const Dx big_a = r2a2 * r2a2 - a2 * dlt * s2;
    const Dx sd = sig * dlt;
    const Dx neg_2ma = -(two_m * a), pos_2ma = two_m * a;
    g[0] = -big_a / sd;
    g[1] = neg_2ma * r / sd;
    g[2] = dlt / sig;
    g[3] = dx_const(1.0) / sig;
    g[4] = (dlt - a2 * s2) / (sd * s2);
    const Dx sig_r = 2.0 * r;
    const Dx sig_th = -a2 * sin2th;
    const Dx dlt_r = 2.0 * (r - dx_const(m));
    const Dx a_r = 4.0 * r * r2a2 - 2.0 * a2 * (r - dx_const(m)) * s2;
    const Dx a_th = -a2 * dlt * sin2th;
    const Dx sd_r = sig_r * dlt + sig * dlt_r;
    const Dx sd_th = sig_th * dlt;
    const Dx sd2 = sd * sd;
    const Dx sig2 = sig * sig;
    g[5] = -(a_r * sd - big_a * sd_r) / sd2;
    g[6] = neg_2ma * (sd - r * sd_r) / sd2;
    g[7] = (dlt_r * sig - dlt * sig_r) / sig2;
    g[8] = -sig_r / sig2;
    g[9] = -sig_r / (sig2 * s2) + a2 * sd_r / sd2;
    g[10] = -(a_th * sd - big_a * sd_th) / sd2;
    g[11] = pos_2ma * r * sd_th / sd2;
    g[12] = -dlt * sig_th / sig2;
    g[13] = -sig_th / sig2;
    g[14] = -(sig_th * s2 + sig * sin2th) / (sig2 * (s2 * s2)) + a2 * sd_th / sd2;
}

// numba_kernel.py:158-197 on the full 8-component state.
void rhs_dx(double m, const Dx& a, const Dx* y, Dx* f) {
    Dx g[15];
    metric_dx(m, a, y[IDX_R], y[IDX_TH], g);
    const Dx p_t = y[IDX_PT], p_r = y[IDX_PR], p_th = y[IDX_PTH], p_ph = y[IDX_PPH];
    Dx gphph = g[4], r_gphph = g[9], t_gphph = g[14];
    if (p_ph.v == 0.0) gphph = r_gphph = t_gphph = dx_const(0.0);
    f[IDX_T] = g[0] * p_t + g[1] * p_ph;
    f[IDX_R] = g[2] * p_r;
    f[IDX_TH] = g[3] * p_th;
    f[IDX_PH] = g[1] * p_t + gphph * p_ph;
    f[IDX_PT] = dx_const(0.0);
    f[IDX_PR] = -0.5 * (g[5] * p_t * p_t + 2.0 * g[6] * p_t * p_ph + g[7] * p_r * p_r + g[8] * p_th * p_th +
                        r_gphph * p_ph * p_ph);
    f[IDX_PTH] = -0.5 * (g[10] * p_t * p_t + 2.0 * g[11] * p_t * p_ph + g[12] * p_r * p_r + g[13] * p_th * p_th +
                         t_gphph * p_ph * p_ph);
    f[IDX_PPH] = dx_const(0.0);
}

// photon_from_constants / carter_potentials (initial_conditions.py:50-112)
// for the Bardeen constants of (alpha, beta) (rayset.py:158-165), E = 1.
void bardeen_dx(double m, double r0, const Dx& alpha, const Dx& beta, const Dx& a, const Dx& th_o, Dx* y) {
    const Dx sin_o = dx_sin(th_o), cos_o = dx_cos(th_o);
    const Dx xi = -(alpha * sin_o);
    const Dx cot = cos_o / sin_o;
    const Dx eta = beta * beta - a * a * cos_o * cos_o + xi * xi * cot * cot;
    const Dx rr = dx_const(r0 * r0);
    const Dx dlt = rr - dx_const(2.0 * m * r0) + a * a;
    const Dx term1 = (rr + a * a) - a * xi;
    const Dx d = xi - a;
    const Dx big_r = term1 * term1 - dlt * (eta + d * d);
    const Dx big_theta = eta - cos_o * cos_o * (xi * xi / (sin_o * sin_o) - a * a);
    const double sign_theta = beta.v >= 0.0 ? -1.0 : 1.0;
    y[IDX_T] = dx_const(0.0);
    y[IDX_R] = dx_const(r0);
    y[IDX_TH] = th_o;
    y[IDX_PH] = dx_const(0.0);
    y[IDX_PT] = dx_const(-1.0);
    y[IDX_PR] = -(dx_sqrt(big_r) / dlt);
    y[IDX_PTH] = sign_theta * dx_sqrt(big_theta);
    y[IDX_PPH] = xi;
}

// One Dormand-Prince 5(4) step (numba_kernel.py:201-238, y_new only) on
// dual states with the tableau A (dense 7 x 7) and a fixed step h.
void dp5_dx(double m, const Dx& a, const double* A, const Dx* y, double h, Dx* y_new) {
    Dx k[6][S], ys[S];
    rhs_dx(m, a, y, k[0]);
    for (int st = 1; st < 6; ++st) {
        for (int i = 0; i < S; ++i) {
            Dx acc = dx_const(0.0);
            for (int j = 0; j < st; ++j) acc = acc + A[st * 7 + j] * k[j][i];
            ys[i] = y[i] + h * acc;
        }
        rhs_dx(m, a, ys, k[st]);
    }
    for (int i = 0; i < S; ++i) {
        Dx acc = dx_const(0.0);
        for (int j = 0; j < 6; ++j) acc = acc + A[6 * 7 + j] * k[j][i];
        y_new[i] = y[i] + h * acc;
    }
}


// Escape direction at r = infinity: (theta, phi, p_theta) integrated from
// the final state in u = 1/r down to u = 0 (classical RK4, 16 steps), with
// the geodesic equations divided by du/dlambda and written in
// Sigma^ = u^2 Sigma, Delta^ = u^2 Delta, A^ = u^4 A so that they stay finite
// at u = 0; p_r > 0 from the null condition.
constexpr int TAIL_STEPS = 16;

void tail_rhs_dx(double m, const Dx& a, const Dx& u, const Dx& th, const Dx& pth, const Dx& pt, const Dx& pph,
                 Dx* d) {
    const Dx one = dx_const(1.0);
    const Dx s = dx_sin(th), c = dx_cos(th);
    const Dx s2 = s * s, c2 = c * c;
    const Dx a2 = a * a, u2 = u * u;
    const Dx sig = one + a2 * u2 * c2;
    const Dx dlt = one - 2.0 * m * u + a2 * u2;
    const Dx big_a = (one + a2 * u2) * (one + a2 * u2) - a2 * u2 * dlt * s2;
    const Dx gph = dlt - a2 * u2 * s2;
    const Dx pr2 = (big_a * pt * pt + 4.0 * m * a * u * u2 * pt * pph - u2 * dlt * pth * pth - u2 * gph * pph * pph / s2) /
                   (dlt * dlt);
    const Dx pr = dx_sqrt(pr2);
    d[0] = -(pth / (dlt * pr));
    d[1] = (2.0 * m * a * u * pt - gph * pph / s2) / (dlt * dlt * pr);
    const Dx sc = s * c, sig2 = sig * sig;
    const Dx t1 = 2.0 * a2 * sc * (dlt * sig - big_a) / (sig2 * dlt);
    const Dx t2 = -(4.0 * m * a * a2 * u * u2 * sc) / (sig2 * dlt);
    const Dx t3 = 2.0 * a2 * sc * dlt / sig2;
    const Dx t4 = 2.0 * a2 * u2 * sc / sig2;
    const Dx t5 = 2.0 * sc * (a2 * u2 * s2 - sig) / (sig2 * s2 * s2) - 2.0 * a2 * a2 * u2 * u2 * sc / (sig2 * dlt);
    d[2] = 0.5 * sig * (t1 * pt * pt + 2.0 * t2 * pt * pph + t3 * pr2 + t4 * pth * pth + t5 * pph * pph) / (dlt * pr);
}

void escape_tail_dx(double m, const Dx& a, const Dx* y, Dx& th_inf, Dx& ph_inf) {
    Dx u = dx_const(1.0) / y[IDX_R];
    const Dx h = (-1.0 / TAIL_STEPS) * u;
    Dx z[3] = {y[IDX_TH], y[IDX_PH], y[IDX_PTH]};
    const Dx pt = y[IDX_PT], pph = y[IDX_PPH];
    for (int k = 0; k < TAIL_STEPS; ++k) {
        Dx k1[3], k2[3], k3[3], k4[3], zz[3];
        tail_rhs_dx(m, a, u, z[0], z[2], pt, pph, k1);
        for (int i = 0; i < 3; ++i) zz[i] = z[i] + 0.5 * (h * k1[i]);
        tail_rhs_dx(m, a, u + 0.5 * h, zz[0], zz[2], pt, pph, k2);
        for (int i = 0; i < 3; ++i) zz[i] = z[i] + 0.5 * (h * k2[i]);
        tail_rhs_dx(m, a, u + 0.5 * h, zz[0], zz[2], pt, pph, k3);
        for (int i = 0; i < 3; ++i) zz[i] = z[i] + h * k3[i];
        const Dx u_next = (k == TAIL_STEPS - 1) ? dx_const(0.0) : u + h;
        tail_rhs_dx(m, a, u_next, zz[0], zz[2], pt, pph, k4);
        for (int i = 0; i < 3; ++i) z[i] = z[i] + (1.0 / 6.0) * (h * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]));
        u = u_next;
    }
    th_inf = z[0];
    ph_inf = z[1];
}

}  // namespace

extern "C" {

// Escape directions at infinity of the points (alpha_i, beta_i) of one
// image and their derivatives.  geom..tab_e are the packed kernel constants
// (FP64 DP5, no events).  Outputs: state[n]; theta[n], phi[n], the escape
// direction at r = infinity of the exact integration's final state (NaN
// unless escaped); jac[8 n] = (d theta / d x, d phi / d x) for
// x = (alpha, beta, spin, theta_o), NaN unless escaped.
// replay_err[n] (may be NULL): max |final state of the variational pass -
// exact final state| over the integrated components, a consistency check.
int kro_escape_variational(double mass, double spin, double r0, double theta_o, const double* alpha,
                           const double* beta, int64_t n, const double* geom, const double* consts,
                           const double* params, const int64_t* ipar, const double* tab_a, const double* tab_e,
                           const double* rk4_b, const double* rk4_c, int64_t* state, double* theta, double* phi,
                           double* jac, double* replay_err, int n_threads) {
    if (ipar[I_ADAPTIVE] != 1 || ipar[I_DISK] != 0) return 2;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    const double a_val = spin * mass;
    int bad = 0;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 4) reduction(| : bad)
    for (int64_t i = 0; i < n; ++i) {
        double y[S], f[S];
        const int sign = beta[i] >= 0.0 ? -1 : 1;
        if (kro_bardeen_state(mass, a_val, r0, theta_o, alpha[i], beta[i], sign, y)) bad = 1;
        double lam, mn, me, ml, mq, one = 1.0, zero = 0.0;
        int64_t ns, nr, st;
        uint8_t hit = 0;
        std::vector<double> steps;
        RayArgs args{geom, params, ipar, &lam, &ns, &nr, &st, &mn, &me, &ml, &mq,
                     &one, &zero, &zero, &one, &one, &one, &one, &hit, &steps};
        double cst[N_CONSTS];
        for (int c = 0; c < N_CONSTS; ++c) cst[c] = consts[c];
        integrate_ray<double>(0, args, cst, tab_a, tab_e, rk4_b, rk4_c, y, f, nullptr);
        state[i] = st;
        theta[i] = NAN;
        phi[i] = NAN;
        for (int c = 0; c < 8; ++c) jac[i * 8 + c] = NAN;
        if (replay_err) replay_err[i] = NAN;
        if (st != ESCAPED) continue;
        Dx yd[S], yn[S];
        const Dx ad = mass * dx_var(spin, 2);
        bardeen_dx(mass, r0, dx_var(alpha[i], 0), dx_var(beta[i], 1), ad, dx_var(theta_o, 3), yd);
        for (double h : steps) {
            dp5_dx(mass, ad, tab_a, yd, h, yn);
            for (int c = 0; c < S; ++c) yd[c] = yn[c];
        }
        Dx yexact[S], th_inf, ph_inf;
        for (int c = 0; c < S; ++c) yexact[c] = yd[c];
        escape_tail_dx(mass, ad, yexact, th_inf, ph_inf);
        // The values come from the exact final state (the dual pass's values
        // differ from it by rounding only).
        Dx yv[S], thv, phv;
        for (int c = 0; c < S; ++c) yv[c] = dx_const(y[c]);
        escape_tail_dx(mass, dx_const(a_val), yv, thv, phv);
        theta[i] = thv.v;
        phi[i] = phv.v;
        for (int j = 0; j < NX; ++j) {
            jac[i * 8 + j] = th_inf.d[j];
            jac[i * 8 + 4 + j] = ph_inf.d[j];
        }
        if (replay_err) {
            double e = 0.0;
            for (int c = 0; c < S; ++c) e = pymax(e, std::fabs(yd[c].v - y[c]) / pymax(std::fabs(y[c]), 1.0));
            replay_err[i] = e;
        }
    }
    return bad;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ZAMO camera (src/kerrray/raytracing/camera.py:225-381), restated with the
// full 4 x 4 matrices of the reference:
//   metric_components / inverse_metric_components  geometry/metric.py:147-195
//   _assemble                                      geometry/metric.py:198-209
//   zamo_tetrad                                    camera.py:225-262
//   _null_pt                                       camera.py:265-295
//   initial_states                                 camera.py:298-381
// The two matrix products p_local @ e and p_up @ g are evaluated as the BLAS
// double kernel NumPy calls evaluates them: for each output component, an
// accumulator starting at zero receives fma(x_k, m_kj, acc) for k = 0..3 in
// order.  Per-image quantities follow NumPy's float64 rules: x ** 2 of an
// array is x * x; a ** 2 of the Python float a is glibc pow(a, 2.0).
namespace {

struct ZamoFrame {
    double e[4][4], g[4][4], gi[4][4];
};

// Returns 0, or 1 where zamo_tetrad raises.
int zamo_frame(double m, double a, double r, double th, ZamoFrame* F) {
    std::memset(F, 0, sizeof(*F));
    const double s2 = std::sin(th) * std::sin(th);
    const double c = std::cos(th);
    const double a_sq = std::pow(a, 2.0);
    const double sig = r * r + a_sq * (c * c);
    const double dlt = r * r - 2.0 * m * r + a_sq;
    const double g_tt = -(1.0 - 2.0 * m * r / sig);
    const double g_tphi = -2.0 * m * a * r * s2 / sig;
    const double g_rr = sig / dlt;
    const double g_thth = sig + 0.0 * dlt;
    const double g_phph = (r * r + a_sq + 2.0 * m * a_sq * r * s2 / sig) * s2;
    const double big_a = (r * r + a_sq) * (r * r + a_sq) - a_sq * dlt * s2;
    const double sd = sig * dlt;
    const double gtt = -big_a / sd, gtphi = -2.0 * m * a * r / sd, grr = dlt / sig, gthth = 1.0 / sig + 0.0 * dlt,
                 gphph = (dlt - a_sq * s2) / (sd * s2);
    const double comps[2][5] = {{g_tt, g_tphi, g_rr, g_thth, g_phph}, {gtt, gtphi, grr, gthth, gphph}};
    double(*mats[2])[4] = {F->g, F->gi};
    for (int k = 0; k < 2; ++k) {
        mats[k][0][0] = comps[k][0];
        mats[k][0][3] = mats[k][3][0] = comps[k][1];
        mats[k][1][1] = comps[k][2];
        mats[k][2][2] = comps[k][3];
        mats[k][3][3] = comps[k][4];
    }
    const double omega = -g_tphi / g_phph;
    const double lapse2 = -(g_tt - g_tphi * g_tphi / g_phph);
    if (!(lapse2 > 0.0)) return 1;
    const double lapse = std::sqrt(lapse2);
    F->e[0][0] = 1.0 / lapse;
    F->e[0][3] = omega / lapse;
    F->e[1][1] = 1.0 / std::sqrt(g_rr);
    F->e[2][2] = 1.0 / std::sqrt(g_thth);
    F->e[3][3] = 1.0 / std::sqrt(g_phph);
    return 0;
}

// out_j = sum_k x_k M_kj, accumulated as BLAS does (see above).
void row_times(const double* x, const double (*M)[4], double* out) {
    for (int j = 0; j < 4; ++j) {
        double acc = 0.0;
        for (int k = 0; k < 4; ++k) acc = std::fma(x[k], M[k][j], acc);
        out[j] = acc;
    }
}

}  // namespace

extern "C" {

// The ZAMO frame at (r, theta) as 4 x 4 row-major arrays; 0, or 1 where the
// tetrad does not exist.
int kro_zamo_frame(double mass, double spin, double r, double theta, double* tetrad, double* metric,
                   double* inverse_metric) {
    ZamoFrame F;
    const int bad = zamo_frame(mass, spin * mass, r, theta, &F);
    std::memcpy(tetrad, F.e, sizeof(F.e));
    std::memcpy(metric, F.g, sizeof(F.g));
    std::memcpy(inverse_metric, F.gi, sizeof(F.gi));
    return bad;
}

// initial_states(cam, st, reverse) for n image points.  Returns a bit set:
// 1 the frame does not exist, 2 a projection is not null to 1e-10 or no null
// p_t exists, 4 a point has alpha^2 + beta^2 >= r_o^2.
int kro_zamo_states(double mass, double spin, double r_o, double theta_o, double phi_o, const double* alpha,
                    const double* beta, int64_t n, int reverse, double* Y) {
    ZamoFrame F;
    int bad = zamo_frame(mass, spin * mass, r_o, theta_o, &F);
    const double r_o_sq = std::pow(r_o, 2.0);
    for (int64_t i = 0; i < n; ++i) {
        const double rho2 = (alpha[i] * alpha[i] + beta[i] * beta[i]) / r_o_sq;
        if (!(rho2 < 1.0)) bad |= 4;
        const double local[4] = {1.0, std::sqrt(1.0 - rho2), beta[i] / r_o, -alpha[i] / r_o};
        double up[4], p[4];
        row_times(local, F.e, up);
        row_times(up, F.g, p);
        const double residual = F.gi[0][0] * (p[0] * p[0]) + 2.0 * F.gi[0][3] * p[0] * p[3] +
                                F.gi[1][1] * (p[1] * p[1]) + F.gi[2][2] * (p[2] * p[2]) + F.gi[3][3] * (p[3] * p[3]);
        if (!(std::fabs(residual) < 1e-10)) bad |= 2;
        if (reverse)
            for (double& v : p) v = -v;
        // _null_pt
        const double qa = F.gi[0][0];
        const double qb = 2.0 * F.gi[0][3] * p[3];
        const double qc = F.gi[1][1] * (p[1] * p[1]) + F.gi[2][2] * (p[2] * p[2]) + F.gi[3][3] * (p[3] * p[3]);
        if (!(qa < 0.0)) bad |= 2;
        const double disc = qb * qb - 4.0 * qa * qc;
        if (disc < -1e-10 * (qb * qb + std::fabs(4.0 * qa * qc))) bad |= 2;
        const double q = -0.5 * (qb + std::copysign(std::sqrt(pymax(disc, 0.0)), qb));
        const double root_1 = q / qa;
        const double root_2 = q != 0.0 ? qc / q : root_1;
        double* y = Y + i * S;
        y[IDX_T] = 0.0;
        y[IDX_R] = r_o;
        y[IDX_TH] = theta_o;
        y[IDX_PH] = phi_o;
        y[IDX_PT] = std::fabs(root_1 - p[0]) <= std::fabs(root_2 - p[0]) ? root_1 : root_2;
        y[IDX_PR] = p[1];
        y[IDX_PTH] = p[2];
        y[IDX_PPH] = p[3];
    }
    return bad;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Disk shading (render_disk's post-processing, renderer.py:190-205):
//   keplerian_angular_velocity  photons/orbits.py:281-292
//   emitter_four_velocity       physics/redshift.py:63-90 (metric at theta = pi/2)
//   _frequency_ratio            physics/redshift.py:117-128 (sum over the four
//                               components of p_mu u^mu, in index order)
//   emissivity, observed_intensity  physics/accretion.py:180-196
// Returns 0, or 1 where the reference raises ValueError.
extern "C" int kro_disk_shade(double mass, double spin, double r_in, double r_out, double p_index, int n_pow,
                              int prograde, const double* Y, int64_t n, double* g_out, double* intensity_out) {
    const double a = spin * mass;
    const int s = prograde ? 1 : -1;
    const int sign_a = a < 0.0 ? -1 : 1;
    const double th = 0.5 * M_PI;
    const double s2 = std::sin(th) * std::sin(th), c2 = std::cos(th) * std::cos(th);
    const double a_sq = std::pow(a, 2.0);
    int bad = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double* y = Y + i * S;
        for (int c = 0; c < S; ++c)
            if (!std::isfinite(y[c])) bad = 1;
        if (std::fabs(y[IDX_TH] - 0.5 * M_PI) > 1e-6) bad = 1;
        const double r = y[IDX_R];
        const double omega = (s * sign_a) * std::sqrt(mass) / (std::pow(r, 1.5) + s * std::fabs(a) * std::sqrt(mass));
        const double sig = r * r + a_sq * c2;
        const double g_tt = -(1.0 - 2.0 * mass * r / sig);
        const double g_tphi = -2.0 * mass * a * r * s2 / sig;
        const double g_phph = (r * r + a_sq + 2.0 * mass * a_sq * r * s2 / sig) * s2;
        const double radicand = -(g_tt + 2.0 * omega * g_tphi + omega * omega * g_phph);
        if (!std::isfinite(radicand) || radicand <= 0.0) bad = 1;
        const double u_t = 1.0 / std::sqrt(radicand);
        const double u[4] = {u_t, 0.0, 0.0, omega * u_t};
        double p_dot_u = 0.0;
        for (int k = 0; k < 4; ++k) p_dot_u += y[IDX_PT + k] * u[k];
        const double g = y[IDX_PT] / p_dot_u;
        if (!std::isfinite(g) || g <= 0.0) bad = 1;
        const double emitted = (r >= r_in && r <= r_out) ? std::pow(r / r_in, -p_index) : 0.0;
        g_out[i] = g;
        intensity_out[i] = emitted * std::pow(g, static_cast<double>(n_pow));
    }
    return bad;
}
