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
//
// segment 0 (Jacobian tape): the escape direction has two components, so
// every photon's whole Jacobian d(theta_inf, phi_inf)/d(alpha, beta, spin,
// theta_o) is accumulated once, when the tape is built:
//   1. the recording kernel stores every accepted step's six stage states;
//   2. a kernel with one lane per (photon, step) turns them into the step's
//      increment D (kr_adjoint.h), all steps of all photons at once;
//   3. per photon, both rows of the escape tail's Jacobian are carried back
//      through the D_k^T and the initial state.
// A vector-Jacobian product is then J^T g per photon.  Recorded photons are
// sorted by step count and taken 64 to a block; a block's step buffers hold
// (its longest photon's steps) units of 64 slots, entry e of lane l in unit u
// at [(u * entries + e) * 64 + l], so every access of a wavefront is one
// contiguous 512-byte line.  HBM: 632 B per accepted step while the tape is
// built, 64 B per photon kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
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

// The Jacobian tape (segment 0) while it is built.
struct JacTapeDev {
    const double* alpha;
    const double* beta;
    const int64_t* list;        // recorded photon -> point index, most steps first
    const int64_t* steps;       // accepted steps of each recorded photon
    const int64_t* block_unit;  // first unit of each block of 64 recorded photons
    const int32_t* unit_block;  // block of each unit
    double* stg;                // 6 stage states (6 * NA entries) per slot, while the tape is built
    double* hs;                 // step size per slot, while the tape is built
    double* D;                  // JW entries per slot
    double* pt;                 // p_t and p_phi of each recorded photon
    double* pph;
    int64_t n_rec, n_units;
};

constexpr int STG = 6 * NA;  // stage-state entries per slot

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

// Stage states of one attempt into the slot of the step it becomes if
// accepted (a rejected attempt's are overwritten by the next attempt's).
struct TapeStages {
    double* slot;  // entry e at slot[e * 64]
    bool on;
    KR_HD void operator()(int s, const double (&ys)[NA]) const {
        if (!on) return;
#pragma unroll
        for (int c = 0; c < NA; ++c) slot[(s * NA + c) * 64] = ys[c];
    }
};

// Re-integrates the recorded photons of one block (one per lane) exactly as
// trace_kernel does and stores every accepted step's stage states and size.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void record_stages_kernel(
    Params P, BardeenImage B, JacTapeDev T, const double* Yfinal, double* theta_inf, double* phi_inf,
    int32_t* flags) {
    const int lane = threadIdx.x;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 64 + lane;
    if (t >= T.n_rec) return;
    const int64_t i = T.list[t];
    double y0[8];
    bardeen_state(B, T.alpha[i], T.beta[i], y0);
    T.pt[t] = y0[4];
    T.pph[t] = y0[7];
    RayState<double> R;
    ray_init<double, true>(P, y0, R);
    const int64_t cap = T.steps[t];
    const int64_t u0 = T.block_unit[blockIdx.x];
    const double lambda_max = P.params[P_LAMMAX];
    while (R.state == RUNNING) {
        const double remaining = lambda_max - R.lam;
        const double h_eff = (R.h >= remaining) ? remaining : R.h;
        const int64_t k = R.n_steps;
        const bool on = k < cap;
        const TapeStages sink{T.stg + ((u0 + (on ? k : 0)) * STG) * 64 + lane, on};
        ray_attempt_with<double, true, false>(P, R, ad::StageStepper<TapeStages>{sink});
        if (R.n_steps != k && on) T.hs[(u0 + k) * 64 + lane] = h_eff;
    }
    bool same = R.n_steps == cap && R.state == ESCAPED;
#pragma unroll
    for (int c = 0; c < NA; ++c) same = same && (R.y[c] == Yfinal[i * 8 + full_index(c)]);
    if (!same) atomicOr(flags, TAPE_FLAG_REPLAY);
    double th, ph;
    ad::escape_tail(R.y[1], R.y[2], R.y[3], R.y[5], static_cast<double>(R.p_t), static_cast<double>(R.p_ph),
                    P.geom[G_A], P.geom[G_MASS], th, ph);
    theta_inf[i] = th;
    phi_inf[i] = ph;
}

// Columns [C0, C0 + NC) of the increment D of every recorded step, one lane
// per (photon, step), from the step's stage states.
template <int C0, int NC>
__global__ __launch_bounds__(256) void jacobian_kernel(double m, double a, JacTapeDev T) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t u = g >> 6;
    const int lane = static_cast<int>(g & 63);
    if (u >= T.n_units) return;
    const int64_t b = T.unit_block[u];
    const int64_t k = u - T.block_unit[b];
    const int64_t t = b * 64 + lane;
    if (t >= T.n_rec || k >= T.steps[t]) return;
    const double* src = T.stg + (u * STG) * 64 + lane;
    double* dst = T.D + (u * ad::JW) * 64 + lane;
    auto stage = [&](int s, double (&Y)[NA]) {
#pragma unroll
        for (int c = 0; c < NA; ++c) Y[c] = src[(s * NA + c) * 64];
    };
    auto out = [&](int r, int c, double v) { dst[(r * ad::JC + c) * 64] = v; };
    ad::step_increment<C0, NC>(m, a, T.pt[t], T.pph[t], T.hs[u * 64 + lane], stage, out);
}

// d/d(alpha, beta, spin, theta_o) from the cotangent of the initial state
// y_0(alpha, beta, a, theta_o) (and of a), into g[0..3].
KR_HD void initial_state_gradient(const BardeenImage& B, double mass, double a, double alpha, double beta,
                                  const double (&ybar)[NA], double ptbar, double pphbar, double abar, double* g) {
    ad::Dual<4> yd[8];
    ad::bardeen_state_of<4>(ad::seed<4>(alpha, 0), ad::seed<4>(beta, 1), ad::seed<4>(a, 2),
                            ad::seed<4>(B.theta_o, 3), mass, B.r_o, yd);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        double v = ptbar * yd[4].d[j] + pphbar * yd[7].d[j];
#pragma unroll
        for (int c = 0; c < NA; ++c) v += ybar[c] * yd[full_index(c)].d[j];
        if (j == 2) v = (v + abar) * mass;  // d/d(spin) = M d/da
        g[j] = v;
    }
}

// Jacobian of each recorded photon's escape direction with respect to its
// final state and p_t, p_phi, a: cot[(o * (NA + 3) + k) * n_rec + t] for
// output o (theta_inf, phi_inf).
__global__ __launch_bounds__(64) void tail_jacobian_kernel(double mass, double spin, JacTapeDev T,
                                                           const double* Yfinal, double* cot) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (t >= T.n_rec) return;
    const int64_t i = T.list[t];
    double yN[NA], J[2][NA + 3];
#pragma unroll
    for (int c = 0; c < NA; ++c) yN[c] = Yfinal[i * 8 + full_index(c)];
    ad::escape_tail_jacobian(mass, spin * mass, yN, T.pt[t], T.pph[t], J);
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int k = 0; k < NA + 3; ++k) cot[(o * (NA + 3) + k) * T.n_rec + t] = J[o][k];
}

// d(theta_inf, phi_inf)/d(alpha, beta, spin, theta_o) of every recorded
// photon: both rows of the tail Jacobian carried back through the step
// increments (D_k^T, last step first) and then through the initial state.
// One lane per recorded photon, the blocks of the tape.
__global__ __launch_bounds__(64) void photon_jacobian_kernel(BardeenImage B, double mass, double spin, JacTapeDev T,
                                                             const double* cot, double* Jp) {
    const int lane = threadIdx.x;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 64 + lane;
    if (t >= T.n_rec) return;
    const int64_t i = T.list[t];
    const double a = spin * mass;
    double ybar[2][NA], ptbar[2], pphbar[2], abar[2];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        const double* c = cot + o * (NA + 3) * T.n_rec + t;
#pragma unroll
        for (int k = 0; k < NA; ++k) ybar[o][k] = c[k * T.n_rec];
        ptbar[o] = c[NA * T.n_rec];
        pphbar[o] = c[(NA + 1) * T.n_rec];
        abar[o] = c[(NA + 2) * T.n_rec];
    }
    // Double-buffered and branch-free: step k - 1's increment is loading
    // while step k's is applied.  Every load is issued (at a clamped, valid
    // step) and the update of a step past the end is masked, so the loop has
    // no branch and its wait can leave the newer 42 loads in flight (the
    // counter of outstanding loads holds at most 63).
    const double* d = T.D + (T.block_unit[blockIdx.x] * ad::JW) * 64 + lane;
    constexpr int64_t STEP = ad::JW * 64;
    const int64_t n = T.steps[t];
    double ring[2][ad::JW];
    const int64_t k0 = n > 0 ? n - 1 : 0;
#pragma unroll
    for (int e = 0; e < ad::JW; ++e) ring[0][e] = d[k0 * STEP + e * 64];
    for (int64_t k = n - 1; k >= 0; k -= 2) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t kk = k - j;
            const int64_t next = kk - 1 > 0 ? kk - 1 : 0;
#pragma unroll
            for (int e = 0; e < ad::JW; ++e) ring[1 - j][e] = d[next * STEP + e * 64];
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of this step's arithmetic
            const bool live = kk >= 0;
#pragma unroll
            for (int o = 0; o < 2; ++o)
                ad::increment_vjp([&](int e) { return live ? ring[j][e] : 0.0; }, ybar[o], ptbar[o], pphbar[o],
                                  abar[o]);
        }
    }
#pragma unroll
    for (int o = 0; o < 2; ++o)
        initial_state_gradient(B, mass, a, T.alpha[i], T.beta[i], ybar[o], ptbar[o], pphbar[o], abar[o],
                               Jp + (t * 2 + o) * 4);
}

// Vector-Jacobian product from the photon Jacobians: grad[i] =
// g_theta[i] J[0] + g_phi[i] J[1] for recorded photon t = point i.
__global__ void contract_kernel(const int64_t* list, const double* Jp, int64_t n_rec, const double* g_theta,
                                const double* g_phi, double* grad) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_rec) return;
    const int64_t i = list[t];
    const double gt = g_theta[i], gp = g_phi[i];
#pragma unroll
    for (int j = 0; j < 4; ++j) grad[i * 4 + j] = gt * Jp[(t * 2) * 4 + j] + gp * Jp[(t * 2 + 1) * 4 + j];
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
    initial_state_gradient(B, mass, a, alpha, beta, ybar, ptbar, pphbar, abar, grad + i * 4);
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
    double record_ms = 0.0, jacobian_ms = 0.0;
    // The Jacobian tape (segment 0): Jp kept, the rest while building.
    int64_t n_units = 0;
    int64_t *steps = nullptr, *block_unit = nullptr;
    int32_t* unit_block = nullptr;
    double *stg = nullptr, *D = nullptr, *pt = nullptr, *pph = nullptr, *cot = nullptr, *Jp = nullptr;

    TapeDev dev() const { return TapeDev{alpha, beta, list, h_off, c_off, hs, ckpt, n_rec, segment}; }
    JacTapeDev jdev() const {
        return JacTapeDev{alpha, beta, list, steps, block_unit, unit_block, stg, hs, D, pt, pph, n_rec, n_units};
    }
};

namespace {

void free_tape(kr_tape_s* t) {
    if (!t) return;
    double* d[] = {t->alpha, t->beta, t->Y, t->hs, t->ckpt, t->seg, t->stg, t->D, t->pt, t->pph, t->cot, t->Jp};
    for (double* p : d)
        if (p) (void)hipFree(p);
    int64_t* q[] = {t->list, t->h_off, t->c_off, t->steps, t->block_unit};
    for (int64_t* p : q)
        if (p) (void)hipFree(p);
    if (t->unit_block) (void)hipFree(t->unit_block);
    if (t->flags) (void)hipFree(t->flags);
    delete t;
}

template <typename T>
kr_status dev_alloc(T** p, int64_t count) {
    KR_HIP(hipMalloc(reinterpret_cast<void**>(p), static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
    return KR_OK;
}

}  // namespace

namespace {

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    host::fail(KR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

// The Jacobian tape of the escaped photons `list` (step counts from the
// trace in `steps`, indexed by point).
kr_status create_jacobian_tape(kr_context* h, const Params& P, const BardeenImage& B, const kr_image* img, int64_t n,
                               std::vector<int64_t> list, const std::vector<int64_t>& steps, const double* d_alpha,
                               const double* d_beta, const kr_outputs* d_out, double* d_theta_inf,
                               double* d_phi_inf, kr_tape* out) {
    hipStream_t s = h->stream;
    std::stable_sort(list.begin(), list.end(), [&](int64_t x, int64_t y) { return steps[x] > steps[y]; });
    const int64_t n_rec = static_cast<int64_t>(list.size());
    const int64_t n_blocks = (n_rec + 63) / 64;
    std::vector<int64_t> rec_steps(static_cast<size_t>(n_rec)), block_unit(static_cast<size_t>(n_blocks + 1), 0);
    int64_t total = 0;
    for (int64_t t = 0; t < n_rec; ++t) {
        rec_steps[t] = steps[list[t]];
        total += rec_steps[t];
    }
    for (int64_t b = 0; b < n_blocks; ++b) block_unit[b + 1] = block_unit[b] + rec_steps[b * 64];
    const int64_t n_units = block_unit[n_blocks];
    std::vector<int32_t> unit_block(static_cast<size_t>(n_units));
    for (int64_t b = 0; b < n_blocks; ++b)
        for (int64_t u = block_unit[b]; u < block_unit[b + 1]; ++u) unit_block[u] = static_cast<int32_t>(b);
    // Device memory while building: 632 B per slot.
    const double need = static_cast<double>(n_units) * 64.0 * 8.0 * (ad::JW + STG + 1);
    size_t free_b = 0, total_b = 0;
    KR_HIP(hipMemGetInfo(&free_b, &total_b));
    if (need > 0.9 * static_cast<double>(free_b))
        return host::fail(KR_ERR_DEVICE, "building the Jacobian tape needs " + std::to_string(need / 1e9) + " GB of " +
                                             std::to_string(free_b / 1e9) + " GB free; use segment >= 1");

    kr_tape_s* T = new kr_tape_s();
    T->h = h;
    T->P = P;
    T->B = B;
    T->mass = img->mass;
    T->spin = img->spin;
    T->n = n;
    T->n_rec = n_rec;
    T->total_steps = total;
    T->segment = 0;
    T->n_units = n_units;
    auto bail = [&](kr_status e) {
        free_tape(T);
        return e;
    };
    kr_status st;
    if ((st = dev_alloc(&T->alpha, n)) != KR_OK || (st = dev_alloc(&T->beta, n)) != KR_OK ||
        (st = dev_alloc(&T->Y, n * 8)) != KR_OK || (st = dev_alloc(&T->list, n_rec)) != KR_OK ||
        (st = dev_alloc(&T->steps, n_rec)) != KR_OK || (st = dev_alloc(&T->block_unit, n_blocks + 1)) != KR_OK ||
        (st = dev_alloc(&T->unit_block, n_units)) != KR_OK || (st = dev_alloc(&T->pt, n_rec)) != KR_OK ||
        (st = dev_alloc(&T->pph, n_rec)) != KR_OK || (st = dev_alloc(&T->stg, n_units * 64 * STG)) != KR_OK ||
        (st = dev_alloc(&T->hs, n_units * 64)) != KR_OK || (st = dev_alloc(&T->D, n_units * 64 * ad::JW)) != KR_OK ||
        (st = dev_alloc(&T->cot, n_rec * 2 * (NA + 3))) != KR_OK || (st = dev_alloc(&T->Jp, n_rec * 8)) != KR_OK ||
        (st = dev_alloc(&T->flags, 1)) != KR_OK)
        return bail(st);
    if (n > 0) {
        if (!hip_ok(hipMemcpyAsync(T->alpha, d_alpha, n * 8, hipMemcpyDeviceToDevice, s), "copy alpha") ||
            !hip_ok(hipMemcpyAsync(T->beta, d_beta, n * 8, hipMemcpyDeviceToDevice, s), "copy beta") ||
            !hip_ok(hipMemcpyAsync(T->Y, d_out->Y, n * 64, hipMemcpyDeviceToDevice, s), "copy Y"))
            return bail(KR_ERR_DEVICE);
        const dim3 g(static_cast<unsigned>((n + 255) / 256));
        hipLaunchKernelGGL(fill_nan_kernel, g, dim3(256), 0, s, d_theta_inf, n);
        hipLaunchKernelGGL(fill_nan_kernel, g, dim3(256), 0, s, d_phi_inf, n);
        if (!hip_ok(hipGetLastError(), "fill directions")) return bail(KR_ERR_DEVICE);
    }
    if (n_rec > 0) {
        if (!hip_ok(hipMemcpyAsync(T->list, list.data(), n_rec * 8, hipMemcpyHostToDevice, s), "copy list") ||
            !hip_ok(hipMemcpyAsync(T->steps, rec_steps.data(), n_rec * 8, hipMemcpyHostToDevice, s), "copy steps") ||
            !hip_ok(hipMemcpyAsync(T->block_unit, block_unit.data(), (n_blocks + 1) * 8, hipMemcpyHostToDevice, s),
                    "copy block_unit") ||
            !hip_ok(hipMemcpyAsync(T->unit_block, unit_block.data(), n_units * 4, hipMemcpyHostToDevice, s),
                    "copy unit_block"))
            return bail(KR_ERR_DEVICE);
    }
    hipLaunchKernelGGL(fill_flag_kernel, dim3(1), dim3(64), 0, s, T->flags);
    if (!hip_ok(hipGetLastError(), "fill flags")) return bail(KR_ERR_DEVICE);

    // Record, then the step increments.
    const double a = img->spin * img->mass;
    if (!hip_ok(hipEventRecord(h->ev0, s), "event")) return bail(KR_ERR_DEVICE);
    if (n_rec > 0) {
        hipLaunchKernelGGL(record_stages_kernel, dim3(static_cast<unsigned>(n_blocks)), dim3(64), 0, s, P, B,
                           T->jdev(), T->Y, d_theta_inf, d_phi_inf, T->flags);
        if (!hip_ok(hipGetLastError(), "record_stages_kernel")) return bail(KR_ERR_DEVICE);
    }
    if (!hip_ok(hipEventRecord(h->ev1, s), "event")) return bail(KR_ERR_DEVICE);
    if (n_units > 0) {
        const dim3 g(static_cast<unsigned>((n_units * 64 + 255) / 256));
        // Three passes over the steps, each spill-free: the state columns,
        // then p_t and p_phi, then a (whose column alone needs the metric's
        // a-derivatives, a third of the dual arithmetic).
        hipLaunchKernelGGL((jacobian_kernel<0, 4>), g, dim3(256), 0, s, img->mass, a, T->jdev());
        hipLaunchKernelGGL((jacobian_kernel<4, 2>), g, dim3(256), 0, s, img->mass, a, T->jdev());
        hipLaunchKernelGGL((jacobian_kernel<6, 1>), g, dim3(256), 0, s, img->mass, a, T->jdev());
        if (!hip_ok(hipGetLastError(), "jacobian_kernel")) return bail(KR_ERR_DEVICE);
    }
    if (n_rec > 0) {
        const dim3 g(static_cast<unsigned>(n_blocks));
        hipLaunchKernelGGL(tail_jacobian_kernel, g, dim3(64), 0, s, img->mass, img->spin, T->jdev(), T->Y, T->cot);
        hipLaunchKernelGGL(photon_jacobian_kernel, g, dim3(64), 0, s, B, img->mass, img->spin, T->jdev(), T->cot,
                           T->Jp);
        if (!hip_ok(hipGetLastError(), "photon_jacobian_kernel")) return bail(KR_ERR_DEVICE);
    }
    hipEvent_t ev2 = nullptr;
    if (!hip_ok(hipEventCreate(&ev2), "event")) return bail(KR_ERR_DEVICE);
    const bool rec2 = hip_ok(hipEventRecord(ev2, s), "event") && hip_ok(hipStreamSynchronize(s), "record");
    float ms1 = 0.0f, ms2 = 0.0f;
    const bool el = rec2 && hip_ok(hipEventElapsedTime(&ms1, h->ev0, h->ev1), "elapsed") &&
                    hip_ok(hipEventElapsedTime(&ms2, h->ev1, ev2), "elapsed");
    (void)hipEventDestroy(ev2);
    if (!el) return bail(KR_ERR_DEVICE);
    T->record_ms = ms1;
    T->jacobian_ms = ms2;
    // Everything but the photon Jacobians has been folded into them.
    double** transient[] = {&T->stg, &T->hs, &T->D, &T->cot};
    for (double** p : transient) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    int32_t flags = 0;
    if (!hip_ok(hipMemcpy(&flags, T->flags, 4, hipMemcpyDeviceToHost), "read flags")) return bail(KR_ERR_DEVICE);
    if (flags & TAPE_FLAG_REPLAY)
        return bail(host::fail(KR_ERR_DEVICE, "re-integration of a recorded photon differs from its trace"));
    *out = T;
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
    if (segment < 0 || segment > 4096) return host::fail(KR_ERR_INVALID, "segment must be in [0, 4096]");
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
        c_off.push_back(c_off.back() + (segment > 0 ? (steps[i] + segment - 1) / segment : 0));
    }
    if (segment == 0) return create_jacobian_tape(h, P, B, img, n, std::move(list), steps, d_alpha, d_beta, d_out,
                                                  d_theta_inf, d_phi_inf, out);
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

extern "C" double kr_tape_jacobian_ms(kr_tape t) { return t ? t->jacobian_ms : -1.0; }

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
    if (t->n_rec > 0 && t->segment == 0) {
        hipLaunchKernelGGL(contract_kernel, dim3(static_cast<unsigned>((t->n_rec + 255) / 256)), dim3(256), 0, s,
                           t->list, t->Jp, t->n_rec, d_g_theta, d_g_phi, d_grad);
        KR_HIP(hipGetLastError());
    } else if (t->n_rec > 0) {
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
