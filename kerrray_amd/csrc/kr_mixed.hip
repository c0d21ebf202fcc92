// Planned image traces: longest-first scheduling, FP32 bulk with FP64 repair.
//
// One image, four kernels on one stream (no host synchronisation between):
//  1. plan: every pixel is put in one of three classes with the analytic
//     shadow edge (polar radii of Bardeen's curve about a centre, given by
//     the caller): the edge band (within `band_px` + 1 pixels of the edge)
//     with the polar strip, inside, outside.  A work list is written in that
//     order -- the edge band holds the longest photons (near-critical
//     orbits) and the inside the next longest (captured photons slow down at
//     the horizon), so the slowest photons start first.  Band and strip
//     photons are flagged FP64.
//  2. trace: a persistent kernel walks the work list with lane refill (as
//     kr_batch.hip).  Flagged photons, and all photons when the FP32 bulk is
//     off, take the exact FP64 attempt of the reference; the others take the
//     mixed attempt of kr_mixed.h.  With split roles a wavefront traces
//     photons of one precision only.  Each pixel records the precision used.
//  3. verify: any pixel within `band_px` pixels (Chebyshev) of a
//     captured/not-captured transition of the traced mask, or whose photon
//     ended neither captured nor escaped, that was not traced in FP64 goes
//     on a repair list.
//  4. repair: the listed pixels are traced again with the exact FP64 kernel
//     and their outputs overwritten; the FP32 pixels near any of them whose
//     class changed go on the next round's list, for up to REPAIR_ROUNDS
//     rounds (kr_repair.h).  A class change in the last round raises
//     FLAG_REPAIR.
// With the FP32 bulk off this is the FP64 trace with a better schedule: the
// results equal kr_trace_image's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "kerrray_hip.h"
#include "kr_device.h"
#include "kr_host.h"
#include "kr_mixed.h"
#include "kr_repair.h"

namespace kr {

// Work-list entry: pixel index | bin << BIN_SHIFT | BAND_FLAG | FP64_FLAG.
constexpr int64_t FP64_FLAG = int64_t(1) << 62;  // trace this photon with the exact FP64 attempt
constexpr int64_t BAND_FLAG = int64_t(1) << 61;  // photon of bin 0 (edge band or polar strip)
constexpr int BIN_SHIFT = 57;                     // bins 0..15 in bits 57..60
constexpr int64_t INDEX_MASK = (int64_t(1) << BIN_SHIFT) - 1;

struct PlanCounters {
    unsigned long long seg[NBINS];     // pixels per bin
    unsigned long long cursor[NBINS];  // scatter cursors
    unsigned long long inside;         // pixels inside the analytic edge, outside bin 0
    unsigned long long work;           // work-list counter of the persistent kernel
    unsigned long long work_fp32;      // second counter: the FP32 part of the list when roles are split
    unsigned long long fp64_attempts;
    // Length of the repair list of each round; listed[REPAIR_ROUNDS] counts
    // the pixels the last round would have queued for one more.
    unsigned long long listed[REPAIR_ROUNDS + 1];
};
constexpr int PLAN_COUNTER_WORDS = static_cast<int>(sizeof(PlanCounters) / 8);
static_assert(PLAN_COUNTER_WORDS <= 64, "planned_reset_kernel clears the counters with one wavefront");

// Work-list order: bin 0 (FP64: edge band and polar strip), then the other
// pixels inside the analytic edge, then those outside, each in no particular
// order.  This starts the longest photons first and mixes long and short
// photons among those in flight; ordering every pixel by its distance bin
// instead puts only long photons in flight at once and measures slower.
__device__ __forceinline__ int order_class(int bin, bool inside) { return bin == 0 ? 0 : (inside ? 1 : 2); }

// The plan kernels run 1024-thread workgroups and count in LDS, so the
// global counters take one atomic per workgroup and class: same-address
// atomics from every wavefront serialise at the L2 and cost more than the
// classification itself.
constexpr int PLAN_BLOCK = 1024;

// Classifies every pixel (bin | inside << 4 into cls) and counts the bins and
// the inside pixels outside bin 0.
__global__ __launch_bounds__(PLAN_BLOCK) void plan_count_kernel(PlanDev pl, BardeenImage B, PlanCounters* c,
                                                                uint8_t* cls) {
    __shared__ unsigned int counts[NBINS + 1];
    if (threadIdx.x < NBINS + 1) counts[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * PLAN_BLOCK + threadIdx.x;
    if (idx < B.n * B.n) {
        const PixelPlan q = plan_pixel(pl, B, idx);
        cls[idx] = static_cast<uint8_t>(q.bin | (q.inside ? 0x10 : 0));
        atomicAdd(&counts[q.bin], 1u);
        if (q.bin > 0 && q.inside) atomicAdd(&counts[NBINS], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NBINS + 1 && counts[threadIdx.x] != 0u) {
        unsigned long long* target = threadIdx.x < NBINS ? &c->seg[threadIdx.x] : &c->inside;
        atomicAdd(target, static_cast<unsigned long long>(counts[threadIdx.x]));
    }
}

// Longest-first classes of given initial states, from their constants of
// motion.  A photon with constants (E, L_z, Q) circles near the spherical
// photon orbits when its radial potential per E^2,
//   R(r) = (r^2 + a^2 - a xi)^2 - Delta(r) (eta + (xi - a)^2),
// xi = L_z / E, eta = Q / E^2, nearly has a double root outside the horizon:
// its minimum over r, scaled by (r^2 + a^2)^2, is then near 0 (for a = 0 the
// scaled minimum is 1 - b^2 / 27 at r = 3 M).  Bin 0: |minimum| <= band, the
// near-critical photons; inside: minimum > 0 (no radial turning point, the
// photon falls in); outside: the rest.  The ratios are invariant under the
// reversal of the momentum, so backward-traced camera rows class alike.
// Only the order of the work list depends on this; no photon's arithmetic.
struct RowPlanDev {
    double m, a;
    double r_lo, r_hi;  // radii sampled for the minimum: r_+ .. 5 M
    double band;
};

__device__ __forceinline__ PixelPlan plan_row(const RowPlanDev& pl, const double* y) {
    const double e = -y[4], lz = y[7], th = y[2], pth = y[6];
    double s, c;
    sincos(th, &s, &c);
    const double a = pl.a, a2 = a * a;
    const double lz2s2 = lz != 0.0 ? lz * lz / (s * s) : 0.0;
    const double q = pth * pth + c * c * (lz2s2 - a2 * e * e);
    if (!(e != 0.0) || !isfinite(q)) return PixelPlan{1, false};
    const double xi = lz / e, eta = q / (e * e);
    const double k = eta + (xi - a) * (xi - a), w = a * xi;
    constexpr int NR = 32;
    double lowest = INFINITY;
    for (int i = 0; i < NR; ++i) {
        const double r = pl.r_lo + (pl.r_hi - pl.r_lo) * (static_cast<double>(i) / (NR - 1));
        const double r2a2 = r * r + a2;
        const double u = r2a2 - w;
        const double R = u * u - (r * r - 2.0 * pl.m * r + a2) * k;
        lowest = fmin(lowest, R / (r2a2 * r2a2));
    }
    const double d = fabs(lowest);
    int bin = 0;
    if (!(d <= pl.band)) {
        bin = 1;
        for (double edge = 2.0 * pl.band; bin < NBINS - 1 && d > edge; edge *= 2.0) ++bin;
    }
    return PixelPlan{bin, lowest > 0.0};
}

// plan_count_kernel for given rows.
__global__ __launch_bounds__(PLAN_BLOCK) void plan_rows_count_kernel(RowPlanDev pl, const double* Y0, int64_t n,
                                                                     PlanCounters* c, uint8_t* cls) {
    __shared__ unsigned int counts[NBINS + 1];
    if (threadIdx.x < NBINS + 1) counts[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * PLAN_BLOCK + threadIdx.x;
    if (idx < n) {
        const PixelPlan q = plan_row(pl, Y0 + idx * 8);
        cls[idx] = static_cast<uint8_t>(q.bin | (q.inside ? 0x10 : 0));
        atomicAdd(&counts[q.bin], 1u);
        if (q.bin > 0 && q.inside) atomicAdd(&counts[NBINS], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NBINS + 1 && counts[threadIdx.x] != 0u) {
        unsigned long long* target = threadIdx.x < NBINS ? &c->seg[threadIdx.x] : &c->inside;
        atomicAdd(target, static_cast<unsigned long long>(counts[threadIdx.x]));
    }
}

// Writes the work list from the class bytes: each workgroup reserves one
// contiguous run per order class.
__global__ __launch_bounds__(PLAN_BLOCK) void plan_scatter_kernel(int64_t total, PlanCounters* c, const uint8_t* cls,
                                                                  int64_t* work, int fp32_bulk, int longest_first) {
    __shared__ unsigned int local[3];
    __shared__ unsigned long long base[3];
    if (threadIdx.x < 3) local[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * PLAN_BLOCK + threadIdx.x;
    const bool valid = idx < total;
    const int code = valid ? cls[idx] : 0;
    const int bin = code & 15;
    const int k = order_class(bin, (code & 0x10) != 0);
    unsigned int slot = 0u;
    if (valid && longest_first) slot = atomicAdd(&local[k], 1u);
    __syncthreads();
    if (longest_first && threadIdx.x < 3 && local[threadIdx.x] != 0u)
        base[threadIdx.x] = atomicAdd(&c->cursor[threadIdx.x], static_cast<unsigned long long>(local[threadIdx.x]));
    __syncthreads();
    if (!valid) return;
    int64_t pos = idx;
    if (longest_first) {
        const unsigned long long first = k == 0 ? 0ull : (k == 1 ? c->seg[0] : c->seg[0] + c->inside);
        pos = static_cast<int64_t>(first + base[k] + slot);
    }
    work[pos] = idx | (static_cast<int64_t>(bin) << BIN_SHIFT) | ((bin == 0 || !fp32_bulk) ? FP64_FLAG : 0) |
                (bin == 0 ? BAND_FLAG : 0);
}

// WAVES: wavefronts per SIMD the register allocation is made for (1: 512
// VGPRs, 2: 256, 3: 168, 4: 128).  At 2 and 3 the stage slopes k1..k4 live in
// LDS (192 B per lane: 147 KB per CU at 3 waves/SIMD); at 1 they fit in
// registers, at 4 they stay private (scratch).  With one wavefront per SIMD
// the longest photons never share their SIMD's issue slots.
//
// PARK (LDS slopes only): the DP5 error partial sums wait in LDS during the
// attempt's last right-hand side (dp5_step_stored), which frees six more
// values per lane: the FP64 kernel at two wavefronts per SIMD then spills
// nothing.  It pays where throughput decides and costs where the longest
// photon's latency does (one LDS round trip more per attempt).
//
// split (mixed form only; 1 or 2): the first ceil(bin-0 pixels / 64)
// wavefronts (ceil(bin-0 pixels / 32) with pairs, split == 2) of
// the grid take work only from the bin-0 part of the list (the FP64
// photons), the others only from the rest (FP32); a wavefront of the first
// kind joins the FP32 part once the FP64 part is handed out and its own
// photons are done.  A wavefront then never holds lanes of both precisions,
// which would run both attempts one after the other and slow the longest
// photons, the FP64 ones, down.
//
// ROWS: the photons are given initial states Y0 (n, 8), not image pixels
// (FP64 only; B is then unused).  DISK (ROWS only): the equatorial disk event
// of numba_kernel.py:439-454 ends photons that cross the disk, and every
// finished photon's event row goes to ev_Y / ev_hit.
template <int BLOCK, bool MIXED, int WAVES, bool PARK, bool ROWS = false, bool DISK = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void planned_trace_kernel(
    Params P_arg, BardeenImage B_arg, MixedCfg M, const int64_t* work, PlanCounters* c, int refill_interval,
    int band_priority, int split, uint8_t* used_fp64, RowsOut<double> out, int32_t* flags, const double* Y0,
    int64_t total, double* ev_Y, uint8_t* ev_hit) {
    static_assert(!(ROWS && MIXED), "given rows are traced in FP64");
    static_assert(!DISK || (ROWS && !MIXED), "the disk event runs on given rows in FP64");
    // At one wavefront per SIMD there are registers to spare but scalar
    // registers are short: the launch constants, uniform values the compiler
    // would hold in scalar registers (and spill to vector lanes and reload,
    // instructions a lone wavefront waits for), are moved to vector registers.
    // The empty asm statements only pin where the values live.
    Params P = P_arg;
    BardeenImage B = B_arg;
    if (WAVES == 1) {
#pragma unroll
        for (int k = 0; k < 7; ++k) asm volatile("" : "+v"(P.geom[k]));
#pragma unroll
        for (int k = 0; k < 21; ++k) asm volatile("" : "+v"(P.consts[k]));
#pragma unroll
        for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(P.params[k]));
    }
    if (WAVES == 1 && !MIXED) {
        // The image's initial-state scalars (read when a lane refills); the
        // pixel count and camera kind stay scalar.  (Not in the mixed kernel,
        // which holds both attempts and has no vector registers to spare:
        // 512^2, a=0.9, FP32 form 3.04 against 3.09 ms, FP64 form 4.44
        // against 4.38 ms with them.)
        asm volatile("" : "+v"(B.r_o), "+v"(B.theta_o), "+v"(B.sin_o), "+v"(B.a), "+v"(B.a2c2), "+v"(B.cot2));
        asm volatile("" : "+v"(B.r2a2), "+v"(B.dlt), "+v"(B.c2), "+v"(B.s2), "+v"(B.aaEE), "+v"(B.fov));
        asm volatile("" : "+v"(B.Z.r_o), "+v"(B.Z.theta_o), "+v"(B.Z.phi_o), "+v"(B.Z.r_o2), "+v"(B.Z.e00),
                     "+v"(B.Z.e03), "+v"(B.Z.e11));
        asm volatile("" : "+v"(B.Z.e22), "+v"(B.Z.e33), "+v"(B.Z.g_tt), "+v"(B.Z.g_tph), "+v"(B.Z.g_rr),
                     "+v"(B.Z.g_thth), "+v"(B.Z.g_phph));
        asm volatile("" : "+v"(B.Z.gi_tt), "+v"(B.Z.gi_tph), "+v"(B.Z.gi_rr), "+v"(B.Z.gi_thth),
                     "+v"(B.Z.gi_phph), "+v"(B.Z.sign));
    }
    constexpr bool LDS_SLOPES = WAVES == 2 || WAVES == 3;
    static_assert(LDS_SLOPES || !PARK, "the error sums are parked only in LDS");
    __shared__ double slopes[LDS_SLOPES ? 4 * NA * BLOCK : 1];
    double private_slopes[LDS_SLOPES ? 1 : 4 * NA];
    const SlopeStore<double> K = LDS_SLOPES ? SlopeStore<double>{&slopes[threadIdx.x], BLOCK}
                                            : SlopeStore<double>{private_slopes, 1};
    const int lane = threadIdx.x & 63;
    // The part of the work list this wavefront draws from, [lo, lo + count).
    int64_t lo = 0, count = total;
    unsigned long long* cursor = &c->work;
    const int64_t seg0 = MIXED && split ? static_cast<int64_t>(c->seg[0]) : 0;
    bool fp64_role = false;
    // split == 2: the FP64 band photons are traced by two lanes each (both
    // carry the same state; TrigPair shares out the trigonometry), 32 per
    // wavefront.
    const bool odd = (lane & 1) != 0;
    bool pair = false;
    if (MIXED && split) {
        const int64_t wave = (static_cast<int64_t>(blockIdx.x) * BLOCK + threadIdx.x) >> 6;
        const int64_t per_wave = split == 2 ? 32 : 64;
        fp64_role = wave < (seg0 + per_wave - 1) / per_wave;
        pair = fp64_role && split == 2;
        if (fp64_role) {
            count = seg0;
        } else {
            lo = seg0;
            count -= seg0;
            cursor = &c->work_fp32;
        }
    }
    const Geom32 G = geom32(P);
    RayState<double> R;
    R.state = RUNNING;
    bool busy = false, exhausted = false, exact = true;
    int bin = NBINS;
    int64_t idx = -1;
    int64_t fp64_att = 0;
    int since = refill_interval;
    while (true) {
        const bool any_busy = __any(busy);
        if (MIXED && fp64_role && exhausted && !any_busy) {
            // The FP64 part is handed out and this wavefront's photons are
            // done: it joins the FP32 part instead of leaving its slot idle.
            fp64_role = false;
            pair = false;
            exhausted = false;
            lo = seg0;
            count = total - seg0;
            cursor = &c->work_fp32;
        }
        if (!exhausted && (since >= refill_interval || !any_busy)) {
            since = 0;
            // In pair mode both lanes of a pair are idle together; the even
            // ones stand for their pairs.
            const uint64_t idle = __ballot(!busy) & (pair ? 0x5555555555555555ull : ~0ull);
            const int n_idle = __popcll(idle);
            if (n_idle > 0) {
                const int leader = __ffsll(static_cast<unsigned long long>(idle)) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(cursor, static_cast<unsigned long long>(n_idle));
                base = __shfl(base, leader);
                if (base + static_cast<unsigned long long>(n_idle) >= static_cast<unsigned long long>(count))
                    exhausted = true;
                if (!busy) {
                    const int me = pair ? (lane & ~1) : lane;
                    const uint64_t below = (me == 0) ? 0ull : (idle & (~0ull >> (64 - me)));
                    const int64_t pos = static_cast<int64_t>(base) + __popcll(below);
                    if (pos < count) {
                        const int64_t w = work[lo + pos];
                        idx = w & INDEX_MASK;
                        exact = !MIXED || (w & FP64_FLAG) != 0;
                        bin = static_cast<int>((w >> BIN_SHIFT) & 15);
                        double y0[8];
                        if (ROWS) {
#pragma unroll
                            for (int k = 0; k < 8; ++k) y0[k] = Y0[idx * 8 + k];
                        } else {
                            const int64_t i = idx / B.n, j = idx - i * B.n;
                            const int32_t bad =
                                image_state(B, pixel_centre(B.fov, j, B.n), -pixel_centre(B.fov, i, B.n), y0);
                            if (bad && flags) atomicOr(flags, bad);
                        }
                        if (MIXED && pair)
                            ray_init<double, true>(P, y0, R, TrigPair{odd});
                        else
                            ray_init<double, true>(P, y0, R);
                        fp64_att = 0;
                        if (R.state == RUNNING) {
                            busy = true;
                        } else if (!(pair && odd)) {
                            store_row(out, idx, R);
                            if (DISK) store_event(ev_Y, ev_hit, idx, R);
                            used_fp64[idx] = 1;
                        }
                    }
                }
            }
            // Photons of the lowest bins are the longest and set the end of
            // the launch: a wavefront holding one issues ahead of the others on
            // its SIMD (priority 3 for bin 0, 2 for bin 1, 1 for bin 2).
            if (band_priority) {
                if (__any(busy && bin == 0))
                    __builtin_amdgcn_s_setprio(3);
                else if (__any(busy && bin == 1))
                    __builtin_amdgcn_s_setprio(2);
                else if (__any(busy && bin == 2))
                    __builtin_amdgcn_s_setprio(1);
                else
                    __builtin_amdgcn_s_setprio(0);
            }
            continue;
        }
        if (!any_busy) break;
        if (busy) {
            if (MIXED && !exact) {
                ray_attempt_with<double, true, false>(P, R, MixedStepper<PARK>{G, M, &fp64_att, K});
            } else if (MIXED && pair) {
                ray_attempt_with<double, true, false>(P, R, StoredStepper<double, true, PARK, TrigPair>{K, {odd}});
            } else {
                ray_attempt_with<double, true, DISK>(P, R, StoredStepper<double, true, PARK>{K});
            }
            if (R.state != RUNNING && pair && odd) {
                busy = false;  // the even lane of the pair writes the row
            } else if (R.state != RUNNING) {
                store_row(out, idx, R);
                if (DISK) store_event(ev_Y, ev_hit, idx, R);
                used_fp64[idx] = exact ? 1 : 0;
                const int64_t n_fp64 = exact ? R.n_steps + R.n_rej : fp64_att;
                if (n_fp64) atomicAdd(&c->fp64_attempts, static_cast<unsigned long long>(n_fp64));
                busy = false;
            }
        }
        ++since;
    }
}

// The longest-first trace of given rows in the float32 working dtype (the
// planned kernels hold FP64 states): one photon per lane with its state in
// registers; each wavefront refills its idle lanes from the work list every
// `refill_interval` attempts, or at once when all are idle.  A photon runs
// ray_init and ray_attempt as integrate_ray does, so every row (and disk
// event) equals the plain kernel's bit for bit.  (Also instantiable for
// FP64 and RK4; RK4 photons of one image take about as many steps each, and
// the plain kernel traces them as fast.)
template <typename T, bool ADAPTIVE, bool DISK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void rows_refill_kernel(
    Params P, const int64_t* work, PlanCounters* c, int refill_interval, RowsOut<T> out, T* ev_Y, uint8_t* ev_hit,
    const double* Y0, int64_t total) {
    const int lane = threadIdx.x & 63;
    RayState<T> R;
    R.state = RUNNING;
    bool busy = false, exhausted = false;
    int64_t idx = -1;
    int since = refill_interval;
    while (true) {
        const bool any_busy = __any(busy);
        if (!exhausted && (since >= refill_interval || !any_busy)) {
            since = 0;
            const uint64_t idle = __ballot(!busy);
            const int n_idle = __popcll(idle);
            if (n_idle > 0) {
                const int leader = __ffsll(static_cast<unsigned long long>(idle)) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&c->work, static_cast<unsigned long long>(n_idle));
                base = __shfl(base, leader);
                if (base + static_cast<unsigned long long>(n_idle) >= static_cast<unsigned long long>(total))
                    exhausted = true;
                if (!busy) {
                    const uint64_t below = lane == 0 ? 0ull : (idle & (~0ull >> (64 - lane)));
                    const int64_t pos = static_cast<int64_t>(base) + __popcll(below);
                    if (pos < total) {
                        idx = work[pos] & INDEX_MASK;
                        double y0[8];
#pragma unroll
                        for (int k = 0; k < 8; ++k) y0[k] = Y0[idx * 8 + k];
                        ray_init<T, ADAPTIVE>(P, y0, R);
                        if (R.state == RUNNING) {
                            busy = true;
                        } else {
                            store_row(out, idx, R);
                            if (DISK) store_event(ev_Y, ev_hit, idx, R);
                        }
                    }
                }
            }
            continue;
        }
        if (!any_busy) break;
        if (busy) {
            ray_attempt<T, ADAPTIVE, DISK>(P, R);
            if (R.state != RUNNING) {
                store_row(out, idx, R);
                if (DISK) store_event(ev_Y, ev_hit, idx, R);
                busy = false;
            }
        }
        ++since;
    }
}

// The verify pass: round 0's repair list (kr_repair.h).  Each lane marks
// its own pixel queued; no lane reads another's mark.
__global__ void verify_kernel(const int64_t* state, uint8_t* marks, int64_t n, int band_px, int32_t* list,
                              PlanCounters* c) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    if (verify_pixel(state, marks, n, band_px, idx)) {
        marks[idx] = MARK_QUEUED;
        list[atomicAdd(&c->listed[0], 1ull)] = static_cast<int32_t>(idx);
    }
}

// Repair round `round`: re-traces its list in exact FP64 and queues, for the
// next round, the unmarked pixels within band_px of every pixel whose class
// changed.  Pushes past the last round only count, and raise FLAG_REPAIR.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void repair_kernel(
    Params P, BardeenImage B, int round, int band_px, const int32_t* list, int32_t* next, PlanCounters* c,
    uint8_t* marks, RowsOut<double> out, int32_t* flags) {
    const int64_t count = static_cast<int64_t>(c->listed[round]);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < count; p += stride) {
        const int64_t idx = list[p];
        const int64_t i = idx / B.n, j = idx - i * B.n;
        const bool was_captured = out.state[idx] == CAPTURED;
        double y0[8];
        image_state(B, pixel_centre(B.fov, j, B.n), -pixel_centre(B.fov, i, B.n), y0);
        RayState<double> R;
        ray_init<double, true>(P, y0, R);
        while (R.state == RUNNING) ray_attempt<double, true, false>(P, R);
        store_row(out, idx, R);
        mark_or(marks, idx, MARK_FP64);
        if ((R.state == CAPTURED) == was_captured) continue;
        for_each_neighbour(B.n, band_px, idx, [&](int64_t pos) {
            if (mark_or(marks, pos, MARK_QUEUED) != 0) return;  // FP64 already, or listed before
            const unsigned long long slot = atomicAdd(&c->listed[round + 1], 1ull);
            if (round + 1 < REPAIR_ROUNDS)
                next[slot] = static_cast<int32_t>(pos);
            else
                atomicOr(flags, FLAG_REPAIR);
        });
    }
}

__global__ void planned_reset_kernel(PlanCounters* c, int32_t* flags) {
    if (blockIdx.x == 0 && threadIdx.x < PLAN_COUNTER_WORDS)
        reinterpret_cast<unsigned long long*>(c)[threadIdx.x] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0 && flags) *flags = 0;
}

// ORs the trace's flag word into the caller's.
__global__ void flags_out_kernel(const int32_t* flags, int32_t* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *flags) atomicOr(out, *flags);
}

}  // namespace kr

using namespace kr;

namespace {

struct Work {
    PlanCounters* counters;
    int32_t* flags;
    double* radii;
    int64_t* work;
    int32_t* lists[2];  // repair lists of even and odd rounds
    uint8_t* used_fp64; // per-pixel marks (kr_repair.h); the plan's class bytes before the trace
};

void fill_stats(const PlanCounters& pc, int32_t flags, kr_planned_stats* stats) {
    stats->planned_band = static_cast<int64_t>(pc.seg[0]);
    stats->planned_inside = static_cast<int64_t>(pc.inside);
    int64_t planned = 0;
    for (int b = 0; b < NBINS; ++b) planned += static_cast<int64_t>(pc.seg[b]);
    stats->planned_outside = planned - stats->planned_band - stats->planned_inside;
    stats->repaired = 0;
    stats->repair_rounds = 0;
    for (int r = 0; r < REPAIR_ROUNDS; ++r) {
        stats->repaired += static_cast<int64_t>(pc.listed[r]);
        if (pc.listed[r]) stats->repair_rounds = r + 1;
    }
    stats->unrepaired = static_cast<int64_t>(pc.listed[REPAIR_ROUNDS]);
    stats->fp64_attempts = static_cast<int64_t>(pc.fp64_attempts);
    stats->flags = flags;
    stats->reserved = 0;
}

// Workspace layout, 256-byte aligned parts: counters, flag word, radii, work
// list (i64 per pixel), two repair lists (i32 per pixel each), marks (u8).
constexpr size_t part(int64_t bytes) { return (static_cast<size_t>(bytes) + 255) & ~static_cast<size_t>(255); }
constexpr size_t FLAGS_OFFSET = part(sizeof(PlanCounters));

constexpr size_t workspace_bytes(int64_t pixels, int32_t n_angles) {
    return part(sizeof(PlanCounters)) + part(64) + part(static_cast<int64_t>(n_angles) * 8) + part(pixels * 8) +
           2 * part(pixels * 4) + part(pixels);
}

Work carve_work(void* base, int64_t pixels, int32_t n_angles) {
    host::Carver c;
    char* b = static_cast<char*>(base);
    Work w;
    w.counters = reinterpret_cast<PlanCounters*>(b + c.take(sizeof(PlanCounters)));
    w.flags = reinterpret_cast<int32_t*>(b + c.take(64));
    w.radii = reinterpret_cast<double*>(b + c.take(static_cast<int64_t>(n_angles) * 8));
    w.work = reinterpret_cast<int64_t*>(b + c.take(pixels * 8));
    w.lists[0] = reinterpret_cast<int32_t*>(b + c.take(pixels * 4));
    w.lists[1] = reinterpret_cast<int32_t*>(b + c.take(pixels * 4));
    w.used_fp64 = reinterpret_cast<uint8_t*>(b + c.take(pixels));
    return w;
}

kr_status check_opts(const kr_planned_opts* o) {
    if (!o) return host::fail(KR_ERR_INVALID, "options are NULL");
    if (o->band_px < 0 || o->band_px > 64) return host::fail(KR_ERR_INVALID, "band_px must be in [0, 64]");
    if (o->refill_interval < 1 || o->refill_interval > 4096)
        return host::fail(KR_ERR_INVALID, "refill_interval must be in [1, 4096]");
    if (o->block_size != 64 && o->block_size != 128 && o->block_size != 256)
        return host::fail(KR_ERR_INVALID, "block_size must be 64, 128 or 256");
    if (o->waves_per_simd < 0 || o->waves_per_simd > 4)
        return host::fail(KR_ERR_INVALID, "waves_per_simd must be 0 (automatic), 1, 2, 3 or 4");
    if (!(o->axis_lz >= 0.0)) return host::fail(KR_ERR_INVALID, "axis_lz must be >= 0");
    if (o->split_roles < 0 || o->split_roles > 3)
        return host::fail(KR_ERR_INVALID, "split_roles must be 0, 1, 2 or 3");
    if (!(o->horizon_margin >= 0.0) || !(o->axis_sin >= 0.0) || !(o->turn_drdl >= 0.0) || !(o->turn_rmax >= 0.0))
        return host::fail(KR_ERR_INVALID, "switch thresholds must be >= 0");
    if (!(o->grid_fraction >= 0.0 && o->grid_fraction <= 1.0))
        return host::fail(KR_ERR_INVALID, "grid_fraction must be 0 (automatic) or in (0, 1]");
    return KR_OK;
}

template <int BLOCK, bool MIXED, int WAVES, bool PARK = false>
void launch_planned_t(int cus, double frac, hipStream_t s, const Params& P, const BardeenImage& B,
                      const MixedCfg& M, const Work& w, int refill, int band_priority, int split,
                      const RowsOut<double>& o) {
    // One resident wavefront per SIMD slot, or the share frac of them.
    const int slots = cus * (4 * WAVES * 64 / BLOCK);
    const int grid = std::max(1, static_cast<int>(slots * frac));
    hipLaunchKernelGGL((planned_trace_kernel<BLOCK, MIXED, WAVES, PARK>), dim3(grid), dim3(BLOCK), 0, s, P, B, M, w.work,
                       w.counters, refill, band_priority, split, w.used_fp64, o, w.flags, nullptr, B.n * B.n, nullptr,
                       nullptr);
}

template <int BLOCK, bool MIXED>
void launch_planned_w(int waves, bool park, int cus, double frac, hipStream_t s, const Params& P, const BardeenImage& B,
                      const MixedCfg& M, const Work& w, int refill, int band_priority, int split,
                      const RowsOut<double>& o) {
    if (waves == 1)
        launch_planned_t<BLOCK, MIXED, 1>(cus, frac, s, P, B, M, w, refill, band_priority, split, o);
    else if (waves == 3)
        launch_planned_t<BLOCK, MIXED, 3>(cus, frac, s, P, B, M, w, refill, band_priority, split, o);
    else if (waves == 4)
        launch_planned_t<BLOCK, MIXED, 4>(cus, frac, s, P, B, M, w, refill, band_priority, split, o);
    else if (park)
        launch_planned_t<BLOCK, MIXED, 2, true>(cus, frac, s, P, B, M, w, refill, band_priority, split, o);
    else
        launch_planned_t<BLOCK, MIXED, 2>(cus, frac, s, P, B, M, w, refill, band_priority, split, o);
}

kr_status enqueue_planned(kr_context* h, const Params& P, const BardeenImage& B, const kr_planned_opts* o,
                          double ca, double cb, int32_t n_angles, const Work& w, const kr_outputs& d_out,
                          hipStream_t s) {
    const int64_t pixels = B.n * B.n;
    int cus = 0;
    KR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    const unsigned pblocks = static_cast<unsigned>((pixels + 255) / 256);
    const unsigned plan_blocks = static_cast<unsigned>((pixels + PLAN_BLOCK - 1) / PLAN_BLOCK);
    PlanDev pl{ca, cb, w.radii, n_angles, (o->band_px + 1.0) * (2.0 * B.fov / static_cast<double>(B.n)), o->axis_lz};
    hipLaunchKernelGGL(planned_reset_kernel, dim3(1), dim3(64), 0, s, w.counters, w.flags);
    KR_HIP(hipGetLastError());
    // The class bytes live in used_fp64 until the trace overwrites it.
    hipLaunchKernelGGL(plan_count_kernel, dim3(plan_blocks), dim3(PLAN_BLOCK), 0, s, pl, B, w.counters, w.used_fp64);
    KR_HIP(hipGetLastError());
    hipLaunchKernelGGL(plan_scatter_kernel, dim3(plan_blocks), dim3(PLAN_BLOCK), 0, s, pixels, w.counters,
                       w.used_fp64, w.work, o->fp32_bulk ? 1 : 0, o->longest_first ? 1 : 0);
    KR_HIP(hipGetLastError());
    const double r_plus = P.geom[G_MASS] + std::sqrt(P.geom[G_MASS] * P.geom[G_MASS] - P.geom[G_A2]);
    MixedCfg M{r_plus + o->horizon_margin, o->axis_sin, o->turn_rmax, o->turn_drdl};
    RowsOut<double> bo{static_cast<double*>(d_out.Y), d_out.lam, d_out.n_steps, d_out.n_rejected, d_out.state,
                   d_out.max_null_error, d_out.max_energy_drift, d_out.max_lz_drift, d_out.max_carter_drift};
    const bool mixed = o->fp32_bulk != 0;
    const int refill = o->refill_interval, prio = o->band_priority ? 1 : 0;
    // Split roles rely on the band coming first in the work list.
    int split = o->longest_first ? o->split_roles : 0;
    // Automatic occupancy: up to 2^19 pixels both forms are bounded by their
    // longest FP64 photons, which run fastest alone on their SIMD (one
    // wavefront per SIMD, 512^2, a=0.9: split FP32 form 3.29 against 4.1 ms,
    // FP64 form 4.4-4.6 against 4.85-4.9 ms; the FP32 form without split
    // roles mixes precisions in a wavefront and stays at two); above,
    // throughput decides and two wavefronts per SIMD are faster (1024^2,
    // a=0.99: FP64 12.6 against 15.3 ms, FP32 form 6.2 against 8.7 ms).  At
    // two, the error sums are parked in LDS above 2^19 pixels (1024^2,
    // a=0.99: FP64 14.4 against 15.0 ms, FP32 form 7.5 against 8.2 ms, with
    // the library pow) and not below (512^2 FP64: 4.9 against 5.25 ms).
    const bool large = pixels > (int64_t(1) << 19);
    const int waves =
        o->waves_per_simd != 0 ? o->waves_per_simd : ((mixed ? split != 0 : true) && !large ? 1 : 2);
    // Automatic grid: at one wavefront per SIMD, up to 2^18 pixels and rtol
    // >= 1e-6, the split FP32 form's bulk finishes well before its longest
    // FP64 photon, and launching 7/8 of the slots lowers the power drawn
    // meanwhile, so the clock that photon runs at rises (512^2, a=0.9:
    // 3.50 against 3.68 ms, 3/4: 3.47 ms, 5/8: the bulk becomes the bound,
    // 3.78 ms; a=0.99, i=85: 4.24 against 4.43 ms; 362^2: equal).  At tighter
    // tolerances the FP32 bulk takes many more attempts and bounds the time
    // (512^2, a=0, rtol 1e-8: 7.84 against 7.29 ms; 256^2, rtol 1e-7: 4.53
    // against 4.41 ms), and above 2^18 pixels throughput decides: all slots.
    // Pairs (split 2) where the longest FP64 photon bounds the launch, at one
    // wavefront per SIMD (512^2, a=0.9: 3.11 against 3.51 ms, the same rows);
    // where throughput decides the band's lanes are worth more one photon each
    // (1024^2, a=0.99, two per SIMD: 6.59 against 6.21 ms), and so they are
    // in the FP64 form, whose bulk bounds it (512^2: band pairs with the rest
    // one photon per lane measured 4.82 ms at one wavefront per SIMD, 4.95 at
    // two with band priority, against 4.50 unsplit), and at tight tolerances,
    // where the FP32 bulk takes many more attempts (256^2, a=0, rtol 1e-8:
    // 5.49 paired against 5.14 ms; 362^2, a=0.5, rtol 1e-7: 3.88 against
    // 4.69; 512^2, a=0.99, i=85, rtol 1e-6: 3.77 against 4.10).  3: split,
    // never paired.
    if (split == 1 && waves == 1 && P.consts[C_RTOL] >= 1e-7) split = 2;
    if (split == 3) split = 1;
    const bool small = pixels <= (int64_t(1) << 18);
    const bool loose = P.consts[C_RTOL] >= 1e-6;
    // Paired band photons take twice the band's wavefronts from the FP32 part:
    // 15/16 of the slots then (512^2, a=0.9, Bardeen image: 3.07 against 3.11
    // ms at 7/8 and 3.15 at all; ZAMO camera: 2.94 against 3.13 and 2.79).
    const double auto_frac = split == 2 ? 0.9375 : 0.875;
    const double frac =
        o->grid_fraction > 0.0 ? o->grid_fraction : (mixed && split && waves == 1 && small && loose ? auto_frac : 1.0);
    switch (o->block_size) {
        case 64:
            mixed ? launch_planned_w<64, true>(waves, large, cus, frac, s, P, B, M, w, refill, prio, split, bo)
                  : launch_planned_w<64, false>(waves, large, cus, frac, s, P, B, M, w, refill, prio, split, bo);
            break;
        case 128:
            mixed ? launch_planned_w<128, true>(waves, large, cus, frac, s, P, B, M, w, refill, prio, split, bo)
                  : launch_planned_w<128, false>(waves, large, cus, frac, s, P, B, M, w, refill, prio, split, bo);
            break;
        default:
            mixed ? launch_planned_w<256, true>(waves, large, cus, frac, s, P, B, M, w, refill, prio, split, bo)
                  : launch_planned_w<256, false>(waves, large, cus, frac, s, P, B, M, w, refill, prio, split, bo);
    }
    KR_HIP(hipGetLastError());
    hipLaunchKernelGGL(verify_kernel, dim3(pblocks), dim3(256), 0, s, d_out.state, w.used_fp64, B.n, o->band_px,
                       w.lists[0], w.counters);
    KR_HIP(hipGetLastError());
    for (int round = 0; round < REPAIR_ROUNDS; ++round) {
        hipLaunchKernelGGL(repair_kernel, dim3(static_cast<unsigned>(cus * 2)), dim3(256), 0, s, P, B, round,
                           o->band_px, w.lists[round & 1], w.lists[(round + 1) & 1], w.counters, w.used_fp64, bo,
                           w.flags);
        KR_HIP(hipGetLastError());
    }
    return KR_OK;
}

// Near-critical band of the row planner, in the scaled radial potential:
// |1 - b^2 / 27| <= 0.05 is |b - 3 sqrt 3| <~ 0.13 M for a = 0.
constexpr double ROW_BAND = 0.05;

kr_status check_common(kr_handle h, const kr_params* p, const kr_image* img, const kr_planned_opts* o,
                       int32_t n_angles, Params* P, BardeenImage* B) {
    kr_status st = host::bardeen_image(img, true, B);
    if (st != KR_OK) return st;
    if ((st = host::validate(h, p, B->n * B->n, P)) != KR_OK) return st;
    if (p->dtype != KR_FLOAT64 || P->ipar[I_ADAPTIVE] != 1 || P->ipar[I_DISK] != 0)
        return host::fail(KR_ERR_UNSUPPORTED, "planned traces are float64 rk45 without a disk event");
    if ((st = check_opts(o)) != KR_OK) return st;
    if (n_angles < 8 || n_angles > (1 << 20)) return host::fail(KR_ERR_INVALID, "n_angles must be in [8, 2^20]");
    if (B->n * B->n > (int64_t(1) << 31)) return host::fail(KR_ERR_INVALID, "planned traces take at most 2^31 pixels");
    return KR_OK;
}

// The workspace of kr_trace_image_planned_noalloc: a static device array,
// one per device, sized for the largest image and edge table it accepts.
// Host code never uses the array's name as an address (on the host it is
// only a stand-in); kr_create looks up each device's own copy once, with
// hipGetSymbolAddress, and the entry point passes that address on.
constexpr size_t NOALLOC_SCRATCH_BYTES = workspace_bytes(KR_NOALLOC_MAX_PIXELS, KR_NOALLOC_MAX_ANGLES);
static_assert(NOALLOC_SCRATCH_BYTES < (size_t(18) << 20), "the static scratch is documented as 17 MiB");

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[MAX_DEVICES] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

size_t kr::host::rows_workspace_bytes(int64_t n) { return workspace_bytes(n, 0); }

// One launch of the planned kernel over given rows: WAVES wavefronts per SIMD
// on every slot, with or without the disk event.
template <int BLOCK, int WAVES, bool PARK, bool DISK>
static void launch_rows(int cus, const Params& P, const Work& w, const RowsOut<double>& o, const double* d_Y0,
                        int64_t n, const kr_outputs& d_out, hipStream_t s) {
    const BardeenImage none{};
    const MixedCfg M{};
    hipLaunchKernelGGL((planned_trace_kernel<BLOCK, false, WAVES, PARK, true, DISK>),
                       dim3(cus * (4 * WAVES * 64 / BLOCK)), dim3(BLOCK), 0, s, P, none, M, w.work, w.counters, 8, 0,
                       0, w.used_fp64, o, w.flags, d_Y0, n, DISK ? static_cast<double*>(d_out.event_Y) : nullptr,
                       DISK ? d_out.event_hit : nullptr);
}

// One launch of rows_refill_kernel, two wavefronts per SIMD on every slot.
template <typename T, bool ADAPTIVE, bool DISK>
static void launch_refill(int cus, const Params& P, const Work& w, const kr_outputs& d_out, const double* d_Y0,
                          int64_t n, hipStream_t s) {
    RowsOut<T> o{static_cast<T*>(d_out.Y), d_out.lam, d_out.n_steps, d_out.n_rejected, d_out.state,
                 d_out.max_null_error, d_out.max_energy_drift, d_out.max_lz_drift, d_out.max_carter_drift};
    hipLaunchKernelGGL((rows_refill_kernel<T, ADAPTIVE, DISK>), dim3(cus * (4 * 2 * 64 / 256)), dim3(256), 0, s, P,
                       w.work, w.counters, 8, o, DISK ? static_cast<T*>(d_out.event_Y) : nullptr,
                       DISK ? d_out.event_hit : nullptr, d_Y0, n);
}

// The Dormand-Prince trace of given rows (kr_integrate_batch[_dev], and the
// disk render): the row plan, then the persistent trace of the longest-first
// work list, in FP64 on planned_trace_kernel (exactly as the planned image
// trace runs it), in float32 on rows_refill_kernel.  Every row (and disk
// event) equals the plain kernel's.
kr_status kr::host::enqueue_rows_planned(kr_context* h, const Params& P, int dtype, const double* d_Y0, int64_t n,
                                         const kr_outputs& d_out, void* d_work, hipStream_t s) {
    if (n <= 0) return KR_OK;
    if (P.ipar[I_ADAPTIVE] == 0) return host::fail(KR_ERR_UNSUPPORTED, "the row planner takes Dormand-Prince only");
    if (P.ipar[I_DISK] && (!d_out.event_Y || !d_out.event_hit))
        return host::fail(KR_ERR_INVALID, "the disk event needs event_Y and event_hit");
    int cus = 0;
    KR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    Work w = carve_work(d_work, n, 0);
    const double m = P.geom[G_MASS];
    const double r_plus = m + std::sqrt(m * m - P.geom[G_A2]);
    RowPlanDev pl{m, P.geom[G_A], r_plus, 5.0 * m, ROW_BAND};
    const unsigned plan_blocks = static_cast<unsigned>((n + PLAN_BLOCK - 1) / PLAN_BLOCK);
    hipLaunchKernelGGL(planned_reset_kernel, dim3(1), dim3(64), 0, s, w.counters, w.flags);
    KR_HIP(hipGetLastError());
    hipLaunchKernelGGL(plan_rows_count_kernel, dim3(plan_blocks), dim3(PLAN_BLOCK), 0, s, pl, d_Y0, n, w.counters,
                       w.used_fp64);
    KR_HIP(hipGetLastError());
    hipLaunchKernelGGL(plan_scatter_kernel, dim3(plan_blocks), dim3(PLAN_BLOCK), 0, s, n, w.counters, w.used_fp64,
                       w.work, 0, 1);
    KR_HIP(hipGetLastError());
    if (dtype != KR_FLOAT64) {
        if (P.ipar[I_DISK])
            launch_refill<float, true, true>(cus, P, w, d_out, d_Y0, n, s);
        else
            launch_refill<float, true, false>(cus, P, w, d_out, d_Y0, n, s);
        KR_HIP(hipGetLastError());
        return KR_OK;
    }
    RowsOut<double> o{static_cast<double*>(d_out.Y), d_out.lam, d_out.n_steps, d_out.n_rejected, d_out.state,
                      d_out.max_null_error, d_out.max_energy_drift, d_out.max_lz_drift, d_out.max_carter_drift};
    // The planned image trace's automatic occupancy: one wavefront per SIMD up
    // to 2^19 rows, two with the error sums parked in LDS above.
    constexpr int BLOCK = 128;
    const bool two = n > (int64_t(1) << 19);
    if (P.ipar[I_DISK]) {
        if (two)
            launch_rows<BLOCK, 2, true, true>(cus, P, w, o, d_Y0, n, d_out, s);
        else
            launch_rows<BLOCK, 1, false, true>(cus, P, w, o, d_Y0, n, d_out, s);
    } else {
        if (two)
            launch_rows<BLOCK, 2, true, false>(cus, P, w, o, d_Y0, n, d_out, s);
        else
            launch_rows<BLOCK, 1, false, false>(cus, P, w, o, d_Y0, n, d_out, s);
    }
    KR_HIP(hipGetLastError());
    return KR_OK;
}

kr_status kr::host::planned_scratch_setup(int device, void** address) {
    *address = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return fail(KR_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        int current = -1;
        KR_HIP(hipGetDevice(&current));
        if (current != device) return fail(KR_ERR_INVALID, "planned_scratch_setup: the device must be current");
        void* ptr = nullptr;
        KR_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(KR_ERR_DEVICE, "hipGetSymbolAddress returned no address for the planned scratch");
        g_scratch_address[device] = ptr;
    }
    *address = g_scratch_address[device];
    return KR_OK;
}

extern "C" size_t kr_planned_workspace_bytes(int64_t pixels, int32_t n_angles) {
    return workspace_bytes(pixels, n_angles);
}

extern "C" kr_status kr_trace_image_planned_dev(kr_handle h, const kr_params* p, const kr_image* img,
                                                const kr_planned_opts* o, double centre_alpha, double centre_beta,
                                                const double* d_radii, int32_t n_angles, const kr_outputs* d_out,
                                                void* d_work, size_t work_bytes, void* stream) {
    host::clear_error();
    Params P;
    BardeenImage B;
    kr_status st = check_common(h, p, img, o, n_angles, &P, &B);
    if (st != KR_OK) return st;
    if ((st = host::check_outputs(d_out, false)) != KR_OK) return st;
    if (!d_radii || !d_work) return host::fail(KR_ERR_INVALID, "the radii or workspace pointer is NULL");
    if (work_bytes < workspace_bytes(B.n * B.n, n_angles)) return host::fail(KR_ERR_INVALID, "workspace too small");
    KR_HIP(hipSetDevice(h->device));
    Work w = carve_work(d_work, B.n * B.n, n_angles);
    hipStream_t s = static_cast<hipStream_t>(stream);
    KR_HIP(hipMemcpyAsync(w.radii, d_radii, static_cast<size_t>(n_angles) * 8, hipMemcpyDeviceToDevice, s));
    return enqueue_planned(h, P, B, o, centre_alpha, centre_beta, n_angles, w, *d_out, s);
}

extern "C" kr_status kr_trace_image_planned_noalloc(kr_handle h, const kr_params* p, const kr_image* img,
                                                    const kr_planned_opts* o, double centre_alpha,
                                                    double centre_beta, const double* d_radii, int32_t n_angles,
                                                    const kr_outputs* d_out, int32_t* d_flags, void* stream) {
    host::clear_error();
    Params P;
    BardeenImage B;
    kr_status st = check_common(h, p, img, o, n_angles, &P, &B);
    if (st != KR_OK) return st;
    if ((st = host::check_outputs(d_out, false)) != KR_OK) return st;
    if (!d_radii) return host::fail(KR_ERR_INVALID, "the radii pointer is NULL");
    const int64_t pixels = B.n * B.n;
    if (pixels > KR_NOALLOC_MAX_PIXELS || n_angles > KR_NOALLOC_MAX_ANGLES)
        return host::fail(KR_ERR_INVALID, "kr_trace_image_planned_noalloc takes at most KR_NOALLOC_MAX_PIXELS (2^20) "
                                          "pixels and KR_NOALLOC_MAX_ANGLES (8192) radii");
    if (!h->scratch) return host::fail(KR_ERR_DEVICE, "the handle has no planned scratch address");
    KR_HIP(hipSetDevice(h->device));
    Work w = carve_work(h->scratch, pixels, n_angles);
    hipStream_t s = static_cast<hipStream_t>(stream);
    KR_HIP(hipMemcpyAsync(w.radii, d_radii, static_cast<size_t>(n_angles) * 8, hipMemcpyDeviceToDevice, s));
    if ((st = enqueue_planned(h, P, B, o, centre_alpha, centre_beta, n_angles, w, *d_out, s)) != KR_OK) return st;
    if (d_flags) {
        hipLaunchKernelGGL(flags_out_kernel, dim3(1), dim3(64), 0, s, w.flags, d_flags);
        KR_HIP(hipGetLastError());
    }
    return KR_OK;
}

extern "C" kr_status kr_trace_image_planned(kr_handle h, const kr_params* p, const kr_image* img,
                                            const kr_planned_opts* o, double centre_alpha, double centre_beta,
                                            const double* radii, int32_t n_angles, const kr_outputs* out,
                                            kr_planned_stats* stats, double* kernel_ms) {
    host::clear_error();
    Params P;
    BardeenImage B;
    kr_status st = check_common(h, p, img, o, n_angles, &P, &B);
    if (st != KR_OK) return st;
    if ((st = host::check_outputs(out, false)) != KR_OK) return st;
    if (!radii) return host::fail(KR_ERR_INVALID, "the radii array is NULL");
    KR_HIP(hipSetDevice(h->device));
    const int64_t n = B.n * B.n;
    host::Carver c;
    const size_t oY = c.take(n * 64), olam = c.take(n * 8), ost = c.take(n * 8), orej = c.take(n * 8),
                 ostate = c.take(n * 8), om0 = c.take(n * 8), om1 = c.take(n * 8), om2 = c.take(n * 8),
                 om3 = c.take(n * 8), owork = c.take(static_cast<int64_t>(workspace_bytes(n, n_angles)));
    if ((st = host::reserve(h, c.total)) != KR_OK) return st;
    char* base = static_cast<char*>(h->buf);
    kr_outputs d{};
    d.Y = base + oY;
    d.lam = reinterpret_cast<double*>(base + olam);
    d.n_steps = reinterpret_cast<int64_t*>(base + ost);
    d.n_rejected = reinterpret_cast<int64_t*>(base + orej);
    d.state = reinterpret_cast<int64_t*>(base + ostate);
    d.max_null_error = reinterpret_cast<double*>(base + om0);
    d.max_energy_drift = reinterpret_cast<double*>(base + om1);
    d.max_lz_drift = reinterpret_cast<double*>(base + om2);
    d.max_carter_drift = reinterpret_cast<double*>(base + om3);
    Work w = carve_work(base + owork, n, n_angles);
    hipStream_t s = h->stream;
    KR_HIP(hipMemcpyAsync(w.radii, radii, static_cast<size_t>(n_angles) * 8, hipMemcpyHostToDevice, s));
    KR_HIP(hipEventRecord(h->ev0, s));
    if ((st = enqueue_planned(h, P, B, o, centre_alpha, centre_beta, n_angles, w, d, s)) != KR_OK) return st;
    KR_HIP(hipEventRecord(h->ev1, s));
    PlanCounters pc;
    int32_t flags = 0;
    KR_HIP(hipMemcpyAsync(&pc, w.counters, sizeof(pc), hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(&flags, w.flags, 4, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->Y, d.Y, n * 64, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->lam, d.lam, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->n_steps, d.n_steps, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->n_rejected, d.n_rejected, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->state, d.state, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->max_null_error, d.max_null_error, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->max_energy_drift, d.max_energy_drift, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->max_lz_drift, d.max_lz_drift, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(out->max_carter_drift, d.max_carter_drift, n * 8, hipMemcpyDeviceToHost, s));
    KR_HIP(hipStreamSynchronize(s));
    if (stats) fill_stats(pc, flags, stats);
    if (flags) return host::flag_status(flags);
    if (kernel_ms) {
        float ms = 0.f;
        KR_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *kernel_ms = ms;
    }
    return KR_OK;
}

extern "C" kr_status kr_planned_read_stats(const void* d_work, size_t work_bytes, kr_planned_stats* stats,
                                           void* stream) {
    host::clear_error();
    if (!d_work || !stats) return host::fail(KR_ERR_INVALID, "workspace or stats is NULL");
    if (work_bytes < FLAGS_OFFSET + 4) return host::fail(KR_ERR_INVALID, "workspace is too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    PlanCounters pc;
    int32_t flags = 0;
    KR_HIP(hipMemcpyAsync(&pc, d_work, sizeof(pc), hipMemcpyDeviceToHost, s));
    KR_HIP(hipMemcpyAsync(&flags, static_cast<const char*>(d_work) + FLAGS_OFFSET, 4, hipMemcpyDeviceToHost, s));
    KR_HIP(hipStreamSynchronize(s));
    fill_stats(pc, flags, stats);
    return KR_OK;
}
