// Persistent, work-stealing integration of a ragged batch of Bardeen images.
//
// One launch integrates every pixel of every image of the batch (mixed
// spins, inclinations, resolutions and tolerances).  The grid is sized to
// stay resident; each lane holds one photon at a time.  Every
// `refill_interval` attempts a wavefront counts its idle lanes, takes that
// many photon indices from a global counter with one atomic, and starts new
// photons in those lanes, so finished photons are replaced instead of
// leaving lanes idle until the slowest photon of the wavefront ends.  A
// wavefront exits when the counter is exhausted and none of its lanes is
// busy.  Per-photon arithmetic is that of the plain kernel (kr_device.h), so
// every image's results equal the plain per-image kernel's bit for bit.
//
// Up to 128 images, each image's integration constants (and the first photon
// of each image, for the lookup) are copied into LDS at kernel start; a lane
// reads its own job's constants from there at every attempt, and the image
// scalars from the job table in global memory when it starts a photon.  A
// larger batch reads the constants from the job table too, where the L2
// keeps them.  A batch of one image uses the uniform-constant variant
// (constants in kernel arguments, scalar registers).
//
// At two wavefronts per SIMD, when the device's LDS allows, the stage slopes
// k1..k4 of each lane's Dormand-Prince attempt live in LDS too (and the
// error sums are parked there during the last right-hand side), as in the
// planned kernel: 192 B per lane, which takes the attempt out of scratch.
//
// Photons are handed out in a processing order of the images that need not
// be the job order: with kr_tune.order = 1 the images whose photons take the
// most attempts (smallest rtol, then most pixels) go first, so their longest
// photons do not start at the end of the launch.  Each image writes its rows
// where the job order puts them either way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kerrray_hip.h"
#include "kr_device.h"
#include "kr_host.h"
#include "kr_mixed.h"
#include "kr_repair.h"

namespace kr {

struct JobDev {
    Params P;
    BardeenImage B;
    int64_t first;      // first photon of this image in the processing order
    int64_t count;      // n * n
    int64_t out_first;  // first output row of this image (images in job order)
    PlanDev plan;       // precision form: the image's analytic edge, band and polar strip
    double r_plus;      // precision form: outer horizon (the FP32 stages carry r - r_+)
};

// Counters of the precision form of a batch.
struct BatchCounters {
    unsigned long long fp64_attempts;  // attempts taken in exact FP64
    unsigned long long listed[REPAIR_ROUNDS + 1];
};

constexpr int MAX_LDS_JOBS = 128;

template <typename T>
using BatchOut = RowsOut<T>;

__device__ __forceinline__ int find_job(const int64_t* firsts, int n_jobs, int64_t idx) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (firsts[mid] <= idx)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// A launch integrates the photons at positions [begin, begin + count) of the
// batch's work order: photon work[p] of the processing order (work == NULL:
// photon p).
struct BatchArgs {
    const JobDev* jobs;     // device copy of the job table
    const int64_t* firsts;  // jobs[k].first, k = 0 .. n_jobs - 1
    int n_jobs;
    int refill_interval;
    const int32_t* work;  // precision form: exact photons first, then the FP32 ones
    int64_t begin;
    int64_t count;
    unsigned long long* counter;
    int32_t* flags;
    uint8_t* marks;          // precision form: MARK_FP64 per output row (NULL otherwise)
    BatchCounters* counters; // precision form
};

// Where a launch finds each image's first photon and integration constants:
// a batch of one image in the kernel arguments (scalar registers); up to
// MAX_LDS_JOBS images copied into LDS at kernel start; more images read from
// the job table in global memory (L2) at every attempt.
enum : int { TABLE_ARGS = 0, TABLE_LDS = 1, TABLE_GLOBAL = 2 };

int table_of(int n_jobs) { return n_jobs == 1 ? TABLE_ARGS : n_jobs <= MAX_LDS_JOBS ? TABLE_LDS : TABLE_GLOBAL; }

// Dynamic LDS of batch_kernel, in 8-byte words: the stage slopes (4 NA BLOCK
// doubles, LDS_SLOPES only), then with TABLE_LDS the first photon of every
// job (n_jobs), then every job's Params.
constexpr size_t slope_bytes(int block) { return static_cast<size_t>(4 * NA) * block * sizeof(double); }

size_t batch_lds_bytes(int table, bool lds_slopes, int block, int n_jobs) {
    return (lds_slopes ? slope_bytes(block) : 0) +
           (table == TABLE_LDS ? static_cast<size_t>(n_jobs) * (sizeof(int64_t) + sizeof(Params)) : 0);
}

// The precision form runs a batch as two launches over one work order: the
// photons of every image's edge band or polar strip (bin 0 of the image plan,
// kr_mixed.h) and every photon of an image below fp32_min_rtol first, in the
// exact FP64 kernel (MIXED = false), then all the others in the MIXED kernel:
// FP32 right-hand sides with the FP64 update, error norm and step control.
// So no wavefront carries both attempts, and the exact photons run in the
// kernel that holds only the exact attempt.  Rows record which (marks), for
// the verify/repair pass.
template <typename T, int TABLE, int BLOCK, int WAVES, bool LDS_SLOPES, bool MIXED = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void batch_kernel(
    BatchArgs a, JobDev uniform_job, BatchOut<T> out) {
    static_assert(!MIXED || sizeof(T) == sizeof(double), "the precision form integrates in FP64 and FP32");
    static_assert(sizeof(Params) % sizeof(uint64_t) == 0, "Params is copied in 8-byte words");
    constexpr bool UNIFORM = TABLE == TABLE_ARGS;
    extern __shared__ double lds[];
    int64_t* lds_firsts = reinterpret_cast<int64_t*>(lds + (LDS_SLOPES ? 4 * NA * BLOCK : 0));
    Params* params = reinterpret_cast<Params*>(lds_firsts + (TABLE == TABLE_LDS ? a.n_jobs : 0));
    if (TABLE == TABLE_LDS) {
        constexpr int PW = static_cast<int>(sizeof(Params) / sizeof(uint64_t));
        for (int k = threadIdx.x; k < a.n_jobs; k += BLOCK) lds_firsts[k] = a.firsts[k];
        uint64_t* dst = reinterpret_cast<uint64_t*>(params);
        for (int w = threadIdx.x; w < a.n_jobs * PW; w += BLOCK)
            dst[w] = reinterpret_cast<const uint64_t*>(&a.jobs[w / PW].P)[w % PW];
        __syncthreads();
    }
    const int64_t* firsts = TABLE == TABLE_LDS ? lds_firsts : a.firsts;
    // Stage slopes of an exact attempt: in LDS at two wavefronts per SIMD,
    // else private (only the precision form's occasional exact attempts of a
    // bulk photon use them there; the exact photons keep theirs in registers).
    T private_slopes[LDS_SLOPES ? 1 : 4 * NA];
    const SlopeStore<T> K = LDS_SLOPES ? SlopeStore<T>{reinterpret_cast<T*>(lds) + threadIdx.x, BLOCK}
                                       : SlopeStore<T>{private_slopes, 1};
    const int lane = threadIdx.x & 63;
    RayState<T> R;
    R.state = RUNNING;
    bool busy = false;
    bool exhausted = false;  // wave-uniform: the counter has run past the end
    Geom32 G{};
    MixedCfg M{};
    int64_t fp64_att = 0;
    int64_t idx = -1;
    int job = 0;
    int since = a.refill_interval;
    while (true) {
        const bool any_busy = __any(busy);
        if (!exhausted && (since >= a.refill_interval || !any_busy)) {
            since = 0;
            const uint64_t idle = __ballot(!busy);
            const int n_idle = __popcll(idle);
            if (n_idle > 0) {
                const int leader = __ffsll(static_cast<unsigned long long>(idle)) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(a.counter, static_cast<unsigned long long>(n_idle));
                base = __shfl(base, leader);
                if (base + static_cast<unsigned long long>(n_idle) >= static_cast<unsigned long long>(a.count))
                    exhausted = true;
                if (!busy) {
                    const uint64_t below = (lane == 0) ? 0ull : (idle & (~0ull >> (64 - lane)));
                    const int64_t mine = static_cast<int64_t>(base) + __popcll(below);
                    if (mine < a.count) {
                        const int64_t pos = a.begin + mine;
                        idx = a.work ? static_cast<int64_t>(a.work[pos]) : pos;  // in processing order
                        job = UNIFORM ? 0 : find_job(firsts, a.n_jobs, idx);
                        const JobDev& J = UNIFORM ? uniform_job : a.jobs[job];
                        const int64_t local = idx - J.first;
                        const int64_t i = local / J.B.n, j = local - i * J.B.n;
                        idx = J.out_first + local;  // from here on: the output row
                        double y0[8];
                        const int32_t bad = image_state(J.B, pixel_centre(J.B.fov, j, J.B.n),
                                                        -pixel_centre(J.B.fov, i, J.B.n), y0);
                        if (bad && a.flags) atomicOr(a.flags, bad);
                        ray_init<T, true>(UNIFORM ? uniform_job.P : TABLE == TABLE_LDS ? params[job] : J.P, y0, R);
                        if (MIXED) {
                            G = geom32(J.P);
                            M = MixedCfg{J.r_plus, 0.0, 0.0, 0.0};
                            fp64_att = 0;
                        }
                        if (R.state == RUNNING) {
                            busy = true;
                        } else {
                            store_row(out, idx, R);
                            if (a.marks) a.marks[idx] = MARK_FP64;
                        }
                    }
                }
            }
            continue;  // re-evaluate: new photons may have terminated at once
        }
        if (!any_busy) break;  // exhausted and nothing in flight
        if (busy) {
            const Params& P = UNIFORM ? uniform_job.P : TABLE == TABLE_LDS ? params[job] : a.jobs[job].P;
            if constexpr (MIXED) {
                if (LDS_SLOPES)
                    ray_attempt_with<double, true, false>(P, R, MixedStepper<true>{G, M, &fp64_att, K});
                else
                    ray_attempt_with<double, true, false>(P, R, MixedStepper<false>{G, M, &fp64_att, K});
            } else if (LDS_SLOPES) {
                ray_attempt_with<T, true, false>(P, R, StoredStepper<T, true, true>{K});
            } else {
                ray_attempt<T, true, false>(P, R);
            }
            if (R.state != RUNNING) {
                store_row(out, idx, R);
                if (a.marks) {
                    a.marks[idx] = MIXED ? 0 : MARK_FP64;
                    const int64_t n64 = MIXED ? fp64_att : R.n_steps + R.n_rej;
                    if (n64) atomicAdd(&a.counters->fp64_attempts, static_cast<unsigned long long>(n64));
                }
                busy = false;
            }
        }
        ++since;
    }
}

// Clears the launches' cursors (counter[0 .. n_counters)), the flags and the
// precision counters.
__global__ void batch_reset_kernel(unsigned long long* counter, int n_counters, int32_t* flags,
                                   BatchCounters* counters) {
    constexpr int words = static_cast<int>(sizeof(BatchCounters) / 8);
    if (blockIdx.x == 0 && threadIdx.x < words && counters)
        reinterpret_cast<unsigned long long*>(counters)[threadIdx.x] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < n_counters) counter[threadIdx.x] = 0ull;
    if (threadIdx.x == 0 && blockIdx.x == 0 && flags) *flags = 0;
}

// One byte per photon of the batch in processing order: 1 if its image's plan
// traces it in exact FP64 (bin 0), else 0.  Run once, when a precision batch
// is created, to build the work order.
__global__ void batch_bins_kernel(const JobDev* jobs, int n_jobs, int64_t total, uint8_t* exact) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= total) return;
    int lo = 0, hi = n_jobs - 1;  // the job table is in processing order: jobs[k].first ascends
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    const JobDev& J = jobs[lo];
    exact[p] = plan_pixel(J.plan, J.B, p - J.first).bin == 0 ? 1 : 0;
}

// The job owning output row `row`: jobs are in processing order, rows in job
// order; out_firsts[k] is the first row of the k-th job in job order and
// order_job[k] its index in the job table.
__device__ __forceinline__ const JobDev& row_job(const JobDev* jobs, const int64_t* out_firsts, const int* order_job,
                                                 int n_jobs, int64_t row) {
    return jobs[order_job[find_job(out_firsts, n_jobs, row)]];
}

// Verify pass of the precision form over every row of the batch: each pixel
// against its own image's neighbours (kr_repair.h).
__global__ void batch_verify_kernel(const JobDev* jobs, const int64_t* out_firsts, const int* order_job, int n_jobs,
                                    int64_t total, const int64_t* state, uint8_t* marks, int band_px, int32_t* list,
                                    BatchCounters* c) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= total) return;
    const JobDev& J = row_job(jobs, out_firsts, order_job, n_jobs, row);
    const int64_t local = row - J.out_first;
    if (verify_pixel(state + J.out_first, marks + J.out_first, J.B.n, band_px, local)) {
        marks[row] = MARK_QUEUED;
        list[atomicAdd(&c->listed[0], 1ull)] = static_cast<int32_t>(row);
    }
}

// Repair round `round` of the precision form (repair_kernel of kr_mixed.hip
// over the rows of a batch).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void batch_repair_kernel(
    const JobDev* jobs, const int64_t* out_firsts, const int* order_job, int n_jobs, int round, int band_px,
    const int32_t* list, int32_t* next, BatchCounters* c, uint8_t* marks, BatchOut<double> out, int32_t* flags) {
    const int64_t count = static_cast<int64_t>(c->listed[round]);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < count; p += stride) {
        const int64_t row = list[p];
        const JobDev& J = row_job(jobs, out_firsts, order_job, n_jobs, row);
        const int64_t local = row - J.out_first;
        const int64_t i = local / J.B.n, j = local - i * J.B.n;
        const bool was_captured = out.state[row] == CAPTURED;
        double y0[8];
        image_state(J.B, pixel_centre(J.B.fov, j, J.B.n), -pixel_centre(J.B.fov, i, J.B.n), y0);
        RayState<double> R;
        ray_init<double, true>(J.P, y0, R);
        while (R.state == RUNNING) ray_attempt<double, true, false>(J.P, R);
        store_row(out, row, R);
        mark_or(marks, row, MARK_FP64);
        if ((R.state == CAPTURED) == was_captured) continue;
        for_each_neighbour(J.B.n, band_px, local, [&](int64_t pos) {
            const int64_t g = J.out_first + pos;
            if (mark_or(marks, g, MARK_QUEUED) != 0) return;
            const unsigned long long slot = atomicAdd(&c->listed[round + 1], 1ull);
            if (round + 1 < REPAIR_ROUNDS)
                next[slot] = static_cast<int32_t>(g);
            else
                atomicOr(flags, FLAG_REPAIR);
        });
    }
}

using BatchFn = void (*)(BatchArgs, JobDev, BatchOut<double>);

template <int TABLE, int BLOCK, bool MIXED>
BatchFn pick_waves(int waves, bool lds_slopes) {
    if (waves >= 3) return batch_kernel<double, TABLE, BLOCK, 3, false, MIXED>;
    return lds_slopes ? batch_kernel<double, TABLE, BLOCK, 2, true, MIXED>
                      : batch_kernel<double, TABLE, BLOCK, 2, false, MIXED>;
}

template <int TABLE>
BatchFn pick_table(int block, int waves, bool lds_slopes, bool mixed) {
    if (mixed) {
        if (block == 64) return pick_waves<TABLE, 64, true>(waves, lds_slopes);
        if (block == 128) return pick_waves<TABLE, 128, true>(waves, lds_slopes);
        return pick_waves<TABLE, 256, true>(waves, lds_slopes);
    }
    if (block == 64) return pick_waves<TABLE, 64, false>(waves, lds_slopes);
    if (block == 128) return pick_waves<TABLE, 128, false>(waves, lds_slopes);
    return pick_waves<TABLE, 256, false>(waves, lds_slopes);
}

BatchFn pick(int table, int block, int waves, bool lds_slopes, bool mixed = false) {
    if (table == TABLE_ARGS) return pick_table<TABLE_ARGS>(block, waves, lds_slopes, mixed);
    if (table == TABLE_LDS) return pick_table<TABLE_LDS>(block, waves, lds_slopes, mixed);
    return pick_table<TABLE_GLOBAL>(block, waves, lds_slopes, mixed);
}

}  // namespace kr

using namespace kr;

struct kr_batch_s {
    kr_context* h = nullptr;
    void* dmem = nullptr;  // job table, counter, flags
    JobDev* d_jobs = nullptr;
    unsigned long long* d_counter = nullptr;
    int32_t* d_flags = nullptr;
    int64_t* d_firsts = nullptr;  // jobs[k].first in processing order
    std::vector<JobDev> jobs;
    int64_t total = 0;
    kr_tune tune{};
    kr_outputs out{};
    int cus = 0;
    bool lds_slopes = false;  // stage slopes in LDS (two wavefronts per SIMD, where LDS allows)
    size_t lds_bytes = 0;     // dynamic LDS per block
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // Precision form (kr_batch_create_precision).
    bool precision = false;
    int band_px = 0;
    void* pmem = nullptr;  // radii, marks, repair lists, counters, row tables
    uint8_t* d_marks = nullptr;
    int32_t* d_lists[2] = {nullptr, nullptr};
    BatchCounters* d_counters = nullptr;
    int64_t* d_out_firsts = nullptr;
    int* d_order_job = nullptr;
    int32_t* d_work = nullptr;  // work order: the n_exact exact photons, then the FP32 ones
    int64_t n_exact = 0;
    hipStream_t side = nullptr;  // precision form: the FP32 launch runs here, beside the exact one
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

kr_status check_tune(const kr_tune* t) {
    if (!t) return host::fail(KR_ERR_INVALID, "tune is NULL");
    if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
        return host::fail(KR_ERR_INVALID, "block_size must be 64, 128 or 256");
    if (t->waves_per_simd < 1 || t->waves_per_simd > 8) return host::fail(KR_ERR_INVALID, "waves_per_simd in [1, 8]");
    if (t->refill_interval < 1 || t->refill_interval > 4096)
        return host::fail(KR_ERR_INVALID, "refill_interval in [1, 4096]");
    if (t->order != 0 && t->order != 1) return host::fail(KR_ERR_INVALID, "order must be 0 or 1");
    return KR_OK;
}

kr_status build_jobs(kr_context* h, const kr_job* jobs, int32_t n_jobs, int order, std::vector<JobDev>& out,
                     int64_t& total) {
    if (!jobs || n_jobs < 1) return host::fail(KR_ERR_INVALID, "need at least one job");
    out.resize(n_jobs);
    total = 0;
    for (int k = 0; k < n_jobs; ++k) {
        JobDev& J = out[k];
        std::memset(&J, 0, sizeof(J));
        const int64_t n = jobs[k].image.resolution;
        kr_status st = host::validate(h, &jobs[k].params, n * n, &J.P);
        if (st != KR_OK) return st;
        if (jobs[k].params.dtype != KR_FLOAT64 || J.P.ipar[I_ADAPTIVE] != 1 || J.P.ipar[I_DISK] != 0)
            return host::fail(KR_ERR_UNSUPPORTED, "batched images must be float64 rk45 without a disk event");
        if ((st = host::bardeen_image(&jobs[k].image, true, &J.B)) != KR_OK) return st;
        J.out_first = total;
        J.count = n * n;
        total += n * n;
    }
    if (order == 1)  // smallest rtol first, then most pixels; ties keep the job order
        std::stable_sort(out.begin(), out.end(), [](const JobDev& x, const JobDev& y) {
            if (x.P.consts[C_RTOL] != y.P.consts[C_RTOL]) return x.P.consts[C_RTOL] < y.P.consts[C_RTOL];
            return x.count > y.count;
        });
    int64_t first = 0;
    for (JobDev& J : out) {
        J.first = first;
        first += J.count;
    }
    return KR_OK;
}

hipError_t enqueue(kr_batch_s* b, hipStream_t s) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, b->d_counter, 2, b->d_flags, b->d_counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int cus = b->cus;
    const int block = b->tune.block_size;
    const int waves_per_cu = 4 * b->tune.waves_per_simd;
    const int blocks_per_cu = (waves_per_cu * 64 + block - 1) / block;
    const unsigned grid = static_cast<unsigned>(cus * blocks_per_cu);
    const int n_jobs = static_cast<int>(b->jobs.size());
    BatchOut<double> o{static_cast<double*>(b->out.Y), b->out.lam, b->out.n_steps, b->out.n_rejected, b->out.state,
                       b->out.max_null_error, b->out.max_energy_drift, b->out.max_lz_drift, b->out.max_carter_drift};
    // The exact part (every photon, without the precision form) on `s`, and the FP32 part forked onto the
    // batch's side stream (joined before the verify pass), so that the FP32 wavefronts fill the slots the exact
    // launch leaves while its longest photons finish.  Launches over an empty range are left out; each draws
    // from a cursor of its own.  With both parts the exact launch is sized to its own photons (at most every
    // slot), which leaves the rest of the chip to the FP32 launch from the start.
    const int64_t parts[3] = {0, b->precision ? b->n_exact : b->total, b->total};
    const bool fork = parts[1] > parts[0] && parts[2] > parts[1];
    if (fork) {
        if ((e = hipEventRecord(b->ev_fork, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(b->side, b->ev_fork, 0)) != hipSuccess) return e;
    }
    for (int part = 0; part < 2; ++part) {
        if (parts[part + 1] == parts[part]) continue;
        const bool mixed = part == 1;
        const int64_t count = parts[part + 1] - parts[part];
        unsigned g = grid;
        if (fork && !mixed) {
            const int64_t blocks = (count + block - 1) / block;
            if (blocks < static_cast<int64_t>(g)) g = static_cast<unsigned>(blocks);
        }
        BatchArgs args{b->d_jobs, b->d_firsts, n_jobs, b->tune.refill_interval, b->d_work, parts[part], count,
                       b->d_counter + part, b->d_flags, b->d_marks, b->d_counters};
        BatchFn fn = pick(table_of(n_jobs), block, b->tune.waves_per_simd, b->lds_slopes, mixed);
        hipLaunchKernelGGL(fn, dim3(g), dim3(block), b->lds_bytes, fork && mixed ? b->side : s, args, b->jobs[0], o);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (fork) {
        if ((e = hipEventRecord(b->ev_join, b->side)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s, b->ev_join, 0)) != hipSuccess) return e;
    }
    if (!b->precision) return hipSuccess;
    hipLaunchKernelGGL(batch_verify_kernel, dim3(static_cast<unsigned>((b->total + 255) / 256)), dim3(256), 0, s,
                       b->d_jobs, b->d_out_firsts, b->d_order_job, n_jobs, b->total, b->out.state, b->d_marks,
                       b->band_px, b->d_lists[0], b->d_counters);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int round = 0; round < REPAIR_ROUNDS; ++round) {
        hipLaunchKernelGGL(batch_repair_kernel, dim3(static_cast<unsigned>(cus * 2)), dim3(256), 0, s, b->d_jobs,
                           b->d_out_firsts, b->d_order_job, n_jobs, round, b->band_px, b->d_lists[round & 1],
                           b->d_lists[(round + 1) & 1], b->d_counters, b->d_marks, o, b->d_flags);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

extern "C" kr_status kr_batch_create(kr_handle h, const kr_job* jobs, int32_t n_jobs, const kr_tune* tune,
                                     const kr_outputs* d_out, kr_batch* out) {
    host::clear_error();
    if (!h) return host::fail(KR_ERR_INVALID, "handle is NULL");
    if (!out) return host::fail(KR_ERR_INVALID, "out is NULL");
    *out = nullptr;
    kr_status st = check_tune(tune);
    if (st != KR_OK) return st;
    if (!d_out || !d_out->Y || !d_out->lam || !d_out->n_steps || !d_out->n_rejected || !d_out->state ||
        !d_out->max_null_error || !d_out->max_energy_drift || !d_out->max_lz_drift || !d_out->max_carter_drift)
        return host::fail(KR_ERR_INVALID, "an output array is NULL");
    kr_batch_s* b = new (std::nothrow) kr_batch_s();
    if (!b) return host::fail(KR_ERR_DEVICE, "out of host memory");
    b->h = h;
    if ((st = build_jobs(h, jobs, n_jobs, tune->order, b->jobs, b->total)) != KR_OK) {
        delete b;
        return st;
    }
    b->tune = *tune;
    b->out = *d_out;
    hipError_t e = hipSetDevice(h->device);
    const size_t table = b->jobs.size() * sizeof(JobDev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, h->device);
    // Device block: the job table, then (256-byte aligned) the launches' cursors and the flags word (512 B),
    // then the first photon of every job.
    const size_t o_counter = (table + 255) & ~size_t(255);
    if (e == hipSuccess) e = hipMalloc(&b->dmem, o_counter + 512 + b->jobs.size() * sizeof(int64_t));
    if (e == hipSuccess) {
        b->d_jobs = static_cast<JobDev*>(b->dmem);
        b->d_counter = reinterpret_cast<unsigned long long*>(static_cast<char*>(b->dmem) + o_counter);
        b->d_flags = reinterpret_cast<int32_t*>(b->d_counter + 32);
        b->d_firsts = reinterpret_cast<int64_t*>(static_cast<char*>(b->dmem) + o_counter + 512);
        std::vector<int64_t> firsts(b->jobs.size());
        for (size_t k = 0; k < b->jobs.size(); ++k) firsts[k] = b->jobs[k].first;
        e = hipMemcpy(b->d_jobs, b->jobs.data(), table, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_firsts, firsts.data(), firsts.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    }
    // Stage slopes in LDS at two wavefronts per SIMD when a block's share
    // fits the device's per-block and per-CU LDS.
    int lds_block = 0, lds_cu = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->device);
    if (e == hipSuccess) {
        const int table = table_of(n_jobs);
        const int block = tune->block_size;
        const size_t blocks_per_cu = static_cast<size_t>((4 * tune->waves_per_simd * 64 + block - 1) / block);
        const size_t with = batch_lds_bytes(table, true, block, n_jobs);
        b->lds_slopes = tune->waves_per_simd == 2 && with <= static_cast<size_t>(lds_block) &&
                        with * blocks_per_cu <= static_cast<size_t>(lds_cu);
        if (b->lds_slopes && with > 65536) {
            BatchFn fn = pick(table, block, 2, true);
            b->lds_slopes = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                static_cast<int>(with)) == hipSuccess;
            (void)hipGetLastError();
        }
        b->lds_bytes = batch_lds_bytes(table, b->lds_slopes, block, n_jobs);
    }
    if (e != hipSuccess) {
        if (b->dmem) (void)hipFree(b->dmem);
        delete b;
        return host::fail(KR_ERR_DEVICE, std::string("kr_batch_create: ") + hipGetErrorString(e));
    }
    *out = b;
    return KR_OK;
}

extern "C" int64_t kr_batch_photons(kr_batch b) { return b ? b->total : -1; }

extern "C" kr_status kr_batch_run(kr_batch b, void* stream) {
    host::clear_error();
    if (!b) return host::fail(KR_ERR_INVALID, "batch is NULL");
    KR_HIP(hipSetDevice(b->h->device));
    KR_HIP(enqueue(b, static_cast<hipStream_t>(stream)));
    return KR_OK;
}

extern "C" kr_status kr_batch_graph_launch(kr_batch b, void* stream) {
    host::clear_error();
    if (!b) return host::fail(KR_ERR_INVALID, "batch is NULL");
    KR_HIP(hipSetDevice(b->h->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!b->exec) {
        // Capture on a private stream so the caller's stream is never in capture mode.
        // A graph left by an earlier capture whose instantiation failed is released first.
        if (b->graph) {
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        hipStream_t cap;
        KR_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t e2 = enqueue(b, cap);
            e = hipStreamEndCapture(cap, &b->graph);
            if (e2 != hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess) return host::fail(KR_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    }
    KR_HIP(hipGraphLaunch(b->exec, s));
    return KR_OK;
}

extern "C" kr_status kr_batch_flags(kr_batch b, int32_t* flags) {
    host::clear_error();
    if (!b || !flags) return host::fail(KR_ERR_INVALID, "batch or flags is NULL");
    KR_HIP(hipSetDevice(b->h->device));
    KR_HIP(hipMemcpy(flags, b->d_flags, 4, hipMemcpyDeviceToHost));
    return KR_OK;
}

extern "C" kr_status kr_batch_destroy(kr_batch b) {
    host::clear_error();
    if (!b) return KR_OK;
    (void)hipSetDevice(b->h->device);
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->dmem) (void)hipFree(b->dmem);
    if (b->pmem) (void)hipFree(b->pmem);
    delete b;
    return KR_OK;
}

extern "C" kr_status kr_batch_create_precision(kr_handle h, const kr_job* jobs, int32_t n_jobs, const kr_tune* tune,
                                               const kr_batch_precision* prec, const kr_outputs* d_out,
                                               kr_batch* out) {
    host::clear_error();
    if (!out) return host::fail(KR_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!prec) return host::fail(KR_ERR_INVALID, "precision options are NULL");
    if (prec->band_px < 0 || prec->band_px > 64) return host::fail(KR_ERR_INVALID, "band_px must be in [0, 64]");
    if (prec->n_angles < 8 || prec->n_angles > 8192) return host::fail(KR_ERR_INVALID, "n_angles must be in [8, 8192]");
    if (!(prec->axis_lz >= 0.0)) return host::fail(KR_ERR_INVALID, "axis_lz must be >= 0");
    if (!(prec->fp32_min_rtol >= 0.0)) return host::fail(KR_ERR_INVALID, "fp32_min_rtol must be >= 0");
    if (!prec->centre_alpha || !prec->centre_beta || !prec->radii)
        return host::fail(KR_ERR_INVALID, "centre or radii array is NULL");
    kr_batch b = nullptr;
    kr_status st = kr_batch_create(h, jobs, n_jobs, tune, d_out, &b);
    if (st != KR_OK) return st;
    if (b->total >= (int64_t(1) << 31)) {
        kr_batch_destroy(b);
        return host::fail(KR_ERR_INVALID, "the precision form of a batch takes fewer than 2^31 photons");
    }
    // Job-order tables for the row -> job lookup, and each job's plan.
    const int nj = n_jobs;
    std::vector<int64_t> out_firsts(nj);
    std::vector<int> order_job(nj);
    for (int k = 0; k < nj; ++k) {
        // the job table is in processing order; out_first is monotone in job order
        int pos = 0;
        for (int m = 0; m < nj; ++m)
            if (b->jobs[m].out_first < b->jobs[k].out_first) ++pos;
        out_firsts[pos] = b->jobs[k].out_first;
        order_job[pos] = k;
    }
    host::Carver c;
    const size_t o_radii = c.take(static_cast<int64_t>(nj) * prec->n_angles * 8), o_marks = c.take(b->total),
                 o_l0 = c.take(b->total * 4), o_l1 = c.take(b->total * 4), o_cnt = c.take(sizeof(BatchCounters)),
                 o_of = c.take(nj * 8), o_oj = c.take(nj * 4), o_work = c.take(b->total * 4);
    hipError_t e = hipMalloc(&b->pmem, c.total);
    if (e == hipSuccess) {
        char* base = static_cast<char*>(b->pmem);
        double* d_radii = reinterpret_cast<double*>(base + o_radii);
        b->d_marks = reinterpret_cast<uint8_t*>(base + o_marks);
        b->d_lists[0] = reinterpret_cast<int32_t*>(base + o_l0);
        b->d_lists[1] = reinterpret_cast<int32_t*>(base + o_l1);
        b->d_counters = reinterpret_cast<BatchCounters*>(base + o_cnt);
        b->d_out_firsts = reinterpret_cast<int64_t*>(base + o_of);
        b->d_order_job = reinterpret_cast<int*>(base + o_oj);
        b->d_work = reinterpret_cast<int32_t*>(base + o_work);
        // The caller's arrays are in job order; the table is in processing order.
        for (int k = 0; k < nj; ++k) {
            JobDev& J = b->jobs[order_job[k]];
            const double m = J.P.geom[G_MASS];
            J.r_plus = m + std::sqrt(m * m - J.P.geom[G_A2]);
            // An image below fp32_min_rtol is all band: every photon exact.
            const bool fp32 = J.P.consts[C_RTOL] >= prec->fp32_min_rtol;
            J.plan = PlanDev{prec->centre_alpha[k], prec->centre_beta[k], d_radii + static_cast<int64_t>(k) * prec->n_angles,
                             prec->n_angles,
                             fp32 ? (prec->band_px + 1.0) * (2.0 * J.B.fov / static_cast<double>(J.B.n)) : INFINITY,
                             prec->axis_lz};
        }
        e = hipMemcpy(d_radii, prec->radii, static_cast<size_t>(nj) * prec->n_angles * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_out_firsts, out_firsts.data(), nj * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->d_order_job, order_job.data(), nj * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_jobs, b->jobs.data(), b->jobs.size() * sizeof(JobDev), hipMemcpyHostToDevice);
        // The work order: each photon's bin from the device plan (the marks array serves as the buffer), then
        // the exact photons and the FP32 ones, each part in processing order.
        if (e == hipSuccess) {
            hipLaunchKernelGGL(batch_bins_kernel, dim3(static_cast<unsigned>((b->total + 255) / 256)), dim3(256), 0,
                               nullptr, b->d_jobs, nj, b->total, b->d_marks);
            e = hipGetLastError();
        }
        std::vector<uint8_t> exact;
        std::vector<int32_t> work;
        if (e == hipSuccess) {
            try {
                exact.resize(static_cast<size_t>(b->total));
                work.resize(static_cast<size_t>(b->total));
            } catch (const std::bad_alloc&) {
                e = hipErrorOutOfMemory;
            }
        }
        if (e == hipSuccess) e = hipMemcpy(exact.data(), b->d_marks, static_cast<size_t>(b->total), hipMemcpyDeviceToHost);
        if (e == hipSuccess) {
            int64_t n0 = 0;
            for (int64_t p = 0; p < b->total; ++p)
                if (exact[p]) work[n0++] = static_cast<int32_t>(p);
            b->n_exact = n0;
            for (int64_t p = 0; p < b->total; ++p)
                if (!exact[p]) work[n0++] = static_cast<int32_t>(p);
            e = hipMemcpy(b->d_work, work.data(), static_cast<size_t>(b->total) * 4, hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        kr_batch_destroy(b);
        return host::fail(KR_ERR_DEVICE, std::string("kr_batch_create_precision: ") + hipGetErrorString(e));
    }
    // The side stream (default priority) and the events that fork the FP32 launch onto it and join it back.
    e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        kr_batch_destroy(b);
        return host::fail(KR_ERR_DEVICE, std::string("kr_batch_create_precision: ") + hipGetErrorString(e));
    }
    b->precision = true;
    b->band_px = prec->band_px;
    // The precision kernel needs the device's LDS limit raised as the FP64 one does.
    if (b->lds_slopes && b->lds_bytes > 65536) {
        BatchFn fn = pick(table_of(static_cast<int>(b->jobs.size())), b->tune.block_size, 2, true, true);
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(b->lds_bytes)) != hipSuccess) {
            (void)hipGetLastError();
            kr_batch_destroy(b);
            return host::fail(KR_ERR_DEVICE, "kr_batch_create_precision: the LDS request was refused");
        }
    }
    *out = b;
    return KR_OK;
}

extern "C" kr_status kr_batch_stats(kr_batch b, kr_planned_stats* stats) {
    host::clear_error();
    if (!b || !stats) return host::fail(KR_ERR_INVALID, "batch or stats is NULL");
    if (!b->precision) return host::fail(KR_ERR_INVALID, "the batch has no precision form");
    KR_HIP(hipSetDevice(b->h->device));
    BatchCounters c{};
    int32_t flags = 0;
    KR_HIP(hipDeviceSynchronize());
    KR_HIP(hipMemcpy(&c, b->d_counters, sizeof(c), hipMemcpyDeviceToHost));
    KR_HIP(hipMemcpy(&flags, b->d_flags, 4, hipMemcpyDeviceToHost));
    std::memset(stats, 0, sizeof(*stats));
    stats->planned_band = b->n_exact;
    stats->planned_inside = -1;
    stats->planned_outside = -1;
    for (int r = 0; r < REPAIR_ROUNDS; ++r) {
        stats->repaired += static_cast<int64_t>(c.listed[r]);
        if (c.listed[r]) stats->repair_rounds = r + 1;
    }
    stats->unrepaired = static_cast<int64_t>(c.listed[REPAIR_ROUNDS]);
    stats->fp64_attempts = static_cast<int64_t>(c.fp64_attempts);
    stats->flags = flags;
    return KR_OK;
}
