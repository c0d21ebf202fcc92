/*
 * kerrray_hip.h - C ABI of the MI355X (gfx950) null-geodesic backend.
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns
 * a kr_status; kr_last_error() gives the message of the most recent failure
 * on the calling thread.  No entry point aborts or throws across the ABI.
 *
 * The reference this replaces is the Python/Numba "numba" backend of KerrRay
 * (src/kerrray/raytracing/backends/numba_backend.py); each declaration below
 * cites the reference interface it stands in for.
 *
 * Conventions
 *  - Photon state y = [t, r, theta, phi, p_t, p_r, p_theta, p_phi]
 *    (src/kerrray/geodesics/state.py:35), float64 inputs.
 *  - "_dev" entry points take device pointers and a hipStream_t (as void*)
 *    and only enqueue work; the others take caller-owned host arrays,
 *    copy through library-owned device memory and return after the work is
 *    done.
 *  - A handle owns one device, one stream and a grow-only device buffer; it
 *    is not re-entrant (one call at a time per handle).  The _dev calls of
 *    one handle share its device workspace: enqueue them on one stream, or
 *    synchronise between calls made on different streams.
 */
#ifndef KERRRAY_HIP_H
#define KERRRAY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KR_ABI_VERSION 7

typedef enum kr_status {
    KR_OK = 0,
    KR_ERR_INVALID = 1,     /* bad argument: maps to ValueError */
    KR_ERR_DEVICE = 2,      /* HIP runtime failure / no device: BackendUnavailable */
    KR_ERR_RADICAND = 3,    /* negative Carter radicand in an initial condition: ValueError */
    KR_ERR_UNSUPPORTED = 4, /* option the kernels do not implement: ValueError */
    KR_ERR_NOT_NULL = 5,    /* a ZAMO camera's photon is not null: RuntimeError (camera.py:290-291,366-371) */
    KR_ERR_REPAIR = 6       /* the precision form's FP64 mask repair hit its round limit: RuntimeError */
} kr_status;

typedef struct kr_context* kr_handle;

/* Working dtype of the integration (IntegratorOptions.dtype,
 * src/kerrray/geodesics/integrators.py:95). */
enum { KR_FLOAT64 = 0, KR_FLOAT32 = 1 };

/* Kernel constants, packed exactly as the reference's host wrapper packs
 * them: geom/consts/params/ipar = _kernel_constants
 * (numba_backend.py:87-115, layouts numba_kernel.py:89-107; consts hold the
 * float64 values that the reference then casts to the working dtype) and
 * the dense tableaus of _tableaus (numba_backend.py:118-128).  Only the
 * Dormand-Prince 5(4) and classical RK4 tableaus of
 * src/kerrray/geodesics/tableaus.py:59-86 are accepted. */
typedef struct kr_params {
    double geom[7];
    double consts[21];
    double params[3];
    int64_t ipar[4];
    double tab_a[49];
    double tab_e[7];
    double rk4_b[4];
    double rk4_c[4];
    int32_t dtype; /* KR_FLOAT64 or KR_FLOAT32 */
    int32_t reserved;
} kr_params;

/* Output arrays of a batch: the BatchResult fields
 * (src/kerrray/photons/trajectories.py:100-128).  Y and event_Y hold the
 * working dtype; event_Y/event_hit may be NULL unless ipar[3] (disk) is set. */
typedef struct kr_outputs {
    void* Y;                  /* (n, 8) */
    double* lam;              /* (n) */
    int64_t* n_steps;         /* (n) */
    int64_t* n_rejected;      /* (n) */
    int64_t* state;           /* (n) TerminationState codes */
    double* max_null_error;   /* (n) */
    double* max_energy_drift; /* (n) */
    double* max_lz_drift;     /* (n) */
    double* max_carter_drift; /* (n) */
    void* event_Y;            /* (n, 8) or NULL */
    uint8_t* event_hit;       /* (n) or NULL */
} kr_outputs;

/* A square image at the observer: pixel centres fov (2j+1-n)/n
 * (src/kerrray/raytracing/camera.py:206-222), alpha along columns,
 * beta = -centre_i along rows (row 0 on top).  The camera sets the photon of
 * each image point (alpha, beta):
 *  - KR_CAMERA_BARDEEN: the photon of src/kerrray/benchmarks/rayset.py:158-165
 *    with E = 1, inward, sign_theta = -1 for beta >= 0 and +1 otherwise;
 *  - KR_CAMERA_ZAMO: the photon of initial_states
 *    (src/kerrray/raytracing/camera.py:298-381) seen by a zero-angular-
 *    momentum observer at (r_o, theta_o, phi_o), reversed for backward
 *    tracing.  theta_o is Camera.inclination_rad, i.e. already clamped off
 *    the axis. */
#define KR_CAMERA_BARDEEN 0
#define KR_CAMERA_ZAMO 1
typedef struct kr_image {
    double mass;
    double spin;        /* a / M */
    double r_o;         /* observer radius */
    double theta_o;     /* observer polar angle, radians */
    double fov;         /* half-width of the image plane */
    int64_t resolution; /* n: the image has n x n pixels, row-major */
    int32_t camera;     /* KR_CAMERA_BARDEEN or KR_CAMERA_ZAMO */
    int32_t reserved;   /* 0 */
    double phi_o;       /* observer azimuth, radians (ZAMO) */
} kr_image;

/* ABI version (KR_ABI_VERSION). */
int kr_abi_version(void);

/* Number of visible HIP devices, or a negative value on runtime failure. */
int kr_device_count(void);

/* Open a handle on `device`.  Stands in for the availability check of
 * NumbaBackend.__init__ (numba_backend.py:223-226): failure here is the
 * backend being unavailable. */
kr_status kr_create(int device, kr_handle* out);
kr_status kr_destroy(kr_handle h);

/* Message of the last failure on this thread ("" if none). */
const char* kr_last_error(void);

/* Backend.integrate_batch over caller-supplied initial states:
 * integrate_batch_numba (numba_backend.py:131-207) ->
 * integrate_rays_kernel (numba_kernel.py:466-474).  Y0 is (n, 8) float64
 * host memory; outputs are host arrays.  kernel_ms (may be NULL) receives
 * the device time of the integration kernel.  rk45 batches of at least
 * 2^15 rows (FP64) or 2^17 rows (FP64 with the disk event, float32) run on
 * the persistent longest-first kernels, the others one photon per lane; the
 * rows are the same bit for bit either way. */
kr_status kr_integrate_batch(kr_handle h, const kr_params* p, const double* Y0, int64_t n, const kr_outputs* out,
                             double* kernel_ms);

/* Same on device pointers, enqueued on `stream` (hipStream_t). */
kr_status kr_integrate_batch_dev(kr_handle h, const kr_params* p, const double* d_Y0, int64_t n,
                                 const kr_outputs* d_out, void* stream);

/* Device-side error flags of the image and point entry points: the kernel
 * ORs KR_FLAG_RADICAND into *d_flags (when d_flags is not NULL) if some
 * initial condition had a Carter radicand below -1e-12 x its scale, the
 * case photon_from_constants rejects with ValueError
 * (src/kerrray/geodesics/initial_conditions.py:71-74).  The caller zeroes
 * the flag word before the launch. */
#define KR_FLAG_RADICAND 1
/* ZAMO cameras: a photon whose tetrad projection is not null to 1e-10, or
 * for which no null p_t exists (RuntimeError in camera.py:290-291,366-371),
 * and an image point with alpha^2 + beta^2 >= r_o^2, which is no direction
 * on the observer's sky (ValueError in camera.py:342-349). */
#define KR_FLAG_NOT_NULL 2
#define KR_FLAG_SKY 4
/* Disk shading: a crossing state that is not finite or not equatorial, with
 * no timelike circular orbit at its radius, or with a non-positive frequency
 * ratio (ValueError in physics/redshift.py:108-151). */
#define KR_FLAG_DISK 8
/* Mask edge: a direction whose starting point is outside the captured region
 * (ValueError in raytracing/boundary.py:163-164). */
#define KR_FLAG_EDGE 16
/* Planned traces: the FP64 verify/repair of the mask still changed a
 * pixel's class in its last allowed round, so pixels near that change may
 * keep their FP32 class (KR_ERR_REPAIR on the host entry points). */
#define KR_FLAG_REPAIR 32

/* The initial states of every pixel of `img`, row-major (n*n, 8) float64
 * into d_Y0: initial_states(cam, st, reverse) (camera.py:298-381) for a ZAMO
 * camera, the Bardeen photons for a Bardeen image (reverse must be 1 there:
 * they are already the inward rays).  reverse = 0 gives the ZAMO camera's
 * physical, future-directed photon.  KR_FLAG_* bits go to *d_flags. */
kr_status kr_image_states_dev(kr_handle h, const kr_image* img, int32_t reverse, double* d_Y0, int32_t* d_flags,
                              void* stream);

/* The initial states of n image points (d_alpha[k], d_beta[k]) of the
 * camera in `img` (resolution and fov are ignored), (n, 8) float64 into
 * d_Y0: image_plane_states(cam, st, alpha, beta)
 * (src/kerrray/benchmarks/cpu.py:140-175) for a ZAMO camera, with the
 * construction of initial_states (so the states of pixel centres are those
 * of kr_image_states_dev bit for bit; the reference's two constructions
 * agree to 1e-12), or the Bardeen photons for a Bardeen camera (reverse must
 * be 1).  KR_FLAG_* bits go to *d_flags. */
kr_status kr_point_states_dev(kr_handle h, const kr_image* img, int32_t reverse, const double* d_alpha,
                              const double* d_beta, int64_t n, double* d_Y0, int32_t* d_flags, void* stream);

/* ---------------------------------------------------------------------
 * Accretion-disk images: render_disk (src/kerrray/raytracing/renderer.py:
 * 163-216).  The trace stops photons at the disk plane (DISK_HIT, the event
 * of numba_kernel.py:439-454, enabled by ipar[3] = 1 with consts[18:20] =
 * r_in, r_out); each crossing is then shaded with the redshift factor of
 * the Keplerian gas (physics/redshift.py:63-151) and I_obs = g^n I_em(r)
 * (physics/accretion.py:180-196).
 * ------------------------------------------------------------------- */
typedef struct kr_disk {
    double r_in, r_out;       /* emitting annulus (M); must equal consts[18], consts[19] */
    double emissivity_index;  /* p of I_em = (r / r_in)^(-p) */
    int32_t redshift_power;   /* n of I_obs = g^n I_em: 3 or 4 */
    int32_t prograde;         /* 1: gas orbits with a L_z > 0; 0: retrograde */
} kr_disk;

/* Shades n crossing rows: event_Y (n, 8) in the params' working dtype
 * (float64 or float32 values are read as float64), event_hit (n) uint8.
 * Writes intensity (0 off the disk), g and r_hit (NaN off the disk), float64
 * device arrays; KR_FLAG_DISK into *d_flags where the reference raises. */
kr_status kr_disk_shade_dev(kr_handle h, const kr_disk* disk, double mass, double spin, int32_t dtype,
                            const void* d_event_Y, const uint8_t* d_event_hit, int64_t n, double* d_intensity,
                            double* d_g, double* d_r_hit, int32_t* d_flags, void* stream);

/* render_disk for every pixel of `img` (ZAMO or Bardeen camera): the trace
 * of kr_trace_image with the disk event, then the shading above.  Host
 * arrays: the BatchResult fields in `out` (event_Y, event_hit included) and
 * intensity, g, r_hit (n*n each).  kernel_ms (may be NULL) receives the
 * device time of trace and shading.  rk45 renders of at least 2^17 pixels
 * (FP64 or float32) trace the image's initial states on the persistent
 * longest-first kernels (the rows and crossing states of kr_trace_image bit
 * for bit). */
kr_status kr_render_disk(kr_handle h, const kr_params* p, const kr_image* img, const kr_disk* disk,
                         const kr_outputs* out, double* intensity, double* g, double* r_hit, double* kernel_ms);
/* Same on device arrays, enqueued on `stream`. */
kr_status kr_render_disk_dev(kr_handle h, const kr_params* p, const kr_image* img, const kr_disk* disk,
                             const kr_outputs* d_out, double* d_intensity, double* d_g, double* d_r_hit,
                             int32_t* d_flags, void* stream);

/* Sub-pixel shadow edge of a captured mask: extract_boundary_from_mask
 * (src/kerrray/raytracing/boundary.py:116-170) for the directions
 * (cos[k], sin[k]) from (centre_alpha, centre_beta).  d_mask is rows x cols
 * bytes (1 = captured) of a regular grid whose first pixel centre is
 * (alpha0, beta0), pixel size d (alpha along columns, beta down the rows);
 * the mask's bilinear interpolant (zero outside the image) is sampled every
 * d / 8 up to `reach` and radii[k] is the linear 0.5 crossing after the last
 * sample at or above 0.5.  A direction whose first sample is below 0.5 gets
 * NaN and KR_FLAG_EDGE in *d_flags.  One lane per direction. */
kr_status kr_mask_edge_dev(kr_handle h, const uint8_t* d_mask, int64_t rows, int64_t cols, double alpha0,
                           double beta0, double d, double centre_alpha, double centre_beta, const double* d_cos,
                           const double* d_sin, int64_t n_angles, double reach, double* d_radii, int32_t* d_flags,
                           void* stream);

/* The ZAMO frame at (r_o, theta_o) of `img` as 4 x 4 row-major float64
 * arrays: the tetrad rows e_(a)^mu (zamo_tetrad, camera.py:225-262), the
 * metric g_mu_nu and its inverse g^mu_nu (geometry/metric.py:147-219), in
 * the values the ZAMO kernels use.  Host only. */
kr_status kr_zamo_frame(const kr_image* img, double* tetrad, double* metric, double* inverse_metric);

/* compute_shadow's trace (src/kerrray/raytracing/shadow.py:179-211 ->
 * rays.py:72-117) for a Bardeen image: initial conditions are generated
 * on the device, one pixel per lane, outputs row-major over the image. */
kr_status kr_trace_image(kr_handle h, const kr_params* p, const kr_image* img, const kr_outputs* out,
                         double* kernel_ms);
kr_status kr_trace_image_dev(kr_handle h, const kr_params* p, const kr_image* img, const kr_outputs* d_out,
                             int32_t* d_flags, void* stream);

/* Bardeen photons at arbitrary image points (alpha[k], beta[k]) of the
 * observer in `img` (resolution and fov are ignored); device pointers. */
kr_status kr_trace_points_dev(kr_handle h, const kr_params* p, const kr_image* img, const double* d_alpha,
                              const double* d_beta, int64_t n, const kr_outputs* d_out, int32_t* d_flags,
                              void* stream);

/* Shadow-edge search along polar directions about (centre_alpha,
 * centre_beta): the loop of scripts/study_edge_bisection.py:53-68 on the
 * device.  Direction k has unit vector (cos_angles[k], sin_angles[k]); its
 * bracket [lo[k], hi[k]] (inside at lo, outside at hi) is updated in place
 * for `iterations` rounds.  n_probe = 1 is plain bisection (mid =
 * 0.5 (lo + hi); escaped -> hi = mid, else lo = mid); n_probe > 1 traces
 * n_probe evenly spaced interior points per round and keeps the sub-bracket
 * around the first escaping one (multi-probe form of
 * src/kerrray/validation/kerr_bisection.py:12-18,230-246).  failures
 * counts traced photons that neither escaped nor were captured;
 * nonmonotone counts rounds whose probes were not inside...outside ordered.
 * Host arrays.  FP64 rk45 searches whose directions all fit on the chip at
 * one wavefront per SIMD (n_probe <= 256) run every round in one launch;
 * others launch once per round.  The brackets are the same either way. */
kr_status kr_bisect_edge(kr_handle h, const kr_params* p, const kr_image* img, double centre_alpha,
                         double centre_beta, const double* cos_angles, const double* sin_angles, int64_t n_angles,
                         double* lo, double* hi, int32_t iterations, int32_t n_probe, int64_t* failures,
                         int64_t* nonmonotone, double* kernel_ms);

/* ---------------------------------------------------------------------
 * Critical impact parameter of equatorial photons: the bracketed search of
 * bisect_critical_impacts (src/kerrray/validation/kerr_bisection.py:177-300)
 * on the device.  Each search keeps a bracket [lo, hi], CAPTURED at lo and
 * ESCAPED at hi, for photons launched inward in the equatorial plane from
 * r0 with E = 1, Q = 0 and L_z = lz_sign[k] * b (photon_from_constants,
 * src/kerrray/geodesics/initial_conditions.py:83-112, with sign_r = -1 and
 * sign_theta = +1).  The endpoints are classified first; then every round
 * traces n_probe evenly spaced interior impact parameters of every open
 * search in one launch and narrows the brackets in a second (lo = the
 * largest captured probe, hi = the smallest escaped one), until each
 * bracket is no wider than tol or has taken max_iterations rounds.
 * ------------------------------------------------------------------- */
typedef struct kr_impact_opts {
    double r0;             /* launch radius (outside the outer horizon) */
    double theta;          /* launch polar angle (pi / 2 for the reference's searches) */
    double cos2_theta;     /* math.cos(theta) ** 2, as the reference computes it */
    double sin2_theta;     /* math.sin(theta) ** 2 */
    double tol;            /* bracket width at which a search stops (> 0) */
    double capture_radius; /* r_+ + horizon_epsilon: the relabelling radius below */
    int64_t max_iterations;
    int32_t n_probe;       /* interior probes per search and round, 1 .. 1024 */
    int32_t horizon_crossing_as_capture; /* 1: OUT_OF_DOMAIN with finite r <= capture_radius counts as CAPTURED */
} kr_impact_opts;

/* What a search run found.  error: 0 none; 1 an endpoint was not CAPTURED
 * at lo / ESCAPED at hi; 2 a probe ended neither CAPTURED nor ESCAPED; 3 a
 * probe escaped below a captured one (the reference raises RuntimeError for
 * each).  On an error, the brackets are those before the failing round, and
 * probe_b / probe_state / was_open hold that round (round 0: the endpoint
 * check, two probes per search). */
typedef struct kr_impact_result {
    int64_t rays;      /* photons traced, endpoints included */
    int64_t max_steps; /* largest accepted-step count of any of them */
    int64_t crossings; /* photons relabelled CAPTURED */
    int32_t error;
    int32_t error_round;
} kr_impact_result;

/* n_search searches (1 .. 1024) of one spacetime (p's geometry).  lo, hi
 * (in: the starting bracket; out: the final one) and iterations (out) have
 * n_search entries; probe_b and probe_state have n_search * max(2, n_probe),
 * was_open n_search (written on an error only).  Host arrays; kernel_ms may
 * be NULL. */
kr_status kr_impact_search(kr_handle h, const kr_params* p, const kr_impact_opts* o, int32_t n_search,
                           const double* lz_sign, double* lo, double* hi, int64_t* iterations, double* probe_b,
                           int64_t* probe_state, uint8_t* was_open, kr_impact_result* res, double* kernel_ms);

/* ---------------------------------------------------------------------
 * Ragged batches of images in one persistent launch.  No counterpart in the
 * reference, which traces one image at a time (compute_shadow ->
 * trace_rays, src/kerrray/raytracing/rays.py:72-117); each image here is
 * integrated exactly as kr_trace_image integrates it.
 * ------------------------------------------------------------------- */

/* One image of a batch: its integration constants (float64, rk45, no disk
 * event) and its Bardeen image. */
typedef struct kr_job {
    kr_params params;
    kr_image image;
} kr_job;

/* Launch shape of the persistent kernel: threads per block (64, 128, 256),
 * resident wavefronts per SIMD the grid is sized for (the register budget
 * is set for 2 or >= 3), and the number of integration attempts between two
 * refills of a wavefront's idle lanes. */
typedef struct kr_tune {
    int32_t block_size;
    int32_t waves_per_simd;
    int32_t refill_interval;
    int32_t order; /* 0: images in job order; 1: smallest rtol first, then most pixels (outputs stay in job order) */
} kr_tune;

typedef struct kr_batch_s* kr_batch;

/* Uploads the job table and binds the device output arrays: rows are the
 * images' pixels in job order, each image row-major.  Up to 128 images the
 * launch keeps every image's constants in LDS; a larger batch reads them
 * from the job table in device memory.  At most 2^31 - 1 images. */
kr_status kr_batch_create(kr_handle h, const kr_job* jobs, int32_t n_jobs, const kr_tune* tune,
                          const kr_outputs* d_out, kr_batch* out);
/* Total photons (output rows) of the batch, or -1. */
int64_t kr_batch_photons(kr_batch b);
/* Enqueue one run (counter reset + persistent kernel) on `stream`. */
kr_status kr_batch_run(kr_batch b, void* stream);
/* Same, as one hipGraph launch (captured on first use, then replayed). */
kr_status kr_batch_graph_launch(kr_batch b, void* stream);
/* KR_FLAG_* bits of the last completed run (synchronous read). */
kr_status kr_batch_flags(kr_batch b, int32_t* flags);
kr_status kr_batch_destroy(kr_batch b);

/* ---------------------------------------------------------------------
 * Planned image traces: longest-first scheduling and the FP32 bulk with
 * FP64 repair.  No counterpart in the reference (whose float32 mode is a
 * separate working dtype, numba_backend.py:16-20); the FP64 form gives
 * exactly kr_trace_image's results.
 * ------------------------------------------------------------------- */

/* Options of a planned trace.
 *  band_px: pixels within band_px (Chebyshev) of the capture boundary are
 *    integrated in FP64: those within band_px + 1 pixels of the analytic
 *    edge from the start (with those of the polar strip, see axis_lz), and
 *    after the trace every pixel within band_px of a captured/not-captured
 *    transition of the mask, or whose photon ended neither captured nor
 *    escaped, that was not (repeated to a fixed point, KR_REPAIR_ROUNDS).
 *  fp32_bulk: 0 = every photon in FP64 (the longest-first schedule only);
 *    1 = other photons take FP32 right-hand sides with the FP64 update,
 *    error norm and step control, switching an attempt to FP64 when
 *    r < r_+ + horizon_margin, |sin theta| < axis_sin, or r < turn_rmax and
 *    |dr/dlambda| < turn_drdl.
 *  refill_interval, block_size: launch shape of the persistent kernel.
 *  longest_first: order of the work list: the FP64 band and polar strip
 *    first, then the other pixels inside the analytic edge, then those
 *    outside (the longest photons start first); 0: row-major.
 *  band_priority: raise the issue priority of wavefronts holding a photon of
 *    the nearest distance classes (the longest ones), which otherwise finish
 *    last.
 *  waves_per_simd: resident wavefronts per SIMD the persistent kernel is
 *    compiled for; more hide more FP64 latency, fewer registers each.  0:
 *    one up to 2^19 pixels for the FP64 form and the split FP32 form (their
 *    longest FP64 photons then run alone on their SIMD), two otherwise.
 *  split_roles (fp32_bulk only): the first ceil(band / 64) wavefronts trace
 *    only the FP64 band and strip, the others only FP32 photons, so no
 *    wavefront runs both attempts.  Needs longest_first (which puts the band
 *    first in the work list); ignored without it.  1: at one wavefront per
 *    SIMD and rtol >= 1e-7 each FP64 photon is carried by two lanes
 *    (ceil(band / 32) wavefronts), which share out its trigonometry: the
 *    same rows, a shorter longest photon; 2: always paired; 3: never paired.
 *  grid_fraction: share of the resident-wavefront slots (waves_per_simd per
 *    SIMD) the persistent kernel is launched on, in (0, 1].  0: 7/8 for the
 *    split FP32 form at one wavefront per SIMD up to 2^18 pixels and rtol
 *    >= 1e-6 (15/16 with paired FP64 photons; its FP32 photons then finish
 *    well before the longest FP64 photon, and fewer active wavefronts let the
 *    clock those photons run at rise), 1 otherwise. */
typedef struct kr_planned_opts {
    double horizon_margin;
    double axis_sin;
    double turn_rmax;
    double turn_drdl;
    int32_t band_px;
    int32_t fp32_bulk;
    int32_t refill_interval;
    int32_t block_size;
    int32_t longest_first; /* 1: work list band/strip -> inside -> outside; 0: row-major */
    int32_t band_priority; /* 1: wavefronts tracing the nearest photons issue first on their SIMD */
    int32_t waves_per_simd; /* 0 automatic, or 1-4: occupancy the trace kernel's registers are allocated for */
    int32_t split_roles;    /* 1-3: wavefronts trace FP64 or FP32 photons, never both (fp32_bulk only) */
    double axis_lz; /* pixels with |alpha sin theta_o| < axis_lz are integrated in FP64 (near-polar photons) */
    double grid_fraction; /* 0 automatic, or (0, 1]: share of the wavefront slots launched */
} kr_planned_opts;

typedef struct kr_planned_stats {
    int64_t planned_band;    /* pixels scheduled first and integrated in FP64 (band and polar strip) */
    int64_t planned_inside;  /* other pixels inside the analytic edge */
    int64_t planned_outside; /* other pixels outside it */
    int64_t repaired;        /* pixels the mask check sent back to FP64, over all repair rounds */
    int64_t fp64_attempts;   /* integration attempts taken in FP64 */
    int64_t repair_rounds;   /* repair rounds that had pixels to re-trace (0 .. KR_REPAIR_ROUNDS) */
    int64_t unrepaired;      /* pixels queued for a round past the limit (KR_FLAG_REPAIR when > 0) */
    int32_t flags;           /* KR_FLAG_* bits of the trace (initial states, repair limit) */
    int32_t reserved;
} kr_planned_stats;

/* Verify/repair of a planned trace: after the trace, every pixel not traced
 * in FP64 that failed or lies within band_px of a captured/not-captured
 * transition is re-traced in FP64 (round 1); each later round re-traces the
 * FP32 pixels within band_px of a pixel whose class the previous round
 * changed, until a round changes no class, for at most this many rounds. */
#define KR_REPAIR_ROUNDS 4

/* Device workspace of kr_trace_image_planned_dev for an image of `pixels`
 * pixels and an edge table of n_angles radii. */
size_t kr_planned_workspace_bytes(int64_t pixels, int32_t n_angles);

/* The analytic edge is given as radii[k] from (centre_alpha, centre_beta)
 * along the angle 2 pi k / n_angles (e.g. Bardeen's curve).  Host arrays;
 * stats and kernel_ms may be NULL. */
kr_status kr_trace_image_planned(kr_handle h, const kr_params* p, const kr_image* img, const kr_planned_opts* o,
                                 double centre_alpha, double centre_beta, const double* radii, int32_t n_angles,
                                 const kr_outputs* out, kr_planned_stats* stats, double* kernel_ms);
/* Same on device pointers (radii, outputs, workspace), enqueued on `stream`. */
kr_status kr_trace_image_planned_dev(kr_handle h, const kr_params* p, const kr_image* img, const kr_planned_opts* o,
                                     double centre_alpha, double centre_beta, const double* d_radii,
                                     int32_t n_angles, const kr_outputs* d_out, void* d_work, size_t work_bytes,
                                     void* stream);
/* Counters and KR_FLAG_* bits of the last planned trace that used the
 * workspace d_work, read after the work enqueued on `stream` (synchronises
 * that stream). */
kr_status kr_planned_read_stats(const void* d_work, size_t work_bytes, kr_planned_stats* stats, void* stream);

/* The precision form on a batch (FP32 bulk, FP64 where needed, as
 * kr_trace_image_planned with fp32_bulk = 1): each image's photons within
 * band_px + 1 pixels of its analytic edge, or with |alpha sin theta_o| <
 * axis_lz, take the exact FP64 attempt, the others FP32 right-hand sides
 * with the FP64 update, error norm and step control.  The two kinds run as
 * two persistent launches side by side (the FP32 one on a stream the batch
 * owns, forked from and joined back to the caller's stream); after them
 * every image is verified and repaired in FP64 to a fixed point
 * (KR_REPAIR_ROUNDS), all in the same kr_batch_run / graph launch, so the
 * caller's stream sees one ordered run.  Image k
 * (in job order) has its edge as n_angles radii radii[k * n_angles + j]
 * about (centre_alpha[k], centre_beta[k]) along the angles 2 pi j /
 * n_angles.  Host arrays, copied at creation. */
typedef struct kr_batch_precision {
    int32_t band_px;
    int32_t n_angles;
    double axis_lz;
    double fp32_min_rtol; /* images with a smaller rtol are traced in FP64 throughout: FP32 stages hold their
                             error estimate near FP32 round-off, which tighter tolerances cannot accept */
    const double* centre_alpha;
    const double* centre_beta;
    const double* radii;
} kr_batch_precision;
kr_status kr_batch_create_precision(kr_handle h, const kr_job* jobs, int32_t n_jobs, const kr_tune* tune,
                                    const kr_batch_precision* prec, const kr_outputs* d_out, kr_batch* out);
/* Counters of the last run of a precision batch (synchronises the device):
 * FP64 band photons, repaired pixels and rounds, FP64 attempts, flags
 * (planned_inside and planned_outside are -1). */
kr_status kr_batch_stats(kr_batch b, kr_planned_stats* stats);

/* kr_trace_image_planned_dev without a caller workspace and without any
 * device allocation: the plan, work list, repair lists and counters live in
 * a static device array of the library (17 MiB on every device the library
 * runs on, held until the process ends), whose address kr_create finds once
 * per device.  No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Images of at most
 * KR_NOALLOC_MAX_PIXELS pixels and edge tables of at most
 * KR_NOALLOC_MAX_ANGLES radii; larger ones are rejected with
 * KR_ERR_INVALID.  Every call on a device shares that device's array, so
 * calls that overlap on one device (other streams, other handles, other
 * threads) race; calls in one stream are ordered and safe.  The trace's
 * KR_FLAG_* bits are ORed into *d_flags (caller-zeroed device word; may be
 * NULL).  Results equal kr_trace_image_planned_dev's bit for bit. */
#define KR_NOALLOC_MAX_PIXELS (1 << 20)
#define KR_NOALLOC_MAX_ANGLES 8192
kr_status kr_trace_image_planned_noalloc(kr_handle h, const kr_params* p, const kr_image* img,
                                         const kr_planned_opts* o, double centre_alpha, double centre_beta,
                                         const double* d_radii, int32_t n_angles, const kr_outputs* d_out,
                                         int32_t* d_flags, void* stream);

/* ---------------------------------------------------------------------
 * Differentiation form: escape-direction gradients by the discrete adjoint
 * of Dormand-Prince 5(4).  No counterpart in the reference (which has no
 * derivatives of the ray map); inputs are image-plane points as for
 * kr_trace_points_dev.
 * ------------------------------------------------------------------- */
typedef struct kr_tape_s* kr_tape;

/* Traces the n points (alpha, beta) of image `img` in FP64 (adaptive DP5,
 * no events) into d_out, exactly as kr_trace_points_dev, and records every
 * escaped photon's accepted steps:
 *   segment 0: the Jacobian tape.  Every accepted step's stage states are
 *     recorded and turned, one step per lane, into the step's Jacobian,
 *     through which every photon's 2 x 4 Jacobian d(theta_inf, phi_inf)/
 *     d(alpha, beta, spin, theta_o) is accumulated (632 B of device memory
 *     per accepted step while building, 64 B per photon kept); a
 *     vector-Jacobian product is then J^T g per photon.  Fails with
 *     KR_ERR_DEVICE if the device lacks the memory.
 *   segment 1..4096: the step sizes and a checkpoint of the state every
 *     `segment` accepted steps (56 B per step); each vector-Jacobian product
 *     recomputes the steps from the checkpoints and sweeps them backwards.
 * d_theta_inf / d_phi_inf (n each) receive
 * the escape direction at r = infinity: (theta, phi) carried from the final
 * state to u = 1/r = 0 along the geodesic (NaN for photons that did not
 * escape).  The map differentiated later is the sequence of accepted steps
 * with those step sizes frozen, followed by that tail.  p's geometry must be
 * the image's (mass, a = spin * mass).  Synchronous; the tape owns copies of
 * alpha, beta and the final states. */
kr_status kr_tape_create(kr_handle h, const kr_params* p, const kr_image* img, const double* d_alpha,
                         const double* d_beta, int64_t n, int32_t segment, const kr_outputs* d_out,
                         double* d_theta_inf, double* d_phi_inf, kr_tape* tape);
/* Photons on the tape (the escaped ones), accepted steps recorded, the
 * recording kernel's time and the Jacobian accumulation kernels' time (0 for
 * segment >= 1); -1 for a NULL tape. */
int64_t kr_tape_recorded(kr_tape t);
int64_t kr_tape_steps(kr_tape t);
double kr_tape_record_ms(kr_tape t);
double kr_tape_jacobian_ms(kr_tape t);
/* Vector-Jacobian product: for every point i, d_grad[4 i + j] =
 * g_theta[i] d theta_inf / d x_j + g_phi[i] d phi_inf / d x_j with
 * x = (alpha, beta, spin, theta_o); 0 for photons that did not escape.
 * Device arrays, enqueued on `stream` (NULL is the default stream, as for
 * the other _dev calls).  Calls on one tape must not overlap (they share
 * its work buffers). */
kr_status kr_tape_vjp(kr_tape t, const double* d_g_theta, const double* d_g_phi, double* d_grad, void* stream);
/* Flags of the tape's runs (synchronises the device): bit 8 = a
 * re-integration differed from the trace, bit 9 = a recomputed final state
 * differed from the traced one. */
kr_status kr_tape_flags(kr_tape t, int32_t* flags);
kr_status kr_tape_destroy(kr_tape t);

/* Escape direction at r = infinity of n FP64 result rows (d_Y (n, 8), d_state
 * (n) as written by the trace entry points, for the geometry in p): theta
 * and phi carried from each escaped photon's final state along its geodesic
 * to u = 1/r = 0; NaN for rows that did not escape.  Enqueued on `stream`. */
kr_status kr_escape_directions_dev(kr_handle h, const kr_params* p, const double* d_Y, const int64_t* d_state,
                                   int64_t n, double* d_theta_inf, double* d_phi_inf, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* KERRRAY_HIP_H */
