/*
 * kete_hip.h -- C ABI of libkete_hip.so, the MI355X (gfx950) implementation of
 * kete's batched orbit-propagation path.
 *
 * The reference (Caltech-IPAC/kete 1.1.3) is Rust with no FFI of its own; each
 * entry point below replaces the per-object loop body of one of its batch
 * entry points, called once per batch instead of once per object from rayon.
 * Paths are relative to the kete source tree.
 *
 * Conventions
 *  - Host-pointer entry points copy inputs to the device, run, copy results
 *    back and synchronise before returning (kete's calls are synchronous).
 *    The *_dev variants take device pointers and run on the context's stream;
 *    they return once the work is enqueued unless `sync` is non-zero.
 *  - States are structure-of-arrays: xyz[0..2] position (AU), xyz[3..5]
 *    velocity (AU/day), one array per component, object index fastest.
 *  - Frames use kete's SPK numbering: 1 = Equatorial (J2000), 17 = Ecliptic.
 *    Centers are NAIF ids (0 = solar-system barycentre, 10 = Sun).
 *  - Per-object status mirrors kete_core::errors::Error (errors.rs:14-32):
 *    0 ok, 1 Convergence, 2 ValueError, 3 DAFLimits, 4 UnknownFrame,
 *    5 IOError, 6 Impact.  A failed object's outputs are NaN, as the pyo3
 *    bindings return State::new_nan (src/kete/rust/propagation.rs:115-147).
 *  - The return value is an API status: KH_OK or a negative KH_E_* code; the
 *    message is available from kh_last_error().
 */
#ifndef KETE_HIP_H
#define KETE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kh_ctx kh_ctx;

enum {
    KH_OK = 0,
    KH_E_ARG = -1,      /* invalid argument                            */
    KH_E_HIP = -2,      /* HIP runtime failure                         */
    KH_E_NO_SPK = -3,   /* no ephemeris uploaded / body missing        */
    KH_E_NO_DEVICE = -4 /* no GPU visible                              */
};

enum { KH_FRAME_EQUATORIAL = 1, KH_FRAME_ECLIPTIC = 17 };

/* Per-object work counters of the Radau path. */
typedef struct {
    int32_t n_force; /* force evaluations                          */
    int32_t n_sweep; /* predictor-corrector sweeps (incl. failed)  */
    int32_t n_acc;   /* accepted steps                             */
    int32_t n_fail;  /* failed step attempts                       */
} kh_counters;

/* JPL comet / dust non-gravitational model (propagation/nongrav.rs:12-61).
 * kind: 0 none, 1 JplComet (a1..k), 2 Dust (beta).  The JplComet time delay
 * `dt` must be 0 (a non-zero delay is rejected with KH_E_ARG). */
typedef struct {
    int32_t kind;
    double a1, a2, a3, alpha, r_0, m, n, k, dt, beta;
} kh_nongrav;

/* One SPK Type 2 segment (spice/spk_segments.rs:316-397).  `records` is the
 * segment's DAF array without its 4-double trailer: n_rec records of rsize
 * doubles [mid, radius, x[n], y[n], z[n]] (km, TDB seconds past J2000).
 * jd_start/jd_end are the summary times converted with spice_jds_to_jd
 * (spice/mod.rs:25-27); intlen_s is the trailer's INTLEN. */
typedef struct {
    int64_t obj_id;
    int64_t center_id;
    int32_t frame;
    int32_t rsize;
    int64_t n_rec;
    double jd_start;
    double jd_end;
    double intlen_s;
    const double *records;
} kh_spk2_seg;

/* One SPK Type 21 segment (spice/spk_segments.rs:958-1102): extended
 * modified difference arrays, as JPL Horizons delivers small bodies.
 * `records` holds n_rec records of 4*maxdim+11 doubles in kete's layout
 * [t_ref, G[maxdim], x vx y vy z vz, DT[3][maxdim], KQMAX1, KQ[3]] (km, TDB
 * seconds past J2000; spk_segments.rs:1010-1018), `epochs` the n_rec final
 * epochs that follow them in the DAF array; maxdim is the array trailer's
 * MAXDIM (spk_segments.rs:971-986).  jd_start/jd_end are the summary times
 * converted with spice_jds_to_jd (spice/mod.rs:25-27).  Centers 10 and 0 and
 * frames 1 and 17 are accepted. */
typedef struct {
    int64_t obj_id;
    int64_t center_id;
    int32_t frame;
    int32_t maxdim;
    int64_t n_rec;
    double jd_start;
    double jd_end;
    const double *records;
    const double *epochs;
} kh_spk21_seg;

/* Device selection and stream ownership.  Replaces nothing in kete (its
 * state is the process-global LOADED_SPK, spice/spk.rs:368-377); one context
 * per host thread. */
int32_t kh_ctx_create(int32_t device, kh_ctx **out);
int32_t kh_ctx_destroy(kh_ctx *ctx);
const char *kh_last_error(void);
/* hipStream_t of the context (as void*), for callers that enqueue around it. */
void *kh_ctx_stream(kh_ctx *ctx);

/* Upload the Type 2 ephemeris.  Replaces SpkCollection::load_file for the
 * planet segments kete's spk_accel reads (spice/spk.rs:307-329,65-82): the
 * host side walks SpkCollection.planet_segments and passes them here. */
int32_t kh_spk_upload_type2(kh_ctx *ctx, const kh_spk2_seg *segs, int64_t nseg);

/* Upload the Type 21 segments (replacing those uploaded before; nseg may be
 * 0), in load order.  Replaces SpkCollection::load_file for the segments it
 * files under SpkCollection.segments (spice/spk.rs:307-329) and the
 * From<DafArray> view of SpkSegmentType21 (spice/spk_segments.rs:971-986).
 * Segments of kete's five MASSES asteroids (20000001, 20000002, 20000004,
 * 20000010, 20000704; constants/gravity.rs:103-134) are the perturbers of
 * include_asteroids; every object's are readable with kh_spk_states.
 * Records whose difference order exceeds 16, and records kete could not
 * evaluate (a zero step size, spk_segments.rs:1037-1042), are rejected here
 * with KH_E_ARG.  At most 32 segments. */
int32_t kh_spk_upload_type21(kh_ctx *ctx, const kh_spk21_seg *segs, int64_t nseg);

/* State of `obj_id` relative to `center` in `frame` at m epochs, evaluated on
 * the device (SpkCollection::try_get_state, spice/spk.rs:87-96).
 * out: [6][m]; status: [m]. */
int32_t kh_spk_states(kh_ctx *ctx, int64_t obj_id, int64_t center, int32_t frame,
                      int64_t m, const double *jd, double *const out[6], int32_t *status);

/* Two-body propagation of n states to m target epochs.  Replaces the loop
 * body of propagation_kepler_py (src/kete/rust/kepler.rs:68-108): center to
 * the Sun, analytic_2_body (src/kete_core/src/propagation/kepler.rs:254-322),
 * optional one-step light delay against observer_pos, center back.
 * center must be 10 or 0.  out: [6][m][n]; status: [m][n]. */
int32_t kh_propagate_two_body(kh_ctx *ctx, int64_t n, const double *jd0,
                              const double *const xyz[6], int64_t center, int32_t frame,
                              int64_t m, const double *jd_targets,
                              const double *observer_pos, double *const out[6],
                              int32_t *status);
/* Device-pointer variant; newton_iters (nullable, [m][n]) receives the
 * Newton iterations each solve took (both solves with an observer). */
int32_t kh_propagate_two_body_dev(kh_ctx *ctx, int64_t n, const double *jd0,
                                  const double *const xyz[6], int64_t center, int32_t frame,
                                  int64_t m, const double *jd_targets,
                                  const double *observer_pos, double *const out[6],
                                  int32_t *status, int32_t *newton_iters, int32_t sync);

/* N-body propagation with the Gauss-Radau 15 integrator and kete's SPK force
 * model.  Replaces the loop body of propagation_n_body_spk_py
 * (src/kete/rust/propagation.rs:70-151) around propagate_n_body_spk
 * (src/kete_core/src/propagation/mod.rs:70-118).
 * jd_final: n entries, or 1 entry broadcast when jd_final_stride is 0.
 * include_asteroids: non-zero sums kete's MASSES (PLANETS and five asteroids
 * read from the Type 21 segments) instead of PLANETS
 * (propagation/mod.rs:82-88); without their segments every object ends
 * with DAFLimits, as in kete.
 * epsilon: Radau EPSILON; 1e-6 reproduces kete (radau.rs:71).
 * mode: 0 plain launch (one lane per object), 1 persistent launch with a
 * device work queue, 2 persistent with three lanes per object (one per
 * coordinate), 3 plain with three lanes per object, -1 let the library
 * choose (0).  Modes 0-2 hand out the objects longest first (|span| times
 * mean motion), so the slowest arcs start at once.
 * ng, counters, impact_id, impact_jd are nullable. */
int32_t kh_propagate_n_body(kh_ctx *ctx, int64_t n, const double *jd0,
                            const double *const xyz[6], int64_t center, int32_t frame,
                            const double *jd_final, int32_t jd_final_stride,
                            int32_t include_asteroids, const kh_nongrav *ng, double epsilon,
                            int32_t mode,
                            double *const out[6], int32_t *status, int64_t *impact_id,
                            double *impact_jd, kh_counters *counters);
/* Device-pointer variant; ng points to device memory (nullable), counters is
 * a device array of n kh_counters (nullable), impact_id is int64 on device. */
int32_t kh_propagate_n_body_dev(kh_ctx *ctx, int64_t n, const double *jd0,
                                const double *const xyz[6], int64_t center, int32_t frame,
                                const double *jd_final, int32_t jd_final_stride,
                                int32_t include_asteroids, const kh_nongrav *ng,
                                double epsilon, int32_t mode,
                                double *const out[6], int32_t *status, int64_t *impact_id,
                                double *impact_jd, kh_counters *counters, int32_t sync);

/* Largest batch of kh_propagate_n_body_dev_noalloc. */
#define KH_NOALLOC_MAX_N 2097152

/* kh_propagate_n_body_dev that allocates no device memory.  No kete
 * counterpart: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Arguments,
 * kernels and results are those of kh_propagate_n_body_dev; the scratch of
 * the longest-first object order (at most KH_NOALLOC_MAX_N objects, 12.6 MB)
 * is a device array built into the library, one per GPU, held until the
 * process ends.  Calls on one GPU share that array, so calls that overlap
 * there (several contexts or threads on different streams) race; calls on
 * one context's stream are ordered.  n > KH_NOALLOC_MAX_N returns KH_E_ARG.
 * The context itself (created by kh_ctx_create, with the ephemeris already
 * uploaded) is set up beforehand; this call allocates nothing. */
int32_t kh_propagate_n_body_dev_noalloc(kh_ctx *ctx, int64_t n, const double *jd0,
                                        const double *const xyz[6], int64_t center,
                                        int32_t frame, const double *jd_final,
                                        int32_t jd_final_stride, int32_t include_asteroids,
                                        const kh_nongrav *ng, double epsilon, int32_t mode,
                                        double *const out[6], int32_t *status,
                                        int64_t *impact_id, double *impact_jd,
                                        kh_counters *counters, int32_t sync);

/* Edge normals of a rectangular field (OnSkyRectangle::new,
 * src/kete_core/src/fov/patches.rs:122-155; RectangleFOV's constructor
 * src/kete/rust/fovs/definitions.rs:401-418 converts degrees to radians
 * first).  Angles in radians; normals: [4][3].  Host-only, no device work. */
void kh_fov_rect_normals(const double pointing[3], double rotation, double lon_width,
                         double lat_width, double normals[12]);

/* Visibility of n objects in rectangular fields with light time.  Replaces
 * fov_checks_py (src/kete/rust/fovs/checks.rs:24-110) and the per-field
 * FovLike::check_visible (src/kete_core/src/fov/fov_like.rs:134-192):
 * fields are sorted by epoch (stably), grouped within 2*dt_limit, the
 * objects re-propagated with the N-body integrator on kete's big/small step
 * schedule, and every (object, field) pair checked.  States: n objects at
 * epoch jd0, Sun-centred; fields: epoch, observer state [6][n_fields]
 * (Sun-centred) and normals [n_fields][12]; everything in the ecliptic frame
 * (17), the frame kete converts fields to (collection.rs:23-36).
 * Outputs, ordered by (field epoch order, object index) as kete returns them:
 * the original field index, the object index, and the light-time-corrected
 * state [7][cap] (jd, pos, vel).  *n_hits receives the count; a count above
 * cap returns KH_E_ARG with *n_hits set to the size needed.
 * include_asteroids: kete's flag of the same name (checks.rs:25-31): every
 * N-body propagation sums MASSES instead of PLANETS. */
int32_t kh_fov_state_check(kh_ctx *ctx, int64_t n, double jd0, const double *const xyz[6],
                           int32_t frame, int64_t n_fields, const double *field_jd,
                           const double *const obs[6], const double *normals, double dt_limit,
                           int32_t include_asteroids,
                           int64_t cap, int64_t *n_hits, int64_t *hit_field, int64_t *hit_obj,
                           double *const hit_state[7]);

/* SPK objects in rectangular fields with light time.  Replaces
 * fov_spk_checks_py (src/kete/rust/fovs/checks.rs:123-142) and the per-field
 * FovLike::check_spks (src/kete_core/src/fov/fov_like.rs:196-226): for every
 * field (sorted by epoch, stably) and object id, the object's state from the
 * uploaded ephemeris (Type 21 segments or the planet tree) at the field
 * epoch, Sun-centred in the field frame (ecliptic, 17); ids with no state
 * there are skipped; then the two-body light-time step and the field test.
 * Fields as for kh_fov_state_check.  Outputs in kete's order (field epoch
 * order, then id order): the original field index, the index into obj_ids,
 * and the state [7][cap] (jd, pos, vel).  *n_hits receives the count; a
 * count above cap returns KH_E_ARG with *n_hits set to the size needed. */
int32_t kh_fov_spk_check(kh_ctx *ctx, int64_t n_ids, const int64_t *obj_ids, int32_t frame,
                         int64_t n_fields, const double *field_jd, const double *const obs[6],
                         const double *normals, int64_t cap, int64_t *n_hits,
                         int64_t *hit_field, int64_t *hit_obj, double *const hit_state[7]);

/* RA, Dec, dRA/dt*cos(Dec), dDec/dt (radians, radians/day) and V magnitude
 * of n detections.  Replaces SimultaneousStates::ra_dec_with_rates
 * (src/kete_core/src/simult_states.rs:86-120) and HGParams::apparent_mag
 * (src/kete_core/src/flux/reflected.rs:255-268).  state and obs: [6][n],
 * Sun-centred, in `frame`; h_mag: [n]; out: [5][n]. */
int32_t kh_radec_vmag(kh_ctx *ctx, int64_t n, const double *const state[6],
                      const double *const obs[6], int32_t frame, const double *h_mag,
                      double g_param, double *const out[5]);

/* Astrometric observables of n orbits at m epochs each and their partial
 * derivatives with respect to 9 parameters: x, y, z, vx, vy, vz at jd0
 * (Sun-centred, in `frame`) and A1, A2, A3 of the JplComet non-grav model.
 * kete has no derivatives of propagate_n_body and no counterpart of this
 * entry point; its observable chain is FovLike::check_n_body's
 * (src/kete_core/src/fov/fov_like.rs:89-110): propagate_n_body_spk to each
 * epoch (a new integrator per epoch, as successive propagate_n_body calls
 * make), one two-body step back by the light time to the observer, then
 * RA/Dec of the observer-to-object vector in the equatorial frame
 * (src/kete_core/src/simult_states.rs:86-120).
 * ng: [n], nullable; kind 0 is read as JplComet with the new_asteroid
 * defaults and A1 = A2 = A3 = 0; kind 2 (dust) and JplComet with dt != 0
 * are rejected.  jd_obs and obs[3]: [m][n]; obs is the observer position,
 * Sun-centred, in `frame`.  force: 0 for kete's PLANETS model from the
 * uploaded ephemeris, 1 for central_accel
 * (src/kete_core/src/propagation/acceleration.rs:66-80: the Sun fixed at
 * the origin, no ephemeris needed).  restart: 1 starts a new integrator at
 * each epoch (B = 0, h = 0.1 d), as successive propagate_n_body calls do; 0
 * integrates once through all epochs, shortening the step that reaches each
 * epoch and keeping the integrator state.  epsilon: Radau EPSILON (1e-6 as
 * kete).
 * Outputs: radec [2][m][n] (radians); jac [2][9][m][n] (dRA, dDec per
 * parameter); state [6][m][n] and dstate [6][9][m][n] (nullable; dstate
 * needs state): the geometric Sun-centred state at each epoch, in `frame`,
 * and its partials; status [n] (kete's error codes; NaN outputs from the
 * failing epoch on); counters [n] (nullable). */
int32_t kh_radec_partials(kh_ctx *ctx, int64_t n, const double *jd0,
                          const double *const xyz[6], int32_t frame, const kh_nongrav *ng,
                          int32_t m, const double *jd_obs, const double *const obs[3],
                          int32_t force, int32_t restart, double epsilon, double *const radec[2],
                          double *const jac[2], double *const state[6],
                          double *const dstate[6], int32_t *status, kh_counters *counters);
/* Device-pointer variant (all arrays in device memory).  Non-grav models are
 * checked per object on the device: a dust model, a JplComet delay or an
 * unknown kind ends that object with ValueError (2) instead of KH_E_ARG. */
int32_t kh_radec_partials_dev(kh_ctx *ctx, int64_t n, const double *jd0,
                              const double *const xyz[6], int32_t frame, const kh_nongrav *ng,
                              int32_t m, const double *jd_obs, const double *const obs[3],
                              int32_t force, int32_t restart, double epsilon,
                              double *const radec[2],
                              double *const jac[2], double *const state[6],
                              double *const dstate[6], int32_t *status, kh_counters *counters,
                              int32_t sync);

/* propagate_n_body_long: particles integrated together with kete's
 * SIMPLE_PLANETS (Sun, Mercury, Venus, Earth-Moon barycentre, Mars .. Neptune)
 * as one N-body system, no ephemeris during the integration.  Replaces
 * propagation_n_body_py (src/kete/rust/propagation.rs:202-246): the states
 * in consecutive batches of batch_size, each batch with its own copy of the
 * planets integrated by propagate_n_body_vec
 * (src/kete_core/src/propagation/mod.rs:142-237) with vec_accel
 * (src/kete_core/src/propagation/acceleration.rs:197-249).  States: n
 * objects at jd0, center 10 (Sun), `frame`; planet_xyz: [9][6] Sun-centred
 * equatorial at jd0, or null for the ephemeris' states.  out: [6][n]
 * Sun-centred EQUATORIAL at jd_final, as kete returns them; planets_out
 * ([9][6], nullable): the last batch's planets.  *status: kete's error for
 * the call (0 ok; 2 ValueError for an empty batch or a center other than 10;
 * 1 Convergence or 6 Impact of the first failing batch; 3 DAFLimits when the
 * ephemeris lacks a planet), out undefined unless 0.  batch_size 1 .. 55
 * (kete's default is 10). */
int32_t kh_propagate_n_body_long(kh_ctx *ctx, int64_t n, double jd0, const double *const xyz[6],
                                 int64_t center, int32_t frame, double jd_final,
                                 const double *planet_xyz, const kh_nongrav *ng,
                                 int32_t batch_size, double *const out[6], double *planets_out,
                                 int32_t *status);
/* kh_propagate_n_body_long on device pointers (xyz, ng, out; planet_xyz stays
 * a host array or null): every batch's status in batch_status[nb] and, when
 * not null, its planets at jd_final in batch_planets[nb][9][6] and its
 * integrator work in batch_counters[nb]; nb = ceil(n / batch_size).  An empty
 * batch or a center other than 10 is KH_E_ARG here; kete's call-level error
 * is the first non-zero batch_status. */
int32_t kh_propagate_n_body_long_dev(kh_ctx *ctx, int64_t n, double jd0,
                                     const double *const xyz[6], int64_t center, int32_t frame,
                                     double jd_final, const double *planet_xyz,
                                     const kh_nongrav *ng, int32_t batch_size,
                                     double *const out[6], double *batch_planets,
                                     int32_t *batch_status, kh_counters *batch_counters,
                                     int32_t sync);

/* MOID (minimum orbit intersection distance, AU) of n orbit pairs.  Replaces
 * moid (src/kete_core/src/propagation/kepler.rs:325-404) called per pair, and
 * moid_py (src/kete/rust/propagation.rs:27-39) for its default second state:
 * pair i is state a (jd_a[i], xyz_a[.][i]) and state b (jd_b[i], or jd_a[i]
 * when jd_b is null; xyz_b[.][i]), both Sun-centred in `frame`; with xyz_b
 * null, b is the Earth (399) at jd_a[i] from the uploaded ephemeris.
 * moid[i]: the result (NaN on error); status[i]: 0, or 1 Convergence when a
 * two-body propagation fails (kete's moid returns that error; a failure
 * inside the Nelder-Mead search, where kete panics, is reported the same
 * way), 3 DAFLimits when the Earth is not covered.  Host pointers. */
int32_t kh_moid(kh_ctx *ctx, int64_t n, const double *jd_a, const double *const xyz_a[6],
                int32_t frame, const double *jd_b, const double *const xyz_b[6], double *moid,
                int32_t *status);
/* kh_moid on device pointers; optional outputs: grid_min[n] (the closest of
 * the 100 x 100 grid positions), iters[n] (Nelder-Mead iterations, kete's
 * count), work[2][n] (Kepler solves and their Newton iterations per pair).
 * sync = 0 returns with the work queued on the library's stream. */
int32_t kh_moid_dev(kh_ctx *ctx, int64_t n, const double *jd_a, const double *const xyz_a[6],
                    int32_t frame, const double *jd_b, const double *const xyz_b[6],
                    double *moid, int32_t *status, double *grid_min, int32_t *iters,
                    int32_t *work, int32_t sync);

/* Launch shape picked by kh_autotune_n_body. */
typedef struct {
    int32_t mode;       /* 0-3 as kh_propagate_n_body's mode */
    int32_t block;      /* objects (mode 0/1) or lanes (2/3) per block */
    int32_t window;     /* records per segment in the LDS ephemeris cache (0: none) */
    int32_t n_tried;    /* shapes measured */
    int64_t latency_k;  /* most objects the latency kernel takes (mode 0) */
    int64_t n_sample;   /* objects the shapes were timed on */
    double ms;          /* best time of one call on the sample */
    double default_ms;  /* the library default's time on the sample */
} kh_tune_result;

/* Autotuner (no kete counterpart; kete has one CPU code path).  Times the
 * N-body call (device pointers, arguments as kh_propagate_n_body_dev) on
 * every (n/sample)-th object of the batch (all when sample is 0) over
 * launch shapes -- threads per object and plain or persistent launch
 * (mode), objects per block, the LDS ephemeris window, the latency kernel's
 * share -- by coordinate descent from the default, and reports the fastest.
 * With apply != 0 the context uses it for every later mode -1 call until
 * kh_autotune_clear.  Results of a call do not depend on the shape beyond
 * last-ulp differences between the one- and three-lane kernels.  Allocates
 * and frees a scratch buffer of 20 x 8 x sample bytes. */
int32_t kh_autotune_n_body(kh_ctx *ctx, int64_t n, const double *jd0, const double *const xyz[6],
                           int64_t center, int32_t frame, const double *jd_final,
                           int32_t jd_final_stride, int32_t include_asteroids, int64_t sample,
                           int32_t apply, kh_tune_result *out);
int32_t kh_autotune_clear(kh_ctx *ctx);

/* Launch geometry last used by kh_propagate_n_body*: threads per block,
 * blocks, and mode actually run. */
int32_t kh_last_launch(kh_ctx *ctx, int32_t *block, int32_t *grid, int32_t *mode);

/* Average duration in milliseconds of the last n kernel launches of the
 * N-body path, measured with HIP events on the context's stream. */
double kh_last_kernel_ms(kh_ctx *ctx);

/* Prefilter phase of the last kh_fov_state_check on this context (a
 * measurement hook; kete has no counterpart): total milliseconds of the
 * timed prefilter launches (HIP events on the context's stream, the first
 * 4096 launches of the call), how many were timed, and the (state, field)
 * pairs they scanned.  KH_E_ARG when no FOV call has completed. */
int32_t kh_fov_last_prefilter(kh_ctx *ctx, double *ms, int64_t *launches, int64_t *pairs);

#ifdef __cplusplus
}
#endif
#endif
