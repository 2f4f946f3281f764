/*
 * kete_oracle.h -- CPU restatement of kete's batched orbit-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in kete_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function restates a Rust function of kete 1.1.3 (Caltech-IPAC/kete);
 * the citation (path:line, relative to the kete source root) is given next to
 * each declaration and again at the definition.
 */
#ifndef KETE_ORACLE_H
#define KETE_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-object status, mirrors kete_core::errors::Error (errors.rs:14-32). */
enum {
    KO_OK = 0,
    KO_CONVERGENCE = 1,
    KO_VALUE_ERROR = 2,
    KO_DAF_LIMITS = 3,
    KO_UNKNOWN_FRAME = 4,
    KO_IO_ERROR = 5,
    KO_IMPACT = 6,
};

/* Frames as kete numbers them in SPK files (spk_segments.rs:139-143). */
enum { KO_FRAME_EQUATORIAL = 1, KO_FRAME_ECLIPTIC = 17 };

typedef struct {
    double jd;
    double pos[3];
    double vel[3];
    int32_t frame;
    int64_t center;
} ko_state;

/* Non-gravitational model (propagation/nongrav.rs:12-61). kind 0 = none,
 * 1 = JplComet, 2 = Dust. */
typedef struct {
    int32_t kind;
    double a1, a2, a3, alpha, r_0, m, n, k, dt;
    double beta;
} ko_nongrav;

/* Work counters per propagated object (not in the reference; used for the
 * flop model of the Radau path). */
typedef struct {
    int64_t n_force;   /* force evaluations                        */
    int64_t n_sweep;   /* predictor-corrector sweeps (incl failed) */
    int64_t n_acc;     /* accepted steps                           */
    int64_t n_fail;    /* failed step attempts                     */
} ko_counters;

/* ---- constants ---------------------------------------------------------- */
double ko_const(const char *name);

/* ---- fitting/newton.rs:18-65 (test hook on x^2-1) ----------------------- */
int ko_newton_square_root_test(double start, double atol, double *root);

/* ---- propagation/kepler.rs ---------------------------------------------- */
int ko_analytic_2_body(double dt, const double pos[3], const double vel[3],
                       double pos_out[3], double vel_out[3]);
int ko_compute_eccentric_anomaly(double ecc, double mean_anom, double peri_dist,
                                 double *out);
int ko_elements_to_state(double epoch, double ecc, double incl, double lon_asc,
                         double peri_arg, double peri_dist, double peri_time,
                         double pos[3], double vel[3]);

/* ---- spice --------------------------------------------------------------- */
/* Load a DAF/SPK file into the oracle's global collection (spk.rs:307-329). */
int ko_spk_load_file(const char *path);
void ko_spk_reset(void);
int ko_spk_n_segments(void);
int ko_spk_segment_info(int idx, int64_t *obj_id, int64_t *center_id,
                        int32_t *frame, int32_t *seg_type, double *jd_start,
                        double *jd_end);
/* try_get_state(id, jd, center, frame) (spk.rs:87-96). */
int ko_spk_get_state(int64_t id, double jd, int64_t center, int32_t frame,
                     double pos[3], double vel[3]);
/* SpkSegmentType21::try_get_pos_vel on one record (spk_segments.rs:1007-1100):
 * rec is kete's record layout with n = MAXDIM, js in TDB seconds past J2000;
 * position in AU, velocity in AU/day. */
int ko_spk21_record_eval(const double *rec, int64_t n, double js, double pos[3],
                         double vel[3]);
/* chebyshev_evaluate_both (interpolation.rs:25-77). */
int ko_chebyshev_evaluate_both(double x, const double *cx, const double *cy,
                               const double *cz, int n, double val[3],
                               double der[3]);
/* State center change through the loaded SPK (spk.rs:99-126). */
int ko_change_center(ko_state *s, int64_t new_center);
int ko_change_frame(ko_state *s, int32_t new_frame);

/* ---- forces -------------------------------------------------------------- */
/* spk_accel with PLANETS (acceleration.rs:121-168). */
int ko_spk_accel(double t, const double pos[3], const double vel[3],
                 const ko_nongrav *ng, int exact_eval, double acc[3],
                 int64_t *impact_id);
/* spk_accel with PLANETS, or with MASSES (the planets, the Moon and five
 * asteroids, constants/gravity.rs:43-134) when include_asteroids. */
int ko_spk_accel_ext(double t, const double pos[3], const double vel[3],
                     const ko_nongrav *ng, int exact_eval, int include_asteroids,
                     double acc[3], int64_t *impact_id);

/* ---- integrator ---------------------------------------------------------- */
/* RadauIntegrator::integrate with central_accel (radau.rs:160-218,
 * acceleration.rs:66-80). */
int ko_radau_central(double t0, double t1, const double pos[3],
                     const double vel[3], double pos_out[3], double vel_out[3],
                     ko_counters *cnt);
/* propagate_n_body_spk (propagation/mod.rs:70-118). */
/* Test hook: Radau with one PLANETS body (naif_id) at rest at the origin,
 * GravParams::add_acceleration as the force; equatorial states. */
int ko_radau_one_body(int64_t naif_id, double t0, double t1, const double pos[3],
                      const double vel[3], double pos_out[3], double vel_out[3]);
int ko_propagate_n_body(ko_state *s, double jd_final, const ko_nongrav *ng,
                        ko_counters *cnt, int64_t *impact_id, double *impact_jd);
/* ... with kete's include_extended flag (propagation/mod.rs:73,82-88). */
int ko_propagate_n_body_ext(ko_state *s, double jd_final, int include_asteroids,
                            const ko_nongrav *ng, ko_counters *cnt, int64_t *impact_id,
                            double *impact_jd);
/* propagate_two_body (kepler.rs:306-322). */
int ko_propagate_two_body(ko_state *s, double jd_final);

/* ---- batch drivers (mirror the pyo3 bindings' rayon loops) --------------- */
/* propagation_kepler_py (src/kete/rust/kepler.rs:68-108), SoA in/out.
 * in: jd0[n], xyz[6][n], center, frame; targets[m]; observer (nullable).
 * out: out[6][m][n], status[m][n]. */
int ko_batch_two_body(int64_t n, const double *jd0, const double *const *xyz,
                      int64_t center, int32_t frame, int64_t m,
                      const double *targets, const double *observer_pos,
                      double *const *out, int32_t *status, int n_threads);
/* propagation_n_body_spk_py (src/kete/rust/propagation.rs:70-151).
 * jd_final has n entries (or 1 if jd_final_stride == 0). */
int ko_batch_n_body(int64_t n, const double *jd0, const double *const *xyz,
                    int64_t center, int32_t frame, const double *jd_final,
                    int jd_final_stride, const ko_nongrav *ng /* [n] or NULL */,
                    double *const *out, int32_t *status, int64_t *impact_id,
                    double *impact_jd, ko_counters *cnt /* [n] or NULL */,
                    int n_threads);
/* ... with propagate_n_body's include_asteroids (propagation.rs:71-72). */
int ko_batch_n_body_ext(int64_t n, const double *jd0, const double *const *xyz,
                        int64_t center, int32_t frame, const double *jd_final,
                        int jd_final_stride, int include_asteroids, const ko_nongrav *ng,
                        double *const *out, int32_t *status, int64_t *impact_id,
                        double *impact_jd, ko_counters *cnt, int n_threads);

/* propagate_n_body_long (src/kete/rust/propagation.rs:202-246): the states
 * (all at jd0, Sun-centred, in `frame`) in consecutive batches of
 * batch_size, each integrated together with kete's SIMPLE_PLANETS (Sun,
 * Mercury, Venus, Earth-Moon barycentre, Mars..Neptune; gravity.rs:137-154)
 * as one Radau state (propagate_n_body_vec, propagation/mod.rs:142-237;
 * vec_accel, acceleration.rs:197-249).  planet_xyz: [9][6] Sun-centred
 * equatorial at jd0, or NULL to read them from the loaded SPK.  out: [6][n]
 * Sun-centred EQUATORIAL at jd_final (kete returns that frame); planets_out
 * ([9][6], nullable): the last batch's planets.  Returns the first failing
 * batch's status (kete fails the whole call). */
int ko_propagate_n_body_long(int64_t n, double jd0, const double *const *xyz, int32_t frame,
                             double jd_final, const double *planet_xyz, const ko_nongrav *ng,
                             int64_t batch_size, double *const *out, double *planets_out,
                             int n_threads);

/* ---- FOV ----------------------------------------------------------------- */
/* OnSkyRectangle::new (fov/patches.rs:122-155): normals[4][3]. */
void ko_rect_normals(const double pointing[3], double rotation, double lon_w,
                     double lat_w, double normals[12]);
/* SphericalPolygon::contains (fov/patches.rs:214-234): returns 1 if inside,
 * else 0 and the distance in *dist. */
int ko_rect_contains(const double normals[12], const double v[3], double *dist);
/* FovLike::check_visible (fov/fov_like.rs:134-192) for one GenericRectangle.
 * states: n objects in SoA (xyz[6]) at jd[n], center 10, frame of the FOV.
 * Writes up to n visible indices, and their final states (out[6][n]).
 * Returns the number visible. */
int64_t ko_fov_check_visible(const double normals[12], double obs_jd,
                             const double obs_pos[3], const double obs_vel[3],
                             int32_t frame, int64_t n, const double *jd,
                             const double *const *xyz, double dt_limit,
                             int64_t *vis_idx, double *const *vis_state,
                             int n_threads);
/* fov_checks_py (src/kete/rust/fovs/checks.rs:24-110).  Fields are given
 * sorted by jd.  Outputs (field index, object index, state) triples, ordered
 * by field then object index.  Returns the count, or -1 if cap is too small. */
int64_t ko_fov_state_check(int64_t n, double jd0, const double *const *xyz,
                           int32_t frame, int64_t n_fields,
                           const double *field_jd, const double *obs_xyz /*[n_fields][6]*/,
                           const double *normals /*[n_fields][12]*/,
                           double dt_limit, int64_t cap, int64_t *hit_field,
                           int64_t *hit_obj, double *hit_state /*[cap][7] jd,pos,vel*/,
                           int n_threads);
/* ... with fov_state_check's include_asteroids (checks.rs:25-31). */
int64_t ko_fov_state_check_ext(int64_t n, double jd0, const double *const *xyz,
                               int32_t frame, int64_t n_fields, const double *field_jd,
                               const double *obs_xyz, const double *normals, double dt_limit,
                               int include_asteroids, int64_t cap, int64_t *hit_field,
                               int64_t *hit_obj, double *hit_state, int n_threads);

/* fov_spk_checks_py (src/kete/rust/fovs/checks.rs:123-142) with
 * FovLike::check_spks (fov/fov_like.rs:196-226) per field; fields in epoch
 * order; hits (field, id index, state [cap][7]) by field then id order.
 * Returns the count, or -1 if cap is too small. */
int64_t ko_fov_spk_check(int64_t n_ids, const int64_t *ids, int32_t frame, int64_t n_fields,
                         const double *field_jd, const double *obs_xyz, const double *normals,
                         int64_t cap, int64_t *hit_field, int64_t *hit_obj, double *hit_state);

/* ---- MOID (propagation/kepler.rs:325-404) --------------------------------- */
/* kete's moid of two states: both to ecliptic, each moved to its perihelion,
 * a 100 x 100 grid of two-body positions (steps of period/50, or 6 days for
 * open orbits), then Nelder-Mead (argmin 0.10) from the closest grid pair.
 * grid_min: the grid's smallest distance; iters: Nelder-Mead iterations. */
/* CometElements::from_state's perihelion time and orbital_period (elements.rs:50-163,287-293). */
void ko_comet_elements_peri(double epoch, const double pos[3], const double vel[3],
                            double *peri_time, double *period);
int ko_moid(const ko_state *a, const ko_state *b, double *moid, double *grid_min,
            int32_t *iters);
/* A batch: a[i] = (jd_a[i], xyz_a[.][i]) and b[i] = (jd_b[i] or jd_a[i],
 * xyz_b[.][i]) or, with xyz_b NULL, the Earth at jd_a[i] from the loaded
 * SPK (moid_py, src/kete/rust/propagation.rs:27-39).  Sun-centred, frame. */
void ko_moid_batch(int64_t n, const double *jd_a, const double *const *xyz_a, int32_t frame,
                   const double *jd_b, const double *const *xyz_b, double *out,
                   double *grid_min, int32_t *status, int32_t *iters, int n_threads);

/* ---- post-detection epilogue --------------------------------------------- */
/* ra_dec_with_rates for one state (simult_states.rs:86-120), equatorial. */
void ko_ra_dec_with_rates(const double pos[3], const double vel[3],
                          const double obs_pos[3], const double obs_vel[3],
                          double out[4]);
/* HGParams::apparent_mag (flux/reflected.rs:255-268). */
double ko_hg_apparent_mag(double h, double g, const double sun2obj[3],
                          const double sun2obs[3]);
double ko_hg_phase_curve_correction(double g, double phase);

/* The orbit-fit observable chain (N-body arc to each epoch, light-time step,
 * RA/Dec) for n objects at m epochs; jd_obs [m][n], obs [3*m][n] (observer,
 * Sun-centred, frame), radec [2*m][n], state [6*m][n] (nullable).  mode 0:
 * kete's adaptive steps; 1: record each object's step attempts into
 * sched_* [n][cap] (sched_n [n] receives the counts); 2: replay them. */
int ko_chain_observe_batch(int64_t n, const double *jd0, const double *const *xyz,
                           int32_t frame, const ko_nongrav *ng, int32_t m,
                           const double *jd_obs, const double *obs, int restart, int mode,
                           int64_t cap, double *sched_h, int32_t *sched_sweeps,
                           int32_t *sched_acc, int64_t *sched_n, double *radec, double *state,
                           int32_t *status, int n_threads);

/* Test hook: record the FOV re-propagation schedule ((kind, epoch) pairs). */
void ko_fov_trace_set(double *buf, int64_t cap);
int64_t ko_fov_trace_len(void);

/* Radau tolerance used by every oracle integration (1e-6, kete's EPSILON,
 * by default); returns the previous value; eps <= 0 only reads it. */
double ko_set_radau_epsilon(double eps);

/* RK45Integrator::integrate (propagation/runge_kutta.rs:119-147) on kete's
 * two test equations (kind 0: y' = -y, kind 1: y' = 2). */
int ko_rk45_test(int kind, double y0, double t0, double t1, double tol, double *y1,
                 int64_t *n_eval);
/* compute_state_transition (propagation/state_transition.rs:52-112); fix
 * bits: 1 the true accel_grad, 2 the acceleration in scaled time with the
 * central mass (0 = kete's arithmetic).  stm: row-major, the matrix kete
 * returns (the transpose of d x(jd1) / d x(jd0)). */
int ko_compute_state_transition(const double pos[3], const double vel[3], double jd0,
                                double jd1, double central_mass, int fix, double out_pos[3],
                                double out_vel[3], double stm[36], int64_t *n_eval);

#ifdef __cplusplus
}
#endif
#endif
