/*
 * kete_oracle.c -- CPU restatement of kete 1.1.3's orbit-propagation hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  The
 * product (kete_amd/, libkete_hip.so) never links or calls this file.
 *
 * Built with -O3 -ffp-contract=off so that every multiply and add rounds the
 * way kete's Rust does (Rust never fuses a*b+c on its own).  libm calls are
 * glibc's, as kete's are on Linux.  Each function cites the Rust it restates
 * (paths relative to the kete source tree, src/kete_core/src unless noted).
 *
 * Parity pins (checks that a plausible mistake here would fail) live in
 * tests/test_oracle_*.py: kete's own known-answer tests, closed forms and
 * invariants.  Functions with no such pin are listed in the oracle section of
 * the design notes as "parity unpinned".
 */
#define _GNU_SOURCE
#include "kete_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Constants: constants/gravity.rs:9-40,43-148; constants/universal.rs:11-23 */
/* ------------------------------------------------------------------------ */
#define GMS 0.00029591220828411956
#define GMS_SQRT 0.01720209894996
#define SUN_J2 2.2e-7
#define EARTH_J2 0.00108262998905
#define JUPITER_J2 0.014696572
#define C_AU_PER_DAY 173.14463267424034
#define C_AU_PER_DAY_INV 0.005775518331436995
#define C_AU_PER_DAY_INV_SQUARED 3.33566119967647e-05
#define AU_KM 149597870.7
#define OBLIQUITY 0.40909280422232897 /* frames/definitions.rs:135 */
#define TAU 6.283185307179586
#define PI_ 3.141592653589793
#define PARABOLIC_BETA 1e-10          /* propagation/kepler.rs:120 */
#define PARABOLIC_ECC_LIMIT 1e-3      /* propagation/kepler.rs:17 */

typedef struct {
    int64_t naif_id;
    double mass;
    double radius;
} grav_params;

/* MASSES (constants/gravity.rs:43-134): the Sun, the planets and the Moon
 * (PLANETS = MASSES[0..10], gravity.rs:145-148), then Ceres, Pallas, Vesta,
 * Hygiea and Interamnia. */
#define N_PLANETS 10
#define N_MASSES 15
static const grav_params MASSES[N_MASSES] = {
    {10, GMS, 0.004654758765894654},
    {1, 1.66012082548908e-07 * GMS, 1.63139354095098e-05},
    {2, 2.44783828779694e-06 * GMS, 4.04537843465442e-05},
    {399, 3.00348961546514e-06 * GMS, 4.33036684926559e-05},
    {301, 3.69430335010988e-08 * GMS, 1.16138016662292e-05},
    {4, 3.22715608291416e-07 * GMS, 2.27021279387769e-05},
    {5, 0.0009547919099414246 * GMS, 0.00047789450254521576},
    {6, 0.00028588567002459455 * GMS, 0.0004028666966848747},
    {7, 4.36624961322212e-05 * GMS, 0.0001708513622580592},
    {8, 5.15138377265457e-05 * GMS, 0.0001655371154958558},
    {20000001, 4.7191659919436e-10 * GMS, 6.276827e-6},
    {20000002, 1.0259143964958e-10 * GMS, 3.415824e-6},
    {20000004, 1.3026452498655e-10 * GMS, 3.512082e-6},
    {20000010, 4.3953391300849e-11 * GMS, 2.894426e-6},
    {20000704, 1.911e-11 * GMS, 2.219283e-6},
};
#define PLANETS MASSES

double ko_const(const char *name) {
    if (!strcmp(name, "GMS")) return GMS;
    if (!strcmp(name, "GMS_SQRT")) return GMS_SQRT;
    if (!strcmp(name, "C_AU_PER_DAY")) return C_AU_PER_DAY;
    if (!strcmp(name, "C_AU_PER_DAY_INV")) return C_AU_PER_DAY_INV;
    if (!strcmp(name, "AU_KM")) return AU_KM;
    if (!strcmp(name, "OBLIQUITY")) return OBLIQUITY;
    return NAN;
}

/* ------------------------------------------------------------------------ */
/* Small vector helpers (nalgebra semantics: norm = sqrt(x*x+y*y+z*z)).      */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }
static inline void cross3(const double *a, const double *b, double *o) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}
/* Rust f64::rem_euclid. */
static inline double rem_euclid(double x, double m) {
    double r = fmod(x, m);
    return r < 0.0 ? r + fabs(m) : r;
}
/* powi(x, -3) as LLVM expands it: 1 / (x*x*x). */
static inline double powi_m3(double x) { return 1.0 / (x * x * x); }

/* Ecliptic <-> equatorial rotation about x by the obliquity
 * (frames/definitions.rs:155-158,224-237): R = [[1,0,0],[0,c,-s],[0,s,c]]. */
static double ROT_C, ROT_S;
static pthread_once_t rot_once = PTHREAD_ONCE_INIT;
static void rot_init(void) {
    ROT_S = sin(OBLIQUITY);
    ROT_C = cos(OBLIQUITY);
}
static inline void ecl_to_equ(const double *v, double *o) {
    pthread_once(&rot_once, rot_init);
    double x = v[0], y = v[1], z = v[2];
    o[0] = x;
    o[1] = ROT_C * y + (-ROT_S) * z;
    o[2] = ROT_S * y + ROT_C * z;
}
static inline void equ_to_ecl(const double *v, double *o) {
    pthread_once(&rot_once, rot_init);
    double x = v[0], y = v[1], z = v[2];
    o[0] = x;
    o[1] = ROT_C * y + ROT_S * z;
    o[2] = (-ROT_S) * y + ROT_C * z;
}

/* ------------------------------------------------------------------------ */
/* fitting/newton.rs:18-65                                                   */
/* ------------------------------------------------------------------------ */
typedef double (*scalar_fn)(double, const void *);

static int newton_raphson(scalar_fn func, scalar_fn der, const void *ctx,
                          double start, double atol, double *out) {
    double x = start;
    if (fabs(der(x, ctx)) < 1e-12) x += 0.1;
    for (int it = 0; it < 100; ++it) {
        double f = func(x, ctx);
        if (fabs(f) < atol) { *out = x; return KO_OK; }
        double d = der(x, ctx);
        if (fabs(d) < 1e-12) return KO_CONVERGENCE;
        if (!isfinite(d) || !isfinite(f)) return KO_CONVERGENCE;
        x -= 0.5 * f / d;
        d = der(x, ctx);
        if (fabs(d) < 1e-3) f = func(x, ctx);
        x -= 0.5 * f / d;
    }
    return KO_CONVERGENCE;
}

static double sq_f(double x, const void *c) { (void)c; return 1.0 * x * x - 1.0; }
static double sq_d(double x, const void *c) { (void)c; return 2.0 * x; }
/* fitting/newton.rs:72-78 known-answer hook. */
int ko_newton_square_root_test(double start, double atol, double *root) {
    return newton_raphson(sq_f, sq_d, NULL, start, atol, root);
}

/* ------------------------------------------------------------------------ */
/* propagation/kepler.rs:120-297 -- universal-variable Kepler solver         */
/* ------------------------------------------------------------------------ */
typedef struct {
    double r0, rv0, beta, b_sqrt, dt;
} kep_ctx;

static double kep_ell_f(double x, const void *c) {
    const kep_ctx *k = c;
    double g1 = sin(k->b_sqrt * x) / k->b_sqrt;
    double sh = sin(k->b_sqrt * x / 2.0);
    double g2 = 2.0 * pow(sh, 2.0) / k->beta;
    double g3 = (x - g1) / k->beta;
    return k->r0 * g1 + k->rv0 * g2 + GMS * g3 - k->dt;
}
static double kep_ell_d(double x, const void *c) {
    const kep_ctx *k = c;
    double g2d = sin(k->b_sqrt * x), g1d = cos(k->b_sqrt * x);
    double g3d = (1.0 - g1d) / k->beta;
    return k->r0 * g1d + k->rv0 * g2d + GMS * g3d;
}
static double kep_hyp_f(double x, const void *c) {
    const kep_ctx *k = c;
    double g1 = sinh(k->b_sqrt * x) / k->b_sqrt;
    double g2 = (1.0 - cosh(k->b_sqrt * x)) / k->beta;
    double g3 = (x - g1) / k->beta;
    return k->r0 * g1 + k->rv0 * g2 + GMS * g3 - k->dt;
}
static double kep_hyp_d(double x, const void *c) {
    const kep_ctx *k = c;
    double g2d = sinh(k->b_sqrt * x), g1d = cosh(k->b_sqrt * x);
    double g3d = (1.0 - g1d) / k->beta;
    return k->r0 * g1d + k->rv0 * g2d + GMS * g3d;
}

/* kepler.rs:129-140 */
static double g_1(double s, double beta) {
    if (fabs(beta) < PARABOLIC_BETA) return s;
    double bs = sqrt(fabs(beta));
    if (beta >= 0.0) return sin(bs * s) / bs;
    return sinh(bs * s) / bs;
}
/* kepler.rs:148-161 */
static double g_2(double s, double beta) {
    if (fabs(beta) < PARABOLIC_BETA) return s * s / 2.0;
    double bs = sqrt(fabs(beta)) * s;
    if (beta >= 0.0) return 2.0 * pow(sin(bs / 2.0), 2.0) / beta;
    return (1.0 - cosh(bs)) / beta;
}

/* kepler.rs:179-233 */
static int solve_kepler_universal(double dt, double r0, double v0, double rv0,
                                  double *s_out, double *beta_out) {
    double beta = 2.0 * GMS / r0 - v0 * v0;
    double b_sqrt = sqrt(fabs(beta));
    int st;
    double res = 0.0;
    if (beta >= PARABOLIC_BETA) {
        double period = GMS * powi_m3(b_sqrt);
        dt = fmod(dt, period * TAU);
        kep_ctx k = {r0, rv0, beta, b_sqrt, dt};
        double guess = (1.0 / b_sqrt) * dt / period;
        st = newton_raphson(kep_ell_f, kep_ell_d, &k, guess, 1e-11, &res);
    } else if (fabs(beta) < PARABOLIC_BETA) {
        double a0 = -6.0 * dt / GMS;
        double a1 = 6.0 * r0 / GMS;
        double a2 = 3.0 * rv0 / GMS;
        double p = (3.0 * a1 - a2 * a2) / 3.0;
        double q = (9.0 * a1 * a2 - 27.0 * a0 - 2.0 * (a2 * a2 * a2)) / 27.0;
        double w = pow(q / 2.0 + sqrt(q * q / 4.0 + (p * p * p) / 27.0), 1.0 / 3.0);
        res = w - p / (3.0 * w) - a2 / 3.0;
        st = KO_OK;
    } else {
        kep_ctx k = {r0, rv0, beta, b_sqrt, dt};
        st = newton_raphson(kep_hyp_f, kep_hyp_d, &k, GMS_SQRT * fabs(beta) * dt,
                            1e-11, &res);
    }
    if (st != KO_OK) return KO_CONVERGENCE;
    *s_out = res;
    *beta_out = beta;
    return KO_OK;
}

/* kepler.rs:254-297 */
static int analytic_2_body_depth(double time, const double *pos, const double *vel,
                                 int depth, double *po, double *vo) {
    if (depth >= 10) return KO_CONVERGENCE;
    depth += 1;
    if (fabs(time) < 1e-10) {
        memcpy(po, pos, 3 * sizeof(double));
        memcpy(vo, vel, 3 * sizeof(double));
        return KO_OK;
    }
    double r0 = norm3(pos);
    double v0 = norm3(vel);
    double rv0 = dot3(pos, vel);
    if (isnan(rv0) || isinf(rv0)) return KO_CONVERGENCE;
    double s, beta;
    if (solve_kepler_universal(time, r0, v0, rv0, &s, &beta) == KO_OK) {
        double g1 = g_1(s, beta);
        double g2 = g_2(s, beta);
        double f = 1.0 - GMS / r0 * g2;
        double g = r0 * g1 + rv0 * g2;
        double np[3];
        for (int i = 0; i < 3; ++i) np[i] = pos[i] * f + vel[i] * g;
        double nr0 = norm3(np);
        double f_dot = -GMS / (nr0 * r0) * g1;
        double g_dot = 1.0 - (GMS / nr0) * g2;
        for (int i = 0; i < 3; ++i) vo[i] = pos[i] * f_dot + vel[i] * g_dot;
        memcpy(po, np, 3 * sizeof(double));
        return KO_OK;
    }
    double ip[3], iv[3];
    int st = analytic_2_body_depth(0.5 * time, pos, vel, depth, ip, iv);
    if (st) return st;
    return analytic_2_body_depth(time * 0.5, ip, iv, depth, po, vo);
}

int ko_analytic_2_body(double dt, const double pos[3], const double vel[3],
                       double pos_out[3], double vel_out[3]) {
    return analytic_2_body_depth(dt, pos, vel, 0, pos_out, vel_out);
}

/* kepler.rs:27-60 (used for the element->state conversion). */
typedef struct { double ecc, m; } ecc_ctx;
static double ea_ell_f(double e, const void *c) {
    const ecc_ctx *k = c;
    return -k->ecc * sin(e) + e - rem_euclid(k->m, TAU);
}
static double ea_ell_d(double e, const void *c) {
    const ecc_ctx *k = c;
    return 1.0 - 1.0 * k->ecc * cos(e);
}
static double ea_hyp_f(double e, const void *c) {
    const ecc_ctx *k = c;
    return k->ecc * sinh(e) - e - k->m;
}
static double ea_hyp_d(double e, const void *c) {
    const ecc_ctx *k = c;
    return -1.0 + k->ecc * cosh(e);
}
int ko_compute_eccentric_anomaly(double ecc, double mean_anom, double peri_dist,
                                 double *out) {
    if (!isfinite(ecc)) return KO_VALUE_ERROR;
    if (ecc < 0.0) return KO_VALUE_ERROR;
    if (ecc < 1e-6) { *out = mean_anom; return KO_OK; }
    if (ecc < 1.0 - PARABOLIC_ECC_LIMIT) {
        ecc_ctx k = {ecc, mean_anom};
        double r;
        int st = newton_raphson(ea_ell_f, ea_ell_d, &k, rem_euclid(mean_anom, TAU),
                                1e-11, &r);
        if (st) return st;
        *out = rem_euclid(r, TAU);
        return KO_OK;
    }
    if (ecc < 1.0 + PARABOLIC_ECC_LIMIT) {
        double p = 2.0 * peri_dist;
        double q = 6.0 * mean_anom;
        double w = cbrt(0.5 * (q + sqrt(q * q + 4.0 * (p * p * p))));
        *out = w - p / w;
        return KO_OK;
    }
    ecc_ctx k = {ecc, mean_anom};
    return newton_raphson(ea_hyp_f, ea_hyp_d, &k, asinh(mean_anom / ecc), 1e-11, out);
}

/* elements.rs:190-264 (CometElements::to_pos_vel). */
int ko_elements_to_state(double epoch, double ecc, double incl, double lon_asc,
                         double peri_arg, double peri_dist, double peri_time,
                         double pos[3], double vel[3]) {
    int elliptical = ecc < 1.0 - PARABOLIC_ECC_LIMIT;
    int hyperbolic = ecc > 1.0 + PARABOLIC_ECC_LIMIT;
    int parabolic = !elliptical && !hyperbolic;
    double semi_major = parabolic ? 0.0 : peri_dist / (1.0 - ecc);
    double mean_motion = parabolic ? GMS_SQRT : pow(fabs(semi_major), -1.5) * GMS_SQRT;
    double mean_anom = mean_motion * (epoch - peri_time);
    double ea;
    int st = ko_compute_eccentric_anomaly(ecc, mean_anom, peri_dist, &ea);
    if (st) return st;
    double x, y, xd, yd;
    if (elliptical) {
        double se = sin(ea), ce = cos(ea);
        double e_dot = pow(semi_major, 1.5) * (1.0 - ecc * ce);
        double b = semi_major * sqrt(1.0 - ecc * ecc);
        x = semi_major * (ce - ecc);
        y = b * se;
        xd = -semi_major / e_dot * se * GMS_SQRT;
        yd = b / e_dot * ce * GMS_SQRT;
    } else if (hyperbolic) {
        double sh = sinh(ea), ch = cosh(ea);
        double b = -semi_major * sqrt(ecc * ecc - 1.0);
        double h_dot = pow(fabs(semi_major), 1.5) * (1.0 - ecc * ch);
        x = semi_major * (ch - ecc);
        y = b * sh;
        xd = -semi_major / h_dot * sh * GMS_SQRT;
        yd = -b / h_dot * ch * GMS_SQRT;
    } else {
        double d_dot = peri_dist + ea * ea / 2.0;
        x = peri_dist - ea * ea / 2.0;
        y = sqrt(2.0 * peri_dist) * ea;
        xd = -ea / d_dot * GMS_SQRT;
        yd = sqrt(2.0 * peri_dist) / d_dot * GMS_SQRT;
    }
    double s_w = sin(peri_arg), c_w = cos(peri_arg);
    double s_o = sin(lon_asc), c_o = cos(lon_asc);
    double s_i = sin(incl), c_i = cos(incl);
    double px = c_w * c_o - s_w * s_o * c_i;
    double py = c_w * s_o + s_w * c_o * c_i;
    double pz = s_w * s_i;
    double qx = -s_w * c_o - c_w * s_o * c_i;
    double qy = -s_w * s_o + c_w * c_o * c_i;
    double qz = c_w * s_i;
    pos[0] = x * px + y * qx; pos[1] = x * py + y * qy; pos[2] = x * pz + y * qz;
    vel[0] = xd * px + yd * qx; vel[1] = xd * py + yd * qy; vel[2] = xd * pz + yd * qz;
    return KO_OK;
}

/* ------------------------------------------------------------------------ */
/* spice: DAF reader (spice/daf.rs:148-351), SPK collection (spice/spk.rs),  */
/* Type 2 segments (spice/spk_segments.rs:105-177,316-397), Type 21         */
/* segments (spice/spk_segments.rs:958-1102).                                */
/* ------------------------------------------------------------------------ */
#define MAX_SEGS 256
typedef struct {
    int64_t obj_id, center_id;
    int32_t frame; /* 1 equatorial, 17 ecliptic, else unknown */
    int32_t seg_type;
    double jd_start, jd_end;
    double *data;
    int64_t n_data;
    /* type 2: n_coef Chebyshev coefficients per axis; type 21: MAXDIM */
    double jd_step;
    int64_t n_coef, record_len, n_rec;
} spk_seg;

static spk_seg SEGS[MAX_SEGS];
static int N_SEGS = 0;
/* Planet segments (ids 0..=1000) are scanned first, in load order
 * (spk.rs:65-82,318-325); PLANET_ORDER lists them, OTHER_ORDER the rest. */
static int PLANET_ORDER[MAX_SEGS], N_PLANET = 0;
static int OTHER_ORDER[MAX_SEGS], N_OTHER = 0;

void ko_spk_reset(void) {
    for (int i = 0; i < N_SEGS; ++i) free(SEGS[i].data);
    N_SEGS = N_PLANET = N_OTHER = 0;
}

static double rd_f64(const unsigned char *b, int le) {
    unsigned char t[8];
    if (le) memcpy(t, b, 8);
    else for (int i = 0; i < 8; ++i) t[i] = b[7 - i];
    double v;
    memcpy(&v, t, 8);
    return v;
}
static int32_t rd_i32(const unsigned char *b, int le) {
    unsigned char t[4];
    if (le) memcpy(t, b, 4);
    else for (int i = 0; i < 4; ++i) t[i] = b[3 - i];
    int32_t v;
    memcpy(&v, t, 4);
    return v;
}

/* spice/mod.rs:25-34 */
static inline double spice_jds_to_jd(double s) { return s / 86400.0 + 2451545.0; }
static inline double jd_to_spice_jd(double jd) { return (jd - 2451545.0) * 86400.0; }

int ko_spk_load_file(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return KO_IO_ERROR;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char *buf = malloc(sz);
    if (!buf || fread(buf, 1, sz, f) != (size_t)sz) { fclose(f); free(buf); return KO_IO_ERROR; }
    fclose(f);
    int rc = KO_OK;
    if (sz < 1024 || memcmp(buf, "DAF/SPK", 7) != 0) { rc = KO_IO_ERROR; goto done; }
    int le;
    if (!memcmp(buf + 88, "LTL-IEEE", 8)) le = 1;
    else if (!memcmp(buf + 88, "BIG-IEEE", 8)) le = 0;
    else { rc = KO_IO_ERROR; goto done; }
    int32_t nd = rd_i32(buf + 8, le), ni = rd_i32(buf + 12, le);
    int32_t fward = rd_i32(buf + 76, le);
    if (nd != 2 || ni != 6) { rc = KO_IO_ERROR; goto done; }
    int summary_size = nd + (ni + 1) / 2;
    int next = fward;
    while (next != 0) {
        long off = 1024L * (next - 1);
        if (off + 1024 > sz) { rc = KO_IO_ERROR; goto done; }
        const unsigned char *rec = buf + off;
        next = (int)rd_f64(rec, le);
        int nsum = (int)rd_f64(rec + 16, le);
        for (int s = 0; s < nsum; ++s) {
            const unsigned char *sb = rec + 24 + s * summary_size * 8;
            double fl[2] = {rd_f64(sb, le), rd_f64(sb + 8, le)};
            int32_t in[6];
            for (int k = 0; k < 6; ++k) in[k] = rd_i32(sb + 16 + 4 * k, le);
            if (N_SEGS >= MAX_SEGS) { rc = KO_IO_ERROR; goto done; }
            spk_seg *g = &SEGS[N_SEGS];
            memset(g, 0, sizeof(*g));
            g->jd_start = spice_jds_to_jd(fl[0]);
            g->jd_end = spice_jds_to_jd(fl[1]);
            g->obj_id = in[0];
            g->center_id = in[1];
            g->frame = in[2];
            g->seg_type = in[3];
            int64_t a0 = in[4], a1 = in[5];
            g->n_data = a1 - a0 + 1;
            if (a0 < 1 || 8 * a1 > sz || g->n_data < 4) { rc = KO_IO_ERROR; goto done; }
            g->data = malloc(8 * g->n_data);
            for (int64_t k = 0; k < g->n_data; ++k)
                g->data[k] = rd_f64(buf + 8 * (a0 - 1 + k), le);
            if (g->seg_type == 2) {
                /* From<DafArray> for SpkSegmentType2 (spk_segments.rs:379-397) */
                g->record_len = (int64_t)g->data[g->n_data - 2];
                g->jd_step = g->data[g->n_data - 3];
                g->n_coef = (g->record_len - 2) / 3;
                g->n_rec = (int64_t)g->data[g->n_data - 1];
                if (3 * g->n_coef + 2 != g->record_len) { rc = KO_IO_ERROR; goto done; }
            } else if (g->seg_type == 21) {
                /* From<DafArray> for SpkSegmentType21 (spk_segments.rs:971-986):
                 * trailer [MAXDIM, N]; records of 4*MAXDIM+11 doubles, then the
                 * N final epochs.  kete does not check the sizes; the oracle
                 * refuses arrays too short for them. */
                g->n_rec = (int64_t)g->data[g->n_data - 1];
                g->n_coef = (int64_t)g->data[g->n_data - 2];
                g->record_len = 4 * g->n_coef + 11;
                if (g->n_rec < 1 || g->n_coef < 1 ||
                    g->n_rec * (g->record_len + 1) + 2 > g->n_data) { rc = KO_IO_ERROR; goto done; }
            }
            if (g->obj_id >= 0 && g->obj_id <= 1000) PLANET_ORDER[N_PLANET++] = N_SEGS;
            else OTHER_ORDER[N_OTHER++] = N_SEGS;
            N_SEGS++;
        }
    }
done:
    free(buf);
    return rc;
}

int ko_spk_n_segments(void) { return N_SEGS; }
int ko_spk_segment_info(int idx, int64_t *obj_id, int64_t *center_id,
                        int32_t *frame, int32_t *seg_type, double *jd_start,
                        double *jd_end) {
    if (idx < 0 || idx >= N_SEGS) return KO_VALUE_ERROR;
    *obj_id = SEGS[idx].obj_id; *center_id = SEGS[idx].center_id;
    *frame = SEGS[idx].frame; *seg_type = SEGS[idx].seg_type;
    *jd_start = SEGS[idx].jd_start; *jd_end = SEGS[idx].jd_end;
    return KO_OK;
}

/* interpolation.rs:25-77: forward T/U recurrences, ascending order. */
int ko_chebyshev_evaluate_both(double x, const double *cx, const double *cy,
                               const double *cz, int n, double val[3],
                               double der[3]) {
    if (n < 2) return KO_IO_ERROR;
    double x2 = 2.0 * x;
    val[0] = cx[0] + cx[1] * x;
    val[1] = cy[0] + cy[1] * x;
    val[2] = cz[0] + cz[1] * x;
    double second_t = 1.0, last_t = x, next_t;
    double second_u = 1.0, last_u = x2, next_u;
    der[0] = cx[1]; der[1] = cy[1]; der[2] = cz[1];
    for (int idx = 2; idx < n; ++idx) {
        next_t = x2 * last_t - second_t;
        val[0] += cx[idx] * next_t;
        val[1] += cy[idx] * next_t;
        val[2] += cz[idx] * next_t;
        second_t = last_t;
        last_t = next_t;
        next_u = x2 * last_u - second_u;
        der[0] += cx[idx] * last_u * (double)idx;
        der[1] += cy[idx] * last_u * (double)idx;
        der[2] += cz[idx] * last_u * (double)idx;
        second_u = last_u;
        last_u = next_u;
    }
    return KO_OK;
}

/* SpkSegment::contains (spk_segments.rs:143-145). */
static inline int seg_contains(const spk_seg *g, double jd) {
    return jd >= g->jd_start && jd <= g->jd_end;
}

/* SpkSegmentType21::try_get_pos_vel (spk_segments.rs:1007-1100) on one
 * record: the modified-difference-array (MDA) evaluation of NAIF's SPKE21.
 * rec: [t_ref, G[n], x vx y vy z vz, DT[3][n], KQMAX1, KQ[3]] in km, s.
 * Position in AU, velocity in AU/day.  Returns KO_IO_ERROR for a record kete
 * would reject (a zero step size) or could not index (orders outside
 * 1..n, which kete does not check). */
#define MDA_MAX 64
int ko_spk21_record_eval(const double *rec, int64_t n, double js, double pos[3],
                         double vel[3]) {
    const double t_ref = rec[0];
    const double *g = rec + 1;
    const double *ref = rec + n + 1;
    const double *dd = rec + n + 7;
    const int64_t kqmax1 = (int64_t)rec[4 * n + 7];
    int64_t kq[3];
    for (int i = 0; i < 3; ++i) kq[i] = (int64_t)rec[4 * n + 8 + i];
    if (n > MDA_MAX || kqmax1 < 2 || kqmax1 > n + 1) return KO_IO_ERROR;
    for (int i = 0; i < 3; ++i)
        if (kq[i] < 0 || kq[i] > kqmax1 - 1) return KO_IO_ERROR;
    const double dt = js - t_ref;
    double fc[MDA_MAX], wc[MDA_MAX], w[MDA_MAX + 1];
    double tp = dt;
    for (int64_t j = 0; j < kqmax1 - 2; ++j) {
        if (g[j] == 0.0) return KO_IO_ERROR;
        fc[j] = tp / g[j];
        wc[j] = dt / g[j];
        tp = dt + g[j];
    }
    for (int64_t j = 0; j < kqmax1; ++j) w[j] = 1.0 / ((double)j + 1.0);
    /* integration passes KS = KQMAX1-1 .. 2 */
    int64_t ks = kqmax1 - 1, jx = 0, ks1 = ks - 1;
    while (ks >= 2) {
        jx += 1;
        for (int64_t j = 0; j < jx; ++j) w[j + ks] = fc[j] * w[j + ks1] - wc[j] * w[j + ks];
        ks = ks1;
        ks1 -= 1;
    }
    for (int i = 0; i < 3; ++i) {
        double sum = 0.0;
        for (int64_t j = kq[i]; j >= 1; --j) sum += dd[i * n + j - 1] * w[j + ks - 1];
        pos[i] = (ref[2 * i] + dt * (sum * dt + ref[2 * i + 1])) / AU_KM;
    }
    /* one more pass (KS = 1) for the velocity */
    for (int64_t j = 0; j < jx; ++j) w[j + ks] = fc[j] * w[j + ks1] - wc[j] * w[j + ks];
    ks -= 1;
    for (int i = 0; i < 3; ++i) {
        double sum = 0.0;
        for (int64_t j = kq[i]; j >= 1; --j) sum += dd[i * n + j - 1] * w[j + ks - 1];
        vel[i] = (ref[2 * i + 1] + dt * sum) / AU_KM * 86400.0;
    }
    return KO_OK;
}

/* Record selection of SpkSegmentType21::try_get_pos_vel
 * (spk_segments.rs:1023-1030): the first record whose final epoch is >= the
 * time (slice::binary_search_by over sorted epochs).  A time after the last
 * final epoch -- reachable at jd == jd_end through the JD <-> seconds round
 * trip -- is clamped to the last record, where kete reads past the array. */
static int64_t t21_record_index(const spk_seg *g, double js) {
    const double *ep = g->data + g->n_rec * g->record_len;
    int64_t lo = 0, hi = g->n_rec;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (ep[mid] < js) lo = mid + 1;
        else hi = mid;
    }
    return lo < g->n_rec ? lo : g->n_rec - 1;
}

/* SpkSegment::try_get_state + SpkSegmentType2::try_get_pos_vel
 * (spk_segments.rs:150-177,355-376).  The record index is clamped to the
 * last record: kete reads one record past the end at jd == jd_end. */
static int seg_get_state(const spk_seg *g, double jd, double *pos, double *vel) {
    if (jd < g->jd_start || jd > g->jd_end) return KO_DAF_LIMITS;
    if (g->seg_type == 21) {
        double js = jd_to_spice_jd(jd);
        int64_t ri = t21_record_index(g, js);
        return ko_spk21_record_eval(g->data + ri * g->record_len, g->n_coef, js, pos, vel);
    }
    if (g->seg_type != 2) return KO_IO_ERROR;
    double js = jd_to_spice_jd(jd);
    double j0 = jd_to_spice_jd(g->jd_start);
    double fi = floor((js - j0) / g->jd_step);
    int64_t ri = fi > 0.0 ? (int64_t)fi : 0;
    if (ri >= g->n_rec) ri = g->n_rec - 1;
    const double *r = g->data + ri * g->record_len;
    double t_mid = r[0], t_step = r[1];
    double t = (js - t_mid) / t_step;
    double t_step_scaled = 86400.0 / t_step / AU_KM;
    double p[3], v[3];
    int n = (int)g->n_coef;
    int st = ko_chebyshev_evaluate_both(t, r + 2, r + 2 + n, r + 2 + 2 * n, n, p, v);
    if (st) return st;
    for (int i = 0; i < 3; ++i) {
        pos[i] = p[i] / AU_KM;
        vel[i] = v[i] * t_step_scaled;
    }
    return KO_OK;
}

typedef struct {
    int64_t id;      /* Desig::Naif id */
    ko_state s;
} naif_state;

/* SpkCollection::try_get_raw_state (spk.rs:65-82). */
static int raw_state(int64_t id, double jd, naif_state *out) {
    for (int k = 0; k < N_PLANET; ++k) {
        const spk_seg *g = &SEGS[PLANET_ORDER[k]];
        if (g->obj_id == id && seg_contains(g, jd)) goto found_g;
        continue;
    found_g:
        out->id = id;
        out->s.jd = jd;
        out->s.frame = g->frame;
        out->s.center = g->center_id;
        return seg_get_state(g, jd, out->s.pos, out->s.vel);
    }
    for (int k = 0; k < N_OTHER; ++k) {
        const spk_seg *g = &SEGS[OTHER_ORDER[k]];
        if (g->obj_id == id && seg_contains(g, jd)) {
            out->id = id;
            out->s.jd = jd;
            out->s.frame = g->frame;
            out->s.center = g->center_id;
            return seg_get_state(g, jd, out->s.pos, out->s.vel);
        }
    }
    return KO_DAF_LIMITS;
}

/* State::try_change_frame_mut (state.rs:146-210) for the two inertial frames
 * this path uses. */
int ko_change_frame(ko_state *s, int32_t target) {
    if (s->frame == target) return KO_OK;
    double p[3], v[3];
    if (s->frame == KO_FRAME_EQUATORIAL) { equ_to_ecl(s->pos, p); equ_to_ecl(s->vel, v); }
    else if (s->frame == KO_FRAME_ECLIPTIC) { memcpy(p, s->pos, 24); memcpy(v, s->vel, 24); }
    else return KO_UNKNOWN_FRAME;
    if (target == KO_FRAME_EQUATORIAL) { ecl_to_equ(p, s->pos); ecl_to_equ(v, s->vel); }
    else if (target == KO_FRAME_ECLIPTIC) { memcpy(s->pos, p, 24); memcpy(s->vel, v, 24); }
    else return KO_UNKNOWN_FRAME;
    s->frame = target;
    return KO_OK;
}

/* State::try_change_center (state.rs:242-281). */
static int state_apply_center(ko_state *self, naif_state other) {
    if (self->jd != other.s.jd) return KO_VALUE_ERROR;
    int64_t other_id = other.id;
    if (self->center != other.s.center && self->center != other_id) return KO_VALUE_ERROR;
    if (self->center == other.s.center) {
        /* try_flip_center_id (state.rs:215-229) */
        int64_t t = other.id;
        other.id = other.s.center;
        other.s.center = t;
        for (int i = 0; i < 3; ++i) { other.s.pos[i] = -other.s.pos[i]; other.s.vel[i] = -other.s.vel[i]; }
    }
    if (self->frame != other.s.frame) {
        int st = ko_change_frame(&other.s, self->frame);
        if (st) return st;
    }
    self->center = other.s.center;
    for (int i = 0; i < 3; ++i) { self->pos[i] += other.s.pos[i]; self->vel[i] += other.s.vel[i]; }
    return KO_OK;
}

/* find_path (spk.rs:210-234): breadth-first search over the segment graph
 * (unit edge weights), returning the via-segment ids after the start. */
static int find_path(int64_t start, int64_t goal, int64_t *path, int *n_path) {
    enum { MAXN = 512 };
    int64_t node[MAXN], via[MAXN];
    int parent[MAXN];
    int head = 0, tail = 0;
    node[tail] = start; via[tail] = INT64_MIN; parent[tail] = -1; tail++;
    while (head < tail) {
        int cur = head++;
        if (node[cur] == goal) {
            int64_t rev[64];
            int nr = 0;
            for (int k = cur; parent[k] >= 0 && nr < 64; k = parent[k]) rev[nr++] = via[k];
            for (int k = 0; k < nr; ++k) path[k] = rev[nr - 1 - k];
            *n_path = nr;
            return KO_OK;
        }
        for (int sidx = 0; sidx < N_SEGS; ++sidx) {
            const spk_seg *g = &SEGS[sidx];
            int64_t nb = INT64_MIN, vv = 0;
            if (g->obj_id == node[cur]) { nb = g->center_id; vv = g->obj_id; }
            else if (g->center_id == node[cur]) { nb = g->obj_id; vv = g->obj_id; }
            else continue;
            int seen = 0;
            for (int k = 0; k < tail; ++k) if (node[k] == nb) { seen = 1; break; }
            if (seen || tail >= MAXN) continue;
            node[tail] = nb; via[tail] = vv; parent[tail] = cur; tail++;
        }
    }
    return KO_DAF_LIMITS;
}

/* SpkCollection::try_change_center (spk.rs:99-126). */
int ko_change_center(ko_state *s, int64_t nc) {
    int64_t a = s->center;
    naif_state r;
    int st;
#define APPLY(id) do { st = raw_state((id), s->jd, &r); if (st) return st; \
                       st = state_apply_center(s, r); if (st) return st; } while (0)
    if (a == nc) return KO_OK;
    if (nc == 0 && a <= 10) { APPLY(a); return KO_OK; }
    if (a == 0 && nc == 10) { APPLY(10); return KO_OK; }
    if (nc == 10 && a < 10) { APPLY(a); APPLY(10); return KO_OK; }
    if (a == 10 && nc > 1 && nc < 10) { APPLY(10); APPLY(nc); return KO_OK; }
    int64_t path[64];
    int np = 0;
    st = find_path(a, nc, path, &np);
    if (st) return st;
    for (int k = 0; k < np; ++k) APPLY(path[k]);
#undef APPLY
    return KO_OK;
}

/* SpkCollection::try_get_state (spk.rs:87-96). */
static int spk_get_state(int64_t id, double jd, int64_t center, int32_t frame,
                         ko_state *out) {
    naif_state r;
    int st = raw_state(id, jd, &r);
    if (st) return st;
    *out = r.s;
    if (out->center != center) {
        st = ko_change_center(out, center);
        if (st) return st;
    }
    if (out->frame != frame) {
        st = ko_change_frame(out, frame);
        if (st) return st;
    }
    return KO_OK;
}

int ko_spk_get_state(int64_t id, double jd, int64_t center, int32_t frame,
                     double pos[3], double vel[3]) {
    ko_state s;
    int st = spk_get_state(id, jd, center, frame, &s);
    if (st) return st;
    memcpy(pos, s.pos, 24);
    memcpy(vel, s.vel, 24);
    return KO_OK;
}

/* ------------------------------------------------------------------------ */
/* Force model: constants/gravity.rs:171-253, propagation/nongrav.rs:114-157 */
/* propagation/acceleration.rs:66-80,121-168                                 */
/* ------------------------------------------------------------------------ */
/* gravity.rs:223-235 */
static void j2_correction(const double *p, double radius, double j2, double mass,
                          double *o) {
    double r = norm3(p);
    double zr = p[2] / r;
    double z_squared = 5.0 * (zr * zr);
    double rr = radius / r;
    double coef = 1.5 * j2 * mass * (rr * rr) * powi_m3(r);
    o[0] = p[0] * coef * (z_squared - 1.0);
    o[1] = p[1] * coef * (z_squared - 1.0);
    o[2] = p[2] * coef * (z_squared - 3.0);
}

/* gravity.rs:239-253 */
static void apply_gr_correction(double *acc, const double *p, const double *v,
                                double mass) {
    double r_v = 4.0 * dot3(p, v);
    double rel_v2 = dot3(v, v);
    double r = norm3(p);
    double gr_const = mass * C_AU_PER_DAY_INV_SQUARED * powi_m3(r);
    double c = 4. * mass / r - rel_v2;
    for (int i = 0; i < 3; ++i) acc[i] += gr_const * (c * p[i] + r_v * v[i]);
}

/* GravParams::add_acceleration (gravity.rs:171-216). */
static void grav_add_acceleration(const grav_params *gp, double *acc,
                                  const double *p, const double *v) {
    double mass = gp->mass;
    double k = mass * powi_m3(norm3(p));
    for (int i = 0; i < 3; ++i) acc[i] -= p[i] * k;
    if (gp->naif_id == 5 || gp->naif_id == 10) {
        apply_gr_correction(acc, p, v, mass);
        double pe[3], j[3], je[3];
        equ_to_ecl(p, pe);
        j2_correction(pe, gp->radius, gp->naif_id == 5 ? JUPITER_J2 : SUN_J2, mass, j);
        ecl_to_equ(j, je);
        for (int i = 0; i < 3; ++i) acc[i] += je[i];
    } else if (gp->naif_id == 399) {
        double j[3];
        j2_correction(p, gp->radius, EARTH_J2, mass, j);
        for (int i = 0; i < 3; ++i) acc[i] += j[i];
    }
}

/* NonGravModel::add_acceleration (nongrav.rs:114-157). */
static void nongrav_add_acceleration(const ko_nongrav *ng, double *acc,
                                     const double *pos_in, const double *vel) {
    if (ng->kind == 2) {
        double n = norm3(pos_in);
        double pn[3] = {pos_in[0] / n, pos_in[1] / n, pos_in[2] / n};
        double r_dot = dot3(pn, vel);
        double norm2_inv = 1.0 / dot3(pos_in, pos_in);
        double scaling = GMS * ng->beta * norm2_inv;
        for (int i = 0; i < 3; ++i)
            acc[i] += scaling * ((1.0 - r_dot * C_AU_PER_DAY_INV_SQUARED) * pn[i] -
                                 vel[i] * C_AU_PER_DAY_INV_SQUARED);
    } else if (ng->kind == 1) {
        double pos[3] = {pos_in[0], pos_in[1], pos_in[2]};
        double n = norm3(pos);
        double pn[3] = {pos[0] / n, pos[1] / n, pos[2] / n};
        double vd = dot3(vel, pn);
        double t[3] = {vel[0] - pn[0] * vd, vel[1] - pn[1] * vd, vel[2] - pn[2] * vd};
        double tn = norm3(t);
        for (int i = 0; i < 3; ++i) t[i] /= tn;
        double nv[3];
        cross3(t, pn, nv);
        double nn = norm3(nv);
        for (int i = 0; i < 3; ++i) nv[i] /= nn;
        if (ng->dt != 0.0) {
            double pp[3], vv[3];
            if (ko_analytic_2_body(-ng->dt, pos, vel, pp, vv) == KO_OK) memcpy(pos, pp, 24);
            else { for (int i = 0; i < 3; ++i) acc[i] = NAN; return; }
        }
        double rr0 = norm3(pos) / ng->r_0;
        double scale = ng->alpha * pow(rr0, -ng->m) * pow(1.0 + pow(rr0, ng->n), -ng->k);
        for (int i = 0; i < 3; ++i) acc[i] += pn[i] * (scale * ng->a1);
        for (int i = 0; i < 3; ++i) acc[i] += t[i] * (scale * ng->a2);
        for (int i = 0; i < 3; ++i) acc[i] += nv[i] * (scale * ng->a3);
    }
}

/* spk_accel (acceleration.rs:121-168) with massive_obj = PLANETS, or MASSES
 * when include_asteroids (propagation/mod.rs:82-88). */
int ko_spk_accel_ext(double t, const double pos[3], const double vel[3],
                     const ko_nongrav *ng, int exact_eval, int include_asteroids,
                     double acc[3], int64_t *impact_id) {
    acc[0] = acc[1] = acc[2] = 0.0;
    const int n_bodies = include_asteroids ? N_MASSES : N_PLANETS;
    for (int b = 0; b < n_bodies; ++b) {
        const grav_params *gp = &PLANETS[b];
        ko_state st;
        int rc = spk_get_state(gp->naif_id, t, 0, KO_FRAME_EQUATORIAL, &st);
        if (rc) return rc;
        double rp[3], rv[3];
        for (int i = 0; i < 3; ++i) { rp[i] = pos[i] - st.pos[i]; rv[i] = vel[i] - st.vel[i]; }
        if (exact_eval) {
            double r = norm3(rp);
            if (r <= gp->radius) {
                if (impact_id) *impact_id = gp->naif_id;
                return KO_IMPACT;
            }
        }
        grav_add_acceleration(gp, acc, rp, rv);
        if (gp->naif_id == 10 && ng && ng->kind) nongrav_add_acceleration(ng, acc, rp, rv);
    }
    return KO_OK;
}

int ko_spk_accel(double t, const double pos[3], const double vel[3],
                 const ko_nongrav *ng, int exact_eval, double acc[3],
                 int64_t *impact_id) {
    return ko_spk_accel_ext(t, pos, vel, ng, exact_eval, 0, acc, impact_id);
}

/* ------------------------------------------------------------------------ */
/* Gauss-Radau 15 (propagation/radau.rs:24-346)                              */
/* ------------------------------------------------------------------------ */
static const double H_SP[8] = {
    0.0, 0.05626256053692215, 0.18024069173689236, 0.3526247171131696,
    0.5471536263305554, 0.7342101772154105, 0.8853209468390958, 0.9775206135612875,
};
static double W_VEC[7], U_VEC[7], C_MAT[7][7];
static double WPOW[8][7], UPOW[8][7];
static pthread_once_t radau_once = PTHREAD_ONCE_INIT;
static void radau_tables(void) {
    for (int i = 0; i < 7; ++i) {
        W_VEC[i] = 1.0 / (double)((i + 2) * (i + 3));
        U_VEC[i] = 1.0 / (double)(i + 2);
    }
    memset(C_MAT, 0, sizeof(C_MAT));
    for (int i = 0; i < 7; ++i) C_MAT[i][i] = 1.0;
    for (int i = 0; i < 7; ++i) {
        if (i > 0) C_MAT[i][0] = -H_SP[i] * C_MAT[i - 1][0];
        for (int j = 1; j < i; ++j) C_MAT[i][j] = C_MAT[i - 1][j - 1] - H_SP[i] * C_MAT[i - 1][j];
    }
    /* gauss_radau_frac.powi(1 + idx) * w  (radau.rs:243-256).  powi by
     * repeated squaring as compiler-rt's __powidf2 does. */
    for (int j = 1; j < 8; ++j)
        for (int k = 0; k < 7; ++k) {
            int e = k + 1;
            double a = H_SP[j], r = 1.0;
            while (1) { if (e & 1) r *= a; e /= 2; if (!e) break; a *= a; }
            WPOW[j][k] = r * W_VEC[k];
            UPOW[j][k] = r * U_VEC[k];
        }
}

#define MIN_RATIO 0.25
/* kete's EPSILON (radau.rs:71) unless a caller sets another tolerance for a
 * labelled run (set before a batch call, not during one). */
static double RADAU_EPSILON = 1e-6;
double ko_set_radau_epsilon(double eps) {
    const double old = RADAU_EPSILON;
    if (eps > 0.0) RADAU_EPSILON = eps;
    return old;
}
#define MIN_STEP 0.005

typedef int (*force_fn)(double t, const double *x, const double *v, int exact,
                        double *a, void *ctx);

/* Step schedule of an integration (test hook, no kete counterpart): mode 1
 * records each step attempt (its step size, the sweeps it took, whether it
 * was accepted); mode 2 replays a recording -- the same step sizes and sweep
 * counts whatever the convergence test says -- so that the computed map is a
 * smooth function of the initial state and finite differences of it
 * converge to the derivative of the map with its step sequence held fixed. */
typedef struct {
    int mode;            /* 0 off, 1 record, 2 replay */
    double *h;           /* [cap] step size of each attempt */
    int32_t *sweeps;     /* [cap] sweeps of each attempt */
    int32_t *acc;        /* [cap] 1: accepted */
    int64_t n, cap, pos;
    int bad;             /* record overflow, or replay past its end */
} radau_sched;

static void sched_note(radau_sched *S, double h, int sweeps, int acc) {
    if (!S) return;
    if (S->mode == 1) {
        if (S->n < S->cap) {
            S->h[S->n] = h;
            S->sweeps[S->n] = sweeps;
            S->acc[S->n] = acc;
            S->n++;
        } else {
            S->bad = 1;
        }
    } else if (S->mode == 2) {
        S->pos++;
    }
}

typedef struct {
    force_fn f;
    void *ctx;
    double t;
    double x[3], v[3], a[3];
    double B[7][3];
    ko_counters *cnt;
    radau_sched *sched;  /* NULL: kete's adaptive integrator */
} radau_t;

static inline double max3(const double *v) { return fmax(fmax(v[0], v[1]), v[2]); }

/* RadauIntegrator::step (radau.rs:225-346). */
static int radau_step(radau_t *R, double h, double *next_h) {
    double G[7][3];
    double xs[3], vs[3], ev[3];
    memset(G, 0, sizeof(G));
    memcpy(ev, R->a, 24);
    radau_sched *S = R->sched;
    const int replay = S && S->mode == 2;
    int rsweeps = 10, racc = 0;
    if (replay) {
        if (S->pos >= S->n) { S->bad = 1; return KO_CONVERGENCE; }
        rsweeps = S->sweeps[S->pos];
        racc = S->acc[S->pos];
    }
    for (int sweep = 0; sweep < 10; ++sweep) {
        if (R->cnt) R->cnt->n_sweep++;
        double bprev[3] = {R->B[6][0], R->B[6][1], R->B[6][2]};
        for (int j = 1; j < 8; ++j) {
            double s = H_SP[j];
            for (int d = 0; d < 3; ++d) {
                double bw = 0.0, bu = 0.0;
                for (int k = 0; k < 7; ++k) bw += R->B[k][d] * WPOW[j][k];
                for (int k = 0; k < 7; ++k) bu += R->B[k][d] * UPOW[j][k];
                xs[d] = R->x[d] + s * R->v[d] * h + (s * s) * (h * h) * (R->a[d] / 2.0 + bw);
                vs[d] = R->v[d] + s * h * (R->a[d] + bu);
            }
            if (R->cnt) R->cnt->n_force++;
            int rc = R->f(R->t + s * h, xs, vs, 0, ev, R->ctx);
            if (rc) return rc;
            double gk[3];
            for (int d = 0; d < 3; ++d) gk[d] = (ev[d] - R->a[d]) / s;
            for (int z = 1; z < j; ++z)
                for (int d = 0; d < 3; ++d) gk[d] = (gk[d] - G[z - 1][d]) / (s - H_SP[z]);
            memcpy(G[j - 1], gk, 24);
        }
        /* B = G * C  (radau.rs:313) */
        for (int c = 0; c < 7; ++c)
            for (int d = 0; d < 3; ++d) {
                double acc = G[0][d] * C_MAT[0][c];
                for (int k = 1; k < 7; ++k) acc += G[k][d] * C_MAT[k][c];
                R->B[c][d] = acc;
            }
        double ratio[3], fmx[3];
        for (int d = 0; d < 3; ++d) {
            fmx[d] = fabs(ev[d]) + 1e-6;
            ratio[d] = fabs(R->B[6][d] - bprev[d]) / fmx[d];
        }
        if (replay ? (racc && sweep + 1 == rsweeps) : max3(ratio) < 1e-14) {
            sched_note(S, h, sweep + 1, 1);
            for (int d = 0; d < 3; ++d) {
                double bw = 0.0, bu = 0.0;
                for (int k = 0; k < 7; ++k) bw += R->B[k][d] * W_VEC[k];
                for (int k = 0; k < 7; ++k) bu += R->B[k][d] * U_VEC[k];
                R->x[d] += R->v[d] * h + (h * h) * (R->a[d] * 0.5 + bw);
                R->v[d] += h * (R->a[d] + bu);
            }
            R->t += h;
            if (R->cnt) { R->cnt->n_force++; R->cnt->n_acc++; }
            int rc = R->f(R->t, R->x, R->v, 1, R->a, R->ctx);
            if (rc) return rc;
            double b6[3] = {fabs(R->B[6][0]), fabs(R->B[6][1]), fabs(R->B[6][2])};
            double ratio7 = pow(RADAU_EPSILON * max3(fmx) / max3(b6), 1.0 / 7.0);
            if (ratio7 < MIN_RATIO) ratio7 = MIN_RATIO;
            else if (ratio7 > 1.0 / MIN_RATIO) ratio7 = 1.0 / MIN_RATIO;
            *next_h = h * ratio7;
            return KO_OK;
        }
    }
    sched_note(S, h, 10, 0);
    return KO_CONVERGENCE;
}

/* RadauIntegrator::new + integrate (radau.rs:119-218). */
static int radau_integrate(force_fn f, void *ctx, const double *x0, const double *v0,
                           double t0, double tf, double *xo, double *vo,
                           ko_counters *cnt) {
    pthread_once(&radau_once, radau_tables);
    radau_t R;
    memset(&R, 0, sizeof(R));
    R.f = f; R.ctx = ctx; R.t = t0; R.cnt = cnt;
    memcpy(R.x, x0, 24);
    memcpy(R.v, v0, 24);
    if (cnt) cnt->n_force++;
    int rc = f(t0, R.x, R.v, 1, R.a, ctx);
    if (rc) return rc;
    if (fabs(tf - t0) < 1e-10) { memcpy(xo, R.x, 24); memcpy(vo, R.v, 24); return KO_OK; }
    double h = copysign(0.1, tf - R.t);
    int fails = 0;
    for (;;) {
        if (fabs(R.t - tf) <= fabs(h)) h = tf - R.t;
        double nh;
        rc = radau_step(&R, h, &nh);
        if (rc == KO_OK) {
            h = nh;
            if (fabs(R.t - tf) < 1e-12) { memcpy(xo, R.x, 24); memcpy(vo, R.v, 24); return KO_OK; }
            fails = 0;
        } else if (rc == KO_IMPACT || rc == KO_DAF_LIMITS) {
            return rc;
        } else {
            if (cnt) cnt->n_fail++;
            fails++;
            h *= 0.7;
            if (fails > 10) return KO_CONVERGENCE;
        }
        if (fabs(h) < MIN_STEP) h = copysign(MIN_STEP, h);
    }
}

/* central_accel (acceleration.rs:66-80). */
static int central_force(double t, const double *x, const double *v, int exact,
                         double *a, void *ctx) {
    (void)t; (void)v; (void)exact; (void)ctx;
    double k = powi_m3(norm3(x)) * GMS;
    for (int i = 0; i < 3; ++i) a[i] = -x[i] * k;
    return KO_OK;
}

int ko_radau_central(double t0, double t1, const double pos[3],
                     const double vel[3], double pos_out[3], double vel_out[3],
                     ko_counters *cnt) {
    return radau_integrate(central_force, NULL, pos, vel, t0, t1, pos_out, vel_out, cnt);
}

/* Test hook (no kete counterpart): kete's integrator with one PLANETS body
 * at rest at the origin, its GravParams::add_acceleration the whole force
 * (Newton, and the GR and J2 terms of the Sun, Jupiter and the Earth), so
 * the terms can be pinned by the orbital effects they cause. */
static int one_body_force(double t, const double *x, const double *v, int exact,
                          double *a, void *ctx) {
    (void)t; (void)exact;
    a[0] = a[1] = a[2] = 0.0;
    grav_add_acceleration((const grav_params *)ctx, a, x, v);
    return KO_OK;
}

int ko_radau_one_body(int64_t naif_id, double t0, double t1, const double pos[3],
                      const double vel[3], double pos_out[3], double vel_out[3]) {
    for (int b = 0; b < N_PLANETS; ++b)
        if (PLANETS[b].naif_id == naif_id)
            return radau_integrate(one_body_force, (void *)&PLANETS[b], pos, vel, t0, t1,
                                   pos_out, vel_out, NULL);
    return KO_VALUE_ERROR;
}

typedef struct {
    const ko_nongrav *ng;
    int include_asteroids;
    int64_t impact_id;
    double impact_jd;
} spk_ctx;

static int spk_force(double t, const double *x, const double *v, int exact,
                     double *a, void *ctx) {
    spk_ctx *c = ctx;
    int64_t id = 0;
    int rc = ko_spk_accel_ext(t, x, v, c->ng, exact, c->include_asteroids, a, &id);
    if (rc == KO_IMPACT) { c->impact_id = id; c->impact_jd = t; }
    return rc;
}

/* propagate_n_body_spk (propagation/mod.rs:70-118). */
int ko_propagate_n_body_ext(ko_state *s, double jd_final, int include_asteroids,
                            const ko_nongrav *ng, ko_counters *cnt, int64_t *impact_id,
                            double *impact_jd) {
    int64_t center = s->center;
    int32_t frame = s->frame;
    int rc = ko_change_center(s, 0);
    if (rc) return rc;
    rc = ko_change_frame(s, KO_FRAME_EQUATORIAL);
    if (rc) return rc;
    spk_ctx c = {ng, include_asteroids, 0, 0.0};
    double xo[3], vo[3];
    rc = radau_integrate(spk_force, &c, s->pos, s->vel, s->jd, jd_final, xo, vo, cnt);
    if (rc) {
        if (rc == KO_IMPACT) {
            if (impact_id) *impact_id = c.impact_id;
            if (impact_jd) *impact_jd = c.impact_jd;
        }
        return rc;
    }
    s->jd = jd_final;
    memcpy(s->pos, xo, 24);
    memcpy(s->vel, vo, 24);
    s->frame = KO_FRAME_EQUATORIAL;
    s->center = 0;
    rc = ko_change_center(s, center);
    if (rc) return rc;
    return ko_change_frame(s, frame);
}

int ko_propagate_n_body(ko_state *s, double jd_final, const ko_nongrav *ng,
                        ko_counters *cnt, int64_t *impact_id, double *impact_jd) {
    return ko_propagate_n_body_ext(s, jd_final, 0, ng, cnt, impact_id, impact_jd);
}

/* propagate_two_body (kepler.rs:306-322). */
int ko_propagate_two_body(ko_state *s, double jd_final) {
    double po[3], vo[3];
    int rc = ko_analytic_2_body(jd_final - s->jd, s->pos, s->vel, po, vo);
    if (rc) return rc;
    s->jd = jd_final;
    memcpy(s->pos, po, 24);
    memcpy(s->vel, vo, 24);
    return KO_OK;
}

/* ------------------------------------------------------------------------ */
/* Thread pool: dynamic chunks over objects, as rayon's par_iter is fed     */
/* 1000-object chunks by the pyo3 bindings (src/kete/rust/propagation.rs:98) */
/* ------------------------------------------------------------------------ */
typedef void (*range_fn)(int64_t lo, int64_t hi, void *ctx);
typedef struct {
    atomic_long next;
    int64_t n, chunk;
    range_fn fn;
    void *ctx;
} pool_job;

static void *pool_worker(void *arg) {
    pool_job *j = arg;
    for (;;) {
        int64_t lo = atomic_fetch_add(&j->next, j->chunk);
        if (lo >= j->n) break;
        int64_t hi = lo + j->chunk < j->n ? lo + j->chunk : j->n;
        j->fn(lo, hi, j->ctx);
    }
    return NULL;
}

static void parallel_for(int64_t n, int64_t chunk, int n_threads, range_fn fn, void *ctx) {
    pool_job j;
    atomic_init(&j.next, 0);
    j.n = n; j.chunk = chunk; j.fn = fn; j.ctx = ctx;
    if (n_threads <= 1) { pool_worker(&j); return; }
    pthread_t th[512];
    if (n_threads > 512) n_threads = 512;
    for (int i = 0; i < n_threads; ++i) pthread_create(&th[i], NULL, pool_worker, &j);
    for (int i = 0; i < n_threads; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------------ */
/* The astrometric observable chain of an orbit fit (test infrastructure for */
/* the partials kernel; kete has no derivatives): an N-body arc to each       */
/* observation epoch, one two-body step back by the light time, RA/Dec of     */
/* the observer-to-object vector -- the chain of FovLike::check_n_body        */
/* (fov/fov_like.rs:89-110, simult_states.rs:86-120).                        */
/*  restart = 1: a new integrator at each epoch (RadauIntegrator::new and     */
/*    integrate, radau.rs:119-218: B = 0, an exact evaluation, h = 0.1 d),   */
/*    what successive propagate_n_body calls do, kept in barycentric          */
/*    equatorial coordinates between arcs;                                    */
/*  restart = 0: one integration through all epochs: the step that would      */
/*    pass an epoch is shortened to land on it, B carries on, and the next    */
/*    arc starts from the step size proposed before the shortening.           */
/* With a schedule (mode 1 or 2) the step sequence is recorded or replayed.  */
/* ------------------------------------------------------------------------ */
static int chain_observe(double jd0, const double *x6, int32_t frame, const ko_nongrav *ng,
                         int32_t m, const double *jd_obs, const double *obs, int64_t ostride,
                         int restart, radau_sched *S, double *radec, double *state,
                         int64_t sstride) {
    pthread_once(&radau_once, radau_tables);
    ko_state s0 = {jd0, {x6[0], x6[1], x6[2]}, {x6[3], x6[4], x6[5]}, frame, 10};
    int finite = isfinite(jd0);
    for (int k = 0; k < 6; ++k) finite &= isfinite(x6[k]);
    if (!finite) return KO_VALUE_ERROR;
    int rc = ko_change_center(&s0, 0);
    if (!rc) rc = ko_change_frame(&s0, KO_FRAME_EQUATORIAL);
    if (rc) return rc;
    spk_ctx c = {ng, 0, 0, 0.0};
    radau_t R;
    memset(&R, 0, sizeof(R));
    R.f = spk_force; R.ctx = &c; R.t = jd0; R.sched = S;
    memcpy(R.x, s0.pos, 24);
    memcpy(R.v, s0.vel, 24);
    rc = spk_force(R.t, R.x, R.v, 1, R.a, &c);
    if (rc) return rc;
    const int replay = S && S->mode == 2;
    double h = 0.0, h_prop = 0.0;
    int clamped = 0;
    for (int32_t j = 0; j < m; ++j) {
        const double tf = jd_obs[j * ostride];
        int steps = 1;
        if (j > 0 && restart) {
            memset(R.B, 0, sizeof(R.B));
            rc = spk_force(R.t, R.x, R.v, 1, R.a, &c);
            if (rc) return rc;
        }
        if (j == 0 || restart) {
            if (fabs(tf - R.t) < 1e-10) steps = 0;
            else h = copysign(0.1, tf - R.t);
        } else if (fabs(tf - R.t) < 1e-10) {
            rc = spk_force(R.t, R.x, R.v, 1, R.a, &c);   /* the same epoch again */
            if (rc) return rc;
            steps = 0;
        } else {
            h = copysign(fmax(fabs(clamped ? h_prop : h), MIN_STEP), tf - R.t);
        }
        int fails = 0;
        while (steps) {
            if (replay) {
                if (S->pos >= S->n) { S->bad = 1; return KO_CONVERGENCE; }
                h = S->h[S->pos];
            }
            h_prop = h;
            clamped = fabs(R.t - tf) <= fabs(h);
            if (clamped && !replay) h = tf - R.t;
            double nh;
            rc = radau_step(&R, h, &nh);
            if (rc == KO_OK) {
                h = nh;
                if (fabs(R.t - tf) < 1e-12) break;
                fails = 0;
            } else if (rc == KO_IMPACT || rc == KO_DAF_LIMITS || (S && S->bad)) {
                return rc;
            } else {
                fails++;
                h *= 0.7;
                if (fails > 10) return KO_CONVERGENCE;
            }
            if (fabs(h) < MIN_STEP) h = copysign(MIN_STEP, h);
        }
        /* observables at tf: Sun-centred state, light time, RA/Dec */
        double sp[3], sv[3], xh[3], vh[3], ob[3];
        rc = ko_spk_get_state(10, tf, 0, KO_FRAME_EQUATORIAL, sp, sv);
        if (rc) return rc;
        for (int k = 0; k < 3; ++k) {
            xh[k] = R.x[k] - sp[k];
            vh[k] = R.v[k] - sv[k];
            ob[k] = obs[(3 * j + k) * ostride];
        }
        if (frame == KO_FRAME_ECLIPTIC) {
            ko_state o = {tf, {ob[0], ob[1], ob[2]}, {0, 0, 0}, KO_FRAME_ECLIPTIC, 10};
            ko_change_frame(&o, KO_FRAME_EQUATORIAL);
            memcpy(ob, o.pos, 24);
        }
        double d[3] = {xh[0] - ob[0], xh[1] - ob[1], xh[2] - ob[2]};
        double po[3], vo[3], rd[4], zero[3] = {0, 0, 0};
        rc = ko_analytic_2_body(-(norm3(d) * C_AU_PER_DAY_INV), xh, vh, po, vo);
        if (rc) return rc;
        ko_ra_dec_with_rates(po, vo, ob, zero, rd);
        radec[(2 * j) * ostride] = rd[0];
        radec[(2 * j + 1) * ostride] = rd[1];
        if (state) {
            ko_state o = {tf, {xh[0], xh[1], xh[2]}, {vh[0], vh[1], vh[2]}, KO_FRAME_EQUATORIAL, 10};
            ko_change_frame(&o, frame);
            for (int k = 0; k < 3; ++k) {
                state[(6 * j + k) * sstride] = o.pos[k];
                state[(6 * j + 3 + k) * sstride] = o.vel[k];
            }
        }
    }
    return KO_OK;
}

typedef struct {
    int64_t n;
    const double *jd0;
    const double *const *xyz;
    int32_t frame;
    const ko_nongrav *ng;
    int32_t m;
    const double *jd_obs, *obs;
    int restart, mode;
    int64_t cap;
    double *sh;
    int32_t *ssw, *sacc;
    int64_t *sn;
    double *radec, *state;
    int32_t *status;
} chain_ctx;

static void chain_range(int64_t lo, int64_t hi, void *vctx) {
    chain_ctx *c = vctx;
    for (int64_t i = lo; i < hi; ++i) {
        double x6[6];
        for (int k = 0; k < 6; ++k) x6[k] = c->xyz[k][i];
        radau_sched S = {c->mode, c->sh ? c->sh + i * c->cap : NULL,
                         c->ssw ? c->ssw + i * c->cap : NULL,
                         c->sacc ? c->sacc + i * c->cap : NULL,
                         c->mode == 2 ? c->sn[i] : 0, c->cap, 0, 0};
        for (int32_t j = 0; j < c->m; ++j) {
            c->radec[(2 * j) * c->n + i] = NAN;
            c->radec[(2 * j + 1) * c->n + i] = NAN;
            for (int k = 0; c->state && k < 6; ++k) c->state[(6 * j + k) * c->n + i] = NAN;
        }
        int rc = chain_observe(c->jd0[i], x6, c->frame, c->ng ? &c->ng[i] : NULL, c->m,
                               c->jd_obs + i, c->obs + i, c->n, c->restart,
                               c->mode ? &S : NULL, c->radec + i, c->state ? c->state + i : NULL,
                               c->n);
        if (!rc && S.bad) rc = KO_VALUE_ERROR;
        c->status[i] = rc;
        if (c->mode == 1) c->sn[i] = S.n;
    }
}

int ko_chain_observe_batch(int64_t n, const double *jd0, const double *const *xyz,
                           int32_t frame, const ko_nongrav *ng, int32_t m,
                           const double *jd_obs, const double *obs, int restart, int mode,
                           int64_t cap, double *sched_h, int32_t *sched_sweeps,
                           int32_t *sched_acc, int64_t *sched_n, double *radec, double *state,
                           int32_t *status, int n_threads) {
    chain_ctx c = {n, jd0, xyz, frame, ng, m, jd_obs, obs, restart, mode, cap, sched_h,
                   sched_sweeps, sched_acc, sched_n, radec, state, status};
    parallel_for(n, 8, n_threads, chain_range, &c);
    return KO_OK;
}

/* ---- propagation_kepler_py (src/kete/rust/kepler.rs:68-108) ------------ */
typedef struct {
    int64_t n, m;
    const double *jd0;
    const double *const *xyz;
    int64_t center;
    int32_t frame;
    const double *targets;
    const double *obs;
    double *const *out;
    int32_t *status;
} tb_ctx;

static void tb_range(int64_t lo, int64_t hi, void *vctx) {
    tb_ctx *c = vctx;
    for (int64_t q = lo; q < hi; ++q) {
        int64_t e = q / c->n, i = q % c->n;
        double jd = c->targets[e];
        ko_state s = {c->jd0[i], {c->xyz[0][i], c->xyz[1][i], c->xyz[2][i]},
                      {c->xyz[3][i], c->xyz[4][i], c->xyz[5][i]}, c->frame, c->center};
        int rc = ko_change_center(&s, 10);
        if (!rc) rc = ko_propagate_two_body(&s, jd);
        if (!rc && c->obs) {
            double d[3] = {s.pos[0] - c->obs[0], s.pos[1] - c->obs[1], s.pos[2] - c->obs[2]};
            double delay = -norm3(d) / C_AU_PER_DAY;
            rc = ko_propagate_two_body(&s, s.jd + delay);
        }
        if (!rc) rc = ko_change_center(&s, c->center);
        int64_t o = e * c->n + i;
        c->status[o] = rc;
        for (int k = 0; k < 3; ++k) {
            c->out[k][o] = rc ? NAN : s.pos[k];
            c->out[3 + k][o] = rc ? NAN : s.vel[k];
        }
    }
}

int ko_batch_two_body(int64_t n, const double *jd0, const double *const *xyz,
                      int64_t center, int32_t frame, int64_t m,
                      const double *targets, const double *observer_pos,
                      double *const *out, int32_t *status, int n_threads) {
    tb_ctx c = {n, m, jd0, xyz, center, frame, targets, observer_pos, out, status};
    parallel_for(n * m, 1000, n_threads, tb_range, &c);
    return KO_OK;
}

/* ---- propagation_n_body_spk_py (src/kete/rust/propagation.rs:70-151) --- */
typedef struct {
    const double *jd0;
    const double *const *xyz;
    int64_t center;
    int32_t frame;
    const double *jd_final;
    int stride;
    const ko_nongrav *ng;
    double *const *out;
    int32_t *status;
    int64_t *impact_id;
    double *impact_jd;
    ko_counters *cnt;
    int include_asteroids;
} nb_ctx;

static void nb_range(int64_t lo, int64_t hi, void *vctx) {
    nb_ctx *c = vctx;
    for (int64_t i = lo; i < hi; ++i) {
        ko_state s = {c->jd0[i], {c->xyz[0][i], c->xyz[1][i], c->xyz[2][i]},
                      {c->xyz[3][i], c->xyz[4][i], c->xyz[5][i]}, c->frame, c->center};
        double jf = c->jd_final[c->stride ? i : 0];
        ko_counters tmp = {0, 0, 0, 0};
        int64_t iid = 0;
        double ijd = NAN;
        int rc;
        int finite = isfinite(s.jd);
        for (int k = 0; k < 3; ++k) finite &= isfinite(s.pos[k]) && isfinite(s.vel[k]);
        if (!finite) rc = KO_VALUE_ERROR;
        else rc = ko_propagate_n_body_ext(&s, jf, c->include_asteroids,
                                          c->ng ? &c->ng[i] : NULL, &tmp, &iid, &ijd);
        c->status[i] = rc;
        if (c->impact_id) c->impact_id[i] = rc == KO_IMPACT ? iid : 0;
        if (c->impact_jd) c->impact_jd[i] = rc == KO_IMPACT ? ijd : NAN;
        if (c->cnt) c->cnt[i] = tmp;
        for (int k = 0; k < 3; ++k) {
            c->out[k][i] = rc ? NAN : s.pos[k];
            c->out[3 + k][i] = rc ? NAN : s.vel[k];
        }
    }
}

int ko_batch_n_body_ext(int64_t n, const double *jd0, const double *const *xyz,
                        int64_t center, int32_t frame, const double *jd_final,
                        int jd_final_stride, int include_asteroids, const ko_nongrav *ng,
                        double *const *out, int32_t *status, int64_t *impact_id,
                        double *impact_jd, ko_counters *cnt, int n_threads) {
    nb_ctx c = {jd0, xyz, center, frame, jd_final, jd_final_stride, ng, out,
                status, impact_id, impact_jd, cnt, include_asteroids};
    parallel_for(n, 1000, n_threads, nb_range, &c);
    return KO_OK;
}

int ko_batch_n_body(int64_t n, const double *jd0, const double *const *xyz,
                    int64_t center, int32_t frame, const double *jd_final,
                    int jd_final_stride, const ko_nongrav *ng,
                    double *const *out, int32_t *status, int64_t *impact_id,
                    double *impact_jd, ko_counters *cnt, int n_threads) {
    return ko_batch_n_body_ext(n, jd0, xyz, center, frame, jd_final, jd_final_stride, 0, ng,
                               out, status, impact_id, impact_jd, cnt, n_threads);
}

/* ------------------------------------------------------------------------ */
/* propagate_n_body_long (src/kete/rust/propagation.rs:202-246):            */
/* propagate_n_body_vec (propagation/mod.rs:142-237) with vec_accel          */
/* (acceleration.rs:197-249) over SIMPLE_PLANETS (gravity.rs:137-154), one   */
/* RadauIntegrator over the whole batch state (radau.rs:119-346, any dim).  */
/* ------------------------------------------------------------------------ */
#define N_SIMPLE 9
static const grav_params SIMPLE_PLANETS[N_SIMPLE] = {
    {10, GMS, 0.004654758765894654},
    {1, 1.66012082548908e-07 * GMS, 1.63139354095098e-05},
    {2, 2.44783828779694e-06 * GMS, 4.04537843465442e-05},
    /* EM_BARY (gravity.rs:137-143): Earth + Moon mass, Earth-Moon distance */
    {3, (3.00348961546514e-06 + 3.69430335010988e-08) * GMS, 385000.0 / AU_KM},
    {4, 3.22715608291416e-07 * GMS, 2.27021279387769e-05},
    {5, 0.0009547919099414246 * GMS, 0.00047789450254521576},
    {6, 0.00028588567002459455 * GMS, 0.0004028666966848747},
    {7, 4.36624961322212e-05 * GMS, 0.0001708513622580592},
    {8, 5.15138377265457e-05 * GMS, 0.0001655371154958558},
};

typedef int (*vec_force_fn)(double t, const double *x, const double *v, int exact,
                            double *a, void *ctx);

typedef struct {
    int64_t dim;
    vec_force_fn f;
    void *ctx;
    double t;
    double *x, *v, *a;
    double *B;   /* [7][dim] */
    double *G;   /* [7][dim] */
    double *xs, *vs, *ev, *bprev, *gk;
} radau_vec_t;

/* RadauIntegrator::step (radau.rs:225-346) for a state of any dimension:
 * convergence and the step ratio take maxima over all components. */
static int radau_vec_step(radau_vec_t *R, double h, double *next_h) {
    const int64_t D = R->dim;
    memset(R->G, 0, sizeof(double) * 7 * D);
    memcpy(R->ev, R->a, sizeof(double) * D);
    for (int sweep = 0; sweep < 10; ++sweep) {
        for (int64_t d = 0; d < D; ++d) R->bprev[d] = R->B[6 * D + d];
        for (int j = 1; j < 8; ++j) {
            double s = H_SP[j];
            for (int64_t d = 0; d < D; ++d) {
                double bw = 0.0, bu = 0.0;
                for (int k = 0; k < 7; ++k) bw += R->B[k * D + d] * WPOW[j][k];
                for (int k = 0; k < 7; ++k) bu += R->B[k * D + d] * UPOW[j][k];
                R->xs[d] = R->x[d] + s * R->v[d] * h + (s * s) * (h * h) * (R->a[d] / 2.0 + bw);
                R->vs[d] = R->v[d] + s * h * (R->a[d] + bu);
            }
            int rc = R->f(R->t + s * h, R->xs, R->vs, 0, R->ev, R->ctx);
            if (rc) return rc;
            for (int64_t d = 0; d < D; ++d) {
                double g = (R->ev[d] - R->a[d]) / s;
                for (int z = 1; z < j; ++z) g = (g - R->G[(z - 1) * D + d]) / (s - H_SP[z]);
                R->G[(j - 1) * D + d] = g;
            }
        }
        for (int c = 0; c < 7; ++c)
            for (int64_t d = 0; d < D; ++d) {
                double acc = R->G[d] * C_MAT[0][c];
                for (int k = 1; k < 7; ++k) acc += R->G[k * D + d] * C_MAT[k][c];
                R->B[c * D + d] = acc;
            }
        /* nalgebra's max: a fold of f64::max from the first component, so
         * NaN components are skipped and only an all-NaN vector gives NaN */
        double rmax = NAN, fmax_ = NAN, bmax = NAN;
        for (int64_t d = 0; d < D; ++d) {
            double fm = fabs(R->ev[d]) + 1e-6;
            double r = fabs(R->B[6 * D + d] - R->bprev[d]) / fm;
            rmax = fmax(rmax, r);
            fmax_ = fmax(fmax_, fm);
            bmax = fmax(bmax, fabs(R->B[6 * D + d]));
        }
        if (rmax < 1e-14) {
            for (int64_t d = 0; d < D; ++d) {
                double bw = 0.0, bu = 0.0;
                for (int k = 0; k < 7; ++k) bw += R->B[k * D + d] * W_VEC[k];
                for (int k = 0; k < 7; ++k) bu += R->B[k * D + d] * U_VEC[k];
                R->x[d] += R->v[d] * h + (h * h) * (R->a[d] * 0.5 + bw);
                R->v[d] += h * (R->a[d] + bu);
            }
            R->t += h;
            int rc = R->f(R->t, R->x, R->v, 1, R->a, R->ctx);
            if (rc) return rc;
            double ratio7 = pow(RADAU_EPSILON * fmax_ / bmax, 1.0 / 7.0);
            if (ratio7 < MIN_RATIO) ratio7 = MIN_RATIO;
            else if (ratio7 > 1.0 / MIN_RATIO) ratio7 = 1.0 / MIN_RATIO;
            *next_h = h * ratio7;
            return KO_OK;
        }
    }
    return KO_CONVERGENCE;
}

/* RadauIntegrator::new + integrate (radau.rs:119-218), any dimension. */
static int radau_vec_integrate(vec_force_fn f, void *ctx, int64_t D, double *x, double *v,
                               double t0, double tf) {
    pthread_once(&radau_once, radau_tables);
    radau_vec_t R;
    memset(&R, 0, sizeof(R));
    R.dim = D; R.f = f; R.ctx = ctx; R.t = t0;
    double *mem = calloc((size_t)D * 21, sizeof(double));
    if (!mem) return KO_VALUE_ERROR;
    R.x = x; R.v = v;
    R.a = mem; R.B = mem + D; R.G = mem + 8 * D; R.xs = mem + 15 * D; R.vs = mem + 16 * D;
    R.ev = mem + 17 * D; R.bprev = mem + 18 * D; R.gk = mem + 19 * D;
    int rc = f(t0, R.x, R.v, 1, R.a, ctx);
    if (rc) { free(mem); return rc; }
    if (fabs(tf - t0) < 1e-10) { free(mem); return KO_OK; }
    double h = copysign(0.1, tf - R.t);
    int fails = 0;
    for (;;) {
        if (fabs(R.t - tf) <= fabs(h)) h = tf - R.t;
        double nh;
        rc = radau_vec_step(&R, h, &nh);
        if (rc == KO_OK) {
            h = nh;
            if (fabs(R.t - tf) < 1e-12) { free(mem); return KO_OK; }
            fails = 0;
        } else if (rc == KO_IMPACT || rc == KO_DAF_LIMITS) {
            free(mem);
            return rc;
        } else {
            fails++;
            h *= 0.7;
            if (fails > 10) { free(mem); return KO_CONVERGENCE; }
        }
        if (fabs(h) < MIN_STEP) h = copysign(MIN_STEP, h);
    }
}

typedef struct {
    int64_t n_obj;            /* massive + particles */
    const ko_nongrav *ng;     /* per particle (kind 0: none) */
    int64_t impact_id;
    double impact_jd;
} vec_ctx;

/* vec_accel (acceleration.rs:197-249).  kete applies a particle's non-grav
 * model only for idx > n_massive, so the first particle of a batch never
 * gets its model; restated as is. */
static int vec_force(double t, const double *x, const double *v, int exact, double *a,
                     void *vctx) {
    vec_ctx *c = vctx;
    for (int64_t i = 0; i < 3 * c->n_obj; ++i) a[i] = 0.0;
    for (int64_t idx = 0; idx < c->n_obj; ++idx) {
        for (int idy = 0; idy < N_SIMPLE; ++idy) {
            if (idx == idy) continue;
            const grav_params *gp = &SIMPLE_PLANETS[idy];
            double w[3] = {0.0, 0.0, 0.0}, rp[3], rv[3];
            for (int k = 0; k < 3; ++k) {
                rp[k] = x[3 * idx + k] - x[3 * idy + k];
                rv[k] = v[3 * idx + k] - v[3 * idy + k];
            }
            if (exact && norm3(rp) <= gp->radius) {
                c->impact_id = gp->naif_id;
                c->impact_jd = t;
                return KO_IMPACT;
            }
            grav_add_acceleration(gp, w, rp, rv);
            if (gp->naif_id == 10 && idx > N_SIMPLE) {
                const ko_nongrav *ng = c->ng ? &c->ng[idx - N_SIMPLE] : NULL;
                if (ng && ng->kind) nongrav_add_acceleration(ng, w, rp, rv);
            }
            for (int k = 0; k < 3; ++k) a[3 * idx + k] += w[k];
        }
    }
    return KO_OK;
}

/* propagate_n_body_vec (propagation/mod.rs:142-237) for one batch: m
 * particles at jd0 (Sun-centred equatorial), the 9 SIMPLE_PLANETS states
 * planets[9][6] (Sun-centred equatorial at jd0); out: particles then planets,
 * Sun-centred equatorial at jd_final. */
static int long_batch(int64_t m, double jd0, const double *px, const ko_nongrav *ng,
                      const double *planets, double jd_final, double *out_part,
                      double *out_planets) {
    const int64_t n = N_SIMPLE + m, D = 3 * n;
    double *x = malloc(sizeof(double) * 2 * D);
    if (!x) return KO_VALUE_ERROR;
    double *v = x + D;
    for (int b = 0; b < N_SIMPLE; ++b)
        for (int k = 0; k < 3; ++k) { x[3 * b + k] = planets[6 * b + k]; v[3 * b + k] = planets[6 * b + 3 + k]; }
    for (int64_t i = 0; i < m; ++i)
        for (int k = 0; k < 3; ++k) {
            x[3 * (N_SIMPLE + i) + k] = px[6 * i + k];
            v[3 * (N_SIMPLE + i) + k] = px[6 * i + 3 + k];
        }
    vec_ctx c = {n, ng, 0, 0.0};
    int rc = radau_vec_integrate(vec_force, &c, D, x, v, jd0, jd_final);
    if (rc == KO_OK) {
        /* states relative to the Sun's final state */
        for (int64_t i = 0; i < n; ++i) {
            double *o = i < N_SIMPLE ? out_planets + 6 * i : out_part + 6 * (i - N_SIMPLE);
            for (int k = 0; k < 3; ++k) {
                o[k] = x[3 * i + k] - x[k];
                o[3 + k] = v[3 * i + k] - v[k];
            }
        }
    }
    free(x);
    return rc;
}

typedef struct {
    int64_t n, batch;
    double jd0, jd_final;
    const double *px;        /* [n][6] equatorial Sun-centred */
    const ko_nongrav *ng;    /* [n] or NULL */
    const double *planets;   /* [9][6] */
    double *out;             /* [n][6] */
    double *planets_out;     /* [n_batches][9][6] */
    int32_t *status;         /* [n_batches] */
} long_ctx;

static void long_range(int64_t lo, int64_t hi, void *vctx) {
    long_ctx *c = vctx;
    for (int64_t b = lo; b < hi; ++b) {
        const int64_t i0 = b * c->batch;
        const int64_t m = i0 + c->batch <= c->n ? c->batch : c->n - i0;
        c->status[b] = long_batch(m, c->jd0, c->px + 6 * i0, c->ng ? c->ng + i0 : NULL,
                                  c->planets, c->jd_final, c->out + 6 * i0,
                                  c->planets_out + 54 * b);
    }
}

int ko_propagate_n_body_long(int64_t n, double jd0, const double *const *xyz, int32_t frame,
                             double jd_final, const double *planet_xyz, const ko_nongrav *ng,
                             int64_t batch_size, double *const *out, double *planets_out,
                             int n_threads) {
    if (n <= 0) return KO_VALUE_ERROR;        /* "State vector is empty" */
    if (batch_size <= 0) return KO_VALUE_ERROR;
    double planets[N_SIMPLE * 6];
    if (planet_xyz) {
        memcpy(planets, planet_xyz, sizeof(planets));
    } else {
        for (int b = 0; b < N_SIMPLE; ++b)
            if (ko_spk_get_state(SIMPLE_PLANETS[b].naif_id, jd0, 10, KO_FRAME_EQUATORIAL,
                                 planets + 6 * b, planets + 6 * b + 3) != KO_OK)
                return KO_DAF_LIMITS;
    }
    double *px = malloc(sizeof(double) * 6 * n);
    double *po = malloc(sizeof(double) * 6 * n);
    const int64_t nb = (n + batch_size - 1) / batch_size;
    double *pl = malloc(sizeof(double) * 54 * nb);
    int32_t *st = malloc(sizeof(int32_t) * nb);
    for (int64_t i = 0; i < n; ++i) {
        ko_state s = {jd0, {xyz[0][i], xyz[1][i], xyz[2][i]}, {xyz[3][i], xyz[4][i], xyz[5][i]},
                      frame, 10};
        int rc = ko_change_frame(&s, KO_FRAME_EQUATORIAL);
        if (rc) { free(px); free(po); free(pl); free(st); return rc; }
        memcpy(px + 6 * i, s.pos, 24);
        memcpy(px + 6 * i + 3, s.vel, 24);
    }
    long_ctx c = {n, batch_size, jd0, jd_final, px, ng, planets, po, pl, st};
    parallel_for(nb, 1, n_threads, long_range, &c);
    int rc = KO_OK;
    for (int64_t b = 0; b < nb && rc == KO_OK; ++b) rc = st[b];
    if (rc == KO_OK) {
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 6; ++k) out[k][i] = po[6 * i + k];
        if (planets_out) memcpy(planets_out, pl + 54 * (nb - 1), sizeof(double) * 54);
    }
    free(px); free(po); free(pl); free(st);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* FOV: fov/patches.rs, fov/fov_like.rs, fov/generic.rs                      */
/* ------------------------------------------------------------------------ */
/* rotate_around (frames/definitions.rs:356-363): Rodrigues rotation matrix
 * from a normalised axis, identity for a zero angle. */
static void rotate_around(const double *v, const double *axis, double angle, double *o) {
    if (!(angle != 0.0)) { memcpy(o, v, 24); return; }
    double n = norm3(axis);
    double ux = axis[0] / n, uy = axis[1] / n, uz = axis[2] / n;
    double sqx = ux * ux, sqy = uy * uy, sqz = uz * uz;
    double s = sin(angle), c = cos(angle);
    double omc = 1.0 - c;
    double m[3][3] = {
        {sqx + (1.0 - sqx) * c, ux * uy * omc - uz * s, ux * uz * omc + uy * s},
        {ux * uy * omc + uz * s, sqy + (1.0 - sqy) * c, uy * uz * omc - ux * s},
        {ux * uz * omc - uy * s, uy * uz * omc + ux * s, sqz + (1.0 - sqz) * c},
    };
    double x = v[0], y = v[1], z = v[2];
    for (int i = 0; i < 3; ++i) o[i] = m[i][0] * x + m[i][1] * y + m[i][2] * z;
}

/* OnSkyRectangle::new (fov/patches.rs:122-155). */
void ko_rect_normals(const double pointing[3], double rotation, double lon_w,
                     double lat_w, double normals[12]) {
    const double zax[3] = {0.0, 0.0, 1.0};
    double up[3], left[3], nleft[3], nup[3];
    rotate_around(zax, pointing, -rotation, up);
    cross3(pointing, up, left);
    cross3(pointing, left, up);
    rotate_around(left, up, -lon_w / 2.0, normals + 0);
    rotate_around(up, left, lat_w / 2.0, normals + 3);
    for (int i = 0; i < 3; ++i) { nleft[i] = -left[i]; nup[i] = -up[i]; }
    rotate_around(nleft, up, lon_w / 2.0, normals + 6);
    rotate_around(nup, left, -lat_w / 2.0, normals + 9);
}

/* SphericalPolygon::contains (fov/patches.rs:214-234). */
int ko_rect_contains(const double nrm[12], const double v[3], double *dist) {
    double closest = -INFINITY;
    for (int e = 0; e < 4; ++e) {
        double d = dot3(v, nrm + 3 * e);
        if (isnan(d)) { *dist = d; return 0; }
        if (signbit(d) && fabs(d) > closest) closest = fabs(d);
    }
    if (isfinite(closest)) { *dist = fmin(closest, norm3(v)); return 0; }
    *dist = 0.0;
    return 1;
}

typedef struct {
    const double *nrm;
    double jd;
    double pos[3], vel[3];
    int32_t frame;
    int include_asteroids;  /* check_visible's flag (fov_like.rs:137) */
} fov_t;

/* check_linear (fov_like.rs:43-66); only the Contains result is used. */
static int check_linear(const fov_t *f, const ko_state *s, double *dist) {
    double rel[3];
    for (int i = 0; i < 3; ++i) rel[i] = s->pos[i] - f->pos[i];
    double dt = f->jd - s->jd - norm3(rel) * C_AU_PER_DAY_INV;
    double nr[3];
    for (int i = 0; i < 3; ++i) nr[i] = (s->pos[i] + s->vel[i] * dt) - f->pos[i];
    return ko_rect_contains(f->nrm, nr, dist);
}

/* check_two_body (fov_like.rs:71-85).  Returns status; *inside, *dist, *out. */
static int check_two_body(const fov_t *f, const ko_state *s, int *inside,
                          double *dist, ko_state *out) {
    ko_state fs = *s;
    int rc = ko_propagate_two_body(&fs, f->jd);
    if (rc) return rc;
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = fs.pos[i] - f->pos[i];
    double dt = -norm3(d) * C_AU_PER_DAY_INV;
    rc = ko_propagate_two_body(&fs, f->jd + dt);
    if (rc) return rc;
    for (int i = 0; i < 3; ++i) d[i] = fs.pos[i] - f->pos[i];
    *inside = ko_rect_contains(f->nrm, d, dist);
    *out = fs;
    return KO_OK;
}

/* check_n_body (fov_like.rs:90-108). */
static int check_n_body(const fov_t *f, const ko_state *s, int *inside,
                        double *dist, ko_state *out) {
    ko_state es = *s;
    int rc = ko_propagate_n_body_ext(&es, f->jd, f->include_asteroids, NULL, NULL, NULL, NULL);
    if (rc) return rc;
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = es.pos[i] - f->pos[i];
    double dt = -norm3(d) * C_AU_PER_DAY_INV;
    rc = ko_propagate_two_body(&es, f->jd + dt);
    if (rc) return rc;
    for (int i = 0; i < 3; ++i) d[i] = es.pos[i] - f->pos[i];
    *inside = ko_rect_contains(f->nrm, d, dist);
    *out = es;
    return KO_OK;
}

/* The per-state closure of FovLike::check_visible (fov_like.rs:140-176).
 * Returns 1 and fills *out when the object is seen. */
static int check_visible_one(const fov_t *f, const ko_state *s, double dt_limit,
                             ko_state *out) {
    double dv[3];
    for (int i = 0; i < 3; ++i) dv[i] = s->vel[i] - f->vel[i];
    double max_dist = norm3(dv) * dt_limit * 2.0;
    int inside;
    double dist;
    if (fabs(s->jd - f->jd) < dt_limit) {
        inside = check_linear(f, s, &dist);
        if (!inside && dist > max_dist) return 0;
        if (check_two_body(f, s, &inside, &dist, out)) return 0;
        return inside;
    }
    ko_state tmp;
    if (check_two_body(f, s, &inside, &dist, &tmp)) return 0;
    if (!inside && dist > max_dist) return 0;
    if (check_n_body(f, s, &inside, &dist, out)) return 0;
    return inside;
}

typedef struct {
    const fov_t *f;
    const double *jd;
    const double *const *xyz;
    double dt_limit;
    int32_t frame;
    int8_t *vis;
    ko_state *res;
} cv_ctx;

static void cv_range(int64_t lo, int64_t hi, void *vctx) {
    cv_ctx *c = vctx;
    for (int64_t i = lo; i < hi; ++i) {
        ko_state s = {c->jd[i], {c->xyz[0][i], c->xyz[1][i], c->xyz[2][i]},
                      {c->xyz[3][i], c->xyz[4][i], c->xyz[5][i]}, c->frame, 10};
        c->vis[i] = (int8_t)check_visible_one(c->f, &s, c->dt_limit, &c->res[i]);
    }
}

int64_t ko_fov_check_visible(const double normals[12], double obs_jd,
                             const double obs_pos[3], const double obs_vel[3],
                             int32_t frame, int64_t n, const double *jd,
                             const double *const *xyz, double dt_limit,
                             int64_t *vis_idx, double *const *vis_state,
                             int n_threads) {
    fov_t f = {normals, obs_jd, {obs_pos[0], obs_pos[1], obs_pos[2]},
               {obs_vel[0], obs_vel[1], obs_vel[2]}, frame, 0};
    int8_t *vis = calloc(n > 0 ? n : 1, 1);
    ko_state *res = calloc(n > 0 ? n : 1, sizeof(ko_state));
    cv_ctx c = {&f, jd, xyz, dt_limit, frame, vis, res};
    parallel_for(n, 1000, n_threads, cv_range, &c);
    int64_t nv = 0;
    for (int64_t i = 0; i < n; ++i)
        if (vis[i]) {
            vis_idx[nv] = i;
            if (vis_state) {
                vis_state[0][nv] = res[i].jd;
                for (int k = 0; k < 3; ++k) {
                    vis_state[1 + k][nv] = res[i].pos[k];
                    vis_state[4 + k][nv] = res[i].vel[k];
                }
            }
            nv++;
        }
    free(vis);
    free(res);
    return nv;
}

/* Test hook: the re-propagation schedule of the last ko_fov_state_check_ext
 * call, as (kind, epoch) pairs in call order -- 0: a field group and its mean
 * epoch, 1: a big step of the big-step copy to the epoch given, 2: the
 * small-step states reset from the big-step copy, 3: a small step to the
 * epoch given.  Off unless a buffer is set. */
static double *g_fov_trace = NULL;
static int64_t g_fov_trace_cap = 0, g_fov_trace_n = 0;

void ko_fov_trace_set(double *buf, int64_t cap) {
    g_fov_trace = buf;
    g_fov_trace_cap = buf ? cap : 0;
    g_fov_trace_n = 0;
}

int64_t ko_fov_trace_len(void) { return g_fov_trace_n; }

static void fov_trace(int kind, double jd) {
    if (g_fov_trace && g_fov_trace_n < g_fov_trace_cap) {
        g_fov_trace[2 * g_fov_trace_n] = kind;
        g_fov_trace[2 * g_fov_trace_n + 1] = jd;
        g_fov_trace_n++;
    }
}

/* fov_checks_py (src/kete/rust/fovs/checks.rs:24-110).  Objects that fail a
 * re-propagation are dropped, as filter_map does there; the surviving
 * objects keep their original index so hits can be matched to inputs. */
typedef struct {
    int64_t n;
    ko_state *st;
    int8_t *alive;
    double jd;
    int include_asteroids;
} nb_prop_ctx;

static void prop_range(int64_t lo, int64_t hi, void *vctx) {
    nb_prop_ctx *c = vctx;
    for (int64_t i = lo; i < hi; ++i) {
        if (!c->alive[i]) continue;
        if (ko_propagate_n_body_ext(&c->st[i], c->jd, c->include_asteroids, NULL, NULL, NULL,
                                    NULL))
            c->alive[i] = 0;
    }
}

typedef struct {
    fov_t *fovs;
    int64_t nf;
    ko_state *st;
    int8_t *alive;
    int64_t n;
    double dt_limit;
    int8_t *vis;     /* [nf][n] */
    ko_state *res;   /* [nf][n] */
} grp_ctx;

static void grp_range(int64_t lo, int64_t hi, void *vctx) {
    grp_ctx *c = vctx;
    for (int64_t q = lo; q < hi; ++q) {
        int64_t fi = q / c->n, i = q % c->n;
        c->vis[q] = 0;
        if (!c->alive[i]) continue;
        c->vis[q] = (int8_t)check_visible_one(&c->fovs[fi], &c->st[i], c->dt_limit, &c->res[q]);
    }
}

int64_t ko_fov_state_check_ext(int64_t n, double jd0, const double *const *xyz,
                               int32_t frame, int64_t n_fields,
                               const double *field_jd, const double *obs_xyz,
                               const double *normals, double dt_limit, int include_asteroids,
                               int64_t cap, int64_t *hit_field, int64_t *hit_obj,
                               double *hit_state, int n_threads) {
    ko_state *states = malloc(sizeof(ko_state) * (n > 0 ? n : 1));
    ko_state *big = malloc(sizeof(ko_state) * (n > 0 ? n : 1));
    int8_t *alive = malloc(n > 0 ? n : 1), *big_alive = malloc(n > 0 ? n : 1);
    for (int64_t i = 0; i < n; ++i) {
        ko_state s = {jd0, {xyz[0][i], xyz[1][i], xyz[2][i]}, {xyz[3][i], xyz[4][i], xyz[5][i]},
                      frame, 10};
        states[i] = big[i] = s;
        alive[i] = big_alive[i] = 1;
    }
    fov_t *fovs = malloc(sizeof(fov_t) * (n_fields > 0 ? n_fields : 1));
    for (int64_t f = 0; f < n_fields; ++f) {
        fovs[f].nrm = normals + 12 * f;
        fovs[f].jd = field_jd[f];
        for (int k = 0; k < 3; ++k) {
            fovs[f].pos[k] = obs_xyz[6 * f + k];
            fovs[f].vel[k] = obs_xyz[6 * f + 3 + k];
        }
        fovs[f].frame = frame;
        fovs[f].include_asteroids = include_asteroids;
    }
    double jd = jd0, big_jd = jd0;
    int64_t nh = 0;
    int64_t f0 = 0;
    while (f0 < n_fields) {
        /* chunk: fields whose jd is within 2*dt_limit of the chunk's first */
        int64_t f1 = f0 + 1;
        while (f1 < n_fields && !(fabs(field_jd[f1] - field_jd[f0]) >= 2.0 * dt_limit)) f1++;
        double jd_mean = (field_jd[f1 - 1] + field_jd[f0]) / 2.0;
        fov_trace(0, jd_mean);
        if (fabs(jd_mean - big_jd) >= dt_limit * 50.0) {
            big_jd = jd_mean;
            nb_prop_ctx pc = {n, big, big_alive, jd, include_asteroids};
            fov_trace(1, pc.jd);
            parallel_for(n, 1000, n_threads, prop_range, &pc);
        }
        if (fabs(jd_mean - jd) >= dt_limit) {
            if (fabs(jd - big_jd) >= dt_limit * 25.0) {
                fov_trace(2, jd);
                memcpy(states, big, sizeof(ko_state) * n);
                memcpy(alive, big_alive, n);
            }
            jd = jd_mean;
            nb_prop_ctx pc = {n, states, alive, jd, include_asteroids};
            fov_trace(3, pc.jd);
            parallel_for(n, 1000, n_threads, prop_range, &pc);
        }
        int64_t nf = f1 - f0;
        int8_t *vis = malloc(nf * n > 0 ? nf * n : 1);
        ko_state *res = malloc(sizeof(ko_state) * (nf * n > 0 ? nf * n : 1));
        grp_ctx gc = {fovs + f0, nf, states, alive, n, dt_limit, vis, res};
        parallel_for(nf * n, 1000, n_threads, grp_range, &gc);
        for (int64_t q = 0; q < nf * n; ++q) {
            if (!vis[q]) continue;
            if (nh >= cap) { nh = -1; free(vis); free(res); goto out; }
            hit_field[nh] = f0 + q / n;
            hit_obj[nh] = q % n;
            double *hs = hit_state + 7 * nh;
            hs[0] = res[q].jd;
            for (int k = 0; k < 3; ++k) { hs[1 + k] = res[q].pos[k]; hs[4 + k] = res[q].vel[k]; }
            nh++;
        }
        free(vis);
        free(res);
        f0 = f1;
    }
out:
    free(states); free(big); free(alive); free(big_alive); free(fovs);
    return nh;
}

/* fov_spk_checks_py (src/kete/rust/fovs/checks.rs:123-142) over fields given
 * in epoch order, each by FovLike::check_spks (fov/fov_like.rs:196-226): the
 * state of every id from the loaded SPK at the field epoch (the observer's
 * center, the Sun, and the field's frame), ids without one skipped; then
 * check_two_body; visible ones in id order. */
int64_t ko_fov_spk_check(int64_t n_ids, const int64_t *ids, int32_t frame, int64_t n_fields,
                         const double *field_jd, const double *obs_xyz, const double *normals,
                         int64_t cap, int64_t *hit_field, int64_t *hit_obj, double *hit_state) {
    int64_t nh = 0;
    for (int64_t f = 0; f < n_fields; ++f) {
        fov_t F = {normals + 12 * f, field_jd[f],
                   {obs_xyz[6 * f], obs_xyz[6 * f + 1], obs_xyz[6 * f + 2]},
                   {obs_xyz[6 * f + 3], obs_xyz[6 * f + 4], obs_xyz[6 * f + 5]}, frame, 0};
        for (int64_t k = 0; k < n_ids; ++k) {
            ko_state s = {F.jd, {0, 0, 0}, {0, 0, 0}, frame, 10};
            if (ko_spk_get_state(ids[k], F.jd, 10, frame, s.pos, s.vel) != KO_OK) continue;
            int inside = 0;
            double dist;
            ko_state out;
            if (check_two_body(&F, &s, &inside, &dist, &out) != KO_OK || !inside) continue;
            if (nh >= cap) return -1;
            hit_field[nh] = f;
            hit_obj[nh] = k;
            double *hs = hit_state + 7 * nh;
            hs[0] = out.jd;
            for (int q = 0; q < 3; ++q) { hs[1 + q] = out.pos[q]; hs[4 + q] = out.vel[q]; }
            nh++;
        }
    }
    return nh;
}

int64_t ko_fov_state_check(int64_t n, double jd0, const double *const *xyz,
                           int32_t frame, int64_t n_fields,
                           const double *field_jd, const double *obs_xyz,
                           const double *normals, double dt_limit, int64_t cap,
                           int64_t *hit_field, int64_t *hit_obj, double *hit_state,
                           int n_threads) {
    return ko_fov_state_check_ext(n, jd0, xyz, frame, n_fields, field_jd, obs_xyz, normals,
                                  dt_limit, 0, cap, hit_field, hit_obj, hit_state, n_threads);
}

/* ------------------------------------------------------------------------ */
/* Epilogue: simult_states.rs:86-120, frames/definitions.rs:337-345,         */
/* flux/reflected.rs:23-33,255-268                                           */
/* ------------------------------------------------------------------------ */
void ko_ra_dec_with_rates(const double pos[3], const double vel[3],
                          const double obs_pos[3], const double obs_vel[3],
                          double out[4]) {
    double p[3], v[3];
    for (int i = 0; i < 3; ++i) { p[i] = pos[i] - obs_pos[i]; v[i] = vel[i] - obs_vel[i]; }
    double d_ra = (p[0] * v[1] - p[1] * v[0]) / (p[0] * p[0] + p[1] * p[1]);
    double r2 = dot3(p, p);
    double d_dec = (v[2] - p[2] * dot3(p, v) / r2) / sqrt(r2 - p[2] * p[2]);
    /* to_lat_lon */
    double r = norm3(p), dec, ra;
    if (r < 1e-10) { dec = 0.0; ra = 0.0; }
    else {
        double zr = p[2] / r;
        if (zr < -1.0) zr = -1.0;
        if (zr > 1.0) zr = 1.0;
        dec = rem_euclid(3.0 * (PI_ / 2.0) - acos(zr), TAU) - PI_;
        ra = rem_euclid(atan2(p[1], p[0]), TAU);
    }
    out[0] = ra; out[1] = dec; out[2] = d_ra * cos(dec); out[3] = d_dec;
}

static double hg_helper(double a, double b, double c, double phase) {
    double phase_s = sin(phase);
    double theta_l = exp(-a * pow(tan(0.5 * phase), b));
    double theta_s = 1.0 - c * phase_s / (0.119 + 1.341 * phase_s - 0.754 * (phase_s * phase_s));
    double t = tan(0.5 * phase);
    double w = exp(-90.56 * (t * t));
    return w * theta_s + (1.0 - w) * theta_l;
}

double ko_hg_phase_curve_correction(double g, double phase) {
    return (1.0 - g) * hg_helper(3.332, 0.631, 0.986, phase) +
           g * hg_helper(1.862, 1.218, 0.238, phase);
}

double ko_hg_apparent_mag(double h, double g, const double sun2obj[3],
                          const double sun2obs[3]) {
    double obj_r = norm3(sun2obj);
    double o2o[3], neg[3];
    for (int i = 0; i < 3; ++i) { o2o[i] = sun2obs[i] - sun2obj[i]; neg[i] = -o2o[i]; }
    double o2o_r = norm3(o2o);
    /* nalgebra angle(): acos of the clamped cosine */
    double n1 = obj_r, n2 = norm3(neg), phase;
    if (n1 == 0.0 || n2 == 0.0) phase = 0.0;
    else {
        double cang = dot3(sun2obj, neg) / (n1 * n2);
        phase = cang > 1.0 ? 0.0 : (cang < -1.0 ? PI_ : acos(cang));
    }
    if (phase > 2.792526803190927) return NAN;
    double corr = log10(ko_hg_phase_curve_correction(g, phase));
    return h + 5.0 * log10(obj_r * o2o_r) - 2.5 * corr;
}

/* ------------------------------------------------------------------------ */
/* MOID (propagation/kepler.rs:325-404; moid_py src/kete/rust/propagation.rs:27-39) */
/* ------------------------------------------------------------------------ */
/* nalgebra Vector::angle: acos of the cosine clamped to [-1, 1], 0 when
 * either vector is zero. */
static double vec_angle(const double *a, const double *b) {
    double prod = dot3(a, b), n1 = norm3(a), n2 = norm3(b);
    if (n1 == 0.0 || n2 == 0.0) return 0.0;
    double c = prod / (n1 * n2);
    return c > 1.0 ? 0.0 : (c < -1.0 ? PI_ : acos(c));
}

/* CometElements::from_pos_vel (elements.rs:62-163), the two quantities moid
 * uses: the perihelion time and orbital_period (elements.rs:278-293). */
static void moid_elements(double epoch, const double *pos, const double *vel,
                          double *peri_time, double *period) {
    double vs[3] = {vel[0] / GMS_SQRT, vel[1] / GMS_SQRT, vel[2] / GMS_SQRT};
    double v_mag2 = dot3(vs, vs);
    double p_mag = norm3(pos);
    double vp_mag = dot3(pos, vs);
    double ecc_vec[3], ang[3];
    for (int k = 0; k < 3; ++k) ecc_vec[k] = (v_mag2 - 1.0 / p_mag) * pos[k] - vp_mag * vs[k];
    cross3(pos, vs, ang);
    double lon[3] = {-ang[1], ang[0], 0.0};
    double ecc = norm3(ecc_vec), ang_mag = norm3(ang), lon_mag = norm3(lon);
    double peri_dist = (ang_mag * ang_mag) / (1.0 + ecc);
    if (lon_mag < 1e-8) { lon[0] = 1.0; lon[1] = 0.0; lon[2] = 0.0; }
    double pt;
    if (fabs(ecc - 1.0) < PARABOLIC_ECC_LIMIT) {
        double ta = vec_angle(ecc_vec, pos);
        if (signbit(vp_mag)) ta = -ta;
        double d = tan(ta / 2.0);
        double dt = (sqrt(2.0) * pow(peri_dist, 1.5) / GMS_SQRT) * (d + (d * d * d) / 3.0);
        pt = epoch - dt;
    } else if (ecc < 1e-6) {
        double a = 1.0 / (2.0 / p_mag - v_mag2);
        double mm = pow(fabs(a), -1.5) * GMS_SQRT;
        double ta = vec_angle(lon, pos);
        double c[3];
        cross3(ang, lon, c);
        if (signbit(dot3(c, pos))) ta = -ta;
        pt = epoch - ta / mm;
    } else {
        double a = 1.0 / (2.0 / p_mag - v_mag2);
        double mm = pow(fabs(a), -1.5) * GMS_SQRT;
        double x_bar = (ang_mag * ang_mag - p_mag) / ecc;
        double y_bar = vp_mag / ecc * ang_mag;
        double b = a * sqrt(fabs(1.0 - ecc * ecc));
        double s_e = y_bar / b;
        double m;
        if (ecc < 1.0) {
            double c_e = x_bar / a + ecc;
            m = atan2(s_e, c_e) - ecc * s_e;
        } else {
            m = -ecc * s_e - asinh(-s_e);
        }
        pt = epoch - m / mm;
    }
    *peri_time = pt;
    /* semi_major: NaN for parabolic; period infinite for a <= 1e-8 */
    double sm = fabs(ecc - 1.0) <= PARABOLIC_ECC_LIMIT ? NAN : peri_dist / (1.0 - ecc);
    *period = sm <= 1e-8 ? INFINITY : TAU * pow(sm, 1.5) / GMS_SQRT;
}

void ko_comet_elements_peri(double epoch, const double pos[3], const double vel[3],
                            double *peri_time, double *period) {
    moid_elements(epoch, pos, vel, peri_time, period);
}

/* MoidCost::cost (kepler.rs:333-344): distance between the two states moved
 * by param = [dt_a, dt_b] days. */
static int moid_cost(const ko_state *a, const ko_state *b, const double prm[2], double *out) {
    ko_state sa = *a, sb = *b;
    int rc = ko_propagate_two_body(&sa, a->jd + prm[0]);
    if (!rc) rc = ko_propagate_two_body(&sb, b->jd + prm[1]);
    if (rc) return rc;
    double d[3] = {sa.pos[0] - sb.pos[0], sa.pos[1] - sb.pos[1], sa.pos[2] - sb.pos[2]};
    *out = norm3(d);
    return KO_OK;
}

/* Nelder-Mead of the argmin crate (argmin ^0.10.0, kete_core/Cargo.toml:13;
 * the crate is not in the kete source tree): alpha 1, gamma 2, rho 0.5,
 * sigma 0.5; vertices kept sorted by cost (stable); per iteration the
 * reflection of the worst vertex through the centroid of the others,
 * accepted when it lies in [best, second-worst); an expansion when it beats
 * the best (the better of the two kept); otherwise an inside contraction
 * toward the worst, kept only if it beats the worst (argmin reaches its
 * shrink step only for NaN costs, which here are an error).  Before every
 * iteration the executor stops when the sample standard deviation of the
 * vertex costs is below f64::EPSILON, then when 1000 iterations are done.
 * The result is the best vertex's cost. */
typedef struct { double p[2]; double c; } nm_vertex;

static void nm_sort(nm_vertex v[3]) {
    for (int i = 1; i < 3; ++i) {
        nm_vertex x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j].c > x.c) { v[j + 1] = v[j]; --j; }
        v[j + 1] = x;
    }
}

static int moid_nelder_mead(const ko_state *a, const ko_state *b, double *best, int32_t *iters) {
    nm_vertex v[3] = {{{-15.0, -15.0}, 0.0}, {{15.0, -15.0}, 0.0}, {{0.0, 15.0}, 0.0}};
    for (int k = 0; k < 3; ++k) {
        int rc = moid_cost(a, b, v[k].p, &v[k].c);
        if (rc) return rc;
        if (isnan(v[k].c)) return KO_CONVERGENCE;
    }
    nm_sort(v);
    int it = 0;
    for (;;) {
        double c0 = ((0.0 + v[0].c) + v[1].c) + v[2].c;
        c0 = c0 / 3.0;
        double ss = 0.0;
        for (int k = 0; k < 3; ++k) ss += (v[k].c - c0) * (v[k].c - c0);
        double s = sqrt(1.0 / (3.0 - 1.0) * ss);
        if (s < 2.220446049250313e-16) break;
        if (it >= 1000) break;
        double x0[2], xr[2], cr;
        for (int d = 0; d < 2; ++d) x0[d] = (v[0].p[d] + v[1].p[d]) * (1.0 / 2.0);
        for (int d = 0; d < 2; ++d) xr[d] = x0[d] + (x0[d] - v[2].p[d]) * 1.0;
        int rc = moid_cost(a, b, xr, &cr);
        if (rc) return rc;
        if (isnan(cr)) return KO_CONVERGENCE;
        if (cr < v[1].c && cr >= v[0].c) {
            v[2].p[0] = xr[0]; v[2].p[1] = xr[1]; v[2].c = cr;
        } else if (cr < v[0].c) {
            double xe[2], ce;
            for (int d = 0; d < 2; ++d) xe[d] = x0[d] + (xr[d] - x0[d]) * 2.0;
            rc = moid_cost(a, b, xe, &ce);
            if (rc) return rc;
            if (isnan(ce)) return KO_CONVERGENCE;
            if (ce < cr) { v[2].p[0] = xe[0]; v[2].p[1] = xe[1]; v[2].c = ce; }
            else { v[2].p[0] = xr[0]; v[2].p[1] = xr[1]; v[2].c = cr; }
        } else {
            double xc[2], cc;
            for (int d = 0; d < 2; ++d) xc[d] = x0[d] + (v[2].p[d] - x0[d]) * 0.5;
            rc = moid_cost(a, b, xc, &cc);
            if (rc) return rc;
            if (isnan(cc)) return KO_CONVERGENCE;
            if (cc < v[2].c) { v[2].p[0] = xc[0]; v[2].p[1] = xc[1]; v[2].c = cc; }
        }
        nm_sort(v);
        ++it;
    }
    *best = v[0].c;
    if (iters) *iters = it;
    return KO_OK;
}

#define MOID_STEPS 50

int ko_moid(const ko_state *a_in, const ko_state *b_in, double *moid, double *grid_min,
            int32_t *iters) {
    ko_state a = *a_in, b = *b_in;
    int rc = ko_change_frame(&a, KO_FRAME_ECLIPTIC);
    if (!rc) rc = ko_change_frame(&b, KO_FRAME_ECLIPTIC);
    if (rc) return rc;
    double pt_a, per_a, pt_b, per_b;
    moid_elements(a.jd, a.pos, a.vel, &pt_a, &per_a);
    rc = ko_propagate_two_body(&a, pt_a);
    if (rc) return rc;
    moid_elements(b.jd, b.pos, b.vel, &pt_b, &per_b);
    rc = ko_propagate_two_body(&b, pt_b);
    if (rc) return rc;
    double step_a = isfinite(per_a) ? per_a / MOID_STEPS : 300.0 / MOID_STEPS;
    double step_b = isfinite(per_b) ? per_b / MOID_STEPS : 300.0 / MOID_STEPS;
    static __thread ko_state ga[2 * MOID_STEPS], gb[2 * MOID_STEPS];
    for (int idx = -MOID_STEPS; idx < MOID_STEPS; ++idx) {
        ga[idx + MOID_STEPS] = a;
        rc = ko_propagate_two_body(&ga[idx + MOID_STEPS], a.jd + idx * step_a);
        if (rc) return rc;
        gb[idx + MOID_STEPS] = b;
        rc = ko_propagate_two_body(&gb[idx + MOID_STEPS], b.jd + idx * step_b);
        if (rc) return rc;
    }
    double best = INFINITY;
    ko_state ba = a, bb = b;
    for (int i = 0; i < 2 * MOID_STEPS; ++i)
        for (int j = 0; j < 2 * MOID_STEPS; ++j) {
            double d[3] = {ga[i].pos[0] - gb[j].pos[0], ga[i].pos[1] - gb[j].pos[1],
                           ga[i].pos[2] - gb[j].pos[2]};
            double r = norm3(d);
            if (r < best) { best = r; ba = ga[i]; bb = gb[j]; }
        }
    if (grid_min) *grid_min = best;
    return moid_nelder_mead(&ba, &bb, moid, iters);
}

typedef struct {
    const double *jd_a, *jd_b;
    const double *const *xa, *const *xb;
    int32_t frame;
    double *out, *grid_min;
    int32_t *status, *iters;
} moid_ctx;

static void moid_range(int64_t lo, int64_t hi, void *vctx) {
    moid_ctx *c = vctx;
    for (int64_t i = lo; i < hi; ++i) {
        ko_state a = {c->jd_a[i], {c->xa[0][i], c->xa[1][i], c->xa[2][i]},
                      {c->xa[3][i], c->xa[4][i], c->xa[5][i]}, c->frame, 10};
        ko_state b = {c->jd_b ? c->jd_b[i] : c->jd_a[i], {0, 0, 0}, {0, 0, 0}, c->frame, 10};
        int rc = KO_OK;
        if (c->xb) {
            for (int k = 0; k < 3; ++k) { b.pos[k] = c->xb[k][i]; b.vel[k] = c->xb[3 + k][i]; }
        } else {
            /* moid_py's default: the Earth at state_a's epoch, Sun-centred */
            rc = ko_spk_get_state(399, a.jd, 10, c->frame, b.pos, b.vel);
        }
        double m = NAN, g = NAN;
        int32_t it = 0;
        if (!rc) rc = ko_moid(&a, &b, &m, &g, &it);
        c->status[i] = rc;
        c->out[i] = rc ? NAN : m;
        if (c->grid_min) c->grid_min[i] = g;
        if (c->iters) c->iters[i] = it;
    }
}

void ko_moid_batch(int64_t n, const double *jd_a, const double *const *xyz_a, int32_t frame,
                   const double *jd_b, const double *const *xyz_b, double *out,
                   double *grid_min, int32_t *status, int32_t *iters, int n_threads) {
    moid_ctx c = {jd_a, jd_b, xyz_a, xyz_b, frame, out, grid_min, status, iters};
    parallel_for(n, 16, n_threads, moid_range, &c);
}

/* ------------------------------------------------------------------------ */
/* RK4(5) Fehlberg and kete's two-body STM                                   */
/* (propagation/runge_kutta.rs:33-147, state_transition.rs:6-112,            */
/*  acceleration.rs:254-290)                                                 */
/* ------------------------------------------------------------------------ */
#define RK_MAX_DIM 42

typedef int (*rk_fn)(double t, const double *y, double *dy, int dim, void *meta);

typedef struct {
    rk_fn f;
    void *meta;
    int dim;
    double t, tf, tol;
    double y[RK_MAX_DIM], der[RK_MAX_DIM];
    int64_t n_eval;
} rk45;

static int rk_eval(rk45 *I, double t, const double *y, double *dy) {
    I->n_eval++;
    return I->f(t, y, dy, I->dim, I->meta);
}

/* RK45Integrator::step (runge_kutta.rs:33-117): the Fehlberg stages with
 * k1 = h * der; a step whose error norm exceeds tol is retried at the
 * shrunk size (kete recurses; a loop here); a zero error doubles the next
 * step. */
static int rk_step(rk45 *I, double h, double *next) {
    const int D = I->dim;
    double k2[RK_MAX_DIM], k3[RK_MAX_DIM], k4[RK_MAX_DIM], k5[RK_MAX_DIM], k6[RK_MAX_DIM];
    double tmp[RK_MAX_DIM];
    for (;;) {
        if (h < 1e-30) return KO_CONVERGENCE;   /* step_size < 1e-30 */
        const double *y = I->y, *d = I->der;
        for (int i = 0; i < D; ++i) tmp[i] = y[i] + d[i] * (h / 4.0);
        if (rk_eval(I, I->t + h / 4.0, tmp, k2)) return KO_CONVERGENCE;
        for (int i = 0; i < D; ++i) k2[i] *= h;
        for (int i = 0; i < D; ++i) tmp[i] = y[i] + d[i] * (3.0 * h / 32.0) + k2[i] * (9.0 / 32.0);
        if (rk_eval(I, I->t + h * (3.0 / 8.0), tmp, k3)) return KO_CONVERGENCE;
        for (int i = 0; i < D; ++i) k3[i] *= h;
        for (int i = 0; i < D; ++i)
            tmp[i] = y[i] + d[i] * (h * 1932.0 / 2197.0) - k2[i] * (7200.0 / 2197.0) +
                     k3[i] * (7296.0 / 2197.0);
        if (rk_eval(I, I->t + h * (12.0 / 13.0), tmp, k4)) return KO_CONVERGENCE;
        for (int i = 0; i < D; ++i) k4[i] *= h;
        for (int i = 0; i < D; ++i)
            tmp[i] = y[i] + d[i] * (h * 439.0 / 216.0) - k2[i] * 8.0 + k3[i] * (3680.0 / 513.0) -
                     k4[i] * (845.0 / 4104.0);
        if (rk_eval(I, I->t + h, tmp, k5)) return KO_CONVERGENCE;
        for (int i = 0; i < D; ++i) k5[i] *= h;
        for (int i = 0; i < D; ++i)
            tmp[i] = y[i] - d[i] * (h * 8.0 / 27.0) + k2[i] * 2.0 - k3[i] * (3544.0 / 2565.0) +
                     k4[i] * (1859.0 / 4104.0) - k5[i] * (11.0 / 40.0);
        if (rk_eval(I, I->t + h * 0.5, tmp, k6)) return KO_CONVERGENCE;
        for (int i = 0; i < D; ++i) k6[i] *= h;
        double err2 = 0.0;
        for (int i = 0; i < D; ++i) {
            const double e = d[i] * (h / -360.) + k3[i] * (128.0 / 4275.0) +
                             k4[i] * (2197.0 / 75240.0) - k5[i] / 50.0 - k6[i] * (2.0 / 55.0);
            err2 += e * e;
        }
        const double trunc_err = sqrt(err2);
        double new_step;
        if (trunc_err == 0.0) {
            new_step = 2.0 * h;
        } else {
            new_step = 0.9 * h * pow(I->tol / trunc_err, 1.0 / 5.0);
            if (trunc_err > I->tol) {
                h = new_step;
                continue;
            }
        }
        for (int i = 0; i < D; ++i)
            I->y[i] += d[i] * (h * 16.0 / 135.0) + k3[i] * (6656.0 / 12825.0) +
                       k4[i] * (28561.0 / 56430.0) - k5[i] * (9.0 / 50.0) + k6[i] * (2.0 / 55.0);
        I->t += h;
        if (rk_eval(I, I->t, I->y, I->der)) return KO_CONVERGENCE;
        *next = new_step;
        return KO_OK;
    }
}

/* RK45Integrator::integrate (runge_kutta.rs:119-147): the first step tries
 * the whole interval; later steps are clipped to land on the final time;
 * the loop ends when |t - tf| <= tol (the same tolerance as the error). */
static int rk_integrate(rk_fn f, void *meta, int dim, const double *y0, double t0, double tf,
                        double tol, double *y_out, int64_t *n_eval) {
    rk45 I;
    I.f = f;
    I.meta = meta;
    I.dim = dim;
    I.t = t0;
    I.tf = tf;
    I.tol = tol;
    I.n_eval = 0;
    memcpy(I.y, y0, dim * sizeof(double));
    if (fabs(t0 - tf) < 1e-30) {
        memcpy(y_out, y0, dim * sizeof(double));
        if (n_eval) *n_eval = 0;
        return KO_OK;
    }
    if (rk_eval(&I, t0, I.y, I.der)) return KO_CONVERGENCE;
    double next = tf - I.t;
    int st = rk_step(&I, next, &next);
    /* kete loops without bound; the cap (far beyond any arc here) turns a
     * time that cannot reach tf within tol into an error instead of a hang */
    int64_t steps = 0;
    while (st == KO_OK && fabs(I.t - I.tf) > tol) {
        if (++steps > 10000000) return KO_CONVERGENCE;
        if (fabs(I.t - I.tf) < fabs(next)) next = I.tf - I.t;
        st = rk_step(&I, next, &next);
    }
    memcpy(y_out, I.y, dim * sizeof(double));
    if (n_eval) *n_eval = I.n_eval;
    return st;
}

/* Test equations: kete's RK45 tests (runge_kutta.rs:158-197), kind 0
 * y' = -y and kind 1 y' = 2; kind 2 y' = 5 t^4, whose solution t^5 + c the
 * fifth-order Fehlberg weights integrate exactly (a wrong weight or node
 * shows). */
static int rk_test_fn(double t, const double *y, double *dy, int dim, void *meta) {
    (void)dim;
    const int kind = *(const int *)meta;
    dy[0] = kind == 0 ? -y[0] : (kind == 1 ? 2.0 : 5.0 * t * t * t * t);
    return 0;
}

int ko_rk45_test(int kind, double y0, double t0, double t1, double tol, double *y1,
                 int64_t *n_eval) {
    return rk_integrate(rk_test_fn, &kind, 1, &y0, t0, t1, tol, y1, n_eval);
}

/* stm_ivp_eqn (state_transition.rs:6-47) in kete's scaled time tau = k t
 * (k = GMS_SQRT): y = (pos, vel / k, STM column-major).  fix bits switch
 * off the two defects of kete's arithmetic (for the cross-check only):
 *   1: accel_grad (acceleration.rs:267-290) returns
 *      (r^2 I - 3 p p^T) / (2 r^5) * m, i.e. -1/2 the Jacobian of -m p / r^3;
 *      with the bit, m (3 p p^T - r^2 I) / r^5;
 *   2: the trajectory's acceleration is central_accel (acceleration.rs:
 *      66-80), -p GMS / r^3, in day units although the equation runs in
 *      tau (and mass_scaling is ignored); with the bit, -m p / r^3. */
typedef struct {
    double mass;
    int fix;
} stm_meta;

static int stm_fn(double t, const double *y, double *dy, int dim, void *meta) {
    (void)t;
    (void)dim;
    const stm_meta *M = (const stm_meta *)meta;
    const double p[3] = {y[0], y[1], y[2]};
    const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double ac = (M->fix & 2) ? M->mass : GMS;
    for (int i = 0; i < 3; ++i) {
        dy[i] = y[3 + i];
        dy[3 + i] = -p[i] * pow(r, -3) * ac;
    }
    /* accel_grad: kete's form, or the true Jacobian */
    double G[3][3];
    const double r2 = r * r, r5 = pow(r, 5);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double e = (i == j ? r2 : 0.0) - 3.0 * p[i] * p[j];
            G[i][j] = (M->fix & 1) ? -e / r5 * M->mass : e / (2.0 * r5) * M->mass;
        }
    /* d STM / d tau = A STM, A = [[0, I], [G, 0]]; STM column-major */
    const double *S = y + 6;
    double *dS = dy + 6;
    for (int c = 0; c < 6; ++c) {
        const double *col = S + 6 * c;
        double *dcol = dS + 6 * c;
        for (int i = 0; i < 3; ++i) dcol[i] = col[3 + i];
        for (int i = 0; i < 3; ++i) dcol[3 + i] = G[i][0] * col[0] + G[i][1] * col[1] + G[i][2] * col[2];
    }
    return 0;
}

/* compute_state_transition (state_transition.rs:52-112): RK45 at tol 1e-12
 * over tau from jd0 k to jd1 k.  Returns the final (pos, vel) and the 6x6
 * matrix kete returns, row-major: diag(1,1,1,1/k,1/k,1/k) S^T
 * diag(1,1,1,k,k,k) -- the transpose of d x(jd1) / d x(jd0) (kete
 * transposes the integrated matrix before scaling). */
int ko_compute_state_transition(const double pos[3], const double vel[3], double jd0,
                                double jd1, double central_mass, int fix, double out_pos[3],
                                double out_vel[3], double stm[36], int64_t *n_eval) {
    double y0[42] = {0}, y1[42];
    for (int i = 0; i < 3; ++i) {
        y0[i] = pos[i];
        y0[3 + i] = vel[i] / GMS_SQRT;
    }
    for (int i = 0; i < 6; ++i) y0[6 + 7 * i] = 1.0;
    stm_meta M = {central_mass, fix};
    const int st = rk_integrate(stm_fn, &M, 42, y0, jd0 * GMS_SQRT, jd1 * GMS_SQRT, 1e-12, y1,
                                n_eval);
    if (st) return st;
    for (int i = 0; i < 3; ++i) {
        out_pos[i] = y1[i];
        out_vel[i] = y1[3 + i] * GMS_SQRT;
    }
    const double sa[6] = {1, 1, 1, 1.0 / GMS_SQRT, 1.0 / GMS_SQRT, 1.0 / GMS_SQRT};
    const double sb[6] = {1, 1, 1, GMS_SQRT, GMS_SQRT, GMS_SQRT};
    /* S column-major: S[row][col] = y1[6 + 6 col + row]; S^T[i][j] = S[j][i] */
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) stm[6 * i + j] = sa[i] * y1[6 + 6 * i + j] * sb[j];
    return KO_OK;
}
