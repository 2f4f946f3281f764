// kh_device.h -- device-side building blocks shared by the kete_hip kernels:
// constants, the SPK Type 2 table, Chebyshev evaluation, the universal Kepler
// solver and the spk_accel force model.  FP64 throughout.
//
// Each routine restates kete 1.1.3 (Caltech-IPAC/kete; paths relative to
// src/kete_core/src) and cites it.  The arithmetic keeps kete's operation
// order; the compiler may contract a*b+c into FMA, which changes results at
// the last-ulp level only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kh {

// constants/gravity.rs:9-40, constants/universal.rs:11-23
constexpr double GMS = 0.00029591220828411956;
constexpr double GMS_SQRT = 0.01720209894996;
constexpr double SUN_J2 = 2.2e-7;
constexpr double EARTH_J2 = 0.00108262998905;
constexpr double JUPITER_J2 = 0.014696572;
constexpr double C_AU_PER_DAY = 173.14463267424034;
constexpr double C_AU_PER_DAY_INV = 0.005775518331436995;
constexpr double C_AU_PER_DAY_INV_SQUARED = 3.33566119967647e-05;
constexpr double AU_KM = 149597870.7;
constexpr double INV_AU_KM = 1.0 / AU_KM;
constexpr double TAU = 6.283185307179586;
constexpr double PARABOLIC_BETA = 1e-10;  // kepler.rs:120

// Per-object status (errors.rs:14-32)
enum : int32_t {
  ST_OK = 0,
  ST_CONVERGENCE = 1,
  ST_VALUE_ERROR = 2,
  ST_DAF_LIMITS = 3,
  ST_UNKNOWN_FRAME = 4,
  ST_IO_ERROR = 5,
  ST_IMPACT = 6,
};

constexpr int MAX_SEGS = 24;
constexpr int MAX_CHAIN = 4;
constexpr int MAX_COEF = 32;

struct SegDev {
  double jd_start, jd_end;  // JD TDB, as kete converts the summary floats
  double j0s;               // jd_to_spice_jd(jd_start)
  double intlen;            // seconds per record
  int64_t off;              // first double of record 0 in the coefficient array
  int32_t n_rec, rsize, n_coef, pad;
};

// A chain of segments whose states sum to a body's (or center's) SSB state.
struct Chain {
  int32_t n;
  int32_t seg[MAX_CHAIN];
};

// Bodies of kete's PLANETS list (gravity.rs:145-148) in its order.
constexpr int N_PLANETS = 10;

// One body of kete's PLANETS list (gravity.rs:43-101,145-148).
// kind: 0 Newtonian only, 1 Newtonian + GR + ecliptic-frame J2 (Sun,
// Jupiter), 2 Newtonian + equatorial J2 (Earth).
struct BodyDev {
  int32_t seg;  // segment of the body w.r.t. its center
  int32_t emb;  // 1: center is the Earth-Moon barycentre segment
  int32_t id;   // NAIF id
  int32_t kind;
  double mass, radius, j2, pad;
};

// SPK Type 21 segments (spk_segments.rs:958-1102).  At upload each record's
// modified-difference line is rewritten as three polynomials in the
// normalised time tau = (t - t_ref) / S, S the largest |t - t_ref| over the
// times the record serves: position = ref + dt (ref_vel + dt P(tau)),
// velocity its derivative, P of degree < T21_NP.  Device record:
// [t_ref, 1/S, x vx y vy z vz, P_x[T21_NP], P_y[T21_NP], P_z[T21_NP]] (km, s).
constexpr int T21_NP = 16;
constexpr int T21_REC = 8 + 3 * T21_NP;
constexpr int MAX_T21_SEGS = 32;
// kete's MASSES after PLANETS: Ceres, Pallas, Vesta, Hygiea, Interamnia
// (gravity.rs:103-134).
constexpr int N_ASTEROIDS = 5;

struct T21SegDev {
  double jd_start, jd_end;  // JD TDB, as kete converts the summary floats
  double t0, inv_bw;        // bucket grid over the final epochs: origin (s), 1/width
  int64_t rec_off;          // first record (in records) in t21_rec
  int64_t ep_off;           // first final epoch in t21_ep
  int64_t tab_off;          // first bucket in t21_tab
  int32_t n_rec, n_bucket;
  int32_t ecliptic;     // frame 17: rotated to equatorial after evaluation
  int32_t sun_centred;  // center 10: the Sun's barycentric state is added
};

// A bucket of the final-epoch grid: rec is a record at or before the first
// record serving any time in the bucket, ep0/ep1 the final epochs of rec and
// rec + 1 (+inf past the last record), so that the lookup usually ends
// without touching the epoch array.
struct T21Bucket {
  double ep0, ep1;
  int32_t rec, pad0;
  int64_t pad1;
};

// A MASSES asteroid: its Type 21 segments t21[seg_begin, seg_end) in load
// order (kete scans an object's segments in load order, spk.rs:72-78) and
// its Newtonian parameters (kind 0).
struct AstDev {
  int32_t seg_begin, seg_end, pad0, pad1;
  BodyDev bd;
};

// Device-resident ephemeris description; kernels take a pointer to it, so the
// per-body and per-segment fields are read with scalar loads.
struct SpkDev {
  const double* __restrict__ coef;
  int32_t n_seg;
  int32_t all14;  // every segment has 14 coefficients
  int32_t shared_grid;  // all segments on one record grid (see accel_shared_grid)
  int32_t emb_seg;
  int32_t sun_seg;
  SegDev seg[MAX_SEGS];
  BodyDev body[N_PLANETS];
  // Type 21 segments of any object, and the MASSES asteroids among them
  const double* __restrict__ t21_rec;   // [records][T21_REC]
  const double* __restrict__ t21_ep;    // final epochs, s past J2000
  const T21Bucket* __restrict__ t21_tab;  // final-epoch buckets
  int32_t n_t21_seg, n_t21_rec;
  T21SegDev t21[MAX_T21_SEGS];
  AstDev ast[N_ASTEROIDS];
};

struct Vec3 {
  double x, y, z;
};
__device__ __forceinline__ Vec3 v3(double a, double b, double c) { return Vec3{a, b, c}; }
__device__ __forceinline__ double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double norm(Vec3 a) { return sqrt(dot(a, a)); }

// Ecliptic <-> equatorial rotation about x by the obliquity
// (frames/definitions.rs:135,155-158,224-237).
constexpr double ROT_C = 0.9174820620691818;  // cos(0.40909280422232897)
constexpr double ROT_S = 0.3977771559319137;  // sin(0.40909280422232897)
__device__ __forceinline__ Vec3 ecl_to_equ(Vec3 v) {
  return Vec3{v.x, ROT_C * v.y + (-ROT_S) * v.z, ROT_S * v.y + ROT_C * v.z};
}
__device__ __forceinline__ Vec3 equ_to_ecl(Vec3 v) {
  return Vec3{v.x, ROT_C * v.y + ROT_S * v.z, (-ROT_S) * v.y + ROT_C * v.z};
}

// ---------------------------------------------------------------------------
// SPK Type 2 (spk_segments.rs:150-177,355-376; interpolation.rs:25-77)
// ---------------------------------------------------------------------------
// Position-only Chebyshev sum, forward T recurrence in ascending order.
template <int N>
__device__ __forceinline__ Vec3 cheb_pos_n(const double* __restrict__ c, double x) {
  double x2 = 2.0 * x;
  double vx = c[0] + c[1] * x, vy = c[N] + c[N + 1] * x, vz = c[2 * N] + c[2 * N + 1] * x;
  double t2 = 1.0, t1 = x;
#pragma unroll
  for (int k = 2; k < N; ++k) {
    double t = x2 * t1 - t2;
    vx += c[k] * t;
    vy += c[N + k] * t;
    vz += c[2 * N + k] * t;
    t2 = t1;
    t1 = t;
  }
  return Vec3{vx, vy, vz};
}

__device__ __forceinline__ Vec3 cheb_pos_dyn(const double* __restrict__ c, int n, double x) {
  double x2 = 2.0 * x;
  double vx = c[0] + c[1] * x, vy = c[n] + c[n + 1] * x, vz = c[2 * n] + c[2 * n + 1] * x;
  double t2 = 1.0, t1 = x;
  for (int k = 2; k < n; ++k) {
    double t = x2 * t1 - t2;
    vx += c[k] * t;
    vy += c[n + k] * t;
    vz += c[2 * n + k] * t;
    t2 = t1;
    t1 = t;
  }
  return Vec3{vx, vy, vz};
}

// Position and derivative (interpolation.rs:25-77): der += c_k * U_{k-1} * k.
__device__ __forceinline__ void cheb_both_dyn(const double* __restrict__ c, int n, double x,
                                              Vec3& p, Vec3& d) {
  double x2 = 2.0 * x;
  double vx = c[0] + c[1] * x, vy = c[n] + c[n + 1] * x, vz = c[2 * n] + c[2 * n + 1] * x;
  double dx = c[1], dy = c[n + 1], dz = c[2 * n + 1];
  double t2 = 1.0, t1 = x, u2 = 1.0, u1 = x2;
  for (int k = 2; k < n; ++k) {
    double t = x2 * t1 - t2;
    vx += c[k] * t;
    vy += c[n + k] * t;
    vz += c[2 * n + k] * t;
    t2 = t1;
    t1 = t;
    double u = x2 * u1 - u2;
    double kk = (double)k;
    dx += c[k] * u1 * kk;
    dy += c[n + k] * u1 * kk;
    dz += c[2 * n + k] * u1 * kk;
    u2 = u1;
    u1 = u;
  }
  p = Vec3{vx, vy, vz};
  d = Vec3{dx, dy, dz};
}

template <int N>
__device__ __forceinline__ void cheb_both_n(const double* __restrict__ c, double x, Vec3& p,
                                            Vec3& d) {
  double x2 = 2.0 * x;
  double vx = c[0] + c[1] * x, vy = c[N] + c[N + 1] * x, vz = c[2 * N] + c[2 * N + 1] * x;
  double dx = c[1], dy = c[N + 1], dz = c[2 * N + 1];
  double t2 = 1.0, t1 = x, u2 = 1.0, u1 = x2;
#pragma unroll
  for (int k = 2; k < N; ++k) {
    double t = x2 * t1 - t2;
    vx += c[k] * t;
    vy += c[N + k] * t;
    vz += c[2 * N + k] * t;
    t2 = t1;
    t1 = t;
    double u = x2 * u1 - u2;
    const double kk = (double)k;
    dx += c[k] * u1 * kk;
    dy += c[N + k] * u1 * kk;
    dz += c[2 * N + k] * u1 * kk;
    u2 = u1;
    u1 = u;
  }
  p = Vec3{vx, vy, vz};
  d = Vec3{dx, dy, dz};
}

// Record index of jd in segment g (spk_segments.rs:355-358), clamped to the
// records present: kete reads one record past the end at jd == jd_end.
__device__ __forceinline__ int32_t record_index(const SegDev& g, double js) {
  double fi = floor((js - g.j0s) / g.intlen);
  return fi > 0.0 ? (int32_t)fmin(fi, (double)(g.n_rec - 1)) : 0;
}

// A window of W consecutive records of every segment, staged in LDS by the
// N-body kernel (all segments must then have 14 coefficients, i.e. 44-double
// records): kh_win_rec is [n_seg][W][44] with word 1 (the record radius)
// stored as its reciprocal, kh_win_w0 the first staged record of each
// segment.  Declared at file scope so that reads compile to ds_read.
extern __shared__ double kh_win_rec[];
__shared__ int32_t kh_win_w0[MAX_SEGS];

// Chebyshev lookups read a 44-double record.  When every active lane of the
// wavefront needs the same record -- the common case, since a wavefront's
// objects advance through time together -- the record address is uniform and
// the coefficients are fetched once per wavefront with scalar loads into
// SGPRs; otherwise each lane loads its own record.
template <bool VEL>
__device__ __forceinline__ int32_t seg_eval(const SpkDev* __restrict__ spk,
                                            const double* __restrict__ coef, int s, double jd,
                                            Vec3& pos, Vec3& vel, int32_t win = 0) {
  const SegDev& g = spk->seg[s];
  const bool in = !(jd < g.jd_start || jd > g.jd_end);
  const double js = (jd - 2451545.0) * 86400.0;
  const int32_t ri = record_index(g, js);
  const bool all14 = spk->all14 != 0;
  Vec3 p, d{0.0, 0.0, 0.0};
  double inv_radius;
  int32_t loc = -1;
  if (win > 0) loc = ri - kh_win_w0[s];
  const int32_t r0 = __builtin_amdgcn_readfirstlane(ri);
  if (win > 0 && __all(loc >= 0 && loc < win)) {
    // every lane's record is in the LDS window
    const double* rec = kh_win_rec + (s * win + loc) * 44;
    inv_radius = rec[1];  // staged as 1 / radius
    const double x = (js - rec[0]) * inv_radius;
    if (VEL) cheb_both_n<14>(rec + 2, x, p, d);
    else p = cheb_pos_n<14>(rec + 2, x);
  } else if (__all(ri == r0) && all14) {
    const double* __restrict__ rec = coef + g.off + (int64_t)r0 * g.rsize;
    inv_radius = 1.0 / rec[1];
    const double x = (js - rec[0]) * inv_radius;
    if (VEL) cheb_both_n<14>(rec + 2, x, p, d);
    else p = cheb_pos_n<14>(rec + 2, x);
  } else if (all14) {
    // per-lane records, fixed length: unrolled, so all loads issue up front
    const double* __restrict__ rec = coef + g.off + (int64_t)ri * 44;
    inv_radius = 1.0 / rec[1];
    const double x = (js - rec[0]) * inv_radius;
    if (VEL) cheb_both_n<14>(rec + 2, x, p, d);
    else p = cheb_pos_n<14>(rec + 2, x);
  } else {
    const double* __restrict__ rec = coef + g.off + (int64_t)ri * g.rsize;
    inv_radius = 1.0 / rec[1];
    const double x = (js - rec[0]) * inv_radius;
    if (VEL) cheb_both_dyn(rec + 2, g.n_coef, x, p, d);
    else p = cheb_pos_dyn(rec + 2, g.n_coef, x);
  }
  if (!in) return ST_DAF_LIMITS;
  // kete divides by AU_KM and by the record radius (spk_segments.rs:366-375);
  // multiplying by reciprocals changes the result by at most an ulp
  pos = Vec3{p.x * INV_AU_KM, p.y * INV_AU_KM, p.z * INV_AU_KM};
  if (VEL) {
    const double sc = 86400.0 * inv_radius * INV_AU_KM;
    vel = Vec3{d.x * sc, d.y * sc, d.z * sc};
  }
  return ST_OK;
}

// Segment position in AU (spk_segments.rs:355-376).
__device__ __forceinline__ int32_t seg_pos(const SpkDev* __restrict__ spk,
                                           const double* __restrict__ coef, int s, double jd,
                                           Vec3& pos, int32_t win = 0) {
  Vec3 unused;
  return seg_eval<false>(spk, coef, s, jd, pos, unused, win);
}

// Segment position (AU) and velocity (AU/day).
__device__ __forceinline__ int32_t seg_pos_vel(const SpkDev* __restrict__ spk,
                                               const double* __restrict__ coef, int s, double jd,
                                               Vec3& pos, Vec3& vel, int32_t win = 0) {
  return seg_eval<true>(spk, coef, s, jd, pos, vel, win);
}

// Where a Type 21 object's state at JD t (js = seconds past J2000) comes
// from: the first of its segments t21[begin, end) containing t (kete scans
// an object's segments in load order, spk.rs:72-78), and in it the first
// record whose final epoch is >= js (slice::binary_search_by,
// spk_segments.rs:1023-1030), clamped to the last record (kete reads past
// the array there).  The segment fields are selected from scalar loads; the
// bucket of js gives a record at or before the answer and the final epochs of
// it and its successor, so one load usually settles the record.
struct T21Ref {
  int32_t rec;    // global record index (valid when found)
  int32_t found;  // 0: no segment contains t (DAFLimits)
  int32_t sun_centred, ecliptic;
};

__device__ __forceinline__ T21Ref t21_find(const SpkDev* __restrict__ spk, int32_t begin,
                                           int32_t end, double t, double js) {
  T21Ref r{0, 0, 0, 0};
  double t0 = 0.0, inv_bw = 0.0;
  int64_t tab_off = 0, ep_off = 0, rec_off = 0;
  int32_t n_rec = 1, n_bucket = 1;
  for (int32_t q = begin; q < end; ++q) {
    const T21SegDev& g = spk->t21[q];
    if (!r.found && !(t < g.jd_start || t > g.jd_end)) {
      r.found = 1;
      t0 = g.t0;
      inv_bw = g.inv_bw;
      tab_off = g.tab_off;
      ep_off = g.ep_off;
      rec_off = g.rec_off;
      n_rec = g.n_rec;
      n_bucket = g.n_bucket;
      r.sun_centred = g.sun_centred;
      r.ecliptic = g.ecliptic;
    }
  }
  if (!r.found) return r;
  const double u = (js - t0) * inv_bw;
  const int32_t c = u > 0.0 ? (int32_t)fmin(u, (double)(n_bucket - 1)) : 0;
  const T21Bucket b = spk->t21_tab[tab_off + c];
  int32_t i = b.rec;
  if (b.ep0 < js && i < n_rec - 1) {
    ++i;
    if (b.ep1 < js) {
      const double* __restrict__ ep = spk->t21_ep + ep_off;
      while (i < n_rec - 1 && ep[i] < js) ++i;
    }
  }
  r.rec = (int32_t)rec_off + i;
  return r;
}

// One Type 21 record at js: position in AU and, if VEL, velocity in AU/day
// (spk_segments.rs:1088-1100 divide by AU_KM and scale by 86400 s/day).
template <bool VEL>
__device__ __forceinline__ void t21_eval(const double* __restrict__ r, double js, Vec3& pos,
                                         Vec3& vel) {
  const double dt = js - r[0];
  const double tau = dt * r[1];
  double p[3], d[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* __restrict__ q = r + 8 + c * T21_NP;
    double acc = q[T21_NP - 1];
#pragma unroll
    for (int m = T21_NP - 2; m >= 0; --m) acc = acc * tau + q[m];
    p[c] = acc;
    if (VEL) {
      double der = (double)(T21_NP - 1) * q[T21_NP - 1];
#pragma unroll
      for (int m = T21_NP - 2; m >= 1; --m) der = der * tau + (double)m * q[m];
      d[c] = der;
    }
  }
  pos = Vec3{(r[2] + dt * (p[0] * dt + r[3])) * INV_AU_KM,
             (r[4] + dt * (p[1] * dt + r[5])) * INV_AU_KM,
             (r[6] + dt * (p[2] * dt + r[7])) * INV_AU_KM};
  if (VEL) {
    const double sc = 86400.0 * INV_AU_KM;
    vel = Vec3{(r[3] + dt * (2.0 * p[0] + dt * d[0] * r[1])) * sc,
               (r[5] + dt * (2.0 * p[1] + dt * d[1] * r[1])) * sc,
               (r[7] + dt * (2.0 * p[2] + dt * d[2] * r[1])) * sc};
  }
}

// The equatorial state of a found Type 21 reference at js, relative to its
// segment's center.
template <bool VEL>
__device__ __forceinline__ void t21_state_at(const SpkDev* __restrict__ spk, const T21Ref& ref,
                                             double js, Vec3& pos, Vec3& vel) {
  Vec3 p, v{0.0, 0.0, 0.0};
  t21_eval<VEL>(spk->t21_rec + (int64_t)ref.rec * T21_REC, js, p, v);
  if (ref.ecliptic) {
    p = ecl_to_equ(p);
    if (VEL) v = ecl_to_equ(v);
  }
  pos = p;
  vel = v;
}

// State of the object whose Type 21 segments are t21[begin, end), at JD t.
template <bool VEL>
__device__ __forceinline__ int32_t t21_state(const SpkDev* __restrict__ spk, int32_t begin,
                                             int32_t end, double t, Vec3& pos, Vec3& vel,
                                             bool& sun_centred) {
  const double js = (t - 2451545.0) * 86400.0;
  const T21Ref ref = t21_find(spk, begin, end, t, js);
  if (!ref.found) return ST_DAF_LIMITS;
  t21_state_at<VEL>(spk, ref, js, pos, vel);
  sun_centred = ref.sun_centred != 0;
  return ST_OK;
}

// Sum of a chain of segment states (the center-chaining of spk.rs:99-126).
__device__ __forceinline__ int32_t chain_state(const SpkDev* __restrict__ spk,
                                               const double* __restrict__ coef, const Chain& ch,
                                               double jd, Vec3& pos, Vec3& vel) {
  pos = Vec3{0, 0, 0};
  vel = Vec3{0, 0, 0};
  for (int k = 0; k < ch.n; ++k) {
    Vec3 p, v;
    int32_t st = seg_pos_vel(spk, coef, ch.seg[k], jd, p, v);
    if (st) return st;
    if (k == 0) {
      pos = p;
      vel = v;
    } else {
      pos = Vec3{pos.x + p.x, pos.y + p.y, pos.z + p.z};
      vel = Vec3{vel.x + v.x, vel.y + v.y, vel.z + v.z};
    }
  }
  return ST_OK;
}

// ---------------------------------------------------------------------------
// sin and cos of one argument for the Kepler iteration: Cody-Waite reduction
// by pi/2 (W. J. Cody and W. Waite, Software Manual for the Elementary
// Functions, 1980) with pi/2 split into three doubles, the Taylor series of
// sin (through r^17) and cos (through r^18) on |r| <= pi/4 (the first omitted
// terms are below 1e-19), and the quadrant.  Arguments of 1e5 and beyond
// (the elliptic iteration's b x stays within a few turns) go to the library.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void kep_sincos(double x, double& s, double& c) {
  if (!(fabs(x) < 1.0e5)) {
    sincos(x, &s, &c);
    return;
  }
  const double k = rint(x * 0.6366197723675814);  // 2 / pi
  double r = fma(-k, 1.5707963267948966, x);      // pi/2 = P1 + P2 + P3
  r = fma(-k, 6.123233995736766e-17, r);
  r = fma(-k, -1.4973849048591698e-33, r);
  const double z = r * r;
  // (-1)^n / (2n+1)!, n = 8 .. 1, and (-1)^n / (2n)!, n = 9 .. 1
  double ps = 2.8114572543455206e-15;
  ps = fma(ps, z, -7.647163731819816e-13);
  ps = fma(ps, z, 1.6059043836821613e-10);
  ps = fma(ps, z, -2.505210838544172e-08);
  ps = fma(ps, z, 2.7557319223985893e-06);
  ps = fma(ps, z, -0.0001984126984126984);
  ps = fma(ps, z, 0.008333333333333333);
  ps = fma(ps, z, -0.16666666666666666);
  const double sr = fma(r * z, ps, r);
  double pc = -1.5619206968586225e-16;
  pc = fma(pc, z, 4.779477332387385e-14);
  pc = fma(pc, z, -1.1470745597729725e-11);
  pc = fma(pc, z, 2.08767569878681e-09);
  pc = fma(pc, z, -2.755731922398589e-07);
  pc = fma(pc, z, 2.48015873015873e-05);
  pc = fma(pc, z, -0.001388888888888889);
  pc = fma(pc, z, 0.041666666666666664);
  pc = fma(pc, z, -0.5);
  const double cr = fma(z, pc, 1.0);
  const int q = (int)(long long)k & 3;
  s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
  c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}
__device__ __forceinline__ double kep_sin(double x) {
  double s, c;
  kep_sincos(x, s, c);
  return s;
}

// ---------------------------------------------------------------------------
// Universal Kepler solver (kepler.rs:120-297, fitting/newton.rs:18-65)
// ---------------------------------------------------------------------------
// Elliptic branch.  sin(b x) and 1 - cos(b x) come from one sin/cos pair at
// the half angle: sin(b x) = 2 sh ch and 1 - cos(b x) = 2 sh^2, sh the
// sin(b x / 2) kete's g_2 already takes (the identities change only the
// rounding of the last bits).  The divisions by b and beta of kete's g1, g2,
// g3 are multiplications by reciprocals formed once per solve (a division is
// a ten-instruction sequence on the FP64 pipe; kete's iteration makes five
// per step): the quotients move by at most an ulp, and the iteration still
// stops at |f| < 1e-11.
struct KepEll {
  double r0, rv0, beta, b, dt, inv_b, inv_beta;
  __device__ __forceinline__ void half(double x, double& s, double& omc) const {
    double sh, ch;
    kep_sincos(b * x / 2.0, sh, ch);
    s = 2.0 * sh * ch;
    omc = 2.0 * (sh * sh);
  }
  __device__ __forceinline__ double f(double x) const {
    double s, omc;
    half(x, s, omc);
    double g1 = s * inv_b;
    double g2 = omc * inv_beta;
    double g3 = (x - g1) * inv_beta;
    return r0 * g1 + rv0 * g2 + GMS * g3 - dt;
  }
  __device__ __forceinline__ double d(double x) const {
    double s, omc;
    half(x, s, omc);
    return r0 * (1.0 - omc) + rv0 * s + GMS * (omc * inv_beta);
  }
  // f and d at one point
  __device__ __forceinline__ void fd(double x, double& fv, double& dv) const {
    double s, omc;
    half(x, s, omc);
    double g1 = s * inv_b;
    double g2 = omc * inv_beta;
    double g3 = (x - g1) * inv_beta;
    fv = r0 * g1 + rv0 * g2 + GMS * g3 - dt;
    dv = r0 * (1.0 - omc) + rv0 * s + GMS * g2;
  }
};

struct KepHyp {
  double r0, rv0, beta, b, dt;
  __device__ __forceinline__ double f(double x) const {
    double g1 = sinh(b * x) / b;
    double g2 = (1.0 - cosh(b * x)) / beta;
    double g3 = (x - g1) / beta;
    return r0 * g1 + rv0 * g2 + GMS * g3 - dt;
  }
  __device__ __forceinline__ double d(double x) const {
    double sh = sinh(b * x), ch = cosh(b * x);
    double g3d = (1.0 - ch) / beta;
    return r0 * ch + rv0 * sh + GMS * g3d;
  }
  // f and d at one point, sharing sinh(b x) and cosh(b x)
  __device__ __forceinline__ void fd(double x, double& fv, double& dv) const {
    double sh = sinh(b * x), ch = cosh(b * x);
    double g1 = sh / b;
    double g3 = (x - g1) / beta;
    fv = r0 * g1 + rv0 * ((1.0 - ch) / beta) + GMS * g3 - dt;
    dv = r0 * ch + rv0 * sh + GMS * ((1.0 - ch) / beta);
  }
};

// newton_raphson (fitting/newton.rs:18-65); returns iterations used or -1.
// The function and its derivative at the top of an iteration come from one
// evaluation (fd) that shares the transcendentals; kete evaluates the
// derivative after the convergence test, so an iteration that converges
// computes it without using it.
template <class F>
__device__ __forceinline__ int newton(const F& fn, double start, double atol, double& out) {
  double x = start;
  if (fabs(fn.d(x)) < 1e-12) x += 0.1;
  for (int it = 0; it < 100; ++it) {
    double f, d;
    fn.fd(x, f, d);
    if (fabs(f) < atol) {
      out = x;
      return it + 1;
    }
    if (fabs(d) < 1e-12) return -1;
    if (!isfinite(d) || !isfinite(f)) return -1;
    x -= 0.5 * f / d;
    d = fn.d(x);
    if (fabs(d) < 1e-3) f = fn.f(x);
    x -= 0.5 * f / d;
  }
  return -1;
}

// solve_kepler_universal (kepler.rs:179-233).  Returns newton iterations
// (0 for the closed-form parabolic branch) or -1 on failure.
__device__ __forceinline__ int solve_kepler_universal(double dt, double r0, double v0, double rv0,
                                                      double& s, double& beta) {
  beta = 2.0 * GMS / r0 - v0 * v0;
  double b = sqrt(fabs(beta));
  if (beta >= PARABOLIC_BETA) {
    double period = GMS * (1.0 / (b * b * b));
    dt = fmod(dt, period * TAU);
    KepEll k{r0, rv0, beta, b, dt, 1.0 / b, 1.0 / beta};
    double guess = (1.0 / b) * dt / period;
    return newton(k, guess, 1e-11, s);
  } else if (fabs(beta) < PARABOLIC_BETA) {
    double a0 = -6.0 * dt / GMS;
    double a1 = 6.0 * r0 / GMS;
    double a2 = 3.0 * rv0 / GMS;
    double p = (3.0 * a1 - a2 * a2) / 3.0;
    double q = (9.0 * a1 * a2 - 27.0 * a0 - 2.0 * (a2 * a2 * a2)) / 27.0;
    double w = pow(q / 2.0 + sqrt(q * q / 4.0 + (p * p * p) / 27.0), 1.0 / 3.0);
    s = w - p / (3.0 * w) - a2 / 3.0;
    return 0;
  }
  KepHyp k{r0, rv0, beta, b, dt};
  return newton(k, GMS_SQRT * fabs(beta) * dt, 1e-11, s);
}

__device__ __forceinline__ double g_1(double s, double beta) {
  if (fabs(beta) < PARABOLIC_BETA) return s;
  double bs = sqrt(fabs(beta));
  if (beta >= 0.0) return kep_sin(bs * s) / bs;
  return sinh(bs * s) / bs;
}
__device__ __forceinline__ double g_2(double s, double beta) {
  if (fabs(beta) < PARABOLIC_BETA) return s * s / 2.0;
  double bs = sqrt(fabs(beta)) * s;
  if (beta >= 0.0) {
    double sh = kep_sin(bs / 2.0);
    return 2.0 * (sh * sh) / beta;
  }
  return (1.0 - cosh(bs)) / beta;
}

// One universal-variable solve and f/g update (kepler.rs:265-290).
__device__ __forceinline__ int kepler_once(double time, Vec3 p, Vec3 v, Vec3& po, Vec3& vo,
                                           int& iters) {
  double r0 = norm(p), v0 = norm(v), rv0 = dot(p, v);
  double s, beta;
  int it = solve_kepler_universal(time, r0, v0, rv0, s, beta);
  if (it < 0) return -1;
  iters += it;
  double g1, g2;
  if (beta >= PARABOLIC_BETA) {
    // g_1 and g_2 from one sin/cos pair at half of sqrt(beta) s
    const double bs = sqrt(fabs(beta));
    double sh, ch;
    kep_sincos(bs * s / 2.0, sh, ch);
    g1 = 2.0 * sh * ch / bs;
    g2 = 2.0 * (sh * sh) / beta;
  } else {
    g1 = g_1(s, beta);
    g2 = g_2(s, beta);
  }
  double f = 1.0 - GMS / r0 * g2;
  double g = r0 * g1 + rv0 * g2;
  Vec3 np{p.x * f + v.x * g, p.y * f + v.y * g, p.z * f + v.z * g};
  double nr0 = norm(np);
  double fd = -GMS / (nr0 * r0) * g1;
  double gd = 1.0 - (GMS / nr0) * g2;
  vo = Vec3{p.x * fd + v.x * gd, p.y * fd + v.y * gd, p.z * fd + v.z * gd};
  po = np;
  return 0;
}

// analytic_2_body (kepler.rs:254-297).  kete recurses on failure into two
// half steps, at most 10 levels deep.  Walked iteratively here: `at` counts
// the part of `time` done and `blk` the current step, both in units of
// time / 512 (the step at depth 10).  After a step the next one is the
// largest block aligned at `at` -- the pending sibling or ancestor half --
// which is the order the recursion takes them in; a failure halves the
// block, and a failure at depth 10 is kete's recursion-depth error.
__device__ __forceinline__ int32_t analytic_2_body(double time, Vec3 pos, Vec3 vel, Vec3& po,
                                                   Vec3& vo, int* iters_out = nullptr) {
  int iters = 0;
  Vec3 p = pos, v = vel;
  int at = 0, blk = 512;
  while (at < 512) {
    const double t = time * ((double)blk * (1.0 / 512.0));  // exact: a power of two
    bool ok = true;
    if (!(fabs(t) < 1e-10)) {
      const double rv0 = dot(p, v);
      if (isnan(rv0) || isinf(rv0)) return ST_CONVERGENCE;
      Vec3 np, nv;
      if (kepler_once(t, p, v, np, nv, iters) == 0) {
        p = np;
        v = nv;
      } else {
        ok = false;
      }
    }
    if (ok) {
      at += blk;
      blk = at & -at;
    } else {
      if (blk == 1) return ST_CONVERGENCE;
      blk >>= 1;
    }
  }
  po = p;
  vo = v;
  if (iters_out) *iters_out = iters;
  return ST_OK;
}

// y = x^(1/7) for the step-size ratio (radau.rs:339-342), by two Newton steps
// from a single-precision seed; inputs outside [3e-5, 3e4] only matter through
// the clamp to [0.25, 4] that follows, so they are clamped first.
__device__ __forceinline__ double root7(double x) {
  if (isnan(x)) return x;
  const double xc = fmin(fmax(x, 3e-5), 3e4);
  double y = (double)exp2f(log2f((float)xc) * (1.0f / 7.0f));
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double y2 = y * y, y3 = y2 * y, y6 = y3 * y3, y7 = y6 * y;
    y = y - (y7 - xc) / (7.0 * y6);
  }
  return y;
}

// ---------------------------------------------------------------------------
// Force model: spk_accel with PLANETS (acceleration.rs:121-168),
// GravParams::add_acceleration (gravity.rs:171-253), JplComet non-grav
// (nongrav.rs:114-157).
// ---------------------------------------------------------------------------
// Same layout as kh_nongrav in kete_hip.h.
struct NonGravDev {
  int32_t kind;  // 0 none, 1 JplComet, 2 Dust
  double a1, a2, a3, alpha, r_0, m, n, k, dt, beta;
};

__device__ __forceinline__ double inv_cube(double r) { return 1.0 / (r * r * r); }

// 1/sqrt(x) for x > 0: v_rsq_f64 then two Newton-Raphson steps
// y <- y (3 - x y^2) / 2, each roughly doubling the correct bits.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-hx, y * y, 0.5);
    y = fma(y, e, y);
  }
  return y;
}

// j2_correction (gravity.rs:223-235)
__device__ __forceinline__ Vec3 j2_correction(Vec3 p, double radius, double j2, double mass) {
  double r = norm(p);
  double zr = p.z / r;
  double z_squared = 5.0 * (zr * zr);
  double rr = radius / r;
  double coef = 1.5 * j2 * mass * (rr * rr) * inv_cube(r);
  return Vec3{p.x * coef * (z_squared - 1.0), p.y * coef * (z_squared - 1.0),
              p.z * coef * (z_squared - 3.0)};
}

// apply_gr_correction (gravity.rs:239-253)
__device__ __forceinline__ void gr_correction(Vec3& a, Vec3 p, Vec3 v, double mass) {
  double r_v = 4.0 * dot(p, v);
  double v2 = dot(v, v);
  double r = norm(p);
  double gc = mass * C_AU_PER_DAY_INV_SQUARED * inv_cube(r);
  double c = 4. * mass / r - v2;
  a.x += gc * (c * p.x + r_v * v.x);
  a.y += gc * (c * p.y + r_v * v.y);
  a.z += gc * (c * p.z + r_v * v.z);
}

__device__ __forceinline__ void nongrav_add(const NonGravDev& ng, Vec3& a, Vec3 p, Vec3 v) {
  if (ng.kind == 2) {
    double n = norm(p);
    Vec3 pn{p.x / n, p.y / n, p.z / n};
    double r_dot = dot(pn, v);
    double norm2_inv = 1.0 / dot(p, p);
    double sc = GMS * ng.beta * norm2_inv;
    double f1 = 1.0 - r_dot * C_AU_PER_DAY_INV_SQUARED;
    a.x += sc * (f1 * pn.x - v.x * C_AU_PER_DAY_INV_SQUARED);
    a.y += sc * (f1 * pn.y - v.y * C_AU_PER_DAY_INV_SQUARED);
    a.z += sc * (f1 * pn.z - v.z * C_AU_PER_DAY_INV_SQUARED);
  } else if (ng.kind == 1) {
    double n = norm(p);
    Vec3 pn{p.x / n, p.y / n, p.z / n};
    double vd = dot(v, pn);
    Vec3 t{v.x - pn.x * vd, v.y - pn.y * vd, v.z - pn.z * vd};
    double tn = norm(t);
    t = Vec3{t.x / tn, t.y / tn, t.z / tn};
    Vec3 nv{t.y * pn.z - t.z * pn.y, t.z * pn.x - t.x * pn.z, t.x * pn.y - t.y * pn.x};
    double nn = norm(nv);
    nv = Vec3{nv.x / nn, nv.y / nn, nv.z / nn};
    double rr0 = n / ng.r_0;
    double scale = ng.alpha * pow(rr0, -ng.m) * pow(1.0 + pow(rr0, ng.n), -ng.k);
    double s1 = scale * ng.a1, s2 = scale * ng.a2, s3 = scale * ng.a3;
    a.x += pn.x * s1; a.y += pn.y * s1; a.z += pn.z * s1;
    a.x += t.x * s2;  a.y += t.y * s2;  a.z += t.z * s2;
    a.x += nv.x * s3; a.y += nv.y * s3; a.z += nv.z * s3;
  }
}

// One body's contribution (GravParams::add_acceleration, gravity.rs:171-216,
// plus the non-grav term attached to the Sun, acceleration.rs:161-165) to the
// acceleration of a massless object at barycentric equatorial (x, v), time t.
// emb is the Earth-Moon barycentre position when the body is given relative
// to it.  On exact evaluations, r <= radius is an Impact.
// The body's barycentric equatorial position bp (and velocity bv for kind
// 1) are given.
__device__ __forceinline__ int32_t add_body_at(const BodyDev& bd, Vec3 bp, Vec3 bv, Vec3 x,
                                               Vec3 v, bool exact, const NonGravDev* ng,
                                               Vec3& a) {
  const double mass = bd.mass;
  Vec3 rp{x.x - bp.x, x.y - bp.y, x.z - bp.z};
  // 1/r from the hardware reciprocal square root refined by two Newton
  // steps (full FP64 accuracy), r = r^2/r, 1/r^3 = (1/r)^3; kete computes
  // sqrt and r.powi(-3) -- the results differ by an ulp or two
  const double r2 = dot(rp, rp);
  const double inv_r = rsqrt_nr(r2);
  const double r = r2 * inv_r;
  if (exact && r <= bd.radius) return ST_IMPACT;
  const double inv_r3 = inv_r * inv_r * inv_r;
  const double k = mass * inv_r3;
  a.x -= rp.x * k;
  a.y -= rp.y * k;
  a.z -= rp.z * k;
  if (bd.kind == 1) {
    // apply_gr_correction (gravity.rs:239-253)
    const Vec3 rv{v.x - bv.x, v.y - bv.y, v.z - bv.z};
    const double r_v = 4.0 * dot(rp, rv);
    const double v2 = dot(rv, rv);
    const double gc = mass * C_AU_PER_DAY_INV_SQUARED * inv_r3;
    const double c = 4. * mass * inv_r - v2;
    a.x += gc * (c * rp.x + r_v * rv.x);
    a.y += gc * (c * rp.y + r_v * rv.y);
    a.z += gc * (c * rp.z + r_v * rv.z);
    // j2_correction in the ecliptic frame and back (gravity.rs:184-201,223-235);
    // |rp| is the same in both frames
    const Vec3 pe = equ_to_ecl(rp);
    const double zr = pe.z * inv_r;
    const double z_squared = 5.0 * (zr * zr);
    const double rr = bd.radius * inv_r;
    const double coef2 = 1.5 * bd.j2 * mass * (rr * rr) * inv_r3;
    const Vec3 j = ecl_to_equ(Vec3{pe.x * coef2 * (z_squared - 1.0),
                                   pe.y * coef2 * (z_squared - 1.0),
                                   pe.z * coef2 * (z_squared - 3.0)});
    a.x += j.x;
    a.y += j.y;
    a.z += j.z;
    if (bd.id == 10 && ng != nullptr) nongrav_add(*ng, a, rp, rv);
  } else if (bd.kind == 2) {
    // Earth J2 in the equatorial frame (gravity.rs:203)
    const double zr = rp.z * inv_r;
    const double z_squared = 5.0 * (zr * zr);
    const double rr = bd.radius * inv_r;
    const double coef2 = 1.5 * bd.j2 * mass * (rr * rr) * inv_r3;
    a.x += rp.x * coef2 * (z_squared - 1.0);
    a.y += rp.y * coef2 * (z_squared - 1.0);
    a.z += rp.z * coef2 * (z_squared - 3.0);
  }
  return ST_OK;
}

// The asteroid terms of spk_accel over MASSES (acceleration.rs:145-160 for
// MASSES[10..15], gravity.rs:103-134; propagation/mod.rs:82-88): each
// asteroid's barycentric position -- its Type 21 state plus, for a
// Sun-centred segment, the Sun's (try_get_state to center 0, spk.rs:87-106)
// -- and a Newtonian term, in MASSES order after the planets.  sun is the
// Sun's barycentric position at t.
__device__ __forceinline__ int32_t add_asteroids(const SpkDev* __restrict__ spk, double t,
                                                 Vec3 sun, Vec3 x, Vec3 v, bool exact, Vec3& a,
                                                 int32_t& impact_id) {
  const double js = (t - 2451545.0) * 86400.0;
#pragma unroll 1
  for (int k = 0; k < N_ASTEROIDS; ++k) {
    const AstDev& ad = spk->ast[k];
    const T21Ref ref = t21_find(spk, ad.seg_begin, ad.seg_end, t, js);
    if (!ref.found) return ST_DAF_LIMITS;
    Vec3 p, unused;
    t21_state_at<false>(spk, ref, js, p, unused);
    if (ref.sun_centred) p = Vec3{p.x + sun.x, p.y + sun.y, p.z + sun.z};
    const int32_t st = add_body_at(ad.bd, p, Vec3{0.0, 0.0, 0.0}, x, v, exact, nullptr, a);
    if (st) {
      impact_id = ad.bd.id;
      return st;
    }
  }
  return ST_OK;
}

__device__ __forceinline__ int32_t add_body(const SpkDev* __restrict__ spk,
                                            const double* __restrict__ coef, const BodyDev& bd,
                                            double t, Vec3 x, Vec3 v, bool exact,
                                            const NonGravDev* ng, Vec3 emb, Vec3& a,
                                            int32_t win) {
  Vec3 bp, bv{0.0, 0.0, 0.0};
  int32_t st;
  if (bd.kind == 1) {
    st = seg_pos_vel(spk, coef, bd.seg, t, bp, bv, win);
  } else {
    st = seg_pos(spk, coef, bd.seg, t, bp, win);
    if (bd.emb) bp = Vec3{bp.x + emb.x, bp.y + emb.y, bp.z + emb.z};
  }
  if (st) return st;
  return add_body_at(bd, bp, bv, x, v, exact, ng, a);
}

// Body states when every segment shares one record grid (same start, record
// length, count, and bitwise-identical record mid-points and radii -- DE-style
// files built on one grid, like the synthetic ephemeris) and the wave's
// records are all in the LDS window: the record index and the normalised time
// are computed once per force evaluation instead of once per segment.
// slot is the record's position in the window (the same for every segment).
// UNI: the wavefront's record is uniform and `base` points at its
// coefficients in segment 0's numbering (coefficients then come through
// scalar loads); otherwise they are read from the LDS window `base` at
// `slot`.
template <bool UNI>
__device__ __forceinline__ int32_t accel_shared_grid(const SpkDev* __restrict__ spk,
                                                     const double* __restrict__ base,
                                                     int32_t slot, int32_t win, double t,
                                                     double xs, double inv_radius, Vec3 x, Vec3 v,
                                                     bool exact, const NonGravDev* ng, Vec3& acc,
                                                     int32_t& impact_id, int32_t n_ast) {
  const double vscale = 86400.0 * inv_radius * INV_AU_KM;
  auto coefs = [&](int s) -> const double* {
    if (UNI) return base + spk->seg[s].off;
    return base + (s * win + slot) * 44 + 2;
  };
  Vec3 emb = cheb_pos_n<14>(coefs(spk->emb_seg), xs);
  emb = Vec3{emb.x * INV_AU_KM, emb.y * INV_AU_KM, emb.z * INV_AU_KM};
  Vec3 a{0.0, 0.0, 0.0};
#pragma unroll 1
  for (int b = 0; b < N_PLANETS; ++b) {
    const BodyDev& bd = spk->body[b];
    // Opaque copy of the normalised time: keeps the compiler from hoisting
    // the 26 loop-invariant Chebyshev T/U values out of the body loop, which
    // would hold ~50 VGPRs for the whole force evaluation.
    double xb = xs;
    asm volatile("" : "+v"(xb));
    Vec3 bp, bv{0.0, 0.0, 0.0};
    if (bd.kind == 1) {
      Vec3 p, d;
      cheb_both_n<14>(coefs(bd.seg), xb, p, d);
      bp = Vec3{p.x * INV_AU_KM, p.y * INV_AU_KM, p.z * INV_AU_KM};
      bv = Vec3{d.x * vscale, d.y * vscale, d.z * vscale};
    } else {
      const Vec3 p = cheb_pos_n<14>(coefs(bd.seg), xb);
      bp = Vec3{p.x * INV_AU_KM, p.y * INV_AU_KM, p.z * INV_AU_KM};
      if (bd.emb) bp = Vec3{bp.x + emb.x, bp.y + emb.y, bp.z + emb.z};
    }
    const int32_t st = add_body_at(bd, bp, bv, x, v, exact, ng, a);
    if (st) {
      impact_id = bd.id;
      return st;
    }
  }
  if (n_ast) {
    const Vec3 sp = cheb_pos_n<14>(coefs(spk->sun_seg), xs);
    const Vec3 sun{sp.x * INV_AU_KM, sp.y * INV_AU_KM, sp.z * INV_AU_KM};
    const int32_t st = add_asteroids(spk, t, sun, x, v, exact, a, impact_id);
    if (st) return st;
  }
  acc = a;
  return ST_OK;
}

#ifndef ILP_BODIES
#define ILP_BODIES 2  // bodies whose evaluations overlap in spk_accel_ilp
#endif

// spk_accel on the shared record grid, arranged for the shortest dependency
// chain rather than the fewest registers (the kernel variant that carries the
// slowest objects runs one wavefront per SIMD, so its latency is exposed):
// the Chebyshev T_k and U_k are formed once for all bodies, and the bodies are
// unrolled so their coefficient loads and sums overlap.  The arithmetic and
// the order of the accumulation into the acceleration are those of the
// per-body path, so the results are bitwise the same.  Coefficients are read
// from HBM/L2 per lane (no LDS).
__device__ __forceinline__ int32_t spk_accel_ilp(const SpkDev* __restrict__ spk,
                                                 const double* __restrict__ coef, double t,
                                                 Vec3 x, Vec3 v, bool exact,
                                                 const NonGravDev* ng, Vec3& acc,
                                                 int32_t& impact_id, int32_t n_ast) {
  const SegDev& g = spk->seg[spk->emb_seg];
  if (t < g.jd_start || t > g.jd_end) return ST_DAF_LIMITS;
  const double js = (t - 2451545.0) * 86400.0;
  const int64_t roff = (int64_t)record_index(g, js) * 44;
  const double* __restrict__ r0 = coef + g.off + roff;
  const double ir = 1.0 / r0[1];
  const double xs = (js - r0[0]) * ir;
  const double vscale = 86400.0 * ir * INV_AU_KM;
  // T_k and U_k (interpolation.rs:25-77), the recurrences of cheb_pos_n /
  // cheb_both_n
  const double x2 = 2.0 * xs;
  double T[14], U[14];
  T[0] = 1.0;
  T[1] = xs;
  U[0] = 1.0;
  U[1] = x2;
#pragma unroll
  for (int k = 2; k < 14; ++k) {
    T[k] = x2 * T[k - 1] - T[k - 2];
    U[k] = x2 * U[k - 1] - U[k - 2];
  }
  auto pos = [&](const double* __restrict__ c) {
    double vx = c[0] + c[1] * xs, vy = c[14] + c[15] * xs, vz = c[28] + c[29] * xs;
#pragma unroll
    for (int k = 2; k < 14; ++k) {
      vx += c[k] * T[k];
      vy += c[14 + k] * T[k];
      vz += c[28 + k] * T[k];
    }
    return Vec3{vx * INV_AU_KM, vy * INV_AU_KM, vz * INV_AU_KM};
  };
  auto vel = [&](const double* __restrict__ c) {
    double dx = c[1], dy = c[15], dz = c[29];
#pragma unroll
    for (int k = 2; k < 14; ++k) {
      const double kk = (double)k;
      dx += c[k] * U[k - 1] * kk;
      dy += c[14 + k] * U[k - 1] * kk;
      dz += c[28 + k] * U[k - 1] * kk;
    }
    return Vec3{dx * vscale, dy * vscale, dz * vscale};
  };
  const Vec3 emb = pos(r0 + 2);
  Vec3 a{0.0, 0.0, 0.0};
#pragma unroll ILP_BODIES
  for (int b = 0; b < N_PLANETS; ++b) {
    const BodyDev& bd = spk->body[b];
    const double* __restrict__ c = coef + spk->seg[bd.seg].off + roff + 2;
    Vec3 bp = pos(c), bv{0.0, 0.0, 0.0};
    if (bd.kind == 1) bv = vel(c);
    if (bd.emb) bp = Vec3{bp.x + emb.x, bp.y + emb.y, bp.z + emb.z};
    const int32_t st = add_body_at(bd, bp, bv, x, v, exact, ng, a);
    if (st) {
      impact_id = bd.id;
      return st;
    }
  }
  if (n_ast) {
    const Vec3 sun = pos(coef + spk->seg[spk->sun_seg].off + roff + 2);
    const int32_t st = add_asteroids(spk, t, sun, x, v, exact, a, impact_id);
    if (st) return st;
  }
  acc = a;
  return ST_OK;
}

// spk_accel: acceleration of a massless body at barycentric equatorial
// (x, v) and TDB JD t, summed over kete's PLANETS in its order, then, with
// n_ast != 0, the five asteroids that complete MASSES.  The Earth-Moon
// barycentre is evaluated once and shared by 399 and 301 (kete evaluates it
// for each, with identical results).
// win > 0: records in an LDS window of win records per segment -- the
// block's (kh_win_rec, first records kh_win_w0) or, with wwin, a wavefront's
// own window wwin whose first record is ww0 in every segment (shared grid
// only; lanes outside it take the per-body path without a window).
__device__ __forceinline__ int32_t spk_accel(const SpkDev* __restrict__ spk,
                                             const double* __restrict__ coef, double t, Vec3 x,
                                             Vec3 v, bool exact, const NonGravDev* ng, Vec3& acc,
                                             int32_t& impact_id, int32_t win = 0,
                                             int32_t n_ast = 0, const double* wwin = nullptr,
                                             int32_t ww0 = 0) {
  if (spk->shared_grid) {
    const SegDev& g = spk->seg[spk->emb_seg];
    if (t < g.jd_start || t > g.jd_end) return ST_DAF_LIMITS;
    const double js = (t - 2451545.0) * 86400.0;
    const int32_t ri = record_index(g, js);
    // a wavefront-uniform record: its coefficients through scalar loads
    // (SGPR operands of the sums; no LDS reads, fewer VGPRs): 72.3 -> 69.3
    // ms per launch at 1e6 objects against the LDS window below
    const int32_t ru = __builtin_amdgcn_readfirstlane(ri);
    if (__all(ri == ru)) {
      const double* rec = coef + g.off + (int64_t)ru * 44;
      const double ir = 1.0 / rec[1];
      return accel_shared_grid<true>(spk, coef + (int64_t)ru * 44 + 2, 0, 0, t,
                                     (js - rec[0]) * ir, ir, x, v, exact, ng, acc, impact_id,
                                     n_ast);
    }
    const int32_t loc = win > 0 ? ri - (wwin ? ww0 : kh_win_w0[spk->emb_seg]) : -1;
    if (win > 0 && __all(loc >= 0 && loc < win)) {
      const double* wr = wwin ? wwin : kh_win_rec;
      const double* r0 = wr + (spk->emb_seg * win + loc) * 44;
      const double ir = r0[1];  // staged as 1 / radius
      return accel_shared_grid<false>(spk, wr, loc, win, t, (js - r0[0]) * ir, ir, x, v, exact,
                                      ng, acc, impact_id, n_ast);
    }
  }
  if (wwin) win = 0;
  Vec3 a{0.0, 0.0, 0.0};
  Vec3 emb;
  int32_t st = seg_pos(spk, coef, spk->emb_seg, t, emb, win);
  if (st) return st;
#pragma unroll 1
  for (int b = 0; b < N_PLANETS; ++b) {
    const BodyDev& bd = spk->body[b];
    st = add_body(spk, coef, bd, t, x, v, exact, ng, emb, a, win);
    if (st) {
      if (st == ST_IMPACT) impact_id = bd.id;
      return st;
    }
  }
  if (n_ast) {
    Vec3 sun;
    st = seg_pos(spk, coef, spk->sun_seg, t, sun, win);
    if (st) return st;
    st = add_asteroids(spk, t, sun, x, v, exact, a, impact_id);
    if (st) return st;
  }
  acc = a;
  return ST_OK;
}

// ---------------------------------------------------------------------------
// Sixteen-lane groups (one DPP row): an object's force sum spread over the
// lanes of a row, one body per lane (the wide partials and latency kernels).
// ---------------------------------------------------------------------------
// Chebyshev position and derivative (interpolation.rs:25-77, forward T/U
// recurrences in ascending order, as cheb_both_n) of a 14-coefficient record
// in LDS, [mid, radius, x[14], y[14], z[14]]; the coefficients are read in
// chunks of four orders so that few of them are held in registers at once.
__device__ __forceinline__ void cheb_both_lds(const double* rec, double x, Vec3& p, Vec3& d) {
  const double* c = rec + 2;
  const double x2 = 2.0 * x;
  double vx = c[0] + c[1] * x, vy = c[14] + c[15] * x, vz = c[28] + c[29] * x;
  double dx = c[1], dy = c[15], dz = c[29];
  double t2 = 1.0, t1 = x, u2 = 1.0, u1 = x2;
#pragma unroll
  for (int k0 = 2; k0 < 14; k0 += 4) {
    // (chunk boundary)
#pragma unroll
    for (int k = k0; k < k0 + 4 && k < 14; ++k) {
      const double t = x2 * t1 - t2;
      vx += c[k] * t;
      vy += c[14 + k] * t;
      vz += c[28 + k] * t;
      t2 = t1;
      t1 = t;
      const double u = x2 * u1 - u2;
      const double kk = (double)k;
      dx += c[k] * u1 * kk;
      dy += c[14 + k] * u1 * kk;
      dz += c[28 + k] * u1 * kk;
      u2 = u1;
      u1 = u;
    }
  }
  p = Vec3{vx, vy, vz};
  d = Vec3{dx, dy, dz};
}

template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int lo2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  const int hi2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi2, lo2);
}

// v of lane N of this lane's DPP row (row_newbcast)
template <int N>
__device__ __forceinline__ double row_bcast(double v) {
  return dpp_row<0x150 + N>(v);
}
template <int N>
__device__ __forceinline__ int row_bcast_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x150 + N, 0xF, 0xF, false);
}

// Sum over the 16 lanes of a row, the same bits in every lane (each step
// adds a commuted pair): quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror.
__device__ __forceinline__ double row_sum(double v) {
  v += dpp_row<0xB1>(v);
  v += dpp_row<0x4E>(v);
  v += dpp_row<0x141>(v);
  v += dpp_row<0x140>(v);
  return v;
}

// v of lane K of this lane's quad (quad_perm [K,K,K,K])
template <int K>
__device__ __forceinline__ double quad_bcast(double v) {
  return dpp_row<K * 0x55>(v);
}
template <int K>
__device__ __forceinline__ int quad_bcast_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, K * 0x55, 0xF, 0xF, false);
}

}  // namespace kh
