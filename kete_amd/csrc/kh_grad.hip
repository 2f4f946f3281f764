// kh_grad.hip -- astrometric observables of N-body orbits and their partial
// derivatives with respect to the initial state and the non-gravitational
// parameters A1, A2, A3 (forward-mode sensitivities of the integrator).
//
// kete has no derivatives of propagate_n_body; the observable chain is the
// one its FOV check uses (fov_like.rs:89-110): Radau to the observation
// epoch (propagate_n_body_spk, propagation/mod.rs:70-118), one two-body step
// back by the light time |r - r_obs|/c (kepler.rs:254-322), then RA/Dec of
// the observer-to-object vector in the equatorial frame
// (simult_states.rs:86-120, frames/definitions.rs:337-345).
//
// Layout and launch
//  - One lane per (object, parameter): 9 lanes per object.  Every lane runs
//    the full nominal integration in dual numbers (value, derivative along its
//    own parameter); all step-size and convergence decisions read the values
//    only, so the 9 lanes of an object take identical decisions and need no
//    communication.  The derivative is that of the computed map with the step
//    sequence held fixed (the step sizes depend on the parameters only at the
//    integrator's error level, ~1e-13 relative).
//  - Observation epochs are visited in the order given.  restart = 1: each arc
//    starts a new integrator (B = 0, h = 0.1 d), as successive
//    propagate_n_body calls do.  restart = 0: one integration through all
//    epochs; a step is shortened to land on each epoch, and the integrator
//    keeps B and the step size it proposed before shortening.
//  - The loop body makes one force evaluation, as in the propagation kernel,
//    so every live lane of a wavefront evaluates the force together.
//  - Light-time step: the value uses the universal Kepler solver; its
//    derivative uses the two-body state transition matrix expanded to third
//    order in the step (|dt| = light time < 0.06 d: relative error < 1e-9).
#include "kh_device.h"
#include "kh_internal.h"

namespace kh {

__constant__ RadauTables RTG;

int32_t upload_grad_tables(const RadauTables& t) {
  return hipMemcpyToSymbol(HIP_SYMBOL(RTG), &t, sizeof(t)) == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Dual numbers: value and one directional derivative.
// ---------------------------------------------------------------------------
struct Dual {
  double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
  const double s = sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
// a^p for a > 0 and a constant exponent
__device__ __forceinline__ Dual dpow(Dual a, double p) {
  const double r = pow(a.v, p);
  return {r, p * r / a.v * a.d};
}

struct DVec3 {
  Dual x, y, z;
};
__device__ __forceinline__ DVec3 operator+(DVec3 a, DVec3 b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ DVec3 operator-(DVec3 a, DVec3 b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ DVec3 operator-(DVec3 a, Vec3 b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ DVec3 operator*(DVec3 a, Dual s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ DVec3 operator*(DVec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ Dual ddot(DVec3 a, DVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Vec3 val(DVec3 a) { return {a.x.v, a.y.v, a.z.v}; }
__device__ __forceinline__ Vec3 der(DVec3 a) { return {a.x.d, a.y.d, a.z.d}; }
__device__ __forceinline__ DVec3 dvec(Vec3 v, Vec3 d) {
  return {{v.x, d.x}, {v.y, d.y}, {v.z, d.z}};
}
__device__ __forceinline__ DVec3 ecl_to_equ(DVec3 a) {
  const Vec3 v = ecl_to_equ(val(a)), d = ecl_to_equ(der(a));
  return dvec(v, d);
}
__device__ __forceinline__ DVec3 equ_to_ecl(DVec3 a) {
  const Vec3 v = equ_to_ecl(val(a)), d = equ_to_ecl(der(a));
  return dvec(v, d);
}

// ---------------------------------------------------------------------------
// Force model in dual numbers: the terms of GravParams::add_acceleration
// (gravity.rs:171-253) and JplComet (nongrav.rs:114-157).
// ---------------------------------------------------------------------------
struct GradNonGrav {
  Dual a1, a2, a3;
  double alpha, r_0, m, n, k;
};

__device__ __forceinline__ void nongrav_dual(const GradNonGrav& ng, DVec3& a, DVec3 p, DVec3 v) {
  const Dual n = dsqrt(ddot(p, p));
  const DVec3 pn{p.x / n, p.y / n, p.z / n};
  const Dual vd = ddot(v, pn);
  DVec3 t = v - pn * vd;
  const Dual tn = dsqrt(ddot(t, t));
  t = DVec3{t.x / tn, t.y / tn, t.z / tn};
  DVec3 nv{t.y * pn.z - t.z * pn.y, t.z * pn.x - t.x * pn.z, t.x * pn.y - t.y * pn.x};
  const Dual nn = dsqrt(ddot(nv, nv));
  nv = DVec3{nv.x / nn, nv.y / nn, nv.z / nn};
  const Dual rr0 = n / ng.r_0;
  const Dual scale = ng.alpha * dpow(rr0, -ng.m) * dpow(1.0 + dpow(rr0, ng.n), -ng.k);
  const Dual s1 = scale * ng.a1, s2 = scale * ng.a2, s3 = scale * ng.a3;
  a = a + pn * s1;
  a = a + t * s2;
  a = a + nv * s3;
}

// One body's terms at relative position rp = x - body, relative velocity
// rv = v - body velocity (kind 1 only).  Returns ST_IMPACT on an exact
// evaluation inside the body's radius.
__device__ __forceinline__ int32_t body_dual(const BodyDev& bd, DVec3 rp, DVec3 rv, bool exact,
                                             const GradNonGrav& ng, DVec3& a) {
  const double mass = bd.mass;
  const Dual r2 = ddot(rp, rp);
  const Dual r = dsqrt(r2);
  if (exact && r.v <= bd.radius) return ST_IMPACT;
  const Dual inv_r = 1.0 / r;
  const Dual inv_r3 = inv_r * inv_r * inv_r;
  a = a - rp * (mass * inv_r3);
  if (bd.kind == 1) {
    const Dual r_v = 4.0 * ddot(rp, rv);
    const Dual v2 = ddot(rv, rv);
    const Dual gc = (mass * C_AU_PER_DAY_INV_SQUARED) * inv_r3;
    const Dual c = (4. * mass) * inv_r - v2;
    a = a + (rp * c + rv * r_v) * gc;
    const DVec3 pe = equ_to_ecl(rp);
    const Dual zr = pe.z * inv_r;
    const Dual z_squared = 5.0 * (zr * zr);
    const Dual rr = bd.radius * inv_r;
    const Dual coef2 = (1.5 * bd.j2 * mass) * (rr * rr) * inv_r3;
    const DVec3 j{pe.x * coef2 * (z_squared - 1.0), pe.y * coef2 * (z_squared - 1.0),
                  pe.z * coef2 * (z_squared - 3.0)};
    a = a + ecl_to_equ(j);
    if (bd.id == 10) nongrav_dual(ng, a, rp, rv);
  } else if (bd.kind == 2) {
    const Dual zr = rp.z * inv_r;
    const Dual z_squared = 5.0 * (zr * zr);
    const Dual rr = bd.radius * inv_r;
    const Dual coef2 = (1.5 * bd.j2 * mass) * (rr * rr) * inv_r3;
    a = a + DVec3{rp.x * coef2 * (z_squared - 1.0), rp.y * coef2 * (z_squared - 1.0),
                  rp.z * coef2 * (z_squared - 3.0)};
  }
  return ST_OK;
}

// spk_accel (acceleration.rs:121-168) over kete's PLANETS, in dual numbers.
// Body states do not depend on the parameters and are evaluated once.
__device__ __forceinline__ int32_t spk_accel_dual(const SpkDev* __restrict__ spk,
                                                  const double* __restrict__ coef, double t,
                                                  DVec3 x, DVec3 v, bool exact,
                                                  const GradNonGrav& ng, DVec3& acc,
                                                  int32_t& impact_id) {
  Vec3 emb;
  int32_t st = seg_pos(spk, coef, spk->emb_seg, t, emb);
  if (st) return st;
  DVec3 a{{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 1
  for (int b = 0; b < N_PLANETS; ++b) {
    const BodyDev& bd = spk->body[b];
    Vec3 bp, bv{0.0, 0.0, 0.0};
    if (bd.kind == 1) {
      st = seg_pos_vel(spk, coef, bd.seg, t, bp, bv);
    } else {
      st = seg_pos(spk, coef, bd.seg, t, bp);
      if (bd.emb) bp = Vec3{bp.x + emb.x, bp.y + emb.y, bp.z + emb.z};
    }
    if (st) return st;
    st = body_dual(bd, x - bp, v - bv, exact, ng, a);
    if (st) {
      impact_id = bd.id;
      return st;
    }
  }
  acc = a;
  return ST_OK;
}

// central_accel (acceleration.rs:66-80): the Sun at the origin, GMS.
__device__ __forceinline__ DVec3 central_dual(DVec3 x) {
  const Dual r = dsqrt(ddot(x, x));
  const Dual k = GMS / (r * r * r);
  return DVec3{-(x.x * k), -(x.y * k), -(x.z * k)};
}

struct GradArgs {
  int64_t n;
  int32_t m;
  int32_t frame;  // input/output frame: 1 equatorial, 17 ecliptic (Sun-centred)
  int32_t force;  // 0: spk_accel with PLANETS; 1: central_accel, no ephemeris
  int32_t restart;  // 1: a new integrator per epoch; 0: one integration
  double epsilon;
  const double* jd0;        // [n]
  const double* xyz[6];     // [n], Sun-centred, input frame
  const NonGravDev* ng;     // [n] or null
  const double* jd_obs;     // [m][n]
  const double* obs[3];     // [m][n] observer position w.r.t. the Sun, input frame
  Chain sun;                // segments summing to the Sun's SSB state
  double* radec[2];         // [m][n]
  double* jac[2];           // [9][m][n]
  double* state[6];         // [m][n] or null: geometric state at each epoch
  double* dstate[6];        // [9][m][n] or null
  int32_t* status;          // [n]
  int32_t* counters;        // [n][4] or null
};

enum : int32_t { GP_INIT = 0, GP_SUB = 1, GP_EXACT = 2 };

// Two-body light-time step: value by the universal Kepler solver, derivative
// by Phi(dt) = [[I + G dt^2/2, dt (I + G dt^2/6)], ...] applied to the
// heliocentric state derivative, G = GMS (3 r r^T / r^2 - I) / r^3, plus the
// end velocity times the derivative of dt itself.
__device__ __forceinline__ int32_t light_time_step(DVec3 xh, DVec3 vh, Dual dt, DVec3& xo) {
  Vec3 po, vo;
  const int32_t st = analytic_2_body(dt.v, val(xh), val(vh), po, vo);
  if (st) return st;
  const Vec3 r = val(xh);
  const double r2 = dot(r, r);
  const double ir = 1.0 / sqrt(r2);
  const double g0 = GMS * ir * ir * ir;
  const Vec3 dx = der(xh), dv = der(vh);
  const double t = dt.v, t2 = t * t;
  // G y = g0 (3 r (r.y) / r^2 - y)
  const double rdx = 3.0 * dot(r, dx) / r2, rdv = 3.0 * dot(r, dv) / r2;
  const Vec3 gdx{g0 * (r.x * rdx - dx.x), g0 * (r.y * rdx - dx.y), g0 * (r.z * rdx - dx.z)};
  const Vec3 gdv{g0 * (r.x * rdv - dv.x), g0 * (r.y * rdv - dv.y), g0 * (r.z * rdv - dv.z)};
  const Vec3 d{dx.x + 0.5 * t2 * gdx.x + t * (dv.x + t2 * gdv.x / 6.0) + vo.x * dt.d,
               dx.y + 0.5 * t2 * gdx.y + t * (dv.y + t2 * gdv.y / 6.0) + vo.y * dt.d,
               dx.z + 0.5 * t2 * gdx.z + t * (dv.z + t2 * gdv.z / 6.0) + vo.z * dt.d};
  xo = dvec(po, d);
  return ST_OK;
}

__device__ __forceinline__ double rem_tau(double x) {
  const double r = fmod(x, 6.283185307179586);
  return r < 0.0 ? r + 6.283185307179586 : r;
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void radau_partials_kernel(const SpkDev* __restrict__ spk, const double* __restrict__ coef,
                           GradArgs p) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.n * 9) return;
  const int64_t o = q / 9;
  const int k = (int)(q - o * 9);
  const int64_t n = p.n;
  const int32_t m = p.m;
  const double nan = __builtin_nan("");
  const bool ecl = p.frame == 17;
  const bool central = p.force == 1;

  // parameters: this lane differentiates along parameter k
  Vec3 x0{p.xyz[0][o], p.xyz[1][o], p.xyz[2][o]};
  Vec3 v0{p.xyz[3][o], p.xyz[4][o], p.xyz[5][o]};
  const Vec3 e0{k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
  const Vec3 e1{k == 3 ? 1.0 : 0.0, k == 4 ? 1.0 : 0.0, k == 5 ? 1.0 : 0.0};
  GradNonGrav ng{{0.0, k == 6 ? 1.0 : 0.0}, {0.0, k == 7 ? 1.0 : 0.0}, {0.0, k == 8 ? 1.0 : 0.0},
                 1.0, 1.0, 2.0, 1.0, 0.0};
  if (p.ng != nullptr && p.ng[o].kind == 1) {
    const NonGravDev& g = p.ng[o];
    ng.a1.v = g.a1;
    ng.a2.v = g.a2;
    ng.a3.v = g.a3;
    ng.alpha = g.alpha;
    ng.r_0 = g.r_0;
    ng.m = g.m;
    ng.n = g.n;
    ng.k = g.k;
  }

  int32_t st = ST_OK;
  int32_t impact = 0;
  double t = p.jd0[o];
  if (!(isfinite(t) && isfinite(x0.x) && isfinite(x0.y) && isfinite(x0.z) && isfinite(v0.x) &&
        isfinite(v0.y) && isfinite(v0.z)))
    st = ST_VALUE_ERROR;

  // to barycentric equatorial (propagation/mod.rs:78-79): add the Sun's SSB
  // state in the input frame, then rotate
  if (st == ST_OK && !central) {
    Vec3 sp, sv;
    st = chain_state(spk, coef, p.sun, t, sp, sv);
    if (ecl) {
      sp = equ_to_ecl(sp);
      sv = equ_to_ecl(sv);
    }
    x0 = Vec3{x0.x + sp.x, x0.y + sp.y, x0.z + sp.z};
    v0 = Vec3{v0.x + sv.x, v0.y + sv.y, v0.z + sv.z};
  }
  DVec3 x = dvec(x0, e0), v = dvec(v0, e1);
  if (ecl) {
    x = ecl_to_equ(x);
    v = ecl_to_equ(v);
  }

  const DVec3 zero3{{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  DVec3 a = zero3;
  DVec3 B[7], G[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) B[i] = G[i] = zero3;
  double h = 0.0, next_h = 0.0, tf = 0.0, h_prop = 0.0;
  bool clamped = false;  // the current step was shortened to land on tf
  double bprev[3] = {0.0, 0.0, 0.0};
  int32_t phase = GP_INIT, sub = 1, sweep = 0, fails = 0, j = 0;
  int4 cnt{0, 0, 0, 0};
  if (st == ST_OK && m > 0) tf = p.jd_obs[o];

  while (st == ST_OK && j < m) {
    // evaluation point
    double te = t;
    DVec3 xe = x, ve = v;
    const bool exact = phase != GP_SUB;
    if (phase == GP_SUB) {
      const double s = RTG.h[sub];
      te = t + s * h;
      Dual bw[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}, bu[3] = {bw[0], bw[1], bw[2]};
#pragma unroll
      for (int c = 0; c < 7; ++c) {
        const double wp = RTG.wpow[sub][c], up = RTG.upow[sub][c];
        bw[0] = bw[0] + B[c].x * wp;
        bw[1] = bw[1] + B[c].y * wp;
        bw[2] = bw[2] + B[c].z * wp;
        bu[0] = bu[0] + B[c].x * up;
        bu[1] = bu[1] + B[c].y * up;
        bu[2] = bu[2] + B[c].z * up;
      }
      const double sh = s * h, s2h2 = (s * s) * (h * h);
      xe = DVec3{x.x + v.x * sh + (a.x / 2.0 + bw[0]) * s2h2,
                 x.y + v.y * sh + (a.y / 2.0 + bw[1]) * s2h2,
                 x.z + v.z * sh + (a.z / 2.0 + bw[2]) * s2h2};
      ve = DVec3{v.x + (a.x + bu[0]) * sh, v.y + (a.y + bu[1]) * sh, v.z + (a.z + bu[2]) * sh};
    }
    DVec3 ev;
    if (central) {
      ev = central_dual(xe);
    } else {
      st = spk_accel_dual(spk, coef, te, xe, ve, exact, ng, ev, impact);
    }
    cnt.x++;
    if (st) break;

    bool start_step = false, arc_done = false;
    if (phase == GP_INIT) {
      a = ev;
      if (fabs(tf - t) < 1e-10) {
        arc_done = true;
      } else {
        h = copysign(0.1, tf - t);
        fails = 0;
        start_step = true;
      }
    } else if (phase == GP_SUB) {
      // G column `sub` (radau.rs:292-308)
      const double ih = RTG.inv_h[sub];
      DVec3 gk{(ev.x - a.x) * ih, (ev.y - a.y) * ih, (ev.z - a.z) * ih};
#pragma unroll
      for (int z = 1; z < 7; ++z) {
        if (z < sub) {
          const double rd = RTG.rdiff[sub][z];
          gk = DVec3{(gk.x - G[z - 1].x) * rd, (gk.y - G[z - 1].y) * rd,
                     (gk.z - G[z - 1].z) * rd};
        }
      }
#pragma unroll
      for (int z = 0; z < 7; ++z)
        if (z == sub - 1) G[z] = gk;
      if (sub < 7) {
        ++sub;
      } else {
        // B = G C (radau.rs:313)
#pragma unroll
        for (int c = 0; c < 7; ++c) {
          DVec3 acc = G[c];
#pragma unroll
          for (int r = c + 1; r < 7; ++r) {
            const double cc = RTG.c[r][c];
            acc = DVec3{acc.x + G[r].x * cc, acc.y + G[r].y * cc, acc.z + G[r].z * cc};
          }
          B[c] = acc;
        }
        const double f0 = fabs(ev.x.v) + 1e-6, f1 = fabs(ev.y.v) + 1e-6,
                     f2 = fabs(ev.z.v) + 1e-6;
        const double ratio = fmax(fmax(fabs(B[6].x.v - bprev[0]) / f0,
                                       fabs(B[6].y.v - bprev[1]) / f1),
                                  fabs(B[6].z.v - bprev[2]) / f2);
        if (ratio < 1e-14) {
          // accept (radau.rs:320-342)
          Dual bw[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}, bu[3] = {bw[0], bw[1], bw[2]};
#pragma unroll
          for (int c = 0; c < 7; ++c) {
            const double w = RTG.w[c], u = RTG.u[c];
            bw[0] = bw[0] + B[c].x * w;
            bw[1] = bw[1] + B[c].y * w;
            bw[2] = bw[2] + B[c].z * w;
            bu[0] = bu[0] + B[c].x * u;
            bu[1] = bu[1] + B[c].y * u;
            bu[2] = bu[2] + B[c].z * u;
          }
          const double h2 = h * h;
          x = DVec3{x.x + v.x * h + (a.x * 0.5 + bw[0]) * h2,
                    x.y + v.y * h + (a.y * 0.5 + bw[1]) * h2,
                    x.z + v.z * h + (a.z * 0.5 + bw[2]) * h2};
          v = DVec3{v.x + (a.x + bu[0]) * h, v.y + (a.y + bu[1]) * h, v.z + (a.z + bu[2]) * h};
          t += h;
          const double fm = fmax(fmax(f0, f1), f2);
          const double bm = fmax(fmax(fabs(B[6].x.v), fabs(B[6].y.v)), fabs(B[6].z.v));
          double r7 = pow(p.epsilon * fm / bm, 1.0 / 7.0);
          r7 = r7 < 0.25 ? 0.25 : (r7 > 4.0 ? 4.0 : r7);
          next_h = h * r7;
          cnt.z++;
          phase = GP_EXACT;
        } else if (sweep < 9) {
          ++sweep;
          sub = 1;
          bprev[0] = B[6].x.v;
          bprev[1] = B[6].y.v;
          bprev[2] = B[6].z.v;
          cnt.y++;
        } else {
          cnt.w++;
          ++fails;
          h *= 0.7;
          if (fails > 10) {
            st = ST_CONVERGENCE;
            break;
          }
          if (fabs(h) < 0.005) h = copysign(0.005, h);
          start_step = true;
        }
      }
    } else {  // GP_EXACT
      a = ev;
      h = next_h;
      if (fabs(t - tf) < 1e-12) {
        arc_done = true;
      } else {
        fails = 0;
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        start_step = true;
      }
    }
    if (start_step) {
      h_prop = h;
      clamped = fabs(t - tf) <= fabs(h);
      if (clamped) h = tf - t;
      sweep = 0;
      sub = 1;
#pragma unroll
      for (int i = 0; i < 7; ++i) G[i] = zero3;
      bprev[0] = B[6].x.v;
      bprev[1] = B[6].y.v;
      bprev[2] = B[6].z.v;
      phase = GP_SUB;
      cnt.y++;
    }
    if (arc_done) {
      // observables at epoch j
      Vec3 sp{0.0, 0.0, 0.0}, sv{0.0, 0.0, 0.0};
      if (!central) {
        st = chain_state(spk, coef, p.sun, tf, sp, sv);
        if (st) break;
      }
      const DVec3 xh = x - sp, vh = v - sv;
      const int64_t jo = (int64_t)j * n + o;
      Vec3 ob{p.obs[0][jo], p.obs[1][jo], p.obs[2][jo]};
      if (ecl) ob = ecl_to_equ(ob);
      const DVec3 rho0 = xh - ob;
      const Dual dt = -(dsqrt(ddot(rho0, rho0)) * C_AU_PER_DAY_INV);
      DVec3 xl;
      st = light_time_step(xh, vh, dt, xl);
      if (st) break;
      const DVec3 rho = xl - ob;
      const Vec3 r = val(rho), dr = der(rho);
      const double rxy2 = r.x * r.x + r.y * r.y;
      const double rn2 = rxy2 + r.z * r.z;
      const double rn = sqrt(rn2);
      double ra = 0.0, dec = 0.0;
      if (!(rn < 1e-10)) {
        const double zr = fmin(fmax(r.z / rn, -1.0), 1.0);
        dec = rem_tau(3.0 * (3.141592653589793 / 2.0) - acos(zr)) - 3.141592653589793;
        ra = rem_tau(atan2(r.y, r.x));
      }
      const double dra = (r.x * dr.y - r.y * dr.x) / rxy2;
      const double ddec = (dr.z * rxy2 - r.z * (r.x * dr.x + r.y * dr.y)) / (rn2 * sqrt(rxy2));
      const int64_t jk = ((int64_t)k * m + j) * n + o;
      if (k == 0) {
        p.radec[0][jo] = ra;
        p.radec[1][jo] = dec;
      }
      p.jac[0][jk] = dra;
      p.jac[1][jk] = ddec;
      if (p.state[0] != nullptr) {
        DVec3 so = xh, vo = vh;
        if (ecl) {
          so = equ_to_ecl(so);
          vo = equ_to_ecl(vo);
        }
        if (k == 0) {
          p.state[0][jo] = so.x.v;
          p.state[1][jo] = so.y.v;
          p.state[2][jo] = so.z.v;
          p.state[3][jo] = vo.x.v;
          p.state[4][jo] = vo.y.v;
          p.state[5][jo] = vo.z.v;
        }
        if (p.dstate[0] != nullptr) {
          p.dstate[0][jk] = so.x.d;
          p.dstate[1][jk] = so.y.d;
          p.dstate[2][jk] = so.z.d;
          p.dstate[3][jk] = vo.x.d;
          p.dstate[4][jk] = vo.y.d;
          p.dstate[5][jk] = vo.z.d;
        }
      }
      ++j;
      if (j < m) {
        tf = p.jd_obs[(int64_t)j * n + o];
        if (p.restart) {
          // a new integrator from this state (RadauIntegrator::new)
#pragma unroll
          for (int i = 0; i < 7; ++i) B[i] = zero3;
          phase = GP_INIT;
        } else if (fabs(tf - t) < 1e-10) {
          phase = GP_EXACT;  // the same epoch again: a zero-length arc
        } else {
          // carry on with the step size the controller proposed: before the
          // last step was shortened to land on the epoch, else after it
          h = copysign(fmax(fabs(clamped ? h_prop : h), 0.005), tf - t);
          fails = 0;
          h_prop = h;
          clamped = fabs(t - tf) <= fabs(h);
          if (clamped) h = tf - t;
          sweep = 0;
          sub = 1;
#pragma unroll
          for (int i = 0; i < 7; ++i) G[i] = zero3;
          bprev[0] = B[6].x.v;
          bprev[1] = B[6].y.v;
          bprev[2] = B[6].z.v;
          phase = GP_SUB;
          cnt.y++;
        }
      }
    }
  }

  // an error leaves this epoch and the later ones undefined
  for (int jj = j; jj < m; ++jj) {
    const int64_t jo = (int64_t)jj * n + o;
    const int64_t jk = ((int64_t)k * m + jj) * n + o;
    if (k == 0) {
      p.radec[0][jo] = nan;
      p.radec[1][jo] = nan;
    }
    p.jac[0][jk] = nan;
    p.jac[1][jk] = nan;
    if (p.state[0] != nullptr) {
      if (k == 0)
        for (int c = 0; c < 6; ++c) p.state[c][jo] = nan;
      if (p.dstate[0] != nullptr)
        for (int c = 0; c < 6; ++c) p.dstate[c][jk] = nan;
    }
  }
  if (k == 0) {
    p.status[o] = st;
    if (p.counters) reinterpret_cast<int4*>(p.counters)[o] = cnt;
  }
}

int32_t launch_radau_partials(const SpkDev* spk, const double* coef, const GradLaunch& g,
                              hipStream_t stream) {
  GradArgs p{};
  p.n = g.n;
  p.m = g.m;
  p.frame = g.frame;
  p.force = g.force;
  p.restart = g.restart;
  p.epsilon = g.epsilon;
  p.jd0 = g.jd0;
  for (int c = 0; c < 6; ++c) {
    p.xyz[c] = g.xyz[c];
    p.state[c] = g.state[c];
    p.dstate[c] = g.dstate[c];
  }
  p.ng = reinterpret_cast<const NonGravDev*>(g.ng);
  p.jd_obs = g.jd_obs;
  for (int c = 0; c < 3; ++c) p.obs[c] = g.obs[c];
  p.sun = g.sun;
  p.radec[0] = g.radec[0];
  p.radec[1] = g.radec[1];
  p.jac[0] = g.jac[0];
  p.jac[1] = g.jac[1];
  p.status = g.status;
  p.counters = g.counters;
  const int64_t lanes = g.n * 9;
  if (lanes <= 0) return 0;
  const unsigned blocks = (unsigned)((lanes + 63) / 64);
  hipLaunchKernelGGL(radau_partials_kernel, dim3(blocks), dim3(64), 0, stream, spk, coef, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace kh
