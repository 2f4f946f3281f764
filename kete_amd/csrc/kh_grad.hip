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
// Two kernels compute the same columns.  The default,
// radau_partials_wide_kernel (below), spreads an object over sixteen lanes:
// one body of the force model and one column of the state per lane, with
// analytic force Jacobians, then partials_observe_kernel forms the
// observables.  radau_partials_kernel (KH_PARTIALS=lane9) is the first
// design, kept as a second implementation to compare against:
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
#include <cstdlib>
#include <cstring>

#include "kh_device.h"
#include "kh_internal.h"
#include "kh_jac.h"

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
  // wide kernel: the states at each epoch, for the observation kernel
  double* ep_state;         // [m][n][10][6]: columns 0..8 (derivatives), 9 (nominal)
  double* ep_sun;           // [m][n][6]: the Sun's barycentric state at the epoch
  int32_t* ep_done;         // [n]: epochs completed
  int32_t* lt_fail;         // [n]: first epoch whose light-time step failed (m: none)
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

// ===========================================================================
// Wide-lane partials kernel: sixteen lanes per object (one DPP row), four
// objects per wavefront.  Lane l of an object's group
//  - carries one column of the integrator state: l = 0..8 the derivative of
//    the state along parameter l (x, y, z, vx, vy, vz, A1, A2, A3), l = 9 the
//    nominal orbit, l = 10..15 a zero column;
//  - evaluates one body of the force model: l = 0..9 kete's PLANETS in order
//    (gravity.rs:145-148), l = 10 the Earth-Moon barycentre segment, which
//    lanes 3 (Earth) and 4 (Moon) add to their positions: the body's Newton,
//    GR and J2 terms and their Jacobians with respect to the object's
//    position and velocity (kh_jac.h).
// Per force evaluation the group sums the acceleration and the Jacobians
// over its lanes (DPP row reductions); the nominal lane adds the non-grav
// term, and each tangent lane forms its column's acceleration
// J_r dx + J_v dv plus the directional derivative of the non-grav term
// (dual numbers along its own column).  The nominal lane's predicted state,
// the Sun's state and the integrator's decisions reach the rest of the group
// by lane shuffles.  A lane thus does the work of one body and one column,
// where the one-lane-per-parameter kernel above evaluates the whole force
// model and the nominal orbit in every lane.  The arithmetic of each column
// is the dual-number kernel's; only the order in which body terms are summed
// differs (a row reduction instead of kete's sequential sum).
// ===========================================================================
constexpr int WL = 16;
constexpr int WL_NOMINAL = 9;
constexpr int WL_EMB = 10;

// v from another lane of the same DPP row (a 64-bit move is two 32-bit ones)
// JplComet non-grav acceleration (nongrav.rs:114-157) a and its derivative
// da along (dp, dv, dA) at position p (ip = 1/|p|), velocity v relative to
// the Sun:
// a = g(r) (A1 r^ + A2 t^ + A3 n^), g = alpha (r/r_0)^-m (1 + (r/r_0)^n)^-k.
// Norms come from reciprocal square roots and kete's divisions by them
// become multiplications (last-bit differences); with kete's default
// exponents (m = 2, k = 0) g = alpha (r/r_0)^-2 needs no pow.
__device__ __forceinline__ void nongrav_dir(const double p[3], const double dp[3],
                                            const double v[3], const double dv[3],
                                            const double A[3], const double dA[3], double alpha,
                                            double r_0, double m, double ne, double k,
                                            double ip, double a[3], double da[3]) {
  auto dot3 = [](const double* x, const double* y) {
    return x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
  };
  const double n = dot3(p, p) * ip;
  const double dn = dot3(p, dp) * ip;
  double pn[3], dpn[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pn[c] = p[c] * ip;
    dpn[c] = (dp[c] - pn[c] * dn) * ip;
  }
  const double vd = dot3(v, pn);
  const double dvd = dot3(dv, pn) + dot3(v, dpn);
  double tv[3], dtv[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    tv[c] = v[c] - pn[c] * vd;
    dtv[c] = dv[c] - dpn[c] * vd - pn[c] * dvd;
  }
  const double it = rsqrt_nr(dot3(tv, tv));
  const double dtn = dot3(tv, dtv) * it;
  double tu[3], dtu[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    tu[c] = tv[c] * it;
    dtu[c] = (dtv[c] - tu[c] * dtn) * it;
  }
  const double nv[3] = {tu[1] * pn[2] - tu[2] * pn[1], tu[2] * pn[0] - tu[0] * pn[2],
                        tu[0] * pn[1] - tu[1] * pn[0]};
  const double dnv[3] = {dtu[1] * pn[2] - dtu[2] * pn[1] + tu[1] * dpn[2] - tu[2] * dpn[1],
                         dtu[2] * pn[0] - dtu[0] * pn[2] + tu[2] * dpn[0] - tu[0] * dpn[2],
                         dtu[0] * pn[1] - dtu[1] * pn[0] + tu[0] * dpn[1] - tu[1] * dpn[0]};
  const double in_ = rsqrt_nr(dot3(nv, nv));
  const double dnn = dot3(nv, dnv) * in_;
  double nu[3], dnu[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    nu[c] = nv[c] * in_;
    dnu[c] = (dnv[c] - nu[c] * dnn) * in_;
  }
  const double rr0 = n / r_0, drr0 = dn / r_0;
  double g, dg;
  if (m == 2.0 && k == 0.0) {
    const double ir = 1.0 / rr0;
    g = alpha * (ir * ir);
    dg = -2.0 * g * drr0 * ir;
  } else {
    const double q = pow(rr0, ne);
    g = alpha * pow(rr0, -m) * pow(1.0 + q, -k);
    dg = g * (-m * drr0 / rr0 - k * ne * q * drr0 / (rr0 * (1.0 + q)));
  }
  double s[3], ds[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    s[c] = g * A[c];
    ds[c] = dg * A[c] + g * dA[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a[c] = pn[c] * s[0] + tu[c] * s[1] + nu[c] * s[2];
    da[c] = dpn[c] * s[0] + pn[c] * ds[0] + dtu[c] * s[1] + tu[c] * ds[1] + dnu[c] * s[2] +
            nu[c] * ds[2];
  }
}

template <bool CENTRAL>
__global__ __launch_bounds__(256) void radau_partials_wide_kernel(
    const SpkDev* __restrict__ spk, const double* __restrict__ coef, GradArgs p) {
  // per-substep factors, indexed by each group's own substep
  __shared__ double s_h[8], s_inv_h[8], s_rdiff[8][8];
  // the Chebyshev record each segment lane last used, per object of the block
  __shared__ __align__(16) double s_rec[256 / WL][WL_EMB + 1][44];
  for (int q = threadIdx.x; q < 64; q += blockDim.x) {
    s_rdiff[q / 8][q % 8] = RTG.rdiff[q / 8][q % 8];
    if (q < 8) {
      s_h[q] = RTG.h[q];
      s_inv_h[q] = RTG.inv_h[q];
    }
  }
  __syncthreads();
  const int64_t gl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t o = gl / WL;
  if (o >= p.n) return;
  const int lane = threadIdx.x & 63;
  const int l = lane & (WL - 1);
  const int g0 = lane & ~(WL - 1);
  const int64_t n = p.n;
  const int32_t m = p.m;
  const double nan = __builtin_nan("");
  const bool ecl = p.frame == 17;
  constexpr bool central = CENTRAL;
  const bool nominal = l == WL_NOMINAL;
  const bool tangent = l < 9;

  // JplComet parameters (new_asteroid defaults and A = 0 for kind 0); a
  // dust model, a JplComet delay or an unknown kind has nothing to
  // differentiate here: the object ends with ValueError
  double a1 = 0.0, a2 = 0.0, a3 = 0.0, alpha = 1.0, r_0 = 1.0, mexp = 2.0, nexp = 1.0,
         kexp = 0.0;
  bool bad_model = false;
  if (p.ng != nullptr) {
    const int32_t kind = p.ng[o].kind;
    bad_model = kind < 0 || kind > 1 || (kind == 1 && p.ng[o].dt != 0.0);
  }
  if (p.ng != nullptr && p.ng[o].kind == 1) {
    const NonGravDev& g = p.ng[o];
    a1 = g.a1;
    a2 = g.a2;
    a3 = g.a3;
    alpha = g.alpha;
    r_0 = g.r_0;
    mexp = g.m;
    nexp = g.n;
    kexp = g.k;
  }
  const double ngA[3] = {a1, a2, a3};
  const double ngdA[3] = {l == 6 ? 1.0 : 0.0, l == 7 ? 1.0 : 0.0, l == 8 ? 1.0 : 0.0};

  // nominal initial state, barycentric equatorial (every lane)
  int32_t st = ST_OK;
  double t = p.jd0[o];
  Vec3 x0{p.xyz[0][o], p.xyz[1][o], p.xyz[2][o]};
  Vec3 v0{p.xyz[3][o], p.xyz[4][o], p.xyz[5][o]};
  if (!(isfinite(t) && isfinite(x0.x) && isfinite(x0.y) && isfinite(x0.z) && isfinite(v0.x) &&
        isfinite(v0.y) && isfinite(v0.z)))
    st = ST_VALUE_ERROR;
  if (bad_model) st = ST_VALUE_ERROR;
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
  if (ecl) {
    x0 = ecl_to_equ(x0);
    v0 = ecl_to_equ(v0);
  }
  // this lane's column: the nominal state, or the seed of parameter l
  Vec3 sx{l == 0 ? 1.0 : 0.0, l == 1 ? 1.0 : 0.0, l == 2 ? 1.0 : 0.0};
  Vec3 sv{l == 3 ? 1.0 : 0.0, l == 4 ? 1.0 : 0.0, l == 5 ? 1.0 : 0.0};
  if (ecl) {
    sx = ecl_to_equ(sx);
    sv = ecl_to_equ(sv);
  }
  double X[3] = {nominal ? x0.x : sx.x, nominal ? x0.y : sx.y, nominal ? x0.z : sx.z};
  double V[3] = {nominal ? v0.x : sv.x, nominal ? v0.y : sv.y, nominal ? v0.z : sv.z};
  double A[3] = {0.0, 0.0, 0.0};
  double B[7][3], G[7][3];
#pragma unroll
  for (int i = 0; i < 7; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) B[i][c] = G[i][c] = 0.0;

  // this lane's body (lanes 0..9) or segment (lane 10); lanes 11..15 none
  const bool is_body = l < N_PLANETS && (!central || l == 0);
  const bool is_seg = !central && l <= WL_EMB;
  const BodyDev bd = spk->body[l < N_PLANETS ? l : 0];
  const int32_t seg = l == WL_EMB ? spk->emb_seg : bd.seg;
  const SegDev sg = spk->seg[seg];
  const bool all14 = spk->all14 != 0;
  const bool cache_rec = all14 && sg.rsize == 44;
  int32_t rec_cached = -1;
  const double mass = central ? GMS : bd.mass;
  const int gr = !central && bd.kind == 1;
  const double j2 = central ? 0.0 : bd.j2;
  const int ecl_j2 = bd.kind == 1;
  // the GR bodies of kete's PLANETS, the Sun (lane 0) and Jupiter (lane 6):
  // the tangent lanes differentiate their GR terms along their own column
  constexpr int JUP = 6;
  const double m_sun = spk->body[0].mass, m_jup = spk->body[JUP].mass;

  double h = 0.0, next_h = 0.0, tf = 0.0, h_prop = 0.0;
  bool clamped = false;
  double bprev[3] = {0.0, 0.0, 0.0};
  int32_t phase = GP_INIT, sub = 1, sweep = 0, fails = 0, j = 0, impact = 0;
  int4 cnt{0, 0, 0, 0};
  if (st == ST_OK && m > 0) tf = p.jd_obs[o];
  Vec3 sunp{0.0, 0.0, 0.0}, sunv{0.0, 0.0, 0.0};

  while (st == ST_OK && j < m) {
    // predicted column state at the evaluation point
    double te = t;
    double P[3] = {X[0], X[1], X[2]}, Q[3] = {V[0], V[1], V[2]};
    const bool exact = phase != GP_SUB;
    if (phase == GP_SUB) {
      const double s = s_h[sub];
      te = t + s * h;
      const double sh = s * h, s2h2 = (s * s) * (h * h);
      double bw[3] = {0.0, 0.0, 0.0}, bu[3] = {0.0, 0.0, 0.0};
      double sp = 1.0;
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        sp *= s;
        const double wq = sp * RTC.w[q], uq = sp * RTC.u[q];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bw[c] += B[q][c] * wq;
          bu[c] += B[q][c] * uq;
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        P[c] = X[c] + V[c] * sh + (A[c] / 2.0 + bw[c]) * s2h2;
        Q[c] = V[c] + (A[c] + bu[c]) * sh;
      }
    }
    const Vec3 xe{row_bcast<WL_NOMINAL>(P[0]), row_bcast<WL_NOMINAL>(P[1]),
                  row_bcast<WL_NOMINAL>(P[2])};
    const Vec3 ve{row_bcast<WL_NOMINAL>(Q[0]), row_bcast<WL_NOMINAL>(Q[1]),
                  row_bcast<WL_NOMINAL>(Q[2])};

    // this lane's segment at te, its record cached in LDS
    int32_t lst = ST_OK;
    Vec3 bp{0.0, 0.0, 0.0}, bv{0.0, 0.0, 0.0};
    if (is_seg) {
      if (te < sg.jd_start || te > sg.jd_end) lst = ST_DAF_LIMITS;
      const double js = (te - 2451545.0) * 86400.0;
      const int32_t ri = record_index(sg, js);
      const double* __restrict__ grec = coef + sg.off + (int64_t)ri * sg.rsize;
      double ir, xs;
      Vec3 pp, dd;
      if (cache_rec) {
        double* slot = &s_rec[threadIdx.x / WL][l][0];
        if (ri != rec_cached) {
          // the record, with word 1 (the radius) stored as its reciprocal
          const double2* __restrict__ src = reinterpret_cast<const double2*>(grec);
#pragma unroll
          for (int q = 0; q < 22; ++q) reinterpret_cast<double2*>(slot)[q] = src[q];
          slot[1] = 1.0 / grec[1];
          rec_cached = ri;
        }
        ir = slot[1];
        xs = (js - slot[0]) * ir;
        cheb_both_lds(slot, xs, pp, dd);
      } else {
        ir = 1.0 / grec[1];
        xs = (js - grec[0]) * ir;
        if (all14) cheb_both_n<14>(grec + 2, xs, pp, dd);
        else cheb_both_dyn(grec + 2, sg.n_coef, xs, pp, dd);
      }
      const double vsc = 86400.0 * ir * INV_AU_KM;
      bp = Vec3{pp.x * INV_AU_KM, pp.y * INV_AU_KM, pp.z * INV_AU_KM};
      bv = Vec3{dd.x * vsc, dd.y * vsc, dd.z * vsc};
    }
    Vec3 jupp{0.0, 0.0, 0.0}, jupv{0.0, 0.0, 0.0};
    if (!central) {
      const Vec3 emb{row_bcast<WL_EMB>(bp.x), row_bcast<WL_EMB>(bp.y),
                     row_bcast<WL_EMB>(bp.z)};
      if (l < N_PLANETS && bd.emb) bp = Vec3{bp.x + emb.x, bp.y + emb.y, bp.z + emb.z};
      sunp = Vec3{row_bcast<0>(bp.x), row_bcast<0>(bp.y), row_bcast<0>(bp.z)};
      sunv = Vec3{row_bcast<0>(bv.x), row_bcast<0>(bv.y), row_bcast<0>(bv.z)};
      jupp = Vec3{row_bcast<JUP>(bp.x), row_bcast<JUP>(bp.y), row_bcast<JUP>(bp.z)};
      jupv = Vec3{row_bcast<JUP>(bv.x), row_bcast<JUP>(bv.y), row_bcast<JUP>(bv.z)};
    }
    // this lane's body: acceleration and the symmetric Jacobian (Newton, J2)
    double ba[3] = {0.0, 0.0, 0.0}, bj[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double inv_r = 0.0;
    if (is_body) {
      const double rp[3] = {xe.x - bp.x, xe.y - bp.y, xe.z - bp.z};
      const double rv[3] = {ve.x - bv.x, ve.y - bv.y, ve.z - bv.z};
      const double r2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
      inv_r = rsqrt_nr(r2);
      if (exact && !central && r2 * inv_r <= bd.radius && lst == ST_OK) lst = ST_IMPACT;
      body_terms_sym(mass, bd.radius, j2, gr, ecl_j2, ROT_C, ROT_S, C_AU_PER_DAY_INV_SQUARED,
                     rp, rv, inv_r, ba, bj);
    }
    // the first failing body of the group, in kete's order
    const uint64_t bad = __ballot(lst != ST_OK);
    const uint32_t gbad = (uint32_t)((bad >> g0) & 0xFFFFull);
    cnt.x++;
    if (gbad) {
      const int first = __builtin_ctz(gbad);
      st = __shfl(lst, g0 + first, 64);
      if (st == ST_IMPACT) impact = spk->body[first < N_PLANETS ? first : 0].id;
      break;
    }
    // group sums; a tangent column's acceleration J dx is accumulated as the
    // sums arrive
    double ev[3];
    {
      double js[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) js[q] = row_sum(bj[q]);
      const double as0 = row_sum(ba[0]), as1 = row_sum(ba[1]), as2 = row_sum(ba[2]);
      ev[0] = nominal ? as0 : js[0] * P[0] + js[1] * P[1] + js[2] * P[2];
      ev[1] = nominal ? as1 : js[1] * P[0] + js[3] * P[1] + js[4] * P[2];
      ev[2] = nominal ? as2 : js[2] * P[0] + js[4] * P[1] + js[5] * P[2];
    }
    // inverse distances to the Sun and Jupiter, from their body lanes
    const double ir_sun = row_bcast<0>(inv_r), ir_jup = row_bcast<JUP>(inv_r);
    if (!central) {
      // derivatives of the GR terms of the Sun and Jupiter along this column
      const double dp[3] = {nominal ? 0.0 : P[0], nominal ? 0.0 : P[1], nominal ? 0.0 : P[2]};
      const double dq[3] = {nominal ? 0.0 : Q[0], nominal ? 0.0 : Q[1], nominal ? 0.0 : Q[2]};
      const Vec3 gb[2][2] = {{sunp, sunv}, {jupp, jupv}};
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const double rp[3] = {xe.x - gb[b][0].x, xe.y - gb[b][0].y, xe.z - gb[b][0].z};
        const double rv[3] = {ve.x - gb[b][1].x, ve.y - gb[b][1].y, ve.z - gb[b][1].z};
        double dg[3];
        gr_dir(b == 0 ? m_sun : m_jup, C_AU_PER_DAY_INV_SQUARED, rp, rv, dp, dq,
               b == 0 ? ir_sun : ir_jup, dg);
#pragma unroll
        for (int c = 0; c < 3; ++c) ev[c] += dg[c];
      }
    }
    if (!central) {
      // the non-grav term on the Sun's (acceleration.rs:161-165) and its
      // derivative along this column (the nominal lane takes the value)
      const double rp[3] = {xe.x - sunp.x, xe.y - sunp.y, xe.z - sunp.z};
      const double rv[3] = {ve.x - sunv.x, ve.y - sunv.y, ve.z - sunv.z};
      const double dp[3] = {nominal ? 0.0 : P[0], nominal ? 0.0 : P[1], nominal ? 0.0 : P[2]};
      const double dq[3] = {nominal ? 0.0 : Q[0], nominal ? 0.0 : Q[1], nominal ? 0.0 : Q[2]};
      double an[3], dan[3];
      nongrav_dir(rp, dp, rv, dq, ngA, ngdA, alpha, r_0, mexp, nexp, kexp, ir_sun, an, dan);
#pragma unroll
      for (int c = 0; c < 3; ++c) ev[c] += nominal ? an[c] : dan[c];
    }

    bool start_step = false, arc_done = false;
    if (phase == GP_INIT) {
      A[0] = ev[0];
      A[1] = ev[1];
      A[2] = ev[2];
      if (fabs(tf - t) < 1e-10) {
        arc_done = true;
      } else {
        h = copysign(0.1, tf - t);
        fails = 0;
        start_step = true;
      }
    } else if (phase == GP_SUB) {
      // G column `sub` (radau.rs:292-308)
      const double ih = s_inv_h[sub];
      double gk[3] = {(ev[0] - A[0]) * ih, (ev[1] - A[1]) * ih, (ev[2] - A[2]) * ih};
#pragma unroll
      for (int z = 1; z < 7; ++z) {
        const double rd = s_rdiff[sub][z];
#pragma unroll
        for (int c = 0; c < 3; ++c) gk[c] = z < sub ? (gk[c] - G[z - 1][c]) * rd : gk[c];
      }
#pragma unroll
      for (int z = 0; z < 7; ++z) {
        const bool put = z == sub - 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) G[z][c] = put ? gk[c] : G[z][c];
      }
      if (sub < 7) {
        ++sub;
      } else {
        // B = G C (radau.rs:313)
#pragma unroll
        for (int q = 0; q < 7; ++q)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double acc = G[q][c];
#pragma unroll
            for (int r = q + 1; r < 7; ++r) acc += G[r][c] * RTC.c[r][q];
            B[q][c] = acc;
          }
        // the nominal column decides (radau.rs:314-319)
        const double f0 = fabs(ev[0]) + 1e-6, f1 = fabs(ev[1]) + 1e-6, f2 = fabs(ev[2]) + 1e-6;
        const double my_ratio = fmax(fmax(fabs(B[6][0] - bprev[0]) / f0,
                                          fabs(B[6][1] - bprev[1]) / f1),
                                     fabs(B[6][2] - bprev[2]) / f2);
        const double ratio = row_bcast<WL_NOMINAL>(my_ratio);
        if (ratio < 1e-14) {
          // accept (radau.rs:320-342)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double bw = 0.0, bu = 0.0;
#pragma unroll
            for (int q = 0; q < 7; ++q) {
              bw += B[q][c] * RTC.w[q];
              bu += B[q][c] * RTC.u[q];
            }
            const double h2 = h * h;
            X[c] = X[c] + V[c] * h + (A[c] * 0.5 + bw) * h2;
            V[c] = V[c] + (A[c] + bu) * h;
          }
          t += h;
          const double fm = fmax(fmax(f0, f1), f2);
          const double bm = fmax(fmax(fabs(B[6][0]), fabs(B[6][1])), fabs(B[6][2]));
          double r7 = root7(p.epsilon * fm / bm);
          r7 = r7 < 0.25 ? 0.25 : (r7 > 4.0 ? 4.0 : r7);
          r7 = row_bcast<WL_NOMINAL>(r7);
          next_h = h * r7;
          cnt.z++;
          phase = GP_EXACT;
        } else if (sweep < 9) {
          ++sweep;
          sub = 1;
          bprev[0] = B[6][0];
          bprev[1] = B[6][1];
          bprev[2] = B[6][2];
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
      A[0] = ev[0];
      A[1] = ev[1];
      A[2] = ev[2];
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
      for (int i = 0; i < 7; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) G[i][c] = 0.0;
      bprev[0] = B[6][0];
      bprev[1] = B[6][1];
      bprev[2] = B[6][2];
      phase = GP_SUB;
      cnt.y++;
    }
    if (arc_done) {
      // the states at epoch j (barycentric equatorial), for the observation
      // kernel: this lane's column, and from the nominal lane the Sun's state
      // (evaluated at te = t = tf)
      const int64_t jo = (int64_t)j * n + o;
      if (l <= WL_NOMINAL) {
        double* dst = p.ep_state + (jo * 10 + l) * 6;
        dst[0] = X[0];
        dst[1] = X[1];
        dst[2] = X[2];
        dst[3] = V[0];
        dst[4] = V[1];
        dst[5] = V[2];
      }
      if (nominal) {
        double* ds = p.ep_sun + jo * 6;
        ds[0] = central ? 0.0 : sunp.x;
        ds[1] = central ? 0.0 : sunp.y;
        ds[2] = central ? 0.0 : sunp.z;
        ds[3] = central ? 0.0 : sunv.x;
        ds[4] = central ? 0.0 : sunv.y;
        ds[5] = central ? 0.0 : sunv.z;
      }
      ++j;
      if (j < m) {
        tf = p.jd_obs[(int64_t)j * n + o];
        if (p.restart) {
#pragma unroll
          for (int i = 0; i < 7; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) B[i][c] = 0.0;
          phase = GP_INIT;
        } else if (fabs(tf - t) < 1e-10) {
          phase = GP_EXACT;
        } else {
          h = copysign(fmax(fabs(clamped ? h_prop : h), 0.005), tf - t);
          fails = 0;
          h_prop = h;
          clamped = fabs(t - tf) <= fabs(h);
          if (clamped) h = tf - t;
          sweep = 0;
          sub = 1;
#pragma unroll
          for (int i = 0; i < 7; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) G[i][c] = 0.0;
          bprev[0] = B[6][0];
          bprev[1] = B[6][1];
          bprev[2] = B[6][2];
          phase = GP_SUB;
          cnt.y++;
        }
      }
    }
  }

  if (nominal) {
    p.ep_done[o] = j;
    p.status[o] = st;
    p.lt_fail[o] = m;
    if (p.counters) reinterpret_cast<int4*>(p.counters)[o] = cnt;
  }
  (void)impact;
}

// Observables at the epochs the wide kernel reached, one lane per (epoch,
// object): the light-time step (value by the universal Kepler solver,
// derivative by light_time_step) and RA/Dec of the observer-to-object vector
// (simult_states.rs:86-120) with their derivatives along the nine columns;
// the geometric states in the input frame when asked for.  A failed
// light-time solve records its epoch for partials_finish_kernel.
__global__ __launch_bounds__(256) void partials_observe_kernel(GradArgs p) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = p.n;
  const int32_t m = p.m;
  if (q >= (int64_t)m * n) return;
  const int32_t j = (int32_t)(q / n);
  const int64_t o = q - (int64_t)j * n;
  const int64_t jo = q;
  const double nan = __builtin_nan("");
  const bool ecl = p.frame == 17;
  bool ok = j < p.ep_done[o];
  Vec3 po{nan, nan, nan}, vo{nan, nan, nan}, ob{0.0, 0.0, 0.0}, xh{0, 0, 0}, vh{0, 0, 0};
  const double* es = p.ep_state + jo * 60;
  double dtv = 0.0;
  if (ok) {
    const double* sn = p.ep_sun + jo * 6;
    xh = Vec3{es[54] - sn[0], es[55] - sn[1], es[56] - sn[2]};
    vh = Vec3{es[57] - sn[3], es[58] - sn[4], es[59] - sn[5]};
    ob = Vec3{p.obs[0][jo], p.obs[1][jo], p.obs[2][jo]};
    if (ecl) ob = ecl_to_equ(ob);
    const Vec3 rho0{xh.x - ob.x, xh.y - ob.y, xh.z - ob.z};
    dtv = -(sqrt(dot(rho0, rho0)) * C_AU_PER_DAY_INV);
    if (analytic_2_body(dtv, xh, vh, po, vo) != ST_OK) {
      ok = false;
      atomicMin(&p.lt_fail[o], j);
    }
  }
  const Vec3 r{po.x - ob.x, po.y - ob.y, po.z - ob.z};
  const double rxy2 = r.x * r.x + r.y * r.y;
  const double rn2 = rxy2 + r.z * r.z;
  const double rn = sqrt(rn2);
  double ra = nan, dec = nan;
  if (ok) {
    ra = 0.0;
    dec = 0.0;
    if (!(rn < 1e-10)) {
      const double zr = fmin(fmax(r.z / rn, -1.0), 1.0);
      dec = rem_tau(3.0 * (3.141592653589793 / 2.0) - acos(zr)) - 3.141592653589793;
      ra = rem_tau(atan2(r.y, r.x));
    }
  }
  p.radec[0][jo] = ra;
  p.radec[1][jo] = dec;
  if (p.state[0] != nullptr) {
    Vec3 so = ok ? xh : Vec3{nan, nan, nan}, sv = ok ? vh : Vec3{nan, nan, nan};
    if (ecl) {
      so = equ_to_ecl(so);
      sv = equ_to_ecl(sv);
    }
    p.state[0][jo] = so.x;
    p.state[1][jo] = so.y;
    p.state[2][jo] = so.z;
    p.state[3][jo] = sv.x;
    p.state[4][jo] = sv.y;
    p.state[5][jo] = sv.z;
  }
  // the light-time step's derivative: Phi(dt) (third order) on the state
  // derivative plus the end velocity times the derivative of dt
  const double r2h = dot(xh, xh);
  const double irh = 1.0 / sqrt(r2h);
  const double g0 = GMS * irh * irh * irh;
  const Vec3 rho0{xh.x - ob.x, xh.y - ob.y, xh.z - ob.z};
  const double irho = 1.0 / sqrt(dot(rho0, rho0));
  const double t = dtv, t2 = t * t;
#pragma unroll 1
  for (int k = 0; k < 9; ++k) {
    const int64_t jk = ((int64_t)k * m + j) * n + o;
    double dra = nan, ddec = nan;
    Vec3 dxo{nan, nan, nan}, dvo{nan, nan, nan};
    if (ok) {
      const Vec3 dx{es[6 * k], es[6 * k + 1], es[6 * k + 2]};
      const Vec3 dv{es[6 * k + 3], es[6 * k + 4], es[6 * k + 5]};
      const double ddt = -(dot(rho0, dx) * irho) * C_AU_PER_DAY_INV;
      const double rdx = 3.0 * dot(xh, dx) / r2h, rdv = 3.0 * dot(xh, dv) / r2h;
      const Vec3 gdx{g0 * (xh.x * rdx - dx.x), g0 * (xh.y * rdx - dx.y), g0 * (xh.z * rdx - dx.z)};
      const Vec3 gdv{g0 * (xh.x * rdv - dv.x), g0 * (xh.y * rdv - dv.y), g0 * (xh.z * rdv - dv.z)};
      const Vec3 dr{dx.x + 0.5 * t2 * gdx.x + t * (dv.x + t2 * gdv.x / 6.0) + vo.x * ddt,
                    dx.y + 0.5 * t2 * gdx.y + t * (dv.y + t2 * gdv.y / 6.0) + vo.y * ddt,
                    dx.z + 0.5 * t2 * gdx.z + t * (dv.z + t2 * gdv.z / 6.0) + vo.z * ddt};
      dra = (r.x * dr.y - r.y * dr.x) / rxy2;
      ddec = (dr.z * rxy2 - r.z * (r.x * dr.x + r.y * dr.y)) / (rn2 * sqrt(rxy2));
      dxo = ecl ? equ_to_ecl(dx) : dx;
      dvo = ecl ? equ_to_ecl(dv) : dv;
    }
    p.jac[0][jk] = dra;
    p.jac[1][jk] = ddec;
    if (p.dstate[0] != nullptr) {
      p.dstate[0][jk] = dxo.x;
      p.dstate[1][jk] = dxo.y;
      p.dstate[2][jk] = dxo.z;
      p.dstate[3][jk] = dvo.x;
      p.dstate[4][jk] = dvo.y;
      p.dstate[5][jk] = dvo.z;
    }
  }
}

// Per object: a failed light-time step is the object's error (the analytic
// two-body error, as a propagate_two_body call would return it) and leaves
// its epoch and the later ones undefined.
__global__ void partials_finish_kernel(GradArgs p) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= p.n) return;
  const int32_t f = p.lt_fail[o];
  if (f >= p.ep_done[o]) return;
  if (p.status[o] == ST_OK) p.status[o] = ST_CONVERGENCE;
  const double nan = __builtin_nan("");
  const int64_t n = p.n;
  const int32_t m = p.m;
  for (int32_t j = f; j < m; ++j) {
    const int64_t jo = (int64_t)j * n + o;
    p.radec[0][jo] = nan;
    p.radec[1][jo] = nan;
    if (p.state[0] != nullptr)
      for (int c = 0; c < 6; ++c) p.state[c][jo] = nan;
    for (int k = 0; k < 9; ++k) {
      const int64_t jk = ((int64_t)k * m + j) * n + o;
      p.jac[0][jk] = nan;
      p.jac[1][jk] = nan;
      if (p.dstate[0] != nullptr)
        for (int c = 0; c < 6; ++c) p.dstate[c][jk] = nan;
    }
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
  if (g.n <= 0) return 0;
  // KH_PARTIALS=lane9: the one-lane-per-parameter dual-number kernel
  const char* sel = getenv("KH_PARTIALS");
  if (sel && strcmp(sel, "lane9") == 0) {
    const int64_t lanes = g.n * 9;
    const unsigned blocks = (unsigned)((lanes + 63) / 64);
    hipLaunchKernelGGL(radau_partials_kernel, dim3(blocks), dim3(64), 0, stream, spk, coef, p);
  } else {
    p.ep_state = g.ep_state;
    p.ep_sun = g.ep_sun;
    p.ep_done = g.ep_done;
    p.lt_fail = g.lt_fail;
    const int64_t lanes = g.n * WL;
    const unsigned blocks = (unsigned)((lanes + 255) / 256);
    if (g.force == 1)
      hipLaunchKernelGGL(radau_partials_wide_kernel<true>, dim3(blocks), dim3(256), 0, stream,
                         spk, coef, p);
    else
      hipLaunchKernelGGL(radau_partials_wide_kernel<false>, dim3(blocks), dim3(256), 0, stream,
                         spk, coef, p);
    const int64_t mn = (int64_t)g.m * g.n;
    if (mn > 0)
      hipLaunchKernelGGL(partials_observe_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256),
                         0, stream, p);
    hipLaunchKernelGGL(partials_finish_kernel, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0,
                       stream, p);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace kh
