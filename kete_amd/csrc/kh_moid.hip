// kh_moid.hip -- kete's MOID (minimum orbit intersection distance) for a batch
// of orbit pairs on gfx950 (propagation/kepler.rs:325-404).
//
// Per pair kete moves both states to their perihelion, propagates each to 100
// epochs a period/50 apart (6 days for open orbits), takes the closest of the
// 100 x 100 position pairs, and runs Nelder-Mead in (dt_a, dt_b) from there.
// Two kernels:
//  - moid_grid_kernel: one wavefront per pair.  The 200 grid propagations are
//    spread over the lanes, the positions go to LDS (x, y, z planes, so the
//    lanes of a row read consecutive doubles), and the 10,000 distances are
//    scanned in row-major order, each lane taking every 64th; the pair kete
//    keeps (the first strict minimum) is the smallest (distance, index) of
//    the lanes' candidates.
//  - moid_nm_kernel: one lane per pair runs the Nelder-Mead iterations (two
//    Kepler solves per cost).  An iteration that leaves the simplex unchanged
//    (a failed contraction) repeats identically until the iteration limit, so
//    the lane stops there with the same result.
#include "kh_device.h"
#include "kh_internal.h"

namespace kh {

constexpr int MOID_N = 100;  // grid epochs per orbit (idx = -50 .. 49)

// nalgebra Vector::angle: acos of the clamped cosine, 0 for a zero vector
__device__ __forceinline__ double vec_angle(Vec3 a, Vec3 b) {
  const double prod = dot(a, b), n1 = norm(a), n2 = norm(b);
  if (n1 == 0.0 || n2 == 0.0) return 0.0;
  const double c = prod / (n1 * n2);
  return c > 1.0 ? 0.0 : (c < -1.0 ? M_PI : acos(c));
}

__device__ __forceinline__ Vec3 cross(Vec3 a, Vec3 b) {
  return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// CometElements::from_pos_vel (elements.rs:62-163): the perihelion time, and
// orbital_period (elements.rs:278-293): infinite for a <= 1e-8 (open orbits),
// NaN for parabolic ones (both take the 6-day grid step).
__device__ void comet_peri(double epoch, Vec3 pos, Vec3 vel, double& peri_time,
                           double& period) {
  const Vec3 vs{vel.x / GMS_SQRT, vel.y / GMS_SQRT, vel.z / GMS_SQRT};
  const double v_mag2 = dot(vs, vs);
  const double p_mag = norm(pos);
  const double vp_mag = dot(pos, vs);
  const double ef = v_mag2 - 1.0 / p_mag;
  const Vec3 ev{ef * pos.x - vp_mag * vs.x, ef * pos.y - vp_mag * vs.y,
                ef * pos.z - vp_mag * vs.z};
  const Vec3 ang = cross(pos, vs);
  Vec3 lon{-ang.y, ang.x, 0.0};
  const double ecc = norm(ev), ang_mag = norm(ang), lon_mag = norm(lon);
  const double peri_dist = (ang_mag * ang_mag) / (1.0 + ecc);
  if (lon_mag < 1e-8) lon = Vec3{1.0, 0.0, 0.0};
  double pt;
  if (fabs(ecc - 1.0) < 1e-3) {  // PARABOLIC_ECC_LIMIT
    double ta = vec_angle(ev, pos);
    if (signbit(vp_mag)) ta = -ta;
    const double d = tan(ta / 2.0);
    pt = epoch - (M_SQRT2 * pow(peri_dist, 1.5) / GMS_SQRT) * (d + (d * d * d) / 3.0);
  } else if (ecc < 1e-6) {
    const double a = 1.0 / (2.0 / p_mag - v_mag2);
    const double mm = pow(fabs(a), -1.5) * GMS_SQRT;
    double ta = vec_angle(lon, pos);
    if (signbit(dot(cross(ang, lon), pos))) ta = -ta;
    pt = epoch - ta / mm;
  } else {
    const double a = 1.0 / (2.0 / p_mag - v_mag2);
    const double mm = pow(fabs(a), -1.5) * GMS_SQRT;
    const double x_bar = (ang_mag * ang_mag - p_mag) / ecc;
    const double y_bar = vp_mag / ecc * ang_mag;
    const double b = a * sqrt(fabs(1.0 - ecc * ecc));
    const double s_e = y_bar / b;
    const double m = ecc < 1.0 ? atan2(s_e, x_bar / a + ecc) - ecc * s_e
                               : -ecc * s_e - asinh(-s_e);
    pt = epoch - m / mm;
  }
  peri_time = pt;
  const double sm = fabs(ecc - 1.0) <= 1e-3 ? __builtin_nan("") : peri_dist / (1.0 - ecc);
  period = sm <= 1e-8 ? __builtin_inf() : 2.0 * M_PI * pow(sm, 1.5) / GMS_SQRT;
}

struct MoidState {
  double jd;
  Vec3 p, v;
};

// propagate_two_body (kepler.rs:306-322) to jd_final
__device__ __forceinline__ int32_t two_body_to(const MoidState& s, double jd_final,
                                               MoidState& o) {
  Vec3 po, vo;
  const int32_t st = analytic_2_body(jd_final - s.jd, s.p, s.v, po, vo);
  o = MoidState{jd_final, po, vo};
  return st;
}

__device__ __forceinline__ void load_pair(const MoidLaunch& a, int64_t i, MoidState& sa,
                                          MoidState& sb) {
  sa = MoidState{a.jd_a[i], Vec3{a.xa[0][i], a.xa[1][i], a.xa[2][i]},
                 Vec3{a.xa[3][i], a.xa[4][i], a.xa[5][i]}};
  sb = MoidState{a.jd_b ? a.jd_b[i] : a.jd_a[i], Vec3{a.xb[0][i], a.xb[1][i], a.xb[2][i]},
                 Vec3{a.xb[3][i], a.xb[4][i], a.xb[5][i]}};
  if (a.frame == 1) {  // State::try_change_frame_mut(Ecliptic)
    sa.p = equ_to_ecl(sa.p);
    sa.v = equ_to_ecl(sa.v);
    sb.p = equ_to_ecl(sb.p);
    sb.v = equ_to_ecl(sb.v);
  }
}

struct MoidLds {
  double ax[MOID_N], ay[MOID_N], az[MOID_N];
  double bx[MOID_N], by[MOID_N], bz[MOID_N];
};

__global__ void __launch_bounds__(256) moid_grid_kernel(MoidLaunch a) {
  __shared__ MoidLds lds[4];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 4 + w;  // this wavefront's pair
  const bool valid = i < a.n;
  MoidLds& L = lds[w];
  int32_t st = valid && a.st_b ? a.st_b[i] : ST_OK;
  MoidState pa{0.0, {1.0, 0.0, 0.0}, {0.0, 0.01, 0.0}}, pb = pa;
  if (valid) load_pair(a, i, pa, pb);
  // each state to its perihelion; grid steps from the periods
  double pt_a, per_a, pt_b, per_b;
  comet_peri(pa.jd, pa.p, pa.v, pt_a, per_a);
  if (st == ST_OK) st = two_body_to(pa, pt_a, pa);
  comet_peri(pb.jd, pb.p, pb.v, pt_b, per_b);
  if (st == ST_OK) st = two_body_to(pb, pt_b, pb);
  const double step_a = isfinite(per_a) ? per_a / 50.0 : 300.0 / 50.0;
  const double step_b = isfinite(per_b) ? per_b / 50.0 : 300.0 / 50.0;
  int32_t gst = ST_OK;
  if (st == ST_OK) {
    for (int k = lane; k < MOID_N; k += 64) {
      MoidState s;
      int32_t r = two_body_to(pa, pa.jd + (double)(k - 50) * step_a, s);
      L.ax[k] = s.p.x;
      L.ay[k] = s.p.y;
      L.az[k] = s.p.z;
      if (r && gst == ST_OK) gst = r;
      r = two_body_to(pb, pb.jd + (double)(k - 50) * step_b, s);
      L.bx[k] = s.p.x;
      L.by[k] = s.p.y;
      L.bz[k] = s.p.z;
      if (r && gst == ST_OK) gst = r;
    }
  }
  __syncthreads();
  // kete returns the first grid propagation's error; all of them are
  // Kepler convergence failures
  gst = __any(gst != ST_OK) ? ST_CONVERGENCE : ST_OK;
  if (st == ST_OK) st = gst;
  double best = __builtin_inf();
  int kbest = 0x7fffffff;
  if (st == ST_OK) {
    for (int k = lane; k < MOID_N * MOID_N; k += 64) {
      const int r = k / MOID_N, c = k - r * MOID_N;
      const double dx = L.ax[r] - L.bx[c], dy = L.ay[r] - L.by[c], dz = L.az[r] - L.bz[c];
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (d < best) {
        best = d;
        kbest = k;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int ok = __shfl_xor(kbest, o, 64);
    if (ob < best || (ob == best && ok < kbest)) {
      best = ob;
      kbest = ok;
    }
  }
  if (!valid || lane > 1) return;
  // the pair kept (the perihelion states when no distance compared)
  MoidState s = lane == 0 ? pa : pb;
  if (st == ST_OK && kbest != 0x7fffffff) {
    const int r = kbest / MOID_N, c = kbest - r * MOID_N;
    const int idx = lane == 0 ? r : c;
    const double step = lane == 0 ? step_a : step_b;
    const int32_t e = two_body_to(s, s.jd + (double)(idx - 50) * step, s);
    if (e) st = e;  // recomputes a grid state, which succeeded above
  }
  double* const* out = lane == 0 ? a.best_a : a.best_b;
  out[0][i] = s.jd;
  out[1][i] = s.p.x;
  out[2][i] = s.p.y;
  out[3][i] = s.p.z;
  out[4][i] = s.v.x;
  out[5][i] = s.v.y;
  out[6][i] = s.v.z;
  if (lane == 0) {
    a.status[i] = st;
    if (a.grid_min) a.grid_min[i] = st == ST_OK ? best : __builtin_nan("");
  }
}

// MoidCost::cost (kepler.rs:333-344); the two propagations share one solve
// site (a loop of two trips) to keep the kernel's registers down
__device__ __forceinline__ int32_t moid_cost(const MoidState& sa, const MoidState& sb,
                                             double pa, double pb, double& c) {
  Vec3 x[2];
  int32_t st = ST_OK;
#pragma unroll 1
  for (int k = 0; k < 2; ++k) {
    const MoidState& s = k == 0 ? sa : sb;
    const double dt = ((k == 0 ? pa : pb) + s.jd) - s.jd;
    Vec3 vo;
    const int32_t r = analytic_2_body(dt, s.p, s.v, x[k], vo);
    if (r && st == ST_OK) st = r;
  }
  const Vec3 d{x[0].x - x[1].x, x[0].y - x[1].y, x[0].z - x[1].z};
  c = norm(d);
  if (st == ST_OK && isnan(c)) st = ST_CONVERGENCE;
  return st;
}

struct Vertex {
  double x, y, c;
};

__device__ __forceinline__ void sort3(Vertex& v0, Vertex& v1, Vertex& v2) {
  // stable sort by cost (Rust's sort_by)
  if (v1.c < v0.c) { const Vertex t = v0; v0 = v1; v1 = t; }
  if (v2.c < v1.c) {
    const Vertex t = v1; v1 = v2; v2 = t;
    if (v1.c < v0.c) { const Vertex u = v0; v0 = v1; v1 = u; }
  }
}

// Nelder-Mead of the argmin crate (0.10): alpha 1, gamma 2, rho 0.5; the
// executor stops before an iteration when the sample standard deviation of
// the vertex costs is below f64::EPSILON or 1000 iterations are done.  Each
// trip of the lane's loop evaluates one cost (two Kepler solves) at the point
// its phase needs -- a vertex of the initial simplex, the reflection, the
// expansion or the contraction -- so lanes in different phases share the one
// evaluation site instead of taking divergent branches that each solve.
enum NmPhase : int32_t { NM_INIT0 = 0, NM_INIT1, NM_INIT2, NM_REFLECT, NM_EXPAND, NM_CONTRACT };

__global__ void __launch_bounds__(256) moid_nm_kernel(MoidLaunch a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < a.n;
  int32_t st = valid ? a.status[i] : ST_CONVERGENCE;
  MoidState sa{0.0, {1.0, 0.0, 0.0}, {0.0, 0.01, 0.0}}, sb = sa;
  if (valid && st == ST_OK) {
    sa = MoidState{a.best_a[0][i], Vec3{a.best_a[1][i], a.best_a[2][i], a.best_a[3][i]},
                   Vec3{a.best_a[4][i], a.best_a[5][i], a.best_a[6][i]}};
    sb = MoidState{a.best_b[0][i], Vec3{a.best_b[1][i], a.best_b[2][i], a.best_b[3][i]},
                   Vec3{a.best_b[4][i], a.best_b[5][i], a.best_b[6][i]}};
  }
  Vertex v0{-15.0, -15.0, 0.0}, v1{15.0, -15.0, 0.0}, v2{0.0, 15.0, 0.0};
  double px = v0.x, py = v0.y;  // the point the next evaluation is at
  double x0 = 0.0, y0 = 0.0, xr = 0.0, yr = 0.0, cr = 0.0;
  int32_t phase = NM_INIT0, it = 0;
  bool done = st != ST_OK;
  while (__any(!done)) {
    if (done) continue;
    double c;
    st = moid_cost(sa, sb, px, py, c);
    if (st != ST_OK) {
      done = true;
      continue;
    }
    bool accepted = false;  // an iteration ended with a new simplex
    if (phase == NM_INIT0) {
      v0.c = c;
      phase = NM_INIT1;
      px = v1.x;
      py = v1.y;
    } else if (phase == NM_INIT1) {
      v1.c = c;
      phase = NM_INIT2;
      px = v2.x;
      py = v2.y;
    } else if (phase == NM_INIT2) {
      v2.c = c;
      sort3(v0, v1, v2);
      accepted = true;  // the executor's checks come before the first iteration
      it = -1;
    } else if (phase == NM_REFLECT) {
      cr = c;
      if (cr < v1.c && cr >= v0.c) {
        v2 = Vertex{xr, yr, cr};
        accepted = true;
      } else if (cr < v0.c) {
        phase = NM_EXPAND;
        px = x0 + (xr - x0) * 2.0;
        py = y0 + (yr - y0) * 2.0;
      } else {
        phase = NM_CONTRACT;
        px = x0 + (v2.x - x0) * 0.5;
        py = y0 + (v2.y - y0) * 0.5;
      }
    } else if (phase == NM_EXPAND) {
      v2 = c < cr ? Vertex{px, py, c} : Vertex{xr, yr, cr};
      accepted = true;
    } else {  // NM_CONTRACT
      if (c < v2.c) {
        v2 = Vertex{px, py, c};
        accepted = true;
      } else {
        // a failed contraction leaves the simplex as it was: every later
        // iteration repeats this one up to the limit
        it = a.max_iter;
        done = true;
        continue;
      }
    }
    if (!accepted) continue;
    if (phase != NM_INIT2) sort3(v0, v1, v2);
    ++it;
    const double c0 = (((0.0 + v0.c) + v1.c) + v2.c) / 3.0;
    const double d0 = v0.c - c0, d1 = v1.c - c0, d2 = v2.c - c0;
    const double s = sqrt(0.5 * (((0.0 + d0 * d0) + d1 * d1) + d2 * d2));
    if (s < 2.220446049250313e-16 || it >= a.max_iter) {
      done = true;
      continue;
    }
    x0 = (v0.x + v1.x) * 0.5;
    y0 = (v0.y + v1.y) * 0.5;
    xr = x0 + (x0 - v2.x);
    yr = y0 + (y0 - v2.y);
    phase = NM_REFLECT;
    px = xr;
    py = yr;
  }
  if (!valid) return;
  a.moid[i] = st == ST_OK ? v0.c : __builtin_nan("");
  a.status[i] = st;
  if (a.iters) a.iters[i] = it;
}

int32_t launch_moid(const MoidLaunch& p, hipStream_t stream) {
  if (p.n <= 0) return 0;
  hipLaunchKernelGGL(moid_grid_kernel, dim3((unsigned)((p.n + 3) / 4)), dim3(256), 0, stream, p);
  if (hipGetLastError() != hipSuccess) return -1;
  hipLaunchKernelGGL(moid_nm_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, stream,
                     p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace kh
