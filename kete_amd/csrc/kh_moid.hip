// kh_moid.hip -- kete's MOID (minimum orbit intersection distance) for a batch
// of orbit pairs on gfx950 (propagation/kepler.rs:325-404).
//
// Per pair kete moves both states to their perihelion, propagates each to 100
// epochs a period/50 apart (6 days for open orbits), takes the closest of the
// 100 x 100 position pairs, and runs Nelder-Mead in (dt_a, dt_b) from there.
// Kernels:
//  - moid_earth_grid_kernel: the Earth's perihelion state and grid at the
//    batch's first epoch, once, when the second orbit is the Earth.
//  - moid_grid_kernel: one wavefront per pair.  The perihelion states take
//    one solve round (half the lanes each), the 200 grid propagations are
//    spread over the lanes (100 when the second orbit is the Earth at the
//    batch's first epoch: that grid is computed once), the states go to LDS,
//    and the 10,000 distances are scanned with each lane owning one or two
//    columns of the table; the pair kept (kete's first strict minimum in
//    row-major order) is the smallest (squared distance, index) of the
//    lanes' candidates.
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
                                               MoidState& o, int* newton = nullptr) {
  Vec3 po, vo;
  const int32_t st = analytic_2_body(jd_final - s.jd, s.p, s.v, po, vo, newton);
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

// grid state k (idx = k - 50) of an orbit from its perihelion state
__device__ __forceinline__ int32_t grid_state(const MoidState& peri, double step, int k,
                                              MoidState& s, int* newton = nullptr) {
  return two_body_to(peri, peri.jd + (double)(k - 50) * step, s, newton);
}

// elements -> grid step (orbital_period / 50, or 300 / 50 days)
__device__ __forceinline__ double grid_step(double period) {
  return isfinite(period) ? period / 50.0 : 300.0 / 50.0;
}

// The Earth's grid at the first pair's epoch, shared by every pair whose
// second orbit is the Earth at that epoch (moid_py's default, one epoch for
// the batch): the same states each such pair would compute.
__global__ void __launch_bounds__(64) moid_earth_grid_kernel(MoidLaunch a) {
  MoidSharedGrid* g = a.sg;
  const int lane = threadIdx.x;
  MoidState sa, sb;
  load_pair(a, 0, sa, sb);
  int32_t st = a.st_b ? a.st_b[0] : ST_OK;
  double pt, per;
  comet_peri(sb.jd, sb.p, sb.v, pt, per);
  MoidState peri = sb;
  if (st == ST_OK) st = two_body_to(sb, pt, peri);
  const double step = grid_step(per);
  int32_t gst = ST_OK;
  if (st == ST_OK) {
    for (int k = lane; k < MOID_N; k += 64) {
      MoidState s;
      const int32_t r = grid_state(peri, step, k, s);
      if (r && gst == ST_OK) gst = r;
      g->g[0][k] = s.p.x;
      g->g[1][k] = s.p.y;
      g->g[2][k] = s.p.z;
      g->g[3][k] = s.v.x;
      g->g[4][k] = s.v.y;
      g->g[5][k] = s.v.z;
    }
  }
  gst = __any(gst != ST_OK) ? ST_CONVERGENCE : ST_OK;
  if (lane == 0) {
    g->key = a.jd_a[0];
    g->st = st == ST_OK ? gst : st;
    g->step = step;
    g->peri_jd = peri.jd;
  }
}

struct MoidLds {
  double a[6][MOID_N];  // x y z vx vy vz of orbit a's grid states
  double b[6][MOID_N];
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
  const MoidSharedGrid* sg = a.sg;
  const bool use_sg = sg != nullptr && valid && a.jd_a[i] == sg->key;
  // both perihelion states in one solve round: lanes 0-31 orbit a, 32-63 b
  double pt_a, per_a, pt_b, per_b;
  comet_peri(pa.jd, pa.p, pa.v, pt_a, per_a);
  comet_peri(pb.jd, pb.p, pb.v, pt_b, per_b);
  const bool half_b = lane >= 32;
  MoidState peri = half_b ? pb : pa;
  int it_peri = 0;
  const int32_t pr =
      st == ST_OK ? two_body_to(peri, half_b ? pt_b : pt_a, peri, &it_peri) : ST_OK;
  const MoidState qa{__shfl(peri.jd, 0, 64),
                     Vec3{__shfl(peri.p.x, 0, 64), __shfl(peri.p.y, 0, 64),
                          __shfl(peri.p.z, 0, 64)},
                     Vec3{__shfl(peri.v.x, 0, 64), __shfl(peri.v.y, 0, 64),
                          __shfl(peri.v.z, 0, 64)}};
  const MoidState qb{__shfl(peri.jd, 32, 64),
                     Vec3{__shfl(peri.p.x, 32, 64), __shfl(peri.p.y, 32, 64),
                          __shfl(peri.p.z, 32, 64)},
                     Vec3{__shfl(peri.v.x, 32, 64), __shfl(peri.v.y, 32, 64),
                          __shfl(peri.v.z, 32, 64)}};
  const int32_t st_pa = __shfl(pr, 0, 64), st_pb = __shfl(pr, 32, 64);
  if (st == ST_OK) st = st_pa;
  if (st == ST_OK) st = use_sg ? sg->st : st_pb;
  const double step_a = grid_step(per_a);
  const double step_b = use_sg ? sg->step : grid_step(per_b);
  // grid states: orbit a's 100 (and b's unless shared), one per lane and round
  int32_t gst = ST_OK;
  int n_solve = 0, n_newton = 0;  // work counters of the grid propagations
  if (st == ST_OK) {
    const int nj = use_sg ? MOID_N : 2 * MOID_N;
    for (int j = lane; j < nj; j += 64) {
      const bool ob = j >= MOID_N;
      const int k = ob ? j - MOID_N : j;
      MoidState s;
      int it = 0;
      const int32_t r = grid_state(ob ? qb : qa, ob ? step_b : step_a, k, s, &it);
      ++n_solve;
      n_newton += it;
      if (r && gst == ST_OK) gst = r;
      double(*dst)[MOID_N] = ob ? L.b : L.a;
      dst[0][k] = s.p.x;
      dst[1][k] = s.p.y;
      dst[2][k] = s.p.z;
      dst[3][k] = s.v.x;
      dst[4][k] = s.v.y;
      dst[5][k] = s.v.z;
    }
    if (use_sg)
      for (int k = lane; k < MOID_N; k += 64) {
        L.b[0][k] = sg->g[0][k];
        L.b[1][k] = sg->g[1][k];
        L.b[2][k] = sg->g[2][k];
      }
  }
  __syncthreads();
  // kete returns the first grid propagation's error; all of them are
  // Kepler convergence failures
  gst = __any(gst != ST_OK) ? ST_CONVERGENCE : ST_OK;
  if (st == ST_OK) st = gst;
  // closest grid pair: lane l owns columns l and l + 64 of the 100 x 100
  // table (b's positions in registers) and walks the rows (a's positions,
  // one LDS broadcast per coordinate), so each lane meets its entries in
  // increasing row-major index; squared distances are compared
  double best = __builtin_inf();
  int kbest = 0x7fffffff;
  if (st == ST_OK) {
    const int c0 = lane, c1 = lane + 64;
    const bool two = c1 < MOID_N;
    const double b0x = L.b[0][c0], b0y = L.b[1][c0], b0z = L.b[2][c0];
    const double b1x = two ? L.b[0][c1] : 0.0, b1y = two ? L.b[1][c1] : 0.0,
                 b1z = two ? L.b[2][c1] : 0.0;
    for (int r = 0; r < MOID_N; ++r) {
      const double ax = L.a[0][r], ay = L.a[1][r], az = L.a[2][r];
      const double dx = ax - b0x, dy = ay - b0y, dz = az - b0z;
      const double d0 = dx * dx + dy * dy + dz * dz;
      if (d0 < best) {
        best = d0;
        kbest = r * MOID_N + c0;
      }
      const double ex = ax - b1x, ey = ay - b1y, ez = az - b1z;
      const double d1 = ex * ex + ey * ey + ez * ez;
      if (two && d1 < best) {
        best = d1;
        kbest = r * MOID_N + c1;
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
  best = sqrt(best);
  if (a.work) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      n_solve += __shfl_xor(n_solve, o, 64);
      n_newton += __shfl_xor(n_newton, o, 64);
    }
    // the perihelion solves the pair needs (lane 0: a, lane 32: b)
    const int it_a = __shfl(it_peri, 0, 64), it_b = __shfl(it_peri, 32, 64);
    n_solve += use_sg ? 1 : 2;
    n_newton += it_a + (use_sg ? 0 : it_b);
  }
  if (!valid || lane > 1) return;
  // the pair kept: grid states from LDS (or the shared Earth grid), the
  // perihelion states when no distance compared
  MoidState s = lane == 0 ? qa : qb;
  if (st == ST_OK && kbest != 0x7fffffff) {
    const int r = kbest / MOID_N, c = kbest - r * MOID_N;
    if (lane == 0) {
      s = MoidState{qa.jd + (double)(r - 50) * step_a, Vec3{L.a[0][r], L.a[1][r], L.a[2][r]},
                    Vec3{L.a[3][r], L.a[4][r], L.a[5][r]}};
    } else if (use_sg) {
      s = MoidState{sg->peri_jd + (double)(c - 50) * step_b,
                    Vec3{sg->g[0][c], sg->g[1][c], sg->g[2][c]},
                    Vec3{sg->g[3][c], sg->g[4][c], sg->g[5][c]}};
    } else {
      s = MoidState{qb.jd + (double)(c - 50) * step_b, Vec3{L.b[0][c], L.b[1][c], L.b[2][c]},
                    Vec3{L.b[3][c], L.b[4][c], L.b[5][c]}};
    }
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
    if (a.work) {
      a.work[i] = n_solve;
      a.work[a.n + i] = n_newton;
    }
  }
}

// MoidCost::cost (kepler.rs:333-344); the two propagations share one solve
// site (a loop of two trips) to keep the kernel's registers down
__device__ __forceinline__ int32_t moid_cost(const MoidState& sa, const MoidState& sb,
                                             double pa, double pb, double& c, int& newton) {
  Vec3 x[2];
  int32_t st = ST_OK;
#pragma unroll 1
  for (int k = 0; k < 2; ++k) {
    const MoidState& s = k == 0 ? sa : sb;
    const double dt = ((k == 0 ? pa : pb) + s.jd) - s.jd;
    Vec3 vo;
    int it = 0;
    const int32_t r = analytic_2_body(dt, s.p, s.v, x[k], vo, &it);
    newton += it;
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
  int n_solve = 0, n_newton = 0;
  bool done = st != ST_OK;
  while (__any(!done)) {
    if (done) continue;
    double c;
    st = moid_cost(sa, sb, px, py, c, n_newton);
    n_solve += 2;
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
  if (a.work) {
    a.work[i] += n_solve;
    a.work[a.n + i] += n_newton;
  }
}

int32_t launch_moid(const MoidLaunch& p, hipStream_t stream) {
  if (p.n <= 0) return 0;
  if (p.sg) {
    hipLaunchKernelGGL(moid_earth_grid_kernel, dim3(1), dim3(64), 0, stream, p);
    if (hipGetLastError() != hipSuccess) return -1;
  }
  hipLaunchKernelGGL(moid_grid_kernel, dim3((unsigned)((p.n + 3) / 4)), dim3(256), 0, stream, p);
  if (hipGetLastError() != hipSuccess) return -1;
  hipLaunchKernelGGL(moid_nm_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, stream,
                     p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace kh
