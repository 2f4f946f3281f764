// kh_jac.h -- the terms of GravParams::add_acceleration
// (src/kete_core/src/constants/gravity.rs:171-253) for one body together with
// their Jacobians with respect to the object's position and velocity, for the
// forward-mode partials of the orbit-fit path.
//
// Notation: rp, rv are the object's position and velocity relative to the
// body (barycentric equatorial, AU and AU/day); m the body's GM.  Matrices
// are row-major 3x3, J[3 i + j] = d a_i / d x_j.
//   Newton:  a = -m rp / r^3,  d a / d rp = (m / r^3) (3 rp rp^T / r^2 - I).
//   GR (apply_gr_correction, gravity.rs:239-253): a = g (c rp + s rv) with
//            g = m / (C^2 r^3), c = 4 m / r - rv.rv, s = 4 rp.rv;
//            d a / d rp = (c rp + s rv)(-3 g rp / r^2)^T
//                         + g (c I - 4 m rp rp^T / r^3 + 4 rv rv^T),
//            d a / d rv = g (-2 rp rv^T + s I + 4 rv rp^T).
//   J2 (j2_correction, gravity.rs:223-235) in the body's frame p = M rp:
//            a_i = k p_i w_i / r^5, w = (z5 - 1, z5 - 1, z5 - 3),
//            z5 = 5 p_z^2 / r^2, k = 1.5 J2 m R^2;
//            d a_i / d p_j = (k / r^5) (w_i delta_ij - 5 p_i w_i p_j / r^2
//                                       + p_i (10 p_z / r^2)(delta_jz - p_z p_j / r^2)),
//            and in equatorial coordinates M^T (d a / d p) M.
// Every expression keeps the operation order of kete's code for the value.
#pragma once
#include <hip/hip_runtime.h>

namespace kh {

struct BodyTerms {
  double a[3];
  double jr[9];
  double jv[9];
};

// M: 0 identity (Earth's J2, equatorial), 1 equatorial -> ecliptic (the Sun's
// and Jupiter's J2, gravity.rs:184-201).  rot_c, rot_s: cos, sin of the
// obliquity.
__host__ __device__ inline void body_terms(double mass, double radius, double j2, int gr,
                                           int ecliptic_j2, double rot_c, double rot_s,
                                           double c_inv2, const double rp[3],
                                           const double rv[3], double inv_r, BodyTerms& o) {
  const double r2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
  const double ir2 = inv_r * inv_r;
  const double inv_r3 = ir2 * inv_r;
  // Newton
  const double kn = mass * inv_r3;
  for (int i = 0; i < 3; ++i) {
    o.a[i] = -(rp[i] * kn);
    for (int j = 0; j < 3; ++j)
      o.jr[3 * i + j] = kn * (3.0 * rp[i] * rp[j] * ir2 - (i == j ? 1.0 : 0.0));
    for (int j = 0; j < 3; ++j) o.jv[3 * i + j] = 0.0;
  }
  (void)r2;
  if (gr) {
    const double r_v = 4.0 * (rp[0] * rv[0] + rp[1] * rv[1] + rp[2] * rv[2]);
    const double v2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    const double g = mass * c_inv2 * inv_r3;
    const double c = 4. * mass * inv_r - v2;
    double u[3];
    for (int i = 0; i < 3; ++i) {
      u[i] = c * rp[i] + r_v * rv[i];
      o.a[i] += g * u[i];
    }
    const double dg = -3.0 * g * ir2;
    const double dc = -4.0 * mass * inv_r3;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double d = i == j ? 1.0 : 0.0;
        o.jr[3 * i + j] += u[i] * dg * rp[j] + g * (c * d + rp[i] * dc * rp[j] + 4.0 * rv[i] * rv[j]);
        o.jv[3 * i + j] += g * (-2.0 * rp[i] * rv[j] + r_v * d + 4.0 * rv[i] * rp[j]);
      }
  }
  if (j2 != 0.0) {
    // p = M rp
    double p[3];
    if (ecliptic_j2) {
      p[0] = rp[0];
      p[1] = rot_c * rp[1] + rot_s * rp[2];
      p[2] = (-rot_s) * rp[1] + rot_c * rp[2];
    } else {
      p[0] = rp[0];
      p[1] = rp[1];
      p[2] = rp[2];
    }
    const double zr = p[2] * inv_r;
    const double z5 = 5.0 * (zr * zr);
    const double rr = radius * inv_r;
    const double coef = 1.5 * j2 * mass * (rr * rr) * inv_r3;
    const double w[3] = {z5 - 1.0, z5 - 1.0, z5 - 3.0};
    double ab[3], jb[9];
    for (int i = 0; i < 3; ++i) ab[i] = p[i] * coef * w[i];
    const double dz = 10.0 * p[2] * ir2;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double d = i == j ? 1.0 : 0.0;
        const double dzj = dz * ((j == 2 ? 1.0 : 0.0) - p[2] * p[j] * ir2);
        jb[3 * i + j] = coef * (w[i] * d - 5.0 * p[i] * w[i] * p[j] * ir2 + p[i] * dzj);
      }
    if (ecliptic_j2) {
      // a = M^T ab, J = M^T jb M; M = [[1,0,0],[0,c,s],[0,-s,c]]
      const double M[9] = {1.0, 0.0, 0.0, 0.0, rot_c, rot_s, 0.0, -rot_s, rot_c};
      double t[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          t[3 * i + j] = jb[3 * i] * M[j] + jb[3 * i + 1] * M[3 + j] + jb[3 * i + 2] * M[6 + j];
      for (int i = 0; i < 3; ++i) {
        o.a[i] += M[i] * ab[0] + M[3 + i] * ab[1] + M[6 + i] * ab[2];
        for (int j = 0; j < 3; ++j)
          o.jr[3 * i + j] += M[i] * t[j] + M[3 + i] * t[3 + j] + M[6 + i] * t[6 + j];
      }
    } else {
      for (int i = 0; i < 3; ++i) {
        o.a[i] += ab[i];
        for (int j = 0; j < 3; ++j) o.jr[3 * i + j] += jb[3 * i + j];
      }
    }
  }
}

// The terms whose Jacobians are symmetric (Newton, J2: gradients of a
// potential) with their Jacobian's six distinct entries js = (xx, xy, xz,
// yy, yz, zz), plus the value of the GR term (gr != 0), whose derivative the
// caller forms along its own direction (gr_dir).
__host__ __device__ inline void body_terms_sym(double mass, double radius, double j2, int gr,
                                               int ecliptic_j2, double rot_c, double rot_s,
                                               double c_inv2, const double rp[3],
                                               const double rv[3], double inv_r, double a[3],
                                               double js[6]) {
  const double ir2 = inv_r * inv_r;
  const double inv_r3 = ir2 * inv_r;
  const double kn = mass * inv_r3;
  a[0] = -(rp[0] * kn);
  a[1] = -(rp[1] * kn);
  a[2] = -(rp[2] * kn);
  const double k3 = 3.0 * kn * ir2;
  js[0] = k3 * rp[0] * rp[0] - kn;
  js[1] = k3 * rp[0] * rp[1];
  js[2] = k3 * rp[0] * rp[2];
  js[3] = k3 * rp[1] * rp[1] - kn;
  js[4] = k3 * rp[1] * rp[2];
  js[5] = k3 * rp[2] * rp[2] - kn;
  if (gr) {
    const double r_v = 4.0 * (rp[0] * rv[0] + rp[1] * rv[1] + rp[2] * rv[2]);
    const double v2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
    const double g = mass * c_inv2 * inv_r3;
    const double c = 4. * mass * inv_r - v2;
    for (int i = 0; i < 3; ++i) a[i] += g * (c * rp[i] + r_v * rv[i]);
  }
  if (j2 != 0.0) {
    double p[3];
    if (ecliptic_j2) {
      p[0] = rp[0];
      p[1] = rot_c * rp[1] + rot_s * rp[2];
      p[2] = (-rot_s) * rp[1] + rot_c * rp[2];
    } else {
      p[0] = rp[0];
      p[1] = rp[1];
      p[2] = rp[2];
    }
    const double zr = p[2] * inv_r;
    const double z5 = 5.0 * (zr * zr);
    const double rr = radius * inv_r;
    const double coef = 1.5 * j2 * mass * (rr * rr) * inv_r3;
    const double w[3] = {z5 - 1.0, z5 - 1.0, z5 - 3.0};
    const double ab[3] = {p[0] * coef * w[0], p[1] * coef * w[1], p[2] * coef * w[2]};
    const double dz = 10.0 * p[2] * ir2;
    // the body-frame Jacobian (symmetric), entries (i, j) with i <= j
    double jb[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        const double d = i == j ? 1.0 : 0.0;
        const double dzj = dz * ((j == 2 ? 1.0 : 0.0) - p[2] * p[j] * ir2);
        jb[i][j] = jb[j][i] = coef * (w[i] * d - 5.0 * p[i] * w[i] * p[j] * ir2 + p[i] * dzj);
      }
    if (ecliptic_j2) {
      // a += M^T ab, J += M^T jb M, M = [[1,0,0],[0,c,s],[0,-s,c]]
      a[0] += ab[0];
      a[1] += rot_c * ab[1] + (-rot_s) * ab[2];
      a[2] += rot_s * ab[1] + rot_c * ab[2];
      // (jb M): columns 1, 2 mix
      double t[3][3];
      for (int i = 0; i < 3; ++i) {
        t[i][0] = jb[i][0];
        t[i][1] = rot_c * jb[i][1] + (-rot_s) * jb[i][2];
        t[i][2] = rot_s * jb[i][1] + rot_c * jb[i][2];
      }
      // M^T (jb M): rows 1, 2 mix
      const double j00 = t[0][0], j01 = t[0][1], j02 = t[0][2];
      const double j11 = rot_c * t[1][1] + (-rot_s) * t[2][1];
      const double j12 = rot_c * t[1][2] + (-rot_s) * t[2][2];
      const double j22 = rot_s * t[1][2] + rot_c * t[2][2];
      js[0] += j00;
      js[1] += j01;
      js[2] += j02;
      js[3] += j11;
      js[4] += j12;
      js[5] += j22;
    } else {
      a[0] += ab[0];
      a[1] += ab[1];
      a[2] += ab[2];
      js[0] += jb[0][0];
      js[1] += jb[0][1];
      js[2] += jb[0][2];
      js[3] += jb[1][1];
      js[4] += jb[1][2];
      js[5] += jb[2][2];
    }
  }
}

// Derivative of the GR term (gravity.rs:239-253) of a body of GM `mass` at
// relative position rp, velocity rv along (dp, dv).
__host__ __device__ inline void gr_dir(double mass, double c_inv2, const double rp[3],
                                       const double rv[3], const double dp[3],
                                       const double dv[3], double inv_r, double da[3]) {
  const double ir2 = inv_r * inv_r;
  const double inv_r3 = ir2 * inv_r;
  const double rdp = rp[0] * dp[0] + rp[1] * dp[1] + rp[2] * dp[2];
  const double r_v = 4.0 * (rp[0] * rv[0] + rp[1] * rv[1] + rp[2] * rv[2]);
  const double dr_v = 4.0 * (dp[0] * rv[0] + dp[1] * rv[1] + dp[2] * rv[2] + rp[0] * dv[0] +
                             rp[1] * dv[1] + rp[2] * dv[2]);
  const double v2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
  const double dv2 = 2.0 * (rv[0] * dv[0] + rv[1] * dv[1] + rv[2] * dv[2]);
  const double g = mass * c_inv2 * inv_r3;
  const double dg = -3.0 * g * rdp * ir2;
  const double c = 4. * mass * inv_r - v2;
  const double dc = -4.0 * mass * rdp * inv_r3 - dv2;
  for (int i = 0; i < 3; ++i)
    da[i] = dg * (c * rp[i] + r_v * rv[i]) + g * (dc * rp[i] + c * dp[i] + dr_v * rv[i] + r_v * dv[i]);
}

}  // namespace kh
