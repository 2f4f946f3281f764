// kh_kepler.hip -- batched universal-variable two-body propagation and SPK
// state evaluation on gfx950.
//
// Two-body: one lane per (target epoch, object) pair, object index fastest so
// that state loads and result stores coalesce.  The lane runs kete's
// propagation_kepler_py loop body (src/kete/rust/kepler.rs:77-107): center to
// the Sun, analytic_2_body (src/kete_core/src/propagation/kepler.rs:254-297),
// the optional one-step light delay (kepler.rs:89-102 of the binding), and
// center back.  The solver is transcendental-bound (sin/cos per Newton
// iteration); the output (48 B per pair) is the only HBM stream.
#include "kh_device.h"
#include "kh_internal.h"

namespace kh {

// Two waves per SIMD: left to itself the compiler may take 256 VGPRs and
// run one wave per SIMD; held to two it fits without spilling, and the second
// wave hides the transcendental latency of the Newton iterations.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
two_body_kernel(const SpkDev* __restrict__ spk, const double* __restrict__ coef, TwoBodyLaunch p) {
  const int64_t total = p.n * p.m;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const double nan = __builtin_nan("");
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    const int64_t e = q / p.n;
    const int64_t i = q - e * p.n;
    const double jd0 = p.jd0[i];
    const double jd = p.targets[e];
    Vec3 x{p.xyz[0][i], p.xyz[1][i], p.xyz[2][i]};
    Vec3 v{p.xyz[3][i], p.xyz[4][i], p.xyz[5][i]};
    int32_t st = ST_OK;
    const bool ecl = p.frame == 17;
    if (!p.center_sun) {
      // SSB -> Sun: add the flipped Sun state (spk.rs:104-107, state.rs:262-280)
      Vec3 sp, sv;
      st = seg_pos_vel(spk, coef, p.sun_seg, jd0, sp, sv);
      if (ecl) {
        sp = equ_to_ecl(sp);
        sv = equ_to_ecl(sv);
      }
      x = Vec3{x.x + (-sp.x), x.y + (-sp.y), x.z + (-sp.z)};
      v = Vec3{v.x + (-sv.x), v.y + (-sv.y), v.z + (-sv.z)};
    }
    Vec3 po, vo;
    int it0 = 0, it1 = 0;
    if (st == ST_OK) st = analytic_2_body(jd - jd0, x, v, po, vo, &it0);
    if (st == ST_OK && p.has_observer) {
      Vec3 d{po.x - p.observer[0], po.y - p.observer[1], po.z - p.observer[2]};
      double delay = -norm(d) / C_AU_PER_DAY;
      Vec3 p2, v2;
      st = analytic_2_body((jd + delay) - jd, po, vo, p2, v2, &it1);
      po = p2;
      vo = v2;
    }
    if (p.iters) p.iters[q] = it0 + it1;
    if (st == ST_OK && !p.center_sun) {
      // Sun -> SSB at the new epoch (spk.rs:102-103)
      double jd_new = jd;
      Vec3 sp, sv;
      st = seg_pos_vel(spk, coef, p.sun_seg, jd_new, sp, sv);
      if (ecl) {
        sp = equ_to_ecl(sp);
        sv = equ_to_ecl(sv);
      }
      po = Vec3{po.x + sp.x, po.y + sp.y, po.z + sp.z};
      vo = Vec3{vo.x + sv.x, vo.y + sv.y, vo.z + sv.z};
    }
    if (st != ST_OK) po = vo = Vec3{nan, nan, nan};
    p.out[0][q] = po.x;
    p.out[1][q] = po.y;
    p.out[2][q] = po.z;
    p.out[3][q] = vo.x;
    p.out[4][q] = vo.y;
    p.out[5][q] = vo.z;
    p.status[q] = st;
  }
}

int32_t launch_two_body(const SpkDev* spk, const double* coef, const TwoBodyLaunch& p,
                        hipStream_t stream) {
  const int64_t total = p.n * p.m;
  if (total <= 0) return 0;
  const int block = 256;
  int64_t grid = (total + block - 1) / block;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(two_body_kernel, dim3((unsigned)grid), dim3(block), 0, stream, spk, coef,
                     p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void gather_states_kernel(int64_t m, int64_t stride, const double* __restrict__ jd0,
                                     const double* x0, const double* x1, const double* x2,
                                     const double* x3, const double* x4, const double* x5,
                                     const double* __restrict__ jf, int32_t jf_stride,
                                     double* o_jd0, double* y0, double* y1, double* y2,
                                     double* y3, double* y4, double* y5, double* o_jf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t j = i * stride;
  o_jd0[i] = jd0[j];
  y0[i] = x0[j];
  y1[i] = x1[j];
  y2[i] = x2[j];
  y3[i] = x3[j];
  y4[i] = x4[j];
  y5[i] = x5[j];
  o_jf[i] = jf[jf_stride ? j : 0];
}

int32_t launch_gather_states(int64_t m, int64_t stride, const double* jd0,
                             const double* const xyz[6], const double* jd_final,
                             int32_t jf_stride, double* o_jd0, double* const o_xyz[6],
                             double* o_jf, hipStream_t stream) {
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gather_states_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                     stream, m, stride, jd0, xyz[0], xyz[1], xyz[2], xyz[3], xyz[4], xyz[5],
                     jd_final, jf_stride, o_jd0, o_xyz[0], o_xyz[1], o_xyz[2], o_xyz[3],
                     o_xyz[4], o_xyz[5], o_jf);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// State of a body chain (or a Type 21 object) relative to a center chain,
// rotated to `frame`.
__global__ void spk_states_kernel(const SpkDev* __restrict__ spk,
                                  const double* __restrict__ coef, Chain body, Chain center,
                                  int32_t t21_begin, int32_t t21_end, Chain sun,
                                  int64_t center_id, int32_t frame, int64_t m, const double* jd, double* o0, double* o1, double* o2,
                                  double* o3, double* o4, double* o5, int32_t* status) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  Vec3 bp, bv, cp, cv;
  int32_t st;
  bool same_center = false;
  if (t21_begin < t21_end) {
    bool sun_centred = false;
    st = t21_state<true>(spk, t21_begin, t21_end, jd[q], bp, bv, sun_centred);
    // kete changes the center only when it differs from the segment's
    // (SpkCollection::try_get_state, spk.rs:87-96)
    same_center = (sun_centred ? 10 : 0) == center_id;
    if (st == ST_OK && sun_centred && !same_center) {
      Vec3 sp, sv;
      st = chain_state(spk, coef, sun, jd[q], sp, sv);
      bp = Vec3{bp.x + sp.x, bp.y + sp.y, bp.z + sp.z};
      bv = Vec3{bv.x + sv.x, bv.y + sv.y, bv.z + sv.z};
    }
  } else {
    st = chain_state(spk, coef, body, jd[q], bp, bv);
  }
  if (same_center) {
    cp = cv = Vec3{0.0, 0.0, 0.0};
  } else if (st == ST_OK) {
    st = chain_state(spk, coef, center, jd[q], cp, cv);
  }
  Vec3 p{bp.x - cp.x, bp.y - cp.y, bp.z - cp.z};
  Vec3 v{bv.x - cv.x, bv.y - cv.y, bv.z - cv.z};
  if (frame == 17) {
    p = equ_to_ecl(p);
    v = equ_to_ecl(v);
  }
  const double nan = __builtin_nan("");
  if (st != ST_OK) p = v = Vec3{nan, nan, nan};
  o0[q] = p.x;
  o1[q] = p.y;
  o2[q] = p.z;
  o3[q] = v.x;
  o4[q] = v.y;
  o5[q] = v.z;
  status[q] = st;
}

int32_t launch_spk_states(const SpkDev* spk, const double* coef, Chain body, Chain center,
                          int32_t t21_begin, int32_t t21_end, Chain sun, int64_t center_id,
                          int32_t frame, int64_t m, const double* jd, double* const out[6],
                          int32_t* status, hipStream_t stream) {
  if (m <= 0) return 0;
  const int block = 256;
  hipLaunchKernelGGL(spk_states_kernel, dim3((unsigned)((m + block - 1) / block)), dim3(block), 0,
                     stream, spk, coef, body, center, t21_begin, t21_end, sun, center_id, frame, m,
                     jd,
                     out[0], out[1], out[2], out[3], out[4], out[5], status);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace kh
