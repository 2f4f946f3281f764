// kh_fov.hip -- field-of-view visibility with light time, and the RA/Dec and
// V-magnitude epilogue, on gfx950.
//
// Restates FovLike::check_visible (src/kete_core/src/fov/fov_like.rs:134-192)
// for rectangular fields (SphericalPolygon::contains, fov/patches.rs:214-234)
// as three data-parallel phases over a group of fields that share one state
// epoch (the grouping of fov_checks_py, src/kete/rust/fovs/checks.rs:24-110):
//  1. prefilter: one lane per object, looping over the group's fields held in
//     LDS; the linear-motion check (fov_like.rs:43-66) against the
//     2|dv|dt_limit margin (:147-150).  Survivors are appended to a candidate
//     list, so the expensive checks run with full wavefronts.
//  2. two-body: one lane per candidate -- two universal-Kepler solves with the
//     light-time step (fov_like.rs:71-85) and the containment test.  Hits are
//     appended with their light-time-corrected state.
//  3. n-body (only for fields more than dt_limit from the state epoch): the
//     candidates are re-propagated with the Radau kernel, then a light-time
//     step and containment (fov_like.rs:90-108) finish them.
#include <hipcub/hipcub.hpp>

#include "kh_device.h"
#include "kh_internal.h"

namespace kh {

// SphericalPolygon::contains: inside iff no edge has a negative (sign bit
// set) dot product; else the largest such |d|, capped by |v|; NaN -> outside.
__device__ __forceinline__ bool rect_contains(const double* __restrict__ nrm, Vec3 v,
                                              double& dist) {
  double closest = -__builtin_inf();
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double d = v.x * nrm[3 * e] + v.y * nrm[3 * e + 1] + v.z * nrm[3 * e + 2];
    if (isnan(d)) {
      dist = d;
      return false;
    }
    if (signbit(d) && fabs(d) > closest) closest = fabs(d);
  }
  if (isfinite(closest)) {
    dist = fmin(closest, norm(v));
    return false;
  }
  dist = 0.0;
  return true;
}

// Candidates collect in a per-block LDS list and reach the global list with
// one atomic per block: a single global counter bumped per wavefront and
// field serialised ~1e5 atomics per launch (5x the kernel's arithmetic).
constexpr int PF_BLOCK_CAND = 1024;

__global__ void __launch_bounds__(256) fov_prefilter_kernel(FovPrefilter p) {
  // the tile's fields (at most FOV_MAX_TILE_FIELDS; dynamic LDS)
  extern __shared__ FieldDev sf[];
  __shared__ int32_t c_field[PF_BLOCK_CAND], c_kind[PF_BLOCK_CAND];
  __shared__ int64_t c_obj[PF_BLOCK_CAND];
  __shared__ uint32_t c_n;
  __shared__ unsigned long long c_base;
  for (int k = threadIdx.x; k < p.nf; k += blockDim.x) sf[k] = p.fields[k];
  if (threadIdx.x == 0) c_n = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool active = i < p.n;
  double jd = 0.0;
  Vec3 x{0, 0, 0}, v{0, 0, 0};
  if (active) {
    jd = p.jd[i];
    x = Vec3{p.x[0][i], p.x[1][i], p.x[2][i]};
    v = Vec3{p.x[3][i], p.x[4][i], p.x[5][i]};
    // objects whose re-propagation failed were dropped by kete (filter_map);
    // their NaN states can never be seen, so they are skipped here.
    active = isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) &&
             isfinite(v.y) && isfinite(v.z) && isfinite(jd);
  }
  for (int f = 0; active && f < p.nf; ++f) {
    // the whole field at once, and no branches below
    const FieldDev F = sf[f];
    const Vec3 dv{v.x - F.obs[3], v.y - F.obs[4], v.z - F.obs[5]};
    const double max_dist = norm(dv) * p.dt_limit * 2.0;
    // check_linear (fov_like.rs:43-66) and SphericalPolygon::contains
    // (patches.rs:214-234), evaluated for every field; only fields within
    // dt_limit of the states' epoch use it (kind 0), the others go to the
    // two-body check unconditionally (kind 1)
    const Vec3 rel{x.x - F.obs[0], x.y - F.obs[1], x.z - F.obs[2]};
    const double dt = F.jd - jd - norm(rel) * C_AU_PER_DAY_INV;
    const Vec3 nr{(x.x + v.x * dt) - F.obs[0], (x.y + v.y * dt) - F.obs[1],
                  (x.z + v.z * dt) - F.obs[2]};
    bool any_nan = false;
    double closest = -__builtin_inf();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double d = nr.x * F.nrm[3 * e] + nr.y * F.nrm[3 * e + 1] + nr.z * F.nrm[3 * e + 2];
      any_nan = any_nan || isnan(d);
      closest = (signbit(d) && fabs(d) > closest) ? fabs(d) : closest;
    }
    // NaN: Outside(NaN), never farther than max_dist; an edge with a
    // negative side: Outside(min(closest, |nr|)); otherwise Inside
    const bool outside_far = !any_nan && isfinite(closest) && fmin(closest, norm(nr)) > max_dist;
    const bool near = fabs(jd - F.jd) < p.dt_limit;
    if (near && outside_far) continue;
    const int32_t kind = near ? 0 : 1;
    const uint32_t k = atomicAdd(&c_n, 1u);
    if (k < (uint32_t)PF_BLOCK_CAND) {
      c_field[k] = f;
      c_kind[k] = kind;
      c_obj[k] = i;
    } else {
      // the block's list is full: straight to the global list
      const unsigned long long slot = atomicAdd(p.n_cand, 1ull);
      if (slot < (unsigned long long)p.cap) {
        p.cand_field[slot] = f;
        p.cand_kind[slot] = kind;
        p.cand_obj[slot] = i;
      }
    }
  }
  __syncthreads();
  const uint32_t nb = c_n < (uint32_t)PF_BLOCK_CAND ? c_n : (uint32_t)PF_BLOCK_CAND;
  if (threadIdx.x == 0) c_base = nb ? atomicAdd(p.n_cand, (unsigned long long)nb) : 0ull;
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) {
    const unsigned long long slot = c_base + k;
    if (slot < (unsigned long long)p.cap) {
      p.cand_field[slot] = c_field[k];
      p.cand_kind[slot] = c_kind[k];
      p.cand_obj[slot] = c_obj[k];
    }
  }
}

// check_two_body (fov_like.rs:71-85) for one candidate; kind 0 -> hits,
// kind 1 -> n-body list when the two-body position is within the margin.
__device__ __forceinline__ void two_body_candidate(const FovTwoBody& p, int64_t q) {
  const int32_t f = p.cand_field[q];
  const int64_t i = p.cand_obj[q];
  const FieldDev& F = p.fields[f];
  const double jd = p.jd[i];
  const Vec3 x{p.x[0][i], p.x[1][i], p.x[2][i]};
  const Vec3 v{p.x[3][i], p.x[4][i], p.x[5][i]};
  Vec3 p1, v1, p2, v2;
  if (analytic_2_body(F.jd - jd, x, v, p1, v1) != ST_OK) return;
  const Vec3 d1{p1.x - F.obs[0], p1.y - F.obs[1], p1.z - F.obs[2]};
  const double dt = -norm(d1) * C_AU_PER_DAY_INV;
  const double jd_obs = F.jd + dt;
  if (analytic_2_body(jd_obs - F.jd, p1, v1, p2, v2) != ST_OK) return;
  const Vec3 rel{p2.x - F.obs[0], p2.y - F.obs[1], p2.z - F.obs[2]};
  double dist;
  const bool inside = rect_contains(F.nrm, rel, dist);
  if (p.cand_kind[q] == 0) {
    if (!inside) return;
    const unsigned long long slot = atomicAdd(p.n_hit, 1ull);
    if (slot < (unsigned long long)p.hit_cap) {
      p.hit_field[slot] = p.field_base + f;
      p.hit_obj[slot] = i;
      p.hit_state[0][slot] = jd_obs;
      p.hit_state[1][slot] = p2.x;
      p.hit_state[2][slot] = p2.y;
      p.hit_state[3][slot] = p2.z;
      p.hit_state[4][slot] = v2.x;
      p.hit_state[5][slot] = v2.y;
      p.hit_state[6][slot] = v2.z;
    }
  } else {
    const Vec3 dvv{v.x - F.obs[3], v.y - F.obs[4], v.z - F.obs[5]};
    const double max_dist = norm(dvv) * p.dt_limit * 2.0;
    if (!inside && dist > max_dist) return;
    const unsigned long long slot = atomicAdd(p.n_nb, 1ull);
    if (slot < (unsigned long long)p.nb_cap) {
      p.nb_field[slot] = p.field_base + f;
      p.nb_obj[slot] = i;
      p.g_jd0[slot] = jd;
      p.g_jdf[slot] = F.jd;
      p.g_x[0][slot] = x.x;
      p.g_x[1][slot] = x.y;
      p.g_x[2][slot] = x.z;
      p.g_x[3][slot] = v.x;
      p.g_x[4][slot] = v.y;
      p.g_x[5][slot] = v.z;
    }
  }
}

// A fixed grid strides over the candidates the prefilter listed: the count is
// read here, on the device, so no host round trip separates the two launches.
// Candidates beyond the list's capacity were not stored; the largest count is
// kept so that the caller can detect the overflow after the last group.
__global__ void __launch_bounds__(256) fov_two_body_kernel(FovTwoBody p) {
  const unsigned long long nc = *p.n_cand;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(p.cand_max, nc);
  const int64_t m = nc < (unsigned long long)p.cand_cap ? (int64_t)nc : p.cand_cap;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += stride)
    two_body_candidate(p, q);
}

// check_n_body tail (fov_like.rs:98-107): light-time step and containment.
__global__ void fov_nb_finish_kernel(FovNBody p) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.m) return;
  if (p.g_status[q] != ST_OK) return;
  const int32_t f = p.nb_field[q];
  const FieldDev& F = p.fields[f];
  const Vec3 x{p.g_out[0][q], p.g_out[1][q], p.g_out[2][q]};
  const Vec3 v{p.g_out[3][q], p.g_out[4][q], p.g_out[5][q]};
  const Vec3 d1{x.x - F.obs[0], x.y - F.obs[1], x.z - F.obs[2]};
  const double dt = -norm(d1) * C_AU_PER_DAY_INV;
  const double jd_obs = F.jd + dt;
  Vec3 p2, v2;
  if (analytic_2_body(jd_obs - F.jd, x, v, p2, v2) != ST_OK) return;
  const Vec3 rel{p2.x - F.obs[0], p2.y - F.obs[1], p2.z - F.obs[2]};
  double dist;
  if (!rect_contains(F.nrm, rel, dist)) return;
  const unsigned long long slot = atomicAdd(p.n_hit, 1ull);
  if (slot < (unsigned long long)p.hit_cap) {
    p.hit_field[slot] = f;
    p.hit_obj[slot] = p.nb_obj[q];
    p.hit_state[0][slot] = jd_obs;
    p.hit_state[1][slot] = p2.x;
    p.hit_state[2][slot] = p2.y;
    p.hit_state[3][slot] = p2.z;
    p.hit_state[4][slot] = v2.x;
    p.hit_state[5][slot] = v2.y;
    p.hit_state[6][slot] = v2.z;
  }
}

// ---------------------------------------------------------------------------
// fov_spk_check: SPK objects in fields (FovLike::check_spks,
// fov/fov_like.rs:196-226): the object's Sun-centred state at the field epoch
// from the ephemeris (try_get_state with the observer's center and frame),
// then check_two_body -- a zero-length two-body step, the light-time step,
// containment.  Objects without a state at the epoch are skipped.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fov_spk_kernel(const SpkDev* __restrict__ spk,
                                                      const double* __restrict__ coef,
                                                      FovSpk p) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.n_obj * p.n_fields) return;
  const int64_t fi = q / p.n_obj, k = q - fi * p.n_obj;
  const SpkObjDesc d = p.obj[k];
  if (d.kind == 0) return;
  const FieldDev& F = p.fields[fi];
  Vec3 pos, vel;
  int32_t st;
  bool sun_centred = false;
  if (d.kind == 1) st = t21_state<true>(spk, d.t21_begin, d.t21_end, F.jd, pos, vel, sun_centred);
  else st = chain_state(spk, coef, d.body, F.jd, pos, vel);
  if (st == ST_OK && !sun_centred) {
    Vec3 sp, sv;
    st = chain_state(spk, coef, p.sun, F.jd, sp, sv);
    pos = Vec3{pos.x - sp.x, pos.y - sp.y, pos.z - sp.z};
    vel = Vec3{vel.x - sv.x, vel.y - sv.y, vel.z - sv.z};
  }
  if (st != ST_OK) return;
  // the fields are ecliptic (kete converts FOVs to it, collection.rs:23-36)
  pos = equ_to_ecl(pos);
  vel = equ_to_ecl(vel);
  Vec3 p1, v1, p2, v2;
  if (analytic_2_body(0.0, pos, vel, p1, v1) != ST_OK) return;
  const Vec3 d1{p1.x - F.obs[0], p1.y - F.obs[1], p1.z - F.obs[2]};
  const double jd_obs = F.jd + (-norm(d1) * C_AU_PER_DAY_INV);
  if (analytic_2_body(jd_obs - F.jd, p1, v1, p2, v2) != ST_OK) return;
  const Vec3 rel{p2.x - F.obs[0], p2.y - F.obs[1], p2.z - F.obs[2]};
  double dist;
  if (!rect_contains(F.nrm, rel, dist)) return;
  const unsigned long long slot = atomicAdd(p.n_hit, 1ull);
  if (slot < (unsigned long long)p.hit_cap) {
    p.hit_field[slot] = (int32_t)fi;
    p.hit_obj[slot] = k;
    p.hit_state[0][slot] = jd_obs;
    p.hit_state[1][slot] = p2.x;
    p.hit_state[2][slot] = p2.y;
    p.hit_state[3][slot] = p2.z;
    p.hit_state[4][slot] = v2.x;
    p.hit_state[5][slot] = v2.y;
    p.hit_state[6][slot] = v2.z;
  }
}

int32_t launch_fov_spk(const SpkDev* spk, const double* coef, const FovSpk& p, hipStream_t s) {
  const int64_t pairs = p.n_obj * p.n_fields;
  if (pairs <= 0) return 0;
  hipLaunchKernelGGL(fov_spk_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s, spk,
                     coef, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Hit ordering: fields in epoch order, then objects by index, as kete returns
// them (the rayon collect over sorted fields, checks.rs:87-104)
// ---------------------------------------------------------------------------
__global__ void fov_hit_keys_kernel(int64_t n, int32_t obj_bits, const int32_t* __restrict__ f,
                                    const int64_t* __restrict__ o, uint64_t* __restrict__ key,
                                    uint32_t* __restrict__ val) {
  const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= n) return;
  key[h] = ((uint64_t)(uint32_t)f[h] << obj_bits) | (uint64_t)o[h];
  val[h] = (uint32_t)h;
}

__global__ void fov_hit_gather_kernel(int64_t n, const uint32_t* __restrict__ perm,
                                      const int32_t* __restrict__ f,
                                      const int64_t* __restrict__ o,
                                      const int64_t* __restrict__ order, const double* s0,
                                      const double* s1, const double* s2, const double* s3,
                                      const double* s4, const double* s5, const double* s6,
                                      int64_t* __restrict__ of, int64_t* __restrict__ oo,
                                      double* d0, double* d1, double* d2, double* d3,
                                      double* d4, double* d5, double* d6) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t h = perm[k];
  of[k] = order[f[h]];
  oo[k] = o[h];
  d0[k] = s0[h];
  d1[k] = s1[h];
  d2[k] = s2[h];
  d3[k] = s3[h];
  d4[k] = s4[h];
  d5[k] = s5[h];
  d6[k] = s6[h];
}

size_t fov_sort_temp_bytes(int64_t n_hits) {
  size_t cub = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, cub, (const uint64_t*)nullptr,
                                           (uint64_t*)nullptr, (const uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)n_hits, 0, 64);
  const size_t a = ((size_t)n_hits * 8 + 255) & ~(size_t)255;
  const size_t b = ((size_t)n_hits * 4 + 255) & ~(size_t)255;
  return 2 * a + 2 * b + cub + 256;
}

int32_t launch_fov_sort_hits(int64_t n_hits, int32_t obj_bits, int32_t field_bits,
                             const int32_t* hit_field, const int64_t* hit_obj,
                             double* const hit_state[7], const int64_t* order, void* temp,
                             size_t temp_bytes, int64_t* out_field, int64_t* out_obj,
                             double* const out_state[7], hipStream_t s) {
  if (n_hits <= 0) return 0;
  if (n_hits >= ((int64_t)1 << 31) || obj_bits + field_bits > 64) return -1;
  const size_t a = ((size_t)n_hits * 8 + 255) & ~(size_t)255;
  const size_t b = ((size_t)n_hits * 4 + 255) & ~(size_t)255;
  char* q = (char*)temp;
  uint64_t* k_in = (uint64_t*)q;
  uint64_t* k_out = (uint64_t*)(q + a);
  uint32_t* v_in = (uint32_t*)(q + 2 * a);
  uint32_t* v_out = (uint32_t*)(q + 2 * a + b);
  void* cub = q + 2 * a + 2 * b;
  size_t cub_bytes = temp_bytes - (2 * a + 2 * b);
  const unsigned g = (unsigned)((n_hits + 255) / 256);
  hipLaunchKernelGGL(fov_hit_keys_kernel, dim3(g), dim3(256), 0, s, n_hits, obj_bits, hit_field,
                     hit_obj, k_in, v_in);
  if (hipcub::DeviceRadixSort::SortPairs(cub, cub_bytes, k_in, k_out, v_in, v_out, (int)n_hits, 0,
                                         obj_bits + field_bits, s) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(fov_hit_gather_kernel, dim3(g), dim3(256), 0, s, n_hits, v_out, hit_field,
                     hit_obj, order, hit_state[0], hit_state[1], hit_state[2], hit_state[3],
                     hit_state[4], hit_state[5], hit_state[6], out_field, out_obj, out_state[0],
                     out_state[1], out_state[2], out_state[3], out_state[4], out_state[5],
                     out_state[6]);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ra_dec_with_rates (simult_states.rs:86-120, to_lat_lon frames/definitions.rs:
// 337-345) and HGParams::apparent_mag (flux/reflected.rs:23-33,255-268).
__device__ __forceinline__ double rem_euclid(double x, double m) {
  const double r = fmod(x, m);
  return r < 0.0 ? r + fabs(m) : r;
}

__device__ __forceinline__ double hg_helper(double a, double b, double c, double phase) {
  const double ps = sin(phase);
  const double th = tan(0.5 * phase);
  const double theta_l = exp(-a * pow(th, b));
  const double theta_s = 1.0 - c * ps / (0.119 + 1.341 * ps - 0.754 * (ps * ps));
  const double w = exp(-90.56 * (th * th));
  return w * theta_s + (1.0 - w) * theta_l;
}

__global__ void radec_vmag_kernel(RadecArgs p) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.n) return;
  Vec3 x{p.st[0][q], p.st[1][q], p.st[2][q]};
  Vec3 v{p.st[3][q], p.st[4][q], p.st[5][q]};
  Vec3 o{p.obs[0][q], p.obs[1][q], p.obs[2][q]};
  Vec3 ov{p.obs[3][q], p.obs[4][q], p.obs[5][q]};
  // V magnitude is frame invariant: use the input frame, Sun-centred.
  double vmag;
  {
    const double obj_r = norm(x);
    const Vec3 o2o{o.x - x.x, o.y - x.y, o.z - x.z};
    const double o2o_r = norm(o2o);
    const Vec3 neg{-o2o.x, -o2o.y, -o2o.z};
    const double n2 = norm(neg);
    double phase;
    if (obj_r == 0.0 || n2 == 0.0) {
      phase = 0.0;
    } else {
      const double cang = dot(x, neg) / (obj_r * n2);
      phase = cang > 1.0 ? 0.0 : (cang < -1.0 ? 3.141592653589793 : acos(cang));
    }
    if (phase > 2.792526803190927) {
      vmag = __builtin_nan("");
    } else {
      const double g = p.g;
      const double corr = log10((1.0 - g) * hg_helper(3.332, 0.631, 0.986, phase) +
                                g * hg_helper(1.862, 1.218, 0.238, phase));
      vmag = p.h[q] + 5.0 * log10(obj_r * o2o_r) - 2.5 * corr;
    }
  }
  if (p.frame == 17) {
    x = ecl_to_equ(x);
    v = ecl_to_equ(v);
    o = ecl_to_equ(o);
    ov = ecl_to_equ(ov);
  }
  const Vec3 d{x.x - o.x, x.y - o.y, x.z - o.z};
  const Vec3 w{v.x - ov.x, v.y - ov.y, v.z - ov.z};
  const double d_ra = (d.x * w.y - d.y * w.x) / (d.x * d.x + d.y * d.y);
  const double r2 = dot(d, d);
  const double d_dec = (w.z - d.z * dot(d, w) / r2) / sqrt(r2 - d.z * d.z);
  const double r = norm(d);
  double dec = 0.0, ra = 0.0;
  if (!(r < 1e-10)) {
    const double zr = fmin(fmax(d.z / r, -1.0), 1.0);
    dec = rem_euclid(3.0 * (3.141592653589793 / 2.0) - acos(zr), 6.283185307179586) -
          3.141592653589793;
    ra = rem_euclid(atan2(d.y, d.x), 6.283185307179586);
  }
  p.out[0][q] = ra;
  p.out[1][q] = dec;
  p.out[2][q] = d_ra * cos(dec);
  p.out[3][q] = d_dec;
  p.out[4][q] = vmag;
}

int32_t launch_fov_prefilter(const FovPrefilter& p, hipStream_t s) {
  if (p.n <= 0 || p.nf <= 0) return 0;
  if (p.nf > FOV_MAX_TILE_FIELDS) return -1;
  hipLaunchKernelGGL(fov_prefilter_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256),
                     (size_t)p.nf * sizeof(FieldDev), s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t launch_fov_two_body(const FovTwoBody& p, hipStream_t s) {
  if (p.cand_cap <= 0) return 0;
  const int64_t g = std::min<int64_t>((p.cand_cap + 255) / 256, 2048);
  hipLaunchKernelGGL(fov_two_body_kernel, dim3((unsigned)g), dim3(256), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t launch_fov_nb_finish(const FovNBody& p, hipStream_t s) {
  if (p.m <= 0) return 0;
  hipLaunchKernelGGL(fov_nb_finish_kernel, dim3((unsigned)((p.m + 255) / 256)), dim3(256), 0, s,
                     p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t launch_radec_vmag(const RadecArgs& p, hipStream_t s) {
  if (p.n <= 0) return 0;
  hipLaunchKernelGGL(radec_vmag_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, s, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace kh
