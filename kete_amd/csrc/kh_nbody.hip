// kh_nbody.hip -- Gauss-Radau 15 N-body propagation on gfx950.
//
// Restates RadauIntegrator::integrate/step (src/kete_core/src/propagation/
// radau.rs:119-346) driving spk_accel (acceleration.rs:121-168) inside
// propagate_n_body_spk (propagation/mod.rs:70-118).
//
// Layout and launch
//  - One lane integrates one object.  Object state is structure-of-arrays in
//    HBM (coalesced loads/stores at start and end); the integrator state
//    (t, h, x, v, a, B[7][3], G[7][3]) lives in VGPRs for the whole arc.
//  - Each lane runs kete's integrator as a state machine whose loop body makes
//    exactly one force evaluation: the initial exact evaluation, one of the 7
//    substeps of a predictor-corrector sweep, or the exact evaluation after an
//    accepted step.  The force model (11 Chebyshev lookups plus the Newton,
//    GR and J2 terms) is >90% of the arithmetic and runs with every live lane
//    of the wavefront active, whatever step, sweep or substep each lane is on.
//  - The bookkeeping around it is branch-free: every lane computes the
//    candidate updates (G column, B = G*C, accepted x and v) and keeps them
//    with selects.  Long-lived state modified under divergent branches would
//    otherwise be held twice by the register allocator.
//  - The Radau tables live in LDS and are indexed per lane by its substep.
//  - Persistent mode: the grid is sized to the machine and a lane that
//    finishes an object takes the next one from a device-side counter, so
//    objects with very different step counts do not leave lanes idle.
#include <cstdlib>
#include <string>
#include <type_traits>

#include "kh_device.h"
#include "kh_internal.h"

namespace kh {

// Radau tables (radau.rs:24-68), filled by the host from the same formulas.
__constant__ RadauTables RT;


struct NBodyArgs {
  int64_t n;
  const double* jd0;
  const double* xyz[6];
  const double* jd_final;
  int32_t jd_final_stride;
  int32_t frame;  // 1 or 17
  Chain center;   // segments whose sum is the center's SSB state
  const NonGravDev* ng;
  double epsilon;
  double* out[6];
  int32_t* status;
  int64_t* impact_id;
  double* impact_jd;
  int32_t* counters;  // [n][4] or null
  unsigned long long* queue;
  int64_t first_dynamic;  // objects below this are assigned by lane id
  unsigned long long* range;  // [4]: bits of min and max epoch over the batch, then of the starting epochs
  int32_t window;             // records per segment staged in LDS (0: none)
  const uint32_t* perm;       // queue slot -> object (longest first), or null
  int64_t slot_begin, slot_end;  // the queue slots this launch takes
  // device-chosen split of the order (head: latency kernel, tail: the rest):
  // with split != null, part 1 takes slots [0, *split), part 2 [*split, n)
  const int64_t* split;
  int32_t part;
  // multi-epoch propagation: every object is integrated through epochs[0..
  // n_epochs) in order without restarting, and its SSB-equatorial state at
  // each is stored in mout[6][n_epochs][n] (NaN from a failing epoch on)
  const double* epochs;
  int32_t n_epochs;
  double* mout[6];
  int32_t n_ast;  // asteroids after the planets in the force list (0 or 5)
  // tail hand-off of the plain planets-only launch (park_lanes = 0: off)
  int32_t park_lanes;
  double park_min_left;
  // plain launch over the longest-first order: each object's state as one
  // 64-byte record [n][8] in object order -- written by the pre-pass
  // (x, v, jd0, final epoch; jd0 NaN for an object the pre-pass ended),
  // read by the integrator through the permutation, overwritten with the
  // result (x, v, impact epoch, status and impact id), turned back into the
  // output columns by the post-pass -- so an object's input and its result
  // each move as one cache line instead of seven to nine scattered words.
  // null: the output columns carry the state.
  double* stage;
};

// Last index of the run of epochs from e that moves away from t in one
// direction (multi-epoch propagation).
__device__ __forceinline__ int32_t multi_run_end(const NBodyArgs& a, int32_t e, double t) {
  const double dir = a.epochs[e] - t;
  int32_t k = e;
  while (k + 1 < a.n_epochs && (a.epochs[k + 1] - a.epochs[k]) * dir > 0.0) ++k;
  return k;
}

// The slot range of this launch (uniform: read once per lane from the
// device-chosen split when there is one).
__device__ __forceinline__ void slot_range(NBodyArgs& a) {
  if (a.split == nullptr) return;
  const int64_t k = *a.split;
  if (a.part == 1) {
    a.slot_begin = 0;
    a.slot_end = k;
  } else {
    a.slot_begin = k;
    a.slot_end = a.n;
  }
}

enum : int32_t { PH_INIT = 0, PH_SUB = 1, PH_EXACT = 2 };

// Convert the input state to SSB-centred equatorial (mod.rs:78-79 via
// spk.rs:99-126 and state.rs:146-281): add the center's chain states in the
// input frame, then rotate to the equatorial frame.
__device__ __forceinline__ int32_t to_ssb_equatorial(const SpkDev* __restrict__ spk,
                                                     const double* __restrict__ coef,
                                                     const NBodyArgs& a, double jd, Vec3& x,
                                                     Vec3& v) {
  const bool ecl = a.frame == 17;
  for (int k = 0; k < a.center.n; ++k) {
    Vec3 p, w;
    int32_t st = seg_pos_vel(spk, coef, a.center.seg[k], jd, p, w);
    if (st) return st;
    if (ecl) {
      p = equ_to_ecl(p);
      w = equ_to_ecl(w);
    }
    x = Vec3{x.x + p.x, x.y + p.y, x.z + p.z};
    v = Vec3{v.x + w.x, v.y + w.y, v.z + w.z};
  }
  if (ecl) {
    x = ecl_to_equ(x);
    v = ecl_to_equ(v);
  }
  return ST_OK;
}

// Back to the input center and frame (mod.rs:115-116): subtract the chain
// from the root down (the find_path order from 0), then rotate.
__device__ __forceinline__ int32_t from_ssb_equatorial(const SpkDev* __restrict__ spk,
                                                       const double* __restrict__ coef,
                                                       const NBodyArgs& a, double jd, Vec3& x,
                                                       Vec3& v) {
  for (int k = a.center.n - 1; k >= 0; --k) {
    Vec3 p, w;
    int32_t st = seg_pos_vel(spk, coef, a.center.seg[k], jd, p, w);
    if (st) return st;
    x = Vec3{x.x + (-p.x), x.y + (-p.y), x.z + (-p.z)};
    v = Vec3{v.x + (-w.x), v.y + (-w.y), v.z + (-w.z)};
  }
  if (a.frame == 17) {
    x = equ_to_ecl(x);
    v = equ_to_ecl(v);
  }
  return ST_OK;
}

__device__ __forceinline__ void store_result(const NBodyArgs& a, int64_t i, int32_t st, Vec3 x,
                                             Vec3 v, int32_t impact, double impact_jd,
                                             int4 cnt) {
  const double nan = __builtin_nan("");
  if (st != ST_OK) x = v = Vec3{nan, nan, nan};
  if (a.stage != nullptr) {
    // the object's record: x, v, impact epoch, status | impact id << 32
    double2* r = reinterpret_cast<double2*>(a.stage + 8 * i);
    const long long tag = (long long)(uint32_t)st |
                          ((long long)(uint32_t)(st == ST_IMPACT ? impact : 0) << 32);
    r[0] = double2{x.x, x.y};
    r[1] = double2{x.z, v.x};
    r[2] = double2{v.y, v.z};
    r[3] = double2{st == ST_IMPACT ? impact_jd : nan, __longlong_as_double(tag)};
    if (a.counters) reinterpret_cast<int4*>(a.counters)[i] = cnt;
    return;
  }
  a.out[0][i] = x.x;
  a.out[1][i] = x.y;
  a.out[2][i] = x.z;
  a.out[3][i] = v.x;
  a.out[4][i] = v.y;
  a.out[5][i] = v.z;
  a.status[i] = st;
  if (a.impact_id) a.impact_id[i] = st == ST_IMPACT ? impact : 0;
  if (a.impact_jd) a.impact_jd[i] = st == ST_IMPACT ? impact_jd : nan;
  if (a.counters) reinterpret_cast<int4*>(a.counters)[i] = cnt;
}

// Pre-pass, one lane per object: non-finite inputs and DAFLimits at the
// center conversion finish here with a NaN state (the pyo3 binding's
// behaviour, src/kete/rust/propagation.rs:113-121); every other object's
// starting state is converted to SSB-centred equatorial and parked in the
// output arrays, where the integrator picks it up.
__global__ void nbody_prepare_kernel(const SpkDev* __restrict__ spk,
                                     const double* __restrict__ coef, NBodyArgs a) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i0 < a.n;
  const int64_t i = valid ? i0 : a.n - 1;
  const double t = a.jd0[i];
  Vec3 x{a.xyz[0][i], a.xyz[1][i], a.xyz[2][i]};
  Vec3 v{a.xyz[3][i], a.xyz[4][i], a.xyz[5][i]};
  const bool finite = isfinite(t) && isfinite(x.x) && isfinite(x.y) && isfinite(x.z) &&
                      isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
  int32_t st = finite ? to_ssb_equatorial(spk, coef, a, t, x, v) : ST_VALUE_ERROR;
  if (valid) {
    if (a.stage != nullptr) {
      // the object's record for the integrator (jd0 NaN: ended here); the
      // columns carry only the objects that end here
      const double nan = __builtin_nan("");
      const double tf = a.jd_final[a.jd_final_stride ? i : 0];
      double2* r = reinterpret_cast<double2*>(a.stage + 8 * i);
      r[0] = double2{x.x, x.y};
      r[1] = double2{x.z, v.x};
      r[2] = double2{v.y, v.z};
      r[3] = double2{st == ST_OK ? t : nan, tf};
      if (st != ST_OK) {
        NBodyArgs b = a;
        b.stage = nullptr;
        store_result(b, i, st, x, v, 0, 0.0, int4{0, 0, 0, 0});
      } else {
        a.status[i] = ST_OK;
      }
    } else {
      store_result(a, i, st, x, v, 0, 0.0, int4{0, 0, 0, 0});
    }
  }
  if (!valid) st = ST_VALUE_ERROR;
  if (a.range != nullptr) {
    // epoch range of the batch (all epochs are positive JDs, so the IEEE bit
    // patterns order like the values); reduced over the wavefront first
    const double tf = a.jd_final[a.jd_final_stride ? i : 0];
    double lo = st == ST_OK ? fmin(t, tf) : __builtin_inf();
    double hi = st == ST_OK ? fmax(t, tf) : 0.0;
    // and the range of the starting epochs (where the window is centred)
    double slo = st == ST_OK ? t : __builtin_inf();
    double shi = st == ST_OK ? t : 0.0;
    for (int off = 32; off > 0; off >>= 1) {
      lo = fmin(lo, __shfl_xor(lo, off));
      hi = fmax(hi, __shfl_xor(hi, off));
      slo = fmin(slo, __shfl_xor(slo, off));
      shi = fmax(shi, __shfl_xor(shi, off));
    }
    // then over the block's wavefronts, so one lane per block updates the
    // four words (same-address atomics serialise)
    __shared__ double red[4][4];
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[w][0] = lo;
      red[w][1] = hi;
      red[w][2] = slo;
      red[w][3] = shi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < nw; ++k) {
        lo = fmin(lo, red[k][0]);
        hi = fmax(hi, red[k][1]);
        slo = fmin(slo, red[k][2]);
        shi = fmax(shi, red[k][3]);
      }
      if (lo <= hi && lo > 0.0) {
        atomicMin(&a.range[0], (unsigned long long)__double_as_longlong(lo));
        atomicMax(&a.range[1], (unsigned long long)__double_as_longlong(hi));
        atomicMin(&a.range[2], (unsigned long long)__double_as_longlong(slo));
        atomicMax(&a.range[3], (unsigned long long)__double_as_longlong(shi));
      }
    }
  }
}

// Epoch range before the pre-pass folds in the objects' own epochs: empty,
// or [lo, hi] (the epochs of a multi-epoch propagation); r[2..3], the range
// of the starting epochs, empty.
__global__ void init_range(unsigned long long* r, double lo, double hi) {
  if (threadIdx.x == 0) {
    r[0] = (unsigned long long)__double_as_longlong(lo);
    r[1] = (unsigned long long)__double_as_longlong(hi);
    r[2] = (unsigned long long)__double_as_longlong(__builtin_inf());
    r[3] = 0ull;
  }
}

// First record of segment g in the launch's LDS window of a.window records:
// the window covers the batch's epoch range from its start when the range
// fits; otherwise it is centred on the starting epochs, where every arc
// begins and the short ones stay (arcs in both directions from a common
// epoch), within the range.
__device__ __forceinline__ int32_t window_first_record(const NBodyArgs& a, const SegDev& g) {
  const double lo = (__longlong_as_double((long long)a.range[0]) - 2451545.0) * 86400.0;
  const double hi = (__longlong_as_double((long long)a.range[1]) - 2451545.0) * 86400.0;
  const double slo = __longlong_as_double((long long)a.range[2]);
  const double shi = __longlong_as_double((long long)a.range[3]);
  const double len = (double)a.window * g.intlen;
  double w = lo;
  if (hi - lo > len && shi >= slo) {
    const double c = (0.5 * (slo + shi) - 2451545.0) * 86400.0;
    w = fmax(lo, fmin(c - 0.5 * len, hi - len));
  }
  return record_index(g, w);
}

// Wavefront windows.  When the batch's epochs span more than the block's LDS
// window covers (long arcs), each wavefront of the plain launch keeps a
// window of its own -- ww_records(a) consecutive records of every segment,
// its share of the block's LDS -- and moves it along with its lanes: when a
// force evaluation has lanes outside it, the window moves to start at the
// record of the lanes' earliest step start (latest, for arcs run backwards)
// if that is further along.  Step starts only advance, so the window only
// moves forward and each record is staged once; evaluations with lanes
// outside the window take the per-body path from HBM.  Shared record grids
// only (one record index for every segment).
__device__ __forceinline__ int32_t ww_records(const NBodyArgs& a) {
  return a.window * 64 / (int32_t)blockDim.x;
}

__device__ __forceinline__ bool ww_mode(const NBodyArgs& a, const SpkDev* __restrict__ spk) {
  if (a.window <= 0 || !spk->shared_grid || !spk->all14 || ww_records(a) < 2) return false;
  const SegDev& g = spk->seg[spk->emb_seg];
  if (g.n_rec < ww_records(a)) return false;
  const double lo = (__longlong_as_double((long long)a.range[0]) - 2451545.0) * 86400.0;
  const double hi = (__longlong_as_double((long long)a.range[1]) - 2451545.0) * 86400.0;
  return hi - lo > (double)a.window * g.intlen;
}

// Stage records [w0, w0 + wr) of every segment into the wavefront's window
// (layout [seg][wr][44], word 1 as its reciprocal), spread over the active
// lanes eight loads at a time; the wavefront's later LDS reads see the stores
// (LDS operations of a wavefront complete in order).
__device__ __forceinline__ void ww_stage(const SpkDev* __restrict__ spk,
                                         const double* __restrict__ coef, double* wrec,
                                         int32_t w0, int32_t wr) {
  const uint64_t act = __ballot(1);
  const int nact = __popcll(act);
  const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
  const int per_seg = wr * 44;
  const int total = spk->n_seg * per_seg;
  constexpr int U = 8;
  for (int base = rank; base < total; base += U * nact) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = base + u * nact;
      if (k < total) {
        const int sg = k / per_seg, rem = k - sg * per_seg;
        v[u] = coef[spk->seg[sg].off + (int64_t)w0 * 44 + rem];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * nact < total) wrec[base + u * nact] = v[u];
  }
  for (int k = rank; k < spk->n_seg * wr; k += nact) {
    const int sg = k / wr, r = k - sg * wr;
    wrec[k * 44 + 1] = 1.0 / coef[spk->seg[sg].off + (int64_t)(w0 + r) * 44 + 1];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Post-pass: back to the input center and frame at jd_final (mod.rs:115-116).
__global__ void nbody_finish_kernel(const SpkDev* __restrict__ spk,
                                    const double* __restrict__ coef, NBodyArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n || a.status[i] != ST_OK) return;
  const double tf = a.jd_final[a.jd_final_stride ? i : 0];
  Vec3 x, v;
  if (a.stage != nullptr) {
    // the integrator's result record; a failed integration keeps its NaNs
    const double2* r = reinterpret_cast<const double2*>(a.stage + 8 * i);
    const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    const long long tag = __double_as_longlong(r3.y);
    const int32_t ist = (int32_t)(uint32_t)(tag & 0xffffffffll);
    x = Vec3{r0.x, r0.y, r1.x};
    v = Vec3{r1.y, r2.x, r2.y};
    if (ist != ST_OK) {
      a.out[0][i] = x.x;
      a.out[1][i] = x.y;
      a.out[2][i] = x.z;
      a.out[3][i] = v.x;
      a.out[4][i] = v.y;
      a.out[5][i] = v.z;
      a.status[i] = ist;
      if (a.impact_id) a.impact_id[i] = (int64_t)(int32_t)(uint32_t)(tag >> 32);
      if (a.impact_jd) a.impact_jd[i] = r3.x;
      return;
    }
    if (a.impact_id) a.impact_id[i] = 0;
    if (a.impact_jd) a.impact_jd[i] = __builtin_nan("");
  } else {
    x = Vec3{a.out[0][i], a.out[1][i], a.out[2][i]};
    v = Vec3{a.out[3][i], a.out[4][i], a.out[5][i]};
  }
  const int32_t st = from_ssb_equatorial(spk, coef, a, tf, x, v);
  if (st != ST_OK) {
    const double nan = __builtin_nan("");
    x = v = Vec3{nan, nan, nan};
    a.status[i] = st;
  }
  a.out[0][i] = x.x;
  a.out[1][i] = x.y;
  a.out[2][i] = x.z;
  a.out[3][i] = v.x;
  a.out[4][i] = v.y;
  a.out[5][i] = v.z;
}

// Post-pass of the multi-epoch propagation: each stored state back to the
// input center and frame at its epoch, in place.  One lane per (epoch, object).
__global__ void nbody_finish_multi_kernel(const SpkDev* __restrict__ spk,
                                          const double* __restrict__ coef, NBodyArgs a) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)a.n_epochs * a.n) return;
  const int64_t e = q / a.n;
  Vec3 x{a.mout[0][q], a.mout[1][q], a.mout[2][q]};
  Vec3 v{a.mout[3][q], a.mout[4][q], a.mout[5][q]};
  if (!isfinite(x.x)) return;
  const int32_t st = from_ssb_equatorial(spk, coef, a, a.epochs[e], x, v);
  if (st != ST_OK) {
    const double nan = __builtin_nan("");
    x = v = Vec3{nan, nan, nan};
  }
  a.mout[0][q] = x.x;
  a.mout[1][q] = x.y;
  a.mout[2][q] = x.z;
  a.mout[3][q] = v.x;
  a.mout[4][q] = v.y;
  a.mout[5][q] = v.z;
}

// Next object whose pre-pass succeeded, from queue slot `slot` on (the slot
// is the object index unless a longest-first order is given); n when none is
// left.  Its SSB-equatorial starting state is read from the output arrays.
template <bool PERSISTENT>
__device__ __forceinline__ int64_t fetch(const NBodyArgs& a, int64_t slot, double& t,
                                         double& tf, Vec3& x, Vec3& v) {
  while (slot < a.slot_end) {
    const int64_t i = a.perm != nullptr ? (int64_t)a.perm[slot] : slot;
    if (a.stage != nullptr) {
      // one 64-byte record (jd0 NaN: the pre-pass ended the object)
      const double2* r = reinterpret_cast<const double2*>(a.stage + 8 * i);
      const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
      if (!isnan(r3.x)) {
        t = r3.x;
        tf = r3.y;
        x = Vec3{r0.x, r0.y, r1.x};
        v = Vec3{r1.y, r2.x, r2.y};
        return i;
      }
    } else if (a.status[i] == ST_OK) {
      t = a.jd0[i];
      tf = a.jd_final[a.jd_final_stride ? i : 0];
      x = Vec3{a.out[0][i], a.out[1][i], a.out[2][i]};
      v = Vec3{a.out[3][i], a.out[4][i], a.out[5][i]};
      return i;
    }
    if (!PERSISTENT) return a.n;
    slot = a.first_dynamic + (int64_t)atomicAdd(a.queue, 1ull);
  }
  return a.n;
}

// ---------------------------------------------------------------------------
// Longest-first order.  The slowest objects bound the launch; handing them
// out first lets the rest fill in around them, and lanes of a wavefront then
// carry objects of similar cost, so few idle while their wavefront finishes.
// The cost key is |span| times the mean of r^-1.5 (the local dynamical rate
// the Radau steps follow) at nine points of the arc on the two-body orbit
// through the input state, times exp(e^2); binned on a log scale
// (eighth-octave buckets) and counting-sorted, highest first.  On the
// default workload the nine points packed wavefronts to 82% of their lanes'
// work where the mean motion alone (the rate at the epoch's semi-major
// axis) packed them to 68%; the eccentricity factor takes it to 85%.  The order
// changes which lane takes an object, never its result.
// ---------------------------------------------------------------------------
// eighth-octave cost levels, each split by the direction of the arc, so that
// the lanes of a wavefront also move through time the same way
constexpr int ORDER_LEVELS = 1024;
constexpr int ORDER_BUCKETS = 2 * ORDER_LEVELS;
constexpr int ORDER_POINTS = 9;

__device__ __forceinline__ int32_t order_bucket(const NBodyArgs& a, int64_t i) {
  const double t = a.jd0[i], tf = a.jd_final[a.jd_final_stride ? i : 0];
  const Vec3 x{a.xyz[0][i], a.xyz[1][i], a.xyz[2][i]};
  const Vec3 v{a.xyz[3][i], a.xyz[4][i], a.xyz[5][i]};
  const double span = tf - t;
  const double r0 = norm(x);
  double sum = 1.0 / (r0 * sqrt(r0));
  // each point from the previous one (short solves converge in fewer Newton
  // iterations than solves from the epoch); after a failed solve the
  // remaining points count as the epoch's rate
  Vec3 xc = x, vc = v;
  bool ok = true;
  const double dt = span / (ORDER_POINTS - 1);
#pragma unroll 1
  for (int k = 1; k < ORDER_POINTS; ++k) {
    double rk = r0;
    if (ok) {
      Vec3 p, w;
      ok = analytic_2_body(dt, xc, vc, p, w) == ST_OK;
      if (ok) {
        xc = p;
        vc = w;
        rk = norm(p);
      }
    }
    sum += 1.0 / (rk * sqrt(rk));
  }
  // eccentric orbits cost more than the nine points show (the steps crowd
  // around perihelion passages between them): times exp(e^2), e from the
  // input state.  Sorting 131,072 objects of the 1e6-object one-year
  // workload and 65,536 of the uneven-span one by their key and cutting the
  // order into wavefronts of 64, the oracle's force-evaluation counts fill
  // 84.8% and 79.8% of the wavefronts' lane time with it, 81.7% and 70.8%
  // without (99.8% and 97.1% sorted by the counts themselves).
  const Vec3 hv{x.y * v.z - x.z * v.y, x.z * v.x - x.x * v.z, x.x * v.y - x.y * v.x};
  const Vec3 ev{(v.y * hv.z - v.z * hv.y) * (1.0 / GMS) - x.x / r0,
                (v.z * hv.x - v.x * hv.z) * (1.0 / GMS) - x.y / r0,
                (v.x * hv.y - v.y * hv.x) * (1.0 / GMS) - x.z / r0};
  const double key = fabs(span) * (sum / ORDER_POINTS) * exp(fmin(dot(ev, ev), 4.0));
  if (!(key > 0.0) || !isfinite(key)) return 0;
  int lv = (int)floor(log2(key) * 8.0) + ORDER_LEVELS / 2;
  lv = lv < 1 ? 1 : (lv > ORDER_LEVELS - 1 ? ORDER_LEVELS - 1 : lv);
  return 2 * lv + (span < 0.0 ? 1 : 0);
}

// Bucket of each object and the bucket histogram (privatised in LDS).  With
// skey, also the radix-sort key of each object (bucket above the arc length,
// log2 |span| in 1/65536 octaves) and its index.
__global__ void order_hist_kernel(NBodyArgs a, uint16_t* __restrict__ bkt,
                                  uint32_t* __restrict__ hist, uint32_t* __restrict__ skey,
                                  uint32_t* __restrict__ sval) {
  __shared__ uint32_t lh[ORDER_BUCKETS];
  for (int k = threadIdx.x; k < ORDER_BUCKETS; k += blockDim.x) lh[k] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) {
    const int32_t b = order_bucket(a, i);
    bkt[i] = (uint16_t)b;
    atomicAdd(&lh[b], 1u);
    if (skey != nullptr) {
      const double sp = fabs(a.jd_final[a.jd_final_stride ? i : 0] - a.jd0[i]);
      const double q = (log2(sp) + 16.0) * 65536.0;
      const uint32_t sk = q > 0.0 ? (q < 2097151.0 ? (uint32_t)q : 2097151u) : 0u;
      skey[i] = ((uint32_t)b << 21) | sk;
      sval[i] = (uint32_t)i;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ORDER_BUCKETS; k += blockDim.x)
    if (lh[k] != 0) atomicAdd(&hist[k], lh[k]);
}

__global__ void order_zero_kernel(uint32_t* __restrict__ hist) {
  for (int k = threadIdx.x; k < ORDER_BUCKETS; k += blockDim.x) hist[k] = 0;
}

// Start of each bucket in the order, highest bucket first (one thread: 1024
// entries).  With `split`, also the head of the order that goes to the
// latency kernel: the objects whose cost key exceeds the average key load of
// one resident lane of the main launch (`lanes`), at most `cap` of them --
// objects that would otherwise outlast everything else.
__global__ void __launch_bounds__(256) order_scan_kernel(const uint32_t* __restrict__ hist,
                                                         uint32_t* __restrict__ off,
                                                         int64_t* __restrict__ split,
                                                         double lanes, int64_t cap) {
  // 256 threads, each over eight consecutive buckets of the descending order
  constexpr int PER = ORDER_BUCKETS / 256;
  static_assert(ORDER_BUCKETS % 256 == 0, "buckets per thread");
  __shared__ uint32_t s_cnt[256];
  __shared__ double s_load[256];
  __shared__ unsigned long long s_brk;
  const int t = threadIdx.x;
  uint32_t c[PER];
  uint32_t tot = 0;
  double ld = 0.0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int b = ORDER_BUCKETS - 1 - (t * PER + j);
    c[j] = hist[b];
    tot += c[j];
    if (b > 1) ld += (double)c[j] * exp2(((b >> 1) - ORDER_LEVELS / 2 + 0.5) / 8.0);
  }
  s_cnt[t] = tot;
  s_load[t] = ld;
  __syncthreads();
  if (t == 0) {
    // exclusive prefix of the threads' counts, and the total key load
    uint32_t run = 0;
    double total = 0.0;
    for (int k = 0; k < 256; ++k) {
      const uint32_t v = s_cnt[k];
      s_cnt[k] = run;
      run += v;
      total += s_load[k];
    }
    s_load[0] = total;
    s_brk = ~0ull;
  }
  __syncthreads();
  uint32_t run = s_cnt[t];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    off[ORDER_BUCKETS - 1 - (t * PER + j)] = run;
    run += c[j];
  }
  if (split == nullptr) return;
  // the head for the latency kernel: from the top, the buckets whose key
  // exceeds a resident lane's share of the load, at most `cap` objects; the
  // first bucket that fails (the lowest two always do) ends it
  const double per_lane = s_load[0] / lanes;
  run = s_cnt[t];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int pos = t * PER + j, b = ORDER_BUCKETS - 1 - pos;
    if (b <= 1 || exp2(((b >> 1) - ORDER_LEVELS / 2) / 8.0) <= per_lane ||
        (int64_t)run + (int64_t)c[j] > cap) {
      atomicMin(&s_brk, ((unsigned long long)pos << 32) | run);
      break;
    }
    run += c[j];
  }
  __syncthreads();
  if (t == 0) *split = (int64_t)(s_brk & 0xffffffffull);
}

// perm[position] = object; positions within a bucket are claimed per block.
__global__ void order_scatter_kernel(int64_t n, const uint16_t* __restrict__ bkt,
                                     uint32_t* __restrict__ off, uint32_t* __restrict__ perm) {
  __shared__ uint32_t cnt[ORDER_BUCKETS];
  __shared__ uint32_t base[ORDER_BUCKETS];
  for (int k = threadIdx.x; k < ORDER_BUCKETS; k += blockDim.x) cnt[k] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t b = 0;
  uint32_t rank = 0;
  if (i < n) {
    b = bkt[i];
    rank = atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ORDER_BUCKETS; k += blockDim.x)
    if (cnt[k] != 0) base[k] = atomicAdd(&off[k], cnt[k]);
  __syncthreads();
  if (i < n) perm[base[b] + rank] = (uint32_t)i;
}

// LDS copy of the per-substep factors, indexed by each lane's substep j.
struct LdsTables {
  double h[8];
  double inv_h[8];
  double rdiff[8][8];
};

constexpr int WIDE_L = 16;
constexpr int WIDE_EMB = 10;

// Parked objects of the plain launch: up to PARK_SLOTS per wavefront (one
// per sixteen-lane row), four wavefronts per block.
constexpr int PARK_SLOTS = 64 / WIDE_L;
__shared__ ParkRec kh_park_rec[4][PARK_SLOTS];
__shared__ int32_t kh_park_n[4];
// first record of each wavefront's window (wavefront windows): the value the
// window holds, for lanes that left the integration loop before its last move
__shared__ int32_t kh_ww_w0[4];

// ---------------------------------------------------------------------------
// One object on sixteen lanes (one DPP row), from `st` to the end of its arc:
// lane l < 10 evaluates body l of kete's PLANETS list (gravity.rs:145-148)
// -- its Chebyshev record, its Newton, GR and J2 terms and, for the Sun, the
// non-grav term (add_body_at) -- and lane 10 the Earth-Moon barycentre
// segment the Earth and Moon lanes add to their positions.  The row sums the
// terms (row_sum: the same bits in every lane).  The integrator is split by
// coordinate: lane c of each quad of the row (c < 3; lane 3 repeats
// coordinate 2) holds x_c, v_c, a_c, B[7]_c and G[7]_c and runs the one-lane
// kernel's arithmetic for that coordinate; the predicted position and
// velocity reach every lane by quad broadcasts, and the convergence test and
// the step-size maximum combine the quad's three coordinates.  The control
// values (t, h, phase, counters) are the same in every lane, so the row takes
// identical decisions without exchanging them.  A force evaluation thus
// costs one body's terms and a third of the integrator's bookkeeping per
// lane instead of the whole force sum: the latency of a slow object's serial
// chain drops, at a sixteenth of the one-lane kernel's objects per wavefront.
// Results differ from the one-lane kernel's only by rounding (the order in
// which the body terms are summed, the compiler's FMA contraction).  With
// AST, lanes 11..15 add the five MASSES asteroids.
// fresh: start at st.t from st.x, st.v (the integrator's first exact
// evaluation); otherwise st is a step boundary (ParkRec) and the object
// continues with its first substep.  Records: WINDOW reads them from the
// LDS window when there (the block's, or with wwin the wavefront's, whose
// first record is ww0; else from HBM); otherwise each lane keeps its record
// in rslot (LDS, 44 doubles), reloaded when the record changes.
// l: lane in the row, g0: the row's first lane in the wavefront.
// ---------------------------------------------------------------------------
// Output column c (< 6) of a multi-epoch propagation, selected without
// indexing the argument array by a lane-dependent value.
__device__ __forceinline__ double* mout_col(const NBodyArgs& a, int c) {
  double* r = a.mout[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) r = c == k ? a.mout[k] : r;
  return r;
}

template <bool WINDOW, bool AST = false, bool MULTI = false>
__device__ __forceinline__ void wide_integrate(const SpkDev* __restrict__ spk,
                                               const double* __restrict__ coef,
                                               const NBodyArgs& a, const LdsTables& lds,
                                               int64_t obj, const ParkRec& st0, bool fresh,
                                               int l, int g0, double* rslot, int32_t winw,
                                               const double* wwin, int32_t ww0) {
  const double eps = a.epsilon;
  double t = st0.t;
  double tf = st0.tf;
  int32_t ej = 0;          // MULTI: the next epoch to store
  double h_want = 0.0;     // MULTI: the step proposed before shortening
  bool shortened = false;  // MULTI: the step was shortened to land on tf
  // this lane's coordinate of the integrator (lane 3 of each quad repeats
  // coordinate 2, so quad reductions see only the three coordinates)
  const int cd = (l & 3) < 3 ? (l & 3) : 2;
  const auto pick = [cd](double c0, double c1, double c2) {
    return cd == 0 ? c0 : (cd == 1 ? c1 : c2);
  };
  double xd = pick(st0.x[0], st0.x[1], st0.x[2]);
  double vd = pick(st0.v[0], st0.v[1], st0.v[2]);
  const NonGravDev* ngp = (a.ng && a.ng[obj].kind != 0) ? &a.ng[obj] : nullptr;

  // this lane's body (lanes 0..9) or segment (lane 10); lanes 11..15 none
  const bool is_body = l < N_PLANETS;
  const bool is_seg = l <= WIDE_EMB;
  const BodyDev bd = spk->body[is_body ? l : 0];
  const int32_t seg = l == WIDE_EMB ? spk->emb_seg : bd.seg;
  const SegDev sg = spk->seg[seg];
  const bool need_vel = is_body && bd.kind == 1;
  const bool all14 = spk->all14 != 0;
  const bool cache_rec = all14 && sg.rsize == 44;
  int32_t rec_cached = -1;

  double h = 0.0, h_next = 0.0, ad = 0.0;
  double B[7], G[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) B[k] = G[k] = 0.0;
  int32_t phase = PH_INIT, j = 1, sweep = 0, fails = 0;
  int32_t n_force = 0, n_sweep = 0, n_acc = 0, n_fail = 0;
  if (!fresh) {
    // the first substep of the step that starts at st0 (the state after
    // `begin` in the one-lane kernel)
    h = st0.h;
    ad = pick(st0.a[0], st0.a[1], st0.a[2]);
#pragma unroll
    for (int k = 0; k < 7; ++k) B[k] = pick(st0.B[k][0], st0.B[k][1], st0.B[k][2]);
    phase = PH_SUB;
    n_force = st0.n_force;
    n_sweep = st0.n_sweep;
    n_acc = st0.n_acc;
    n_fail = st0.n_fail;
    if constexpr (MULTI) {
      ej = st0.ej;
      h_want = st0.h_want;
      shortened = st0.shortened != 0;
    }
  }

  for (;;) {
    const bool sub = phase == PH_SUB;
    // ---- evaluation point: predictor at spacing j (radau.rs:252-278), this
    // lane's coordinate, then the quad's three coordinates in every lane ----
    const double s = lds.h[j];
    double xe_d, ve_d;
    {
      double bw = 0.0, bu = 0.0;
      double sp = 1.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        sp *= s;
        const double wk = sp * RTC.w[k], uk = sp * RTC.u[k];
        bw += B[k] * wk;
        bu += B[k] * uk;
      }
      const double px = xd + s * vd * h + (s * s) * (h * h) * (ad / 2.0 + bw);
      const double pv = vd + s * h * (ad + bu);
      xe_d = sub ? px : xd;
      ve_d = sub ? pv : vd;
    }
    const double xe[3] = {quad_bcast<0>(xe_d), quad_bcast<1>(xe_d), quad_bcast<2>(xe_d)};
    const double ve[3] = {quad_bcast<0>(ve_d), quad_bcast<1>(ve_d), quad_bcast<2>(ve_d)};
    const double te = sub ? t + s * h : t;

    // ---- this lane's share of the force evaluation ----
    n_force++;
    int32_t lst = ST_OK;
    Vec3 bp{0.0, 0.0, 0.0}, bv{0.0, 0.0, 0.0};
    if (is_seg) {
      if (te < sg.jd_start || te > sg.jd_end) lst = ST_DAF_LIMITS;
      const double js = (te - 2451545.0) * 86400.0;
      const int32_t ri = record_index(sg, js);
      const double* __restrict__ grec = coef + sg.off + (int64_t)ri * sg.rsize;
      double ir, xs;
      Vec3 pp, dd{0.0, 0.0, 0.0};
      int32_t loc = -1;
      if constexpr (WINDOW) loc = winw > 0 && all14 ? ri - (wwin ? ww0 : kh_win_w0[seg]) : -1;
      if (WINDOW && loc >= 0 && loc < winw) {
        // the record in the LDS window (the block's, or the wavefront's
        // wwin; word 1 staged as 1 / radius)
        const double* rec = (wwin ? wwin : kh_win_rec) + (seg * winw + loc) * 44;
        ir = rec[1];
        xs = (js - rec[0]) * ir;
        cheb_both_lds(rec, xs, pp, dd);
      } else if (!WINDOW && cache_rec) {
        if (ri != rec_cached) {
          // the record, word 1 (the radius) as its reciprocal
          const double2* __restrict__ src = reinterpret_cast<const double2*>(grec);
#pragma unroll
          for (int q = 0; q < 22; ++q) reinterpret_cast<double2*>(rslot)[q] = src[q];
          rslot[1] = 1.0 / grec[1];
          rec_cached = ri;
        }
        ir = rslot[1];
        xs = (js - rslot[0]) * ir;
        cheb_both_lds(rslot, xs, pp, dd);
      } else {
        ir = 1.0 / grec[1];
        xs = (js - grec[0]) * ir;
        if (all14) cheb_both_n<14>(grec + 2, xs, pp, dd);
        else cheb_both_dyn(grec + 2, sg.n_coef, xs, pp, dd);
      }
      const double vsc = 86400.0 * ir * INV_AU_KM;
      bp = Vec3{pp.x * INV_AU_KM, pp.y * INV_AU_KM, pp.z * INV_AU_KM};
      bv = need_vel ? Vec3{dd.x * vsc, dd.y * vsc, dd.z * vsc} : Vec3{0.0, 0.0, 0.0};
    }
    // the barycentre is read before the bodies (spk_accel), so its failure
    // comes first; Earth and Moon are given relative to it
    const int32_t emb_st = row_bcast_i<WIDE_EMB>(lst);
    const Vec3 emb{row_bcast<WIDE_EMB>(bp.x), row_bcast<WIDE_EMB>(bp.y),
                   row_bcast<WIDE_EMB>(bp.z)};
    Vec3 acc{0.0, 0.0, 0.0};
    // with the asteroids (MASSES), lanes 11..15 take asteroid l - 11: its
    // Type 21 state, plus the Sun's (lane 0's body) for a Sun-centred
    // segment, and its Newtonian term (add_asteroids)
    const bool is_ast = AST && l > WIDE_EMB && l <= WIDE_EMB + N_ASTEROIDS;
    Vec3 sun{0.0, 0.0, 0.0};
    if constexpr (AST) sun = Vec3{row_bcast<0>(bp.x), row_bcast<0>(bp.y), row_bcast<0>(bp.z)};
    if (is_body) {
      if (bd.emb) bp = Vec3{bp.x + emb.x, bp.y + emb.y, bp.z + emb.z};
      if (lst == ST_OK)
        lst = add_body_at(bd, bp, bv, Vec3{xe[0], xe[1], xe[2]}, Vec3{ve[0], ve[1], ve[2]}, !sub,
                          ngp, acc);
    }
    if constexpr (AST) {
      if (is_ast) {
        const AstDev& ad = spk->ast[l - WIDE_EMB - 1];
        const double js = (te - 2451545.0) * 86400.0;
        const T21Ref ref = t21_find(spk, ad.seg_begin, ad.seg_end, te, js);
        if (!ref.found) {
          lst = ST_DAF_LIMITS;
        } else {
          Vec3 p, unused;
          t21_state_at<false>(spk, ref, js, p, unused);
          if (ref.sun_centred) p = Vec3{p.x + sun.x, p.y + sun.y, p.z + sun.z};
          lst = add_body_at(ad.bd, p, Vec3{0.0, 0.0, 0.0}, Vec3{xe[0], xe[1], xe[2]},
                            Vec3{ve[0], ve[1], ve[2]}, !sub, nullptr, acc);
        }
      }
    }
    int32_t st = emb_st, impact = 0;
    {
      const uint64_t bad = __ballot((is_body || is_ast) && lst != ST_OK);
      const uint32_t gbad = (uint32_t)((bad >> g0) & 0xFFFFull);
      if (st == ST_OK && gbad) {
        const int first = __builtin_ctz(gbad);
        st = __shfl(lst, g0 + first, 64);
        if (st == ST_IMPACT)
          impact = first < N_PLANETS ? spk->body[first].id : spk->ast[first - WIDE_EMB - 1].bd.id;
      }
    }
    const double ev[3] = {row_sum(acc.x), row_sum(acc.y), row_sum(acc.z)};
    const double ev_d = pick(ev[0], ev[1], ev[2]);

    // ---- G column j (radau.rs:292-308), kept only on substeps ----
    {
      double gk = (ev_d - ad) * lds.inv_h[j];
#pragma unroll
      for (int z = 1; z < 7; ++z) {
        const double rd = lds.rdiff[j][z];
        gk = z < j ? (gk - G[z - 1]) * rd : gk;
      }
#pragma unroll
      for (int q = 0; q < 7; ++q) G[q] = sub && q == j - 1 ? gk : G[q];
    }
    // ---- end of sweep: B = G*C and the convergence test (radau.rs:313-319),
    // the test's three coordinates combined over the quad ----
    const bool es = sub && j == 7 && st == ST_OK;
    const double bp6 = B[6];
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      double acc2 = G[c];
#pragma unroll
      for (int k = c + 1; k < 7; ++k) acc2 += G[k] * RTC.c[k][c];
      B[c] = es ? acc2 : B[c];
    }
    const int32_t small_d = fabs(B[6] - bp6) < 1e-14 * (fabs(ev_d) + 1e-6) ? 1 : 0;
    const bool small = (quad_bcast_i<0>(small_d) & quad_bcast_i<1>(small_d) &
                        quad_bcast_i<2>(small_d)) != 0;
    const bool conv = es && small;
    // ---- accepted step (radau.rs:320-338) ----
    {
      double bw = 0.0, bu = 0.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        bw += B[k] * RTC.w[k];
        bu += B[k] * RTC.u[k];
      }
      const double nx = xd + (vd * h + (h * h) * (ad * 0.5 + bw));
      const double nv = vd + h * (ad + bu);
      if constexpr (MULTI) {
        // epochs strictly inside the accepted step: the step's interpolant
        // at s = (e - t)/h, this lane's coordinate (stored by lanes 0..2)
        while (conv && ej < a.n_epochs) {
          const double e = a.epochs[ej];
          const double se = (e - t) / h;
          if (!(se > 0.0 && se < 1.0) || fabs(e - (t + h)) < 1e-12) break;
          double bw2 = 0.0, bu2 = 0.0, sp = 1.0;
#pragma unroll
          for (int k = 0; k < 7; ++k) {
            sp *= se;
            bw2 += B[k] * (sp * RTC.w[k]);
            bu2 += B[k] * (sp * RTC.u[k]);
          }
          const double ox = xd + se * vd * h + (se * se) * (h * h) * (ad / 2.0 + bw2);
          const double ov = vd + se * h * (ad + bu2);
          if (l < 3) {
            const int64_t q = (int64_t)ej * a.n + obj;
            mout_col(a, l)[q] = ox;
            mout_col(a, 3 + l)[q] = ov;
          }
          ++ej;
        }
      }
      xd = conv ? nx : xd;
      vd = conv ? nv : vd;
    }
    if (conv) {
      // next step size (radau.rs:339-342)
      const double fm =
          fmax(fmax(fabs(ev[0]) + 1e-6, fabs(ev[1]) + 1e-6), fabs(ev[2]) + 1e-6);
      const double b6 = fabs(B[6]);
      const double bm = fmax(fmax(quad_bcast<0>(b6), quad_bcast<1>(b6)), quad_bcast<2>(b6));
      double ratio = root7(eps * fm / bm);
      ratio = ratio < 0.25 ? 0.25 : (ratio > 4.0 ? 4.0 : ratio);
      h_next = h * ratio;
    }
    t = conv ? t + h : t;
    const bool take_a = st == ST_OK && !sub;
    ad = take_a ? ev_d : ad;

    // ---- scalar control (radau.rs:160-218) ----
    int32_t done = -1;
    bool begin = false;
    if (st != ST_OK) {
      done = st;
    } else if (phase == PH_INIT) {
      if (fabs(tf - t) < 1e-10) {
        done = ST_OK;
      } else {
        h = copysign(0.1, tf - t);
        begin = true;
      }
    } else if (phase == PH_EXACT) {
      h = h_next;
      if (fabs(t - tf) < 1e-12) {
        done = ST_OK;
      } else {
        fails = 0;
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    } else if (j < 7) {
      j = j + 1;
    } else if (conv) {
      n_acc++;
      phase = PH_EXACT;
    } else if (sweep + 1 < 10) {
      sweep++;
      j = 1;
      n_sweep++;
    } else {
      n_fail++;
      fails++;
      h *= 0.7;
      if (fails > 10) {
        done = ST_CONVERGENCE;
      } else {
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    }
    if constexpr (MULTI) {
      // the epochs at the current time, after an accepted step and at the
      // end of a run; then on to the end of the next run (as the one-lane
      // kernel)
      if (done == ST_OK || (conv && done < 0)) {
        while (ej < a.n_epochs && fabs(a.epochs[ej] - t) < 1e-12) {
          if (l < 3) {
            const int64_t q = (int64_t)ej * a.n + obj;
            mout_col(a, l)[q] = xd;
            mout_col(a, 3 + l)[q] = vd;
          }
          ++ej;
        }
      }
      if (done == ST_OK) {
        if (ej < a.n_epochs) {
          tf = a.epochs[multi_run_end(a, ej, t)];
          const double hc = phase == PH_INIT ? 0.1 : (shortened ? h_want : h);
          h = copysign(fmax(fabs(hc), 0.005), tf - t);
          fails = 0;
          done = -1;
          begin = true;
        }
      } else if (done > 0) {
        // a failing epoch and the later ones are undefined
        const double nan = __builtin_nan("");
        if (l < 6)
          for (int32_t e = ej; e < a.n_epochs; ++e) mout_col(a, l)[(int64_t)e * a.n + obj] = nan;
      }
    }
    if (begin) {
      if (MULTI) h_want = h;
      shortened = fabs(t - tf) <= fabs(h);
      if (shortened) h = tf - t;
      sweep = 0;
      j = 1;
      n_sweep++;
      phase = PH_SUB;
    }
    if (done >= 0) {
      const Vec3 xo{quad_bcast<0>(xd), quad_bcast<1>(xd), quad_bcast<2>(xd)};
      const Vec3 vo{quad_bcast<0>(vd), quad_bcast<1>(vd), quad_bcast<2>(vd)};
      if (l == 0)
        store_result(a, obj, done, xo, vo, impact, te, int4{n_force, n_sweep, n_acc, n_fail});
      return;
    }
  }
}

// ILP: the latency variant for the slowest objects -- one wavefront per block
// and per SIMD, the whole register file, the force evaluation arranged for
// the shortest dependency chain (spk_accel_ilp), no LDS window.
// MULTI: the multi-epoch propagation (plain launch only).  The epochs are
// taken in runs that go one way in time from the current time; each run is
// one integration to its last epoch, and the states at the epochs inside it
// come from the accepted steps' interpolants.
// AST: the force list is MASSES (the planets and five asteroids) instead of
// PLANETS; a separate instantiation, so the planets-only kernels carry no
// asteroid code.
// WW: the plain planets-only launch of a batch with per-object final
// epochs, whose wavefronts keep windows of their own when its epochs span
// more than the block window (ww_mode); a separate instantiation, so the
// common-epoch kernel carries no such code.
template <bool PERSISTENT, bool ILP = false, bool MULTI = false, bool AST = false,
          bool WW = false>
__global__ void __launch_bounds__(ILP ? 64 : 256)
__attribute__((amdgpu_waves_per_eu(ILP ? 1 : 2, ILP ? 1 : 8)))
radau_nbody_kernel(const SpkDev* __restrict__ spk, const double* __restrict__ coef,
                   NBodyArgs a) {
  __shared__ LdsTables lds;
  for (int k = threadIdx.x; k < 8 * 8; k += blockDim.x) {
    const int r = k / 8, c = k % 8;
    lds.rdiff[r][c] = RT.rdiff[r][c];
    if (r == 0) {
      lds.h[c] = RT.h[c];
      lds.inv_h[c] = RT.inv_h[c];
    }
  }
  // Stage a window of a.window records per segment (a one-year arc needs 13
  // of 32 days), placed by window_first_record -- or, for the plain launch
  // over a longer range of epochs, give each wavefront a window of its own
  // (ww_mode), staged as its lanes go.
  constexpr bool WWABLE = WW && !PERSISTENT && !ILP && !MULTI;
  const bool ww = WWABLE && ww_mode(a, spk);
  if (a.window > 0 && !ww) {
    if (threadIdx.x < spk->n_seg)
      kh_win_w0[threadIdx.x] = window_first_record(a, spk->seg[threadIdx.x]);
    __syncthreads();
    const int per_seg = a.window * 44;
    for (int k = threadIdx.x; k < spk->n_seg * per_seg; k += blockDim.x) {
      const int sg = k / per_seg, rem = k - sg * per_seg;
      const SegDev& g = spk->seg[sg];
      const int r = kh_win_w0[sg] + rem / 44;
      // word 1 of a record (the radius) is staged as its reciprocal
      const double w = r < g.n_rec ? coef[g.off + (int64_t)r * 44 + rem % 44] : 1.0;
      kh_win_rec[k] = rem % 44 == 1 ? 1.0 / w : w;
    }
  }
  __syncthreads();
  const int32_t winw = ww ? ww_records(a) : a.window;
  double* const wwin = ww ? kh_win_rec + (threadIdx.x >> 6) * (spk->n_seg * winw * 44) : nullptr;
  int32_t ww0 = -(1 << 30);  // first record of the wavefront's window (none yet)

  const double eps = a.epsilon;
  slot_range(a);
  int64_t obj = a.slot_begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double t = 0.0, tf = 0.0, h = 0.0, h_next = 0.0;
  Vec3 x{0, 0, 0}, v{0, 0, 0}, a0{0, 0, 0};
  double B[7][3], G[7][3];
#pragma unroll
  for (int k = 0; k < 7; ++k) B[k][0] = B[k][1] = B[k][2] = G[k][0] = G[k][1] = G[k][2] = 0.0;
  int32_t phase = PH_INIT, j = 1, sweep = 0, fails = 0;
  int32_t n_force = 0, n_sweep = 0, n_acc = 0, n_fail = 0;
  obj = fetch<PERSISTENT>(a, obj, t, tf, x, v);
  int32_t ej = 0;                // MULTI: the next epoch to store
  double h_want = 0.0;           // MULTI: the step proposed before shortening
  bool shortened = false;        // MULTI: the step was shortened to land on tf
  if (MULTI) tf = a.epochs[multi_run_end(a, 0, t)];

  // the plain launch hands its wavefronts' tails on
  constexpr bool PARKABLE = !PERSISTENT && !ILP;
  if constexpr (PARKABLE)
    if ((threadIdx.x & 63) == 0) {
      kh_park_n[threadIdx.x >> 6] = 0;
      kh_ww_w0[threadIdx.x >> 6] = ww0;
    }
  while (obj < a.n) {
    const bool sub = phase == PH_SUB;
    // live lanes of the wavefront (every lane with an object is here)
    const uint64_t live_mask = PARKABLE ? __ballot(1) : 0ull;
    // ---- evaluation point: predictor at spacing j (radau.rs:252-278) ----
    const double s = lds.h[j];
    double xe[3], ve[3];
    {
      const double xs[3] = {x.x, x.y, x.z}, vs[3] = {v.x, v.y, v.z};
      const double as[3] = {a0.x, a0.y, a0.z};
      // bw = sum_k B_k s^(k+1) w_k, bu = sum_k B_k s^(k+1) u_k, k ascending
      double bw[3] = {0.0, 0.0, 0.0}, bu[3] = {0.0, 0.0, 0.0};
      double sp = 1.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        sp *= s;
        const double wk = sp * RTC.w[k], uk = sp * RTC.u[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          bw[d] += B[k][d] * wk;
          bu[d] += B[k][d] * uk;
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double px = xs[d] + s * vs[d] * h + (s * s) * (h * h) * (as[d] / 2.0 + bw[d]);
        double pv = vs[d] + s * h * (as[d] + bu[d]);
        xe[d] = sub ? px : xs[d];
        ve[d] = sub ? pv : vs[d];
      }
    }
    const double te = sub ? t + s * h : t;

    // ---- the wavefront's window follows its lanes ----
    if constexpr (WWABLE) {
      if (ww) {
        const SegDev& g = spk->seg[spk->emb_seg];
        const int32_t ri = record_index(g, (te - 2451545.0) * 86400.0);
        if (__ballot(ri < ww0 || ri >= ww0 + winw) != 0ull) {
          const int32_t rs = record_index(g, (t - 2451545.0) * 86400.0);
          const bool fwd = __builtin_amdgcn_readfirstlane(tf >= t ? 1 : 0) != 0;
          int32_t w = fwd ? __ockl_wfred_min_i32(rs) : __ockl_wfred_max_i32(rs) - winw + 1;
          w = w < 0 ? 0 : (w > g.n_rec - winw ? g.n_rec - winw : w);
          if (ww0 < 0 || (fwd ? w > ww0 : w < ww0)) {
            ww0 = w;
            kh_ww_w0[threadIdx.x >> 6] = w;
            ww_stage(spk, coef, wwin, w, winw);
          }
        }
      }
    }

    // ---- the one force evaluation of this iteration ----
    Vec3 acc;
    int32_t impact = 0;
    n_force++;
    const NonGravDev* ngp = (a.ng && a.ng[obj].kind != 0) ? &a.ng[obj] : nullptr;
    int32_t st;
    if constexpr (ILP)
      st = spk_accel_ilp(spk, coef, te, Vec3{xe[0], xe[1], xe[2]}, Vec3{ve[0], ve[1], ve[2]},
                         !sub, ngp, acc, impact, AST ? N_ASTEROIDS : 0);
    else
      st = spk_accel(spk, coef, te, Vec3{xe[0], xe[1], xe[2]}, Vec3{ve[0], ve[1], ve[2]}, !sub,
                     ngp, acc, impact, winw, AST ? N_ASTEROIDS : 0, wwin, ww0);
    const double ev[3] = {acc.x, acc.y, acc.z};

    // ---- G column j (radau.rs:292-308), kept only on substeps ----
    {
      const double as[3] = {a0.x, a0.y, a0.z};
      double gk[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) gk[d] = (ev[d] - as[d]) * lds.inv_h[j];
#pragma unroll
      for (int z = 1; z < 7; ++z) {
        const double rd = lds.rdiff[j][z];
#pragma unroll
        for (int d = 0; d < 3; ++d) gk[d] = z < j ? (gk[d] - G[z - 1][d]) * rd : gk[d];
      }
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        const bool put = sub && q == j - 1;
#pragma unroll
        for (int d = 0; d < 3; ++d) G[q][d] = put ? gk[d] : G[q][d];
      }
    }
    // ---- end of sweep: B = G*C and the convergence test (radau.rs:313-319) ----
    const bool es = sub && j == 7 && st == ST_OK;
    // B6 of the previous sweep (radau.rs:233); B is rewritten column by
    // column so no second copy of it is ever live.
    const double bp6[3] = {B[6][0], B[6][1], B[6][2]};
#pragma unroll
    for (int c = 0; c < 7; ++c) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double acc2 = G[c][d];
#pragma unroll
        for (int k = c + 1; k < 7; ++k) acc2 += G[k][d] * RTC.c[k][c];
        B[c][d] = es ? acc2 : B[c][d];
      }
    }
    // max_d |dB6_d| / (|a_d| + 1e-6) < 1e-14, tested per coordinate as
    // |dB6_d| < 1e-14 (|a_d| + 1e-6): no divides (the two forms can differ
    // only for ratios within an ulp of the threshold)
    double fmx[3];
    bool small = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      fmx[d] = fabs(ev[d]) + 1e-6;
      small = small && fabs(B[6][d] - bp6[d]) < 1e-14 * fmx[d];
    }
    const bool conv = es && small;
    // ---- accepted step (radau.rs:320-338) ----
    {
      double xs[3] = {x.x, x.y, x.z}, vs[3] = {v.x, v.y, v.z};
      const double as[3] = {a0.x, a0.y, a0.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double bw = 0.0, bu = 0.0;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          bw += B[k][d] * RTC.w[k];
          bu += B[k][d] * RTC.u[k];
        }
        double nx = xs[d] + (vs[d] * h + (h * h) * (as[d] * 0.5 + bw));
        double nv = vs[d] + h * (as[d] + bu);
        xs[d] = conv ? nx : xs[d];
        vs[d] = conv ? nv : vs[d];
      }
      if (MULTI && conv) {
        // epochs strictly inside the accepted step: the step's own
        // interpolant (the predictor of radau.rs:252-278 at s = (e - t)/h)
        while (ej < a.n_epochs) {
          const double e = a.epochs[ej];
          const double se = (e - t) / h;
          if (!(se > 0.0 && se < 1.0) || fabs(e - (t + h)) < 1e-12) break;
          double ox[3], ov[3];
          double bw[3] = {0.0, 0.0, 0.0}, bu[3] = {0.0, 0.0, 0.0};
          double sp = 1.0;
#pragma unroll
          for (int k = 0; k < 7; ++k) {
            sp *= se;
            const double wk = sp * RTC.w[k], uk = sp * RTC.u[k];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              bw[d] += B[k][d] * wk;
              bu[d] += B[k][d] * uk;
            }
          }
          const double xv[3] = {x.x, x.y, x.z}, vv[3] = {v.x, v.y, v.z};
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            ox[d] = xv[d] + se * vv[d] * h + (se * se) * (h * h) * (as[d] / 2.0 + bw[d]);
            ov[d] = vv[d] + se * h * (as[d] + bu[d]);
          }
          const int64_t q = (int64_t)ej * a.n + obj;
          a.mout[0][q] = ox[0];
          a.mout[1][q] = ox[1];
          a.mout[2][q] = ox[2];
          a.mout[3][q] = ov[0];
          a.mout[4][q] = ov[1];
          a.mout[5][q] = ov[2];
          ++ej;
        }
      }
      x = Vec3{xs[0], xs[1], xs[2]};
      v = Vec3{vs[0], vs[1], vs[2]};
    }
    if (conv) {
      // next step size (radau.rs:339-342)
      double fm = fmax(fmax(fmx[0], fmx[1]), fmx[2]);
      double bm = fmax(fmax(fabs(B[6][0]), fabs(B[6][1])), fabs(B[6][2]));
      double ratio = root7(eps * fm / bm);
      ratio = ratio < 0.25 ? 0.25 : (ratio > 4.0 ? 4.0 : ratio);
      h_next = h * ratio;
    }
    t = conv ? t + h : t;
    const bool take_a = st == ST_OK && !sub;
    a0 = Vec3{take_a ? acc.x : a0.x, take_a ? acc.y : a0.y, take_a ? acc.z : a0.z};

    // ---- scalar control (radau.rs:160-218) ----
    int32_t done = -1;
    bool begin = false, next_step = false;
    if (st != ST_OK) {
      done = st;  // Impact and DAFLimits abort the integration (radau.rs:207-208)
    } else if (phase == PH_INIT) {
      if (fabs(tf - t) < 1e-10) {
        done = ST_OK;
      } else {
        h = copysign(0.1, tf - t);
        begin = true;
      }
    } else if (phase == PH_EXACT) {
      h = h_next;
      if (fabs(t - tf) < 1e-12) {
        done = ST_OK;
      } else {
        fails = 0;
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
        next_step = true;
      }
    } else if (j < 7) {
      j = j + 1;
    } else if (conv) {
      n_acc++;
      phase = PH_EXACT;
    } else if (sweep + 1 < 10) {
      sweep++;
      j = 1;
      n_sweep++;
    } else {
      // the step failed to converge (radau.rs:345): retry at 0.7 h
      n_fail++;
      fails++;
      h *= 0.7;
      if (fails > 10) {
        done = ST_CONVERGENCE;
      } else {
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    }
    if (MULTI && (done == ST_OK || (conv && done < 0))) {
      // the epochs at the current time: after an accepted step (one that
      // ends on an epoch) and at the end of a run
      while (ej < a.n_epochs && fabs(a.epochs[ej] - t) < 1e-12) {
        const int64_t q = (int64_t)ej * a.n + obj;
        a.mout[0][q] = x.x;
        a.mout[1][q] = x.y;
        a.mout[2][q] = x.z;
        a.mout[3][q] = v.x;
        a.mout[4][q] = v.y;
        a.mout[5][q] = v.z;
        ++ej;
      }
    }
    if (MULTI && done == ST_OK) {
      // end of a run of epochs in one direction: integrate on to the end of
      // the next run
      if (ej < a.n_epochs) {
        tf = a.epochs[multi_run_end(a, ej, t)];
        // the controller's proposal: before the last step was shortened to
        // land on the epoch, else after it
        const double hc = phase == PH_INIT ? 0.1 : (shortened ? h_want : h);
        h = copysign(fmax(fabs(hc), 0.005), tf - t);
        fails = 0;
        done = -1;
        begin = true;
      }
    } else if (MULTI && done > 0) {
      // a failing epoch and the later ones are undefined
      const double nan = __builtin_nan("");
      for (int32_t e = ej; e < a.n_epochs; ++e)
        for (int c = 0; c < 6; ++c) a.mout[c][(int64_t)e * a.n + obj] = nan;
    }
    if (begin) {
      if (MULTI) h_want = h;
      shortened = fabs(t - tf) <= fabs(h);
      if (shortened) h = tf - t;
      sweep = 0;
      j = 1;
      n_sweep++;
      phase = PH_SUB;
    }

    // ---- tail hand-off: with few live lanes left in the wavefront, an
    // object with much work to go leaves at the start of a step, its state
    // parked in LDS, and the wavefront finishes the parked objects sixteen
    // lanes each once its other lanes are done (the remaining work is
    // estimated from the evaluations so far and the fraction of the arc
    // covered) ----
    bool parked = false;
    if constexpr (PARKABLE) {
      if (next_step && a.park_lanes > 0 && __popcll(live_mask) <= a.park_lanes) {
        const double t0 = a.jd0[obj];
        const double covered = (t - t0) / (tf - t0);
        if (covered > 0.0 && (double)n_force * (1.0 - covered) > a.park_min_left * covered) {
          const int w = threadIdx.x >> 6;
          const int k = atomicAdd(&kh_park_n[w], 1);
          if (k < PARK_SLOTS) {
            ParkRec& r = kh_park_rec[w][k];
            r.obj = obj;
            r.n_force = n_force;
            r.n_sweep = n_sweep;
            r.n_acc = n_acc;
            r.n_fail = n_fail;
            r.t = t;
            r.tf = tf;
            r.h = h;
            r.x[0] = x.x;
            r.x[1] = x.y;
            r.x[2] = x.z;
            r.v[0] = v.x;
            r.v[1] = v.y;
            r.v[2] = v.z;
            r.a[0] = a0.x;
            r.a[1] = a0.y;
            r.a[2] = a0.z;
#pragma unroll
            for (int q = 0; q < 7; ++q)
#pragma unroll
              for (int d = 0; d < 3; ++d) r.B[q][d] = B[q][d];
            r.ej = ej;
            r.shortened = shortened ? 1 : 0;
            r.h_want = h_want;
            parked = true;
          }
        }
      }
    }

    // ---- finished (or parked) objects: store, then take the next one (the
    // plain launch has one object per lane: the lane is done) ----
    if (done >= 0 || parked) {
      // SSB-equatorial result; the post-pass converts it back
      if (!parked)
        store_result(a, obj, done, x, v, impact, te, int4{n_force, n_sweep, n_acc, n_fail});
      if constexpr (!PERSISTENT) break;
      int64_t nxt = a.n;
      if (PERSISTENT) nxt = a.first_dynamic + (int64_t)atomicAdd(a.queue, 1ull);
      double nt, ntf;
      Vec3 nx, nv;
      obj = fetch<PERSISTENT>(a, nxt, nt, ntf, nx, nv);
      t = nt;
      tf = MULTI ? a.epochs[multi_run_end(a, 0, nt)] : ntf;
      x = nx;
      v = nv;
      phase = PH_INIT;
      j = 1;
      sweep = 0;
      fails = 0;
      ej = 0;
      n_force = n_sweep = n_acc = n_fail = 0;
    }
    if constexpr (PERSISTENT) {
      const bool fresh = done >= 0 || parked;
#pragma unroll
      for (int k = 0; k < 7; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) B[k][d] = fresh ? 0.0 : B[k][d];
    }
  }
  // ---- the wavefront's parked objects, sixteen lanes each ----
  if constexpr (PARKABLE) {
    const int w = threadIdx.x >> 6;
    const int g = (threadIdx.x & 63) / WIDE_L;
    const int np = kh_park_n[w];
    // the window as the wavefront's last move left it (a lane that ended
    // earlier holds an older first record)
    if (WWABLE && ww) ww0 = kh_ww_w0[w];
    if (g < np) {
      const ParkRec& r = kh_park_rec[w][g];
      wide_integrate<true, AST, MULTI>(spk, coef, a, lds, r.obj, r, false,
                                       threadIdx.x & (WIDE_L - 1), g * WIDE_L, nullptr, winw,
                                       wwin, ww0);
    }
  }
}

// ---------------------------------------------------------------------------
// Three lanes per object.  Lane c of a group owns coordinate c of the
// integrator (x_c, v_c, a_c, B[7]_c, G[7]_c), so its state is a third of the
// one-lane kernel's and more wavefronts fit on a SIMD.  The force evaluation
// is split across the group by body -- lane 0: Sun, Mercury, Venus, Mars;
// lane 1: Jupiter, Saturn, Uranus, Neptune; lane 2: the Earth-Moon
// barycentre, Earth, Moon -- and the three partial accelerations are summed
// with wavefront shuffles.  Convergence and step-size maxima are reduced over
// the group, so the three lanes always take the same control path.  A
// wavefront holds 21 groups; its last lane idles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gshfl(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int gshfl(int v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ double group_max(double v, int base) {
  return fmax(fmax(gshfl(v, base), gshfl(v, base + 1)), gshfl(v, base + 2));
}

// kete index of the body in slot k of lane c (-1: none); PLANETS order is
// 10, 1, 2, 399, 301, 4, 5, 6, 7, 8 (gravity.rs:145-148)
__device__ __forceinline__ int slot_body(int c, int k) {
  const int t0[4] = {0, 1, 2, 5};
  const int t1[4] = {6, 7, 8, 9};
  const int t2[4] = {3, 4, -1, -1};
  return c == 0 ? t0[k] : (c == 1 ? t1[k] : t2[k]);
}

template <bool PERSISTENT>
__device__ __forceinline__ int64_t fetch3(const NBodyArgs& a, int64_t slot, int c, int base,
                                          double& t, double& tf, double& x, double& v) {
  while (slot < a.slot_end) {
    const int64_t i = a.perm != nullptr ? (int64_t)a.perm[slot] : slot;
    if (a.status[i] == ST_OK) {
      t = a.jd0[i];
      tf = a.jd_final[a.jd_final_stride ? i : 0];
      x = a.out[c][i];
      v = a.out[3 + c][i];
      return i;
    }
    if (!PERSISTENT) return a.n;
    unsigned long long k = 0;
    if (c == 0) k = atomicAdd(a.queue, 1ull);
    const int lo = gshfl((int)(k & 0xffffffffull), base);
    const int hi = gshfl((int)(k >> 32), base);
    slot = a.first_dynamic + (int64_t)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
  }
  return a.n;
}

template <bool PERSISTENT, bool AST = false>
__global__ void __launch_bounds__(256) radau3_nbody_kernel(const SpkDev* __restrict__ spk,
                                                           const double* __restrict__ coef,
                                                           NBodyArgs a) {
  __shared__ LdsTables lds;
  for (int k = threadIdx.x; k < 8 * 8; k += blockDim.x) {
    const int r = k / 8, cc = k % 8;
    lds.rdiff[r][cc] = RT.rdiff[r][cc];
    if (r == 0) {
      lds.h[cc] = RT.h[cc];
      lds.inv_h[cc] = RT.inv_h[cc];
    }
  }
  if (a.window > 0) {
    if (threadIdx.x < spk->n_seg)
      kh_win_w0[threadIdx.x] = window_first_record(a, spk->seg[threadIdx.x]);
    __syncthreads();
    const int per_seg = a.window * 44;
    for (int k = threadIdx.x; k < spk->n_seg * per_seg; k += blockDim.x) {
      const int sg = k / per_seg, rem = k - sg * per_seg;
      const SegDev& g = spk->seg[sg];
      const int r = kh_win_w0[sg] + rem / 44;
      // word 1 of a record (the radius) is staged as its reciprocal
      const double w = r < g.n_rec ? coef[g.off + (int64_t)r * 44 + rem % 44] : 1.0;
      kh_win_rec[k] = rem % 44 == 1 ? 1.0 / w : w;
    }
  }
  __syncthreads();
  const int32_t winw = a.window;

  const int lane = threadIdx.x & 63;
  const int gw = lane / 3;
  const int c = lane - 3 * gw;
  const int base = 3 * gw;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const double eps = a.epsilon;
  int64_t obj = gw < 21 ? a.slot_begin + wave * 21 + gw : a.n;
  double t = 0.0, tf = 0.0, h = 0.0, h_next = 0.0, x = 0.0, v = 0.0, a0 = 0.0;
  double B[7], G[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) B[k] = G[k] = 0.0;
  int32_t phase = PH_INIT, j = 1, sweep = 0, fails = 0;
  int32_t n_force = 0, n_sweep = 0, n_acc = 0, n_fail = 0;
  if (gw < 21) obj = fetch3<PERSISTENT>(a, obj, c, base, t, tf, x, v);

  while (obj < a.n) {
    const bool sub = phase == PH_SUB;
    // ---- predictor for this lane's coordinate (radau.rs:252-278) ----
    const double s = lds.h[j];
    double xe, ve;
    {
      double bw = 0.0, bu = 0.0, sp = 1.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        sp *= s;
        bw += B[k] * (sp * RTC.w[k]);
        bu += B[k] * (sp * RTC.u[k]);
      }
      const double px = x + s * v * h + (s * s) * (h * h) * (a0 / 2.0 + bw);
      const double pv = v + s * h * (a0 + bu);
      xe = sub ? px : x;
      ve = sub ? pv : v;
    }
    const double te = sub ? t + s * h : t;
    const Vec3 X{gshfl(xe, base), gshfl(xe, base + 1), gshfl(xe, base + 2)};
    const Vec3 V{gshfl(ve, base), gshfl(ve, base + 1), gshfl(ve, base + 2)};

    // ---- this lane's share of the force evaluation ----
    n_force++;
    const NonGravDev* ngp = (a.ng && a.ng[obj].kind != 0) ? &a.ng[obj] : nullptr;
    Vec3 part{0.0, 0.0, 0.0};
    int key = 0x7fffffff;  // (kete index << 4) | status of this lane's first failure
    Vec3 emb{0.0, 0.0, 0.0};
    if (c == 2) {
      const int32_t se = seg_pos(spk, coef, spk->emb_seg, te, emb, winw);
      if (se) key = (3 << 4) | se;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int b = slot_body(c, k);
      if (b >= 0 && key == 0x7fffffff) {
        const int32_t sb = add_body(spk, coef, spk->body[b], te, X, V, !sub, ngp, emb, part, winw);
        if (sb) key = (b << 4) | sb;
      }
    }
    // the MASSES asteroids (kete indices 10..14) on the lane with the fewest bodies
    if (AST && c == 2 && key == 0x7fffffff) {
      Vec3 sun;
      const int32_t ss = seg_pos(spk, coef, spk->sun_seg, te, sun, winw);
      if (ss) key = ss;  // the Sun (index 0) was read first
#pragma unroll 1
      for (int k = 0; k < N_ASTEROIDS && key == 0x7fffffff; ++k) {
        const AstDev& ad = spk->ast[k];
        Vec3 ap, unused;
        bool sun_centred = false;
        int32_t sa = t21_state<false>(spk, ad.seg_begin, ad.seg_end, te, ap, unused, sun_centred);
        if (sa == ST_OK) {
          if (sun_centred) ap = Vec3{ap.x + sun.x, ap.y + sun.y, ap.z + sun.z};
          sa = add_body_at(ad.bd, ap, Vec3{0.0, 0.0, 0.0}, X, V, !sub, nullptr, part);
        }
        if (sa) key = ((N_PLANETS + k) << 4) | sa;
      }
    }
    const int kmin = min(min(gshfl(key, base), gshfl(key, base + 1)), gshfl(key, base + 2));
    const int32_t st = kmin == 0x7fffffff ? ST_OK : (kmin & 15);
    // a_c = sum over the group of the partials' component c
    const double p0 = c == 0 ? part.x : (c == 1 ? part.y : part.z);
    const double q1 = c == 0 ? part.z : (c == 1 ? part.x : part.y);  // = part[(c+2)%3]
    const double q2 = c == 0 ? part.y : (c == 1 ? part.z : part.x);  // = part[(c+1)%3]
    const int l1 = base + (c == 2 ? 0 : c + 1);
    const int l2 = base + (c == 0 ? 2 : c - 1);
    const double ev = p0 + gshfl(q1, l1) + gshfl(q2, l2);

    // ---- G column j (radau.rs:292-308) ----
    {
      double gk = (ev - a0) * lds.inv_h[j];
#pragma unroll
      for (int z = 1; z < 7; ++z) {
        const double rd = lds.rdiff[j][z];
        gk = z < j ? (gk - G[z - 1]) * rd : gk;
      }
#pragma unroll
      for (int q = 0; q < 7; ++q) G[q] = (sub && q == j - 1) ? gk : G[q];
    }
    // ---- end of sweep: B = G*C, convergence (radau.rs:313-319) ----
    const bool es = sub && j == 7 && st == ST_OK;
    const double bp6 = B[6];
#pragma unroll
    for (int cc = 0; cc < 7; ++cc) {
      double acc2 = G[cc];
#pragma unroll
      for (int k = cc + 1; k < 7; ++k) acc2 += G[k] * RTC.c[k][cc];
      B[cc] = es ? acc2 : B[cc];
    }
    const double fmx = fabs(ev) + 1e-6;
    const double rmax = group_max(fabs(B[6] - bp6) / fmx, base);
    const bool conv = es && rmax < 1e-14;
    // ---- accepted step (radau.rs:320-338) ----
    {
      double bw = 0.0, bu = 0.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        bw += B[k] * RTC.w[k];
        bu += B[k] * RTC.u[k];
      }
      const double nx = x + (v * h + (h * h) * (a0 * 0.5 + bw));
      const double nv = v + h * (a0 + bu);
      x = conv ? nx : x;
      v = conv ? nv : v;
    }
    if (conv) {
      const double fm = group_max(fmx, base);
      const double bm = group_max(fabs(B[6]), base);
      double ratio = root7(eps * fm / bm);
      ratio = ratio < 0.25 ? 0.25 : (ratio > 4.0 ? 4.0 : ratio);
      h_next = h * ratio;
    }
    t = conv ? t + h : t;
    a0 = (st == ST_OK && !sub) ? ev : a0;

    // ---- scalar control (radau.rs:160-218), identical in the group ----
    int32_t done = -1;
    bool begin = false;
    if (st != ST_OK) {
      done = st;
    } else if (phase == PH_INIT) {
      if (fabs(tf - t) < 1e-10) {
        done = ST_OK;
      } else {
        h = copysign(0.1, tf - t);
        begin = true;
      }
    } else if (phase == PH_EXACT) {
      h = h_next;
      if (fabs(t - tf) < 1e-12) {
        done = ST_OK;
      } else {
        fails = 0;
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    } else if (j < 7) {
      j = j + 1;
    } else if (conv) {
      n_acc++;
      phase = PH_EXACT;
    } else if (sweep + 1 < 10) {
      sweep++;
      j = 1;
      n_sweep++;
    } else {
      n_fail++;
      fails++;
      h *= 0.7;
      if (fails > 10) {
        done = ST_CONVERGENCE;
      } else {
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    }
    if (begin) {
      if (fabs(t - tf) <= fabs(h)) h = tf - t;
      sweep = 0;
      j = 1;
      n_sweep++;
      phase = PH_SUB;
    }

    if (done >= 0) {
      const double nan = __builtin_nan("");
      const bool ok = done == ST_OK;
      a.out[c][obj] = ok ? x : nan;
      a.out[3 + c][obj] = ok ? v : nan;
      if (c == 0) {
        a.status[obj] = done;
        if (a.impact_id)
          a.impact_id[obj] = done != ST_IMPACT ? 0
                             : (kmin >> 4) < N_PLANETS ? spk->body[kmin >> 4].id
                                                       : spk->ast[(kmin >> 4) - N_PLANETS].bd.id;
        if (a.impact_jd) a.impact_jd[obj] = done == ST_IMPACT ? te : nan;
        if (a.counters)
          reinterpret_cast<int4*>(a.counters)[obj] = int4{n_force, n_sweep, n_acc, n_fail};
      }
      int64_t nxt = a.n;
      if (PERSISTENT) {
        unsigned long long k = 0;
        if (c == 0) k = atomicAdd(a.queue, 1ull);
        const int lo = gshfl((int)(k & 0xffffffffull), base);
        const int hi = gshfl((int)(k >> 32), base);
        nxt = a.first_dynamic + (int64_t)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
      }
      double nt, ntf, nx2, nv2;
      obj = fetch3<PERSISTENT>(a, nxt, c, base, nt, ntf, nx2, nv2);
      t = nt;
      tf = ntf;
      x = nx2;
      v = nv2;
      phase = PH_INIT;
      j = 1;
      sweep = 0;
      fails = 0;
      n_force = n_sweep = n_acc = n_fail = 0;
    }
    const bool fresh = done >= 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) B[k] = fresh ? 0.0 : B[k];
  }
}

__global__ void zero_u64(unsigned long long* p) { *p = 0ull; }

// ---------------------------------------------------------------------------
// Wide latency variant for the slowest objects (the head of the longest-
// first order, beside the main launch): four objects per wavefront, each on
// sixteen lanes (wide_integrate), each lane's record cached in LDS.  One
// wavefront per block and at most 256 VGPRs, so a SIMD holds one of these
// waves beside a wave of the main launch (the main launch's blocks need 256
// VGPRs per wave and the LDS window, a block here 16 KB of LDS).
// ---------------------------------------------------------------------------
constexpr int WIDE_BLOCK = 64;
__global__ void __launch_bounds__(WIDE_BLOCK) __attribute__((amdgpu_waves_per_eu(2, 8)))
radau_wide_kernel(const SpkDev* __restrict__ spk, const double* __restrict__ coef, NBodyArgs a) {
  __shared__ LdsTables lds;
  __shared__ __align__(16) double s_rec[WIDE_BLOCK / WIDE_L][WIDE_EMB + 1][44];
  for (int k = threadIdx.x; k < 8 * 8; k += blockDim.x) {
    const int r = k / 8, c = k % 8;
    lds.rdiff[r][c] = RT.rdiff[r][c];
    if (r == 0) {
      lds.h[c] = RT.h[c];
      lds.inv_h[c] = RT.inv_h[c];
    }
  }
  __syncthreads();
  slot_range(a);
  const int64_t slot = a.slot_begin + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WIDE_L;
  if (slot >= a.slot_end) return;
  const int64_t obj = a.perm != nullptr ? (int64_t)a.perm[slot] : slot;
  ParkRec st0;
  if (a.stage != nullptr) {
    const double* r = a.stage + 8 * obj;
    if (isnan(r[6])) return;
    for (int k = 0; k < 3; ++k) {
      st0.x[k] = r[k];
      st0.v[k] = r[3 + k];
    }
    st0.t = r[6];
    st0.tf = r[7];
  } else {
    if (a.status[obj] != ST_OK) return;
    st0.t = a.jd0[obj];
    st0.tf = a.jd_final[a.jd_final_stride ? obj : 0];
    for (int k = 0; k < 3; ++k) {
      st0.x[k] = a.out[k][obj];
      st0.v[k] = a.out[3 + k][obj];
    }
  }
  const int l = threadIdx.x & (WIDE_L - 1);
  const int g0 = (threadIdx.x & 63) & ~(WIDE_L - 1);
  double* const rslot = &s_rec[threadIdx.x / WIDE_L][l <= WIDE_EMB ? l : 0][0];
  wide_integrate<false>(spk, coef, a, lds, obj, st0, true, l, g0, rslot, 0, nullptr, 0);
}

// Blocks of `kernel` that are resident at once on the whole device: the
// persistent launch must not exceed it, or the blocks beyond it start only
// after the first ones have drained the queue.
template <typename K>
static int64_t resident_grid(K kernel, int block, size_t lds_bytes, int32_t n_cu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds_bytes) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (int64_t)per_cu * (n_cu > 0 ? n_cu : 1);
}

// Calls f with std::true_type when the force list includes the asteroids,
// std::false_type otherwise (selects the kernel instantiation).
template <class F>
static void with_ast(int32_t n_ast, F&& f) {
  if (n_ast) f(std::true_type{});
  else f(std::false_type{});
}

int32_t launch_nbody(const SpkDev* spk, const double* coef, const NBodyLaunch& p,
                     hipStream_t stream, int32_t* used_block, int32_t* used_grid,
                     int32_t* used_mode, hipEvent_t ev0, hipEvent_t ev1) {
  NBodyArgs a;
  a.n = p.n;
  a.jd0 = p.jd0;
  for (int k = 0; k < 6; ++k) {
    a.xyz[k] = p.xyz[k];
    a.out[k] = p.out[k];
  }
  a.jd_final = p.jd_final;
  a.jd_final_stride = p.jd_final_stride;
  a.frame = p.frame;
  a.center = p.center;
  a.ng = reinterpret_cast<const NonGravDev*>(p.ng);
  a.epsilon = p.epsilon;
  a.status = p.status;
  a.impact_id = p.impact_id;
  a.impact_jd = p.impact_jd;
  a.counters = p.counters;
  a.queue = p.queue;
  a.range = p.range;
  a.perm = nullptr;
  a.slot_begin = 0;
  a.slot_end = p.n;
  a.split = nullptr;
  a.part = 0;
  a.window = p.range != nullptr ? p.window : 0;
  const size_t lds_bytes = a.window > 0 ? (size_t)p.n_seg * a.window * 44 * sizeof(double) : 0;
  const int block = p.block > 0 ? p.block : 256;
  const unsigned g1 = (unsigned)((p.n + 255) / 256);
  a.n_ast = p.n_ast;
  a.epochs = p.epochs;
  a.n_epochs = p.n_epochs;
  for (int k = 0; k < 6; ++k) a.mout[k] = p.mout[k];
  const bool multi = p.epochs != nullptr && p.n_epochs > 0;
  // tail hand-off: the plain launch
  a.park_lanes = p.mode == 0 ? p.park_lanes : 0;
  a.park_min_left = p.park_min_left;
  // object records: the plain launch over the longest-first order
  a.stage = p.stage != nullptr && p.mode == 0 && !multi && p.order_buf != nullptr &&
                    p.n < (int64_t)1 << 31
                ? p.stage
                : nullptr;
  if (a.range)
    hipLaunchKernelGGL(init_range, dim3(1), dim3(64), 0, stream, a.range,
                       multi ? p.ep_lo : __builtin_inf(), multi ? p.ep_hi : 0.0);
  hipLaunchKernelGGL(nbody_prepare_kernel, dim3(g1), dim3(256), 0, stream, spk, coef, a);
  if (p.order_buf != nullptr && p.mode != 3 && p.n < (int64_t)1 << 31) {
    // longest-first order of the persistent queue
    uint32_t* hist = static_cast<uint32_t*>(p.order_buf);
    uint32_t* off = hist + ORDER_BUCKETS;
    uint32_t* perm = off + ORDER_BUCKETS;
    uint16_t* bkt = reinterpret_cast<uint16_t*>(perm + p.n);
    hipLaunchKernelGGL(order_zero_kernel, dim3(1), dim3(256), 0, stream, hist);
    // per-object final epochs with a sort buffer: radix sort by cost bucket
    // and arc length (kh_sort.hip) instead of the counting sort's scatter
    const bool rsort = p.sort_buf != nullptr && p.jd_final_stride != 0 && !multi;
    hipLaunchKernelGGL(order_hist_kernel, dim3(g1), dim3(256), 0, stream, a, bkt, hist,
                       rsort ? order_sort_keys(p.sort_buf, p.n) : nullptr,
                       rsort ? order_sort_vals(p.sort_buf, p.n) : nullptr);
    int64_t* split = p.mode == 0 && p.latency_k > 0 ? p.split : nullptr;
    const double lanes = (double)resident_grid(radau_nbody_kernel<false>, block, lds_bytes,
                                               p.n_cu) * block;
    // KH_LATENCY_SHARE scales the resident lane's share above which an
    // object goes to the latency kernel (1 by default)
    double share = 1.0;
    if (const char* e = getenv("KH_LATENCY_SHARE")) share = atof(e) > 0.0 ? atof(e) : 1.0;
    hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(256), 0, stream, hist, off, split,
                       lanes / share, p.latency_k);
    if (rsort) {
      if (launch_order_sort(p.n, p.sort_buf, p.sort_bytes, perm, stream)) return -1;
    } else {
      hipLaunchKernelGGL(order_scatter_kernel, dim3(g1), dim3(256), 0, stream, p.n, bkt, off,
                         perm);
    }
    a.perm = perm;
  }
  int64_t grid;
  if (p.mode == 2 || p.mode == 3) {
    // three lanes per object, 21 objects per wavefront
    const int64_t waves_needed = (p.n + 20) / 21;
    const int64_t blocks_needed = (waves_needed * 64 + block - 1) / block;
    const bool persistent = p.mode == 2;
    grid = persistent ? (p.grid > 0 ? p.grid
                                    : resident_grid(radau3_nbody_kernel<true>, block, lds_bytes,
                                                    p.n_cu))
                      : blocks_needed;
    if (grid > blocks_needed) grid = blocks_needed;
    a.first_dynamic = grid * (block / 64) * 21;
    if (persistent) hipLaunchKernelGGL(zero_u64, dim3(1), dim3(1), 0, stream, p.queue);
    if (ev0) (void)hipEventRecord(ev0, stream);
    with_ast(p.n_ast, [&](auto ast) {
      constexpr bool A = decltype(ast)::value;
      if (persistent)
        hipLaunchKernelGGL((radau3_nbody_kernel<true, A>), dim3((unsigned)grid), dim3(block),
                           lds_bytes, stream, spk, coef, a);
      else
        hipLaunchKernelGGL((radau3_nbody_kernel<false, A>), dim3((unsigned)grid), dim3(block),
                           lds_bytes, stream, spk, coef, a);
    });
    if (ev1) (void)hipEventRecord(ev1, stream);
  } else if (p.mode == 1) {
    grid = p.grid > 0 ? p.grid
                      : resident_grid(radau_nbody_kernel<true>, block, lds_bytes, p.n_cu);
    if (grid * block > p.n) grid = (p.n + block - 1) / block;
    a.first_dynamic = grid * block;
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(1), 0, stream, p.queue);
    if (ev0) (void)hipEventRecord(ev0, stream);
    with_ast(p.n_ast, [&](auto ast) {
      constexpr bool A = decltype(ast)::value;
      hipLaunchKernelGGL((radau_nbody_kernel<true, false, false, A>), dim3((unsigned)grid),
                         dim3(block), lds_bytes, stream, spk, coef, a);
    });
    if (ev1) (void)hipEventRecord(ev1, stream);
  } else {
    a.first_dynamic = p.n;
    // the slowest objects (the head of the longest-first order) go to the
    // latency variant on the second stream, concurrently with the rest
    const bool hybrid = !multi && a.perm != nullptr && p.latency_k > 0 && p.split &&
                        p.stream2 && p.fork && p.join;
    // the sixteen-lane latency kernel goes on the launch stream and the main
    // kernel on the second one, behind the fork wait: the latency kernel's
    // one-wavefront blocks are then dispatched first and hold their place on
    // the SIMDs beside the main launch's (uneven arcs: 340 ms, 383 ms with
    // the main kernel dispatched first); the one-lane variant goes on the
    // second stream
    const bool wide = p.wide_latency && p.n_ast == 0;
    hipStream_t s_lat = stream, s_main = stream;
    if (ev0) (void)hipEventRecord(ev0, stream);
    if (hybrid) {
      // up to latency_k slots; the kernel reads how many from *split
      NBodyArgs al = a;
      al.split = p.split;
      al.part = 1;
      al.window = 0;
      const int64_t k_max = p.latency_k < p.n ? p.latency_k : p.n;
      (void)hipEventRecord(p.fork, stream);
      (void)hipStreamWaitEvent(p.stream2, p.fork, 0);
      if (wide)
        s_main = p.stream2;
      else
        s_lat = p.stream2;
      if (wide)
        hipLaunchKernelGGL(radau_wide_kernel,
                           dim3((unsigned)((k_max * WIDE_L + WIDE_BLOCK - 1) / WIDE_BLOCK)),
                           dim3(WIDE_BLOCK), 0, s_lat, spk, coef, al);
      else
        with_ast(p.n_ast, [&](auto ast) {
          constexpr bool A = decltype(ast)::value;
          hipLaunchKernelGGL((radau_nbody_kernel<false, true, false, A>),
                             dim3((unsigned)((k_max + 63) / 64)), dim3(64), 0, s_lat, spk,
                             coef, al);
        });
      a.split = p.split;
      a.part = 2;
    }
    grid = (p.n + block - 1) / block;
    with_ast(p.n_ast, [&](auto ast) {
      constexpr bool A = decltype(ast)::value;
      if (multi)
        hipLaunchKernelGGL((radau_nbody_kernel<false, false, true, A>), dim3((unsigned)grid),
                           dim3(block), lds_bytes, s_main, spk, coef, a);
      else if (p.jd_final_stride != 0 && !A)
        hipLaunchKernelGGL((radau_nbody_kernel<false, false, false, false, true>),
                           dim3((unsigned)grid), dim3(block), lds_bytes, s_main, spk, coef, a);
      else
        hipLaunchKernelGGL((radau_nbody_kernel<false, false, false, A>), dim3((unsigned)grid),
                           dim3(block), lds_bytes, s_main, spk, coef, a);
    });
    if (hybrid) (void)hipEventRecord(p.join, p.stream2);
    if (hybrid) (void)hipStreamWaitEvent(stream, p.join, 0);
    if (ev1) (void)hipEventRecord(ev1, stream);
  }
  hipLaunchKernelGGL(nbody_finish_kernel, dim3(g1), dim3(256), 0, stream, spk, coef, a);
  if (multi) {
    const int64_t nq = (int64_t)p.n_epochs * p.n;
    hipLaunchKernelGGL(nbody_finish_multi_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256),
                       0, stream, spk, coef, a);
  }
  if (used_block) *used_block = block;
  if (used_grid) *used_grid = (int32_t)grid;
  if (used_mode) *used_mode = p.mode;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// propagate_n_body_long (src/kete/rust/propagation.rs:202-246): each batch of
// particles is integrated with kete's SIMPLE_PLANETS as one Radau state
// (propagate_n_body_vec, propagation/mod.rs:142-237), the bodies pulling on
// each other and on the particles (vec_accel, acceleration.rs:197-249).
// One lane per body of the batch, in kete's order (the nine planets, then the
// particles); a batch takes a group of G = 9 + batch lanes, packed
// floor(64 / G) groups to a wavefront (three of 19 lanes at kete's default
// batch of 10), or a whole wavefront when G > 32.  Every integrator decision
// is the group's: convergence is the largest |dB6|/(|a|+1e-6) over all its
// components, the step ratio uses the largest |a|+1e-6 and |B6|
// (radau.rs:313-342), and an impact anywhere ends the batch.  The massive
// bodies' positions at each evaluation point reach the other lanes by
// shuffles.
// ---------------------------------------------------------------------------
// SIMPLE_PLANETS (gravity.rs:150-154, EM_BARY gravity.rs:137-143):
// id, mass, radius, J2 and the terms of GravParams::add_acceleration.
__constant__ BodyDev kSimplePlanets[9] = {
    {-1, 0, 10, 1, GMS, 0.004654758765894654, SUN_J2, 0.0},
    {-1, 0, 1, 0, 1.66012082548908e-07 * GMS, 1.63139354095098e-05, 0.0, 0.0},
    {-1, 0, 2, 0, 2.44783828779694e-06 * GMS, 4.04537843465442e-05, 0.0, 0.0},
    {-1, 0, 3, 0, (3.00348961546514e-06 + 3.69430335010988e-08) * GMS, 385000.0 / AU_KM, 0.0,
     0.0},
    {-1, 0, 4, 0, 3.22715608291416e-07 * GMS, 2.27021279387769e-05, 0.0, 0.0},
    {-1, 0, 5, 1, 0.0009547919099414246 * GMS, 0.00047789450254521576, JUPITER_J2, 0.0},
    {-1, 0, 6, 0, 0.00028588567002459455 * GMS, 0.0004028666966848747, 0.0, 0.0},
    {-1, 0, 7, 0, 4.36624961322212e-05 * GMS, 0.0001708513622580592, 0.0, 0.0},
    {-1, 0, 8, 0, 5.15138377265457e-05 * GMS, 0.0001655371154958558, 0.0, 0.0},
};

// Maximum over a group of lanes [base, base + G) (G <= 64; li = lane - base):
// an inclusive max-scan from the group's first lane (each step takes the
// lane `o` below when it is in the group), read at the group's last lane.
// fmax skips NaN as nalgebra's max does.
__device__ __forceinline__ double group_maxd(double v, int li, int last) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(v, o, 64);
    v = li >= o ? fmax(v, u) : v;
  }
  return __shfl(v, last, 64);
}
__device__ __forceinline__ int group_maxi(int v, int li, int last) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    v = li >= o ? max(v, u) : v;
  }
  return __shfl(v, last, 64);
}

__global__ void __launch_bounds__(256) radau_long_kernel(LongLaunch a) {
  __shared__ LdsTables lds;
  for (int k = threadIdx.x; k < 8 * 8; k += blockDim.x) {
    const int r = k / 8, cc = k % 8;
    lds.rdiff[r][cc] = RT.rdiff[r][cc];
    if (r == 0) {
      lds.h[cc] = RT.h[cc];
      lds.inv_h[cc] = RT.inv_h[cc];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int gl = 9 + a.batch;            // lanes per group
  const int gpw = 64 / gl;               // groups per wavefront
  const int slot = min(lane / gl, gpw);  // gpw: the wavefront's spare lanes
  const int li = lane - slot * gl;       // index of the body in its batch
  const int base = lane - li;            // first lane of the group
  const int last = min(base + gl - 1, 63);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t gid = wave * gpw + slot;
  const int64_t n_batches = (a.n + a.batch - 1) / a.batch;
  const int64_t p0 = gid * a.batch;     // first particle of the batch
  const int m = (slot < gpw && gid < n_batches) ? (int)min((int64_t)a.batch, a.n - p0) : 0;
  // this lane carries a body of a batch (groups past the last batch, m == 0,
  // carry none: their planet lanes must not load or store)
  const bool body = m > 0 && li < 9 + m;
  const int64_t pi = p0 + (li - 9);     // particle index (li >= 9)
  Vec3 x{0, 0, 0}, v{0, 0, 0}, a0{0, 0, 0};
  if (body && li < 9) {
    const double* ps = a.planets + 6 * li;
    x = Vec3{ps[0], ps[1], ps[2]};
    v = Vec3{ps[3], ps[4], ps[5]};
  } else if (body) {
    x = Vec3{a.xyz[0][pi], a.xyz[1][pi], a.xyz[2][pi]};
    v = Vec3{a.xyz[3][pi], a.xyz[4][pi], a.xyz[5][pi]};
    if (a.frame == 17) {  // State::try_change_frame_mut to Equatorial
      x = ecl_to_equ(x);
      v = ecl_to_equ(v);
    }
  }
  // kete's vec_accel applies a particle's non-grav model for idx > n_massive
  const NonGravDev* ngp = nullptr;
  if (body && li > 9 && a.ng) {
    const NonGravDev* g = reinterpret_cast<const NonGravDev*>(a.ng) + pi;
    if (g->kind != 0) ngp = g;
  }
  double B[7][3], G[7][3];
#pragma unroll
  for (int k = 0; k < 7; ++k) B[k][0] = B[k][1] = B[k][2] = G[k][0] = G[k][1] = G[k][2] = 0.0;
  double t = a.jd0, h = 0.0, h_next = 0.0;
  const double tf = a.jd_final;
  int32_t phase = PH_INIT, j = 1, sweep = 0, fails = 0;
  int32_t n_force = 0, n_sweep = 0, n_acc = 0, n_fail = 0;
  bool live = m > 0;  // the group's batch is still integrating
  int32_t status = ST_OK;
  while (__any(live)) {
    const bool sub = phase == PH_SUB;
    const double s = lds.h[j];
    double xe[3], ve[3];
    {
      const double xs[3] = {x.x, x.y, x.z}, vs[3] = {v.x, v.y, v.z};
      const double as[3] = {a0.x, a0.y, a0.z};
      double bw[3] = {0.0, 0.0, 0.0}, bu[3] = {0.0, 0.0, 0.0};
      double sp = 1.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        sp *= s;
        const double wk = sp * RTC.w[k], uk = sp * RTC.u[k];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          bw[d] += B[k][d] * wk;
          bu[d] += B[k][d] * uk;
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double px = xs[d] + s * vs[d] * h + (s * s) * (h * h) * (as[d] / 2.0 + bw[d]);
        const double pv = vs[d] + s * h * (as[d] + bu[d]);
        xe[d] = sub ? px : xs[d];
        ve[d] = sub ? pv : vs[d];
      }
    }
    // ---- vec_accel at the evaluation point ----
    const bool exact = !sub;
    Vec3 acc{0.0, 0.0, 0.0};
    int32_t st = ST_OK;
    n_force += live ? 1 : 0;
#pragma unroll 1
    for (int b = 0; b < 9; ++b) {
      const Vec3 bp{__shfl(xe[0], base + b, 64), __shfl(xe[1], base + b, 64),
                    __shfl(xe[2], base + b, 64)};
      // velocities enter only the GR term of the Sun and Jupiter (kind 1);
      // the branch is uniform, so every lane takes part in the shuffles
      Vec3 bv{0.0, 0.0, 0.0};
      if (kSimplePlanets[b].kind == 1)
        bv = Vec3{__shfl(ve[0], base + b, 64), __shfl(ve[1], base + b, 64),
                  __shfl(ve[2], base + b, 64)};
      if (b == li || !body) continue;
      Vec3 w{0.0, 0.0, 0.0};
      const int32_t sb = add_body_at(kSimplePlanets[b], bp, bv, Vec3{xe[0], xe[1], xe[2]},
                                     Vec3{ve[0], ve[1], ve[2]}, exact, ngp, w);
      if (sb && st == ST_OK) st = sb;
      acc = Vec3{acc.x + w.x, acc.y + w.y, acc.z + w.z};
    }
    const int32_t gst = group_maxi(live ? st : 0, li, last);
    const double ev[3] = {acc.x, acc.y, acc.z};
    // ---- G column j (radau.rs:292-308) ----
    {
      const double as[3] = {a0.x, a0.y, a0.z};
      double gk[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) gk[d] = (ev[d] - as[d]) * lds.inv_h[j];
#pragma unroll
      for (int z = 1; z < 7; ++z) {
        const double rd = lds.rdiff[j][z];
#pragma unroll
        for (int d = 0; d < 3; ++d) gk[d] = z < j ? (gk[d] - G[z - 1][d]) * rd : gk[d];
      }
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        const bool put = sub && q == j - 1;
#pragma unroll
        for (int d = 0; d < 3; ++d) G[q][d] = put ? gk[d] : G[q][d];
      }
    }
    // ---- end of sweep: B = G*C, the group's convergence test ----
    const bool es = sub && j == 7 && gst == ST_OK;
    const double bp6[3] = {B[6][0], B[6][1], B[6][2]};
#pragma unroll
    for (int c = 0; c < 7; ++c) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double acc2 = G[c][d];
#pragma unroll
        for (int k = c + 1; k < 7; ++k) acc2 += G[k][d] * RTC.c[k][c];
        B[c][d] = es ? acc2 : B[c][d];
      }
    }
    double rmax = 0.0, fmx = 0.0, bmx = 0.0;
    if (body) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double fm = fabs(ev[d]) + 1e-6;
        rmax = fmax(rmax, fabs(B[6][d] - bp6[d]) / fm);
        fmx = fmax(fmx, fm);
        bmx = fmax(bmx, fabs(B[6][d]));
      }
    }
    rmax = group_maxd(rmax, li, last);
    fmx = group_maxd(fmx, li, last);
    bmx = group_maxd(bmx, li, last);
    const bool conv = es && rmax < 1e-14;
    // ---- accepted step (radau.rs:320-338) ----
    {
      double xs[3] = {x.x, x.y, x.z}, vs[3] = {v.x, v.y, v.z};
      const double as[3] = {a0.x, a0.y, a0.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double bw = 0.0, bu = 0.0;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          bw += B[k][d] * RTC.w[k];
          bu += B[k][d] * RTC.u[k];
        }
        const double nx = xs[d] + (vs[d] * h + (h * h) * (as[d] * 0.5 + bw));
        const double nv = vs[d] + h * (as[d] + bu);
        xs[d] = conv ? nx : xs[d];
        vs[d] = conv ? nv : vs[d];
      }
      x = Vec3{xs[0], xs[1], xs[2]};
      v = Vec3{vs[0], vs[1], vs[2]};
    }
    if (conv) {
      double ratio = root7(1e-6 * fmx / bmx);
      ratio = ratio < 0.25 ? 0.25 : (ratio > 4.0 ? 4.0 : ratio);
      h_next = h * ratio;
    }
    t = conv ? t + h : t;
    const bool take_a = gst == ST_OK && !sub;
    a0 = Vec3{take_a ? acc.x : a0.x, take_a ? acc.y : a0.y, take_a ? acc.z : a0.z};
    // ---- control (radau.rs:160-218), the same for the whole group ----
    int32_t done = -1;
    bool begin = false;
    if (gst != ST_OK) {
      done = gst;
    } else if (phase == PH_INIT) {
      if (fabs(tf - t) < 1e-10) {
        done = ST_OK;
      } else {
        h = copysign(0.1, tf - t);
        begin = true;
      }
    } else if (phase == PH_EXACT) {
      h = h_next;
      if (fabs(t - tf) < 1e-12) {
        done = ST_OK;
      } else {
        fails = 0;
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    } else if (j < 7) {
      j = j + 1;
    } else if (conv) {
      n_acc += live ? 1 : 0;
      phase = PH_EXACT;
    } else if (sweep + 1 < 10) {
      sweep++;
      j = 1;
      n_sweep += live ? 1 : 0;
    } else {
      n_fail += live ? 1 : 0;
      fails++;
      h *= 0.7;
      if (fails > 10) {
        done = ST_CONVERGENCE;
      } else {
        if (fabs(h) < 0.005) h = copysign(0.005, h);
        begin = true;
      }
    }
    if (begin) {
      if (fabs(t - tf) <= fabs(h)) h = tf - t;
      sweep = 0;
      j = 1;
      n_sweep += live ? 1 : 0;
      phase = PH_SUB;
    }
    if (live && done >= 0) {
      live = false;
      status = done;
    }
  }
  // states relative to the Sun's final state (propagation/mod.rs:224-233)
  const Vec3 sx{__shfl(x.x, base, 64), __shfl(x.y, base, 64), __shfl(x.z, base, 64)};
  const Vec3 sv{__shfl(v.x, base, 64), __shfl(v.y, base, 64), __shfl(v.z, base, 64)};
  if (!body) return;
  const bool ok = status == ST_OK;
  const double nan = __builtin_nan("");
  const Vec3 rx = ok ? Vec3{x.x - sx.x, x.y - sx.y, x.z - sx.z} : Vec3{nan, nan, nan};
  const Vec3 rv = ok ? Vec3{v.x - sv.x, v.y - sv.y, v.z - sv.z} : Vec3{nan, nan, nan};
  if (li < 9) {
    if (a.planets_out) {
      double* po = a.planets_out + (gid * 9 + li) * 6;
      po[0] = rx.x; po[1] = rx.y; po[2] = rx.z;
      po[3] = rv.x; po[4] = rv.y; po[5] = rv.z;
    }
  } else {
    a.out[0][pi] = rx.x;
    a.out[1][pi] = rx.y;
    a.out[2][pi] = rx.z;
    a.out[3][pi] = rv.x;
    a.out[4][pi] = rv.y;
    a.out[5][pi] = rv.z;
  }
  if (li == 0) {
    a.status[gid] = status;
    if (a.counters) {
      int32_t* c = a.counters + 4 * gid;
      c[0] = n_force; c[1] = n_sweep; c[2] = n_acc; c[3] = n_fail;
    }
  }
}

int32_t launch_nbody_long(const LongLaunch& p, hipStream_t stream) {
  if (p.n <= 0 || p.batch <= 0 || p.batch > 55) return -1;
  const int64_t nb = (p.n + p.batch - 1) / p.batch;
  const int64_t gpw = 64 / (9 + p.batch);
  const int64_t waves = (nb + gpw - 1) / gpw;
  const unsigned grid = (unsigned)((waves + 3) / 4);
  hipLaunchKernelGGL(radau_long_kernel, dim3(grid), dim3(256), 0, stream, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t upload_radau_tables(const RadauTables& t) {
  return hipMemcpyToSymbol(HIP_SYMBOL(RT), &t, sizeof(RadauTables)) == hipSuccess ? 0 : -1;
}

}  // namespace kh
