// kh_internal.h -- structures shared between the host side of libkete_hip
// (kh_api.cpp) and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kh_device.h"

namespace kh {

// Gauss-Radau tables (radau.rs:24-68) plus precomputed factors:
// wpow[j][k] = h_j^(k+1) * w_k, upow[j][k] = h_j^(k+1) * u_k (radau.rs:243-256),
// inv_h[j] = 1/h_j and rdiff[j][z] = 1/(h_j - h_z) for the G update.
struct RadauTables {
  double h[8];
  double w[7];
  double u[7];
  double c[7][7];
  double wpow[8][7];
  double upow[8][7];
  double inv_h[8];
  double rdiff[8][8];
};

void fill_radau_tables(RadauTables& t);

// The same tables as compile-time constants, for the fixed-index uses (C, W,
// U), which then become instruction literals instead of registers.
constexpr RadauTables make_tables() {
  RadauTables t{};
  const double H[8] = {0.0, 0.05626256053692215, 0.18024069173689236, 0.3526247171131696,
                       0.5471536263305554, 0.7342101772154105, 0.8853209468390958,
                       0.9775206135612875};
  for (int i = 0; i < 8; ++i) t.h[i] = H[i];
  for (int i = 0; i < 7; ++i) {
    t.w[i] = 1.0 / (double)((i + 2) * (i + 3));
    t.u[i] = 1.0 / (double)(i + 2);
  }
  for (int i = 0; i < 7; ++i) t.c[i][i] = 1.0;
  for (int i = 0; i < 7; ++i) {
    if (i > 0) t.c[i][0] = -H[i] * t.c[i - 1][0];
    for (int j = 1; j < i; ++j) t.c[i][j] = t.c[i - 1][j - 1] - H[i] * t.c[i - 1][j];
  }
  return t;
}
constexpr RadauTables RTC = make_tables();
int32_t upload_radau_tables(const RadauTables& t);

struct NBodyLaunch {
  int64_t n;
  const double* jd0;
  const double* xyz[6];
  const double* jd_final;
  int32_t jd_final_stride;
  int32_t frame;
  Chain center;
  const void* ng;
  double epsilon;
  double* out[6];
  int32_t* status;
  int64_t* impact_id;
  double* impact_jd;
  int32_t* counters;
  unsigned long long* queue;
  int32_t mode;   // 0 plain, 1 persistent
  int32_t block;  // threads per block (0: default)
  int32_t grid;   // persistent grid size in blocks (0: default)
  unsigned long long* range;  // [4] epoch-range scratch, or null: no LDS window
  int32_t window;             // records per segment staged in LDS
  int32_t n_seg;
  void* order_buf;  // scratch for the longest-first queue order (order_buf_bytes), or null
  int32_t n_cu;     // compute units of the device (sizes the persistent grid)
  // mode 0: the first latency_k objects of the order run in the latency
  // kernel on stream2 (fork/join events order it with `stream`)
  int64_t latency_k;
  // the latency kernel: sixteen lanes per object (planets only) instead of
  // one lane per object
  int32_t wide_latency;
  int64_t* split;  // device int64: head of the order the latency kernel takes
  hipStream_t stream2;
  hipEvent_t fork, join;
  // multi-epoch propagation (mode 0 only): n_epochs shared epochs in
  // integration order, spanning [ep_lo, ep_hi]; states at each in mout
  // [6][n_epochs][n], in the input center and frame
  const double* epochs;
  int32_t n_epochs;
  double ep_lo, ep_hi;
  double* mout[6];
  // force list: 0 PLANETS, N_ASTEROIDS MASSES (include_asteroids)
  int32_t n_ast;
  // tail hand-off of the plain launch (park_lanes = 0: off):
  // when a wavefront is down to park_lanes (<= 4) live lanes, an object that
  // by its progress so far has more than park_min_left force evaluations to
  // go is parked at its next step boundary, and the wavefront finishes its
  // parked objects on sixteen lanes each once its other lanes are done
  int32_t park_lanes;
  double park_min_left;
  // [n][8] doubles, 64-byte aligned: the object records of the plain launch
  // over the longest-first order (stage_buf_bytes), or null
  double* stage;
  // order_sort_temp_bytes(n) bytes: the order is radix-sorted by cost bucket
  // and arc length (per-object final epochs), or null: counting-sorted by
  // cost bucket
  void* sort_buf;
  size_t sort_bytes;
};

// Shared epochs of a multi-epoch propagation (device array, integration
// order) and the outputs [6][n_epochs][n].
struct MultiEpochs {
  const double* epochs;
  int32_t n_epochs;
  double lo, hi;  // smallest and largest epoch
  double* out[6];
};

// An object parked by the plain N-body launch at a step boundary: the
// integrator state a new step starts from (radau.rs:183-215 between two
// calls of step): time, final time, the step about to be tried, position,
// velocity, acceleration, the B table carried over as the predictor, and the
// work counters so far.  B is [k][d].
struct ParkRec {
  int64_t obj;
  int32_t n_force, n_sweep, n_acc, n_fail;
  double t, tf, h;
  double x[3], v[3], a[3];
  double B[7][3];
  // multi-epoch propagation: the next epoch to store, whether the step was
  // shortened to land on the run's end, the step proposed before that
  int32_t ej, shortened;
  double h_want;
};

// Bytes of NBodyLaunch::order_buf for n objects.
constexpr size_t order_buf_bytes(int64_t n) { return 2 * 2048 * 4 + (size_t)n * 6 + 256; }
// Radix sort of the order (kh_sort.hip): NBodyLaunch::sort_buf holds the keys
// and object indices the order pre-pass writes, the sorted keys and hipcub's
// scratch, order_sort_temp_bytes(n) in all.
constexpr size_t order_sort_span(int64_t n) { return ((size_t)n * 4 + 255) & ~(size_t)255; }
inline uint32_t* order_sort_keys(void* buf, int64_t n) { return static_cast<uint32_t*>(buf); }
inline uint32_t* order_sort_vals(void* buf, int64_t n) {
  return reinterpret_cast<uint32_t*>(static_cast<char*>(buf) + order_sort_span(n));
}
size_t order_sort_temp_bytes(int64_t n);
int32_t launch_order_sort(int64_t n, void* buf, size_t bytes, uint32_t* perm, hipStream_t s);
// Bytes of NBodyLaunch::stage for n objects (one 64-byte record each).
constexpr size_t stage_buf_bytes(int64_t n) { return (size_t)n * 64; }

// Pre-pass, integrator and post-pass on `stream`; ev0/ev1 (nullable) bracket
// the integrator kernel alone.
// propagate_n_body_long: batches of `batch` particles, each integrated with
// the nine SIMPLE_PLANETS as one Radau state (one lane per body, a group of
// 32 or 64 lanes per batch).
struct LongLaunch {
  int64_t n;          // particles
  int32_t batch;      // particles per batch (<= 55)
  int32_t frame;      // input frame (1 or 17)
  double jd0, jd_final;
  const double* xyz[6];     // particles at jd0, Sun-centred, in `frame`
  const double* planets;    // [9][6] Sun-centred equatorial at jd0
  const void* ng;           // [n] kh_nongrav or null
  double* out[6];           // particles at jd_final, Sun-centred equatorial
  double* planets_out;      // [n_batches][9][6] or null
  int32_t* status;          // [n_batches]
  int32_t* counters;        // [n_batches][4] or null
};
int32_t launch_nbody_long(const LongLaunch& p, hipStream_t stream);

// kete moid for n orbit pairs (kh_moid.hip): state a (jd_a, xa) and state b
// (jd_b or jd_a, xb), Sun-centred in `frame`; st_b: the b states' lookup
// status or null.  best_a/best_b [7][n] (jd, pos, vel) carry the grid's
// closest pair from the first kernel to the Nelder-Mead one.
// The Earth's MOID grid at one epoch, shared by the pairs at that epoch.
struct MoidSharedGrid {
  double key;      // the epoch (jd_a of the pairs served)
  int32_t st, pad;
  double step;     // grid step (days)
  double peri_jd;  // epoch of the perihelion state
  double g[6][100];  // grid states x y z vx vy vz
};
struct MoidLaunch {
  int64_t n;
  int32_t frame;
  int32_t max_iter;
  const double* jd_a;
  const double* xa[6];
  const double* jd_b;
  const double* xb[6];
  const int32_t* st_b;
  double* best_a[7];
  double* best_b[7];
  double* moid;
  double* grid_min;  // or null
  int32_t* status;
  int32_t* iters;    // or null
  MoidSharedGrid* sg;  // the second orbit is the Earth at jd_a: its grid at jd_a[0]
  int32_t* work;       // [2][n] Kepler solves and Newton iterations per pair, or null
};
int32_t launch_moid(const MoidLaunch& p, hipStream_t stream);

int32_t launch_nbody(const SpkDev* spk, const double* coef, const NBodyLaunch& p,
                     hipStream_t stream, int32_t* used_block, int32_t* used_grid,
                     int32_t* used_mode, hipEvent_t ev0, hipEvent_t ev1);

struct TwoBodyLaunch {
  int64_t n, m;
  const double* jd0;
  const double* xyz[6];
  const double* targets;
  int32_t has_observer;
  double observer[3];
  int32_t center_sun;  // 1: center 10 (no conversion); 0: center 0 (SSB)
  int32_t sun_seg;
  int32_t frame;
  double* out[6];
  int32_t* status;
  int32_t* iters;  // Newton iterations per solve [m][n], or null
};

int32_t launch_two_body(const SpkDev* spk, const double* coef, const TwoBodyLaunch& p,
                        hipStream_t stream);

// Every stride-th object of a batch (m of them) into contiguous arrays;
// jd_final broadcast when jf_stride is 0.
int32_t launch_gather_states(int64_t m, int64_t stride, const double* jd0,
                             const double* const xyz[6], const double* jd_final,
                             int32_t jf_stride, double* o_jd0, double* const o_xyz[6],
                             double* o_jf, hipStream_t stream);

// State of `body` relative to `center` (NAIF id center_id) at m epochs.
// With t21_begin < t21_end the body is the object whose Type 21 segments are
// spk.t21[t21_begin, t21_end) instead, and `sun` the Sun's chain (added for
// Sun-centred segments when the center differs).
int32_t launch_spk_states(const SpkDev* spk, const double* coef, Chain body, Chain center,
                          int32_t t21_begin, int32_t t21_end, Chain sun, int64_t center_id,
                          int32_t frame, int64_t m, const double* jd, double* const out[6],
                          int32_t* status, hipStream_t stream);

// One rectangular field: epoch, observer state (center 10, same frame as the
// objects) and the four edge normals (fov/patches.rs:122-155).
struct FieldDev {
  double jd;
  double obs[6];
  double nrm[12];
  double pad;
};

// fields one prefilter launch stages in LDS (160 B each); the host splits
// larger groups into tiles of this size
constexpr int FOV_MAX_TILE_FIELDS = 256;

struct FovPrefilter {
  int64_t n;
  const double* jd;
  const double* x[6];
  const FieldDev* fields;
  int32_t nf;
  double dt_limit;
  unsigned long long* n_cand;
  int64_t cap;
  int32_t* cand_field;
  int32_t* cand_kind;
  int64_t* cand_obj;
};

struct FovTwoBody {
  const unsigned long long* n_cand;  // the prefilter's count, read on the device
  unsigned long long* cand_max;      // largest count seen (overflow check)
  int64_t cand_cap;
  const double* jd;
  const double* x[6];
  const FieldDev* fields;
  double dt_limit;
  const int32_t* cand_field;
  const int32_t* cand_kind;
  const int64_t* cand_obj;
  unsigned long long* n_hit;
  int64_t hit_cap;
  int32_t field_base;  // sorted index of fields[0]; hits record field_base + f
  int32_t* hit_field;
  int64_t* hit_obj;
  double* hit_state[7];
  // n-body re-checks, appended over all groups: the field (sorted index),
  // the object, and the state the re-check starts from with its epoch and
  // the field's epoch as the target
  unsigned long long* n_nb;
  int64_t nb_cap;
  int32_t* nb_field;
  int64_t* nb_obj;
  double* g_jd0;
  double* g_jdf;
  double* g_x[6];
};

struct FovNBody {
  int64_t m;
  const FieldDev* fields;  // all fields, in sorted order
  const int32_t* nb_field;
  const int64_t* nb_obj;
  const double* g_out[6];
  const int32_t* g_status;
  unsigned long long* n_hit;
  int64_t hit_cap;
  int32_t* hit_field;
  int64_t* hit_obj;
  double* hit_state[7];
};

struct RadecArgs {
  int64_t n;
  const double* st[6];
  const double* obs[6];
  const double* h;
  double g;
  int32_t frame;
  double* out[5];
};

int32_t launch_fov_prefilter(const FovPrefilter& p, hipStream_t s);
int32_t launch_fov_two_body(const FovTwoBody& p, hipStream_t s);
int32_t launch_fov_nb_finish(const FovNBody& p, hipStream_t s);
int32_t launch_radec_vmag(const RadecArgs& p, hipStream_t s);

// Where an SPK object's state comes from (fov_spk_check): kind 0 not in
// the ephemeris, 1 its Type 21 segments t21[t21_begin, t21_end), 2 the
// planet-tree chain `body` to the barycentre.
struct SpkObjDesc {
  int32_t kind, t21_begin, t21_end, pad;
  Chain body;
};

struct FovSpk {
  int64_t n_obj;           // objects per field (pairs: n_fields x n_obj, objects fastest)
  int64_t n_fields;
  const SpkObjDesc* obj;
  Chain sun;               // the Sun's chain (states are Sun-centred)
  const FieldDev* fields;  // sorted by epoch
  unsigned long long* n_hit;
  int64_t hit_cap;
  int32_t* hit_field;      // sorted field index
  int64_t* hit_obj;        // object index
  double* hit_state[7];
};
int32_t launch_fov_spk(const SpkDev* spk, const double* coef, const FovSpk& p, hipStream_t s);

// Hits in kete's order -- sorted field index, then object index -- by a
// device radix sort on packed keys, gathered into out_* with the original
// field index order[sorted index].  temp/temp_bytes: scratch of at least
// fov_sort_temp_bytes(n_hits) bytes.
size_t fov_sort_temp_bytes(int64_t n_hits);
int32_t launch_fov_sort_hits(int64_t n_hits, int32_t obj_bits, int32_t field_bits,
                             const int32_t* hit_field, const int64_t* hit_obj,
                             double* const hit_state[7], const int64_t* order, void* temp,
                             size_t temp_bytes, int64_t* out_field, int64_t* out_obj,
                             double* const out_state[7], hipStream_t s);

// Observables and partial derivatives (kh_grad.hip).
struct GradLaunch {
  int64_t n;
  int32_t m;
  int32_t frame;
  int32_t force;  // 0 spk_accel (PLANETS), 1 central_accel
  int32_t restart;
  double epsilon;
  const double* jd0;
  const double* xyz[6];
  const void* ng;
  const double* jd_obs;
  const double* obs[3];
  Chain sun;
  double* radec[2];
  double* jac[2];
  double* state[6];
  double* dstate[6];
  int32_t* status;
  int32_t* counters;
  // working memory of the wide kernel (partials_buf_bytes)
  double* ep_state;  // [m][n][10][6]
  double* ep_sun;    // [m][n][6]
  int32_t* ep_done;  // [n]
  int32_t* lt_fail;  // [n]
};

// Bytes of the wide partials kernel's working memory for n objects, m epochs.
constexpr size_t partials_buf_bytes(int64_t n, int32_t m) {
  return ((size_t)m * (size_t)n * 66 * 8 + 255) / 256 * 256 + 2 * (((size_t)n * 4 + 255) / 256 * 256);
}

int32_t upload_grad_tables(const RadauTables& t);
int32_t launch_radau_partials(const SpkDev* spk, const double* coef, const GradLaunch& g,
                              hipStream_t stream);

}  // namespace kh
