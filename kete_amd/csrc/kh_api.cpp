// kh_api.cpp -- host side of libkete_hip.so: the C ABI declared in
// include/kete_hip.h.  Owns the device, the stream, the uploaded ephemeris and
// the staging buffers; validates arguments; launches the kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kh_ctx.h"

using namespace kh;

// Scratch of kh_propagate_n_body_dev_noalloc: the longest-first order of at
// most KH_NOALLOC_MAX_N objects (histogram, offsets, permutation, buckets).
// A device array built into the library, so that entry point allocates
// nothing; every GPU the library runs on holds its own copy until the process
// ends.  Host code never passes this name to a kernel or copy: it uses the
// device address hipGetSymbolAddress returns, taken once per device at
// context creation (noalloc_scratch below).
constexpr size_t kNoallocScratchBytes = order_buf_bytes(KH_NOALLOC_MAX_N);
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per device
}  // namespace

static thread_local std::string g_err;

int32_t kh_fail(int32_t code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define fail kh_fail

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

namespace kh {
void fill_radau_tables(RadauTables& t) {
  // GAUSS_RADAU_SPACINGS (radau.rs:24-33)
  static const double H[8] = {0.0,
                              0.05626256053692215,
                              0.18024069173689236,
                              0.3526247171131696,
                              0.5471536263305554,
                              0.7342101772154105,
                              0.8853209468390958,
                              0.9775206135612875};
  std::memset(&t, 0, sizeof(t));
  for (int i = 0; i < 8; ++i) t.h[i] = H[i];
  // W_VEC, U_VEC, C_MAT (radau.rs:35-68)
  for (int i = 0; i < 7; ++i) {
    t.w[i] = 1.0 / (double)((i + 2) * (i + 3));
    t.u[i] = 1.0 / (double)(i + 2);
  }
  for (int i = 0; i < 7; ++i) t.c[i][i] = 1.0;
  for (int i = 0; i < 7; ++i) {
    if (i > 0) t.c[i][0] = -H[i] * t.c[i - 1][0];
    for (int j = 1; j < i; ++j) t.c[i][j] = t.c[i - 1][j - 1] - H[i] * t.c[i - 1][j];
  }
  for (int j = 1; j < 8; ++j) {
    for (int k = 0; k < 7; ++k) {
      // powi(1 + k) by repeated squaring
      int e = k + 1;
      double a = H[j], r = 1.0;
      while (true) {
        if (e & 1) r *= a;
        e /= 2;
        if (!e) break;
        a *= a;
      }
      t.wpow[j][k] = r * t.w[k];
      t.upow[j][k] = r * t.u[k];
    }
    t.inv_h[j] = 1.0 / H[j];
    for (int z = 1; z < j; ++z) t.rdiff[j][z] = 1.0 / (H[j] - H[z]);
  }
}
}  // namespace kh

extern "C" {

const char* kh_last_error(void) { return g_err.c_str(); }

int32_t kh_ctx_create(int32_t device, kh_ctx** out) {
  if (!out) return fail(KH_E_ARG, "out is null");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(KH_E_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= n) return fail(KH_E_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  kh_ctx* c = new kh_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return fail(KH_E_HIP, "hipGetDeviceProperties failed");
  }
  c->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipMalloc(&c->d_queue, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->d_counts, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->d_range, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->d_split, sizeof(int64_t)) != hipSuccess) {
    delete c;
    return fail(KH_E_HIP, "stream/event/queue allocation failed");
  }

  RadauTables t;
  fill_radau_tables(t);
  if (upload_radau_tables(t) != 0 || upload_grad_tables(t) != 0) {
    delete c;
    return fail(KH_E_HIP, "upload of Radau tables failed");
  }
  // device address of the no-alloc scratch on this device (current above)
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < kMaxDevices && !g_scratch_dev[device]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        delete c;
        return fail(KH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(e));
      }
      g_scratch_dev[device] = p;
    }
    c->noalloc_scratch = device < kMaxDevices ? g_scratch_dev[device] : nullptr;
  }
  *out = c;
  return KH_OK;
}

int32_t kh_ctx_destroy(kh_ctx* c) {
  if (!c) return KH_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_coef) (void)hipFree(c->d_coef);
  if (c->d_t21) (void)hipFree(c->d_t21);
  if (c->d_spk) (void)hipFree(c->d_spk);
  if (c->dbuf) (void)hipFree(c->dbuf);
  if (c->obuf) (void)hipFree(c->obuf);
  if (c->fbuf) (void)hipFree(c->fbuf);
  if (c->hbuf) (void)hipFree(c->hbuf);
  if (c->mbuf) (void)hipFree(c->mbuf);
  if (c->lbuf) (void)hipFree(c->lbuf);
  if (c->pbuf) (void)hipFree(c->pbuf);
  if (c->d_counts) (void)hipFree(c->d_counts);
  if (c->d_range) (void)hipFree(c->d_range);
  if (c->d_queue) (void)hipFree(c->d_queue);
  if (c->d_split) (void)hipFree(c->d_split);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipEvent_t e : c->fov_ev) (void)hipEventDestroy(e);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return KH_OK;
}

void* kh_ctx_stream(kh_ctx* c) { return c ? (void*)c->stream : nullptr; }

static int find_seg(const kh_ctx* c, int64_t obj) {
  for (size_t k = 0; k < c->seg_obj.size(); ++k)
    if (c->seg_obj[k] == obj) return (int)k;
  return -1;
}

// Segments whose states sum to the SSB state of `id` (center chaining of
// spice/spk.rs:99-126 for the Type 2 planet tree).
bool kh_build_chain(const kh_ctx* c, int64_t id, Chain& ch) {
  ch.n = 0;
  int64_t cur = id;
  while (cur != 0) {
    int s = find_seg(c, cur);
    if (s < 0 || ch.n >= MAX_CHAIN) return false;
    ch.seg[ch.n++] = s;
    cur = c->seg_center[s];
  }
  return true;
}

int32_t kh_spk_upload_type2(kh_ctx* c, const kh_spk2_seg* segs, int64_t nseg) {
  if (!c || !segs || nseg <= 0) return fail(KH_E_ARG, "null context or no segments");
  if (nseg > MAX_SEGS) return fail(KH_E_ARG, "too many segments (max 24)");
  HIPCHK(hipSetDevice(c->device));
  SpkDev spk = c->spk;  // keeps the Type 21 part
  spk.coef = nullptr;
  spk.n_seg = spk.all14 = spk.shared_grid = spk.emb_seg = spk.sun_seg = 0;
  std::memset(spk.seg, 0, sizeof(spk.seg));
  std::memset(spk.body, 0, sizeof(spk.body));
  int64_t total = 0;
  bool all14 = true;
  c->seg_obj.clear();
  c->seg_center.clear();
  for (int64_t k = 0; k < nseg; ++k) {
    const kh_spk2_seg& s = segs[k];
    if (!s.records || s.n_rec <= 0 || s.rsize < 8 || (s.rsize - 2) % 3 != 0 ||
        (s.rsize - 2) / 3 > MAX_COEF)
      return fail(KH_E_ARG, "segment " + std::to_string(k) + ": bad record layout");
    if (s.frame != KH_FRAME_EQUATORIAL)
      return fail(KH_E_ARG, "segment " + std::to_string(k) + ": only frame 1 (equatorial) is supported");
    if (!(s.intlen_s > 0.0) || !(s.jd_end >= s.jd_start))
      return fail(KH_E_ARG, "segment " + std::to_string(k) + ": bad time range");
    SegDev& g = spk.seg[k];
    g.jd_start = s.jd_start;
    g.jd_end = s.jd_end;
    g.j0s = (s.jd_start - 2451545.0) * 86400.0;  // jd_to_spice_jd (spice/mod.rs:30-34)
    g.intlen = s.intlen_s;
    g.off = total;
    g.n_rec = (int32_t)s.n_rec;
    g.rsize = s.rsize;
    g.n_coef = (s.rsize - 2) / 3;
    if (g.n_coef != 14) all14 = false;
    total += s.n_rec * s.rsize;
    c->seg_obj.push_back(s.obj_id);
    c->seg_center.push_back(s.center_id);
  }
  spk.n_seg = (int32_t)nseg;
  spk.all14 = all14 ? 1 : 0;
  // one record grid for all segments: same start, length, count, and
  // bitwise-identical record mid-points and radii
  bool shared = all14;
  for (int64_t k = 1; shared && k < nseg; ++k) {
    const SegDev &a0 = spk.seg[0], &ak = spk.seg[k];
    if (ak.jd_start != a0.jd_start || ak.jd_end != a0.jd_end || ak.j0s != a0.j0s ||
        ak.intlen != a0.intlen || ak.n_rec != a0.n_rec) {
      shared = false;
      break;
    }
    for (int64_t r = 0; r < segs[k].n_rec; ++r) {
      const double* p0 = segs[0].records + r * segs[0].rsize;
      const double* pk = segs[k].records + r * segs[k].rsize;
      if (std::memcmp(p0, pk, 2 * sizeof(double)) != 0) {
        shared = false;
        break;
      }
    }
  }
  spk.shared_grid = shared ? 1 : 0;
  if (const char* sg = getenv("KH_SHARED_GRID")) spk.shared_grid = shared && atoi(sg) != 0;
  // PLANETS (gravity.rs:43-101,145-148): id, mass, radius, J2 and the
  // special terms of GravParams::add_acceleration (gravity.rs:171-216).
  struct PB { int64_t id; double mass, radius, j2; int kind; };
  const double GMS_ = GMS;
  const PB pl[N_PLANETS] = {
      {10, GMS_, 0.004654758765894654, SUN_J2, 1},
      {1, 1.66012082548908e-07 * GMS_, 1.63139354095098e-05, 0.0, 0},
      {2, 2.44783828779694e-06 * GMS_, 4.04537843465442e-05, 0.0, 0},
      {399, 3.00348961546514e-06 * GMS_, 4.33036684926559e-05, EARTH_J2, 2},
      {301, 3.69430335010988e-08 * GMS_, 1.16138016662292e-05, 0.0, 0},
      {4, 3.22715608291416e-07 * GMS_, 2.27021279387769e-05, 0.0, 0},
      {5, 0.0009547919099414246 * GMS_, 0.00047789450254521576, JUPITER_J2, 1},
      {6, 0.00028588567002459455 * GMS_, 0.0004028666966848747, 0.0, 0},
      {7, 4.36624961322212e-05 * GMS_, 0.0001708513622580592, 0.0, 0},
      {8, 5.15138377265457e-05 * GMS_, 0.0001655371154958558, 0.0, 0},
  };
  spk.emb_seg = -1;
  for (int b = 0; b < N_PLANETS; ++b) {
    int s = find_seg(c, pl[b].id);
    if (s < 0) return fail(KH_E_NO_SPK, "no segment for body " + std::to_string(pl[b].id));
    BodyDev& bd = spk.body[b];
    bd.seg = s;
    bd.emb = 0;
    bd.id = (int32_t)pl[b].id;
    bd.kind = pl[b].kind;
    bd.mass = pl[b].mass;
    bd.radius = pl[b].radius;
    bd.j2 = pl[b].j2;
    int64_t cen = c->seg_center[s];
    if (cen == 3) {
      int e = find_seg(c, 3);
      if (e < 0 || c->seg_center[e] != 0)
        return fail(KH_E_NO_SPK, "no segment for the Earth-Moon barycentre (3) w.r.t. 0");
      bd.emb = 1;
      spk.emb_seg = e;
    } else if (cen != 0) {
      return fail(KH_E_NO_SPK, "body " + std::to_string(pl[b].id) + " must be w.r.t. 0 or 3");
    }
  }
  if (spk.emb_seg < 0) spk.emb_seg = spk.body[0].seg;
  spk.sun_seg = spk.body[0].seg;
  if (c->d_coef) {
    HIPCHK(hipFree(c->d_coef));
    c->d_coef = nullptr;
  }
  HIPCHK(hipMalloc(&c->d_coef, total * sizeof(double)));
  for (int64_t k = 0; k < nseg; ++k)
    HIPCHK(hipMemcpy(c->d_coef + spk.seg[k].off, segs[k].records,
                     segs[k].n_rec * segs[k].rsize * sizeof(double), hipMemcpyHostToDevice));
  spk.coef = c->d_coef;
  if (!c->d_spk) HIPCHK(hipMalloc(&c->d_spk, sizeof(SpkDev)));
  HIPCHK(hipMemcpy(c->d_spk, &spk, sizeof(SpkDev), hipMemcpyHostToDevice));
  c->spk = spk;
  c->has_spk = true;
  return KH_OK;
}

// ---------------------------------------------------------------------------
// SPK Type 21
// ---------------------------------------------------------------------------
// kete's MASSES after PLANETS (gravity.rs:103-134): id, mass (GM in
// AU^3/day^2), radius (AU).
static const struct {
  int64_t id;
  double mass, radius;
} kAsteroids[N_ASTEROIDS] = {
    {20000001, 4.7191659919436e-10 * GMS, 6.276827e-6},
    {20000002, 1.0259143964958e-10 * GMS, 3.415824e-6},
    {20000004, 1.3026452498655e-10 * GMS, 3.512082e-6},
    {20000010, 4.3953391300849e-11 * GMS, 2.894426e-6},
    {20000704, 1.911e-11 * GMS, 2.219283e-6},
};

// Rewrites one Type 21 record (kete's layout, spk_segments.rs:1010-1018) as
// the device record: the modified-difference line of SpkSegmentType21::
// try_get_pos_vel (spk_segments.rs:1032-1089) is a polynomial in dt; its
// coefficients, in tau = dt / S with S the largest |dt| the record serves
// ([lo, hi] in seconds), are formed here once in long double by running the
// same W recursion on polynomials instead of numbers.  Returns false (why set)
// for records kete cannot evaluate (a zero step size) or this layout cannot
// hold (an order above T21_NP).
static bool t21_prepare(const double* rec, int n, double lo, double hi, double* out,
                        std::string& why) {
  typedef long double L;
  constexpr int MAXD = 32;
  const int K = (int)rec[4 * n + 7];  // KQMAX1
  int kq[3];
  for (int c = 0; c < 3; ++c) kq[c] = (int)rec[4 * n + 8 + c];
  if (n > 25 || K < 2 || K > n + 1) {
    why = "MAXDIM or KQMAX1 out of range";
    return false;
  }
  for (int c = 0; c < 3; ++c)
    if (kq[c] < 0 || kq[c] > K - 1 || kq[c] > T21_NP) {
      why = "difference order out of range (at most " + std::to_string(T21_NP) + ")";
      return false;
    }
  const double* g = rec + 1;
  for (int j = 0; j < K - 2; ++j)
    if (g[j] == 0.0) {
      why = "zero step size";
      return false;
    }
  const double t_ref = rec[0];
  double S = std::fmax(std::fabs(lo - t_ref), std::fabs(hi - t_ref));
  if (!(S > 0.0) || !std::isfinite(S)) S = 1.0;
  // w[j] as polynomials in tau; initial values 1/(j+1)
  static thread_local L w[MAXD][MAXD];
  for (int j = 0; j < K; ++j) {
    for (int d = 0; d < MAXD; ++d) w[j][d] = 0.0L;
    w[j][0] = 1.0L / (L)(j + 1);
  }
  // FC(j) = (G(j-1) + S tau) / G(j), WC(j) = S tau / G(j)  (G(-1) = 0)
  auto pass = [&](int ks, int jx) {
    for (int j = 0; j < jx; ++j) {
      const L gj = (L)g[j];
      const L f0 = (j > 0 ? (L)g[j - 1] : 0.0L) / gj, f1 = (L)S / gj, w1 = (L)S / gj;
      L nw[MAXD];
      const L* a = w[j + ks - 1];
      const L* b = w[j + ks];
      nw[0] = f0 * a[0];
      for (int d = 1; d < MAXD; ++d) nw[d] = f0 * a[d] + f1 * a[d - 1] - w1 * b[d - 1];
      for (int d = 0; d < MAXD; ++d) w[j + ks][d] = nw[d];
    }
  };
  int ks = K - 1, jx = 0;
  while (ks >= 2) {
    jx += 1;
    pass(ks, jx);
    ks -= 1;
  }
  // position sums use w[1..kq] (ks = 1): P_c = sum_j DT_c[j-1] w[j]
  const double* ref = rec + n + 1;
  const double* dd = rec + n + 7;
  out[0] = t_ref;
  out[1] = 1.0 / S;
  for (int k = 0; k < 6; ++k) out[2 + k] = ref[k];
  for (int c = 0; c < 3; ++c) {
    L P[MAXD];
    for (int d = 0; d < MAXD; ++d) P[d] = 0.0L;
    for (int j = kq[c]; j >= 1; --j)
      for (int d = 0; d < MAXD; ++d) P[d] += (L)dd[c * n + j - 1] * w[j][d];
    for (int d = T21_NP; d < MAXD; ++d)
      if (P[d] != 0.0L) {
        why = "internal: polynomial degree above the order";
        return false;
      }
    for (int d = 0; d < T21_NP; ++d) out[8 + c * T21_NP + d] = (double)P[d];
  }
  return true;
}

int32_t kh_spk_upload_type21(kh_ctx* c, const kh_spk21_seg* segs, int64_t nseg) {
  if (!c || (nseg > 0 && !segs) || nseg < 0) return fail(KH_E_ARG, "null context or segments");
  if (nseg > MAX_T21_SEGS)
    return fail(KH_E_ARG, "too many Type 21 segments (max " + std::to_string(MAX_T21_SEGS) + ")");
  HIPCHK(hipSetDevice(c->device));
  // device order: grouped by object (first appearance), load order within
  std::vector<int64_t> order;
  {
    std::vector<int64_t> ids;
    for (int64_t k = 0; k < nseg; ++k) {
      bool seen = false;
      for (int64_t x : ids) seen = seen || x == segs[k].obj_id;
      if (!seen) ids.push_back(segs[k].obj_id);
    }
    for (int64_t id : ids)
      for (int64_t k = 0; k < nseg; ++k)
        if (segs[k].obj_id == id) order.push_back(k);
  }
  int64_t n_rec_tot = 0, n_tab_tot = 0;
  std::vector<int32_t> nb(nseg);
  for (int64_t k = 0; k < nseg; ++k) {
    const kh_spk21_seg& s = segs[k];
    const std::string tag = "Type 21 segment " + std::to_string(k) + ": ";
    if (!s.records || !s.epochs || s.n_rec <= 0 || s.n_rec > (1 << 24) || s.maxdim < 1 ||
        s.maxdim > 25)
      return fail(KH_E_ARG, tag + "bad record layout");
    if (s.frame != KH_FRAME_EQUATORIAL && s.frame != KH_FRAME_ECLIPTIC)
      return fail(KH_E_ARG, tag + "frame must be 1 or 17");
    if (s.center_id != 10 && s.center_id != 0)
      return fail(KH_E_ARG, tag + "center must be the Sun (10) or the barycentre (0)");
    if (!(s.jd_end >= s.jd_start)) return fail(KH_E_ARG, tag + "bad time range");
    for (int64_t r = 1; r < s.n_rec; ++r)
      if (!(s.epochs[r] >= s.epochs[r - 1])) return fail(KH_E_ARG, tag + "final epochs unsorted");
    nb[k] = (int32_t)std::min<int64_t>(std::max<int64_t>(2 * s.n_rec, 1), 1 << 20);
    n_rec_tot += s.n_rec;
    n_tab_tot += nb[k];
  }
  if (n_rec_tot >= ((int64_t)1 << 31) / T21_REC) return fail(KH_E_ARG, "too many Type 21 records");
  std::vector<double> rec_h((size_t)n_rec_tot * T21_REC), ep_h((size_t)n_rec_tot);
  std::vector<T21Bucket> tab_h((size_t)n_tab_tot);
  SpkDev spk = c->spk;
  std::memset(spk.t21, 0, sizeof(spk.t21));
  std::memset(spk.ast, 0, sizeof(spk.ast));
  std::vector<int64_t> obj(nseg);
  int64_t ro = 0, to = 0;
  for (int64_t q = 0; q < nseg; ++q) {
    const int64_t k = order[q];
    const kh_spk21_seg& s = segs[k];
    const int n = s.maxdim, rl = 4 * s.maxdim + 11;
    T21SegDev& g = spk.t21[q];
    g.jd_start = s.jd_start;
    g.jd_end = s.jd_end;
    g.rec_off = ro;
    g.ep_off = ro;
    g.tab_off = to;
    g.n_rec = (int32_t)s.n_rec;
    g.n_bucket = nb[k];
    g.ecliptic = s.frame == KH_FRAME_ECLIPTIC;
    g.sun_centred = s.center_id == 10;
    const double seg_lo = (s.jd_start - 2451545.0) * 86400.0;  // jd_to_spice_jd
    for (int64_t r = 0; r < s.n_rec; ++r) {
      const double lo = r == 0 ? std::fmin(seg_lo, s.epochs[0]) : s.epochs[r - 1];
      std::string why;
      if (!t21_prepare(s.records + r * rl, n, lo, s.epochs[r], &rec_h[(ro + r) * T21_REC], why))
        return fail(KH_E_ARG, "Type 21 segment " + std::to_string(k) + " record " +
                                  std::to_string(r) + ": " + why);
      ep_h[ro + r] = s.epochs[r];
    }
    // buckets over [t0, last epoch]: entry b holds the first record whose
    // final epoch is >= t0 + b*width - width*1e-6 (a lower bound for any time
    // that rounds into bucket b) and the final epochs of it and the next
    const double t0 = std::fmin(seg_lo, s.epochs[0]);
    const double t1 = s.epochs[s.n_rec - 1];
    double bw = (t1 - t0) / nb[k];
    if (!(bw > 0.0)) {
      bw = 1.0;
      g.n_bucket = nb[k] = 1;
    }
    g.t0 = t0;
    g.inv_bw = 1.0 / bw;
    int64_t i = 0;
    for (int32_t b = 0; b < g.n_bucket; ++b) {
      const double x = t0 + b * bw - bw * 1e-6;
      while (i < s.n_rec - 1 && s.epochs[i] < x) ++i;
      T21Bucket& e = tab_h[to + b];
      e.rec = (int32_t)i;
      e.ep0 = s.epochs[i];
      e.ep1 = i + 1 < s.n_rec ? s.epochs[i + 1] : INFINITY;
      e.pad0 = 0;
      e.pad1 = 0;
    }
    obj[q] = s.obj_id;
    ro += s.n_rec;
    to += g.n_bucket;
  }
  spk.n_t21_seg = (int32_t)nseg;
  spk.n_t21_rec = (int32_t)n_rec_tot;
  for (int a = 0; a < N_ASTEROIDS; ++a) {
    AstDev& ad = spk.ast[a];
    ad.seg_begin = ad.seg_end = 0;
    for (int64_t q = 0; q < nseg; ++q)
      if (obj[q] == kAsteroids[a].id) {
        if (ad.seg_end == 0) ad.seg_begin = (int32_t)q;
        ad.seg_end = (int32_t)q + 1;
      }
    ad.bd.seg = -1;
    ad.bd.emb = 0;
    ad.bd.id = (int32_t)kAsteroids[a].id;
    ad.bd.kind = 0;
    ad.bd.mass = kAsteroids[a].mass;
    ad.bd.radius = kAsteroids[a].radius;
    ad.bd.j2 = 0.0;
  }
  if (c->d_t21) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(c->d_t21));
    c->d_t21 = nullptr;
  }
  spk.t21_rec = nullptr;
  spk.t21_ep = nullptr;
  spk.t21_tab = nullptr;
  if (nseg > 0) {
    const size_t brec = align256(rec_h.size() * 8), bep = align256(ep_h.size() * 8);
    const size_t btab = tab_h.size() * sizeof(T21Bucket);
    HIPCHK(hipMalloc(&c->d_t21, brec + bep + btab));
    char* base = (char*)c->d_t21;
    HIPCHK(hipMemcpy(base, rec_h.data(), rec_h.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(base + brec, ep_h.data(), ep_h.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(base + brec + bep, tab_h.data(), btab, hipMemcpyHostToDevice));
    spk.t21_rec = (const double*)base;
    spk.t21_ep = (const double*)(base + brec);
    spk.t21_tab = (const T21Bucket*)(base + brec + bep);
  }
  if (!c->d_spk) HIPCHK(hipMalloc(&c->d_spk, sizeof(SpkDev)));
  HIPCHK(hipMemcpy(c->d_spk, &spk, sizeof(SpkDev), hipMemcpyHostToDevice));
  c->spk = spk;
  c->t21_obj = obj;
  return KH_OK;
}

// Grow-only device buffer.
int32_t kh_grow(void** buf, size_t* size, size_t bytes) {
  if (bytes > *size) {
    if (*buf) HIPCHK(hipFree(*buf));
    *buf = nullptr;
    *size = 0;
    HIPCHK(hipMalloc(buf, bytes));
    *size = bytes;
  }
  return KH_OK;
}

// Staging buffer for the host-pointer entry points.
static int32_t staging(kh_ctx* c, size_t bytes, char** out) {
  int32_t r = kh_grow(&c->dbuf, &c->dbuf_size, bytes);
  *out = (char*)c->dbuf;
  return r;
}

int32_t kh_spk_states(kh_ctx* c, int64_t obj_id, int64_t center, int32_t frame, int64_t m,
                      const double* jd, double* const out[6], int32_t* status) {
  if (!c || !jd || !out || !status || m < 0) return fail(KH_E_ARG, "null argument");
  if (!c->has_spk) return fail(KH_E_NO_SPK, "no ephemeris uploaded");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  Chain body{}, cen{}, sun{};
  // an object with Type 21 segments is read from them (kete keeps an
  // object's segments in load order and takes the first covering the time,
  // spk.rs:72-78); the planet tree otherwise
  int32_t t21_begin = 0, t21_end = 0;
  for (size_t q = 0; q < c->t21_obj.size(); ++q)
    if (c->t21_obj[q] == obj_id) {
      if (t21_end == 0) t21_begin = (int32_t)q;
      t21_end = (int32_t)q + 1;
    }
  if (t21_end > t21_begin) {
    if (!kh_build_chain(c, 10, sun)) return fail(KH_E_NO_SPK, "no path from the Sun to 0");
  } else if (!kh_build_chain(c, obj_id, body)) {
    return fail(KH_E_NO_SPK, "no path from object to 0");
  }
  if (!kh_build_chain(c, center, cen)) return fail(KH_E_NO_SPK, "no path from center to 0");
  if (m == 0) return KH_OK;
  HIPCHK(hipSetDevice(c->device));
  size_t bj = align256(m * 8), bo = align256(m * 8), bs = align256(m * 4);
  char* base;
  int32_t r = staging(c, bj + 6 * bo + bs, &base);
  if (r) return r;
  double* d_jd = (double*)base;
  double* d_out[6];
  for (int k = 0; k < 6; ++k) d_out[k] = (double*)(base + bj + k * bo);
  int32_t* d_st = (int32_t*)(base + bj + 6 * bo);
  HIPCHK(hipMemcpyAsync(d_jd, jd, m * 8, hipMemcpyHostToDevice, c->stream));
  if (launch_spk_states(c->d_spk, c->d_coef, body, cen, t21_begin, t21_end, sun, center, frame, m,
                        d_jd, d_out, d_st, c->stream))
    return fail(KH_E_HIP, "spk_states launch failed");
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync(out[k], d_out[k], m * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(status, d_st, m * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

static int32_t check_two_body_args(kh_ctx* c, int64_t n, const double* jd0,
                                   const double* const xyz[6], int64_t center, int32_t frame,
                                   int64_t m, const double* t, double* const out[6],
                                   int32_t* status) {
  if (!c) return fail(KH_E_ARG, "null context");
  if (n < 0 || m < 0) return fail(KH_E_ARG, "negative size");
  if ((n > 0 && (!jd0 || !xyz)) || (m > 0 && !t) || (n * m > 0 && (!out || !status)))
    return fail(KH_E_ARG, "null array");
  for (int k = 0; n > 0 && k < 6; ++k)
    if (!xyz[k] || (m > 0 && !out[k])) return fail(KH_E_ARG, "null state column");
  if (center != 10 && center != 0) return fail(KH_E_ARG, "center must be 10 or 0");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  if (center == 0 && !c->has_spk) return fail(KH_E_NO_SPK, "center 0 needs an ephemeris");
  return KH_OK;
}

int32_t kh_propagate_two_body_dev(kh_ctx* c, int64_t n, const double* jd0,
                                  const double* const xyz[6], int64_t center, int32_t frame,
                                  int64_t m, const double* jd_targets,
                                  const double* observer_pos, double* const out[6],
                                  int32_t* status, int32_t* newton_iters, int32_t sync) {
  int32_t r = check_two_body_args(c, n, jd0, xyz, center, frame, m, jd_targets, out, status);
  if (r) return r;
  if (n == 0 || m == 0) return KH_OK;
  HIPCHK(hipSetDevice(c->device));
  TwoBodyLaunch p{};
  p.n = n;
  p.m = m;
  p.jd0 = jd0;
  for (int k = 0; k < 6; ++k) {
    p.xyz[k] = xyz[k];
    p.out[k] = out[k];
  }
  p.targets = jd_targets;
  p.has_observer = observer_pos != nullptr;
  if (observer_pos)
    for (int k = 0; k < 3; ++k) p.observer[k] = observer_pos[k];
  p.center_sun = center == 10;
  p.sun_seg = c->has_spk ? c->spk.sun_seg : 0;
  p.frame = frame;
  p.status = status;
  p.iters = newton_iters;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (launch_two_body(c->d_spk, c->d_coef, p, c->stream)) return fail(KH_E_HIP, "two-body launch failed");
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  if (sync) HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

int32_t kh_propagate_two_body(kh_ctx* c, int64_t n, const double* jd0,
                              const double* const xyz[6], int64_t center, int32_t frame,
                              int64_t m, const double* jd_targets, const double* observer_pos,
                              double* const out[6], int32_t* status) {
  int32_t r = check_two_body_args(c, n, jd0, xyz, center, frame, m, jd_targets, out, status);
  if (r) return r;
  if (n == 0 || m == 0) return KH_OK;
  HIPCHK(hipSetDevice(c->device));
  size_t bn = align256(n * 8), bm = align256(m * 8), bo = align256(n * m * 8),
         bs = align256(n * m * 4);
  char* base;
  r = staging(c, 7 * bn + bm + 6 * bo + bs, &base);
  if (r) return r;
  double* d_jd0 = (double*)base;
  const double* d_xyz[6];
  double* d_out[6];
  for (int k = 0; k < 6; ++k) {
    d_xyz[k] = (double*)(base + (1 + k) * bn);
    d_out[k] = (double*)(base + 7 * bn + bm + k * bo);
  }
  double* d_t = (double*)(base + 7 * bn);
  int32_t* d_st = (int32_t*)(base + 7 * bn + bm + 6 * bo);
  HIPCHK(hipMemcpyAsync(d_jd0, jd0, n * 8, hipMemcpyHostToDevice, c->stream));
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync((void*)d_xyz[k], xyz[k], n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_t, jd_targets, m * 8, hipMemcpyHostToDevice, c->stream));
  r = kh_propagate_two_body_dev(c, n, d_jd0, d_xyz, center, frame, m, d_t, observer_pos, d_out,
                                d_st, nullptr, 0);
  if (r) return r;
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync(out[k], d_out[k], n * m * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(status, d_st, n * m * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

static int32_t check_nbody_args(kh_ctx* c, int64_t n, const double* jd0,
                                const double* const xyz[6], int64_t center, int32_t frame,
                                const double* jd_final, double epsilon, int32_t mode,
                                double* const out[6], int32_t* status, Chain& ch) {
  if (!c) return fail(KH_E_ARG, "null context");
  if (!c->has_spk) return fail(KH_E_NO_SPK, "no ephemeris uploaded");
  if (n < 0) return fail(KH_E_ARG, "negative size");
  if (n > 0 && (!jd0 || !xyz || !jd_final || !out || !status)) return fail(KH_E_ARG, "null array");
  for (int k = 0; n > 0 && k < 6; ++k)
    if (!xyz[k] || !out[k]) return fail(KH_E_ARG, "null state column");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  if (!(epsilon > 0.0)) return fail(KH_E_ARG, "epsilon must be positive");
  if (mode < -1 || mode > 3) return fail(KH_E_ARG, "mode must be -1 .. 3");
  if (!kh_build_chain(c, center, ch))
    return fail(KH_E_NO_SPK, "center " + std::to_string(center) + " is not in the ephemeris");
  return KH_OK;
}

// The N-body launch behind kh_propagate_n_body_dev and the multi-epoch
// propagation (me != null).
static int32_t nbody_dev(kh_ctx* c, int64_t n, const double* jd0, const double* const xyz[6],
                         int64_t center, int32_t frame, const double* jd_final,
                         int32_t jd_final_stride, int32_t include_asteroids,
                         const kh_nongrav* ng, double epsilon,
                         int32_t mode, double* const out[6], int32_t* status, int64_t* impact_id,
                         double* impact_jd, kh_counters* counters, int32_t sync,
                         const kh::MultiEpochs* me, void* order_scratch = nullptr) {
  Chain ch;
  int32_t r = check_nbody_args(c, n, jd0, xyz, center, frame, jd_final, epsilon, mode, out, status,
                               ch);
  if (r) return r;
  if (n == 0) return KH_OK;
  HIPCHK(hipSetDevice(c->device));
  NBodyLaunch p{};
  p.n = n;
  p.jd0 = jd0;
  for (int k = 0; k < 6; ++k) {
    p.xyz[k] = xyz[k];
    p.out[k] = out[k];
  }
  p.jd_final = jd_final;
  p.jd_final_stride = jd_final_stride ? 1 : 0;
  p.frame = frame;
  p.center = ch;
  p.ng = ng;
  p.epsilon = epsilon;
  p.status = status;
  p.impact_id = impact_id;
  p.impact_jd = impact_jd;
  p.counters = reinterpret_cast<int32_t*>(counters);
  p.queue = c->d_queue;
  p.mode = mode < 0 ? 0 : mode;
  p.n_ast = include_asteroids ? N_ASTEROIDS : 0;
  p.block = 256;
  const bool tuned = mode < 0 && c->tune.valid && !me;
  p.grid = 0;
  // LDS window of 16 records (62 KB for 11 segments) when all segments have
  // 14 coefficients and the window fits next to a second block on the CU
  p.n_seg = c->spk.n_seg;
  const size_t win_bytes = (size_t)c->spk.n_seg * 16 * 44 * sizeof(double);
  p.range = (c->spk.all14 && win_bytes <= 64 * 1024) ? c->d_range : nullptr;
  p.window = p.range ? 16 : 0;
  if (const char* w = getenv("KH_LDS_WINDOW")) {
    p.window = atoi(w);
    if (p.window <= 0 || (size_t)c->spk.n_seg * p.window * 44 * 8 > 64 * 1024) p.range = nullptr;
  }
  p.n_cu = c->n_cu;
  if (const char* g = getenv("KH_GRID")) p.grid = atoi(g);
  // tail hand-off of the plain launch (KH_PARK=0 turns it off): a wavefront
  // down to KH_PARK_LANES live lanes (default 4, at most 4) parks the objects
  // with more than KH_PARK_LEFT force evaluations to go (by their rate so
  // far; default 0: every object still running) and finishes them sixteen
  // lanes each.  It catches the slow objects no cost key predicts (close
  // approaches): 1e6 objects over a year, epsilon 1e-9: 471 -> 212 ms
  // (149-150 ms with the rows split by coordinate); epsilon 1e-6 with the
  // one-lane latency kernel: 71.0 -> 69.9 ms, without it 67.5 ms.  With the
  // split rows the threshold measured 62.1-62.6 / 61.1-61.3 / 60.6-60.2 /
  // 59.8-60.0 ms at 200 / 50 / 10 / 0 evaluations (epsilon 1e-6), 148-150
  // ms at 1e-9 for all of 0-200, 400: 192 ms.
  {
    const char* pk = getenv("KH_PARK");
    if (!(pk && atoi(pk) == 0)) {
      p.park_lanes = 4;
      p.park_min_left = 0.0;
      if (const char* e = getenv("KH_PARK_LANES")) p.park_lanes = atoi(e);
      if (const char* e = getenv("KH_PARK_LEFT")) p.park_min_left = atof(e);
      p.park_lanes = p.park_lanes < 0 ? 0 : (p.park_lanes > 4 ? 4 : p.park_lanes);
    }
  }
  // plain launch on the shared record grid: the slowest objects by the cost
  // key run in the latency kernel beside the main launch, those whose cost
  // outweighs a lane's share of the batch, at most a cap (KH_LATENCY_K sets
  // it).  The latency kernel: sixteen lanes per object for batches of uneven
  // arcs (per-object final epochs; planets only), whose heavy tail it
  // shortens (2e6 objects over 1-3650 d with the tail hand-off: 316 ms with
  // up to 32 x n_cu objects in it, 317-320 ms with 64-256 x n_cu (the key
  // threshold then decides), 340 ms with 16, 379 ms with 8, 640 ms with
  // none; more objects by a lower threshold, 600 ms); the one-lane variant,
  // up to 8 x n_cu objects, for per-object final epochs with
  // include_asteroids.  For a common final epoch the main launch alone is
  // faster: planets only, its tail hand-off takes the slow objects (67.5 vs
  // 69.9 ms with the one-lane variant beside it); with include_asteroids the
  // cost key's head is the eccentric NEOs, which the one-lane variant holds
  // longer than the main launch takes (191 vs 168 ms).
  // KH_LATENCY=wide or =ilp chooses.
  {
    const char* lv = getenv("KH_LATENCY");
    bool wide = jd_final_stride != 0;
    if (lv && std::string(lv) == "wide") wide = true;
    if (lv && std::string(lv) == "ilp") wide = false;
    p.wide_latency = wide && !include_asteroids;
    p.latency_k = !c->spk.shared_grid || jd_final_stride == 0 ? 0
                  : p.wide_latency                             ? (int64_t)c->n_cu * 32
                                                               : (int64_t)c->n_cu * 8;
  }
  if (const char* k = getenv("KH_LATENCY_K")) p.latency_k = c->spk.shared_grid ? atoll(k) : 0;
  if (tuned) {
    p.mode = c->tune.mode;
    p.block = c->tune.block;
    p.latency_k = c->spk.shared_grid ? c->tune.latency_k : 0;
    p.window = c->tune.window;
    p.range = (c->spk.all14 && p.window > 0 &&
               (size_t)c->spk.n_seg * p.window * 44 * 8 <= 64 * 1024)
                  ? c->d_range
                  : nullptr;
    if (!p.range) p.window = 0;
  }
  p.split = c->d_split;
  p.stream2 = c->stream2;
  p.fork = c->fork;
  p.join = c->join;
  if (me) {
    p.mode = 0;
    p.epochs = me->epochs;
    p.n_epochs = me->n_epochs;
    p.ep_lo = me->lo;
    p.ep_hi = me->hi;
    for (int k = 0; k < 6; ++k) p.mout[k] = me->out[k];
  }
  if (p.mode != 3) {
    // longest-first order of the objects (one lane per object); KH_ORDER=0
    // turns it off
    const char* ord = getenv("KH_ORDER");
    if (!(ord && atoi(ord) == 0)) {
      // the allocating entry points also stage each object as one 64-byte
      // record after the order, and per-object final epochs the radix sort's
      // keys and scratch after those (the no-alloc entry point's scratch
      // holds the order only: counting sort, no records)
      const size_t ob = align256(order_buf_bytes(n));
      const size_t sb = align256(stage_buf_bytes(n));
      const size_t rb = jd_final_stride != 0 && !me ? order_sort_temp_bytes(n) : 0;
      if (order_scratch) {
        p.order_buf = order_scratch;
      } else if (kh_grow(&c->obuf, &c->obuf_size, ob + sb + rb) == KH_OK) {
        p.order_buf = c->obuf;
        p.stage = reinterpret_cast<double*>(static_cast<char*>(c->obuf) + ob);
        if (rb) {
          p.sort_buf = static_cast<char*>(c->obuf) + ob + sb;
          p.sort_bytes = rb;
        }
      } else {
        return KH_E_HIP;
      }
    }
  }
  if (launch_nbody(c->d_spk, c->d_coef, p, c->stream, &c->last_block, &c->last_grid,
                   &c->last_mode, c->ev0, c->ev1))
    return fail(KH_E_HIP, "n-body launch failed");
  c->timed = true;
  if (sync) HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

int32_t kh_propagate_n_body_dev(kh_ctx* c, int64_t n, const double* jd0,
                                const double* const xyz[6], int64_t center, int32_t frame,
                                const double* jd_final, int32_t jd_final_stride,
                                int32_t include_asteroids, const kh_nongrav* ng, double epsilon,
                                int32_t mode, double* const out[6], int32_t* status,
                                int64_t* impact_id, double* impact_jd, kh_counters* counters,
                                int32_t sync) {
  return nbody_dev(c, n, jd0, xyz, center, frame, jd_final, jd_final_stride, include_asteroids,
                   ng, epsilon, mode, out, status, impact_id, impact_jd, counters, sync, nullptr);
}

int32_t kh_propagate_n_body_dev_noalloc(kh_ctx* c, int64_t n, const double* jd0,
                                        const double* const xyz[6], int64_t center,
                                        int32_t frame, const double* jd_final,
                                        int32_t jd_final_stride, int32_t include_asteroids,
                                        const kh_nongrav* ng, double epsilon, int32_t mode,
                                        double* const out[6], int32_t* status,
                                        int64_t* impact_id, double* impact_jd,
                                        kh_counters* counters, int32_t sync) {
  if (n > KH_NOALLOC_MAX_N)
    return fail(KH_E_ARG, "kh_propagate_n_body_dev_noalloc takes at most " +
                              std::to_string(KH_NOALLOC_MAX_N) + " objects, got " +
                              std::to_string(n));
  if (!c) return fail(KH_E_ARG, "null context");
  if (!c->noalloc_scratch)
    return fail(KH_E_ARG, "no scratch on device " + std::to_string(c->device));
  return nbody_dev(c, n, jd0, xyz, center, frame, jd_final, jd_final_stride, include_asteroids,
                   ng, epsilon, mode, out, status, impact_id, impact_jd, counters, sync, nullptr,
                   c->noalloc_scratch);
}

int32_t kh_propagate_n_body(kh_ctx* c, int64_t n, const double* jd0,
                            const double* const xyz[6], int64_t center, int32_t frame,
                            const double* jd_final, int32_t jd_final_stride,
                            int32_t include_asteroids, const kh_nongrav* ng, double epsilon,
                            int32_t mode,
                            double* const out[6], int32_t* status, int64_t* impact_id,
                            double* impact_jd, kh_counters* counters) {
  Chain ch;
  int32_t r = check_nbody_args(c, n, jd0, xyz, center, frame, jd_final, epsilon, mode, out, status,
                               ch);
  if (r) return r;
  if (n == 0) return KH_OK;
  if (ng)
    for (int64_t i = 0; i < n; ++i)
      if (ng[i].kind == 1 && ng[i].dt != 0.0)
        return fail(KH_E_ARG, "JplComet non-grav with a time delay (dt != 0) is not supported");
  HIPCHK(hipSetDevice(c->device));
  const int64_t nf = jd_final_stride ? n : 1;
  size_t bn = align256(n * 8), bf = align256(nf * 8), bng = align256(n * sizeof(kh_nongrav)),
         bs = align256(n * 4), bc = align256(n * sizeof(kh_counters));
  char* base;
  r = staging(c, 7 * bn + bf + (ng ? bng : 0) + 6 * bn + bs + 2 * bn + bc, &base);
  if (r) return r;
  char* q = base;
  double* d_jd0 = (double*)q;
  q += bn;
  const double* d_xyz[6];
  for (int k = 0; k < 6; ++k, q += bn) d_xyz[k] = (double*)q;
  double* d_jf = (double*)q;
  q += bf;
  kh_nongrav* d_ng = nullptr;
  if (ng) {
    d_ng = (kh_nongrav*)q;
    q += bng;
  }
  double* d_out[6];
  for (int k = 0; k < 6; ++k, q += bn) d_out[k] = (double*)q;
  int32_t* d_st = (int32_t*)q;
  q += bs;
  int64_t* d_iid = (int64_t*)q;
  q += bn;
  double* d_ijd = (double*)q;
  q += bn;
  kh_counters* d_cnt = (kh_counters*)q;
  HIPCHK(hipMemcpyAsync(d_jd0, jd0, n * 8, hipMemcpyHostToDevice, c->stream));
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync((void*)d_xyz[k], xyz[k], n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_jf, jd_final, nf * 8, hipMemcpyHostToDevice, c->stream));
  if (ng) HIPCHK(hipMemcpyAsync(d_ng, ng, n * sizeof(kh_nongrav), hipMemcpyHostToDevice, c->stream));
  r = kh_propagate_n_body_dev(c, n, d_jd0, d_xyz, center, frame, d_jf, jd_final_stride,
                              include_asteroids, d_ng, epsilon, mode, d_out, d_st, d_iid, d_ijd,
                              counters ? d_cnt : nullptr, 0);
  if (r) return r;
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync(out[k], d_out[k], n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(status, d_st, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (impact_id) HIPCHK(hipMemcpyAsync(impact_id, d_iid, n * 8, hipMemcpyDeviceToHost, c->stream));
  if (impact_jd) HIPCHK(hipMemcpyAsync(impact_jd, d_ijd, n * 8, hipMemcpyDeviceToHost, c->stream));
  if (counters)
    HIPCHK(hipMemcpyAsync(counters, d_cnt, n * sizeof(kh_counters), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

// ---------------------------------------------------------------------------
// propagate_n_body_long
// ---------------------------------------------------------------------------
// SIMPLE_PLANETS at jd0 (Sun-centred equatorial) into the device array
// d_pl[9][6]: given by the caller (host), or read from the ephemeris
// (try_get_state(id, jd_init, 10, Equatorial), propagation/mod.rs:163-175)
// by the state kernel writing straight into d_pl.  *status: DAFLimits when a
// planet is not covered.
static int32_t long_planets(kh_ctx* c, double jd0, const double* planet_xyz, double* d_pl,
                            int32_t* status) {
  *status = ST_OK;
  if (planet_xyz) {
    // the caller's array may be gone before queued work reads it: wait
    HIPCHK(hipMemcpyAsync(d_pl, planet_xyz, 54 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return KH_OK;
  }
  if (!c->has_spk) return fail(KH_E_NO_SPK, "no ephemeris uploaded");
  static const int64_t ids[9] = {10, 1, 2, 3, 4, 5, 6, 7, 8};
  Chain cen{};
  if (!kh_build_chain(c, 10, cen)) return fail(KH_E_NO_SPK, "no path from the Sun to 0");
  double* d_jd = d_pl + 54;
  int32_t* d_pst = (int32_t*)(d_pl + 55);
  HIPCHK(hipMemcpyAsync(d_jd, &jd0, 8, hipMemcpyHostToDevice, c->stream));
  for (int b = 0; b < 9; ++b) {
    Chain body{};
    if (!kh_build_chain(c, ids[b], body)) return fail(KH_E_NO_SPK, "no path from a planet to 0");
    double* o[6];
    for (int k = 0; k < 6; ++k) o[k] = d_pl + 6 * b + k;
    Chain none{};
    if (launch_spk_states(c->d_spk, c->d_coef, body, cen, 0, 0, none, 10, 1, 1, d_jd, o,
                          d_pst + b, c->stream))
      return fail(KH_E_HIP, "spk_states launch failed");
  }
  int32_t pst[9];
  HIPCHK(hipMemcpyAsync(pst, d_pst, sizeof(pst), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < 9; ++b)
    if (pst[b] != ST_OK) {
      *status = pst[b];
      break;
    }
  return KH_OK;
}

int32_t kh_propagate_n_body_long_dev(kh_ctx* c, int64_t n, double jd0,
                                     const double* const xyz[6], int64_t center, int32_t frame,
                                     double jd_final, const double* planet_xyz,
                                     const kh_nongrav* ng, int32_t batch_size,
                                     double* const out[6], double* batch_planets,
                                     int32_t* batch_status, kh_counters* batch_counters,
                                     int32_t sync) {
  if (!c || !batch_status) return fail(KH_E_ARG, "null argument");
  if (batch_size < 1 || batch_size > 55)
    return fail(KH_E_ARG, "batch_size must be 1 .. 55 (a batch and its nine planets share one "
                          "wavefront)");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  if (n <= 0) return fail(KH_E_ARG, "empty batch");
  if (center != 10) return fail(KH_E_ARG, "center must be 10 (the Sun)");
  if (!xyz || !out) return fail(KH_E_ARG, "null array");
  for (int k = 0; k < 6; ++k)
    if (!xyz[k] || !out[k]) return fail(KH_E_ARG, "null state column");
  HIPCHK(hipSetDevice(c->device));
  // planets [54], their epoch, their lookup statuses [9]: 476 bytes
  int32_t r = kh_grow(&c->lbuf, &c->lbuf_size, 512);
  if (r) return r;
  double* d_pl = (double*)c->lbuf;
  int32_t pst = ST_OK;
  r = long_planets(c, jd0, planet_xyz, d_pl, &pst);
  if (r) return r;
  const int64_t nb = (n + batch_size - 1) / batch_size;
  if (pst != ST_OK) {
    // every batch fails as kete's try_get_state does for the call
    std::vector<int32_t> st(nb, pst);
    HIPCHK(hipMemcpyAsync(batch_status, st.data(), nb * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return KH_OK;
  }
  LongLaunch p{};
  p.n = n;
  p.batch = batch_size;
  p.frame = frame;
  p.jd0 = jd0;
  p.jd_final = jd_final;
  for (int k = 0; k < 6; ++k) {
    p.xyz[k] = xyz[k];
    p.out[k] = out[k];
  }
  p.planets = d_pl;
  p.ng = ng;
  p.planets_out = batch_planets;
  p.status = batch_status;
  p.counters = (int32_t*)batch_counters;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (launch_nbody_long(p, c->stream)) return fail(KH_E_HIP, "n-body long launch failed");
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  if (sync) HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

int32_t kh_propagate_n_body_long(kh_ctx* c, int64_t n, double jd0, const double* const xyz[6],
                                 int64_t center, int32_t frame, double jd_final,
                                 const double* planet_xyz, const kh_nongrav* ng,
                                 int32_t batch_size, double* const out[6], double* planets_out,
                                 int32_t* status) {
  if (!c || !status) return fail(KH_E_ARG, "null argument");
  *status = ST_OK;
  if (batch_size < 1 || batch_size > 55)
    return fail(KH_E_ARG, "batch_size must be 1 .. 55 (a batch and its nine planets share one "
                          "wavefront)");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  // propagate_n_body_vec's own checks (propagation/mod.rs:148-160,190-208)
  if (n <= 0) {
    *status = ST_VALUE_ERROR;  // "State vector is empty"
    return KH_OK;
  }
  if (!xyz || !out) return fail(KH_E_ARG, "null array");
  for (int k = 0; k < 6; ++k)
    if (!xyz[k] || !out[k]) return fail(KH_E_ARG, "null state column");
  if (center != 10) {
    *status = ST_VALUE_ERROR;  // "Center of all states must be 10 (the Sun)."
    return KH_OK;
  }
  if (ng)
    for (int64_t i = 0; i < n; ++i)
      if (ng[i].kind == 1 && ng[i].dt != 0.0)
        return fail(KH_E_ARG, "JplComet non-grav with a time delay (dt != 0) is not supported");
  HIPCHK(hipSetDevice(c->device));
  const int64_t nb = (n + batch_size - 1) / batch_size;
  const size_t bn = align256(n * 8), bpo = align256(nb * 54 * 8), bs = align256(nb * 4),
               bng = align256(n * sizeof(kh_nongrav));
  char* base;
  int32_t r = staging(c, 12 * bn + bpo + bs + (ng ? bng : 0), &base);
  if (r) return r;
  char* q = base;
  const double* d_x[6];
  double* d_o[6];
  for (int k = 0; k < 6; ++k, q += bn) d_x[k] = (double*)q;
  for (int k = 0; k < 6; ++k, q += bn) d_o[k] = (double*)q;
  double* d_plo = (double*)q;
  q += bpo;
  int32_t* d_st = (int32_t*)q;
  q += bs;
  kh_nongrav* d_ng = ng ? (kh_nongrav*)q : nullptr;
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync((void*)d_x[k], xyz[k], n * 8, hipMemcpyHostToDevice, c->stream));
  if (ng) HIPCHK(hipMemcpyAsync(d_ng, ng, n * sizeof(kh_nongrav), hipMemcpyHostToDevice, c->stream));
  r = kh_propagate_n_body_long_dev(c, n, jd0, d_x, center, frame, jd_final, planet_xyz, d_ng,
                                   batch_size, d_o, d_plo, d_st, nullptr, 0);
  if (r) return r;
  std::vector<int32_t> st(nb);
  HIPCHK(hipMemcpyAsync(st.data(), d_st, nb * 4, hipMemcpyDeviceToHost, c->stream));
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync(out[k], d_o[k], n * 8, hipMemcpyDeviceToHost, c->stream));
  if (planets_out)
    HIPCHK(hipMemcpyAsync(planets_out, d_plo + (nb - 1) * 54, 54 * 8, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // kete collects the batches into one Result: the first failure fails the call
  for (int64_t b = 0; b < nb; ++b)
    if (st[b] != ST_OK) {
      *status = st[b];
      break;
    }
  return KH_OK;
}

// ---------------------------------------------------------------------------
// MOID (propagation/kepler.rs:325-404; moid_py src/kete/rust/propagation.rs:27-39)
// ---------------------------------------------------------------------------
int32_t kh_moid_dev(kh_ctx* c, int64_t n, const double* jd_a, const double* const xyz_a[6],
                    int32_t frame, const double* jd_b, const double* const xyz_b[6],
                    double* moid, int32_t* status, double* grid_min, int32_t* iters,
                    int32_t* work, int32_t sync) {
  if (!c) return fail(KH_E_ARG, "null context");
  if (n < 0) return fail(KH_E_ARG, "negative size");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  if (n == 0) return KH_OK;
  if (!jd_a || !xyz_a || !moid || !status) return fail(KH_E_ARG, "null array");
  for (int k = 0; k < 6; ++k)
    if (!xyz_a[k] || (xyz_b && !xyz_b[k])) return fail(KH_E_ARG, "null state column");
  Chain earth{}, cen{};
  if (!xyz_b) {
    // the Earth at each state_a epoch, Sun-centred in the states' frame
    if (!c->has_spk) return fail(KH_E_NO_SPK, "no ephemeris uploaded");
    if (!kh_build_chain(c, 399, earth) || !kh_build_chain(c, 10, cen))
      return fail(KH_E_NO_SPK, "no path from the Earth or the Sun to 0");
  }
  HIPCHK(hipSetDevice(c->device));
  const size_t bd = align256(n * 8), bi = align256(n * 4);
  const size_t bsg = align256(sizeof(MoidSharedGrid));
  int32_t r = kh_grow(&c->mbuf, &c->mbuf_size, 20 * bd + bi + bsg);
  if (r) return r;
  char* q = (char*)c->mbuf;
  MoidLaunch p{};
  p.n = n;
  p.frame = frame;
  p.max_iter = 1000;
  p.jd_a = jd_a;
  p.jd_b = jd_b;
  for (int k = 0; k < 7; ++k, q += bd) p.best_a[k] = (double*)q;
  for (int k = 0; k < 7; ++k, q += bd) p.best_b[k] = (double*)q;
  for (int k = 0; k < 6; ++k) p.xa[k] = xyz_a[k];
  if (xyz_b) {
    for (int k = 0; k < 6; ++k) p.xb[k] = xyz_b[k];
  } else {
    double* eo[6];
    for (int k = 0; k < 6; ++k, q += bd) {
      eo[k] = (double*)q;
      p.xb[k] = eo[k];
    }
    int32_t* est = (int32_t*)q;
    q += bi;
    p.st_b = est;
    p.sg = (MoidSharedGrid*)q;
    Chain none{};
    if (launch_spk_states(c->d_spk, c->d_coef, earth, cen, 0, 0, none, 10, frame, n, jd_a, eo, est,
                          c->stream))
      return fail(KH_E_HIP, "spk_states launch failed");
  }
  p.moid = moid;
  p.grid_min = grid_min;
  p.status = status;
  p.iters = iters;
  p.work = work;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (launch_moid(p, c->stream)) return fail(KH_E_HIP, "moid launch failed");
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  if (sync) HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

int32_t kh_moid(kh_ctx* c, int64_t n, const double* jd_a, const double* const xyz_a[6],
                int32_t frame, const double* jd_b, const double* const xyz_b[6], double* moid,
                int32_t* status) {
  if (!c) return fail(KH_E_ARG, "null context");
  if (n <= 0) return n == 0 ? KH_OK : fail(KH_E_ARG, "negative size");
  if (!jd_a || !xyz_a || !moid || !status) return fail(KH_E_ARG, "null array");
  for (int k = 0; k < 6; ++k)
    if (!xyz_a[k] || (xyz_b && !xyz_b[k])) return fail(KH_E_ARG, "null state column");
  HIPCHK(hipSetDevice(c->device));
  const size_t bd = align256(n * 8), bi = align256(n * 4);
  char* base;
  int32_t r = staging(c, 15 * bd + bi, &base);
  if (r) return r;
  char* q = base;
  double* d_jda = (double*)q;
  q += bd;
  double* d_jdb = jd_b ? (double*)q : nullptr;
  q += bd;
  const double* d_xa[6];
  const double* d_xb[6];
  for (int k = 0; k < 6; ++k, q += bd) d_xa[k] = (const double*)q;
  for (int k = 0; k < 6; ++k, q += bd) d_xb[k] = (const double*)q;
  double* d_moid = (double*)q;
  q += bd;
  int32_t* d_st = (int32_t*)q;
  HIPCHK(hipMemcpyAsync(d_jda, jd_a, n * 8, hipMemcpyHostToDevice, c->stream));
  if (jd_b) HIPCHK(hipMemcpyAsync(d_jdb, jd_b, n * 8, hipMemcpyHostToDevice, c->stream));
  for (int k = 0; k < 6; ++k) {
    HIPCHK(hipMemcpyAsync((void*)d_xa[k], xyz_a[k], n * 8, hipMemcpyHostToDevice, c->stream));
    if (xyz_b)
      HIPCHK(hipMemcpyAsync((void*)d_xb[k], xyz_b[k], n * 8, hipMemcpyHostToDevice, c->stream));
  }
  r = kh_moid_dev(c, n, d_jda, d_xa, frame, d_jdb, xyz_b ? d_xb : nullptr, d_moid, d_st, nullptr,
                  nullptr, nullptr, 0);
  if (r) return r;
  HIPCHK(hipMemcpyAsync(moid, d_moid, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(status, d_st, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: launch shape of the N-body path measured on the caller's batch
// ---------------------------------------------------------------------------
int32_t kh_autotune_n_body(kh_ctx* c, int64_t n, const double* jd0, const double* const xyz[6],
                           int64_t center, int32_t frame, const double* jd_final,
                           int32_t jd_final_stride, int32_t include_asteroids, int64_t sample,
                           int32_t apply, kh_tune_result* out) {
  if (!c || !out) return fail(KH_E_ARG, "null argument");
  std::memset(out, 0, sizeof(*out));
  if (!c->has_spk) return fail(KH_E_NO_SPK, "no ephemeris uploaded");
  if (n <= 0 || !jd0 || !xyz || !jd_final) return fail(KH_E_ARG, "empty batch or null array");
  for (int k = 0; k < 6; ++k)
    if (!xyz[k]) return fail(KH_E_ARG, "null state column");
  HIPCHK(hipSetDevice(c->device));
  // the objects tuned on: every stride-th of the batch (all when sample is 0
  // or >= n), so the sample keeps the batch's mix of orbits and spans
  const int64_t m = (sample <= 0 || sample >= n) ? n : sample;
  const int64_t stride = n / m;
  const size_t b = align256(m * 8);
  void* buf = nullptr;
  HIPCHK(hipMalloc(&buf, 20 * b));
  char* q = (char*)buf;
  double* s_jd0 = (double*)q;
  double* s_jf = (double*)(q + b);
  const double* s_xyz[6];
  double* s_in[6];
  double* s_out[6];
  for (int k = 0; k < 6; ++k) {
    s_in[k] = (double*)(q + (2 + k) * b);
    s_xyz[k] = s_in[k];
    s_out[k] = (double*)(q + (8 + k) * b);
  }
  int32_t* s_st = (int32_t*)(q + 14 * b);
  int32_t r = launch_gather_states(m, stride, jd0, xyz, jd_final, jd_final_stride, s_jd0, s_in,
                                   s_jf, c->stream)
                  ? fail(KH_E_HIP, "gather launch failed")
                  : KH_OK;
  struct Cand {
    int32_t mode, block, window;
    int64_t latency_k;
  };
  const bool shared = c->spk.shared_grid != 0;
  const int32_t win_max = c->spk.all14 ? (int32_t)std::min<int64_t>(
                                             16, 64 * 1024 / ((int64_t)c->spk.n_seg * 352))
                                       : 0;
  const int64_t lat_default = shared ? (int64_t)c->n_cu * 8 : 0;
  Cand best{0, 256, win_max, lat_default};
  double best_ms = INFINITY, default_ms = INFINITY;
  int32_t tried = 0;
  const bool keep = c->tune.valid;
  const auto saved = c->tune;
  auto run = [&](const Cand& cd) -> double {
    c->tune.valid = true;
    c->tune.mode = cd.mode;
    c->tune.block = cd.block;
    c->tune.window = cd.window;
    c->tune.latency_k = cd.latency_k;
    double ms = INFINITY;
    // one untimed run (first use of a shape), then the timed one
    for (int rep = 0; rep < 2 && r == KH_OK; ++rep) {
      if (hipEventRecord(c->ev0, c->stream) != hipSuccess) r = fail(KH_E_HIP, "event record");
      if (r == KH_OK)
        r = nbody_dev(c, m, s_jd0, s_xyz, center, frame, s_jf, 1, include_asteroids, nullptr,
                      1e-6, -1, s_out, s_st, nullptr, nullptr, nullptr, 0, nullptr);
      hipEvent_t e1 = nullptr;
      if (r == KH_OK && hipEventCreate(&e1) == hipSuccess &&
          hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess) {
        float f = 0.f;
        if (hipEventElapsedTime(&f, c->ev0, e1) == hipSuccess) ms = f;
      } else if (r == KH_OK) {
        r = fail(KH_E_HIP, "timing event failed");
      }
      if (e1) (void)hipEventDestroy(e1);
    }
    ++tried;
    return ms;
  };
  auto consider = [&](const Cand& cd) {
    const double ms = run(cd);
    if (ms < best_ms) {
      best_ms = ms;
      best = cd;
    }
  };
  // coordinate descent from the default shape: threads per object and
  // launch kind, then objects per block, then the LDS window, then the
  // latency-kernel head
  default_ms = run(best);
  best_ms = default_ms;
  const Cand base = best;
  for (int32_t md : {1, 2, 3}) {
    Cand cd = base;
    cd.mode = md;
    consider(cd);
  }
  if (best.mode <= 1)
    for (int32_t bl : {64, 128}) {
      Cand cd = best;
      cd.block = bl;
      consider(cd);
    }
  if (win_max > 0) {
    for (int32_t w : {0, 8, 12}) {
      if (w >= win_max) continue;
      Cand cd = best;
      cd.window = w;
      consider(cd);
    }
  }
  if (best.mode == 0 && shared)
    for (int64_t lk : {(int64_t)0, (int64_t)c->n_cu * 2, (int64_t)c->n_cu * 16,
                       (int64_t)c->n_cu * 32}) {
      Cand cd = best;
      cd.latency_k = lk;
      consider(cd);
    }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  c->tune = saved;
  c->tune.valid = keep;
  if (r != KH_OK) return r;
  out->mode = best.mode;
  out->block = best.block;
  out->window = best.window;
  out->latency_k = best.latency_k;
  out->ms = best_ms;
  out->default_ms = default_ms;
  out->n_tried = tried;
  out->n_sample = m;
  if (apply) {
    c->tune.valid = true;
    c->tune.mode = best.mode;
    c->tune.block = best.block;
    c->tune.window = best.window;
    c->tune.latency_k = best.latency_k;
  }
  return KH_OK;
}

int32_t kh_autotune_clear(kh_ctx* c) {
  if (!c) return fail(KH_E_ARG, "null context");
  c->tune.valid = false;
  return KH_OK;
}

int32_t kh_last_launch(kh_ctx* c, int32_t* block, int32_t* grid, int32_t* mode) {
  if (!c) return fail(KH_E_ARG, "null context");
  if (block) *block = c->last_block;
  if (grid) *grid = c->last_grid;
  if (mode) *mode = c->last_mode;
  return KH_OK;
}

double kh_last_kernel_ms(kh_ctx* c) {
  if (!c || !c->timed) return -1.0;
  if (hipEventSynchronize(c->ev1) != hipSuccess) return -1.0;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) return -1.0;
  return (double)ms;
}

}  // extern "C"

int32_t kh_propagate_n_body_epochs_dev(kh_ctx* c, int64_t n, const double* jd0,
                                       const double* const xyz[6], int64_t center, int32_t frame,
                                       const kh::MultiEpochs& me, int32_t include_asteroids,
                                       double epsilon, double* const out[6], int32_t* status,
                                       int32_t sync) {
  if (me.n_epochs <= 0 || !me.epochs) return fail(KH_E_ARG, "no epochs");
  for (int k = 0; k < 6; ++k)
    if (!me.out[k]) return fail(KH_E_ARG, "null epoch output column");
  return nbody_dev(c, n, jd0, xyz, center, frame, me.epochs, 0, include_asteroids, nullptr,
                   epsilon, 0, out, status, nullptr, nullptr, nullptr, sync, &me);
}

