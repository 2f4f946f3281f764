// kh_fov_host.cpp -- host driver of the field-of-view path: the grouping and
// re-propagation schedule of fov_checks_py (src/kete/rust/fovs/checks.rs:
// 24-110) around the device phases in kh_fov.hip, the rectangle constructor
// (OnSkyRectangle::new, src/kete_core/src/fov/patches.rs:122-155) and the
// RA/Dec + V-magnitude epilogue.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "kh_ctx.h"

using namespace kh;

namespace {

// rotate_around (frames/definitions.rs:356-363): Rodrigues rotation about the
// normalised axis; identity for a zero angle.
void rotate_around(const double* v, const double* axis, double angle, double* o) {
  if (!(angle != 0.0)) {
    std::memcpy(o, v, 3 * sizeof(double));
    return;
  }
  const double n = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  const double ux = axis[0] / n, uy = axis[1] / n, uz = axis[2] / n;
  const double sqx = ux * ux, sqy = uy * uy, sqz = uz * uz;
  const double s = std::sin(angle), c = std::cos(angle), omc = 1.0 - c;
  const double m[3][3] = {
      {sqx + (1.0 - sqx) * c, ux * uy * omc - uz * s, ux * uz * omc + uy * s},
      {ux * uy * omc + uz * s, sqy + (1.0 - sqy) * c, uy * uz * omc - ux * s},
      {ux * uz * omc - uy * s, uy * uz * omc + ux * s, sqz + (1.0 - sqz) * c},
  };
  const double x = v[0], y = v[1], z = v[2];
  for (int i = 0; i < 3; ++i) o[i] = m[i][0] * x + m[i][1] * y + m[i][2] * z;
}

void cross(const double* a, const double* b, double* o) {
  const double x = a[1] * b[2] - a[2] * b[1];
  const double y = a[2] * b[0] - a[0] * b[2];
  const double z = a[0] * b[1] - a[1] * b[0];
  o[0] = x;
  o[1] = y;
  o[2] = z;
}

__global__ void fill_f64(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void zero_counts(unsigned long long* p) {
  if (threadIdx.x < 4) p[threadIdx.x] = 0ull;
}

// per group: the candidate count restarts; hits, n-body re-checks and the
// largest candidate count accumulate over the groups
__global__ void zero_list_counts(unsigned long long* p) {
  if (threadIdx.x == 0) p[0] = 0ull;
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// fields per prefilter launch (the LDS tile of fov_prefilter_kernel)
constexpr int64_t FOV_TILE_FIELDS = kh::FOV_MAX_TILE_FIELDS;
// prefilter launches timed per call (event pairs held by the context)
constexpr int64_t FOV_TIMED_LAUNCHES = 4096;

}  // namespace

extern "C" {

void kh_fov_rect_normals(const double pointing[3], double rotation, double lon_width,
                         double lat_width, double normals[12]) {
  const double zax[3] = {0.0, 0.0, 1.0};
  double up[3], left[3], nleft[3], nup[3];
  rotate_around(zax, pointing, -rotation, up);
  cross(pointing, up, left);
  cross(pointing, left, up);
  rotate_around(left, up, -lon_width / 2.0, normals + 0);
  rotate_around(up, left, lat_width / 2.0, normals + 3);
  for (int i = 0; i < 3; ++i) {
    nleft[i] = -left[i];
    nup[i] = -up[i];
  }
  rotate_around(nleft, up, lon_width / 2.0, normals + 6);
  rotate_around(nup, left, -lat_width / 2.0, normals + 9);
}

int32_t kh_fov_state_check(kh_ctx* c, int64_t n, double jd0, const double* const xyz[6],
                           int32_t frame, int64_t n_fields, const double* field_jd,
                           const double* const obs[6], const double* normals, double dt_limit,
                           int32_t include_asteroids, int64_t cap, int64_t* n_hits,
                           int64_t* hit_field, int64_t* hit_obj, double* const hit_state[7]) {
  if (!c || !n_hits) return kh_fail(KH_E_ARG, "null argument");
  *n_hits = 0;
  if (n < 0 || n_fields < 0 || cap < 0) return kh_fail(KH_E_ARG, "negative size");
  if (frame != KH_FRAME_ECLIPTIC)
    return kh_fail(KH_E_ARG, "fields and states must be ecliptic (kete converts FOVs to it)");
  if (!(dt_limit > 0.0)) return kh_fail(KH_E_ARG, "dt_limit must be positive");
  if (!c->has_spk) return kh_fail(KH_E_NO_SPK, "no ephemeris uploaded");
  if (n == 0 || n_fields == 0) return KH_OK;
  if (!xyz || !field_jd || !obs || !normals || (cap > 0 && (!hit_field || !hit_obj || !hit_state)))
    return kh_fail(KH_E_ARG, "null array");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;

  // fovs.sort_by(jd) (collection.rs:23-36): a stable sort
  std::vector<int64_t> order(n_fields);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return field_jd[a] < field_jd[b]; });
  std::vector<FieldDev> fields(n_fields);
  for (int64_t k = 0; k < n_fields; ++k) {
    const int64_t f = order[k];
    FieldDev& F = fields[k];
    F.jd = field_jd[f];
    for (int q = 0; q < 6; ++q) F.obs[q] = obs[q][f];
    for (int q = 0; q < 12; ++q) F.nrm[q] = normals[12 * f + q];
    F.pad = 0.0;
  }
  // chunks within 2*dt_limit of their first field (checks.rs:34-57)
  std::vector<std::pair<int64_t, int64_t>> chunks;
  for (int64_t k = 0; k < n_fields;) {
    int64_t e = k + 1;
    while (e < n_fields && !(std::fabs(fields[e].jd - fields[k].jd) >= 2.0 * dt_limit)) ++e;
    chunks.push_back({k, e});
    k = e;
  }
  // a chunk's fields are checked in tiles of at most FOV_TILE_FIELDS (the
  // prefilter stages one tile in LDS); the tiles share the chunk's states
  int64_t n_tiles = 0;
  for (auto& ch : chunks) n_tiles += (ch.second - ch.first + FOV_TILE_FIELDS - 1) / FOV_TILE_FIELDS;
  const int64_t n_timed = std::min<int64_t>(n_tiles, FOV_TIMED_LAUNCHES);
  while ((int64_t)c->fov_ev.size() < 2 * n_timed) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->fov_ev.push_back(e);
  }
  c->fov_pf_ms = -1.0;
  c->fov_pf_launches = c->fov_pf_pairs = 0;

  // The states each chunk is checked against.  kete re-propagates all states
  // at each chunk (small steps, checks.rs:74-85) and every 50 dt_limit from a
  // second copy (big steps, :63-72), restarting the integrator each time.  By
  // default here the states are integrated once through the same small-step
  // epochs (kh_propagate_n_body_epochs_dev): the same states to the
  // integrator's error level (~1e-12 AU), far below the edge tolerance of the
  // visible sets, at a fraction of the work.  KH_FOV_SCHEDULE=kete runs
  // kete's schedule with restarts.
  const char* sched = getenv("KH_FOV_SCHEDULE");
  const bool continuous = !(sched && std::string(sched) == "kete");
  std::vector<double> epochs;             // distinct small-step epochs, in order
  std::vector<int64_t> chunk_epoch(chunks.size());
  {
    double jd_s = jd0;
    for (size_t k = 0; k < chunks.size(); ++k) {
      const double jd_mean =
          (fields[chunks[k].second - 1].jd + fields[chunks[k].first].jd) / 2.0;
      if (std::fabs(jd_mean - jd_s) >= dt_limit) {
        jd_s = jd_mean;
        epochs.push_back(jd_s);
      }
      chunk_epoch[k] = (int64_t)epochs.size() - 1;  // -1: still at jd0
    }
  }
  // epochs per multi-epoch launch: at most ~4 GiB of stored states
  const int64_t win_epochs =
      continuous ? std::max<int64_t>(1, std::min<int64_t>((int64_t)epochs.size(),
                                                          (4ll << 30) / (48 * n)))
                 : 0;

  // device memory: states (cur, big) with per-object epochs, fields, lists
  int64_t cand_cap = std::max<int64_t>(1 << 20, 8 * n);
  int64_t nb_cap = std::max<int64_t>(1 << 16, 2 * n);  // all groups together
  // KH_FOV_LIST_CAP sets both lists' first capacity (a caller short of device
  // memory; the lists grow, with a rerun, when a call overflows them)
  if (const char* lc = getenv("KH_FOV_LIST_CAP")) {
    const long long v = atoll(lc);
    if (v > 0) cand_cap = nb_cap = (int64_t)v;
  }
  const int64_t hit_cap = std::max<int64_t>(cap, 1);
  for (int attempt = 0; attempt < 3; ++attempt) {
    const size_t bn = al(n * 8), bs = al(n * 4);
    const size_t bw = al((size_t)win_epochs * n * 8);
    const size_t need = 2 * (7 * bn) + bs + al(n_fields * sizeof(FieldDev)) +
                        al(cand_cap * 4) * 2 + al(cand_cap * 8) + al(nb_cap * 4) +
                        al(nb_cap * 8) + al(nb_cap * 8) * 14 + al(nb_cap * 4) +
                        al(hit_cap * 4) + al(hit_cap * 8) * 8 + al(8) +
                        (continuous ? 6 * bw + bn + al(win_epochs * 8) + 6 * bn : 0);
    int32_t r = kh_grow(&c->fbuf, &c->fbuf_size, need);
    if (r) return r;
    char* q = (char*)c->fbuf;
    auto take = [&](size_t b) {
      char* p = q;
      q += b;
      return p;
    };
    double* cur_jd = (double*)take(bn);
    double* cur[6];
    for (int k = 0; k < 6; ++k) cur[k] = (double*)take(bn);
    double* big_jd = (double*)take(bn);
    double* big[6];
    for (int k = 0; k < 6; ++k) big[k] = (double*)take(bn);
    int32_t* st = (int32_t*)take(bs);
    FieldDev* d_fields = (FieldDev*)take(al(n_fields * sizeof(FieldDev)));
    int32_t* cand_field = (int32_t*)take(al(cand_cap * 4));
    int32_t* cand_kind = (int32_t*)take(al(cand_cap * 4));
    int64_t* cand_obj = (int64_t*)take(al(cand_cap * 8));
    int32_t* nb_field = (int32_t*)take(al(nb_cap * 4));
    int64_t* nb_obj = (int64_t*)take(al(nb_cap * 8));
    double* g_jd0 = (double*)take(al(nb_cap * 8));
    double* g_jdf = (double*)take(al(nb_cap * 8));
    double* g_x[6];
    double* g_out[6];
    for (int k = 0; k < 6; ++k) g_x[k] = (double*)take(al(nb_cap * 8));
    for (int k = 0; k < 6; ++k) g_out[k] = (double*)take(al(nb_cap * 8));
    int32_t* g_st = (int32_t*)take(al(nb_cap * 4));
    int32_t* h_field = (int32_t*)take(al(hit_cap * 4));
    int64_t* h_obj = (int64_t*)take(al(hit_cap * 8));
    double* h_state[7];
    for (int k = 0; k < 7; ++k) h_state[k] = (double*)take(al(hit_cap * 8));
    double* d_scalar = (double*)take(al(8));
    double* m_out[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double* ep_jd = nullptr;    // the epoch of the states a chunk uses, per object
    double* d_epochs = nullptr;  // this window's epochs
    double* fin[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (continuous) {
      for (int k = 0; k < 6; ++k) m_out[k] = (double*)take(bw);
      ep_jd = (double*)take(bn);
      d_epochs = (double*)take(al(win_epochs * 8));
      for (int k = 0; k < 6; ++k) fin[k] = (double*)take(bn);
    }

    HIPCHK(hipMemcpyAsync(d_fields, fields.data(), n_fields * sizeof(FieldDev),
                          hipMemcpyHostToDevice, s));
    for (int k = 0; k < 6; ++k) {
      HIPCHK(hipMemcpyAsync(cur[k], xyz[k], n * 8, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(big[k], cur[k], n * 8, hipMemcpyDeviceToDevice, s));
    }
    const unsigned gb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(fill_f64, dim3(gb), dim3(256), 0, s, cur_jd, n, jd0);
    hipLaunchKernelGGL(fill_f64, dim3(gb), dim3(256), 0, s, big_jd, n, jd0);
    hipLaunchKernelGGL(zero_counts, dim3(1), dim3(64), 0, s, c->d_counts);
    unsigned long long* d_nhit = c->d_counts + 1;

    // propagate states (in place) to jd_target, then set their epochs
    auto propagate = [&](double* const* x, double* xjd, double jd_target) -> int32_t {
      hipLaunchKernelGGL(fill_f64, dim3(1), dim3(64), 0, s, d_scalar, (int64_t)1, jd_target);
      const double* xin[6] = {x[0], x[1], x[2], x[3], x[4], x[5]};
      int32_t rr = kh_propagate_n_body_dev(c, n, xjd, xin, 10, frame, d_scalar, 0, include_asteroids,
                                           nullptr, 1e-6,
                                           -1, x, st, nullptr, nullptr, nullptr, 0);
      if (rr) return rr;
      hipLaunchKernelGGL(fill_f64, dim3(gb), dim3(256), 0, s, xjd, n, jd_target);
      return KH_OK;
    };

    double jd = jd0, big_epoch_ref = jd0;
    bool overflow = false;
    int64_t tile = 0;  // prefilter launches so far (timed while < n_timed)
    int64_t w0 = -1, w1 = -1;  // epochs [w0, w1) are stored in m_out
    for (size_t ci_ = 0; ci_ < chunks.size(); ++ci_) {
      auto& ch = chunks[ci_];
      const int64_t f0 = ch.first, f1 = ch.second;
      const double jd_mean = (fields[f1 - 1].jd + fields[f0].jd) / 2.0;
      double* sx[6];     // states this chunk is checked against
      double* sjd = cur_jd;
      for (int k = 0; k < 6; ++k) sx[k] = cur[k];
      if (continuous) {
        const int64_t e = chunk_epoch[ci_];
        if (e >= 0) {
          if (e >= w1) {
            // next window of epochs: from the states at the last stored
            // epoch (or the inputs), one launch through the window
            if (w1 > 0) {
              const int64_t last = (w1 - 1 - w0) * n;
              for (int k = 0; k < 6; ++k)
                HIPCHK(hipMemcpyAsync(cur[k], m_out[k] + last, n * 8, hipMemcpyDeviceToDevice,
                                      s));
              hipLaunchKernelGGL(fill_f64, dim3(gb), dim3(256), 0, s, cur_jd, n,
                                 epochs[w1 - 1]);
            }
            w0 = e;
            w1 = std::min<int64_t>((int64_t)epochs.size(), w0 + win_epochs);
            HIPCHK(hipMemcpyAsync(d_epochs, epochs.data() + w0, (w1 - w0) * 8,
                                  hipMemcpyHostToDevice, s));
            MultiEpochs me{};
            me.epochs = d_epochs;
            me.n_epochs = (int32_t)(w1 - w0);
            me.lo = *std::min_element(epochs.begin() + w0, epochs.begin() + w1);
            me.hi = *std::max_element(epochs.begin() + w0, epochs.begin() + w1);
            for (int k = 0; k < 6; ++k) me.out[k] = m_out[k];
            const double* xin[6] = {cur[0], cur[1], cur[2], cur[3], cur[4], cur[5]};
            r = kh_propagate_n_body_epochs_dev(c, n, cur_jd, xin, 10, frame, me, include_asteroids,
                                               1e-6, fin, st,
                                               0);
            if (r) return r;
          }
          for (int k = 0; k < 6; ++k) sx[k] = m_out[k] + (e - w0) * n;
          hipLaunchKernelGGL(fill_f64, dim3(gb), dim3(256), 0, s, ep_jd, n, epochs[e]);
          sjd = ep_jd;
        }
      } else {
        // big steps (checks.rs:63-72): the big-step states move to the
        // current small-step epoch
        if (std::fabs(jd_mean - big_epoch_ref) >= dt_limit * 50.0) {
          big_epoch_ref = jd_mean;
          r = propagate(big, big_jd, jd);
          if (r) return r;
        }
        // small steps (checks.rs:74-85)
        if (std::fabs(jd_mean - jd) >= dt_limit) {
          if (std::fabs(jd - big_epoch_ref) >= dt_limit * 25.0) {
            for (int k = 0; k < 6; ++k)
              HIPCHK(hipMemcpyAsync(cur[k], big[k], n * 8, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(cur_jd, big_jd, n * 8, hipMemcpyDeviceToDevice, s));
          }
          jd = jd_mean;
          r = propagate(cur, cur_jd, jd);
          if (r) return r;
        }
      }
      for (int64_t t0 = f0; t0 < f1; t0 += FOV_TILE_FIELDS, ++tile) {
        const int64_t t1 = std::min<int64_t>(f1, t0 + FOV_TILE_FIELDS);
        // phase 1: prefilter over (object, field) pairs of this tile
        hipLaunchKernelGGL(zero_list_counts, dim3(1), dim3(64), 0, s, c->d_counts);
        FovPrefilter pf{};
        pf.n = n;
        pf.jd = sjd;
        for (int k = 0; k < 6; ++k) pf.x[k] = sx[k];
        pf.fields = d_fields + t0;
        pf.nf = (int32_t)(t1 - t0);
        pf.dt_limit = dt_limit;
        pf.n_cand = c->d_counts;
        pf.cap = cand_cap;
        pf.cand_field = cand_field;
        pf.cand_kind = cand_kind;
        pf.cand_obj = cand_obj;
        if (tile < n_timed) HIPCHK(hipEventRecord(c->fov_ev[2 * tile], s));
        if (launch_fov_prefilter(pf, s)) return kh_fail(KH_E_HIP, "prefilter launch failed");
        if (tile < n_timed) HIPCHK(hipEventRecord(c->fov_ev[2 * tile + 1], s));
        if (tile < n_timed) c->fov_pf_pairs += n * (t1 - t0);
        // phase 2: two-body checks of the candidates; those whose field is
        // more than dt_limit from the states' epoch join the n-body list
        FovTwoBody tb{};
        tb.n_cand = c->d_counts;
        tb.cand_max = c->d_counts + 3;
        tb.cand_cap = cand_cap;
        tb.jd = sjd;
        for (int k = 0; k < 6; ++k) tb.x[k] = sx[k];
        tb.fields = d_fields + t0;
        tb.field_base = (int32_t)t0;
        tb.dt_limit = dt_limit;
        tb.cand_field = cand_field;
        tb.cand_kind = cand_kind;
        tb.cand_obj = cand_obj;
        tb.n_hit = d_nhit;
        tb.hit_cap = hit_cap;
        tb.hit_field = h_field;
        tb.hit_obj = h_obj;
        for (int k = 0; k < 7; ++k) tb.hit_state[k] = h_state[k];
        tb.n_nb = c->d_counts + 2;
        tb.nb_cap = nb_cap;
        tb.nb_field = nb_field;
        tb.nb_obj = nb_obj;
        tb.g_jd0 = g_jd0;
        tb.g_jdf = g_jdf;
        for (int k = 0; k < 6; ++k) tb.g_x[k] = g_x[k];
        if (launch_fov_two_body(tb, s)) return kh_fail(KH_E_HIP, "two-body launch failed");
      }
    }
    // the groups ran without a host round trip; the list sizes are checked
    // once, here, and an overflow reruns the call with larger lists
    unsigned long long counts[4];
    HIPCHK(hipMemcpyAsync(counts, c->d_counts, sizeof(counts), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if ((int64_t)counts[3] > cand_cap) {
      cand_cap = (int64_t)counts[3] * 2;
      overflow = true;
    }
    if ((int64_t)counts[2] > nb_cap) {
      nb_cap = (int64_t)counts[2] * 2;
      overflow = true;
    }
    // phase 3: n-body re-checks of all groups in one launch (fields more than
    // dt_limit from their group's epoch), from the states gathered in phase 2
    if (!overflow && counts[2] > 0) {
      FovNBody nb{};
      nb.m = (int64_t)counts[2];
      nb.fields = d_fields;
      nb.nb_field = nb_field;
      nb.nb_obj = nb_obj;
      for (int k = 0; k < 6; ++k) nb.g_out[k] = g_out[k];
      nb.g_status = g_st;
      nb.n_hit = d_nhit;
      nb.hit_cap = hit_cap;
      nb.hit_field = h_field;
      nb.hit_obj = h_obj;
      for (int k = 0; k < 7; ++k) nb.hit_state[k] = h_state[k];
      const double* gx[6] = {g_x[0], g_x[1], g_x[2], g_x[3], g_x[4], g_x[5]};
      r = kh_propagate_n_body_dev(c, nb.m, g_jd0, gx, 10, frame, g_jdf, 1, include_asteroids,
                                  nullptr, 1e-6, -1,
                                  g_out, g_st, nullptr, nullptr, nullptr, 0);
      if (r) return r;
      if (launch_fov_nb_finish(nb, s)) return kh_fail(KH_E_HIP, "n-body finish failed");
    }
    if (overflow) {
      c->fov_pf_pairs = 0;
      continue;
    }
    {
      float ms = 0.0f;
      double sum = 0.0;
      for (int64_t k = 0; k < n_timed; ++k) {
        HIPCHK(hipEventElapsedTime(&ms, c->fov_ev[2 * k], c->fov_ev[2 * k + 1]));
        sum += ms;
      }
      c->fov_pf_ms = sum;
      c->fov_pf_launches = n_timed;
    }
    unsigned long long total = 0;
    HIPCHK(hipMemcpyAsync(&total, d_nhit, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *n_hits = (int64_t)total;
    if ((int64_t)total > cap) return kh_fail(KH_E_ARG, "hit capacity too small: need " +
                                                           std::to_string(total));
    // hits in (sorted field, object index) order, the order kete returns:
    // a device radix sort on packed keys, then one gather into the outputs
    if (total) {
      auto bits = [](int64_t v) {
        int32_t b = 1;
        while (b < 63 && ((int64_t)1 << b) <= v) ++b;
        return b;
      };
      const int32_t obj_bits = bits(n - 1), field_bits = bits(n_fields - 1);
      const int64_t m = (int64_t)total;
      const size_t tmp = fov_sort_temp_bytes(m);
      const size_t b8 = al(m * 8);
      r = kh_grow(&c->hbuf, &c->hbuf_size, al(n_fields * 8) + tmp + 9 * b8);
      if (r) return r;
      char* q2 = (char*)c->hbuf;
      int64_t* d_order = (int64_t*)q2;
      q2 += al(n_fields * 8);
      void* d_tmp = q2;
      q2 += tmp;
      int64_t* o_field = (int64_t*)q2;
      int64_t* o_obj = (int64_t*)(q2 + b8);
      double* o_state[7];
      for (int k = 0; k < 7; ++k) o_state[k] = (double*)(q2 + (2 + k) * b8);
      HIPCHK(hipMemcpyAsync(d_order, order.data(), n_fields * 8, hipMemcpyHostToDevice, s));
      if (launch_fov_sort_hits(m, obj_bits, field_bits, h_field, h_obj, h_state, d_order, d_tmp,
                               tmp, o_field, o_obj, o_state, s))
        return kh_fail(KH_E_HIP, "hit ordering failed");
      HIPCHK(hipMemcpyAsync(hit_field, o_field, m * 8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(hit_obj, o_obj, m * 8, hipMemcpyDeviceToHost, s));
      for (int k = 0; k < 7; ++k)
        HIPCHK(hipMemcpyAsync(hit_state[k], o_state[k], m * 8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
    }
    return KH_OK;
  }
  return kh_fail(KH_E_HIP, "candidate lists overflowed twice");
}

int32_t kh_fov_last_prefilter(kh_ctx* c, double* ms, int64_t* launches, int64_t* pairs) {
  if (!c || !ms || !launches || !pairs) return kh_fail(KH_E_ARG, "null argument");
  if (c->fov_pf_ms < 0.0) return kh_fail(KH_E_ARG, "no completed fov_state_check call");
  *ms = c->fov_pf_ms;
  *launches = c->fov_pf_launches;
  *pairs = c->fov_pf_pairs;
  return KH_OK;
}

int32_t kh_fov_spk_check(kh_ctx* c, int64_t n_ids, const int64_t* obj_ids, int32_t frame,
                         int64_t n_fields, const double* field_jd, const double* const obs[6],
                         const double* normals, int64_t cap, int64_t* n_hits, int64_t* hit_field,
                         int64_t* hit_obj, double* const hit_state[7]) {
  if (!c || !n_hits) return kh_fail(KH_E_ARG, "null argument");
  *n_hits = 0;
  if (n_ids < 0 || n_fields < 0 || cap < 0) return kh_fail(KH_E_ARG, "negative size");
  if (frame != KH_FRAME_ECLIPTIC)
    return kh_fail(KH_E_ARG, "fields must be ecliptic (kete converts FOVs to it)");
  if (!c->has_spk) return kh_fail(KH_E_NO_SPK, "no ephemeris uploaded");
  if (n_ids == 0 || n_fields == 0) return KH_OK;
  if (!obj_ids || !field_jd || !obs || !normals ||
      (cap > 0 && (!hit_field || !hit_obj || !hit_state)))
    return kh_fail(KH_E_ARG, "null array");
  if (n_ids * n_fields >= ((int64_t)1 << 40)) return kh_fail(KH_E_ARG, "too many pairs");
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // fields in epoch order (into_sorted_vec_fov: a stable sort by jd)
  std::vector<int64_t> order(n_fields);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return field_jd[a] < field_jd[b]; });
  std::vector<FieldDev> fields(n_fields);
  for (int64_t k = 0; k < n_fields; ++k) {
    const int64_t f = order[k];
    FieldDev& F = fields[k];
    F.jd = field_jd[f];
    for (int q = 0; q < 6; ++q) F.obs[q] = obs[q][f];
    for (int q = 0; q < 12; ++q) F.nrm[q] = normals[12 * f + q];
    F.pad = 0.0;
  }
  // where each object's state comes from (SpkCollection::try_get_raw_state
  // scans an object's segments, spk.rs:65-82): Type 21 segments, else the
  // planet tree; objects in neither are never visible (filter_map)
  std::vector<SpkObjDesc> desc(n_ids);
  for (int64_t k = 0; k < n_ids; ++k) {
    SpkObjDesc& d = desc[k];
    std::memset(&d, 0, sizeof(d));
    for (size_t q = 0; q < c->t21_obj.size(); ++q)
      if (c->t21_obj[q] == obj_ids[k]) {
        if (d.kind == 0) d.t21_begin = (int32_t)q;
        d.kind = 1;
        d.t21_end = (int32_t)q + 1;
      }
    if (d.kind == 0 && kh_build_chain(c, obj_ids[k], d.body)) d.kind = 2;
  }
  Chain sun{};
  if (!kh_build_chain(c, 10, sun)) return kh_fail(KH_E_NO_SPK, "no path from the Sun to 0");
  const int64_t hit_cap = std::max<int64_t>(cap, 1);
  const size_t bd = al(n_ids * sizeof(SpkObjDesc)), bf = al(n_fields * sizeof(FieldDev));
  const size_t need = bd + bf + al(8) + al(hit_cap * 4) + al(hit_cap * 8) * 8;
  int32_t r = kh_grow(&c->fbuf, &c->fbuf_size, need);
  if (r) return r;
  char* q = (char*)c->fbuf;
  SpkObjDesc* d_desc = (SpkObjDesc*)q;
  FieldDev* d_fields = (FieldDev*)(q + bd);
  unsigned long long* d_n = (unsigned long long*)(q + bd + bf);
  int32_t* h_field = (int32_t*)(q + bd + bf + al(8));
  char* q2 = q + bd + bf + al(8) + al(hit_cap * 4);
  int64_t* h_obj = (int64_t*)q2;
  double* h_state[7];
  for (int k = 0; k < 7; ++k) h_state[k] = (double*)(q2 + (1 + k) * al(hit_cap * 8));
  HIPCHK(hipMemcpyAsync(d_desc, desc.data(), n_ids * sizeof(SpkObjDesc), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_fields, fields.data(), n_fields * sizeof(FieldDev),
                        hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(zero_counts, dim3(1), dim3(64), 0, s, d_n);
  FovSpk p{};
  p.n_obj = n_ids;
  p.n_fields = n_fields;
  p.obj = d_desc;
  p.sun = sun;
  p.fields = d_fields;
  p.n_hit = d_n;
  p.hit_cap = hit_cap;
  p.hit_field = h_field;
  p.hit_obj = h_obj;
  for (int k = 0; k < 7; ++k) p.hit_state[k] = h_state[k];
  if (launch_fov_spk(c->d_spk, c->d_coef, p, s)) return kh_fail(KH_E_HIP, "fov_spk launch failed");
  unsigned long long total = 0;
  HIPCHK(hipMemcpyAsync(&total, d_n, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_hits = (int64_t)total;
  if ((int64_t)total > cap)
    return kh_fail(KH_E_ARG, "hit capacity too small: need " + std::to_string(total));
  if (total) {
    auto bits = [](int64_t v) {
      int32_t b = 1;
      while (b < 63 && ((int64_t)1 << b) <= v) ++b;
      return b;
    };
    const int64_t m = (int64_t)total;
    const size_t tmp = fov_sort_temp_bytes(m);
    const size_t b8 = al(m * 8);
    r = kh_grow(&c->hbuf, &c->hbuf_size, al(n_fields * 8) + tmp + 9 * b8);
    if (r) return r;
    char* h = (char*)c->hbuf;
    int64_t* d_order = (int64_t*)h;
    h += al(n_fields * 8);
    void* d_tmp = h;
    h += tmp;
    int64_t* o_field = (int64_t*)h;
    int64_t* o_obj = (int64_t*)(h + b8);
    double* o_state[7];
    for (int k = 0; k < 7; ++k) o_state[k] = (double*)(h + (2 + k) * b8);
    HIPCHK(hipMemcpyAsync(d_order, order.data(), n_fields * 8, hipMemcpyHostToDevice, s));
    if (launch_fov_sort_hits(m, bits(n_ids - 1), bits(n_fields - 1), h_field, h_obj, h_state,
                             d_order, d_tmp, tmp, o_field, o_obj, o_state, s))
      return kh_fail(KH_E_HIP, "hit ordering failed");
    HIPCHK(hipMemcpyAsync(hit_field, o_field, m * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hit_obj, o_obj, m * 8, hipMemcpyDeviceToHost, s));
    for (int k = 0; k < 7; ++k)
      HIPCHK(hipMemcpyAsync(hit_state[k], o_state[k], m * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return KH_OK;
}

int32_t kh_radec_vmag(kh_ctx* c, int64_t n, const double* const state[6],
                      const double* const obs[6], int32_t frame, const double* h_mag,
                      double g_param, double* const out[5]) {
  if (!c || n < 0) return kh_fail(KH_E_ARG, "bad argument");
  if (n == 0) return KH_OK;
  if (!state || !obs || !h_mag || !out) return kh_fail(KH_E_ARG, "null array");
  if (frame != 1 && frame != 17) return kh_fail(KH_E_ARG, "frame must be 1 or 17");
  HIPCHK(hipSetDevice(c->device));
  const size_t bn = al(n * 8);
  int32_t r = kh_grow(&c->fbuf, &c->fbuf_size, 18 * bn);
  if (r) return r;
  double* base = (double*)c->fbuf;
  RadecArgs a{};
  a.n = n;
  for (int k = 0; k < 6; ++k) {
    double* ds = (double*)((char*)base + k * bn);
    double* dobs = (double*)((char*)base + (6 + k) * bn);
    HIPCHK(hipMemcpyAsync(ds, state[k], n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dobs, obs[k], n * 8, hipMemcpyHostToDevice, c->stream));
    a.st[k] = ds;
    a.obs[k] = dobs;
  }
  double* dh = (double*)((char*)base + 12 * bn);
  HIPCHK(hipMemcpyAsync(dh, h_mag, n * 8, hipMemcpyHostToDevice, c->stream));
  a.h = dh;
  a.g = g_param;
  a.frame = frame;
  for (int k = 0; k < 5; ++k) a.out[k] = (double*)((char*)base + (13 + k) * bn);
  if (launch_radec_vmag(a, c->stream)) return kh_fail(KH_E_HIP, "radec launch failed");
  for (int k = 0; k < 5; ++k)
    HIPCHK(hipMemcpyAsync(out[k], a.out[k], n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

}  // extern "C"
