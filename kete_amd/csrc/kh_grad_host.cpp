// kh_grad_host.cpp -- C ABI of the observable/partials path (kh_grad.hip):
// argument checks, staging of host arrays, launch.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kh_ctx.h"

using namespace kh;
#define fail kh_fail

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int32_t check_ng(const kh_nongrav* ng, int64_t n) {
  for (int64_t i = 0; ng && i < n; ++i) {
    if (ng[i].kind == 2)
      return fail(KH_E_ARG, "dust non-grav models have no A1..A3 parameters to differentiate");
    if (ng[i].kind == 1 && ng[i].dt != 0.0)
      return fail(KH_E_ARG, "JplComet non-grav with a time delay (dt != 0) is not supported");
    if (ng[i].kind < 0 || ng[i].kind > 2) return fail(KH_E_ARG, "unknown non-grav kind");
  }
  return KH_OK;
}

int32_t check_args(kh_ctx* c, int64_t n, const double* jd0, const double* const xyz[6],
                   int32_t frame, int32_t m, const double* jd_obs, const double* const obs[3],
                   int32_t force, double epsilon, double* const radec[2], double* const jac[2],
                   double* const state[6], double* const dstate[6], int32_t* status,
                   Chain& sun) {
  if (!c) return fail(KH_E_ARG, "null context");
  if (n < 0 || m < 0) return fail(KH_E_ARG, "negative size");
  if (frame != 1 && frame != 17) return fail(KH_E_ARG, "frame must be 1 or 17");
  if (force != 0 && force != 1) return fail(KH_E_ARG, "force must be 0 (PLANETS) or 1 (central)");
  if (!(epsilon > 0.0)) return fail(KH_E_ARG, "epsilon must be positive");
  if (n == 0) return KH_OK;
  if (!jd0 || !xyz || !status || !radec || !jac || (m > 0 && (!jd_obs || !obs)))
    return fail(KH_E_ARG, "null array");
  for (int k = 0; k < 6; ++k)
    if (!xyz[k]) return fail(KH_E_ARG, "null state column");
  for (int k = 0; m > 0 && k < 3; ++k)
    if (!obs[k]) return fail(KH_E_ARG, "null observer column");
  if (!radec[0] || !radec[1] || !jac[0] || !jac[1]) return fail(KH_E_ARG, "null output column");
  if (dstate && !state) return fail(KH_E_ARG, "dstate needs state");
  for (int k = 0; k < 6; ++k) {
    if (state && !state[k]) return fail(KH_E_ARG, "null state output column");
    if (dstate && !dstate[k]) return fail(KH_E_ARG, "null dstate output column");
  }
  sun.n = 0;
  if (force == 0) {
    if (!c->has_spk) return fail(KH_E_NO_SPK, "no ephemeris uploaded");
    if (!kh_build_chain(c, 10, sun)) return fail(KH_E_NO_SPK, "the Sun (10) is not in the ephemeris");
  }
  return KH_OK;
}

int32_t run_partials(kh_ctx* c, int64_t n, const double* jd0, const double* const xyz[6],
                     int32_t frame, const kh_nongrav* ng, int32_t m, const double* jd_obs,
                     const double* const obs[3], int32_t force, int32_t restart, double epsilon,
                     double* const radec[2], double* const jac[2], double* const state[6],
                     double* const dstate[6], int32_t* status, kh_counters* counters,
                     const Chain& sun, int32_t sync) {
  GradLaunch g{};
  g.n = n;
  g.m = m;
  g.frame = frame;
  g.force = force;
  g.restart = restart ? 1 : 0;
  g.epsilon = epsilon;
  g.jd0 = jd0;
  for (int k = 0; k < 6; ++k) {
    g.xyz[k] = xyz[k];
    g.state[k] = state ? state[k] : nullptr;
    g.dstate[k] = dstate ? dstate[k] : nullptr;
  }
  g.ng = ng;
  g.jd_obs = jd_obs;
  for (int k = 0; k < 3; ++k) g.obs[k] = m > 0 ? obs[k] : nullptr;
  g.sun = sun;
  g.radec[0] = radec[0];
  g.radec[1] = radec[1];
  g.jac[0] = jac[0];
  g.jac[1] = jac[1];
  g.status = status;
  g.counters = reinterpret_cast<int32_t*>(counters);
  // working memory: the states at every epoch, for the observation kernel
  {
    const size_t mn = (size_t)m * (size_t)n;
    int32_t r = kh_grow(&c->pbuf, &c->pbuf_size, partials_buf_bytes(n, m));
    if (r) return r;
    char* q = (char*)c->pbuf;
    g.ep_state = (double*)q;
    g.ep_sun = g.ep_state + mn * 60;
    q += (mn * 66 * 8 + 255) / 256 * 256;
    g.ep_done = (int32_t*)q;
    q += ((size_t)n * 4 + 255) / 256 * 256;
    g.lt_fail = (int32_t*)q;
  }
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (launch_radau_partials(c->d_spk, c->d_coef, g, c->stream))
    return fail(KH_E_HIP, "partials launch failed");
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  if (sync) HIPCHK(hipStreamSynchronize(c->stream));
  return KH_OK;
}

}  // namespace

extern "C" {

int32_t kh_radec_partials_dev(kh_ctx* c, int64_t n, const double* jd0,
                              const double* const xyz[6], int32_t frame, const kh_nongrav* ng,
                              int32_t m, const double* jd_obs, const double* const obs[3],
                              int32_t force, int32_t restart, double epsilon, double* const radec[2],
                              double* const jac[2], double* const state[6],
                              double* const dstate[6], int32_t* status, kh_counters* counters,
                              int32_t sync) {
  Chain sun;
  int32_t r = check_args(c, n, jd0, xyz, frame, m, jd_obs, obs, force, epsilon, radec, jac, state,
                         dstate, status, sun);
  if (r) return r;
  if (n == 0) return KH_OK;
  HIPCHK(hipSetDevice(c->device));
  // device models are checked per object in the kernel (ValueError), with
  // no host round trip; the one-lane-per-parameter kernel checks on the host
  const char* sel = getenv("KH_PARTIALS");
  if (ng && sel && strcmp(sel, "lane9") == 0) {
    std::vector<kh_nongrav> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), ng, n * sizeof(kh_nongrav), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    r = check_ng(h.data(), n);
    if (r) return r;
  }
  return run_partials(c, n, jd0, xyz, frame, ng, m, jd_obs, obs, force, restart, epsilon, radec,
                      jac, state, dstate, status, counters, sun, sync);
}

int32_t kh_radec_partials(kh_ctx* c, int64_t n, const double* jd0, const double* const xyz[6],
                          int32_t frame, const kh_nongrav* ng, int32_t m, const double* jd_obs,
                          const double* const obs[3], int32_t force, int32_t restart, double epsilon,
                          double* const radec[2], double* const jac[2], double* const state[6],
                          double* const dstate[6], int32_t* status, kh_counters* counters) {
  Chain sun;
  int32_t r = check_args(c, n, jd0, xyz, frame, m, jd_obs, obs, force, epsilon, radec, jac, state,
                         dstate, status, sun);
  if (r) return r;
  r = check_ng(ng, n);
  if (r) return r;
  if (n == 0) return KH_OK;
  HIPCHK(hipSetDevice(c->device));
  const size_t mn = (size_t)m * (size_t)n;
  const size_t bn = align256(n * 8), bmn = align256(mn * 8), bng = align256(n * sizeof(kh_nongrav)),
               bs = align256(n * 4), bc = align256(n * sizeof(kh_counters));
  const size_t bytes = 7 * bn + (ng ? bng : 0) + 4 * bmn + 2 * bmn + 18 * bmn +
                       (state ? 6 * bmn : 0) + (dstate ? 54 * bmn : 0) + bs + bc;
  r = kh_grow(&c->dbuf, &c->dbuf_size, bytes);
  if (r) return r;
  char* q = (char*)c->dbuf;
  auto take = [&](size_t b) {
    char* p = q;
    q += b;
    return p;
  };
  double* d_jd0 = (double*)take(bn);
  const double* d_xyz[6];
  for (int k = 0; k < 6; ++k) d_xyz[k] = (const double*)take(bn);
  kh_nongrav* d_ng = ng ? (kh_nongrav*)take(bng) : nullptr;
  double* d_jo = (double*)take(bmn);
  const double* d_obs[3];
  for (int k = 0; k < 3; ++k) d_obs[k] = (const double*)take(bmn);
  double* d_radec[2];
  for (int k = 0; k < 2; ++k) d_radec[k] = (double*)take(bmn);
  double* d_jac[2];
  for (int k = 0; k < 2; ++k) d_jac[k] = (double*)take(9 * bmn);
  double* d_state[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double* d_dstate[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (state)
    for (int k = 0; k < 6; ++k) d_state[k] = (double*)take(bmn);
  if (dstate)
    for (int k = 0; k < 6; ++k) d_dstate[k] = (double*)take(9 * bmn);
  int32_t* d_st = (int32_t*)take(bs);
  kh_counters* d_cnt = (kh_counters*)take(bc);
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(d_jd0, jd0, n * 8, hipMemcpyHostToDevice, s));
  for (int k = 0; k < 6; ++k)
    HIPCHK(hipMemcpyAsync((void*)d_xyz[k], xyz[k], n * 8, hipMemcpyHostToDevice, s));
  if (ng) HIPCHK(hipMemcpyAsync(d_ng, ng, n * sizeof(kh_nongrav), hipMemcpyHostToDevice, s));
  if (m > 0) {
    HIPCHK(hipMemcpyAsync(d_jo, jd_obs, mn * 8, hipMemcpyHostToDevice, s));
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemcpyAsync((void*)d_obs[k], obs[k], mn * 8, hipMemcpyHostToDevice, s));
  }
  r = run_partials(c, n, d_jd0, d_xyz, frame, d_ng, m, d_jo, d_obs, force, restart, epsilon,
                   d_radec,
                   d_jac, state ? d_state : nullptr, dstate ? d_dstate : nullptr, d_st,
                   counters ? d_cnt : nullptr, sun, 0);
  if (r) return r;
  for (int k = 0; k < 2; ++k) {
    HIPCHK(hipMemcpyAsync(radec[k], d_radec[k], mn * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(jac[k], d_jac[k], 9 * mn * 8, hipMemcpyDeviceToHost, s));
  }
  for (int k = 0; state && k < 6; ++k)
    HIPCHK(hipMemcpyAsync(state[k], d_state[k], mn * 8, hipMemcpyDeviceToHost, s));
  for (int k = 0; dstate && k < 6; ++k)
    HIPCHK(hipMemcpyAsync(dstate[k], d_dstate[k], 9 * mn * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(status, d_st, n * 4, hipMemcpyDeviceToHost, s));
  if (counters)
    HIPCHK(hipMemcpyAsync(counters, d_cnt, n * sizeof(kh_counters), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return KH_OK;
}

}  // extern "C"
