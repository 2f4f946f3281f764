// kh_ctx.h -- the library context and host helpers shared by the C ABI
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/kete_hip.h"
#include "kh_internal.h"

int32_t kh_fail(int32_t code, const std::string& msg);

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return kh_fail(KH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct kh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  kh::SpkDev spk{};            // host copy
  kh::SpkDev* d_spk = nullptr;  // device copy read by the kernels
  bool has_spk = false;
  double* d_coef = nullptr;
  std::vector<int64_t> seg_obj, seg_center;
  // Type 21 segments (device copies in d_t21), object ids in device order
  std::vector<int64_t> t21_obj;
  void* d_t21 = nullptr;
  void* dbuf = nullptr;
  size_t dbuf_size = 0;
  void* obuf = nullptr;  // queue-order scratch of the N-body launch
  size_t obuf_size = 0;
  void* noalloc_scratch = nullptr;  // this device's built-in order scratch (no-alloc entry)
  void* fbuf = nullptr;  // FOV working memory
  size_t fbuf_size = 0;
  void* hbuf = nullptr;  // FOV hit ordering
  size_t hbuf_size = 0;
  void* mbuf = nullptr;  // MOID working memory
  size_t mbuf_size = 0;
  void* lbuf = nullptr;  // propagate_n_body_long: planets at jd0
  size_t lbuf_size = 0;
  void* pbuf = nullptr;  // partials: states at the epochs
  size_t pbuf_size = 0;
  unsigned long long* d_queue = nullptr;
  unsigned long long* d_range = nullptr;  // epoch ranges of the last N-body batch [4]
  unsigned long long* d_counts = nullptr;  // FOV list counters [4]
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // FOV prefilter timing of the last kh_fov_state_check: event pairs around
  // each prefilter launch (the first FOV_TIMED_LAUNCHES of a call)
  std::vector<hipEvent_t> fov_ev;
  double fov_pf_ms = -1.0;
  int64_t fov_pf_launches = 0, fov_pf_pairs = 0;
  hipStream_t stream2 = nullptr;  // the N-body latency kernel runs here
  int64_t* d_split = nullptr;     // head of the N-body order it takes
  hipEvent_t fork = nullptr, join = nullptr;
  // launch shape chosen by kh_autotune_n_body, used by mode -1 calls
  struct {
    bool valid = false;
    int32_t mode = 0, block = 256, window = 16;
    int64_t latency_k = 0;
  } tune;
  bool timed = false;
  int32_t last_block = 0, last_grid = 0, last_mode = 0;
  int n_cu = 0;
};

extern "C" {
bool kh_build_chain(const kh_ctx* c, int64_t id, kh::Chain& ch);
int32_t kh_grow(void** buf, size_t* size, size_t bytes);
}

// Multi-epoch N-body propagation on the device (the FOV driver's states at
// its group epochs): every object from its jd0 through me.epochs in order,
// one integration without restarts; me.out holds the states in the input
// center and frame, NaN from a failing epoch on; out/status the final state.
int32_t kh_propagate_n_body_epochs_dev(kh_ctx* c, int64_t n, const double* jd0,
                                       const double* const xyz[6], int64_t center, int32_t frame,
                                       const kh::MultiEpochs& me, int32_t include_asteroids,
                                       double epsilon,
                                       double* const out[6], int32_t* status, int32_t sync);
