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
};

int32_t launch_nbody(const SpkDev* spk, const NBodyLaunch& p, hipStream_t stream,
                     int32_t* used_block, int32_t* used_grid, int32_t* used_mode);

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
};

int32_t launch_two_body(const SpkDev* spk, const TwoBodyLaunch& p, hipStream_t stream);

int32_t launch_spk_states(const SpkDev* spk, Chain body, Chain center, int32_t frame, int64_t m,
                          const double* jd, double* const out[6], int32_t* status,
                          hipStream_t stream);

}  // namespace kh
