// Host-side objects of the learned warm start: relaxation sample buffers and
// the network (device parameters, optimiser state, BF16 inference fragments).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ecogen_amd.h"
#include "ecg_mlp.hpp"

struct ecg_samples {
  float* buf = nullptr;                 // cap * SAMPLE_FLOATS on the device
  unsigned long long* seen = nullptr;   // device counter of candidate cells
  long long cap = 0;
  int stride = 1;                       // of the current attachment
  long long base = 0;                   // records held before the current attachment
  long long uploaded = -1;              // record count set by ecg_samples_upload (-1: from the counter)
};

struct ecg_mlp {
  float* prm = nullptr;       // FP32 master parameters (MLP_NPARAM)
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grad = nullptr;      // reduced gradient
  float* gpart = nullptr;     // per-block gradient partials
  float* lpart = nullptr;     // per-block loss partials
  double* loss_acc = nullptr; // device loss accumulator
  int gpart_blocks = 0;
  uint2* frag = nullptr;      // BF16 A fragments (MLP_NFRAG * 64)
  float* bias = nullptr;      // FP32 biases (MLP_NBIAS)
  long long adam_t = 0;
  hipStream_t stream = nullptr;
};

namespace ecg {
int api_fail(int code, const std::string& msg);
// Number of records a sample buffer holds (synchronises the device).
long long samples_count(ecg_samples* s);
}  // namespace ecg
