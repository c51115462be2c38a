// Property table handle shared by the table kernels and the MLP surrogate.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/chapsim_hip.h"

// column order of chs_props::col and of the device copy
enum { CHS_PC_T = 0, CHS_PC_D, CHS_PC_M, CHS_PC_K, CHS_PC_B, CHS_PC_CP, CHS_PC_H, CHS_PC_RH };

struct chs_props {
  int n = 0;
  double dhmin = 0, dhmax = 0;
  double ref[7] = {0, 0, 0, 0, 0, 0, 0};
  std::vector<double> col[8];  // t, d, m, k, b, cp, h, rhoh (non-dimensional)
  double* dev = nullptr;       // 8 columns of n, same order
  int* flag = nullptr;         // out-of-range counter
  void* io = nullptr;          // host-memory staging
  size_t io_bytes = 0;
  hipStream_t stream = nullptr;
  ~chs_props() {
    if (dev) (void)hipFree(dev);
    if (flag) (void)hipFree(flag);
    if (io) (void)hipFree(io);
    if (stream) (void)hipStreamDestroy(stream);
  }
};
