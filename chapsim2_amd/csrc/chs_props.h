// Property table handle shared by the table kernels and the MLP surrogate.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/chapsim_hip.h"

// column order of chs_props::col and of the device copy
enum { CHS_PC_T = 0, CHS_PC_D, CHS_PC_M, CHS_PC_K, CHS_PC_B, CHS_PC_CP, CHS_PC_H, CHS_PC_RH };

struct chs_mlp;

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
  chs_mlp* surrogate = nullptr;  // used by CHS_PROPS_MLP (not owned)
  size_t lds_block = 65536;      // LDS bytes a workgroup may use on the table's device
  int cus = 0;                   // compute units of that device
  bool h_rising = false;         // the h column strictly increases (bisections on h may be skipped)
  ~chs_props() {
    if (dev) (void)hipFree(dev);
    if (flag) (void)hipFree(flag);
    if (io) (void)hipFree(io);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace chs {
// Per-cell properties from rho*h for device arrays on stream s (no host
// synchronisation): the table path of chs_props_from_rhoh, or with the
// surrogate the same clamp and bisection on rho*h for the first h and the
// MLP for the properties from h (twice, as the table path).  Outputs may be
// null; the out-of-range count is added to *flag (device int).  With r32,
// both paths also store h, m, k, t rounded to FP32 into r32[0..3] (each
// possibly null)
std::string props_refresh_dh(const chs_props* p, int mode, const double* rhoh, long n, double* h, double* d,
                             double* m, double* k, double* t, int* flag, hipStream_t s,
                             float* const* r32 = nullptr);
std::string mlp_refresh_dh(chs_mlp* m, const chs_props* p, const double* rhoh, long n, double* h, double* d,
                           double* mu, double* k, double* t, int* flag, hipStream_t s,
                           float* const* r32 = nullptr);
}  // namespace chs
