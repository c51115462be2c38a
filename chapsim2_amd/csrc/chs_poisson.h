// Pressure Poisson solver: real FFT in x, complex FFT in z, tridiagonal
// solve in y per (kx, kz) mode, and the inverse transforms.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "chs_exact.h"
#include "chs_geom.h"

namespace chs {

class Poisson {
 public:
  // kind: 2 = FFT(x,z)+TDMA(y) with the 2nd-order y operator of
  // src/poisson_fishpack.f90:213-350 (y walls or periodic y is rejected).
  // kind 3 = 3-D FFT path with the modified wavenumbers of the iacc
  // staggered scheme: triply periodic, poisson_000
  // (src/poisson_1stderivcomp.f90:1016-1195); y walls on a uniform mesh,
  // poisson_010 (:1457-1816, istret = 0; a cosine transform along y).
  static std::string create(const Mesh& m, int kind, int iacc, std::unique_ptr<Poisson>* out);
  ~Poisson();
  // In-place solve on a device array of (nx, ny, nz), Fortran order.
  // `in_scale` multiplies the right-hand side before the forward transform.
  std::string solve(double* rhs, hipStream_t s, double in_scale = 1.0);
  std::string solve(float* rhs, hipStream_t s, float in_scale = 1.0f);
  // Exact discrete projection operator (see chs_exact.h): upload the Fourier
  // eigenvalues and the eigen-decomposition of the wall-normal block; then
  // solve_exact inverts D G with FFT(x,z) and two dense products along y.
  std::string setup_exact(const ExactOps& o);
  std::string solve_exact(double* rhs, hipStream_t s, double in_scale = 1.0);
  std::string solve_exact(float* rhs, hipStream_t s, float in_scale = 1.0f);
  bool has_exact() const { return V64 != nullptr && !dct_y; }
  int nx = 0, ny = 0, nz = 0, nxh = 0;
  int kind = 2;
  // kind 3 with y walls: the exact-solve tables hold the cosine transform
  bool dct_y = false;

 private:
  static std::string create_3d(const Mesh& m, int iacc, std::unique_ptr<Poisson>* out);
  static std::string create_010(const Mesh& m, int iacc, std::unique_ptr<Poisson>* out);
  // times the launch-shape variants of this solver's FP64 transforms once
  // per shape (the choices are cached for the process)
  std::string tune(bool ydir);
  template <typename T> std::string solve_t(T* rhs, hipStream_t s, T in_scale);
  template <typename T> std::string solve_3d_t(T* rhs, hipStream_t s, T in_scale);
  std::vector<int> rad_x, rad_y, rad_z;  // radix plans of nx/2, ny and nz
  double* tw_y64 = nullptr; float* tw_y32 = nullptr;   // twiddles of length ny (3-D path)
  double* ly64 = nullptr; float* ly32 = nullptr;       // 3-D path: ky'^2 (lx/lz hold kx'^2, kz'^2)
  // device tables
  double* tw_x64 = nullptr; float* tw_x32 = nullptr;   // twiddles of length nx/2 (cos, sin pairs)
  double* tw_z64 = nullptr; float* tw_z32 = nullptr;   // twiddles of length nz
  double* rot64 = nullptr; float* rot32 = nullptr;     // exp(-2 pi i k / nx), k = 0..nx/2
  double* ya64 = nullptr; double* yc64 = nullptr;      // y tridiagonal off-diagonals (ny)
  float* ya32 = nullptr; float* yc32 = nullptr;
  double* fd64 = nullptr; double* fz64 = nullptr;      // per-mode factors (nxh, nz, ny)
  float* fd32 = nullptr; float* fz32 = nullptr;
  void* work = nullptr;                                // complex work (nxh, ny, nz)
  size_t work_bytes = 0;
  template <typename T> std::string solve_exact_t(T* rhs, hipStream_t s, T in_scale);
  // exact operator tables
  double *V64 = nullptr, *Vi64 = nullptr, *lam64 = nullptr, *lx64 = nullptr, *lz64 = nullptr;
  float *V32 = nullptr, *Vi32 = nullptr, *lam32 = nullptr, *lx32 = nullptr, *lz32 = nullptr;
  int null_idx = -1;
  // even / odd blocks of a mirror-symmetric channel (null when not symmetric)
  double *Ve64 = nullptr, *Vo64 = nullptr, *Vie64 = nullptr, *Vio64 = nullptr, *lame64 = nullptr, *lamo64 = nullptr;
  float *Ve32 = nullptr, *Vo32 = nullptr, *Vie32 = nullptr, *Vio32 = nullptr, *lame32 = nullptr, *lamo32 = nullptr;
  int null_even = -1;
  void* work2 = nullptr;                               // second complex work array
};

// Radix plan of n over {4, 2, 3, 5}; empty if n has another prime factor.
std::vector<int> radix_plan(int n);

}  // namespace chs
