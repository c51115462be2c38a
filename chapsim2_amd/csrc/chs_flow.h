// Device-resident flow solver of one channel / periodic box.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "chs_handle.h"
#include "chs_poisson.h"

namespace chs {

// block partials used by the max reductions (the partial buffer holds at
// least this many entries)
constexpr int kMaxPartials = 512;

// A device array with its Fortran extents.
template <typename T>
struct Field {
  T* p = nullptr;
  Ext3 e{{0, 0, 0}};
  long size() const { return (long)e.n[0] * e.n[1] * e.n[2]; }
};

class Flow {
 public:
  static std::string create(Handle* H, std::unique_ptr<Flow>* out);
  ~Flow();

  std::string set_state(const double* qx, const double* qy, const double* qz, const double* pr, int mem);
  std::string get_state(double* qx, double* qy, double* qz, double* pr, int mem);
  void update_re(int iter);
  std::string substep(int isub);
  std::string momentum_rhs(int isub);  // leaves dt-scaled rhs in acc_x of each component
  std::string divergence(double* div_dev);
  std::string mass_check(double out[3]);
  std::string copy_rhs(double* mx, double* my, double* mz, int mem);

  Handle* H = nullptr;
  int nx, ny, nz, px, py, pz;
  bool wall_y = true;
  double rre = 1.0, dt = 0.0, sigma2p = 1.0;
  double tg[4], tz[4], ta[4];
  int ibc[3][5][2];     // [dir][u,v,w,p,T][side]
  int mcov[3][3][2], mtau[3][3][2];
  double vol = 1.0;     // sum of the cell volumes (x-pencil qx cells)
  std::vector<double> dyw;  // per-j volume weight dy (size ny)

  // state
  Field<double> qx, qy, qz, pr, pcor, mx0, my0, mz0;
  // first-level intermediates
  Field<double> qxix, qxdx, qxdy, qxiy, qxiz, qxdz;
  Field<double> qydx, qyix, qydy, qyiy, qydz, qyiz;
  Field<double> qzdx, qzix, qziy, qzdy, qziz, qzdz, div;
  // accumulators per momentum component: [comp][x-part, y-part, z-part, pfc]
  Field<double> acc[3][4];
  // wall planes (nx,4,nz) and plane scratch
  double *fby_u = nullptr, *fby_v = nullptr, *fby_w = nullptr, *fby_p = nullptr, *fby_pn = nullptr;
  double *fby_conv = nullptr, *fby_visc = nullptr, *fby_div = nullptr;
  double *plane_a = nullptr, *plane_b = nullptr, *plane_c = nullptr, *plane_d = nullptr;
  double *partial = nullptr, *bulk = nullptr;
  long npartial = 0;  // entries of partial (at least kMaxPartials)
  double* scratch = nullptr;
  long scratch_n = 0;
  double* dyw_dev = nullptr;

 private:
  std::string op(int dir, int opk, const int ibc2[2], const Field<double>& A, Field<double>& out,
                 const double* fbc, double sA = 1.0, const Field<double>* B = nullptr, double cB = 1.0,
                 const Field<double>* Cm = nullptr, double sOut = 1.0, int accumulate = 0);
  std::string alloc(Field<double>& f, int e1, int e2, int e3);
  std::vector<void*> owned;
};

}  // namespace chs
