// Mesh of the solver: uniform x and z, uniform or tanh-stretched y.
// Follows src/geometry.f90:151-269 (Buildup_grid_mapping_1D_tanh) and
// :410-593 (Buildup_geometry_mesh_info).
#pragma once

#include <vector>

#include "chs_plan.h"

namespace chs {

struct Mesh {
  int nc[3] = {0, 0, 0};
  int np[3] = {0, 0, 0};
  bool periodic[3] = {true, true, true};
  bool stretched = false;
  int istret = 0;
  double rstret = 0.0;
  double lxx = 1, lyb = 0, lyt = 1, lzz = 1;
  double h[3] = {1, 1, 1}, h1r[3] = {1, 1, 1}, h2r[3] = {1, 1, 1};
  std::vector<double> yp, yc, ymp, ymc;  // y nodes (nc+1), centres (nc), ds/dy at both

  void build(const int nc_[3], bool periodic_y, double lxx_, double lyb_, double lyt_, double lzz_, int istret_,
             double rstret_);
  LineGeom line_geom(int dir) const;
};

}  // namespace chs
