#include "chs_geom.h"

#include <cmath>

namespace chs {

namespace {
// Two-sided / one-sided tanh map of the computational coordinate eta to y and
// its derivative ds/dy (src/geometry.f90:151-269).
void tanh_map(const Mesh& m, bool node, int n, std::vector<double>& y, std::vector<double>& mp) {
  y.assign(n, 0.0);
  mp.assign(n, 1.0);
  double shift = node ? 0.0 : 1.0 / double(n) * 0.5;
  double step = node ? 1.0 / double(n - 1) : 1.0 / double(n);
  double gamma = 1.0, delta = 0.0, ymin = 0.0, ymax = 0.0;
  switch (m.istret) {
    case 2: gamma = 0.5; delta = 0.5; ymin = -1.0; ymax = 1.0; break;
    case 3: gamma = -1.0; delta = 1.0; ymin = 0.0; ymax = 1.0; break;
    case 4: gamma = 1.0; delta = 0.0; ymin = 0.0; ymax = 1.0; break;
    default: break;
  }
  const double beta = m.rstret * 20.0;
  const double mm = std::tanh(beta * gamma);
  const double ff = (m.lyt - m.lyb) / (ymax - ymin);
  for (int j = 0; j < n; ++j) {
    double eta = (j == 0) ? shift : shift + double(j) * step;
    double s = std::tanh(beta * (eta - delta)) / mm;
    double q = 1.0 - s * s * mm * mm;
    q = (q < 1.0e-20 && q > -1.0e-20) ? 1.0 : mm / beta / q;
    y[j] = (s - ymin) * ff + m.lyb;
    mp[j] = q / ff;
  }
}
}  // namespace

void Mesh::build(const int nc_[3], bool periodic_y, double lxx_, double lyb_, double lyt_, double lzz_, int istret_,
                 double rstret_) {
  for (int d = 0; d < 3; ++d) nc[d] = nc_[d];
  periodic[0] = true;
  periodic[1] = periodic_y;
  periodic[2] = true;
  for (int d = 0; d < 3; ++d) np[d] = periodic[d] ? nc[d] : nc[d] + 1;
  lxx = lxx_; lyb = lyb_; lyt = lyt_; lzz = lzz_;
  istret = istret_; rstret = rstret_;
  stretched = (istret != 0);
  h[0] = lxx / double(nc[0]);
  h[2] = lzz / double(nc[2]);
  if (stretched) {
    h[1] = 1.0 / double(nc[1]);
    tanh_map(*this, true, nc[1] + 1, yp, ymp);
    tanh_map(*this, false, nc[1], yc, ymc);
  } else {
    h[1] = (lyt - lyb) / double(nc[1]);
    yp.resize(nc[1] + 1);
    yc.resize(nc[1]);
    for (int j = 0; j <= nc[1]; ++j) yp[j] = double(j) * h[1] + lyb;
    for (int j = 0; j < nc[1]; ++j) yc[j] = double(j) * h[1] + h[1] * 0.5 + lyb;
    ymp.assign(nc[1] + 1, 1.0);
    ymc.assign(nc[1], 1.0);
  }
  for (int d = 0; d < 3; ++d) {
    h2r[d] = 1.0 / (h[d] * h[d]);
    h1r[d] = 1.0 / h[d];
  }
}

LineGeom Mesh::line_geom(int dir) const {
  LineGeom g;
  g.dir = dir;
  g.h = h[dir];
  g.h1r = h1r[dir];
  g.nc = nc[dir];
  g.np = np[dir];
  if (dir == 1) {
    g.stretched = stretched;
    g.yp = yp;
    g.yc = yc;
    g.ymp = ymp;
    g.ymc = ymc;
  }
  return g;
}

}  // namespace chs
