// Host set-up of the exact discrete projection (operators and the eigen-
// decomposition of the wall-normal block).
#include "chs_exact.h"

#include <algorithm>
#include <cmath>

namespace chs {

namespace {

constexpr double kEps = 2.220446049250313e-16;

// Householder reduction to upper Hessenberg form (eigenvalues only).
void hessenberg(int n, std::vector<double>& H) {
  std::vector<double> v(n);
  for (int k = 0; k < n - 2; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += H[(size_t)i * n + k] * H[(size_t)i * n + k];
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (H[(size_t)(k + 1) * n + k] > 0) alpha = -alpha;
    double vn = 0.0;
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    for (int i = k + 1; i < n; ++i) v[i] = H[(size_t)i * n + k];
    v[k + 1] -= alpha;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double f = 2.0 / vn;
    for (int j = 0; j < n; ++j) {  // H = P H
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * H[(size_t)i * n + j];
      s *= f;
      for (int i = k + 1; i < n; ++i) H[(size_t)i * n + j] -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {  // H = H P
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += H[(size_t)i * n + j] * v[j];
      s *= f;
      for (int j = k + 1; j < n; ++j) H[(size_t)i * n + j] -= s * v[j];
    }
    for (int i = k + 2; i < n; ++i) H[(size_t)i * n + k] = 0.0;
  }
}

// Shifted QR (Wilkinson shift, Givens rotations) on the Hessenberg matrix.
std::string hqr_real(int n, std::vector<double> H, std::vector<double>* lam) {
  lam->assign(n, 0.0);
  auto h = [&](int i, int j) -> double& { return H[(size_t)i * n + j]; };
  double anorm = 0.0;
  for (double x : H) anorm = std::max(anorm, std::fabs(x));
  int hi = n - 1, iter = 0;
  std::vector<double> cs(n), sn(n);
  while (hi >= 0) {
    if (hi == 0) {
      (*lam)[0] = h(0, 0);
      break;
    }
    int l = hi;
    while (l > 0) {
      double s = std::fabs(h(l - 1, l - 1)) + std::fabs(h(l, l));
      if (s == 0.0) s = anorm;
      if (std::fabs(h(l, l - 1)) <= kEps * s) {
        h(l, l - 1) = 0.0;
        break;
      }
      --l;
    }
    if (l == hi) {
      (*lam)[hi] = h(hi, hi);
      --hi;
      iter = 0;
      continue;
    }
    const double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
    const double p = 0.5 * (a - d), disc = p * p + b * c;
    double mu = d;
    if (disc >= 0.0) {
      const double r = std::sqrt(disc);
      const double den = p + (p >= 0.0 ? r : -r);
      mu = (den != 0.0) ? d - b * c / den : d;
    }
    if (iter == 10 || iter == 20) mu = d + 0.75 * std::fabs(c);  // exceptional shifts
    if (++iter > 60) return "eigenvalues: QR iteration did not converge (complex pair?)";
    for (int k = l; k <= hi; ++k) h(k, k) -= mu;
    for (int k = l; k < hi; ++k) {
      const double x = h(k, k), y = h(k + 1, k);
      const double r = std::hypot(x, y);
      const double cc = (r == 0.0) ? 1.0 : x / r, ss = (r == 0.0) ? 0.0 : y / r;
      cs[k] = cc;
      sn[k] = ss;
      for (int j = k; j <= hi; ++j) {
        const double t1 = h(k, j), t2 = h(k + 1, j);
        h(k, j) = cc * t1 + ss * t2;
        h(k + 1, j) = -ss * t1 + cc * t2;
      }
    }
    for (int k = l; k < hi; ++k) {
      const double cc = cs[k], ss = sn[k];
      const int ie = std::min(k + 2, hi);
      for (int i = l; i <= ie; ++i) {
        const double t1 = h(i, k), t2 = h(i, k + 1);
        h(i, k) = cc * t1 + ss * t2;
        h(i, k + 1) = -ss * t1 + cc * t2;
      }
    }
    for (int k = l; k <= hi; ++k) h(k, k) += mu;
  }
  return "";
}

// LU with partial pivoting in place; tiny pivots are replaced by tiny*norm.
void lu(int n, std::vector<double>& M, std::vector<int>& piv, double tiny) {
  piv.resize(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(M[(size_t)i * n + k]) > std::fabs(M[(size_t)p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(M[(size_t)p * n + j], M[(size_t)k * n + j]);
    if (std::fabs(M[(size_t)k * n + k]) < tiny) M[(size_t)k * n + k] = (M[(size_t)k * n + k] < 0 ? -tiny : tiny);
    const double inv = 1.0 / M[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = M[(size_t)i * n + k] * inv;
      M[(size_t)i * n + k] = f;
      if (f == 0.0) continue;
      for (int j = k + 1; j < n; ++j) M[(size_t)i * n + j] -= f * M[(size_t)k * n + j];
    }
  }
}

void lu_solve(int n, const std::vector<double>& M, const std::vector<int>& piv, std::vector<double>& x) {
  for (int k = 0; k < n; ++k) {
    if (piv[k] != k) std::swap(x[k], x[piv[k]]);
    for (int i = k + 1; i < n; ++i) x[i] -= M[(size_t)i * n + k] * x[k];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int j = i + 1; j < n; ++j) s -= M[(size_t)i * n + j] * x[j];
    x[i] = s / M[(size_t)i * n + i];
  }
}

std::string invert(int n, std::vector<double> A, std::vector<double>* inv) {
  std::vector<double>& I = *inv;
  I.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int p = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)p * n + col])) p = r;
    if (std::fabs(A[(size_t)p * n + col]) < 1e-300) return "eigenvectors: singular basis";
    if (p != col)
      for (int k = 0; k < n; ++k) {
        std::swap(A[(size_t)p * n + k], A[(size_t)col * n + k]);
        std::swap(I[(size_t)p * n + k], I[(size_t)col * n + k]);
      }
    const double f = 1.0 / A[(size_t)col * n + col];
    for (int k = 0; k < n; ++k) {
      A[(size_t)col * n + k] *= f;
      I[(size_t)col * n + k] *= f;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      const double g = A[(size_t)r * n + col];
      if (g == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        A[(size_t)r * n + k] -= g * A[(size_t)col * n + k];
        I[(size_t)r * n + k] -= g * I[(size_t)col * n + k];
      }
    }
  }
  return "";
}

}  // namespace

std::string eig_real(int n, const std::vector<double>& A, std::vector<double>* lam, std::vector<double>* V,
                     std::vector<double>* Vi, double* resid) {
  if (n < 1 || A.size() != (size_t)n * n) return "eigen: bad matrix";
  double anorm = 0.0;
  for (double x : A) anorm = std::max(anorm, std::fabs(x));
  if (anorm == 0.0) return "eigen: zero matrix";
  std::vector<double> H = A;
  hessenberg(n, H);
  std::vector<double> ev;
  std::string e = hqr_real(n, H, &ev);
  if (!e.empty()) return e;
  std::sort(ev.begin(), ev.end());
  // right eigenvectors by inverse iteration on A - sigma I
  V->assign((size_t)n * n, 0.0);
  std::vector<double> M((size_t)n * n), x(n);
  std::vector<int> piv;
  for (int q = 0; q < n; ++q) {
    const double sigma = ev[q] + 1e-13 * anorm;
    M = A;
    for (int i = 0; i < n; ++i) M[(size_t)i * n + i] -= sigma;
    lu(n, M, piv, kEps * anorm);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.25 * std::sin(1.0 + 0.7 * i);
    // eigenvalues already handled that (nearly) coincide with this one: the
    // iterate is kept orthogonal to their vectors, so a cluster (e.g. the
    // even/odd pair of wall modes of a symmetric channel) yields a basis of
    // its invariant subspace instead of the same vector twice
    std::vector<int> near;
    for (int p = 0; p < q; ++p)
      if (std::fabs(ev[p] - ev[q]) <= 1e-10 * anorm) near.push_back(p);
    for (int it = 0; it < 4; ++it) {
      lu_solve(n, M, piv, x);
      for (int pass = 0; pass < 2; ++pass)
        for (int p : near) {
          double dot = 0.0, nn = 0.0;
          for (int i = 0; i < n; ++i) {
            dot += (*V)[(size_t)i * n + p] * x[i];
            nn += (*V)[(size_t)i * n + p] * (*V)[(size_t)i * n + p];
          }
          for (int i = 0; i < n; ++i) x[i] -= dot / nn * (*V)[(size_t)i * n + p];
        }
      double nr = 0.0;
      for (double t : x) nr += t * t;
      nr = std::sqrt(nr);
      if (!(nr > 0.0) || !std::isfinite(nr)) return "eigen: inverse iteration failed";
      for (double& t : x) t /= nr;
    }
    // fix the sign: largest component positive
    int im = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(x[i]) > std::fabs(x[im])) im = i;
    const double sg = x[im] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) (*V)[(size_t)i * n + q] = sg * x[i];
  }
  // Refinement: with D = V^-1 A V, the first-order correction V <- V (I + E),
  // E_ij = D_ij / (lam_j - lam_i) (i != j, outside clusters), removes the
  // off-diagonal part quadratically.
  std::vector<double> AV((size_t)n * n), D((size_t)n * n), T((size_t)n * n);
  auto matmul = [&](const std::vector<double>& X, const std::vector<double>& Y, std::vector<double>& Z) {
    std::fill(Z.begin(), Z.end(), 0.0);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < n; ++k) {
        const double a = X[(size_t)i * n + k];
        if (a == 0.0) continue;
        for (int j = 0; j < n; ++j) Z[(size_t)i * n + j] += a * Y[(size_t)k * n + j];
      }
  };
  lam->assign(n, 0.0);
  double off = 0.0;
  for (int pass = 0; pass < 4; ++pass) {
    e = invert(n, *V, Vi);
    if (!e.empty()) return e;
    matmul(A, *V, AV);
    matmul(*Vi, AV, D);
    off = 0.0;
    for (int i = 0; i < n; ++i) {
      (*lam)[i] = D[(size_t)i * n + i];
      for (int j = 0; j < n; ++j)
        if (i != j) off = std::max(off, std::fabs(D[(size_t)i * n + j]));
    }
    if (off <= 1e-15 * anorm || pass == 3) break;
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        if (i == j) continue;
        const double gap = (*lam)[j] - (*lam)[i];
        if (std::fabs(gap) > 1e-10 * anorm) T[(size_t)i * n + j] = D[(size_t)i * n + j] / gap;
      }
    matmul(*V, T, AV);  // V E
    for (size_t t = 0; t < V->size(); ++t) (*V)[t] += AV[t];
    for (int j = 0; j < n; ++j) {
      double nr = 0.0;
      for (int i = 0; i < n; ++i) nr += (*V)[(size_t)i * n + j] * (*V)[(size_t)i * n + j];
      nr = std::sqrt(nr);
      for (int i = 0; i < n; ++i) (*V)[(size_t)i * n + j] /= nr;
    }
  }
  *resid = off / anorm;
  if (!(*resid <= 1e-10)) return "eigen: decomposition residual too large (non-real or clustered spectrum)";
  return "";
}

void flow_bc_kinds(bool wall_y, int ibc[3][5][2]) {
  for (int d = 0; d < 3; ++d)
    for (int v = 0; v < 5; ++v)
      for (int s = 0; s < 2; ++s) ibc[d][v][s] = BC_PERIODIC;
  if (wall_y)
    for (int s = 0; s < 2; ++s) {
      for (int v = 0; v < 3; ++v) ibc[1][v][s] = BC_DIRICHLET;
      ibc[1][4][s] = BC_DIRICHLET;
      ibc[1][3][s] = BC_NEUMANN;
    }
}

std::string build_exact_ops(const Mesh& m, int acc, const int ibc[3][5][2], bool wall_y, ExactOps* out) {
  ExactOps& o = *out;
  o.nx = m.nc[0]; o.ny = m.nc[1]; o.nz = m.nc[2];
  o.nxh = o.nx / 2 + 1;
  if (!m.periodic[0] || !m.periodic[2]) return "exact projection: x and z must be periodic";
  if (!wall_y) return "exact projection: needs walls in y";
  const double pi = 2.0 * std::asin(1.0);
  std::string e;
  // circulant x and z blocks
  for (int dir : {0, 2}) {
    const int n = m.nc[dir];
    const LineGeom g = m.line_geom(dir);
    LinePlan G, D;
    if (!(e = make_line_plan(g, D1_C2P, acc, ibc[dir][3], false, n, n, &G)).empty()) return e;
    if (!(e = make_line_plan(g, D1_P2C, acc, ibc[dir][dir], false, n, n, &D)).empty()) return e;
    std::vector<double> e0(n, 0.0), t(n), col(n);
    e0[0] = 1.0;
    apply_plan_host(G, e0.data(), nullptr, t.data());
    apply_plan_host(D, t.data(), nullptr, col.data());
    double cmax = 0.0, asym = 0.0;
    for (int j = 0; j < n; ++j) {
      cmax = std::max(cmax, std::fabs(col[j]));
      asym = std::max(asym, std::fabs(col[j] - col[(n - j) % n]));
    }
    if (asym > 1e-12 * cmax) return "exact projection: periodic D*G is not symmetric";
    const int nk = (dir == 0) ? o.nxh : n;
    std::vector<double>& lv = (dir == 0) ? o.lx : o.lz;
    lv.assign(nk, 0.0);
    for (int k = 0; k < nk; ++k) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += col[j] * std::cos(2.0 * pi * double((long)j * k % n) / double(n));
      lv[k] = s;
    }
  }
  // wall-normal block
  const int ny = o.ny;
  const LineGeom gy = m.line_geom(1);
  LinePlan G, D;
  if (!(e = make_line_plan(gy, D1_C2P, acc, ibc[1][3], true, ny, ny + 1, &G)).empty()) return e;
  if (!(e = make_line_plan(gy, D1_P2C, acc, ibc[1][1], false, ny + 1, ny, &D)).empty()) return e;
  o.Ly.assign((size_t)ny * ny, 0.0);
  std::vector<double> ek(ny), gf(ny + 1), col(ny);
  const double fbc0[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < ny; ++k) {
    std::fill(ek.begin(), ek.end(), 0.0);
    ek[k] = 1.0;
    apply_plan_host(G, ek.data(), fbc0, gf.data());
    gf[0] = 0.0;   // wall faces keep their boundary value
    gf[ny] = 0.0;
    apply_plan_host(D, gf.data(), nullptr, col.data());
    for (int i = 0; i < ny; ++i) o.Ly[(size_t)i * ny + k] = col[i];
  }
  // mirror symmetry
  double lmx = 0.0, asym = 0.0;
  for (int i = 0; i < ny; ++i)
    for (int k = 0; k < ny; ++k) {
      lmx = std::max(lmx, std::fabs(o.Ly[(size_t)i * ny + k]));
      asym = std::max(asym, std::fabs(o.Ly[(size_t)i * ny + k] - o.Ly[(size_t)(ny - 1 - i) * ny + (ny - 1 - k)]));
    }
  if (ny % 2 == 0 && asym <= 1e-12 * lmx) {
    const int h = ny / 2;
    std::vector<double> Le((size_t)h * h), Lo((size_t)h * h);
    for (int j = 0; j < h; ++j)
      for (int k = 0; k < h; ++k) {
        const double a = o.Ly[(size_t)j * ny + k], b = o.Ly[(size_t)j * ny + (ny - 1 - k)];
        Le[(size_t)j * h + k] = a + b;
        Lo[(size_t)j * h + k] = a - b;
      }
    double re = 0.0, ro = 0.0;
    if (!(e = eig_real(h, Le, &o.lame, &o.Ve, &o.Vie, &re)).empty()) return e;
    if (!(e = eig_real(h, Lo, &o.lamo, &o.Vo, &o.Vio, &ro)).empty()) return e;
    o.null_even = 0;
    for (int i = 1; i < h; ++i)
      if (std::fabs(o.lame[i]) < std::fabs(o.lame[o.null_even])) o.null_even = i;
    double lm = 0.0;
    for (double l : o.lame) lm = std::max(lm, std::fabs(l));
    for (double l : o.lamo) lm = std::max(lm, std::fabs(l));
    double omin = 1e300;
    for (double l : o.lamo) omin = std::min(omin, std::fabs(l));
    if (std::fabs(o.lame[o.null_even]) <= 1e-9 * lm && omin > 1e-9 * lm) o.sym = true;
  }
  if (!(e = eig_real(ny, o.Ly, &o.lam, &o.V, &o.Vi, &o.eig_resid)).empty()) return e;
  o.null_idx = 0;
  for (int i = 1; i < ny; ++i)
    if (std::fabs(o.lam[i]) < std::fabs(o.lam[o.null_idx])) o.null_idx = i;
  double lmax = 0.0;
  for (double l : o.lam) lmax = std::max(lmax, std::fabs(l));
  if (std::fabs(o.lam[o.null_idx]) > 1e-9 * lmax) return "exact projection: wall-normal operator has no null mode";
  return "";
}

}  // namespace chs
