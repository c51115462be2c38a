// Exact discrete projection: the pressure-correction operator that the
// divergence and gradient line operators actually compose to.
//
// The reference corrects the velocity with the compact gradient of a pressure
// increment obtained from a 2nd-order Poisson operator
// (src/poisson_fishpack.f90:213-350, src/eq_momentum.f90 Correct_massflux),
// which differs from D*G of the 4th/6th-order schemes; the discrete divergence
// left after one projection then grows from sub-step to sub-step.  This
// solver inverts L = D_x G_x + D_y G_y + D_z G_z itself:
//  - x and z are periodic and uniform, so D G is circulant and the Fourier
//    transform diagonalises it: eigenvalue lx(kx), lz(kz) from the first
//    column of the composed line operators;
//  - y (walls) is a dense ny x ny matrix Ly = D_y G_y with the wall faces held
//    at their boundary value (the corrected wall-normal velocity is reset to
//    the wall value afterwards); it is diagonalised once on the host,
//    Ly = V diag(lam) V^-1 (its spectrum is real and V is well conditioned),
//    so each Fourier mode is solved by phi = V diag(1/(lam + lx + lz)) V^-1 r,
//    two dense ny x ny products over all modes.
// The constant vector spans the null space of Ly; the (kx, kz) = (0, 0)
// mode drops its component along it (the pressure gauge).
#pragma once

#include <string>
#include <vector>

#include "chs_geom.h"

namespace chs {

struct ExactOps {
  int nx = 0, ny = 0, nz = 0, nxh = 0;
  std::vector<double> lx;   // nxh: eigenvalue of D_x G_x on Fourier mode kx
  std::vector<double> lz;   // nz
  std::vector<double> Ly;   // ny * ny, row-major
  std::vector<double> lam;  // ny eigenvalues of Ly (ascending)
  std::vector<double> V;    // ny * ny, row-major, unit 2-norm columns
  std::vector<double> Vi;   // inverse of V
  int null_idx = -1;        // eigenvalue index of the constant mode
  double eig_resid = 0.0;   // max |Vi Ly V - diag(lam)| / max |Ly|
  // Mirror-symmetric channel (Ly commutes with the reflection j -> ny-1-j,
  // ny even): Ly splits into an even and an odd block of size ny/2 acting on
  // r_e(j) = r(j) + r(ny-1-j) and r_o(j) = r(j) - r(ny-1-j); each block is
  // diagonalised on its own (half the work per solve; the near-degenerate
  // even/odd pairs of wall modes fall into different blocks).
  bool sym = false;
  std::vector<double> Ve, Vie, lame, Vo, Vio, lamo;  // (ny/2)^2, (ny/2)
  int null_even = -1;       // constant mode in the even block
};

// A = V diag(lam) V^-1 for a real matrix with real simple eigenvalues
// (Hessenberg reduction, shifted QR for the eigenvalues, inverse iteration for
// the vectors, Gauss-Jordan inverse); fails when the spectrum is not real or
// the decomposition residual exceeds 1e-10.
std::string eig_real(int n, const std::vector<double>& A, std::vector<double>* lam, std::vector<double>* V,
                     std::vector<double>* Vi, double* resid);

// Boundary kinds [dir][u, v, w, p, T][side] of the channel set-up
// (src/bc_general.f90:137-245): periodic x and z; no-slip walls in y with
// Dirichlet velocity and temperature and Neumann pressure, or periodic y.
void flow_bc_kinds(bool wall_y, int ibc[3][5][2]);

// Operators of the channel's projection: divergence D1_P2C of (u, v, w) with
// the velocity boundary kinds, gradient D1_C2P of p with the pressure kinds
// (zero Neumann data at walls); ibc[dir][var][side] with var 0..3 = u, v, w, p.
std::string build_exact_ops(const Mesh& m, int acc, const int ibc[3][5][2], bool wall_y, ExactOps* out);

}  // namespace chs
