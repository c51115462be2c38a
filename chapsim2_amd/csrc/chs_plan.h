// Host-side planning of the compact line operators: scheme coefficient
// tables, pre-factored tridiagonal LHS, and the affine "recipes" of the four
// boundary rows of each line.  The HIP kernels consume the resulting LinePlan.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace chs {

// Boundary kinds, src/modules.f90:223-234.
enum Bc : int { BC_INTERIOR = 0, BC_PERIODIC = 1, BC_SYMMETRIC = 2, BC_ASYMMETRIC = 3,
                BC_DIRICHLET = 4, BC_NEUMANN = 5, BC_INTRPL = 6 };
// Accuracy, src/modules.f90:248-251.
enum Acc : int { CD2 = 1, CD4 = 2, CP4 = 3, CP6 = 4 };
// Operator kinds: the six 1-D families of src/basics_operations2.f90.
enum Op : int { MIDP_C2P = 0, MIDP_P2C = 1, D1_C2C = 2, D1_P2P = 3, D1_C2P = 4, D1_P2C = 5 };
// Coefficient families (table index).
enum Fam : int { F_D1C2C = 0, F_D1P2P = 1, F_D1C2P = 2, F_D1P2C = 3, F_M1C2P = 4, F_M1P2C = 5 };

Fam family_of(Op op);
bool is_derivative(Op op);
bool input_at_centre(Op op);
bool output_at_centre(Op op);

// Coefficient of the scheme tables (row 1..5, k 1..3 for LHS or 1..6 for RHS).
double coeff(Fam f, bool rhs, int row, int k, int bc, int acc);

// Grid facts a line plan needs (one direction).
struct LineGeom {
  int dir = 0;                 // 0 x, 1 y, 2 z
  double h = 1.0, h1r = 1.0;   // spacing and 1/spacing (computational for y)
  bool stretched = false;      // y only
  std::vector<double> yp, yc;  // y node / centre coordinates (y only)
  std::vector<double> ymp, ymc;// ds/dy at nodes / centres (y only)
  int nc = 0, np = 0;          // cells / points in this direction
};

constexpr int kMaxRowTerms = 12;
constexpr int kChunk = 16;      // rows per thread in the partitioned solve
constexpr int kChunkLong = 32;  // rows per thread for lines longer than kChunk * kMaxChunks
constexpr int kMaxChunks = 16;  // chunks (waves) per line

// One boundary row: rhs = (sum_t w[t] * f[idx[t]] + sum_m u[m] * fbc[m]) * dd
struct RowRecipe {
  int nterm = 0;
  int idx[kMaxRowTerms] = {0};
  double w[kMaxRowTerms] = {0};
  double u[4] = {0, 0, 0, 0};
};

struct LinePlan {
  int dir = 0;
  Op op = D1_C2P;
  int acc = CP6;
  int ibc[2] = {0, 0};         // after the reduction to IBC_INTRPL when fbc is absent
  bool has_fbc = false;
  int n_in = 0, n_out = 0;
  // bulk stencil: c1*(f[r+p1] + s*f[r+p2]) + c2*(f[r+p3] + s*f[r+p4]), rows 2..n_out-3
  double c1 = 0, c2 = 0, dd = 1.0;
  int p1 = 0, p2 = 0, p3 = 0, p4 = 0;
  double sgn = 1.0;
  RowRecipe rows[4];           // rows 0, 1, n_out-2, n_out-1
  // solve: 0 explicit, 1 Thomas, 2 cyclic Thomas (Sherman-Morrison)
  int solve = 0;
  double b0 = 1.0;
  std::vector<double> a, c, d; // pre-factored (src/basics_algorithms.f90:14-34)
  std::vector<double> pw;      // cyclic: weights giving the basic solution's first entry
  std::vector<double> x1;      // cyclic: RHS-independent correction vector
  double cyc_an = 0, cyc_cn = 0, cyc_den = 1.0;
  // Partitioned solve used by the GPU kernels: the line is split into P chunks
  // of at most kChunk rows; each chunk is solved locally (ld/lc local Thomas
  // factors) and corrected with precomputed spike vectors sv (coupling to the
  // previous chunk's last unknown) and sw (to the next chunk's first unknown),
  // whose values follow from the 2P x 2P reduced system (inverse rinv).
  // Chunk q holds rows [q*n_out/P, (q+1)*n_out/P) with P = ceil(n_out/part_ch),
  // part_ch = kChunk up to kChunk * kMaxChunks rows and kChunkLong beyond;
  // the GPU kernels use this form when P <= kMaxChunks.
  int part_P = 1;
  int part_ch = kChunk;
  std::vector<int> part_start;     // P+1 chunk boundaries
  std::vector<double> a_raw, b_raw, c_raw;
  std::vector<double> ld, lc, sv, sw, rinv;
  // post-processing (y derivatives): metric factor per output row, Neumann overwrite
  std::vector<double> metric;
  bool neu_lo = false, neu_hi = false;
};

// Build a plan; returns empty string on success or an error message.
// chunk_rows: rows per chunk of the partitioned evaluation (kChunk or
// kChunkLong); 0 takes kChunk up to kChunk * kMaxChunks rows, else kChunkLong.
std::string make_line_plan(const LineGeom& g, Op op, int acc, const int ibc0[2], bool has_fbc,
                           int n_in, int n_out, LinePlan* out, int chunk_rows = 0);

// Evaluate the explicit RHS of one line on the host (reference row formulas,
// src/basics_operations2.f90:1980-2719), before the 1/h scaling.
void line_rhs_host(const LineGeom& g, Op op, int acc, const int ibc[2], const double* fi, int n_in,
                   double* fo, int n_out, const double* fbc);

// Apply a plan to one line on the host, with the same arithmetic sequence as
// the GPU kernel (bulk stencil, boundary recipes, forward/backward sweeps).
void apply_plan_host(const LinePlan& p, const double* f, const double* fbc, double* out);

// Pre-factored LHS of one family for a line of n rows (Buildup_TDMA_LHS_array).
void build_lhs(Fam f, int n, int bc1, int bc2, int acc, std::vector<double>& a,
               std::vector<double>& b, std::vector<double>& c, std::vector<double>& d);

}  // namespace chs
