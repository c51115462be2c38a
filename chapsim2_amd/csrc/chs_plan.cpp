// Host-side planning of compact line operators.
//
// Coefficient tables follow src/basics_operations2.f90:361-1509
// (Prepare_compact_coefficients, Gaitonde & Visbal 1998 tables 2.1-2.13),
// the tridiagonal LHS follows :1529-1588 and src/basics_algorithms.f90:14-34,
// and the boundary-row rules follow the branch logic of the six
// Prepare_TDMA_*_RHS_array routines (:1980-2719) with ghost cells built as in
// buildup_ghost_cells_C/P (:1803-1968).  The rules are re-expressed here as a
// small per-operator table of row modes (stencil / one-sided / assigned),
// evaluated on a ghost-extended line; the resulting affine row recipes are what
// the GPU kernels apply.
#include "chs_plan.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>

namespace chs {

namespace {

constexpr int NROW = 5, NBC = 7, NACC = 4;
constexpr double kBig = 1.0e20;

struct Tab {
  double f[6][NROW][3][NBC][NACC];
  double r[6][NROW][6][NBC][NACC];
};

Tab g_tab;
std::once_flag g_tab_once;

inline double& L(Fam f, int row, int k, int bc, int acc) { return g_tab.f[f][row - 1][k - 1][bc][acc - 1]; }
inline double& R(Fam f, int row, int k, int bc, int acc) { return g_tab.r[f][row - 1][k - 1][bc][acc - 1]; }

// Interior scheme of one family and accuracy: alpha, and the two bulk RHS
// weights as stored in the tables (already multiplied by 1/2, 1/4 or 1/3).
struct Interior { double alpha, r1, r2; };
// One-sided boundary scheme: LHS off-diagonal and up to six RHS weights.
struct OneSided { double alpha; double w[6]; };

Interior interior(Fam f, int acc) {
  switch (f) {
    case F_D1C2C:
    case F_D1P2P: {  // Table 2.1
      static const double al[5] = {0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 3.0};
      static const double a[5] = {0, 1.0, 4.0 / 3.0, 3.0 / 2.0, 14.0 / 9.0};
      static const double b[5] = {0, 0.0, -1.0 / 3.0, 0.0, 1.0 / 9.0};
      return {al[acc], a[acc] * 0.5, b[acc] * 0.25};
    }
    case F_D1C2P:
    case F_D1P2C: {  // Table 2.10
      double al, a, b;
      if (acc == CD2) { al = 0.0; a = 1.0; b = 0.0; }
      else if (acc == CD4 || acc == CP4) {
        al = (acc == CD4) ? 0.0 : 1.0 / 22.0;
        a = (9.0 - 6.0 * al) / 8.0;
        b = (-1.0 + 22.0 * al) / 8.0;
      } else { al = 9.0 / 62.0; a = 63.0 / 62.0; b = 17.0 / 62.0; }
      return {al, a, b * (1.0 / 3.0)};
    }
    default: {  // Table 2.7, interpolation
      double al, a, b;
      if (acc == CD2) { al = 0.0; a = 1.0; b = 0.0; }
      else if (acc == CD4 || acc == CP4) {
        al = (acc == CD4) ? 0.0 : 1.0 / 6.0;
        a = 9.0 / 8.0 + 5.0 / 4.0 * al;
        b = -1.0 / 8.0 + 3.0 / 4.0 * al;
      } else {
        al = 3.0 / 10.0;
        a = 5.0 * (15.0 + 14.0 * al) / 64.0;
        b = (-25.0 + 126.0 * al) / 128.0;
      }
      return {al, a * 0.5, b * 0.5};
    }
  }
}

// Row-1 one-sided closure (rows 1/5) of each family.
OneSided closure1(Fam f, int acc) {
  OneSided s{};
  switch (f) {
    case F_D1C2C:
    case F_D1P2P:  // Table 2.2
      if (acc == CD2) s = {0.0, {-3.0 / 2.0, 2.0, -1.0 / 2.0, 0, 0, 0}};
      else if (acc == CD4) s = {0.0, {-25.0 / 12.0, 4.0, -3.0, 4.0 / 3.0, -1.0 / 4.0, 0}};
      else if (acc == CP4) s = {3.0, {-17.0 / 6.0, 3.0 / 2.0, 3.0 / 2.0, -1.0 / 6.0, 0, 0}};
      else s = {5.0, {-197.0 / 60.0, -5.0 / 12.0, 5.0, -5.0 / 3.0, 5.0 / 12.0, -1.0 / 20.0}};
      break;
    case F_D1C2P: {  // Table 2.12
      if (acc == CD2) { s = {0.0, {-2.0 - 0.0, 3.0 + 0.0, -1.0, 0, 0, 0}}; }
      else if (acc == CD4 || acc == CP4) {
        double q = 0.0;
        s = {q, {(-93.0 - 22.0 * q) / 24.0, (229.0 + 17.0 * q) / 24.0, (-75.0 + 3.0 * q) / 8.0,
                 (111.0 - 5.0 * q) / 24.0, (-22.0 + 1.0 * q) / 24.0, 0}};
      } else {
        double q = 1689.0 / 71.0;
        s = {q, {(-3043.0 - 563.0 * q) / 640.0, (5353.0 + 201.0 * q) / 384.0, (-3489.0 + 143.0 * q) / 192.0,
                 (859.0 - 37.0 * q) / 64.0, (-2041.0 + 87.0 * q) / 384.0, (1689.0 - 71.0 * q) / 1920.0}};
      }
      break;
    }
    case F_D1P2C: {  // Table 2.11
      if (acc == CD2) { s = {0.0, {-1.0, 1.0, 0, 0, 0, 0}}; }
      else if (acc == CD4 || acc == CP4) {
        double q = 0.0;
        s = {q, {(-22.0 + q) / 24.0, (17.0 - 27.0 * q) / 24.0, (3.0 + 9.0 * q) / 8.0, (-5.0 - q) / 24.0,
                 1.0 / 24.0, 0}};
      } else {
        double q = 71.0 / 9.0;
        s = {q, {(-1689.0 + 71.0 * q) / 1920.0, (67.0 - 141.0 * q) / 128.0, (143.0 + 207.0 * q) / 192.0,
                 (-111.0 + 1.0 * q) / 192.0, (29.0 - 3.0 * q) / 128.0, (-71.0 + 9.0 * q) / 1920.0}};
      }
      break;
    }
    case F_M1C2P: {  // extrapolation, :1220-1255
      if (acc == CD2) { s = {0.0, {(3.0 + 0.0) / 2.0, (-1.0 + 0.0) / 2.0, 0, 0, 0, 0}}; }
      else if (acc == CD4 || acc == CP4) {
        double q = (acc == CD4) ? 0.0 : 5.0;
        s = {q, {(35.0 + 5.0 * q) / 16.0, (-35.0 + 15.0 * q) / 16.0, (21.0 - 5.0 * q) / 16.0, (-5.0 + q) / 16.0, 0,
                 0}};
      } else {
        double q = 9.0;
        s = {q, {(693.0 + 63.0 * q) / 256.0, (-1155.0 + 315.0 * q) / 256.0, (693.0 - 105.0 * q) / 128.0,
                 (-495.0 + 63.0 * q) / 128.0, (385.0 - 45.0 * q) / 256.0, (-63.0 + 7.0 * q) / 256.0}};
      }
      break;
    }
    case F_M1P2C: {  // Table 2.8
      if (acc == CD2) { s = {0.0, {0.5, 0.5, 0.0, 0, 0, 0}}; }
      else if (acc == CD4 || acc == CP4) {
        double q = (acc == CD4) ? 0.0 : 1.0;
        s = {q, {(5.0 - q) / 16.0, (15.0 + 9.0 * q) / 16.0, (-5.0 + 9.0 * q) / 16.0, (1.0 - q) / 16.0, 0, 0}};
      } else {
        double q = 7.0 / 3.0;
        s = {q, {(63.0 - 7.0 * q) / 256.0, (315.0 + 105.0 * q) / 256.0, (-7.0 + 7.0 * q) * 15.0 / 128.0,
                 (63.0 - 35.0 * q) / 128.0, (-15.0 + 7.0 * q) * 3.0 / 256.0, (7.0 - 3.0 * q) / 256.0}};
      }
      break;
    }
  }
  return s;
}

// Row-2 one-sided closure (rows 2/4); only C2C/P2P, C2P and interpolation C2P have one.
bool closure2(Fam f, int acc, OneSided* s) {
  switch (f) {
    case F_D1C2C:
    case F_D1P2P:  // Tables 2.3/2.6
      if (acc == CD2) *s = {0.0, {-1.0 / 2.0, 0.0, 1.0 / 2.0, 0, 0, 0}};
      else if (acc == CD4) *s = {0.0, {-1.0 / 4.0, -5.0 / 6.0, 3.0 / 2.0, -1.0 / 2.0, 1.0 / 12.0, 0}};
      else if (acc == CP4) *s = {1.0 / 4.0, {-3.0 / 4.0, 0.0, 3.0 / 4.0, 0, 0, 0}};
      else *s = {2.0 / 11.0, {-20.0 / 33.0, -35.0 / 132.0, 34.0 / 33.0, -7.0 / 33.0, 2.0 / 33.0, -1.0 / 132.0}};
      return true;
    case F_D1C2P:  // Table 2.13
      if (acc == CD2) *s = {0.0, {-1.0, 1.0, 0, 0, 0, 0}};
      else if (acc == CD4 || acc == CP4) {
        double q = (acc == CD4) ? 0.0 : 1.0 / 22.0;
        *s = {q, {(-11.0 - 46.0 * q) / 12.0, (17.0 + 202.0 * q) / 24.0, (3.0 - 66.0 * q) / 8.0,
                  (-5.0 + 110.0 * q) / 24.0, (1.0 - 22.0 * q) / 24.0, 0}};
      } else {
        *s = {31.0 / 818.0, {-5195.0 / 4908.0, 4957.0 / 4908.0, 119.0 / 1227.0, -85.0 / 1227.0, 119.0 / 4908.0,
                             -17.0 / 4908.0}};
      }
      return true;
    case F_M1C2P: {
      if (acc == CD2) *s = {0.0, {(1.0 + 2.0 * 0.0) / 2.0, (1.0 + 2.0 * 0.0) / 2.0, 0, 0, 0, 0}};
      else if (acc == CD4 || acc == CP4) {
        double q = (acc == CD4) ? 0.0 : 1.0 / 6.0;
        *s = {q, {(5.0 + 34.0 * q) / 16.0, (15.0 - 26.0 * q) / 16.0, (-5.0 + 30.0 * q) / 16.0,
                  (1.0 - 6.0 * q) / 16.0, 0, 0}};
      } else {
        double q = 9.0;
        *s = {q, {(63.0 + 686.0 * q) / 256.0, (315.0 - 1050.0 * q) / 256.0, (-105.0 + 798.0 * q) / 128.0,
                  (63.0 - 530.0 * q) / 128.0, (-45.0 + 406.0 * q) / 256.0, (7.0 - 66.0 * q) / 256.0}};
      }
      return true;
    }
    default:
      return false;
  }
}

void set_row(Fam f, int row, int bc, int acc, double l1, double l2, double l3) {
  L(f, row, 1, bc, acc) = l1; L(f, row, 2, bc, acc) = l2; L(f, row, 3, bc, acc) = l3;
}
void copy_row(Fam f, int row, int bcd, int accd, int rows, int bcs, int accs, bool lhs, bool rhs) {
  if (lhs) for (int k = 1; k <= 3; ++k) L(f, row, k, bcd, accd) = L(f, rows, k, bcs, accs);
  if (rhs) for (int k = 1; k <= 6; ++k) R(f, row, k, bcd, accd) = R(f, rows, k, bcs, accs);
}

void build_tables() {
  for (auto& x : g_tab.f) for (auto& y : x) for (auto& z : y) for (auto& w : z) for (auto& v : w) v = kBig;
  for (auto& x : g_tab.r) for (auto& y : x) for (auto& z : y) for (auto& w : z) for (auto& v : w) v = kBig;
  const bool deriv[6] = {true, true, true, true, false, false};
  for (int fi = 0; fi < 6; ++fi) {
    Fam f = Fam(fi);
    for (int acc = 1; acc <= NACC; ++acc) {
      Interior in = interior(f, acc);
      // periodic: identical interior rows 1..5
      for (int row = 1; row <= 5; ++row) {
        set_row(f, row, BC_PERIODIC, acc, in.alpha, 1.0, in.alpha);
        R(f, row, 1, BC_PERIODIC, acc) = in.r1;
        R(f, row, 2, BC_PERIODIC, acc) = in.r2;
        if (f == F_D1C2P || f == F_D1P2C || f == F_M1C2P || f == F_M1P2C) R(f, row, 3, BC_PERIODIC, acc) = 0.0;
      }
      // symmetric / antisymmetric reflections: edge rows differ, rest periodic
      double sym_b = 1.0, sym_c = 0.0, asy_b = 1.0, asy_c = 0.0;
      switch (f) {
        case F_D1C2C: sym_b = 1.0 - in.alpha; sym_c = in.alpha; asy_b = 1.0 + in.alpha; asy_c = in.alpha; break;
        case F_D1P2P: {
          double a6 = interior(F_D1C2C, acc).alpha;
          sym_b = 1.0; sym_c = 0.0; asy_b = 1.0; asy_c = 2.0 * a6;
          break;
        }
        case F_D1C2P: sym_b = 1.0; sym_c = 0.0; asy_b = 1.0; asy_c = 2.0 * in.alpha; break;
        case F_D1P2C: sym_b = 1.0 - in.alpha; sym_c = in.alpha; asy_b = 1.0 + in.alpha; asy_c = in.alpha; break;
        case F_M1C2P: sym_b = 1.0; sym_c = in.alpha + in.alpha; asy_b = 1.0; asy_c = 0.0; break;
        case F_M1P2C: sym_b = 1.0 + in.alpha; sym_c = in.alpha; asy_b = 1.0 - in.alpha; asy_c = in.alpha; break;
      }
      for (int bc : {BC_SYMMETRIC, BC_ASYMMETRIC}) {
        double bb = (bc == BC_SYMMETRIC) ? sym_b : asy_b, cc = (bc == BC_SYMMETRIC) ? sym_c : asy_c;
        set_row(f, 1, bc, acc, 0.0, bb, cc);
        set_row(f, 5, bc, acc, cc, bb, 0.0);
        for (int row = 2; row <= 4; ++row) copy_row(f, row, bc, acc, row, BC_PERIODIC, acc, true, false);
        for (int row = 1; row <= 5; ++row) copy_row(f, row, bc, acc, row, BC_PERIODIC, acc, false, true);
      }
    }
    // interior: periodic, with the edge rows of compact schemes from CD4
    for (int acc = 1; acc <= NACC; ++acc)
      for (int row = 1; row <= 5; ++row) {
        int src = ((acc == CP4 || acc == CP6) && (row == 1 || row == 5)) ? CD4 : acc;
        copy_row(f, row, BC_INTERIOR, acc, row, BC_PERIODIC, src, true, true);
      }
    // one-sided closures (IBC_INTRPL)
    for (int acc = 1; acc <= NACC; ++acc) {
      OneSided s1 = closure1(f, acc);
      double mir = deriv[fi] ? -1.0 : 1.0;
      set_row(f, 1, BC_INTRPL, acc, 0.0, 1.0, s1.alpha);
      set_row(f, 5, BC_INTRPL, acc, s1.alpha, 1.0, 0.0);
      for (int k = 1; k <= 6; ++k) {
        R(f, 1, k, BC_INTRPL, acc) = s1.w[k - 1];
        R(f, 5, k, BC_INTRPL, acc) = mir * s1.w[k - 1];
      }
      OneSided s2;
      if (closure2(f, acc, &s2)) {
        set_row(f, 2, BC_INTRPL, acc, s2.alpha, 1.0, s2.alpha);
        set_row(f, 4, BC_INTRPL, acc, s2.alpha, 1.0, s2.alpha);
        for (int k = 1; k <= 6; ++k) {
          R(f, 2, k, BC_INTRPL, acc) = s2.w[k - 1];
          R(f, 4, k, BC_INTRPL, acc) = mir * s2.w[k - 1];
        }
        copy_row(f, 3, BC_INTRPL, acc, 3, BC_PERIODIC, acc, true, true);
      } else {
        for (int row = 2; row <= 4; ++row) copy_row(f, row, BC_INTRPL, acc, row, BC_PERIODIC, acc, true, true);
      }
    }
    // Dirichlet / Neumann through ghost cells (bc_ghost_cd = .true.)
    for (int acc = 1; acc <= NACC; ++acc) {
      for (int row = 1; row <= 5; ++row) {
        copy_row(f, row, BC_DIRICHLET, acc, row, BC_INTERIOR, acc, true, true);
        copy_row(f, row, BC_NEUMANN, acc, row, BC_INTERIOR, acc, true, true);
      }
      // rows that are assigned directly from the boundary value
      int assigned_bc = -1;
      if (f == F_D1P2P || f == F_D1C2P) assigned_bc = BC_NEUMANN;
      if (f == F_M1C2P) assigned_bc = BC_DIRICHLET;
      if (assigned_bc >= 0)
        for (int row : {1, 5}) {
          set_row(f, row, assigned_bc, acc, 0.0, 1.0, 0.0);
          for (int k = 1; k <= 6; ++k) R(f, row, k, assigned_bc, acc) = 0.0;
        }
    }
  }
}

const Tab& tab() {
  std::call_once(g_tab_once, build_tables);
  return g_tab;
}

// ---------------------------------------------------------------------------
// Row modes of the boundary rows of each operator.
// ---------------------------------------------------------------------------
enum RowMode { RM_STENCIL, RM_ONESIDED, RM_ASSIGN };
struct RowRule {
  RowMode mode;
  int coef_row;  // table row (1..5)
  int coef_bc;   // table bc
  int start;     // one-sided: first input index (1-based)
  int step;      // one-sided: +1 / -1
  int side;      // assign: fbc side (0/1)
  bool div_dd;   // assign: value is fbc/dd
};

inline bool in(int bc, std::initializer_list<int> s) {
  for (int v : s) if (bc == v) return true;
  return false;
}

// rows[0..3] = output rows 1, 2, n-1, n (1-based), following the is_bc* logic
// of the reference routines (src/basics_operations2.f90:1980-2719).
void row_rules(Op op, const int ibc[2], int n, RowRule rows[4]) {
  const int b0 = ibc[0], b1 = ibc[1];
  auto S = [](int r, int bc) { return RowRule{RM_STENCIL, r, bc, 0, 0, 0, false}; };
  auto I = [](int r, int start, int step) { return RowRule{RM_ONESIDED, r, BC_INTRPL, start, step, 0, false}; };
  auto A = [](int side, bool dd) { return RowRule{RM_ASSIGN, 0, 0, 0, 0, side, dd}; };
  const auto REF4 = {BC_INTERIOR, BC_SYMMETRIC, BC_ASYMMETRIC};
  (void)REF4;
  switch (op) {
    case MIDP_P2C: {
      bool bc1 = in(b0, {BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET, BC_NEUMANN});
      bool bc5 = in(b1, {BC_INTERIOR, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET, BC_NEUMANN});
      rows[0] = bc1 ? S(1, b0) : I(1, 1, +1);
      rows[1] = S(3, BC_PERIODIC);
      rows[2] = (b1 == BC_PERIODIC) ? S(4, BC_PERIODIC) : S(4, b1);
      rows[3] = bc5 ? S(5, b1) : (b1 == BC_PERIODIC ? S(5, BC_PERIODIC) : I(5, n + 1, -1));
      break;
    }
    case MIDP_C2P: {
      bool bc1 = in(b0, {BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_NEUMANN});
      bool bc5 = in(b1, {BC_INTERIOR, BC_SYMMETRIC, BC_ASYMMETRIC, BC_NEUMANN});
      bool bc2 = bc1 || b0 == BC_DIRICHLET;
      bool bc4 = bc5 || b1 == BC_DIRICHLET;
      rows[0] = bc1 ? S(1, b0) : (b0 == BC_DIRICHLET ? A(0, false) : I(1, 1, +1));
      rows[1] = bc2 ? S(2, b0) : I(2, 1, +1);
      rows[2] = bc4 ? S(4, b1) : (b1 == BC_PERIODIC ? S(4, BC_PERIODIC) : I(4, n - 1, -1));
      rows[3] = bc5 ? S(5, b1)
                    : (b1 == BC_PERIODIC ? S(5, BC_PERIODIC) : (b1 == BC_DIRICHLET ? A(1, false) : I(5, n - 1, -1)));
      break;
    }
    case D1_C2C: {
      auto isbc = [](int b) {
        return in(b, {BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET, BC_NEUMANN});
      };
      rows[0] = isbc(b0) ? S(1, b0) : I(1, 1, +1);
      rows[1] = isbc(b0) ? S(2, b0) : I(2, 1, +1);
      rows[2] = isbc(b1) ? S(4, b1) : I(4, n, -1);
      rows[3] = isbc(b1) ? S(5, b1) : I(5, n, -1);
      break;
    }
    case D1_P2P: {
      auto is1 = [](int b) { return in(b, {BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET}); };
      auto is2 = [&](int b) { return is1(b) || b == BC_NEUMANN; };
      rows[0] = is1(b0) ? S(1, b0) : (b0 == BC_NEUMANN ? A(0, true) : I(1, 1, +1));
      rows[1] = is2(b0) ? S(2, b0) : I(2, 1, +1);
      rows[2] = is2(b1) ? S(4, b1) : I(4, n, -1);
      rows[3] = is1(b1) ? S(5, b1) : (b1 == BC_NEUMANN ? A(1, true) : I(5, n, -1));
      break;
    }
    case D1_C2P: {
      bool bc1 = in(b0, {BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET});
      bool bc2 = bc1 || in(b0, {BC_DIRICHLET, BC_NEUMANN});
      bool bc5 = in(b1, {BC_INTERIOR, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET});
      // The reference decides the n-1 row from the first side's kind here
      // (src/basics_operations2.f90:2548-2552); kept as is.
      bool bc4 = bc5 || in(b0, {BC_DIRICHLET, BC_NEUMANN});
      rows[0] = bc1 ? S(1, b0) : (b0 == BC_NEUMANN ? A(0, true) : I(1, 1, +1));
      rows[1] = bc2 ? S(2, b0) : I(2, 1, +1);
      rows[2] = bc4 ? S(4, b1) : (b1 == BC_PERIODIC ? S(4, BC_PERIODIC) : I(4, n - 1, -1));
      rows[3] = bc5 ? S(5, b1)
                    : (b1 == BC_PERIODIC ? S(5, BC_PERIODIC) : (b1 == BC_NEUMANN ? A(1, true) : I(5, n - 1, -1)));
      break;
    }
    case D1_P2C: {
      bool bc1 = in(b0, {BC_INTERIOR, BC_PERIODIC, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET, BC_NEUMANN});
      bool bc5 = in(b1, {BC_INTERIOR, BC_SYMMETRIC, BC_ASYMMETRIC, BC_DIRICHLET, BC_NEUMANN});
      rows[0] = bc1 ? S(1, b0) : I(1, 1, +1);
      rows[1] = S(3, BC_PERIODIC);
      rows[2] = S(4, BC_PERIODIC);
      rows[3] = bc5 ? S(5, b1) : (b1 == BC_PERIODIC ? S(5, BC_PERIODIC) : I(5, n + 1, -1));
      break;
    }
  }
}

// bulk stencil offsets (1-based input index relative to output row i) and sign
void bulk_offsets(Op op, int* p1, int* p2, int* p3, int* p4, double* sgn) {
  switch (op) {
    case MIDP_C2P: *p1 = 0; *p2 = -1; *p3 = 1; *p4 = -2; *sgn = 1.0; break;
    case MIDP_P2C: *p1 = 0; *p2 = 1; *p3 = -1; *p4 = 2; *sgn = 1.0; break;
    case D1_C2C:
    case D1_P2P: *p1 = 1; *p2 = -1; *p3 = 2; *p4 = -2; *sgn = -1.0; break;
    case D1_C2P: *p1 = 0; *p2 = -1; *p3 = 1; *p4 = -2; *sgn = -1.0; break;
    case D1_P2C: *p1 = 1; *p2 = 0; *p3 = 2; *p4 = -1; *sgn = -1.0; break;
  }
}

// ghost values g[0..3] = f(-1), f(0), f(n+1), f(n+2) (1-based), from
// buildup_ghost_cells_C / _P.
void ghosts(bool centre_input, const double* fi, int n, const int ibc[2], const double* fbc,
            const double dp[4], double g[4]) {
  auto F = [&](int k) { return fi[k - 1]; };
  // start side
  switch (ibc[0]) {
    case BC_INTERIOR: g[1] = fbc[0]; g[0] = fbc[2]; break;
    case BC_PERIODIC: g[1] = F(n); g[0] = F(n - 1); break;
    case BC_SYMMETRIC: g[1] = centre_input ? F(1) : F(2); g[0] = centre_input ? F(2) : F(3); break;
    case BC_ASYMMETRIC: g[1] = centre_input ? -F(1) : -F(2); g[0] = centre_input ? -F(2) : -F(3); break;
    case BC_DIRICHLET:
      if (centre_input) { g[1] = 2.0 * fbc[0] - F(1); g[0] = 2.0 * fbc[0] - F(2); }
      else if (fbc) { g[1] = 2.0 * fbc[0] - F(2); g[0] = 2.0 * fbc[0] - F(3); }
      else { g[1] = 2.0 * F(1) - F(2); g[0] = 2.0 * F(1) - F(3); }
      break;
    case BC_NEUMANN:
      if (centre_input) { g[1] = F(1) - fbc[0] * dp[0]; g[0] = F(2) - fbc[0] * dp[2]; }
      else { g[1] = F(2) - fbc[0] * dp[0]; g[0] = F(3) - fbc[0] * dp[2]; }
      break;
    default: g[1] = kBig; g[0] = kBig;
  }
  // end side
  switch (ibc[1]) {
    case BC_INTERIOR: g[2] = fbc[1]; g[3] = fbc[3]; break;
    case BC_PERIODIC: g[2] = F(1); g[3] = F(2); break;
    case BC_SYMMETRIC: g[2] = centre_input ? F(n) : F(n - 1); g[3] = centre_input ? F(n - 1) : F(n - 2); break;
    case BC_ASYMMETRIC: g[2] = centre_input ? -F(n) : -F(n - 1); g[3] = centre_input ? -F(n - 1) : -F(n - 2); break;
    case BC_DIRICHLET:
      if (centre_input) { g[2] = 2.0 * fbc[1] - F(n); g[3] = 2.0 * fbc[1] - F(n - 1); }
      else if (fbc) { g[2] = 2.0 * fbc[1] - F(n - 1); g[3] = 2.0 * fbc[1] - F(n - 2); }
      else { g[2] = 2.0 * F(n) - F(n - 1); g[3] = 2.0 * F(n) - F(n - 2); }
      break;
    case BC_NEUMANN:
      if (centre_input) { g[2] = F(n) + fbc[1] * dp[1]; g[3] = F(n - 1) + fbc[1] * dp[3]; }
      else { g[2] = F(n - 1) + fbc[1] * dp[1]; g[3] = F(n - 2) + fbc[1] * dp[3]; }
      break;
    default: g[2] = kBig; g[3] = kBig;
  }
}

void ghost_spacing(const LineGeom& g, Op op, double dp[4]) {
  bool cin = input_at_centre(op);
  if (g.dir == 1) {
    const int nc = g.nc, np = g.np;
    const auto& yp = g.yp;
    const auto& yc = g.yc;
    if (cin) {
      dp[0] = (yc[0] - yp[0]) * 2.0;
      dp[2] = (yc[1] - yp[0]) * 2.0;
      dp[1] = (yp[np - 1] - yc[nc - 1]) * 2.0;
      dp[3] = (yp[np - 1] - yc[nc - 2]) * 2.0;
    } else {
      dp[0] = (yp[1] - yp[0]) * 2.0;
      dp[2] = (yp[2] - yp[0]) * 2.0;
      dp[1] = (yp[np - 1] - yp[np - 2]) * 2.0;
      dp[3] = (yp[np - 1] - yp[np - 3]) * 2.0;
    }
  } else {
    double h = g.h;
    if (cin) { dp[0] = h; dp[2] = h * 2.0; dp[1] = h; dp[3] = h * 2.0; }
    else { dp[0] = h * 2.0; dp[2] = h * 4.0; dp[1] = h * 2.0; dp[3] = h * 4.0; }
  }
}

}  // namespace

Fam family_of(Op op) {
  static const Fam m[6] = {F_M1C2P, F_M1P2C, F_D1C2C, F_D1P2P, F_D1C2P, F_D1P2C};
  return m[op];
}
bool is_derivative(Op op) { return op >= D1_C2C; }
bool input_at_centre(Op op) { return op == MIDP_C2P || op == D1_C2C || op == D1_C2P; }
bool output_at_centre(Op op) { return op == MIDP_P2C || op == D1_C2C || op == D1_P2C; }

double coeff(Fam f, bool rhs, int row, int k, int bc, int acc) {
  const Tab& t = tab();
  return rhs ? t.r[f][row - 1][k - 1][bc][acc - 1] : t.f[f][row - 1][k - 1][bc][acc - 1];
}

void build_lhs(Fam f, int n, int bc1, int bc2, int acc, std::vector<double>& a, std::vector<double>& b,
               std::vector<double>& c, std::vector<double>& d) {
  a.assign(n, 0.0); b.assign(n, 0.0); c.assign(n, 0.0); d.assign(n, 0.0);
  for (int k = 1; k <= 3; ++k) {
    std::vector<double>& v = (k == 1) ? a : (k == 2) ? b : c;
    v[0] = coeff(f, false, 1, k, bc1, acc);
    v[1] = coeff(f, false, 2, k, bc1, acc);
    for (int i = 2; i < n - 2; ++i) v[i] = coeff(f, false, 3, k, BC_PERIODIC, acc);
    v[n - 2] = coeff(f, false, 4, k, bc2, acc);
    v[n - 1] = coeff(f, false, 5, k, bc2, acc);
  }
  int m = (bc1 == BC_PERIODIC && bc2 == BC_PERIODIC) ? n - 1 : n;
  c[0] = c[0] / b[0];
  for (int i = 1; i < m; ++i) {
    d[i] = 1.0 / (b[i] - a[i] * c[i - 1]);
    if (i < m - 1) c[i] = c[i] * d[i];
  }
}

void line_rhs_host(const LineGeom& g, Op op, int acc, const int ibc[2], const double* fi, int n_in, double* fo,
                   int n_out, const double* fbc) {
  const Fam fam = family_of(op);
  double dp[4];
  ghost_spacing(g, op, dp);
  double gh[4];
  ghosts(input_at_centre(op), fi, n_in, ibc, fbc, dp, gh);
  auto FE = [&](int k) -> double {  // 1-based ghost-extended input
    if (k == -1) return gh[0];
    if (k == 0) return gh[1];
    if (k == n_in + 1) return gh[2];
    if (k == n_in + 2) return gh[3];
    return fi[k - 1];
  };
  int p1, p2, p3, p4;
  double sgn;
  bulk_offsets(op, &p1, &p2, &p3, &p4, &sgn);
  auto stencil = [&](int i, int row, int bc) {
    double c1 = coeff(fam, true, row, 1, bc, acc), c2 = coeff(fam, true, row, 2, bc, acc);
    return c1 * (FE(i + p1) + sgn * FE(i + p2)) + c2 * (FE(i + p3) + sgn * FE(i + p4));
  };
  const double dd = is_derivative(op) ? g.h1r : 1.0;
  for (int i = 1; i <= n_out; ++i) fo[i - 1] = 0.0;
  for (int i = 3; i <= n_out - 2; ++i) fo[i - 1] = stencil(i, 3, BC_PERIODIC);
  RowRule rr[4];
  row_rules(op, ibc, n_out, rr);
  const int rows_out[4] = {1, 2, n_out - 1, n_out};
  for (int q = 0; q < 4; ++q) {
    const RowRule& r = rr[q];
    int i = rows_out[q];
    double v = 0.0;
    if (r.mode == RM_STENCIL) {
      v = stencil(i, r.coef_row, r.coef_bc);
    } else if (r.mode == RM_ONESIDED) {
      for (int k = 1; k <= 6; ++k) v = v + coeff(fam, true, r.coef_row, k, BC_INTRPL, acc) * FE(r.start + (k - 1) * r.step);
    } else {
      v = fbc[r.side];
      if (r.div_dd) v = v / dd;
    }
    fo[i - 1] = v;
  }
}

std::string make_line_plan(const LineGeom& g, Op op, int acc, const int ibc0[2], bool has_fbc, int n_in, int n_out,
                           LinePlan* P, int chunk_rows) {
  if (n_out < 8 || n_in < 8) return "line too short (need at least 8 points)";
  if (acc < CD2 || acc > CP6) return "invalid accuracy";
  for (int s = 0; s < 2; ++s)
    if (ibc0[s] < BC_INTERIOR || ibc0[s] > BC_INTRPL) return "invalid boundary kind";
  LinePlan& p = *P;
  p = LinePlan();
  p.dir = g.dir; p.op = op; p.acc = acc; p.has_fbc = has_fbc; p.n_in = n_in; p.n_out = n_out;
  p.ibc[0] = ibc0[0]; p.ibc[1] = ibc0[1];
  // reduce_bc_to_interp when no boundary values are passed
  if (!has_fbc)
    for (int s = 0; s < 2; ++s) {
      int b = p.ibc[s];
      bool red = (b == BC_INTERIOR || b == BC_NEUMANN) ||
                 (b == BC_DIRICHLET && (op == MIDP_C2P || op == D1_C2C || op == D1_C2P));
      if (red) p.ibc[s] = BC_INTRPL;
    }
  if ((p.ibc[0] == BC_INTERIOR || p.ibc[1] == BC_INTERIOR) && !has_fbc) return "interior bc needs fbc";
  // shape consistency: periodic lines keep their length
  bool per = (p.ibc[0] == BC_PERIODIC);
  if (per && n_in != n_out) return "periodic line lengths differ";
  if (!per) {
    int expect = (op == MIDP_C2P || op == D1_C2P) ? n_in + 1 : (op == MIDP_P2C || op == D1_P2C) ? n_in - 1 : n_in;
    if (n_out != expect) return "line lengths inconsistent with operator";
  }
  const Fam fam = family_of(op);
  p.dd = is_derivative(op) ? g.h1r : 1.0;
  p.c1 = coeff(fam, true, 3, 1, BC_PERIODIC, acc);
  p.c2 = coeff(fam, true, 3, 2, BC_PERIODIC, acc);
  bulk_offsets(op, &p.p1, &p.p2, &p.p3, &p.p4, &p.sgn);
  // probe the four boundary rows as affine maps of (f, fbc)
  {
    std::vector<int> probe;
    for (int k = 0; k < std::min(8, n_in); ++k) probe.push_back(k);
    for (int k = std::max(8, n_in - 8); k < n_in; ++k) probe.push_back(k);
    std::vector<double> fi(n_in, 0.0), fo(n_out, 0.0);
    double fbc[4] = {0, 0, 0, 0};
    const double* fb = has_fbc ? fbc : nullptr;
    const int rows_out[4] = {0, 1, n_out - 2, n_out - 1};
    for (int q = 0; q < 4; ++q) p.rows[q] = RowRecipe();
    // the rows must be linear in (f, fbc): a constant term means a ghost value
    // the reference leaves undefined for this pair of boundary kinds
    line_rhs_host(g, op, acc, p.ibc, fi.data(), n_in, fo.data(), n_out, fb);
    for (int q = 0; q < 4; ++q)
      if (fo[rows_out[q]] != 0.0) return "boundary kinds combination uses an undefined ghost value";
    for (int k : probe) {
      fi[k] = 1.0;
      line_rhs_host(g, op, acc, p.ibc, fi.data(), n_in, fo.data(), n_out, fb);
      fi[k] = 0.0;
      for (int q = 0; q < 4; ++q) {
        double w = fo[rows_out[q]];
        if (w != 0.0) {
          RowRecipe& r = p.rows[q];
          if (r.nterm >= kMaxRowTerms) return "boundary row has too many terms";
          r.idx[r.nterm] = k;
          r.w[r.nterm] = w;
          r.nterm++;
        }
      }
    }
    if (has_fbc)
      for (int m = 0; m < 4; ++m) {
        fbc[m] = 1.0;
        line_rhs_host(g, op, acc, p.ibc, fi.data(), n_in, fo.data(), n_out, fb);
        fbc[m] = 0.0;
        for (int q = 0; q < 4; ++q) p.rows[q].u[m] = fo[rows_out[q]];
      }
  }
  // LHS
  if (acc == CP4 || acc == CP6) {
    std::vector<double> b;
    build_lhs(fam, n_out, p.ibc[0], p.ibc[1], acc, p.a, b, p.c, p.d);
    p.b0 = b[0];
    if (per) {
      p.solve = 2;
      const int n = n_out, m = n - 1;
      // RHS-independent correction vector (Solve_TDMA_cyclic, src/basics_algorithms.f90:68-90)
      std::vector<double> x1(n, 0.0);
      x1[0] = -p.a[0];
      x1[m - 1] = -p.c[m - 1];
      x1[0] = x1[0] / b[0];
      for (int i = 1; i < m; ++i) x1[i] = (x1[i] - p.a[i] * x1[i - 1]) * p.d[i];
      for (int i = m - 2; i >= 0; --i) x1[i] = x1[i] - p.c[i] * x1[i + 1];
      p.x1 = x1;
      p.cyc_an = p.a[n - 1];
      p.cyc_cn = p.c[n - 1];
      p.cyc_den = b[n - 1] + p.c[n - 1] * x1[0] + p.a[n - 1] * x1[m - 1];
      // weights of the basic solution's first entry in terms of the forward sweep
      p.pw.assign(n, 0.0);
      p.pw[0] = 1.0;
      for (int i = 1; i < m; ++i) p.pw[i] = -p.c[i - 1] * p.pw[i - 1];
    } else {
      p.solve = 1;
    }
  }
  // chunk boundaries of the partitioned evaluation (all plans)
  {
    const int n = n_out;
    p.part_ch = (n <= kChunk * kMaxChunks) ? kChunk : kChunkLong;
    if (chunk_rows == kChunkLong) p.part_ch = kChunkLong;
    else if (chunk_rows != 0 && chunk_rows != kChunk) return "chunk rows must be kChunk or kChunkLong";
    int P = (n + p.part_ch - 1) / p.part_ch;
    if (P < 1) P = 1;
    p.part_P = P;
    p.part_start.resize(P + 1);
    for (int q = 0; q <= P; ++q) p.part_start[q] = (int)((long)q * n / P);
    for (int q = 0; q < P; ++q)
      if (p.part_start[q + 1] - p.part_start[q] < 2) return "chunk shorter than two rows";
  }
  // partitioned form of the same tridiagonal system
  if (p.solve) {
    const int n = n_out;
    std::vector<double> a(n), b(n), c(n);
    for (int i = 0; i < n; ++i) {
      int row = (i == 0) ? 1 : (i == 1) ? 2 : (i == n - 2) ? 4 : (i == n - 1) ? 5 : 3;
      int bc = (i < 2) ? p.ibc[0] : (i >= n - 2) ? p.ibc[1] : BC_PERIODIC;
      a[i] = coeff(fam, false, row, 1, bc, acc);
      b[i] = coeff(fam, false, row, 2, bc, acc);
      c[i] = coeff(fam, false, row, 3, bc, acc);
    }
    if (!per) {
      a[0] = 0.0;      // no coupling outside the line
      c[n - 1] = 0.0;
    }
    p.a_raw = a; p.b_raw = b; p.c_raw = c;
    const int P = p.part_P;
    p.ld.assign(n, 0.0); p.lc.assign(n, 0.0); p.sv.assign(n, 0.0); p.sw.assign(n, 0.0);
    for (int q = 0; q < P; ++q) {
      const int s0 = p.part_start[q], e0 = p.part_start[q + 1];
      // local factors
      for (int i = s0; i < e0; ++i) {
        if (i == s0) {
          p.ld[i] = 1.0 / b[i];
          p.lc[i] = c[i] * p.ld[i];
        } else {
          p.ld[i] = 1.0 / (b[i] - a[i] * p.lc[i - 1]);
          p.lc[i] = c[i] * p.ld[i];
        }
      }
      auto local_solve = [&](std::vector<double>& x) {
        for (int i = s0; i < e0; ++i) x[i] = (i == s0) ? x[i] * p.ld[i] : (x[i] - a[i] * x[i - 1]) * p.ld[i];
        for (int i = e0 - 2; i >= s0; --i) x[i] = x[i] - p.lc[i] * x[i + 1];
      };
      std::vector<double> v(n, 0.0), w(n, 0.0);
      v[s0] = a[s0];
      w[e0 - 1] = c[e0 - 1];
      local_solve(v);
      local_solve(w);
      for (int i = s0; i < e0; ++i) { p.sv[i] = v[i]; p.sw[i] = w[i]; }
    }
    // reduced system on (F_q, L_q): F_q + sv[s]L_{q-1} + sw[s]F_{q+1} = yF_q, same for L_q
    const int M = 2 * P;
    std::vector<double> R((size_t)M * M, 0.0);
    for (int q = 0; q < P; ++q) {
      const int s0 = p.part_start[q], e0 = p.part_start[q + 1];
      const int qm = (q + P - 1) % P, qp = (q + 1) % P;
      for (int t = 0; t < 2; ++t) {
        const int row = 2 * q + t, i = t ? e0 - 1 : s0;
        R[(size_t)row * M + row] += 1.0;
        if (per || q > 0) R[(size_t)row * M + 2 * qm + 1] += p.sv[i];
        if (per || q < P - 1) R[(size_t)row * M + 2 * qp] += p.sw[i];
      }
    }
    // invert R by Gauss-Jordan with partial pivoting
    std::vector<double> I((size_t)M * M, 0.0);
    for (int i = 0; i < M; ++i) I[(size_t)i * M + i] = 1.0;
    for (int col = 0; col < M; ++col) {
      int piv = col;
      for (int r = col + 1; r < M; ++r)
        if (std::fabs(R[(size_t)r * M + col]) > std::fabs(R[(size_t)piv * M + col])) piv = r;
      if (std::fabs(R[(size_t)piv * M + col]) < 1e-300) return "singular reduced system";
      if (piv != col)
        for (int k = 0; k < M; ++k) {
          std::swap(R[(size_t)piv * M + k], R[(size_t)col * M + k]);
          std::swap(I[(size_t)piv * M + k], I[(size_t)col * M + k]);
        }
      const double inv = 1.0 / R[(size_t)col * M + col];
      for (int k = 0; k < M; ++k) { R[(size_t)col * M + k] *= inv; I[(size_t)col * M + k] *= inv; }
      for (int r = 0; r < M; ++r) {
        if (r == col) continue;
        const double f = R[(size_t)r * M + col];
        if (f == 0.0) continue;
        for (int k = 0; k < M; ++k) {
          R[(size_t)r * M + k] -= f * R[(size_t)col * M + k];
          I[(size_t)r * M + k] -= f * I[(size_t)col * M + k];
        }
      }
    }
    p.rinv = I;
  }
  // y metric and Neumann overwrite
  if (g.dir == 1 && is_derivative(op)) {
    if (g.stretched) {
      const auto& m = output_at_centre(op) ? g.ymc : g.ymp;
      p.metric.assign(m.begin(), m.begin() + n_out);
    }
    if (op == D1_P2P || op == D1_C2P) {
      p.neu_lo = (p.ibc[0] == BC_NEUMANN && has_fbc);
      p.neu_hi = (p.ibc[1] == BC_NEUMANN && has_fbc);
    }
  }
  return "";
}

void apply_plan_host(const LinePlan& p, const double* f, const double* fbc, double* out) {
  const int n_in = p.n_in, n_out = p.n_out;
  auto clampi = [&](int k) { return k < 0 ? 0 : (k >= n_in ? n_in - 1 : k); };
  auto special = [&](int q) {
    const RowRecipe& R = p.rows[q];
    double v = 0.0;
    for (int t = 0; t < R.nterm; ++t) v = v + R.w[t] * f[R.idx[t]];
    if (fbc)
      for (int m = 0; m < 4; ++m) v = v + R.u[m] * fbc[m];
    return v * p.dd;
  };
  std::vector<double> y(n_out);
  for (int r = 0; r < n_out; ++r) {
    if (r < 2) y[r] = special(r);
    else if (r >= n_out - 2) y[r] = special(r - n_out + 4);
    else
      y[r] = (p.c1 * (f[clampi(r + p.p1)] + p.sgn * f[clampi(r + p.p2)]) +
              p.c2 * (f[clampi(r + p.p3)] + p.sgn * f[clampi(r + p.p4)])) * p.dd;
  }
  if (p.solve) {
    const int P = p.part_P, M = 2 * P;
    std::vector<double> Y(M);
    for (int q = 0; q < P; ++q) {
      const int s0 = p.part_start[q], e0 = p.part_start[q + 1];
      for (int i = s0; i < e0; ++i) y[i] = (i == s0) ? y[i] * p.ld[i] : (y[i] - p.a_raw[i] * y[i - 1]) * p.ld[i];
      for (int i = e0 - 2; i >= s0; --i) y[i] = y[i] - p.lc[i] * y[i + 1];
      Y[2 * q] = y[s0];
      Y[2 * q + 1] = y[e0 - 1];
    }
    std::vector<double> U(M, 0.0);
    for (int r = 0; r < M; ++r) {
      double acc = 0.0;
      for (int k = 0; k < M; ++k) acc = acc + p.rinv[(size_t)r * M + k] * Y[k];
      U[r] = acc;
    }
    for (int q = 0; q < P; ++q) {
      const int s0 = p.part_start[q], e0 = p.part_start[q + 1];
      const double Lm = U[2 * ((q + P - 1) % P) + 1], Fp = U[2 * ((q + 1) % P)];
      for (int i = s0; i < e0; ++i) y[i] = y[i] - p.sv[i] * Lm - p.sw[i] * Fp;
    }
  }
  for (int r = 0; r < n_out; ++r) {
    double val = y[r];
    if (!p.metric.empty()) val = val * p.metric[r];
    if (r == 0 && p.neu_lo) val = fbc[0];
    if (r == n_out - 1 && p.neu_hi) val = fbc[1];
    out[r] = val;
  }
}

}  // namespace chs
