/*
 * Oracle part 1: compact-scheme coefficient tables, TDMA, ghost cells, RHS
 * stencils, 1-D line drivers and 3-D drivers.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows (reference paths relative to the CHAPSim2 tree):
 *   tables      src/basics_operations2.f90:361-1509  (Prepare_compact_coefficients)
 *   LHS         src/basics_operations2.f90:1529-1588 (Buildup_TDMA_LHS_array)
 *   TDMA        src/basics_algorithms.f90:14-109
 *   ghosts      src/basics_operations2.f90:1803-1968
 *   RHS rows    src/basics_operations2.f90:1980-2719
 *   1-D drivers src/basics_operations2.f90:2739-3933
 *   3-D drivers src/basics_operations2.f90:3952-4790
 *
 * Pinning (tests/test_oracle_ops.py, test_oracle_tdma.py, test_oracle_pins.py):
 *   TDMA      - the reference's two known-answer vectors
 *               (src/basics_algorithms.f90:122-127, 168-181) and dense solves.
 *   tables    - periodic rows reproduce the published transfer functions of
 *               all 24 (operator, accuracy) pairs.
 *   closures  - every operator is exact on linear data at every row of a
 *               uniform wall-bounded line (Dirichlet, Neumann, interpolation
 *               closures), and converges on a smooth non-polynomial profile
 *               at every row, on uniform and tanh-stretched lines, at first
 *               order or better (derivatives) and second order or better
 *               (interpolations) for all three closure kinds.  The Neumann
 *               rows of the compact P2P / C2P derivatives on stretched lines
 *               are pinned for the zero wall gradient only (the reference
 *               enters the physical gradient unscaled, consistent only then).
 *   3-D       - the 3-D drivers equal the 1-D driver applied line by line.
 */
#include "chs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NL 5
#define NS 3
#define NBC 7
#define NACC 4
#define MAXP 1.0e20

static const double ONE_THIRD = 1.0 / 3.0;

/* [family][row][k][bc][acc] with 0-based storage, 1-based macros */
static double TF[6][NL][NS][NBC][NACC];
static double TR[6][NL][2 * NS][NBC][NACC];
static int tables_ready = 0;

#define LF(f, r, k, bc, n) TF[f][(r)-1][(k)-1][bc][(n)-1]
#define LR(f, r, k, bc, n) TR[f][(r)-1][(k)-1][bc][(n)-1]

enum { D1C2C = ORC_FAM_D1C2C, D1P2P = ORC_FAM_D1P2P, D1C2P = ORC_FAM_D1C2P,
       D1P2C = ORC_FAM_D1P2C, M1C2P = ORC_FAM_M1C2P, M1P2C = ORC_FAM_M1P2C };
enum { INT_ = ORC_BC_INTERIOR, PER = ORC_BC_PERIODIC, SYM = ORC_BC_SYMMETRIC,
       ASY = ORC_BC_ASYMMETRIC, DIR = ORC_BC_DIRICHLET, NEU = ORC_BC_NEUMANN,
       ITP = ORC_BC_INTRPL };

static void copy_rows(int f, int bcdst, int bcsrc, int r0, int r1, int accdst,
                      int accsrc, int do_lhs, int do_rhs) {
  for (int r = r0; r <= r1; ++r) {
    if (do_lhs)
      for (int k = 1; k <= NS; ++k) LF(f, r, k, bcdst, accdst) = LF(f, r, k, bcsrc, accsrc);
    if (do_rhs)
      for (int k = 1; k <= 2 * NS; ++k) LR(f, r, k, bcdst, accdst) = LR(f, r, k, bcsrc, accsrc);
  }
}

/* rows 1 (given) and 5 (mirrored) of the LHS */
static void lhs_edge(int f, int bc, int n, double a1, double b1, double c1) {
  LF(f, 1, 1, bc, n) = a1;
  LF(f, 1, 2, bc, n) = b1;
  LF(f, 1, 3, bc, n) = c1;
  LF(f, 5, 1, bc, n) = c1;
  LF(f, 5, 2, bc, n) = b1;
  LF(f, 5, 3, bc, n) = a1;
}

/* symmetric/asymmetric/periodic-style bc: edge rows given, rows 2:4 and rhs
 * copied from periodic (src/basics_operations2.f90:441-462 and analogues) */
static void set_reflect(int f, int bc, int n, double b1, double c1) {
  lhs_edge(f, bc, n, 0.0, b1, c1);
  copy_rows(f, bc, PER, 2, 4, n, n, 1, 0);
  copy_rows(f, bc, PER, 1, 5, n, n, 0, 1);
}

/* interior bc: periodic, with rows 1,5 from CD4 for compact schemes */
static void set_interior(int f) {
  for (int n = 1; n <= NACC; ++n) copy_rows(f, INT_, PER, 1, 5, n, n, 1, 1);
  for (int n = ORC_CP4; n <= ORC_CP6; ++n) {
    copy_rows(f, INT_, PER, 1, 1, n, ORC_CD4, 1, 1);
    copy_rows(f, INT_, PER, 5, 5, n, ORC_CD4, 1, 1);
  }
}

/* one-sided row r (1 or 2) plus its mirror (5 or 4); sgn = -1 for
 * derivatives (antisymmetric mirror), +1 for interpolation */
static void set_intrpl_row(int f, int r, int n, double al, const double w[6],
                           double sgn, int centred_lhs) {
  int rm = (r == 1) ? 5 : 4;
  if (centred_lhs) {
    LF(f, r, 1, ITP, n) = al;
    LF(f, r, 2, ITP, n) = 1.0;
    LF(f, r, 3, ITP, n) = al;
  } else {
    LF(f, r, 1, ITP, n) = 0.0;
    LF(f, r, 2, ITP, n) = 1.0;
    LF(f, r, 3, ITP, n) = al;
  }
  for (int k = 1; k <= 6; ++k) LR(f, r, k, ITP, n) = w[k - 1];
  LF(f, rm, 1, ITP, n) = LF(f, r, 3, ITP, n);
  LF(f, rm, 2, ITP, n) = LF(f, r, 2, ITP, n);
  LF(f, rm, 3, ITP, n) = LF(f, r, 1, ITP, n);
  for (int k = 1; k <= 6; ++k) LR(f, rm, k, ITP, n) = sgn * LR(f, r, k, ITP, n);
}

static void build_tables(void) {
  double al[NACC + 1], a[NACC + 1], b[NACC + 1];
  for (int f = 0; f < 6; ++f)
    for (int r = 0; r < NL; ++r)
      for (int bc = 0; bc < NBC; ++bc)
        for (int n = 0; n < NACC; ++n) {
          for (int k = 0; k < NS; ++k) TF[f][r][k][bc][n] = MAXP;
          for (int k = 0; k < 2 * NS; ++k) TR[f][r][k][bc][n] = MAXP;
        }

  /* ---------------- 1st derivative, collocated (C2C / P2P) -------------- */
  /* Gaitonde & Visbal (1998) Table 2.1; src/basics_operations2.f90:409-437 */
  al[ORC_CD2] = 0.0;        a[ORC_CD2] = 1.0;        b[ORC_CD2] = 0.0;
  al[ORC_CD4] = 0.0;        a[ORC_CD4] = 4.0 / 3.0;  b[ORC_CD4] = -1.0 / 3.0;
  al[ORC_CP4] = 1.0 / 4.0;  a[ORC_CP4] = 3.0 / 2.0;  b[ORC_CP4] = 0.0;
  al[ORC_CP6] = 1.0 / 3.0;  a[ORC_CP6] = 14.0 / 9.0; b[ORC_CP6] = 1.0 / 9.0;
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      LF(D1C2C, r, 1, PER, n) = al[n];
      LF(D1C2C, r, 2, PER, n) = 1.0;
      LF(D1C2C, r, 3, PER, n) = al[n];
      LR(D1C2C, r, 1, PER, n) = a[n] * 0.5;
      LR(D1C2C, r, 2, PER, n) = b[n] * 0.25;
    }
  for (int n = 1; n <= NACC; ++n) {
    set_reflect(D1C2C, SYM, n, 1.0 - al[n], al[n]);
    set_reflect(D1C2C, ASY, n, 1.0 + al[n], al[n]);
  }
  set_interior(D1C2C);
  {
    /* Table 2.2, src/basics_operations2.f90:497-553 */
    double w1[NACC + 1][6] = {{0}}, w2[NACC + 1][6] = {{0}};
    double al1[NACC + 1] = {0}, al2[NACC + 1] = {0};
    al1[ORC_CD2] = 0.0;
    w1[ORC_CD2][0] = -3.0 / 2.0; w1[ORC_CD2][1] = 2.0; w1[ORC_CD2][2] = -1.0 / 2.0;
    al1[ORC_CD4] = 0.0;
    w1[ORC_CD4][0] = -25.0 / 12.0; w1[ORC_CD4][1] = 4.0; w1[ORC_CD4][2] = -3.0;
    w1[ORC_CD4][3] = 4.0 / 3.0; w1[ORC_CD4][4] = -1.0 / 4.0;
    al1[ORC_CP4] = 3.0;
    w1[ORC_CP4][0] = -17.0 / 6.0; w1[ORC_CP4][1] = 3.0 / 2.0; w1[ORC_CP4][2] = 3.0 / 2.0;
    w1[ORC_CP4][3] = -1.0 / 6.0;
    al1[ORC_CP6] = 5.0;
    w1[ORC_CP6][0] = -197.0 / 60.0; w1[ORC_CP6][1] = -5.0 / 12.0; w1[ORC_CP6][2] = 5.0;
    w1[ORC_CP6][3] = -5.0 / 3.0; w1[ORC_CP6][4] = 5.0 / 12.0; w1[ORC_CP6][5] = -1.0 / 20.0;
    /* Tables 2.3/2.6, src/basics_operations2.f90:566-613 */
    al2[ORC_CD2] = 0.0;
    w2[ORC_CD2][0] = -1.0 / 2.0; w2[ORC_CD2][1] = 0.0; w2[ORC_CD2][2] = 1.0 / 2.0;
    al2[ORC_CD4] = 0.0;
    w2[ORC_CD4][0] = -1.0 / 4.0; w2[ORC_CD4][1] = -5.0 / 6.0; w2[ORC_CD4][2] = 3.0 / 2.0;
    w2[ORC_CD4][3] = -1.0 / 2.0; w2[ORC_CD4][4] = 1.0 / 12.0;
    al2[ORC_CP4] = 1.0 / 4.0;
    w2[ORC_CP4][0] = -3.0 / 4.0; w2[ORC_CP4][1] = 0.0; w2[ORC_CP4][2] = 3.0 / 4.0;
    al2[ORC_CP6] = 2.0 / 11.0;
    w2[ORC_CP6][0] = -20.0 / 33.0; w2[ORC_CP6][1] = -35.0 / 132.0; w2[ORC_CP6][2] = 34.0 / 33.0;
    w2[ORC_CP6][3] = -7.0 / 33.0; w2[ORC_CP6][4] = 2.0 / 33.0; w2[ORC_CP6][5] = -1.0 / 132.0;
    for (int n = 1; n <= NACC; ++n) {
      set_intrpl_row(D1C2C, 1, n, al1[n], w1[n], -1.0, 0);
      copy_rows(D1C2C, ITP, PER, 3, 3, n, n, 1, 1);
      set_intrpl_row(D1C2C, 2, n, al2[n], w2[n], -1.0, 1);
    }
  }
  /* Dirichlet / Neumann via ghost cells (bc_ghost_cd), :621-626 */
  for (int n = 1; n <= NACC; ++n) {
    copy_rows(D1C2C, DIR, INT_, 1, 5, n, n, 1, 1);
    copy_rows(D1C2C, NEU, INT_, 1, 5, n, n, 1, 1);
  }

  /* P2P, src/basics_operations2.f90:646-733 */
  for (int n = 1; n <= NACC; ++n) copy_rows(D1P2P, PER, PER, 1, 5, n, n, 0, 0);
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      for (int k = 1; k <= NS; ++k) LF(D1P2P, r, k, PER, n) = LF(D1C2C, r, k, PER, n);
      for (int k = 1; k <= 2 * NS; ++k) LR(D1P2P, r, k, PER, n) = LR(D1C2C, r, k, PER, n);
    }
  for (int n = 1; n <= NACC; ++n) {
    set_reflect(D1P2P, SYM, n, 1.0, 0.0);
    set_reflect(D1P2P, ASY, n, 1.0, 2.0 * al[n]);
  }
  set_interior(D1P2P);
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      for (int k = 1; k <= NS; ++k) LF(D1P2P, r, k, ITP, n) = LF(D1C2C, r, k, ITP, n);
      for (int k = 1; k <= 2 * NS; ++k) LR(D1P2P, r, k, ITP, n) = LR(D1C2C, r, k, ITP, n);
    }
  for (int n = 1; n <= NACC; ++n) {
    lhs_edge(D1P2P, NEU, n, 0.0, 1.0, 0.0);
    for (int k = 1; k <= 2 * NS; ++k) {
      LR(D1P2P, 1, k, NEU, n) = 0.0;
      LR(D1P2P, 5, k, NEU, n) = 0.0;
    }
    copy_rows(D1P2P, DIR, INT_, 1, 5, n, n, 1, 1);
    copy_rows(D1P2P, NEU, INT_, 2, 4, n, n, 1, 1);
  }

  /* ---------------- 1st derivative, staggered C2P / P2C ----------------- */
  /* Table 2.10, src/basics_operations2.f90:743-770 */
  al[ORC_CD2] = 0.0; a[ORC_CD2] = 1.0; b[ORC_CD2] = 0.0;
  al[ORC_CD4] = 0.0;
  a[ORC_CD4] = (9.0 - 6.0 * al[ORC_CD4]) / 8.0;
  b[ORC_CD4] = (-1.0 + 22.0 * al[ORC_CD4]) / 8.0;
  al[ORC_CP4] = 1.0 / 22.0;
  a[ORC_CP4] = (9.0 - 6.0 * al[ORC_CP4]) / 8.0;
  b[ORC_CP4] = (-1.0 + 22.0 * al[ORC_CP4]) / 8.0;
  al[ORC_CP6] = 9.0 / 62.0; a[ORC_CP6] = 63.0 / 62.0; b[ORC_CP6] = 17.0 / 62.0;
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      LF(D1C2P, r, 1, PER, n) = al[n];
      LF(D1C2P, r, 2, PER, n) = 1.0;
      LF(D1C2P, r, 3, PER, n) = al[n];
      LR(D1C2P, r, 1, PER, n) = a[n];
      LR(D1C2P, r, 2, PER, n) = b[n] * ONE_THIRD;
      LR(D1C2P, r, 3, PER, n) = 0.0;
    }
  for (int n = 1; n <= NACC; ++n) {
    set_reflect(D1C2P, SYM, n, 1.0, 0.0);
    set_reflect(D1C2P, ASY, n, 1.0, 2.0 * al[n]);
  }
  set_interior(D1C2P);
  {
    /* Table 2.12, :835-867 */
    double w1[NACC + 1][6] = {{0}}, w2[NACC + 1][6] = {{0}};
    double al1[NACC + 1] = {0}, al2[NACC + 1] = {0};
    al1[ORC_CD2] = 0.0;
    w1[ORC_CD2][0] = -2.0 - al1[ORC_CD2]; w1[ORC_CD2][1] = 3.0 + al1[ORC_CD2]; w1[ORC_CD2][2] = -1.0;
    for (int n = ORC_CD4; n <= ORC_CP4; ++n) {
      al1[n] = 0.0;
      w1[n][0] = (-93.0 - 22.0 * al1[n]) / 24.0;
      w1[n][1] = (229.0 + 17.0 * al1[n]) / 24.0;
      w1[n][2] = (-75.0 + 3.0 * al1[n]) / 8.0;
      w1[n][3] = (111.0 - 5.0 * al1[n]) / 24.0;
      w1[n][4] = (-22.0 + 1.0 * al1[n]) / 24.0;
    }
    al1[ORC_CP6] = 1689.0 / 71.0;
    w1[ORC_CP6][0] = (-3043.0 - 563.0 * al1[ORC_CP6]) / 640.0;
    w1[ORC_CP6][1] = (5353.0 + 201.0 * al1[ORC_CP6]) / 384.0;
    w1[ORC_CP6][2] = (-3489.0 + 143.0 * al1[ORC_CP6]) / 192.0;
    w1[ORC_CP6][3] = (859.0 - 37.0 * al1[ORC_CP6]) / 64.0;
    w1[ORC_CP6][4] = (-2041.0 + 87.0 * al1[ORC_CP6]) / 384.0;
    w1[ORC_CP6][5] = (1689.0 - 71.0 * al1[ORC_CP6]) / 1920.0;
    /* Table 2.13, :897-935 */
    al2[ORC_CD2] = 0.0; w2[ORC_CD2][0] = -1.0; w2[ORC_CD2][1] = 1.0;
    for (int n = ORC_CD4; n <= ORC_CP4; ++n) {
      al2[n] = (n == ORC_CD4) ? 0.0 : 1.0 / 22.0;
      w2[n][0] = (-11.0 - 46.0 * al2[n]) / 12.0;
      w2[n][1] = (17.0 + 202.0 * al2[n]) / 24.0;
      w2[n][2] = (3.0 - 66.0 * al2[n]) / 8.0;
      w2[n][3] = (-5.0 + 110.0 * al2[n]) / 24.0;
      w2[n][4] = (1.0 - 22.0 * al2[n]) / 24.0;
    }
    al2[ORC_CP6] = 31.0 / 818.0;
    w2[ORC_CP6][0] = -5195.0 / 4908.0; w2[ORC_CP6][1] = 4957.0 / 4908.0;
    w2[ORC_CP6][2] = 119.0 / 1227.0; w2[ORC_CP6][3] = -85.0 / 1227.0;
    w2[ORC_CP6][4] = 119.0 / 4908.0; w2[ORC_CP6][5] = -17.0 / 4908.0;
    for (int n = 1; n <= NACC; ++n) {
      set_intrpl_row(D1C2P, 1, n, al1[n], w1[n], -1.0, 0);
      copy_rows(D1C2P, ITP, PER, 3, 3, n, n, 1, 1);
      set_intrpl_row(D1C2P, 2, n, al2[n], w2[n], -1.0, 1);
    }
  }
  for (int n = 1; n <= NACC; ++n) {
    lhs_edge(D1C2P, NEU, n, 0.0, 1.0, 0.0);
    for (int k = 1; k <= 2 * NS; ++k) {
      LR(D1C2P, 1, k, NEU, n) = 0.0;
      LR(D1C2P, 5, k, NEU, n) = 0.0;
    }
    copy_rows(D1C2P, DIR, INT_, 1, 5, n, n, 1, 1);
    copy_rows(D1C2P, NEU, INT_, 2, 4, n, n, 1, 1);
  }

  /* P2C, :988-1102 (alpha/a/b still the staggered-derivative ones) */
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      for (int k = 1; k <= NS; ++k) LF(D1P2C, r, k, PER, n) = LF(D1C2P, r, k, PER, n);
      for (int k = 1; k <= 2 * NS; ++k) LR(D1P2C, r, k, PER, n) = LR(D1C2P, r, k, PER, n);
    }
  for (int n = 1; n <= NACC; ++n) {
    set_reflect(D1P2C, SYM, n, 1.0 - al[n], al[n]);
    set_reflect(D1P2C, ASY, n, 1.0 + al[n], al[n]);
  }
  set_interior(D1P2C);
  {
    /* Table 2.11, :1045-1072 */
    double w1[NACC + 1][6] = {{0}};
    double al1[NACC + 1] = {0};
    al1[ORC_CD2] = 0.0; w1[ORC_CD2][0] = -1.0; w1[ORC_CD2][1] = 1.0;
    for (int n = ORC_CD4; n <= ORC_CP4; ++n) {
      al1[n] = 0.0;
      w1[n][0] = (-22.0 + al1[n]) / 24.0;
      w1[n][1] = (17.0 - 27.0 * al1[n]) / 24.0;
      w1[n][2] = (3.0 + 9.0 * al1[n]) / 8.0;
      w1[n][3] = (-5.0 - al1[n]) / 24.0;
      w1[n][4] = 1.0 / 24.0;
    }
    al1[ORC_CP6] = 71.0 / 9.0;
    w1[ORC_CP6][0] = (-1689.0 + 71.0 * al1[ORC_CP6]) / 1920.0;
    w1[ORC_CP6][1] = (67.0 - 141.0 * al1[ORC_CP6]) / 128.0;
    w1[ORC_CP6][2] = (143.0 + 207.0 * al1[ORC_CP6]) / 192.0;
    w1[ORC_CP6][3] = (-111.0 + 1.0 * al1[ORC_CP6]) / 192.0;
    w1[ORC_CP6][4] = (29.0 - 3.0 * al1[ORC_CP6]) / 128.0;
    w1[ORC_CP6][5] = (-71.0 + 9.0 * al1[ORC_CP6]) / 1920.0;
    for (int n = 1; n <= NACC; ++n) {
      set_intrpl_row(D1P2C, 1, n, al1[n], w1[n], -1.0, 0);
      copy_rows(D1P2C, ITP, PER, 2, 4, n, n, 1, 1);
    }
  }
  for (int n = 1; n <= NACC; ++n) {
    copy_rows(D1P2C, DIR, INT_, 1, 5, n, n, 1, 1);
    copy_rows(D1P2C, NEU, INT_, 1, 5, n, n, 1, 1);
  }

  /* ---------------- mid-point interpolation C2P / P2C ------------------- */
  /* Table 2.7, src/basics_operations2.f90:1145-1170 */
  {
    double cc[NACC + 1];
    al[ORC_CD2] = 0.0; a[ORC_CD2] = 1.0; b[ORC_CD2] = 0.0; cc[ORC_CD2] = 0.0;
    al[ORC_CD4] = 0.0;
    a[ORC_CD4] = 9.0 / 8.0 + 5.0 / 4.0 * al[ORC_CD4];
    b[ORC_CD4] = -1.0 / 8.0 + 3.0 / 4.0 * al[ORC_CD4];
    cc[ORC_CD4] = 0.0;
    al[ORC_CP4] = 1.0 / 6.0;
    a[ORC_CP4] = 9.0 / 8.0 + 5.0 / 4.0 * al[ORC_CP4];
    b[ORC_CP4] = -1.0 / 8.0 + 3.0 / 4.0 * al[ORC_CP4];
    cc[ORC_CP4] = 0.0;
    al[ORC_CP6] = 3.0 / 10.0;
    a[ORC_CP6] = 5.0 * (15.0 + 14.0 * al[ORC_CP6]) / 64.0;
    b[ORC_CP6] = (-25.0 + 126.0 * al[ORC_CP6]) / 128.0;
    cc[ORC_CP6] = (3.0 - 10.0 * al[ORC_CP6]) / 128.0;
    (void)cc;
  }
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      LF(M1C2P, r, 1, PER, n) = al[n];
      LF(M1C2P, r, 2, PER, n) = 1.0;
      LF(M1C2P, r, 3, PER, n) = al[n];
      LR(M1C2P, r, 1, PER, n) = a[n] * 0.5;
      LR(M1C2P, r, 2, PER, n) = b[n] * 0.5;
      LR(M1C2P, r, 3, PER, n) = 0.0;
    }
  for (int n = 1; n <= NACC; ++n) {
    set_reflect(M1C2P, SYM, n, 1.0, al[n] + al[n]);
    set_reflect(M1C2P, ASY, n, 1.0, 0.0);
  }
  set_interior(M1C2P);
  {
    /* extrapolation rows, :1220-1307 */
    double w1[NACC + 1][6] = {{0}}, w2[NACC + 1][6] = {{0}};
    double al1[NACC + 1] = {0}, al2[NACC + 1] = {0};
    al1[ORC_CD2] = 0.0;
    w1[ORC_CD2][0] = (3.0 + al1[ORC_CD2]) / 2.0;
    w1[ORC_CD2][1] = (-1.0 + al1[ORC_CD2]) / 2.0;
    for (int n = ORC_CD4; n <= ORC_CP4; ++n) {
      al1[n] = (n == ORC_CD4) ? 0.0 : 5.0;
      w1[n][0] = (35.0 + 5.0 * al1[n]) / 16.0;
      w1[n][1] = (-35.0 + 15.0 * al1[n]) / 16.0;
      w1[n][2] = (21.0 - 5.0 * al1[n]) / 16.0;
      w1[n][3] = (-5.0 + al1[n]) / 16.0;
    }
    al1[ORC_CP6] = 9.0;
    w1[ORC_CP6][0] = (693.0 + 63.0 * al1[ORC_CP6]) / 256.0;
    w1[ORC_CP6][1] = (-1155.0 + 315.0 * al1[ORC_CP6]) / 256.0;
    w1[ORC_CP6][2] = (693.0 - 105.0 * al1[ORC_CP6]) / 128.0;
    w1[ORC_CP6][3] = (-495.0 + 63.0 * al1[ORC_CP6]) / 128.0;
    w1[ORC_CP6][4] = (385.0 - 45.0 * al1[ORC_CP6]) / 256.0;
    w1[ORC_CP6][5] = (-63.0 + 7.0 * al1[ORC_CP6]) / 256.0;
    al2[ORC_CD2] = 0.0;
    w2[ORC_CD2][0] = (1.0 + 2.0 * al2[ORC_CD2]) / 2.0;
    w2[ORC_CD2][1] = (1.0 + 2.0 * al2[ORC_CD2]) / 2.0;
    for (int n = ORC_CD4; n <= ORC_CP4; ++n) {
      al2[n] = (n == ORC_CD4) ? 0.0 : 1.0 / 6.0;
      w2[n][0] = (5.0 + 34.0 * al2[n]) / 16.0;
      w2[n][1] = (15.0 - 26.0 * al2[n]) / 16.0;
      w2[n][2] = (-5.0 + 30.0 * al2[n]) / 16.0;
      w2[n][3] = (1.0 - 6.0 * al2[n]) / 16.0;
    }
    al2[ORC_CP6] = 9.0;
    w2[ORC_CP6][0] = (63.0 + 686.0 * al2[ORC_CP6]) / 256.0;
    w2[ORC_CP6][1] = (315.0 - 1050.0 * al2[ORC_CP6]) / 256.0;
    w2[ORC_CP6][2] = (-105.0 + 798.0 * al2[ORC_CP6]) / 128.0;
    w2[ORC_CP6][3] = (63.0 - 530.0 * al2[ORC_CP6]) / 128.0;
    w2[ORC_CP6][4] = (-45.0 + 406.0 * al2[ORC_CP6]) / 256.0;
    w2[ORC_CP6][5] = (7.0 - 66.0 * al2[ORC_CP6]) / 256.0;
    for (int n = 1; n <= NACC; ++n) {
      set_intrpl_row(M1C2P, 1, n, al1[n], w1[n], +1.0, 0);
      copy_rows(M1C2P, ITP, PER, 3, 3, n, n, 1, 1);
      set_intrpl_row(M1C2P, 2, n, al2[n], w2[n], +1.0, 1);
    }
  }
  for (int n = 1; n <= NACC; ++n) {
    lhs_edge(M1C2P, DIR, n, 0.0, 1.0, 0.0);
    for (int k = 1; k <= 2 * NS; ++k) {
      LR(M1C2P, 1, k, DIR, n) = 0.0;
      LR(M1C2P, 5, k, DIR, n) = 0.0;
    }
    copy_rows(M1C2P, NEU, INT_, 1, 5, n, n, 1, 1);
    copy_rows(M1C2P, DIR, INT_, 2, 4, n, n, 1, 1);
  }

  /* P2C, :1334-1502 */
  for (int n = 1; n <= NACC; ++n)
    for (int r = 1; r <= 5; ++r) {
      for (int k = 1; k <= NS; ++k) LF(M1P2C, r, k, PER, n) = LF(M1C2P, r, k, PER, n);
      for (int k = 1; k <= 2 * NS; ++k) LR(M1P2C, r, k, PER, n) = LR(M1C2P, r, k, PER, n);
    }
  for (int n = 1; n <= NACC; ++n) {
    set_reflect(M1P2C, SYM, n, 1.0 + al[n], al[n]);
    set_reflect(M1P2C, ASY, n, 1.0 - al[n], al[n]);
  }
  set_interior(M1P2C);
  {
    /* Table 2.8, :1427-1460 */
    double w1[NACC + 1][6] = {{0}};
    double al1[NACC + 1] = {0};
    al1[ORC_CD2] = 0.0; w1[ORC_CD2][0] = 0.5; w1[ORC_CD2][1] = 0.5; w1[ORC_CD2][2] = 0.0;
    for (int n = ORC_CD4; n <= ORC_CP4; ++n) {
      al1[n] = (n == ORC_CD4) ? 0.0 : 1.0;
      w1[n][0] = (5.0 - al1[n]) / 16.0;
      w1[n][1] = (15.0 + 9.0 * al1[n]) / 16.0;
      w1[n][2] = (-5.0 + 9.0 * al1[n]) / 16.0;
      w1[n][3] = (1.0 - al1[n]) / 16.0;
    }
    al1[ORC_CP6] = 7.0 / 3.0;
    w1[ORC_CP6][0] = (63.0 - 7.0 * al1[ORC_CP6]) / 256.0;
    w1[ORC_CP6][1] = (315.0 + 105.0 * al1[ORC_CP6]) / 256.0;
    w1[ORC_CP6][2] = (-7.0 + 7.0 * al1[ORC_CP6]) * 15.0 / 128.0;
    w1[ORC_CP6][3] = (63.0 - 35.0 * al1[ORC_CP6]) / 128.0;
    w1[ORC_CP6][4] = (-15.0 + 7.0 * al1[ORC_CP6]) * 3.0 / 256.0;
    w1[ORC_CP6][5] = (7.0 - 3.0 * al1[ORC_CP6]) / 256.0;
    for (int n = 1; n <= NACC; ++n) {
      set_intrpl_row(M1P2C, 1, n, al1[n], w1[n], +1.0, 0);
      copy_rows(M1P2C, ITP, PER, 2, 4, n, n, 1, 1);
    }
  }
  for (int n = 1; n <= NACC; ++n) {
    copy_rows(M1P2C, NEU, INT_, 1, 5, n, n, 1, 1);
    copy_rows(M1P2C, DIR, INT_, 1, 5, n, n, 1, 1);
  }
  tables_ready = 1;
}

static void ensure_tables(void) {
  if (!tables_ready) {
#pragma omp critical(orc_tables)
    {
      if (!tables_ready) build_tables();
    }
  }
}

double orc_coeff(int fam, int is_rhs, int row, int k, int bc, int acc) {
  ensure_tables();
  return is_rhs ? LR(fam, row, k, bc, acc) : LF(fam, row, k, bc, acc);
}

/* ---------------------------------------------------------------------- */
/* TDMA, src/basics_algorithms.f90:14-109                                  */
/* ---------------------------------------------------------------------- */
void orc_tdma_preprocess(const double *a, const double *b, double *c,
                         double *d, int n) {
  c[0] = c[0] / b[0];
  for (int i = 1; i < n; ++i) {
    d[i] = 1.0 / (b[i] - a[i] * c[i - 1]);
    if (i < n - 1) c[i] = c[i] * d[i];
  }
}

static void tdma_basic(double *x, const double *a, const double *b,
                       const double *c, const double *d, int n) {
  x[0] = x[0] / b[0];
  for (int i = 1; i < n; ++i) x[i] = (x[i] - a[i] * x[i - 1]) * d[i];
  for (int i = n - 2; i >= 0; --i) x[i] = x[i] - c[i] * x[i + 1];
}

void orc_tdma_solve(int peri, double *x, const double *a, const double *b,
                    const double *c, const double *d, int n) {
  if (!peri) {
    tdma_basic(x, a, b, c, d, n);
    return;
  }
  double *x1 = (double *)calloc((size_t)n, sizeof(double));
  tdma_basic(x, a, b, c, d, n - 1);
  x1[0] = -a[0];
  x1[n - 2] = -c[n - 2];
  tdma_basic(x1, a, b, c, d, n - 1);
  x[n - 1] = (x[n - 1] - c[n - 1] * x[0] - a[n - 1] * x[n - 2]) /
             (b[n - 1] + c[n - 1] * x1[0] + a[n - 1] * x1[n - 2]);
  for (int i = 0; i < n - 1; ++i) x[i] = x[i] + x1[i] * x[n - 1];
  free(x1);
}

void orc_lhs(int fam, int n, int bc1, int bc2, int acc, double *a, double *b,
             double *c, double *d) {
  ensure_tables();
  for (int i = 0; i < n; ++i) a[i] = b[i] = c[i] = d[i] = 0.0;
  for (int k = 1; k <= 3; ++k) {
    double *v = (k == 1) ? a : (k == 2) ? b : c;
    v[0] = LF(fam, 1, k, bc1, acc);
    v[1] = LF(fam, 2, k, bc1, acc);
    for (int i = 2; i < n - 2; ++i) v[i] = LF(fam, 3, k, PER, acc);
    v[n - 2] = LF(fam, 4, k, bc2, acc);
    v[n - 1] = LF(fam, 5, k, bc2, acc);
  }
  if (bc1 == PER && bc2 == PER)
    orc_tdma_preprocess(a, b, c, d, n - 1);
  else
    orc_tdma_preprocess(a, b, c, d, n);
}

/* ---------------------------------------------------------------------- */
/* ghost cells, src/basics_operations2.f90:1803-1968                       */
/* g[0..3] = fc(-1), fc(0), fc(1), fc(2)                                   */
/* ---------------------------------------------------------------------- */
#define FI(k) fi[(k)-1]
static void ghost_c(const double *fi, int nc, const int ibc[2], double g[4],
                    const double *fbc, const double *dp) {
  switch (ibc[0]) {
  case INT_: g[1] = fbc[0]; g[0] = fbc[2]; break;
  case PER: g[1] = FI(nc); g[0] = FI(nc - 1); break;
  case SYM: g[1] = FI(1); g[0] = FI(2); break;
  case ASY: g[1] = -FI(1); g[0] = -FI(2); break;
  case DIR: g[1] = 2.0 * fbc[0] - FI(1); g[0] = 2.0 * fbc[0] - FI(2); break;
  case NEU: g[1] = FI(1) - fbc[0] * dp[0]; g[0] = FI(2) - fbc[0] * dp[2]; break;
  default: g[1] = MAXP; g[0] = MAXP;
  }
  switch (ibc[1]) {
  case INT_: g[2] = fbc[1]; g[3] = fbc[3]; break;
  case PER: g[2] = FI(1); g[3] = FI(2); break;
  case SYM: g[2] = FI(nc); g[3] = FI(nc - 1); break;
  case ASY: g[2] = -FI(nc); g[3] = -FI(nc - 1); break;
  case DIR: g[2] = 2.0 * fbc[1] - FI(nc); g[3] = 2.0 * fbc[1] - FI(nc - 1); break;
  case NEU: g[2] = FI(nc) + fbc[1] * dp[1]; g[3] = FI(nc - 1) + fbc[1] * dp[3]; break;
  default: g[2] = MAXP; g[3] = MAXP;
  }
}

static void ghost_p(const double *fi, int np, const int ibc[2], double g[4],
                    const double *fbc, const double *dp) {
  switch (ibc[0]) {
  case INT_: g[1] = fbc[0]; g[0] = fbc[2]; break;
  case PER: g[1] = FI(np); g[0] = FI(np - 1); break;
  case SYM: g[1] = FI(2); g[0] = FI(3); break;
  case ASY: g[1] = -FI(2); g[0] = -FI(3); break;
  case DIR:
    if (fbc) { g[1] = 2.0 * fbc[0] - FI(2); g[0] = 2.0 * fbc[0] - FI(3); }
    else { g[1] = 2.0 * FI(1) - FI(2); g[0] = 2.0 * FI(1) - FI(3); }
    break;
  case NEU: g[1] = FI(2) - fbc[0] * dp[0]; g[0] = FI(3) - fbc[0] * dp[2]; break;
  default: g[1] = MAXP; g[0] = MAXP;
  }
  switch (ibc[1]) {
  case INT_: g[2] = fbc[1]; g[3] = fbc[3]; break;
  case PER: g[2] = FI(1); g[3] = FI(2); break;
  case SYM: g[2] = FI(np - 1); g[3] = FI(np - 2); break;
  case ASY: g[2] = -FI(np - 1); g[3] = -FI(np - 2); break;
  case NEU: g[2] = FI(np - 1) + fbc[1] * dp[1]; g[3] = FI(np - 2) + fbc[1] * dp[3]; break;
  case DIR:
    if (fbc) { g[2] = 2.0 * fbc[1] - FI(np - 1); g[3] = 2.0 * fbc[1] - FI(np - 2); }
    else { g[2] = 2.0 * FI(np) - FI(np - 1); g[3] = 2.0 * FI(np) - FI(np - 2); }
    break;
  default: g[2] = MAXP; g[3] = MAXP;
  }
}
#define GM1 g[0]
#define G0 g[1]
#define G1 g[2]
#define G2 g[3]
#define CO(r, k, bc) LR(fam, r, k, bc, acc)
#define FO(k) fo[(k)-1]

static int is_reflect4(int bc) {
  return bc == INT_ || bc == SYM || bc == ASY;
}

/* six-point one-sided sum with coefficients of row r */
static double intrpl6(int fam, int acc, int r, const double *fi, int s, int step) {
  double v = 0.0;
  v = CO(r, 1, ITP) * FI(s) + CO(r, 2, ITP) * FI(s + step) +
      CO(r, 3, ITP) * FI(s + 2 * step) + CO(r, 4, ITP) * FI(s + 3 * step) +
      CO(r, 5, ITP) * FI(s + 4 * step) + CO(r, 6, ITP) * FI(s + 5 * step);
  return v;
}

/* Prepare_TDMA_interp_P2C_RHS_array, :1980-2071 */
static void rhs_midp_p2c(const double *fi, int nin, double *fo, int nc,
                         int acc, const int ibc[2], const double *fbc,
                         const double *dp) {
  const int fam = M1P2C;
  double g[4];
  (void)nin;
  for (int i = 1; i <= nc; ++i) FO(i) = 0.0;
  for (int i = 2; i <= nc - 2; ++i)
    FO(i) = CO(3, 1, PER) * (FI(i) + FI(i + 1)) + CO(3, 2, PER) * (FI(i - 1) + FI(i + 2));
  ghost_p(fi, nin, ibc, g, fbc, dp);
  int bc1 = is_reflect4(ibc[0]) || ibc[0] == PER || ibc[0] == DIR || ibc[0] == NEU;
  int bc5 = is_reflect4(ibc[1]) || ibc[1] == DIR || ibc[1] == NEU;
  int i = 1;
  if (bc1)
    FO(i) = CO(1, 1, ibc[0]) * (FI(i) + FI(i + 1)) + CO(1, 2, ibc[0]) * (G0 + FI(i + 2));
  else
    FO(i) = intrpl6(fam, acc, 1, fi, i, 1);
  i = nc;
  if (bc5)
    FO(i) = CO(5, 1, ibc[1]) * (FI(i) + FI(i + 1)) + CO(5, 2, ibc[1]) * (FI(i - 1) + G1);
  else if (ibc[1] == PER)
    FO(i) = CO(5, 1, PER) * (FI(i) + G1) + CO(5, 2, PER) * (FI(i - 1) + G2);
  else
    FO(nc) = CO(5, 1, ITP) * FI(i + 1) + CO(5, 2, ITP) * FI(i) + CO(5, 3, ITP) * FI(i - 1) +
             CO(5, 4, ITP) * FI(i - 2) + CO(5, 5, ITP) * FI(i - 3) + CO(5, 6, ITP) * FI(i - 4);
  i = nc - 1;
  if (ibc[1] == PER)
    FO(i) = CO(4, 1, PER) * (FI(i) + FI(i + 1)) + CO(4, 2, PER) * (FI(i - 1) + G1);
  else
    FO(i) = CO(4, 1, ibc[1]) * (FI(i) + FI(i + 1)) + CO(4, 2, ibc[1]) * (FI(i - 1) + FI(i + 2));
}

/* Prepare_TDMA_interp_C2P_RHS_array, :2084-2218 */
static void rhs_midp_c2p(const double *fi, int nin, double *fo, int np,
                         int acc, const int ibc[2], const double *fbc,
                         const double *dp) {
  const int fam = M1C2P;
  double g[4];
  for (int i = 1; i <= np; ++i) FO(i) = 0.0;
  for (int i = 3; i <= np - 2; ++i)
    FO(i) = CO(3, 1, PER) * (FI(i) + FI(i - 1)) + CO(3, 2, PER) * (FI(i + 1) + FI(i - 2));
  ghost_c(fi, nin, ibc, g, fbc, dp);
  int bc1 = is_reflect4(ibc[0]) || ibc[0] == PER || ibc[0] == NEU;
  int bc5 = is_reflect4(ibc[1]) || ibc[1] == NEU;
  int bc2 = bc1 || ibc[0] == DIR;
  int bc4 = bc5 || ibc[1] == DIR;
  int i = 1;
  if (bc1)
    FO(i) = CO(1, 1, ibc[0]) * (FI(i) + G0) + CO(1, 2, ibc[0]) * (FI(i + 1) + GM1);
  else if (ibc[0] == DIR)
    FO(i) = fbc[0];
  else
    FO(i) = intrpl6(fam, acc, 1, fi, i, 1);
  i = 2;
  if (bc2)
    FO(i) = CO(2, 1, ibc[0]) * (FI(i) + FI(i - 1)) + CO(2, 2, ibc[0]) * (FI(i + 1) + G0);
  else
    FO(i) = intrpl6(fam, acc, 2, fi, i - 1, 1);
  i = np;
  if (bc5)
    FO(i) = CO(5, 1, ibc[1]) * (G1 + FI(i - 1)) + CO(5, 2, ibc[1]) * (G2 + FI(i - 2));
  else if (ibc[1] == PER)
    FO(i) = CO(5, 1, PER) * (FI(i) + FI(i - 1)) + CO(5, 2, PER) * (G1 + FI(i - 2));
  else if (ibc[1] == DIR)
    FO(i) = fbc[1];
  else
    FO(i) = intrpl6(fam, acc, 5, fi, i - 1, -1);
  i = np - 1;
  if (bc4)
    FO(i) = CO(4, 1, ibc[1]) * (FI(i) + FI(i - 1)) + CO(4, 2, ibc[1]) * (G1 + FI(i - 2));
  else if (ibc[1] == PER)
    FO(i) = CO(4, 1, PER) * (FI(i) + FI(i - 1)) + CO(4, 2, PER) * (FI(i + 1) + FI(i - 2));
  else
    FO(i) = intrpl6(fam, acc, 4, fi, i, -1);
}

/* Prepare_TDMA_1deri_C2C_RHS_array, :2236-2338 */
static void rhs_1der_c2c(const double *fi, int nin, double *fo, int nc, int acc,
                         double dd, const int ibc[2], const double *fbc,
                         const double *dp) {
  const int fam = D1C2C;
  double g[4];
  int isbc[2];
  for (int i = 1; i <= nc; ++i) FO(i) = 0.0;
  for (int i = 3; i <= nc - 2; ++i)
    FO(i) = CO(3, 1, PER) * (FI(i + 1) - FI(i - 1)) + CO(3, 2, PER) * (FI(i + 2) - FI(i - 2));
  ghost_c(fi, nin, ibc, g, fbc, dp);
  for (int s = 0; s < 2; ++s)
    isbc[s] = is_reflect4(ibc[s]) || ibc[s] == PER || ibc[s] == DIR || ibc[s] == NEU;
  int i = 1;
  if (isbc[0])
    FO(i) = CO(1, 1, ibc[0]) * (FI(i + 1) - G0) + CO(1, 2, ibc[0]) * (FI(i + 2) - GM1);
  else
    FO(i) = intrpl6(fam, acc, 1, fi, i, 1);
  i = 2;
  if (isbc[0])
    FO(i) = CO(2, 1, ibc[0]) * (FI(i + 1) - FI(i - 1)) + CO(2, 2, ibc[0]) * (FI(i + 2) - G0);
  else
    FO(i) = intrpl6(fam, acc, 2, fi, i - 1, 1);
  i = nc;
  if (isbc[1])
    FO(i) = CO(5, 1, ibc[1]) * (G1 - FI(i - 1)) + CO(5, 2, ibc[1]) * (G2 - FI(i - 2));
  else
    FO(i) = intrpl6(fam, acc, 5, fi, i, -1);
  i = nc - 1;
  if (isbc[1])
    FO(i) = CO(4, 1, ibc[1]) * (FI(i + 1) - FI(i - 1)) + CO(4, 2, ibc[1]) * (G1 - FI(i - 2));
  else
    FO(i) = intrpl6(fam, acc, 4, fi, i + 1, -1);
  for (int k = 1; k <= nc; ++k) FO(k) = FO(k) * dd;
}

/* Prepare_TDMA_1deri_P2P_RHS_array, :2357-2468 */
static void rhs_1der_p2p(const double *fi, int nin, double *fo, int np, int acc,
                         double dd, const int ibc[2], const double *fbc,
                         const double *dp) {
  const int fam = D1P2P;
  double g[4];
  int b1[2], b2[2];
  for (int i = 1; i <= np; ++i) FO(i) = 0.0;
  for (int i = 3; i <= np - 2; ++i)
    FO(i) = CO(3, 1, PER) * (FI(i + 1) - FI(i - 1)) + CO(3, 2, PER) * (FI(i + 2) - FI(i - 2));
  ghost_p(fi, nin, ibc, g, fbc, dp);
  for (int s = 0; s < 2; ++s) {
    b1[s] = is_reflect4(ibc[s]) || ibc[s] == PER || ibc[s] == DIR;
    b2[s] = b1[s] || ibc[s] == DIR || ibc[s] == NEU;
  }
  int i = 1;
  if (b1[0])
    FO(i) = CO(1, 1, ibc[0]) * (FI(i + 1) - G0) + CO(1, 2, ibc[0]) * (FI(i + 2) - GM1);
  else if (ibc[0] == NEU)
    FO(i) = fbc[0] / dd;
  else
    FO(i) = intrpl6(fam, acc, 1, fi, i, 1);
  i = 2;
  if (b2[0])
    FO(i) = CO(2, 1, ibc[0]) * (FI(i + 1) - FI(i - 1)) + CO(2, 2, ibc[0]) * (FI(i + 2) - G0);
  else
    FO(i) = intrpl6(fam, acc, 2, fi, i - 1, 1);
  i = np;
  if (b1[1])
    FO(i) = CO(5, 1, ibc[1]) * (G1 - FI(i - 1)) + CO(5, 2, ibc[1]) * (G2 - FI(i - 2));
  else if (ibc[1] == NEU)
    FO(i) = fbc[1] / dd;
  else
    FO(i) = intrpl6(fam, acc, 5, fi, i, -1);
  i = np - 1;
  if (b2[1])
    FO(i) = CO(4, 1, ibc[1]) * (FI(i + 1) - FI(i - 1)) + CO(4, 2, ibc[1]) * (G1 - FI(i - 2));
  else
    FO(i) = intrpl6(fam, acc, 4, fi, i + 1, -1);
  for (int k = 1; k <= np; ++k) FO(k) = FO(k) * dd;
}

/* Prepare_TDMA_1deri_C2P_RHS_array, :2488-2613 (including its use of
 * ibc(1) when deciding the np-1 row, :2548-2552) */
static void rhs_1der_c2p(const double *fi, int nin, double *fo, int np, int acc,
                         double dd, const int ibc[2], const double *fbc,
                         const double *dp) {
  const int fam = D1C2P;
  double g[4];
  for (int i = 1; i <= np; ++i) FO(i) = 0.0;
  for (int i = 3; i <= np - 2; ++i)
    FO(i) = CO(3, 1, PER) * (FI(i) - FI(i - 1)) + CO(3, 2, PER) * (FI(i + 1) - FI(i - 2));
  ghost_c(fi, nin, ibc, g, fbc, dp);
  int bc1 = is_reflect4(ibc[0]) || ibc[0] == PER || ibc[0] == DIR;
  int bc2 = bc1 || ibc[0] == DIR || ibc[0] == NEU;
  int bc5 = is_reflect4(ibc[1]) || ibc[1] == DIR;
  int bc4 = bc5 || ibc[0] == DIR || ibc[0] == NEU;
  int i = 1;
  if (bc1)
    FO(i) = CO(1, 1, ibc[0]) * (FI(i) - G0) + CO(1, 2, ibc[0]) * (FI(i + 1) - GM1);
  else if (ibc[0] == NEU)
    FO(i) = fbc[0] / dd;
  else
    FO(i) = intrpl6(fam, acc, 1, fi, i, 1);
  i = 2;
  if (bc2)
    FO(i) = CO(2, 1, ibc[0]) * (FI(i) - FI(i - 1)) + CO(2, 2, ibc[0]) * (FI(i + 1) - G0);
  else
    FO(i) = intrpl6(fam, acc, 2, fi, i - 1, 1);
  i = np;
  if (bc5)
    FO(i) = CO(5, 1, ibc[1]) * (G1 - FI(i - 1)) + CO(5, 2, ibc[1]) * (G2 - FI(i - 2));
  else if (ibc[1] == PER)
    FO(i) = CO(5, 1, PER) * (FI(i) - FI(i - 1)) + CO(5, 2, PER) * (G1 - FI(i - 2));
  else if (ibc[1] == NEU)
    FO(i) = fbc[1] / dd;
  else
    FO(i) = intrpl6(fam, acc, 5, fi, i - 1, -1);
  i = np - 1;
  if (bc4)
    FO(i) = CO(4, 1, ibc[1]) * (FI(i) - FI(i - 1)) + CO(4, 2, ibc[1]) * (G1 - FI(i - 2));
  else if (ibc[1] == PER)
    FO(i) = CO(4, 1, PER) * (FI(i) - FI(i - 1)) + CO(4, 2, PER) * (FI(i + 1) - FI(i - 2));
  else
    FO(i) = intrpl6(fam, acc, 4, fi, i, -1);
  for (int k = 1; k <= np; ++k) FO(k) = FO(k) * dd;
}

/* Prepare_TDMA_1deri_P2C_RHS_array, :2626-2719 */
static void rhs_1der_p2c(const double *fi, int nin, double *fo, int nc, int acc,
                         double dd, const int ibc[2], const double *fbc,
                         const double *dp) {
  const int fam = D1P2C;
  double g[4];
  for (int i = 1; i <= nc; ++i) FO(i) = 0.0;
  for (int i = 2; i <= nc - 2; ++i)
    FO(i) = CO(3, 1, PER) * (FI(i + 1) - FI(i)) + CO(3, 2, PER) * (FI(i + 2) - FI(i - 1));
  ghost_p(fi, nin, ibc, g, fbc, dp);
  int bc1 = is_reflect4(ibc[0]) || ibc[0] == PER || ibc[0] == DIR || ibc[0] == NEU;
  int bc5 = is_reflect4(ibc[1]) || ibc[1] == DIR || ibc[1] == NEU;
  int i = 1;
  if (bc1)
    FO(i) = CO(1, 1, ibc[0]) * (FI(i + 1) - FI(i)) + CO(1, 2, ibc[0]) * (FI(i + 2) - G0);
  else
    FO(i) = intrpl6(fam, acc, 1, fi, i, 1);
  i = nc;
  if (bc5)
    FO(i) = CO(5, 1, ibc[1]) * (FI(i + 1) - FI(i)) + CO(5, 2, ibc[1]) * (G1 - FI(i - 1));
  else if (ibc[1] == PER)
    FO(i) = CO(5, 1, PER) * (G1 - FI(i)) + CO(5, 2, PER) * (G2 - FI(i - 1));
  else
    FO(i) = intrpl6(fam, acc, 5, fi, i + 1, -1);
  i = nc - 1;
  if (ibc[1] == PER)
    FO(i) = CO(4, 1, PER) * (FI(i + 1) - FI(i)) + CO(4, 2, PER) * (G1 - FI(i - 1));
  else
    FO(i) = CO(4, 1, PER) * (FI(i + 1) - FI(i)) + CO(4, 2, PER) * (FI(i + 2) - FI(i - 1));
  for (int k = 1; k <= nc; ++k) FO(k) = FO(k) * dd;
}

/* ---------------------------------------------------------------------- */
/* 1-D line drivers, :2739-3933                                            */
/* ---------------------------------------------------------------------- */
static const int op_family[6] = {M1C2P, M1P2C, D1C2C, D1P2P, D1C2P, D1P2C};
/* input located at cell centres? (selects ghost spacing set) */
static const int op_in_centre[6] = {1, 0, 1, 0, 1, 0};
/* output located at cell centres? (selects the y metric) */
static const int op_out_centre[6] = {0, 1, 1, 0, 0, 1};

/* BC kinds that are reduced to IBC_INTRPL when no fbc is passed
 * (reduce_bc_to_interp call sites of each 1-D driver) */
static int reduces(int op, int bc) {
  if (bc == INT_ || bc == NEU) return 1;
  if (bc == DIR && (op == ORC_MIDP_C2P || op == ORC_1DER_C2C || op == ORC_1DER_C2P))
    return 1;
  return 0;
}

void orc_op1d(const orc_dom *dm, int dir, int op, int iacc, const int ibc0[2],
              const double *fi, int nin, double *fo, int nout,
              const double *fbc) {
  ensure_tables();
  int ibc[2] = {ibc0[0], ibc0[1]};
  if (!fbc)
    for (int s = 0; s < 2; ++s)
      if (reduces(op, ibc[s])) ibc[s] = ITP;
  double dp[4];
  if (dir == 1) {
    const int nc = dm->nc[1], np = dm->np[1];
    const double *yp = dm->yp, *yc = dm->yc;
    if (op_in_centre[op]) {
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
    const double h = dm->h[dir];
    if (op_in_centre[op]) {
      dp[0] = h; dp[2] = h * 2.0; dp[1] = h; dp[3] = h * 2.0;
    } else {
      dp[0] = h * 2.0; dp[2] = h * 4.0; dp[1] = h * 2.0; dp[3] = h * 4.0;
    }
  }
  const double dd = dm->h1r[dir];
  const int acc = iacc;
  switch (op) {
  case ORC_MIDP_C2P: rhs_midp_c2p(fi, nin, fo, nout, acc, ibc, fbc, dp); break;
  case ORC_MIDP_P2C: rhs_midp_p2c(fi, nin, fo, nout, acc, ibc, fbc, dp); break;
  case ORC_1DER_C2C: rhs_1der_c2c(fi, nin, fo, nout, acc, dd, ibc, fbc, dp); break;
  case ORC_1DER_P2P: rhs_1der_p2p(fi, nin, fo, nout, acc, dd, ibc, fbc, dp); break;
  case ORC_1DER_C2P: rhs_1der_c2p(fi, nin, fo, nout, acc, dd, ibc, fbc, dp); break;
  case ORC_1DER_P2C: rhs_1der_p2c(fi, nin, fo, nout, acc, dd, ibc, fbc, dp); break;
  }
  if (iacc == ORC_CP4 || iacc == ORC_CP6) {
    double *a = (double *)malloc(sizeof(double) * 4 * (size_t)nout);
    double *b = a + nout, *c = b + nout, *d = c + nout;
    orc_lhs(op_family[op], nout, ibc[0], ibc[1], iacc, a, b, c, d);
    orc_tdma_solve(ibc[0] == PER, fo, a, b, c, d, nout);
    free(a);
  }
  if (dir == 1 && op >= ORC_1DER_C2C) {
    if (dm->stretching) {
      const double *m = op_out_centre[op] ? dm->ymc : dm->ymp;
      for (int i = 0; i < nout; ++i) fo[i] = fo[i] * m[i];
    }
    if (op == ORC_1DER_P2P || op == ORC_1DER_C2P) {
      if (ibc[0] == NEU && fbc) fo[0] = fbc[0];
      if (ibc[1] == NEU && fbc) fo[nout - 1] = fbc[1];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* 3-D drivers, :3952-4790                                                 */
/* ---------------------------------------------------------------------- */
int orc_op3d(const orc_dom *dm, int dir, int op, int iacc, const int ibc[2],
             const double *fi3, const int si[3], double *fo3, const int so[3],
             const double *fbc) {
  for (int d = 0; d < 3; ++d)
    if (d != dir && si[d] != so[d]) return -1;
  const int nin = si[dir], nout = so[dir];
  const int n1 = si[0], n2 = si[1], n3 = si[2];
  const int m1 = so[0], m2 = so[1];
  long nlines = (dir == 0) ? (long)n2 * n3 : (dir == 1) ? (long)n1 * n3 : (long)n1 * n2;
#pragma omp parallel
  {
    double *fi = (double *)malloc(sizeof(double) * (size_t)nin);
    double *fo = (double *)malloc(sizeof(double) * (size_t)nout);
    double fb[4];
#pragma omp for schedule(static)
    for (long L = 0; L < nlines; ++L) {
      long ib, ob, istr, ostr;
      long fb_base = 0, fb_str = 0;
      if (dir == 0) {
        long j = L % n2, k = L / n2;
        ib = (j + (long)n2 * k) * n1; istr = 1;
        ob = (j + (long)m2 * k) * m1; ostr = 1;
        fb_base = 4 * (j + (long)n2 * k); fb_str = 1;
      } else if (dir == 1) {
        long i = L % n1, k = L / n1;
        ib = i + (long)n1 * n2 * k; istr = n1;
        ob = i + (long)m1 * m2 * k; ostr = m1;
        fb_base = i + (long)n1 * 4 * k; fb_str = n1;
      } else {
        long i = L % n1, j = L / n1;
        ib = i + (long)n1 * j; istr = (long)n1 * n2;
        ob = i + (long)m1 * j; ostr = (long)m1 * m2;
        fb_base = i + (long)n1 * j; fb_str = (long)n1 * n2;
      }
      for (int t = 0; t < nin; ++t) fi[t] = fi3[ib + istr * t];
      if (fbc)
        for (int t = 0; t < 4; ++t) fb[t] = fbc[fb_base + fb_str * t];
      orc_op1d(dm, dir, op, iacc, ibc, fi, nin, fo, nout, fbc ? fb : NULL);
      for (int t = 0; t < nout; ++t) fo3[ob + ostr * t] = fo[t];
    }
    free(fi);
    free(fo);
  }
  (void)n3;
  return 0;
}
