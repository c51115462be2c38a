/*
 * Oracle part 3: thermal-property table path (TEST INFRASTRUCTURE ONLY).
 *
 * Follows:
 *   table build   src/input_thermo.f90:617-693 (read, sort by T, reference
 *                 state from T, non-dimensionalisation, rho*h, dhmin/dhmax)
 *   sort          src/input_thermo.f90:512-556 (selection by minimum T)
 *   from T        src/input_thermo.f90:121-154 (bisection on T, linear weights;
 *                 reference state) and :213-241 (the same on the
 *                 non-dimensional table: ftp_refresh_thermal_properties_
 *                 from_T_undim, wall properties of isothermal walls)
 *   from H        src/input_thermo.f90:345-374 (bisection on h)
 *   from rho*h    src/input_thermo.f90:393-446 (clamp to [dhmin, dhmax],
 *                 bisection on rho*h, h interpolated, properties from h,
 *                 h = rho*h / d, properties from h again)
 *
 * Pinning (tests/test_oracle_props.py): table rows are reproduced exactly at
 * the table's own h / rho*h values; linear data is interpolated exactly;
 * the reference state equals the table row at T0 when T0 is a table
 * temperature.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chs_oracle.h"

static const double MINP = 1.0e-20; /* src/modules.f90:141 */

/* bisection of src/input_thermo.f90:134-143 on column `col` for value v */
static void bisect(const double *col, int n, double v, int *i1o, int *i2o) {
  int i1 = 0, i2 = n - 1;
  while (i2 - i1 > 1) {
    int im = i1 + (i2 - i1) / 2;
    double d1 = col[i1] - v, dm = col[im] - v;
    if (d1 * dm > MINP) i1 = im;
    else i2 = im;
  }
  *i1o = i1;
  *i2o = i2;
}

int orc_props_build(const double *raw, int nlist, double ref_t0, orc_props *p) {
  if (nlist < 2 || nlist > ORC_PROPS_MAX) return -1;
  memset(p, 0, sizeof(*p));
  p->n = nlist;
  /* columns P H T D M K CP B; rhoh = d*h */
  for (int i = 0; i < nlist; ++i) {
    p->h[i] = raw[8 * i + 1];
    p->t[i] = raw[8 * i + 2];
    p->d[i] = raw[8 * i + 3];
    p->m[i] = raw[8 * i + 4];
    p->k[i] = raw[8 * i + 5];
    p->cp[i] = raw[8 * i + 6];
    p->b[i] = raw[8 * i + 7];
    p->rhoh[i] = p->d[i] * p->h[i];
  }
  /* selection sort by t */
  double *cols[8] = {p->t, p->d, p->m, p->k, p->b, p->cp, p->h, p->rhoh};
  for (int i = 0; i < nlist; ++i) {
    int kmin = i;
    for (int j = i + 1; j < nlist; ++j)
      if (p->t[j] < p->t[kmin]) kmin = j;
    for (int q = 0; q < 8; ++q) {
      double buf = cols[q][i];
      cols[q][i] = cols[q][kmin];
      cols[q][kmin] = buf;
    }
  }
  p->dimensional = ref_t0 > 0.0;
  if (!p->dimensional) {
    /* rows already non-dimensional: no reference state, no scaling */
    int imn = 0, imx = 0;
    for (int i = 1; i < nlist; ++i) {
      if (p->rhoh[i] < p->rhoh[imn]) imn = i;
      if (p->rhoh[i] > p->rhoh[imx]) imx = i;
    }
    p->dhmin = p->rhoh[imn];
    p->dhmax = p->rhoh[imx];
    return 0;
  }
  /* reference state from T0 (dimensional table) */
  int i1, i2;
  bisect(p->t, nlist, ref_t0, &i1, &i2);
  double w1 = (p->t[i2] - ref_t0) / (p->t[i2] - p->t[i1]), w2 = 1.0 - w1;
  p->ref[0] = ref_t0;
  p->ref[1] = w1 * p->d[i1] + w2 * p->d[i2];
  p->ref[2] = w1 * p->m[i1] + w2 * p->m[i2];
  p->ref[3] = w1 * p->k[i1] + w2 * p->k[i2];
  p->ref[4] = w1 * p->h[i1] + w2 * p->h[i2];
  p->ref[5] = w1 * p->b[i1] + w2 * p->b[i2];
  p->ref[6] = w1 * p->cp[i1] + w2 * p->cp[i2];
  /* non-dimensionalise */
  for (int i = 0; i < nlist; ++i) {
    p->t[i] = p->t[i] / p->ref[0];
    p->d[i] = p->d[i] / p->ref[1];
    p->m[i] = p->m[i] / p->ref[2];
    p->k[i] = p->k[i] / p->ref[3];
    p->b[i] = p->b[i] / p->ref[5];
    p->cp[i] = p->cp[i] / p->ref[6];
    p->h[i] = (p->h[i] - p->ref[4]) / p->ref[0] / p->ref[6];
    p->rhoh[i] = p->d[i] * p->h[i];
  }
  int imin = 0, imax = 0;
  for (int i = 1; i < nlist; ++i) {
    if (p->rhoh[i] < p->rhoh[imin]) imin = i;
    if (p->rhoh[i] > p->rhoh[imax]) imax = i;
  }
  p->dhmin = p->rhoh[imin];
  p->dhmax = p->rhoh[imax];
  return 0;
}

/* ftp_refresh_thermal_properties_from_H: out = d, m, k, t, b, cp */
void orc_props_from_h(const orc_props *p, double h, double out[6]) {
  int i1, i2;
  bisect(p->h, p->n, h, &i1, &i2);
  double w1 = (p->h[i2] - h) / (p->h[i2] - p->h[i1]), w2 = 1.0 - w1;
  out[0] = w1 * p->d[i1] + w2 * p->d[i2];
  out[1] = w1 * p->m[i1] + w2 * p->m[i2];
  out[2] = w1 * p->k[i1] + w2 * p->k[i2];
  out[3] = w1 * p->t[i1] + w2 * p->t[i2];
  out[4] = w1 * p->b[i1] + w2 * p->b[i2];
  out[5] = w1 * p->cp[i1] + w2 * p->cp[i2];
}

/* ftp_refresh_thermal_properties_from_DH (table state): out = h, d, m, k,
 * t, b, cp; returns 1 when rhoh was outside [dhmin, dhmax] (clamped) */
int orc_props_from_dh(const orc_props *p, double rhoh, double out[7]) {
  int clamped = 0;
  if (rhoh < p->dhmin) { rhoh = p->dhmin; clamped = 1; }
  else if (rhoh > p->dhmax) { rhoh = p->dhmax; clamped = 1; }
  int i1, i2;
  bisect(p->rhoh, p->n, rhoh, &i1, &i2);
  double w1 = (p->rhoh[i2] - rhoh) / (p->rhoh[i2] - p->rhoh[i1]), w2 = 1.0 - w1;
  double h = w1 * p->h[i1] + w2 * p->h[i2];
  double pr[6];
  orc_props_from_h(p, h, pr);
  h = rhoh / pr[0];
  orc_props_from_h(p, h, pr);
  out[0] = h;
  for (int q = 0; q < 6; ++q) out[1 + q] = pr[q];
  return clamped;
}

/* ftp_refresh_thermal_properties_from_T_undim, table branch
 * (src/input_thermo.f90:213-241): out = h, d, m, k, t, b, cp */
void orc_props_from_t(const orc_props *p, double t, double out[7]) {
  int i1, i2;
  bisect(p->t, p->n, t, &i1, &i2);
  double w1 = (p->t[i2] - t) / (p->t[i2] - p->t[i1]), w2 = 1.0 - w1;
  out[1] = w1 * p->d[i1] + w2 * p->d[i2];
  out[2] = w1 * p->m[i1] + w2 * p->m[i2];
  out[3] = w1 * p->k[i1] + w2 * p->k[i2];
  out[0] = w1 * p->h[i1] + w2 * p->h[i2];
  out[5] = w1 * p->b[i1] + w2 * p->b[i2];
  out[6] = w1 * p->cp[i1] + w2 * p->cp[i2];
  out[4] = t;
}
