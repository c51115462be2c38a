/*
 * Oracle part 2: mesh, boundary tables, FFTPACK-convention transforms, the
 * FISHPACK FFT(x,z)+TDMA(y) Poisson solve, the 3-D FFT Poisson solve for the
 * triply periodic case, and the momentum RHS / RK3 / projection sub-step.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Follows:
 *   mesh        src/geometry.f90:151-269, 410-593
 *   bc tables   src/bc_general.f90:137-245, 455-757
 *   RFFTF/RFFTB lib/fishpack4.1/fftpack.f:129-240 (transform definitions)
 *   Poisson     src/poisson_fishpack.f90:20-57 (TRID0), 60-157, 213-350, 354-490
 *   3-D Poisson src/poisson_1stderivcomp.f90:1016-1195 (poisson_000), 2343-2541
 *   momentum    src/eq_momentum2.f90:35-71, 101-1571, 1577-1718, 1799-1932
 *   divergence  src/eq_continuity.f90:91-161, 337-397
 *   bulk sums   src/tools_general.f90:1202-1268
 *   Re          src/tools_solver.f90:30-47
 *   RK3         src/input_general.f90:842-888
 *
 * Pinning (tests/test_oracle_pins.py, test_oracle_fft_poisson.py,
 * test_oracle_diagnostics.py):
 *   mesh      - closed form of the tanh map; yMapping against the cosh form
 *               of dy/deta and a finite difference of the nodes.
 *   RFFTF/B   - numpy.fft on random data (the FFTPACK definitions).
 *   Poisson   - the assembled 2nd-order operator applied to the solution
 *               reproduces the RHS (dense check on small grids) and the
 *               singular-pivot gauge; the 3-D path divides a Fourier mode by
 *               the published CP6 modified wavenumbers.
 *   exact projection - zero divergence after one projection, idempotence,
 *               equality with FISHPACK for CD2.
 *   momentum RHS - manufactured fields on the stretched channel: every
 *               component converges to the continuous stress-form right-hand
 *               side (first order at the wall rows, the x/z design order in
 *               the core), per field and all together; y-mirror symmetry;
 *               x/z translation invariance; plane Couette flow with a moving
 *               wall is an exact steady state.
 *   RK3, initial step - the 2-D Taylor-Green mode (in each coordinate
 *               plane) is multiplied by the third-order stability polynomial
 *               1 + z + z^2/2 + z^3/6 per step, z = -2 nu (T/h)^2 dt, to
 *               round-off; the initial step by 1 + z; the accumulated
 *               pressure balances the convective term.
 *   forcing   - constant-flux forcing keeps the volume-weighted bulk
 *               velocity to round-off.
 *   walls     - the wall-normal velocity on the wall nodes equals the wall
 *               values after a sub-step.
 *   diagnostics - CFL, bulk values and mass-check windows by closed forms.
 */
#include "chs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXP 1.0e20
static const double ONE_THIRD = 1.0 / 3.0;

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* mesh: src/geometry.f90:151-269 (tanh), 410-593                          */
/* ---------------------------------------------------------------------- */
static void map_tanh(const orc_dom *dm, int is_node, int n, double *y,
                     double *mp) {
  double eta_shift, eta_delta;
  if (is_node) {
    eta_shift = 0.0;
    eta_delta = 1.0 / (double)(n - 1);
  } else {
    eta_shift = 1.0 / ((double)n) * 0.5;
    eta_delta = 1.0 / (double)n;
  }
  double gamma = 1.0, delta = 0.0, ymin = 0.0, ymax = 0.0;
  if (dm->istret == 2) { gamma = 0.5; delta = 0.5; ymin = -1.0; ymax = 1.0; }
  else if (dm->istret == 3) { gamma = -1.0; delta = 1.0; ymin = 0.0; ymax = 1.0; }
  else if (dm->istret == 4) { gamma = 1.0; delta = 0.0; ymin = 0.0; ymax = 1.0; }
  double beta = dm->rstret * 20.0;
  double mm = tanh(beta * gamma);
  double eta1 = 0.0 + eta_shift;
  for (int j = 0; j < n; ++j) {
    double eta = (j == 0) ? eta1 : eta1 + (double)j * eta_delta;
    double yy = tanh(beta * (eta - delta)) / mm;
    double m = 1.0 - yy * yy * mm * mm;
    if (m < 1.0e-20 && m > -1.0e-20)
      m = 1.0;
    else
      m = mm / beta / m;
    double ff = (dm->lyt - dm->lyb) / (ymax - ymin);
    y[j] = (yy - ymin) * ff + dm->lyb;
    mp[j] = m / ff;
  }
}

int orc_dom_init(orc_dom *dm, const int nc[3], const int periodic[3],
                 double lxx, double lyb, double lyt, double lzz, int istret,
                 double rstret) {
  memset(dm, 0, sizeof(*dm));
  for (int d = 0; d < 3; ++d) {
    dm->nc[d] = nc[d];
    dm->periodic[d] = periodic[d];
    dm->npgeo[d] = nc[d] + 1;
    dm->np[d] = periodic[d] ? nc[d] : nc[d] + 1;
  }
  dm->lxx = lxx; dm->lyb = lyb; dm->lyt = lyt; dm->lzz = lzz;
  dm->istret = istret; dm->rstret = rstret;
  dm->stretching = (istret != 0);
  dm->h[0] = lxx / (double)nc[0];
  dm->h[2] = lzz / (double)nc[2];
  dm->h[1] = (lyt - lyb) / (double)nc[1];
  int npg = dm->npgeo[1], ncy = nc[1];
  dm->yp = (double *)calloc((size_t)npg, sizeof(double));
  dm->yc = (double *)calloc((size_t)ncy, sizeof(double));
  dm->ymp = (double *)malloc(sizeof(double) * (size_t)npg);
  dm->ymc = (double *)malloc(sizeof(double) * (size_t)ncy);
  for (int j = 0; j < npg; ++j) dm->ymp[j] = 1.0;
  for (int j = 0; j < ncy; ++j) dm->ymc[j] = 1.0;
  if (dm->stretching) {
    dm->h[1] = 1.0 / (double)ncy;
    map_tanh(dm, 1, npg, dm->yp, dm->ymp);
    map_tanh(dm, 0, ncy, dm->yc, dm->ymc);
  } else {
    for (int j = 0; j < npg; ++j) dm->yp[j] = (double)j * dm->h[1] + lyb;
    for (int j = 0; j < ncy; ++j)
      dm->yc[j] = (double)j * dm->h[1] + dm->h[1] * 0.5 + lyb;
  }
  for (int d = 0; d < 3; ++d) {
    dm->h2r[d] = 1.0 / (dm->h[d] * dm->h[d]);
    dm->h1r[d] = 1.0 / dm->h[d];
  }
  return 0;
}

void orc_dom_free(orc_dom *dm) {
  free(dm->yp); free(dm->yc); free(dm->ymp); free(dm->ymc);
  dm->yp = dm->yc = dm->ymp = dm->ymc = NULL;
}

/* ---------------------------------------------------------------------- */
/* complex FFT (mixed radix 2/3/5, O(n^2) DFT for other factors)           */
/* ---------------------------------------------------------------------- */
typedef struct { double re, im; } cpx;

static void fft_rec(cpx *x, int n, int stride, cpx *out, int sign, cpx *tmp) {
  if (n == 1) { out[0] = x[0]; return; }
  int p = (n % 4 == 0) ? 2 : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
  int m = n / p;
  if (p == n) {
    for (int k = 0; k < n; ++k) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < n; ++j) {
        double ang = sign * 2.0 * M_PI * (double)((long)j * k % n) / (double)n;
        double c = cos(ang), s = sin(ang);
        cpx v = x[(long)j * stride];
        sr += v.re * c - v.im * s;
        si += v.re * s + v.im * c;
      }
      out[k].re = sr; out[k].im = si;
    }
    return;
  }
  for (int r = 0; r < p; ++r) fft_rec(x + (long)r * stride, m, stride * p, tmp + (long)r * m, sign, out);
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; ++r) {
        long e = ((long)r * (k + (long)q * m)) % n;
        double ang = sign * 2.0 * M_PI * (double)e / (double)n;
        double c = cos(ang), s = sin(ang);
        cpx v = tmp[(long)r * m + k];
        sr += v.re * c - v.im * s;
        si += v.re * s + v.im * c;
      }
      out[k + (long)q * m].re = sr;
      out[k + (long)q * m].im = si;
    }
  }
}

static void cfft(cpx *x, int n, int sign) {
  cpx *out = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  fft_rec(x, n, 1, out, sign, tmp);
  memcpy(x, out, sizeof(cpx) * (size_t)n);
  free(out);
  free(tmp);
}

/* RFFTF: r(1)=sum, r(2k)=sum r cos, r(2k+1)=-sum r sin, r(n)=alternating */
void orc_rfftf(int n, double *r) {
  cpx *x = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int j = 0; j < n; ++j) { x[j].re = r[j]; x[j].im = 0.0; }
  cfft(x, n, -1);
  r[0] = x[0].re;
  for (int k = 1; 2 * k < n; ++k) {
    r[2 * k - 1] = x[k].re;
    r[2 * k] = x[k].im;
  }
  if (n % 2 == 0) r[n - 1] = x[n / 2].re;
  free(x);
}

void orc_rfftb(int n, double *r) {
  cpx *x = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  x[0].re = r[0]; x[0].im = 0.0;
  for (int k = 1; 2 * k < n; ++k) {
    x[k].re = r[2 * k - 1]; x[k].im = r[2 * k];
    x[n - k].re = r[2 * k - 1]; x[n - k].im = -r[2 * k];
  }
  if (n % 2 == 0) { x[n / 2].re = r[n - 1]; x[n / 2].im = 0.0; }
  cfft(x, n, +1);
  for (int j = 0; j < n; ++j) r[j] = x[j].re;
  free(x);
}

/* ---------------------------------------------------------------------- */
/* flow state                                                             */
/* ---------------------------------------------------------------------- */
enum { VU = 0, VV = 1, VW = 2, VP = 3, VT = 4 };
enum { BINT = 0, BPER = 1, BSYM = 2, BASY = 3, BDIR = 4, BNEU = 5 };

struct orc_flow {
  orc_flow_cfg cfg;
  orc_dom dm;
  int ibc[3][5][2]; /* [dir][var][side] calculation BCs */
  int ftp[3][2];
  double fconst[3][5][2];
  int mcov[3][3][2]; /* [mom component][dir][side] */
  int mtau[3][3][2];
  int freq[3];
  double rre, dt, sigma2p;
  double tg[4], tz[4], ta[4];
  int nx, ny, nz, px, py, pz;
  double *qx, *qy, *qz, *pres, *pcor, *mx0, *my0, *mz0;
  double *fbx[4], *fby[4], *fbz[4]; /* constant-bc arrays for u,v,w,p */
  /* Poisson */
  double *xrt, *zrt, *pa, *pb, *pc;
  double *kx2, *ky2, *kz2; /* modified wavenumbers for the 3-D path */
  /* exact projection: wall-normal block D_y G_y (row-major), eigenvalues of
   * the periodic blocks per FFTPACK coefficient, left null vector of ly */
  double *ly, *lxe, *lze, *lnull;
  /* variable-property (thermo) state: src/eq_energy.f90, src/para_conversion.f90 */
  int thermo;
  const orc_props *props;
  double wall[2][7];   /* h, d, m, k, t, b, cp of the isothermal y walls */
  double prandtl, rprren;
  double *gx, *gy, *gz;                       /* conservative variables rho u */
  double *dens, *visc, *rhoh, *hent, *temp, *kcond, *erhs0, *densm1, *densm2;
  double *fby_ftp[4];  /* (nx,4,nz) wall arrays of h, k, m, d (rows 3,4 = rows 1,2) */
};

/* build_bc_symm_operation, src/bc_general.f90:455-510; returns JBC_PROD */
static void bc_symm(const int ibc[2], const int *jbc, int mbc[2][3]) {
  for (int i = 0; i < 2; ++i) {
    mbc[i][0] = mbc[i][1] = mbc[i][2] = ibc[i];
    if (jbc) {
      if (ibc[i] == BSYM && jbc[i] == BSYM) mbc[i][2] = BSYM;
      else if (ibc[i] == BSYM && jbc[i] == BASY) mbc[i][2] = BASY;
      else if (ibc[i] == BASY && jbc[i] == BSYM) mbc[i][2] = BASY;
      else if (ibc[i] == BASY && jbc[i] == BASY) mbc[i][2] = BSYM;
      else if (ibc[i] != jbc[i]) {
        if (ibc[i] == BDIR) mbc[i][0] = mbc[i][1] = mbc[i][2] = ibc[i];
        if (jbc[i] == BDIR) mbc[i][0] = mbc[i][1] = mbc[i][2] = jbc[i];
      } else {
        mbc[i][0] = mbc[i][1] = mbc[i][2] = ibc[i];
      }
    } else {
      if (ibc[i] == BSYM) { mbc[i][1] = BASY; }
      else if (ibc[i] == BASY) { mbc[i][1] = BSYM; mbc[i][2] = BSYM; }
    }
  }
}

/* conv: prod(ibc_a, ibc_b); tau: prod(ftp, grad(ibc_q)) */
static void bc_prod(const int a[2], const int b[2], int out[2]) {
  int m[2][3];
  bc_symm(a, b, m);
  out[0] = m[0][2]; out[1] = m[1][2];
}
static void bc_tau(const int ftp[2], const int q[2], int out[2]) {
  int m[2][3], g[2];
  bc_symm(q, NULL, m);
  g[0] = m[0][1]; g[1] = m[1][1];
  bc_symm(ftp, g, m);
  out[0] = m[0][2]; out[1] = m[1][2];
}

static double *fbc_const_array(int dir, int n1, int n2, int n3,
                               const double v[2]) {
  /* layouts (4,n2,n3) / (n1,4,n3) / (n1,n2,4) */
  long sz = (dir == 0) ? 4L * n2 * n3 : (dir == 1) ? (long)n1 * 4 * n3 : (long)n1 * n2 * 4;
  double *a = (double *)malloc(sizeof(double) * (size_t)sz);
  for (long t = 0; t < sz; ++t) {
    long m;
    if (dir == 0) m = t % 4;
    else if (dir == 1) m = (t / n1) % 4;
    else m = t / ((long)n1 * n2);
    a[t] = v[m % 2];
  }
  return a;
}

/* ---------------------------------------------------------------------- */
/* Exact projection (no counterpart in the reference).  The composed        */
/* operator L = D_x G_x + D_y G_y + D_z G_z of the divergence (D1_P2C of     */
/* the velocity, src/eq_continuity.f90:91-161) and of the correction        */
/* gradient (D1_C2P of the pressure increment, src/eq_momentum2.f90         */
/* :1577-1660, wall faces then reset to the wall velocity) is built by       */
/* applying the oracle's own line operators to unit vectors: a circulant     */
/* first column in x and z (diagonal in the FFTPACK real Fourier basis),     */
/* a dense ny x ny block in y.  Each real Fourier coefficient line is solved */
/* by LU with partial pivoting of ly + (lx + lz) I; the (0,0) line replaces  */
/* its last equation by phi(ny-1) = 0 and is then shifted so that the left  */
/* null vector of ly is orthogonal to it (the gauge).                       */
/* ---------------------------------------------------------------------- */
static int dense_solve(int n, double *M, double *x) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(M[(size_t)i * n + k]) > fabs(M[(size_t)p * n + k])) p = i;
    if (M[(size_t)p * n + k] == 0.0) return -1;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        double t = M[(size_t)p * n + j];
        M[(size_t)p * n + j] = M[(size_t)k * n + j];
        M[(size_t)k * n + j] = t;
      }
      double t = x[p]; x[p] = x[k]; x[k] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      double fct = M[(size_t)i * n + k] / M[(size_t)k * n + k];
      if (fct == 0.0) continue;
      for (int j = k + 1; j < n; ++j) M[(size_t)i * n + j] -= fct * M[(size_t)k * n + j];
      x[i] -= fct * x[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double sum = x[i];
    for (int j = i + 1; j < n; ++j) sum -= M[(size_t)i * n + j] * x[j];
    x[i] = sum / M[(size_t)i * n + i];
  }
  return 0;
}

static void exact_setup(orc_flow *f) {
  const orc_dom *dm = &f->dm;
  int nx = f->nx, ny = f->ny, nz = f->nz, acc = f->cfg.iacc;
  double pi = 2.0 * asin(1.0);
  f->ly = (double *)calloc((size_t)ny * ny, sizeof(double));
  f->lnull = (double *)calloc((size_t)ny, sizeof(double));
  double *e = (double *)calloc((size_t)(ny + 1), sizeof(double));
  double *g = (double *)calloc((size_t)(ny + 1), sizeof(double));
  double *col = (double *)calloc((size_t)(ny + 1), sizeof(double));
  double fb0[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < ny; ++k) {
    for (int j = 0; j < ny; ++j) e[j] = (j == k) ? 1.0 : 0.0;
    orc_op1d(dm, 1, ORC_1DER_C2P, acc, f->ibc[1][VP], e, ny, g, ny + 1, fb0);
    g[0] = 0.0;
    g[ny] = 0.0;
    orc_op1d(dm, 1, ORC_1DER_P2C, acc, f->ibc[1][VV], g, ny + 1, col, ny, NULL);
    for (int i = 0; i < ny; ++i) f->ly[(size_t)i * ny + k] = col[i];
  }
  /* left null vector: ly^T l = 0 with l(ny-1) = 1 */
  double *M = (double *)malloc(sizeof(double) * (size_t)ny * ny);
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < ny; ++j) M[(size_t)i * ny + j] = f->ly[(size_t)j * ny + i];
  for (int j = 0; j < ny; ++j) {
    M[(size_t)(ny - 1) * ny + j] = (j == ny - 1) ? 1.0 : 0.0;
    f->lnull[j] = (j == ny - 1) ? 1.0 : 0.0;
  }
  dense_solve(ny, M, f->lnull);
  free(M);
  for (int pass = 0; pass < 2; ++pass) {
    int dir = pass ? 2 : 0, n = pass ? nz : nx;
    double *e0 = (double *)calloc((size_t)n, sizeof(double));
    double *t = (double *)calloc((size_t)n, sizeof(double));
    double *c = (double *)calloc((size_t)n, sizeof(double));
    double *lv = (double *)calloc((size_t)n, sizeof(double));
    e0[0] = 1.0;
    orc_op1d(dm, dir, ORC_1DER_C2P, acc, f->ibc[dir][VP], e0, n, t, n, NULL);
    orc_op1d(dm, dir, ORC_1DER_P2C, acc, f->ibc[dir][dir], t, n, c, n, NULL);
    for (int i = 0; i < n; ++i) {
      int m = (i + 1) / 2; /* FFTPACK coefficient i holds wavenumber (i+1)/2 */
      double sum = 0.0;
      for (int j = 0; j < n; ++j) sum += c[j] * cos(2.0 * pi * (double)(((long)j * m) % n) / (double)n);
      lv[i] = sum;
    }
    if (pass) f->lze = lv; else f->lxe = lv;
    free(e0); free(t); free(c);
  }
  free(e); free(g); free(col);
}

orc_flow *orc_flow_create(const orc_flow_cfg *cfg) {
  orc_flow *f = (orc_flow *)calloc(1, sizeof(orc_flow));
  f->cfg = *cfg;
  int per[3] = {1, cfg->wall_y ? 0 : 1, 1};
  orc_dom_init(&f->dm, cfg->nc, per, cfg->lxx, cfg->lyb, cfg->lyt, cfg->lzz,
               cfg->istret, cfg->rstret);
  for (int d = 0; d < 3; ++d)
    for (int v = 0; v < 5; ++v)
      for (int s = 0; s < 2; ++s) {
        f->ibc[d][v][s] = BPER;
        f->fconst[d][v][s] = 0.0;
      }
  if (cfg->wall_y) {
    for (int s = 0; s < 2; ++s) {
      for (int v = 0; v < 3; ++v) {
        f->ibc[1][v][s] = BDIR;
        f->fconst[1][v][s] = cfg->uwall[s][v];
      }
      f->ibc[1][VT][s] = BDIR;
      /* config_calc_basic_ibc: pressure Neumann where qy is Dirichlet */
      f->ibc[1][VP][s] = BNEU;
      f->fconst[1][VP][s] = 0.0;
    }
  }
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      f->ftp[d][s] = (f->ibc[d][VT][s] == BNEU) ? BDIR : f->ibc[d][VT][s];
  /* config_calc_eqs_ibc, src/bc_general.f90:512-757 (Cartesian) */
  {
    int (*B)[5][2] = f->ibc;
    bc_prod(B[0][VU], B[0][VU], f->mcov[0][0]);
    bc_prod(B[1][VV], B[1][VU], f->mcov[0][1]);
    bc_prod(B[2][VW], B[2][VU], f->mcov[0][2]);
    bc_tau(f->ftp[0], B[0][VU], f->mtau[0][0]);
    bc_tau(f->ftp[1], B[1][VU], f->mtau[0][1]);
    bc_tau(f->ftp[2], B[2][VU], f->mtau[0][2]);
    bc_prod(B[0][VU], B[0][VV], f->mcov[1][0]);
    bc_prod(B[1][VV], B[1][VV], f->mcov[1][1]);
    bc_prod(B[2][VW], B[2][VV], f->mcov[1][2]);
    bc_tau(f->ftp[0], B[0][VV], f->mtau[1][0]);
    bc_tau(f->ftp[1], B[1][VV], f->mtau[1][1]);
    bc_tau(f->ftp[2], B[2][VV], f->mtau[1][2]);
    bc_prod(B[0][VU], B[0][VW], f->mcov[2][0]);
    bc_prod(B[1][VV], B[1][VW], f->mcov[2][1]);
    bc_prod(B[2][VW], B[2][VW], f->mcov[2][2]);
    bc_tau(f->ftp[0], B[0][VW], f->mtau[2][0]);
    bc_tau(f->ftp[1], B[1][VW], f->mtau[2][1]);
    bc_tau(f->ftp[2], B[2][VW], f->mtau[2][2]);
    for (int d = 0; d < 3; ++d) {
      f->freq[d] = 0;
      for (int v = 0; v < 3; ++v)
        for (int s = 0; s < 2; ++s)
          if (B[d][v][s] == BDIR || B[d][v][s] == BINT) f->freq[d] = 1;
    }
  }
  f->sigma2p = 1.0;
  f->dt = cfg->dt;
  f->tg[0] = 1.0; f->tg[1] = 8.0 / 15.0; f->tg[2] = 5.0 / 12.0; f->tg[3] = 3.0 * 0.25;
  f->tz[0] = 0.0; f->tz[1] = 0.0; f->tz[2] = -17.0 / 60.0; f->tz[3] = -5.0 / 12.0;
  for (int i = 0; i < 4; ++i) f->ta[i] = f->tg[i] + f->tz[i];
  const orc_dom *dm = &f->dm;
  f->nx = dm->nc[0]; f->ny = dm->nc[1]; f->nz = dm->nc[2];
  f->px = dm->np[0]; f->py = dm->np[1]; f->pz = dm->np[2];
  long npcc = (long)f->px * f->ny * f->nz, ncpc = (long)f->nx * f->py * f->nz;
  long nccp = (long)f->nx * f->ny * f->pz, nccc = (long)f->nx * f->ny * f->nz;
  f->qx = (double *)calloc((size_t)npcc, sizeof(double));
  f->qy = (double *)calloc((size_t)ncpc, sizeof(double));
  f->qz = (double *)calloc((size_t)nccp, sizeof(double));
  f->pres = (double *)calloc((size_t)nccc, sizeof(double));
  f->pcor = (double *)calloc((size_t)nccc, sizeof(double));
  f->mx0 = (double *)calloc((size_t)npcc, sizeof(double));
  f->my0 = (double *)calloc((size_t)ncpc, sizeof(double));
  f->mz0 = (double *)calloc((size_t)nccp, sizeof(double));
  /* constant fbc arrays (initialise_fbc_flow_given): u on pcc, v on cpc,
   * w on ccp, p on ccc */
  const int sh[4][3] = {{f->px, f->ny, f->nz}, {f->nx, f->py, f->nz},
                        {f->nx, f->ny, f->pz}, {f->nx, f->ny, f->nz}};
  for (int v = 0; v < 4; ++v) {
    f->fbx[v] = fbc_const_array(0, sh[v][0], sh[v][1], sh[v][2], f->fconst[0][v]);
    f->fby[v] = fbc_const_array(1, sh[v][0], sh[v][1], sh[v][2], f->fconst[1][v]);
    f->fbz[v] = fbc_const_array(2, sh[v][0], sh[v][1], sh[v][2], f->fconst[2][v]);
  }
  /* Poisson setup */
  if (cfg->poisson == 2) {
    /* fishpack_fft_init, src/poisson_fishpack.f90:213-350 (x, z periodic) */
    int nx = f->nx, ny = f->ny, nz = f->nz;
    f->xrt = (double *)calloc((size_t)nx, sizeof(double));
    f->zrt = (double *)calloc((size_t)nz, sizeof(double));
    for (int pass = 0; pass < 2; ++pass) {
      int L = pass ? nz : nx;
      double C1 = pass ? dm->h2r[2] : dm->h2r[0];
      double *rt = pass ? f->zrt : f->xrt;
      double pi = 2.0 * asin(1.0);
      double scal = (double)L;
      double dx = pi / (2.0 * scal);
      rt[0] = 0.0;
      rt[L - 1] = -4.0 * C1;
      for (int i = 3; i <= L; i += 2) {
        double s = sin((double)(i - 1) * dx);
        rt[i - 2] = -4.0 * C1 * (s * s);
        rt[i - 1] = rt[i - 2];
      }
    }
    f->pa = (double *)calloc((size_t)ny, sizeof(double));
    f->pb = (double *)calloc((size_t)ny, sizeof(double));
    f->pc = (double *)calloc((size_t)ny, sizeof(double));
    for (int j = 0; j < ny; ++j) {
      f->pa[j] = dm->h2r[1] * dm->ymc[j] * dm->ymp[j] * 1.0 * 1.0;
      f->pc[j] = dm->h2r[1] * dm->ymc[j] * dm->ymp[j + 1] * 1.0 * 1.0;
    }
    for (int j = 0; j < ny; ++j) f->pb[j] = -(f->pa[j] + f->pc[j]);
    if (!dm->periodic[1]) {
      f->pb[0] = f->pb[0] + f->pa[0];
      f->pa[0] = 0.0;
      f->pb[ny - 1] = f->pb[ny - 1] + f->pc[ny - 1];
      f->pc[ny - 1] = 0.0;
    }
  }
  if (cfg->poisson == 3) {
    /* waves, src/poisson_1stderivcomp.f90:2405-2480 (bcx = bcy = bcz = 0,
     * istret = 0), coefficients of :190-200; with y walls (bcy = 1,
     * poisson_010) the y wavenumbers follow :2465-2479 below */
    double pi = 2.0 * asin(1.0);
    double alc = orc_coeff(ORC_FAM_D1C2P, 0, 3, 1, BPER, cfg->iacc);
    double ar = orc_coeff(ORC_FAM_D1C2P, 1, 3, 1, BPER, cfg->iacc);
    double br = orc_coeff(ORC_FAM_D1C2P, 1, 3, 2, BPER, cfg->iacc);
    double len[3] = {dm->lxx, dm->lyt - dm->lyb, dm->lzz};
    double **k2[3] = {&f->kx2, &f->ky2, &f->kz2};
    for (int d = 0; d < 3; ++d) {
      int n = dm->nc[d];
      double h = dm->h[d];
      double ac = ar / h, bc = br / h;
      double *t = (double *)calloc((size_t)n, sizeof(double));
      for (int i = 1; i <= n / 2 + 1; ++i) {
        double w = 2.0 * pi * (double)(i - 1) / (double)n;
        double wp = ac * 2.0 * h * sin(w * 0.5) + bc * 2.0 * h * sin(1.5 * w);
        wp = wp / (1.0 + 2.0 * alc * cos(w));
        double kk = (double)n * wp / len[d];
        t[i - 1] = kk * kk;
      }
      for (int i = n / 2 + 2; i <= n; ++i) t[i - 1] = t[n - i + 1];
      if (d == 1 && !dm->periodic[1]) {
        /* bcy = 1 (:2465-2479): w = pi (j-1) / nym over all ny modes, nym =
         * the y cell count (np_geo(2) - 1, :150-156), yk2(1) = 0 */
        for (int j = 1; j <= n; ++j) {
          double w = pi * (double)(j - 1) / (double)n;
          double wp = ac * 2.0 * h * sin(w * 0.5) + bc * 2.0 * h * sin(1.5 * w);
          wp = wp / (1.0 + 2.0 * alc * cos(w));
          double kk = (double)n * wp / len[d];
          t[j - 1] = kk * kk;
        }
        t[0] = 0.0;
      }
      *k2[d] = t;
    }
  }
  if (cfg->projection == 1) exact_setup(f);
  orc_flow_update_re(f, 0);
  return f;
}

void orc_flow_destroy(orc_flow *f) {
  if (!f) return;
  free(f->qx); free(f->qy); free(f->qz); free(f->pres); free(f->pcor);
  free(f->mx0); free(f->my0); free(f->mz0);
  for (int v = 0; v < 4; ++v) { free(f->fbx[v]); free(f->fby[v]); free(f->fbz[v]); }
  free(f->xrt); free(f->zrt); free(f->pa); free(f->pb); free(f->pc);
  free(f->kx2); free(f->ky2); free(f->kz2);
  free(f->ly); free(f->lxe); free(f->lze); free(f->lnull);
  free(f->gx); free(f->gy); free(f->gz); free(f->dens); free(f->visc); free(f->rhoh);
  free(f->hent); free(f->temp); free(f->kcond); free(f->erhs0); free(f->densm1); free(f->densm2);
  for (int v = 0; v < 4; ++v) free(f->fby_ftp[v]);
  orc_dom_free(&f->dm);
  free(f);
}

const orc_dom *orc_flow_dom(const orc_flow *f) { return &f->dm; }

void orc_flow_set_state(orc_flow *f, const double *qx, const double *qy,
                        const double *qz, const double *pr) {
  memcpy(f->qx, qx, sizeof(double) * (size_t)f->px * f->ny * f->nz);
  memcpy(f->qy, qy, sizeof(double) * (size_t)f->nx * f->py * f->nz);
  memcpy(f->qz, qz, sizeof(double) * (size_t)f->nx * f->ny * f->pz);
  memcpy(f->pres, pr, sizeof(double) * (size_t)f->nx * f->ny * f->nz);
  memset(f->pcor, 0, sizeof(double) * (size_t)f->nx * f->ny * f->nz);
}

void orc_flow_get_state(const orc_flow *f, double *qx, double *qy, double *qz,
                        double *pr) {
  memcpy(qx, f->qx, sizeof(double) * (size_t)f->px * f->ny * f->nz);
  memcpy(qy, f->qy, sizeof(double) * (size_t)f->nx * f->py * f->nz);
  memcpy(qz, f->qz, sizeof(double) * (size_t)f->nx * f->ny * f->pz);
  memcpy(pr, f->pres, sizeof(double) * (size_t)f->nx * f->ny * f->nz);
}

void orc_flow_get_rhs0(const orc_flow *f, double *mx, double *my, double *mz) {
  memcpy(mx, f->mx0, sizeof(double) * (size_t)f->px * f->ny * f->nz);
  memcpy(my, f->my0, sizeof(double) * (size_t)f->nx * f->py * f->nz);
  memcpy(mz, f->mz0, sizeof(double) * (size_t)f->nx * f->ny * f->pz);
}

/* Update_Re, src/tools_solver.f90:30-47 */
void orc_flow_update_re(orc_flow *f, int iter) {
  if (iter < f->cfg.init_re_to)
    f->rre = 1.0 / f->cfg.reninit;
  else
    f->rre = 1.0 / f->cfg.ren;
  /* Update_PrGr, src/tools_solver.f90:63: rPrRen = rre k0 / (m0 cp0) */
  if (f->thermo) f->rprren = f->rre / f->prandtl;
}

/* ---------------------------------------------------------------------- */
/* helpers                                                                */
/* ---------------------------------------------------------------------- */
static double *arr(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

static void op(orc_flow *f, int dir, int kind, const int ibc[2],
               const double *in, int a1, int a2, int a3, double *out, int b1,
               int b2, int b3, const double *fbc) {
  int si[3] = {a1, a2, a3}, so[3] = {b1, b2, b3};
  orc_op3d(&f->dm, dir, kind, f->cfg.iacc, ibc, in, si, out, so, fbc);
}

/* volumetric average, src/tools_general.f90:1202-1268 (x-pencil arrays) */
double orc_vol_average_x(orc_flow *f, const double *var, int n1) {
  const orc_dom *dm = &f->dm;
  double vol = 0.0, fo = 0.0;
  double dx = dm->h[0], dy = dm->h[1], dz = dm->h[2];
  for (int j = 0; j < f->ny; ++j) {
    if (dm->stretching) dy = dm->h[1] / dm->ymc[j];
    for (int k = 0; k < f->nz; ++k)
      for (int i = 0; i < n1; ++i) {
        fo = fo + var[i + (long)n1 * (j + (long)f->ny * k)] * dy * dx * dz;
        vol = vol + dy * dx * dz;
      }
  }
  return fo / vol;
}

/* enforce_velo_from_fbc, src/bc_dirichlet.f90:461-515 (y walls), on qy or,
 * with variable properties, on gy (the wall values of g are the wall
 * velocities times the wall density: zero on the stationary walls the
 * thermo path supports) */
static void enforce_velo(orc_flow *f) {
  double *vy = f->thermo ? f->gy : f->qy;
  if (f->ibc[1][VV][0] == BDIR || f->ibc[1][VV][1] == BDIR) {
    for (int k = 0; k < f->nz; ++k)
      for (int i = 0; i < f->nx; ++i) {
        if (f->ibc[1][VV][0] == BDIR)
          vy[i + (long)f->nx * (0 + (long)f->py * k)] = f->fby[VV][i + (long)f->nx * (0 + 4L * k)];
        if (f->ibc[1][VV][1] == BDIR)
          vy[i + (long)f->nx * ((f->py - 1) + (long)f->py * k)] = f->fby[VV][i + (long)f->nx * (1 + 4L * k)];
      }
  }
}

/* Get_divergence_flow, src/eq_continuity.f90:91-161: of q, or of g = rho u
 * with variable properties (:114-117) */
void orc_divergence(orc_flow *f, double *div) {
  int nx = f->nx, ny = f->ny, nz = f->nz;
  long n = (long)nx * ny * nz;
  const double *ux = f->thermo ? f->gx : f->qx, *uy = f->thermo ? f->gy : f->qy;
  const double *uz = f->thermo ? f->gz : f->qz;
  double *d0 = arr(n);
  for (long t = 0; t < n; ++t) div[t] = 0.0;
  op(f, 0, ORC_1DER_P2C, f->ibc[0][VU], ux, f->px, ny, nz, d0, nx, ny, nz, f->fbx[VU]);
  for (long t = 0; t < n; ++t) div[t] = div[t] + d0[t];
  op(f, 1, ORC_1DER_P2C, f->ibc[1][VV], uy, nx, f->py, nz, d0, nx, ny, nz, NULL);
  for (long t = 0; t < n; ++t) div[t] = div[t] + d0[t];
  op(f, 2, ORC_1DER_P2C, f->ibc[2][VW], uz, nx, ny, f->pz, d0, nx, ny, nz, NULL);
  for (long t = 0; t < n; ++t) div[t] = div[t] + d0[t];
  free(d0);
}

/* Check_element_mass_conservation windows, src/eq_continuity.f90:386-389 */
void orc_mass_check(orc_flow *f, double out[3]) {
  int nx = f->nx, ny = f->ny, nz = f->nz;
  double *div = arr((long)nx * ny * nz);
  orc_divergence(f, div);
  int lo[3] = {1, 5, nx - 5}, hi[3] = {4, nx - 4, nx};
  for (int w = 0; w < 3; ++w) {
    double m = 0.0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = lo[w]; i <= hi[w]; ++i) {
          double v = fabs(div[(i - 1) + (long)nx * (j + (long)ny * k)]);
          if (v > m) m = v;
        }
    out[w] = m;
  }
  free(div);
}

/* Check_cfl_convection and Check_cfl_diffusion, src/tools_solver.f90:220-375
 * (Cartesian, constant properties).  out[0] = max |u h1r(1) dt + v h1r(2)
 * yMappingcc(j) dt + w h1r(3) dt| over cells, with u, v, w the P2C
 * interpolations of qx, qy, qz (accumulated in the reference's order, x
 * then y then z); out[1] = max_j (h2r(1) + h2r(3) + 1 / (yp(j+1) - yp(j))^2)
 * 2 dt rre. */
void orc_cfl(orc_flow *f, double out[2]) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Nccc = (long)nx * ny * nz;
  int (*B)[5][2] = f->ibc;
  const orc_dom *dm = &f->dm;
  const double dt = f->dt;
  double *u = arr(Nccc), *v = arr(Nccc), *w = arr(Nccc);
  op(f, 0, ORC_MIDP_P2C, B[0][VU], f->qx, px, ny, nz, u, nx, ny, nz, f->fbx[VU]);
  op(f, 1, ORC_MIDP_P2C, B[1][VV], f->qy, nx, py, nz, v, nx, ny, nz, f->fby[VV]);
  op(f, 2, ORC_MIDP_P2C, B[2][VW], f->qz, nx, ny, pz, w, nx, ny, nz, f->fbz[VW]);
  double m = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long t = i + (long)nx * (j + (long)ny * k);
        double cx = u[t] * dm->h1r[0] * dt;
        double cy = v[t] * dm->h1r[1] * dt;
        if (dm->stretching) cy = cy * dm->ymc[j];
        double c = (cx + cy) + w[t] * dm->h1r[2] * dt;
        if (fabs(c) > m) m = fabs(c);
      }
  out[0] = m;
  double d = 0.0;
  for (int j = 0; j < ny; ++j) {
    double dyi = 1.0 / (dm->yp[j + 1] - dm->yp[j]);
    double r = dm->h2r[0] + dm->h2r[2] + dyi * dyi;
    if (r > d) d = r;
  }
  out[1] = d * 2.0 * dt * f->rre;
  free(u);
  free(v);
  free(w);
}

/* Bulk monitor values of write_monitor_bulk (src/io_monitor.f90:207-229):
 * out[0] = volumetric average of 1/2 (u^2 + v^2 + w^2) with u, v, w the P2C
 * interpolations of qx, qy, qz to the cell centres, out[1] = volumetric
 * average of qx (Get_volumetric_average_3d_for_var_xcx). */
void orc_bulk(orc_flow *f, double out[2]) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Nccc = (long)nx * ny * nz;
  int (*B)[5][2] = f->ibc;
  double *u = arr(Nccc), *v = arr(Nccc), *w = arr(Nccc);
  op(f, 0, ORC_MIDP_P2C, B[0][VU], f->qx, px, ny, nz, u, nx, ny, nz, f->fbx[VU]);
  op(f, 1, ORC_MIDP_P2C, B[1][VV], f->qy, nx, py, nz, v, nx, ny, nz, f->fby[VV]);
  op(f, 2, ORC_MIDP_P2C, B[2][VW], f->qz, nx, ny, pz, w, nx, ny, nz, f->fbz[VW]);
  for (long t = 0; t < Nccc; ++t) u[t] = 0.5 * (u[t] * u[t] + v[t] * v[t] + w[t] * w[t]);
  out[0] = orc_vol_average_x(f, u, nx);
  out[1] = orc_vol_average_x(f, f->qx, px);
  free(u);
  free(v);
  free(w);
}

/* ---------------------------------------------------------------------- */
/* Poisson: FISHPACK path, src/poisson_fishpack.f90:354-490                */
/* ---------------------------------------------------------------------- */
static void trid0(int ng, double *t, const double *a, const double *bb,
                  const double *c, double *d) {
  int nr = ng, mm1 = nr - 1;
  double z = 1.0 / bb[0];
  d[0] = c[0] * z;
  t[0] = t[0] * z;
  for (int i = 1; i < mm1; ++i) {
    z = 1.0 / (bb[i] - a[i] * d[i - 1]);
    d[i] = c[i] * z;
    t[i] = (t[i] - a[i] * t[i - 1]) * z;
  }
  z = bb[nr - 1] - a[nr - 1] * d[mm1 - 1];
  if (fabs(z) > 1.0e-10)
    t[nr - 1] = (t[nr - 1] - a[nr - 1] * t[mm1 - 1]) / z;
  else
    t[nr - 1] = 0.0;
  for (int ip = 1; ip <= mm1; ++ip) {
    int i = nr - ip - 1;
    t[i] = t[i] - d[i] * t[i + 1];
  }
}

static void poisson_fishpack(orc_flow *f, double *rhs) {
  int nx = f->nx, ny = f->ny, nz = f->nz;
#pragma omp parallel
  {
    int nmax = nx > nz ? nx : nz;
    nmax = nmax > ny ? nmax : ny;
    double *t = (double *)malloc(sizeof(double) * (size_t)nmax);
    double *bb = (double *)malloc(sizeof(double) * (size_t)ny);
    double *d = (double *)malloc(sizeof(double) * (size_t)ny);
#pragma omp for schedule(static)
    for (long L = 0; L < (long)ny * nz; ++L) {
      double *row = rhs + L * nx;
      orc_rfftf(nx, row);
    }
#pragma omp for schedule(static)
    for (long L = 0; L < (long)nx * ny; ++L) {
      long i = L % nx, j = L / nx;
      for (int k = 0; k < nz; ++k) t[k] = rhs[i + (long)nx * (j + (long)ny * k)];
      orc_rfftf(nz, t);
      for (int k = 0; k < nz; ++k) rhs[i + (long)nx * (j + (long)ny * k)] = t[k];
    }
#pragma omp for schedule(static)
    for (long L = 0; L < (long)nx * nz; ++L) {
      long i = L % nx, k = L / nx;
      for (int j = 0; j < ny; ++j) {
        bb[j] = f->pb[j] + f->xrt[i] / (1.0 * 1.0) + f->zrt[k];
        t[j] = rhs[i + (long)nx * (j + (long)ny * k)];
      }
      trid0(ny, t, f->pa, bb, f->pc, d);
      for (int j = 0; j < ny; ++j) rhs[i + (long)nx * (j + (long)ny * k)] = t[j];
    }
#pragma omp for schedule(static)
    for (long L = 0; L < (long)nx * ny; ++L) {
      long i = L % nx, j = L / nx;
      for (int k = 0; k < nz; ++k) t[k] = rhs[i + (long)nx * (j + (long)ny * k)];
      orc_rfftb(nz, t);
      for (int k = 0; k < nz; ++k) rhs[i + (long)nx * (j + (long)ny * k)] = t[k];
    }
#pragma omp for schedule(static)
    for (long L = 0; L < (long)ny * nz; ++L) orc_rfftb(nx, rhs + L * nx);
    free(t);
    free(bb);
    free(d);
  }
  double sx = (double)nx, sz = (double)nz;
  long n = (long)nx * ny * nz;
  for (long t = 0; t < n; ++t) rhs[t] = rhs[t] / sx / sz;
}

static void poisson_exact(orc_flow *f, double *rhs) {
  int nx = f->nx, ny = f->ny, nz = f->nz;
#pragma omp parallel
  {
    int nmax = nx > nz ? nx : nz;
    nmax = nmax > ny ? nmax : ny;
    double *t = (double *)malloc(sizeof(double) * (size_t)nmax);
    double *M = (double *)malloc(sizeof(double) * (size_t)ny * ny);
#pragma omp for schedule(static)
    for (long L = 0; L < (long)ny * nz; ++L) orc_rfftf(nx, rhs + L * nx);
#pragma omp for schedule(static)
    for (long L = 0; L < (long)nx * ny; ++L) {
      long i = L % nx, j = L / nx;
      for (int k = 0; k < nz; ++k) t[k] = rhs[i + (long)nx * (j + (long)ny * k)];
      orc_rfftf(nz, t);
      for (int k = 0; k < nz; ++k) rhs[i + (long)nx * (j + (long)ny * k)] = t[k];
    }
#pragma omp for schedule(dynamic, 4)
    for (long L = 0; L < (long)nx * nz; ++L) {
      long i = L % nx, k = L / nx;
      double lam = f->lxe[i] + f->lze[k];
      for (int a = 0; a < ny; ++a)
        for (int b = 0; b < ny; ++b) M[(size_t)a * ny + b] = f->ly[(size_t)a * ny + b] + (a == b ? lam : 0.0);
      for (int j = 0; j < ny; ++j) t[j] = rhs[i + (long)nx * (j + (long)ny * k)];
      int gauge = (i == 0 && k == 0);
      if (gauge) {
        for (int b = 0; b < ny; ++b) M[(size_t)(ny - 1) * ny + b] = (b == ny - 1) ? 1.0 : 0.0;
        t[ny - 1] = 0.0;
      }
      dense_solve(ny, M, t);
      if (gauge) {
        double lp = 0.0, l1 = 0.0;
        for (int j = 0; j < ny; ++j) { lp += f->lnull[j] * t[j]; l1 += f->lnull[j]; }
        for (int j = 0; j < ny; ++j) t[j] -= lp / l1;
      }
      for (int j = 0; j < ny; ++j) rhs[i + (long)nx * (j + (long)ny * k)] = t[j];
    }
#pragma omp for schedule(static)
    for (long L = 0; L < (long)nx * ny; ++L) {
      long i = L % nx, j = L / nx;
      for (int k = 0; k < nz; ++k) t[k] = rhs[i + (long)nx * (j + (long)ny * k)];
      orc_rfftb(nz, t);
      for (int k = 0; k < nz; ++k) rhs[i + (long)nx * (j + (long)ny * k)] = t[k];
    }
#pragma omp for schedule(static)
    for (long L = 0; L < (long)ny * nz; ++L) orc_rfftb(nx, rhs + L * nx);
    free(t);
    free(M);
  }
  double sx = (double)nx, sz = (double)nz;
  long n = (long)nx * ny * nz;
  for (long q = 0; q < n; ++q) rhs[q] = rhs[q] / sx / sz;
}

/* poisson_000, src/poisson_1stderivcomp.f90:1016-1195, triply periodic:
 * forward 3-D transform scaled by 1/(nx ny nz); each coefficient divided by
 * -(kx2 + ky2 + kz2) (real and imaginary parts alike, kxyz = (1,1) k2);
 * zero where k2 < 1e-16; inverse transform.  The abxyz half-cell rotations
 * multiply by exp(-i theta) before and exp(+i theta) after the real
 * division, and the sign flips for the upper half of the modes are applied
 * twice, so both cancel. */
static void poisson_000(orc_flow *f, double *rhs) {
  int nx = f->nx, ny = f->ny, nz = f->nz;
  long n = (long)nx * ny * nz;
  cpx *c = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (long t = 0; t < n; ++t) { c[t].re = rhs[t]; c[t].im = 0.0; }
  for (int pass = 0; pass < 2; ++pass) {
    int sign = pass ? 1 : -1;
    if (pass) {
      double sc = 1.0 / (double)nx / (double)ny / (double)nz;
      for (long t = 0; t < n; ++t) {
        long i = t % nx, j = (t / nx) % ny, k = t / ((long)nx * ny);
        double kk = f->kx2[i] + f->ky2[j] + f->kz2[k];
        double re = c[t].re * sc, im = c[t].im * sc;
        if (kk < 1.0e-16) { re = 0.0; im = 0.0; }
        else { re = re / (-kk); im = im / (-kk); }
        c[t].re = re; c[t].im = im;
      }
    }
#pragma omp parallel
    {
      int nmax = nx > ny ? nx : ny;
      nmax = nmax > nz ? nmax : nz;
      cpx *line = (cpx *)malloc(sizeof(cpx) * (size_t)nmax);
#pragma omp for schedule(static)
      for (long L = 0; L < (long)ny * nz; ++L) cfft(c + L * nx, nx, sign);
#pragma omp for schedule(static)
      for (long L = 0; L < (long)nx * nz; ++L) {
        long i = L % nx, k = L / nx;
        for (int j = 0; j < ny; ++j) line[j] = c[i + (long)nx * (j + (long)ny * k)];
        cfft(line, ny, sign);
        for (int j = 0; j < ny; ++j) c[i + (long)nx * (j + (long)ny * k)] = line[j];
      }
#pragma omp for schedule(static)
      for (long L = 0; L < (long)nx * ny; ++L) {
        for (int k = 0; k < nz; ++k) line[k] = c[L + (long)nx * ny * k];
        cfft(line, nz, sign);
        for (int k = 0; k < nz; ++k) c[L + (long)nx * ny * k] = line[k];
      }
      free(line);
    }
  }
  for (long t = 0; t < n; ++t) rhs[t] = c[t].re;
  free(c);
}

/* poisson_010, src/poisson_1stderivcomp.f90:1457-1816, istret = 0: x and z
 * periodic, y walls (bcy = 1) on a uniform y mesh, ny = ny_global - 1 = the
 * y cell count (decomp_2d_init on np = nc + 1 for the walled direction,
 * src/domain_decomposition.f90:150, src/geometry.f90:435-444; :818-826).
 * Steps, in the reference's order (1-based indices in the comments):
 *  1. y reorder (:1490-1500): b(j) = r(2j-1) for j <= ny/2, b(j) = r(2ny-2j+2)
 *     above;
 *  2. decomp_2d_fft_3d with PHYSICAL_IN_Z: r2c along z (modes 0..nz/2),
 *     c2c along y, c2c along x (e^{-i...}); / (nx ny nz) (:1503-1511);
 *  3. z post-processing: c * (bz - i az), az/bz = sin/cos(pi (k-1)/nz)
 *     (abxyz :2343-2401, bcz = 0) (:1524-1536);
 *  4. x post-processing: c * (bx - i ax), then -c for i > nx/2+1 (:1539-1552);
 *  5. y post-processing (:1557-1578): X(1) = V(1); X(j) = 1/2 [e^{-i th_j}
 *     V(j) + e^{+i th_j} V(ny-j+2)], th_j = pi (j-1) / (2 ny) (ay, by with
 *     bcy = 1);
 *  6. istret = 0 division (:1591-1612): the real part by -Re(kxyz), the
 *     imaginary part by -Im(kxyz), kxyz = (1,1) (xk2 + yk2 + zk2); zero
 *     where both are below 1e-16;
 *  7. backward y post-processing (:1719-1740): V(1) = X(1); V(j) =
 *     (by + i ay)(X(j) - i X(ny-j+2)) written out as in the reference;
 *  8. x, then z post-processing back (:1757-1786): c * (b + i a), the x sign
 *     flip again;
 *  9. inverse transform c2c x, c2c y, c2r z (e^{+i...}, unnormalised), and
 *     the y un-reorder r(2j-1) = b(j), r(2j) = b(ny-j+1) (:1792-1808).
 * The c2r along z rebuilds modes nz/2+1..nz-1 as conjugates of the stored
 * ones and takes the real part, as a c2r transform does. */
static void poisson_010(orc_flow *f, double *rhs) {
  const int nx = f->nx, ny = f->ny, nz = f->nz, nzh = nz / 2 + 1;
  const double pi = 2.0 * asin(1.0);
  const long nsp = (long)nx * ny * nzh;
  cpx *c = (cpx *)malloc(sizeof(cpx) * (size_t)nsp);
  double *b = (double *)malloc(sizeof(double) * (size_t)nx * ny * nz);
  /* 1. reorder */
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      for (int j = 1; j <= ny / 2; ++j)
        b[i + (long)nx * ((j - 1) + (long)ny * k)] = rhs[i + (long)nx * ((2 * (j - 1)) + (long)ny * k)];
      for (int j = ny / 2 + 1; j <= ny; ++j)
        b[i + (long)nx * ((j - 1) + (long)ny * k)] = rhs[i + (long)nx * ((2 * ny - 2 * j + 1) + (long)ny * k)];
    }
  int nmax = nx > ny ? nx : ny;
  nmax = nmax > nz ? nmax : nz;
  cpx *line = (cpx *)malloc(sizeof(cpx) * (size_t)nmax);
  /* 2. r2c z, c2c y, c2c x; c indexed (i, j, k) with k < nzh */
  for (long L = 0; L < (long)nx * ny; ++L) {
    for (int k = 0; k < nz; ++k) { line[k].re = b[L + (long)nx * ny * k]; line[k].im = 0.0; }
    cfft(line, nz, -1);
    for (int k = 0; k < nzh; ++k) c[L + (long)nx * ny * k] = line[k];
  }
  for (long L = 0; L < (long)nx * nzh; ++L) {
    long i = L % nx, k = L / nx;
    for (int j = 0; j < ny; ++j) line[j] = c[i + (long)nx * (j + (long)ny * k)];
    cfft(line, ny, -1);
    for (int j = 0; j < ny; ++j) c[i + (long)nx * (j + (long)ny * k)] = line[j];
  }
  for (long L = 0; L < (long)ny * nzh; ++L) cfft(c + L * nx, nx, -1);
  const double sc = 1.0 / (double)nx / (double)ny / (double)nz;
  for (long t = 0; t < nsp; ++t) { c[t].re *= sc; c[t].im *= sc; }
  /* 3. z, 4. x post-processing */
  for (long t = 0; t < nsp; ++t) {
    long i = t % nx, k = t / ((long)nx * ny);
    double az = sin((double)k * pi / (double)nz), bz = cos((double)k * pi / (double)nz);
    double t1 = c[t].re, t2 = c[t].im;
    c[t].re = t1 * bz + t2 * az;
    c[t].im = t2 * bz - t1 * az;
    double ax = sin((double)i * pi / (double)nx), bx = cos((double)i * pi / (double)nx);
    t1 = c[t].re; t2 = c[t].im;
    c[t].re = t1 * bx + t2 * ax;
    c[t].im = t2 * bx - t1 * ax;
    if (i + 1 > nx / 2 + 1) { c[t].re = -c[t].re; c[t].im = -c[t].im; }
  }
  /* 5. y post-processing, 6. division, 7. backward y post-processing */
  cpx *v = (cpx *)malloc(sizeof(cpx) * (size_t)ny);
  cpx *x = (cpx *)malloc(sizeof(cpx) * (size_t)ny);
  for (long L = 0; L < (long)nx * nzh; ++L) {
    long i = L % nx, k = L / nx;
    for (int j = 0; j < ny; ++j) v[j] = c[i + (long)nx * (j + (long)ny * k)];
    x[0] = v[0];
    for (int j = 2; j <= ny; ++j) {
      double ay = sin((double)(j - 1) * pi * 0.5 / (double)ny), by = cos((double)(j - 1) * pi * 0.5 / (double)ny);
      double t1 = v[j - 1].re, t2 = v[j - 1].im, t3 = v[ny - j + 1].re, t4 = v[ny - j + 1].im;
      double x1 = t1 * by, x2 = t1 * ay, x3 = t2 * by, x4 = t2 * ay;
      double x5 = t3 * by, x6 = t3 * ay, x7 = t4 * by, x8 = t4 * ay;
      x[j - 1].re = 0.5 * (x1 + x4 + x5 - x8);
      x[j - 1].im = 0.5 * (-x2 + x3 + x6 + x7);
    }
    for (int j = 0; j < ny; ++j) {
      double kk = f->kx2[i] + f->ky2[j] + f->kz2[k];
      if (fabs(kk) < 1.0e-16) { x[j].re = 0.0; x[j].im = 0.0; }
      else { x[j].re = x[j].re / (-kk); x[j].im = x[j].im / (-kk); }
    }
    v[0] = x[0];
    for (int j = 2; j <= ny; ++j) {
      double ay = sin((double)(j - 1) * pi * 0.5 / (double)ny), by = cos((double)(j - 1) * pi * 0.5 / (double)ny);
      double t1 = x[j - 1].re, t2 = x[j - 1].im, t3 = x[ny - j + 1].re, t4 = x[ny - j + 1].im;
      double x1 = t1 * by, x2 = t1 * ay, x3 = t2 * by, x4 = t2 * ay;
      double x5 = t3 * by, x6 = t3 * ay, x7 = t4 * by, x8 = t4 * ay;
      v[j - 1].re = x1 - x4 + x6 + x7;
      v[j - 1].im = -(-x2 - x3 + x5 - x8);
    }
    for (int j = 0; j < ny; ++j) c[i + (long)nx * (j + (long)ny * k)] = v[j];
  }
  free(v);
  free(x);
  /* 8. x, z post-processing back */
  for (long t = 0; t < nsp; ++t) {
    long i = t % nx, k = t / ((long)nx * ny);
    double ax = sin((double)i * pi / (double)nx), bx = cos((double)i * pi / (double)nx);
    double t1 = c[t].re, t2 = c[t].im;
    c[t].re = t1 * bx - t2 * ax;
    c[t].im = t2 * bx + t1 * ax;
    if (i + 1 > nx / 2 + 1) { c[t].re = -c[t].re; c[t].im = -c[t].im; }
    double az = sin((double)k * pi / (double)nz), bz = cos((double)k * pi / (double)nz);
    t1 = c[t].re; t2 = c[t].im;
    c[t].re = t1 * bz - t2 * az;
    c[t].im = t2 * bz + t1 * az;
  }
  /* 9. c2c x, c2c y, c2r z, un-reorder */
  for (long L = 0; L < (long)ny * nzh; ++L) cfft(c + L * nx, nx, 1);
  for (long L = 0; L < (long)nx * nzh; ++L) {
    long i = L % nx, k = L / nx;
    for (int j = 0; j < ny; ++j) line[j] = c[i + (long)nx * (j + (long)ny * k)];
    cfft(line, ny, 1);
    for (int j = 0; j < ny; ++j) c[i + (long)nx * (j + (long)ny * k)] = line[j];
  }
  for (long L = 0; L < (long)nx * ny; ++L) {
    for (int k = 0; k < nzh; ++k) line[k] = c[L + (long)nx * ny * k];
    line[0].im = 0.0;
    if (nz % 2 == 0) line[nz / 2].im = 0.0;
    for (int k = nzh; k < nz; ++k) { line[k].re = line[nz - k].re; line[k].im = -line[nz - k].im; }
    cfft(line, nz, 1);
    for (int k = 0; k < nz; ++k) b[L + (long)nx * ny * k] = line[k].re;
  }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i)
      for (int j = 1; j <= ny / 2; ++j) {
        rhs[i + (long)nx * ((2 * j - 2) + (long)ny * k)] = b[i + (long)nx * ((j - 1) + (long)ny * k)];
        rhs[i + (long)nx * ((2 * j - 1) + (long)ny * k)] = b[i + (long)nx * ((ny - j) + (long)ny * k)];
      }
  free(line);
  free(b);
  free(c);
}

void orc_poisson_solve(orc_flow *f, double *rhs) {
  if (f->cfg.projection == 1) poisson_exact(f, rhs);
  else if (f->cfg.poisson == 2) poisson_fishpack(f, rhs);
  else if (f->cfg.poisson == 3 && f->dm.periodic[1]) poisson_000(f, rhs);
  else if (f->cfg.poisson == 3) poisson_010(f, rhs);
}

/* ---------------------------------------------------------------------- */
/* momentum RHS, src/eq_momentum2.f90:101-1571 (Cartesian, flow only)      */
/* ---------------------------------------------------------------------- */
static double *fill_maxp(long n) {
  double *a = (double *)malloc(sizeof(double) * (size_t)n);
  for (long t = 0; t < n; ++t) a[t] = MAXP;
  return a;
}

void orc_momentum_rhs(orc_flow *f, int isub, double *mx, double *my,
                      double *mz) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Npcc = (long)px * ny * nz, Ncpc = (long)nx * py * nz, Nccp = (long)nx * ny * pz;
  long Nccc = (long)nx * ny * nz, Nppc = (long)px * py * nz, Npcp = (long)px * ny * pz;
  long Ncpp = (long)nx * py * pz;
  int (*B)[5][2] = f->ibc;
  const double rre = f->rre;
  (void)isub;
  /* first-level operators */
  double *qxix = arr(Nccc), *qxdx = arr(Nccc), *qxdy = arr(Nppc), *qxiy = arr(Nppc);
  double *qxiz = arr(Npcp), *qxdz = arr(Npcp);
  op(f, 0, ORC_MIDP_P2C, B[0][VU], f->qx, px, ny, nz, qxix, nx, ny, nz, f->fbx[VU]);
  op(f, 0, ORC_1DER_P2C, B[0][VU], f->qx, px, ny, nz, qxdx, nx, ny, nz, f->fbx[VU]);
  op(f, 1, ORC_1DER_C2P, B[1][VU], f->qx, px, ny, nz, qxdy, px, py, nz, f->fby[VU]);
  op(f, 1, ORC_MIDP_C2P, B[1][VU], f->qx, px, ny, nz, qxiy, px, py, nz, f->fby[VU]);
  op(f, 2, ORC_MIDP_C2P, B[2][VU], f->qx, px, ny, nz, qxiz, px, ny, pz, f->fbz[VU]);
  op(f, 2, ORC_1DER_C2P, B[2][VU], f->qx, px, ny, nz, qxdz, px, ny, pz, f->fbz[VU]);
  double *qxdx_cpc = NULL;
  if (f->freq[1]) {
    qxdx_cpc = arr(Ncpc);
    op(f, 0, ORC_1DER_P2C, B[0][VU], qxiy, px, py, nz, qxdx_cpc, nx, py, nz, NULL);
  }
  double *qydx = arr(Nppc), *qyix = arr(Nppc), *qydy = arr(Nccc), *qyiy = arr(Nccc);
  double *qydz = arr(Ncpp), *qyiz = arr(Ncpp);
  op(f, 0, ORC_1DER_C2P, B[0][VV], f->qy, nx, py, nz, qydx, px, py, nz, f->fbx[VV]);
  op(f, 0, ORC_MIDP_C2P, B[0][VV], f->qy, nx, py, nz, qyix, px, py, nz, f->fbx[VV]);
  op(f, 1, ORC_1DER_P2C, B[1][VV], f->qy, nx, py, nz, qydy, nx, ny, nz, NULL);
  op(f, 1, ORC_MIDP_P2C, B[1][VV], f->qy, nx, py, nz, qyiy, nx, ny, nz, f->fby[VV]);
  op(f, 2, ORC_1DER_C2P, B[2][VV], f->qy, nx, py, nz, qydz, nx, py, pz, f->fbz[VV]);
  op(f, 2, ORC_MIDP_C2P, B[2][VV], f->qy, nx, py, nz, qyiz, nx, py, pz, f->fbz[VV]);
  double *qydy_cpc = NULL;
  if (f->freq[1]) {
    double *t = arr(Nccc);
    qydy_cpc = arr(Ncpc);
    op(f, 1, ORC_MIDP_P2C, B[1][VV], f->qy, nx, py, nz, t, nx, ny, nz, f->fby[VV]);
    op(f, 1, ORC_1DER_C2P, B[1][VV], t, nx, ny, nz, qydy_cpc, nx, py, nz, f->fby[VV]);
    free(t);
  }
  double *qzdx = arr(Npcp), *qzix = arr(Npcp), *qziy = arr(Ncpp), *qzdy = arr(Ncpp);
  double *qziz = arr(Nccc), *qzdz = arr(Nccc);
  op(f, 0, ORC_1DER_C2P, B[0][VW], f->qz, nx, ny, pz, qzdx, px, ny, pz, f->fbx[VW]);
  op(f, 0, ORC_MIDP_C2P, B[0][VW], f->qz, nx, ny, pz, qzix, px, ny, pz, f->fbx[VW]);
  op(f, 1, ORC_MIDP_C2P, B[1][VW], f->qz, nx, ny, pz, qziy, nx, py, pz, f->fby[VW]);
  op(f, 1, ORC_1DER_C2P, B[1][VW], f->qz, nx, ny, pz, qzdy, nx, py, pz, f->fby[VW]);
  op(f, 2, ORC_MIDP_P2C, B[2][VW], f->qz, nx, ny, pz, qziz, nx, ny, nz, f->fbz[VW]);
  op(f, 2, ORC_1DER_P2C, B[2][VW], f->qz, nx, ny, pz, qzdz, nx, ny, nz, NULL);
  double *qzdz_cpc = NULL;
  if (f->freq[1]) {
    double *t = arr(Ncpp);
    qzdz_cpc = arr(Ncpc);
    op(f, 1, ORC_MIDP_C2P, B[1][VW], f->qz, nx, ny, pz, t, nx, py, pz, f->fby[VW]);
    op(f, 2, ORC_1DER_P2C, B[2][VW], t, nx, py, pz, qzdz_cpc, nx, py, nz, NULL);
    free(t);
  }
  /* divergence */
  double *div = arr(Nccc);
  for (long t = 0; t < Nccc; ++t) div[t] = 0.0;
  for (long t = 0; t < Nccc; ++t) div[t] = div[t] + qxdx[t];
  for (long t = 0; t < Nccc; ++t) div[t] = div[t] + qydy[t];
  for (long t = 0; t < Nccc; ++t) div[t] = div[t] + qzdz[t];
  long nfy = (long)nx * 4 * nz;
  double *fbdiv_y = NULL;
  if (f->freq[1]) {
    fbdiv_y = arr(nfy);
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i)
        for (int s = 0; s < 2; ++s) {
          long j = s ? py - 1 : 0;
          long src = i + (long)nx * (j + (long)py * k);
          double v = qxdx_cpc[src] + qydy_cpc[src] + qzdz_cpc[src];
          fbdiv_y[i + (long)nx * (s + 4L * k)] = v;
          fbdiv_y[i + (long)nx * (s + 2 + 4L * k)] = v;
        }
  }
  /* variable properties (:576-703): interpolations of g = rho u, the
   * advecting factor of each convective product, and of the viscosity;
   * with constant properties g is q and mu is 1 */
  const int th = f->thermo;
  double *gxix = qxix, *gxiy = qxiy, *gxiz = qxiz, *gyix = qyix, *gyiy = qyiy, *gyiz = qyiz;
  double *gzix = qzix, *gziy = qziy, *gziz = qziz;
  double *mu_c = NULL, *muixy = NULL, *muixz = NULL, *muiyz = NULL, *fbmu_y = NULL;
  if (th) {
    gxix = arr(Nccc); gxiy = arr(Nppc); gxiz = arr(Npcp);
    gyix = arr(Nppc); gyiy = arr(Nccc); gyiz = arr(Ncpp);
    gzix = arr(Npcp); gziy = arr(Ncpp); gziz = arr(Nccc);
    op(f, 0, ORC_MIDP_P2C, B[0][VU], f->gx, px, ny, nz, gxix, nx, ny, nz, f->fbx[VU]);
    op(f, 1, ORC_MIDP_C2P, B[1][VU], f->gx, px, ny, nz, gxiy, px, py, nz, f->fby[VU]);
    op(f, 2, ORC_MIDP_C2P, B[2][VU], f->gx, px, ny, nz, gxiz, px, ny, pz, f->fbz[VU]);
    op(f, 0, ORC_MIDP_C2P, B[0][VV], f->gy, nx, py, nz, gyix, px, py, nz, f->fbx[VV]);
    op(f, 1, ORC_MIDP_P2C, B[1][VV], f->gy, nx, py, nz, gyiy, nx, ny, nz, f->fby[VV]);
    op(f, 2, ORC_MIDP_C2P, B[2][VV], f->gy, nx, py, nz, gyiz, nx, py, pz, f->fbz[VV]);
    op(f, 0, ORC_MIDP_C2P, B[0][VW], f->gz, nx, ny, pz, gzix, px, ny, pz, f->fbx[VW]);
    op(f, 1, ORC_MIDP_C2P, B[1][VW], f->gz, nx, ny, pz, gziy, nx, py, pz, f->fby[VW]);
    op(f, 2, ORC_MIDP_P2C, B[2][VW], f->gz, nx, ny, pz, gziz, nx, ny, nz, f->fbz[VW]);
    /* mu_ccc; muiy_cpc (y, wall viscosity); muixy_ppc = x of muiy_cpc;
     * muixz_pcp = z of x; muiyz_cpp = z of muiy_cpc; wall rows of muiy_cpc
     * as fbcy_mu_c4c (:648-698) */
    mu_c = f->visc;
    double *muiy = arr(Ncpc), *mupcc = arr(Npcc);
    muixy = arr(Nppc); muixz = arr(Npcp); muiyz = arr(Ncpp);
    op(f, 1, ORC_MIDP_C2P, f->ftp[1], f->visc, nx, ny, nz, muiy, nx, py, nz, f->fby_ftp[2]);
    op(f, 0, ORC_MIDP_C2P, f->ftp[0], muiy, nx, py, nz, muixy, px, py, nz, NULL);
    op(f, 0, ORC_MIDP_C2P, f->ftp[0], f->visc, nx, ny, nz, mupcc, px, ny, nz, NULL);
    op(f, 2, ORC_MIDP_C2P, f->ftp[2], mupcc, px, ny, nz, muixz, px, ny, pz, NULL);
    op(f, 2, ORC_MIDP_C2P, f->ftp[2], muiy, nx, py, nz, muiyz, nx, py, pz, NULL);
    fbmu_y = arr((long)nx * 4 * nz);
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i)
        for (int s = 0; s < 2; ++s) {
          double v = muiy[i + (long)nx * ((s ? py - 1 : 0) + (long)py * k)];
          fbmu_y[i + (long)nx * (s + 4L * k)] = v;
          fbmu_y[i + (long)nx * (s + 2 + 4L * k)] = v;
        }
    free(muiy); free(mupcc);
  }
#define MU(a, t) ((a) ? (a)[t] : 1.0)
  double *maxp_x = fill_maxp(4L * (ny > py ? ny : py) * (nz > pz ? nz : pz));
  double *maxp_y = fill_maxp((long)(nx > px ? nx : px) * 4 * (nz > pz ? nz : pz));
  double *maxp_z = fill_maxp((long)(nx > px ? nx : px) * (ny > py ? ny : py) * 4);
  double *tc = arr(Nccc), *tppc = arr(Nppc), *tpcp = arr(Npcp), *tcpp = arr(Ncpp);
  double *r1 = arr(Npcc > Ncpc ? (Npcc > Nccp ? Npcc : Nccp) : (Ncpc > Nccp ? Ncpc : Nccp));
  double *acc_y = arr(Npcc > Ncpc ? (Npcc > Nccp ? Npcc : Nccp) : (Ncpc > Nccp ? Ncpc : Nccp));
  double *acc_z = arr(Npcc > Ncpc ? (Npcc > Nccp ? Npcc : Nccp) : (Ncpc > Nccp ? Ncpc : Nccp));
  double *pfc = arr(Npcc > Ncpc ? (Npcc > Nccp ? Npcc : Nccp) : (Ncpc > Nccp ? Ncpc : Nccp));
  double *mpres = arr(Nccc);
  for (long t = 0; t < Nccc; ++t) mpres[t] = -f->pres[t];
  double *nfbx = arr(4L * ny * nz), *nfby = arr((long)nx * 4 * nz), *nfbz = arr((long)nx * ny * 4);
  for (long t = 0; t < 4L * ny * nz; ++t) nfbx[t] = -f->fbx[VP][t];
  for (long t = 0; t < (long)nx * 4 * nz; ++t) nfby[t] = -f->fby[VP][t];
  for (long t = 0; t < (long)nx * ny * 4; ++t) nfbz[t] = -f->fbz[VP][t];

  /* ---------------- x-momentum ---------------- */
  for (long t = 0; t < Nccc; ++t) tc[t] = -gxix[t] * qxix[t];
  op(f, 0, ORC_1DER_C2P, f->mcov[0][0], tc, nx, ny, nz, r1, px, ny, nz, maxp_x);
  for (long t = 0; t < Npcc; ++t) mx[t] = 0.0 + r1[t];
  for (long t = 0; t < Nppc; ++t) tppc[t] = -gyix[t] * qxiy[t];
  op(f, 1, ORC_1DER_P2C, f->mcov[0][1], tppc, px, py, nz, r1, px, ny, nz, NULL);
  for (long t = 0; t < Npcc; ++t) acc_y[t] = 0.0 + r1[t];
  for (long t = 0; t < Npcp; ++t) tpcp[t] = -gzix[t] * qxiz[t];
  op(f, 2, ORC_1DER_P2C, f->mcov[0][2], tpcp, px, ny, pz, r1, px, ny, nz, NULL);
  for (long t = 0; t < Npcc; ++t) acc_z[t] = 0.0 + r1[t];
  op(f, 0, ORC_1DER_C2P, B[0][VP], mpres, nx, ny, nz, pfc, px, ny, nz, nfbx);
  for (long t = 0; t < Nccc; ++t) tc[t] = (2.0 * MU(mu_c, t)) * (qxdx[t] - ONE_THIRD * div[t]);
  op(f, 0, ORC_1DER_C2P, f->mtau[0][0], tc, nx, ny, nz, r1, px, ny, nz, maxp_x);
  for (long t = 0; t < Npcc; ++t) mx[t] = mx[t] + r1[t] * rre;
  for (long t = 0; t < Nppc; ++t) tppc[t] = (qxdy[t] + qydx[t]) * MU(muixy, t);
  op(f, 1, ORC_1DER_P2C, f->mtau[0][1], tppc, px, py, nz, r1, px, ny, nz, NULL);
  for (long t = 0; t < Npcc; ++t) acc_y[t] = acc_y[t] + r1[t] * rre;
  for (long t = 0; t < Npcp; ++t) tpcp[t] = (qxdz[t] + qzdx[t]) * MU(muixz, t);
  op(f, 2, ORC_1DER_P2C, f->mtau[0][2], tpcp, px, ny, pz, r1, px, ny, nz, NULL);
  for (long t = 0; t < Npcc; ++t) acc_z[t] = acc_z[t] + r1[t] * rre;
  for (long t = 0; t < Npcc; ++t) mx[t] = mx[t] + acc_y[t];
  for (long t = 0; t < Npcc; ++t) mx[t] = mx[t] + acc_z[t];
  /* Calculate_momentum_fractional_step */
  for (long t = 0; t < Npcc; ++t) {
    double cur = mx[t], last = f->mx0[t];
    double tot = f->tg[isub] * cur + f->tz[isub] * last;
    f->mx0[t] = cur;
    tot = tot + f->ta[isub] * pfc[t];
    mx[t] = f->dt * tot;
  }

  /* ---------------- y-momentum ---------------- */
  for (long t = 0; t < Nppc; ++t) tppc[t] = -gxiy[t] * qyix[t];
  op(f, 0, ORC_1DER_P2C, f->mcov[1][0], tppc, px, py, nz, r1, nx, py, nz, NULL);
  for (long t = 0; t < Ncpc; ++t) my[t] = 0.0 + r1[t];
  double *fbc_y = maxp_y, *fbc_own = NULL;
  if (f->freq[1]) {
    fbc_own = arr(nfy);
    for (long t = 0; t < nfy; ++t) fbc_own[t] = -f->fby[VV][t] * f->fby[VV][t];
    fbc_y = fbc_own;
  }
  for (long t = 0; t < Nccc; ++t) tc[t] = -qyiy[t] * gyiy[t];
  op(f, 1, ORC_1DER_C2P, f->mcov[1][1], tc, nx, ny, nz, r1, nx, py, nz, fbc_y);
  for (long t = 0; t < Ncpc; ++t) acc_y[t] = 0.0 + r1[t];
  for (long t = 0; t < Ncpp; ++t) tcpp[t] = -qyiz[t] * gziy[t];
  op(f, 2, ORC_1DER_P2C, f->mcov[1][2], tcpp, nx, py, pz, r1, nx, py, nz, NULL);
  for (long t = 0; t < Ncpc; ++t) acc_z[t] = 0.0 + r1[t];
  {
    double *pf = arr(Ncpc);
    op(f, 1, ORC_1DER_C2P, B[1][VP], mpres, nx, ny, nz, pf, nx, py, nz, nfby);
    for (long t = 0; t < Ncpc; ++t) pfc[t] = 0.0 + pf[t];
    free(pf);
  }
  for (long t = 0; t < Nppc; ++t) tppc[t] = (qxdy[t] + qydx[t]) * MU(muixy, t);
  op(f, 0, ORC_1DER_P2C, f->mtau[1][0], tppc, px, py, nz, r1, nx, py, nz, NULL);
  for (long t = 0; t < Ncpc; ++t) my[t] = my[t] + r1[t] * rre;
  fbc_y = maxp_y;
  if (f->freq[1]) {
    for (int k = 0; k < nz; ++k)
      for (int i = 0; i < nx; ++i)
        for (int s = 0; s < 2; ++s) {
          long j = s ? py - 1 : 0;
          double v = qydy_cpc[i + (long)nx * (j + (long)py * k)];
          fbc_own[i + (long)nx * (s + 4L * k)] = v;
          fbc_own[i + (long)nx * (s + 2 + 4L * k)] = v;
        }
    for (long t = 0; t < nfy; ++t)
      fbc_own[t] = (fbc_own[t] - ONE_THIRD * fbdiv_y[t]) * 2.0 * MU(fbmu_y, t);
    fbc_y = fbc_own;
  }
  for (long t = 0; t < Nccc; ++t) tc[t] = (qydy[t] - ONE_THIRD * div[t]) * 2.0 * MU(mu_c, t);
  op(f, 1, ORC_1DER_C2P, f->mtau[1][1], tc, nx, ny, nz, r1, nx, py, nz, fbc_y);
  for (long t = 0; t < Ncpc; ++t) acc_y[t] = acc_y[t] + r1[t] * rre;
  for (long t = 0; t < Ncpp; ++t) tcpp[t] = (qzdy[t] + qydz[t] - 0.0) * MU(muiyz, t);
  op(f, 2, ORC_1DER_P2C, f->mtau[1][2], tcpp, nx, py, pz, r1, nx, py, nz, NULL);
  for (long t = 0; t < Ncpc; ++t) acc_z[t] = acc_z[t] + r1[t] * rre;
  for (long t = 0; t < Ncpc; ++t) my[t] = my[t] + acc_y[t];
  for (long t = 0; t < Ncpc; ++t) my[t] = my[t] + acc_z[t];
  for (long t = 0; t < Ncpc; ++t) {
    double cur = my[t], last = f->my0[t];
    double tot = f->tg[isub] * cur + f->tz[isub] * last;
    f->my0[t] = cur;
    tot = tot + f->ta[isub] * pfc[t];
    my[t] = f->dt * tot;
  }

  /* ---------------- z-momentum ---------------- */
  for (long t = 0; t < Npcp; ++t) tpcp[t] = -gxiz[t] * qzix[t];
  op(f, 0, ORC_1DER_P2C, f->mcov[2][0], tpcp, px, ny, pz, r1, nx, ny, pz, NULL);
  for (long t = 0; t < Nccp; ++t) mz[t] = 0.0 + r1[t];
  for (long t = 0; t < Ncpp; ++t) tcpp[t] = -gyiz[t] * qziy[t];
  op(f, 1, ORC_1DER_P2C, f->mcov[2][1], tcpp, nx, py, pz, r1, nx, ny, pz, NULL);
  for (long t = 0; t < Nccp; ++t) acc_y[t] = 0.0 + r1[t];
  for (long t = 0; t < Nccc; ++t) tc[t] = -gziz[t] * qziz[t];
  op(f, 2, ORC_1DER_C2P, f->mcov[2][2], tc, nx, ny, nz, r1, nx, ny, pz, maxp_z);
  for (long t = 0; t < Nccp; ++t) acc_z[t] = 0.0 + r1[t];
  {
    double *pf = arr(Nccp);
    op(f, 2, ORC_1DER_C2P, B[2][VP], mpres, nx, ny, nz, pf, nx, ny, pz, nfbz);
    for (long t = 0; t < Nccp; ++t) pfc[t] = 0.0 + pf[t];
    free(pf);
  }
  for (long t = 0; t < Npcp; ++t) tpcp[t] = (qxdz[t] + qzdx[t]) * MU(muixz, t);
  op(f, 0, ORC_1DER_P2C, f->mtau[2][0], tpcp, px, ny, pz, r1, nx, ny, pz, NULL);
  for (long t = 0; t < Nccp; ++t) mz[t] = mz[t] + r1[t] * rre;
  for (long t = 0; t < Ncpp; ++t) tcpp[t] = (qzdy[t] + qydz[t]) * MU(muiyz, t);
  op(f, 1, ORC_1DER_P2C, f->mtau[2][1], tcpp, nx, py, pz, r1, nx, ny, pz, NULL);
  for (long t = 0; t < Nccp; ++t) acc_y[t] = acc_y[t] + r1[t] * rre;
  for (long t = 0; t < Nccc; ++t) tc[t] = ((qzdz[t] - ONE_THIRD * div[t]) + 0.0) * 2.0 * MU(mu_c, t);
#undef MU
  op(f, 2, ORC_1DER_C2P, f->mtau[2][2], tc, nx, ny, nz, r1, nx, ny, pz, maxp_z);
  for (long t = 0; t < Nccp; ++t) acc_z[t] = acc_z[t] + r1[t] * rre;
  for (long t = 0; t < Nccp; ++t) mz[t] = mz[t] + acc_y[t];
  for (long t = 0; t < Nccp; ++t) mz[t] = mz[t] + acc_z[t];
  for (long t = 0; t < Nccp; ++t) {
    double cur = mz[t], last = f->mz0[t];
    double tot = f->tg[isub] * cur + f->tz[isub] * last;
    f->mz0[t] = cur;
    tot = tot + f->ta[isub] * pfc[t];
    mz[t] = f->dt * tot;
  }
  /* constant-flux forcing (after the fractional step), :1547-1562 */
  if (f->cfg.idriven == 1) {
    double bulk = orc_vol_average_x(f, mx, px);
    for (long t = 0; t < Npcc; ++t) mx[t] = mx[t] - bulk;
  }
  free(qxix); free(qxdx); free(qxdy); free(qxiy); free(qxiz); free(qxdz);
  free(qydx); free(qyix); free(qydy); free(qyiy); free(qydz); free(qyiz);
  free(qzdx); free(qzix); free(qziy); free(qzdy); free(qziz); free(qzdz);
  free(qxdx_cpc); free(qydy_cpc); free(qzdz_cpc); free(div); free(fbdiv_y);
  free(maxp_x); free(maxp_y); free(maxp_z); free(tc); free(tppc); free(tpcp);
  free(tcpp); free(r1); free(acc_y); free(acc_z); free(pfc); free(mpres);
  free(nfbx); free(nfby); free(nfbz); free(fbc_own);
  if (th) {
    free(gxix); free(gxiy); free(gxiz); free(gyix); free(gyiy); free(gyiz);
    free(gzix); free(gziy); free(gziz); free(muixy); free(muixz); free(muiyz); free(fbmu_y);
  }
}

/* the projection half of Solve_momentum_eq (:1866-1921): Poisson RHS,
 * solve, pressure update, velocity correction, wall enforcement */
void orc_project(orc_flow *f, int isub) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Npcc = (long)px * ny * nz, Ncpc = (long)nx * py * nz, Nccp = (long)nx * ny * pz;
  long Nccc = (long)nx * ny * nz;
  double *mx = arr(Npcc), *my = arr(Ncpc), *mz = arr(Nccp);
  double *div = arr(Nccc);
  for (long t = 0; t < Nccc; ++t) f->pcor[t] = 0.0;
  orc_divergence(f, div);
  for (long t = 0; t < Nccc; ++t) f->pcor[t] = f->pcor[t] + div[t];
  double coeff = 1.0 / (f->ta[isub] * f->sigma2p * f->dt);
  for (long t = 0; t < Nccc; ++t) f->pcor[t] = f->pcor[t] * coeff;
  orc_poisson_solve(f, f->pcor);
  for (long t = 0; t < Nccc; ++t) f->pres[t] = f->pres[t] + f->pcor[t];
  double s = f->dt * f->ta[isub] * f->sigma2p;
  op(f, 0, ORC_1DER_C2P, f->ibc[0][VP], f->pcor, nx, ny, nz, mx, px, ny, nz, f->fbx[VP]);
  for (long t = 0; t < Npcc; ++t) f->qx[t] = f->qx[t] - s * mx[t];
  op(f, 1, ORC_1DER_C2P, f->ibc[1][VP], f->pcor, nx, ny, nz, my, nx, py, nz, f->fby[VP]);
  for (long t = 0; t < Ncpc; ++t) f->qy[t] = f->qy[t] - s * my[t];
  op(f, 2, ORC_1DER_C2P, f->ibc[2][VP], f->pcor, nx, ny, nz, mz, nx, ny, pz, f->fbz[VP]);
  for (long t = 0; t < Nccp; ++t) f->qz[t] = f->qz[t] - s * mz[t];
  enforce_velo(f);
  free(mx); free(my); free(mz); free(div);
}

/* convert_primary_conservative, src/para_conversion.f90:5-227: density
 * interpolated to each velocity location (midp C2P with the property BCs and
 * the wall density), g = q d (IQ2G) or q = g / d (IG2Q).  The walls are
 * stationary, so the wall values of g and q are zero and stay so. */
static void convert_qg(orc_flow *f, int to_g) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Npcc = (long)px * ny * nz, Ncpc = (long)nx * py * nz, Nccp = (long)nx * ny * pz;
  double *dx = arr(Npcc), *dy = arr(Ncpc), *dz = arr(Nccp);
  op(f, 0, ORC_MIDP_C2P, f->ftp[0], f->dens, nx, ny, nz, dx, px, ny, nz, NULL);
  op(f, 1, ORC_MIDP_C2P, f->ftp[1], f->dens, nx, ny, nz, dy, nx, py, nz, f->fby_ftp[3]);
  op(f, 2, ORC_MIDP_C2P, f->ftp[2], f->dens, nx, ny, nz, dz, nx, ny, pz, NULL);
  if (to_g) {
    for (long t = 0; t < Npcc; ++t) f->gx[t] = f->qx[t] * dx[t];
    for (long t = 0; t < Ncpc; ++t) f->gy[t] = f->qy[t] * dy[t];
    for (long t = 0; t < Nccp; ++t) f->gz[t] = f->qz[t] * dz[t];
  } else {
    for (long t = 0; t < Npcc; ++t) f->qx[t] = f->gx[t] / dx[t];
    for (long t = 0; t < Ncpc; ++t) f->qy[t] = f->gy[t] / dy[t];
    for (long t = 0; t < Nccp; ++t) f->qz[t] = f->gz[t] / dz[t];
  }
  free(dx); free(dy); free(dz);
}

/* Solve_momentum_eq, src/eq_momentum2.f90:1799-1932; with variable
 * properties the right-hand side advances g, the divergence and the
 * correction act on g, the Poisson right-hand side carries drhodt
 * (Calculate_drhodt, src/eq_continuity.f90:26-78) and q = g / d closes the
 * sub-step (:1915) */
void orc_solve_momentum_eq(orc_flow *f, int isub) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Npcc = (long)px * ny * nz, Ncpc = (long)nx * py * nz, Nccp = (long)nx * ny * pz;
  long Nccc = (long)nx * ny * nz;
  double *ux = f->thermo ? f->gx : f->qx, *uy = f->thermo ? f->gy : f->qy;
  double *uz = f->thermo ? f->gz : f->qz;
  double *mx = arr(Npcc), *my = arr(Ncpc), *mz = arr(Nccp);
  orc_momentum_rhs(f, isub, mx, my, mz);
  for (long t = 0; t < Npcc; ++t) ux[t] = ux[t] + mx[t];
  for (long t = 0; t < Ncpc; ++t) uy[t] = uy[t] + my[t];
  for (long t = 0; t < Nccp; ++t) uz[t] = uz[t] + mz[t];
  enforce_velo(f);
  /* solve_poisson, :1663-1718 */
  double *div = arr(Nccc);
  for (long t = 0; t < Nccc; ++t) f->pcor[t] = 0.0;
  if (f->thermo) {
    /* Calculate_drhodt, default (Euler) branch, :70-72.  Update_thermal_
     * properties sets dDensm1 = dDens right after refreshing the density
     * (src/eq_energy.f90:44-45), so this difference is exactly zero */
    for (long t = 0; t < Nccc; ++t) f->pcor[t] = f->dens[t] - f->densm1[t];
    for (long t = 0; t < Nccc; ++t) f->pcor[t] = f->pcor[t] / f->dt;
  }
  orc_divergence(f, div);
  for (long t = 0; t < Nccc; ++t) f->pcor[t] = f->pcor[t] + div[t];
  double coeff = 1.0 / (f->ta[isub] * f->sigma2p * f->dt);
  for (long t = 0; t < Nccc; ++t) f->pcor[t] = f->pcor[t] * coeff;
  orc_poisson_solve(f, f->pcor);
  for (long t = 0; t < Nccc; ++t) f->pres[t] = f->pres[t] + f->pcor[t];
  /* Correct_massflux, :1577-1660 */
  double s = f->dt * f->ta[isub] * f->sigma2p;
  op(f, 0, ORC_1DER_C2P, f->ibc[0][VP], f->pcor, nx, ny, nz, mx, px, ny, nz, f->fbx[VP]);
  for (long t = 0; t < Npcc; ++t) ux[t] = ux[t] - s * mx[t];
  op(f, 1, ORC_1DER_C2P, f->ibc[1][VP], f->pcor, nx, ny, nz, my, nx, py, nz, f->fby[VP]);
  for (long t = 0; t < Ncpc; ++t) uy[t] = uy[t] - s * my[t];
  op(f, 2, ORC_1DER_C2P, f->ibc[2][VP], f->pcor, nx, ny, nz, mz, nx, ny, pz, f->fbz[VP]);
  for (long t = 0; t < Nccp; ++t) uz[t] = uz[t] - s * mz[t];
  enforce_velo(f);
  if (f->thermo) convert_qg(f, 0);
  free(mx); free(my); free(mz); free(div);
}

/* ---------------------------------------------------------------------- */
/* Variable properties (config: heated channel, isothermal walls)          */
/* ---------------------------------------------------------------------- */
/* Update_thermal_properties, src/eq_energy.f90:14-47 (isothermal walls: the
 * Neumann wall branches :51-143 do not apply): every cell from rho*h by
 * ftp_refresh_thermal_properties_from_DH, then dDensm2 = dDensm1,
 * dDensm1 = dDens.  Returns the number of clamped cells. */
static long refresh_props(orc_flow *f) {
  long n = (long)f->nx * f->ny * f->nz, bad = 0;
  for (long t = 0; t < n; ++t) {
    double o[7];
    bad += orc_props_from_dh(f->props, f->rhoh[t], o);
    f->hent[t] = o[0];
    f->temp[t] = o[4];
    f->kcond[t] = o[3];
    f->dens[t] = o[1];
    f->visc[t] = o[2];
  }
  for (long t = 0; t < n; ++t) f->densm2[t] = f->densm1[t];
  for (long t = 0; t < n; ++t) f->densm1[t] = f->dens[t];
  return bad;
}

/* Set up the variable-property path: property table, wall temperatures of
 * the isothermal y walls (non-dimensional; wall properties from T by the
 * table, bc_dirichlet.f90:433-437 with input_thermo.f90:213-241), the
 * reference Prandtl number m0 cp0 / k0 (1/(Re Pr) = rPrRen,
 * tools_solver.f90:63) and a uniform initial state from T (initialise_
 * thermal_properties, input_thermo.f90:1070-1103).  Stationary y walls and
 * x, z periodic only.  Returns 0, or -1 when the configuration is outside
 * that. */
int orc_thermo_enable(orc_flow *f, const orc_props *p, const double t_wall[2], double prandtl, double t_init) {
  if (!f->cfg.wall_y) return -1;
  for (int s2 = 0; s2 < 2; ++s2)
    for (int v = 0; v < 3; ++v)
      if (f->cfg.uwall[s2][v] != 0.0) return -1;
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Nccc = (long)nx * ny * nz;
  f->thermo = 1;
  f->props = p;
  f->prandtl = prandtl;
  for (int s2 = 0; s2 < 2; ++s2) orc_props_from_t(p, t_wall[s2], f->wall[s2]);
  f->gx = arr((long)px * ny * nz); f->gy = arr((long)nx * py * nz); f->gz = arr((long)nx * ny * pz);
  f->dens = arr(Nccc); f->visc = arr(Nccc); f->rhoh = arr(Nccc); f->hent = arr(Nccc);
  f->temp = arr(Nccc); f->kcond = arr(Nccc); f->erhs0 = arr(Nccc); f->densm1 = arr(Nccc);
  f->densm2 = arr(Nccc);
  /* wall arrays (nx,4,nz) of h, k, m, d (rows 3,4 repeat rows 1,2) */
  const int col[4] = {0, 3, 2, 1};
  for (int v = 0; v < 4; ++v) {
    f->fby_ftp[v] = arr((long)nx * 4 * nz);
    for (int k = 0; k < nz; ++k)
      for (int r = 0; r < 4; ++r)
        for (int i = 0; i < nx; ++i) f->fby_ftp[v][i + (long)nx * (r + 4L * k)] = f->wall[r % 2][col[v]];
  }
  double o[7];
  orc_props_from_t(p, t_init, o);
  for (long t = 0; t < Nccc; ++t) {
    f->dens[t] = o[1];
    f->visc[t] = o[2];
    f->rhoh[t] = o[1] * o[0];
    f->hent[t] = o[0];
    f->kcond[t] = o[3];
    f->temp[t] = o[4];
    f->densm1[t] = o[1];
    f->densm2[t] = o[1];
  }
  orc_flow_update_re(f, 0);
  convert_qg(f, 1);
  return 0;
}

/* rho*h given per cell: properties refreshed from it and g = q d of the
 * current velocity.  Returns the number of clamped cells. */
long orc_thermo_set_rhoh(orc_flow *f, const double *rhoh) {
  long n = (long)f->nx * f->ny * f->nz;
  memcpy(f->rhoh, rhoh, sizeof(double) * (size_t)n);
  long bad = refresh_props(f);
  convert_qg(f, 1);
  return bad;
}

/* state of the variable-property path: rho*h, h, T, k, d, mu (ccc each) and
 * g (pcc, cpc, ccp); any pointer may be NULL */
void orc_thermo_get(const orc_flow *f, double *rhoh, double *h, double *t, double *k, double *d, double *m,
                    double *gx, double *gy, double *gz) {
  size_t n = (size_t)f->nx * f->ny * f->nz;
  if (rhoh) memcpy(rhoh, f->rhoh, sizeof(double) * n);
  if (h) memcpy(h, f->hent, sizeof(double) * n);
  if (t) memcpy(t, f->temp, sizeof(double) * n);
  if (k) memcpy(k, f->kcond, sizeof(double) * n);
  if (d) memcpy(d, f->dens, sizeof(double) * n);
  if (m) memcpy(m, f->visc, sizeof(double) * n);
  if (gx) memcpy(gx, f->gx, sizeof(double) * (size_t)f->px * f->ny * f->nz);
  if (gy) memcpy(gy, f->gy, sizeof(double) * (size_t)f->nx * f->py * f->nz);
  if (gz) memcpy(gz, f->gz, sizeof(double) * (size_t)f->nx * f->ny * f->pz);
}

/* Compute_energy_rhs, src/eq_energy.f90:178-396, with Calculate_energy_
 * fractional_step (:148-176): out = dt (gamma rhs + zeta rhs0), rhs0 = rhs.
 * rhs = -d(g_i h_i)/dx_i + rPrRen d(k_i dT/dx_i)/dx_i with h and k
 * interpolated to the faces (midp C2P, property BCs, wall h / k), dT/dx_i
 * by 1der C2P with the T BCs (wall T, get_fbcy_iTh: rows 1,2 only), the
 * flux's wall values taken from its own wall rows (extract_dirichlet_fbcy)
 * before the outer P2C derivative.  The reference guards that last step
 * with is_fbcx_velo_required (:356); on a channel with periodic x that is
 * false and the wall values passed would be MAXP (1e20), turning the wall
 * rows of the y diffusion into ~1e19: the y-wall test
 * (is_fbcy_velo_required) is followed here.  Sums: the x part, then
 * (y part + z part) as in the pencil order of :386-391. */
void orc_energy_rhs(orc_flow *f, int isub, double *rhs) {
  int nx = f->nx, ny = f->ny, nz = f->nz, px = f->px, py = f->py, pz = f->pz;
  long Npcc = (long)px * ny * nz, Ncpc = (long)nx * py * nz, Nccp = (long)nx * ny * pz;
  long Nccc = (long)nx * ny * nz;
  int (*B)[5][2] = f->ibc;
  int econv[3][2], edifu[3][2];
  for (int d = 0; d < 3; ++d) {
    bc_prod(B[d][d], f->ftp[d], econv[d]);
    bc_tau(f->ftp[d], B[d][VT], edifu[d]);
  }
  double *hpcc = arr(Npcc), *hcpc = arr(Ncpc), *hccp = arr(Nccp);
  double *kpcc = arr(Npcc), *kcpc = arr(Ncpc), *kccp = arr(Nccp);
  op(f, 0, ORC_MIDP_C2P, f->ftp[0], f->hent, nx, ny, nz, hpcc, px, ny, nz, NULL);
  op(f, 1, ORC_MIDP_C2P, f->ftp[1], f->hent, nx, ny, nz, hcpc, nx, py, nz, f->fby_ftp[0]);
  op(f, 2, ORC_MIDP_C2P, f->ftp[2], f->hent, nx, ny, nz, hccp, nx, ny, pz, NULL);
  op(f, 0, ORC_MIDP_C2P, f->ftp[0], f->kcond, nx, ny, nz, kpcc, px, ny, nz, NULL);
  op(f, 1, ORC_MIDP_C2P, f->ftp[1], f->kcond, nx, ny, nz, kcpc, nx, py, nz, f->fby_ftp[1]);
  op(f, 2, ORC_MIDP_C2P, f->ftp[2], f->kcond, nx, ny, nz, kccp, nx, ny, pz, NULL);
  double *ex = arr(Nccc), *ey = arr(Nccc), *ez = arr(Nccc), *r = arr(Nccc);
  double *apcc = arr(Npcc), *acpc = arr(Ncpc), *accp = arr(Nccp);
  /* convection */
  for (long t = 0; t < Npcc; ++t) apcc[t] = -f->gx[t] * hpcc[t];
  op(f, 0, ORC_1DER_P2C, econv[0], apcc, px, ny, nz, r, nx, ny, nz, NULL);
  for (long t = 0; t < Nccc; ++t) ex[t] = 0.0 + r[t];
  for (long t = 0; t < Ncpc; ++t) acpc[t] = -f->gy[t] * hcpc[t];
  op(f, 1, ORC_1DER_P2C, econv[1], acpc, nx, py, nz, r, nx, ny, nz, NULL);
  for (long t = 0; t < Nccc; ++t) ey[t] = 0.0 + r[t];
  for (long t = 0; t < Nccp; ++t) accp[t] = -f->gz[t] * hccp[t];
  op(f, 2, ORC_1DER_P2C, econv[2], accp, nx, ny, pz, r, nx, ny, nz, NULL);
  for (long t = 0; t < Nccc; ++t) ez[t] = 0.0 + r[t];
  /* diffusion */
  long nfy = (long)nx * 4 * nz;
  double *fbt = arr(nfy), *fbq = arr(nfy);
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i)
      for (int s2 = 0; s2 < 2; ++s2) fbt[i + (long)nx * (s2 + 4L * k)] = (B[1][VT][s2] == BDIR) ? f->wall[s2][4] : 0.0;
  op(f, 0, ORC_1DER_C2P, B[0][VT], f->temp, nx, ny, nz, apcc, px, ny, nz, NULL);
  for (long t = 0; t < Npcc; ++t) apcc[t] = apcc[t] * kpcc[t];
  op(f, 0, ORC_1DER_P2C, edifu[0], apcc, px, ny, nz, r, nx, ny, nz, NULL);
  for (long t = 0; t < Nccc; ++t) ex[t] = ex[t] + r[t] * f->rprren;
  op(f, 1, ORC_1DER_C2P, B[1][VT], f->temp, nx, ny, nz, acpc, nx, py, nz, fbt);
  for (long t = 0; t < Ncpc; ++t) acpc[t] = acpc[t] * kcpc[t];
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i)
      for (int s2 = 0; s2 < 2; ++s2) {
        double v = acpc[i + (long)nx * ((s2 ? py - 1 : 0) + (long)py * k)];
        fbq[i + (long)nx * (s2 + 4L * k)] = v;
        fbq[i + (long)nx * (s2 + 2 + 4L * k)] = v;
      }
  op(f, 1, ORC_1DER_P2C, edifu[1], acpc, nx, py, nz, r, nx, ny, nz, fbq);
  for (long t = 0; t < Nccc; ++t) ey[t] = ey[t] + r[t] * f->rprren;
  op(f, 2, ORC_1DER_C2P, B[2][VT], f->temp, nx, ny, nz, accp, nx, ny, pz, NULL);
  for (long t = 0; t < Nccp; ++t) accp[t] = accp[t] * kccp[t];
  op(f, 2, ORC_1DER_P2C, edifu[2], accp, nx, ny, pz, r, nx, ny, nz, NULL);
  for (long t = 0; t < Nccc; ++t) ez[t] = ez[t] + r[t] * f->rprren;
  for (long t = 0; t < Nccc; ++t) ey[t] = ey[t] + ez[t];
  for (long t = 0; t < Nccc; ++t) rhs[t] = ex[t] + ey[t];
  /* fractional step */
  for (long t = 0; t < Nccc; ++t) {
    double cur = rhs[t], last = f->erhs0[t];
    double tot = f->tg[isub] * cur + f->tz[isub] * last;
    f->erhs0[t] = cur;
    rhs[t] = f->dt * tot;
  }
  free(hpcc); free(hcpc); free(hccp); free(kpcc); free(kcpc); free(kccp);
  free(ex); free(ey); free(ez); free(r); free(apcc); free(acpc); free(accp); free(fbt); free(fbq);
}

/* Solve_energy_eq, src/eq_energy.f90:398-458: (isub 3: density history
 * shifted), energy right-hand side, rho*h += rhs, properties refreshed.
 * Returns the number of clamped cells of the refresh. */
long orc_solve_energy_eq(orc_flow *f, int isub) {
  long n = (long)f->nx * f->ny * f->nz;
  if (isub == 3) {
    for (long t = 0; t < n; ++t) f->densm2[t] = f->densm1[t];
    for (long t = 0; t < n; ++t) f->densm1[t] = f->dens[t];
  }
  double *r = arr(n);
  orc_energy_rhs(f, isub, r);
  for (long t = 0; t < n; ++t) f->rhoh[t] = f->rhoh[t] + r[t];
  free(r);
  return refresh_props(f);
}

/* one time step of Solve_eqs_iteration, src/chapsim.f90:243-303: per
 * sub-step the energy equation (variable properties) then the momentum
 * equation (:300-304) */
void orc_flow_step(orc_flow *f, int iter) {
  orc_flow_update_re(f, iter);
  for (int isub = 1; isub <= 3; ++isub) {
    if (f->thermo) orc_solve_energy_eq(f, isub);
    orc_solve_momentum_eq(f, isub);
  }
}
