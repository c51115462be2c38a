/*
 * CPU oracle for the CHAPSim2 per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference solver's algorithms (Fortran, weiwangstfc/CHAPSim2), written to
 * check the HIP product against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links it.
 *
 * Every function cites the reference routine it follows (file:line).  Pinning
 * of each part (golden vectors, closed forms, brute force) is listed in the
 * header comment of each source file; parts without such a check say
 * "parity unpinned" there.
 */
#ifndef CHS_ORACLE_H
#define CHS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary kinds, src/modules.f90:223-234 */
enum { ORC_BC_INTERIOR = 0, ORC_BC_PERIODIC = 1, ORC_BC_SYMMETRIC = 2,
       ORC_BC_ASYMMETRIC = 3, ORC_BC_DIRICHLET = 4, ORC_BC_NEUMANN = 5,
       ORC_BC_INTRPL = 6 };
/* accuracies, src/modules.f90:248-251 */
enum { ORC_CD2 = 1, ORC_CD4 = 2, ORC_CP4 = 3, ORC_CP6 = 4 };
/* operator kinds (the six 1-D families of src/basics_operations2.f90) */
enum { ORC_MIDP_C2P = 0, ORC_MIDP_P2C = 1, ORC_1DER_C2C = 2,
       ORC_1DER_P2P = 3, ORC_1DER_C2P = 4, ORC_1DER_P2C = 5 };
/* coefficient families for orc_coeff */
enum { ORC_FAM_D1C2C = 0, ORC_FAM_D1P2P = 1, ORC_FAM_D1C2P = 2,
       ORC_FAM_D1P2C = 3, ORC_FAM_M1C2P = 4, ORC_FAM_M1P2C = 5 };

typedef struct orc_dom {
  int nc[3];        /* cells */
  int np[3];        /* computational points (nc if periodic else nc+1) */
  int npgeo[3];     /* nc+1 */
  int periodic[3];
  int stretching;   /* y stretched */
  int istret;
  double rstret;
  double lxx, lyb, lyt, lzz;
  double h[3], h1r[3], h2r[3];
  double *yp, *yc;  /* node / centre coordinates in y */
  double *ymp, *ymc;/* yMappingpt(:,1), yMappingcc(:,1) = ds/dy */
} orc_dom;

/* ---- tables / TDMA ---------------------------------------------------- */
double orc_coeff(int fam, int is_rhs, int row, int k, int bc, int acc);
void orc_tdma_preprocess(const double *a, const double *b, double *c,
                         double *d, int n);
void orc_tdma_solve(int peri, double *x, const double *a, const double *b,
                    const double *c, const double *d, int n);
void orc_lhs(int fam, int n, int bc1, int bc2, int acc, double *a, double *b,
             double *c, double *d);

/* ---- domain ----------------------------------------------------------- */
int orc_dom_init(orc_dom *dm, const int nc[3], const int periodic[3],
                 double lxx, double lyb, double lyt, double lzz, int istret,
                 double rstret);
void orc_dom_free(orc_dom *dm);

/* ---- operators -------------------------------------------------------- */
/* One 1-D line, reference Get_{x,y,z}_*_1D.  fbc may be NULL (absent). */
void orc_op1d(const orc_dom *dm, int dir, int op, int iacc, const int ibc[2],
              const double *fi, int nin, double *fo, int nout,
              const double *fbc);
/* 3-D driver, reference Get_{x,y,z}_*_3D.  Arrays are Fortran order (x
 * fastest).  shp_in/shp_out are the three extents; fbc (may be NULL) has the
 * per-direction layout (4,n2,n3) / (n1,4,n3) / (n1,n2,4). */
int orc_op3d(const orc_dom *dm, int dir, int op, int iacc, const int ibc[2],
             const double *fi, const int shp_in[3], double *fo,
             const int shp_out[3], const double *fbc);

/* ---- FFTPACK-convention real transforms -------------------------------- */
void orc_rfftf(int n, double *r);
void orc_rfftb(int n, double *r);

/* ---- flow solver ------------------------------------------------------ */
typedef struct orc_flow orc_flow;

typedef struct orc_flow_cfg {
  int nc[3];
  double lxx, lyb, lyt, lzz;
  int istret;          /* 0 uniform, 2 two-sided tanh */
  double rstret;
  int wall_y;          /* 1: y walls (u,v,w Dirichlet; p Neumann); 0: periodic y */
  double uwall[2][3];  /* wall velocities bottom/top (u,v,w) */
  int iacc;
  double dt;
  double ren, reninit;
  int init_re_to;
  int idriven;         /* 0 none, 1 x mass flux */
  int poisson;         /* 2 = FISHPACK 2-D FFT + y TDMA; 3 = 3-D FFT (triply periodic) */
  int projection;      /* 0 reference correction; 1 exact D*G inverse (y walls only) */
} orc_flow_cfg;

orc_flow *orc_flow_create(const orc_flow_cfg *cfg);
void orc_flow_destroy(orc_flow *f);
const orc_dom *orc_flow_dom(const orc_flow *f);
void orc_flow_set_state(orc_flow *f, const double *qx, const double *qy,
                        const double *qz, const double *pr);
void orc_flow_get_state(const orc_flow *f, double *qx, double *qy, double *qz,
                        double *pr);
void orc_flow_get_rhs0(const orc_flow *f, double *mx, double *my, double *mz);
void orc_flow_update_re(orc_flow *f, int iter);
void orc_solve_momentum_eq(orc_flow *f, int isub);
void orc_project(orc_flow *f, int isub);
void orc_flow_step(orc_flow *f, int iter);
void orc_divergence(orc_flow *f, double *div);
void orc_mass_check(orc_flow *f, double out[3]);
void orc_cfl(orc_flow *f, double out[2]);
void orc_bulk(orc_flow *f, double out[2]);
void orc_momentum_rhs(orc_flow *f, int isub, double *mx, double *my,
                      double *mz);
void orc_poisson_solve(orc_flow *f, double *rhs);
double orc_vol_average_x(orc_flow *f, const double *var, int nx);
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- thermal-property table (src/input_thermo.f90:121-446, 512-693) ---- */
#define ORC_PROPS_MAX 16384
typedef struct orc_props {
  int n;
  double t[ORC_PROPS_MAX], d[ORC_PROPS_MAX], m[ORC_PROPS_MAX], k[ORC_PROPS_MAX];
  double b[ORC_PROPS_MAX], cp[ORC_PROPS_MAX], h[ORC_PROPS_MAX], rhoh[ORC_PROPS_MAX];
  double ref[7]; /* T0, d, m, k, h, b, cp of the reference state (dimensional) */
  double dhmin, dhmax;
  int dimensional; /* 0: rows were given non-dimensional (ref_t0 <= 0) */
} orc_props;
/* raw: nlist rows of (P, H, T, D, M, K, CP, B), dimensional */
int orc_props_build(const double *raw, int nlist, double ref_t0, orc_props *p);
void orc_props_from_h(const orc_props *p, double h, double out[6]);
int orc_props_from_dh(const orc_props *p, double rhoh, double out[7]);
void orc_props_from_t(const orc_props *p, double t, double out[7]);

/* variable-property path (heated channel with isothermal y walls) */
int orc_thermo_enable(orc_flow *f, const orc_props *p, const double t_wall[2], double prandtl,
                      double t_init);
long orc_thermo_set_rhoh(orc_flow *f, const double *rhoh);
void orc_thermo_get(const orc_flow *f, double *rhoh, double *h, double *t, double *k, double *d, double *m,
                    double *gx, double *gy, double *gz);
void orc_energy_rhs(orc_flow *f, int isub, double *rhs);
long orc_solve_energy_eq(orc_flow *f, int isub);

#ifdef __cplusplus
}
#endif
#endif
