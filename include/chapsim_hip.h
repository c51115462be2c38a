/*
 * chapsim_hip.h -- C ABI of the MI355X (gfx950) implementation of the
 * CHAPSim2 per-timestep hot path.
 *
 * The reference (weiwangstfc/CHAPSim2) is a Fortran + MPI code with no plugin
 * API; its hot path is a set of Fortran module procedures.  Each entry point
 * below names the reference procedure it replaces (file:line in the CHAPSim2
 * tree).  A Fortran host binds these through iso_c_binding (see
 * chapsim2_amd/fortran/chapsim_hip.f90).
 *
 * Conventions
 *  - All arrays are Fortran order (x fastest), FP64, with the staggered
 *    extents of the reference: qx (np1,nc2,nc3), qy (nc1,np2,nc3),
 *    qz (nc1,nc2,np3), p (nc1,nc2,nc3); np = nc for a periodic direction,
 *    nc+1 otherwise (src/geometry.f90:434-446).
 *  - Boundary values fbc use the reference's per-direction layouts
 *    (4,n2,n3) / (n1,4,n3) / (n1,n2,4) (src/basics_operations2.f90:3952-4790).
 *  - `mem` = CHS_MEM_HOST for host pointers (copied in and out) or
 *    CHS_MEM_DEVICE for device pointers on the handle's GPU.
 *  - Every function returns 0 on success or a negative CHS_E* code; the
 *    message of the last error of the calling thread is chs_last_error().
 *    The library never aborts and never prints.
 *  - One host thread per handle; calls are synchronous.
 */
#ifndef CHAPSIM_HIP_H
#define CHAPSIM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CHS_ABI_VERSION 2

enum {
  CHS_OK = 0,
  CHS_EINVAL = -1,   /* bad argument, shape or boundary kind */
  CHS_ENOMEM = -2,   /* device allocation failed */
  CHS_EHIP = -3,     /* HIP runtime error */
  CHS_ENAN = -4,     /* non-finite value detected */
  CHS_EUNSUP = -5,   /* configuration outside the supported set */
  CHS_ERANGE = -6    /* input outside the tabulated range (values were clamped) */
};

enum { CHS_MEM_HOST = 0, CHS_MEM_DEVICE = 1 };

/* boundary kinds, src/modules.f90:223-234 */
enum { CHS_BC_INTERIOR = 0, CHS_BC_PERIODIC = 1, CHS_BC_SYMMETRIC = 2, CHS_BC_ASYMMETRIC = 3,
       CHS_BC_DIRICHLET = 4, CHS_BC_NEUMANN = 5, CHS_BC_INTRPL = 6 };
/* accuracies, src/modules.f90:248-251 */
enum { CHS_CD2 = 1, CHS_CD4 = 2, CHS_CP4 = 3, CHS_CP6 = 4 };
/* operator kinds of chs_op3d (the 18 Get_* procedures are kind x direction) */
enum { CHS_MIDP_C2P = 0, CHS_MIDP_P2C = 1, CHS_1DER_C2C = 2, CHS_1DER_P2P = 3,
       CHS_1DER_C2P = 4, CHS_1DER_P2C = 5 };
/* precision of the step-level solver */
enum { CHS_PREC_FP64 = 0, CHS_PREC_MIXED = 1 };
/* Poisson path, src/modules.f90:169-173 */
enum { CHS_POISSON_FISHPACK = 2, CHS_POISSON_3DFFT = 3 };
/* pressure correction of a sub-step:
 *  REFERENCE - one solve with the configured Poisson path (the 2nd-order
 *              operator of src/poisson_fishpack.f90, or the 3-D FFT path
 *              with modified wavenumbers), as the reference does;
 *  EXACT     - inverts the composed operator D*G of the configured compact
 *              schemes (FFT in x and z, eigen-basis of the wall-normal block),
 *              so one correction leaves max|div u| at round-off.  No
 *              counterpart in the reference. */
enum { CHS_PROJ_REFERENCE = 0, CHS_PROJ_EXACT = 1 };

/* Domain/solver description (bind(C) on the Fortran side).  Replaces the
 * t_domain/t_flow set-up of src/geometry.f90:410-593, src/bc_general.f90
 * :137-757 and src/input_general.f90:842-888. */
typedef struct chs_desc {
  int nc[3];            /* cells in x, y, z */
  int periodic_y;       /* 1: periodic y (x and z are always periodic) */
  double lxx, lyb, lyt, lzz;
  int istret;           /* 0 uniform; 2 two-sided, 3 bottom, 4 top tanh */
  double rstret;        /* tanh beta = 20 * rstret (src/geometry.f90:237) */
  int iacc;             /* CHS_CD2 .. CHS_CP6 */
  double dt;
  double ren, reninit;  /* Reynolds numbers (src/tools_solver.f90:30-47) */
  int init_re_to;
  int idriven;          /* 0 none, 1 constant x mass flux */
  int poisson;          /* CHS_POISSON_* */
  double uwall[2][3];   /* y-wall velocities (bottom, top) x (u, v, w) */
  int precision;        /* CHS_PREC_* */
  int projection;       /* CHS_PROJ_* */
} chs_desc_t;

typedef struct chs_handle chs_handle_t;

/* Library / handle lifecycle.  Replaces Prepare_LHS_coeffs_for_operations
 * (src/basics_operations2.f90:1600-1800), initialise_fft
 * (src/poisson_interface.f90:102-116) and Allocate_flow_variables. */
int chs_abi_version(void);
const char *chs_last_error(void);
int chs_create(const chs_desc_t *desc, chs_handle_t **out);
int chs_destroy(chs_handle_t *h);
int chs_sync(chs_handle_t *h);

/* One compact operator on a 3-D array.  Replaces the 18
 * Get_{x,y,z}_{midp_C2P,midp_P2C,1der_C2C,1der_P2P,1der_C2P,1der_P2C}_3D
 * (fi3d, fo3d, dm, iacc, ibc, fbc2d) of src/basics_operations2.f90:270-312,
 * bodies :3952-4790.  dir = 0,1,2 for x,y,z; fbc may be NULL ("absent"). */
int chs_op3d(chs_handle_t *h, int dir, int op, int iacc, const int ibc[2], const double *fi,
             const int shp_in[3], double *fo, const int shp_out[3], const double *fbc, int mem);

/* Pressure Poisson solve in place on a (nc1,nc2,nc3) array.  Replaces
 * solve_fft_poisson (src/poisson_interface.f90:119-149) ->
 * fishpack_fft_simple (src/poisson_fishpack.f90:354-490), poisson_000
 * (src/poisson_1stderivcomp.f90:1016-1195, triply periodic) or poisson_010
 * (:1457-1816, y walls on a uniform mesh, even nc2; CHS_POISSON_3DFFT). */
int chs_poisson(chs_handle_t *h, double *rhs, int mem);

/* Device-resident state exchange (the reference keeps qx, qy, qz, pres in
 * t_flow; restart files src/io_restart.f90:112-171 hold the same arrays). */
int chs_set_state(chs_handle_t *h, const double *qx, const double *qy, const double *qz,
                  const double *pr, int mem);
int chs_get_state(chs_handle_t *h, double *qx, double *qy, double *qz, double *pr, int mem);

/* Set 1/Re for iteration iter.  Replaces Update_Re (src/tools_solver.f90:30-47). */
int chs_update_re(chs_handle_t *h, int iter);

/* One RK3 sub-step (isub = 1..3) or the initial Euler step (isub = 0).
 * Replaces Solve_momentum_eq(fl, dm, isub) (src/eq_momentum2.f90:1799-1932). */
int chs_substep(chs_handle_t *h, int isub);

/* nsteps full time steps starting at iteration iter0+1 (Update_Re, three
 * sub-steps).  Replaces the RK3 loop of Solve_eqs_iteration
 * (src/chapsim.f90:243-303). */
int chs_run(chs_handle_t *h, int iter0, int nsteps);

/* Momentum right-hand sides of sub-step isub after the fractional-step
 * combination and forcing (the mx_rhs/my_rhs/mz_rhs of
 * Compute_momentum_rhs, src/eq_momentum2.f90:101-1571).  Advances rhs0. */
int chs_momentum_rhs(chs_handle_t *h, int isub, double *mx, double *my, double *mz, int mem);

/* max|div u| over the inlet, bulk and outlet x windows.  Replaces
 * Check_element_mass_conservation (src/eq_continuity.f90:337-397). */
int chs_mass_check(chs_handle_t *h, double out[3]);

/* out[0]: convective CFL number max|u dt/dx + v dt/dy + w dt/dz| of the
 * current velocity (P2C interpolations to cell centres, y spacing through the
 * stretching map); out[1]: diffusion number.  Replaces Check_cfl_convection
 * and Check_cfl_diffusion (src/tools_solver.f90:220-375; Cartesian, constant
 * properties), which print these values; the caller prints or checks them. */
int chs_cfl(chs_handle_t *h, double out[2]);

/* out[0]: volumetric average of 1/2 (u^2 + v^2 + w^2) (velocity interpolated
 * to the cell centres), out[1]: volumetric average of qx.  The bulk values
 * the reference's monitor writes (write_monitor_bulk,
 * src/io_monitor.f90:207-229, with Get_volumetric_average_3d_for_var_xcx,
 * src/tools_general.f90:1202-1268). */
int chs_bulk(chs_handle_t *h, double out[2]);

/* out = {max qx, min qx, max qy, min qy, max qz, min qz} of the current
 * state: the per-step check of Find_max_min_3d (src/tools_general.f90:
 * 1071-1113, called at src/chapsim.f90:329-331), whose comparisons skip NaN
 * and start from -1e20 / 1e20.  Returns CHS_ENAN (with out filled) when any
 * velocity value is NaN or infinite, so a caller can stop a run that has
 * blown up. */
int chs_velocity_range(chs_handle_t *h, double out[6]);

/* chs_velocity_range with no allocation of any kind on the device (no
 * counterpart in the reference): for callers that cannot allocate, such as
 * tight stepping loops and harnesses that forbid hipMalloc.  The block
 * partials of the reduction live in a static device array of the library
 * (3 x 512 doubles per device, kept until the process ends) and come back
 * through a host stack buffer; the field size is not limited by it.  Same
 * results and return codes as chs_velocity_range (CHS_ENAN with out filled
 * when a value is NaN or infinite).  Every handle on one device shares the
 * array: calls that overlap on one device race. */
int chs_velocity_range_noalloc(chs_handle_t *h, double out[6]);

/* Divergence of the current velocity, (nc1,nc2,nc3).  Replaces
 * Get_divergence_flow (src/eq_continuity.f90:91-161). */
int chs_divergence(chs_handle_t *h, double *div, int mem);

/* ---- batched channels (no counterpart in the reference, which runs one
 * channel per job) ---------------------------------------------------------
 * n independent instances, each created from its own descriptor (uneven
 * grids allowed).  chs_batch_run advances every instance nsteps RK3 steps
 * (Update_Re + three sub-steps each); instances run on their own streams,
 * and with use_graph the whole RK3 step of all instances is one captured
 * graph (captured after one eager step, re-captured when 1/Re changes).
 * chs_batch_instance gives the per-instance handle for chs_set_state,
 * chs_get_state, chs_update_re, chs_substep and chs_mass_check; it stays
 * owned by the batch. */
typedef struct chs_batch chs_batch_t;
int chs_batch_create(const chs_desc_t *descs, int n, chs_batch_t **out);
int chs_batch_destroy(chs_batch_t *b);
chs_handle_t *chs_batch_instance(chs_batch_t *b, int i);
int chs_batch_run(chs_batch_t *b, int iter0, int nsteps, int use_graph);

/* ---- thermal properties from a tabulated equation of state ------------- */

typedef struct chs_props chs_props_t;
typedef struct chs_mlp chs_mlp_t;
/* property evaluation: the table's interpolation, or the MLP surrogate
 * attached to the table (chs_props_set_surrogate) for the properties from h */
enum { CHS_PROPS_TABLE = 0, CHS_PROPS_MLP = 1 };

/* Build the non-dimensional property table from nrows dimensional rows of
 * (P, H, T, D, M, K, CP, B) (the columns of the reference's NIST tables):
 * sort by T, reference state by interpolation at ref_t0, scaling, rho*h and
 * its range.  Replaces buildup_property_relations_from_table
 * (src/input_thermo.f90:617-693) with ftplist_sort_t_small2big (:512-556)
 * and ftp_get_thermal_properties_dimensional_from_T (:121-154). */
int chs_props_create(const double *table, int nrows, double ref_t0, chs_props_t **out);
int chs_props_destroy(chs_props_t *p);
/* dhmin, dhmax (non-dimensional rho*h range) and the dimensional reference
 * state ref = (T0, d, m, k, h, b, cp); any pointer may be NULL. */
int chs_props_info(const chs_props_t *p, double *dhmin, double *dhmax, double ref[7]);

/* Per-cell properties from rho*h: clamp to [dhmin, dhmax], bisection on
 * rho*h, h by linear interpolation, properties from h, h = rho*h / d,
 * properties from h again.  Replaces ftp_refresh_thermal_properties_from_DH
 * (src/input_thermo.f90:393-446, with ftp_refresh_thermal_properties_from_H
 * :345-374) applied to every cell as Update_thermal_properties does
 * (src/eq_energy.f90:14-47).  Any output may be NULL.  Returns CHS_ERANGE
 * (after writing the clamped results) when an input was out of range, where
 * the reference stops with an error. */
int chs_props_from_rhoh(chs_props_t *p, const double *rhoh, long n, double *h, double *d, double *m, double *k,
                        double *t, double *b, double *cp, int mode, int mem);
/* Attach (or with NULL detach) the MLP surrogate that mode CHS_PROPS_MLP of
 * chs_props_from_rhoh and of the flow solver's property refresh use: the
 * first h still comes from the table's rho*h column, the properties from h
 * (both evaluations of ftp_refresh_thermal_properties_from_H) from the
 * network; b and cp are not given in that mode.  No counterpart in the
 * reference. */
int chs_props_set_surrogate(chs_props_t *p, chs_mlp_t *m);

/* Per-cell properties from h: bisection on h and linear interpolation, with
 * no clamping (the end segments extrapolate).  Replaces
 * ftp_refresh_thermal_properties_from_H (src/input_thermo.f90:345-374).
 * A table created with ref_t0 <= 0 is taken as already non-dimensional. */
int chs_props_from_h(chs_props_t *p, const double *h, long n, double *d, double *m, double *k, double *t,
                     double *b, double *cp, int mem);

/* ---- MLP surrogate of the property table (no counterpart in the reference)
 * h -> (T, rho, mu, kappa, cp); 1 -> 64 -> 64 -> 5, tanh; BF16 weights, FP32
 * accumulation; trained on the GPU on nsamples samples of the table's own
 * interpolation in h (uniform over the table's h range).  Parameters
 * (4613, FP32 master copy): W1[64], b1[64], W2[64][64] (row = output unit),
 * b2[64], W3[5][64], b3[5]; inputs and outputs are normalised with
 * norm = (x centre, x scale, 5 output means, 5 output scales). */
int chs_mlp_create(chs_props_t *table, int nsamples, unsigned seed, chs_mlp_t **out);
int chs_mlp_destroy(chs_mlp_t *m);
int chs_mlp_info(const chs_mlp_t *m, int *nparams, float norm[12]);
int chs_mlp_get_params(chs_mlp_t *m, float *p32, unsigned short *p16);
int chs_mlp_set_params(chs_mlp_t *m, const float *p32);
/* Mean squared error over B given samples (normalised inputs x[B], targets
 * t[B][5]) and its gradient with respect to the parameters, evaluated with
 * the BF16 weights. */
int chs_mlp_grad(chs_mlp_t *m, const float *x, const float *t, int B, float *grad, float *loss);
/* steps Adam steps on random minibatches of the training set; loss of the
 * last minibatch. */
int chs_mlp_train(chs_mlp_t *m, int steps, int batch, float lr, float *loss);
/* Surrogate evaluation for n enthalpies; any output may be NULL. */
int chs_mlp_eval(chs_mlp_t *m, const double *h, long n, double *T, double *rho, double *mu, double *kappa,
                 double *cp, int mem);

/* ---- variable-property flow (heated channel) --------------------------
 * Isothermal y walls (Dirichlet T, stationary), x and z periodic: the
 * energy equation and the variable-density / variable-viscosity momentum
 * terms of the reference's thermo path.  Each sub-step of chs_substep /
 * chs_run then follows src/chapsim.f90:300-304: Solve_energy_eq
 * (src/eq_energy.f90:398-458: Compute_energy_rhs :178-396, rho*h += rhs,
 * Update_thermal_properties :14-47) and Solve_momentum_eq with g = rho u
 * (src/eq_momentum2.f90:576-703 and the thermo branches of :798-1507,
 * 1799-1932; convert_primary_conservative, src/para_conversion.f90:5-227).
 * The initial sub-step (isub 0) is momentum only, as at src/chapsim.f90:135.
 * FP64 and mixed precision (CHS_PREC_MIXED: the momentum and energy
 * right-hand sides in FP32; the state, properties and projection FP64). */
typedef struct chs_thermo_desc {
  chs_props_t *props;  /* property table (kept by the caller while the handle lives) */
  int mode;            /* CHS_PROPS_TABLE or CHS_PROPS_MLP */
  double t_wall[2];    /* non-dimensional wall temperatures, y bottom / top */
  double prandtl;      /* reference Prandtl number mu0 cp0 / k0: 1/(Re Pr) = rPrRen (src/tools_solver.f90:63) */
  double t_init;       /* uniform initial temperature (initialise_thermal_properties,
                          src/input_thermo.f90:1070-1103) */
} chs_thermo_desc_t;
/* Enable the path on a flow handle (walls and wall properties from T by the
 * table, src/bc_dirichlet.f90:433-437; uniform initial state; g from the
 * current velocity).  CHS_EUNSUP for moving walls or no y walls. */
int chs_thermo_enable(chs_handle_t *h, const chs_thermo_desc_t *d);
/* rho*h per cell (nc1,nc2,nc3): properties refreshed from it, g = rho u of
 * the current velocity.  CHS_ERANGE when cells were clamped. */
int chs_thermo_set_rhoh(chs_handle_t *h, const double *rhoh, int mem);
/* switch the property refresh between table and surrogate */
int chs_thermo_set_mode(chs_handle_t *h, int mode);
/* rho*h, h, T, k, rho, mu (cell centres) and gx, gy, gz (the velocity
 * locations); any pointer may be NULL */
int chs_thermo_get(chs_handle_t *h, double *rhoh, double *hh, double *t, double *k, double *d, double *m,
                   double *gx, double *gy, double *gz, int mem);
/* dt-scaled energy right-hand side of sub-step isub (Compute_energy_rhs with
 * Calculate_energy_fractional_step, which also stores it as the next
 * sub-step's rhs0), cell centres */
int chs_energy_rhs(chs_handle_t *h, int isub, double *rhs, int mem);

#ifdef __cplusplus
}
#endif
#endif
