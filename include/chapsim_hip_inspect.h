/*
 * chapsim_hip_inspect.h -- host-only inspection entry points of
 * libchapsim_hip.  They touch no GPU: they expose the operator planner (the
 * scheme tables of src/basics_operations2.f90:361-1509 and the per-line plans
 * the kernels execute) so a build without a GPU can be checked.
 *
 * These have no counterpart in the reference; they are for test harnesses
 * and for maintainers verifying a port of the tables.
 */
#ifndef CHAPSIM_HIP_INSPECT_H
#define CHAPSIM_HIP_INSPECT_H

#include "chapsim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scheme table entry: family 0..5 (d1 C2C, d1 P2P, d1 C2P, d1 P2C, midp C2P,
 * midp P2C), is_rhs 0/1, row 1..5, k 1..3 (LHS) or 1..6 (RHS), bc 0..6,
 * acc 1..4. */
double chs_inspect_coeff(int fam, int is_rhs, int row, int k, int bc, int acc);

/* Apply the planned operator to one line on the host (same arithmetic as the
 * GPU kernel).  The mesh is built from desc; fbc may be NULL.  Returns 0 or a
 * negative CHS_E* code (message in chs_last_error()). */
int chs_inspect_op1d(const chs_desc_t *desc, int dir, int op, int iacc, const int ibc[2],
                     const double *fi, int n_in, double *fo, int n_out, const double *fbc);

/* Operators of the exact projection for the channel described by desc (host
 * only): ly = D_y G_y (ny*ny, row-major) with the wall faces held, its
 * eigenvalues lam (ny, ascending), the Fourier eigenvalues lx (nx/2+1) and lz
 * (nz) of the periodic D G blocks, and the decomposition residual
 * max|V^-1 Ly V - diag(lam)| / max|Ly|.  Any output pointer may be NULL. */
int chs_inspect_projection(const chs_desc_t *desc, double *ly, double *lam, double *lx, double *lz,
                           double *resid);

/* Launch timing of one kernel class during subsequent solver and chs_op3d
 * calls, with HIP events on the handle's stream: class 0/1/2 = x/y/z line
 * operators, 3 = whole chs_poisson / chs_poisson_f32 solves (counted as five
 * passes that each read and write the field).  While a class is profiled,
 * sub-steps issue every kernel on the handle's stream, one at a time (they
 * otherwise overlap independent operators on side streams), so each timed
 * launch has the GPU to itself.  chs_profile_end returns the summed
 * event time, the number of timed launches and their algorithmic bytes
 * (arrays read plus arrays written, 8 bytes per FP64 and 4 per FP32
 * element). */
int chs_profile_begin(chs_handle_t *h, int kernel_class);
int chs_profile_end(chs_handle_t *h, double *ms_total, long long *launches, double *alg_bytes);

/* FP32 forms of chs_op3d and chs_poisson, on device memory only: the same
 * operator plans and kernels in single precision, as the mixed-precision
 * solver runs them.  For the operator benchmark and precision tests; no
 * counterpart in the reference, which is FP64 throughout.  fbc may be NULL;
 * rhs is solved in place (cell-centred, nx*ny*nz). */
int chs_op3d_f32(chs_handle_t *h, int dir, int op, int iacc, const int ibc[2], const float *fi,
                 const int shp_in[3], float *fo, const int shp_out[3], const float *fbc);
int chs_poisson_f32(chs_handle_t *h, float *rhs);

/* Launch shapes chosen so far by the autotuner (FFT lines per workgroup,
 * x-line tile), one "family/shape=choice (us per variant)" line each, for
 * every handle of the process.  Writes at most len-1 characters and a
 * terminating 0; returns the number of tuned shapes.  No counterpart in the
 * reference, which has no tunable launch shapes. */
int chs_tune_report(char *buf, int len);

/* While variant >= 0, every autotuned launch whose family key starts with
 * prefix (e.g. "line_x/") takes launch shape `variant` instead of the timed
 * choice (the next shape that applies when it does not); variant < 0 restores
 * tuning.  Process-wide.  For tests that check every launch shape against
 * the oracle; no counterpart in the reference. */
int chs_tune_force(const char *prefix, int variant);

/* Presets the launch shape of one tuned key (a "key" of chs_tune_report) as
 * if it had been timed, so that a run under a profiler takes the shapes an
 * earlier run chose.  No counterpart in the reference. */
int chs_tune_set(const char *key, int choice);

/* Selects how the surrogate evaluates the network in the property refresh
 * (chs_props_from_rhoh and the flow solver's refresh in mode
 * CHS_PROPS_MLP): on the matrix cores (on != 0, the default: BF16
 * activations, FP32 accumulation) or with vector FMAs on FP32 activations.
 * For tests that compare the two; no counterpart in the reference. */
int chs_mlp_set_matrix_cores(chs_mlp_t *m, int on);

#ifdef __cplusplus
}
#endif
#endif
