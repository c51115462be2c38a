// Device-resident flow solver of one channel / periodic box.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "chs_handle.h"
#include "chs_poisson.h"
#include "chs_props.h"

namespace chs {

// block partials used by the max reductions (the partial buffer holds at
// least this many entries)
constexpr int kMaxPartials = 512;

// A device array with its Fortran extents.
template <typename T>
struct Field {
  T* p = nullptr;
  Ext3 e{{0, 0, 0}};
  long size() const { return (long)e.n[0] * e.n[1] * e.n[2]; }
};

// Everything the momentum right-hand side reads and writes, in the precision
// T it is computed in.  For T = double the velocity / pressure fields alias
// the solver state; for T = float (mixed form) they are rounded copies made
// at the start of each sub-step.
template <typename T>
struct RhsWork {
  Field<T> qx, qy, qz, pr;
  Field<T> mx0, my0, mz0;  // previous sub-step's right-hand side (rhs0)
  // first-level intermediates
  Field<T> qxix, qxdx, qxdy, qxiy, qxiz, qxdz;
  Field<T> qydx, qyix, qydy, qyiy, qydz, qyiz;
  Field<T> qzdx, qzix, qziy, qzdy, qziz, qzdz, div;
  // accumulators per momentum component: [comp][x-part, y-part, z-part, pfc]
  Field<T> acc[3][4];
  // variable properties: g = rho u and mu (FP64: the state itself; FP32:
  // rounded copies), the interpolated g of each convective product, the
  // interpolated viscosities and the wall viscosity (nx,4,nz)
  Field<T> gx, gy, gz, visc;
  Field<T> gxix, gxiy, gxiz, gyix, gyiy, gyiz, gzix, gziy, gziz;
  Field<T> muiy, mupcc, muixy, muixz, muiyz;
  T* fby_mu = nullptr;
  T* fby_mwall = nullptr;  // wall viscosity of the y interpolation (rows 3,4 = rows 1,2)
  // wall data (nx,4,nz) and wall-plane scratch
  T *fby_u = nullptr, *fby_v = nullptr, *fby_w = nullptr, *fby_p = nullptr, *fby_pn = nullptr;
  T *fby_conv = nullptr, *fby_visc = nullptr, *fby_div = nullptr;
  T *plane_a = nullptr, *plane_b = nullptr, *plane_c = nullptr, *plane_d = nullptr;
  T* scratch = nullptr;
  T* scratch_side[2] = {nullptr, nullptr};  // the same for operators on the side streams
  long scratch_n = 0;
  bool ready = false;
};

// The energy right-hand side's operands in the precision T its operators
// run in: h, k and T at the cell centres (FP64: the Flow's own fields; FP32:
// rounded copies, stored by the property refresh), the x, y, z
// parts of the right-hand side, h, k and dT/dx at the faces, and the wall
// rows (nx,4,nz) of h, k, T and of the conductive flux k dT/dy.
template <typename T>
struct EnergyWork {
  Field<T> hent, kcond, temp;
  Field<T> ex, ey, ez;
  Field<T> hpcc, hcpc, hccp, kpcc, kcpc, kccp, tpcc, tcpc, tccp;
  T *fby_h = nullptr, *fby_k = nullptr, *fby_t = nullptr, *fby_q = nullptr;
};

class Flow {
 public:
  static std::string create(Handle* H, std::unique_ptr<Flow>* out);
  ~Flow();

  std::string set_state(const double* qx, const double* qy, const double* qz, const double* pr, int mem);
  std::string get_state(double* qx, double* qy, double* qz, double* pr, int mem);
  void update_re(int iter);
  std::string substep(int isub);
  // leaves the dt-scaled rhs in acc_x of each component; with fuse it
  // also applies the state update (the stored rhs is then only valid for x
  // of a forced flow)
  std::string momentum_rhs(int isub, bool fuse = false);
  std::string divergence(double* div_dev);
  std::string mass_check(double out[3]);
  std::string cfl(double out[2]);   // convective CFL and diffusion numbers
  std::string bulk_monitor(double out[2]);  // bulk mean kinetic energy and bulk qx
  // max and min of qx, qy, qz (out[2c], out[2c+1]) and the count of
  // non-finite values among them
  std::string velocity_range(double out[6], long* nonfinite);
  // the same with the block partials in the library's static per-device
  // array (range_scratch) and the host copy on the stack: no allocation
  std::string velocity_range_noalloc(double out[6], long* nonfinite);
  std::string cell_velocity();      // P2C interpolations of qx, qy, qz into w64.qxix, qyiy, qziz

  // ---- variable properties (heated channel, isothermal y walls) ----
  // set-up: property table (and mode TABLE / MLP), wall temperatures,
  // reference Prandtl number, uniform initial state from t_init; g from q
  std::string thermo_enable(const chs_thermo_desc_t& d);
  // rho*h per cell (FP64, ccc): properties refreshed, g = q rho
  std::string thermo_set_rhoh(const double* rhoh, int mem, long* clamped);
  std::string thermo_get(double* out[9], int mem);
  // energy sub-step (Solve_energy_eq): rhs, rho*h += rhs, property refresh
  std::string energy_substep(int isub);
  // energy right-hand side only, dt-scaled (into eres)
  std::string energy_rhs(int isub);
  // q <-> g with the density interpolated to each velocity location
  std::string convert_qg(bool to_g);
  bool thermo = false;
  int props_mode = CHS_PROPS_TABLE;
  const chs_props* props = nullptr;
  double wall_prop[2][7] = {};  // h, d, m, k, t, b, cp of each y wall
  double prandtl = 1.0, rprren = 0.0;
  Field<double> rhoh, hent, temp, kcond, dens, eres, erhs0, dface[3];
  double* fby_d = nullptr;
  // the energy right-hand side's work: FP64, and FP32 in the mixed form
  // (there the energy right-hand side is FP32 like the momentum one)
  EnergyWork<double> e64;
  EnergyWork<float> e32;
  // mixed form: the FP32 copies of g (w32) were made by the energy sub-step
  // of this sub-step, so the momentum right-hand side need not remake them
  bool g32_fresh = false;
  // mixed form: the FP32 copies of h, k, T (e32) and mu (w32) equal the
  // rounded FP64 fields (the last property refresh stored them)
  bool props32_fresh = false;
  int* props_flag = nullptr;
  std::string copy_rhs(double* mx, double* my, double* mz, int mem);

  Handle* H = nullptr;
  int nx, ny, nz, px, py, pz;
  bool wall_y = true;
  bool mixed = false;   // momentum and energy right-hand sides in FP32 (CHS_PREC_MIXED)
  // mixed form: the FP32 copies of qx, qy, qz, pr (w32) equal the rounded
  // state.  A sub-step leaves them so (its corrections and pressure update
  // store both); set_state clears it.  The batched form captures its steps
  // as if the copies were fresh and, before a replay, rounds the state of
  // any instance whose copies are not (refresh_q32).
  bool q32_fresh = false;
  std::string refresh_q32();  // w32.qx, qy, qz, pr = the rounded state (mixed form)
  double rre = 1.0, dt = 0.0, sigma2p = 1.0;
  double tg[4], tz[4], ta[4];
  int ibc[3][5][2];     // [dir][u,v,w,p,T][side]
  int mcov[3][3][2], mtau[3][3][2];
  double vol = 1.0;     // sum of the cell volumes (x-pencil qx cells)
  std::vector<double> dyw;  // per-j volume weight dy (size ny)

  // FP64 state, pressure increment and divergence
  Field<double> qx, qy, qz, pr, pcor, div;
  RhsWork<double> w64;
  RhsWork<float> w32;
  double *partial = nullptr, *bulk = nullptr;
  long npartial = 0;  // entries of partial (at least kMaxPartials)
  // device address of this device's static range partials (3 kMaxPartials
  // doubles, shared by every handle on the device)
  double* range_scratch = nullptr;
  double* dyw_dev = nullptr;
  double* ymc_dev = nullptr;  // ds/dy at cell centres (stretched y only; made on first use)

  // Independent operators of a sub-step run on three streams (lane 0 = the
  // handle's stream, lanes 1 and 2 = side streams), forked from and joined
  // back into the handle's stream with events, so that their launches
  // overlap and fill each other's partial waves.  Off until the first
  // sub-step has planned and tuned every operator one at a time, while the
  // handle's stream is being captured into a graph (the batched form's
  // captured steps already run their instances on separate streams), and
  // while launch profiling is active (each launch is then timed alone).
  bool concurrent = false;

 private:
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  bool lanes_on = false;  // concurrent, and the handle's stream is not capturing
  void set_lanes();
  hipStream_t lane_stream(int lane) const { return (lane == 0 || !lanes_on) ? H->stream : side[lane - 1]; }
  std::string fork();  // the side streams wait for the work issued so far on the handle's stream
  std::string join();  // the handle's stream waits for the work issued so far on the side streams

  template <typename T>
  std::string op(RhsWork<T>& w, int dir, int opk, const int ibc2[2], const Field<T>& A, Field<T>& out,
                 const T* fbc, T sA = T(1), const Field<T>* B = nullptr, T cB = T(1), const Field<T>* Cm = nullptr,
                 T sOut = T(1), int accumulate = 0, const Field<T>* D = nullptr, T cD = T(0),
                 const Field<T>* E = nullptr, T cE = T(0), int lane = 0, float* out32 = nullptr);
  // the interpolation and the first derivative of A between the same
  // locations (plans opa with bca / fbca and opb with bcb / fbcb) in one
  // pass where the line kernels allow it, else one after the other
  template <typename T>
  std::string op2(RhsWork<T>& w, int lane, int dir, const Field<T>& A, int opa, const int bca[2], Field<T>& outa,
                  const T* fbca, int opb, const int bcb[2], Field<T>& outb, const T* fbcb);
  // out = d(-ca * cb) + r d(ts * (ta + tc * tb)): a convective and a viscous
  // operator of one momentum component along one direction
  template <typename T>
  std::string cov_tau(RhsWork<T>& w, int dir, int opk, const int bc_cov[2], const int bc_tau[2], Field<T>& out,
                      const T* fbc_cov, const T* fbc_tau, const Field<T>& ca, const Field<T>& cb,
                      const Field<T>& ta, T ts, const Field<T>* tb, T tc, int lane = 0);
  template <typename T> std::string momentum_rhs_t(RhsWork<T>& w, int isub, bool fuse);
  template <typename T> std::string thermo_first_level(RhsWork<T>& w);
  template <typename T> std::string alloc_thermo_work(RhsWork<T>& w);
  std::string refresh_props();
  // the x, y, z parts of the energy right-hand side into e.ex, e.ey, e.ez
  // (the FP32 form first rounds g, h, k and T into w and e)
  std::string energy_terms();
  template <typename T> std::string energy_terms_t(RhsWork<T>& w, EnergyWork<T>& e);
  // rhs = dt (tg cur + tz rhs0), rhs0 = cur, with cur = ex + ey + ez; into
  // res and / or added to rho h
  std::string energy_combine(int isub, double* res, double* rhoh_acc);
  template <typename T> std::string alloc_energy_work(EnergyWork<T>& e, bool own_centres);
  template <typename T> std::string alloc_work(RhsWork<T>& w, bool alias_state);
  template <typename T> std::string alloc(Field<T>& f, int e1, int e2, int e3);
  std::vector<void*> owned;
};

}  // namespace chs
