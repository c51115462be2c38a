// extern "C" entry points of libchapsim_hip (declared in include/chapsim_hip.h).
#include <cmath>
#include <cstring>
#include <thread>
#include <string>

#include "chs_flow.h"
#include "chs_handle.h"
#include "chs_exact.h"
#include "chs_poisson.h"
#include "chs_tune.h"

namespace chs {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

std::string DevBuf::ensure(size_t b) {
  if (b <= bytes) return "";
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) return std::string("device allocation: ") + hipGetErrorString(e);
  bytes = b;
  return "";
}

bool Handle::prof_pair(hipEvent_t* a, hipEvent_t* b) {
  if (prof_used + 2 > prof_ev.size()) {
    if (prof_ev.size() >= 16384) return false;
    for (int i = 0; i < 256; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return false;
      prof_ev.push_back(e);
    }
  }
  *a = prof_ev[prof_used];
  *b = prof_ev[prof_used + 1];
  prof_used += 2;
  return true;
}

Handle::~Handle() {
  for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
  flow.reset();
  poisson.reset();
  plans64.clear();
  plans32.clear();
  if (stream) (void)hipStreamDestroy(stream);
}

template <typename T>
static std::map<PlanKey, std::unique_ptr<DevPlan<T>>>& plan_map(Handle* h);
template <>
std::map<PlanKey, std::unique_ptr<DevPlan<double>>>& plan_map<double>(Handle* h) { return h->plans64; }
template <>
std::map<PlanKey, std::unique_ptr<DevPlan<float>>>& plan_map<float>(Handle* h) { return h->plans32; }

// Rows per chunk of a plan: FP32 lines of 129-160 rows take 32-row chunks
// (5 instead of 9 chunks of 16: a y derivative at 256x129x256 in FP32
// 26.6 -> 18.7 us); FP64 ones keep 16-row chunks (46.0 against 40.2 us with
// 32, the registers of the longer window), as does everything else.
static int plan_chunk_rows(size_t tsize, int n_out) {
  if (tsize == 4 && n_out > kChunk * 8 && n_out <= 160) return kChunkLong;
  return 0;
}

template <typename T>
std::string Handle::plan(int dir, int op, int iacc, const int ibc[2], bool has_fbc, int n_in, int n_out,
                         const DevPlan<T>** out) {
  PlanKey key{dir, op, iacc, ibc[0], ibc[1], has_fbc ? 1 : 0, n_in, n_out};
  auto& mp = plan_map<T>(this);
  auto it = mp.find(key);
  if (it != mp.end()) {
    *out = it->second.get();
    return "";
  }
  LinePlan lp;
  std::string e = make_line_plan(mesh.line_geom(dir), Op(op), iacc, ibc, has_fbc, n_in, n_out, &lp,
                                 plan_chunk_rows(sizeof(T), n_out));
  if (!e.empty()) return e;
  std::unique_ptr<DevPlan<T>> dp(new DevPlan<T>());
  e = upload_plan<T>(lp, dp.get());
  if (!e.empty()) return e;
  *out = dp.get();
  mp[key] = std::move(dp);
  return "";
}
template std::string Handle::plan<double>(int, int, int, const int[2], bool, int, int, const DevPlan<double>**);
template std::string Handle::plan<float>(int, int, int, const int[2], bool, int, int, const DevPlan<float>**);

}  // namespace chs

using chs::set_error;

static int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

#define API_CHECK_HANDLE(h) \
  if (!(h)) return fail(CHS_EINVAL, "null handle")

// One Poisson solve on the handle's stream; under launch profiling class 3
// it is bracketed by events and counted as five field passes (forward x and
// z transforms, the y solves, inverse z and x), each reading and writing the
// n cells once.
template <typename T>
static std::string poisson_timed(chs::Handle& h, T* rhs, size_t n) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = (h.prof_class == 3) && h.prof_pair(&e0, &e1);
  if (timed) {
    (void)hipEventRecord(e0, h.stream);
    h.prof_bytes += 10.0 * sizeof(T) * (double)n;
  }
  std::string e = h.poisson->solve(rhs, h.stream);
  if (timed) (void)hipEventRecord(e1, h.stream);
  return e;
}

extern "C" {

int chs_abi_version(void) { return CHS_ABI_VERSION; }

const char* chs_last_error(void) { return chs::g_err.c_str(); }

int chs_create(const chs_desc_t* desc, chs_handle_t** out) {
  if (!desc || !out) return fail(CHS_EINVAL, "null argument");
  *out = nullptr;
  for (int d = 0; d < 3; ++d)
    if (desc->nc[d] < 8) return fail(CHS_EINVAL, "each direction needs at least 8 cells");
  if (desc->iacc < CHS_CD2 || desc->iacc > CHS_CP6) return fail(CHS_EINVAL, "invalid accuracy");
  if (desc->istret != 0 && desc->istret != 2 && desc->istret != 3 && desc->istret != 4)
    return fail(CHS_EINVAL, "invalid istret");
  if (desc->precision != CHS_PREC_FP64 && desc->precision != CHS_PREC_MIXED)
    return fail(CHS_EINVAL, "invalid precision");
  if (desc->projection != CHS_PROJ_REFERENCE && desc->projection != CHS_PROJ_EXACT)
    return fail(CHS_EINVAL, "invalid projection");
  if (desc->projection == CHS_PROJ_EXACT && desc->periodic_y)
    return fail(CHS_EUNSUP, "the exact projection is built for walls in y");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CHS_EHIP, "no HIP device");
  chs_handle_t* H = new chs_handle_t();
  chs::Handle& h = H->h;
  h.desc = *desc;
  (void)hipGetDevice(&h.device);
  if (hipStreamCreate(&h.stream) != hipSuccess) {
    delete H;
    return fail(CHS_EHIP, "stream creation failed");
  }
  h.mesh.build(desc->nc, desc->periodic_y != 0, desc->lxx, desc->lyb, desc->lyt, desc->lzz, desc->istret,
               desc->rstret);
  if (!h.mesh.periodic[1] && desc->poisson != CHS_POISSON_FISHPACK && desc->poisson != CHS_POISSON_3DFFT) {
    delete H;
    return fail(CHS_EUNSUP, "unknown Poisson path");
  }
  if (!h.mesh.periodic[1] && desc->poisson == CHS_POISSON_3DFFT && desc->projection == CHS_PROJ_EXACT) {
    delete H;
    return fail(CHS_EUNSUP, "the exact projection replaces the Poisson solve: use it with the FISHPACK set-up");
  }
  if (h.mesh.periodic[1] && desc->poisson != CHS_POISSON_3DFFT) {
    delete H;
    return fail(CHS_EUNSUP, "periodic y needs the 3-D FFT Poisson path");
  }
  std::string e;
  if (desc->poisson == CHS_POISSON_FISHPACK || desc->poisson == CHS_POISSON_3DFFT) {
    e = chs::Poisson::create(h.mesh, desc->poisson, desc->iacc, &h.poisson);
    if (!e.empty()) {
      delete H;
      return fail(CHS_EUNSUP, e);
    }
    e = chs::Flow::create(&h, &h.flow);
    if (!e.empty()) {
      delete H;
      return fail(CHS_ENOMEM, e);
    }
    if (desc->projection == CHS_PROJ_EXACT) {
      chs::ExactOps o;
      e = chs::build_exact_ops(h.mesh, desc->iacc, h.flow->ibc, h.flow->wall_y, &o);
      if (e.empty()) e = h.poisson->setup_exact(o);
      if (!e.empty()) {
        delete H;
        return fail(CHS_EUNSUP, e);
      }
    }
  }
  *out = H;
  return CHS_OK;
}

int chs_destroy(chs_handle_t* h) {
  if (!h) return CHS_OK;
  (void)hipStreamSynchronize(h->h.stream);
  delete h;
  return CHS_OK;
}

int chs_sync(chs_handle_t* h) {
  API_CHECK_HANDLE(h);
  hipError_t e = hipStreamSynchronize(h->h.stream);
  if (e != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(e));
  return CHS_OK;
}

int chs_op3d(chs_handle_t* H, int dir, int op, int iacc, const int ibc[2], const double* fi, const int shp_in[3],
             double* fo, const int shp_out[3], const double* fbc, int mem) {
  API_CHECK_HANDLE(H);
  chs::Handle& h = H->h;
  if (dir < 0 || dir > 2 || op < 0 || op > 5 || !ibc || !fi || !fo || !shp_in || !shp_out)
    return fail(CHS_EINVAL, "invalid argument");
  for (int d = 0; d < 3; ++d)
    if (shp_in[d] <= 0 || shp_out[d] <= 0) return fail(CHS_EINVAL, "non-positive extent");
  const chs::DevPlan<double>* P = nullptr;
  std::string e = h.plan<double>(dir, op, iacc, ibc, fbc != nullptr, shp_in[dir], shp_out[dir], &P);
  if (!e.empty()) return fail(CHS_EINVAL, e);
  chs::Ext3 ein{{shp_in[0], shp_in[1], shp_in[2]}}, eout{{shp_out[0], shp_out[1], shp_out[2]}};
  const size_t nin = (size_t)shp_in[0] * shp_in[1] * shp_in[2];
  const size_t nout = (size_t)shp_out[0] * shp_out[1] * shp_out[2];
  const size_t nfbc = (dir == 0) ? 4ul * shp_in[1] * shp_in[2]
                      : (dir == 1) ? (size_t)shp_in[0] * 4 * shp_in[2]
                                   : (size_t)shp_in[0] * shp_in[1] * 4;
  const double *din = fi, *dfbc = fbc;
  double* dout = fo;
  if ((e = h.io[2].ensure(sizeof(double) * nout)).size()) return fail(CHS_ENOMEM, e);
  double* scratch = h.io[2].as<double>();
  if (mem == CHS_MEM_HOST) {
    if ((e = h.io[0].ensure(sizeof(double) * nin)).size()) return fail(CHS_ENOMEM, e);
    if ((e = h.io[1].ensure(sizeof(double) * nout)).size()) return fail(CHS_ENOMEM, e);
    if ((e = h.io[3].ensure(sizeof(double) * nfbc)).size()) return fail(CHS_ENOMEM, e);
    if (hipMemcpyAsync(h.io[0].p, fi, sizeof(double) * nin, hipMemcpyHostToDevice, h.stream) != hipSuccess)
      return fail(CHS_EHIP, "copy in");
    if (fbc && hipMemcpyAsync(h.io[3].p, fbc, sizeof(double) * nfbc, hipMemcpyHostToDevice, h.stream) != hipSuccess)
      return fail(CHS_EHIP, "copy fbc");
    din = h.io[0].as<double>();
    dout = h.io[1].as<double>();
    dfbc = fbc ? h.io[3].as<double>() : nullptr;
  } else if (mem != CHS_MEM_DEVICE) {
    return fail(CHS_EINVAL, "invalid mem kind");
  }
  chs::LineIn<double> in;
  in.A = din;
  chs::LineOut<double> o;
  o.out = dout;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = (h.prof_class == dir) && h.prof_pair(&e0, &e1);
  if (timed) {
    (void)hipEventRecord(e0, h.stream);
    h.prof_bytes += 8.0 * ((double)nin + (double)nout);
  }
  e = chs::launch_line_op<double>(*P, ein, eout, in, dfbc, o, scratch, h.stream);
  if (timed) (void)hipEventRecord(e1, h.stream);
  if (!e.empty()) return fail(CHS_EINVAL, e);
  if (mem == CHS_MEM_HOST &&
      hipMemcpyAsync(fo, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, h.stream) != hipSuccess)
    return fail(CHS_EHIP, "copy out");
  hipError_t he = hipStreamSynchronize(h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

int chs_poisson(chs_handle_t* H, double* rhs, int mem) {
  API_CHECK_HANDLE(H);
  chs::Handle& h = H->h;
  if (!h.poisson) return fail(CHS_EUNSUP, "no Poisson solver for this configuration");
  const size_t n = (size_t)h.mesh.nc[0] * h.mesh.nc[1] * h.mesh.nc[2];
  double* d = rhs;
  std::string e;
  if (mem == CHS_MEM_HOST) {
    if ((e = h.io[0].ensure(sizeof(double) * n)).size()) return fail(CHS_ENOMEM, e);
    if (hipMemcpyAsync(h.io[0].p, rhs, sizeof(double) * n, hipMemcpyHostToDevice, h.stream) != hipSuccess)
      return fail(CHS_EHIP, "copy in");
    d = h.io[0].as<double>();
  }
  e = poisson_timed(h, d, n);
  if (!e.empty()) return fail(CHS_EHIP, e);
  if (mem == CHS_MEM_HOST &&
      hipMemcpyAsync(rhs, d, sizeof(double) * n, hipMemcpyDeviceToHost, h.stream) != hipSuccess)
    return fail(CHS_EHIP, "copy out");
  hipError_t he = hipStreamSynchronize(h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

int chs_op3d_f32(chs_handle_t* H, int dir, int op, int iacc, const int ibc[2], const float* fi, const int shp_in[3],
                 float* fo, const int shp_out[3], const float* fbc) {
  API_CHECK_HANDLE(H);
  chs::Handle& h = H->h;
  if (dir < 0 || dir > 2 || op < 0 || op > 5 || !ibc || !fi || !fo || !shp_in || !shp_out)
    return fail(CHS_EINVAL, "invalid argument");
  for (int d = 0; d < 3; ++d)
    if (shp_in[d] <= 0 || shp_out[d] <= 0) return fail(CHS_EINVAL, "non-positive extent");
  const chs::DevPlan<float>* P = nullptr;
  std::string e = h.plan<float>(dir, op, iacc, ibc, fbc != nullptr, shp_in[dir], shp_out[dir], &P);
  if (!e.empty()) return fail(CHS_EINVAL, e);
  chs::Ext3 ein{{shp_in[0], shp_in[1], shp_in[2]}}, eout{{shp_out[0], shp_out[1], shp_out[2]}};
  const size_t nin = (size_t)shp_in[0] * shp_in[1] * shp_in[2];
  const size_t nout = (size_t)shp_out[0] * shp_out[1] * shp_out[2];
  if ((e = h.io[2].ensure(sizeof(float) * nout)).size()) return fail(CHS_ENOMEM, e);
  chs::LineIn<float> in;
  in.A = fi;
  chs::LineOut<float> o;
  o.out = fo;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = (h.prof_class == dir) && h.prof_pair(&e0, &e1);
  if (timed) {
    (void)hipEventRecord(e0, h.stream);
    h.prof_bytes += 4.0 * ((double)nin + (double)nout);
  }
  e = chs::launch_line_op<float>(*P, ein, eout, in, fbc, o, h.io[2].as<float>(), h.stream);
  if (timed) (void)hipEventRecord(e1, h.stream);
  if (!e.empty()) return fail(CHS_EINVAL, e);
  hipError_t he = hipStreamSynchronize(h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

int chs_poisson_f32(chs_handle_t* H, float* rhs) {
  API_CHECK_HANDLE(H);
  chs::Handle& h = H->h;
  if (!h.poisson) return fail(CHS_EUNSUP, "no Poisson solver for this configuration");
  if (!rhs) return fail(CHS_EINVAL, "invalid argument");
  const size_t n = (size_t)h.mesh.nc[0] * h.mesh.nc[1] * h.mesh.nc[2];
  std::string e = poisson_timed(h, rhs, n);
  if (!e.empty()) return fail(CHS_EHIP, e);
  hipError_t he = hipStreamSynchronize(h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

#define FLOW_OR_FAIL(H)                                                         \
  API_CHECK_HANDLE(H);                                                          \
  if (!(H)->h.flow) return fail(CHS_EUNSUP, "no flow solver for this configuration"); \
  chs::Flow& F = *(H)->h.flow

#define RET(expr)                                          \
  do {                                                     \
    std::string e_ = (expr);                               \
    if (!e_.empty()) return fail(CHS_EHIP, e_);            \
  } while (0)

int chs_set_state(chs_handle_t* H, const double* qx, const double* qy, const double* qz, const double* pr,
                  int mem) {
  FLOW_OR_FAIL(H);
  if (!qx || !qy || !qz || !pr) return fail(CHS_EINVAL, "null field");
  RET(F.set_state(qx, qy, qz, pr, mem));
  return CHS_OK;
}

int chs_get_state(chs_handle_t* H, double* qx, double* qy, double* qz, double* pr, int mem) {
  FLOW_OR_FAIL(H);
  if (!qx || !qy || !qz || !pr) return fail(CHS_EINVAL, "null field");
  RET(F.get_state(qx, qy, qz, pr, mem));
  return CHS_OK;
}

int chs_update_re(chs_handle_t* H, int iter) {
  FLOW_OR_FAIL(H);
  F.update_re(iter);
  return CHS_OK;
}

int chs_substep(chs_handle_t* H, int isub) {
  FLOW_OR_FAIL(H);
  if (isub < 0 || isub > 3) return fail(CHS_EINVAL, "isub must be 0..3");
  RET(F.substep(isub));
  hipError_t he = hipStreamSynchronize(H->h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

int chs_run(chs_handle_t* H, int iter0, int nsteps) {
  FLOW_OR_FAIL(H);
  if (F.thermo && hipMemsetAsync(F.props_flag, 0, sizeof(int), H->h.stream) != hipSuccess)
    return fail(CHS_EHIP, "flag reset");
  for (int it = iter0 + 1; it <= iter0 + nsteps; ++it) {
    F.update_re(it);
    for (int isub = 1; isub <= 3; ++isub) RET(F.substep(isub));
  }
  int nbad = 0;
  if (F.thermo &&
      hipMemcpyAsync(&nbad, F.props_flag, sizeof(int), hipMemcpyDeviceToHost, H->h.stream) != hipSuccess)
    return fail(CHS_EHIP, "flag read");
  hipError_t he = hipStreamSynchronize(H->h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  // the reference stops when rho*h leaves the table (src/input_thermo.f90:402-413)
  if (nbad) return fail(CHS_ERANGE, "rho*h left the property table (" + std::to_string(nbad) + " cell refreshes clamped)");
  return CHS_OK;
}

int chs_thermo_enable(chs_handle_t* H, const chs_thermo_desc_t* d) {
  FLOW_OR_FAIL(H);
  if (!d) return fail(CHS_EINVAL, "invalid argument");
  const std::string e = F.thermo_enable(*d);
  if (!e.empty()) return fail(e.find("not supported") != std::string::npos || e.find("need y walls") != std::string::npos
                                  ? CHS_EUNSUP : CHS_EINVAL, e);
  return CHS_OK;
}

int chs_thermo_set_rhoh(chs_handle_t* H, const double* rhoh, int mem) {
  FLOW_OR_FAIL(H);
  if (!rhoh || (mem != CHS_MEM_HOST && mem != CHS_MEM_DEVICE)) return fail(CHS_EINVAL, "invalid argument");
  long nb = 0;
  RET(F.thermo_set_rhoh(rhoh, mem, &nb));
  if (nb) return fail(CHS_ERANGE, "rho*h is out of the table's range (" + std::to_string(nb) + " cells clamped)");
  return CHS_OK;
}

int chs_thermo_set_mode(chs_handle_t* H, int mode) {
  FLOW_OR_FAIL(H);
  if (!F.thermo) return fail(CHS_EINVAL, "variable properties not enabled");
  if (mode != CHS_PROPS_TABLE && mode != CHS_PROPS_MLP) return fail(CHS_EINVAL, "unknown property mode");
  if (mode == CHS_PROPS_MLP && !F.props->surrogate) return fail(CHS_EINVAL, "no surrogate attached to the table");
  F.props_mode = mode;
  return CHS_OK;
}

int chs_thermo_get(chs_handle_t* H, double* rhoh, double* hh, double* t, double* k, double* d, double* m, double* gx,
                   double* gy, double* gz, int mem) {
  FLOW_OR_FAIL(H);
  if (mem != CHS_MEM_HOST && mem != CHS_MEM_DEVICE) return fail(CHS_EINVAL, "invalid mem kind");
  double* out[9] = {rhoh, hh, t, k, d, m, gx, gy, gz};
  RET(F.thermo_get(out, mem));
  return CHS_OK;
}

int chs_energy_rhs(chs_handle_t* H, int isub, double* rhs, int mem) {
  FLOW_OR_FAIL(H);
  if (!F.thermo) return fail(CHS_EINVAL, "variable properties not enabled");
  if (!rhs || isub < 1 || isub > 3 || (mem != CHS_MEM_HOST && mem != CHS_MEM_DEVICE))
    return fail(CHS_EINVAL, "invalid argument");
  RET(F.energy_rhs(isub));
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (hipMemcpyAsync(rhs, F.eres.p, sizeof(double) * F.eres.size(), k, H->h.stream) != hipSuccess)
    return fail(CHS_EHIP, "copy out");
  hipError_t he = hipStreamSynchronize(H->h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched channels: independent instances advanced together, each on its own
// stream, optionally as one captured graph per RK3 step (fork from a main
// stream to every instance stream, join back)
// host threads issuing the instances of an eager batch step
constexpr int kBatchThreads = 8;

struct chs_batch {
  std::vector<chs_handle_t*> inst;
  hipStream_t main = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<double> rre_captured;
  bool warmed = false;
  ~chs_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (main) (void)hipStreamDestroy(main);
    for (chs_handle_t* h : inst) chs_destroy(h);
  }
};

static std::string batch_capture(chs_batch* b) {
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  if (hipStreamBeginCapture(b->main, hipStreamCaptureModeRelaxed) != hipSuccess) return "begin capture";
  std::string err;
  if (hipEventRecord(b->fork, b->main) != hipSuccess) err = "fork record";
  for (size_t i = 0; i < b->inst.size() && err.empty(); ++i) {
    chs::Handle& h = b->inst[i]->h;
    if (hipStreamWaitEvent(h.stream, b->fork, 0) != hipSuccess) { err = "fork wait"; break; }
    // the captured step starts from fresh FP32 copies of the state (a replay
    // follows a step, which leaves them so, or chs_batch_run's rounding of
    // an instance whose state was set); the flag is restored after capture
    const bool fresh = h.flow->q32_fresh;
    h.flow->q32_fresh = true;
    for (int isub = 1; isub <= 3 && err.empty(); ++isub) err = h.flow->substep(isub);
    h.flow->q32_fresh = fresh;
    if (err.empty() && hipEventRecord(b->join[i], h.stream) != hipSuccess) err = "join record";
    if (err.empty() && hipStreamWaitEvent(b->main, b->join[i], 0) != hipSuccess) err = "join wait";
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->main, &g);
  if (!err.empty()) {
    if (g) (void)hipGraphDestroy(g);
    return "capture: " + err;
  }
  if (ec != hipSuccess || !g) return std::string("end capture: ") + hipGetErrorString(ec);
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) return "graph instantiate";
  b->rre_captured.resize(b->inst.size());
  for (size_t i = 0; i < b->inst.size(); ++i) b->rre_captured[i] = b->inst[i]->h.flow->rre;
  return "";
}

extern "C" {

int chs_batch_create(const chs_desc_t* descs, int n, chs_batch_t** out) {
  if (!descs || !out || n < 1) return fail(CHS_EINVAL, "invalid argument");
  *out = nullptr;
  chs_batch* b = new chs_batch();
  for (int i = 0; i < n; ++i) {
    chs_handle_t* h = nullptr;
    const int rc = chs_create(&descs[i], &h);
    if (rc != CHS_OK) {
      const std::string msg = chs::g_err;
      delete b;
      return fail(rc, "instance " + std::to_string(i) + ": " + msg);
    }
    b->inst.push_back(h);
    if (!h->h.flow) {
      delete b;
      return fail(CHS_EUNSUP, "instance " + std::to_string(i) + ": no flow solver for this configuration");
    }
  }
  b->join.resize(n, nullptr);
  bool ok = hipStreamCreate(&b->main) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < n && ok; ++i) ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    delete b;
    return fail(CHS_EHIP, "stream / event creation failed");
  }
  *out = b;
  return CHS_OK;
}

int chs_batch_destroy(chs_batch_t* b) {
  delete b;
  return CHS_OK;
}

chs_handle_t* chs_batch_instance(chs_batch_t* b, int i) {
  if (!b || i < 0 || i >= (int)b->inst.size()) return nullptr;
  return b->inst[i];
}

int chs_batch_run(chs_batch_t* b, int iter0, int nsteps, int use_graph) {
  if (!b || nsteps < 0) return fail(CHS_EINVAL, "invalid argument");
  for (chs_handle_t* h : b->inst)
    if (h->h.prof_class >= 0 && use_graph) return fail(CHS_EINVAL, "launch profiling is active on an instance");
  for (int it = iter0 + 1; it <= iter0 + nsteps; ++it) {
    bool recapture = !b->exec;
    for (size_t i = 0; i < b->inst.size(); ++i) {
      chs::Flow& F = *b->inst[i]->h.flow;
      F.update_re(it);
      if (b->exec && F.rre != b->rre_captured[i]) recapture = true;
    }
    if (!b->warmed) {
      // first step: every instance on its own stream, one after the other
      // (plans are created and launch shapes timed here)
      for (chs_handle_t* h : b->inst)
        for (int isub = 1; isub <= 3; ++isub) RET(h->h.flow->substep(isub));
      b->warmed = true;
      continue;
    }
    if (!use_graph) {
      // eager step: the instances' launches issued from several host threads
      // (each instance from one thread, on its own streams)
      const int n = (int)b->inst.size();
      const int nt = n < kBatchThreads ? n : kBatchThreads;
      std::vector<std::string> errs(nt);
      std::vector<std::thread> th;
      for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
          for (int i = t; i < n && errs[t].empty(); i += nt)
            for (int isub = 1; isub <= 3 && errs[t].empty(); ++isub) errs[t] = b->inst[i]->h.flow->substep(isub);
        });
      for (std::thread& x : th) x.join();
      for (const std::string& e : errs)
        if (!e.empty()) return fail(CHS_EHIP, e);
      continue;
    }
    if (recapture) {
      for (chs_handle_t* h : b->inst)
        if (hipStreamSynchronize(h->h.stream) != hipSuccess) return fail(CHS_EHIP, "synchronise");
      const std::string e = batch_capture(b);
      if (!e.empty()) return fail(CHS_EHIP, e);
    }
    // mixed instances whose FP32 copies of the state are stale (a set_state
    // since the last step) are rounded before the replay
    for (chs_handle_t* h : b->inst) {
      chs::Flow& F = *h->h.flow;
      if (F.mixed && !F.q32_fresh) {
        RET(F.refresh_q32());
        if (hipStreamSynchronize(h->h.stream) != hipSuccess) return fail(CHS_EHIP, "synchronise");
      }
    }
    if (hipGraphLaunch(b->exec, b->main) != hipSuccess) return fail(CHS_EHIP, "graph launch");
    for (chs_handle_t* h : b->inst) h->h.flow->q32_fresh = h->h.flow->mixed;
  }
  if (hipStreamSynchronize(b->main) != hipSuccess) return fail(CHS_EHIP, "synchronise");
  for (chs_handle_t* h : b->inst)
    if (hipStreamSynchronize(h->h.stream) != hipSuccess) return fail(CHS_EHIP, "synchronise");
  return CHS_OK;
}

int chs_momentum_rhs(chs_handle_t* H, int isub, double* mx, double* my, double* mz, int mem) {
  FLOW_OR_FAIL(H);
  if (isub < 0 || isub > 3) return fail(CHS_EINVAL, "isub must be 0..3");
  RET(F.momentum_rhs(isub));
  RET(F.copy_rhs(mx, my, mz, mem));
  return CHS_OK;
}

int chs_cfl(chs_handle_t* H, double out[2]) {
  FLOW_OR_FAIL(H);
  if (!out) return fail(CHS_EINVAL, "invalid argument");
  RET(F.cfl(out));
  return CHS_OK;
}

int chs_bulk(chs_handle_t* H, double out[2]) {
  FLOW_OR_FAIL(H);
  if (!out) return fail(CHS_EINVAL, "invalid argument");
  RET(F.bulk_monitor(out));
  return CHS_OK;
}

int chs_velocity_range(chs_handle_t* H, double out[6]) {
  FLOW_OR_FAIL(H);
  if (!out) return fail(CHS_EINVAL, "invalid argument");
  long bad = 0;
  RET(F.velocity_range(out, &bad));
  if (bad > 0) return fail(CHS_ENAN, "velocity has " + std::to_string(bad) + " non-finite values");
  return CHS_OK;
}

int chs_velocity_range_noalloc(chs_handle_t* H, double out[6]) {
  FLOW_OR_FAIL(H);
  if (!out) return fail(CHS_EINVAL, "invalid argument");
  long bad = 0;
  RET(F.velocity_range_noalloc(out, &bad));
  if (bad > 0) return fail(CHS_ENAN, "velocity has non-finite values");
  return CHS_OK;
}

int chs_mass_check(chs_handle_t* H, double out[3]) {
  FLOW_OR_FAIL(H);
  RET(F.mass_check(out));
  return CHS_OK;
}

int chs_divergence(chs_handle_t* H, double* div, int mem) {
  FLOW_OR_FAIL(H);
  RET(F.divergence(F.div.p));
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (hipMemcpyAsync(div, F.div.p, sizeof(double) * F.div.size(), k, H->h.stream) != hipSuccess)
    return fail(CHS_EHIP, "copy out");
  hipError_t he = hipStreamSynchronize(H->h.stream);
  if (he != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(he));
  return CHS_OK;
}

}  // extern "C"

#include "../../include/chapsim_hip_inspect.h"

extern "C" {

double chs_inspect_coeff(int fam, int is_rhs, int row, int k, int bc, int acc) {
  if (fam < 0 || fam > 5 || row < 1 || row > 5 || k < 1 || k > (is_rhs ? 6 : 3) || bc < 0 || bc > 6 || acc < 1 ||
      acc > 4)
    return NAN;
  return chs::coeff(chs::Fam(fam), is_rhs != 0, row, k, bc, acc);
}

int chs_inspect_op1d(const chs_desc_t* desc, int dir, int op, int iacc, const int ibc[2], const double* fi, int n_in,
                     double* fo, int n_out, const double* fbc) {
  if (!desc || !ibc || !fi || !fo || dir < 0 || dir > 2 || op < 0 || op > 5) return fail(CHS_EINVAL, "invalid argument");
  chs::Mesh m;
  m.build(desc->nc, desc->periodic_y != 0, desc->lxx, desc->lyb, desc->lyt, desc->lzz, desc->istret, desc->rstret);
  chs::LinePlan p;
  std::string e = chs::make_line_plan(m.line_geom(dir), chs::Op(op), iacc, ibc, fbc != nullptr, n_in, n_out, &p);
  if (!e.empty()) return fail(CHS_EINVAL, e);
  chs::apply_plan_host(p, fi, fbc, fo);
  return CHS_OK;
}

int chs_inspect_projection(const chs_desc_t* desc, double* ly, double* lam, double* lx, double* lz, double* resid) {
  if (!desc) return fail(CHS_EINVAL, "invalid argument");
  chs::Mesh m;
  m.build(desc->nc, desc->periodic_y != 0, desc->lxx, desc->lyb, desc->lyt, desc->lzz, desc->istret, desc->rstret);
  int ibc[3][5][2];
  chs::flow_bc_kinds(!m.periodic[1], ibc);
  chs::ExactOps o;
  std::string e = chs::build_exact_ops(m, desc->iacc, ibc, !m.periodic[1], &o);
  if (!e.empty()) return fail(CHS_EINVAL, e);
  if (ly) std::copy(o.Ly.begin(), o.Ly.end(), ly);
  if (lam) std::copy(o.lam.begin(), o.lam.end(), lam);
  if (lx) std::copy(o.lx.begin(), o.lx.end(), lx);
  if (lz) std::copy(o.lz.begin(), o.lz.end(), lz);
  if (resid) *resid = o.eig_resid;
  return CHS_OK;
}

}  // extern "C"

extern "C" {

int chs_profile_begin(chs_handle_t* H, int kernel_class) {
  API_CHECK_HANDLE(H);
  if (kernel_class < 0 || kernel_class > 3) return fail(CHS_EINVAL, "kernel class must be 0..3");
  chs::Handle& h = H->h;
  (void)hipStreamSynchronize(h.stream);
  h.prof_class = kernel_class;
  h.prof_used = 0;
  h.prof_bytes = 0.0;
  return CHS_OK;
}

int chs_tune_report(char* buf, int len) {
  const std::string r = chs::tune_report();
  if (buf && len > 0) {
    const size_t n = r.size() < (size_t)(len - 1) ? r.size() : (size_t)(len - 1);
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  int lines = 0;
  for (char c : r) lines += (c == '\n');
  return lines;
}

int chs_tune_set(const char* key, int choice) {
  if (!key || choice < 0) return fail(CHS_EINVAL, "invalid argument");
  chs::tune_set(std::string(key), choice);
  return CHS_OK;
}

int chs_tune_force(const char* prefix, int variant) {
  chs::tune_force(prefix ? std::string(prefix) : std::string(), variant);
  return CHS_OK;
}

int chs_profile_end(chs_handle_t* H, double* ms_total, long long* launches, double* alg_bytes) {
  API_CHECK_HANDLE(H);
  chs::Handle& h = H->h;
  hipError_t e = hipStreamSynchronize(h.stream);
  if (e != hipSuccess) return fail(CHS_EHIP, hipGetErrorString(e));
  double tot = 0.0;
  for (size_t i = 0; i + 1 < h.prof_used; i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, h.prof_ev[i], h.prof_ev[i + 1]) != hipSuccess) return fail(CHS_EHIP, "event time");
    tot += ms;
  }
  if (ms_total) *ms_total = tot;
  if (launches) *launches = (long long)(h.prof_used / 2);
  if (alg_bytes) *alg_bytes = h.prof_bytes;
  h.prof_class = -1;
  h.prof_used = 0;
  return CHS_OK;
}

}  // extern "C"
