// Thermal properties from the tabulated equation of state, per cell.
//
// Host: the table build of buildup_property_relations_from_table
// (src/input_thermo.f90:617-693): selection sort by T (:512-556), reference
// state by bisection in T (:121-154), non-dimensionalisation, rho*h and its
// range.  Device: one thread per cell runs ftp_refresh_thermal_properties_from_DH
// (:393-446) -- clamp, bisection on rho*h, h, properties from h (:345-374),
// h = rho*h / d, properties from h -- with the bisected columns (rho*h, h, d)
// staged in LDS when they fit.  Floating-point contraction is off in the
// kernel so every product and sum rounds as in the reference's expressions.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "chs_props.h"

namespace chs {
void set_error(const std::string& msg);
}


namespace {

enum { CT = 0, CD, CM, CK, CB, CCP, CH, CRH };
constexpr double kMinP = 1.0e-20;  // src/modules.f90:141
constexpr double kNodeGap = 1.0e-10;  // sqrt(kMinP)
constexpr int kPropsLds = 2048;    // rows of the one column k_props_from_h stages (x3, 48 KB)
constexpr int kPropsBlock = 1024;  // threads of the persistent from-rho*h kernel

int fail(int code, const std::string& msg) {
  chs::set_error(msg);
  return code;
}

void bisect_host(const std::vector<double>& c, double v, int* i1o, int* i2o) {
  int i1 = 0, i2 = (int)c.size() - 1;
  while (i2 - i1 > 1) {
    const int im = i1 + (i2 - i1) / 2;
    const double d1 = c[i1] - v, dm = c[im] - v;
    if (d1 * dm > kMinP) i1 = im;
    else i2 = im;
  }
  *i1o = i1;
  *i2o = i2;
}

#pragma clang fp contract(off)

__device__ __forceinline__ void bisect_dev(const double* c, int n, double v, int* i1o, int* i2o) {
  int i1 = 0, i2 = n - 1;
  while (i2 - i1 > 1) {
    const int im = i1 + (i2 - i1) / 2;
    const double d1 = c[i1] - v, dm = c[im] - v;
    if (d1 * dm > kMinP) i1 = im;
    else i2 = im;
  }
  *i1o = i1;
  *i2o = i2;
}

// One cell of ftp_refresh_thermal_properties_from_DH: clamp rho*h to the
// table, h from the rho*h column, d from h, h = rho*h / d, then every
// property from that h (src/thermo_info.f90 ftp_refresh_thermal_properties_
// from_DH).  crh, ch and cd may point into LDS; the other columns are read
// from tab.
__device__ __forceinline__ void props_dh_cell(const double* crh, const double* ch, const double* cd,
                                              const double* __restrict__ tab, int n, int hrise, double dhmin,
                                              double dhmax,
                                              double v, long c, double* __restrict__ oh, double* __restrict__ od,
                                              double* __restrict__ om, double* __restrict__ ok,
                                              double* __restrict__ ot, double* __restrict__ ob,
                                              double* __restrict__ ocp, int* __restrict__ flag,
                                              float* __restrict__ fh, float* __restrict__ fm,
                                              float* __restrict__ fk, float* __restrict__ ft) {
  if (v < dhmin) {
    v = dhmin;
    atomicAdd(flag, 1);
  } else if (v > dhmax) {
    v = dhmax;
    atomicAdd(flag, 1);
  }
  int i1, i2;
  bisect_dev(crh, n, v, &i1, &i2);
  double w1 = (crh[i2] - v) / (crh[i2] - crh[i1]);
  double w2 = 1.0 - w1;
  double h = w1 * ch[i1] + w2 * ch[i2];
  // properties from h (first pass needs only d).  On a strictly increasing
  // h column the bisection for a value h with ch[k] + kNodeGap < h <
  // ch[k+1] ends in (k, k+1): every sign product it tests is then at least
  // (h - ch[k])^2 > kMinP.  So where h lies so inside the interval already
  // found, that interval is the bisection's result and the search is skipped.
  if (!(hrise && ch[i1] + kNodeGap < h && h < ch[i2])) bisect_dev(ch, n, h, &i1, &i2);
  w1 = (ch[i2] - h) / (ch[i2] - ch[i1]);
  w2 = 1.0 - w1;
  const double d1 = w1 * cd[i1] + w2 * cd[i2];
  h = v / d1;
  if (!(hrise && ch[i1] + kNodeGap < h && h < ch[i2])) bisect_dev(ch, n, h, &i1, &i2);
  w1 = (ch[i2] - h) / (ch[i2] - ch[i1]);
  w2 = 1.0 - w1;
  auto interp = [&](int col) { return w1 * tab[(long)col * n + i1] + w2 * tab[(long)col * n + i2]; };
  if (oh) oh[c] = h;
  if (od) od[c] = w1 * cd[i1] + w2 * cd[i2];
  if (om || fm) {
    const double m = interp(CM);
    if (om) om[c] = m;
    if (fm) fm[c] = (float)m;
  }
  if (ok || fk) {
    const double k = interp(CK);
    if (ok) ok[c] = k;
    if (fk) fk[c] = (float)k;
  }
  if (ot || ft) {
    const double t = interp(CT);
    if (ot) ot[c] = t;
    if (ft) ft[c] = (float)t;
  }
  if (fh) fh[c] = (float)h;
  if (ob) ob[c] = interp(CB);
  if (ocp) ocp[c] = interp(CCP);
}

// Table in global memory (any size): one thread per cell.
__global__ void __launch_bounds__(256) k_props_from_dh(const double* __restrict__ tab, int n, int hrise, double dhmin,
                                                       double dhmax, const double* __restrict__ rhoh_in, long N,
                                                       double* __restrict__ oh, double* __restrict__ od,
                                                       double* __restrict__ om, double* __restrict__ ok,
                                                       double* __restrict__ ot, double* __restrict__ ob,
                                                       double* __restrict__ ocp, int* __restrict__ flag,
                                                       float* __restrict__ fh, float* __restrict__ fm,
                                                       float* __restrict__ fk, float* __restrict__ ft) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  props_dh_cell(tab + (long)CRH * n, tab + (long)CH * n, tab + (long)CD * n, tab, n, hrise, dhmin, dhmax, rhoh_in[c], c,
                oh, od, om, ok, ot, ob, ocp, flag, fh, fm, fk, ft);
}

// The bisection columns (rho*h and h, and d when NC is 3: NC n doubles)
// staged in LDS once per workgroup; a grid of about two workgroups per CU
// then walks the cells with a grid stride, so the table crosses L2 once per
// workgroup instead of once per few hundred cells.  Every thread reaches the
// barrier: the staging loop runs before any thread leaves.
template <int NC>
__global__ void __launch_bounds__(kPropsBlock) k_props_from_dh_lds(
    const double* __restrict__ tab, int n, int hrise, double dhmin, double dhmax, const double* __restrict__ rhoh_in,
    long N,
    double* __restrict__ oh, double* __restrict__ od, double* __restrict__ om, double* __restrict__ ok,
    double* __restrict__ ot, double* __restrict__ ob, double* __restrict__ ocp, int* __restrict__ flag,
    float* __restrict__ fh, float* __restrict__ fm, float* __restrict__ fk, float* __restrict__ ft) {
  extern __shared__ double sh[];
  const double* crh = tab + (long)CRH * n;
  const double* ch = tab + (long)CH * n;
  const double* cd = tab + (long)CD * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sh[i] = crh[i];
    sh[n + i] = ch[i];
    if (NC == 3) sh[2 * n + i] = cd[i];
  }
  __syncthreads();
  const double* sd = NC == 3 ? sh + 2 * n : cd;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += stride)
    props_dh_cell(sh, sh + n, sd, tab, n, hrise, dhmin, dhmax, rhoh_in[c], c, oh, od, om, ok, ot, ob, ocp, flag, fh,
                  fm, fk, ft);
}

// ftp_refresh_thermal_properties_from_H per cell (no clamping, as the
// reference: outside the table the end segment extrapolates)
template <bool LDS>
__global__ void __launch_bounds__(256) k_props_from_h(const double* __restrict__ tab, int n,
                                                      const double* __restrict__ h_in, long N,
                                                      double* __restrict__ od, double* __restrict__ om,
                                                      double* __restrict__ ok, double* __restrict__ ot,
                                                      double* __restrict__ ob, double* __restrict__ ocp) {
  extern __shared__ double sh[];
  const double* ch = tab + (long)CH * n;
  if (LDS) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) sh[i] = ch[i];
    __syncthreads();
    ch = sh;
  }
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const double h = h_in[c];
  int i1, i2;
  bisect_dev(ch, n, h, &i1, &i2);
  const double w1 = (ch[i2] - h) / (ch[i2] - ch[i1]);
  const double w2 = 1.0 - w1;
  auto interp = [&](int col) { return w1 * tab[(long)col * n + i1] + w2 * tab[(long)col * n + i2]; };
  if (od) od[c] = interp(CD);
  if (om) om[c] = interp(CM);
  if (ok) ok[c] = interp(CK);
  if (ot) ot[c] = interp(CT);
  if (ob) ob[c] = interp(CB);
  if (ocp) ocp[c] = interp(CCP);
}

#pragma clang fp contract(on)

// Launches the from-rho*h kernel for n cells: the persistent LDS form with as
// many bisection columns as fit a workgroup's LDS (three or two), else the
// global form.
// o holds the seven outputs (h, d, m, k, t, b, cp), each possibly null; f
// the FP32 roundings of h, m, k, t (or null).
hipError_t launch_from_dh(const chs_props* p, const double* rhoh, long n, double* const o[7], int* flag,
                          hipStream_t s, float* const* f = nullptr) {
  float* const fz[4] = {nullptr, nullptr, nullptr, nullptr};
  if (!f) f = fz;
  const size_t col = sizeof(double) * (size_t)p->n;
  const int hrise = p->h_rising ? 1 : 0;
  const int nc = 3 * col <= p->lds_block ? 3 : 2 * col <= p->lds_block ? 2 : 0;
  if (nc) {
    const long want = (n + kPropsBlock - 1) / kPropsBlock;
    const long cap = 2L * (p->cus > 0 ? p->cus : 256);
    const unsigned grid = (unsigned)(want < cap ? want : cap);
    if (nc == 3)
      hipLaunchKernelGGL(k_props_from_dh_lds<3>, dim3(grid), dim3(kPropsBlock), 3 * col, s, p->dev, p->n, hrise, p->dhmin,
                         p->dhmax, rhoh, n, o[0], o[1], o[2], o[3], o[4], o[5], o[6], flag, f[0], f[1], f[2], f[3]);
    else
      hipLaunchKernelGGL(k_props_from_dh_lds<2>, dim3(grid), dim3(kPropsBlock), 2 * col, s, p->dev, p->n, hrise, p->dhmin,
                         p->dhmax, rhoh, n, o[0], o[1], o[2], o[3], o[4], o[5], o[6], flag, f[0], f[1], f[2], f[3]);
  } else {
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_props_from_dh, dim3(grid), dim3(256), 0, s, p->dev, p->n, hrise, p->dhmin, p->dhmax, rhoh, n,
                       o[0], o[1], o[2], o[3], o[4], o[5], o[6], flag, f[0], f[1], f[2], f[3]);
  }
  return hipGetLastError();
}

}  // namespace

namespace chs {
std::string props_refresh_dh(const chs_props* p, int mode, const double* rhoh, long n, double* h, double* d,
                             double* m, double* k, double* t, int* flag, hipStream_t s, float* const* r32) {
  if (n <= 0) return "";
  if (mode == CHS_PROPS_MLP) {
    if (!p->surrogate) return "properties: no surrogate attached to the table";
    return mlp_refresh_dh(p->surrogate, p, rhoh, n, h, d, m, k, t, flag, s, r32);
  }
  if (mode != CHS_PROPS_TABLE) return "properties: unknown mode";
  double* const o[7] = {h, d, m, k, t, nullptr, nullptr};
  const hipError_t e = launch_from_dh(p, rhoh, n, o, flag, s, r32);
  if (e != hipSuccess) return std::string("property kernel: ") + hipGetErrorString(e);
  return "";
}
}  // namespace chs

extern "C" {

int chs_props_set_surrogate(chs_props_t* p, chs_mlp_t* m) {
  if (!p) return fail(CHS_EINVAL, "null handle");
  p->surrogate = m;
  return CHS_OK;
}

int chs_props_create(const double* table, int nrows, double ref_t0, chs_props_t** out) {
  if (!table || !out || nrows < 2) return fail(CHS_EINVAL, "invalid argument");
  *out = nullptr;
  chs_props* p = new chs_props();
  p->n = nrows;
  for (auto& c : p->col) c.resize(nrows);
  for (int i = 0; i < nrows; ++i) {
    const double* r = table + 8 * (size_t)i;
    p->col[CH][i] = r[1];
    p->col[CT][i] = r[2];
    p->col[CD][i] = r[3];
    p->col[CM][i] = r[4];
    p->col[CK][i] = r[5];
    p->col[CCP][i] = r[6];
    p->col[CB][i] = r[7];
    p->col[CRH][i] = r[3] * r[1];
  }
  // selection sort by T (the first minimum wins, as minloc does)
  for (int i = 0; i < nrows; ++i) {
    int kmin = i;
    for (int j = i + 1; j < nrows; ++j)
      if (p->col[CT][j] < p->col[CT][kmin]) kmin = j;
    if (kmin != i)
      for (auto& c : p->col) std::swap(c[i], c[kmin]);
  }
  if (ref_t0 > 0.0) {
    int i1, i2;
    bisect_host(p->col[CT], ref_t0, &i1, &i2);
    const double w1 = (p->col[CT][i2] - ref_t0) / (p->col[CT][i2] - p->col[CT][i1]), w2 = 1.0 - w1;
    auto in = [&](int c) {
      const double a = w1 * p->col[c][i1];
      const double b = w2 * p->col[c][i2];
      return a + b;
    };
    p->ref[0] = ref_t0;
    p->ref[1] = in(CD);
    p->ref[2] = in(CM);
    p->ref[3] = in(CK);
    p->ref[4] = in(CH);
    p->ref[5] = in(CB);
    p->ref[6] = in(CCP);
    for (int i = 0; i < nrows; ++i) {
      p->col[CT][i] = p->col[CT][i] / p->ref[0];
      p->col[CD][i] = p->col[CD][i] / p->ref[1];
      p->col[CM][i] = p->col[CM][i] / p->ref[2];
      p->col[CK][i] = p->col[CK][i] / p->ref[3];
      p->col[CB][i] = p->col[CB][i] / p->ref[5];
      p->col[CCP][i] = p->col[CCP][i] / p->ref[6];
      p->col[CH][i] = (p->col[CH][i] - p->ref[4]) / p->ref[0] / p->ref[6];
      p->col[CRH][i] = p->col[CD][i] * p->col[CH][i];
    }
  }
  // ref_t0 <= 0: the rows are already non-dimensional (ref stays zero)
  int imin = 0, imax = 0;
  for (int i = 1; i < nrows; ++i) {
    if (p->col[CRH][i] < p->col[CRH][imin]) imin = i;
    if (p->col[CRH][i] > p->col[CRH][imax]) imax = i;
  }
  p->dhmin = p->col[CRH][imin];
  p->dhmax = p->col[CRH][imax];
  p->h_rising = true;
  for (int i = 1; i < nrows; ++i)
    if (!(p->col[CH][i] > p->col[CH][i - 1])) p->h_rising = false;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete p;
    return fail(CHS_EHIP, "no HIP device");
  }
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      p->lds_block = (size_t)v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) p->cus = v;
  }
  std::vector<double> flat(8 * (size_t)nrows);
  for (int c = 0; c < 8; ++c) std::memcpy(flat.data() + (size_t)c * nrows, p->col[c].data(), sizeof(double) * nrows);
  if (hipStreamCreate(&p->stream) != hipSuccess || hipMalloc(&p->dev, sizeof(double) * flat.size()) != hipSuccess ||
      hipMalloc(&p->flag, sizeof(int)) != hipSuccess ||
      hipMemcpy(p->dev, flat.data(), sizeof(double) * flat.size(), hipMemcpyHostToDevice) != hipSuccess) {
    delete p;
    return fail(CHS_ENOMEM, "device allocation failed");
  }
  *out = p;
  return CHS_OK;
}

int chs_props_destroy(chs_props_t* p) {
  if (!p) return CHS_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  delete p;
  return CHS_OK;
}

int chs_props_info(const chs_props_t* p, double* dhmin, double* dhmax, double ref[7]) {
  if (!p) return fail(CHS_EINVAL, "null handle");
  if (dhmin) *dhmin = p->dhmin;
  if (dhmax) *dhmax = p->dhmax;
  if (ref)
    for (int i = 0; i < 7; ++i) ref[i] = p->ref[i];
  return CHS_OK;
}

int chs_props_from_rhoh(chs_props_t* p, const double* rhoh, long n, double* h, double* d, double* m, double* k,
                        double* t, double* b, double* cp, int mode, int mem) {
  if (!p || !rhoh || n < 0) return fail(CHS_EINVAL, "invalid argument");
  if (mode != CHS_PROPS_TABLE && mode != CHS_PROPS_MLP) return fail(CHS_EUNSUP, "unknown property mode");
  if (mode == CHS_PROPS_MLP && !p->surrogate) return fail(CHS_EINVAL, "no surrogate attached to the table");
  if (mode == CHS_PROPS_MLP && (b || cp)) return fail(CHS_EUNSUP, "the surrogate gives h, d, m, k and t only");
  if (mem != CHS_MEM_HOST && mem != CHS_MEM_DEVICE) return fail(CHS_EINVAL, "invalid mem kind");
  if (n == 0) return CHS_OK;
  double* outs[7] = {h, d, m, k, t, b, cp};
  const double* din = rhoh;
  double* dout[7] = {h, d, m, k, t, b, cp};
  if (mem == CHS_MEM_HOST) {
    int nout = 0;
    for (double* o : outs) nout += o != nullptr;
    const size_t need = sizeof(double) * (size_t)n * (1 + nout);
    if (need > p->io_bytes) {
      if (p->io) (void)hipFree(p->io);
      p->io = nullptr;
      p->io_bytes = 0;
      if (hipMalloc(&p->io, need) != hipSuccess) return fail(CHS_ENOMEM, "staging allocation failed");
      p->io_bytes = need;
    }
    double* base = static_cast<double*>(p->io);
    if (hipMemcpyAsync(base, rhoh, sizeof(double) * n, hipMemcpyHostToDevice, p->stream) != hipSuccess)
      return fail(CHS_EHIP, "copy in");
    din = base;
    int slot = 1;
    for (int q = 0; q < 7; ++q) dout[q] = outs[q] ? base + (size_t)n * slot++ : nullptr;
  }
  if (hipMemsetAsync(p->flag, 0, sizeof(int), p->stream) != hipSuccess) return fail(CHS_EHIP, "flag reset");
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (mode == CHS_PROPS_MLP) {
    const std::string e = chs::mlp_refresh_dh(p->surrogate, p, din, n, dout[0], dout[1], dout[2], dout[3], dout[4],
                                              p->flag, p->stream);
    if (!e.empty()) return fail(CHS_EHIP, e);
  } else if (launch_from_dh(p, din, n, dout, p->flag, p->stream) != hipSuccess) {
    return fail(CHS_EHIP, "property kernel launch");
  }
  if (mem == CHS_MEM_HOST)
    for (int q = 0; q < 7; ++q)
      if (outs[q] && hipMemcpyAsync(outs[q], dout[q], sizeof(double) * n, hipMemcpyDeviceToHost, p->stream) !=
                         hipSuccess)
        return fail(CHS_EHIP, "copy out");
  int nbad = 0;
  if (hipMemcpyAsync(&nbad, p->flag, sizeof(int), hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess)
    return fail(CHS_EHIP, "synchronise");
  if (nbad) return fail(CHS_ERANGE, "rho*h is out of range (" + std::to_string(nbad) + " cells clamped)");
  return CHS_OK;
}

int chs_props_from_h(chs_props_t* p, const double* h, long n, double* d, double* m, double* k, double* t,
                     double* b, double* cp, int mem) {
  if (!p || !h || n < 0) return fail(CHS_EINVAL, "invalid argument");
  if (mem != CHS_MEM_HOST && mem != CHS_MEM_DEVICE) return fail(CHS_EINVAL, "invalid mem kind");
  if (n == 0) return CHS_OK;
  double* outs[6] = {d, m, k, t, b, cp};
  const double* din = h;
  double* dout[6] = {d, m, k, t, b, cp};
  if (mem == CHS_MEM_HOST) {
    int nout = 0;
    for (double* o : outs) nout += o != nullptr;
    const size_t need = sizeof(double) * (size_t)n * (1 + nout);
    if (need > p->io_bytes) {
      if (p->io) (void)hipFree(p->io);
      p->io = nullptr;
      p->io_bytes = 0;
      if (hipMalloc(&p->io, need) != hipSuccess) return fail(CHS_ENOMEM, "staging allocation failed");
      p->io_bytes = need;
    }
    double* base = static_cast<double*>(p->io);
    if (hipMemcpyAsync(base, h, sizeof(double) * n, hipMemcpyHostToDevice, p->stream) != hipSuccess)
      return fail(CHS_EHIP, "copy in");
    din = base;
    int slot = 1;
    for (int q = 0; q < 6; ++q) dout[q] = outs[q] ? base + (size_t)n * slot++ : nullptr;
  }
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (p->n <= 3 * kPropsLds)
    hipLaunchKernelGGL(k_props_from_h<true>, dim3(grid), dim3(256), sizeof(double) * (size_t)p->n, p->stream, p->dev,
                       p->n, din, n, dout[0], dout[1], dout[2], dout[3], dout[4], dout[5]);
  else
    hipLaunchKernelGGL(k_props_from_h<false>, dim3(grid), dim3(256), 0, p->stream, p->dev, p->n, din, n, dout[0],
                       dout[1], dout[2], dout[3], dout[4], dout[5]);
  if (hipGetLastError() != hipSuccess) return fail(CHS_EHIP, "property kernel launch");
  if (mem == CHS_MEM_HOST)
    for (int q = 0; q < 6; ++q)
      if (outs[q] && hipMemcpyAsync(outs[q], dout[q], sizeof(double) * n, hipMemcpyDeviceToHost, p->stream) !=
                         hipSuccess)
        return fail(CHS_EHIP, "copy out");
  if (hipStreamSynchronize(p->stream) != hipSuccess) return fail(CHS_EHIP, "synchronise");
  return CHS_OK;
}

}  // extern "C"
