// Device-resident momentum RHS, RK3 fractional step, projection and
// boundary enforcement (Cartesian, constant properties, x and z periodic).
//
// Sequence of one sub-step follows Solve_momentum_eq
// (src/eq_momentum2.f90:1799-1932):
//   Compute_momentum_rhs (:101-1571) -> q += rhs -> enforce_velo_from_fbc
//   -> solve_poisson (:1663-1718) -> p += phi -> Correct_massflux (:1577-1660)
//   -> enforce_velo_from_fbc.
// Elementwise products of the RHS are fused into the loads of the operator
// that consumes them; each directional contribution accumulates in place in
// its own array, and the three are summed in the reference's order by the
// fractional-step kernel (Calculate_momentum_fractional_step, :35-71).
//
// Precision forms: FP64 everywhere, or mixed (CHS_PREC_MIXED): the momentum
// right-hand side (44 line operators, products, fractional-step combination)
// runs in FP32 on rounded copies of the state, while the state update, the
// bulk-flux sum, the divergence, the pressure solve and the velocity
// correction stay FP64.
#include "chs_flow.h"
#include "chs_exact.h"

#include <cmath>
#include <cstring>
#include <mutex>

namespace chs {

// ---------------------------------------------------------------------------
// Static block partials of the allocation-free velocity range
// (chs_velocity_range_noalloc): 3 x kMaxPartials doubles (12 KiB) per device,
// held by each device the library runs on until the process ends.  In host
// code the array's name is not a device address: kernels and copies take the
// address hipGetSymbolAddress returns for the current device, looked up once
// per device when a handle's flow solver is created.  Every handle on a
// device shares the array, so calls that overlap on one device race; each
// call overwrites all of it before reading it back.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[3 * kMaxPartials];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// Device address of `scratch` on the current device (looked up on the first
// call for that device).
static std::string range_scratch_address(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return std::string("hipGetDevice: ") + hipGetErrorString(e);
  if (dev < 0 || dev >= kMaxScratchDevices) return "device index beyond the static scratch table";
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    double* p = nullptr;
    e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return "";
}

#define FCHK(x)                                                             \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) return std::string(#x ": ") + hipGetErrorString(e_); \
  } while (0)
#define SCHK(x)                      \
  do {                               \
    std::string s_ = (x);            \
    if (!s_.empty()) return s_;      \
  } while (0)

static const double ONE_THIRD = 1.0 / 3.0;
enum { VU = 0, VV = 1, VW = 2, VP = 3, VT = 4 };

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_div3(T* __restrict__ d, const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
                       long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    T v = T(0);
    v = v + a[i];
    v = v + b[i];
    v = v + c[i];
    d[i] = v;
  }
}

// rhs = dt * (tg*((ax + ay) + az) + tz*rhs0 + ta*pfc); rhs0 = (ax + ay) + az.
// Optionally accumulates the x-pencil volume integral of rhs per block (FP64).
// With q given, the state update of k_update is applied in place of storing
// rhs: q += rhs, then the wall planes j = 0 and n2 - 1 are set from fbw
// (extents n1 x n2 x n / (n1 n2), as q's).
template <typename T>
__global__ void k_combine(T* __restrict__ ax, const T* __restrict__ ay, const T* __restrict__ az,
                          const T* __restrict__ pfc, T* __restrict__ r0, long n, T tg, T tz, T ta, T dt, int n1,
                          int n2, const double* __restrict__ dyw, double dxdz_x, double dxdz_z,
                          double* __restrict__ partial, double* __restrict__ q, const double* __restrict__ fbw) {
  __shared__ double red[256];
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double w = 0.0;
  if (i < n) {
    T cur = ax[i];
    cur = cur + ay[i];
    cur = cur + az[i];
    T tot = tg * cur + tz * r0[i];
    r0[i] = cur;
    tot = tot + ta * pfc[i];
    T rv = dt * tot;
    if (q) {
      double v = q[i] + (double)rv;
      if (fbw) {
        long i1 = i % n1, rest = i / n1;
        long j = rest % n2, k = rest / n2;
        if (j == 0) v = fbw[i1 + (long)n1 * (0 + 4 * k)];
        else if (j == n2 - 1) v = fbw[i1 + (long)n1 * (1 + 4 * k)];
      }
      q[i] = v;
    } else {
      ax[i] = rv;
    }
    if (partial) {
      int j = (int)((i / n1) % n2);
      w = (double)rv * dyw[j] * dxdz_x * dxdz_z;
    }
  }
  if (partial) {
    red[threadIdx.x] = w;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
  }
}

// Sum the block partials in a fixed order; bulk = sum / vol.  One workgroup
// of 1024 threads: 32 partials per thread loaded at once (one memory round
// trip for up to 32 768 block partials), summed per thread, across each wave
// by lane exchanges, then the 16 wave sums by one wave.
__global__ void __launch_bounds__(1024) k_finish_sum(const double* __restrict__ partial, int np, double vol,
                                                     double* __restrict__ bulk) {
  __shared__ double red[16];
  double s = 0.0;
  int i = threadIdx.x;
  for (; i + 31 * 1024 < np; i += 32 * 1024) {
    double v[32];
#pragma unroll
    for (int u = 0; u < 32; ++u) v[u] = partial[i + u * 1024];
#pragma unroll
    for (int u = 0; u < 32; ++u) s = s + v[u];
  }
  for (; i < np; i += 1024) s = s + partial[i];
#pragma unroll
  for (int k = 32; k > 0; k >>= 1) s = s + __shfl_xor(s, k);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x < 64) {
    double t = threadIdx.x < 16 ? red[threadIdx.x] : 0.0;
#pragma unroll
    for (int k = 8; k > 0; k >>= 1) t = t + __shfl_xor(t, k);
    if (threadIdx.x == 0) bulk[0] = t / vol;
  }
}

// q = q + (rhs - bulk) in FP64; y walls of qy set from fby (enforce_velo_from_fbc)
template <typename T>
__global__ void k_update(double* __restrict__ q, const T* __restrict__ rhs, const double* __restrict__ bulk, long n,
                         int n1, int n2, const double* __restrict__ fbw) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = (double)rhs[i];
  if (bulk) r = r - bulk[0];
  double v = q[i] + r;
  if (fbw) {
    long i1 = i % n1, rest = i / n1;
    long j = rest % n2, k = rest / n2;
    if (j == 0) v = fbw[i1 + (long)n1 * (0 + 4 * k)];
    else if (j == n2 - 1) v = fbw[i1 + (long)n1 * (1 + 4 * k)];
  }
  q[i] = v;
}

template <typename T>
__global__ void k_sub_bulk(double* __restrict__ o, const T* __restrict__ a, const double* __restrict__ bulk, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (double)a[i] - (bulk ? bulk[0] : 0.0);
}

template <typename TO, typename TI>
__global__ void k_cvt(TO* __restrict__ o, const TI* __restrict__ a, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (TO)a[i];
}

// a += b (and, with a32, its FP32 rounding into a32)
__global__ void k_add(double* __restrict__ a, const double* __restrict__ b, long n, float* __restrict__ a32) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = a[i] + b[i];
    a[i] = v;
    if (a32) a32[i] = (float)v;
  }
}

__global__ void k_enforce_walls(double* __restrict__ q, int n1, int n2, int n3, const double* __restrict__ fbw,
                                float* __restrict__ q32) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n1 * n3) return;
  long i = t % n1, k = t / n1;
  const long o0 = i + (long)n1 * (0 + (long)n2 * k), o1 = i + (long)n1 * ((n2 - 1) + (long)n2 * k);
  const double v0 = fbw[i + (long)n1 * (0 + 4 * k)], v1 = fbw[i + (long)n1 * (1 + 4 * k)];
  q[o0] = v0;
  q[o1] = v1;
  if (q32) {
    q32[o0] = (float)v0;
    q32[o1] = (float)v1;
  }
}

// planes (n1,2,n3) <- a(:,0,:), a(:,n2-1,:)
template <typename T>
__global__ void k_extract_walls(T* __restrict__ pl, const T* __restrict__ a, int n1, int n2, int n3) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n1 * n3) return;
  long i = t % n1, k = t / n1;
  pl[i + (long)n1 * (0 + 2 * k)] = a[i + (long)n1 * (0 + (long)n2 * k)];
  pl[i + (long)n1 * (1 + 2 * k)] = a[i + (long)n1 * ((n2 - 1) + (long)n2 * k)];
}

// div and visc wall values (nx,4,nz) from the three wall-plane derivatives
// (times the wall viscosity mu (nx,4,nz) with variable properties, else 1)
template <typename T>
__global__ void k_wall_fbc(T* __restrict__ fdiv, T* __restrict__ fvisc, const T* __restrict__ dxw,
                           const T* __restrict__ dyw, const T* __restrict__ dzw, int n1, int n3,
                           const T* __restrict__ mu) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n1 * n3) return;
  long i = t % n1, k = t / n1;
  for (int s = 0; s < 2; ++s) {
    long p = i + (long)n1 * (s + 2 * k);
    T dv = dxw[p] + dyw[p] + dzw[p];
    T vs = (dyw[p] - T(ONE_THIRD) * dv) * T(2) * (mu ? mu[i + (long)n1 * (s + 4 * k)] : T(1));
    fdiv[i + (long)n1 * (s + 4 * k)] = dv;
    fdiv[i + (long)n1 * (s + 2 + 4 * k)] = dv;
    fvisc[i + (long)n1 * (s + 4 * k)] = vs;
    fvisc[i + (long)n1 * (s + 2 + 4 * k)] = vs;
  }
}

// ---- variable-property kernels ----
template <typename T>
__global__ void k_fill(T* __restrict__ o, T v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = v;
}
// wall rows j = 0 and n2 - 1 of a (times b when given) into (n1,4,n3):
// rows 1,2 and again rows 3,4 (extract_dirichlet_fbcy, src/bc_dirichlet.f90:42-56)
template <typename T>
__global__ void k_wall_rows4(T* __restrict__ o, const T* __restrict__ a, const T* __restrict__ b, int n1, int n2,
                             int n3) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n1 * n3) return;
  long i = t % n1, k = t / n1;
  for (int s = 0; s < 2; ++s) {
    const long src = i + (long)n1 * ((s ? n2 - 1 : 0) + (long)n2 * k);
    const T v = b ? a[src] * b[src] : a[src];
    o[i + (long)n1 * (s + 4 * k)] = v;
    o[i + (long)n1 * (s + 2 + 4 * k)] = v;
  }
}
// g = q d (IQ2G) or q = g / d (IG2Q), src/para_conversion.f90:41-85
__global__ void k_q2g(double* __restrict__ g, const double* __restrict__ q, const double* __restrict__ d, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = q[i] * d[i];
}
__global__ void k_g2q(double* __restrict__ q, const double* __restrict__ g, const double* __restrict__ d, long n,
                      float* __restrict__ q32) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = g[i] / d[i];
    q[i] = v;
    if (q32) q32[i] = (float)v;
  }
}
// energy: rhs = ex + (ey + ez) (the pencil order of src/eq_energy.f90:386-391),
// Calculate_energy_fractional_step (:148-176): res = dt (tg rhs + tz rhs0),
// rhs0 = rhs; with update, rho*h += res (Solve_energy_eq, :440)
template <typename TI>
__global__ void k_energy_combine(const TI* __restrict__ ex, const TI* __restrict__ ey, const TI* __restrict__ ez,
                                 double* __restrict__ r0, double* __restrict__ res, double* __restrict__ rhoh, long n,
                                 double tg, double tz, double dt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double yz = (double)ey[i] + (double)ez[i];
  const double cur = (double)ex[i] + yz;
  double tot = tg * cur + tz * r0[i];
  r0[i] = cur;
  const double rv = dt * tot;
  if (res) res[i] = rv;
  if (rhoh) rhoh[i] = rhoh[i] + rv;
}

template <typename T>
__global__ void k_neg_prod(T* __restrict__ o, const T* __restrict__ a, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = -a[i] * a[i];
}
template <typename T>
__global__ void k_neg(T* __restrict__ o, const T* __restrict__ a, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = -a[i];
}

// maximum that propagates NaN (fmax would drop it and hide a blow-up)
__device__ __forceinline__ double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

// max |div| over x windows [lo, hi] (0-based inclusive), per block partials
__global__ void k_maxabs_xwin(const double* __restrict__ d, int n1, long n, int lo, int hi,
                              double* __restrict__ partial) {
  __shared__ double red[256];
  double m = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x = (int)(i % n1);
    if (x >= lo && x <= hi) m = nanmax(m, fabs(d[i]));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = nanmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Block maxima of |u h1r_x dt + v h1r_y dt ymc(j) + w h1r_z dt| over cell
// centres (x fastest), the sum formed in the order of Check_cfl_convection
// (src/tools_solver.f90:301-344); ymc may be null (uniform y).
__global__ void k_cfl_max(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                          int n1, int n2, long n, double hx, double hy, double hz, double dt,
                          const double* __restrict__ ymc, double* __restrict__ partial) {
  __shared__ double red[256];
  double m = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int j = (int)((i / n1) % n2);
    const double cx = u[i] * hx * dt;
    double cy = v[i] * hy * dt;
    if (ymc) cy = cy * ymc[j];
    const double c = (cx + cy) + w[i] * hz * dt;
    m = nanmax(m, fabs(c));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = nanmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Block maxima and minima of one field and its count of non-finite values
// (partial[b], partial[nb + b], partial[2 nb + b] for block b of nb).  The
// comparisons skip NaN and start from -1e20 / 1e20, as the loops of
// Find_max_min_3d (src/tools_general.f90:1071-1113) do; max and min are
// exact, so the block order does not matter.
__global__ void k_range(const double* __restrict__ a, long n, double* __restrict__ partial) {
  __shared__ double smax[256], smin[256], scnt[256];
  double mx = -1.0e20, mn = 1.0e20, cnt = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double v = a[i];
    if (v > mx) mx = v;
    if (v < mn) mn = v;
    if (!isfinite(v)) cnt += 1.0;
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      if (smax[threadIdx.x + s] > smax[threadIdx.x]) smax[threadIdx.x] = smax[threadIdx.x + s];
      if (smin[threadIdx.x + s] < smin[threadIdx.x]) smin[threadIdx.x] = smin[threadIdx.x + s];
      scnt[threadIdx.x] = scnt[threadIdx.x] + scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = smax[0];
    partial[gridDim.x + blockIdx.x] = smin[0];
    partial[2 * gridDim.x + blockIdx.x] = scnt[0];
  }
}

// Block sums of f(i) dyw(j) over an x-pencil array (x fastest, j = row
// along y), f = a (b, c null) or 1/2 (a^2 + b^2 + c^2): the volume integrals
// of Get_volumetric_average_3d_for_var_xcx (src/tools_general.f90:1202-1268)
// before the dx dz factor.
__global__ void k_vol_sum(const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ c,
                          int n1, int n2, long n, const double* __restrict__ dyw, double* __restrict__ partial) {
  __shared__ double red[256];
  double sum = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int j = (int)((i / n1) % n2);
    const double f = b ? 0.5 * (a[i] * a[i] + b[i] * b[i] + c[i] * c[i]) : a[i];
    sum += f * dyw[j];
  }
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

static inline dim3 grid1(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

// ---------------------------------------------------------------------------
// bc tables (src/bc_general.f90:137-245, 455-757), Cartesian subset
// ---------------------------------------------------------------------------
static void bc_symm(const int ibc[2], const int* jbc, int mbc[2][3]) {
  for (int i = 0; i < 2; ++i) {
    mbc[i][0] = mbc[i][1] = mbc[i][2] = ibc[i];
    if (jbc) {
      const int a = ibc[i], b = jbc[i];
      if (a == BC_SYMMETRIC && b == BC_SYMMETRIC) mbc[i][2] = BC_SYMMETRIC;
      else if (a == BC_SYMMETRIC && b == BC_ASYMMETRIC) mbc[i][2] = BC_ASYMMETRIC;
      else if (a == BC_ASYMMETRIC && b == BC_SYMMETRIC) mbc[i][2] = BC_ASYMMETRIC;
      else if (a == BC_ASYMMETRIC && b == BC_ASYMMETRIC) mbc[i][2] = BC_SYMMETRIC;
      else if (a != b) {
        if (a == BC_DIRICHLET) mbc[i][0] = mbc[i][1] = mbc[i][2] = a;
        if (b == BC_DIRICHLET) mbc[i][0] = mbc[i][1] = mbc[i][2] = b;
      }
    } else {
      if (ibc[i] == BC_SYMMETRIC) mbc[i][1] = BC_ASYMMETRIC;
      else if (ibc[i] == BC_ASYMMETRIC) { mbc[i][1] = BC_SYMMETRIC; mbc[i][2] = BC_SYMMETRIC; }
    }
  }
}
static void prod_bc(const int a[2], const int b[2], int o[2]) {
  int m[2][3];
  bc_symm(a, b, m);
  o[0] = m[0][2]; o[1] = m[1][2];
}
static void tau_bc(const int ftp[2], const int q[2], int o[2]) {
  int m[2][3], g[2];
  bc_symm(q, nullptr, m);
  g[0] = m[0][1]; g[1] = m[1][1];
  bc_symm(ftp, g, m);
  o[0] = m[0][2]; o[1] = m[1][2];
}

// ---------------------------------------------------------------------------
Flow::~Flow() {
  for (hipStream_t st : side)
    if (st) (void)hipStreamSynchronize(st);
  for (hipEvent_t e : {ev_fork, ev_join[0], ev_join[1]})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : side)
    if (st) (void)hipStreamDestroy(st);
  for (void* p : owned) (void)hipFree(p);
}

template <typename T>
std::string Flow::alloc(Field<T>& f, int e1, int e2, int e3) {
  f.e = Ext3{{e1, e2, e3}};
  FCHK(hipMalloc(&f.p, sizeof(T) * (size_t)f.size()));
  FCHK(hipMemset(f.p, 0, sizeof(T) * (size_t)f.size()));
  owned.push_back(f.p);
  return "";
}

// Work arrays of the right-hand side in precision T; with alias_state the
// velocity / pressure fields are the FP64 state itself.
template <typename T>
std::string Flow::alloc_work(RhsWork<T>& w, bool alias_state) {
  const int nx_ = nx, ny_ = ny, nz_ = nz, px_ = px, py_ = py, pz_ = pz;
  if (alias_state) {
    if constexpr (sizeof(T) == sizeof(double)) {
      w.qx = qx; w.qy = qy; w.qz = qz; w.pr = pr;
    } else {
      return "work: only FP64 work arrays may alias the state";
    }
  } else {
    SCHK(alloc(w.qx, px_, ny_, nz_)); SCHK(alloc(w.qy, nx_, py_, nz_)); SCHK(alloc(w.qz, nx_, ny_, pz_));
    SCHK(alloc(w.pr, nx_, ny_, nz_));
  }
  SCHK(alloc(w.mx0, px_, ny_, nz_)); SCHK(alloc(w.my0, nx_, py_, nz_)); SCHK(alloc(w.mz0, nx_, ny_, pz_));
  SCHK(alloc(w.qxix, nx_, ny_, nz_)); SCHK(alloc(w.qxdx, nx_, ny_, nz_)); SCHK(alloc(w.qxdy, px_, py_, nz_));
  SCHK(alloc(w.qxiy, px_, py_, nz_)); SCHK(alloc(w.qxiz, px_, ny_, pz_)); SCHK(alloc(w.qxdz, px_, ny_, pz_));
  SCHK(alloc(w.qydx, px_, py_, nz_)); SCHK(alloc(w.qyix, px_, py_, nz_)); SCHK(alloc(w.qydy, nx_, ny_, nz_));
  SCHK(alloc(w.qyiy, nx_, ny_, nz_)); SCHK(alloc(w.qydz, nx_, py_, pz_)); SCHK(alloc(w.qyiz, nx_, py_, pz_));
  SCHK(alloc(w.qzdx, px_, ny_, pz_)); SCHK(alloc(w.qzix, px_, ny_, pz_)); SCHK(alloc(w.qziy, nx_, py_, pz_));
  SCHK(alloc(w.qzdy, nx_, py_, pz_)); SCHK(alloc(w.qziz, nx_, ny_, nz_)); SCHK(alloc(w.qzdz, nx_, ny_, nz_));
  SCHK(alloc(w.div, nx_, ny_, nz_));
  const int cext[3][3] = {{px_, ny_, nz_}, {nx_, py_, nz_}, {nx_, ny_, pz_}};
  for (int c = 0; c < 3; ++c)
    for (int q = 0; q < 4; ++q) SCHK(alloc(w.acc[c][q], cext[c][0], cext[c][1], cext[c][2]));
  long big = 0;
  for (auto* fl : {&w.qx, &w.qy, &w.qz, &w.qxdy, &w.qydz, &w.qxiz}) big = std::max(big, fl->size());
  w.scratch_n = big;
  for (T** p : {&w.scratch, &w.scratch_side[0], &w.scratch_side[1]}) {
    FCHK(hipMalloc(p, sizeof(T) * (size_t)big));
    owned.push_back(*p);
  }
  // wall values (nx,4,nz) for u, v, w, p; constant walls as initialise_fbc_flow_given
  const chs_desc_t& d = H->desc;
  const long nf = (long)nx_ * 4 * nz_;
  auto wall = [&](T** p, double v0, double v1) -> std::string {
    std::vector<T> h(nf);
    for (long t = 0; t < nf; ++t) h[t] = T(((t / nx_) % 4) % 2 == 0 ? v0 : v1);
    FCHK(hipMalloc(p, sizeof(T) * nf));
    FCHK(hipMemcpy(*p, h.data(), sizeof(T) * nf, hipMemcpyHostToDevice));
    owned.push_back(*p);
    return "";
  };
  SCHK(wall(&w.fby_u, d.uwall[0][0], d.uwall[1][0]));
  SCHK(wall(&w.fby_v, d.uwall[0][1], d.uwall[1][1]));
  SCHK(wall(&w.fby_w, d.uwall[0][2], d.uwall[1][2]));
  SCHK(wall(&w.fby_p, 0.0, 0.0));
  SCHK(wall(&w.fby_pn, -0.0, -0.0));
  SCHK(wall(&w.fby_conv, 0.0, 0.0));
  SCHK(wall(&w.fby_visc, 0.0, 0.0));
  SCHK(wall(&w.fby_div, 0.0, 0.0));
  for (T** p : {&w.plane_a, &w.plane_b, &w.plane_c, &w.plane_d}) {
    FCHK(hipMalloc(p, sizeof(T) * (size_t)std::max(px_, nx_) * 2 * std::max(pz_, nz_)));
    owned.push_back(*p);
  }
  // the conv-y fbc of the y momentum is -fbcy_qy * fbcy_qy
  hipLaunchKernelGGL(k_neg_prod<T>, grid1(nf), dim3(256), 0, H->stream, w.fby_conv, w.fby_v, nf);
  hipLaunchKernelGGL(k_neg<T>, grid1(nf), dim3(256), 0, H->stream, w.fby_pn, w.fby_p, nf);
  w.ready = true;
  return "";
}

std::string Flow::create(Handle* H, std::unique_ptr<Flow>* out) {
  std::unique_ptr<Flow> F(new Flow());
  Flow& f = *F;
  f.H = H;
  const Mesh& m = H->mesh;
  const chs_desc_t& d = H->desc;
  f.nx = m.nc[0]; f.ny = m.nc[1]; f.nz = m.nc[2];
  f.px = m.np[0]; f.py = m.np[1]; f.pz = m.np[2];
  f.wall_y = !m.periodic[1];
  f.mixed = (d.precision == CHS_PREC_MIXED);
  f.dt = d.dt;
  f.sigma2p = 1.0;
  f.tg[0] = 1.0; f.tg[1] = 8.0 / 15.0; f.tg[2] = 5.0 / 12.0; f.tg[3] = 3.0 * 0.25;
  f.tz[0] = 0.0; f.tz[1] = 0.0; f.tz[2] = -17.0 / 60.0; f.tz[3] = -5.0 / 12.0;
  for (int i = 0; i < 4; ++i) f.ta[i] = f.tg[i] + f.tz[i];
  flow_bc_kinds(f.wall_y, f.ibc);
  int ftp[3][2];
  for (int dd = 0; dd < 3; ++dd)
    for (int s = 0; s < 2; ++s) ftp[dd][s] = (f.ibc[dd][VT][s] == BC_NEUMANN) ? BC_DIRICHLET : f.ibc[dd][VT][s];
  for (int c = 0; c < 3; ++c) {
    // convection: product of the advecting (diagonal) velocity and component c
    prod_bc(f.ibc[0][VU], f.ibc[0][c], f.mcov[c][0]);
    prod_bc(f.ibc[1][VV], f.ibc[1][c], f.mcov[c][1]);
    prod_bc(f.ibc[2][VW], f.ibc[2][c], f.mcov[c][2]);
    for (int dd = 0; dd < 3; ++dd) tau_bc(ftp[dd], f.ibc[dd][c], f.mtau[c][dd]);
  }
  f.update_re(0);
  // volume weights (Get_volumetric_average_3d_for_var_xcx)
  f.dyw.assign(f.ny, m.h[1]);
  if (m.stretched)
    for (int j = 0; j < f.ny; ++j) f.dyw[j] = m.h[1] / m.ymc[j];
  {
    double vol = 0.0;
    const double dx = m.h[0], dz = m.h[2];
    for (int j = 0; j < f.ny; ++j)
      for (int k = 0; k < f.nz; ++k)
        for (int i = 0; i < f.px; ++i) vol = vol + f.dyw[j] * dx * dz;
    f.vol = vol;
  }
  const int nx = f.nx, ny = f.ny, nz = f.nz, px = f.px, py = f.py, pz = f.pz;
  SCHK(f.alloc(f.qx, px, ny, nz)); SCHK(f.alloc(f.qy, nx, py, nz)); SCHK(f.alloc(f.qz, nx, ny, pz));
  SCHK(f.alloc(f.pr, nx, ny, nz)); SCHK(f.alloc(f.pcor, nx, ny, nz)); SCHK(f.alloc(f.div, nx, ny, nz));
  // the FP64 work set is always present (projection and its wall data use it)
  SCHK(f.alloc_work(f.w64, true));
  if (f.mixed) SCHK(f.alloc_work(f.w32, false));
  {
    long nblk = std::max<long>((f.qx.size() + 255) / 256, 3 * kMaxPartials);
    f.npartial = nblk;
    FCHK(hipMalloc(&f.partial, sizeof(double) * (size_t)(nblk + 8)));
    f.owned.push_back(f.partial);
    FCHK(hipMalloc(&f.bulk, sizeof(double) * 8));
    f.owned.push_back(f.bulk);
    FCHK(hipMalloc(&f.dyw_dev, sizeof(double) * (size_t)ny));
    FCHK(hipMemcpy(f.dyw_dev, f.dyw.data(), sizeof(double) * ny, hipMemcpyHostToDevice));
    f.owned.push_back(f.dyw_dev);
  }
  SCHK(range_scratch_address(&f.range_scratch));
  // side streams (default priority) and the events that fork and join them
  for (hipStream_t* st : {&f.side[0], &f.side[1]}) FCHK(hipStreamCreate(st));
  for (hipEvent_t* ev : {&f.ev_fork, &f.ev_join[0], &f.ev_join[1]})
    FCHK(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  FCHK(hipStreamSynchronize(H->stream));
  *out = std::move(F);
  return "";
}

void Flow::update_re(int iter) {
  const chs_desc_t& d = H->desc;
  rre = (iter < d.init_re_to) ? 1.0 / d.reninit : 1.0 / d.ren;
  // Update_PrGr, src/tools_solver.f90:63: rPrRen = rre k0 / (m0 cp0)
  rprren = rre / prandtl;
}

template <typename T>
std::string Flow::op(RhsWork<T>& w, int dir, int opk, const int ibc2[2], const Field<T>& A, Field<T>& out,
                     const T* fbc, T sA, const Field<T>* B, T cB, const Field<T>* Cm, T sOut, int accumulate,
                     const Field<T>* D, T cD, const Field<T>* E, T cE, int lane, float* out32) {
  const DevPlan<T>* P = nullptr;
  SCHK(H->plan<T>(dir, opk, H->desc.iacc, ibc2, fbc != nullptr, A.e.n[dir], out.e.n[dir], &P));
  LineIn<T> in;
  in.A = A.p;
  in.B = B ? B->p : nullptr;
  in.Cm = Cm ? Cm->p : nullptr;
  in.sA = sA;
  in.cB = cB;
  in.D = D ? D->p : nullptr;
  in.cD = cD;
  in.E = (D && E) ? E->p : nullptr;
  in.cE = cE;
  LineOut<T> o;
  o.out = out.p;
  o.sOut = sOut;
  o.accumulate = accumulate;
  o.out32 = out32;
  if (out.size() > w.scratch_n) return "flow: scratch too small";
  const hipStream_t s = lane_stream(lane);
  T* const scr = (lane == 0 || !lanes_on) ? w.scratch : w.scratch_side[lane - 1];
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = (H->prof_class == dir) && H->prof_pair(&e0, &e1);
  if (timed) {
    (void)hipEventRecord(e0, s);
    const int nin_arrays = 1 + (B ? 1 : 0) + (Cm && Cm->p != A.p ? 1 : 0) + (D ? 1 : 0) + (D && E ? 1 : 0);
    H->prof_bytes += (double)sizeof(T) *
                     ((double)A.size() * nin_arrays + (double)out.size() * (1 + (accumulate ? 1 : 0))) +
                 (out32 ? 4.0 * (double)out.size() : 0.0);
  }
  std::string r = launch_line_op<T>(*P, A.e, out.e, in, fbc, o, scr, s);
  if (timed) (void)hipEventRecord(e1, s);
  return r;
}

void Flow::set_lanes() {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  // (launch profiling times each kernel on its own, so it keeps one stream)
  lanes_on = concurrent && side[0] && H->prof_class < 0 && hipStreamIsCapturing(H->stream, &cs) == hipSuccess &&
             cs == hipStreamCaptureStatusNone;
}

template <typename T>
std::string Flow::op2(RhsWork<T>& w, int lane, int dir, const Field<T>& A, int opa, const int bca[2], Field<T>& outa,
                      const T* fbca, int opb, const int bcb[2], Field<T>& outb, const T* fbcb) {
  const DevPlan<T>* PA = nullptr;
  const DevPlan<T>* PB = nullptr;
  SCHK(H->plan<T>(dir, opa, H->desc.iacc, bca, fbca != nullptr, A.e.n[dir], outa.e.n[dir], &PA));
  SCHK(H->plan<T>(dir, opb, H->desc.iacc, bcb, fbcb != nullptr, A.e.n[dir], outb.e.n[dir], &PB));
  bool same = true;
  for (int k = 0; k < 3; ++k) same = same && outa.e.n[k] == outb.e.n[k];
  const hipStream_t s = lane_stream(lane);
  LineOut<T> oa, ob;
  oa.out = outa.p;
  ob.out = outb.p;
  T* const scr = (lane == 0 || !lanes_on) ? w.scratch : w.scratch_side[lane - 1];
  int passes = 0;
  if (same) SCHK(launch_line_op2<T>(*PA, *PB, A.e, outa.e, A.p, fbca, fbcb, oa, ob, scr, s, &passes, true));
  // not one pass: the two operators one by one (each launch timed on its own
  // when profiling)
  if (passes != 1) {
    auto one = [&](int opk, const int* bc, Field<T>& out, const T* fbc) {
      return op<T>(w, dir, opk, bc, A, out, fbc, T(1), nullptr, T(1), nullptr, T(1), 0, nullptr, T(0), nullptr,
                   T(0), lane);
    };
    SCHK(one(opa, bca, outa, fbca));
    return one(opb, bcb, outb, fbcb);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timed = (H->prof_class == dir) && H->prof_pair(&e0, &e1);
  if (timed) {
    (void)hipEventRecord(e0, s);
    H->prof_bytes += (double)sizeof(T) * ((double)A.size() + (double)outa.size() + (double)outb.size());
  }
  std::string r = launch_line_op2<T>(*PA, *PB, A.e, outa.e, A.p, fbca, fbcb, oa, ob, scr, s, &passes);
  if (timed) (void)hipEventRecord(e1, s);
  return r;
}

std::string Flow::fork() {
  if (!lanes_on) return "";
  // the side streams are never forked inside a graph capture (set_lanes
  // turns the lanes off while the handle's stream captures); should that
  // change, fail here instead of building the nested capture
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(H->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return "flow: side-stream lanes cannot be used while the stream is captured";
  if (hipEventRecord(ev_fork, H->stream) != hipSuccess) return "flow: fork record";
  for (hipStream_t st : side)
    if (hipStreamWaitEvent(st, ev_fork, 0) != hipSuccess) return "flow: fork wait";
  return "";
}

std::string Flow::join() {
  if (!lanes_on) return "";
  for (int i = 0; i < 2; ++i) {
    if (hipEventRecord(ev_join[i], side[i]) != hipSuccess) return "flow: join record";
    if (hipStreamWaitEvent(H->stream, ev_join[i], 0) != hipSuccess) return "flow: join wait";
  }
  return "";
}

std::string Flow::set_state(const double* hx, const double* hy, const double* hz, const double* hp, int mem) {
  q32_fresh = false;
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  FCHK(hipMemcpyAsync(qx.p, hx, sizeof(double) * qx.size(), k, H->stream));
  FCHK(hipMemcpyAsync(qy.p, hy, sizeof(double) * qy.size(), k, H->stream));
  FCHK(hipMemcpyAsync(qz.p, hz, sizeof(double) * qz.size(), k, H->stream));
  FCHK(hipMemcpyAsync(pr.p, hp, sizeof(double) * pr.size(), k, H->stream));
  FCHK(hipMemsetAsync(pcor.p, 0, sizeof(double) * pcor.size(), H->stream));
  if (thermo) SCHK(convert_qg(true));
  FCHK(hipStreamSynchronize(H->stream));
  return "";
}

std::string Flow::get_state(double* hx, double* hy, double* hz, double* hp, int mem) {
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  FCHK(hipMemcpyAsync(hx, qx.p, sizeof(double) * qx.size(), k, H->stream));
  FCHK(hipMemcpyAsync(hy, qy.p, sizeof(double) * qy.size(), k, H->stream));
  FCHK(hipMemcpyAsync(hz, qz.p, sizeof(double) * qz.size(), k, H->stream));
  FCHK(hipMemcpyAsync(hp, pr.p, sizeof(double) * pr.size(), k, H->stream));
  FCHK(hipStreamSynchronize(H->stream));
  return "";
}

std::string Flow::copy_rhs(double* mx, double* my, double* mz, int mem) {
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  const bool forced = (H->desc.idriven == 1);
  double* out[3] = {mx, my, mz};
  double* tmp = w64.scratch;
  for (int c = 0; c < 3; ++c) {
    const long n = w64.acc[c][0].size();
    // the x right-hand side after the constant-flux correction (:1547-1551)
    const double* b = (c == 0 && forced) ? bulk : nullptr;
    if (mixed)
      hipLaunchKernelGGL(k_sub_bulk<float>, grid1(n), dim3(256), 0, H->stream, tmp, w32.acc[c][0].p, b, n);
    else
      hipLaunchKernelGGL(k_sub_bulk<double>, grid1(n), dim3(256), 0, H->stream, tmp, w64.acc[c][0].p, b, n);
    FCHK(hipMemcpyAsync(out[c], tmp, sizeof(double) * n, k, H->stream));
    FCHK(hipStreamSynchronize(H->stream));
  }
  return "";
}

// The convective and viscous contributions of one momentum component along
// one direction (src/eq_momentum2.f90:798-1507 applies them as two
// operators).  When both use the same boundary kinds and neither has wall
// values, they share one line plan and, the operator being linear, are
// applied as one operator on the summed input
//   -ca * cb + r ts ta + r ts tc tb.
template <typename T>
std::string Flow::cov_tau(RhsWork<T>& w, int dir, int opk, const int bc_cov[2], const int bc_tau[2], Field<T>& out,
                          const T* fbc_cov, const T* fbc_tau, const Field<T>& ca, const Field<T>& cb,
                          const Field<T>& ta, T ts, const Field<T>* tb, T tc, int lane) {
  const T r = T(rre), one = T(1), m1 = T(-1);
  if (!fbc_cov && !fbc_tau && bc_cov[0] == bc_tau[0] && bc_cov[1] == bc_tau[1])
    return op<T>(w, dir, opk, bc_cov, ca, out, nullptr, m1, nullptr, one, &cb, one, 0, &ta, r * ts, tb,
                 tb ? r * ts * tc : T(0), lane);
  SCHK(op<T>(w, dir, opk, bc_cov, ca, out, fbc_cov, m1, nullptr, one, &cb, one, 0, nullptr, T(0), nullptr, T(0),
             lane));
  return op<T>(w, dir, opk, bc_tau, ta, out, fbc_tau, ts, tb, tc, nullptr, r, 1, nullptr, T(0), nullptr, T(0), lane);
}

// Compute_momentum_rhs + Calculate_momentum_fractional_step + forcing.
// Result: dt-scaled rhs in w.acc[c][0]; bulk (if forced) in this->bulk.
template <typename T>
std::string Flow::momentum_rhs_t(RhsWork<T>& w, int isub, bool fuse) {
  hipStream_t s = H->stream;
  const int(*B)[5][2] = ibc;
  const T* fu = wall_y ? w.fby_u : nullptr;
  const T* fv = wall_y ? w.fby_v : nullptr;
  const T* fw = wall_y ? w.fby_w : nullptr;
  const T one = T(1), m1 = T(-1);
  // one operator of the plain form on a lane
  auto op1 = [&](int lane, int dir, int opk, const int* bc, const Field<T>& A, Field<T>& out, const T* fbc,
                 T sA = T(1)) {
    return op<T>(w, dir, opk, bc, A, out, fbc, sA, nullptr, T(1), nullptr, T(1), 0, nullptr, T(0), nullptr, T(0),
                 lane);
  };
  // ---- first-level operators, src/eq_momentum2.f90:347-506 (one lane per
  // velocity component; the interpolation and the derivative of a component
  // along one direction share one pass over it) ----
  SCHK(fork());
  SCHK(op2<T>(w, 0, 0, w.qx, MIDP_P2C, B[0][VU], w.qxix, nullptr, D1_P2C, B[0][VU], w.qxdx, nullptr));
  SCHK(op2<T>(w, 0, 1, w.qx, D1_C2P, B[1][VU], w.qxdy, fu, MIDP_C2P, B[1][VU], w.qxiy, fu));
  SCHK(op2<T>(w, 0, 2, w.qx, MIDP_C2P, B[2][VU], w.qxiz, nullptr, D1_C2P, B[2][VU], w.qxdz, nullptr));
  SCHK(op2<T>(w, 1, 0, w.qy, D1_C2P, B[0][VV], w.qydx, nullptr, MIDP_C2P, B[0][VV], w.qyix, nullptr));
  SCHK(op2<T>(w, 1, 1, w.qy, D1_P2C, B[1][VV], w.qydy, nullptr, MIDP_P2C, B[1][VV], w.qyiy, fv));
  SCHK(op2<T>(w, 1, 2, w.qy, D1_C2P, B[2][VV], w.qydz, nullptr, MIDP_C2P, B[2][VV], w.qyiz, nullptr));
  SCHK(op2<T>(w, 2, 0, w.qz, D1_C2P, B[0][VW], w.qzdx, nullptr, MIDP_C2P, B[0][VW], w.qzix, nullptr));
  SCHK(op2<T>(w, 2, 1, w.qz, MIDP_C2P, B[1][VW], w.qziy, fw, D1_C2P, B[1][VW], w.qzdy, fw));
  SCHK(op2<T>(w, 2, 2, w.qz, MIDP_P2C, B[2][VW], w.qziz, nullptr, D1_P2C, B[2][VW], w.qzdz, nullptr));
  if (thermo) SCHK(thermo_first_level<T>(w));
  SCHK(join());
  hipLaunchKernelGGL(k_div3<T>, grid1(w.div.size()), dim3(256), 0, s, w.div.p, w.qxdx.p, w.qydy.p, w.qzdz.p,
                     w.div.size());
  // ---- y-wall values of the divergence and of the y-normal stress ----
  if (wall_y) {
    const long nf2 = (long)nx * nz;
    // qxdx at the walls: x-derivative of the wall planes of qxiy
    hipLaunchKernelGGL(k_extract_walls<T>, grid1((long)px * nz), dim3(256), 0, s, w.plane_a, w.qxiy.p, px, py, nz);
    Field<T> pa{w.plane_a, Ext3{{px, 2, nz}}}, pb{w.plane_b, Ext3{{nx, 2, nz}}};
    SCHK(op<T>(w, 0, D1_P2C, B[0][VU], pa, pb, nullptr));
    // qydy at the walls: y-derivative C2P of qyiy (Dirichlet rows are decoupled)
    {
      const DevPlan<T>* P = nullptr;
      SCHK(H->plan<T>(1, D1_C2P, H->desc.iacc, B[1][VV], true, ny, py, &P));
      SCHK(launch_wall_rows<T>(*P, w.qyiy.e, w.qyiy.p, w.fby_v, w.plane_c, s));
    }
    // qzdz at the walls: z-derivative of the wall planes of qziy
    hipLaunchKernelGGL(k_extract_walls<T>, grid1((long)nx * pz), dim3(256), 0, s, w.plane_a, w.qziy.p, nx, py, pz);
    Field<T> pa2{w.plane_a, Ext3{{nx, 2, pz}}}, pd{w.plane_d, Ext3{{nx, 2, nz}}};
    SCHK(op<T>(w, 2, D1_P2C, B[2][VW], pa2, pd, nullptr));
    hipLaunchKernelGGL(k_wall_fbc<T>, grid1(nf2), dim3(256), 0, s, w.fby_div, w.fby_visc, w.plane_b, w.plane_c,
                       w.plane_d, nx, nz, (const T*)(thermo ? w.fby_mu : nullptr));
  }
  const T third = T(-ONE_THIRD), two = T(2);
  const T* fconv = wall_y ? w.fby_conv : nullptr;
  const T* fvisc = wall_y ? w.fby_visc : nullptr;
  // ---- second-level operators, one lane per momentum component; each
  // lane ends with its fractional-step combine ----
  SCHK(fork());
  Field<T>* X = w.acc[0];
  Field<T>* Y = w.acc[1];
  Field<T>* Z = w.acc[2];
  if (!thermo) {
  // x-momentum, :798-948
  SCHK(cov_tau<T>(w, 0, D1_C2P, mcov[0][0], mtau[0][0], X[0], nullptr, nullptr, w.qxix, w.qxix, w.qxdx, two,
                  &w.div, third, 0));
  SCHK(cov_tau<T>(w, 1, D1_P2C, mcov[0][1], mtau[0][1], X[1], nullptr, nullptr, w.qyix, w.qxiy, w.qxdy, one,
                  &w.qydx, one, 0));
  SCHK(cov_tau<T>(w, 2, D1_P2C, mcov[0][2], mtau[0][2], X[2], nullptr, nullptr, w.qzix, w.qxiz, w.qxdz, one,
                  &w.qzdx, one, 0));
  // y-momentum, :960-1235
  SCHK(cov_tau<T>(w, 0, D1_P2C, mcov[1][0], mtau[1][0], Y[0], nullptr, nullptr, w.qxiy, w.qyix, w.qxdy, one,
                  &w.qydx, one, 1));
  SCHK(cov_tau<T>(w, 1, D1_C2P, mcov[1][1], mtau[1][1], Y[1], fconv, fvisc, w.qyiy, w.qyiy, w.qydy, two, &w.div,
                  third, 1));
  SCHK(cov_tau<T>(w, 2, D1_P2C, mcov[1][2], mtau[1][2], Y[2], nullptr, nullptr, w.qyiz, w.qziy, w.qzdy, one,
                  &w.qydz, one, 1));
  // z-momentum, :1248-1480
  SCHK(cov_tau<T>(w, 0, D1_P2C, mcov[2][0], mtau[2][0], Z[0], nullptr, nullptr, w.qxiz, w.qzix, w.qxdz, one,
                  &w.qzdx, one, 2));
  SCHK(cov_tau<T>(w, 1, D1_P2C, mcov[2][1], mtau[2][1], Z[1], nullptr, nullptr, w.qyiz, w.qziy, w.qzdy, one,
                  &w.qydz, one, 2));
  SCHK(cov_tau<T>(w, 2, D1_C2P, mcov[2][2], mtau[2][2], Z[2], nullptr, nullptr, w.qziz, w.qziz, w.qzdz, two,
                  &w.div, third, 2));
  } else {
    // variable properties: the convective product takes interpolated g as
    // its first factor and every viscous operand is multiplied by the
    // interpolated viscosity (the is_thermo branches of :805-1480), so the
    // convective and viscous operators stay two launches each
    const T r = T(rre);
    auto ct = [&](int lane, int dir, int opk, const int* bcc, const int* bct, Field<T>& out, const T* fbcc,
                  const T* fbct, const Field<T>& ca, const Field<T>& cb, const Field<T>& ta, T ts, const Field<T>* tb,
                  T tc, const Field<T>& mu) -> std::string {
      SCHK(op<T>(w, dir, opk, bcc, ca, out, fbcc, m1, nullptr, one, &cb, one, 0, nullptr, T(0), nullptr, T(0), lane));
      return op<T>(w, dir, opk, bct, ta, out, fbct, ts, tb, tc, &mu, r, 1, nullptr, T(0), nullptr, T(0), lane);
    };
    SCHK(ct(0, 0, D1_C2P, mcov[0][0], mtau[0][0], X[0], nullptr, nullptr, w.gxix, w.qxix, w.qxdx, two, &w.div, third,
            w.visc));
    SCHK(ct(0, 1, D1_P2C, mcov[0][1], mtau[0][1], X[1], nullptr, nullptr, w.gyix, w.qxiy, w.qxdy, one, &w.qydx, one,
            w.muixy));
    SCHK(ct(0, 2, D1_P2C, mcov[0][2], mtau[0][2], X[2], nullptr, nullptr, w.gzix, w.qxiz, w.qxdz, one, &w.qzdx, one,
            w.muixz));
    SCHK(ct(1, 0, D1_P2C, mcov[1][0], mtau[1][0], Y[0], nullptr, nullptr, w.gxiy, w.qyix, w.qxdy, one, &w.qydx, one,
            w.muixy));
    SCHK(ct(1, 1, D1_C2P, mcov[1][1], mtau[1][1], Y[1], fconv, fvisc, w.qyiy, w.gyiy, w.qydy, two, &w.div, third,
            w.visc));
    SCHK(ct(1, 2, D1_P2C, mcov[1][2], mtau[1][2], Y[2], nullptr, nullptr, w.qyiz, w.gziy, w.qzdy, one, &w.qydz, one,
            w.muiyz));
    SCHK(ct(2, 0, D1_P2C, mcov[2][0], mtau[2][0], Z[0], nullptr, nullptr, w.gxiz, w.qzix, w.qxdz, one, &w.qzdx, one,
            w.muixz));
    SCHK(ct(2, 1, D1_P2C, mcov[2][1], mtau[2][1], Z[1], nullptr, nullptr, w.gyiz, w.qziy, w.qzdy, one, &w.qydz, one,
            w.muiyz));
    SCHK(ct(2, 2, D1_C2P, mcov[2][2], mtau[2][2], Z[2], nullptr, nullptr, w.gziz, w.qziz, w.qzdz, two, &w.div, third,
            w.visc));
  }
  SCHK(op1(0, 0, D1_C2P, B[0][VP], w.pr, X[3], nullptr, m1));
  SCHK(op1(1, 1, D1_C2P, B[1][VP], w.pr, Y[3], wall_y ? w.fby_pn : nullptr, m1));
  SCHK(op1(2, 2, D1_C2P, B[2][VP], w.pr, Z[3], nullptr, m1));
  // fractional step (and the forcing's block partials of component x)
  const Mesh& m = H->mesh;
  const bool forced = (H->desc.idriven == 1);
  // with fuse, components whose update needs no bulk correction are
  // updated here (x is, when the flow is forced: the correction needs the
  // whole volume's sum first)
  Field<T>* rhs0[3] = {&w.mx0, &w.my0, &w.mz0};
  // the updated state: q, or g = rho u with variable properties
  double* qs[3] = {thermo ? w64.gx.p : qx.p, thermo ? w64.gy.p : qy.p, thermo ? w64.gz.p : qz.p};
  for (int c = 0; c < 3; ++c) {
    long n = w.acc[c][0].size();
    double* part = (c == 0 && forced) ? partial : nullptr;
    double* qf = (fuse && !(c == 0 && forced)) ? qs[c] : nullptr;
    const double* fbw = (qf && c == 1 && wall_y) ? w64.fby_v : nullptr;
    hipLaunchKernelGGL(k_combine<T>, grid1(n), dim3(256), 0, lane_stream(c), w.acc[c][0].p, w.acc[c][1].p,
                       w.acc[c][2].p, w.acc[c][3].p, rhs0[c]->p, n, T(tg[isub]), T(tz[isub]), T(ta[isub]), T(dt),
                       w.acc[c][0].e.n[0], w.acc[c][0].e.n[1], dyw_dev, m.h[0], m.h[2], part, qf, fbw);
  }
  if (forced) {
    int nb = (int)((w.acc[0][0].size() + 255) / 256);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, lane_stream(0), partial, nb, vol, bulk);
    if (fuse)
      hipLaunchKernelGGL(k_update<T>, grid1(qx.size()), dim3(256), 0, lane_stream(0), qs[0], w.acc[0][0].p,
                         (const double*)bulk, qx.size(), px, ny, (const double*)nullptr);
  }
  SCHK(join());
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("momentum_rhs: ") + hipGetErrorString(e);
  return "";
}

std::string Flow::refresh_q32() {
  if (!mixed) return "";
  hipStream_t s = H->stream;
  hipLaunchKernelGGL((k_cvt<float, double>), grid1(qx.size()), dim3(256), 0, s, w32.qx.p, qx.p, qx.size());
  hipLaunchKernelGGL((k_cvt<float, double>), grid1(qy.size()), dim3(256), 0, s, w32.qy.p, qy.p, qy.size());
  hipLaunchKernelGGL((k_cvt<float, double>), grid1(qz.size()), dim3(256), 0, s, w32.qz.p, qz.p, qz.size());
  hipLaunchKernelGGL((k_cvt<float, double>), grid1(pr.size()), dim3(256), 0, s, w32.pr.p, pr.p, pr.size());
  q32_fresh = true;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("refresh_q32: ") + hipGetErrorString(e);
  return "";
}

std::string Flow::momentum_rhs(int isub, bool fuse) {
  set_lanes();
  if (!mixed) return momentum_rhs_t<double>(w64, isub, fuse);
  // rounded copies of the state for the FP32 right-hand side, unless the
  // last sub-step left them
  if (!q32_fresh) SCHK(refresh_q32());
  if (thermo) {
    // g and mu change every sub-step (projection, property refresh); the
    // energy sub-step before this right-hand side has rounded g already
    hipStream_t s = H->stream;
    for (auto pr2 : {std::make_pair(&w32.gx, &w64.gx), std::make_pair(&w32.gy, &w64.gy),
                     std::make_pair(&w32.gz, &w64.gz), std::make_pair(&w32.visc, &w64.visc)})
      if (pr2.first == &w32.visc ? !props32_fresh : !g32_fresh)
        hipLaunchKernelGGL((k_cvt<float, double>), grid1(pr2.second->size()), dim3(256), 0, s, pr2.first->p,
                           pr2.second->p, pr2.second->size());
    g32_fresh = false;
  }
  const std::string r = momentum_rhs_t<float>(w32, isub, fuse);
  if (fuse) q32_fresh = false;  // the state was updated
  return r;
}

std::string Flow::divergence(double* d) {
  // of q, or of g = rho u with variable properties (src/eq_continuity.f90:114-117)
  Field<double> D{d, Ext3{{nx, ny, nz}}};
  const Field<double>& ux = thermo ? w64.gx : qx;
  const Field<double>& uy = thermo ? w64.gy : qy;
  const Field<double>& uz = thermo ? w64.gz : qz;
  SCHK(op<double>(w64, 0, D1_P2C, ibc[0][VU], ux, D, nullptr));
  SCHK(op<double>(w64, 1, D1_P2C, ibc[1][VV], uy, D, nullptr, 1.0, nullptr, 1.0, nullptr, 1.0, 1));
  SCHK(op<double>(w64, 2, D1_P2C, ibc[2][VW], uz, D, nullptr, 1.0, nullptr, 1.0, nullptr, 1.0, 1));
  return "";
}

std::string Flow::substep(int isub) {
  hipStream_t s = H->stream;
  // variable properties: the energy equation first (src/chapsim.f90:300-304;
  // the initial sub-step is momentum only, :135)
  if (thermo && isub > 0) SCHK(energy_substep(isub));
  SCHK(momentum_rhs(isub, true));
  // Poisson RHS: div(q) / (alpha * sigma2 * dt), then the solve
  SCHK(divergence(pcor.p));
  const double coeff = 1.0 / (ta[isub] * sigma2p * dt);
  if (H->desc.projection == CHS_PROJ_EXACT) SCHK(H->poisson->solve_exact(pcor.p, s, coeff));
  else SCHK(H->poisson->solve(pcor.p, s, coeff));
  // the mixed form's FP32 copies of the new pressure and velocity are
  // stored by the kernels that make them
  hipLaunchKernelGGL(k_add, grid1(pr.size()), dim3(256), 0, s, pr.p, pcor.p, pr.size(), mixed ? w32.pr.p : nullptr);
  // Correct_massflux, one lane per velocity component (on g with variable
  // properties, :1609-1613, then q = g / rho, :1915)
  const double sc = dt * ta[isub] * sigma2p;
  Field<double>& ux = thermo ? w64.gx : qx;
  Field<double>& uy = thermo ? w64.gy : qy;
  Field<double>& uz = thermo ? w64.gz : qz;
  const bool st32 = mixed && !thermo;
  SCHK(fork());
  SCHK(op<double>(w64, 0, D1_C2P, ibc[0][VP], pcor, ux, nullptr, 1.0, nullptr, 1.0, nullptr, -sc, 1, nullptr, 0.0,
                  nullptr, 0.0, 0, st32 ? w32.qx.p : nullptr));
  SCHK(op<double>(w64, 1, D1_C2P, ibc[1][VP], pcor, uy, wall_y ? w64.fby_p : nullptr, 1.0, nullptr, 1.0, nullptr,
                  -sc, 1, nullptr, 0.0, nullptr, 0.0, 1, st32 ? w32.qy.p : nullptr));
  SCHK(op<double>(w64, 2, D1_C2P, ibc[2][VP], pcor, uz, nullptr, 1.0, nullptr, 1.0, nullptr, -sc, 1, nullptr, 0.0,
                  nullptr, 0.0, 2, st32 ? w32.qz.p : nullptr));
  if (wall_y)
    hipLaunchKernelGGL(k_enforce_walls, grid1((long)nx * nz), dim3(256), 0, lane_stream(1), uy.p, nx, py, nz,
                       w64.fby_v, st32 ? w32.qy.p : nullptr);
  SCHK(join());
  if (thermo) SCHK(convert_qg(false));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("substep: ") + hipGetErrorString(e);
  q32_fresh = mixed;
  // every operator of a sub-step is planned and tuned now: later sub-steps
  // may overlap independent operators
  concurrent = (side[0] != nullptr);
  return "";
}

std::string Flow::cell_velocity() {
  // the P2C interpolations of the first-level operators (their work arrays
  // are free between sub-steps)
  SCHK(op<double>(w64, 0, MIDP_P2C, ibc[0][VU], qx, w64.qxix, nullptr));
  SCHK(op<double>(w64, 1, MIDP_P2C, ibc[1][VV], qy, w64.qyiy, wall_y ? w64.fby_v : nullptr));
  return op<double>(w64, 2, MIDP_P2C, ibc[2][VW], qz, w64.qziz, nullptr);
}

std::string Flow::bulk_monitor(double out[2]) {
  SCHK(cell_velocity());
  const int nb = kMaxPartials;
  std::vector<double> h(nb);
  const double* src[2] = {w64.qxix.p, qx.p};
  const int n1[2] = {nx, px};
  for (int q = 0; q < 2; ++q) {
    const long n = (long)n1[q] * ny * nz;
    hipLaunchKernelGGL(k_vol_sum, dim3(nb), dim3(256), 0, H->stream, src[q], q == 0 ? w64.qyiy.p : nullptr,
                       q == 0 ? w64.qziz.p : nullptr, n1[q], ny, n, dyw_dev, partial);
    FCHK(hipMemcpyAsync(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost, H->stream));
    FCHK(hipStreamSynchronize(H->stream));
    double sum = 0.0;
    for (double v : h) sum += v;
    // volume of the array's cells: n1 nz dx dz sum_j dy_j
    double sy = 0.0;
    for (int j = 0; j < ny; ++j) sy += dyw[j];
    out[q] = sum / ((double)n1[q] * nz * sy);
  }
  return "";
}

std::string Flow::velocity_range(double out[6], long* nonfinite) {
  const int nb = kMaxPartials;
  std::vector<double> h(3 * (size_t)nb);
  const Field<double>* q[3] = {&qx, &qy, &qz};
  long bad = 0;
  for (int c = 0; c < 3; ++c) {
    hipLaunchKernelGGL(k_range, dim3(nb), dim3(256), 0, H->stream, q[c]->p, q[c]->size(), partial);
    FCHK(hipMemcpyAsync(h.data(), partial, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, H->stream));
    FCHK(hipStreamSynchronize(H->stream));
    double mx = -1.0e20, mn = 1.0e20;
    for (int b = 0; b < nb; ++b) {
      if (h[b] > mx) mx = h[b];
      if (h[nb + b] < mn) mn = h[nb + b];
      bad += (long)h[2 * nb + b];
    }
    out[2 * c] = mx;
    out[2 * c + 1] = mn;
  }
  *nonfinite = bad;
  return "";
}

std::string Flow::velocity_range_noalloc(double out[6], long* nonfinite) {
  if (!range_scratch) return "velocity range: static scratch not set up";
  constexpr int nb = kMaxPartials;
  double h[3 * nb];
  const Field<double>* q[3] = {&qx, &qy, &qz};
  long bad = 0;
  for (int c = 0; c < 3; ++c) {
    // every block writes its three slots, so nothing is cleared first
    hipLaunchKernelGGL(k_range, dim3(nb), dim3(256), 0, H->stream, q[c]->p, q[c]->size(), range_scratch);
    FCHK(hipMemcpyAsync(h, range_scratch, sizeof(h), hipMemcpyDeviceToHost, H->stream));
    FCHK(hipStreamSynchronize(H->stream));
    double mx = -1.0e20, mn = 1.0e20;
    for (int b = 0; b < nb; ++b) {
      if (h[b] > mx) mx = h[b];
      if (h[nb + b] < mn) mn = h[nb + b];
      bad += (long)h[2 * nb + b];
    }
    out[2 * c] = mx;
    out[2 * c + 1] = mn;
  }
  *nonfinite = bad;
  return "";
}

std::string Flow::cfl(double out[2]) {
  const Mesh& m = H->mesh;
  const bool stretched = m.stretched;
  if (stretched && !ymc_dev) {
    FCHK(hipMalloc(&ymc_dev, sizeof(double) * (size_t)ny));
    owned.push_back(ymc_dev);
    FCHK(hipMemcpy(ymc_dev, m.ymc.data(), sizeof(double) * (size_t)ny, hipMemcpyHostToDevice));
  }
  SCHK(cell_velocity());
  const int nb = kMaxPartials;
  hipLaunchKernelGGL(k_cfl_max, dim3(nb), dim3(256), 0, H->stream, w64.qxix.p, w64.qyiy.p, w64.qziz.p, nx, ny,
                     w64.qxix.size(), m.h1r[0], m.h1r[1], m.h1r[2], dt, stretched ? ymc_dev : nullptr, partial);
  std::vector<double> h(nb);
  FCHK(hipMemcpyAsync(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost, H->stream));
  FCHK(hipStreamSynchronize(H->stream));
  double mx = 0.0;
  for (double v : h) mx = (v > mx || v != v) ? v : mx;
  out[0] = mx;
  // diffusion number, Cartesian constant properties (:220-259)
  double d = 0.0;
  for (int j = 0; j < ny; ++j) {
    const double dyi = 1.0 / (m.yp[j + 1] - m.yp[j]);
    const double r = m.h2r[0] + m.h2r[2] + dyi * dyi;
    if (r > d) d = r;
  }
  out[1] = d * 2.0 * dt * rre;
  return "";
}

std::string Flow::mass_check(double out[3]) {
  SCHK(divergence(div.p));
  const int lo[3] = {0, 4, nx - 6}, hi[3] = {3, nx - 5, nx - 1};
  const int nb = kMaxPartials;
  std::vector<double> h(nb);
  for (int w = 0; w < 3; ++w) {
    hipLaunchKernelGGL(k_maxabs_xwin, dim3(nb), dim3(256), 0, H->stream, div.p, nx, div.size(), lo[w], hi[w],
                       partial);
    FCHK(hipMemcpyAsync(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost, H->stream));
    FCHK(hipStreamSynchronize(H->stream));
    double mx = 0.0;
    for (double v : h) mx = (v > mx || v != v) ? v : mx;
    out[w] = mx;
  }
  return "";
}

// ---------------------------------------------------------------------------
// Variable properties (heated channel with isothermal, stationary y walls)
// ---------------------------------------------------------------------------
namespace {
// table branch of ftp_refresh_thermal_properties_from_T_undim
// (src/input_thermo.f90:213-241): out = h, d, m, k, t, b, cp
void props_from_t_host(const chs_props* p, double t, double out[7]) {
  const std::vector<double>& c = p->col[CHS_PC_T];
  int i1 = 0, i2 = p->n - 1;
  while (i2 - i1 > 1) {
    const int im = i1 + (i2 - i1) / 2;
    const double d1 = c[i1] - t, dm = c[im] - t;
    if (d1 * dm > 1.0e-20) i1 = im;
    else i2 = im;
  }
  const double w1 = (c[i2] - t) / (c[i2] - c[i1]), w2 = 1.0 - w1;
  auto in = [&](int col) {
    const double a = w1 * p->col[col][i1];
    const double b = w2 * p->col[col][i2];
    return a + b;
  };
  out[1] = in(CHS_PC_D);
  out[2] = in(CHS_PC_M);
  out[3] = in(CHS_PC_K);
  out[0] = in(CHS_PC_H);
  out[5] = in(CHS_PC_B);
  out[6] = in(CHS_PC_CP);
  out[4] = t;
}
}  // namespace

template <typename T>
std::string Flow::alloc_thermo_work(RhsWork<T>& w) {
  SCHK(alloc(w.gx, px, ny, nz)); SCHK(alloc(w.gy, nx, py, nz)); SCHK(alloc(w.gz, nx, ny, pz));
  SCHK(alloc(w.visc, nx, ny, nz));
  SCHK(alloc(w.gxix, nx, ny, nz)); SCHK(alloc(w.gxiy, px, py, nz)); SCHK(alloc(w.gxiz, px, ny, pz));
  SCHK(alloc(w.gyix, px, py, nz)); SCHK(alloc(w.gyiy, nx, ny, nz)); SCHK(alloc(w.gyiz, nx, py, pz));
  SCHK(alloc(w.gzix, px, ny, pz)); SCHK(alloc(w.gziy, nx, py, pz)); SCHK(alloc(w.gziz, nx, ny, nz));
  SCHK(alloc(w.muiy, nx, py, nz)); SCHK(alloc(w.mupcc, px, ny, nz)); SCHK(alloc(w.muixy, px, py, nz));
  SCHK(alloc(w.muixz, px, ny, pz)); SCHK(alloc(w.muiyz, nx, py, pz));
  const long nf = (long)nx * 4 * nz;
  std::vector<T> hm(nf);
  for (long t = 0; t < nf; ++t) hm[t] = T(wall_prop[(t / nx) % 2][2]);
  for (T** q : {&w.fby_mu, &w.fby_mwall}) {
    FCHK(hipMalloc(q, sizeof(T) * nf));
    owned.push_back(*q);
  }
  FCHK(hipMemcpy(w.fby_mwall, hm.data(), sizeof(T) * nf, hipMemcpyHostToDevice));
  return "";
}

// wall rows (nx,4,nz) of wall property col (h, d, m, k, t, b, cp): rows 1,2
// the two walls, rows 3,4 the same again (rows34) or zero
template <typename T>
static std::string upload_wall_rows(T** q, int nx, int nz, const double wp[2][7], int col, bool rows34,
                                    std::vector<void*>& owned) {
  const long nf = (long)nx * 4 * nz;
  std::vector<T> hb(nf);
  for (long t = 0; t < nf; ++t) {
    const long r = (t / nx) % 4;
    hb[t] = (!rows34 && r >= 2) ? T(0) : T(wp[r % 2][col]);
  }
  FCHK(hipMalloc(q, sizeof(T) * nf));
  owned.push_back(*q);
  FCHK(hipMemcpy(*q, hb.data(), sizeof(T) * nf, hipMemcpyHostToDevice));
  return "";
}

// The energy right-hand side's work in precision T; own_centres allocates
// h, k, T at the cell centres (the FP32 copies), else the caller sets them.
// Wall rows: h and k of the walls in rows 1,2 and again in rows 3,4; T of
// the walls in rows 1,2 and zero in rows 3,4 (get_fbcy_iTh,
// src/bc_general.f90:784-805); the wall flux is scratch.
template <typename T>
std::string Flow::alloc_energy_work(EnergyWork<T>& e, bool own_centres) {
  if (own_centres)
    for (Field<T>* f : {&e.hent, &e.kcond, &e.temp}) SCHK(alloc(*f, nx, ny, nz));
  for (Field<T>* f : {&e.ex, &e.ey, &e.ez}) SCHK(alloc(*f, nx, ny, nz));
  SCHK(alloc(e.hpcc, px, ny, nz)); SCHK(alloc(e.hcpc, nx, py, nz)); SCHK(alloc(e.hccp, nx, ny, pz));
  SCHK(alloc(e.kpcc, px, ny, nz)); SCHK(alloc(e.kcpc, nx, py, nz)); SCHK(alloc(e.kccp, nx, ny, pz));
  SCHK(alloc(e.tpcc, px, ny, nz)); SCHK(alloc(e.tcpc, nx, py, nz)); SCHK(alloc(e.tccp, nx, ny, pz));
  SCHK(upload_wall_rows(&e.fby_h, nx, nz, wall_prop, 0, true, owned));
  SCHK(upload_wall_rows(&e.fby_k, nx, nz, wall_prop, 3, true, owned));
  SCHK(upload_wall_rows(&e.fby_t, nx, nz, wall_prop, 4, false, owned));
  FCHK(hipMalloc(&e.fby_q, sizeof(T) * (size_t)nx * 4 * nz));
  owned.push_back(e.fby_q);
  return "";
}

// interpolated g of the convective products and the viscosity interpolations
// (src/eq_momentum2.f90:576-703), issued on the three lanes of the
// first-level operators
template <typename T>
std::string Flow::thermo_first_level(RhsWork<T>& w) {
  const int(*B)[5][2] = ibc;
  int ftp[3][2];
  for (int d = 0; d < 3; ++d)
    for (int s2 = 0; s2 < 2; ++s2) ftp[d][s2] = (ibc[d][VT][s2] == BC_NEUMANN) ? BC_DIRICHLET : ibc[d][VT][s2];
  const T* fu = wall_y ? w.fby_u : nullptr;
  const T* fv = wall_y ? w.fby_v : nullptr;
  const T* fw = wall_y ? w.fby_w : nullptr;
  auto ip = [&](int lane, int dir, int opk, const int* bc, const Field<T>& A, Field<T>& out, const T* fbc) {
    return op<T>(w, dir, opk, bc, A, out, fbc, T(1), nullptr, T(1), nullptr, T(1), 0, nullptr, T(0), nullptr, T(0),
                 lane);
  };
  SCHK(ip(0, 0, MIDP_P2C, B[0][VU], w.gx, w.gxix, nullptr));
  SCHK(ip(0, 1, MIDP_C2P, B[1][VU], w.gx, w.gxiy, fu));
  SCHK(ip(0, 2, MIDP_C2P, B[2][VU], w.gx, w.gxiz, nullptr));
  SCHK(ip(1, 0, MIDP_C2P, B[0][VV], w.gy, w.gyix, nullptr));
  SCHK(ip(1, 1, MIDP_P2C, B[1][VV], w.gy, w.gyiy, fv));
  SCHK(ip(1, 2, MIDP_C2P, B[2][VV], w.gy, w.gyiz, nullptr));
  SCHK(ip(2, 0, MIDP_C2P, B[0][VW], w.gz, w.gzix, nullptr));
  SCHK(ip(2, 1, MIDP_C2P, B[1][VW], w.gz, w.gziy, fw));
  SCHK(ip(2, 2, MIDP_P2C, B[2][VW], w.gz, w.gziz, nullptr));
  // muiy_cpc (y, wall viscosity), muixy_ppc = x of it, muiyz_cpp = z of it,
  // its wall rows as fbcy_mu_c4c; mu_pcc and muixz_pcp = z of mu_pcc
  SCHK(ip(1, 1, MIDP_C2P, ftp[1], w.visc, w.muiy, wall_y ? w.fby_mwall : nullptr));
  SCHK(ip(1, 0, MIDP_C2P, ftp[0], w.muiy, w.muixy, nullptr));
  SCHK(ip(1, 2, MIDP_C2P, ftp[2], w.muiy, w.muiyz, nullptr));
  if (wall_y)
    hipLaunchKernelGGL(k_wall_rows4<T>, grid1((long)nx * nz), dim3(256), 0, lane_stream(1), w.fby_mu, w.muiy.p,
                       (const T*)nullptr, nx, py, nz);
  SCHK(ip(2, 0, MIDP_C2P, ftp[0], w.visc, w.mupcc, nullptr));
  SCHK(ip(2, 2, MIDP_C2P, ftp[2], w.mupcc, w.muixz, nullptr));
  return "";
}

std::string Flow::thermo_enable(const chs_thermo_desc_t& d) {
  if (thermo) return "variable properties: already enabled";
  if (!wall_y) return "variable properties: need y walls";
  const chs_desc_t& hd = H->desc;
  for (int s2 = 0; s2 < 2; ++s2)
    for (int v = 0; v < 3; ++v)
      if (hd.uwall[s2][v] != 0.0) return "variable properties: moving walls are not supported";
  if (!d.props) return "variable properties: no property table";
  if (d.mode != CHS_PROPS_TABLE && d.mode != CHS_PROPS_MLP) return "variable properties: unknown mode";
  if (d.mode == CHS_PROPS_MLP && !d.props->surrogate) return "variable properties: no surrogate attached";
  if (!(d.prandtl > 0.0)) return "variable properties: the Prandtl number must be positive";
  props = d.props;
  props_mode = d.mode;
  prandtl = d.prandtl;
  for (int s2 = 0; s2 < 2; ++s2) props_from_t_host(props, d.t_wall[s2], wall_prop[s2]);
  double p0[7];
  props_from_t_host(props, d.t_init, p0);
  // FP64 fields
  for (Field<double>* f : {&rhoh, &hent, &temp, &kcond, &dens, &eres, &erhs0}) SCHK(alloc(*f, nx, ny, nz));
  SCHK(alloc(dface[0], px, ny, nz)); SCHK(alloc(dface[1], nx, py, nz)); SCHK(alloc(dface[2], nx, ny, pz));
  SCHK(alloc_thermo_work<double>(w64));
  if (mixed) SCHK(alloc_thermo_work<float>(w32));
  e64.hent = hent; e64.kcond = kcond; e64.temp = temp;
  SCHK(alloc_energy_work(e64, false));
  if (mixed) SCHK(alloc_energy_work(e32, true));
  // the density's wall rows (nx,4,nz; rows 3,4 repeat rows 1,2)
  SCHK(upload_wall_rows(&fby_d, nx, nz, wall_prop, 1, true, owned));
  FCHK(hipMalloc(&props_flag, sizeof(int)));
  owned.push_back(props_flag);
  FCHK(hipMemset(props_flag, 0, sizeof(int)));
  // initialise_thermal_properties: uniform state from t_init
  hipStream_t s = H->stream;
  const long n = rhoh.size();
  hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, s, dens.p, p0[1], n);
  hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, s, w64.visc.p, p0[2], n);
  hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, s, rhoh.p, p0[1] * p0[0], n);
  hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, s, hent.p, p0[0], n);
  hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, s, kcond.p, p0[3], n);
  hipLaunchKernelGGL(k_fill<double>, grid1(n), dim3(256), 0, s, temp.p, p0[4], n);
  props32_fresh = false;  // the fills left the FP32 roundings unset
  thermo = true;
  update_re(0);
  SCHK(convert_qg(true));
  FCHK(hipStreamSynchronize(s));
  return "";
}

std::string Flow::refresh_props() {
  // Update_thermal_properties, src/eq_energy.f90:14-47 (dDensm1 = dDens
  // after the refresh, so Calculate_drhodt is zero and is not formed).  The
  // mixed form also gets the FP32 roundings of h, mu, k, T that its energy
  // and momentum right-hand sides read, stored by the refresh kernel itself
  // (table or surrogate).
  float* const r32[4] = {e32.hent.p, w32.visc.p, e32.kcond.p, e32.temp.p};
  SCHK(props_refresh_dh(props, props_mode, rhoh.p, rhoh.size(), hent.p, dens.p, w64.visc.p, kcond.p, temp.p,
                        props_flag, H->stream, mixed ? r32 : nullptr));
  props32_fresh = mixed;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("refresh_props: ") + hipGetErrorString(e);
  return "";
}

std::string Flow::thermo_set_rhoh(const double* r, int mem, long* clamped) {
  if (!thermo) return "variable properties: not enabled";
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  FCHK(hipMemcpyAsync(rhoh.p, r, sizeof(double) * rhoh.size(), k, H->stream));
  FCHK(hipMemsetAsync(props_flag, 0, sizeof(int), H->stream));
  SCHK(refresh_props());
  SCHK(convert_qg(true));
  int nb = 0;
  FCHK(hipMemcpyAsync(&nb, props_flag, sizeof(int), hipMemcpyDeviceToHost, H->stream));
  FCHK(hipStreamSynchronize(H->stream));
  *clamped = nb;
  return "";
}

std::string Flow::thermo_get(double* out[9], int mem) {
  if (!thermo) return "variable properties: not enabled";
  const Field<double>* src[9] = {&rhoh, &hent, &temp, &kcond, &dens, &w64.visc, &w64.gx, &w64.gy, &w64.gz};
  hipMemcpyKind k = (mem == CHS_MEM_HOST) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  for (int q = 0; q < 9; ++q)
    if (out[q]) FCHK(hipMemcpyAsync(out[q], src[q]->p, sizeof(double) * src[q]->size(), k, H->stream));
  FCHK(hipStreamSynchronize(H->stream));
  return "";
}

std::string Flow::convert_qg(bool to_g) {
  // convert_primary_conservative (src/para_conversion.f90:5-227): density at
  // each velocity location by midp C2P with the property BCs and the wall
  // density; stationary walls keep zero wall values of q and g
  int ftp[3][2];
  for (int d = 0; d < 3; ++d)
    for (int s2 = 0; s2 < 2; ++s2) ftp[d][s2] = (ibc[d][VT][s2] == BC_NEUMANN) ? BC_DIRICHLET : ibc[d][VT][s2];
  SCHK(op<double>(w64, 0, MIDP_C2P, ftp[0], dens, dface[0], nullptr));
  SCHK(op<double>(w64, 1, MIDP_C2P, ftp[1], dens, dface[1], wall_y ? fby_d : nullptr));
  SCHK(op<double>(w64, 2, MIDP_C2P, ftp[2], dens, dface[2], nullptr));
  Field<double>* q[3] = {&qx, &qy, &qz};
  Field<double>* g[3] = {&w64.gx, &w64.gy, &w64.gz};
  Field<float>* q32[3] = {&w32.qx, &w32.qy, &w32.qz};
  for (int c = 0; c < 3; ++c) {
    const long n = q[c]->size();
    if (to_g)
      hipLaunchKernelGGL(k_q2g, grid1(n), dim3(256), 0, H->stream, g[c]->p, q[c]->p, dface[c].p, n);
    else
      hipLaunchKernelGGL(k_g2q, grid1(n), dim3(256), 0, H->stream, q[c]->p, g[c]->p, dface[c].p, n,
                         mixed ? q32[c]->p : nullptr);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("convert_qg: ") + hipGetErrorString(e);
  return "";
}

std::string Flow::energy_combine(int isub, double* res, double* rhoh_acc) {
  const long n = rhoh.size();
  if (mixed)
    hipLaunchKernelGGL(k_energy_combine<float>, grid1(n), dim3(256), 0, H->stream, e32.ex.p, e32.ey.p, e32.ez.p,
                       erhs0.p, res, rhoh_acc, n, tg[isub], tz[isub], dt);
  else
    hipLaunchKernelGGL(k_energy_combine<double>, grid1(n), dim3(256), 0, H->stream, e64.ex.p, e64.ey.p, e64.ez.p,
                       erhs0.p, res, rhoh_acc, n, tg[isub], tz[isub], dt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("energy_combine: ") + hipGetErrorString(e);
  return "";
}

std::string Flow::energy_rhs(int isub) {
  SCHK(energy_terms());
  return energy_combine(isub, eres.p, nullptr);
}

std::string Flow::energy_terms() {
  set_lanes();
  if (!mixed) return energy_terms_t<double>(w64, e64);
  // rounded copies of g (kept for this sub-step's momentum right-hand side)
  // and, unless the last property refresh stored them, of h, k, T
  hipStream_t s = H->stream;
  for (auto pr2 : {std::make_pair(&w32.gx, &w64.gx), std::make_pair(&w32.gy, &w64.gy),
                   std::make_pair(&w32.gz, &w64.gz), std::make_pair(&e32.hent, &hent),
                   std::make_pair(&e32.kcond, &kcond), std::make_pair(&e32.temp, &temp)})
    if (!props32_fresh || pr2.first == &w32.gx || pr2.first == &w32.gy || pr2.first == &w32.gz)
      hipLaunchKernelGGL((k_cvt<float, double>), grid1(pr2.second->size()), dim3(256), 0, s, pr2.first->p,
                         pr2.second->p, pr2.second->size());
  return energy_terms_t<float>(w32, e32);
}

template <typename T>
std::string Flow::energy_terms_t(RhsWork<T>& w, EnergyWork<T>& e) {
  // Compute_energy_rhs (src/eq_energy.f90:178-396): convection -d(g_i h_i),
  // diffusion rPrRen d(k_i dT/dx_i); x on lane 0, y on lane 1, z on lane 2.
  // The y flux takes its own wall rows as wall values (extract_dirichlet_
  // fbcy) before the outer derivative: the reference guards that with
  // is_fbcx_velo_required (:356), which is false on this channel and would
  // pass MAXP wall values (wall rows ~1e19); the y-wall test is followed.
  int ftp[3][2], econv[3][2], edifu[3][2];
  for (int d = 0; d < 3; ++d)
    for (int s2 = 0; s2 < 2; ++s2) ftp[d][s2] = (ibc[d][VT][s2] == BC_NEUMANN) ? BC_DIRICHLET : ibc[d][VT][s2];
  for (int d = 0; d < 3; ++d) {
    prod_bc(ibc[d][d], ftp[d], econv[d]);
    tau_bc(ftp[d], ibc[d][VT], edifu[d]);
  }
  const T m1 = T(-1), r = T(rprren), one = T(1);
  auto o = [&](int lane, int dir, int opk, const int* bc, const Field<T>& A, Field<T>& out, const T* fbc, T sA,
               const Field<T>* Cm, T sOut, int acc) {
    return op<T>(w, dir, opk, bc, A, out, fbc, sA, nullptr, one, Cm, sOut, acc, nullptr, T(0), nullptr, T(0), lane);
  };
  SCHK(fork());
  // x
  SCHK(o(0, 0, MIDP_C2P, ftp[0], e.hent, e.hpcc, nullptr, one, nullptr, one, 0));
  SCHK(o(0, 0, MIDP_C2P, ftp[0], e.kcond, e.kpcc, nullptr, one, nullptr, one, 0));
  SCHK(o(0, 0, D1_P2C, econv[0], w.gx, e.ex, nullptr, m1, &e.hpcc, one, 0));
  SCHK(o(0, 0, D1_C2P, ibc[0][VT], e.temp, e.tpcc, nullptr, one, nullptr, one, 0));
  SCHK(o(0, 0, D1_P2C, edifu[0], e.tpcc, e.ex, nullptr, one, &e.kpcc, r, 1));
  // y
  SCHK(o(1, 1, MIDP_C2P, ftp[1], e.hent, e.hcpc, e.fby_h, one, nullptr, one, 0));
  SCHK(o(1, 1, MIDP_C2P, ftp[1], e.kcond, e.kcpc, e.fby_k, one, nullptr, one, 0));
  SCHK(o(1, 1, D1_P2C, econv[1], w.gy, e.ey, nullptr, m1, &e.hcpc, one, 0));
  SCHK(o(1, 1, D1_C2P, ibc[1][VT], e.temp, e.tcpc, e.fby_t, one, nullptr, one, 0));
  hipLaunchKernelGGL(k_wall_rows4<T>, grid1((long)nx * nz), dim3(256), 0, lane_stream(1), e.fby_q, e.tcpc.p,
                     (const T*)e.kcpc.p, nx, py, nz);
  SCHK(o(1, 1, D1_P2C, edifu[1], e.tcpc, e.ey, e.fby_q, one, &e.kcpc, r, 1));
  // z
  SCHK(o(2, 2, MIDP_C2P, ftp[2], e.hent, e.hccp, nullptr, one, nullptr, one, 0));
  SCHK(o(2, 2, MIDP_C2P, ftp[2], e.kcond, e.kccp, nullptr, one, nullptr, one, 0));
  SCHK(o(2, 2, D1_P2C, econv[2], w.gz, e.ez, nullptr, m1, &e.hccp, one, 0));
  SCHK(o(2, 2, D1_C2P, ibc[2][VT], e.temp, e.tccp, nullptr, one, nullptr, one, 0));
  SCHK(o(2, 2, D1_P2C, edifu[2], e.tccp, e.ez, nullptr, one, &e.kccp, r, 1));
  SCHK(join());
  return "";
}

std::string Flow::energy_substep(int isub) {
  // Solve_energy_eq, src/eq_energy.f90:398-458 (the isub 3 density history
  // shift only feeds Calculate_drhodt, which is zero)
  SCHK(energy_terms());
  g32_fresh = mixed;
  SCHK(energy_combine(isub, nullptr, rhoh.p));
  return refresh_props();
}

}  // namespace chs
