// MLP surrogate of the property table: h -> (T, rho, mu, kappa, cp).
//
// Network 1 -> 64 -> 64 -> 5 with tanh hidden activations.  The weights used
// by the forward and backward passes are BF16 (round-to-nearest-even copies
// of FP32 master weights); every product is accumulated in FP32.  Training
// data are samples of the table interpolation itself
// (ftp_refresh_thermal_properties_from_H, src/input_thermo.f90:345-374, run on
// the GPU over the table handle), so the table stays both the fallback and
// the yardstick of the surrogate.
//
// Kernels: a sampling kernel (uniform h, table interpolation, normalised
// targets); an inference kernel (one thread per sample, weights staged in LDS
// as FP32 values of the BF16 copies, read as wave-wide broadcasts); a per-sample
// backward kernel that recomputes the forward pass and writes the layer
// deltas transposed ([unit][sample]); deterministic batch reductions for the
// weight gradients (one thread per gradient entry, fixed summation order);
// and an Adam update of the FP32 master weights followed by the BF16 rounding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "chs_props.h"

namespace chs {
void set_error(const std::string& msg);
}

namespace {

constexpr int H1 = 64, H2 = 64, NO = 5;
constexpr int OW1 = 0, OB1 = OW1 + H1, OW2 = OB1 + H1, OB2 = OW2 + H2 * H1, OW3 = OB2 + H2, OB3 = OW3 + NO * H2;
constexpr int NP = OB3 + NO;  // 4613 parameters
constexpr int kMaxBatch = 65536;
// table columns predicted, in output order
constexpr int kOutCol[NO] = {CHS_PC_T, CHS_PC_D, CHS_PC_M, CHS_PC_K, CHS_PC_CP};

int fail(int code, const std::string& msg) {
  chs::set_error(msg);
  return code;
}

__host__ __device__ inline uint32_t mix32(uint32_t x) {  // integer hash (lowbias32)
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFU + ((u >> 16) & 1U);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct Norm {
  float xc, xs;      // x = (h - xc) / xs
  float mu[NO], sd[NO];  // y_o = (v_o - mu_o) / sd_o
};

__device__ inline void bisect(const double* c, int n, double v, int* i1o, int* i2o) {
  int i1 = 0, i2 = n - 1;
  while (i2 - i1 > 1) {
    const int im = i1 + (i2 - i1) / 2;
    const double d1 = c[i1] - v, dm = c[im] - v;
    if (d1 * dm > 1.0e-20) i1 = im;
    else i2 = im;
  }
  *i1o = i1;
  *i2o = i2;
}

// training set: h uniform in [h0, h1], targets by table interpolation in h
__global__ void k_sample(const double* __restrict__ tab, int n, double h0, double h1, int N, uint32_t seed, Norm nm,
                         float* __restrict__ xs, float* __restrict__ ts) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const uint32_t r = mix32(mix32((uint32_t)s * 2654435761U) ^ seed);
  const double u = (double)(r >> 8) * (1.0 / 16777216.0);
  const double h = h0 + (h1 - h0) * u;
  const double* ch = tab + (long)CHS_PC_H * n;
  int i1, i2;
  bisect(ch, n, h, &i1, &i2);
  const double w1 = (ch[i2] - h) / (ch[i2] - ch[i1]), w2 = 1.0 - w1;
  xs[s] = (float)((h - nm.xc) / nm.xs);
  for (int o = 0; o < NO; ++o) {
    const double* c = tab + (long)kOutCol[o] * n;
    const double v = w1 * c[i1] + w2 * c[i2];
    ts[(long)o * N + s] = (float)((v - nm.mu[o]) / nm.sd[o]);
  }
}

__device__ inline void stage_params(const uint16_t* __restrict__ p16, float* sp) {
  for (int i = threadIdx.x; i < NP; i += blockDim.x) sp[i] = bf2f(p16[i]);
  __syncthreads();
}

// double-precision inputs h -> denormalised double outputs (API path)
__global__ void __launch_bounds__(256) k_eval(const uint16_t* __restrict__ p16, Norm nm, const double* __restrict__ h,
                                              long N, double* __restrict__ o0, double* __restrict__ o1,
                                              double* __restrict__ o2, double* __restrict__ o3,
                                              double* __restrict__ o4) {
  __shared__ float sp[NP];
  stage_params(p16, sp);
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N) return;
  const float xv = (float)((h[b] - nm.xc) / nm.xs);
  float a1[H1];
#pragma unroll
  for (int i = 0; i < H1; ++i) a1[i] = tanhf(sp[OW1 + i] * xv + sp[OB1 + i]);
  float out[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) out[o] = sp[OB3 + o];
  for (int j = 0; j < H2; ++j) {
    float z = sp[OB2 + j];
#pragma unroll
    for (int i = 0; i < H1; ++i) z += sp[OW2 + j * H1 + i] * a1[i];
    const float a2 = tanhf(z);
#pragma unroll
    for (int o = 0; o < NO; ++o) out[o] += sp[OW3 + o * H2 + j] * a2;
  }
  double* dst[NO] = {o0, o1, o2, o3, o4};
#pragma unroll
  for (int o = 0; o < NO; ++o)
    if (dst[o]) dst[o][b] = (double)out[o] * nm.sd[o] + nm.mu[o];
}

// Per-sample backward pass of the mean squared error over B samples x 5
// outputs.  Samples are either given (given != 0: sample b of xs / ts) or
// drawn at random from the training set of N samples.  Writes transposed layer quantities for the
// batch reductions and one loss partial per block.
__global__ void __launch_bounds__(128) k_bwd(const uint16_t* __restrict__ p16, const float* __restrict__ xs,
                                             const float* __restrict__ ts, int N, int B, uint32_t seed,
                                             int given, float* __restrict__ xb, float* __restrict__ a1T,
                                             float* __restrict__ a2T, float* __restrict__ dz1T,
                                             float* __restrict__ dz2T, float* __restrict__ dyT,
                                             float* __restrict__ lossp) {
  __shared__ float sp[NP];
  __shared__ float red[128];
  stage_params(p16, sp);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float lsum = 0.0f;
  if (b < B) {
    int s = b;
    if (!given) s = (int)(mix32(mix32((uint32_t)b * 2246822519U) ^ seed) % (uint32_t)N);
    const float xv = xs[s];
    float a1[H1];
#pragma unroll
    for (int i = 0; i < H1; ++i) a1[i] = tanhf(sp[OW1 + i] * xv + sp[OB1 + i]);
    float out[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) out[o] = sp[OB3 + o];
    for (int j = 0; j < H2; ++j) {
      float z = sp[OB2 + j];
#pragma unroll
      for (int i = 0; i < H1; ++i) z += sp[OW2 + j * H1 + i] * a1[i];
      const float a2 = tanhf(z);
      a2T[(long)j * B + b] = a2;
#pragma unroll
      for (int o = 0; o < NO; ++o) out[o] += sp[OW3 + o * H2 + j] * a2;
    }
    const float sc = 2.0f / (float)(B * NO);
    float dy[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      const float e = out[o] - ts[(long)o * (given ? B : N) + s];
      lsum += e * e;
      dy[o] = sc * e;
      dyT[(long)o * B + b] = dy[o];
    }
    float da1[H1];
#pragma unroll
    for (int i = 0; i < H1; ++i) da1[i] = 0.0f;
    for (int j = 0; j < H2; ++j) {
      const float a2 = a2T[(long)j * B + b];
      float da2 = 0.0f;
#pragma unroll
      for (int o = 0; o < NO; ++o) da2 += sp[OW3 + o * H2 + j] * dy[o];
      const float dz2 = da2 * (1.0f - a2 * a2);
      dz2T[(long)j * B + b] = dz2;
#pragma unroll
      for (int i = 0; i < H1; ++i) da1[i] += sp[OW2 + j * H1 + i] * dz2;
    }
#pragma unroll
    for (int i = 0; i < H1; ++i) {
      a1T[(long)i * B + b] = a1[i];
      dz1T[(long)i * B + b] = da1[i] * (1.0f - a1[i] * a1[i]);
    }
    xb[b] = xv;
  }
  red[threadIdx.x] = lsum;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// G[r][c] = sum_b U[r][b] * V[c][b] (V null: V = 1) in a fixed order: one
// workgroup of 256 threads per gradient entry (r, c) = blockIdx.x; thread t
// sums b = t, t + 256, ... (coalesced across the workgroup), then a fixed
// tree in LDS adds the 256 partial sums.
constexpr int kOuterThreads = 256;
__global__ void __launch_bounds__(kOuterThreads) k_outer(const float* __restrict__ U, int R, const float* __restrict__ V,
                                                         int C, int B, float* __restrict__ G) {
  __shared__ float red[kOuterThreads];
  const int e = blockIdx.x;
  const int r = e / C, c = e - r * C;
  const float* u = U + (long)r * B;
  const float* v = V ? V + (long)c * B : nullptr;
  float acc = 0.0f;
  for (int b = threadIdx.x; b < B; b += kOuterThreads) acc += u[b] * (v ? v[b] : 1.0f);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kOuterThreads / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) G[e] = red[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ p16, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  p16[i] = f2bf(pi);
}

__global__ void k_round(const float* __restrict__ p, uint16_t* __restrict__ q, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = f2bf(p[i]);
}

}  // namespace

struct chs_mlp {
  chs_props* tab = nullptr;
  hipStream_t stream = nullptr;
  Norm nm{};
  float *p32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
  uint16_t* p16 = nullptr;
  float *xs = nullptr, *ts = nullptr;
  int nsamp = 0;
  float *xb = nullptr, *a1T = nullptr, *a2T = nullptr, *dz1T = nullptr, *dz2T = nullptr, *dyT = nullptr;
  float *lossp = nullptr, *tin = nullptr;
  int step = 0;
  uint32_t seed = 1;
  bool matrix_cores = true;  // property refresh on the matrix cores (k_mlp_from_dh_mfma)
  void* io = nullptr;
  size_t io_bytes = 0;
  std::vector<void*> owned;
  ~chs_mlp() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* q : owned) (void)hipFree(q);
    if (io) (void)hipFree(io);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

template <typename T>
bool dalloc(chs_mlp* m, T** p, size_t n) {
  if (hipMalloc(p, sizeof(T) * (n ? n : 1)) != hipSuccess) return false;
  m->owned.push_back(*p);
  return true;
}

// gradients of the batch (already staged by k_bwd) into m->g32; loss sum
int reduce_grads(chs_mlp* m, int B, int nblk, float* loss) {
  hipStream_t s = m->stream;
  auto outer = [&](const float* U, int R, const float* V, int C, float* G) {
    hipLaunchKernelGGL(k_outer, dim3(R * C), dim3(kOuterThreads), 0, s, U, R, V, C, B, G);
  };
  outer(m->dz1T, H1, m->xb, 1, m->g32 + OW1);
  outer(m->dz1T, H1, nullptr, 1, m->g32 + OB1);
  outer(m->dz2T, H2, m->a1T, H1, m->g32 + OW2);
  outer(m->dz2T, H2, nullptr, 1, m->g32 + OB2);
  outer(m->dyT, NO, m->a2T, H2, m->g32 + OW3);
  outer(m->dyT, NO, nullptr, 1, m->g32 + OB3);
  if (loss) {
    std::vector<float> h(nblk);
    if (hipMemcpyAsync(h.data(), m->lossp, sizeof(float) * nblk, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(CHS_EHIP, "loss copy");
    double t = 0.0;
    for (float v : h) t += v;
    *loss = (float)(t / ((double)B * NO));
  }
  if (hipGetLastError() != hipSuccess) return fail(CHS_EHIP, "gradient kernels");
  return CHS_OK;
}

}  // namespace

namespace {

// ---- matrix-core form of the network (inference) ----
// v_mfma_f32_16x16x16_bf16: D(16x16) = A(16x16) B(16x16) + C, one wave; lane
// l holds A[l%16][4(l/16)+0..3], B[4(l/16)+0..3][l%16] and
// D[4(l/16)+0..3][l%16].  The 64x64 layer is 4 (output blocks) x 4 (input
// blocks) of these per 16 cells: A = a block of W2 (rows: output units,
// columns: input units), B = the BF16 first-layer activations (rows: input
// units, columns: cells), D = the second-layer pre-activations.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// tanh x = 1 - 2 / (1 + 2^{2 log2(e) x}) with the hardware base-2
// exponential and reciprocal (v_exp_f32, v_rcp_f32, each ~1 ulp: absolute
// error ~3e-7, below the BF16 rounding of the activations).  Five
// instructions; a written divide (also __fdividef) compiles to the
// correctly rounded IEEE sequence of fifteen.  Saturates to +-1 for large
// |x| (exp2 overflow gives a zero reciprocal).
__device__ __forceinline__ float tanh_fast(float x) {
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(2.8853900817779268f * x));
}

// BF16 of a finite float, rounded to nearest even by the hardware
// conversion (v_cvt_pk_bf16_f32; the same bits as f2bf for finite values)
__device__ __forceinline__ short bf16_rne(float x) { return __builtin_bit_cast(short, (__bf16)x); }

// out[0..4] (normalised) of the 16 cells whose normalised inputs the lanes
// with l%16 = c hold in xv; every lane ends with the outputs of cell l%16.
// wa: this lane's 16 A operands (W2 blocks), sp: the FP32 parameters in LDS.
__device__ __forceinline__ void mlp16_mfma(const bf16x4 wa[16], const float* __restrict__ sp, float xv, float out[NO]) {
  const int q = (threadIdx.x & 63) >> 4;
  bf16x4 bv[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int u = 16 * kk + 4 * q + r;
      bv[kk][r] = bf16_rne(tanh_fast(sp[OW1 + u] * xv + sp[OB1 + u]));
    }
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) out[o] = 0.0f;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[4 * mt + kk], bv[kk], acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int u = 16 * mt + 4 * q + r;
      const float a2 = tanh_fast(acc[r] + sp[OB2 + u]);
#pragma unroll
      for (int o = 0; o < NO; ++o) out[o] += sp[OW3 + o * H2 + u] * a2;
    }
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    out[o] += __shfl_xor(out[o], 16);
    out[o] += __shfl_xor(out[o], 32);
    out[o] += sp[OB3 + o];
  }
}

// k_mlp_from_dh on the matrix cores: 64 cells per wave (one per lane for the
// clamp and the bisection), the network evaluated for them in four passes of
// 16 cells, twice (h from the table, then h = rho*h / rho).
__global__ void __launch_bounds__(256) k_mlp_from_dh_mfma(const uint16_t* __restrict__ p16, Norm nm,
                                                          const double* __restrict__ tab, int nt, double dhmin,
                                                          double dhmax, const double* __restrict__ rhoh, long N,
                                                          double* __restrict__ oh, double* __restrict__ od,
                                                          double* __restrict__ om, double* __restrict__ ok,
                                                          double* __restrict__ ot, int* __restrict__ flag,
                                                          float* __restrict__ fh, float* __restrict__ fm,
                                                          float* __restrict__ fk, float* __restrict__ ft) {
  __shared__ float sp[NP];
  stage_params(p16, sp);
  const int lane = threadIdx.x & 63;
  const int ci = lane & 15, q = lane >> 4;
  bf16x4 wa[16];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      wa[4 * mt + kk] = *reinterpret_cast<const bf16x4*>(p16 + OW2 + (16 * mt + ci) * H1 + 16 * kk + 4 * q);
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = c < N;
  double v = live ? rhoh[c] : dhmin;
  if (live && v < dhmin) {
    v = dhmin;
    atomicAdd(flag, 1);
  } else if (live && v > dhmax) {
    v = dhmax;
    atomicAdd(flag, 1);
  }
  const double* crh = tab + (long)CHS_PC_RH * nt;
  const double* ch = tab + (long)CHS_PC_H * nt;
  int i1, i2;
  bisect(crh, nt, v, &i1, &i2);
  const double w1 = (crh[i2] - v) / (crh[i2] - crh[i1]), w2 = 1.0 - w1;
  double h = w1 * ch[i1] + w2 * ch[i2];
  float mine[NO];
  for (int pass = 0; pass < 2; ++pass) {
    const float x = (float)((h - nm.xc) / nm.xs);
#pragma unroll 1
    for (int p = 0; p < 4; ++p) {
      float out[NO];
      mlp16_mfma(wa, sp, __shfl(x, 16 * p + ci), out);
      if (q == p)
#pragma unroll
        for (int o = 0; o < NO; ++o) mine[o] = out[o];
    }
    if (pass == 0) h = v / ((double)mine[1] * nm.sd[1] + nm.mu[1]);
  }
  if (!live) return;
  const double vt = (double)mine[0] * nm.sd[0] + nm.mu[0], vm = (double)mine[2] * nm.sd[2] + nm.mu[2],
               vk = (double)mine[3] * nm.sd[3] + nm.mu[3];
  if (oh) oh[c] = h;
  if (ot) ot[c] = vt;
  if (od) od[c] = (double)mine[1] * nm.sd[1] + nm.mu[1];
  if (om) om[c] = vm;
  if (ok) ok[c] = vk;
  if (fh) fh[c] = (float)h;
  if (fm) fm[c] = (float)vm;
  if (fk) fk[c] = (float)vk;
  if (ft) ft[c] = (float)vt;
}

// Properties from rho*h with the surrogate (flow solver's refresh, mode
// CHS_PROPS_MLP): clamp to [dhmin, dhmax] (counted), bisection on the
// table's rho*h column and linear interpolation for the first h (as
// ftp_refresh_thermal_properties_from_DH, src/input_thermo.f90:393-446),
// then the network for the properties from h, h = rho*h / rho, and the
// network again; fh, fm, fk, ft (each possibly null) receive h, mu, kappa
// and T rounded to FP32.  One thread per cell; parameters in LDS as in k_eval.
__global__ void __launch_bounds__(256) k_mlp_from_dh(const uint16_t* __restrict__ p16, Norm nm,
                                                     const double* __restrict__ tab, int nt, double dhmin,
                                                     double dhmax, const double* __restrict__ rhoh, long N,
                                                     double* __restrict__ oh, double* __restrict__ od,
                                                     double* __restrict__ om, double* __restrict__ ok,
                                                     double* __restrict__ ot, int* __restrict__ flag,
                                                     float* __restrict__ fh, float* __restrict__ fm,
                                                     float* __restrict__ fk, float* __restrict__ ft) {
  __shared__ float sp[NP];
  stage_params(p16, sp);
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  double v = rhoh[c];
  if (v < dhmin) {
    v = dhmin;
    atomicAdd(flag, 1);
  } else if (v > dhmax) {
    v = dhmax;
    atomicAdd(flag, 1);
  }
  const double* crh = tab + (long)CHS_PC_RH * nt;
  const double* ch = tab + (long)CHS_PC_H * nt;
  int i1, i2;
  bisect(crh, nt, v, &i1, &i2);
  const double w1 = (crh[i2] - v) / (crh[i2] - crh[i1]), w2 = 1.0 - w1;
  double h = w1 * ch[i1] + w2 * ch[i2];
  float out[NO];
  for (int pass = 0; pass < 2; ++pass) {
    const float xv = (float)((h - nm.xc) / nm.xs);
    float a1[H1];
#pragma unroll
    for (int i = 0; i < H1; ++i) a1[i] = tanhf(sp[OW1 + i] * xv + sp[OB1 + i]);
#pragma unroll
    for (int o = 0; o < NO; ++o) out[o] = sp[OB3 + o];
    for (int j = 0; j < H2; ++j) {
      float z = sp[OB2 + j];
#pragma unroll
      for (int i = 0; i < H1; ++i) z += sp[OW2 + j * H1 + i] * a1[i];
      const float a2 = tanhf(z);
#pragma unroll
      for (int o = 0; o < NO; ++o) out[o] += sp[OW3 + o * H2 + j] * a2;
    }
    if (pass == 0) h = v / ((double)out[1] * nm.sd[1] + nm.mu[1]);
  }
  const double vt = (double)out[0] * nm.sd[0] + nm.mu[0], vm = (double)out[2] * nm.sd[2] + nm.mu[2],
               vk = (double)out[3] * nm.sd[3] + nm.mu[3];
  if (oh) oh[c] = h;
  if (ot) ot[c] = vt;
  if (od) od[c] = (double)out[1] * nm.sd[1] + nm.mu[1];
  if (om) om[c] = vm;
  if (ok) ok[c] = vk;
  if (fh) fh[c] = (float)h;
  if (fm) fm[c] = (float)vm;
  if (fk) fk[c] = (float)vk;
  if (ft) ft[c] = (float)vt;
}

}  // namespace

namespace chs {
std::string mlp_refresh_dh(chs_mlp* m, const chs_props* p, const double* rhoh, long n, double* h, double* d,
                           double* mu, double* k, double* t, int* flag, hipStream_t s, float* const* r32) {
  if (n <= 0) return "";
  float* const f[4] = {r32 ? r32[0] : nullptr, r32 ? r32[1] : nullptr, r32 ? r32[2] : nullptr,
                       r32 ? r32[3] : nullptr};
  if (m->matrix_cores)
    hipLaunchKernelGGL(k_mlp_from_dh_mfma, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->p16, m->nm, p->dev,
                       p->n, p->dhmin, p->dhmax, rhoh, n, h, d, mu, k, t, flag, f[0], f[1], f[2], f[3]);
  else
    hipLaunchKernelGGL(k_mlp_from_dh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->p16, m->nm, p->dev,
                       p->n, p->dhmin, p->dhmax, rhoh, n, h, d, mu, k, t, flag, f[0], f[1], f[2], f[3]);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("surrogate kernel: ") + hipGetErrorString(e);
  return "";
}
}  // namespace chs

extern "C" {

int chs_mlp_create(chs_props_t* table, int nsamples, unsigned seed, chs_mlp_t** out) {
  if (!table || !out || nsamples < 1) return fail(CHS_EINVAL, "invalid argument");
  *out = nullptr;
  chs_mlp* m = new chs_mlp();
  m->tab = table;
  m->seed = seed ? seed : 1U;
  m->nsamp = nsamples;
  const std::vector<double>& hc = table->col[CHS_PC_H];
  double h0 = hc[0], h1 = hc[0];
  for (double v : hc) { h0 = std::min(h0, v); h1 = std::max(h1, v); }
  m->nm.xc = (float)(0.5 * (h0 + h1));
  m->nm.xs = (float)(0.5 * (h1 - h0));
  for (int o = 0; o < NO; ++o) {
    const std::vector<double>& c = table->col[kOutCol[o]];
    double mu = 0.0, sd = 0.0;
    for (double v : c) mu += v;
    mu /= (double)c.size();
    for (double v : c) sd += (v - mu) * (v - mu);
    sd = std::sqrt(sd / (double)c.size());
    m->nm.mu[o] = (float)mu;
    m->nm.sd[o] = (float)(sd > 0.0 ? sd : 1.0);
  }
  // Glorot-uniform initial weights from the seed, zero biases
  std::vector<float> p(NP, 0.0f);
  uint32_t st = mix32(m->seed ^ 0x5bd1e995U);
  auto uni = [&](float lim) {
    st = mix32(st + 0x9e3779b9U);
    return lim * (2.0f * (float)(st >> 8) * (1.0f / 16777216.0f) - 1.0f);
  };
  for (int i = 0; i < H1; ++i) p[OW1 + i] = uni(std::sqrt(6.0f / (1 + H1)));
  for (int i = 0; i < H2 * H1; ++i) p[OW2 + i] = uni(std::sqrt(6.0f / (H1 + H2)));
  for (int i = 0; i < NO * H2; ++i) p[OW3 + i] = uni(std::sqrt(6.0f / (H2 + NO)));
  bool ok = hipStreamCreate(&m->stream) == hipSuccess && dalloc(m, &m->p32, NP) && dalloc(m, &m->g32, NP) &&
            dalloc(m, &m->m32, NP) && dalloc(m, &m->v32, NP) && dalloc(m, &m->p16, NP) &&
            dalloc(m, &m->xs, nsamples) && dalloc(m, &m->ts, (size_t)NO * nsamples) && dalloc(m, &m->xb, kMaxBatch) &&
            dalloc(m, &m->a1T, (size_t)H1 * kMaxBatch) && dalloc(m, &m->a2T, (size_t)H2 * kMaxBatch) &&
            dalloc(m, &m->dz1T, (size_t)H1 * kMaxBatch) && dalloc(m, &m->dz2T, (size_t)H2 * kMaxBatch) &&
            dalloc(m, &m->dyT, (size_t)NO * kMaxBatch) && dalloc(m, &m->lossp, kMaxBatch / 128 + 1) &&
            dalloc(m, &m->tin, (size_t)NO * kMaxBatch);
  if (!ok) {
    delete m;
    return fail(CHS_ENOMEM, "device allocation failed");
  }
  hipStream_t s = m->stream;
  ok = hipMemcpyAsync(m->p32, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice, s) == hipSuccess &&
       hipMemsetAsync(m->m32, 0, sizeof(float) * NP, s) == hipSuccess &&
       hipMemsetAsync(m->v32, 0, sizeof(float) * NP, s) == hipSuccess;
  hipLaunchKernelGGL(k_round, dim3((NP + 255) / 256), dim3(256), 0, s, m->p32, m->p16, NP);
  hipLaunchKernelGGL(k_sample, dim3((nsamples + 255) / 256), dim3(256), 0, s, table->dev, table->n, h0, h1, nsamples,
                     m->seed, m->nm, m->xs, m->ts);
  if (!ok || hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    delete m;
    return fail(CHS_EHIP, "initialisation kernels");
  }
  *out = m;
  return CHS_OK;
}

int chs_mlp_destroy(chs_mlp_t* m) {
  delete m;
  return CHS_OK;
}

int chs_mlp_info(const chs_mlp_t* m, int* nparams, float norm[12]) {
  if (!m) return fail(CHS_EINVAL, "null handle");
  if (nparams) *nparams = NP;
  if (norm) {
    norm[0] = m->nm.xc;
    norm[1] = m->nm.xs;
    for (int o = 0; o < NO; ++o) {
      norm[2 + o] = m->nm.mu[o];
      norm[7 + o] = m->nm.sd[o];
    }
  }
  return CHS_OK;
}

int chs_mlp_get_params(chs_mlp_t* m, float* p32, unsigned short* p16) {
  if (!m) return fail(CHS_EINVAL, "null handle");
  if (p32 && hipMemcpyAsync(p32, m->p32, sizeof(float) * NP, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    return fail(CHS_EHIP, "copy");
  if (p16 && hipMemcpyAsync(p16, m->p16, sizeof(uint16_t) * NP, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
    return fail(CHS_EHIP, "copy");
  if (hipStreamSynchronize(m->stream) != hipSuccess) return fail(CHS_EHIP, "synchronise");
  return CHS_OK;
}

int chs_mlp_set_params(chs_mlp_t* m, const float* p32) {
  if (!m || !p32) return fail(CHS_EINVAL, "invalid argument");
  if (hipMemcpyAsync(m->p32, p32, sizeof(float) * NP, hipMemcpyHostToDevice, m->stream) != hipSuccess)
    return fail(CHS_EHIP, "copy");
  hipLaunchKernelGGL(k_round, dim3((NP + 255) / 256), dim3(256), 0, m->stream, m->p32, m->p16, NP);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return fail(CHS_EHIP, "synchronise");
  return CHS_OK;
}

int chs_mlp_grad(chs_mlp_t* m, const float* x, const float* t, int B, float* grad, float* loss) {
  if (!m || !x || !t || !grad || B < 1 || B > kMaxBatch) return fail(CHS_EINVAL, "invalid argument");
  hipStream_t s = m->stream;
  // stage the given samples: inputs in xb, targets as [o][b]
  std::vector<float> tt((size_t)NO * B);
  for (int b = 0; b < B; ++b)
    for (int o = 0; o < NO; ++o) tt[(size_t)o * B + b] = t[(size_t)b * NO + o];
  if (hipMemcpyAsync(m->tin, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(m->xb, x, sizeof(float) * B, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(CHS_EHIP, "copy in");
  const int nblk = (B + 127) / 128;
  hipLaunchKernelGGL(k_bwd, dim3(nblk), dim3(128), 0, s, m->p16, m->xb, m->tin, B, B, 0U, 1, m->xb, m->a1T, m->a2T,
                     m->dz1T, m->dz2T, m->dyT, m->lossp);
  int rc = reduce_grads(m, B, nblk, loss);
  if (rc != CHS_OK) return rc;
  if (hipMemcpyAsync(grad, m->g32, sizeof(float) * NP, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(CHS_EHIP, "copy out");
  return CHS_OK;
}

int chs_mlp_train(chs_mlp_t* m, int steps, int batch, float lr, float* loss) {
  if (!m || steps < 0 || batch < 1 || batch > kMaxBatch || !(lr > 0.0f)) return fail(CHS_EINVAL, "invalid argument");
  hipStream_t s = m->stream;
  const int nblk = (batch + 127) / 128;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; ++it) {
    ++m->step;
    const uint32_t sd = mix32(m->seed * 747796405U + (uint32_t)m->step);
    hipLaunchKernelGGL(k_bwd, dim3(nblk), dim3(128), 0, s, m->p16, m->xs, m->ts, m->nsamp, batch, sd, 0, m->xb,
                       m->a1T, m->a2T, m->dz1T, m->dz2T, m->dyT, m->lossp);
    int rc = reduce_grads(m, batch, nblk, (loss && it == steps - 1) ? loss : nullptr);
    if (rc != CHS_OK) return rc;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, s, m->p32, m->g32, m->m32, m->v32, m->p16, NP,
                       lr, b1, b2, eps, c1, c2);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return fail(CHS_EHIP, "training");
  return CHS_OK;
}

int chs_mlp_set_matrix_cores(chs_mlp_t* m, int on) {
  if (!m) return fail(CHS_EINVAL, "null handle");
  m->matrix_cores = on != 0;
  return CHS_OK;
}

int chs_mlp_eval(chs_mlp_t* m, const double* h, long n, double* T, double* rho, double* mu, double* kappa,
                 double* cp, int mem) {
  if (!m || !h || n < 0) return fail(CHS_EINVAL, "invalid argument");
  if (mem != CHS_MEM_HOST && mem != CHS_MEM_DEVICE) return fail(CHS_EINVAL, "invalid mem kind");
  if (n == 0) return CHS_OK;
  hipStream_t s = m->stream;
  double* outs[NO] = {T, rho, mu, kappa, cp};
  const double* din = h;
  double* dout[NO] = {T, rho, mu, kappa, cp};
  if (mem == CHS_MEM_HOST) {
    int nout = 0;
    for (double* o : outs) nout += o != nullptr;
    const size_t need = sizeof(double) * (size_t)n * (1 + nout);
    if (need > m->io_bytes) {
      if (m->io) (void)hipFree(m->io);
      m->io = nullptr;
      m->io_bytes = 0;
      if (hipMalloc(&m->io, need) != hipSuccess) return fail(CHS_ENOMEM, "staging allocation failed");
      m->io_bytes = need;
    }
    double* base = static_cast<double*>(m->io);
    if (hipMemcpyAsync(base, h, sizeof(double) * n, hipMemcpyHostToDevice, s) != hipSuccess)
      return fail(CHS_EHIP, "copy in");
    din = base;
    int slot = 1;
    for (int o = 0; o < NO; ++o) dout[o] = outs[o] ? base + (size_t)n * slot++ : nullptr;
  }
  hipLaunchKernelGGL(k_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m->p16, m->nm, din, n, dout[0],
                     dout[1], dout[2], dout[3], dout[4]);
  if (hipGetLastError() != hipSuccess) return fail(CHS_EHIP, "inference launch");
  if (mem == CHS_MEM_HOST)
    for (int o = 0; o < NO; ++o)
      if (outs[o] && hipMemcpyAsync(outs[o], dout[o], sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess)
        return fail(CHS_EHIP, "copy out");
  if (hipStreamSynchronize(s) != hipSuccess) return fail(CHS_EHIP, "synchronise");
  return CHS_OK;
}

}  // extern "C"
