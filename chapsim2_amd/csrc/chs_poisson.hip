// FFT(x,z) + TDMA(y) pressure Poisson solve on gfx950.
//
// Follows src/poisson_fishpack.f90:354-490 (fishpack_fft_simple): forward real
// FFT along x, forward FFT along z, a tridiagonal solve along y per wavenumber
// pair with eigenvalues -4/h^2 sin^2(pi m/N) (fishpack_root_1D, :60-157) and the
// 2nd-order y operator of fishpack_fft_init (:213-350), then the inverse
// transforms and the 1/(nx*nz) scaling.  FFTPACK's half-complex real-to-real
// layout is replaced by the equivalent complex spectrum (kx = 0..nx/2, all kz):
// each FFTPACK coefficient pair shares one eigenvalue, so the solves coincide.
//
// Transforms are LDS-resident Stockham FFTs (radix 4/2/3/5), a workgroup per
// tile of lines; the x transform packs the real line into nx/2 complex points.
// The per-mode tridiagonal factors (TRID0's D and pivots, including its
// singular-pivot rule |Z| <= 1e-10 => last unknown 0) are computed once on the
// host in TRID0's order and read by the y kernel.
#include "chs_poisson.h"

#include <cmath>
#include <cstring>

#include "chs_plan.h"
#include "chs_tune.h"

namespace chs {

#define PCHK(x)                                                             \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) return std::string(#x ": ") + hipGetErrorString(e_); \
  } while (0)

std::vector<int> radix_plan(int n) {
  std::vector<int> r;
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  while (n % 3 == 0) { r.push_back(3); n /= 3; }
  while (n % 5 == 0) { r.push_back(5); n /= 5; }
  if (n != 1) r.clear();
  return r;
}

template <typename T>
struct cx {
  T re, im;
};
template <typename T>
__device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ inline cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ inline cx<T> cmul(cx<T> a, cx<T> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <typename T>
__device__ inline cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
// multiply by s*i (s = +-1)
template <typename T>
__device__ inline cx<T> cmuli(cx<T> a, T s) { return {-s * a.im, s * a.re}; }

constexpr int kTileC = 4096;  // complex points per workgroup tile
constexpr int kThreads = 256;

template <typename T, int R>
__device__ inline void small_dft(cx<T>* v, T s) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cx<T> t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 3) {
    const T h = T(0.5), q = T(0.86602540378443864676);
    cx<T> t = cadd(v[1], v[2]), u = cmuli(csub(v[1], v[2]), s * q);
    cx<T> m = csub(v[0], cscale(t, h));
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, u);
    v[2] = csub(m, u);
  } else {  // R == 5
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    cx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    cx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    cx<T> m1 = cadd(v[0], cadd(cscale(a1, c1), cscale(a2, c2)));
    cx<T> m2 = cadd(v[0], cadd(cscale(a1, c2), cscale(a2, c1)));
    cx<T> n1 = cmuli(cadd(cscale(b1, s1), cscale(b2, s2)), s);
    cx<T> n2 = cmuli(csub(cscale(b1, s2), cscale(b2, s1)), s);
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

// One Stockham stage over `nl` lines of Nc points held in LDS (in place: all
// inputs are read into registers before the barrier, outputs written after).
template <typename T, int R>
__device__ void stage(cx<T>* buf, int Nc, int nl, int Ns, const cx<T>* tw, T s) {
  constexpr int QMAX = (16 + R - 1) / R;
  const int nb = Nc / R;
  const int total = nl * nb;
  const int tstep = Nc / (Ns * R);
  cx<T> v[QMAX][R];
  int bidx[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    int b = threadIdx.x + q * kThreads;
    bidx[q] = b;
    if (b < total) {
      int line = b / nb, j = b % nb;
      int k = j % Ns;
      const cx<T>* src = buf + line * Nc + j;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> x = src[r * nb];
        if (r > 0 && k > 0) {
          cx<T> w = tw[r * k * tstep];
          w.im = s < T(0) ? w.im : -w.im;
          x = cmul(x, w);
        }
        v[q][r] = x;
      }
      small_dft<T, R>(v[q], s);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    int b = bidx[q];
    if (b < total) {
      int line = b / nb, j = b % nb;
      int k = j % Ns;
      cx<T>* dst = buf + line * Nc + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

struct Radices {
  int n;
  int r[12];
};

template <typename T>
__device__ void fft_tile(cx<T>* buf, int Nc, int nl, const Radices& rp, const cx<T>* tw, T s) {
  int Ns = 1;
  for (int q = 0; q < rp.n; ++q) {
    switch (rp.r[q]) {
      case 2: stage<T, 2>(buf, Nc, nl, Ns, tw, s); break;
      case 3: stage<T, 3>(buf, Nc, nl, Ns, tw, s); break;
      case 4: stage<T, 4>(buf, Nc, nl, Ns, tw, s); break;
      default: stage<T, 5>(buf, Nc, nl, Ns, tw, s); break;
    }
    Ns *= rp.r[q];
  }
}

// x forward: real lines (nx) -> complex spectrum kx = 0..nx/2 into W.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_fft_x_fwd(const T* __restrict__ in, cx<T>* __restrict__ W, int nx,
                                                       long nlines, Radices rp, const cx<T>* tw, const cx<T>* rot,
                                                       T in_scale) {
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int N = nx / 2, nxh = N + 1;
  const int lpb = kTileC / N;
  const long L0 = (long)blockIdx.x * lpb;
  const int nl = (int)((nlines - L0) < lpb ? (nlines - L0) : lpb);
  for (int t = threadIdx.x; t < nl * N; t += kThreads) {
    int line = t / N, n = t % N;
    const T* src = in + (L0 + line) * nx;
    buf[t] = {src[2 * n] * in_scale, src[2 * n + 1] * in_scale};
  }
  __syncthreads();
  fft_tile<T>(buf, N, nl, rp, tw, T(-1));
  for (int t = threadIdx.x; t < nl * nxh; t += kThreads) {
    int line = t / nxh, k = t % nxh;
    cx<T> zk = buf[line * N + (k % N)];
    cx<T> zn = buf[line * N + ((N - k) % N)];
    zn.im = -zn.im;  // conj(Z[N-k])
    cx<T> e = cscale(cadd(zk, zn), T(0.5));
    cx<T> d = csub(zk, zn);
    cx<T> o = {d.im * T(0.5), -d.re * T(0.5)};  // (Z - conj)/(2i)
    cx<T> x = cadd(e, cmul(rot[k], o));
    W[(L0 + line) * nxh + k] = x;
  }
}

// x inverse: complex spectrum kx = 0..nx/2 -> real lines, times `scale`.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_fft_x_inv(const cx<T>* __restrict__ W, T* __restrict__ out, int nx,
                                                       long nlines, Radices rp, const cx<T>* tw, const cx<T>* rot,
                                                       T scale) {
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int N = nx / 2, nxh = N + 1;
  const int lpb = kTileC / N;
  const long L0 = (long)blockIdx.x * lpb;
  const int nl = (int)((nlines - L0) < lpb ? (nlines - L0) : lpb);
  for (int t = threadIdx.x; t < nl * N; t += kThreads) {
    int line = t / N, k = t % N;
    const cx<T>* src = W + (L0 + line) * nxh;
    cx<T> xk = src[k], xn = src[N - k];
    xn.im = -xn.im;  // conj(X[N-k])
    cx<T> e = cadd(xk, xn);
    cx<T> r = rot[k];
    r.im = -r.im;  // exp(+2 pi i k / nx)
    cx<T> o = cmul(csub(xk, xn), r);
    buf[t] = {e.re - o.im, e.im + o.re};  // e + i*o
  }
  __syncthreads();
  fft_tile<T>(buf, N, nl, rp, tw, T(1));
  for (int t = threadIdx.x; t < nl * N; t += kThreads) {
    int line = t / N, n = t % N;
    T* dst = out + (L0 + line) * nx;
    cx<T> z = buf[t];
    dst[2 * n] = z.re * scale;
    dst[2 * n + 1] = z.im * scale;
  }
}

// Complex transform (either sign) in place along one strided dimension of the
// spectrum W(nxh, ny, nz): lines of n points at stride ls, for every kx (tiles
// of consecutive kx, so loads and stores are contiguous along kx) and every
// value of the remaining index (count no, stride os).  z: n = nz,
// ls = nxh*ny, os = nxh; y: n = ny, ls = nxh, os = nxh*ny.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_fft_s(cx<T>* __restrict__ W, int nxh, int n, long ls, int no, long os,
                                                   Radices rp, const cx<T>* tw, T s) {
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int lpb = kTileC / n;
  const int ntile = (nxh + lpb - 1) / lpb;
  const int tile = blockIdx.x % ntile, o = blockIdx.x / ntile;
  if (o >= no) return;
  const int kx0 = tile * lpb;
  const int nl = (nxh - kx0) < lpb ? (nxh - kx0) : lpb;
  const long base = kx0 + os * o;
  for (int t = threadIdx.x; t < nl * n; t += kThreads) {
    int m = t / nl, line = t % nl;
    buf[line * n + m] = W[base + line + ls * m];
  }
  __syncthreads();
  fft_tile<T>(buf, n, nl, rp, tw, s);
  for (int t = threadIdx.x; t < nl * n; t += kThreads) {
    int m = t / nl, line = t % nl;
    W[base + line + ls * m] = buf[line * n + m];
  }
}

// ---------------------------------------------------------------------------
// Two-pass transforms for N = A * B (A, B <= 32).  With n = a + A b
// and k = k2 + B k1 (a, k1 < A; b, k2 < B):
//   X[k2 + B k1] = sum_a W_A^(a k1) [ W_N^(a k2) sum_b x[a + A b] W_B^(b k2) ].
// A threads per line: thread a loads x[a + A b] for all b and does the
// B-point DFT in registers, applies W_N^(a k2) (table, index a k2 < N), and
// writes to LDS (rows of A + 1 entries, so the A-point reads of pass 2 hit
// distinct banks); after one barrier, thread a does the A-point DFTs for
// k2 = a, a + A, ... < B in registers.  The register DFTs of power-of-two
// length are radix-2 with compile-time twiddles W_32^m = (-i)^(m/8)
// W_32^(m%8) from the first octant below, exactly symmetric; other lengths
// (3, 5, 6, 9, 10, 12, 15) are direct sums over the N-point twiddle table.
__device__ constexpr double kOctC[8] = {1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452,
                                        0.7071067811865476, 0.5555702330196022, 0.3826834323650898,
                                        0.19509032201612825};
__device__ constexpr double kOctS[8] = {0.0, 0.19509032201612825, 0.3826834323650898, 0.5555702330196022,
                                        0.7071067811865476, 0.8314696123025452, 0.9238795325112867,
                                        0.9807852804032304};

// v * W_32^m, W = exp(-2 pi i / 32) (forward) or its conjugate (inverse)
template <typename T, bool FWD>
__device__ __forceinline__ cx<T> tw32(cx<T> v, int m) {
  m &= 31;
  const int q = m >> 3, r = m & 7;
  cx<T> t = v;
  if (r != 0) {
    const T c = T(kOctC[r]), sn = FWD ? T(-kOctS[r]) : T(kOctS[r]);
    t = {v.re * c - v.im * sn, v.re * sn + v.im * c};
  }
  // times (-i)^q (forward) or (+i)^q (inverse)
  if (q == 1) t = FWD ? cx<T>{t.im, -t.re} : cx<T>{-t.im, t.re};
  if (q == 2) t = {-t.re, -t.im};
  if (q == 3) t = FWD ? cx<T>{-t.im, t.re} : cx<T>{t.im, -t.re};
  return t;
}

__host__ __device__ constexpr int bitrev_c(int i, int R) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// In-register DFT of R = 2^p <= 32 points (radix-2, decimation in time).
template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft_reg(cx<T> (&v)[R]) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = bitrev_c(i, R);
    if (j > i) {
      const cx<T> t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int k = 0; k < len / 2; ++k) {
        const cx<T> u = v[i + k];
        const cx<T> t = tw32<T, FWD>(v[i + k + len / 2], k * (32 / len));
        v[i + k] = cadd(u, t);
        v[i + k + len / 2] = csub(u, t);
      }
    }
  }
}

// In-register DFT of R points: radix-2 for powers of two, else the direct
// sum X[k] = sum_j v[j] W_R^(j k) with W_R^m = twN[m S] read from the
// transform's N-point table (S = N / R).
template <typename T, int R, bool FWD>
__device__ __forceinline__ void dft_any(cx<T> (&v)[R], const cx<T>* __restrict__ twN, int S) {
  if constexpr ((R & (R - 1)) == 0) {
    dft_reg<T, R, FWD>(v);
  } else {
    cx<T> out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      cx<T> acc = v[0];
#pragma unroll
      for (int j = 1; j < R; ++j) {
        cx<T> w = twN[((j * k) % R) * S];
        if (!FWD) w.im = -w.im;
        acc = cadd(acc, cmul(v[j], w));
      }
      out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
}

// Pass-1 twiddle v[k2] *= W_N^(a k2) (table of exp(-2 pi i q / N)).
template <typename T, int B, bool FWD>
__device__ __forceinline__ void twiddle_rows(cx<T> (&v)[B], int a, const cx<T>* __restrict__ twN) {
#pragma unroll
  for (int k2 = 1; k2 < B; ++k2) {
    cx<T> w = twN[a * k2];
    if (!FWD) w.im = -w.im;
    v[k2] = cmul(v[k2], w);
  }
}

template <int A, int B>
struct Fft2 {
  static constexpr int N = A * B;
  static constexpr int LSL = B * (A + 1) + 1;  // LDS entries per line
  static constexpr int G = (B + A - 1) / A;    // pass-2 columns per thread
};

// Strided complex transform (as k_fft_s) in two passes; LPW consecutive kx
// lines per workgroup, thread = l + LPW * a.
template <typename T, int A, int B, int LPW, bool FWD>
__global__ void __launch_bounds__(LPW * A) k_fft2_s(cx<T>* __restrict__ W, int nxh, long ls, long os,
                                                   const cx<T>* __restrict__ twN) {
  using F = Fft2<A, B>;
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int ntile = (nxh + LPW - 1) / LPW;
  const int tile = blockIdx.x % ntile, o = blockIdx.x / ntile;
  const int l = threadIdx.x % LPW, a = threadIdx.x / LPW;
  const int kx = tile * LPW + l;
  const bool ok = kx < nxh;
  const long base = (long)(ok ? kx : nxh - 1) + os * o;
  cx<T> v[B];
#pragma unroll
  for (int b = 0; b < B; ++b) v[b] = W[base + ls * (a + A * b)];
  dft_any<T, B, FWD>(v, twN, A);
  twiddle_rows<T, B, FWD>(v, a, twN);
  cx<T>* lb = buf + l * F::LSL;
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) lb[k2 * (A + 1) + a] = v[k2];
  __syncthreads();
#pragma unroll
  for (int g = 0; g < F::G; ++g) {
    const int k2 = a + A * g;
    if (F::G * A != B && k2 >= B) break;
    cx<T> u[A];
#pragma unroll
    for (int j = 0; j < A; ++j) u[j] = lb[k2 * (A + 1) + j];
    dft_any<T, A, FWD>(u, twN, B);
    if (ok) {
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) W[base + ls * (k2 + B * k1)] = u[k1];
    }
  }
}

// x forward (real lines of 2N -> spectrum kx = 0..N) in two passes; LPW
// lines per workgroup, thread = a + A * l (lanes along the line).
template <typename T, int A, int B, int LPW>
__global__ void __launch_bounds__(LPW * A) k_fft2_x_fwd(const T* __restrict__ in, cx<T>* __restrict__ W,
                                                       long nlines, const cx<T>* __restrict__ twN,
                                                       const cx<T>* __restrict__ rot, T in_scale) {
  using F = Fft2<A, B>;
  constexpr int N = F::N, nxh = N + 1;
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int a = threadIdx.x % A, l = threadIdx.x / A;
  const long L = (long)blockIdx.x * LPW + l;
  const bool ok = L < nlines;
  const T* src = in + (ok ? L : nlines - 1) * (2L * N);
  cx<T> v[B];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int n = a + A * b;
    v[b] = {src[2 * n] * in_scale, src[2 * n + 1] * in_scale};
  }
  dft_any<T, B, true>(v, twN, A);
  twiddle_rows<T, B, true>(v, a, twN);
  cx<T>* lb = buf + l * F::LSL;
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) lb[k2 * (A + 1) + a] = v[k2];
  __syncthreads();
  cx<T> z[F::G * A];  // z[g * A + k1] = Z[(a + A g) + B k1]
#pragma unroll
  for (int g = 0; g < F::G; ++g) {
    const int k2 = a + A * g;
    cx<T> u[A];
#pragma unroll
    for (int j = 0; j < A; ++j) u[j] = (k2 < B) ? lb[k2 * (A + 1) + j] : cx<T>{T(0), T(0)};
    dft_any<T, A, true>(u, twN, B);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) z[g * A + k1] = u[k1];
  }
  __syncthreads();  // pass-2 reads done: the line buffer takes Z in natural order
#pragma unroll
  for (int g = 0; g < F::G; ++g)
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1)
      if (a + A * g < B) lb[(a + A * g) + B * k1] = z[g * A + k1];
  __syncthreads();
  if (!ok) return;
  cx<T>* dst = W + L * nxh;
  for (int k = a; k < nxh; k += A) {
    const cx<T> zk = lb[k % N];
    cx<T> zn = lb[(N - k) % N];
    zn.im = -zn.im;  // conj(Z[N-k])
    const cx<T> e = cscale(cadd(zk, zn), T(0.5));
    const cx<T> d = csub(zk, zn);
    const cx<T> od = {d.im * T(0.5), -d.re * T(0.5)};  // (Z - conj)/(2i)
    dst[k] = cadd(e, cmul(rot[k], od));
  }
}

// x inverse (spectrum kx = 0..N -> real lines of 2N, times `scale`) in two
// passes; LPW lines per workgroup, thread = a + A * l.
template <typename T, int A, int B, int LPW>
__global__ void __launch_bounds__(LPW * A) k_fft2_x_inv(const cx<T>* __restrict__ W, T* __restrict__ out,
                                                       long nlines, const cx<T>* __restrict__ twN,
                                                       const cx<T>* __restrict__ rot, T scale) {
  using F = Fft2<A, B>;
  constexpr int N = F::N, nxh = N + 1;
  extern __shared__ unsigned char smem[];
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
  const int a = threadIdx.x % A, l = threadIdx.x / A;
  const long L = (long)blockIdx.x * LPW + l;
  const bool ok = L < nlines;
  const cx<T>* src = W + (ok ? L : nlines - 1) * nxh;
  cx<T>* lb = buf + l * F::LSL;
  for (int k = a; k < N; k += A) {
    const cx<T> xk = src[k];
    cx<T> xn = src[N - k];
    xn.im = -xn.im;  // conj(X[N-k])
    const cx<T> e = cadd(xk, xn);
    cx<T> r = rot[k];
    r.im = -r.im;  // exp(+2 pi i k / nx)
    const cx<T> od = cmul(csub(xk, xn), r);
    lb[k] = {e.re - od.im, e.im + od.re};  // e + i*o
  }
  __syncthreads();
  cx<T> v[B];
#pragma unroll
  for (int b = 0; b < B; ++b) v[b] = lb[a + A * b];
  dft_any<T, B, false>(v, twN, A);
  twiddle_rows<T, B, false>(v, a, twN);
  __syncthreads();  // every thread has read its pass-1 input
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) lb[k2 * (A + 1) + a] = v[k2];
  __syncthreads();
  T* dst = out + L * (2L * N);
#pragma unroll
  for (int g = 0; g < F::G; ++g) {
    const int k2 = a + A * g;
    if (F::G * A != B && k2 >= B) break;
    cx<T> u[A];
#pragma unroll
    for (int j = 0; j < A; ++j) u[j] = lb[k2 * (A + 1) + j];
    dft_any<T, A, false>(u, twN, B);
    if (ok) {
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) {
        const int n = k2 + B * k1;
        dst[2 * n] = u[k1].re * scale;
        dst[2 * n + 1] = u[k1].im * scale;
      }
    }
  }
}

// Launchers: the two-pass kernels for N in {36, 40, 50, 64, 72, 80, 90, 96,
// 100, 128, 144, 180, 192, 256, 512}, the general Stockham kernels otherwise.
// The lines per workgroup of a two-pass kernel (variant v: 4, 8 or 16) or the
// general kernel (v = 3) are chosen by timing the variants once per shape
// (chs_tune.h); untuned shapes use 16 (8 for N = 512).
template <typename T, int A, int B>
static size_t fft2_lds(int lpw) { return sizeof(cx<T>) * (size_t)lpw * Fft2<A, B>::LSL; }

static int fft2_default_variant(int n) { return n >= 512 ? 1 : 2; }

static std::string fft_key(const char* what, size_t tsize, int n, long nl, long no) {
  return std::string(what) + (tsize == 8 ? "/f64/" : "/f32/") + std::to_string(n) + "/" + std::to_string(nl) + "x" +
         std::to_string(no);
}

template <typename T, int A, int B, int LPW>
static bool fft2_s_launch(cx<T>* W, int nxh, long ls, int no, long os, const cx<T>* tw, T sgn, hipStream_t s) {
  dim3 g((unsigned)(((nxh + LPW - 1) / LPW) * no));
  const size_t sh = fft2_lds<T, A, B>(LPW);
  if (sgn < T(0))
    hipLaunchKernelGGL((k_fft2_s<T, A, B, LPW, true>), g, dim3(LPW * A), sh, s, W, nxh, ls, os, tw);
  else
    hipLaunchKernelGGL((k_fft2_s<T, A, B, LPW, false>), g, dim3(LPW * A), sh, s, W, nxh, ls, os, tw);
  return hipGetLastError() == hipSuccess;
}

template <typename T, int A, int B>
static bool fft2_s_lpw(int v, cx<T>* W, int nxh, long ls, int no, long os, const cx<T>* tw, T sgn, hipStream_t s) {
  switch (v) {
    case 0: return fft2_s_launch<T, A, B, 4>(W, nxh, ls, no, os, tw, sgn, s);
    case 1: return fft2_s_launch<T, A, B, 8>(W, nxh, ls, no, os, tw, sgn, s);
    case 2: return fft2_s_launch<T, A, B, 16>(W, nxh, ls, no, os, tw, sgn, s);
  }
  return false;
}

// Two-pass strided transform of length n with variant v; false when n is not
// a two-pass length.
template <typename T>
static bool fft2_s_variant(int v, cx<T>* W, int nxh, int n, long ls, int no, long os, const cx<T>* tw, T sgn,
                           hipStream_t s) {
  switch (n) {
    case 36: return fft2_s_lpw<T, 6, 6>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 40: return fft2_s_lpw<T, 5, 8>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 50: return fft2_s_lpw<T, 5, 10>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 64: return fft2_s_lpw<T, 8, 8>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 72: return fft2_s_lpw<T, 8, 9>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 80: return fft2_s_lpw<T, 8, 10>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 90: return fft2_s_lpw<T, 9, 10>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 96: return fft2_s_lpw<T, 8, 12>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 100: return fft2_s_lpw<T, 10, 10>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 128: return fft2_s_lpw<T, 8, 16>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 144: return fft2_s_lpw<T, 12, 12>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 180: return fft2_s_lpw<T, 12, 15>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 192: return fft2_s_lpw<T, 12, 16>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 256: return fft2_s_lpw<T, 16, 16>(v, W, nxh, ls, no, os, tw, sgn, s);
    case 512: return fft2_s_lpw<T, 16, 32>(v, W, nxh, ls, no, os, tw, sgn, s);
  }
  return false;
}

template <typename T, int A, int B, int LPW>
static bool fft2_x_launch(bool fwd, const T* in, cx<T>* W, T* out, long nlines, const cx<T>* tw, const cx<T>* rot,
                          T sc, hipStream_t s) {
  dim3 g((unsigned)((nlines + LPW - 1) / LPW));
  const size_t sh = fft2_lds<T, A, B>(LPW);
  if (fwd)
    hipLaunchKernelGGL((k_fft2_x_fwd<T, A, B, LPW>), g, dim3(LPW * A), sh, s, in, W, nlines, tw, rot, sc);
  else
    hipLaunchKernelGGL((k_fft2_x_inv<T, A, B, LPW>), g, dim3(LPW * A), sh, s, W, out, nlines, tw, rot, sc);
  return hipGetLastError() == hipSuccess;
}

template <typename T, int A, int B>
static bool fft2_x_lpw(int v, bool fwd, const T* in, cx<T>* W, T* out, long nlines, const cx<T>* tw,
                       const cx<T>* rot, T sc, hipStream_t s) {
  switch (v) {
    case 0: return fft2_x_launch<T, A, B, 4>(fwd, in, W, out, nlines, tw, rot, sc, s);
    case 1: return fft2_x_launch<T, A, B, 8>(fwd, in, W, out, nlines, tw, rot, sc, s);
    case 2: return fft2_x_launch<T, A, B, 16>(fwd, in, W, out, nlines, tw, rot, sc, s);
  }
  return false;
}

template <typename T>
static bool fft2_x_variant(int v, bool fwd, const T* in, cx<T>* W, T* out, int nx, long nlines, const cx<T>* tw,
                           const cx<T>* rot, T sc, hipStream_t s) {
  switch (nx / 2) {
    case 36: return fft2_x_lpw<T, 6, 6>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 40: return fft2_x_lpw<T, 5, 8>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 50: return fft2_x_lpw<T, 5, 10>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 64: return fft2_x_lpw<T, 8, 8>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 72: return fft2_x_lpw<T, 8, 9>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 80: return fft2_x_lpw<T, 8, 10>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 90: return fft2_x_lpw<T, 9, 10>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 96: return fft2_x_lpw<T, 8, 12>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 100: return fft2_x_lpw<T, 10, 10>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 128: return fft2_x_lpw<T, 8, 16>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 144: return fft2_x_lpw<T, 12, 12>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 180: return fft2_x_lpw<T, 12, 15>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 192: return fft2_x_lpw<T, 12, 16>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 256: return fft2_x_lpw<T, 16, 16>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
    case 512: return fft2_x_lpw<T, 16, 32>(v, fwd, in, W, out, nlines, tw, rot, sc, s);
  }
  return false;
}

// the general Stockham kernels (any 2,3,5-smooth length of one tile)
template <typename T>
static void fft_s_general(cx<T>* W, int nxh, int n, long ls, int no, long os, const Radices& rp, const cx<T>* tw,
                          T sgn, hipStream_t s) {
  const int lpb = kTileC / n;
  dim3 g((unsigned)(((nxh + lpb - 1) / lpb) * no));
  hipLaunchKernelGGL(k_fft_s<T>, g, dim3(kThreads), sizeof(cx<T>) * (size_t)lpb * n, s, W, nxh, n, ls, no, os, rp,
                     tw, sgn);
}

template <typename T>
static void fft_x_general(bool fwd, const T* in, cx<T>* W, T* out, int nx, long nlines, const Radices& rp,
                          const cx<T>* tw, const cx<T>* rot, T sc, hipStream_t s) {
  const int N = nx / 2;
  const int lpb = kTileC / N;
  dim3 g((unsigned)((nlines + lpb - 1) / lpb));
  const size_t sh = sizeof(cx<T>) * (size_t)lpb * N;
  if (fwd)
    hipLaunchKernelGGL(k_fft_x_fwd<T>, g, dim3(kThreads), sh, s, in, W, nx, nlines, rp, tw, rot, sc);
  else
    hipLaunchKernelGGL(k_fft_x_inv<T>, g, dim3(kThreads), sh, s, W, out, nx, nlines, rp, tw, rot, sc);
}

// variant 3 (timed with the two-pass ones) is the general kernel
template <typename T>
static void launch_fft_s(cx<T>* W, int nxh, int n, long ls, int no, long os, const Radices& rp, const cx<T>* tw,
                         T sgn, hipStream_t s) {
  int v = forced_or_tuned(fft_key(sgn < T(0) ? "fft_s+" : "fft_s-", sizeof(T), n, nxh, no), 4);
  if (v < 0) v = fft2_default_variant(n);
  if (v < 3 && fft2_s_variant<T>(v, W, nxh, n, ls, no, os, tw, sgn, s)) return;
  fft_s_general<T>(W, nxh, n, ls, no, os, rp, tw, sgn, s);
}

template <typename T>
static void launch_fft_x(bool fwd, const T* in, cx<T>* W, T* out, int nx, long nlines, const Radices& rp,
                         const cx<T>* tw, const cx<T>* rot, T sc, hipStream_t s) {
  const int N = nx / 2;
  int v = forced_or_tuned(fft_key(fwd ? "fft_x+" : "fft_x-", sizeof(T), N, nlines, 1), 4);
  if (v < 0) v = fft2_default_variant(N);
  if (v < 3 && fft2_x_variant<T>(v, fwd, in, W, out, nx, nlines, tw, rot, sc, s)) return;
  fft_x_general<T>(fwd, in, W, out, nx, nlines, rp, tw, rot, sc, s);
}

// Times the two-pass variants of every transform this solver launches (on
// `tmp`, an array as large as the work array, and the work array itself;
// their contents are overwritten by every solve).
template <typename T>
static void tune_transforms(int nx, int ny, int nz, int nxh, bool ydir, cx<T>* work, cx<T>* tmp, const cx<T>* twx,
                            const cx<T>* twy, const cx<T>* twz, const cx<T>* rot, const Radices& rx,
                            const Radices& ry, const Radices& rz) {
  hipStream_t s = nullptr;
  const long nlx = (long)ny * nz;
  T* real = reinterpret_cast<T*>(tmp);
  auto ok = [] { return hipGetLastError() == hipSuccess; };
  for (int fwd = 1; fwd >= 0; --fwd) {
    tune(fft_key(fwd ? "fft_x+" : "fft_x-", sizeof(T), nx / 2, nlx, 1), 4, s, [&](int v) {
      if (v == 3) {
        fft_x_general<T>(fwd != 0, real, work, real, nx, nlx, rx, twx, rot, T(1), s);
        return ok();
      }
      return fft2_x_variant<T>(v, fwd != 0, real, work, real, nx, nlx, twx, rot, T(1), s);
    });
    const T sgn = fwd ? T(-1) : T(1);
    tune(fft_key(fwd ? "fft_s+" : "fft_s-", sizeof(T), nz, nxh, ny), 4, s, [&](int v) {
      if (v == 3) {
        fft_s_general<T>(work, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, sgn, s);
        return ok();
      }
      return fft2_s_variant<T>(v, work, nxh, nz, (long)nxh * ny, ny, (long)nxh, twz, sgn, s);
    });
    if (ydir)
      tune(fft_key(fwd ? "fft_s+" : "fft_s-", sizeof(T), ny, nxh, nz), 4, s, [&](int v) {
        if (v == 3) {
          fft_s_general<T>(work, nxh, ny, (long)nxh, nz, (long)nxh * ny, ry, twy, sgn, s);
          return ok();
        }
        return fft2_s_variant<T>(v, work, nxh, ny, (long)nxh, nz, (long)nxh * ny, twy, sgn, s);
      });
  }
}

// Triply periodic solve in spectral space: phi = -rhs / (kx'^2 + ky'^2 + kz'^2)
// with the compact schemes' modified wavenumbers; the mean mode is set to 0.
template <typename T>
__global__ void k_div_k2(cx<T>* __restrict__ W, int nxh, int ny, int nz, const T* __restrict__ k2x,
                         const T* __restrict__ k2y, const T* __restrict__ k2z) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)nxh * ny * nz;
  if (i >= n) return;
  const int kx = (int)(i % nxh);
  const long r = i / nxh;
  const int ky = (int)(r % ny), kz = (int)(r / ny);
  const T kk = k2x[kx] + k2y[ky] + k2z[kz];
  cx<T> v = W[i];
  if (kk < T(1e-16)) {
    v.re = T(0);
    v.im = T(0);
  } else {
    v.re = v.re / (-kk);
    v.im = v.im / (-kk);
  }
  W[i] = v;
}

// y tridiagonal per (kx, kz) with host-built TRID0 factors.
template <typename T>
__global__ void __launch_bounds__(64) k_ytdma(cx<T>* __restrict__ W, int nxh, int ny, int nz, const double* __restrict__ ya,
                                             const T* __restrict__ fd, const T* __restrict__ fz) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nm = (long)nxh * nz;
  if (m >= nm) return;
  const long kx = m % nxh, kz = m / nxh;
  const long js = nxh;                 // stride along j in W
  const long base = kx + (long)nxh * ny * kz;
  const long fs = nm;                  // stride along j in the factor tables
  cx<T> t = W[base];
  T z = fz[m];
  t = cscale(t, z);
  W[base] = t;
  for (int j = 1; j < ny - 1; ++j) {
    const T aj = T(ya[j]);
    cx<T> v = W[base + js * j];
    z = fz[m + fs * j];
    t = cscale(csub(v, cscale(t, aj)), z);
    W[base + js * j] = t;
  }
  {
    const int j = ny - 1;
    const T aj = T(ya[j]);
    const T zl = fz[m + fs * j];  // the pivot itself for the last row
    cx<T> v = W[base + js * j];
    if (fabs((double)zl) > 1.0e-10) {
      v = csub(v, cscale(t, aj));
      t = {v.re / zl, v.im / zl};
    } else {
      t = {T(0), T(0)};
    }
    W[base + js * j] = t;
  }
  for (int j = ny - 2; j >= 0; --j) {
    cx<T> v = W[base + js * j];
    const T dj = fd[m + fs * j];
    t = csub(v, cscale(t, dj));
    W[base + js * j] = t;
  }
}

// Wall-normal mixing of the spectrum by a dense ny x ny matrix:
//   out(i, g) = s(i, g) * sum_j M(i, j) in(j, g)
// for every real column g = (kz, c) of the complex spectrum (c = 2 kx + re/im),
// with s = 1/(lam_i + lx(kx) + lz(kz)) when lam is given (0 for the constant
// mode of (kx, kz) = (0, 0)) and s = 1 otherwise.  A workgroup computes a
// 128-row x 32-column tile; the j loop runs over 16-wide slabs of M and of the
// input staged in LDS, and each thread accumulates a 4 x 4 block.
constexpr int kMixR = 128, kMixC = 32, kMixK = 16;

template <typename T>
__global__ void __launch_bounds__(256) k_ymix(const T* __restrict__ M, const T* __restrict__ in, T* __restrict__ out,
                                              int ny, int ncr, long ncol, const T* __restrict__ lam,
                                              const T* __restrict__ lx, const T* __restrict__ lz, int null_idx) {
  __shared__ T As[kMixK][kMixR + 1];  // +1: the transposed stores of M hit distinct banks
  __shared__ T Bs[kMixK][kMixC];
  const int tid = threadIdx.x;
  const int tx = tid & 7, ty = tid >> 3;
  const int r0 = blockIdx.y * kMixR;
  const long g0 = (long)blockIdx.x * kMixC;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
  for (int k0 = 0; k0 < ny; k0 += kMixK) {
    for (int e = tid; e < kMixR * kMixK; e += 256) {
      const int r = e / kMixK, k = e - r * kMixK;
      const int gi = r0 + r, gk = k0 + k;
      As[k][r] = (gi < ny && gk < ny) ? M[(long)gi * ny + gk] : T(0);
    }
    for (int e = tid; e < kMixK * kMixC; e += 256) {
      const int k = e / kMixC, c = e - k * kMixC;
      const long g = g0 + c;
      const int gk = k0 + k;
      T v = T(0);
      if (g < ncol && gk < ny) {
        const long kz = g / ncr, cc = g - kz * ncr;
        v = in[cc + (long)ncr * (gk + (long)ny * kz)];
      }
      Bs[k][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMixK; ++k) {
      T a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = As[k][ty * 4 + q];
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = Bs[k][tx * 4 + q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = acc[p][q] + a[p] * b[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int i = r0 + ty * 4 + p;
    if (i >= ny) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long g = g0 + tx * 4 + q;
      if (g >= ncol) continue;
      const long kz = g / ncr, cc = g - kz * ncr;
      const long kx = cc >> 1;
      T v = acc[p][q];
      if (lam) {
        if (i == null_idx && kx == 0 && kz == 0) v = T(0);
        else v = v / (lam[i] + lx[kx] + lz[kz]);
      }
      out[cc + (long)ncr * (i + (long)ny * kz)] = v;
    }
  }
}

// Mirror-symmetric form of the wall-normal mixing (ny even, h = ny/2).
// STAGE 1: fold each column into r_e(j) = r(j) + r(ny-1-j) and
//          r_o(j) = r(j) - r(ny-1-j), multiply by the block inverses
//          Vie / Vio and scale by 1/(lam + lx + lz); the result is stored as
//          rows [0, h) even, [h, ny) odd.
// STAGE 2: multiply the even / odd rows by Ve / Vo and unfold:
//          phi(j) = (e(j) + o(j))/2, phi(ny-1-j) = (e(j) - o(j))/2.
// A workgroup covers 64 rows of each half (row tile t) and 32 columns; each
// thread accumulates a 4 x 4 block of one half.
constexpr int kSymR = 64;

template <typename T, int STAGE>
__global__ void __launch_bounds__(256) k_ymix_sym(const T* __restrict__ Me, const T* __restrict__ Mo,
                                                  const T* __restrict__ in, T* __restrict__ out, int ny, int ncr,
                                                  long ncol, const T* __restrict__ lame, const T* __restrict__ lamo,
                                                  const T* __restrict__ lx, const T* __restrict__ lz, int null_even) {
  constexpr int AS = kMixK * (2 * kSymR + 1);
  constexpr int BS = 2 * kMixK * kMixC;
  constexpr int CS = 2 * kSymR * (kMixC + 1);
  constexpr int SM = (AS + BS > CS) ? AS + BS : CS;
  __shared__ T smem[SM];
  T* As = smem;        // [k][2*kSymR + 1]
  T* Bs = smem + AS;   // [half][k][kMixC]
  const int h = ny / 2;
  const int tid = threadIdx.x;
  const int tx = tid & 7, ty = tid >> 3;
  const int hf = ty >> 4;            // 0 even, 1 odd
  const int rl = (ty & 15) * 4;      // first local row of this thread
  const int j0 = blockIdx.y * kSymR;  // first row of the tile within a half
  const long g0 = (long)blockIdx.x * kMixC;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
  for (int k0 = 0; k0 < h; k0 += kMixK) {
    for (int e = tid; e < 2 * kSymR * kMixK; e += 256) {
      const int r = e / kMixK, k = e - r * kMixK;
      const int half = r / kSymR, gi = j0 + (r - half * kSymR), gk = k0 + k;
      const T* M = half ? Mo : Me;
      As[k * (2 * kSymR + 1) + r] = (gi < h && gk < h) ? M[(long)gi * h + gk] : T(0);
    }
    for (int e = tid; e < 2 * kMixK * kMixC; e += 256) {
      const int half = e / (kMixK * kMixC), rem = e - half * (kMixK * kMixC);
      const int k = rem / kMixC, c = rem - k * kMixC;
      const long g = g0 + c;
      const int gk = k0 + k;
      T v = T(0);
      if (g < ncol && gk < h) {
        const long kz = g / ncr, cc = g - kz * ncr;
        const long base = cc + (long)ncr * (long)ny * kz;
        if (STAGE == 1) {
          const T a = in[base + (long)ncr * gk], b = in[base + (long)ncr * (ny - 1 - gk)];
          v = half ? a - b : a + b;
        } else {
          v = in[base + (long)ncr * (half * h + gk)];
        }
      }
      Bs[(half * kMixK + k) * kMixC + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMixK; ++k) {
      T a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = As[k * (2 * kSymR + 1) + hf * kSymR + rl + q];
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = Bs[(hf * kMixK + k) * kMixC + tx * 4 + q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = acc[p][q] + a[p] * b[q];
    }
    __syncthreads();
  }
  if (STAGE == 1) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int i = j0 + rl + p;
      if (i >= h) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const long g = g0 + tx * 4 + q;
        if (g >= ncol) continue;
        const long kz = g / ncr, cc = g - kz * ncr;
        const long kx = cc >> 1;
        T v = acc[p][q];
        if (hf == 0 && i == null_even && kx == 0 && kz == 0) v = T(0);
        else v = v / ((hf ? lamo[i] : lame[i]) + lx[kx] + lz[kz]);
        out[cc + (long)ncr * ((long)(hf * h + i) + (long)ny * kz)] = v;
      }
    }
  } else {
    T* Cs = smem;  // [half][row][kMixC + 1]; the K loop ended with a barrier
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) Cs[(hf * kSymR + rl + p) * (kMixC + 1) + tx * 4 + q] = acc[p][q];
    __syncthreads();
    for (int e = tid; e < kSymR * kMixC; e += 256) {
      const int r = e / kMixC, c = e - r * kMixC;
      const int j = j0 + r;
      const long g = g0 + c;
      if (j >= h || g >= ncol) continue;
      const long kz = g / ncr, cc = g - kz * ncr;
      const T ev = Cs[r * (kMixC + 1) + c], od = Cs[(kSymR + r) * (kMixC + 1) + c];
      const long base = cc + (long)ncr * (long)ny * kz;
      out[base + (long)ncr * j] = T(0.5) * (ev + od);
      out[base + (long)ncr * (ny - 1 - j)] = T(0.5) * (ev - od);
    }
  }
}

// Both stages of the mirror-symmetric mixing in one pass (h = ny/2 <= 64):
// a workgroup holds the folded even/odd halves of 64 columns in LDS, applies
// Vie / Vio, the scaling 1/(lam + lx + lz) (null mode set to 0), then Ve / Vo,
// and unfolds into the output, so the spectrum is read and written once.
// The 64 x 64 blocks stream through LDS in K-tiles of 16 (the next tile
// prefetched into registers); each thread accumulates 8 rows x 4 columns of
// one half.
constexpr int kFuC = 64, kFuK = 16, kFuH = 64;

template <typename T>
__global__ void __launch_bounds__(256) k_ymix_fused(const T* __restrict__ Vie, const T* __restrict__ Vio,
                                                   const T* __restrict__ Ve, const T* __restrict__ Vo,
                                                   const T* in, T* out, int ny, int ncr,
                                                   long ncol, const T* __restrict__ lame, const T* __restrict__ lamo,
                                                   const T* __restrict__ lx, const T* __restrict__ lz, int null_even) {
  // in and out may be the same array: a workgroup reads all of its columns
  // before it writes any of them
  constexpr int BW = kFuC + 1;        // row stride of the column block (odd: transposed writes spread)
  constexpr int AW = 2 * kFuH + 4;    // row stride of the matrix tile
  extern __shared__ unsigned char smem_raw[];
  T* Bs = reinterpret_cast<T*>(smem_raw);  // [half][row][BW]
  T* As = Bs + 2 * kFuH * BW;               // [k][AW]: rows 0..63 even block, 64..127 odd block
  const int h = ny / 2;
  const int tid = threadIdx.x;
  const int hf = tid >> 7, rg = (tid >> 4) & 7, cg = tid & 15;
  const int r0 = rg * 8, c0 = cg * 4;
  const long g0 = (long)blockIdx.x * kFuC;
  // ---- fold the 64 columns: e(j) = x(j) + x(ny-1-j), o(j) = x(j) - x(ny-1-j)
  {
    const int c = tid & 63;
    const long g = g0 + c;
    const bool cok = g < ncol;
    const long gc = cok ? g : ncol - 1;
    const long kz = gc / ncr, cc = gc - kz * ncr;
    const long base = cc + (long)ncr * (long)ny * kz;
    constexpr int NJ = kFuH / 4;
    T a[NJ], b[NJ];
#pragma unroll
    for (int u = 0; u < NJ; ++u) {
      const int j = (tid >> 6) + 4 * u;
      const int jc = j < h ? j : h - 1;
      a[u] = in[base + (long)ncr * jc];
      b[u] = in[base + (long)ncr * (ny - 1 - jc)];
    }
#pragma unroll
    for (int u = 0; u < NJ; ++u) {
      const int j = (tid >> 6) + 4 * u;
      const bool ok = cok && j < h;
      Bs[(0 * kFuH + j) * BW + c] = ok ? a[u] + b[u] : T(0);
      Bs[(1 * kFuH + j) * BW + c] = ok ? a[u] - b[u] : T(0);
    }
  }
  // one K-tile of the two blocks: As[k][row] = M_half[row][k0 + k]
  auto load_tile = [&](const T* Me, const T* Mo, int k0, T (&v)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;  // 0 .. 2047 = 128 rows x 16 k
      const int row = e >> 4, k = e & 15;
      const int half = row >> 6, i = row & 63, gk = k0 + k;
      const T* M = half ? Mo : Me;
      v[u] = (i < h && gk < h) ? M[(long)i * h + gk] : T(0);
    }
  };
  auto store_tile = [&](const T (&v)[8]) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      const int row = e >> 4, k = e & 15;
      As[k * AW + row] = v[u];
    }
  };
  auto matmul = [&](const T* Me, const T* Mo, T (&acc)[8][4]) {
#pragma unroll
    for (int p = 0; p < 8; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][q] = T(0);
    T v[8];
    load_tile(Me, Mo, 0, v);
    for (int k0 = 0; k0 < h; k0 += kFuK) {
      __syncthreads();  // previous tile consumed (and, first time, the column block written)
      store_tile(v);
      __syncthreads();
      if (k0 + kFuK < h) load_tile(Me, Mo, k0 + kFuK, v);
#pragma unroll
      for (int k = 0; k < kFuK; ++k) {
        T a[8], b[4];
#pragma unroll
        for (int p = 0; p < 8; ++p) a[p] = As[k * AW + hf * kFuH + r0 + p];
#pragma unroll
        for (int q = 0; q < 4; ++q) b[q] = Bs[(hf * kFuH + k0 + k) * BW + c0 + q];
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[p][q] = acc[p][q] + a[p] * b[q];
      }
    }
    __syncthreads();  // every thread has read the column block
  };
  T acc[8][4];
  // ---- stage 1: Vi, then the eigenvalue scaling
  matmul(Vie, Vio, acc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long g = g0 + c0 + q;
    const long gc = g < ncol ? g : ncol - 1;
    const long kz = gc / ncr, kx = (gc - kz * ncr) >> 1;
    const T shift = lx[kx] + lz[kz];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int i = r0 + p;
      T y = T(0);
      if (i < h) {
        const bool null_mode = (hf == 0 && i == null_even && kx == 0 && kz == 0);
        y = null_mode ? T(0) : acc[p][q] / ((hf ? lamo[i] : lame[i]) + shift);
      }
      Bs[(hf * kFuH + i) * BW + c0 + q] = y;
    }
  }
  // ---- stage 2: V (matmul's first barrier orders these writes before use)
  matmul(Ve, Vo, acc);
#pragma unroll
  for (int p = 0; p < 8; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) Bs[(hf * kFuH + r0 + p) * BW + c0 + q] = acc[p][q];
  __syncthreads();
  // ---- unfold: phi(j) = (e(j) + o(j))/2, phi(ny-1-j) = (e(j) - o(j))/2
  {
    const int c = tid & 63;
    const long g = g0 + c;
    if (g < ncol) {
      const long kz = g / ncr, cc = g - kz * ncr;
      const long base = cc + (long)ncr * (long)ny * kz;
#pragma unroll
      for (int u = 0; u < kFuH / 4; ++u) {
        const int j = (tid >> 6) + 4 * u;
        if (j >= h) break;
        const T ev = Bs[j * BW + c], od = Bs[(kFuH + j) * BW + c];
        out[base + (long)ncr * j] = T(0.5) * (ev + od);
        out[base + (long)ncr * (ny - 1 - j)] = T(0.5) * (ev - od);
      }
    }
  }
}

// FP64 mirror-symmetric mixing on the matrix cores (h = ny/2 <= 64), both
// stages in one pass.  v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16]
// and B[l/16][l%16], and register v of lane l holds D[4v + l/16][l%16].
// Workgroup = 8 waves; wave w owns rows 16*(w%4) .. +16 of half w/4 (even,
// odd) and reads that row block of Vi (then of V) from L2 one K-step ahead
// of its products, so that the kernel fits 128 registers and two workgroups
// share a CU (one's loads and barriers overlap the other's products; holding
// the whole row block in registers allowed one workgroup per CU and was
// slower).  Per tile of 64 columns: fold into LDS, D = Vi B (16 K-steps x 4
// column blocks), scale by 1/(lam + lx + lz), back to LDS, D = V Y, back to
// LDS, unfold into the output (in place: a tile is read before it is
// written, and no other workgroup touches it).
typedef double chs_d4 __attribute__((ext_vector_type(4)));

// a / d for d finite, normal and non-zero: the hardware reciprocal (v_rcp_f64)
// refined by two Newton steps, the quotient then corrected once by its
// residual (within an ulp of the rounded quotient); 9 instructions where a
// written divide compiles to the IEEE sequence of about 15 FP64 ones
__device__ __forceinline__ double div_nr(double a, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  const double y = a * r;
  return fma(fma(-d, y, a), r, y);
}
constexpr int kMmC = 64;            // columns per tile
constexpr int kMmW = kMmC + 1;      // LDS row stride
constexpr int kMmWaves = 8;

__global__ void __launch_bounds__(64 * kMmWaves, 4) k_ymix_mfma(const double* __restrict__ Vie,
                                                             const double* __restrict__ Vio,
                                                             const double* __restrict__ Ve,
                                                             const double* __restrict__ Vo, double* W, int ny,
                                                             int ncr, long ncol, const double* __restrict__ lame,
                                                             const double* __restrict__ lamo,
                                                             const double* __restrict__ lx,
                                                             const double* __restrict__ lz, int null_even) {
  __shared__ double Bs[2 * 64 * kMmW];  // [half][row][column]
  const int h = ny / 2;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int half = w >> 2, rb = w & 3;
  const int li = lane & 15, lk = lane >> 4;
  // D += M[16 rb + li][4 ks + lk] B[4 ks + lk][column] over the 16 K-steps
  // and the 4 column blocks, M read from L2 (zero outside h x h)
  const int ia = 16 * rb + li;
  auto products = [&](const double* M, chs_d4 (&acc)[4]) {
    const double* arow = M + (long)(ia < h ? ia : 0) * h + lk;
#pragma unroll 4
    for (int ks = 0; ks < 16; ++ks) {
      const double av = (ia < h && 4 * ks + lk < h) ? arow[4 * ks] : 0.0;
      const double* brow = Bs + half * 64 * kMmW + (4 * ks + lk) * kMmW + li;
      double bv[4];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) bv[cb] = brow[16 * cb];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[cb], acc[cb], 0, 0, 0);
    }
  };
  const double* Mi = half ? Vio : Vie;
  const double* Mv = half ? Vo : Ve;
  const double* lam = half ? lamo : lame;
  double* Bh = Bs + half * 64 * kMmW;
  const long ntile = (ncol + kMmC - 1) / kMmC;
  // the fold's inputs: rows j and ny-1-j of column g0 + tid % 64, j = tid / 64
  // + 8u (clamped to valid rows and columns, so every address is in range)
  double pa[8], pb[8];
  auto fold_load = [&](long g0n) {
    const long g = g0n + (tid & 63);
    const long gc = g < ncol ? g : ncol - 1;
    const long kz = gc / ncr, cc = gc - kz * ncr;
    const long base = cc + (long)ncr * (long)ny * kz;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = (tid >> 6) + 8 * u;
      const int jc = j < h ? j : h - 1;
      pa[u] = W[base + (long)ncr * jc];
      pb[u] = W[base + (long)ncr * (ny - 1 - jc)];
    }
  };
  for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long g0 = tile * kMmC;
    fold_load(g0);
    // ---- fold: 64 columns x 64 row pairs, 8 per thread
    {
      const int c = tid & 63;
      const bool cok = g0 + c < ncol;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = (tid >> 6) + 8 * u;
        const bool ok = cok && j < h;
        Bs[j * kMmW + c] = ok ? pa[u] + pb[u] : 0.0;
        Bs[(64 + j) * kMmW + c] = ok ? pa[u] - pb[u] : 0.0;
      }
    }
    __syncthreads();
    // ---- D = Vi B over this wave's 16 rows and the 64 columns
    chs_d4 acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = chs_d4{0.0, 0.0, 0.0, 0.0};
    products(Mi, acc);
    __syncthreads();  // every wave has read the folded block
    // ---- scale by 1/(lam + lx + lz) (null mode to 0) and store Y
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int c = 16 * cb + li;
      const long g = g0 + c;
      const long gc = g < ncol ? g : ncol - 1;
      const long kz = gc / ncr, kx = (gc - kz * ncr) >> 1;
      const double shift = lx[kx] + lz[kz];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * rb + 4 * v + lk;
        double y = 0.0;
        if (i < h) {
          const bool null_mode = (half == 0 && i == null_even && kx == 0 && kz == 0);
          y = null_mode ? 0.0 : div_nr(acc[cb][v], lam[i] + shift);
        }
        Bh[i * kMmW + c] = y;
      }
    }
    __syncthreads();
    // ---- D = V Y
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb] = chs_d4{0.0, 0.0, 0.0, 0.0};
    products(Mv, acc);
    __syncthreads();  // every wave has read Y
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) Bh[(16 * rb + 4 * v + lk) * kMmW + 16 * cb + li] = acc[cb][v];
    __syncthreads();
    // ---- unfold: phi(j) = (e(j) + o(j))/2, phi(ny-1-j) = (e(j) - o(j))/2
    {
      const int c = tid & 63;
      const long g = g0 + c;
      if (g < ncol) {
        const long kz = g / ncr, cc = g - kz * ncr;
        const long base = cc + (long)ncr * (long)ny * kz;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = (tid >> 6) + 8 * u;
          if (j < h) {
            const double ev = Bs[j * kMmW + c], od = Bs[(64 + j) * kMmW + c];
            W[base + (long)ncr * j] = 0.5 * (ev + od);
            W[base + (long)ncr * (ny - 1 - j)] = 0.5 * (ev - od);
          }
        }
      }
    }
    __syncthreads();  // the next tile's fold overwrites the block
  }
}

Poisson::~Poisson() {
  for (void* p : {(void*)tw_x64, (void*)tw_x32, (void*)tw_z64, (void*)tw_z32, (void*)rot64, (void*)rot32,
                  (void*)ya64, (void*)yc64, (void*)ya32, (void*)yc32, (void*)fd64, (void*)fz64, (void*)fd32,
                  (void*)fz32, work, (void*)V64, (void*)Vi64, (void*)lam64, (void*)lx64, (void*)lz64,
                  (void*)V32, (void*)Vi32, (void*)lam32, (void*)lx32, (void*)lz32, work2, (void*)Ve64,
                  (void*)Vo64, (void*)Vie64, (void*)Vio64, (void*)lame64, (void*)lamo64, (void*)Ve32, (void*)Vo32,
                  (void*)Vie32, (void*)Vio32, (void*)lame32, (void*)lamo32, (void*)tw_y64, (void*)tw_y32,
                  (void*)ly64, (void*)ly32})
    if (p) (void)hipFree(p);
}

template <typename T>
static std::string upv(const std::vector<double>& v, T** d) {
  std::vector<T> t(v.begin(), v.end());
  PCHK(hipMalloc(d, sizeof(T) * t.size()));
  PCHK(hipMemcpy(*d, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice));
  return "";
}

std::string Poisson::create(const Mesh& m, int kind, int iacc, std::unique_ptr<Poisson>* out) {
  if (kind != 2 && kind != 3) return "poisson: unknown kind";
  if (kind == 3 && !m.periodic[1]) return create_010(m, iacc, out);
  if (kind == 3) return create_3d(m, iacc, out);
  std::unique_ptr<Poisson> p(new Poisson());
  p->kind = 2;
  p->nx = m.nc[0]; p->ny = m.nc[1]; p->nz = m.nc[2];
  if (p->nx % 2) return "poisson: nx must be even";
  p->nxh = p->nx / 2 + 1;
  p->rad_x = radix_plan(p->nx / 2);
  p->rad_z = radix_plan(p->nz);
  if (p->rad_x.empty() || p->rad_z.empty()) return "poisson: nx/2 and nz must be 2,3,5-smooth";
  if (p->rad_x.size() > 12 || p->rad_z.size() > 12) return "poisson: transform too long";
  if (p->nx / 2 > kTileC || p->nz > kTileC) return "poisson: transform too long for one tile";
  if (p->ny < 3) return "poisson: ny too small";
  const double pi = 2.0 * std::asin(1.0);
  // twiddles exp(-2 pi i m / Nc)
  auto twid = [&](int Nc) {
    std::vector<double> t(2 * Nc);
    for (int q = 0; q < Nc; ++q) {
      double a = 2.0 * pi * double(q) / double(Nc);
      t[2 * q] = std::cos(a);
      t[2 * q + 1] = -std::sin(a);
    }
    return t;
  };
  std::vector<double> twx = twid(p->nx / 2), twz = twid(p->nz), rot(2 * p->nxh);
  for (int k = 0; k < p->nxh; ++k) {
    double a = 2.0 * pi * double(k) / double(p->nx);
    rot[2 * k] = std::cos(a);
    rot[2 * k + 1] = -std::sin(a);
  }
  std::string e;
  if (!(e = upv(twx, &p->tw_x64)).empty() || !(e = upv(twx, &p->tw_x32)).empty()) return e;
  if (!(e = upv(twz, &p->tw_z64)).empty() || !(e = upv(twz, &p->tw_z32)).empty()) return e;
  if (!(e = upv(rot, &p->rot64)).empty() || !(e = upv(rot, &p->rot32)).empty()) return e;
  // eigenvalues (fishpack_root_1D, LP = 1): 0, -4/h^2 sin^2(pi m/N) pairs, -4/h^2
  auto roots = [&](int L, double C1) {
    std::vector<double> rt(L, 0.0);
    double dx = pi / (2.0 * double(L));
    rt[0] = 0.0;
    rt[L - 1] = -4.0 * C1;
    for (int i = 3; i <= L; i += 2) {
      double s = std::sin(double(i - 1) * dx);
      rt[i - 2] = -4.0 * C1 * (s * s);
      rt[i - 1] = rt[i - 2];
    }
    return rt;
  };
  std::vector<double> xrt = roots(p->nx, m.h2r[0]), zrt = roots(p->nz, m.h2r[2]);
  // complex mode k of a length-L transform -> FFTPACK root index
  auto mode_root = [](const std::vector<double>& rt, int L, int k) {
    int mm = k <= L / 2 ? k : L - k;
    if (mm == 0) return rt[0];
    if (2 * mm == L) return rt[L - 1];
    return rt[2 * mm - 1];
  };
  // y operator (fishpack_fft_init, :312-323; Cartesian: rp = rc = 1)
  const int ny = p->ny;
  std::vector<double> a(ny), b(ny), c(ny);
  for (int j = 0; j < ny; ++j) {
    a[j] = m.h2r[1] * m.ymc[j] * m.ymp[j] * 1.0 * 1.0;
    c[j] = m.h2r[1] * m.ymc[j] * m.ymp[j + 1] * 1.0 * 1.0;
  }
  for (int j = 0; j < ny; ++j) b[j] = -(a[j] + c[j]);
  if (!m.periodic[1]) {
    b[0] = b[0] + a[0];
    a[0] = 0.0;
    b[ny - 1] = b[ny - 1] + c[ny - 1];
    c[ny - 1] = 0.0;
  } else {
    return "poisson: periodic y needs the 3-D FFT path";
  }
  if (!(e = upv(a, &p->ya64)).empty() || !(e = upv(c, &p->yc64)).empty()) return e;
  // TRID0 factors per mode: fz[j] = 1/(bb_j - a_j d_{j-1}) for j < ny-1 and the
  // raw last pivot at j = ny-1; fd[j] = c_j * z_j.
  const long nm = (long)p->nxh * p->nz;
  std::vector<double> fd((size_t)nm * ny), fz((size_t)nm * ny);
  std::vector<double> bb(ny), d(ny);
  for (int kz = 0; kz < p->nz; ++kz)
    for (int kx = 0; kx < p->nxh; ++kx) {
      const double xr = mode_root(xrt, p->nx, kx), zr = mode_root(zrt, p->nz, kz);
      for (int j = 0; j < ny; ++j) bb[j] = b[j] + xr / (1.0 * 1.0) + zr;
      const long mi = kx + (long)p->nxh * kz;
      double z = 1.0 / bb[0];
      d[0] = c[0] * z;
      fz[mi] = z;
      fd[mi] = d[0];
      for (int j = 1; j < ny - 1; ++j) {
        z = 1.0 / (bb[j] - a[j] * d[j - 1]);
        d[j] = c[j] * z;
        fz[mi + nm * j] = z;
        fd[mi + nm * j] = d[j];
      }
      z = bb[ny - 1] - a[ny - 1] * d[ny - 2];
      fz[mi + nm * (ny - 1)] = z;
      fd[mi + nm * (ny - 1)] = 0.0;
    }
  if (!(e = upv(fd, &p->fd64)).empty() || !(e = upv(fz, &p->fz64)).empty()) return e;
  if (!(e = upv(fd, &p->fd32)).empty() || !(e = upv(fz, &p->fz32)).empty()) return e;
  p->work_bytes = sizeof(double) * 2 * (size_t)p->nxh * ny * p->nz;
  PCHK(hipMalloc(&p->work, p->work_bytes));
  if (!(e = p->tune(false)).empty()) return e;
  *out = std::move(p);
  return "";
}

// Squared modified wavenumber of the staggered first derivative of accuracy
// iacc (interior periodic coefficients alpha, a, b per unit spacing) at
// phase w on n cells over length len (waves, src/poisson_1stderivcomp.f90:
// 2405-2480):  (n/len * (2a sin(w/2) + 2b sin(3w/2)) / (1 + 2 alpha cos w))^2.
static double mod_k2(int iacc, double w, int n, double len) {
  const double al = coeff(F_D1C2P, false, 3, 1, BC_PERIODIC, iacc);
  const double ar = coeff(F_D1C2P, true, 3, 1, BC_PERIODIC, iacc);
  const double br = coeff(F_D1C2P, true, 3, 2, BC_PERIODIC, iacc);
  const double wp = (2.0 * ar * std::sin(w * 0.5) + 2.0 * br * std::sin(1.5 * w)) / (1.0 + 2.0 * al * std::cos(w));
  const double kk = double(n) * wp / len;
  return kk * kk;
}

// Channel path of the 3-D FFT library (poisson_010,
// src/poisson_1stderivcomp.f90:1457-1816, istret = 0: x and z periodic, y
// walls on a uniform mesh).  The reference reorders y, transforms in 3-D,
// turns the y transform into a cosine transform by its post-processing
// (:1557-1578), divides by -(kx'^2 + ky'^2 + kz'^2) with the walled
// wavenumbers ky' at w = pi j / ny (:2465-2479), and undoes every step; its
// half-cell rotations in x and z are applied before and undone after a
// real division, so they cancel.  Per (kx, kz) mode that is
//   phi = C^-1 diag(1 / -(kx'^2 + ky'^2 + kz'^2)) C r
// with C the cosine transform (DCT-II) along y, which is the exact solve's
// pipeline (FFT x, FFT z, two dense products along y, inverse FFTs) with
// the cosine basis in place of the eigenvectors of D_y G_y:
//   V[j][k] = cos(pi k (2j+1) / (2 ny)), Vi = diag(1/ny, 2/ny, ...) V^T,
// eigenvalue -ky'^2(k), and the same split into even (k even) and odd
// (k odd) halves, whose first-half blocks are inverted by their own
// orthogonality (sum over j < ny/2 of a product = ny/4, ny/2 for k = 0).
// The (0, 0, 0) mode is set to zero as the reference does for kxyz < 1e-16.
std::string Poisson::create_010(const Mesh& m, int iacc, std::unique_ptr<Poisson>* out) {
  if (!m.periodic[0] || !m.periodic[2]) return "poisson: the channel 3-D FFT path needs periodic x and z";
  if (m.stretched)
    return "poisson: the 3-D FFT path on a stretched channel (3FMD mapping with the pentadiagonal y solve) "
           "is not built";
  const int ny = m.nc[1];
  if (ny % 2) return "poisson: the channel 3-D FFT path needs an even ny (the reference's y reorder)";
  std::unique_ptr<Poisson> p;
  std::string e = create(m, 2, iacc, &p);  // the FFT(x, z) tables of the FISHPACK path
  if (!e.empty()) return e;
  p->kind = 3;
  p->dct_y = true;
  const double pi = 2.0 * std::asin(1.0);
  ExactOps o;
  o.nx = p->nx; o.ny = ny; o.nz = p->nz; o.nxh = p->nxh;
  o.lx.resize(o.nxh);
  for (int k = 0; k < o.nxh; ++k) o.lx[k] = -mod_k2(iacc, 2.0 * pi * k / o.nx, o.nx, m.lxx);
  o.lz.resize(o.nz);
  for (int k = 0; k < o.nz; ++k) {
    const int kk = k <= o.nz / 2 ? k : o.nz - k;
    o.lz[k] = -mod_k2(iacc, 2.0 * pi * kk / o.nz, o.nz, m.lzz);
  }
  const double ly = m.lyt - m.lyb;
  o.lam.resize(ny);
  for (int k = 0; k < ny; ++k) o.lam[k] = k == 0 ? 0.0 : -mod_k2(iacc, pi * k / ny, ny, ly);
  auto cb = [&](int j, int k) { return std::cos(pi * double((long)k * (2 * j + 1) % (4 * ny)) / (2.0 * ny)); };
  o.V.resize((size_t)ny * ny);
  o.Vi.resize((size_t)ny * ny);
  for (int j = 0; j < ny; ++j)
    for (int k = 0; k < ny; ++k) {
      o.V[(size_t)j * ny + k] = cb(j, k);
      o.Vi[(size_t)k * ny + j] = cb(j, k) * (k == 0 ? 1.0 : 2.0) / ny;
    }
  o.null_idx = 0;
  const int h = ny / 2;
  o.Ve.resize((size_t)h * h); o.Vie.resize((size_t)h * h); o.lame.resize(h);
  o.Vo.resize((size_t)h * h); o.Vio.resize((size_t)h * h); o.lamo.resize(h);
  for (int q = 0; q < h; ++q) {
    o.lame[q] = o.lam[2 * q];
    o.lamo[q] = o.lam[2 * q + 1];
    for (int j = 0; j < h; ++j) {
      const double ce = cb(j, 2 * q), co = cb(j, 2 * q + 1);
      o.Ve[(size_t)j * h + q] = ce;
      o.Vo[(size_t)j * h + q] = co;
      o.Vie[(size_t)q * h + j] = ce / (q == 0 ? 0.5 * ny : 0.25 * ny);
      o.Vio[(size_t)q * h + j] = co / (0.25 * ny);
    }
  }
  o.null_even = 0;
  o.sym = true;
  if (!(e = p->setup_exact(o)).empty()) return e;
  *out = std::move(p);
  return "";
}

// Triply periodic path (poisson_000, src/poisson_1stderivcomp.f90:1016-1195,
// wavenumbers of waves :2405-2480 with the coefficients of :190-200): real FFT
// in x, complex FFTs in y and z, division by minus the sum of the squared
// modified wavenumbers of the staggered first derivative
//   k'(w) = n/L * (2a sin(w/2) + 2b sin(3w/2)) / (1 + 2 alpha cos w),
// the mean mode set to zero, inverse transforms and 1/(nx ny nz).  The
// reference's half-cell shift factors (abxyz) are applied before and undone
// after a real scalar division, so they cancel and are not formed here.
std::string Poisson::create_3d(const Mesh& m, int iacc, std::unique_ptr<Poisson>* out) {
  if (!m.periodic[0] || !m.periodic[1] || !m.periodic[2]) return "poisson: the 3-D FFT path needs periodic x, y, z";
  if (m.stretched) return "poisson: the 3-D FFT path needs a uniform y mesh";
  std::unique_ptr<Poisson> p(new Poisson());
  p->kind = 3;
  p->nx = m.nc[0]; p->ny = m.nc[1]; p->nz = m.nc[2];
  if (p->nx % 2) return "poisson: nx must be even";
  p->nxh = p->nx / 2 + 1;
  p->rad_x = radix_plan(p->nx / 2);
  p->rad_y = radix_plan(p->ny);
  p->rad_z = radix_plan(p->nz);
  if (p->rad_x.empty() || p->rad_y.empty() || p->rad_z.empty()) return "poisson: nx/2, ny, nz must be 2,3,5-smooth";
  if (p->rad_x.size() > 12 || p->rad_y.size() > 12 || p->rad_z.size() > 12) return "poisson: transform too long";
  if (p->nx / 2 > kTileC || p->ny > kTileC || p->nz > kTileC) return "poisson: transform too long for one tile";
  const double pi = 2.0 * std::asin(1.0);
  auto twid = [&](int Nc) {
    std::vector<double> t(2 * Nc);
    for (int q = 0; q < Nc; ++q) {
      double a = 2.0 * pi * double(q) / double(Nc);
      t[2 * q] = std::cos(a);
      t[2 * q + 1] = -std::sin(a);
    }
    return t;
  };
  std::vector<double> twx = twid(p->nx / 2), twy = twid(p->ny), twz = twid(p->nz), rot(2 * p->nxh);
  for (int k = 0; k < p->nxh; ++k) {
    double a = 2.0 * pi * double(k) / double(p->nx);
    rot[2 * k] = std::cos(a);
    rot[2 * k + 1] = -std::sin(a);
  }
  // modified wavenumbers squared per complex mode index
  const double al = coeff(F_D1C2P, false, 3, 1, BC_PERIODIC, iacc);
  const double ar = coeff(F_D1C2P, true, 3, 1, BC_PERIODIC, iacc);
  const double br = coeff(F_D1C2P, true, 3, 2, BC_PERIODIC, iacc);
  const double len[3] = {m.lxx, m.lyt - m.lyb, m.lzz};
  std::vector<double> k2[3];
  for (int d = 0; d < 3; ++d) {
    const int n = m.nc[d];
    const double h = len[d] / n;
    const double ac = ar / h, bc = br / h;
    std::vector<double> half(n / 2 + 1);
    for (int i = 0; i <= n / 2; ++i) {
      const double w = 2.0 * pi * double(i) / double(n);
      double wp = ac * 2.0 * h * std::sin(w * 0.5) + bc * 2.0 * h * std::sin(1.5 * w);
      wp = wp / (1.0 + 2.0 * al * std::cos(w));
      const double kk = double(n) * wp / len[d];
      half[i] = kk * kk;
    }
    const int nk = (d == 0) ? p->nxh : n;
    k2[d].resize(nk);
    for (int i = 0; i < nk; ++i) k2[d][i] = half[i <= n / 2 ? i : n - i];
  }
  std::string e;
  if (!(e = upv(twx, &p->tw_x64)).empty() || !(e = upv(twx, &p->tw_x32)).empty()) return e;
  if (!(e = upv(twy, &p->tw_y64)).empty() || !(e = upv(twy, &p->tw_y32)).empty()) return e;
  if (!(e = upv(twz, &p->tw_z64)).empty() || !(e = upv(twz, &p->tw_z32)).empty()) return e;
  if (!(e = upv(rot, &p->rot64)).empty() || !(e = upv(rot, &p->rot32)).empty()) return e;
  if (!(e = upv(k2[0], &p->lx64)).empty() || !(e = upv(k2[0], &p->lx32)).empty()) return e;
  if (!(e = upv(k2[1], &p->ly64)).empty() || !(e = upv(k2[1], &p->ly32)).empty()) return e;
  if (!(e = upv(k2[2], &p->lz64)).empty() || !(e = upv(k2[2], &p->lz32)).empty()) return e;
  p->work_bytes = sizeof(double) * 2 * (size_t)p->nxh * p->ny * p->nz;
  PCHK(hipMalloc(&p->work, p->work_bytes));
  if (!(e = p->tune(true)).empty()) return e;
  *out = std::move(p);
  return "";
}

std::string Poisson::tune(bool ydir) {
  void* tmp = nullptr;
  PCHK(hipMalloc(&tmp, work_bytes));
  Radices rx{}, ry{}, rz{};
  rx.n = (int)rad_x.size();
  for (int q = 0; q < rx.n; ++q) rx.r[q] = rad_x[q];
  ry.n = (int)rad_y.size();
  for (int q = 0; q < ry.n; ++q) ry.r[q] = rad_y[q];
  rz.n = (int)rad_z.size();
  for (int q = 0; q < rz.n; ++q) rz.r[q] = rad_z[q];
  tune_transforms<double>(nx, ny, nz, nxh, ydir, reinterpret_cast<cx<double>*>(work), reinterpret_cast<cx<double>*>(tmp),
                          reinterpret_cast<const cx<double>*>(tw_x64), reinterpret_cast<const cx<double>*>(tw_y64),
                          reinterpret_cast<const cx<double>*>(tw_z64), reinterpret_cast<const cx<double>*>(rot64), rx,
                          ry, rz);
  const hipError_t sync = hipDeviceSynchronize();
  (void)hipFree(tmp);
  if (sync != hipSuccess) return std::string("poisson tuning: ") + hipGetErrorString(sync);
  return "";
}

template <typename T>
std::string Poisson::solve_3d_t(T* rhs, hipStream_t s, T in_scale) {
  Radices rx{}, ry{}, rz{};
  rx.n = (int)rad_x.size();
  for (int q = 0; q < rx.n; ++q) rx.r[q] = rad_x[q];
  ry.n = (int)rad_y.size();
  for (int q = 0; q < ry.n; ++q) ry.r[q] = rad_y[q];
  rz.n = (int)rad_z.size();
  for (int q = 0; q < rz.n; ++q) rz.r[q] = rad_z[q];
  const bool dbl = sizeof(T) == 8;
  auto pk = [&](void* d, void* f) { return dbl ? d : f; };
  const cx<T>* twx = reinterpret_cast<const cx<T>*>(pk(tw_x64, tw_x32));
  const cx<T>* twy = reinterpret_cast<const cx<T>*>(pk(tw_y64, tw_y32));
  const cx<T>* twz = reinterpret_cast<const cx<T>*>(pk(tw_z64, tw_z32));
  const cx<T>* rot = reinterpret_cast<const cx<T>*>(pk(rot64, rot32));
  const T* k2x = reinterpret_cast<const T*>(pk(lx64, lx32));
  const T* k2y = reinterpret_cast<const T*>(pk(ly64, ly32));
  const T* k2z = reinterpret_cast<const T*>(pk(lz64, lz32));
  cx<T>* W = reinterpret_cast<cx<T>*>(work);
  const long nlx = (long)ny * nz;
  launch_fft_x<T>(true, rhs, W, nullptr, nx, nlx, rx, twx, rot, in_scale, s);
  launch_fft_s<T>(W, nxh, ny, (long)nxh, nz, (long)nxh * ny, ry, twy, T(-1), s);
  launch_fft_s<T>(W, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, T(-1), s);
  const long nm = (long)nxh * ny * nz;
  hipLaunchKernelGGL(k_div_k2<T>, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, W, nxh, ny, nz, k2x, k2y, k2z);
  launch_fft_s<T>(W, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, T(1), s);
  launch_fft_s<T>(W, nxh, ny, (long)nxh, nz, (long)nxh * ny, ry, twy, T(1), s);
  const T scale = T(1.0 / (double(nx) * double(ny) * double(nz)));
  launch_fft_x<T>(false, nullptr, W, rhs, nx, nlx, rx, twx, rot, scale, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("poisson launch: ") + hipGetErrorString(e);
  return "";
}

template <typename T>
std::string Poisson::solve_t(T* rhs, hipStream_t s, T in_scale) {
  if (kind == 3 && dct_y) return solve_exact_t<T>(rhs, s, in_scale);
  if (kind == 3) return solve_3d_t<T>(rhs, s, in_scale);
  Radices rx{}, rz{};
  rx.n = (int)rad_x.size();
  for (int q = 0; q < rx.n; ++q) rx.r[q] = rad_x[q];
  rz.n = (int)rad_z.size();
  for (int q = 0; q < rz.n; ++q) rz.r[q] = rad_z[q];
  const bool dbl = sizeof(T) == 8;
  const cx<T>* twx = reinterpret_cast<const cx<T>*>(dbl ? (void*)tw_x64 : (void*)tw_x32);
  const cx<T>* twz = reinterpret_cast<const cx<T>*>(dbl ? (void*)tw_z64 : (void*)tw_z32);
  const cx<T>* rot = reinterpret_cast<const cx<T>*>(dbl ? (void*)rot64 : (void*)rot32);
  const T* fd = reinterpret_cast<const T*>(dbl ? (void*)fd64 : (void*)fd32);
  const T* fz = reinterpret_cast<const T*>(dbl ? (void*)fz64 : (void*)fz32);
  cx<T>* W = reinterpret_cast<cx<T>*>(work);
  const long nlx = (long)ny * nz;
  launch_fft_x<T>(true, rhs, W, nullptr, nx, nlx, rx, twx, rot, in_scale, s);
  launch_fft_s<T>(W, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, T(-1), s);
  const long nm = (long)nxh * nz;
  hipLaunchKernelGGL(k_ytdma<T>, dim3((unsigned)((nm + 63) / 64)), dim3(64), 0, s, W, nxh, ny, nz, ya64, fd, fz);
  launch_fft_s<T>(W, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, T(1), s);
  const T scale = T(1.0 / (double(nx) * double(nz)));
  launch_fft_x<T>(false, nullptr, W, rhs, nx, nlx, rx, twx, rot, scale, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("poisson launch: ") + hipGetErrorString(e);
  return "";
}

std::string Poisson::setup_exact(const ExactOps& o) {
  if (o.nx != nx || o.ny != ny || o.nz != nz || o.nxh != nxh) return "poisson: exact operators of another grid";
  if ((int)o.lx.size() != nxh || (int)o.lz.size() != nz || (int)o.lam.size() != ny) return "poisson: bad exact tables";
  std::string e;
  if (!(e = upv(o.V, &V64)).empty() || !(e = upv(o.V, &V32)).empty()) return e;
  if (!(e = upv(o.Vi, &Vi64)).empty() || !(e = upv(o.Vi, &Vi32)).empty()) return e;
  if (!(e = upv(o.lam, &lam64)).empty() || !(e = upv(o.lam, &lam32)).empty()) return e;
  if (!(e = upv(o.lx, &lx64)).empty() || !(e = upv(o.lx, &lx32)).empty()) return e;
  if (!(e = upv(o.lz, &lz64)).empty() || !(e = upv(o.lz, &lz32)).empty()) return e;
  null_idx = o.null_idx;
  if (o.sym) {
    if (!(e = upv(o.Ve, &Ve64)).empty() || !(e = upv(o.Ve, &Ve32)).empty()) return e;
    if (!(e = upv(o.Vo, &Vo64)).empty() || !(e = upv(o.Vo, &Vo32)).empty()) return e;
    if (!(e = upv(o.Vie, &Vie64)).empty() || !(e = upv(o.Vie, &Vie32)).empty()) return e;
    if (!(e = upv(o.Vio, &Vio64)).empty() || !(e = upv(o.Vio, &Vio32)).empty()) return e;
    if (!(e = upv(o.lame, &lame64)).empty() || !(e = upv(o.lame, &lame32)).empty()) return e;
    if (!(e = upv(o.lamo, &lamo64)).empty() || !(e = upv(o.lamo, &lamo32)).empty()) return e;
    null_even = o.null_even;
  }
  PCHK(hipMalloc(&work2, work_bytes));
  return "";
}

template <typename T>
std::string Poisson::solve_exact_t(T* rhs, hipStream_t s, T in_scale) {
  if (!V64) return "poisson: exact operators not set up";
  Radices rx{}, rz{};
  rx.n = (int)rad_x.size();
  for (int q = 0; q < rx.n; ++q) rx.r[q] = rad_x[q];
  rz.n = (int)rad_z.size();
  for (int q = 0; q < rz.n; ++q) rz.r[q] = rad_z[q];
  const bool dbl = sizeof(T) == 8;
  const cx<T>* twx = reinterpret_cast<const cx<T>*>(dbl ? (void*)tw_x64 : (void*)tw_x32);
  const cx<T>* twz = reinterpret_cast<const cx<T>*>(dbl ? (void*)tw_z64 : (void*)tw_z32);
  const cx<T>* rot = reinterpret_cast<const cx<T>*>(dbl ? (void*)rot64 : (void*)rot32);
  const T* V = reinterpret_cast<const T*>(dbl ? (void*)V64 : (void*)V32);
  const T* Vi = reinterpret_cast<const T*>(dbl ? (void*)Vi64 : (void*)Vi32);
  const T* lam = reinterpret_cast<const T*>(dbl ? (void*)lam64 : (void*)lam32);
  const T* lxv = reinterpret_cast<const T*>(dbl ? (void*)lx64 : (void*)lx32);
  const T* lzv = reinterpret_cast<const T*>(dbl ? (void*)lz64 : (void*)lz32);
  cx<T>* W = reinterpret_cast<cx<T>*>(work);
  T* Wr = reinterpret_cast<T*>(work);
  T* W2r = reinterpret_cast<T*>(work2);
  const long nlx = (long)ny * nz;
  launch_fft_x<T>(true, rhs, W, nullptr, nx, nlx, rx, twx, rot, in_scale, s);
  launch_fft_s<T>(W, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, T(-1), s);
  const int ncr = 2 * nxh;
  const long ncol = (long)ncr * nz;
  if (Ve64 && dbl && ny / 2 <= 64) {
    // one workgroup per tile of 64 columns (two share a CU)
    const long ntile = (ncol + kMmC - 1) / kMmC;
    const unsigned nwg = (unsigned)ntile;
    hipLaunchKernelGGL(k_ymix_mfma, dim3(nwg), dim3(64 * kMmWaves), 0, s, Vie64, Vio64, Ve64, Vo64,
                       reinterpret_cast<double*>(Wr), ny, ncr, ncol, lame64, lamo64,
                       reinterpret_cast<const double*>(lxv), reinterpret_cast<const double*>(lzv), null_even);
  } else if (Ve64 && ny / 2 <= kFuH) {
    auto pick = [&](double* d, float* f) { return reinterpret_cast<const T*>(dbl ? (void*)d : (void*)f); };
    const size_t sh = sizeof(T) * ((size_t)2 * kFuH * (kFuC + 1) + (size_t)kFuK * (2 * kFuH + 4));
    hipLaunchKernelGGL(k_ymix_fused<T>, dim3((unsigned)((ncol + kFuC - 1) / kFuC)), dim3(256), sh, s,
                       pick(Vie64, Vie32), pick(Vio64, Vio32), pick(Ve64, Ve32), pick(Vo64, Vo32), Wr, Wr, ny, ncr,
                       ncol, pick(lame64, lame32), pick(lamo64, lamo32), lxv, lzv, null_even);
  } else if (Ve64) {
    auto pick = [&](double* d, float* f) { return reinterpret_cast<const T*>(dbl ? (void*)d : (void*)f); };
    const int h = ny / 2;
    dim3 gs((unsigned)((ncol + kMixC - 1) / kMixC), (unsigned)((h + kSymR - 1) / kSymR));
    hipLaunchKernelGGL((k_ymix_sym<T, 1>), gs, dim3(256), 0, s, pick(Vie64, Vie32), pick(Vio64, Vio32), Wr, W2r,
                       ny, ncr, ncol, pick(lame64, lame32), pick(lamo64, lamo32), lxv, lzv, null_even);
    hipLaunchKernelGGL((k_ymix_sym<T, 2>), gs, dim3(256), 0, s, pick(Ve64, Ve32), pick(Vo64, Vo32), W2r, Wr, ny,
                       ncr, ncol, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, -1);
  } else {
    dim3 gm((unsigned)((ncol + kMixC - 1) / kMixC), (unsigned)((ny + kMixR - 1) / kMixR));
    hipLaunchKernelGGL(k_ymix<T>, gm, dim3(256), 0, s, Vi, Wr, W2r, ny, ncr, ncol, lam, lxv, lzv, null_idx);
    hipLaunchKernelGGL(k_ymix<T>, gm, dim3(256), 0, s, V, W2r, Wr, ny, ncr, ncol, (const T*)nullptr,
                       (const T*)nullptr, (const T*)nullptr, -1);
  }
  launch_fft_s<T>(W, nxh, nz, (long)nxh * ny, ny, (long)nxh, rz, twz, T(1), s);
  const T scale = T(1.0 / (double(nx) * double(nz)));
  launch_fft_x<T>(false, nullptr, W, rhs, nx, nlx, rx, twx, rot, scale, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("poisson launch: ") + hipGetErrorString(e);
  return "";
}

std::string Poisson::solve_exact(double* rhs, hipStream_t s, double in_scale) {
  return solve_exact_t<double>(rhs, s, in_scale);
}
std::string Poisson::solve_exact(float* rhs, hipStream_t s, float in_scale) {
  return solve_exact_t<float>(rhs, s, in_scale);
}

std::string Poisson::solve(double* rhs, hipStream_t s, double in_scale) { return solve_t<double>(rhs, s, in_scale); }
std::string Poisson::solve(float* rhs, hipStream_t s, float in_scale) { return solve_t<float>(rhs, s, in_scale); }

}  // namespace chs
