// Compact line operators on gfx950.
//
// Partitioned kernels (k_line_yz, k_line_xu, k_line_x; lines of up to
// kChunkLong*kMaxChunks rows).  A line of n rows is cut into P chunks of at
// most CH rows (CH = kChunk up to kChunk*kMaxChunks rows, kChunkLong beyond)
// and one thread owns one chunk, so a launch carries P times more threads
// than lines (a 256x128x256 grid has only 32768-65536 lines per direction,
// which is too few for 256 CUs).  Each thread evaluates the explicit RHS of
// its rows from a register window, runs the chunk-local Thomas sweeps with
// host-factored coefficients, publishes its first and last values through
// LDS, and corrects its rows with the spike vectors and two rows of the
// inverse of the 2P x 2P reduced system (the same algebra for wall-bounded
// and periodic lines; a periodic line's corner entries close the reduced
// system cyclically).
//  - y and z lines: a workgroup is 64 neighbouring lines x P chunks; wave q of
//    the workgroup holds chunk q of 64 lines, so every load and store of a
//    wavefront is a contiguous 64-element segment along x and the plan's
//    per-row coefficients are wave-uniform.
//  - x lines (contiguous in memory) of up to kChunk*kMaxChunks rows: a
//    workgroup stages 64 whole lines in LDS with coalesced loads (applying
//    the input transform), wave q works on chunk q of all 64 lines from LDS,
//    and the results go back through LDS to coalesced stores (k_line_xu; or
//    256/P lines with skewed rows, k_line_x).  Longer x lines take the y/z
//    kernel with x addressing (lanes one line apart, rows contiguous).
//
// Single-thread-per-line kernel (k_line; lines beyond kChunkLong*kMaxChunks rows).  For y- and z-lines consecutive threads own
// consecutive x positions, so every load and store of a wavefront is a
// contiguous 64-element segment.  The explicit RHS (src/basics_operations2.f90
// :1980-2719) is evaluated on the fly with a 5-point register window; the four
// boundary rows use the affine recipes built on the host.  The forward Thomas
// sweep (src/basics_algorithms.f90:36-66) writes its intermediate values to a
// scratch array laid out like the output; the backward sweep reads them back
// (the round trip is served by L2/MALL) and writes the final value through the
// fused output transform.  Periodic lines use the Sherman-Morrison form of
// Solve_TDMA_cyclic (:68-90) with the RHS-independent correction vector
// precomputed, and the basic solution's first entry accumulated during the
// forward sweep so the correction is applied in the same backward pass.
#include "chs_lineop.h"
#include "chs_tune.h"

#include <cmath>
#include <vector>

namespace chs {

#define CHS_CHECK(x)                                                        \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) return std::string(#x ": ") + hipGetErrorString(e_); \
  } while (0)

template <typename T>
DevPlan<T>::~DevPlan() {
  for (void* p : {(void*)rows, (void*)a, (void*)c, (void*)d, (void*)pw, (void*)x1, (void*)metric, (void*)rinv,
                  (void*)swin, (void*)co})
    if (p) (void)hipFree(p);
}

template <typename T>
static std::string up(const std::vector<double>& v, T** dst) {
  if (v.empty()) { *dst = nullptr; return ""; }
  std::vector<T> t(v.begin(), v.end());
  CHS_CHECK(hipMalloc(dst, sizeof(T) * t.size()));
  CHS_CHECK(hipMemcpy(*dst, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice));
  return "";
}

// Line index read for window position k, -2 <= k <= n + CH + 1 (CH the chunk rows):
// wrapped on periodic lines (one period suffices, as periodic lines on the
// partitioned kernels have n >= CH + 2), clamped to the line otherwise.
__host__ __device__ __forceinline__ int window_index(int k, int n, bool wrap) {
  if (wrap) return k < 0 ? k + n : (k >= n ? k - n : k);
  return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
}

template <typename T>
std::string upload_plan(const LinePlan& p, DevPlan<T>* dp) {
  dp->host = p;
  DevRow<T> r[4];
  for (int q = 0; q < 4; ++q) {
    r[q].nterm = p.rows[q].nterm;
    for (int t = 0; t < kMaxRowTerms; ++t) {
      r[q].idx[t] = p.rows[q].idx[t];
      r[q].w[t] = T(p.rows[q].w[t]);
    }
    for (int m = 0; m < 4; ++m) r[q].u[m] = T(p.rows[q].u[m]);
  }
  CHS_CHECK(hipMalloc(&dp->rows, sizeof(r)));
  CHS_CHECK(hipMemcpy(dp->rows, r, sizeof(r), hipMemcpyHostToDevice));
  std::string e;
  if (!(e = up(p.a, &dp->a)).empty()) return e;
  if (!(e = up(p.c, &dp->c)).empty()) return e;
  if (!(e = up(p.d, &dp->d)).empty()) return e;
  if (!(e = up(p.pw, &dp->pw)).empty()) return e;
  if (!(e = up(p.x1, &dp->x1)).empty()) return e;
  if (!(e = up(p.metric, &dp->metric)).empty()) return e;
  if (!(e = up(p.rinv, &dp->rinv)).empty()) return e;
  // dense window weights of the four boundary rows
  {
    const int W = p.part_ch + 4;
    const bool wrap = (p.ibc[0] == BC_PERIODIC);
    std::vector<double> sw(4 * W, 0.0);
    bool ok = p.part_P >= 1 && (int)p.part_start.size() == p.part_P + 1 && (!wrap || p.n_in >= p.part_ch + 2);
    for (int q = 0; q < 4 && ok; ++q) {
      const int s0 = (q < 2) ? 0 : p.part_start[p.part_P - 1];
      for (int t = 0; t < p.rows[q].nterm; ++t) {
        // the window slot holding line index idx (as the kernels index it)
        int slot = -1;
        for (int u = 0; u < W && slot < 0; ++u)
          if (window_index(s0 - 2 + u, p.n_in, wrap) == p.rows[q].idx[t]) slot = u;
        if (slot < 0) { ok = false; break; }
        sw[q * W + slot] += p.rows[q].w[t];
      }
    }
    // self-check of the slot mapping on a deterministic test line: the dense
    // weights applied to the clamped window must give the recipe's value
    if (ok) {
      std::vector<double> v(p.n_in);
      for (int k = 0; k < p.n_in; ++k) v[k] = 1.0 + 0.37 * k - 0.011 * k * (double)k;
      for (int q = 0; q < 4 && ok; ++q) {
        const int s0 = (q < 2) ? 0 : p.part_start[p.part_P - 1];
        double direct = 0.0, dense = 0.0, mag = 0.0;
        for (int t = 0; t < p.rows[q].nterm; ++t) {
          direct += p.rows[q].w[t] * v[p.rows[q].idx[t]];
          mag += std::fabs(p.rows[q].w[t] * v[p.rows[q].idx[t]]);
        }
        for (int t = 0; t < W; ++t) dense += sw[q * W + t] * v[window_index(s0 - 2 + t, p.n_in, wrap)];
        if (std::fabs(direct - dense) > 1e-12 * (mag + 1e-300)) ok = false;
      }
    }
    dp->win_ok = ok;
    if (ok && !(e = up(sw, &dp->swin)).empty()) return e;
  }
  // chunk coefficient rows: CH = part_ch rows per chunk, kCo entries per row,
  // paired for 16-byte reads: (ld local pivot inverse, a sub-diagonal; a = 0
  // on the chunk's first row), (the two spike vectors), (lc local
  // super-diagonal, 0 on the chunk's last row; the output metric, 1 when
  // absent); rows past the chunk's length are identity rows
  const int CH = p.part_ch;
  if (p.part_P >= 1 && (int)p.part_start.size() == p.part_P + 1 && p.part_P * CH >= p.n_out) {
    std::vector<double> co((size_t)p.part_P * CH * kCo, 0.0);
    const bool solve_data = !p.ld.empty();
    for (int q = 0; q < p.part_P; ++q) {
      const int s0 = p.part_start[q], len = p.part_start[q + 1] - s0;
      if (len > CH || s0 != (q * p.n_out) / p.part_P) return "line plan: chunk layout differs from the kernels'";
      for (int i = 0; i < CH; ++i) {
        double* c = &co[((size_t)q * CH + i) * kCo];
        const int r = s0 + i;
        c[CO_LD] = 1.0;
        c[CO_MET] = 1.0;
        if (i >= len) continue;
        if (solve_data) {
          c[CO_LD] = p.ld[r];
          c[CO_AR] = (i > 0) ? p.a_raw[r] : 0.0;
          c[CO_LC] = (i < len - 1) ? p.lc[r] : 0.0;
          c[CO_SV] = p.sv[r];
          c[CO_SW] = p.sw[r];
        }
        if (!p.metric.empty()) c[CO_MET] = p.metric[r];
      }
    }
    if (!(e = up(co, &dp->co)).empty()) return e;
  }
  return "";
}

static inline int lds_pos_host(int k) { return k + (k >> 4); }

template <typename T>
struct KArgs {
  int n_in, n_out;
  int e1, e2;      // input extents 1,2
  int o1, o2;      // output extents 1,2
  long nlines;
  const T* A; const T* B; const T* Cm; T sA, cB;
  const T* D; const T* E; T cD, cE;
  const T* fbc;
  T c1, c2, dd, sgn;
  const DevRow<T>* rows;
  int solve; T b0;
  const T* a; const T* c; const T* d; const T* pw; const T* x1;
  T an, cn, den;
  const T* metric; int neu_lo, neu_hi;
  T* out; T sOut; int accumulate;
  T* scratch;
  // partitioned form
  int P, per, ls, nl;
  const T* rinv;
  const T* swin;  // 4 x (CH + 4) boundary-row window weights (CH = the plan's chunk rows)
  int wrap;       // periodic line: windows wrap around instead of clamping
  const T* co;    // P x CH coefficient rows of kCo entries (see upload_plan)
};

// bulk offsets (p1,p2,p3,p4) per operator kind, relative to the output row
template <int OPK> struct Offs;
template <> struct Offs<MIDP_C2P> { static constexpr int p1 = 0, p2 = -1, p3 = 1, p4 = -2; };
template <> struct Offs<MIDP_P2C> { static constexpr int p1 = 0, p2 = 1, p3 = -1, p4 = 2; };
template <> struct Offs<D1_C2C> { static constexpr int p1 = 1, p2 = -1, p3 = 2, p4 = -2; };
template <> struct Offs<D1_P2P> { static constexpr int p1 = 1, p2 = -1, p3 = 2, p4 = -2; };
template <> struct Offs<D1_C2P> { static constexpr int p1 = 0, p2 = -1, p3 = 1, p4 = -2; };
template <> struct Offs<D1_P2C> { static constexpr int p1 = 1, p2 = 0, p3 = 2, p4 = -1; };

template <typename T, int DIR, int OPK>
__global__ void __launch_bounds__(64) k_line(KArgs<T> P) {
  const long L = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= P.nlines) return;
  long ib, is, ob, os, fb, fs;
  if (DIR == 0) {
    long j = L % P.e2, k = L / P.e2;
    ib = (j + (long)P.e2 * k) * P.e1; is = 1;
    ob = (j + (long)P.o2 * k) * P.o1; os = 1;
    fb = 4 * (j + (long)P.e2 * k); fs = 1;
  } else if (DIR == 1) {
    long i = L % P.e1, k = L / P.e1;
    ib = i + (long)P.e1 * P.e2 * k; is = P.e1;
    ob = i + (long)P.o1 * P.o2 * k; os = P.o1;
    fb = i + (long)P.e1 * 4 * k; fs = P.e1;
  } else {
    long i = L % P.e1, j = L / P.e1;
    ib = i + (long)P.e1 * j; is = (long)P.e1 * P.e2;
    ob = i + (long)P.o1 * j; os = (long)P.o1 * P.o2;
    fb = i + (long)P.e1 * j; fs = (long)P.e1 * P.e2;
  }
  const int n_in = P.n_in, n_out = P.n_out;
  auto F = [&](int k) -> T {
    const long o = ib + is * k;
    T t = P.A[o];
    if (P.B) t = t + P.cB * P.B[o];
    t = P.sA * t;
    if (P.Cm) t = t * P.Cm[o];
    if (P.D) t = t + P.cD * P.D[o];
    if (P.E) t = t + P.cE * P.E[o];
    return t;
  };
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc)
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[fb + fs * m];
  auto special = [&](int q) -> T {
    const DevRow<T>& R = P.rows[q];
    T v = T(0);
    for (int t = 0; t < R.nterm; ++t) v = v + R.w[t] * F(R.idx[t]);
    if (P.fbc)
      for (int m = 0; m < 4; ++m) v = v + R.u[m] * fb4[m];
    return v * P.dd;
  };
  auto emit = [&](int r, T val) {
    if (P.metric) val = val * P.metric[r];
    if (r == 0 && P.neu_lo) val = fb4[0];
    if (r == n_out - 1 && P.neu_hi) val = fb4[1];
    const long o = ob + os * r;
    T prev = P.accumulate ? P.out[o] : T(0);
    P.out[o] = prev + val * P.sOut;
  };
  using O = Offs<OPK>;
  const int m = (P.solve == 2) ? n_out - 1 : n_out;
  T* sc = P.scratch;
  T y = T(0), S = T(0), rn = T(0);
  T w0, w1, w2, w3, w4;  // f[r-2 .. r+2]
  auto clampi = [&](int k) { return k < 0 ? 0 : (k >= n_in ? n_in - 1 : k); };
  w0 = F(0); w1 = F(clampi(1)); w2 = F(clampi(2)); w3 = F(clampi(3)); w4 = F(clampi(4));
  auto win = [&](int off) -> T {
    switch (off) {
      case -2: return w0;
      case -1: return w1;
      case 0: return w2;
      case 1: return w3;
      default: return w4;
    }
  };
  for (int r = 0; r < n_out; ++r) {
    T rv;
    if (r < 2) {
      rv = special(r);
    } else if (r >= n_out - 2) {
      rv = special(r - n_out + 4);
    } else {
      rv = (P.c1 * (win(O::p1) + P.sgn * win(O::p2)) + P.c2 * (win(O::p3) + P.sgn * win(O::p4))) * P.dd;
      w0 = w1; w1 = w2; w2 = w3; w3 = w4;
      w4 = F(clampi(r + 3));
    }
    if (P.solve == 0) {
      emit(r, rv);
    } else if (r < m) {
      y = (r == 0) ? rv / P.b0 : (rv - P.a[r] * y) * P.d[r];
      if (P.solve == 2) S = S + P.pw[r] * y;
      if (r < m - 1) sc[ob + os * r] = y;
    } else {
      rn = rv;
    }
  }
  if (P.solve == 0) return;
  T xn = T(0);
  if (P.solve == 2) {
    xn = (rn - P.cn * S - P.an * y) / P.den;
    emit(n_out - 1, xn);
  }
  T xnext = y;
  emit(m - 1, (P.solve == 2) ? xnext + P.x1[m - 1] * xn : xnext);
  for (int r = m - 2; r >= 0; --r) {
    const T xb = sc[ob + os * r] - P.c[r] * xnext;
    xnext = xb;
    emit(r, (P.solve == 2) ? xb + P.x1[r] * xn : xb);
  }
}

// ---------------------------------------------------------------------------
// partitioned kernels

// Chunk rows [s0, s0+len) of one line: explicit RHS into y[] from the
// window f[t] = input at line index s0-2+t (clamped to the line).  The four
// boundary rows use their dense window weights (plans whose boundary recipes
// reach outside the window run on the one-thread-per-line kernel instead);
// they are formed first, so that the interior rows can then consume the
// window in order (a window value is dead once the row two below it is
// done, and y[] takes its registers).
template <typename T, int OPK, int CH>
__device__ __forceinline__ void chunk_rhs(const KArgs<T>& P, int q, int len, const T (&f)[CH + 4],
                                          const T (&fb4)[4], T (&y)[CH]) {
  using O = Offs<OPK>;
  constexpr int W = CH + 4;
  auto special = [&](int rq) -> T {
    const DevRow<T>& R = P.rows[rq];
    T v = T(0);
    const T* wv = P.swin + rq * W;
#pragma unroll
    for (int t = 0; t < W; ++t) v = v + wv[t] * f[t];
    if (P.fbc)
      for (int m = 0; m < 4; ++m) v = v + R.u[m] * fb4[m];
    return v * P.dd;
  };
  const bool first = q == 0, last = q == P.P - 1;
  T v0 = T(0), v1 = T(0), v2 = T(0), v3 = T(0);
  if (first) {
    v0 = special(0);
    v1 = special(1);
  }
  if (last) {
    v2 = special(2);
    v3 = special(3);
  }
#pragma unroll
  for (int i = 0; i < CH; ++i)
    y[i] = (P.c1 * (f[i + 2 + O::p1] + P.sgn * f[i + 2 + O::p2]) +
            P.c2 * (f[i + 2 + O::p3] + P.sgn * f[i + 2 + O::p4])) * P.dd;
  if (first) {
    y[0] = v0;
    y[1] = v1;
  }
  if (last) {
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      if (i == len - 2) y[i] = v2;
      if (i == len - 1) y[i] = v3;
    }
  }
}

// Element at byte offset `boff` from `base`: the y/z kernels address a line
// as a wave-uniform base pointer plus a 32-bit per-lane byte offset.
template <typename T>
__device__ __forceinline__ T ld_at(const T* base, unsigned boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void st_at(T* base, unsigned boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

// Input transform sA * (A + cB * B) * Cm of `nv` elements: f holds A, g holds
// the array loaded beside it (B if present, else Cm).  A third array (Cm when
// B is also present) is fetched by `third`.
template <typename T, int NV, class TH>
__device__ __forceinline__ void transform_in(const KArgs<T>& P, T (&f)[NV], const T (&g)[NV], TH&& third) {
  if (P.B) {
#pragma unroll
    for (int t = 0; t < NV; ++t) f[t] = f[t] + P.cB * g[t];
  }
#pragma unroll
  for (int t = 0; t < NV; ++t) f[t] = P.sA * f[t];
  if (P.Cm) {
    if (!P.B) {
#pragma unroll
      for (int t = 0; t < NV; ++t) f[t] = f[t] * g[t];
    } else {
      T h[NV];
#pragma unroll
      for (int t = 0; t < NV; ++t) h[t] = third(t);
#pragma unroll
      for (int t = 0; t < NV; ++t) f[t] = f[t] * h[t];
    }
  }
}

// Window rows [T0, T1) of one line chunk: line indices s0-2+t (wrapped on
// periodic lines, clamped otherwise, so every address is valid); the row
// offsets are wave-uniform.  IN = 0: A only; 1: A and the second input array
// (B, or Cm when B is absent), all loads issued before any use; 2: as 1 (the
// second array optional), then the additive arrays D (and E) as a second
// batch.
template <typename T, int IN, int CH, int T0, int T1>
__device__ __forceinline__ void load_rows(const KArgs<T>& P, long base, unsigned ri, long is, int s0,
                                          T (&f)[CH + 4]) {
  constexpr int NB = T1 - T0;
  long ro[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t) ro[t] = base + is * window_index(s0 - 2 + T0 + t, P.n_in, P.wrap);
  T* fw = f + T0;
  T(&fv)[NB] = *reinterpret_cast<T(*)[NB]>(fw);
  const T* S = P.B ? P.B : P.Cm;
  if (IN == 1 || (IN == 2 && S)) {
    T g[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      fv[t] = ld_at(P.A + ro[t], ri);
      g[t] = ld_at(S + ro[t], ri);
    }
    transform_in<T, NB>(P, fv, g, [&](int t) { return ld_at(P.Cm + ro[t], ri); });
  } else {
#pragma unroll
    for (int t = 0; t < NB; ++t) fv[t] = P.sA * ld_at(P.A + ro[t], ri);
  }
  if (IN == 2) {
    // The additive arrays are loaded after the first batch has been
    // combined (FP32: together; FP64: one after the other): their lane
    // offset passes through an empty asm that reads the window's last value,
    // so the compiler cannot hoist these loads next to the earlier batch
    // (registers for the window and all arrays together spill).
    unsigned rj = ri;
    asm volatile("" : "+v"(rj) : "v"(fv[NB - 1]));
    if (sizeof(T) == 4) {
      // FP32: both arrays in one batch
      T d[NB], e[NB];
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        d[t] = ld_at(P.D + ro[t], rj);
        e[t] = P.E ? ld_at(P.E + ro[t], rj) : T(0);
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) fv[t] = fv[t] + P.cD * d[t] + P.cE * e[t];
    } else {
      // FP64: one array at a time
      {
        T d[NB];
#pragma unroll
        for (int t = 0; t < NB; ++t) d[t] = ld_at(P.D + ro[t], rj);
#pragma unroll
        for (int t = 0; t < NB; ++t) fv[t] = fv[t] + P.cD * d[t];
      }
      if (P.E) {
        unsigned rk = rj;
        asm volatile("" : "+v"(rk) : "v"(fv[NB - 1]));
        T e[NB];
#pragma unroll
        for (int t = 0; t < NB; ++t) e[t] = ld_at(P.E + ro[t], rk);
#pragma unroll
        for (int t = 0; t < NB; ++t) fv[t] = fv[t] + P.cE * e[t];
      }
    }
  }
}

// Window of one line chunk of CH rows: rows s0-2 .. s0+CH+1.  FP64 windows
// of long chunks (CH > kChunk) are loaded in two halves, the second half's
// loads waiting (through an empty asm on its lane offset) until the first
// half is combined, so the raw input arrays of the whole window never hold
// registers at the same time.
template <typename T, int IN, int CH>
__device__ __forceinline__ void load_window(const KArgs<T>& P, long base, unsigned ri, long is, int s0,
                                            T (&f)[CH + 4]) {
  constexpr int W = CH + 4;
  if constexpr (sizeof(T) == 8 && CH > kChunk) {
    constexpr int H = W / 2;
    load_rows<T, IN, CH, 0, H>(P, base, ri, is, s0, f);
    unsigned rh = ri;
    asm volatile("" : "+v"(rh) : "v"(f[H - 1]));
    load_rows<T, IN, CH, H, W>(P, base, rh, is, s0, f);
  } else {
    load_rows<T, IN, CH, 0, W>(P, base, ri, is, s0, f);
  }
}

// Chunk-local Thomas sweeps over all CH rows of a chunk whose
// coefficient rows start at c; returns the chunk's first and last values.
// The coefficient rows past the chunk's length are identity rows (and the
// chunk's last row does not couple to the next), so the sweeps need no row
// guards.
template <typename T, int CH>
__device__ __forceinline__ void chunk_solve(const T* c, int len, T (&y)[CH], T& yf, T& yl) {
  y[0] = y[0] * c[CO_LD];
#pragma unroll
  for (int i = 1; i < CH; ++i) y[i] = (y[i] - c[i * kCo + CO_AR] * y[i - 1]) * c[i * kCo + CO_LD];
#pragma unroll
  for (int i = CH - 2; i >= 0; --i) y[i] = y[i] - c[i * kCo + CO_LC] * y[i + 1];
  yf = y[0];
  yl = y[0];
#pragma unroll
  for (int i = 1; i < CH; ++i)
    if (i == len - 1) yl = y[i];
}

// Reduced-system values coupling chunk q to its neighbours (rinv: the
// 2P x 2P inverse; ex(k) reads the published first/last value k of the
// thread's line), applied through the spike vectors of the chunk's rows c.
template <typename T, int CH, class EX>
__device__ __forceinline__ void chunk_couple(const KArgs<T>& P, int q, const T* rinv, const T* c, EX&& ex,
                                             T (&y)[CH]) {
  const int Pn = P.P, M = 2 * Pn;
  const int qm = (q + Pn - 1) % Pn, qp = (q + 1) % Pn;
  const T* rL = rinv + (2 * qm + 1) * M;
  const T* rF = rinv + (2 * qp) * M;
  T uL = T(0), uF = T(0);
  for (int k = 0; k < M; ++k) {
    const T v = ex(k);
    uL = uL + rL[k] * v;
    uF = uF + rF[k] * v;
  }
  if (!P.per && q == 0) uL = T(0);
  if (!P.per && q == Pn - 1) uF = T(0);
#pragma unroll
  for (int i = 0; i < CH; ++i) y[i] = y[i] - c[i * kCo + CO_SV] * uL - c[i * kCo + CO_SW] * uF;
}

// Row i of a chunk (line row r): metric scaling and Neumann overrides.
template <typename T>
__device__ __forceinline__ T chunk_post(const KArgs<T>& P, const T* c, int i, int r, T val, const T (&fb4)[4]) {
  val = val * c[i * kCo + CO_MET];
  if (r == 0 && P.neu_lo) val = fb4[0];
  if (r == P.n_out - 1 && P.neu_hi) val = fb4[1];
  return val;
}

// Per-plan coefficient data a workgroup of P waves copies into LDS: the chunk
// rows (P x CH x kCo) and the reduced-system inverse (2P x 2P), at most
// stage_max(CH) = CH kCo / 64 + 1 values per thread (P <= kMaxChunks).
__host__ __device__ constexpr int stage_max(int ch) { return ch * kCo / 64 + 1; }
constexpr int kStageMax = stage_max(kChunk);
static_assert(64 * kMaxChunks * stage_max(kChunkLong) >= kMaxChunks * kChunkLong * kCo + 4 * kMaxChunks * kMaxChunks &&
                  64 * kMaxChunks * kStageMax >= kMaxChunks * kChunk * kCo + 4 * kMaxChunks * kMaxChunks &&
                  64 * kStageMax >= kChunk * kCo + 4,
              "per-plan staging exceeds stage_max values per thread");

// First elements of line L in the input, output and fbc arrays.
template <int DIR, typename T>
__device__ __forceinline__ void yz_offsets(const KArgs<T>& P, int L, long& ib, long& ob, long& fb) {
  if (DIR == 0) {  // x lines (long lines only): contiguous rows, lanes one line apart
    ib = (long)L * P.e1;
    ob = (long)L * P.o1;
    fb = 4L * L;
  } else if (DIR == 1) {
    const int i = L % P.e1, k = L / P.e1;
    ib = i + (long)P.e1 * P.e2 * k;
    ob = i + (long)P.o1 * P.o2 * k;
    fb = i + (long)P.e1 * 4 * k;
  } else {
    ib = L;  // x varies fastest and the (x, y) extents of input and output agree
    ob = L;
    fb = L;
  }
}

// One y/z workgroup: 64 neighbouring lines x P chunks of CH rows, wave q
// holding chunk q of the 64 lines; IN as in load_window.  CH = kChunkLong
// also serves x lines longer than a 64-line LDS tile (DIR = 0: the lanes
// are 64 consecutive lines, each reading its own contiguous rows).
// Single-input operators ask for 5 waves per SIMD (at most 96 registers), so
// that two workgroups of 9 waves (lines of 129 rows) share a CU.
template <typename T, int DIR, int OPK, int IN, int CH>
__global__ void __launch_bounds__(64 * kMaxChunks)
    __attribute__((amdgpu_waves_per_eu(IN == 0 && CH == kChunk ? 5 : 4))) k_line_yz(KArgs<T> P) {
  __shared__ T ex[2 * kMaxChunks * 64];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sco = reinterpret_cast<T*>(smem_raw);  // P*CH*kCo chunk rows, then the 2P x 2P inverse
  // loads of the per-plan coefficients first (written to LDS once the
  // chunk's window loads are in flight); plans without a line solve (the
  // explicit schemes) have no inverse
  constexpr int NS = stage_max(CH);
  const int nst = P.P * CH * kCo + (P.rinv ? 4 * P.P * P.P : 0);
  T cv[NS];
  {
    const int nco = P.P * CH * kCo;
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int idx = threadIdx.x + u * 64 * P.P;
      cv[u] = (idx < nco) ? P.co[idx] : (idx < nst ? P.rinv[idx - nco] : T(0));
    }
  }
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int L0 = blockIdx.x * 64;  // nlines < 2^31 (checked at launch)
  const int Lr = L0 + lane;
  const bool live = Lr < P.nlines;
  const int L = live ? Lr : (int)P.nlines - 1;
  // block base (wave-uniform) + per-lane byte offset (< 2^32, checked at launch)
  long ib, ob, fb, ib0, ob0, fb0;
  yz_offsets<DIR>(P, L, ib, ob, fb);
  yz_offsets<DIR>(P, L0, ib0, ob0, fb0);
  const unsigned ri = (unsigned)((ib - ib0) * (long)sizeof(T));
  const unsigned ro = (unsigned)((ob - ob0) * (long)sizeof(T));
  const long is = (DIR == 0) ? 1L : (DIR == 1) ? (long)P.e1 : (long)P.e1 * P.e2;
  const long os = (DIR == 0) ? 1L : (DIR == 1) ? (long)P.o1 : (long)P.o1 * P.o2;
  const int n_out = P.n_out, Pn = P.P;
  const int s0 = (q * n_out) / Pn, len = ((q + 1) * n_out) / Pn - s0;
  auto stage = [&]() {
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int idx = threadIdx.x + u * 64 * Pn;
      if (idx < nst) sco[idx] = cv[u];
    }
  };
  // long chunks (FP64: the window and y[] fill the registers) store the
  // staged coefficients before their window loads
  if constexpr (CH > kChunk) stage();
  T f[CH + 4];
  load_window<T, IN, CH>(P, ib0, ri, is, s0, f);
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc && (q == 0 || q == Pn - 1))
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[fb + is * m];
  if constexpr (CH <= kChunk) stage();
  __syncthreads();
  const T* c = sco + q * CH * kCo;
  const T* rinv = sco + Pn * CH * kCo;
  T y[CH];
  chunk_rhs<T, OPK>(P, q, len, f, fb4, y);
  if (P.solve) {
    T yf, yl;
    chunk_solve(c, len, y, yf, yl);
    ex[lane + 64 * (2 * q)] = yf;
    ex[lane + 64 * (2 * q + 1)] = yl;
    __syncthreads();
    chunk_couple<T>(P, q, rinv, c, [&](int k) { return ex[lane + 64 * k]; }, y);
  }
  if (!live) return;
  T* const Ob = P.out + ob0;
  // accumulating outputs read the previous values in two batches of half a
  // chunk (fewer registers held across the loads)
  constexpr int H = CH / 2;
#pragma unroll
  for (int h = 0; h < CH; h += H) {
    T prev[H];
#pragma unroll
    for (int i = 0; i < H; ++i) prev[i] = T(0);
    if (P.accumulate) {
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const int r = (s0 + h + i < n_out) ? s0 + h + i : n_out - 1;
        prev[i] = ld_at<T>(Ob + os * r, ro);
      }
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
      if (h + i < len) {
        const int r = s0 + h + i;
        const T val = chunk_post<T>(P, c, h + i, r, y[h + i], fb4);
        st_at<T>(Ob + os * r, ro, prev[i] + val * P.sOut);
      }
    }
  }
}

__device__ __forceinline__ int lds_pos(int k) { return k + (k >> 4); }

template <typename T, int OPK>
__global__ void __launch_bounds__(256) k_line_x(KArgs<T> P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int NL = P.nl, LS = P.ls, Pn = P.P;
  T* exb = tile + (long)NL * LS;
  const int tid = threadIdx.x;
  const int nthr = NL * Pn;
  const int ll = tid / Pn, q = tid - ll * Pn;
  const long L0 = (long)blockIdx.x * NL;
  const int n_in = P.n_in, n_out = P.n_out;
  // stage NL lines (input transform applied)
  // (an x line L starts at L * e1 in the input and at L * o1 in the output)
  for (int e = tid; e < NL * n_in; e += nthr) {
    const int l = e / n_in, k = e - l * n_in;
    const long L = L0 + l;
    T v = T(0);
    if (L < P.nlines) {
      const long o = L * P.e1 + k;
      v = P.A[o];
      if (P.B) v = v + P.cB * P.B[o];
      v = P.sA * v;
      if (P.Cm) v = v * P.Cm[o];
      if (P.D) v = v + P.cD * P.D[o];
      if (P.E) v = v + P.cE * P.E[o];
    }
    tile[(long)l * LS + lds_pos(k)] = v;
  }
  __syncthreads();
  const long L = L0 + ll;
  const bool live = L < P.nlines;
  const T* row = tile + (long)ll * LS;
  const int s0 = (q * n_out) / Pn, len = ((q + 1) * n_out) / Pn - s0;
  T f[kChunk + 4];
#pragma unroll
  for (int t = 0; t < kChunk + 4; ++t) f[t] = row[lds_pos(window_index(s0 - 2 + t, n_in, P.wrap))];
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc && live && (q == 0 || q == Pn - 1))
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[4 * L + m];
  T y[kChunk];
  chunk_rhs<T, OPK>(P, q, len, f, fb4, y);
  if (P.solve) {
    T yf, yl;
    chunk_solve(P.co + q * kChunk * kCo, len, y, yf, yl);
    T* ex = exb + (long)ll * 2 * Pn;
    ex[2 * q] = yf;
    ex[2 * q + 1] = yl;
    __syncthreads();
    chunk_couple<T>(P, q, P.rinv, P.co + q * kChunk * kCo, [&](int k) { return ex[k]; }, y);
  }
  __syncthreads();  // every thread has finished reading its input rows
  T* orow = tile + (long)ll * LS;
#pragma unroll
  for (int i = 0; i < kChunk; ++i)
    if (i < len) orow[lds_pos(s0 + i)] = chunk_post<T>(P, P.co + q * kChunk * kCo, i, s0 + i, y[i], fb4);
  __syncthreads();
  for (int e = tid; e < NL * n_out; e += nthr) {
    const int l = e / n_out, k = e - l * n_out;
    const long Lw = L0 + l;
    if (Lw >= P.nlines) continue;
    const long o = Lw * P.o1 + k;
    const T val = tile[(long)l * LS + lds_pos(k)];
    const T prev = P.accumulate ? P.out[o] : T(0);
    P.out[o] = prev + val * P.sOut;
  }
}

// x lines with a wave-uniform chunk: a workgroup stages 64 whole lines in LDS
// (row stride odd, so the 64 lanes reading one column of their own lines hit
// distinct banks); wave q of the workgroup processes chunk q of all 64 lines,
// which keeps the plan's per-row coefficients in scalar registers.
template <typename T, int OPK, int IN>
__global__ void __launch_bounds__(64 * kMaxChunks) k_line_xu(KArgs<T> P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int LS = P.ls, Pn = P.P;
  // after the windows are read, the start of the tile holds the exchanged
  // first/last chunk values (64 x 2P)
  T* exb = tile;
  // per-plan coefficients (chunk rows, then the reduced-system inverse) after
  // the tile; their loads are issued first and written once staging is done
  T* sco = tile + (long)64 * LS;
  const int nst = Pn * kChunk * kCo + (P.rinv ? 4 * Pn * Pn : 0);
  T cv[kStageMax];
  {
    const int nco = Pn * kChunk * kCo;
#pragma unroll
    for (int u = 0; u < kStageMax; ++u) {
      const int idx = threadIdx.x + u * 64 * Pn;
      cv[u] = (idx < nco) ? P.co[idx] : (idx < nst ? P.rinv[idx - nco] : T(0));
    }
  }
  const int tid = threadIdx.x, nthr = 64 * Pn;
  const int lane = tid & 63;
  const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
  const long L0 = (long)blockIdx.x * 64;
  const int n_in = P.n_in, n_out = P.n_out;
  // staging: element e = l * n_in + k of the workgroup's 64 lines, thread
  // tid taking e = tid, tid + nthr, ...; (l, k) advance by a fixed step, so
  // no division per element.  Batches of UB elements, all loads of a batch
  // issued before any use; the line is clamped so every address is valid.
  constexpr int UB = 8;
  const long Lmax = P.nlines - 1;
  const int lmax = (int)((Lmax - L0) < 63 ? (Lmax - L0) : 63);  // last line of the workgroup
  const T* const Ab = P.A + L0 * P.e1;
  const T* const Bb = P.B ? P.B + L0 * P.e1 : nullptr;
  const T* const Cb = P.Cm ? P.Cm + L0 * P.e1 : nullptr;
  const T* const Db = P.D ? P.D + L0 * P.e1 : nullptr;
  const T* const Eb = P.E ? P.E + L0 * P.e1 : nullptr;
  {
    const int dl = nthr / n_in, dk = nthr - dl * n_in;
    int l = tid / n_in, k = tid - l * n_in;
    const int tin = 64 * n_in;
    for (int b0 = 0; b0 < tin; b0 += nthr * UB) {
      unsigned o[UB];
      int ll[UB], kk[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        ll[u] = l;
        kk[u] = k;
        o[u] = (unsigned)(((l < lmax ? l : lmax) * P.e1 + k) * (int)sizeof(T));
        k += dk;
        l += dl;
        if (k >= n_in) { k -= n_in; ++l; }
      }
      T v[UB];
      const T* S = Bb ? Bb : Cb;
      if (IN == 1 || (IN == 2 && S)) {
        T g[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
          v[u] = ld_at(Ab, o[u]);
          g[u] = ld_at(S, o[u]);
        }
        transform_in<T, UB>(P, v, g, [&](int u) { return ld_at(Cb, o[u]); });
      } else {
#pragma unroll
        for (int u = 0; u < UB; ++u) v[u] = P.sA * ld_at(Ab, o[u]);
      }
      if (IN == 2) {
        T d[UB], e[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
          d[u] = ld_at(Db, o[u]);
          e[u] = Eb ? ld_at(Eb, o[u]) : T(0);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) v[u] = v[u] + P.cD * d[u] + P.cE * e[u];
      }
#pragma unroll
      for (int u = 0; u < UB; ++u)
        if (ll[u] < 64) tile[ll[u] * LS + kk[u]] = v[u];
    }
  }
#pragma unroll
  for (int u = 0; u < kStageMax; ++u) {
    const int idx = threadIdx.x + u * 64 * Pn;
    if (idx < nst) sco[idx] = cv[u];
  }
  __syncthreads();
  const T* c = sco + q * kChunk * kCo;
  const T* rinv = sco + Pn * kChunk * kCo;
  const long L = L0 + lane;
  const bool live = L < P.nlines;
  T* row = tile + lane * LS;
  const int s0 = (q * n_out) / Pn, len = ((q + 1) * n_out) / Pn - s0;
  T f[kChunk + 4];
#pragma unroll
  for (int t = 0; t < kChunk + 4; ++t) f[t] = row[window_index(s0 - 2 + t, n_in, P.wrap)];
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc && live && (q == 0 || q == Pn - 1))
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[4 * L + m];
  T y[kChunk];
  chunk_rhs<T, OPK>(P, q, len, f, fb4, y);
  if (P.solve) {
    T yf, yl;
    chunk_solve(c, len, y, yf, yl);
    __syncthreads();  // every wave has read its window: the tile's start takes the exchange
    exb[lane + 64 * (2 * q)] = yf;
    exb[lane + 64 * (2 * q + 1)] = yl;
    __syncthreads();
    chunk_couple<T>(P, q, rinv, c, [&](int k) { return exb[lane + 64 * k]; }, y);
  }
  __syncthreads();  // every wave has finished reading its window and the exchange
#pragma unroll
  for (int i = 0; i < kChunk; ++i)
    if (i < len) row[s0 + i] = chunk_post<T>(P, c, i, s0 + i, y[i], fb4);
  __syncthreads();
  {
    const int dl = nthr / n_out, dk = nthr - dl * n_out;
    int l = tid / n_out, k = tid - l * n_out;
    const int tout = 64 * n_out;
    T* const Ob = P.out + L0 * P.o1;
    for (int b0 = 0; b0 < tout; b0 += nthr * UB) {
      unsigned o[UB];
      bool ok[UB];
      T val[UB], prev[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        ok[u] = (l <= lmax);
        const int lc = l < lmax ? l : lmax;
        o[u] = (unsigned)((lc * P.o1 + k) * (int)sizeof(T));
        val[u] = tile[lc * LS + k];
        prev[u] = T(0);
        k += dk;
        l += dl;
        if (k >= n_out) { k -= n_out; ++l; }
      }
      if (P.accumulate) {
#pragma unroll
        for (int u = 0; u < UB; ++u) prev[u] = ld_at<T>(Ob, o[u]);
      }
#pragma unroll
      for (int u = 0; u < UB; ++u)
        if (ok[u]) st_at<T>(Ob, o[u], prev[u] + val[u] * P.sOut);
    }
  }
}

template <typename T, int IN>
static void dispatch_xu_s(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (op) {
    case MIDP_C2P: hipLaunchKernelGGL((k_line_xu<T, MIDP_C2P, IN>), g, b, shm, s, a); break;
    case MIDP_P2C: hipLaunchKernelGGL((k_line_xu<T, MIDP_P2C, IN>), g, b, shm, s, a); break;
    case D1_C2C: hipLaunchKernelGGL((k_line_xu<T, D1_C2C, IN>), g, b, shm, s, a); break;
    case D1_P2P: hipLaunchKernelGGL((k_line_xu<T, D1_P2P, IN>), g, b, shm, s, a); break;
    case D1_C2P: hipLaunchKernelGGL((k_line_xu<T, D1_C2P, IN>), g, b, shm, s, a); break;
    case D1_P2C: hipLaunchKernelGGL((k_line_xu<T, D1_P2C, IN>), g, b, shm, s, a); break;
  }
}

// number of input arrays beyond A, as the kernels' IN parameter
template <typename T>
static int input_kind(const KArgs<T>& a) { return a.D ? 2 : ((a.B || a.Cm) ? 1 : 0); }

template <typename T>
static void dispatch_xu(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (input_kind(a)) {
    case 0: dispatch_xu_s<T, 0>(op, g, b, shm, s, a); break;
    case 1: dispatch_xu_s<T, 1>(op, g, b, shm, s, a); break;
    default: dispatch_xu_s<T, 2>(op, g, b, shm, s, a); break;
  }
}

// LDS bytes a workgroup may use, queried once per device.
static size_t max_lds_per_block() {
  static size_t cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (!cached[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
    cached[dev] = (size_t)v;
  }
  return cached[dev];
}

template <typename T, int DIR, int IN, int CH = kChunk>
static void dispatch_yz(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (op) {
    case MIDP_C2P: hipLaunchKernelGGL((k_line_yz<T, DIR, MIDP_C2P, IN, CH>), g, b, shm, s, a); break;
    case MIDP_P2C: hipLaunchKernelGGL((k_line_yz<T, DIR, MIDP_P2C, IN, CH>), g, b, shm, s, a); break;
    case D1_C2C: hipLaunchKernelGGL((k_line_yz<T, DIR, D1_C2C, IN, CH>), g, b, shm, s, a); break;
    case D1_P2P: hipLaunchKernelGGL((k_line_yz<T, DIR, D1_P2P, IN, CH>), g, b, shm, s, a); break;
    case D1_C2P: hipLaunchKernelGGL((k_line_yz<T, DIR, D1_C2P, IN, CH>), g, b, shm, s, a); break;
    case D1_P2C: hipLaunchKernelGGL((k_line_yz<T, DIR, D1_P2C, IN, CH>), g, b, shm, s, a); break;
  }
}

// lines of kChunkLong-row chunks (more than kChunk * kMaxChunks rows), any
// direction
template <typename T, int DIR>
static void dispatch_long(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (input_kind(a)) {
    case 0: dispatch_yz<T, DIR, 0, kChunkLong>(op, g, b, shm, s, a); break;
    case 1: dispatch_yz<T, DIR, 1, kChunkLong>(op, g, b, shm, s, a); break;
    default: dispatch_yz<T, DIR, 2, kChunkLong>(op, g, b, shm, s, a); break;
  }
}

template <typename T, int DIR>
static void dispatch_part(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  if constexpr (DIR == 0) {
    switch (op) {
      case MIDP_C2P: hipLaunchKernelGGL((k_line_x<T, MIDP_C2P>), g, b, shm, s, a); break;
      case MIDP_P2C: hipLaunchKernelGGL((k_line_x<T, MIDP_P2C>), g, b, shm, s, a); break;
      case D1_C2C: hipLaunchKernelGGL((k_line_x<T, D1_C2C>), g, b, shm, s, a); break;
      case D1_P2P: hipLaunchKernelGGL((k_line_x<T, D1_P2P>), g, b, shm, s, a); break;
      case D1_C2P: hipLaunchKernelGGL((k_line_x<T, D1_C2P>), g, b, shm, s, a); break;
      case D1_P2C: hipLaunchKernelGGL((k_line_x<T, D1_P2C>), g, b, shm, s, a); break;
    }
  } else {
    switch (input_kind(a)) {
      case 0: dispatch_yz<T, DIR, 0>(op, g, b, shm, s, a); break;
      case 1: dispatch_yz<T, DIR, 1>(op, g, b, shm, s, a); break;
      default: dispatch_yz<T, DIR, 2>(op, g, b, shm, s, a); break;
    }
  }
}

template <typename T, int DIR>
static void dispatch_op(int op, dim3 g, dim3 b, hipStream_t s, const KArgs<T>& a) {
  switch (op) {
    case MIDP_C2P: hipLaunchKernelGGL((k_line<T, DIR, MIDP_C2P>), g, b, 0, s, a); break;
    case MIDP_P2C: hipLaunchKernelGGL((k_line<T, DIR, MIDP_P2C>), g, b, 0, s, a); break;
    case D1_C2C: hipLaunchKernelGGL((k_line<T, DIR, D1_C2C>), g, b, 0, s, a); break;
    case D1_P2P: hipLaunchKernelGGL((k_line<T, DIR, D1_P2P>), g, b, 0, s, a); break;
    case D1_C2P: hipLaunchKernelGGL((k_line<T, DIR, D1_C2P>), g, b, 0, s, a); break;
    case D1_P2C: hipLaunchKernelGGL((k_line<T, DIR, D1_P2C>), g, b, 0, s, a); break;
  }
}

template <typename T>
std::string launch_line_op(const DevPlan<T>& D, const Ext3& ein, const Ext3& eout, const LineIn<T>& in, const T* fbc,
                           const LineOut<T>& out, T* scratch, hipStream_t s) {
  const LinePlan& p = D.host;
  const int dir = p.dir;
  if (ein.n[dir] != p.n_in || eout.n[dir] != p.n_out) return "line op: extent along the line differs from plan";
  for (int k = 0; k < 3; ++k)
    if (k != dir && ein.n[k] != eout.n[k]) return "line op: cross extents of input and output differ";
  if (p.has_fbc != (fbc != nullptr)) return "line op: fbc presence differs from plan";
  KArgs<T> a{};
  a.n_in = p.n_in; a.n_out = p.n_out;
  a.e1 = ein.n[0]; a.e2 = ein.n[1];
  a.o1 = eout.n[0]; a.o2 = eout.n[1];
  a.nlines = (dir == 0) ? (long)ein.n[1] * ein.n[2] : (dir == 1) ? (long)ein.n[0] * ein.n[2] : (long)ein.n[0] * ein.n[1];
  a.A = in.A; a.B = in.B; a.Cm = in.Cm; a.sA = in.sA; a.cB = in.cB;
  a.D = in.D; a.E = in.D ? in.E : nullptr; a.cD = in.cD; a.cE = in.cE;
  a.fbc = fbc;
  a.c1 = T(p.c1); a.c2 = T(p.c2); a.dd = T(p.dd); a.sgn = T(p.sgn);
  a.rows = D.rows;
  a.solve = p.solve; a.b0 = T(p.b0);
  a.a = D.a; a.c = D.c; a.d = D.d; a.pw = D.pw; a.x1 = D.x1;
  a.an = T(p.cyc_an); a.cn = T(p.cyc_cn); a.den = T(p.cyc_den);
  a.metric = D.metric; a.neu_lo = p.neu_lo; a.neu_hi = p.neu_hi;
  a.out = out.out; a.sOut = out.sOut; a.accumulate = out.accumulate;
  a.scratch = scratch;
  if (a.nlines <= 0) return "";
  if (a.nlines >= (1L << 31) - 64) return "line op: too many lines";
  if ((long)ein.n[0] * ein.n[1] * ein.n[2] >= (1L << 31) || (long)eout.n[0] * eout.n[1] * eout.n[2] >= (1L << 31))
    return "line op: arrays of 2^31 or more elements";
  if (p.part_P <= kMaxChunks && p.n_out >= 4 && D.win_ok && D.co) {
    a.P = p.part_P;
    a.per = (p.solve == 2);
    a.rinv = D.rinv;
    if (p.solve && !D.rinv) return "line op: plan solves but has no reduced-system inverse";
    a.swin = D.swin;
    a.wrap = (p.ibc[0] == BC_PERIODIC);
    a.co = D.co;
    const int nmax = p.n_in > p.n_out ? p.n_in : p.n_out;
    const int lsu = nmax | 1;  // odd row stride
    // tile of 64 lines (its start later holds the 64 x 2P exchange) and the
    // per-plan coefficients
    const size_t shm_u = sizeof(T) * ((size_t)64 * lsu + (size_t)p.part_P * kChunk * kCo +
                                      4 * (size_t)p.part_P * p.part_P);
    if (p.part_ch == kChunkLong) {
      // long lines: the lane-per-line kernel with chunks of kChunkLong rows;
      // per-lane byte offsets from a workgroup's first line stay below 2^32
      const long span = (dir == 0) ? 64L * (ein.n[0] > eout.n[0] ? ein.n[0] : eout.n[0])
                         : (dir == 1) ? (long)ein.n[0] * (ein.n[1] > eout.n[1] ? ein.n[1] : eout.n[1]) *
                                            (64 / ein.n[0] + 2)
                                      : 64L;
      if (span * (long)sizeof(T) >= (1L << 32)) return "line op: lines too far apart for 32-bit offsets";
      dim3 g((unsigned)((a.nlines + 63) / 64)), b((unsigned)(64 * p.part_P));
      const size_t shm = sizeof(T) * ((size_t)p.part_P * kChunkLong * kCo + 4 * (size_t)p.part_P * p.part_P);
      if (dir == 0) dispatch_long<T, 0>(p.op, g, b, shm, s, a);
      else if (dir == 1) dispatch_long<T, 1>(p.op, g, b, shm, s, a);
      else dispatch_long<T, 2>(p.op, g, b, shm, s, a);
    } else if (dir == 0) {
      // two tile shapes for x lines: 64 whole lines in LDS (k_line_xu) or
      // 256/P lines with skewed rows (k_line_x); timed once per shape when a
      // scratch output is at hand (chs_tune.h), else the 64-line tile
      const bool xu_fits = shm_u <= max_lds_per_block();
      int ls = lds_pos_host(nmax) + 1;
      ls += ((8 - ls) % 32 + 32) % 32;  // row stride = 8 (mod 32) doubles: neighbouring lines in other banks
      // lines per workgroup: at most 256 threads and 48 KiB of LDS
      int nl = 256 / p.part_P;
      const int cap = (int)((48 * 1024) / (sizeof(T) * (size_t)(ls + 2 * p.part_P)));
      if (nl > cap) nl = cap;
      auto run = [&](int v, KArgs<T> ka) -> bool {
        if (v == 0) {
          if (!xu_fits) return false;
          ka.ls = lsu;
          dim3 g((unsigned)((ka.nlines + 63) / 64)), b((unsigned)(64 * p.part_P));
          dispatch_xu<T>(p.op, g, b, shm_u, s, ka);
        } else {
          if (nl < 1) return false;
          ka.ls = ls;
          ka.nl = nl;
          const size_t shm = sizeof(T) * ((size_t)nl * ls + (size_t)nl * 2 * p.part_P);
          dim3 g((unsigned)((ka.nlines + nl - 1) / nl)), b((unsigned)(nl * p.part_P));
          dispatch_part<T, 0>(p.op, g, b, shm, s, ka);
        }
        return hipGetLastError() == hipSuccess;
      };
      int v = -1;
      if (scratch && xu_fits && nl >= 1) {
        KArgs<T> ts = a;
        ts.out = scratch;  // the candidates write the scratch array, never the output
        ts.accumulate = 0;
        const std::string key = std::string("line_x/") + (sizeof(T) == 8 ? "f64/" : "f32/") + std::to_string(p.op) +
                                "/" + std::to_string(p.n_in) + "/" + std::to_string(p.n_out) + "/" +
                                std::to_string(a.nlines) + "/" + std::to_string(input_kind(a));
        v = tune(key, 2, s, [&](int vv) { return run(vv, ts); });
      }
      if (v < 0) v = xu_fits ? 0 : 1;
      if (v == 1 && nl < 1) return "line op: line does not fit in LDS";
      if (!run(v, a)) return std::string("line op launch: ") + hipGetErrorString(hipGetLastError());
    } else {
      // per-lane byte offsets from a workgroup's first line stay below 2^32
      // (a y-line block spans at most 64/e1 + 2 planes of x-y)
      if (dir == 1) {
        const long e2m = ein.n[1] > eout.n[1] ? ein.n[1] : eout.n[1];
        if ((long)ein.n[0] * e2m * (64 / ein.n[0] + 2) * (long)sizeof(T) >= (1L << 32))
          return "line op: x-y plane too large for 32-bit line offsets";
      }
      dim3 g((unsigned)((a.nlines + 63) / 64)), b((unsigned)(64 * p.part_P));
      const size_t shm_yz = sizeof(T) * ((size_t)p.part_P * kChunk * kCo + 4 * (size_t)p.part_P * p.part_P);
      if (dir == 1) dispatch_part<T, 1>(p.op, g, b, shm_yz, s, a);
      else dispatch_part<T, 2>(p.op, g, b, shm_yz, s, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return std::string("line op launch: ") + hipGetErrorString(e);
    return "";
  }
  if (p.solve && !scratch) return "line op: scratch required";
  dim3 b(64), g((unsigned)((a.nlines + 63) / 64));
  if (dir == 0) dispatch_op<T, 0>(p.op, g, b, s, a);
  else if (dir == 1) dispatch_op<T, 1>(p.op, g, b, s, a);
  else dispatch_op<T, 2>(p.op, g, b, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("line op launch: ") + hipGetErrorString(e);
  return "";
}

// First and last output rows of a y-direction plan on every (x, z) line.
template <typename T>
__global__ void __launch_bounds__(64) k_wall_rows(KArgs<T> P, int e3, T* __restrict__ planes) {
  const long L = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= (long)P.e1 * e3) return;
  const long i = L % P.e1, k = L / P.e1;
  const long ib = i + (long)P.e1 * P.e2 * k, is = P.e1;
  const long fb = i + (long)P.e1 * 4 * k, fs = P.e1;
  auto F = [&](int q) -> T { return P.A[ib + is * q]; };
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc)
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[fb + fs * m];
  for (int side = 0; side < 2; ++side) {
    const int q = side ? 3 : 0;
    const int r = side ? P.n_out - 1 : 0;
    const DevRow<T>& R = P.rows[q];
    T v = T(0);
    for (int t = 0; t < R.nterm; ++t) v = v + R.w[t] * F(R.idx[t]);
    if (P.fbc)
      for (int m = 0; m < 4; ++m) v = v + R.u[m] * fb4[m];
    v = v * P.dd;
    if (P.solve) v = side ? v * P.d[r] : v / P.b0;  // decoupled end rows
    if (P.metric) v = v * P.metric[r];
    if (r == 0 && P.neu_lo) v = fb4[0];
    if (r == P.n_out - 1 && P.neu_hi) v = fb4[1];
    planes[i + (long)P.e1 * (side + 2 * k)] = v;
  }
}

template <typename T>
std::string launch_wall_rows(const DevPlan<T>& D, const Ext3& ein, const T* in, const T* fbc, T* planes,
                             hipStream_t s) {
  const LinePlan& p = D.host;
  if (p.dir != 1) return "wall rows: y plans only";
  if (ein.n[1] != p.n_in) return "wall rows: extent differs from plan";
  if (p.has_fbc != (fbc != nullptr)) return "wall rows: fbc presence differs from plan";
  if (p.solve == 2) return "wall rows: periodic line";
  if (p.solve == 1) {
    const int n = p.n_out;
    if (!(p.b0 == 1.0 && p.c[0] == 0.0 && p.a[n - 1] == 0.0 && p.d[n - 1] == 1.0))
      return "wall rows: end rows are coupled to the line";
  }
  KArgs<T> a{};
  a.n_in = p.n_in; a.n_out = p.n_out;
  a.e1 = ein.n[0]; a.e2 = ein.n[1];
  a.A = in;
  a.fbc = fbc;
  a.dd = T(p.dd);
  a.rows = D.rows;
  a.solve = p.solve; a.b0 = T(p.b0);
  a.d = D.d;
  a.metric = D.metric; a.neu_lo = p.neu_lo; a.neu_hi = p.neu_hi;
  const long n = (long)ein.n[0] * ein.n[2];
  if (n <= 0) return "";
  hipLaunchKernelGGL(k_wall_rows<T>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, a, ein.n[2], planes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return std::string("wall rows launch: ") + hipGetErrorString(e);
  return "";
}

template struct DevPlan<double>;
template struct DevPlan<float>;
template std::string upload_plan<double>(const LinePlan&, DevPlan<double>*);
template std::string upload_plan<float>(const LinePlan&, DevPlan<float>*);
template std::string launch_line_op<double>(const DevPlan<double>&, const Ext3&, const Ext3&, const LineIn<double>&,
                                            const double*, const LineOut<double>&, double*, hipStream_t);
template std::string launch_wall_rows<double>(const DevPlan<double>&, const Ext3&, const double*, const double*,
                                              double*, hipStream_t);
template std::string launch_wall_rows<float>(const DevPlan<float>&, const Ext3&, const float*, const float*, float*,
                                             hipStream_t);
template std::string launch_line_op<float>(const DevPlan<float>&, const Ext3&, const Ext3&, const LineIn<float>&,
                                           const float*, const LineOut<float>&, float*, hipStream_t);

}  // namespace chs
