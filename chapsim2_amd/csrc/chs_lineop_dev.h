// Device-side pieces shared by the line-operator translation units: the
// kernel argument block, the chunk helpers of the partitioned kernels and the
// launchers of the x-line kernels (chs_lineop_x.hip).
#pragma once

#include "chs_lineop.h"

namespace chs {

// Line index read for window position k, -2 <= k <= n + CH + 1 (CH the chunk rows):
// wrapped on periodic lines (one period suffices, as periodic lines on the
// partitioned kernels have n >= CH + 2), clamped to the line otherwise.
__host__ __device__ __forceinline__ int window_index(int k, int n, bool wrap) {
  if (wrap) return k < 0 ? k + n : (k >= n ? k - n : k);
  return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
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
  float* out32;   // optional FP32 copy of every value stored to out (same layout)
  T* scratch;
  // partitioned form
  int P, per, ls, nl;
  int vec;        // x tile kernels: lines read and written 16 bytes per thread access
  const T* rinv;
  const T* swin;  // 4 x (CH + 4) boundary-row window weights (CH = the plan's chunk rows)
  int wrap;       // periodic line: windows wrap around instead of clamping
  const T* co;    // P x CH coefficient rows of kCo entries (see upload_plan)
  int co_uni;     // every chunk's rows equal chunk 0's
  int met;        // some row's output metric differs from 1
  int ends;       // the boundary rows have recipes of their own (else the bulk stencil on the wrapped window)
};

// bulk offsets (p1,p2,p3,p4) per operator kind, relative to the output row
template <int OPK> struct Offs;
template <> struct Offs<MIDP_C2P> { static constexpr int p1 = 0, p2 = -1, p3 = 1, p4 = -2; };
template <> struct Offs<MIDP_P2C> { static constexpr int p1 = 0, p2 = 1, p3 = -1, p4 = 2; };
template <> struct Offs<D1_C2C> { static constexpr int p1 = 1, p2 = -1, p3 = 2, p4 = -2; };
template <> struct Offs<D1_P2P> { static constexpr int p1 = 1, p2 = -1, p3 = 2, p4 = -2; };
template <> struct Offs<D1_C2P> { static constexpr int p1 = 0, p2 = -1, p3 = 1, p4 = -2; };
template <> struct Offs<D1_P2C> { static constexpr int p1 = 1, p2 = 0, p3 = 2, p4 = -1; };

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
  const bool first = q == 0 && P.ends, last = q == P.P - 1 && P.ends;
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
// (non-temporal: a line operator's output is written once and read by a
// later kernel, long after it has left the cache; the FP64 step at
// 256x128x256 went from 1289 to 1334 Mpts.steps/s with it)
template <typename T>
__device__ __forceinline__ void st_at(T* base, unsigned boff, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff));
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
  const int qm = q == 0 ? Pn - 1 : q - 1, qp = q == Pn - 1 ? 0 : q + 1;
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
  if (P.met) val = val * c[i * kCo + CO_MET];
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

__device__ __forceinline__ int lds_pos(int k) { return k + (k >> 4); }


// number of input arrays beyond A, as the kernels' IN parameter
template <typename T>
inline int input_kind(const KArgs<T>& a) { return a.D ? 2 : ((a.B || a.Cm) ? 1 : 0); }

// x-line kernels (chs_lineop_x.hip).  Whole lines staged in LDS, NLW (64 or
// 32) lines per workgroup of NLW x P threads, chunks of CH rows (false when
// there is no such kernel: 64 FP64 lines of long chunks):
template <typename T>
bool launch_x_tile(int op, int ch, int nlw, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a);
// Persistent, software-pipelined x tiles (k_line_xp): 32 whole lines per
// tile, a grid of g workgroups looping over the tiles, one input array or
// two (A and one of B, Cm), 16-byte accesses, chunks of kChunk rows (false
// when the operator is not of that form):
template <typename T>
bool launch_x_pipe(int op, int nlw, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a);
// Two operators of one plan geometry on one input (k_line_xu2; pair 0 =
// MIDP_P2C + D1_P2C, 1 = MIDP_C2P + D1_C2P, plans a and a2), whole lines in
// LDS as above (nlw 64 or 32; chunks of kChunk rows only):
template <typename T>
bool launch_x_tile2(int pair, int ch, int nlw, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a,
                    const KArgs<T>& a2);
// 256/P lines per workgroup with skewed LDS rows, chunks of CH rows:
template <typename T>
void launch_x_skew(int op, int ch, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a);

}  // namespace chs
