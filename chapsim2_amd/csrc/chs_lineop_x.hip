// x-line kernels of the partitioned line operators: whole x lines (contiguous
// in memory) are staged in LDS with coalesced loads, every thread of a
// workgroup then evaluates one chunk of one line from LDS (explicit RHS,
// chunk-local Thomas sweeps, reduced-system coupling; the algebra of
// chs_lineop_dev.h), and the results return through LDS to coalesced stores.
#include "chs_lineop_dev.h"

namespace chs {

// 256/P lines per workgroup with skewed LDS rows (lds_pos), thread
// (line tid / P, chunk tid % P).
template <typename T, int OPK, int CH>
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
  T f[CH + 4];
#pragma unroll
  for (int t = 0; t < CH + 4; ++t) f[t] = row[lds_pos(window_index(s0 - 2 + t, n_in, P.wrap))];
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc && live && (q == 0 || q == Pn - 1))
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[4 * L + m];
  T y[CH];
  chunk_rhs<T, OPK>(P, q, len, f, fb4, y);
  if (P.solve) {
    T yf, yl;
    chunk_solve(P.co + q * CH * kCo, len, y, yf, yl);
    T* ex = exb + (long)ll * 2 * Pn;
    ex[2 * q] = yf;
    ex[2 * q + 1] = yl;
    __syncthreads();
    chunk_couple<T>(P, q, P.rinv, P.co + q * CH * kCo, [&](int k) { return ex[k]; }, y);
  }
  __syncthreads();  // every thread has finished reading its input rows
  T* orow = tile + (long)ll * LS;
#pragma unroll
  for (int i = 0; i < CH; ++i)
    if (i < len) orow[lds_pos(s0 + i)] = chunk_post<T>(P, P.co + q * CH * kCo, i, s0 + i, y[i], fb4);
  __syncthreads();
  for (int e = tid; e < NL * n_out; e += nthr) {
    const int l = e / n_out, k = e - l * n_out;
    const long Lw = L0 + l;
    if (Lw >= P.nlines) continue;
    const long o = Lw * P.o1 + k;
    const T val = tile[(long)l * LS + lds_pos(k)];
    const T prev = P.accumulate ? P.out[o] : T(0);
    const T v = prev + val * P.sOut;
    P.out[o] = v;
    if (P.out32) P.out32[o] = (float)v;
  }
}

// G consecutive values of one x line (G = 16 bytes / sizeof(T) where the
// lines' lengths and the arrays allow it, else 1).
template <typename T, int G>
struct alignas(sizeof(T) * G) Grp {
  T v[G];
};
template <typename T, int G>
__device__ __forceinline__ Grp<T, G> ld_grp(const T* base, unsigned boff) {
  return *reinterpret_cast<const Grp<T, G>*>(reinterpret_cast<const char*>(base) + boff);
}
// Stores of finished output are non-temporal: each value is written once
// and read by a later kernel, long after this tile's lines have left the
// cache (FP64 x operator 35.6 -> 34.0 us, FP32 23.8 -> 20.5 us at
// 256x128x256; non-temporal loads of the input were slower).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <typename T, int G>
__device__ __forceinline__ void st_grp(T* base, unsigned boff, const Grp<T, G>& g) {
  char* a = reinterpret_cast<char*>(base) + boff;
  if constexpr (sizeof(T) * G == 16) {
    u32x4 v;
    __builtin_memcpy(&v, &g, 16);
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(a));
  } else {
#pragma unroll
    for (int c = 0; c < G; ++c) __builtin_nontemporal_store(g.v[c], reinterpret_cast<T*>(a) + c);
  }
}

// Stage NLW whole x lines (input transform applied) into tile rows of
// stride LS: group e = l * (n_in / G) + kg of the workgroup's lines, thread
// tid taking e = tid, tid + nthr, ...; (l, kg) advance by a fixed step, so no
// division per group.  Batches of UB groups, all loads of a batch issued
// before any use; the line is clamped so every address is valid.
template <typename T, int IN, int NLW, int G>
__device__ __forceinline__ void stage_tile(const KArgs<T>& P, T* tile, long L0, int lmax) {
  constexpr int UB = G == 1 ? 8 : 4;
  const int LS = P.ls, n_in = P.n_in, ng = n_in / G;
  const int tid = threadIdx.x, nthr = NLW * P.P;
  const T* const Ab = P.A + L0 * P.e1;
  const T* const Bb = P.B ? P.B + L0 * P.e1 : nullptr;
  const T* const Cb = P.Cm ? P.Cm + L0 * P.e1 : nullptr;
  const T* const Db = P.D ? P.D + L0 * P.e1 : nullptr;
  const T* const Eb = P.E ? P.E + L0 * P.e1 : nullptr;
  const int dl = nthr / ng, dk = nthr - dl * ng;
  int l = tid / ng, k = tid - l * ng;
  const int tin = NLW * ng;
  for (int b0 = 0; b0 < tin; b0 += nthr * UB) {
    unsigned o[UB];
    int ll[UB], kk[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      ll[u] = l;
      kk[u] = k * G;
      o[u] = (unsigned)(((l < lmax ? l : lmax) * P.e1 + k * G) * (int)sizeof(T));
      k += dk;
      l += dl;
      if (k >= ng) { k -= ng; ++l; }
    }
    T v[UB * G];
    const T* S = Bb ? Bb : Cb;
    auto put = [&](T (&dst)[UB * G], const T* base) {
      Grp<T, G> g[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) g[u] = ld_grp<T, G>(base, o[u]);
#pragma unroll
      for (int u = 0; u < UB; ++u)
#pragma unroll
        for (int c = 0; c < G; ++c) dst[u * G + c] = g[u].v[c];
    };
    if (IN == 1 || (IN == 2 && S)) {
      T g[UB * G];
      put(v, Ab);
      put(g, S);
      transform_in<T, UB * G>(P, v, g, [&](int t) { return ld_at(Cb, o[t / G] + (unsigned)(sizeof(T) * (t % G))); });
    } else {
      put(v, Ab);
#pragma unroll
      for (int t = 0; t < UB * G; ++t) v[t] = P.sA * v[t];
    }
    if (IN == 2) {
      T d[UB * G], e[UB * G];
      put(d, Db);
      if (Eb) put(e, Eb);
      else
#pragma unroll
        for (int t = 0; t < UB * G; ++t) e[t] = T(0);
#pragma unroll
      for (int t = 0; t < UB * G; ++t) v[t] = v[t] + P.cD * d[t] + P.cE * e[t];
    }
#pragma unroll
    for (int u = 0; u < UB; ++u)
      if (ll[u] < NLW)
#pragma unroll
        for (int c = 0; c < G; ++c) tile[ll[u] * LS + kk[u] + c] = v[u * G + c];
  }
}

// The NLW lines of the tile (rows of stride LS) out to P.out with coalesced
// (accumulating) stores of G values per thread, as stage_tile reads them.
template <typename T, int NLW, int G, bool F32 = false>
__device__ __forceinline__ void store_tile(const KArgs<T>& P, const T* tile, long L0, int lmax) {
  constexpr int UB = G == 1 ? 8 : 4;
  const int LS = P.ls, n_out = P.n_out, ng = n_out / G;
  const int tid = threadIdx.x, nthr = NLW * P.P;
  const int dl = nthr / ng, dk = nthr - dl * ng;
  int l = tid / ng, k = tid - l * ng;
  const int tout = NLW * ng;
  T* const Ob = P.out + L0 * P.o1;
  float* const Ob32 = (F32 && P.out32) ? P.out32 + L0 * P.o1 : nullptr;
  for (int b0 = 0; b0 < tout; b0 += nthr * UB) {
    unsigned o[UB];
    bool ok[UB];
    Grp<T, G> val[UB], prev[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      ok[u] = (l <= lmax);
      const int lc = l < lmax ? l : lmax;
      o[u] = (unsigned)((lc * P.o1 + k * G) * (int)sizeof(T));
#pragma unroll
      for (int c = 0; c < G; ++c) {
        val[u].v[c] = tile[lc * LS + k * G + c];
        prev[u].v[c] = T(0);
      }
      k += dk;
      l += dl;
      if (k >= ng) { k -= ng; ++l; }
    }
    if (P.accumulate) {
#pragma unroll
      for (int u = 0; u < UB; ++u) prev[u] = ld_grp<T, G>(Ob, o[u]);
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      if (ok[u]) {
        Grp<T, G> r;
#pragma unroll
        for (int c = 0; c < G; ++c) r.v[c] = prev[u].v[c] + val[u].v[c] * P.sOut;
        st_grp<T, G>(Ob, o[u], r);
        if (F32 && Ob32) {
          Grp<float, G> r32;
#pragma unroll
          for (int c = 0; c < G; ++c) r32.v[c] = (float)r.v[c];
          st_grp<float, G>(Ob32, o[u] / (unsigned)(sizeof(T) / sizeof(float)), r32);
        }
      }
    }
  }
}

// The (CH + 4)-row window of a chunk starting at line row s0 - 2 from a
// staged line: interior chunks read it straight, the edge chunks through
// window_index (wrapped on periodic lines, clamped otherwise).
template <typename T, int CH>
__device__ __forceinline__ void read_window(const T* row, int s0, int n_in, int wrap, T (&f)[CH + 4]) {
  if (s0 >= 2 && s0 + CH + 1 < n_in) {
#pragma unroll
    for (int t = 0; t < CH + 4; ++t) f[t] = row[s0 - 2 + t];
  } else if (wrap) {
#pragma unroll
    for (int t = 0; t < CH + 4; ++t) {
      const int k = s0 - 2 + t;
      f[t] = row[k < 0 ? k + n_in : (k >= n_in ? k - n_in : k)];
    }
  } else {
#pragma unroll
    for (int t = 0; t < CH + 4; ++t) {
      const int k = s0 - 2 + t;
      f[t] = row[k < 0 ? 0 : (k > n_in - 1 ? n_in - 1 : k)];
    }
  }
}

// The rest of a staged x chunk after its RHS rows y[]: chunk-local solve,
// exchange of first/last values through the start of the tile (NLW x 2P,
// once every window has been read), spike correction, then the chunk's rows
// back into the tile and the NLW lines out with coalesced (accumulating)
// stores.
template <typename T, int CH, int NLW, bool F32 = false>
__device__ __forceinline__ void xu_finish(const KArgs<T>& P, const T* c, const T* rinv, T* tile, int lane, int q,
                                          int s0, int len, long L0, int lmax, T (&y)[CH], const T (&fb4)[4]) {
  const int LS = P.ls;
  T* exb = tile;
  T* row = tile + lane * LS;
  if (P.solve) {
    T yf, yl;
    chunk_solve(c, len, y, yf, yl);
    __syncthreads();  // every wave has read its window: the tile's start takes the exchange
    exb[lane + NLW * (2 * q)] = yf;
    exb[lane + NLW * (2 * q + 1)] = yl;
    __syncthreads();
    chunk_couple<T>(P, q, rinv, c, [&](int k) { return exb[lane + NLW * k]; }, y);
  }
  __syncthreads();  // every wave has finished reading its window and the exchange
#pragma unroll
  for (int i = 0; i < CH; ++i)
    if (i < len) row[s0 + i] = chunk_post<T>(P, c, i, s0 + i, y[i], fb4);
  __syncthreads();
  if (P.vec) store_tile<T, NLW, 16 / sizeof(T), F32>(P, tile, L0, lmax);
  else store_tile<T, NLW, 1, F32>(P, tile, L0, lmax);
}

// Waves per SIMD the tile kernels are compiled for: the 32-line tile of
// 16-row chunks (at most 8 waves) keeps to 128 registers so that two of its
// workgroups share a CU (one loads while the other computes).
__host__ __device__ constexpr int xu_waves(int ch, int nlw) { return (nlw <= 32 && ch == kChunk) ? 4 : 1; }

// NLW whole lines per workgroup of NLW x P threads in LDS (row stride
// ls = n | 1), thread tid on line tid % NLW and chunk tid / NLW.  NLW = 64:
// one chunk per wave, the chunk coefficient rows staged in LDS after the
// tile (wave-uniform reads).  NLW = 32: two chunks per wave and only the
// reduced-system inverse in LDS, the coefficient rows read through the vector
// cache; the tile takes half the LDS of the 64-line form, so lines twice as
// long fit, or two workgroups share a CU.  NLW = 16: four chunks per wave,
// nothing but the tile in LDS (the inverse too is read through the vector
// cache), four workgroups per CU.  IN as in load_window.
template <typename T, int OPK, int IN, int CH, int NLW>
__global__ void __launch_bounds__(NLW * kMaxChunks) __attribute__((amdgpu_waves_per_eu(xu_waves(CH, NLW))))
    k_line_xu(KArgs<T> P) {
  constexpr bool kCoLds = (NLW == 64);
  constexpr bool kRinvLds = (NLW >= 32);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int LS = P.ls, Pn = P.P;
  // after the windows are read, the start of the tile holds the exchanged
  // first/last chunk values (NLW x 2P)
  T* exb = tile;
  // per-plan values after the tile (chunk rows if kCoLds, then the
  // reduced-system inverse); their loads are issued first and written once
  // staging is done
  T* sco = tile + (long)NLW * LS;
  const int nco = kCoLds ? Pn * CH * kCo : 0;
  const int nst = nco + ((kRinvLds && P.rinv) ? 4 * Pn * Pn : 0);
  constexpr int NS = kCoLds ? stage_max(CH) : (kRinvLds ? (4 * kMaxChunks + NLW - 1) / NLW : 1);
  static_assert(!kRinvLds ||
                    NLW * kMaxChunks * NS >= (kCoLds ? kMaxChunks * CH * kCo : 0) + 4 * kMaxChunks * kMaxChunks,
                "per-plan staging exceeds NS values per thread");
  const int tid = threadIdx.x, nthr = NLW * Pn;
  T cv[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int idx = tid + u * nthr;
    cv[u] = (idx < nco) ? P.co[idx] : (idx < nst ? P.rinv[idx - nco] : T(0));
  }
  const int lane = tid & (NLW - 1);
  const int q = kCoLds ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid / NLW;
  const long L0 = (long)blockIdx.x * NLW;
  const int n_in = P.n_in, n_out = P.n_out;
  const long Lmax = P.nlines - 1;
  const int lmax = (int)((Lmax - L0) < NLW - 1 ? (Lmax - L0) : NLW - 1);  // last line of the workgroup
  if (P.vec) stage_tile<T, IN, NLW, 16 / sizeof(T)>(P, tile, L0, lmax);
  else stage_tile<T, IN, NLW, 1>(P, tile, L0, lmax);
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int idx = tid + u * nthr;
    if (idx < nst) sco[idx] = cv[u];
  }
  // equal chunks (co_uni): the one set of chunk rows in LDS after the
  // inverse, read by every chunk instead of its own rows through the cache
  T* scu = sco + nst;
  if (!kCoLds && P.co_uni)
    for (int idx = tid; idx < CH * kCo; idx += nthr) scu[idx] = P.co[idx];
  __syncthreads();
  const T* c = (kCoLds ? sco : P.co) + q * CH * kCo;
  const T* rinv = kRinvLds ? sco + nco : P.rinv;
  const long L = L0 + lane;
  const bool live = L < P.nlines;
  const T* row = tile + lane * LS;
  const int s0 = (q * n_out) / Pn, len = ((q + 1) * n_out) / Pn - s0;
  T f[CH + 4];
  read_window<T, CH>(row, s0, n_in, P.wrap, f);
  T fb4[4] = {T(0), T(0), T(0), T(0)};
  if (P.fbc && live && (q == 0 || q == Pn - 1))
    for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[4 * L + m];
  T y[CH];
  chunk_rhs<T, OPK>(P, q, len, f, fb4, y);
  if (!kCoLds && P.co_uni)
    xu_finish<T, CH, NLW, IN == 0 && sizeof(T) == 8>(P, scu, rinv, tile, lane, q, s0, len, L0, lmax, y, fb4);
  else
    xu_finish<T, CH, NLW, IN == 0 && sizeof(T) == 8>(P, c, rinv, tile, lane, q, s0, len, L0, lmax, y, fb4);
}

// Persistent, software-pipelined form of k_line_xu for lines of 16-row
// chunks read and written 16 bytes per access.  A workgroup of NLW x P
// threads loops over tiles of NLW lines (tile = blockIdx.x, + gridDim.x,
// ...); while one tile is solved and stored, the raw input of its next tile
// is already in flight into registers (JM groups of 16 bytes per thread and
// input array), so the loads overlap the arithmetic and the stores instead of
// forming a phase of their own.  IN = 0 (A) or 1 (A and one of B, Cm).  The
// chunk rows and the reduced-system inverse are copied into LDS after the
// tile once, for the workgroup's life.  The global loads a tile's solve
// waits for (the boundary values fbc) are issued before the next tile's
// prefetch, so waiting for them does not wait for the prefetch.
template <typename T, int OPK, int IN, int NLW>
__global__ void __launch_bounds__(NLW * kMaxChunks) k_line_xp(KArgs<T> P) {
  constexpr int CH = kChunk;
  constexpr int G = 16 / sizeof(T);
  constexpr int JM = CH / G;  // groups per thread and array: ceil(n_in / (G P)) <= JM (host-checked)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int LS = P.ls, Pn = P.P;
  T* sco = tile + (long)NLW * LS;
  const int nco = Pn * CH * kCo;
  const int nst = nco + (P.rinv ? 4 * Pn * Pn : 0);
  const int tid = threadIdx.x, nthr = NLW * Pn;
  for (int idx = tid; idx < nst; idx += nthr) sco[idx] = idx < nco ? P.co[idx] : P.rinv[idx - nco];
  const int lane = tid & (NLW - 1);
  const int q = tid / NLW;
  const int n_in = P.n_in, n_out = P.n_out, ng = n_in / G;
  const long ntile = (P.nlines + NLW - 1) / NLW;
  const T* S = IN ? (P.B ? P.B : P.Cm) : nullptr;
  // this thread's groups e = tid + j nthr of a tile: line lj, first row kj
  int lj[JM], kj[JM];
  {
    const int dl = nthr / ng, dk = nthr - dl * ng;
    int l = tid / ng, k = tid - l * ng;
#pragma unroll
    for (int j = 0; j < JM; ++j) {
      lj[j] = l;
      kj[j] = k * G;
      k += dk;
      l += dl;
      if (k >= ng) { k -= ng; ++l; }
    }
  }
  Grp<T, G> ra[JM], rb[JM];
  // every load of a tile is issued (clamped to the tile's last line, so each
  // address is valid); groups past the tile are not written
  auto issue = [&](long t) {
    const long L0 = t * NLW;
    const int lmax = (int)((P.nlines - 1 - L0) < NLW - 1 ? (P.nlines - 1 - L0) : NLW - 1);
    const T* Ab = P.A + L0 * P.e1;
    const T* Sb = IN ? S + L0 * P.e1 : nullptr;
#pragma unroll
    for (int j = 0; j < JM; ++j) {
      const int lc = lj[j] < lmax ? lj[j] : lmax;
      const unsigned o = (unsigned)((lc * P.e1 + kj[j]) * (int)sizeof(T));
      ra[j] = ld_grp<T, G>(Ab, o);
      if (IN) rb[j] = ld_grp<T, G>(Sb, o);
    }
  };
  long t = blockIdx.x;
  if (t < ntile) issue(t);
  __syncthreads();  // coefficient rows and inverse in LDS
  const T* c = sco + q * CH * kCo;
  const T* rinv = sco + nco;
  const int s0 = (q * n_out) / Pn, len = ((q + 1) * n_out) / Pn - s0;
  T* row = tile + lane * LS;
  for (; t < ntile; t += gridDim.x) {
    const long L0 = t * NLW;
    const int lmax = (int)((P.nlines - 1 - L0) < NLW - 1 ? (P.nlines - 1 - L0) : NLW - 1);
    // the tile, input transform applied (sA (A + cB B) or sA A Cm)
#pragma unroll
    for (int j = 0; j < JM; ++j) {
      if (lj[j] < NLW) {
#pragma unroll
        for (int cc = 0; cc < G; ++cc) {
          T v = ra[j].v[cc];
          if (IN && P.B) v = v + P.cB * rb[j].v[cc];
          v = P.sA * v;
          if (IN && !P.B) v = v * rb[j].v[cc];
          tile[lj[j] * LS + kj[j] + cc] = v;
        }
      }
    }
    const long L = L0 + lane;
    const bool live = L < P.nlines;
    T fb4[4] = {T(0), T(0), T(0), T(0)};
    if (P.fbc && live && (q == 0 || q == Pn - 1))
      for (int m = 0; m < 4; ++m) fb4[m] = P.fbc[4 * L + m];
    __syncthreads();
    // the next tile's input (the current one again on the last pass: every
    // wave issues the same number of loads on every pass)
    const long tn = t + gridDim.x < ntile ? t + gridDim.x : t;
    issue(tn);
    T f[CH + 4];
    read_window<T, CH>(row, s0, n_in, P.wrap, f);
    T y[CH];
    chunk_rhs<T, OPK>(P, q, len, f, fb4, y);
    xu_finish<T, CH, NLW, IN == 0 && sizeof(T) == 8>(P, c, rinv, tile, lane, q, s0, len, L0, lmax, y, fb4);
    __syncthreads();  // the stores have read the tile
  }
}

// Two operators of one line plan geometry on one input (PAIR 0 = MIDP_P2C +
// D1_P2C, 1 = MIDP_C2P + D1_C2P; plans P and Q with the same extents, chunk
// partition and boundary wrap) in one pass over x lines staged as in
// k_line_xu: the tile is loaded once, every thread keeps its window in
// registers, and each operator then runs the RHS, solve and store with its
// own plan and output (the tile is reused for the first operator's results
// once every window is read).  Per-plan values: NLW = 64 both plans' chunk
// rows and inverses in LDS after the tile; NLW = 32 none (read through the
// vector cache), so that two workgroups of FP64 lines of 256 rows share a CU.
template <typename T, int PAIR, int CH, int NLW>
__global__ void __launch_bounds__(NLW * kMaxChunks) __attribute__((amdgpu_waves_per_eu(xu_waves(CH, NLW))))
    k_line_xu2(KArgs<T> P, KArgs<T> Q) {
  constexpr int OPA = PAIR == 0 ? MIDP_P2C : MIDP_C2P;
  constexpr int OPB = PAIR == 0 ? D1_P2C : D1_C2P;
  constexpr bool kCoLds = (NLW == 64);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* tile = reinterpret_cast<T*>(smem_raw);
  const int LS = P.ls, Pn = P.P;
  T* scoP = tile + (long)NLW * LS;
  const int nco = kCoLds ? Pn * CH * kCo : 0;
  const int nstP = kCoLds ? nco + (P.rinv ? 4 * Pn * Pn : 0) : 0;
  const int nstQ = kCoLds ? nco + (Q.rinv ? 4 * Pn * Pn : 0) : 0;
  T* scoQ = scoP + nstP;
  constexpr int NS = kCoLds ? stage_max(CH) : 1;
  static_assert(!kCoLds || NLW * kMaxChunks * NS >= kMaxChunks * CH * kCo + 4 * kMaxChunks * kMaxChunks,
                "per-plan staging exceeds NS values per thread");
  const int tid = threadIdx.x, nthr = NLW * Pn;
  T cvP[NS], cvQ[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int idx = tid + u * nthr;
    cvP[u] = (idx < nco) ? P.co[idx] : (idx < nstP ? P.rinv[idx - nco] : T(0));
    cvQ[u] = (idx < nco) ? Q.co[idx] : (idx < nstQ ? Q.rinv[idx - nco] : T(0));
  }
  const int lane = tid & (NLW - 1);
  const int q = kCoLds ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid / NLW;
  const long L0 = (long)blockIdx.x * NLW;
  const int n_in = P.n_in, n_out = P.n_out;
  const long Lmax = P.nlines - 1;
  const int lmax = (int)((Lmax - L0) < NLW - 1 ? (Lmax - L0) : NLW - 1);
  if (P.vec) stage_tile<T, 0, NLW, 16 / sizeof(T)>(P, tile, L0, lmax);
  else stage_tile<T, 0, NLW, 1>(P, tile, L0, lmax);
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int idx = tid + u * nthr;
    if (idx < nstP) scoP[idx] = cvP[u];
    if (idx < nstQ) scoQ[idx] = cvQ[u];
  }
  __syncthreads();
  const T* cP = (kCoLds ? scoP : P.co) + q * CH * kCo;
  const T* cQ = (kCoLds ? scoQ : Q.co) + q * CH * kCo;
  const T* rP = kCoLds ? scoP + nco : P.rinv;
  const T* rQ = kCoLds ? scoQ + nco : Q.rinv;
  const long L = L0 + lane;
  const bool live = L < P.nlines;
  const T* row = tile + lane * LS;
  const int s0 = (q * n_out) / Pn, len = ((q + 1) * n_out) / Pn - s0;
  T f[CH + 4];
  read_window<T, CH>(row, s0, n_in, P.wrap, f);
  T fbP[4] = {T(0), T(0), T(0), T(0)}, fbQ[4] = {T(0), T(0), T(0), T(0)};
  if (live && (q == 0 || q == Pn - 1)) {
    if (P.fbc)
      for (int m = 0; m < 4; ++m) fbP[m] = P.fbc[4 * L + m];
    if (Q.fbc)
      for (int m = 0; m < 4; ++m) fbQ[m] = Q.fbc[4 * L + m];
  }
  T y[CH];
  chunk_rhs<T, OPA>(P, q, len, f, fbP, y);
  xu_finish<T, CH, NLW>(P, cP, rP, tile, lane, q, s0, len, L0, lmax, y, fbP);
  __syncthreads();  // the first operator's stores have read the tile
  chunk_rhs<T, OPB>(Q, q, len, f, fbQ, y);
  xu_finish<T, CH, NLW>(Q, cQ, rQ, tile, lane, q, s0, len, L0, lmax, y, fbQ);
}

template <typename T, int IN, int CH, int NLW>
static void tile_ops(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (op) {
    case MIDP_C2P: hipLaunchKernelGGL((k_line_xu<T, MIDP_C2P, IN, CH, NLW>), g, b, shm, s, a); break;
    case MIDP_P2C: hipLaunchKernelGGL((k_line_xu<T, MIDP_P2C, IN, CH, NLW>), g, b, shm, s, a); break;
    case D1_C2C: hipLaunchKernelGGL((k_line_xu<T, D1_C2C, IN, CH, NLW>), g, b, shm, s, a); break;
    case D1_P2P: hipLaunchKernelGGL((k_line_xu<T, D1_P2P, IN, CH, NLW>), g, b, shm, s, a); break;
    case D1_C2P: hipLaunchKernelGGL((k_line_xu<T, D1_C2P, IN, CH, NLW>), g, b, shm, s, a); break;
    case D1_P2C: hipLaunchKernelGGL((k_line_xu<T, D1_P2C, IN, CH, NLW>), g, b, shm, s, a); break;
  }
}

template <typename T, int CH, int NLW>
static void tile_in(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (input_kind(a)) {
    case 0: tile_ops<T, 0, CH, NLW>(op, g, b, shm, s, a); break;
    case 1: tile_ops<T, 1, CH, NLW>(op, g, b, shm, s, a); break;
    default: tile_ops<T, 2, CH, NLW>(op, g, b, shm, s, a); break;
  }
}

template <typename T>
bool launch_x_tile(int op, int ch, int nlw, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  if (nlw != 64 && nlw != 32 && nlw != 16) return false;
  if (ch == kChunk) {
    if (nlw == 64) tile_in<T, kChunk, 64>(op, g, b, shm, s, a);
    else if (nlw == 32) tile_in<T, kChunk, 32>(op, g, b, shm, s, a);
    else tile_in<T, kChunk, 16>(op, g, b, shm, s, a);
    return true;
  }
  if (ch != kChunkLong) return false;
  if (nlw == 32) {
    tile_in<T, kChunkLong, 32>(op, g, b, shm, s, a);
    return true;
  }
  if (nlw == 16) {
    // 16 lines of long chunks: an FP64 tile of 512-row lines takes 66 KB, so
    // two workgroups share a CU where one 32-line tile fills it
    tile_in<T, kChunkLong, 16>(op, g, b, shm, s, a);
    return true;
  }
  // 64 lines of long chunks: FP32 only (an FP64 tile exceeds the LDS)
  if constexpr (sizeof(T) == 4) {
    tile_in<T, kChunkLong, 64>(op, g, b, shm, s, a);
    return true;
  }
  return false;
}

template <typename T, int IN, int NLW>
static void pipe_ops(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (op) {
    case MIDP_C2P: hipLaunchKernelGGL((k_line_xp<T, MIDP_C2P, IN, NLW>), g, b, shm, s, a); break;
    case MIDP_P2C: hipLaunchKernelGGL((k_line_xp<T, MIDP_P2C, IN, NLW>), g, b, shm, s, a); break;
    case D1_C2C: hipLaunchKernelGGL((k_line_xp<T, D1_C2C, IN, NLW>), g, b, shm, s, a); break;
    case D1_P2P: hipLaunchKernelGGL((k_line_xp<T, D1_P2P, IN, NLW>), g, b, shm, s, a); break;
    case D1_C2P: hipLaunchKernelGGL((k_line_xp<T, D1_C2P, IN, NLW>), g, b, shm, s, a); break;
    case D1_P2C: hipLaunchKernelGGL((k_line_xp<T, D1_P2C, IN, NLW>), g, b, shm, s, a); break;
  }
}

template <typename T>
bool launch_x_pipe(int op, int nlw, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  constexpr int G = 16 / sizeof(T);
  if (nlw != 32 || !a.vec || a.D || (a.B && a.Cm)) return false;
  if (a.n_in / G > (kChunk / G) * a.P) return false;  // more groups per thread than the kernel holds
  if (a.B || a.Cm) pipe_ops<T, 1, 32>(op, g, b, shm, s, a);
  else pipe_ops<T, 0, 32>(op, g, b, shm, s, a);
  return true;
}

template <typename T, int CH>
static void skew_ops(int op, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  switch (op) {
    case MIDP_C2P: hipLaunchKernelGGL((k_line_x<T, MIDP_C2P, CH>), g, b, shm, s, a); break;
    case MIDP_P2C: hipLaunchKernelGGL((k_line_x<T, MIDP_P2C, CH>), g, b, shm, s, a); break;
    case D1_C2C: hipLaunchKernelGGL((k_line_x<T, D1_C2C, CH>), g, b, shm, s, a); break;
    case D1_P2P: hipLaunchKernelGGL((k_line_x<T, D1_P2P, CH>), g, b, shm, s, a); break;
    case D1_C2P: hipLaunchKernelGGL((k_line_x<T, D1_C2P, CH>), g, b, shm, s, a); break;
    case D1_P2C: hipLaunchKernelGGL((k_line_x<T, D1_P2C, CH>), g, b, shm, s, a); break;
  }
}

template <typename T>
void launch_x_skew(int op, int ch, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a) {
  if (ch == kChunk) skew_ops<T, kChunk>(op, g, b, shm, s, a);
  else skew_ops<T, kChunkLong>(op, g, b, shm, s, a);
}

template <typename T, int CH, int NLW>
static void tile2_pair(int pair, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a, const KArgs<T>& a2) {
  if (pair == 0) hipLaunchKernelGGL((k_line_xu2<T, 0, CH, NLW>), g, b, shm, s, a, a2);
  else hipLaunchKernelGGL((k_line_xu2<T, 1, CH, NLW>), g, b, shm, s, a, a2);
}

template <typename T>
bool launch_x_tile2(int pair, int ch, int nlw, dim3 g, dim3 b, size_t shm, hipStream_t s, const KArgs<T>& a,
                    const KArgs<T>& a2) {
  if (pair != 0 && pair != 1) return false;
  if (ch != kChunk) return false;
  if (nlw == 64) tile2_pair<T, kChunk, 64>(pair, g, b, shm, s, a, a2);
  else if (nlw == 32) tile2_pair<T, kChunk, 32>(pair, g, b, shm, s, a, a2);
  else return false;
  return true;
}

template bool launch_x_tile<double>(int, int, int, dim3, dim3, size_t, hipStream_t, const KArgs<double>&);
template bool launch_x_tile<float>(int, int, int, dim3, dim3, size_t, hipStream_t, const KArgs<float>&);
template bool launch_x_tile2<double>(int, int, int, dim3, dim3, size_t, hipStream_t, const KArgs<double>&,
                                     const KArgs<double>&);
template bool launch_x_tile2<float>(int, int, int, dim3, dim3, size_t, hipStream_t, const KArgs<float>&,
                                    const KArgs<float>&);
template bool launch_x_pipe<double>(int, int, dim3, dim3, size_t, hipStream_t, const KArgs<double>&);
template bool launch_x_pipe<float>(int, int, dim3, dim3, size_t, hipStream_t, const KArgs<float>&);
template void launch_x_skew<double>(int, int, dim3, dim3, size_t, hipStream_t, const KArgs<double>&);
template void launch_x_skew<float>(int, int, dim3, dim3, size_t, hipStream_t, const KArgs<float>&);

}  // namespace chs
