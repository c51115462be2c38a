// Device-side line operator: applies a planned compact operator (explicit
// RHS stencil + pre-factored tridiagonal solve) to every line of a 3-D array
// in one direction, with a fused elementwise input transform and a fused
// accumulating output.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "chs_plan.h"

namespace chs {

template <typename T>
struct DevRow {
  int nterm;
  int idx[kMaxRowTerms];
  T w[kMaxRowTerms];
  T u[4];
};

// Device copy of a LinePlan in precision T.
constexpr int kCo = 8;  // entries per chunk coefficient row
// entries of a chunk coefficient row (pairs share a 16-byte slot)
enum CoEntry : int { CO_LD = 0, CO_AR = 1, CO_SV = 2, CO_SW = 3, CO_LC = 4, CO_MET = 5 };

template <typename T>
struct DevPlan {
  LinePlan host;
  DevRow<T>* rows = nullptr;  // 4 rows
  T* a = nullptr;             // n_out each
  T* c = nullptr;
  T* d = nullptr;
  T* pw = nullptr;
  T* x1 = nullptr;
  T* metric = nullptr;
  // partitioned solve (LinePlan::part_P chunks per line): the 2P x 2P
  // reduced-system inverse and the chunks' coefficient rows (kChunk rows of
  // kCo entries per chunk)
  T* rinv = nullptr;
  T* co = nullptr;
  // boundary rows 0, 1 (first chunk) and n-2, n-1 (last chunk) as dense
  // weights over the chunk's register window of kChunk + 4 inputs
  // (win_ok false when a recipe reaches outside that window)
  T* swin = nullptr;
  bool win_ok = false;
  // every chunk's coefficient rows are those of chunk 0 (equal chunks of a
  // periodic line), and whether any row carries an output metric other than 1
  bool co_uni = false;
  bool has_met = false;
  // the four boundary rows are the bulk stencil on the wrapped window
  // (periodic lines), so the kernels need not form them apart
  bool ends_bulk = false;
  ~DevPlan();
};

template <typename T>
std::string upload_plan(const LinePlan& p, DevPlan<T>* dp);

// Input transform f = sA * (A + cB * B) [* Cm] [+ cD * D [+ cE * E]]; B, Cm, D, E optional.
template <typename T>
struct LineIn {
  const T* A = nullptr;
  const T* B = nullptr;
  const T* Cm = nullptr;
  T sA = T(1);
  T cB = T(1);
  // additive terms: f += cD * D + cE * E (E only together with D); a sum of
  // operators that share one plan is applied as one operator on the summed
  // input
  const T* D = nullptr;
  const T* E = nullptr;
  T cD = T(0);
  T cE = T(0);
};

// Output out = (accumulate ? out : 0) + value * sOut
template <typename T>
struct LineOut {
  T* out = nullptr;
  T sOut = T(1);
  int accumulate = 0;
  float* out32 = nullptr;  // also store each result rounded to FP32 here (single operators)
};

// Extents of the input / output arrays (Fortran order, x fastest).
struct Ext3 {
  int n[3];
};

template <typename T>
std::string launch_line_op(const DevPlan<T>& P, const Ext3& ein, const Ext3& eout, const LineIn<T>& in,
                           const T* fbc, const LineOut<T>& out, T* scratch, hipStream_t s);

// Two operators of one line geometry on the same input in one pass: the
// interpolation and the first derivative between the same locations
// (MIDP_P2C with D1_P2C, or MIDP_C2P with D1_C2P; any order), whose plans
// share the extents, chunk partition and boundary wrap, each with its own
// wall values and output.  *passes: 1 when both ran in one pass, 2 when they
// ran one by one (x lines where that was timed faster; scratch takes the
// timing candidates' output), 0 with no error and nothing launched when the
// pair cannot share a pass (the caller then applies the operators itself).
// check_only: only decide *passes (timing the x shapes once if needed).
template <typename T>
std::string launch_line_op2(const DevPlan<T>& A, const DevPlan<T>& B, const Ext3& ein, const Ext3& eout,
                            const T* in, const T* fbcA, const T* fbcB, const LineOut<T>& outA,
                            const LineOut<T>& outB, T* scratch, hipStream_t s, int* passes,
                            bool check_only = false);

// Evaluate only the first and last output rows of a y-direction plan on every
// (x, z) line, writing planes of extents (e1, 2, e3).  Valid when the plan's
// end rows are decoupled from the line (explicit rows, or LHS rows (0,1,0)),
// as for Dirichlet walls with compact schemes.
template <typename T>
std::string launch_wall_rows(const DevPlan<T>& P, const Ext3& ein, const T* in, const T* fbc, T* planes,
                             hipStream_t s);

}  // namespace chs
