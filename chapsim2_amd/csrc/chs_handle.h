// Handle of the C ABI: mesh, cached device plans, stream, and the
// device-resident flow solver.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/chapsim_hip.h"
#include "chs_geom.h"
#include "chs_lineop.h"

namespace chs {

void set_error(const std::string& msg);

// A growable device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  std::string ensure(size_t b);
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

using PlanKey = std::tuple<int, int, int, int, int, int, int, int>;

class Poisson;
class Flow;

struct Handle {
  chs_desc_t desc{};
  Mesh mesh;
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<PlanKey, std::unique_ptr<DevPlan<double>>> plans64;
  std::map<PlanKey, std::unique_ptr<DevPlan<float>>> plans32;
  DevBuf io[4];
  std::unique_ptr<Poisson> poisson;
  std::unique_ptr<Flow> flow;
  // launch timing of one kernel class (0/1/2: x/y/z line operators) with
  // events on the handle's stream, and the algorithmic bytes of those launches
  int prof_class = -1;
  std::vector<hipEvent_t> prof_ev;
  size_t prof_used = 0;
  double prof_bytes = 0.0;
  ~Handle();
  // event pair for the next timed launch (nullptr pair when the pool is full)
  bool prof_pair(hipEvent_t* a, hipEvent_t* b);

  template <typename T>
  std::string plan(int dir, int op, int iacc, const int ibc[2], bool has_fbc, int n_in, int n_out,
                   const DevPlan<T>** out);
};

}  // namespace chs

struct chs_handle {
  chs::Handle h;
};
