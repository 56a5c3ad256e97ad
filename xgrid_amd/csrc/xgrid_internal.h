// Internal declarations shared by the host sources of libxgrid_hip.so
// (runtime, reductions, batches).  Not part of the C ABI.
#ifndef XGRID_INTERNAL_H
#define XGRID_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "xgrid_abi.h"
#include "xgrid_hip.h"

namespace xgi {

extern thread_local std::string t_err;

inline int fail(int code, const std::string& msg) {
  t_err = msg;
  return code;
}

#define XG_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::xgi::fail(XG_ERR_HIP, std::string(#call) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

struct Device {
  int id = -1;
  hipStream_t stream = nullptr;
  bool ready = false;
  int cus = 0;
  int64_t hbm = 0;
  char name[256] = {0};
  int64_t fuse_max_cells = 0;   // a chained statement walks at most this many cells
};
extern Device g_dev;

inline int check_ready() {
  if (!g_dev.ready) return fail(XG_ERR_STATE, "xg_init has not been called");
  return XG_OK;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// hipMalloc + zero fill on the backend stream
int alloc_zero(size_t bytes, void** out);

// Look up this device's address of the reductions' static scratch (device
// `device` current); called once per device by xg_init.
int reduce_setup(int device);

}  // namespace xgi

// Device plan of one boundary class of a grid.
struct ClassPlan {
  int32_t kind = XG_DOM_EMPTY;
  int64_t count = 0;
  int32_t lo[3] = {0, 0, 0};
  int32_t hi[3] = {1, 1, 1};
  int32_t* d_list = nullptr;
};

struct xg_grid {
  int32_t dims = 1;
  int32_t shape[3] = {1, 1, 1};
  int64_t n = 0;
  int32_t elem = 0;
  std::vector<void*> levels;
  void* scratch = nullptr;
  std::vector<int32_t> h_mask;  // empty: every cell is class 0
  int32_t* d_mask = nullptr;
  uint64_t mask_version = 0;
  std::unordered_map<int32_t, ClassPlan> plans;

  size_t bytes() const { return (size_t)n * (size_t)elem; }
  const ClassPlan& plan(int32_t k) const {
    static const ClassPlan empty;
    auto it = plans.find(k);
    return it == plans.end() ? empty : it->second;
  }
};

#endif
