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
  hipStream_t stream = nullptr;   // every kernel launch
  hipStream_t copy = nullptr;     // grid memory set-up and copies that need not wait for the kernels
  hipMemPool_t pool = nullptr;    // grid and plan memory (dev_alloc / dev_free)
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

// Grid and plan memory: taken from the device's pool on the copy stream (cleared
// there when `zero`) and ready for any work enqueued after the call returns;
// released on the kernel stream behind the work already queued there.  Neither
// waits for the kernel stream.
int dev_alloc(size_t bytes, bool zero, void** out);
void dev_free(void* p);
inline int alloc_zero(size_t bytes, void** out) { return dev_alloc(bytes, true, out); }

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
  bool bound = false;     // a kernel plan has used it (or its level pointer was handed out)
  bool exported = false;  // its level pointer was handed out: callers may use it on their own streams

  size_t bytes() const { return (size_t)n * (size_t)elem; }
  const ClassPlan& plan(int32_t k) const {
    static const ClassPlan empty;
    auto it = plans.find(k);
    return it == plans.end() ? empty : it->second;
  }
};

#endif
