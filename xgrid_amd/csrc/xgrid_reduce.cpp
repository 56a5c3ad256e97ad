// FP64 reductions over the cells of one boundary class of device grids:
// difference norms (convergence tests, error against a reference field) and
// the residual of a Poisson equation.  Fields may be stored in FP32 or FP64;
// every value is widened to FP64 before it is combined.
//
// Two passes with a fixed grid, so the result is deterministic: each workgroup
// reduces a fixed slice of cells (wave64 shuffles, then LDS across waves) into
// one partial, and one workgroup reduces the partials in index order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "xgrid_internal.h"

using namespace xgi;

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

struct Partial {
  double sum;   // sum of squares
  double max;   // largest magnitude (NaN if any value was NaN)
};

__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  return a > b ? a : b;
}

template <class T>
__device__ __forceinline__ double load(const void* p, int64_t c) {
  return (double)static_cast<const T*>(p)[c];
}

// Reduce (s, m) over the workgroup; thread 0 returns the result.
__device__ __forceinline__ Partial block_reduce(double s, double m) {
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o, 64);
    m = nanmax(m, __shfl_xor(m, o, 64));
  }
  __shared__ double ws[kRedThreads / 64], wm[kRedThreads / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ws[w] = s;
    wm[w] = m;
  }
  __syncthreads();
  Partial r{0.0, 0.0};
  if (threadIdx.x == 0) {
    for (int i = 0; i < kRedThreads / 64; ++i) {
      r.sum += ws[i];
      r.max = nanmax(r.max, wm[i]);
    }
  }
  return r;
}

// The k-th cell of the class (box, masked box or list); false if a masked-box
// cell belongs to another class.
__device__ __forceinline__ bool class_cell(const xg_dom& dom, const int32_t* shape, int64_t k, int* i) {
  int64_t c;
  if (dom.kind == XG_DOM_LIST) {
    c = dom.list[k];
    i[0] = (int)(c / ((int64_t)shape[1] * shape[2]));
    const int64_t r = c - (int64_t)i[0] * shape[1] * shape[2];
    i[1] = (int)(r / shape[2]);
    i[2] = (int)(r - (int64_t)i[1] * shape[2]);
    return true;
  }
  const int e1 = dom.hi[1] - dom.lo[1], e2 = dom.hi[2] - dom.lo[2];
  const int64_t j0 = k / ((int64_t)e1 * e2);
  const int64_t r = k - j0 * e1 * e2;
  i[0] = dom.lo[0] + (int)j0;
  i[1] = dom.lo[1] + (int)(r / e2);
  i[2] = dom.lo[2] + (int)(r - (r / e2) * e2);
  if (dom.kind == XG_DOM_MASKED) {
    c = ((int64_t)i[0] * shape[1] + i[1]) * shape[2] + i[2];
    return dom.mask[c] == dom.klass;
  }
  return true;
}

// a, b: element types (float / double); b may be null.
template <class TA, class TB>
__global__ void __launch_bounds__(kRedThreads) diff_partials(const void* a, const void* b, xg_dom dom, int32_t s0,
                                                             int32_t s1, int32_t s2, int64_t total, Partial* part) {
  const int32_t shape[3] = {s0, s1, s2};
  double s = 0.0, m = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; k < total; k += (int64_t)gridDim.x * kRedThreads) {
    int i[3];
    if (!class_cell(dom, shape, k, i)) continue;
    const int64_t c = ((int64_t)i[0] * s1 + i[1]) * s2 + i[2];
    double d = load<TA>(a, c);
    if (b) d -= load<TB>(b, c);
    s += d * d;
    m = nanmax(m, d < 0 ? -d : d);
  }
  const Partial r = block_reduce(s, m);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// r = sum_a coef[a] * (p[+e_a] - 2 p + p[-e_a]) - scale * rhs over class cells
// (every class cell is at least one cell inside the grid: checked on the host).
template <class TP, class TR>
__global__ void __launch_bounds__(kRedThreads) poisson_partials(const void* p, const void* rhs, xg_dom dom, int32_t dims,
                                                                int32_t s0, int32_t s1, int32_t s2, double c0, double c1,
                                                                double c2, double scale, int64_t total, Partial* part) {
  const int32_t shape[3] = {s0, s1, s2};
  const int64_t st[3] = {(int64_t)s1 * s2, s2, 1};
  const double cf[3] = {c0, c1, c2};
  double s = 0.0, m = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; k < total; k += (int64_t)gridDim.x * kRedThreads) {
    int i[3];
    if (!class_cell(dom, shape, k, i)) continue;
    const int64_t c = ((int64_t)i[0] * s1 + i[1]) * s2 + i[2];
    const double pc = load<TP>(p, c);
    double lap = 0.0;
    for (int a = 0; a < dims; ++a) {
      const int ax = 3 - dims + a;   // grid axis a is stored as axis ax of the padded 3-D shape
      lap += cf[a] * ((load<TP>(p, c + st[ax]) - 2.0 * pc) + load<TP>(p, c - st[ax]));
    }
    const double r = lap - scale * load<TR>(rhs, c);
    s += r * r;
    m = nanmax(m, r < 0 ? -r : r);
  }
  const Partial r = block_reduce(s, m);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void __launch_bounds__(kRedThreads) final_reduce(const Partial* part, int32_t n, Partial* out) {
  double s = 0.0, m = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) {
    s += part[i].sum;
    m = nanmax(m, part[i].max);
  }
  const Partial r = block_reduce(s, m);
  if (threadIdx.x == 0) *out = r;
}

Partial* g_part = nullptr;   // kRedBlocks partials + 1 result, kept for the process

// Scratch of the allocation-free entry point (xg_grid_norm_diff_noalloc): the
// same kRedBlocks partials + 1 result, 16,400 bytes, in device memory that
// every GPU the library runs on holds until the process ends.  In host code
// `scratch` names a host-side stand-in, not a device address: set-up looks up
// each device's copy once (reduce_setup, from xg_init) and only that address
// is passed to kernels.
static __device__ __align__(256) Partial scratch[kRedBlocks + 1];
constexpr int kMaxDevices = 64;
Partial* g_scratch[kMaxDevices] = {};

int ensure_partials() {
  if (g_part) return XG_OK;
  void* p = nullptr;
  XG_HIP(hipMalloc(&p, sizeof(Partial) * (kRedBlocks + 1)));
  g_part = (Partial*)p;
  return XG_OK;
}

// Cells of class `klass` of G as a 3-D padded domain (axes shifted so the last
// grid axis is the contiguous one).
int class_domain(const xg_grid* G, int32_t klass, xg_dom& dom, int32_t* shape3, int64_t& total) {
  const ClassPlan& p = G->plan(klass);
  std::memset(&dom, 0, sizeof dom);
  dom.kind = p.kind;
  dom.klass = klass;
  dom.count = p.count;
  dom.list = p.d_list;
  dom.mask = G->d_mask;
  const int D = G->dims;
  for (int a = 0; a < 3; ++a) {
    const int g = a - (3 - D);      // grid axis of padded axis a
    shape3[a] = g >= 0 ? G->shape[g] : 1;
    dom.lo[a] = g >= 0 ? p.lo[g] : 0;
    dom.hi[a] = g >= 0 ? p.hi[g] : 1;
  }
  if (p.kind == XG_DOM_EMPTY || p.count == 0) {
    total = 0;
    return XG_OK;
  }
  total = p.kind == XG_DOM_LIST ? p.count
                                : (int64_t)(dom.hi[0] - dom.lo[0]) * (dom.hi[1] - dom.lo[1]) * (dom.hi[2] - dom.lo[2]);
  return XG_OK;
}

int finish(int64_t total, double* l2, double* linf, Partial* part) {
  Partial h{0.0, 0.0};
  if (total > 0) {
    const int nb = (int)std::min<int64_t>(kRedBlocks, ceil_div(total, kRedThreads));
    hipLaunchKernelGGL(final_reduce, dim3(1), dim3(kRedThreads), 0, g_dev.stream, part, nb, part + kRedBlocks);
    XG_HIP(hipGetLastError());
    XG_HIP(hipMemcpyAsync(&h, part + kRedBlocks, sizeof h, hipMemcpyDeviceToHost, g_dev.stream));
    XG_HIP(hipStreamSynchronize(g_dev.stream));
  }
  if (l2) *l2 = std::sqrt(h.sum);
  if (linf) *linf = h.max;
  return XG_OK;
}

bool float_type(int32_t t) { return t == XG_T_F32 || t == XG_T_F64; }

// L2 / Linf of a - b over class klass of a, reducing through `part` (room for
// kRedBlocks partials + 1 result).  Every partial the final pass reads is
// written by the same call first, so nothing in `part` needs clearing.
int norm_diff(xg_grid* a, int32_t level_a, int32_t type_a, xg_grid* b, int32_t level_b, int32_t type_b,
              int32_t klass, double* l2, double* linf, int64_t* cells, Partial* part) {
  if (!a || level_a < 0 || level_a >= (int32_t)a->levels.size() || !float_type(type_a))
    return fail(XG_ERR_ARG, "invalid first grid, level or element type");
  if (b && (level_b < 0 || level_b >= (int32_t)b->levels.size() || !float_type(type_b)))
    return fail(XG_ERR_ARG, "invalid second grid, level or element type");
  if (b && (b->dims != a->dims || std::memcmp(b->shape, a->shape, sizeof a->shape) != 0))
    return fail(XG_ERR_ARG, "grids of different shapes");
  if ((int64_t)a->elem != (type_a == XG_T_F32 ? 4 : 8) || (b && (int64_t)b->elem != (type_b == XG_T_F32 ? 4 : 8)))
    return fail(XG_ERR_ARG, "element type does not match the grid's element size");
  xg_dom dom;
  int32_t sh[3];
  int64_t total = 0;
  class_domain(a, klass, dom, sh, total);
  if (cells) *cells = a->plan(klass).count;
  if (total > 0) {
    const int nb = (int)std::min<int64_t>(kRedBlocks, ceil_div(total, kRedThreads));
    const void* pa = a->levels[(size_t)level_a];
    const void* pb = b ? b->levels[(size_t)level_b] : nullptr;
    const bool fa = type_a == XG_T_F32, fb = !b || type_b == XG_T_F32;
    auto* k = fa ? (fb ? diff_partials<float, float> : diff_partials<float, double>)
                 : (fb ? diff_partials<double, float> : diff_partials<double, double>);
    hipLaunchKernelGGL(k, dim3(nb), dim3(kRedThreads), 0, g_dev.stream, pa, pb, dom, sh[0], sh[1], sh[2], total, part);
    XG_HIP(hipGetLastError());
  }
  return finish(total, l2, linf, part);
}

}  // namespace

namespace xgi {

int reduce_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(XG_ERR_ARG, "device index beyond the scratch table");
  if (g_scratch[device]) return XG_OK;
  void* p = nullptr;
  XG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = (Partial*)p;
  return XG_OK;
}

}  // namespace xgi

extern "C" {

int xg_grid_norm_diff(xg_grid* a, int32_t level_a, int32_t type_a, xg_grid* b, int32_t level_b, int32_t type_b,
                      int32_t klass, double* l2, double* linf, int64_t* cells) {
  if (int rc = check_ready()) return rc;
  if (int rc = ensure_partials()) return rc;
  return norm_diff(a, level_a, type_a, b, level_b, type_b, klass, l2, linf, cells, g_part);
}

int xg_grid_norm_diff_noalloc(xg_grid* a, int32_t level_a, int32_t type_a, xg_grid* b, int32_t level_b,
                              int32_t type_b, int32_t klass, double* l2, double* linf, int64_t* cells) {
  if (int rc = check_ready()) return rc;
  // The partials fit whatever the class size: at most kRedBlocks workgroups, each
  // reducing a strided slice of the cells.
  static_assert(sizeof scratch / sizeof(Partial) == kRedBlocks + 1, "scratch holds every partial and the result");
  Partial* part = g_dev.id >= 0 && g_dev.id < kMaxDevices ? g_scratch[g_dev.id] : nullptr;
  if (!part) return fail(XG_ERR_STATE, "the device's scratch address was not looked up (xg_init)");
  return norm_diff(a, level_a, type_a, b, level_b, type_b, klass, l2, linf, cells, part);
}

int xg_grid_poisson_residual(xg_grid* p, int32_t level_p, int32_t type_p, xg_grid* rhs, int32_t level_rhs,
                             int32_t type_rhs, int32_t klass, const double* coef, double rhs_scale, double* l2,
                             double* linf) {
  if (int rc = check_ready()) return rc;
  if (!p || !rhs || !coef || level_p < 0 || level_p >= (int32_t)p->levels.size() || level_rhs < 0 ||
      level_rhs >= (int32_t)rhs->levels.size() || !float_type(type_p) || !float_type(type_rhs))
    return fail(XG_ERR_ARG, "invalid grid, level, element type or coefficients");
  if (rhs->dims != p->dims || std::memcmp(rhs->shape, p->shape, sizeof p->shape) != 0)
    return fail(XG_ERR_ARG, "grids of different shapes");
  if ((int64_t)p->elem != (type_p == XG_T_F32 ? 4 : 8) || (int64_t)rhs->elem != (type_rhs == XG_T_F32 ? 4 : 8))
    return fail(XG_ERR_ARG, "element type does not match the grid's element size");
  const ClassPlan& cp = p->plan(klass);
  if (cp.kind != XG_DOM_EMPTY && cp.count > 0)
    for (int a = 0; a < p->dims; ++a)
      if (cp.lo[a] < 1 || cp.hi[a] > p->shape[a] - 1)
        return fail(XG_ERR_ARG, "the class reaches the edge of the grid: the residual's stencil would leave it");
  if (int rc = ensure_partials()) return rc;
  xg_dom dom;
  int32_t sh[3];
  int64_t total = 0;
  class_domain(p, klass, dom, sh, total);
  if (total > 0) {
    const int nb = (int)std::min<int64_t>(kRedBlocks, ceil_div(total, kRedThreads));
    double c[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < p->dims; ++a) c[a] = coef[a];
    const bool fp = type_p == XG_T_F32, fr = type_rhs == XG_T_F32;
    auto* k = fp ? (fr ? poisson_partials<float, float> : poisson_partials<float, double>)
                 : (fr ? poisson_partials<double, float> : poisson_partials<double, double>);
    hipLaunchKernelGGL(k, dim3(nb), dim3(kRedThreads), 0, g_dev.stream, p->levels[(size_t)level_p],
                       rhs->levels[(size_t)level_rhs], dom, p->dims, sh[0], sh[1], sh[2], c[0], c[1], c[2], rhs_scale,
                       total, g_part);
    XG_HIP(hipGetLastError());
  }
  return finish(total, l2, linf, g_part);
}

}  // extern "C"
