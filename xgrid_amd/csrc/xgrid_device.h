/* Device templates for generated stencil statements (gfx950, compiled by hipRTC).
 *
 * A generated statement is a struct S with
 *   typedef <elem> T;  static constexpr int D;  struct Refs; struct Params;
 *   static T value(const Refs&, const Params&, int i0, int i1, int i2);
 * and the templates below decide which cells evaluate it:
 *   run_main      - the cells of one boundary class (box / list / masked box),
 *                   stored in place (reference nest, xgrid/lang/generator.py:289-304,353-361)
 *                   or into a temporary (phase 1 of a two-phase statement);
 *   run_pingpong  - Jacobi: every cell of the grid; class cells computed from
 *                   the old level, other cells copied (replaces the two-phase nest
 *                   with its per-sweep malloc and copy-back, generator.py:312-352);
 *   run_phase2    - phase 2 of a two-phase statement: temporary -> store cells;
 *   run_block     - one workgroup walks a small class (fused boundary chains).
 * Linear indices are row-major; for 1-D and square 2-D grids they coincide with
 * the reference's accessor (generator.py:158-182).
 */
#ifndef XGRID_DEVICE_H
#define XGRID_DEVICE_H

#include "xgrid_abi.h"

#define XG_OV_NONE 0
#define XG_OV_LIMIT 1
#define XG_OV_WRAP 2

/* tile of one workgroup (256 work-items laid out 64 x 4) */
#define XG_TX 64
#define XG_THREADS 256

namespace xg {

template <int D> struct Tile;
template <> struct Tile<1> { static constexpr int X = 1024, Y = 1, Z = 1; };
template <> struct Tile<2> { static constexpr int X = 64, Y = 16, Z = 1; };
template <> struct Tile<3> { static constexpr int X = 64, Y = 4, Z = 4; };

/* The hardware hands workgroup b to XCD (b mod 8).  Give each XCD one
 * contiguous run of tiles so vertical neighbours stay in that XCD's L2. */
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t X = 8u;
  const uint32_t q = nb / X, r = nb % X;
  const uint32_t x = b % X, k = b / X;
  return x * q + (x < r ? x : r) + k;
}

template <int D>
__device__ __forceinline__ int64_t cells(const xg_gview& g) {
  int64_t n = g.shape[0];
  if (D > 1) n *= g.shape[1];
  if (D > 2) n *= g.shape[2];
  return n;
}

template <int D>
__device__ __forceinline__ int64_t lin(const int32_t* shape, int i0, int i1, int i2) {
  if (D == 1) return i0;
  if (D == 2) return (int64_t)i0 * shape[1] + i1;
  return ((int64_t)i0 * shape[1] + i1) * shape[2] + i2;
}

template <int D, int OV>
__device__ __forceinline__ int64_t index(const xg_gview& g, int i0, int i1, int i2) {
  if (OV == XG_OV_WRAP) {
    i0 = (i0 + g.shape[0]) % g.shape[0];
    if (D > 1) i1 = (i1 + g.shape[1]) % g.shape[1];
    if (D > 2) i2 = (i2 + g.shape[2]) % g.shape[2];
  } else if (OV == XG_OV_LIMIT) {
    i0 = i0 < 0 ? 0 : (i0 >= g.shape[0] ? g.shape[0] - 1 : i0);
    if (D > 1) i1 = i1 < 0 ? 0 : (i1 >= g.shape[1] ? g.shape[1] - 1 : i1);
    if (D > 2) i2 = i2 < 0 ? 0 : (i2 >= g.shape[2] ? g.shape[2] - 1 : i2);
  }
  return lin<D>(g.shape, i0, i1, i2);
}

/* Load one cell.  Addresses outside the array (undefined in the reference's
 * C) are clamped to the array so the device never reads out of bounds. */
template <class T, int D, int OV>
__device__ __forceinline__ T ld(const xg_gview& g, int i0, int i1, int i2) {
  int64_t k = index<D, OV>(g, i0, i1, i2);
  const int64_t n = cells<D>(g);
  k = k < 0 ? 0 : (k >= n ? n - 1 : k);
  return static_cast<const T*>(g.ptr)[k];
}

/* Store one cell; a store outside the array is dropped. */
template <class T, int D, int OV>
__device__ __forceinline__ void st(const xg_gview& g, int i0, int i1, int i2, const T& v) {
  const int64_t k = index<D, OV>(g, i0, i1, i2);
  if (k >= 0 && k < cells<D>(g)) static_cast<T*>(g.ptr)[k] = v;
}

template <class S, int OV>
__device__ __forceinline__ void apply(const xg_launch_hdr& H, const typename S::Refs& R,
                                      const typename S::Params& P, int i0, int i1, int i2) {
  typedef typename S::T T;
  const T v = S::value(R, P, i0, i1, i2);
  if (H.dom.flags & XG_F_TO_TMP) {
    static_cast<T*>(H.src.ptr)[lin<S::D>(H.out.shape, i0, i1, i2)] = v;
  } else {
    st<T, S::D, OV>(H.out, i0 + S::SO0, i1 + S::SO1, i2 + S::SO2, v);
  }
}

template <int D>
__device__ __forceinline__ void unlin(const int32_t* shape, int64_t c, int& i0, int& i1, int& i2) {
  if (D == 1) { i0 = (int)c; i1 = 0; i2 = 0; }
  else if (D == 2) { i0 = (int)(c / shape[1]); i1 = (int)(c - (int64_t)i0 * shape[1]); i2 = 0; }
  else {
    const int64_t plane = (int64_t)shape[1] * shape[2];
    i0 = (int)(c / plane);
    const int64_t r = c - (int64_t)i0 * plane;
    i1 = (int)(r / shape[2]);
    i2 = (int)(r - (int64_t)i1 * shape[2]);
  }
}

/* Visit every cell of the box [lo, hi) handled by this workgroup's tile. */
template <int D, class F>
__device__ __forceinline__ void for_tile(const int32_t* lo, const int32_t* hi, F&& f) {
  typedef Tile<D> TL;
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  if (D == 1) {
    const int base = lo[0] + (int)tile * TL::X + tid;
#pragma unroll
    for (int r = 0; r < TL::X / XG_THREADS; ++r) {
      const int i0 = base + r * XG_THREADS;
      if (i0 < hi[0]) f(i0, 0, 0);
    }
  } else if (D == 2) {
    const int nx = (hi[1] - lo[1] + TL::X - 1) / TL::X;
    const int t0 = (int)tile / nx, t1 = (int)tile - t0 * nx;
    const int i1 = lo[1] + t1 * TL::X + (tid & 63);
    const int r0 = lo[0] + t0 * TL::Y + (tid >> 6) * (TL::Y / 4);
    if (i1 < hi[1]) {
#pragma unroll
      for (int r = 0; r < TL::Y / 4; ++r) {
        const int i0 = r0 + r;
        if (i0 < hi[0]) f(i0, i1, 0);
      }
    }
  } else {
    const int nx = (hi[2] - lo[2] + TL::X - 1) / TL::X;
    const int ny = (hi[1] - lo[1] + TL::Y - 1) / TL::Y;
    const int t0 = (int)tile / (nx * ny);
    const int rem = (int)tile - t0 * nx * ny;
    const int t1 = rem / nx, t2 = rem - t1 * nx;
    const int i2 = lo[2] + t2 * TL::X + (tid & 63);
    const int i1 = lo[1] + t1 * TL::Y + (tid >> 6);
    if (i2 < hi[2] && i1 < hi[1]) {
#pragma unroll
      for (int r = 0; r < TL::Z; ++r) {
        const int i0 = lo[0] + t0 * TL::Z + r;
        if (i0 < hi[0]) f(i0, i1, i2);
      }
    }
  }
}

template <class S, int OV>
__device__ void run_main(const xg_launch_hdr& H, const typename S::Refs& R, const typename S::Params& P) {
  constexpr int D = S::D;
  const xg_dom& dom = H.dom;
  if (dom.kind == XG_DOM_LIST) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < dom.count; k += stride) {
      int i0, i1, i2;
      unlin<D>(H.out.shape, dom.list[k], i0, i1, i2);
      apply<S, OV>(H, R, P, i0, i1, i2);
    }
  } else if (dom.kind == XG_DOM_MASKED) {
    const int32_t* mask = dom.mask;
    const int klass = dom.klass;
    const int32_t* shape = H.out.shape;
    for_tile<D>(dom.lo, dom.hi, [&](int i0, int i1, int i2) {
      if (mask[lin<D>(shape, i0, i1, i2)] == klass) apply<S, OV>(H, R, P, i0, i1, i2);
    });
  } else {
    for_tile<D>(dom.lo, dom.hi, [&](int i0, int i1, int i2) { apply<S, OV>(H, R, P, i0, i1, i2); });
  }
}

/* Jacobi sweep: out = scratch, src = old level 0.  Every cell of the grid is
 * written exactly once, so after the sweep the two buffers swap roles. */
template <class S, int OV>
__device__ void run_pingpong(const xg_launch_hdr& H, const typename S::Refs& R, const typename S::Params& P) {
  constexpr int D = S::D;
  typedef typename S::T T;
  const xg_dom& dom = H.dom;
  const int32_t zero[3] = {0, 0, 0};
  const int32_t* shape = H.out.shape;
  const T* __restrict__ src = static_cast<const T*>(H.src.ptr);
  T* __restrict__ out = static_cast<T*>(H.out.ptr);
  for_tile<D>(zero, shape, [&](int i0, int i1, int i2) {
    const int64_t c = lin<D>(shape, i0, i1, i2);
    bool member;
    if (dom.kind == XG_DOM_BOX) {
      member = i0 >= dom.lo[0] && i0 < dom.hi[0];
      if (D > 1) member = member && i1 >= dom.lo[1] && i1 < dom.hi[1];
      if (D > 2) member = member && i2 >= dom.lo[2] && i2 < dom.hi[2];
    } else {
      member = dom.mask[c] == dom.klass;
    }
    out[c] = member ? S::value(R, P, i0, i1, i2) : src[c];
  });
}

/* Phase 2 of a two-phase statement: store the temporary at the shifted cell. */
template <class T, int D, int OV>
__device__ void run_phase2(const xg_launch_hdr& H) {
  const xg_dom& dom = H.dom;
  const T* tmp = static_cast<const T*>(H.src.ptr);
  const int32_t* shape = H.out.shape;
  auto f = [&](int i0, int i1, int i2) {
    st<T, D, OV>(H.out, i0 + dom.so[0], i1 + dom.so[1], i2 + dom.so[2], tmp[lin<D>(shape, i0, i1, i2)]);
  };
  if (dom.kind == XG_DOM_LIST) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < dom.count; k += stride) {
      int i0, i1, i2;
      unlin<D>(shape, dom.list[k], i0, i1, i2);
      f(i0, i1, i2);
    }
  } else if (dom.kind == XG_DOM_MASKED) {
    for_tile<D>(dom.lo, dom.hi, [&](int i0, int i1, int i2) {
      if (dom.mask[lin<D>(shape, i0, i1, i2)] == dom.klass) f(i0, i1, i2);
    });
  } else {
    for_tile<D>(dom.lo, dom.hi, f);
  }
}

/* One workgroup walks the whole (small) class; used to chain several boundary
 * statements in one launch with a workgroup barrier between them. */
template <class S, int OV>
__device__ void run_block(const xg_launch_hdr& H, const typename S::Refs& R, const typename S::Params& P) {
  constexpr int D = S::D;
  const xg_dom& dom = H.dom;
  if (dom.kind == XG_DOM_EMPTY) return;
  if (dom.kind == XG_DOM_LIST) {
    for (int64_t k = threadIdx.x; k < dom.count; k += blockDim.x) {
      int i0, i1, i2;
      unlin<D>(H.out.shape, dom.list[k], i0, i1, i2);
      apply<S, OV>(H, R, P, i0, i1, i2);
    }
    return;
  }
  int32_t ext[3] = {dom.hi[0] - dom.lo[0], D > 1 ? dom.hi[1] - dom.lo[1] : 1, D > 2 ? dom.hi[2] - dom.lo[2] : 1};
  const int64_t vol = (int64_t)ext[0] * ext[1] * ext[2];
  for (int64_t k = threadIdx.x; k < vol; k += blockDim.x) {
    int j0, j1, j2;
    unlin<3>(ext, k, j0, j1, j2);
    const int i0 = dom.lo[0] + j0, i1 = (D > 1 ? dom.lo[1] : 0) + j1, i2 = (D > 2 ? dom.lo[2] : 0) + j2;
    if (dom.kind == XG_DOM_MASKED && dom.mask[lin<D>(H.out.shape, i0, i1, i2)] != dom.klass) continue;
    apply<S, OV>(H, R, P, i0, i1, i2);
  }
}

template <class T>
__device__ __forceinline__ T sq(T x) { return x * x; }

}  // namespace xg

#endif
