/* Device templates for generated stencil statements (gfx950, compiled by hipRTC).
 *
 * A generated statement is a struct S with
 *   typedef <elem> T;  static constexpr int D, SO0..2 (store offset),
 *   RL0..2 / RH0..2 (largest backward / forward load offset per dimension);
 *   struct Refs; struct Params;
 *   template <bool SAFE> static T value(const Refs&, const Params&, int i0, int i1, int i2);
 * and the templates below decide which cells evaluate it:
 *   run_main      - the cells of one boundary class (box / list / masked box),
 *                   stored in place (reference nest, xgrid/lang/generator.py:289-304,353-361)
 *                   or into a temporary (phase 1 of a two-phase statement);
 *   run_pingpong  - Jacobi: every cell of the grid; class cells computed from
 *                   the old level, other cells copied (replaces the two-phase nest
 *                   with its per-sweep malloc and copy-back, generator.py:312-352);
 *   run_phase2    - phase 2 of a two-phase statement: temporary -> store cells;
 *   run_block     - one workgroup walks a small class (chained boundary statements).
 * Linear indices are row-major; for 1-D and square 2-D grids they coincide with
 * the reference's accessor (generator.py:158-182).
 *
 * Each work-item owns NC cells of a workgroup tile.  It evaluates all of them
 * before storing any (the statement never reads a cell it writes: the runtime
 * sends statements that could to the two-phase path), so its loads are all in
 * flight together.  Tiles whose cells and stencil reach lie inside the grid
 * (the common case) read without the per-load clamp/wrap.
 */
#ifndef XGRID_DEVICE_H
#define XGRID_DEVICE_H

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>  /* offline builds (hipcc); hipRTC predefines the same */
#endif
#include "xgrid_abi.h"

#define XG_OV_NONE 0
#define XG_OV_LIMIT 1
#define XG_OV_WRAP 2

#define XG_THREADS 256

namespace xg {

/* Workgroup tile: 256 work-items, NC cells each per chunk, CH chunks in turn.
 *   1-D: 1024 consecutive cells (work-item t: t, t+256, t+512, t+768)
 *   2-D: 4 rows x 64 columns (one cell per work-item)
 *   3-D: XG_TILE3_CHUNKS x 4 planes x 4 rows x 64 columns (work-item (r, c):
 *        planes 0..3 of each chunk, chunk after chunk); column tiles are
 *        aligned to multiples of 64 columns, so the first may start below lo */
template <int D> struct Tile;
template <> struct Tile<1> { static constexpr int E0 = 1024, E1 = 1, E2 = 1, NC = 4, CH = 1; };
template <> struct Tile<2> { static constexpr int E0 = 4, E1 = 64, E2 = 1, NC = 1, CH = 1; };
template <> struct Tile<3> { static constexpr int E0 = 4, E1 = 4, E2 = 64, NC = 4, CH = XG_TILE3_CHUNKS; };

/* Every pointer in a launch header, view or batch table is device memory
 * (global address space).  A pointer read out of memory (a batch
 * table entry) is generic to the compiler, which then emits flat accesses:
 * those wait on the LDS counter too and serialize against LDS traffic.  This
 * cast states the address space, so accesses are global loads and stores. */
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> gp(const void* p) {
  return (gptr<T>)(p);
}
/* K adjacent elements read or written as one access (staged 3-D tiles: K
 * elements of 16 bytes together, only at addresses that are multiples of 16). */
template <class T, int K>
using vecn = T __attribute__((ext_vector_type(K)));
/* The views and parameter blocks a batch table points to are written by the
 * host before the launch and only read by kernels: reading them through the
 * constant address space lets uniform reads be scalar loads into SGPRs, as
 * kernel arguments are. */
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* cp(const void* p) {
  return (const __attribute__((address_space(4))) T*)(p);
}

/* The hardware hands workgroup b to XCD (b mod 8).  Give each XCD one
 * contiguous run of tiles so vertical neighbours stay in that XCD's L2. */
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t X = 8u;
  const uint32_t q = nb / X, r = nb % X;
  const uint32_t x = b % X, k = b / X;
  return x * q + (x < r ? x : r) + k;
}

template <int D>
__device__ __forceinline__ int64_t cells(const int32_t* shape) {
  int64_t n = shape[0];
  if (D > 1) n *= shape[1];
  if (D > 2) n *= shape[2];
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

/* Plane p clamped to the array's planes (staged 3-D loads of padding planes). */
__device__ __forceinline__ int64_t st3_plane(int p, const int32_t* sh) {
  return (int64_t)min(max(p, 0), sh[0] - 1);
}

/* Linear offset of a stencil offset under the strides of `shape`. */
template <int D>
__device__ __forceinline__ int64_t off(const int32_t* shape, int o0, int o1, int o2) {
  if (D == 1) return o0;
  if (D == 2) return (int64_t)o0 * shape[1] + o1;
  return ((int64_t)o0 * shape[1] + o1) * shape[2] + o2;
}

/* Load cell (i0 + o0, i1 + o1, i2 + o2).
 * SAFE: the caller proved the cell lies inside the grid and that every view
 * has the store grid's shape `sh`; `xc` is the linear index of (i0, i1, i2), so
 * the address is xc plus a constant offset times the (scalar) strides.
 * Otherwise the overstep rule applies, and addresses still outside the array
 * (undefined in the reference's C) are clamped so the device never reads out
 * of bounds. */
template <class T, int D, int OV, bool SAFE>
__device__ __forceinline__ T ld(const xg_gview& g, const int32_t* sh, int64_t xc, int o0, int o1, int o2,
                                int i0, int i1, int i2) {
  if (SAFE) return gp<const T>(g.ptr)[xc + off<D>(sh, o0, o1, o2)];
  i0 += o0;
  i1 += o1;
  i2 += o2;
  int64_t k = index<D, OV>(g, i0, i1, i2);
  const int64_t n = cells<D>(g.shape);
  k = k < 0 ? 0 : (k >= n ? n - 1 : k);
  return gp<const T>(g.ptr)[k];
}

/* Write v at dst; NT: with the non-temporal hint (the line is not kept for
 * reuse).  The choice is a template argument: the compiler merges a run-time
 * branch between the two forms into one plain store. */
template <class V, bool NT>
__device__ __forceinline__ void put(gptr<V> dst, const V& v) {
  if constexpr (NT) __builtin_nontemporal_store(v, dst);
  else *dst = v;
}

/* One 16-byte global load per lane written straight into LDS, with no
 * register in between (gfx950's dwordx4 LDS load): lane l's bytes land at
 * dst + 16 l.  dst must be the same for every lane of the wave.  The data is
 * in LDS once the wave's vector-memory counter reaches 0. */
__device__ __forceinline__ void load_lds16(gptr<const void> src, void* dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}

/* Wait for this wave's vector-memory loads, LDS loads above included. */
__device__ __forceinline__ void wait_vm() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

/* Store one cell; a store outside the array is dropped. */
template <class T, int D, int OV, bool NT = false>
__device__ __forceinline__ void st(const xg_gview& g, int i0, int i1, int i2, const T& v) {
  const int64_t k = index<D, OV>(g, i0, i1, i2);
  if (k >= 0 && k < cells<D>(g.shape)) put<T, NT>(gp<T>(g.ptr) + k, v);
}

template <class S, int OV, bool NT = false>
__device__ __forceinline__ void store(const xg_launch_hdr& H, int i0, int i1, int i2, const typename S::T& v) {
  typedef typename S::T T;
  if (H.dom.flags & XG_F_TO_TMP) {
    gp<T>(H.src.ptr)[lin<S::D>(H.out.shape, i0, i1, i2)] = v;
  } else {
    st<T, S::D, OV, NT>(H.out, i0 + S::SO0, i1 + S::SO1, i2 + S::SO2, v);
  }
}

template <class S, int OV>
__device__ __forceinline__ void apply(const xg_launch_hdr& H, const typename S::Refs& R,
                                      const typename S::Params& P, int i0, int i1, int i2) {
  store<S, OV>(H, i0, i1, i2, S::template value<false>(R, P, H.out.shape, i0, i1, i2));
}

/* Coordinates of list entry c.  Lists exist only for grids of at most
 * INT32_MAX cells (the planner uses a masked box otherwise), so the entry and
 * the plane size fit in 32 bits and the divisions are 32-bit ones (a 64-bit
 * division is a long instruction sequence per cell). */
template <int D>
__device__ __forceinline__ void unlin_list(const int32_t* shape, int32_t c, int& i0, int& i1, int& i2) {
  const uint32_t u = (uint32_t)c;
  if (D == 1) { i0 = (int)u; i1 = 0; i2 = 0; }
  else if (D == 2) {
    const uint32_t s1 = (uint32_t)shape[1], q = u / s1;
    i0 = (int)q; i1 = (int)(u - q * s1); i2 = 0;
  } else {
    const uint32_t s2 = (uint32_t)shape[2], plane = (uint32_t)shape[1] * s2;
    const uint32_t q0 = u / plane, r = u - q0 * plane, q1 = r / s2;
    i0 = (int)q0; i1 = (int)q1; i2 = (int)(r - q1 * s2);
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

/* This workgroup's tile of the box [lo, hi): the NC cells of this work-item,
 * whether each is inside the box, and the tile's extent clipped to the box.
 * CPL (3-D only): columns per work-item.  A tile is 64 * CPL columns wide and
 * work-item (r, c) holds columns CPL * c .. CPL * c + CPL - 1 of each of its
 * planes: cell n is plane n / CPL, column CPL * c + n % CPL (statements whose
 * staged loads read 16 bytes of a row per lane: CPL = 2 for 8-byte elements,
 * 4 for 4-byte ones). */
template <int D, int CPL = 1>
struct TileCells {
  static constexpr int NC = Tile<D>::NC * (D == 3 ? CPL : 1);
  static constexpr int E2 = Tile<D>::E2 * (D == 3 ? CPL : 1);
  int c[NC][3];
  bool in[NC];
  int32_t tlo[3], thi[3];
  int32_t col0;  /* 3-D: first column of the aligned tile (work-item c holds column col0 + c % 64) */

  __device__ __forceinline__ TileCells(const int32_t* lo, const int32_t* hi, uint32_t tile, int chunk = 0) {
    typedef Tile<D> TL;
    const int tid = threadIdx.x;
    int t0 = (int)tile, t1 = 0, t2 = 0;
    /* 3-D column tiles start at multiples of their width E2 (64 columns per
     * column a lane holds), not at lo[2]:
     * with a row pitch that is a multiple of the 128-byte cache line, a wave's
     * row segment then covers whole lines instead of straddling one more, and
     * a line is fetched once instead of in parts by two tiles (512^3 FP32:
     * 312 against 352 us).  Columns of the first tile below lo[2] are not in
     * the tile. */
    const int a2 = D > 2 ? (lo[2] & ~(E2 - 1)) : 0;
    if (D == 2) {
      const int n1 = (hi[1] - lo[1] + TL::E1 - 1) / TL::E1;
      t0 = (int)tile / n1;
      t1 = (int)tile - t0 * n1;
    } else if (D == 3) {
      /* Row bands in groups of G; inside a group plane chunks slowest, then
       * bands, then column tiles.  A chunk's halo planes were read by the
       * group's previous chunk G x 64 x 4 x (row bytes) earlier (512 KB at
       * 512^3 FP32), inside the 4 MB L2 of the XCD the group runs on, and a
       * band's halo rows by the band before it, n2 tiles earlier.  With row
       * bands slowest, the reuse distance of a halo row was a whole band
       * (4 MB there): 282 against 312 us with aligned columns.  128-column
       * tiles (FP64 staged) group 32 bands: 512^3 FP64 417 against 421-424 us
       * with 16 and 431 with 8; 256-column FP32 tiles were 1 % slower with 32. */
      constexpr int64_t G = E2 == 128 ? 32 : 16;
      const int64_t n2 = (hi[2] - a2 + E2 - 1) / E2;
      const int64_t n1 = (hi[1] - lo[1] + TL::E1 - 1) / TL::E1;
      const int64_t n0 = (hi[0] - lo[0] + TL::E0 * TL::CH - 1) / (TL::E0 * TL::CH);
      const int64_t per = G * n0 * n2;
      const int64_t g = (int64_t)tile / per;
      const int64_t gsz = min(G, n1 - g * G);
      const int64_t rem = (int64_t)tile - g * per;
      t0 = (int)(rem / (gsz * n2));
      const int64_t r2 = rem - (int64_t)t0 * gsz * n2;
      const int64_t b = r2 / n2;
      t1 = (int)(g * G + b);
      t2 = (int)(r2 - b * n2);
    }
    tlo[0] = lo[0] + t0 * (TL::E0 * TL::CH) + chunk * TL::E0;
    tlo[1] = D > 1 ? lo[1] + t1 * TL::E1 : 0;
    const int c2 = a2 + t2 * E2;      /* first column of the (aligned) 3-D tile */
    col0 = c2;
    tlo[2] = D > 2 ? max(c2, lo[2]) : 0;
    thi[0] = min(tlo[0] + TL::E0, hi[0]);
    thi[1] = D > 1 ? min(tlo[1] + TL::E1, hi[1]) : 1;
    thi[2] = D > 2 ? min(c2 + E2, hi[2]) : 1;
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      if (D == 1) {
        c[n][0] = tlo[0] + tid + n * XG_THREADS;
        c[n][1] = 0;
        c[n][2] = 0;
      } else if (D == 2) {
        c[n][0] = tlo[0] + (tid >> 6) + 4 * n;  /* NC = 1: one row per wave */
        c[n][1] = tlo[1] + (tid & 63);
        c[n][2] = 0;
      } else {
        c[n][0] = tlo[0] + n / CPL;
        c[n][1] = tlo[1] + (tid >> 6);
        c[n][2] = c2 + (tid & 63) * CPL + n % CPL;
      }
      in[n] = c[n][0] < thi[0] && (D < 2 || c[n][1] < thi[1]) && (D < 3 || (c[n][2] >= tlo[2] && c[n][2] < thi[2]));
    }
  }

  /* Move cells outside the tile's clipped extent onto its nearest cell (in[]
   * keeps which are real). */
  __device__ __forceinline__ void clamp_to_tile() {
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      c[n][0] = min(c[n][0], thi[0] - 1);
      if (D > 1) c[n][1] = min(c[n][1], thi[1] - 1);
      if (D > 2) c[n][2] = max(min(c[n][2], thi[2] - 1), tlo[2]);
    }
  }
};

/* Do the tile's cells, the statement's loads and its store all lie inside the
 * grid, with every view of the grid shape (so no clamp or wrap is needed)? */
template <class S>
__device__ __forceinline__ bool tile_safe(const xg_launch_hdr& H, const int32_t* tlo, const int32_t* thi) {
  if (!(H.dom.flags & XG_F_UNIFORM)) return false;
  const int32_t* sh = H.out.shape;
  bool ok = tlo[0] - S::RL0 >= 0 && thi[0] + S::RH0 <= sh[0];
  if (S::D > 1) ok = ok && tlo[1] - S::RL1 >= 0 && thi[1] + S::RH1 <= sh[1];
  if (S::D > 2) ok = ok && tlo[2] - S::RL2 >= 0 && thi[2] + S::RH2 <= sh[2];
  return ok;
}

template <class S, int OV, bool MASKED, bool NT = false>
__device__ __forceinline__ void box_cells(const xg_launch_hdr& H, const typename S::Refs& R,
                                          const typename S::Params& P, uint32_t tile, unsigned char* st3_lds) {
  constexpr int D = S::D;
  typedef typename S::T T;
  typedef TileCells<D, S::CPL> TC;
  for (int ch = 0; ch < Tile<D>::CH; ++ch) {
    TC tc(H.dom.lo, H.dom.hi, tile, ch);
    if (tc.tlo[0] >= H.dom.hi[0]) break;
    constexpr int NC = TC::NC;
    /* A work-item evaluates all its cells without branching, so their loads are
     * in flight together: a cell past the tile's edge is evaluated at the
     * nearest cell inside the tile (a valid address) and not stored. */
    tc.clamp_to_tile();
    if (MASKED) {
#pragma unroll
      for (int n = 0; n < NC; ++n)
        tc.in[n] = tc.in[n] && gp<const int32_t>(H.dom.mask)[lin<D>(H.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2])] == H.dom.klass;
    }
    T v[NC];
    bool pairs = false;   /* store a lane's CPL columns with one 16-byte store */
    if (tile_safe<S>(H, tc.tlo, tc.thi)) {
      if constexpr (D == 3 && S::ST3) {
        /* Staged 3-D tile: each array's columns of the tile's planes are read
         * once into registers; column neighbours come over DPP lane shifts and
         * row neighbours from an LDS copy of the tile's rows.  tile_safe is the
         * same for every work-item of the workgroup, so all of them reach the
         * barrier inside st3_load. */
        typename S::St3 st;
        S::st3_load(st, st3_lds, R, H.out.shape, tc.tlo[0], tc.tlo[1], tc.col0);
#pragma unroll
        for (int n = 0; n < NC; ++n) {
          if (n % S::CPL == 0) S::st3_plane_n(st, n / S::CPL);   /* a plane's neighbours, once */
          v[n] = S::value_st3(st, R, P, H.out.shape, n, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
        }
        if (Tile<D>::CH > 1) __syncthreads();  /* the next chunk rewrites the LDS copy */
        /* In a safe tile every stored cell lies inside the grid, so with no store
         * offset and no scratch destination the store is a plain one at the cell's
         * index, and a lane's CPL adjacent columns (16 bytes) are one store where
         * the base is 16-byte aligned and the row pitch a multiple of CPL. */
        if constexpr (S::CPL * sizeof(T) == 16 && S::SO0 == 0 && S::SO1 == 0 && S::SO2 == 0)
          pairs = !(H.dom.flags & XG_F_TO_TMP) && ((uint64_t)H.out.ptr & 15) == 0 && (H.out.shape[2] % S::CPL) == 0;
      } else {
#pragma unroll
        for (int n = 0; n < NC; ++n) v[n] = S::template value<true>(R, P, H.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
      }
    } else {
#pragma unroll
      for (int n = 0; n < NC; ++n) v[n] = S::template value<false>(R, P, H.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
    }
    if constexpr (D == 3 && S::CPL * sizeof(T) == 16) {
      if (pairs) {
        /* A lane whose CPL cells are only partly stored (the class edge, a hole of a
         * masked class) merges its computed cells into the current values of the
         * others and stores all 16 bytes, so every lane of the wave takes the same
         * single store (with two columns per lane in a box class it stores its one
         * cell instead, below).  No other work-item of the launch writes those cells: they
         * are outside the statement's class, or outside the columns another tile
         * covers (tiles start at multiples of their width), and staged stores run
         * only unshifted and in place. */
        constexpr int K = S::CPL;
        const int col = tc.col0 + (int)(threadIdx.x & 63) * K;
        const bool row_fits = col + K <= H.out.shape[2];
        /* The partly stored lanes' current values are all read before the first
         * store: read one store at a time, the compiler cannot move a read above
         * the store before it (same array), and the wave waited for memory once
         * per pair (the 512^3 FP64 heat update spent 47 of its 419 us there).
         * With two columns per lane and a box class (no holes), such a lane has
         * one cell to store and stores just that cell, with no read at all (FP64
         * heat 390 against 405 us); with four columns its up to three cells
         * cost more as single stores than the read (FP32 234 against 207 us). */
        constexpr bool MERGE = MASKED || K > 2;
        vecn<T, K> old[NC / K];
        if (MERGE && row_fits) {
#pragma unroll
          for (int n = 0; n < NC; n += K) {
            bool all = true, any = false;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              all = all && tc.in[n + k];
              any = any || tc.in[n + k];
            }
            if (any && !all)
              old[n / K] = *(gptr<const vecn<T, K>>)(gp<T>(H.out.ptr) + lin<3>(H.out.shape, tc.c[n][0], tc.c[n][1], col));
          }
        }
#pragma unroll
        for (int n = 0; n < NC; n += K) {
          bool all = true, any = false;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            all = all && tc.in[n + k];
            any = any || tc.in[n + k];
          }
          if (!any) continue;
          if (row_fits && (MERGE || all)) {
            const gptr<vecn<T, K>> dst =
                (gptr<vecn<T, K>>)(gp<T>(H.out.ptr) + lin<3>(H.out.shape, tc.c[n][0], tc.c[n][1], col));
            vecn<T, K> w;
#pragma unroll
            for (int k = 0; k < K; ++k) w[k] = (!MERGE || tc.in[n + k]) ? v[n + k] : old[n / K][k];
            put<vecn<T, K>, NT>(dst, w);
          } else {
#pragma unroll
            for (int k = 0; k < K; ++k)
              if (tc.in[n + k]) store<S, OV, NT>(H, tc.c[n + k][0], tc.c[n + k][1], tc.c[n + k][2], v[n + k]);
          }
        }
        continue;
      }
    }
#pragma unroll
    for (int n = 0; n < NC; ++n)
      if (tc.in[n]) store<S, OV, NT>(H, tc.c[n][0], tc.c[n][1], tc.c[n][2], v[n]);
  }
}

/* Tile `tile` of `ntiles` of the statement's class (a list class: a slice of
 * the list, strided over the tiles). */
template <class S, int OV>
__device__ void run_main(const xg_launch_hdr& H, const typename S::Refs& R, const typename S::Params& P,
                         uint32_t tile, int64_t ntiles) {
  constexpr int D = S::D;
  const xg_dom& dom = H.dom;
  if (dom.kind == XG_DOM_LIST) {
    const int64_t stride = ntiles * blockDim.x;
    for (int64_t k = (int64_t)tile * blockDim.x + threadIdx.x; k < dom.count; k += stride) {
      int i0, i1, i2;
      unlin_list<D>(H.out.shape, gp<const int32_t>(dom.list)[k], i0, i1, i2);
      apply<S, OV>(H, R, P, i0, i1, i2);
    }
  } else if constexpr (D == 3 && S::ST3) {
    /* one LDS copy for the box variants (staged 3-D loads) */
    __shared__ __attribute__((aligned(16))) unsigned char st3_lds[S::ST3_LDS];
    if (dom.kind == XG_DOM_MASKED) box_cells<S, OV, true>(H, R, P, tile, st3_lds);
    else if (dom.flags & XG_F_STREAM) box_cells<S, OV, false, true>(H, R, P, tile, st3_lds);
    else box_cells<S, OV, false>(H, R, P, tile, st3_lds);
  } else if constexpr (D == 3) {
    /* 3-D box tiles of a grid larger than the cache store non-temporally */
    if (dom.kind == XG_DOM_MASKED) box_cells<S, OV, true>(H, R, P, tile, nullptr);
    else if (dom.flags & XG_F_STREAM) box_cells<S, OV, false, true>(H, R, P, tile, nullptr);
    else box_cells<S, OV, false>(H, R, P, tile, nullptr);
  } else if (dom.kind == XG_DOM_MASKED) {
    box_cells<S, OV, true>(H, R, P, tile, nullptr);
  } else {
    box_cells<S, OV, false>(H, R, P, tile, nullptr);
  }
}

/* Two in-place statements over the same box in one tile walk (the runtime
 * pairs them only when their classes are that same box and the first does not
 * read the second's store grid at time 0).  Each work-item evaluates both
 * statements at its cells, then stores both.  FWD: the second reads the first's
 * store at time 0, at its own cell only; it takes the value the first just
 * computed for that cell (SB::value_fwd) instead of loading it. */
template <class SA, class SB, int OV, bool FWD>
__device__ void run_pair(const xg_launch_hdr& HA, const typename SA::Refs& RA, const typename SA::Params& PA,
                         const xg_launch_hdr& HB, const typename SB::Refs& RB, const typename SB::Params& PB,
                         uint32_t tile) {
  constexpr int D = SA::D;
  static_assert(SB::D == D, "paired statements share one tile");
  for (int ch = 0; ch < Tile<D>::CH; ++ch) {
    TileCells<D> tc(HA.dom.lo, HA.dom.hi, tile, ch);
    if (tc.tlo[0] >= HA.dom.hi[0]) break;
    constexpr int NC = TileCells<D>::NC;
    tc.clamp_to_tile();  /* evaluate every cell without branching (see box_cells) */
    typename SA::T va[NC];
    typename SB::T vb[NC];
    if (tile_safe<SA>(HA, tc.tlo, tc.thi) && tile_safe<SB>(HB, tc.tlo, tc.thi)) {
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        va[n] = SA::template value<true>(RA, PA, HA.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
        if constexpr (FWD) vb[n] = SB::template value_fwd<true>(RB, PB, HB.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2], va[n]);
        else vb[n] = SB::template value<true>(RB, PB, HB.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
      }
    } else {
#pragma unroll
      for (int n = 0; n < NC; ++n) {
        va[n] = SA::template value<false>(RA, PA, HA.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
        if constexpr (FWD) vb[n] = SB::template value_fwd<false>(RB, PB, HB.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2], va[n]);
        else vb[n] = SB::template value<false>(RB, PB, HB.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
      }
    }
#pragma unroll
    for (int n = 0; n < NC; ++n)
      if (tc.in[n]) {
        store<SA, OV>(HA, tc.c[n][0], tc.c[n][1], tc.c[n][2], va[n]);
        store<SB, OV>(HB, tc.c[n][0], tc.c[n][1], tc.c[n][2], vb[n]);
      }
  }
}

/* Jacobi sweep: out = scratch, src = old level 0.  Every cell of the grid is
 * written exactly once, so after the sweep the two buffers swap roles. */
template <class S, int OV>
__device__ void run_pingpong(const xg_launch_hdr& H, const typename S::Refs& R, const typename S::Params& P,
                             uint32_t tile) {
  constexpr int D = S::D;
  typedef typename S::T T;
  const xg_dom& dom = H.dom;
  const int32_t zero[3] = {0, 0, 0};
  const int32_t* shape = H.out.shape;
  const gptr<const T> __restrict__ src = gp<const T>(H.src.ptr);
  const gptr<T> __restrict__ out = gp<T>(H.out.ptr);
  for (int ch = 0; ch < Tile<D>::CH; ++ch) {
  TileCells<D> tc(zero, shape, tile, ch);
  if (tc.tlo[0] >= shape[0]) break;
  constexpr int NC = TileCells<D>::NC;
  tc.clamp_to_tile();          /* evaluate every cell without branching (see box_cells) */
  bool member[NC];
  int64_t lc[NC];
#pragma unroll
  for (int n = 0; n < NC; ++n) {
    const int* c = tc.c[n];
    lc[n] = lin<D>(shape, c[0], c[1], c[2]);
    if (dom.kind == XG_DOM_BOX) {
      bool m = c[0] >= dom.lo[0] && c[0] < dom.hi[0];
      if (D > 1) m = m && c[1] >= dom.lo[1] && c[1] < dom.hi[1];
      if (D > 2) m = m && c[2] >= dom.lo[2] && c[2] < dom.hi[2];
      member[n] = m;
    } else {
      member[n] = gp<const int32_t>(dom.mask)[lc[n]] == dom.klass;
    }
  }
  T v[NC];
  if (tile_safe<S>(H, tc.tlo, tc.thi)) {
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const T x = S::template value<true>(R, P, H.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
      v[n] = member[n] ? x : src[lc[n]];
    }
  } else {
#pragma unroll
    for (int n = 0; n < NC; ++n) {
      const T x = S::template value<false>(R, P, H.out.shape, tc.c[n][0], tc.c[n][1], tc.c[n][2]);
      v[n] = member[n] ? x : src[lc[n]];
    }
  }
#pragma unroll
  for (int n = 0; n < NC; ++n)
    if (tc.in[n]) out[lc[n]] = v[n];
  }
}

/* Visit every cell of the box [lo, hi) handled by this workgroup's tile. */
template <int D, class F>
__device__ __forceinline__ void for_tile(const int32_t* lo, const int32_t* hi, uint32_t tile, F&& f) {
  for (int ch = 0; ch < Tile<D>::CH; ++ch) {
    TileCells<D> tc(lo, hi, tile, ch);
    if (tc.tlo[0] >= hi[0]) break;
#pragma unroll
    for (int n = 0; n < TileCells<D>::NC; ++n)
      if (tc.in[n]) f(tc.c[n][0], tc.c[n][1], tc.c[n][2]);
  }
}

/* Phase 2 of a two-phase statement: store the temporary at the shifted cell. */
template <class T, int D, int OV>
__device__ void run_phase2(const xg_launch_hdr& H, uint32_t tile, int64_t ntiles) {
  const xg_dom& dom = H.dom;
  const gptr<const T> tmp = gp<const T>(H.src.ptr);
  const int32_t* shape = H.out.shape;
  auto f = [&](int i0, int i1, int i2) {
    st<T, D, OV>(H.out, i0 + dom.so[0], i1 + dom.so[1], i2 + dom.so[2], tmp[lin<D>(shape, i0, i1, i2)]);
  };
  if (dom.kind == XG_DOM_LIST) {
    const int64_t stride = ntiles * blockDim.x;
    for (int64_t k = (int64_t)tile * blockDim.x + threadIdx.x; k < dom.count; k += stride) {
      int i0, i1, i2;
      unlin_list<D>(shape, gp<const int32_t>(dom.list)[k], i0, i1, i2);
      f(i0, i1, i2);
    }
  } else if (dom.kind == XG_DOM_MASKED) {
    for_tile<D>(dom.lo, dom.hi, tile, [&](int i0, int i1, int i2) {
      if (gp<const int32_t>(dom.mask)[lin<D>(shape, i0, i1, i2)] == dom.klass) f(i0, i1, i2);
    });
  } else {
    for_tile<D>(dom.lo, dom.hi, tile, f);
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
      unlin_list<D>(H.out.shape, gp<const int32_t>(dom.list)[k], i0, i1, i2);
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
    if (dom.kind == XG_DOM_MASKED && gp<const int32_t>(dom.mask)[lin<D>(H.out.shape, i0, i1, i2)] != dom.klass) continue;
    apply<S, OV>(H, R, P, i0, i1, i2);
  }
}

template <class T>
__device__ __forceinline__ T sq(T x) { return x * x; }

/* ---- division by a launch-uniform divisor ---------------------------------
 * A statement's per-cell divisions whose divisor holds no grid load have one
 * divisor per launch.  When that divisor d is +-2^k with 2^-k representable
 * (biased exponent 1..2046 in FP64, 1..254 in FP32; 2^-1023 and 2^-127 are
 * exact subnormals), r = 1/d is exact and a * r is the same real number as
 * a / d for every a, so IEEE multiplication and division round it to the same
 * result, with the same infinities, signed zeros, subnormals and overflow.
 * Kernels test the divisors once per workgroup (dv_ok) and run the
 * instantiation DV = true only when all of them pass; the reciprocal is built
 * from the exponent bits, with no division. */
template <class T> struct is_fp { static constexpr bool value = false; };
template <> struct is_fp<float> { static constexpr bool value = true; };
template <> struct is_fp<double> { static constexpr bool value = true; };

__device__ __forceinline__ bool rcp_exact(double d) {
  uint64_t b;
  __builtin_memcpy(&b, &d, 8);
  const uint64_t e = (b >> 52) & 0x7ffull;
  return (b & 0xfffffffffffffull) == 0 && e >= 1 && e <= 2046;
}
__device__ __forceinline__ bool rcp_exact(float d) {
  uint32_t b;
  __builtin_memcpy(&b, &d, 4);
  const uint32_t e = (b >> 23) & 0xffu;
  return (b & 0x7fffffu) == 0 && e >= 1 && e <= 254;
}
template <class I>
__device__ __forceinline__ bool rcp_exact(I) { return false; }   /* integer divisors: no reciprocal form */

/* 1/d for d passing rcp_exact. */
__device__ __forceinline__ double rcp_pow2(double d) {
  uint64_t b;
  __builtin_memcpy(&b, &d, 8);
  const uint64_t e = (b >> 52) & 0x7ffull;
  const uint64_t r = (b & 0x8000000000000000ull) | (e < 2046 ? (2046 - e) << 52 : 1ull << 51);
  double x;
  __builtin_memcpy(&x, &r, 8);
  return x;
}
__device__ __forceinline__ float rcp_pow2(float d) {
  uint32_t b;
  __builtin_memcpy(&b, &d, 4);
  const uint32_t e = (b >> 23) & 0xffu;
  const uint32_t r = (b & 0x80000000u) | (e < 254 ? (254 - e) << 23 : 1u << 22);
  float x;
  __builtin_memcpy(&x, &r, 4);
  return x;
}

template <bool DV, class A, class B>
__device__ __forceinline__ auto udiv(A a, B b) -> decltype(a / b) {
  typedef decltype(a / b) R;
  if constexpr (DV && is_fp<R>::value && is_fp<B>::value) return (R)a * rcp_pow2((R)b);
  else return a / b;
}

/* A compile-time boolean carried as a value (selects a template instantiation). */
template <bool B>
struct bconst {
  static constexpr bool value = B;
};

/* ---- temporally blocked loops (2-D), LDS regions --------------------------
 * A workgroup owns a core tile of the grid and keeps core + halo (an RH x RW
 * region) of the loop's grid in LDS for several iterations of the loop body.
 * Region cells outside the grid hold 0 and are never valid.  Only cells at
 * least the body's reach inside the region are computed, so a read of the
 * grid at region coordinates (r, c) is always inside the region.  With CHECK,
 * `ok` records whether the value read is exact (valid), so a validation launch
 * can prove that every core cell it writes is exact. */
template <class T, bool CHECK, int RH, int RW>
__device__ __forceinline__ T tb_ld(const T* lds, const uint8_t* vld, bool& ok, int r, int c) {
  const int i = r * RW + c;
  if (CHECK) ok = ok && vld[i];
  return lds[i];
}

/* ---- temporally blocked loops (2-D), register columns --------------------
 * Lane c of a wave holds column c of a 64-column region; a read at column
 * offset o is the value lane c + o holds, moved across lanes with DPP
 * (wave_shl:1 moves lane c + 1 to lane c, wave_shr:1 lane c - 1 to lane c;
 * lanes whose source lies outside the wave get 0).  Every lane of the wave
 * must execute the move (it reads other lanes' registers). */
template <int O>
__device__ __forceinline__ uint32_t lane_shift_u32(uint32_t x) {
  if (O == 0) return x;
  int v = (int)x;
#pragma unroll
  for (int k = 0; k < (O > 0 ? O : -O); ++k)
    v = __builtin_amdgcn_update_dpp(0, v, O > 0 ? 0x130 : 0x138, 0xf, 0xf, false);
  return (uint32_t)v;
}

template <int O, class T>
__device__ __forceinline__ T lane_shift(T x) {
  if (O == 0) return x;
  if (sizeof(T) == 8) {
    uint64_t b;
    __builtin_memcpy(&b, &x, 8);
    const uint32_t lo = lane_shift_u32<O>((uint32_t)b), hi = lane_shift_u32<O>((uint32_t)(b >> 32));
    b = ((uint64_t)hi << 32) | lo;
    T r;
    __builtin_memcpy(&r, &b, 8);
    return r;
  } else if (sizeof(T) == 4) {
    uint32_t b;
    __builtin_memcpy(&b, &x, 4);
    b = lane_shift_u32<O>(b);
    T r;
    __builtin_memcpy(&r, &b, 4);
    return r;
  } else {
    return (T)lane_shift_u32<O>((uint32_t)x);
  }
}

/* ---- batched launches ---------------------------------------------------
 * Global tile g of a batched launch belongs to the entry whose range
 * [tile0, tile0 + ntiles) holds it (entries sorted by tile0).  Workgroups walk
 * the global tiles with a stride of the grid size (the host keeps the grid a
 * multiple of 8, so a workgroup stays on one XCD and the remap keeps each
 * XCD's tiles contiguous). */
template <class E>
__device__ __forceinline__ int find_entry(const E* __restrict__ ents, int32_t n, int64_t g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ents[mid].tile0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class E, class F>
__device__ __forceinline__ void for_batch_tiles(const E* __restrict__ ents, int32_t n, int64_t total, F&& f) {
  for (int64_t p = blockIdx.x; p < total; p += gridDim.x) {
    /* A persistent workgroup's next tile rewrites LDS that its waves may still
     * be reading for the last one (a staged 3-D tile's rows): wait for all of
     * them first.  p is the same for every work-item of the workgroup. */
    if (p != (int64_t)blockIdx.x) __syncthreads();
    const int64_t g = xcd_remap((uint32_t)p, (uint32_t)total);
    const E& e = ents[find_entry(ents, n, g)];
    f(e, (uint32_t)(g - e.tile0));
  }
}

/* Bitwise OR of `v` over the workgroup (all work-items must call it). */
__device__ __forceinline__ uint32_t block_or(uint32_t v) {
  __shared__ uint32_t acc;
  if (threadIdx.x == 0) acc = 0u;
  __syncthreads();
  if (v) atomicOr(&acc, v);
  __syncthreads();
  const uint32_t r = acc;
  __syncthreads();
  return r;
}

}  // namespace xg

#endif
