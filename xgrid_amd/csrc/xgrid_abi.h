/* Launch-argument layout shared by the host runtime (x86-64, gcc/clang) and the
 * device kernels (gfx950, hipRTC).  Both sides are LP64 with natural alignment,
 * so every struct here has one layout; explicit padding keeps it obvious.
 *
 * A statement kernel takes (xg_launch_hdr, xg_gview ref[NR], params):
 *   - hdr: the iteration domain of the store grid's boundary class and the
 *          store target;
 *   - ref: one view per distinct (grid, time level) the statement loads;
 *   - params: the scalars the statement's expression captures.
 */
#ifndef XGRID_ABI_H
#define XGRID_ABI_H

#if defined(__HIPCC_RTC__)
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#else
#include <stdint.h>
#endif

#define XG_MAX_DIMS 3
/* 3-D workgroup tiles: chunks of 4 planes a workgroup does in turn */
#define XG_TILE3_CHUNKS 1
#define XG_MAX_REFS 16

/* domain kinds: how a boundary class's cells are enumerated */
#define XG_DOM_BOX 0     /* the class is exactly the box [lo, hi): no mask read */
#define XG_DOM_LIST 1    /* compacted list of linear cell indices */
#define XG_DOM_MASKED 2  /* the box [lo, hi) with a per-cell mask test */
#define XG_DOM_EMPTY 3   /* the class has no cell: nothing is launched */

/* statement flags */
#define XG_F_PINGPONG 1  /* iterate the whole grid; class cells computed, others copied from src */
#define XG_F_TO_TMP 2    /* phase 1 of a two-phase statement: value -> src[cell] (no store offset) */
#define XG_F_UNIFORM 4   /* every loaded view has the store grid's rank and shape */
#define XG_F_STREAM 8    /* in place, store grid above the last-level cache: 3-D box tiles store non-temporally */

typedef struct xg_gview {
  void* ptr;                 /* element 0 of one time level */
  int32_t shape[XG_MAX_DIMS];
  int32_t pad;
} xg_gview;                  /* 24 bytes */

typedef struct xg_dom {
  int32_t kind;              /* XG_DOM_* */
  int32_t klass;             /* boundary class (mask value) */
  int32_t lo[XG_MAX_DIMS];
  int32_t hi[XG_MAX_DIMS];   /* exclusive */
  int64_t count;             /* cells in the class */
  const int32_t* list;       /* XG_DOM_LIST */
  const int32_t* mask;       /* XG_DOM_MASKED / pingpong membership */
  int32_t flags;             /* XG_F_* */
  int32_t so[XG_MAX_DIMS];   /* store offsets (phase 2) */
} xg_dom;                    /* 4+4+12+12 +(pad 0)+ 8+8+8+4+12 = 72 bytes */

typedef struct xg_launch_hdr {
  xg_dom dom;
  xg_gview out;              /* store target (level 0, the scratch in pingpong) */
  xg_gview src;              /* pingpong source / two-phase temporary */
} xg_launch_hdr;             /* 120 bytes */

/* Temporally blocked loop: several iterations of a loop body of stencil
 * statements on one grid, per workgroup tile, in LDS. */
#define XG_TB_MAX_STMTS 8
typedef struct xg_tb_hdr {
  xg_gview src;                    /* the grid's level 0 before the launch */
  xg_gview dst;                    /* receives every cell after `iters` iterations */
  const int32_t* mask;             /* boundary mask (null: every cell is class 0) */
  int32_t* fail;                   /* validation launch: set when a core cell is not exact */
  int32_t klass[XG_TB_MAX_STMTS];  /* class of each statement */
  int32_t iters;                   /* iterations in this launch (<= the compiled maximum) */
  int32_t uniform;                 /* every loaded view has the grid's shape */
} xg_tb_hdr;                       /* 24 + 24 + 8 + 8 + 32 + 8 = 104 bytes */

/* Batched launches: one launch runs one statement (or blocked loop, or chain
 * of boundary statements) for many independent instances of one kernel.  A
 * device table holds one entry per instance; instance i owns the global tiles
 * [tile0, tile0 + ntiles) of the launch (entries sorted by tile0). */
typedef struct xg_bent {
  xg_launch_hdr hdr;         /* as for a single launch */
  const void* refs;          /* this instance's xg_gview[] (in the table) */
  const void* params;        /* its parameter block (in the table) */
  int64_t tile0;
  int64_t ntiles;
} xg_bent;                   /* 120 + 8 + 8 + 8 + 8 = 152 bytes */

typedef struct xg_btent {
  xg_tb_hdr th;                          /* iters may differ between instances */
  const void* refs[XG_TB_MAX_STMTS];
  const void* params[XG_TB_MAX_STMTS];
  int64_t tile0;
  int64_t ntiles;
} xg_btent;                              /* 104 + 64 + 64 + 16 = 248 bytes */

#endif
