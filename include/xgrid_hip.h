/* xgrid_hip.h - C ABI of libxgrid_hip.so, the MI355X (gfx950) execution backend
 * for xgrid stencil kernels.
 *
 * Every entry point returns XG_OK (0) or a negative XG_ERR_* code; the message
 * of the last failure on the calling thread is xg_last_error().  Nothing
 * aborts and no C++ exception crosses this boundary.  Plain pointers and sizes
 * only.  Citations name the reference (acasta-yhliu/xgrid) interface each entry
 * point replaces.
 */
#ifndef XGRID_HIP_H
#define XGRID_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XG_OK 0
#define XG_ERR_HIP (-1)     /* a HIP runtime call failed */
#define XG_ERR_RTC (-2)     /* hipRTC compilation failed (log in xg_last_error) */
#define XG_ERR_ARG (-3)     /* invalid argument / handle */
#define XG_ERR_MEM (-4)     /* allocation refused (device memory or budget) */
#define XG_ERR_STATE (-5)   /* call out of order (e.g. no device initialised) */
#define XG_ERR_IO (-6)      /* code-object cache file error */

/* statement execution modes (decided from the kernel's source) */
#define XG_MODE_INPLACE 0   /* reference K1/K3 nest, generator.py:289-304,353-361 */
#define XG_MODE_PINGPONG 1  /* reference K2 implicit nest as a buffer swap, generator.py:312-352 */
#define XG_MODE_TWOPHASE 2  /* reference K2 with a shifted store: compute, then store */

#define XG_DESC_MAX_REFS 16
#define XG_DESC_MAX_HAZ 8

typedef struct xg_grid xg_grid;
typedef struct xg_kernel xg_kernel;
typedef struct xg_module xg_module;

/* One stencil statement of a kernel (filled by the emitter). */
typedef struct xg_stmt_desc {
  int32_t store_arg;                         /* index of the store grid among the kernel's grid args */
  int32_t klass;                             /* boundary class (mask value) the statement updates */
  int32_t mode;                              /* XG_MODE_* */
  int32_t dims;
  int32_t elem_bytes;
  int32_t nrefs;                             /* distinct (grid, level) arrays loaded */
  int32_t ref_arg[XG_DESC_MAX_REFS];
  int32_t ref_level[XG_DESC_MAX_REFS];       /* -1: only the grid's shape is read */
  int32_t store_off[3];
  int32_t nhazard;                           /* same-grid time-0 read offsets relative to the store */
  int32_t hazard[XG_DESC_MAX_HAZ][3];
  int32_t param_bytes;
  int32_t overstep;                          /* 0 none, 1 limit, 2 wrap (xgrid/util/init.py:15) */
  int32_t group;                             /* fusion group (-1: none) */
  int32_t group_pos;
  int32_t pair_fwd;                          /* fn_pair forwards this store to the next statement */
  int32_t tile3_cols;                        /* 3-D fn_main: columns per tile (64, or 128 / 256 for staged
                                                loads of 16 bytes per lane); 0 means 64 */
  void* fn_main;                             /* hipFunction_t of the class kernel */
  void* fn_pingpong;                         /* hipFunction_t of the Jacobi sweep (PINGPONG) */
  void* fn_phase2;                           /* hipFunction_t storing a temporary (two-phase) */
  void* fn_main_b;                           /* batched variants (many instances per launch) */
  void* fn_pingpong_b;
  void* fn_phase2_b;
  void* fn_pair;                             /* this statement and the next in one launch (both in
                                                place over one box), or NULL */
} xg_stmt_desc;

/* A run of consecutive statements that may be chained in one workgroup. */
typedef struct xg_group_desc {
  int32_t first;
  int32_t count;
  void* fn;                                  /* hipFunction_t taking (uint32 enable, {hdr, refs, params} x count) */
  void* fn_b;                                /* batched: one workgroup per instance */
} xg_group_desc;

/* A `for` loop whose body is stencil statements on one 2-D grid, which may run
 * temporally blocked: `tmax` iterations per launch on core0 x core1 tiles with
 * a halo of `halo` cells kept in LDS (reference: the same statements as
 * separate nests, xgrid/lang/generator.py:369-374 around :289-361). */
#define XG_LOOP_MAX_STMTS 8
typedef struct xg_loop_desc {
  int32_t nstmts;
  int32_t sids[XG_LOOP_MAX_STMTS];
  int32_t grid_arg;
  int32_t core0, core1, halo, tmax;
  int32_t threads;                           /* work-items per workgroup */
  void* fn;                                  /* hipFunction_t: tmax-bounded blocked iterations */
  void* fn_check;                            /* the same, proving every written core cell exact */
  void* fn_b;                                /* batched variants */
  void* fn_check_b;
  int32_t alt;                               /* plain calls on a grid of at most alt_max_cells cells run
                                                loop plan `alt` (another tile shape of the same loop)
                                                instead; -1: none */
  int32_t alt_pad;
  int64_t alt_max_cells;
} xg_loop_desc;

/* Function table handed to generated host stubs (the stub runs the kernel's
 * scalar control flow and issues one stmt() per executed stencil statement;
 * for a blockable loop it first offers the whole loop to loop(), which returns
 * 1 when it ran all `trip` iterations and 0 when the stub must run them). */
typedef struct xg_api {
  int (*begin)(xg_kernel* k, xg_grid* const* grids, int32_t ngrids, const void* key, int64_t key_bytes);
  int (*stmt)(xg_kernel* k, int32_t sid, const void* params, int64_t param_bytes);
  int (*end)(xg_kernel* k);
  const int32_t* (*shape)(xg_grid* g);
  int (*loop)(xg_kernel* k, int32_t lid, int64_t trip, const void* const* params, int32_t nparams);
} xg_api;

/* ---- errors and device ------------------------------------------------- */
/* Replaces Logger.dead exceptions (xgrid/util/logging.py:58-60). */
const char* xg_last_error(void);
/* Select the device and create the backend's stream.  Replaces nothing in the
 * reference (host only); xgrid.init (xgrid/util/init.py:48-66) calls it. */
int xg_init(int device);
int xg_initialized(void);
int xg_device_name(char* buf, int64_t buf_bytes, int* cus, int64_t* hbm_bytes);
int xg_synchronize(void);
int xg_mem_info(int64_t* free_bytes, int64_t* total_bytes);

/* ---- code objects ------------------------------------------------------ */
/* hipRTC-compile HIP source for gfx950 and write the code object to out_path.
 * Needs no GPU.  Replaces Compiler.compile (xgrid/util/ffi.py:59-94). */
int xg_rtc_compile(const char* src, const char* name, const char* const* opts, int32_t nopts,
                   const char* const* hdr_names, const char* const* hdr_srcs, int32_t nhdr,
                   const char* out_path);
/* Load a code object and look up kernels.  Replaces Library / Library.function
 * (xgrid/util/ffi.py:13-34). */
int xg_module_load(const char* path, xg_module** out);
int xg_module_function(xg_module* m, const char* name, void** out_fn);
int xg_module_unload(xg_module* m);

/* ---- grids: device-resident time-level rings --------------------------- */
/* Storage of one Grid (xgrid/xgrid/__init__.py:22-41): shape, element size,
 * levels allocated on demand. */
int xg_grid_create(int32_t dims, const int32_t* shape, int32_t elem_bytes, xg_grid** out);
/* Releases the grid's memory behind the kernels already queued (no wait for them). */
int xg_grid_destroy(xg_grid* g);
/* Grow with zero-filled levels or truncate to `depth` (Grid._extend_time, :43-47). */
int xg_grid_set_depth(xg_grid* g, int32_t depth);
int32_t xg_grid_depth(xg_grid* g);
/* Rotate: the oldest level becomes level 0 (Grid._op_invoke, :49-54). */
int xg_grid_tick(xg_grid* g);
/* Host <-> device copies of one level (Grid.now / fill, :70-80).  Blocking.  A
 * grid that no kernel call has used yet is uploaded without waiting for the
 * queued kernels; otherwise the upload waits for them (they may read it). */
int xg_grid_upload(xg_grid* g, int32_t level, const void* host);
int xg_grid_download(xg_grid* g, int32_t level, void* host);
/* Copy one level to the host once the work before `event` (xg_event_record) is
 * done, without waiting for kernels queued after it, which keep running.  No
 * reference counterpart: the serving loop reads a finished request's fields
 * while the next tick runs.  The caller ensures that nothing queued after the
 * event writes this grid.  Blocking. */
int xg_grid_download_after(xg_grid* g, int32_t level, void* host, void* event);
/* Install the int32 boundary mask (Grid.boundary, :41) and plan each class. */
int xg_grid_set_mask(xg_grid* g, const int32_t* mask);
/* Plan of one class: kind (XG_DOM_*), cell count, box. */
int xg_grid_class_info(xg_grid* g, int32_t klass, int32_t* kind, int64_t* count, int32_t* lo, int32_t* hi);
/* Device address of one level.  Work the caller queues with it on its own streams
 * is waited for before the grid's memory is released or replaced. */
void* xg_grid_level_ptr(xg_grid* g, int32_t level);

/* ---- kernels: one compiled @xgrid.kernel ------------------------------- */
/* Create the launch plan of a kernel from its statements.  Replaces the
 * native handle Generator(op).result returns (xgrid/lang/operator.py:31-34). */
int xg_kernel_create(int32_t nstmts, const xg_stmt_desc* stmts, int32_t ngroups,
                     const xg_group_desc* groups, int32_t nloops, const xg_loop_desc* loops,
                     int32_t ngrids, xg_kernel** out);
int xg_kernel_destroy(xg_kernel* k);
/* Status of the last call made through the kernel's stub (0 or XG_ERR_*). */
int xg_kernel_status(xg_kernel* k);
/* Options: graphs (replay repeated calls as hipGraphs), fuse (chain small
 * boundary statements in one workgroup), blocked (run blockable loops
 * temporally blocked). */
int xg_kernel_set_options(xg_kernel* k, int32_t graphs, int32_t fuse, int32_t blocked);
/* Launch counters and graph cache statistics. */
int xg_kernel_counters(xg_kernel* k, int64_t* launches, int64_t* fused, int64_t* graph_replays,
                       int64_t* graph_captures);
/* Blocked-loop statistics: launches, iterations run blocked, and the plan
 * state of loop `lid` (1 exact and in use, 0 not yet validated, -1 declined). */
int xg_kernel_loop_info(xg_kernel* k, int32_t lid, int64_t* launches, int64_t* iterations, int32_t* state);
/* Run loop `lid` blocked for `trip` iterations (the launches a call makes:
 * chunks of tmax iterations), `reps` times back to back, on the grids and
 * parameters of its last call, between two HIP events; returns ms per launch.
 * Advances the grid's state. */
int xg_kernel_time_loop(xg_kernel* k, int32_t lid, int32_t trip, int32_t reps, double* ms_per_launch);
/* Bracket every direct launch of statement `sid` with HIP events on the
 * backend stream; xg_kernel_stmt_time returns the summed kernel time. */
int xg_kernel_time_stmt(xg_kernel* k, int32_t sid, int32_t enable);
int xg_kernel_stmt_time(xg_kernel* k, int32_t sid, double* total_ms, int64_t* launches);
/* Launch statement `sid` `reps` times back to back on the grids and with the
 * parameters of its last call, between two HIP events; returns ms per launch.
 * Advances the grids' state like that many extra executions would. */
int xg_kernel_time_isolated(xg_kernel* k, int32_t sid, int32_t reps, double* ms_per_launch);
/* Executions of each statement (stmt() calls, replayed calls included) since
 * the last reset; fills counts[0..n). */
int xg_kernel_stmt_counts(xg_kernel* k, int64_t* counts, int32_t n, int32_t reset);
/* The stub function table (pointers into this library). */
const xg_api* xg_api_table(void);

/* ---- batches: many instances of one kernel per launch ------------------ */
/* No reference counterpart (the reference runs one kernel call at a time,
 * xgrid/lang/operator.py:29-44).  A batch holds independent instances of one
 * compiled kernel - each its own xg_kernel plan (xg_kernel_create with the
 * same descriptors), its own grids of any size and its own arguments - and
 * advances all of them one call per step, with one launch per statement
 * covering the tiles of every instance, blocked loops batched the same way,
 * and whole steps replayed as hipGraphs when they repeat.  Results equal those
 * of calling each instance's kernel on its own. */
typedef struct xg_batch xg_batch;
/* packed_entry: the stub's xg_entry_packed(kernel, api, args); tick: the
 * kernel rotates its grids' time levels on each call (Operator.tick). */
int xg_batch_create(void* packed_entry, int32_t tick, xg_batch** out);
int xg_batch_destroy(xg_batch* b);
/* Add an instance: its plan, its arguments packed as the stub's struct
 * xg_args (grid arguments are xg_grid handles), and its grid arguments in
 * order (ticked once per occurrence before each call).  A grid may belong to
 * one instance only.  Returns a stable id. */
int xg_batch_add(xg_batch* b, xg_kernel* k, const void* args, int64_t arg_bytes, xg_grid* const* tick_grids,
                 int32_t nticks, int64_t* id);
int xg_batch_remove(xg_batch* b, int64_t id);
int xg_batch_size(xg_batch* b);
/* graphs: replay repeated steps; max_blocks: 0 launches one workgroup per
 * tile, else at most this many (rounded down to a multiple of 8) walk the tiles. */
int xg_batch_set_options(xg_batch* b, int32_t graphs, int64_t max_blocks);
/* Advance every instance by `steps` calls (asynchronous, on the backend stream). */
int xg_batch_step(xg_batch* b, int32_t steps);
int xg_batch_counters(xg_batch* b, int64_t* launches, int64_t* replays, int64_t* captures, int64_t* steps,
                      int64_t* validations);

/* ---- FP64 reductions over one boundary class --------------------------- */
/* No reference counterpart: the DSL has no reductions (a grid load is only
 * allowed inside a stencil store, xgrid/lang/generator.py:62-65).  These give
 * FP32-stored solvers an FP64 convergence test and error measure.  Element
 * type codes say how to read each grid's bytes. */
#define XG_T_F32 1
#define XG_T_F64 2
/* l2 = sqrt(sum d^2), linf = max |d| (NaN if any d is NaN) of
 * d = a[level_a] - b[level_b] over the cells of class `klass` of a; b may be
 * NULL (norms of a).  a and b have one shape.  cells = the class's cell count.
 * Synchronises the backend stream. */
int xg_grid_norm_diff(xg_grid* a, int32_t level_a, int32_t type_a, xg_grid* b, int32_t level_b, int32_t type_b,
                      int32_t klass, double* l2, double* linf, int64_t* cells);
/* No counterpart in the reference.  For callers that cannot allocate (tight
 * serving loops, harnesses that forbid hipMalloc): xg_grid_norm_diff with the
 * same arguments and results, but it allocates nothing, not even on its first
 * call.  Its partials live in a static device array of 1024 partials + 1
 * result (16,400 bytes per GPU, held until the process ends; its address is
 * looked up by xg_init).  Any class size fits: at most 1024 workgroups each
 * reduce a strided slice.  Every call on a GPU shares that GPU's array, so
 * calls that overlap on one GPU (several host threads at once) race.
 * Synchronises the backend stream. */
int xg_grid_norm_diff_noalloc(xg_grid* a, int32_t level_a, int32_t type_a, xg_grid* b, int32_t level_b,
                              int32_t type_b, int32_t klass, double* l2, double* linf, int64_t* cells);
/* Norms of the residual r = sum_a coef[a] (p[+e_a] - 2p + p[-e_a]) - rhs_scale * rhs
 * of a Poisson equation over the cells of class `klass` of p, every cell of
 * which must lie at least one cell inside the grid (else XG_ERR_ARG).  coef
 * holds one coefficient per grid axis (axis 0 first).  Synchronises. */
int xg_grid_poisson_residual(xg_grid* p, int32_t level_p, int32_t type_p, xg_grid* rhs, int32_t level_rhs,
                             int32_t type_rhs, int32_t klass, const double* coef, double rhs_scale, double* l2,
                             double* linf);

/* ---- measurement probes ------------------------------------------------ */
/* Record an event on the backend stream (a point in its work); elapsed time
 * between two recorded events (waits for the second); release one.  No
 * reference counterpart: the serving loop measures device busy time with them. */
int xg_event_record(void** out);
int xg_event_elapsed(void* a, void* b, double* ms);
int xg_event_destroy(void* e);
/* Device copy of `bytes` bytes repeated `iters` times; returns GB/s (read+write). */
int xg_copy_probe(int64_t bytes, int32_t iters, double* gbps, double* ms_per_iter);

#ifdef __cplusplus
}
#endif
#endif
