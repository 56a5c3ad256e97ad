// libxgrid_hip.so - host runtime of the MI355X backend (C ABI in include/xgrid_hip.h).
//
// Owns device memory (time-level rings of every Grid, their boundary-mask
// plans, Jacobi scratch buffers), compiles generated HIP with hipRTC, launches
// statement kernels on one stream, chains small boundary statements into one
// workgroup, and replays repeated kernel calls as hipGraphs.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <climits>
#include <mutex>
#include <vector>

#include <unistd.h>

#include "xgrid_internal.h"
#include "xgrid_kernel.h"

namespace xgi {

thread_local std::string t_err;
Device g_dev;

/* A serving loop admits and retires grids while other requests' kernels run:
 * hipFree waits for the whole device, and so did the stream synchronisations
 * around allocation and copies, so each admission or retirement left the
 * device idle until the host queued work again.  Grid and
 * plan memory therefore comes from a stream-ordered pool.  It is taken (and
 * cleared) on the copy stream, which the host then waits for, so it is ready
 * for any work enqueued afterwards on either stream; it is given back on the
 * kernel stream, behind the kernels already queued there that may use it.  The
 * pool never makes the copy stream wait for the kernel stream (internal
 * dependencies off): a block freed behind queued kernels is reused only once
 * those kernels are done. */
int dev_alloc(size_t bytes, bool zero, void** out) {
  void* p = nullptr;
  const size_t n = bytes ? bytes : 1;
  hipError_t e = hipMallocFromPoolAsync(&p, n, g_dev.pool, g_dev.copy);
  if (e != hipSuccess) return fail(XG_ERR_MEM, std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  if (zero) e = hipMemsetAsync(p, 0, n, g_dev.copy);
  if (e == hipSuccess) e = hipStreamSynchronize(g_dev.copy);
  if (e != hipSuccess) {
    (void)hipFreeAsync(p, g_dev.copy);
    return fail(XG_ERR_HIP, std::string("device allocation set-up failed: ") + hipGetErrorString(e));
  }
  *out = p;
  return XG_OK;
}

/* While the kernel stream records a graph (another thread inside a call), a free
 * enqueued there would become part of the graph: such frees wait in a list until
 * the next release made outside a capture. */
void dev_free(void* p) {
  static std::mutex mu;
  static std::vector<void*> held;
  std::lock_guard<std::mutex> lock(mu);
  if (p) held.push_back(p);
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(g_dev.stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return;
  for (void* q : held) (void)hipFreeAsync(q, g_dev.stream);
  held.clear();
}

}  // namespace xgi

using namespace xgi;

namespace {

uint64_t g_mask_version = 1;

void free_plans(xg_grid* g) {
  for (auto& kv : g->plans)
    dev_free(kv.second.d_list);
  g->plans.clear();
}

void default_plan(xg_grid* g) {
  free_plans(g);
  ClassPlan p;
  p.kind = XG_DOM_BOX;
  p.count = g->n;
  for (int d = 0; d < 3; ++d) {
    p.lo[d] = 0;
    p.hi[d] = d < g->dims ? g->shape[d] : 1;
  }
  g->plans[0] = p;
  g->mask_version = g_mask_version++;
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace xgi {

uint64_t fnv1a(const uint8_t* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

int set_status(xg_kernel* k, int code) {
  if (k->status == XG_OK && code != XG_OK) {
    k->status = code;
    k->status_msg = t_err;
  }
  return code;
}

// Does some cell of the statement's class read (at time 0) a cell of the same
// class that another cell of the statement writes?
bool compute_hazard(const xg_stmt_desc& d, const xg_grid* G) {
  if (d.nhazard == 0) return false;
  const ClassPlan& p = G->plan(d.klass);
  if (p.kind == XG_DOM_EMPTY) return false;
  const int D = G->dims;
  auto class_at = [&](int64_t c) -> int32_t { return G->h_mask.empty() ? 0 : G->h_mask[(size_t)c]; };
  auto neighbour = [&](const int* i, const int32_t* o, int64_t* out) -> bool {
    int j[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) {
      j[a] = i[a] + o[a];
      if (d.overstep == 2) j[a] = ((j[a] % G->shape[a]) + G->shape[a]) % G->shape[a];
      else if (d.overstep == 1) j[a] = std::min(std::max(j[a], 0), G->shape[a] - 1);
    }
    int64_t c = j[0];
    if (D > 1) c = c * G->shape[1] + j[1];
    if (D > 2) c = c * G->shape[2] + j[2];
    if (c < 0 || c >= G->n) return false;
    *out = c;
    return true;
  };
  int i[3];
  for (i[0] = p.lo[0]; i[0] < p.hi[0]; ++i[0])
    for (i[1] = p.lo[1]; i[1] < (D > 1 ? p.hi[1] : 1); ++i[1])
      for (i[2] = p.lo[2]; i[2] < (D > 2 ? p.hi[2] : 1); ++i[2]) {
        int64_t c = i[0];
        if (D > 1) c = c * G->shape[1] + i[1];
        if (D > 2) c = c * G->shape[2] + i[2];
        if (class_at(c) != d.klass) continue;
        for (int h = 0; h < d.nhazard; ++h) {
          int64_t nb;
          if (neighbour(i, d.hazard[h], &nb) && nb != c && class_at(nb) == d.klass) return true;
        }
      }
  return false;
}

// Effective execution mode of one statement for the grids bound to this call.
int effective_mode(xg_kernel* k, StmtPlan& sp) {
  const xg_stmt_desc& d = sp.d;
  if (d.mode != XG_MODE_INPLACE) return d.mode;
  xg_grid* G = k->grids[d.store_arg];
  // another argument bound to the same grid and read at level 0 aliases the store
  for (int r = 0; r < d.nrefs; ++r)
    if (d.ref_level[r] == 0 && d.ref_arg[r] != d.store_arg && k->grids[d.ref_arg[r]] == G) return XG_MODE_TWOPHASE;
  if (d.nhazard) {
    if (sp.haz_version != G->mask_version) {
      sp.haz = compute_hazard(d, G);
      sp.haz_version = G->mask_version;
      if (sp.haz)
        fprintf(stderr, "xgrid: statement %d reads cells of its own boundary class at time 0; "
                        "running it compute-then-store\n", (int)(&sp - &k->stmts[0]));
    }
    if (sp.haz) return XG_MODE_TWOPHASE;
  }
  return XG_MODE_INPLACE;
}

void fill_dom(xg_dom& dom, const ClassPlan& p, const xg_grid* G, int32_t klass) {
  std::memset(&dom, 0, sizeof dom);
  dom.kind = p.kind;
  dom.klass = klass;
  for (int d = 0; d < 3; ++d) {
    dom.lo[d] = p.lo[d];
    dom.hi[d] = p.hi[d];
  }
  dom.count = p.count;
  dom.list = p.d_list;
  dom.mask = G->d_mask;
}

void fill_view(xg_gview& v, const xg_grid* G, void* ptr) {
  std::memset(&v, 0, sizeof v);
  v.ptr = ptr;
  for (int d = 0; d < 3; ++d) v.shape[d] = d < G->dims ? G->shape[d] : 1;
}

int32_t stream_flag(const xg_grid* G) { return G->bytes() > kStreamBytes ? XG_F_STREAM : 0; }

// Does every array the statement loads have the store grid's rank and shape?
bool uniform_refs(xg_kernel* k, const xg_stmt_desc& d) {
  const xg_grid* G = k->grids[(size_t)d.store_arg];
  for (int r = 0; r < d.nrefs; ++r) {
    if (d.ref_level[r] < 0) continue;
    const xg_grid* R = k->grids[(size_t)d.ref_arg[r]];
    if (R->dims != G->dims) return false;
    for (int a = 0; a < G->dims; ++a)
      if (R->shape[a] != G->shape[a]) return false;
  }
  return true;
}

int fill_refs(xg_kernel* k, const xg_stmt_desc& d, std::vector<xg_gview>& refs) {
  refs.assign((size_t)std::max(1, d.nrefs), xg_gview{});
  for (int r = 0; r < d.nrefs; ++r) {
    xg_grid* R = k->grids[d.ref_arg[r]];
    void* ptr = nullptr;
    if (d.ref_level[r] >= 0) {
      if (d.ref_level[r] >= (int)R->levels.size())
        return fail(XG_ERR_STATE, "grid has fewer time levels than the statement reads");
      ptr = R->levels[(size_t)d.ref_level[r]];
    }
    fill_view(refs[(size_t)r], R, ptr);
  }
  return XG_OK;
}

// workgroups covering the box [lo, hi) with the tile of a D-dimensional statement
int64_t tiles_for(int D, const int32_t* lo, const int32_t* hi, int cols3) {
  if (D == 1) return ceil_div(hi[0] - lo[0], 1024);
  if (D == 2) return ceil_div(hi[0] - lo[0], 4) * ceil_div(hi[1] - lo[1], 64);
  // column tiles of cols3 (64, 128 or 256) columns start at multiples of cols3 (TileCells in xgrid_device.h)
  return ceil_div(hi[0] - lo[0], 4 * XG_TILE3_CHUNKS) * ceil_div(hi[1] - lo[1], 4) *
         ceil_div(hi[2] - (lo[2] & ~(cols3 - 1)), cols3);
}

int64_t blocks_for_dom(int D, const xg_dom& dom, int cols3) {
  if (dom.kind == XG_DOM_LIST) return std::min<int64_t>(ceil_div(dom.count, kThreads), 16384);
  return tiles_for(D, dom.lo, dom.hi, cols3);
}

// 64 columns per lane-row of a 3-D tile times the columns each lane holds (1, 2 or 4)
int tile3_cols(const xg_stmt_desc& d) { return d.tile3_cols == 128 || d.tile3_cols == 256 ? d.tile3_cols : 64; }

int do_launch(xg_kernel* k, int32_t sid, void* fn, int64_t blocks, int threads, void** args) {
  if (blocks <= 0) return XG_OK;
  if (blocks > 0x7fffffff) return fail(XG_ERR_ARG, "launch grid too large");
  const bool timed = (sid >= 0 && sid == k->timed_sid && k->mode == kDirect);
  EventPair ev{};
  if (timed) {
    XG_HIP(hipEventCreate(&ev.a));
    XG_HIP(hipEventCreate(&ev.b));
    XG_HIP(hipEventRecord(ev.a, g_dev.stream));
  }
  XG_HIP(hipModuleLaunchKernel((hipFunction_t)fn, (unsigned)blocks, 1, 1, (unsigned)threads, 1, 1, 0,
                               g_dev.stream, args, nullptr));
  if (timed) {
    XG_HIP(hipEventRecord(ev.b, g_dev.stream));
    k->events.push_back(ev);
  }
  k->n_launch++;
  return XG_OK;
}

int exec_stmt(xg_kernel* k, int32_t sid, const void* params) {
  StmtPlan& sp = k->stmts[(size_t)sid];
  const xg_stmt_desc& d = sp.d;
  xg_grid* G = k->grids[(size_t)d.store_arg];
  const ClassPlan& p = G->plan(d.klass);
  if (p.kind == XG_DOM_EMPTY || p.count == 0) return XG_OK;
  const int mode = effective_mode(k, sp);
  xg_launch_hdr H;
  std::memset(&H, 0, sizeof H);
  fill_dom(H.dom, p, G, d.klass);
  thread_local std::vector<xg_gview> refs;   // kernel arguments are copied at the launch
  int rc = fill_refs(k, d, refs);
  if (rc) return rc;
  void* args[3] = {&H, refs.data(), const_cast<void*>(params)};
  const int32_t uni = uniform_refs(k, d) ? XG_F_UNIFORM : 0;
  H.dom.flags = uni;
  if (mode == XG_MODE_INPLACE) {
    H.dom.flags |= stream_flag(G);  // 3-D box tiles of large grids store non-temporally
    fill_view(H.out, G, G->levels[0]);
    return do_launch(k, sid, d.fn_main, blocks_for_dom(d.dims, H.dom, tile3_cols(d)), kThreads, args);
  }
  if (!G->scratch) return fail(XG_ERR_STATE, "scratch buffer missing");
  if (mode == XG_MODE_PINGPONG) {
    if (!d.fn_pingpong) return fail(XG_ERR_STATE, "statement has no Jacobi kernel");
    if (H.dom.kind == XG_DOM_LIST) H.dom.kind = XG_DOM_MASKED;  // membership by mask
    H.dom.flags = XG_F_PINGPONG | uni;
    fill_view(H.out, G, G->scratch);
    fill_view(H.src, G, G->levels[0]);
    int32_t zero[3] = {0, 0, 0};
    int32_t full[3] = {G->shape[0], G->dims > 1 ? G->shape[1] : 1, G->dims > 2 ? G->shape[2] : 1};
    rc = do_launch(k, sid, d.fn_pingpong, tiles_for(d.dims, zero, full), kThreads, args);
    if (rc) return rc;
    std::swap(G->levels[0], G->scratch);
    return XG_OK;
  }
  // two-phase: value -> scratch[cell], then scratch[cell] -> store(cell + offset)
  if (!d.fn_phase2) return fail(XG_ERR_STATE, "statement has no phase-2 kernel");
  H.dom.flags = XG_F_TO_TMP | uni;
  fill_view(H.out, G, G->levels[0]);
  fill_view(H.src, G, G->scratch);
  rc = do_launch(k, sid, d.fn_main, blocks_for_dom(d.dims, H.dom, tile3_cols(d)), kThreads, args);
  if (rc) return rc;
  xg_launch_hdr H2 = H;
  H2.dom.flags = 0;
  for (int a = 0; a < 3; ++a) H2.dom.so[a] = d.store_off[a];
  void* args2[1] = {&H2};
  return do_launch(k, -1, d.fn_phase2, blocks_for_dom(d.dims, H2.dom), kThreads, args2);
}

// Chain statements [a, b) of group gi in one workgroup.
int exec_group(xg_kernel* k, int32_t gi, const std::vector<int>& sids, const std::vector<const void*>& params) {
  const xg_group_desc& gd = k->groups[(size_t)gi];
  static const uint8_t zeros[4096] = {0};
  // buffers kept between calls (kernel arguments are copied at the launch)
  thread_local std::vector<xg_launch_hdr> hdrs;
  thread_local std::vector<std::vector<xg_gview>> refs;
  thread_local std::vector<void*> args;
  thread_local std::vector<const void*> par;
  hdrs.resize((size_t)gd.count);
  if (refs.size() < (size_t)gd.count) refs.resize((size_t)gd.count);
  uint32_t en = 0;
  args.clear();
  args.reserve(1 + 3 * (size_t)gd.count);
  args.push_back(&en);
  par.assign((size_t)gd.count, zeros);
  for (size_t i = 0; i < sids.size(); ++i) {
    const int pos = sids[i] - gd.first;
    en |= 1u << pos;
    par[(size_t)pos] = params[i];
  }
  for (int pos = 0; pos < gd.count; ++pos) {
    const xg_stmt_desc& d = k->stmts[(size_t)(gd.first + pos)].d;
    std::memset(&hdrs[(size_t)pos], 0, sizeof(xg_launch_hdr));
    refs[(size_t)pos].assign((size_t)std::max(1, d.nrefs), xg_gview{});
    if (en & (1u << pos)) {
      xg_grid* G = k->grids[(size_t)d.store_arg];
      const ClassPlan& p = G->plan(d.klass);
      fill_dom(hdrs[(size_t)pos].dom, p, G, d.klass);
      if (p.kind == XG_DOM_EMPTY || p.count == 0) hdrs[(size_t)pos].dom.kind = XG_DOM_EMPTY;
      fill_view(hdrs[(size_t)pos].out, G, G->levels[0]);
      int rc = fill_refs(k, d, refs[(size_t)pos]);
      if (rc) return rc;
    }
    args.push_back(&hdrs[(size_t)pos]);
    args.push_back(refs[(size_t)pos].data());
    args.push_back(const_cast<void*>(par[(size_t)pos]));
  }
  k->n_fused += (int64_t)sids.size();
  return do_launch(k, -1, gd.fn, 1, kGroupThreads, args.data());
}

// Statements a and a + 1 run as one launch (the emitter's pair kernel) when both
// run in place over the same box and no argument aliases a store grid the other
// statement reads at time 0 or stores to.  In a forwarding pair the second
// statement's time-0 reads of the first's store argument (at its own cell) take
// the computed value; any other argument bound to that grid still disqualifies.
// A statement being timed runs alone.
bool pairable(xg_kernel* k, int32_t a, int32_t b) {
  if (b != a + 1 || b >= (int32_t)k->stmts.size()) return false;
  StmtPlan& A = k->stmts[(size_t)a];
  StmtPlan& B = k->stmts[(size_t)b];
  if (!A.d.fn_pair || k->timed_sid == a || k->timed_sid == b) return false;
  if (effective_mode(k, A) != XG_MODE_INPLACE || effective_mode(k, B) != XG_MODE_INPLACE) return false;
  xg_grid* GA = k->grids[(size_t)A.d.store_arg];
  xg_grid* GB = k->grids[(size_t)B.d.store_arg];
  if (GA == GB) return false;
  for (int r = 0; r < A.d.nrefs; ++r)
    if (A.d.ref_level[r] == 0 && k->grids[(size_t)A.d.ref_arg[r]] == GB) return false;
  for (int r = 0; r < B.d.nrefs; ++r)
    if (B.d.ref_level[r] == 0 && k->grids[(size_t)B.d.ref_arg[r]] == GA &&
        !(A.d.pair_fwd && B.d.ref_arg[r] == A.d.store_arg))
      return false;
  const ClassPlan& pa = GA->plan(A.d.klass);
  const ClassPlan& pb = GB->plan(B.d.klass);
  if (pa.kind != XG_DOM_BOX || pb.kind != XG_DOM_BOX || pa.count == 0) return false;
  for (int d = 0; d < 3; ++d)
    if (pa.lo[d] != pb.lo[d] || pa.hi[d] != pb.hi[d]) return false;
  return true;
}

int exec_pair(xg_kernel* k, int32_t a, const void* pa, const void* pb) {
  xg_launch_hdr H[2];
  std::vector<xg_gview> refs[2];
  for (int i = 0; i < 2; ++i) {
    const xg_stmt_desc& d = k->stmts[(size_t)(a + i)].d;
    xg_grid* G = k->grids[(size_t)d.store_arg];
    std::memset(&H[i], 0, sizeof(xg_launch_hdr));
    fill_dom(H[i].dom, G->plan(d.klass), G, d.klass);
    int rc = fill_refs(k, d, refs[i]);
    if (rc) return rc;
    H[i].dom.flags = uniform_refs(k, d) ? XG_F_UNIFORM : 0;
    fill_view(H[i].out, G, G->levels[0]);
  }
  void* args[6] = {&H[0], refs[0].data(), const_cast<void*>(pa), &H[1], refs[1].data(), const_cast<void*>(pb)};
  k->n_fused += 2;
  const xg_stmt_desc& d = k->stmts[(size_t)a].d;
  return do_launch(k, -1, d.fn_pair, blocks_for_dom(d.dims, H[0].dom), kThreads, args);
}

int flush(xg_kernel* k) {
  if (k->pending.empty()) return XG_OK;
  // the run takes the pending list's buffer and gives back its own (empty, with capacity):
  // a call allocates nothing here once the buffers have grown
  std::vector<Pending>& run = k->run_buf;
  run.clear();
  run.swap(k->pending);
  const int32_t gi = k->pending_group;
  k->pending_group = -1;
  auto params_of = [&](const Pending& p) -> const void* { return k->arena.data() + p.off; };
  int rc = XG_OK;
  if (gi < 0 || !k->fuse || run.size() < 2 || k->groups[(size_t)gi].fn == nullptr) {
    for (auto& p : run)
      if ((rc = exec_stmt(k, p.sid, params_of(p)))) break;
    k->arena.clear();
    return rc;
  }
  // eligibility: in place, small, no hazard
  thread_local std::vector<char> ok;
  ok.assign(run.size(), 0);
  for (size_t i = 0; i < run.size(); ++i) {
    StmtPlan& sp = k->stmts[(size_t)run[i].sid];
    xg_grid* G = k->grids[(size_t)sp.d.store_arg];
    const ClassPlan& p = G->plan(sp.d.klass);
    ok[i] = effective_mode(k, sp) == XG_MODE_INPLACE && p.count <= g_dev.fuse_max_cells;
  }
  size_t i = 0;
  while (i < run.size() && rc == XG_OK) {
    if (!ok[i]) {
      if (i + 1 < run.size() && !ok[i + 1] && pairable(k, run[i].sid, run[i + 1].sid)) {
        rc = exec_pair(k, run[i].sid, params_of(run[i]), params_of(run[i + 1]));
        i += 2;
        continue;
      }
      rc = exec_stmt(k, run[i].sid, params_of(run[i]));
      ++i;
      continue;
    }
    size_t j = i;
    while (j < run.size() && ok[j]) ++j;
    if (j - i >= 2) {
      thread_local std::vector<int> sids;
      thread_local std::vector<const void*> ps;
      sids.clear();
      ps.clear();
      for (size_t t = i; t < j; ++t) {
        sids.push_back(run[t].sid);
        ps.push_back(params_of(run[t]));
      }
      rc = exec_group(k, gi, sids, ps);
    } else {
      rc = exec_stmt(k, run[i].sid, params_of(run[i]));
    }
    i = j;
  }
  k->arena.clear();
  return rc;
}

int ensure_scratch(xg_kernel* k) {
  // scratch for every grid a statement stores to (Jacobi, two-phase, hazards)
  for (auto& sp : k->stmts) {
    xg_grid* G = k->grids[(size_t)sp.d.store_arg];
    if (!G->scratch) {
      int rc = alloc_zero(G->bytes(), &G->scratch);
      if (rc) return rc;
    }
  }
  return XG_OK;
}

bool loop_blockable(xg_kernel* k, LoopPlan& L) {
  xg_grid* G = k->grids[(size_t)L.d.grid_arg];
  // another argument bound to the loop's grid would miss the LDS copy
  for (int j = 0; j < L.d.nstmts; ++j) {
    const xg_stmt_desc& d = k->stmts[(size_t)L.d.sids[j]].d;
    for (int r = 0; r < d.nrefs; ++r)
      if (d.ref_arg[r] != L.d.grid_arg && k->grids[(size_t)d.ref_arg[r]] == G) return false;
  }
  if (L.grid != (const void*)G || L.mask_version != G->mask_version) {
    L.state = 0;
    L.validated = 0;
    L.grid = G;
    L.mask_version = G->mask_version;
  }
  return true;
}

void fill_tb_hdr(xg_kernel* k, LoopPlan& L, xg_tb_hdr& TH, int iters, int32_t* fail_flag) {
  xg_grid* G = k->grids[(size_t)L.d.grid_arg];
  std::memset(&TH, 0, sizeof TH);
  fill_view(TH.src, G, G->levels[0]);
  fill_view(TH.dst, G, G->scratch);
  TH.mask = G->h_mask.empty() ? nullptr : G->d_mask;
  TH.fail = fail_flag;
  TH.iters = iters;
  TH.uniform = 1;
  for (int j = 0; j < L.d.nstmts; ++j) {
    const xg_stmt_desc& d = k->stmts[(size_t)L.d.sids[j]].d;
    TH.klass[j] = d.klass;
    if (!uniform_refs(k, d)) TH.uniform = 0;
  }
}

// An executable graph may still be running on the stream (replays are
// asynchronous): wait for the stream before destroying it, so its kernel
// arguments are not freed under a launch in flight.
void destroy_entry(GraphEntry& e) {
  if (e.exec) {
    if (g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
    (void)hipGraphExecDestroy(e.exec);
  }
  e.exec = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// stub API: begin / stmt / end
// ---------------------------------------------------------------------------
namespace {

int api_begin(xg_kernel* k, xg_grid* const* grids, int32_t ngrids, const void* key, int64_t key_bytes) {
  k->status = XG_OK;
  k->status_msg.clear();
  k->mode = kDirect;
  k->cur = nullptr;
  k->pending.clear();
  k->arena.clear();
  k->pending_group = -1;
  if (int rc = check_ready()) return set_status(k, rc);
  if (ngrids != k->ngrids) return set_status(k, fail(XG_ERR_ARG, "grid argument count mismatch"));
  k->grids.assign(grids, grids + ngrids);
  for (auto* G : k->grids) {
    if (!G) return set_status(k, fail(XG_ERR_ARG, "null grid handle"));
    G->bound = true;
  }
  if (int rc = ensure_scratch(k)) return set_status(k, rc);
  if (k->record_next) {
    k->mode = kRecord;
    k->rec.clear();
    k->rec_arena.clear();
    return XG_OK;
  }
  if (!k->graphs || k->graphs_auto_off) return XG_OK;
  // key: scalar arguments + state of every bound grid
  auto& kb = k->key_buf;
  kb.assign((const uint8_t*)key, (const uint8_t*)key + key_bytes);
  for (auto* G : k->grids) {
    auto put = [&](const void* p, size_t n) { kb.insert(kb.end(), (const uint8_t*)p, (const uint8_t*)p + n); };
    put(&G, sizeof G);
    for (void* l : G->levels) put(&l, sizeof l);
    put(&G->scratch, sizeof G->scratch);
    put(&G->mask_version, sizeof G->mask_version);
  }
  const uint64_t h = fnv1a(kb.data(), kb.size());
  ++k->tick;
  for (auto& e : k->cache) {
    if (e.hash == h && e.key == kb) {
      e.last_use = k->tick;
      e.seen++;
      k->cur = &e;
      if (e.direct) {
        k->cur = nullptr;
        return XG_OK;
      }
      if (e.exec) {
        k->mode = kReplay;
        k->n_replay++;
        k->graph_hits++;
        hipError_t err = hipGraphLaunch(e.exec, g_dev.stream);
        if (err != hipSuccess) return set_status(k, fail(XG_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(err)));
        return XG_OK;
      }
      k->mode = kCapture;
      k->capture_mark = k->n_launch;
      hipError_t err = hipStreamBeginCapture(g_dev.stream, hipStreamCaptureModeThreadLocal);
      if (err != hipSuccess) {
        k->mode = kDirect;
        k->cur = nullptr;
        return set_status(k, fail(XG_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(err)));
      }
      return XG_OK;
    }
  }
  // first sighting: run directly, remember the key
  k->graph_misses++;
  if (k->graph_hits == 0 && k->graph_misses > 64) {
    k->graphs_auto_off = true;  // every call differs (e.g. a time argument): graphs cannot pay
    for (auto& e : k->cache) destroy_entry(e);
    k->cache.clear();
    return XG_OK;
  }
  if (k->cache.size() >= 16) {
    auto lru = std::min_element(k->cache.begin(), k->cache.end(),
                                [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
    destroy_entry(*lru);
    k->cache.erase(lru);
  }
  GraphEntry e;
  e.key = kb;
  e.hash = h;
  e.seen = 1;
  e.last_use = k->tick;
  k->cache.push_back(std::move(e));
  return XG_OK;
}

int api_stmt(xg_kernel* k, int32_t sid, const void* params, int64_t bytes) {
  if (k->status != XG_OK) return k->status;
  if (sid < 0 || sid >= (int32_t)k->stmts.size()) return set_status(k, fail(XG_ERR_ARG, "bad statement id"));
  k->last_params[(size_t)sid].assign((const uint8_t*)params, (const uint8_t*)params + bytes);
  if (k->mode == kReplay) {
    k->exec_count[(size_t)sid]++;
    return XG_OK;
  }
  const xg_stmt_desc& d = k->stmts[(size_t)sid].d;
  if (bytes != d.param_bytes) return set_status(k, fail(XG_ERR_ARG, "parameter block size mismatch"));
  if (k->mode == kRecord) {
    k->exec_count[(size_t)sid]++;
    Rec r{0, sid, 0, k->rec_arena.size()};
    k->rec_arena.resize(r.off + (size_t)((bytes + 7) & ~7ll), 0);
    std::memcpy(k->rec_arena.data() + r.off, params, (size_t)bytes);
    k->rec.push_back(r);
    return XG_OK;
  }
  k->exec_count[(size_t)sid]++;
  const int32_t gi = d.group;
  const bool extends = !k->pending.empty() && gi >= 0 && gi == k->pending_group &&
                       k->pending.back().sid + 1 == sid;
  if (!extends) {
    int rc = flush(k);
    if (rc) return set_status(k, rc);
  }
  Pending p;
  p.sid = sid;
  p.off = k->arena.size();
  p.bytes = bytes;
  // keep 8-byte alignment of every parameter block
  k->arena.resize(p.off + (size_t)((bytes + 7) & ~7ll), 0);
  std::memcpy(k->arena.data() + p.off, params, (size_t)bytes);
  k->pending.push_back(p);
  k->pending_group = gi;
  if (gi < 0) {
    int rc = flush(k);
    if (rc) return set_status(k, rc);
  }
  return XG_OK;
}

int api_end(xg_kernel* k) {
  int rc = XG_OK;
  if (k->mode == kRecord) {
    k->mode = kDirect;
    k->record_next = false;
    return k->status;
  }
  if (k->mode != kReplay && k->status == XG_OK) {
    rc = flush(k);
    if (rc) set_status(k, rc);
  }
  if (k->mode == kCapture) {
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g_dev.stream, &graph);
    GraphEntry* ent = k->cur;
    if (e != hipSuccess || k->status != XG_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      if (e != hipSuccess) set_status(k, fail(XG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e)));
      k->mode = kDirect;
      return k->status;
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      k->mode = kDirect;
      return set_status(k, fail(XG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)));
    }
    ent->exec = exec;
    ent->post_levels.clear();
    ent->post_scratch.clear();
    for (auto* G : k->grids) {
      ent->post_levels.push_back(G->levels);
      ent->post_scratch.push_back(G->scratch);
    }
    k->n_capture++;
    e = hipGraphLaunch(exec, g_dev.stream);
    k->mode = kDirect;
    if (e != hipSuccess) return set_status(k, fail(XG_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e)));
    if (k->n_launch - k->capture_mark < kGraphMinLaunches) {
      // launching this graph costs more host time than launching its few kernels
      (void)hipStreamSynchronize(g_dev.stream);
      (void)hipGraphExecDestroy(exec);
      ent->exec = nullptr;
      ent->direct = true;
    }
    return k->status;
  }
  if (k->mode == kReplay) {
    GraphEntry* ent = k->cur;
    for (size_t i = 0; i < k->grids.size(); ++i) {
      k->grids[i]->levels = ent->post_levels[i];
      k->grids[i]->scratch = ent->post_scratch[i];
    }
    k->mode = kDirect;
  }
  return k->status;
}

const int32_t* api_shape(xg_grid* g) { return g ? g->shape : nullptr; }

// One launch of a blocked loop: `iters` iterations from level 0 into the scratch.
int launch_blocked(xg_kernel* k, LoopPlan& L, xg_grid* G, const void* const* params, int iters, bool check) {
  xg_tb_hdr TH;
  fill_tb_hdr(k, L, TH, iters, k->d_fail);
  const int m = L.d.nstmts;
  std::vector<std::vector<xg_gview>> refs((size_t)m);
  std::vector<void*> args;
  args.reserve(1 + 2 * (size_t)m);
  args.push_back(&TH);
  for (int j = 0; j < m; ++j) {
    int rc = fill_refs(k, k->stmts[(size_t)L.d.sids[j]].d, refs[(size_t)j]);
    if (rc) return rc;
  }
  for (int j = 0; j < m; ++j) {
    args.push_back(refs[(size_t)j].data());
    args.push_back(const_cast<void*>(params[j]));
  }
  const int64_t blocks = ceil_div(G->shape[0], L.d.core0) * ceil_div(G->shape[1], L.d.core1);
  L.launches++;
  return do_launch(k, -1, check ? L.d.fn_check : L.d.fn, blocks, L.d.threads, args.data());
}

int api_loop(xg_kernel* k, int32_t lid, int64_t trip, const void* const* params, int32_t n) {
  if (k->status != XG_OK) return 1;  // the call already failed: skip the loop
  if (lid < 0 || lid >= (int32_t)k->loops.size() || n != k->loops[(size_t)lid].d.nstmts) {
    set_status(k, fail(XG_ERR_ARG, "bad loop description"));
    return 1;
  }
  if (k->mode != kRecord) {  // plain calls on small grids take the loop's other tile shape
    const xg_loop_desc& d = k->loops[(size_t)lid].d;
    if (d.alt >= 0 && k->grids[(size_t)d.grid_arg]->n <= d.alt_max_cells) lid = d.alt;
  }
  LoopPlan& L = k->loops[(size_t)lid];
  if (k->mode == kRecord) {
    Rec r{1, lid, trip, k->rec_arena.size()};
    for (int j = 0; j < n; ++j) {
      const int32_t sid = L.d.sids[j];
      const int64_t b = k->stmts[(size_t)sid].d.param_bytes;
      const size_t at = k->rec_arena.size();
      k->rec_arena.resize(at + (size_t)((b + 7) & ~7ll), 0);
      std::memcpy(k->rec_arena.data() + at, params[j], (size_t)b);
      k->exec_count[(size_t)sid] += trip;
    }
    k->rec.push_back(r);
    return 1;
  }
  auto account = [&]() {
    for (int j = 0; j < n; ++j) {
      const int32_t sid = L.d.sids[j];
      k->exec_count[(size_t)sid] += trip;
      const int64_t b = k->stmts[(size_t)sid].d.param_bytes;
      k->last_params[(size_t)sid].assign((const uint8_t*)params[j], (const uint8_t*)params[j] + b);
    }
    L.last_params.resize((size_t)n);
    for (int j = 0; j < n; ++j) {
      const int64_t b = k->stmts[(size_t)L.d.sids[j]].d.param_bytes;
      L.last_params[(size_t)j].assign((const uint8_t*)params[j], (const uint8_t*)params[j] + b);
    }
  };
  if (k->mode == kReplay) {  // the graph holds whatever ran when it was captured
    account();
    return 1;
  }
  if (!k->blocked || trip <= 0 || !L.d.fn) return 0;
  xg_grid* G = k->grids[(size_t)L.d.grid_arg];
  if (!loop_blockable(k, L)) return 0;
  if (L.state < 0) return 0;
  // A check proves launches of up to the iterations it ran: the set of exact region
  // cells only shrinks from one iteration to the next, so a proof at T covers fewer
  // iterations, but a launch of more iterations than any check ran needs its own.
  const int T0 = (int)std::min<int64_t>(L.d.tmax, trip);
  const bool check = L.state == 0 || T0 > L.validated;
  if (check && k->mode == kCapture) return 0;  // validation needs a synchronisation
  int rc = flush(k);
  if (rc) {
    set_status(k, rc);
    return 1;
  }
  int64_t done = 0;
  if (check) {
    if (!k->d_fail) {
      void* p = nullptr;
      if (int e = dev_alloc(sizeof(int32_t), true, &p)) {
        set_status(k, e);
        return 1;
      }
      k->d_fail = (int32_t*)p;
    }
    int32_t h_fail = 0;
    const int T = T0;
    if (hipMemsetAsync(k->d_fail, 0, sizeof(int32_t), g_dev.stream) != hipSuccess ||
        (rc = launch_blocked(k, L, G, params, T, true)) != XG_OK ||
        hipMemcpyAsync(&h_fail, k->d_fail, sizeof(int32_t), hipMemcpyDeviceToHost, g_dev.stream) != hipSuccess ||
        hipStreamSynchronize(g_dev.stream) != hipSuccess) {
      set_status(k, rc ? rc : fail(XG_ERR_HIP, "blocked-loop validation failed"));
      return 1;
    }
    if (h_fail) {  // some core cell depends on data outside its tile's halo: the scratch is discarded
      L.state = -1;
      return 0;
    }
    L.state = 1;
    L.validated = std::max(L.validated, T);
    std::swap(G->levels[0], G->scratch);
    done = T;
    L.iterations += T;
  }
  while (done < trip) {
    const int T = (int)std::min<int64_t>(L.d.tmax, trip - done);
    rc = launch_blocked(k, L, G, params, T, false);
    if (rc) {
      set_status(k, rc);
      return 1;
    }
    std::swap(G->levels[0], G->scratch);
    done += T;
    L.iterations += T;
  }
  account();
  return 1;
}


// copy probe kernel (calibration of HBM bandwidth and of byte counters)
__global__ void __launch_bounds__(256) xg_copy_kernel(const double4* __restrict__ in, double4* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

}  // namespace

#if !defined(__HIP_DEVICE_COMPILE__)   // host function pointers: no device copy
namespace xgi {
const xg_api g_api = {api_begin, api_stmt, api_end, api_shape, api_loop};
}  // namespace xgi
#endif

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* xg_last_error(void) { return t_err.c_str(); }

int xg_init(int device) {
  if (g_dev.ready) {
    if (device == g_dev.id) return XG_OK;
    return fail(XG_ERR_STATE, "backend already initialised on another device");
  }
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(XG_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(XG_ERR_ARG, "device index out of range");
  XG_HIP(hipSetDevice(device));
  if (int rc = reduce_setup(device)) return rc;
  hipDeviceProp_t prop;
  XG_HIP(hipGetDeviceProperties(&prop, device));
  XG_HIP(hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking));
  XG_HIP(hipStreamCreateWithFlags(&g_dev.copy, hipStreamNonBlocking));
  hipMemPoolProps props;
  std::memset(&props, 0, sizeof props);
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device;
  XG_HIP(hipMemPoolCreate(&g_dev.pool, &props));
  uint64_t keep = kPoolKeepBytes;   // freed memory the pool keeps for reuse instead of returning it
  XG_HIP(hipMemPoolSetAttribute(g_dev.pool, hipMemPoolAttrReleaseThreshold, &keep));
  int no_waits = 0;
  XG_HIP(hipMemPoolSetAttribute(g_dev.pool, hipMemPoolReuseAllowInternalDependencies, &no_waits));
  g_dev.id = device;
  g_dev.cus = prop.multiProcessorCount;
  g_dev.hbm = (int64_t)prop.totalGlobalMem;
  std::snprintf(g_dev.name, sizeof g_dev.name, "%s (%s)", prop.name, prop.gcnArchName);
  g_dev.fuse_max_cells = kFuseMaxCells;
  if (const char* v = std::getenv("XGRID_FUSE_MAX_CELLS")) g_dev.fuse_max_cells = std::atoll(v);
  g_dev.ready = true;
  return XG_OK;
}

int xg_initialized(void) { return g_dev.ready ? 1 : 0; }

int xg_device_name(char* buf, int64_t n, int* cus, int64_t* hbm) {
  if (int rc = check_ready()) return rc;
  if (buf && n > 0) std::snprintf(buf, (size_t)n, "%s", g_dev.name);
  if (cus) *cus = g_dev.cus;
  if (hbm) *hbm = g_dev.hbm;
  return XG_OK;
}

int xg_synchronize(void) {
  if (int rc = check_ready()) return rc;
  XG_HIP(hipStreamSynchronize(g_dev.stream));
  return XG_OK;
}

int xg_mem_info(int64_t* free_bytes, int64_t* total_bytes) {
  if (int rc = check_ready()) return rc;
  size_t f = 0, t = 0;
  XG_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return XG_OK;
}

int xg_rtc_compile(const char* src, const char* name, const char* const* opts, int32_t nopts,
                   const char* const* hdr_names, const char* const* hdr_srcs, int32_t nhdr,
                   const char* out_path) {
  if (!src || !out_path) return fail(XG_ERR_ARG, "null source or output path");
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src, name ? name : "xgrid.hip", nhdr, hdr_srcs, hdr_names);
  if (r != HIPRTC_SUCCESS) return fail(XG_ERR_RTC, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
  r = hiprtcCompileProgram(prog, nopts, opts);
  size_t log_n = 0;
  hiprtcGetProgramLogSize(prog, &log_n);
  std::string log(log_n + 1, '\0');
  if (log_n) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(XG_ERR_RTC, std::string("hipRTC compilation failed:\n") + log.c_str());
  }
  size_t code_n = 0;
  hiprtcGetCodeSize(prog, &code_n);
  std::vector<char> code(code_n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  // a name of this process and moment: processes sharing a cache may compile the
  // same kernel at once, and each renames its own complete file into place
  std::string tmp = std::string(out_path) + ".tmp" + std::to_string((long long)getpid()) + "." +
                    std::to_string((long long)std::chrono::steady_clock::now().time_since_epoch().count());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return fail(XG_ERR_IO, std::string("cannot write ") + tmp);
  const size_t w = std::fwrite(code.data(), 1, code.size(), f);
  std::fclose(f);
  if (w != code.size() || std::rename(tmp.c_str(), out_path) != 0) {
    std::remove(tmp.c_str());
    return fail(XG_ERR_IO, std::string("cannot write ") + out_path);
  }
  return XG_OK;
}

struct xg_module {
  hipModule_t mod;
};

int xg_module_load(const char* path, xg_module** out) {
  if (int rc = check_ready()) return rc;
  if (!path || !out) return fail(XG_ERR_ARG, "null argument");
  hipModule_t m;
  XG_HIP(hipModuleLoad(&m, path));
  *out = new xg_module{m};
  return XG_OK;
}

int xg_module_function(xg_module* m, const char* name, void** out_fn) {
  if (!m || !name || !out_fn) return fail(XG_ERR_ARG, "null argument");
  hipFunction_t f;
  hipError_t e = hipModuleGetFunction(&f, m->mod, name);
  if (e != hipSuccess) return fail(XG_ERR_ARG, std::string("kernel '") + name + "' not found: " + hipGetErrorString(e));
  *out_fn = (void*)f;
  return XG_OK;
}

int xg_module_unload(xg_module* m) {
  if (!m) return XG_OK;
  if (g_dev.ready) XG_HIP(hipStreamSynchronize(g_dev.stream));  // its kernels may still run
  XG_HIP(hipModuleUnload(m->mod));
  delete m;
  return XG_OK;
}

int xg_grid_create(int32_t dims, const int32_t* shape, int32_t elem_bytes, xg_grid** out) {
  if (int rc = check_ready()) return rc;
  if (dims < 1 || dims > 3 || !shape || !out || elem_bytes <= 0) return fail(XG_ERR_ARG, "invalid grid description");
  auto* g = new xg_grid;
  g->dims = dims;
  g->n = 1;
  for (int d = 0; d < dims; ++d) {
    if (shape[d] <= 0) {
      delete g;
      return fail(XG_ERR_ARG, "grid extents must be positive");
    }
    g->shape[d] = shape[d];
    g->n *= shape[d];
  }
  g->elem = elem_bytes;
  default_plan(g);
  *out = g;
  return XG_OK;
}

int xg_grid_destroy(xg_grid* g) {
  if (!g) return XG_OK;
  // the memory is released behind the kernels queued on the kernel stream; work that
  // a caller queued elsewhere on a handed-out level pointer is waited for first
  if (g->exported && g_dev.ready) (void)hipDeviceSynchronize();
  for (void* l : g->levels) dev_free(l);
  dev_free(g->scratch);
  dev_free(g->d_mask);
  free_plans(g);
  delete g;
  return XG_OK;
}

int xg_grid_set_depth(xg_grid* g, int32_t depth) {
  if (int rc = check_ready()) return rc;
  if (!g || depth < 0) return fail(XG_ERR_ARG, "invalid grid or depth");
  while ((int32_t)g->levels.size() < depth) {
    void* p = nullptr;
    int rc = alloc_zero(g->bytes(), &p);
    if (rc) return rc;
    g->levels.push_back(p);
  }
  if ((int32_t)g->levels.size() > depth) {
    if (g->exported) XG_HIP(hipDeviceSynchronize());
    while ((int32_t)g->levels.size() > depth) {
      dev_free(g->levels.back());
      g->levels.pop_back();
    }
  }
  return XG_OK;
}

int32_t xg_grid_depth(xg_grid* g) { return g ? (int32_t)g->levels.size() : -1; }

int xg_grid_tick(xg_grid* g) {
  if (!g) return fail(XG_ERR_ARG, "null grid");
  if (g->levels.size() > 1) std::rotate(g->levels.rbegin(), g->levels.rbegin() + 1, g->levels.rend());
  return XG_OK;
}

int xg_grid_upload(xg_grid* g, int32_t level, const void* host) {
  if (int rc = check_ready()) return rc;
  if (!g || !host || level < 0 || level >= (int32_t)g->levels.size()) return fail(XG_ERR_ARG, "invalid level");
  // A grid no kernel has used yet is written beside the queued kernels (a request
  // being admitted while others run); otherwise behind them, since they may read it.
  hipStream_t s = g->bound ? g_dev.stream : g_dev.copy;
  XG_HIP(hipMemcpyAsync(g->levels[(size_t)level], host, g->bytes(), hipMemcpyHostToDevice, s));
  XG_HIP(hipStreamSynchronize(s));
  return XG_OK;
}

int xg_grid_download(xg_grid* g, int32_t level, void* host) {
  if (int rc = check_ready()) return rc;
  if (!g || !host || level < 0 || level >= (int32_t)g->levels.size()) return fail(XG_ERR_ARG, "invalid level");
  XG_HIP(hipMemcpyAsync(host, g->levels[(size_t)level], g->bytes(), hipMemcpyDeviceToHost, g_dev.stream));
  XG_HIP(hipStreamSynchronize(g_dev.stream));
  return XG_OK;
}

int xg_grid_download_after(xg_grid* g, int32_t level, void* host, void* event) {
  if (int rc = check_ready()) return rc;
  if (!g || !host || !event || level < 0 || level >= (int32_t)g->levels.size())
    return fail(XG_ERR_ARG, "invalid level or event");
  // on the copy stream: the kernels queued after the event keep the device busy
  XG_HIP(hipStreamWaitEvent(g_dev.copy, (hipEvent_t)event, 0));
  XG_HIP(hipMemcpyAsync(host, g->levels[(size_t)level], g->bytes(), hipMemcpyDeviceToHost, g_dev.copy));
  XG_HIP(hipStreamSynchronize(g_dev.copy));
  return XG_OK;
}

void* xg_grid_level_ptr(xg_grid* g, int32_t level) {
  if (!g || level < 0 || level >= (int32_t)g->levels.size()) return nullptr;
  g->bound = g->exported = true;
  return g->levels[(size_t)level];
}

int xg_grid_set_mask(xg_grid* g, const int32_t* mask) {
  if (int rc = check_ready()) return rc;
  if (!g || !mask) return fail(XG_ERR_ARG, "null grid or mask");
  const int D = g->dims;
  const int32_t s0 = g->shape[0], s1 = D > 1 ? g->shape[1] : 1, s2 = D > 2 ? g->shape[2] : 1;
  struct Stat {
    int64_t count = 0;
    int32_t lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int32_t hi[3] = {-1, -1, -1};
  };
  std::unordered_map<int32_t, Stat> stats;
  int64_t c = 0;
  int32_t last_v = mask[0];
  Stat* last = &stats[last_v];
  for (int32_t i0 = 0; i0 < s0; ++i0)
    for (int32_t i1 = 0; i1 < s1; ++i1)
      for (int32_t i2 = 0; i2 < s2; ++i2, ++c) {
        const int32_t v = mask[c];
        if (v != last_v) {
          last_v = v;
          last = &stats[v];
        }
        Stat& s = *last;
        s.count++;
        const int32_t idx[3] = {i0, i1, i2};
        for (int d = 0; d < 3; ++d) {
          if (idx[d] < s.lo[d]) s.lo[d] = idx[d];
          if (idx[d] > s.hi[d]) s.hi[d] = idx[d];
        }
      }
  // New lists and mask go to new memory (written on the copy stream); the old ones are
  // released behind the queued kernels that may still read them.
  if (g->exported) XG_HIP(hipDeviceSynchronize());
  free_plans(g);
  std::unordered_map<int32_t, std::vector<int32_t>> lists;
  // leaving early (an allocation or copy failed), copies already queued from `lists`
  // finish before it is freed (the guard goes out of scope first)
  struct CopiesDone {
    ~CopiesDone() { (void)hipStreamSynchronize(g_dev.copy); }
  } copies_done;
  for (auto& kv : stats) {
    ClassPlan p;
    p.count = kv.second.count;
    int64_t vol = 1;
    for (int d = 0; d < 3; ++d) {
      p.lo[d] = kv.second.lo[d];
      p.hi[d] = kv.second.hi[d] + 1;
      vol *= (p.hi[d] - p.lo[d]);
    }
    // A list holds int32 linear indices: a grid of more than INT32_MAX cells plans its
    // sparse classes as masked boxes instead (the device unpacks list entries in 32 bits).
    if (p.count == vol) p.kind = XG_DOM_BOX;
    else if (p.count * 4 >= vol || g->n > (int64_t)INT32_MAX) p.kind = XG_DOM_MASKED;
    else {
      p.kind = XG_DOM_LIST;
      lists[kv.first].reserve((size_t)p.count);
    }
    g->plans[kv.first] = p;
  }
  if (!lists.empty()) {
    for (int64_t i = 0; i < g->n; ++i) {
      auto it = lists.find(mask[i]);
      if (it != lists.end()) it->second.push_back((int32_t)i);
    }
    for (auto& kv : lists) {
      ClassPlan& p = g->plans[kv.first];
      void* d = nullptr;
      if (int rc = dev_alloc(kv.second.size() * sizeof(int32_t), false, &d)) return rc;
      p.d_list = (int32_t*)d;
      XG_HIP(hipMemcpyAsync(d, kv.second.data(), kv.second.size() * sizeof(int32_t), hipMemcpyHostToDevice, g_dev.copy));
    }
  }
  dev_free(g->d_mask);
  g->d_mask = nullptr;
  {
    void* d = nullptr;
    if (int rc = dev_alloc((size_t)g->n * sizeof(int32_t), false, &d)) return rc;
    g->d_mask = (int32_t*)d;
  }
  XG_HIP(hipMemcpyAsync(g->d_mask, mask, (size_t)g->n * sizeof(int32_t), hipMemcpyHostToDevice, g_dev.copy));
  XG_HIP(hipStreamSynchronize(g_dev.copy));   // the host lists and the caller's mask are read by then
  g->h_mask.assign(mask, mask + g->n);
  g->mask_version = g_mask_version++;
  return XG_OK;
}

int xg_grid_class_info(xg_grid* g, int32_t klass, int32_t* kind, int64_t* count, int32_t* lo, int32_t* hi) {
  if (!g) return fail(XG_ERR_ARG, "null grid");
  const ClassPlan& p = g->plan(klass);
  if (kind) *kind = p.kind;
  if (count) *count = p.count;
  for (int d = 0; d < 3; ++d) {
    if (lo) lo[d] = p.lo[d];
    if (hi) hi[d] = p.hi[d];
  }
  return XG_OK;
}

int xg_kernel_create(int32_t nstmts, const xg_stmt_desc* stmts, int32_t ngroups, const xg_group_desc* groups,
                     int32_t nloops, const xg_loop_desc* loops, int32_t ngrids, xg_kernel** out) {
  if (nstmts < 0 || ngroups < 0 || nloops < 0 || ngrids < 0 || !out) return fail(XG_ERR_ARG, "invalid kernel description");
  auto* k = new xg_kernel;
  k->ngrids = ngrids;
  for (int i = 0; i < nstmts; ++i) {
    const xg_stmt_desc& d = stmts[i];
    if (d.store_arg < 0 || d.store_arg >= ngrids || d.nrefs < 0 || d.nrefs > XG_DESC_MAX_REFS ||
        d.nhazard < 0 || d.nhazard > XG_DESC_MAX_HAZ || d.dims < 1 || d.dims > 3 || !d.fn_main) {
      delete k;
      return fail(XG_ERR_ARG, "invalid statement description " + std::to_string(i));
    }
    for (int r = 0; r < d.nrefs; ++r)
      if (d.ref_arg[r] < 0 || d.ref_arg[r] >= ngrids) {
        delete k;
        return fail(XG_ERR_ARG, "invalid statement reference " + std::to_string(i));
      }
    StmtPlan sp;
    sp.d = d;
    k->stmts.push_back(sp);
  }
  for (int i = 0; i < ngroups; ++i) {
    if (groups[i].first < 0 || groups[i].count < 2 || groups[i].count > 32 || groups[i].first + groups[i].count > nstmts) {
      delete k;
      return fail(XG_ERR_ARG, "invalid group description");
    }
    k->groups.push_back(groups[i]);
  }
  for (int i = 0; i < nloops; ++i) {
    const xg_loop_desc& d = loops[i];
    bool ok = d.nstmts >= 1 && d.nstmts <= XG_LOOP_MAX_STMTS && d.grid_arg >= 0 && d.grid_arg < ngrids &&
              d.core0 > 0 && d.core1 > 0 && d.tmax >= 1 && d.halo >= 0 && d.threads >= 64 &&
              d.threads <= 1024 && d.threads % 64 == 0;
    for (int j = 0; ok && j < d.nstmts; ++j)
      ok = d.sids[j] >= 0 && d.sids[j] < nstmts && k->stmts[(size_t)d.sids[j]].d.store_arg == d.grid_arg &&
           k->stmts[(size_t)d.sids[j]].d.dims == 2;
    if (ok && d.alt >= 0) {  // the alternative runs the same statements and has no alternative itself
      ok = d.alt < nloops && d.alt != i && loops[d.alt].alt < 0 && loops[d.alt].nstmts == d.nstmts &&
           loops[d.alt].grid_arg == d.grid_arg;
      for (int j = 0; ok && j < d.nstmts; ++j) ok = loops[d.alt].sids[j] == d.sids[j];
    }
    if (!ok) {
      delete k;
      return fail(XG_ERR_ARG, "invalid loop description " + std::to_string(i));
    }
    LoopPlan L;
    L.d = d;
    k->loops.push_back(L);
  }
  k->exec_count.assign((size_t)nstmts, 0);
  k->last_params.assign((size_t)nstmts, std::vector<uint8_t>());
  *out = k;
  return XG_OK;
}

int xg_kernel_destroy(xg_kernel* k) {
  if (!k) return XG_OK;
  for (auto& e : k->cache) destroy_entry(e);
  for (auto& ev : k->events) {
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  dev_free(k->d_fail);
  delete k;
  return XG_OK;
}

int xg_kernel_status(xg_kernel* k) {
  if (!k) return fail(XG_ERR_ARG, "null kernel");
  if (k->status != XG_OK) t_err = k->status_msg;
  return k->status;
}

int xg_kernel_set_options(xg_kernel* k, int32_t graphs, int32_t fuse, int32_t blocked) {
  if (!k) return fail(XG_ERR_ARG, "null kernel");
  k->graphs = graphs != 0;
  k->fuse = fuse != 0;
  k->blocked = blocked != 0;
  // captured graphs embody the old options; setting the options also ends an
  // automatic switch-off (the caller asked for graphs again)
  for (auto& e : k->cache) destroy_entry(e);
  k->cache.clear();
  k->graphs_auto_off = false;
  k->graph_hits = 0;
  k->graph_misses = 0;
  return XG_OK;
}

int xg_kernel_loop_info(xg_kernel* k, int32_t lid, int64_t* launches, int64_t* iterations, int32_t* state) {
  if (!k || lid < 0 || lid >= (int32_t)k->loops.size()) return fail(XG_ERR_ARG, "invalid loop");
  const LoopPlan& L = k->loops[(size_t)lid];
  if (launches) *launches = L.launches;
  if (iterations) *iterations = L.iterations;
  if (state) *state = L.state;
  return XG_OK;
}

int xg_kernel_time_loop(xg_kernel* k, int32_t lid, int32_t trip, int32_t reps, double* ms_per_launch) {
  if (int rc = check_ready()) return rc;
  if (!k || lid < 0 || lid >= (int32_t)k->loops.size() || reps < 1 || trip < 1) return fail(XG_ERR_ARG, "invalid arguments");
  LoopPlan& L = k->loops[(size_t)lid];
  if (L.state != 1 || L.last_params.size() != (size_t)L.d.nstmts || k->grids.empty())
    return fail(XG_ERR_STATE, "the loop has not run blocked yet");
  if (k->mode != kDirect) return fail(XG_ERR_STATE, "a call is in progress");
  xg_grid* G = k->grids[(size_t)L.d.grid_arg];
  std::vector<const void*> ps;
  for (auto& p : L.last_params) ps.push_back(p.data());
  // one trip = the launches a call makes for `trip` iterations (chunks of tmax)
  auto run_trip = [&]() -> int {
    for (int done = 0; done < trip;) {
      const int T = std::min(L.d.tmax, trip - done);
      int rc = launch_blocked(k, L, G, ps.data(), T, false);
      if (rc) return rc;
      std::swap(G->levels[0], G->scratch);
      done += T;
    }
    return XG_OK;
  };
  const int64_t per_trip = ceil_div(trip, L.d.tmax);
  hipEvent_t e0, e1;
  XG_HIP(hipEventCreate(&e0));
  XG_HIP(hipEventCreate(&e1));
  int rc = run_trip();
  if (!rc) {
    XG_HIP(hipEventRecord(e0, g_dev.stream));
    for (int i = 0; i < reps && !rc; ++i) rc = run_trip();
    XG_HIP(hipEventRecord(e1, g_dev.stream));
    XG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    XG_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_launch) *ms_per_launch = (double)ms / ((double)reps * (double)per_trip);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int xg_kernel_counters(xg_kernel* k, int64_t* launches, int64_t* fused, int64_t* replays, int64_t* captures) {
  if (!k) return fail(XG_ERR_ARG, "null kernel");
  if (launches) *launches = k->n_launch;
  if (fused) *fused = k->n_fused;
  if (replays) *replays = k->n_replay;
  if (captures) *captures = k->n_capture;
  return XG_OK;
}

int xg_kernel_time_stmt(xg_kernel* k, int32_t sid, int32_t enable) {
  if (!k) return fail(XG_ERR_ARG, "null kernel");
  k->timed_sid = enable ? sid : -1;
  return XG_OK;
}

int xg_kernel_stmt_time(xg_kernel* k, int32_t sid, double* total_ms, int64_t* launches) {
  if (!k) return fail(XG_ERR_ARG, "null kernel");
  (void)sid;
  XG_HIP(hipStreamSynchronize(g_dev.stream));
  double tot = 0.0;
  for (auto& ev : k->events) {
    float ms = 0.f;
    XG_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
    tot += ms;
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = (int64_t)k->events.size();
  k->events.clear();
  return XG_OK;
}

int xg_kernel_stmt_counts(xg_kernel* k, int64_t* counts, int32_t n, int32_t reset) {
  if (!k || (n > 0 && !counts)) return fail(XG_ERR_ARG, "null argument");
  for (int32_t i = 0; i < n && i < (int32_t)k->exec_count.size(); ++i) counts[i] = k->exec_count[(size_t)i];
  if (reset) std::fill(k->exec_count.begin(), k->exec_count.end(), 0);
  return XG_OK;
}

int xg_kernel_time_isolated(xg_kernel* k, int32_t sid, int32_t reps, double* ms_per_launch) {
  if (int rc = check_ready()) return rc;
  if (!k || sid < 0 || sid >= (int32_t)k->stmts.size() || reps < 1) return fail(XG_ERR_ARG, "invalid arguments");
  if (k->grids.empty() || k->last_params[(size_t)sid].empty())
    return fail(XG_ERR_STATE, "the statement has not run yet");
  if (k->mode != kDirect) return fail(XG_ERR_STATE, "a call is in progress");
  const std::vector<uint8_t> params = k->last_params[(size_t)sid];
  const int32_t saved = k->timed_sid;
  k->timed_sid = -1;
  hipEvent_t e0, e1;
  XG_HIP(hipEventCreate(&e0));
  XG_HIP(hipEventCreate(&e1));
  int rc = exec_stmt(k, sid, params.data());  // warm
  if (!rc) {
    XG_HIP(hipEventRecord(e0, g_dev.stream));
    for (int i = 0; i < reps && !rc; ++i) rc = exec_stmt(k, sid, params.data());
    XG_HIP(hipEventRecord(e1, g_dev.stream));
    XG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    XG_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_launch) *ms_per_launch = (double)ms / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  k->timed_sid = saved;
  return rc;
}

const xg_api* xg_api_table(void) { return &g_api; }

int xg_event_record(void** out) {
  if (int rc = check_ready()) return rc;
  if (!out) return fail(XG_ERR_ARG, "null output");
  hipEvent_t e;
  XG_HIP(hipEventCreate(&e));
  hipError_t err = hipEventRecord(e, g_dev.stream);
  if (err != hipSuccess) {
    (void)hipEventDestroy(e);
    return fail(XG_ERR_HIP, std::string("hipEventRecord: ") + hipGetErrorString(err));
  }
  *out = (void*)e;
  return XG_OK;
}

int xg_event_elapsed(void* a, void* b, double* ms) {
  if (!a || !b || !ms) return fail(XG_ERR_ARG, "null event");
  XG_HIP(hipEventSynchronize((hipEvent_t)b));
  float f = 0.f;
  XG_HIP(hipEventElapsedTime(&f, (hipEvent_t)a, (hipEvent_t)b));
  *ms = (double)f;
  return XG_OK;
}

int xg_event_destroy(void* e) {
  if (e) (void)hipEventDestroy((hipEvent_t)e);
  return XG_OK;
}

int xg_copy_probe(int64_t bytes, int32_t iters, double* gbps, double* ms_per_iter) {
  if (int rc = check_ready()) return rc;
  if (bytes < 32 || iters < 1) return fail(XG_ERR_ARG, "invalid probe size");
  const int64_t n = bytes / 32;
  void *a = nullptr, *b = nullptr;
  int rc = alloc_zero((size_t)n * 32, &a);
  if (rc) return rc;
  rc = alloc_zero((size_t)n * 32, &b);
  if (rc) {
    dev_free(a);
    return rc;
  }
  const int blocks = (int)std::min<int64_t>(ceil_div(n, 256), 256 * 64);
  hipEvent_t e0, e1;
  XG_HIP(hipEventCreate(&e0));
  XG_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(xg_copy_kernel, dim3(blocks), dim3(256), 0, g_dev.stream, (const double4*)a, (double4*)b, n);
  XG_HIP(hipEventRecord(e0, g_dev.stream));
  for (int i = 0; i < iters; ++i)
    hipLaunchKernelGGL(xg_copy_kernel, dim3(blocks), dim3(256), 0, g_dev.stream, (const double4*)a, (double4*)b, n);
  XG_HIP(hipEventRecord(e1, g_dev.stream));
  XG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  XG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  const double per = ms / iters;
  if (ms_per_iter) *ms_per_iter = per;
  if (gbps) *gbps = 2.0 * (double)n * 32.0 / (per * 1e-3) / 1e9;
  return XG_OK;
}

}  // extern "C"
