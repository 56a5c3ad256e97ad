// Batched execution: many independent instances of one compiled kernel (each
// with its own grids, of any size, and its own scalar arguments) advance one
// call per step, with one launch per statement for all of them.
//
// Each step:
//   1. every instance's grids are ticked (the kernel's time-level rotation);
//   2. with graphs on, the step's key (membership, arguments, grid states) is
//      looked up: a known key replays its hipGraph and nothing else runs;
//   3. otherwise each instance's host stub runs with its statements recorded
//      (xg_kernel in record mode) instead of launched;
//   4. instances whose recorded streams agree are walked together: for each
//      statement one batched launch covers the tiles of all of them (a device
//      table of per-instance entries, see xg_bent), consecutive small boundary
//      statements are chained per instance in one workgroup, and a blockable
//      loop runs temporally blocked for all of them (xg_btent) - after a
//      checking launch proves the plan exact for each instance's mask.
// Launches are persistent when a block cap is set: workgroups walk the global
// tile range with a stride of the grid size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "xgrid_kernel.h"

using namespace xgi;

typedef void (*xg_packed_entry)(void* k, const void* api, const void* args);

namespace {

struct Inst {
  int64_t id;
  xg_kernel* k;
  std::vector<uint8_t> args;
  std::vector<xg_grid*> ticks;   // grid arguments in order (a grid passed twice ticks twice)
};

enum LaunchKind { kTiles = 0, kGroup = 1 };

struct PendingLaunch {
  void* fn;
  int kind;
  int threads;
  size_t table_off;    // first entry
  int32_t n;           // entries (kTiles) / statements per instance (kGroup)
  int64_t total;       // tiles (kTiles) / instances (kGroup)
};

struct BatchGraph {
  std::vector<uint8_t> key;
  uint64_t hash = 0;
  int seen = 0;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> tables;
  std::vector<std::vector<void*>> post_levels;
  std::vector<void*> post_scratch;
  uint64_t last_use = 0;
};

// One statement position of a walk: the statement and each walked instance's parameters.
struct Pos {
  int32_t sid;
  std::vector<const void*> par;
};

constexpr size_t kAlign = 16;

}  // namespace

struct xg_batch {
  xg_packed_entry entry = nullptr;
  bool tick = true;
  std::vector<Inst> inst;
  int64_t next_id = 1;
  bool graphs = true;
  int64_t max_blocks = 0;             // 0: one workgroup per tile; else persistent grid of at most this many
  std::vector<uint8_t> arena;          // host image of the device table being built
  std::vector<size_t> fixups;          // offsets of table-relative pointers in the arena
  std::vector<PendingLaunch> launches;
  void* d_table = nullptr;
  size_t d_table_bytes = 0;
  // Pinned staging buffers of the launch tables, used in turn: the host may run
  // kStages calls ahead of the device before it waits for a buffer's last copy.
  static constexpr int kStages = 8;
  void* h_stage[kStages] = {};
  size_t h_stage_bytes[kStages] = {};
  hipEvent_t stage_ev[kStages] = {};
  int stage_i = 0;
  std::vector<BatchGraph> cache;
  uint64_t clock = 0;
  int mode = kDirect;
  std::vector<uint8_t> key_buf;
  int64_t n_launch = 0, n_replay = 0, n_capture = 0, n_steps = 0, n_validations = 0;
};

namespace {

size_t put(xg_batch* b, const void* p, size_t n) {
  const size_t off = (b->arena.size() + kAlign - 1) & ~(kAlign - 1);
  b->arena.resize(off + n);
  if (n) std::memcpy(b->arena.data() + off, p, n);
  return off;
}

// A pointer field at `field_off` in the arena holding the table offset `target`.
void set_rel(xg_batch* b, size_t field_off, size_t target) {
  uint64_t v = (uint64_t)target;
  std::memcpy(b->arena.data() + field_off, &v, sizeof v);
  b->fixups.push_back(field_off);
}

int64_t launch_blocks(const xg_batch* b, int64_t total) {
  if (b->max_blocks <= 0 || total <= b->max_blocks) return total;
  return std::max<int64_t>(8, b->max_blocks / 8 * 8);
}

int issue(xg_batch* b, void* dev) {
  for (const auto& L : b->launches) {
    const void* E = (const uint8_t*)dev + L.table_off;
    int32_t n = L.n;
    int64_t total = L.total;
    hipError_t e;
    if (L.kind == kTiles) {
      void* args[3] = {&E, &n, &total};
      e = hipModuleLaunchKernel((hipFunction_t)L.fn, (unsigned)launch_blocks(b, total), 1, 1, (unsigned)L.threads, 1,
                                1, 0, g_dev.stream, args, nullptr);
    } else {
      void* args[2] = {&E, &n};
      e = hipModuleLaunchKernel((hipFunction_t)L.fn, (unsigned)total, 1, 1, (unsigned)L.threads, 1, 1, 0,
                                g_dev.stream, args, nullptr);
    }
    if (e != hipSuccess) return fail(XG_ERR_HIP, std::string("batched launch: ") + hipGetErrorString(e));
    b->n_launch++;
  }
  return XG_OK;
}

void patch(xg_batch* b, void* dev) {
  for (size_t off : b->fixups) {
    uint64_t v;
    std::memcpy(&v, b->arena.data() + off, sizeof v);
    v += (uint64_t)(uintptr_t)dev;
    std::memcpy(b->arena.data() + off, &v, sizeof v);
  }
}

void reset_table(xg_batch* b) {
  b->arena.clear();
  b->fixups.clear();
  b->launches.clear();
}

// Direct mode: upload the table through a pinned staging buffer, then launch.
int flush_direct(xg_batch* b) {
  if (b->launches.empty()) {
    reset_table(b);
    return XG_OK;
  }
  const size_t bytes = b->arena.size();
  if (bytes > b->d_table_bytes) {
    XG_HIP(hipStreamSynchronize(g_dev.stream));
    if (b->d_table) XG_HIP(hipFree(b->d_table));
    b->d_table = nullptr;
    const size_t cap = std::max<size_t>(bytes * 2, 1 << 20);
    XG_HIP(hipMalloc(&b->d_table, cap));
    b->d_table_bytes = cap;
  }
  const int si = b->stage_i;
  b->stage_i = (si + 1) % xg_batch::kStages;
  if (!b->stage_ev[si]) XG_HIP(hipEventCreateWithFlags(&b->stage_ev[si], hipEventDisableTiming));
  else XG_HIP(hipEventSynchronize(b->stage_ev[si]));   // the last copy out of this buffer is done
  if (bytes > b->h_stage_bytes[si]) {
    if (b->h_stage[si]) XG_HIP(hipHostFree(b->h_stage[si]));
    b->h_stage[si] = nullptr;
    const size_t cap = std::max<size_t>(bytes * 2, 1 << 20);
    XG_HIP(hipHostMalloc(&b->h_stage[si], cap, hipHostMallocDefault));
    b->h_stage_bytes[si] = cap;
  }
  patch(b, b->d_table);
  std::memcpy(b->h_stage[si], b->arena.data(), bytes);
  XG_HIP(hipMemcpyAsync(b->d_table, b->h_stage[si], bytes, hipMemcpyHostToDevice, g_dev.stream));
  XG_HIP(hipEventRecord(b->stage_ev[si], g_dev.stream));
  int rc = issue(b, b->d_table);
  reset_table(b);
  return rc;
}

// ---- entries -------------------------------------------------------------

// Entries are laid out contiguously: build them into a side vector, then copy.
struct EntryList {
  std::vector<xg_bent> ents;
  std::vector<size_t> roff, poff;
  int64_t total = 0;
};

void add_entry(xg_batch* b, EntryList& L, xg_kernel* k, const xg_stmt_desc& d, const xg_launch_hdr& H,
               const void* params, int64_t ntiles, int* rc) {
  std::vector<xg_gview> refs;
  int r = fill_refs(k, d, refs);
  if (r) *rc = r;
  L.roff.push_back(put(b, refs.data(), refs.size() * sizeof(xg_gview)));
  L.poff.push_back(put(b, params, (size_t)std::max<int32_t>(d.param_bytes, 1)));
  xg_bent e;
  std::memset(&e, 0, sizeof e);
  e.hdr = H;
  e.tile0 = L.total;
  e.ntiles = ntiles;
  L.ents.push_back(e);
  L.total += ntiles;
}

int commit_entries(xg_batch* b, EntryList& L, void* fn, int threads) {
  if (L.ents.empty() || L.total == 0) return XG_OK;
  if (!fn) return fail(XG_ERR_STATE, "the kernel has no batched variant of a statement");
  if (L.total > 0x7fffffff) return fail(XG_ERR_ARG, "batched launch has too many tiles");
  const size_t base = put(b, L.ents.data(), L.ents.size() * sizeof(xg_bent));
  for (size_t i = 0; i < L.ents.size(); ++i) {
    set_rel(b, base + i * sizeof(xg_bent) + offsetof(xg_bent, refs), L.roff[i]);
    set_rel(b, base + i * sizeof(xg_bent) + offsetof(xg_bent, params), L.poff[i]);
  }
  b->launches.push_back(PendingLaunch{fn, kTiles, threads, base, (int32_t)L.ents.size(), L.total});
  return XG_OK;
}

// ---- one statement for a set of instances ----------------------------------

int run_stmt(xg_batch* b, int32_t sid, const std::vector<int>& idx, const std::vector<const void*>& par) {
  EntryList inplace, pp, ph1, ph2;
  int rc = XG_OK;
  const xg_stmt_desc* d0 = nullptr;
  for (size_t t = 0; t < idx.size(); ++t) {
    xg_kernel* k = b->inst[(size_t)idx[t]].k;
    StmtPlan& sp = k->stmts[(size_t)sid];
    const xg_stmt_desc& d = sp.d;
    d0 = &d;
    xg_grid* G = k->grids[(size_t)d.store_arg];
    const ClassPlan& p = G->plan(d.klass);
    if (p.kind == XG_DOM_EMPTY || p.count == 0) continue;
    const int mode = effective_mode(k, sp);
    xg_launch_hdr H;
    std::memset(&H, 0, sizeof H);
    fill_dom(H.dom, p, G, d.klass);
    const int32_t uni = uniform_refs(k, d) ? XG_F_UNIFORM : 0;
    H.dom.flags = uni;
    if (mode == XG_MODE_INPLACE) {
      fill_view(H.out, G, G->levels[0]);
      add_entry(b, inplace, k, d, H, par[t], blocks_for_dom(d.dims, H.dom, tile3_cols(d)), &rc);
    } else if (mode == XG_MODE_PINGPONG) {
      if (H.dom.kind == XG_DOM_LIST) H.dom.kind = XG_DOM_MASKED;
      H.dom.flags = XG_F_PINGPONG | uni;
      fill_view(H.out, G, G->scratch);
      fill_view(H.src, G, G->levels[0]);
      int32_t zero[3] = {0, 0, 0};
      int32_t full[3] = {G->shape[0], G->dims > 1 ? G->shape[1] : 1, G->dims > 2 ? G->shape[2] : 1};
      add_entry(b, pp, k, d, H, par[t], tiles_for(d.dims, zero, full), &rc);
      std::swap(G->levels[0], G->scratch);
    } else {
      H.dom.flags = XG_F_TO_TMP | uni;
      fill_view(H.out, G, G->levels[0]);
      fill_view(H.src, G, G->scratch);
      add_entry(b, ph1, k, d, H, par[t], blocks_for_dom(d.dims, H.dom, tile3_cols(d)), &rc);
      xg_launch_hdr H2 = H;
      H2.dom.flags = 0;
      for (int a = 0; a < 3; ++a) H2.dom.so[a] = d.store_off[a];
      add_entry(b, ph2, k, d, H2, par[t], blocks_for_dom(d.dims, H2.dom), &rc);
    }
  }
  if (rc || !d0) return rc;
  if ((rc = commit_entries(b, inplace, d0->fn_main_b, kThreads))) return rc;
  if ((rc = commit_entries(b, pp, d0->fn_pingpong_b, kThreads))) return rc;
  if ((rc = commit_entries(b, ph1, d0->fn_main_b, kThreads))) return rc;
  return commit_entries(b, ph2, d0->fn_phase2_b, kThreads);
}

// Statements [first, first + n) of positions, all of group gi.  Maximal runs of
// statements every instance may chain (in place and small) become one launch
// with a workgroup per instance; the others run one by one.
int run_group_chain(xg_batch* b, int32_t gi, const std::vector<Pos>& pos, size_t first, size_t n,
                    const std::vector<int>& idx);

int run_group(xg_batch* b, int32_t gi, const std::vector<Pos>& pos, size_t first, size_t n, const std::vector<int>& idx) {
  xg_kernel* k0 = b->inst[(size_t)idx[0]].k;
  const xg_group_desc& gd = k0->groups[(size_t)gi];
  std::vector<char> ok(n, k0->fuse && gd.fn_b != nullptr);
  for (size_t j = 0; j < n; ++j)
    for (size_t t = 0; ok[j] && t < idx.size(); ++t) {
      xg_kernel* k = b->inst[(size_t)idx[t]].k;
      StmtPlan& sp = k->stmts[(size_t)pos[first + j].sid];
      const ClassPlan& p = k->grids[(size_t)sp.d.store_arg]->plan(sp.d.klass);
      ok[j] = effective_mode(k, sp) == XG_MODE_INPLACE && p.count <= g_dev.fuse_max_cells;
    }
  size_t j = 0;
  while (j < n) {
    size_t e = j;
    while (e < n && ok[e]) ++e;
    if (e - j >= 2) {
      if (int rc = run_group_chain(b, gi, pos, first + j, e - j, idx)) return rc;
      j = e;
      continue;
    }
    if (int rc = run_stmt(b, pos[first + j].sid, idx, pos[first + j].par)) return rc;
    ++j;
  }
  return XG_OK;
}

int run_group_chain(xg_batch* b, int32_t gi, const std::vector<Pos>& pos, size_t first, size_t n,
                    const std::vector<int>& idx) {
  xg_kernel* k0 = b->inst[(size_t)idx[0]].k;
  const xg_group_desc& gd = k0->groups[(size_t)gi];
  // entries E[i * count + p]: position p of the group for instance i (empty domain: skipped)
  const int32_t count = gd.count;
  std::vector<xg_bent> ents((size_t)count * idx.size());
  std::vector<size_t> roff(ents.size(), 0), poff(ents.size(), 0);
  static const uint8_t zeros[64] = {0};
  const size_t zoff = put(b, zeros, sizeof zeros);
  int rc = XG_OK;
  for (size_t t = 0; t < idx.size(); ++t) {
    xg_kernel* k = b->inst[(size_t)idx[t]].k;
    for (int32_t q = 0; q < count; ++q) {
      xg_bent& e = ents[t * (size_t)count + (size_t)q];
      std::memset(&e, 0, sizeof e);
      e.hdr.dom.kind = XG_DOM_EMPTY;
      roff[t * (size_t)count + (size_t)q] = zoff;
      poff[t * (size_t)count + (size_t)q] = zoff;
    }
    for (size_t j = first; j < first + n; ++j) {
      const int32_t sid = pos[j].sid;
      const xg_stmt_desc& d = k->stmts[(size_t)sid].d;
      const int32_t q = sid - gd.first;
      xg_bent& e = ents[t * (size_t)count + (size_t)q];
      xg_grid* G = k->grids[(size_t)d.store_arg];
      const ClassPlan& p = G->plan(d.klass);
      fill_dom(e.hdr.dom, p, G, d.klass);
      if (p.kind == XG_DOM_EMPTY || p.count == 0) e.hdr.dom.kind = XG_DOM_EMPTY;
      fill_view(e.hdr.out, G, G->levels[0]);
      std::vector<xg_gview> refs;
      if (int r = fill_refs(k, d, refs)) rc = r;
      roff[t * (size_t)count + (size_t)q] = put(b, refs.data(), refs.size() * sizeof(xg_gview));
      poff[t * (size_t)count + (size_t)q] = put(b, pos[j].par[t], (size_t)std::max<int32_t>(d.param_bytes, 1));
    }
  }
  if (rc) return rc;
  const size_t base = put(b, ents.data(), ents.size() * sizeof(xg_bent));
  for (size_t i = 0; i < ents.size(); ++i) {
    set_rel(b, base + i * sizeof(xg_bent) + offsetof(xg_bent, refs), roff[i]);
    set_rel(b, base + i * sizeof(xg_bent) + offsetof(xg_bent, params), poff[i]);
  }
  b->launches.push_back(PendingLaunch{gd.fn_b, kGroup, kGroupThreads, base, count, (int64_t)idx.size()});
  return XG_OK;
}

int run_positions(xg_batch* b, const std::vector<Pos>& pos, const std::vector<int>& idx) {
  xg_kernel* k0 = b->inst[(size_t)idx[0]].k;
  size_t j = 0;
  while (j < pos.size()) {
    const int32_t gi = k0->stmts[(size_t)pos[j].sid].d.group;
    size_t e = j + 1;
    if (gi >= 0)
      while (e < pos.size() && k0->stmts[(size_t)pos[e].sid].d.group == gi && pos[e].sid == pos[e - 1].sid + 1) ++e;
    int rc = (gi >= 0 && e - j >= 2) ? run_group(b, gi, pos, j, e - j, idx) : run_stmt(b, pos[j].sid, idx, pos[j].par);
    if (rc) return rc;
    j = e;
  }
  return XG_OK;
}

// Loop `lid` for the instances idx, `trip` iterations, body parameters par[t][j].
int run_loop(xg_batch* b, int32_t lid, int64_t trip, const std::vector<int>& idx,
             const std::vector<std::vector<const void*>>& par) {
  std::vector<int> blocked, check, plain;            // positions in idx
  for (size_t t = 0; t < idx.size(); ++t) {
    xg_kernel* k = b->inst[(size_t)idx[t]].k;
    LoopPlan& L = k->loops[(size_t)lid];
    const bool can = k->blocked && trip > 0 && L.d.fn_b && L.d.fn_check_b && loop_blockable(k, L) && L.state >= 0;
    // a launch of more iterations than the plan's checks ran needs a check of its own
    // (a check needs a synchronisation, so none is captured)
    const bool unproven = can && (L.state == 0 || std::min<int64_t>(L.d.tmax, trip) > L.validated);
    if (!can || (unproven && b->mode == kCapture))
      plain.push_back((int)t);
    else if (unproven)
      check.push_back((int)t);
    else
      blocked.push_back((int)t);
  }
  std::vector<int64_t> remaining(idx.size(), trip);
  auto tb_launch = [&](const std::vector<int>& set, bool chk) -> int {
    if (set.empty()) return XG_OK;
    xg_kernel* k0 = b->inst[(size_t)idx[(size_t)set[0]]].k;
    const LoopPlan& L0 = k0->loops[(size_t)lid];
    std::vector<xg_btent> ents;
    std::vector<std::vector<size_t>> roff, poff;
    int64_t total = 0;
    for (int t : set) {
      xg_kernel* k = b->inst[(size_t)idx[(size_t)t]].k;
      LoopPlan& L = k->loops[(size_t)lid];
      xg_grid* G = k->grids[(size_t)L.d.grid_arg];
      const int T = (int)std::min<int64_t>(L.d.tmax, remaining[(size_t)t]);
      xg_btent e;
      std::memset(&e, 0, sizeof e);
      fill_tb_hdr(k, L, e.th, T, k->d_fail);
      e.tile0 = total;
      e.ntiles = ceil_div(G->shape[0], L.d.core0) * ceil_div(G->shape[1], L.d.core1);
      total += e.ntiles;
      std::vector<size_t> ro, po;
      for (int j = 0; j < L.d.nstmts; ++j) {
        const xg_stmt_desc& d = k->stmts[(size_t)L.d.sids[j]].d;
        std::vector<xg_gview> refs;
        if (int r = fill_refs(k, d, refs)) return r;
        ro.push_back(put(b, refs.data(), refs.size() * sizeof(xg_gview)));
        po.push_back(put(b, par[(size_t)t][(size_t)j], (size_t)std::max<int32_t>(d.param_bytes, 1)));
      }
      ents.push_back(e);
      roff.push_back(ro);
      poff.push_back(po);
      L.launches++;
    }
    if (total > 0x7fffffff) return fail(XG_ERR_ARG, "batched launch has too many tiles");
    const size_t base = put(b, ents.data(), ents.size() * sizeof(xg_btent));
    for (size_t i = 0; i < ents.size(); ++i)
      for (size_t j = 0; j < roff[i].size(); ++j) {
        set_rel(b, base + i * sizeof(xg_btent) + offsetof(xg_btent, refs) + j * sizeof(void*), roff[i][j]);
        set_rel(b, base + i * sizeof(xg_btent) + offsetof(xg_btent, params) + j * sizeof(void*), poff[i][j]);
      }
    b->launches.push_back(PendingLaunch{chk ? L0.d.fn_check_b : L0.d.fn_b, kTiles, L0.d.threads, base,
                                        (int32_t)ents.size(), total});
    return XG_OK;
  };
  auto advance = [&](const std::vector<int>& set) {
    for (int t : set) {
      xg_kernel* k = b->inst[(size_t)idx[(size_t)t]].k;
      LoopPlan& L = k->loops[(size_t)lid];
      xg_grid* G = k->grids[(size_t)L.d.grid_arg];
      const int64_t T = std::min<int64_t>(L.d.tmax, remaining[(size_t)t]);
      std::swap(G->levels[0], G->scratch);
      remaining[(size_t)t] -= T;
      L.iterations += T;
    }
  };
  if (!check.empty()) {
    // prove the plan on each instance's mask: a checking launch, then read the flags
    for (int t : check) {
      xg_kernel* k = b->inst[(size_t)idx[(size_t)t]].k;
      if (!k->d_fail) {
        void* p = nullptr;
        if (int rc = dev_alloc(sizeof(int32_t), true, &p)) return rc;
        k->d_fail = (int32_t*)p;
      }
      XG_HIP(hipMemsetAsync(k->d_fail, 0, sizeof(int32_t), g_dev.stream));
    }
    int rc = tb_launch(check, true);
    if (!rc) rc = flush_direct(b);
    if (rc) return rc;
    std::vector<int32_t> flags(check.size(), 0);
    for (size_t c = 0; c < check.size(); ++c) {
      xg_kernel* k = b->inst[(size_t)idx[(size_t)check[c]]].k;
      XG_HIP(hipMemcpyAsync(&flags[c], k->d_fail, sizeof(int32_t), hipMemcpyDeviceToHost, g_dev.stream));
    }
    XG_HIP(hipStreamSynchronize(g_dev.stream));
    b->n_validations++;
    std::vector<int> passed;
    for (size_t c = 0; c < check.size(); ++c) {
      xg_kernel* k = b->inst[(size_t)idx[(size_t)check[c]]].k;
      LoopPlan& L = k->loops[(size_t)lid];
      if (flags[c]) {
        L.state = -1;                 // its scratch is discarded; it runs the loop unblocked
        plain.push_back(check[c]);
      } else {
        L.state = 1;
        L.validated = std::max<int>(L.validated, (int)std::min<int64_t>(L.d.tmax, trip));
        passed.push_back(check[c]);
      }
    }
    advance(passed);
    blocked.insert(blocked.end(), passed.begin(), passed.end());
  }
  for (;;) {
    std::vector<int> live;
    for (int t : blocked)
      if (remaining[(size_t)t] > 0) live.push_back(t);
    if (live.empty()) break;
    if (int rc = tb_launch(live, false)) return rc;
    advance(live);
  }
  if (!plain.empty()) {
    std::sort(plain.begin(), plain.end());
    std::vector<int> sub;
    for (int t : plain) sub.push_back(idx[(size_t)t]);
    xg_kernel* k0 = b->inst[(size_t)sub[0]].k;
    const LoopPlan& L0 = k0->loops[(size_t)lid];
    std::vector<Pos> body;
    for (int64_t it = 0; it < trip; ++it)
      for (int j = 0; j < L0.d.nstmts; ++j) {
        Pos p;
        p.sid = L0.d.sids[j];
        for (int t : plain) p.par.push_back(par[(size_t)t][(size_t)j]);
        body.push_back(std::move(p));
      }
    if (int rc = run_positions(b, body, sub)) return rc;
  }
  return XG_OK;
}

// Walk the recorded streams of instances idx (identical shape) and build their launches.
int run_recorded(xg_batch* b, const std::vector<int>& idx) {
  xg_kernel* k0 = b->inst[(size_t)idx[0]].k;
  const size_t nrec = k0->rec.size();
  size_t r = 0;
  while (r < nrec) {
    const Rec& R0 = k0->rec[r];
    if (R0.kind == 0) {
      std::vector<Pos> pos;
      size_t e = r;
      while (e < nrec && k0->rec[e].kind == 0) {
        Pos p;
        p.sid = k0->rec[e].id;
        for (int i : idx) {
          xg_kernel* k = b->inst[(size_t)i].k;
          p.par.push_back(k->rec_arena.data() + k->rec[e].off);
        }
        pos.push_back(std::move(p));
        ++e;
      }
      if (int rc = run_positions(b, pos, idx)) return rc;
      r = e;
      continue;
    }
    std::vector<std::vector<const void*>> par;
    for (int i : idx) {
      xg_kernel* k = b->inst[(size_t)i].k;
      const LoopPlan& L = k->loops[(size_t)R0.id];
      std::vector<const void*> ps;
      size_t off = k->rec[r].off;
      for (int j = 0; j < L.d.nstmts; ++j) {
        ps.push_back(k->rec_arena.data() + off);
        const int64_t bytes = k->stmts[(size_t)L.d.sids[j]].d.param_bytes;
        off += (size_t)((bytes + 7) & ~7ll);
      }
      par.push_back(ps);
    }
    if (int rc = run_loop(b, R0.id, R0.trip, idx, par)) return rc;
    ++r;
  }
  return XG_OK;
}

// One call of every instance, recorded and launched batched (direct or capture build).
int batch_call(xg_batch* b) {
  reset_table(b);
  for (auto& I : b->inst) {
    I.k->record_next = true;
    b->entry(I.k, &g_api, I.args.data());
    I.k->record_next = false;
    if (I.k->status != XG_OK) {
      t_err = I.k->status_msg;
      return I.k->status;
    }
    if (I.k->mode != kDirect) return fail(XG_ERR_STATE, "an instance's call did not finish");
  }
  // instances whose streams agree (statement ids, loop ids and trips) are walked together
  std::map<std::vector<int64_t>, std::vector<int>> parts;
  std::vector<const std::vector<int64_t>*> order;
  for (size_t i = 0; i < b->inst.size(); ++i) {
    std::vector<int64_t> sig;
    for (const Rec& r : b->inst[i].k->rec) {
      sig.push_back(r.kind);
      sig.push_back(r.id);
      sig.push_back(r.trip);
    }
    auto it = parts.find(sig);
    if (it == parts.end()) {
      it = parts.emplace(sig, std::vector<int>()).first;
      order.push_back(&it->first);
    }
    it->second.push_back((int)i);
  }
  for (const auto* sig : order) {
    const std::vector<int>& idx = parts[*sig];
    if (int rc = run_recorded(b, idx)) return rc;
  }
  return XG_OK;
}

void destroy_graph(BatchGraph& g) {
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g.exec = nullptr;
  for (void* t : g.tables) (void)hipFree(t);
  g.tables.clear();
}

void clear_graphs(xg_batch* b) {
  if (b->cache.empty()) return;
  (void)hipStreamSynchronize(g_dev.stream);
  for (auto& g : b->cache) destroy_graph(g);
  b->cache.clear();
}

void build_key(xg_batch* b) {
  auto& kb = b->key_buf;
  kb.clear();
  auto put_k = [&](const void* p, size_t n) { kb.insert(kb.end(), (const uint8_t*)p, (const uint8_t*)p + n); };
  for (auto& I : b->inst) {
    put_k(&I.k, sizeof I.k);
    const int64_t n = (int64_t)I.args.size();
    put_k(&n, sizeof n);
    put_k(I.args.data(), I.args.size());
    const uint8_t opts[2] = {(uint8_t)I.k->blocked, (uint8_t)I.k->fuse};
    put_k(opts, 2);
    for (auto* G : I.k->grids) {
      put_k(&G, sizeof G);
      for (void* l : G->levels) put_k(&l, sizeof l);
      put_k(&G->scratch, sizeof G->scratch);
      put_k(&G->mask_version, sizeof G->mask_version);
    }
    for (auto& L : I.k->loops) put_k(&L.state, sizeof L.state);
  }
}

int finish_capture(xg_batch* b, BatchGraph* g) {
  // the table is uploaded before capture; the graph holds only the launches
  void* dev = nullptr;
  if (!b->launches.empty()) {
    XG_HIP(hipMalloc(&dev, std::max<size_t>(b->arena.size(), 16)));
    g->tables.push_back(dev);
    patch(b, dev);
    XG_HIP(hipMemcpyAsync(dev, b->arena.data(), b->arena.size(), hipMemcpyHostToDevice, g_dev.stream));
    XG_HIP(hipStreamSynchronize(g_dev.stream));
  }
  XG_HIP(hipStreamBeginCapture(g_dev.stream, hipStreamCaptureModeThreadLocal));
  int rc = issue(b, dev);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(g_dev.stream, &graph);
  reset_table(b);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : fail(XG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(XG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  g->exec = exec;
  g->post_levels.clear();
  g->post_scratch.clear();
  for (auto& I : b->inst)
    for (auto* G : I.k->grids) {
      g->post_levels.push_back(G->levels);
      g->post_scratch.push_back(G->scratch);
    }
  b->n_capture++;
  XG_HIP(hipGraphLaunch(exec, g_dev.stream));
  return XG_OK;
}

int batch_step_once(xg_batch* b) {
  if (b->tick)
    for (auto& I : b->inst)
      for (auto* G : I.ticks)
        if (G->levels.size() > 1) std::rotate(G->levels.rbegin(), G->levels.rbegin() + 1, G->levels.rend());
  b->n_steps++;
  BatchGraph* g = nullptr;
  b->mode = kDirect;
  bool recorded_before = true;
  for (auto& I : b->inst) recorded_before = recorded_before && !I.k->grids.empty();
  if (b->graphs && recorded_before) {
    build_key(b);
    const uint64_t h = fnv1a(b->key_buf.data(), b->key_buf.size());
    ++b->clock;
    for (auto& e : b->cache)
      if (e.hash == h && e.key == b->key_buf) {
        g = &e;
        break;
      }
    if (g && g->exec) {
      g->last_use = b->clock;
      XG_HIP(hipGraphLaunch(g->exec, g_dev.stream));
      size_t q = 0;
      for (auto& I : b->inst)
        for (auto* G : I.k->grids) {
          G->levels = g->post_levels[q];
          G->scratch = g->post_scratch[q];
          ++q;
        }
      b->n_replay++;
      return XG_OK;
    }
    if (g) {
      g->last_use = b->clock;
      b->mode = kCapture;
    } else {
      if (b->cache.size() >= 8) {
        auto lru = std::min_element(b->cache.begin(), b->cache.end(),
                                    [](const BatchGraph& x, const BatchGraph& y) { return x.last_use < y.last_use; });
        XG_HIP(hipStreamSynchronize(g_dev.stream));
        destroy_graph(*lru);
        b->cache.erase(lru);
      }
      BatchGraph ng;
      ng.key = b->key_buf;
      ng.hash = h;
      ng.seen = 1;
      ng.last_use = b->clock;
      b->cache.push_back(std::move(ng));
    }
  }
  int rc = batch_call(b);
  if (rc) {
    reset_table(b);
    b->mode = kDirect;
    return rc;
  }
  rc = b->mode == kCapture ? finish_capture(b, g) : flush_direct(b);
  b->mode = kDirect;
  return rc;
}

}  // namespace

extern "C" {

int xg_batch_create(void* packed_entry, int32_t tick, xg_batch** out) {
  if (int rc = check_ready()) return rc;
  if (!packed_entry || !out) return fail(XG_ERR_ARG, "null entry point or output");
  auto* b = new xg_batch;
  b->entry = (xg_packed_entry)packed_entry;
  b->tick = tick != 0;
  *out = b;
  return XG_OK;
}

int xg_batch_destroy(xg_batch* b) {
  if (!b) return XG_OK;
  if (g_dev.ready) (void)hipStreamSynchronize(g_dev.stream);
  for (auto& g : b->cache) destroy_graph(g);
  if (b->d_table) (void)hipFree(b->d_table);
  for (int i = 0; i < xg_batch::kStages; ++i) {
    if (b->h_stage[i]) (void)hipHostFree(b->h_stage[i]);
    if (b->stage_ev[i]) (void)hipEventDestroy(b->stage_ev[i]);
  }
  delete b;
  return XG_OK;
}

int xg_batch_add(xg_batch* b, xg_kernel* k, const void* args, int64_t arg_bytes, xg_grid* const* tick_grids,
                 int32_t nticks, int64_t* id) {
  if (!b || !k || (!args && arg_bytes > 0) || arg_bytes < 0 || nticks < 0 || (nticks > 0 && !tick_grids))
    return fail(XG_ERR_ARG, "invalid batch instance");
  // the batched variants come from a second code object that the host loads on
  // demand: a plan made before it was loaded cannot run in a batch
  for (auto& sp : k->stmts)
    if (!sp.d.fn_main_b) return fail(XG_ERR_STATE, "the kernel's batched variants are not loaded");
  for (auto& g : k->groups)
    if (!g.fn_b) return fail(XG_ERR_STATE, "the kernel's batched variants are not loaded");
  for (auto& I : b->inst) {
    if (I.k == k) return fail(XG_ERR_ARG, "the kernel plan already belongs to this batch");
    for (auto* G : I.ticks)
      for (int32_t t = 0; t < nticks; ++t)
        if (G == tick_grids[t]) return fail(XG_ERR_ARG, "a grid belongs to two instances of the batch");
  }
  clear_graphs(b);
  Inst I;
  I.id = b->next_id++;
  I.k = k;
  I.args.assign((const uint8_t*)args, (const uint8_t*)args + arg_bytes);
  I.ticks.assign(tick_grids, tick_grids + nticks);
  k->grids.clear();           // set by the first recorded call
  b->inst.push_back(std::move(I));
  if (id) *id = b->inst.back().id;
  return XG_OK;
}

int xg_batch_remove(xg_batch* b, int64_t id) {
  if (!b) return fail(XG_ERR_ARG, "null batch");
  auto it = std::find_if(b->inst.begin(), b->inst.end(), [&](const Inst& I) { return I.id == id; });
  if (it == b->inst.end()) return fail(XG_ERR_ARG, "no such instance");
  clear_graphs(b);
  b->inst.erase(it);
  return XG_OK;
}

int xg_batch_size(xg_batch* b) { return b ? (int)b->inst.size() : -1; }

int xg_batch_set_options(xg_batch* b, int32_t graphs, int64_t max_blocks) {
  if (!b || max_blocks < 0) return fail(XG_ERR_ARG, "invalid batch options");
  b->graphs = graphs != 0;
  b->max_blocks = max_blocks;
  clear_graphs(b);
  return XG_OK;
}

int xg_batch_step(xg_batch* b, int32_t steps) {
  if (int rc = check_ready()) return rc;
  if (!b || steps < 0) return fail(XG_ERR_ARG, "invalid batch or step count");
  if (b->inst.empty()) return XG_OK;
  for (int32_t s = 0; s < steps; ++s)
    if (int rc = batch_step_once(b)) return rc;
  return XG_OK;
}

int xg_batch_counters(xg_batch* b, int64_t* launches, int64_t* replays, int64_t* captures, int64_t* steps,
                      int64_t* validations) {
  if (!b) return fail(XG_ERR_ARG, "null batch");
  if (launches) *launches = b->n_launch;
  if (replays) *replays = b->n_replay;
  if (captures) *captures = b->n_capture;
  if (steps) *steps = b->n_steps;
  if (validations) *validations = b->n_validations;
  return XG_OK;
}

}  // extern "C"
