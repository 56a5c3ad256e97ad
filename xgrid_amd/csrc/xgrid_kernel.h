// Internal: the launch plan of one compiled kernel (xg_kernel) and the helpers
// that turn one stencil statement into launch arguments.  Shared by the
// runtime (single calls) and the batch executor (many instances per launch).
#ifndef XGRID_KERNEL_H
#define XGRID_KERNEL_H

#include <cstdint>
#include <string>
#include <vector>

#include "xgrid_internal.h"

namespace xgi {

constexpr int kThreads = 256;
constexpr int kGroupThreads = 1024;
constexpr int64_t kFuseMaxCells = 4096;  // default of g_dev.fuse_max_cells (XGRID_FUSE_MAX_CELLS)
constexpr int64_t kGraphMinLaunches = 4;    // a call with fewer launches runs them directly (cheaper)
// Store grids larger than the 256 MB last-level cache (Infinity Cache) take non-temporal
// stores in 3-D box tiles: no later launch can find their cells in cache anyway.  The
// 512^3 heat update moves the same bytes with them and runs 5-6 % faster (the write
// stream drains faster); smaller grids keep plain stores for the next statement.
constexpr size_t kStreamBytes = (size_t)256 << 20;
// Freed grid memory the device pool keeps for reuse (beyond it, memory goes back at
// the next synchronisation).
constexpr uint64_t kPoolKeepBytes = (uint64_t)16 << 30;

// kRecord: the stub's statements are recorded for a batched launch instead of run
enum CallMode { kDirect = 0, kCapture = 1, kReplay = 2, kRecord = 3 };

struct StmtPlan {
  xg_stmt_desc d;
  uint64_t haz_version = 0;
  bool haz = false;
};

struct Pending {
  int32_t sid;
  size_t off;   // offset of its params in the arena
  int64_t bytes;
};

struct GraphEntry {
  std::vector<uint8_t> key;
  uint64_t hash = 0;
  int seen = 0;
  hipGraphExec_t exec = nullptr;
  bool direct = false;          // the call launches too little for a graph to pay: run it directly
  std::vector<std::vector<void*>> post_levels;
  std::vector<void*> post_scratch;
  uint64_t last_use = 0;
};

struct EventPair {
  hipEvent_t a, b;
};

struct LoopPlan {
  xg_loop_desc d;
  const void* grid = nullptr;   // the grid the plan was validated on
  uint64_t mask_version = 0;    // ... and its mask
  int state = 0;                // 1 exact, 0 unknown, -1 declined
  int validated = 0;            // state 1: launches of at most this many iterations are proven exact
  int64_t launches = 0, iterations = 0;
  std::vector<std::vector<uint8_t>> last_params;
};

// One recorded stub event: a statement (kind 0) or a whole blockable loop (kind 1).
struct Rec {
  int32_t kind;
  int32_t id;         // statement id or loop id
  int64_t trip;       // loop iterations
  size_t off;         // parameter block(s) in the record arena, 8-byte aligned each
};

}  // namespace xgi

struct xg_kernel {
  std::vector<xgi::StmtPlan> stmts;
  std::vector<xg_group_desc> groups;
  int32_t ngrids = 0;
  std::vector<xg_grid*> grids;
  int status = XG_OK;
  std::string status_msg;
  int mode = xgi::kDirect;
  xgi::GraphEntry* cur = nullptr;
  std::vector<xgi::Pending> pending;
  std::vector<uint8_t> arena;
  int32_t pending_group = -1;
  std::vector<xgi::Pending> run_buf;   // flush's working list (kept for its capacity)
  bool graphs = true, fuse = true;
  bool graphs_auto_off = false;
  int64_t graph_misses = 0, graph_hits = 0;
  std::vector<xgi::GraphEntry> cache;
  uint64_t tick = 0;
  int64_t n_launch = 0, n_fused = 0, n_replay = 0, n_capture = 0;
  int64_t capture_mark = 0;         // n_launch when the current capture began
  int32_t timed_sid = -1;
  std::vector<xgi::EventPair> events;
  std::vector<xgi::LoopPlan> loops;
  bool blocked = true;
  int32_t* d_fail = nullptr;        // validation flag of blocked launches
  std::vector<int64_t> exec_count;  // stmt() calls per statement (all modes)
  std::vector<std::vector<uint8_t>> last_params;  // parameters of each statement's last call
  std::vector<uint8_t> key_buf;
  bool record_next = false;         // the next call through the stub is recorded (batches)
  std::vector<xgi::Rec> rec;        // recorded events of the last recorded call
  std::vector<uint8_t> rec_arena;
};

namespace xgi {

extern const xg_api g_api;   // the stub function table
uint64_t fnv1a(const uint8_t* p, size_t n);
int set_status(xg_kernel* k, int code);
int effective_mode(xg_kernel* k, StmtPlan& sp);
void fill_dom(xg_dom& dom, const ClassPlan& p, const xg_grid* G, int32_t klass);
void fill_view(xg_gview& v, const xg_grid* G, void* ptr);
bool uniform_refs(xg_kernel* k, const xg_stmt_desc& d);
int32_t stream_flag(const xg_grid* G);
int fill_refs(xg_kernel* k, const xg_stmt_desc& d, std::vector<xg_gview>& refs);
// Workgroups over the box [lo, hi) for a statement's tile; cols3 = columns of a 3-D tile.
int64_t tiles_for(int D, const int32_t* lo, const int32_t* hi, int cols3 = 64);
int64_t blocks_for_dom(int D, const xg_dom& dom, int cols3 = 64);
// The 3-D tile width of a statement's run_main kernels (fn_main, fn_main_b).
int tile3_cols(const xg_stmt_desc& d);
int ensure_scratch(xg_kernel* k);
// Can loop `lid` run blocked on the bound grids (aliasing), and is its plan for
// the grid's current mask known?  Resets the plan when the grid or mask changed.
bool loop_blockable(xg_kernel* k, LoopPlan& L);
void fill_tb_hdr(xg_kernel* k, LoopPlan& L, xg_tb_hdr& TH, int iters, int32_t* fail_flag);

}  // namespace xgi

#endif
