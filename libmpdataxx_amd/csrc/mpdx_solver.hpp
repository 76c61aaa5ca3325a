// Host-side state of one solver object (one advectee set sharing one advector).
#pragma once

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "mpdx.h"
#include "mpdx_common.hpp"

struct mpdx_fused_state;

namespace mpdx {

// launches k_task_list<T> over n tasks at `tasks` (device) with the block
// prefix table `first` (device) and `blocks` blocks in all
using TaskLaunchFn = void (*)(hipStream_t, const void *tasks, const unsigned *first, int n, unsigned blocks);

// one launch of a solver's time step, recorded as a task of a batched kernel
struct RecOp {
  TaskLaunchFn fn = nullptr;
  unsigned tsize = 0, blocks = 0;
  std::vector<unsigned char> bytes;
};

// The launches of one time step of a solver, recorded instead of launched
// (ensembles batch them across members).  Launches made between begin() and
// end() form one group: they touch disjoint cells and read nothing the
// others write, so they may run in any order or together; a group runs
// after the previous one has finished, as launches on one stream would.
struct TaskRecorder {
  std::vector<std::vector<RecOp>> groups;
  int depth = 0;
  void begin()
  {
    if (depth++ == 0) groups.emplace_back();
  }
  void end() { --depth; }
  RecOp &next()
  {
    if (depth == 0) groups.emplace_back();
    groups.back().emplace_back();
    return groups.back().back();
  }
};

// a batched launch of k_task_list: its kernel, the byte offsets of its
// tasks and of their block prefix table in BatchPlan::dev, tasks, blocks,
// and the earlier batches it must follow (for each sequence it holds tasks
// of, the last earlier batch holding tasks of that sequence)
struct EnsBatch {
  TaskLaunchFn fn;
  size_t toff, foff;
  int n;
  unsigned blocks;
  std::vector<int> deps;
};

// recorded launches of several independent sequences merged into batches
struct BatchPlan {
  void *dev = nullptr;
  std::vector<EnsBatch> batches;
  size_t tasks = 0;
};

}  // namespace mpdx

// per-kernel-class timing with HIP events on the solver's stream
struct mpdx_prof_entry {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0;
  int64_t launches = 0;
};

struct mpdx_solver {
  mpdx_params p;
  mpdx::Layout L;
  int esz = 8;
  int device = 0;            // HIP device current at creation
  std::vector<void *> allocs;
  void *psi[MPDX_MAX_EQNS][2] = {};
  int lev[MPDX_MAX_EQNS] = {};
  void *GC[3] = {};
  void *tmp[2][3] = {};
  // antidiffusive-velocity scratch of equations 1.. when the equations'
  // steps run batched (equation 0 uses tmp); tmp_sel >= 0 selects a set
  void *tmp_eq[MPDX_MAX_EQNS][2][3] = {};
  int tmp_sel = -1;
  void *flux[3] = {};
  void *mono[3] = {};
  void *pmin = nullptr, *pmax = nullptr, *bup = nullptr, *bdn = nullptr;
  bool perpass_ready = false;
  double2 *red_part = nullptr;   // block partials of a reduction (mpdx::red_parts(n0))
  double *red_out = nullptr;
  void *staging = nullptr;
  size_t staging_bytes = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int64_t timestep = 0;
  double time = 0;
  int panic = 0;
  const volatile unsigned char *panic_ext = nullptr;   // caller's flag (mpdx_watch_panic)
  // one event per queued step of advance(): the host waits for the step
  // kInflight back before queueing another, so a panic flag takes effect
  // within kInflight steps
  static constexpr int kInflight = 6;
  hipEvent_t inflight[kInflight] = {};
  int active_path = MPDX_PATH_PERPASS;
  std::string err;
  mpdx_fused_state *fz = nullptr;
  mpdx::TaskRecorder *rec = nullptr;   // set: launches are recorded, not made
  int tile_y = 0;            // interior tile height (0: the element type's default)
  // recorded FP32 interior passes and 2D passes as mixed-variant tasks (one
  // batched launch for every member's pass; MPDX_NO_TASK_MIX=1: one launch
  // per variant)
  bool mix_tasks = true;
  // one captured time step per time-level parity and seam mode (fused
  // path, index parity + 2 seam; + 6: graph_steps steps in one graph); with
  // several equations the step's launches are batched across them (eqplan:
  // the device tables the captured step reads)
  static constexpr int kGraphs = 12;
  hipGraphExec_t graph[kGraphs] = {};
  mpdx::BatchPlan eqplan[kGraphs];
  bool use_graphs = true;
  // steps per captured graph: several for small steps, whose launch
  // overhead would dominate (0: chosen at the first capture; an even number,
  // so the time levels return to the same parity)
  int graph_steps = 0;
  bool batch_eqns = false;   // MPDX_BATCH_EQNS=1: batched equations always (slower beside a boundary shell); velocity scratch per equation
  // batched equations by default where they pay: all-cyclic seam mode (no
  // boundary shell) with the extra velocity scratch in an eighth of the
  // device's memory (MPDX_NO_AUTO_BATCH=1: off)
  bool auto_batch_eqns = true;
  int auto_batch_fits = -1;  // the scratch fits the bound (-1: not tested yet)
  bool rec_tile16 = false;   // recording a batched-equation step: 16-row tiles when tile_y is 0
  // auxiliary stream (boundary shell next to the interior kernel) and its
  // fork / join events
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool use_concurrency = true;
  // with open bounds on both z sides the interior tiles also compute the
  // z-boundary cells (no z slabs in the boundary shell)
  bool use_zbnd = true;
  // with cyclic bounds on both z sides the tiles also compute the
  // z-boundary cells while the fields repeat plane z = 0 at z = n-1, and
  // with cyclic bounds everywhere the x- and y-boundary cells as well while
  // the fields repeat those planes too (seam: 0 no, 1 z, 2 x, y and z;
  // checked on the device at the start of each advance call)
  bool use_seam = true;
  int seam = 0;
  // the seam mode stays valid while only the library's own steps change the
  // fields (they keep the planes equal); setting a field or the advector's
  // source drops it, and once a raw device pointer to a field has been
  // handed out every advance call checks again
  bool seam_known = false, seam_exposed = false;
  bool use_pass2d = true;   // 2D: the pass kernel (MPDX_NO_PASS2D=1: the shell kernels over every cell)
  int *seam_flag = nullptr;
  int ix_chunk = 0;         // output planes per x chunk of the interior kernel (0: automatic)
  // velocity arrays whose tangential z halo (open z bounds) has been
  // zero-filled in full at least once (later fills only refill the bands
  // next to the x and y edges)
  std::vector<const void *> zhalo_zeroed;
  int n_cu = 256;           // compute units of the device
  // explicit forcing (mpdata_rhs): scheme, user update_rhs, forcing arrays
  int rhs_scheme = MPDX_RHS_NONE;
  mpdx_update_rhs_fn rhs_fn = nullptr;
  void *rhs_user = nullptr;
  mpdx_rhs_hook_fn rhs_hook = nullptr;   // the mixed scheme's hooks
  void *rhs_hook_user = nullptr;
  void *rhs[MPDX_MAX_EQNS] = {};
  mpdx_calc_gc_fn calc_gc = nullptr;
  void *calc_gc_user = nullptr;
  // var_dt output (output_common.hpp:32-37,72-140): out_buf[k][v] holds the
  // state saved before step k of a record's interpolation (k = 0 afterwards
  // the interpolated field) for equation out_var[v]; out_cnt is the
  // reference's do_record_cnt, out_t its intrp_times, out_rec record_time
  mpdx_output_fn out_fn = nullptr;
  void *out_user = nullptr;
  double out_freq = 1;
  int out_ord = 1, out_n = 0;
  int out_var[MPDX_MAX_EQNS] = {};
  void *out_buf[2][MPDX_MAX_EQNS] = {};
  int out_cnt = 0;
  double out_t[2] = {}, out_rec = 0;
  int gen_kind = MPDX_GC_NONE;
  double gen_T = 1, gen_scale = 0;
  double *gen_sx2 = nullptr, *gen_sy2 = nullptr;   // device sin^2 tables from index -2
  double *gen_amp = nullptr;   // device amplitude per step of the current advance call
  long gen_amp_cap = 0;
  int *gen_step = nullptr;     // device index into gen_amp
  bool prof_on = false;
  std::vector<std::string> prof_order;
  std::map<std::string, mpdx_prof_entry> prof;
  std::vector<hipEvent_t> ev_pool;
};

namespace mpdx {

inline hipEvent_t prof_event(mpdx_solver *s)
{
  hipEvent_t e = nullptr;
  if (!s->ev_pool.empty()) {
    e = s->ev_pool.back();
    s->ev_pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

// records a start/stop event pair around the kernel launches in its scope
struct ProfScope {
  mpdx_solver *s;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(mpdx_solver *s_, const char *n) : s(s_), name(n)
  {
    if (!s->prof_on) return;
    a = prof_event(s);
    b = prof_event(s);
    if (a) (void)hipEventRecord(a, s->stream);
  }
  ~ProfScope()
  {
    if (!a || !b) return;
    (void)hipEventRecord(b, s->stream);
    auto it = s->prof.find(name);
    if (it == s->prof.end()) {
      s->prof_order.push_back(name);
      it = s->prof.emplace(name, mpdx_prof_entry()).first;
    }
    it->second.pending.emplace_back(a, b);
  }
};

inline void prof_collect(mpdx_solver *s)
{
  for (auto &kv : s->prof) {
    for (auto &pr : kv.second.pending) {
      float ms = 0;
      (void)hipEventSynchronize(pr.second);
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) kv.second.ms += ms;
      kv.second.launches += 1;
      s->ev_pool.push_back(pr.first);
      s->ev_pool.push_back(pr.second);
    }
    kv.second.pending.clear();
  }
}

}  // namespace mpdx

