// capi.hip — host runtime and C ABI of libavbd_hip.so (declarations in
// include/avbd.h). Owns every device buffer of a solver, turns the static rows
// of a flush into contact groups, and issues the substep on one HIP stream in
// the order of Solver6DOF._step_one (src/avbd3d/solver_6dof.py:1232-1464),
// readback path: the pair count, the row reservation count and the
// color-conflict count are read back each substep and drive the same
// regrow / recolor decisions as the reference.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/avbd.h"
#include "avbd_internal.h"

using namespace avbd;

namespace avbd {
void color_bucket(hipStream_t st, int nb, const int* color, unsigned* keys, int* vals,
                  unsigned* keys_tmp, int* vals_tmp, void* ws, const int* gate = nullptr);
}

namespace {

thread_local std::string g_err;
// Bumped by every device (re)allocation of a solver buffer: a captured
// substep graph holds raw pointers, so any reallocation retires it.
unsigned long long g_alloc_gen = 0;

int set_err(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCK(expr)                                                                     \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return set_err(AVBD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define CK(expr)               \
  do {                         \
    int r_ = (expr);           \
    if (r_ != AVBD_OK) return r_; \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      ++g_alloc_gen;
    }
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    release();
    size_t c = count ? count : 1;
    if (hipMalloc(&p, c * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return set_err(AVBD_E_NOMEM, "hipMalloc of " + std::to_string(c * sizeof(T)) + " bytes failed");
    }
    ++g_alloc_gen;
    n = c;
    return AVBD_OK;
  }
  int ensure(size_t count) { return count <= n && p ? AVBD_OK : alloc(count); }
  // Grow to `count`, keeping the first `keep` elements.
  int grow_keep(size_t count, size_t keep, hipStream_t st) {
    if (count <= n && p) return AVBD_OK;
    T* q = nullptr;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess)
      return set_err(AVBD_E_NOMEM, "hipMalloc (grow) failed");
    if (p && keep) {
      hipError_t e = hipMemcpyAsync(q, p, std::min(keep, n) * sizeof(T), hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) return set_err(AVBD_E_HIP, hipGetErrorString(e));
      (void)hipStreamSynchronize(st);
    }
    release();
    p = q;
    n = count;
    ++g_alloc_gen;
    return AVBD_OK;
  }
};

enum ProfClass { kPrimal = 0, kDual, kPrelude, kContacts, kColoring, kCsr, kOther, kNumProf };

// Device counters (dcnt) by slot: 0 pair count, 1-3 emission totals (row
// demand, emitted groups, reserved rows), 4 conflict rows, 5 uncolored
// bodies, 6 coloring error, 8 dataflow error (sticky until the next flush),
// 9 color clamp, 10 residual, 12 substep epoch (stamps of the dataflow
// sweeps). Substeps whose decisions stay on the device (graph capture) also
// keep there: 13 kept pairs, 14 live groups, 15 cache entries, 16 slots of
// the warm-start table as last built, 17 / 18 the recolor gate as a count
// (all / the kept pairs, 0 when the coloring is kept), 19 recolors done, 20
// achieved color count, 32.. per-color body counts, 64.. color starts,
// 100.. first tile of each color. One copy of slots [0, kCounterSlots)
// brings every count to the host.
constexpr int kPairSlot = 0, kGroupsSlot = 2, kRowsSlot = 3, kConflictSlot = 4;
constexpr int kUncoloredSlot = 5;
constexpr int kErrSlot = 8;
constexpr int kEpochSlot = 12;
constexpr int kKeptSlot = 13, kNGroupsSlot = 14, kEntriesSlot = 15, kHashLiveSlot = 16;
constexpr int kGateAllSlot = 17, kGateKeptSlot = 18, kRecolorsSlot = 19, kNColorsSlot = 20;
constexpr int kCountsSlot = 32, kCstartSlot = 64, kTilesSlot = 100;
constexpr int kCounterSlots = 136;
constexpr int kDcntSize = 256;

}  // namespace

struct avbd_solver {
  int device = 0;
  hipStream_t st = nullptr;      // the stream every launch and copy goes to
  hipStream_t own_st = nullptr;  // the solver's own stream (st unless a caller's is set)
  bool flushed = false;

  // bodies
  int nb = 0;
  DBuf<float4> pose[2];
  int cur = 0;
  DBuf<float4> poseI, inert, vel, pvel, bstat, lo, hi;
  DBuf<float> Ilocal, bfric, prio;
  DBuf<int> order_id, order_jp, pos_jp;

  // rows / groups
  int n_static_rows = 0, g_static = 0, row_cap = 0, n_dyn_cap = 0, g_cap = 0;
  int n_floor_groups = 0;
  DBuf<GroupGeo> geo;
  DBuf<GroupSt> gst;
  std::vector<int> h_st_count, h_st_start;
  DBuf<int> st_count, st_start, st_idx;
  int st_total = 0;

  // pairs
  int pair_cap = 0;
  DBuf<int> pa, pb, pov, psat, mfc, mfr, demand, rpc, dscan, succ, sscan, reserve, rscan;
  DBuf<float4> pn, mfoff, mfn;
  DBuf<float> pdepth;
  DBuf<int> bcount, boff;
  DBuf<int> pslots;  // nb x kPartnerSlots: each body's first partners, sorted

  // hash
  int hash_cap = 0, pool_max = 0;
  // Slots of the table as last built: about four times its entries (at most
  // hash_cap, the reference's capacity, which sets the allocation). Lookups
  // find the same entries at any capacity above the entry count, and a table
  // sized to the live contacts keeps the scattered probes in cache.
  int hash_live = 0;
  DBuf<int> hkeys, howner, slot_of;
  DBuf<float4> hstate;

  // CSR
  // body -> group CSR in color-position order (see kernels_graph.hip)
  DBuf<int> tc_work;        // 5 * nb: as-b pair count, as-b group count, as-a start, as-a end, cursor
  DBuf<int> asb_start, asb_list, padc, cpos, pos_of, pos_map;
  DBuf<int2> tcsr;
  // dataflow primal sweep: per-body stamp of the last sweep that moved it;
  // the stamps derive from the substep epoch in dcnt[kEpochSlot]
  DBuf<int> flow_done;
  bool colors_conflict_free = false;  // no two neighbours share a color
  bool conflicts_ready = false;  // this substep's conflict count is in hcnt[4]
  bool csr_built = false;       // a CSR has been built since the last flush
  bool csr_order_stale = false; // recolored since that CSR was built

  // coloring
  DBuf<int> color, color_next, nbr_count, nbr_start, nbr_cursor, nbr, pub, ticket;
  DBuf<unsigned> ckeys, ckeys_tmp;
  DBuf<int> cvals, cvals_tmp;
  int cstart[kMaxColors + 1] = {0};
  int n_active_colors = kMaxColors;

  // workspaces + counters
  DBuf<char> ws_bvh, ws_scan, ws_color;
  size_t ws_scan_elems = 0;
  DBuf<int> dcnt;   // device counters
  int* hcnt = nullptr;  // pinned mirror

  // host mirrors of the last substep
  int n_active_rows = 0, pair_count = 0, pool_count = 0, n_contacts = 0;
  int adj_pairs = 0;  // pairs whose emission this substep feeds the coloring adjacency
  int last_post = 0;

  // batched scenes (avbd_set_scenes): per-body scene index, per-scene recolor
  // flags, and the broad phase's per-scene shifted boxes
  int n_scenes = 0;
  int scene_grid = 1;
  float scene_pitch = 0.0f;
  DBuf<int> scene_of, scene_dirty;
  DBuf<float4> slo, shi;

  // learned warm start: 0 off, 1 log (features at emission, targets at the
  // end of the substep), 2 apply the network to the initial dual state
  int ml_mode = 0;
  int ml_scope = 0;          // 1: only groups with an empty plain start
  bool ml_have_params = false;
  int ml_n = 0;              // groups featurised by the last substep
  DBuf<float> ml_params, ml_feat, ml_targ;
  DBuf<int> ml_mask, hage;   // hage: substeps each cached contact has lived

  // options
  int primal_variant = 6;  // one dataflow launch per sweep
  // dual placement with the dataflow sweep: 0 separate launch, 1 inside the
  // tiles (measured slower: the sweep is VALU-bound), 2 trailing workgroups
  int fuse_dual_mode = 0;
  int broad_lanes = 0;  // AVBD_OPT_BROAD_LANES
  // test of the dataflow give-up path (AVBD_OPT_FLOW_GIVE_UP_TEST): waits
  // poll flow_never (never written) and give up after this many polls
  int flow_test_cap = 0;
  DBuf<int> flow_never;

  // substeps with every count and decision on the device, captured as a HIP
  // graph (AVBD_OPT_GRAPH): one executable graph per pose buffer (the
  // substep alternates them), with the key of what it was captured from
  // state handed from one stage of a direct substep to the next (the stage
  // entry points run them one call at a time); `stage` is the next stage
  struct StageCtx {
    int stage = 0;
    int n_active = 0;
    bool pool_reset = false;
    int n_ml = 0;
    bool recolor = false;
    const int* pos_map = nullptr;
    int n_groups = 0;
    int iterations = 0;   // may shrink when iterating to a residual
    int post = 0;
    float tol = -1.0f;
    float resid = 0.0f;
    bool flow = false, fuse = false, tail = false;
    int flow_bpw = 2;
    int nonempty = 0;
    double pbytes = 0.0;
    float max_lin = 0.0f, max_ang = 0.0f;
    int sweeps_done = 0;
    bool finalized = false;
  } sx;
  bool graph_on = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<char> gkey[2];
  int recolors_seen = 0;     // dcnt[kRecolorsSlot] at the last frame end
  long long syncs = 0;       // device -> host synchronisations of counters
  long long graph_launches = 0;

  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_pending;  // (class, start event index)
  size_t ev_used = 0;
  double prof_ms[kNumProf] = {0};
  long long prof_launch[kNumProf] = {0};
  double prof_bytes[kNumProf] = {0};
};

namespace {

int ensure_scan_ws(avbd_solver* s, size_t n) {
  if (n <= s->ws_scan_elems && s->ws_scan.p) return AVBD_OK;
  size_t want = std::max(n, (size_t)1024);
  const size_t bytes = scan_workspace_bytes((int)want);
  CK(s->ws_scan.alloc(bytes));
  // the single-pass scan needs a zero workspace when first used
  launch_fill_int(s->st, (int*)s->ws_scan.p, (int)(bytes / sizeof(int)), 0);
  s->ws_scan_elems = want;
  return AVBD_OK;
}

int scan(avbd_solver* s, const int* in, int* out, int n, int* total_dev,
         const int* nd = nullptr) {
  CK(ensure_scan_ws(s, (size_t)n));
  exclusive_scan(s->st, in, out, n, total_dev, s->ws_scan.p, s->dcnt.p + kErrSlot, nd);
  return AVBD_OK;
}

// dcnt[kErrSlot]: set by a dataflow primal sweep or a scan whose wait hit
// its spin cap. It is copied to the host at the end of every substep and
// checked after every later sync (readbacks and reads), and it stays set
// until the next flush: the state it left behind is not the reference's.
int check_flow_error(avbd_solver* s) {
  if (s->hcnt[kErrSlot] != 0)
    return set_err(AVBD_E_STATE,
                   "a dataflow wait (primal sweep or scan) did not complete; flush to reset");
  return AVBD_OK;
}

int readback(avbd_solver* s, int first, int count) {
  HIPCK(hipMemcpyAsync(s->hcnt + first, s->dcnt.p + first, count * sizeof(int),
                       hipMemcpyDeviceToHost, s->st));
  HIPCK(hipStreamSynchronize(s->st));
  ++s->syncs;
  return check_flow_error(s);
}

// ---- profiling ----
int prof_begin(avbd_solver* s, int cls) {
  if (!s->prof) return AVBD_OK;
  if (s->ev_used + 2 > s->ev_pool.size()) {
    for (int k = 0; k < 64; ++k) {
      hipEvent_t e;
      HIPCK(hipEventCreate(&e));
      s->ev_pool.push_back(e);
    }
  }
  HIPCK(hipEventRecord(s->ev_pool[s->ev_used], s->st));
  s->ev_pending.push_back({cls, (int)s->ev_used});
  s->ev_used += 2;
  return AVBD_OK;
}
// Close the interval opened by prof_begin; it covered `launches` kernel
// launches of its class (back to back on the solver stream).
int prof_end(avbd_solver* s, double bytes, int launches = 1) {
  if (!s->prof) return AVBD_OK;
  auto& pe = s->ev_pending.back();
  HIPCK(hipEventRecord(s->ev_pool[pe.second + 1], s->st));
  s->prof_bytes[pe.first] += bytes;
  s->prof_launch[pe.first] += launches;
  return AVBD_OK;
}
int prof_collect(avbd_solver* s) {
  if (!s->prof || s->ev_pending.empty()) return AVBD_OK;
  HIPCK(hipStreamSynchronize(s->st));
  for (auto& pe : s->ev_pending) {
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, s->ev_pool[pe.second], s->ev_pool[pe.second + 1]));
    s->prof_ms[pe.first] += ms;
  }
  s->ev_pending.clear();
  s->ev_used = 0;
  return AVBD_OK;
}

int ensure_pairs(avbd_solver* s, int cap) {
  if (cap <= s->pair_cap && s->pa.p) return AVBD_OK;
  CK(s->pa.alloc(cap)); CK(s->pb.alloc(cap)); CK(s->pov.alloc(cap)); CK(s->psat.alloc(cap));
  CK(s->mfc.alloc(cap)); CK(s->mfr.alloc(cap)); CK(s->demand.alloc(cap)); CK(s->rpc.alloc(cap));
  CK(s->dscan.alloc(cap)); CK(s->succ.alloc(cap)); CK(s->sscan.alloc(cap));
  CK(s->reserve.alloc(cap)); CK(s->rscan.alloc(cap));
  CK(s->pn.alloc(cap)); CK(s->mfn.alloc(cap)); CK(s->mfoff.alloc(8 * (size_t)cap));
  CK(s->pdepth.alloc(cap));
  s->pair_cap = cap;
  return ensure_scan_ws(s, (size_t)cap);
}

// _grow_row_pool (solver_6dof.py:1698-1772): row capacity grows; the groups
// written this substep are kept.
int grow_row_pool(avbd_solver* s, int new_dyn, int live_groups) {
  int new_cap = s->n_static_rows + new_dyn;
  if (new_cap <= s->row_cap) return AVBD_OK;
  int new_g = s->g_static + new_dyn;
  CK(s->geo.grow_keep(new_g, live_groups, s->st));
  CK(s->gst.grow_keep(new_g, live_groups, s->st));
  s->g_cap = new_g;
  s->row_cap = new_cap;
  s->n_dyn_cap = new_dyn;
  int new_pool = std::max(s->pool_max, s->pair_cap * 4);
  if (new_pool > s->pool_max) {
    CK(s->slot_of.alloc(new_pool));
    s->pool_max = new_pool;
  }
  return AVBD_OK;
}

// Row-side byte model per launch (algorithmic bytes on a minimal record).
double primal_sweep_bytes(const avbd_solver* s) {
  return (double)s->nb * 120.0 + 2.0 * s->n_contacts * 112.0 + (double)s->n_floor_groups * 76.0;
}
double dual_sweep_bytes(const avbd_solver* s) {
  return (double)s->n_contacts * 136.0 + (double)s->n_floor_groups * 100.0;
}

// _gpu_recolor (solver_6dof.py:680-844)
int recolor(avbd_solver* s, int mode) {
  const int nb = s->nb;
  hipStream_t st = s->st;
  const int nd = s->n_contacts;
  CK(prof_begin(s, kColoring));
  // Neighbour sets of the dynamic contact groups, one entry per pair that
  // emitted a group this substep.
  const int np = nd > 0 ? s->adj_pairs : 0;
  launch_fill_int(st, s->nbr_count.p, nb, 0);
  launch_adj_count_pairs(st, np, s->pa.p, s->pb.p, s->succ.p, s->nbr_count.p);
  CK(scan(s, s->nbr_count.p, s->nbr_start.p, nb, s->nbr_start.p + nb));
  launch_fill_int(st, s->nbr_cursor.p, nb, 0);
  CK(s->nbr.ensure(2 * (size_t)np + 1));
  launch_adj_fill_pairs(st, np, s->pa.p, s->pb.p, s->succ.p, s->nbr_start.p, s->nbr_cursor.p,
                        s->nbr.p);
  // batched scenes: only the scenes marked dirty are recolored; the others
  // keep their colors (no edge crosses scenes)
  const int* scene = s->n_scenes > 0 ? s->scene_of.p : nullptr;
  const int* dirty = s->n_scenes > 0 ? s->scene_dirty.p : nullptr;
  launch_color_reset(st, nb, s->color.p, scene, dirty);
  launch_fill_int(st, s->dcnt.p + 6, 1, 0);
  launch_fill_int(st, s->dcnt.p + 9, 1, 0);
  bool leftover = false;  // bodies still uncolored after the last round take color 0
  if (mode == AVBD_COLOR_JACOBI) {
    for (int round = 0; round < kMaxColors; ++round) {
      launch_color_greedy(st, nb, s->order_id.p, nullptr, s->nbr_start.p, s->nbr.p, s->color.p,
                          s->pub.p, s->dcnt.p + 6, s->ticket.p, 0, s->dcnt.p + 9, scene, dirty);
      launch_fill_int(st, s->dcnt.p + 5, 1, 0);
      launch_color_resolve(st, nb, s->nbr_start.p, s->nbr.p, s->prio.p, s->color.p,
                           s->color_next.p, s->dcnt.p + 5);
      std::swap(s->color.p, s->color_next.p);
      CK(readback(s, 5, 1));
      leftover = s->hcnt[5] != 0;
      if (!leftover) break;
    }
  } else {
    launch_color_greedy(st, nb, s->order_jp.p, s->pos_jp.p, s->nbr_start.p, s->nbr.p, s->color.p,
                        s->pub.p, s->dcnt.p + 6, s->ticket.p, 1, s->dcnt.p + 9, scene, dirty);
  }
  launch_color_finish(st, nb, s->color.p, s->dcnt.p + kCountsSlot);
  color_bucket(st, nb, s->color.p, s->ckeys.p, s->cvals.p, s->ckeys_tmp.p, s->cvals_tmp.p,
               s->ws_color.p);
  CK(prof_end(s, 0.0));
  HIPCK(hipMemcpyAsync(s->hcnt + kCountsSlot, s->dcnt.p + kCountsSlot, kMaxColors * sizeof(int),
                       hipMemcpyDeviceToHost, st));
  HIPCK(hipMemcpyAsync(s->hcnt + 6, s->dcnt.p + 6, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCK(hipMemcpyAsync(s->hcnt + 9, s->dcnt.p + 9, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCK(hipStreamSynchronize(st));
  ++s->syncs;
  if (s->hcnt[6] != 0) return set_err(AVBD_E_STATE, "coloring dataflow did not converge");
  // dcnt[9]: a body found every color taken (or, JP, ran out of rounds)
  s->colors_conflict_free = s->hcnt[9] == 0 && !leftover;
  int mx = 0;
  s->cstart[0] = 0;
  for (int c = 0; c < kMaxColors; ++c) {
    s->cstart[c + 1] = s->cstart[c] + s->hcnt[kCountsSlot + c];
    if (s->hcnt[kCountsSlot + c] > 0) mx = c;
  }
  s->n_active_colors = std::max(1, std::min(mx + 1, kMaxColors));
  return AVBD_OK;
}

// Rows of this substep's contact groups whose bodies share a color, into
// dcnt[4]. Batched scenes: also marks each scene with a conflict in
// scene_dirty.
void launch_conflict_count(avbd_solver* s) {
  launch_fill_int(s->st, s->dcnt.p + 4, 1, 0);
  if (s->n_scenes > 0) launch_fill_int(s->st, s->scene_dirty.p, s->n_scenes, 0);
  // pairs that emitted no group (succ 0) count nothing
  launch_count_conflicts(s->st, s->adj_pairs, s->pa.p, s->pb.p, s->succ.p, s->rpc.p, s->color.p,
                         s->dcnt.p + 4,
                         s->n_scenes > 0 ? s->scene_of.p : nullptr,
                         s->n_scenes > 0 ? s->scene_dirty.p : nullptr);
}

int count_conflicts(avbd_solver* s, int* out) {
  launch_conflict_count(s);
  CK(readback(s, 4, 1));
  *out = s->hcnt[4];
  return AVBD_OK;
}

int build_csr(avbd_solver* s) {
  const int nb = s->nb, nd = s->n_contacts;
  const int kept = nd > 0 ? s->adj_pairs : 0;
  hipStream_t st = s->st;
  CK(prof_begin(s, kCsr));
  int* asb_cnt = s->tc_work.p;
  int* asb_groups = asb_cnt + nb;
  int* start_a = asb_groups + nb;
  int* end_a = start_a + nb;
  int* asb_cur = end_a + nb;
  launch_fill_int(st, s->tc_work.p, 5 * nb, 0);
  CK(s->asb_list.ensure((size_t)std::max(kept, 1)));
  launch_tcsr_pairs(st, kept, s->pa.p, s->pb.p, s->succ.p, asb_cnt, asb_groups, start_a, end_a);
  CK(scan(s, asb_cnt, s->asb_start.p, nb, s->asb_start.p + nb));
  launch_tcsr_asb_fill(st, kept, s->pb.p, s->succ.p, s->asb_start.p, asb_cur, s->asb_list.p);
  launch_tcsr_count(st, nb, s->cvals.p, s->st_count.p, asb_groups, start_a, end_a, s->sscan.p,
                    s->succ.p, s->padc.p, s->pos_of.p);
  CK(scan(s, s->padc.p, s->cpos.p, nb, s->cpos.p + nb));
  CK(s->tcsr.ensure((size_t)s->st_total + 2 * (size_t)nd + 3 * (size_t)nb + 1));
  launch_tcsr_fill(st, nb, s->cvals.p, s->st_start.p, s->st_idx.p, s->asb_start.p, s->asb_list.p,
                   start_a, end_a, s->sscan.p, s->succ.p, s->pa.p, s->pb.p, s->g_static, s->cpos.p,
                   s->tcsr.p);
  CK(prof_end(s, 0.0));
  s->csr_built = true;
  s->csr_order_stale = false;
  return AVBD_OK;
}

int emit_contacts(avbd_solver* s, const avbd_params* p, bool fixed, avbd_report* rep) {
  const int nb = s->nb;
  hipStream_t st = s->st;
  auto t0 = std::chrono::steady_clock::now();
  CK(prof_begin(s, kContacts));
  CK(ensure_pairs(s, std::max(256, 16 * nb)));
  float4* pc = s->pose[s->cur].p;
  launch_aabb(st, nb, pc, s->bstat.p, p->margin, s->lo.p, s->hi.p);
  // batched scenes: the tree is built over per-scene shifted boxes
  const bool batched = s->n_scenes > 0;
  const float4* tlo = s->lo.p;
  const float4* thi = s->hi.p;
  const int* scene = nullptr;
  if (batched) {
    launch_scene_shift(st, nb, s->lo.p, s->hi.p, s->scene_of.p, s->scene_grid, s->scene_pitch,
                       s->slo.p, s->shi.p);
    tlo = s->slo.p;
    thi = s->shi.p;
    scene = s->scene_of.p;
  }
  bvh_build(st, nb, tlo, thi, s->ws_bvh.p, s->ws_bvh.n);
  CK(s->pslots.ensure((size_t)nb * kPartnerSlots));
  bvh_count_pairs(st, nb, s->lo.p, s->hi.p, tlo, thi, scene, s->bstat.p, s->ws_bvh.p, s->bcount.p,
                  s->pslots.p, s->broad_lanes);
  CK(scan(s, s->bcount.p, s->boff.p, nb, s->dcnt.p + 0));
  CK(prof_end(s, 0.0));
  CK(readback(s, 0, 1));
  int total = s->hcnt[0];
  if (!fixed && total > s->pair_cap) {
    CK(ensure_pairs(s, std::max(2 * s->pair_cap, 2 * total)));
    rep->pair_grew = 1;
  }
  s->pair_count = total;
  rep->pair_count = total;
  rep->pair_capacity = s->pair_cap;
  int kept = std::min(total, s->pair_cap);
  s->adj_pairs = kept;
  if (fixed || total > 0) {
    CK(prof_begin(s, kContacts));
    bvh_write_pairs(st, nb, s->lo.p, s->hi.p, tlo, thi, scene, s->bstat.p, s->ws_bvh.p,
                    s->bcount.p, s->pslots.p, s->boff.p, s->pair_cap, s->pa.p, s->pb.p);
    launch_narrowphase(st, kept, pc, s->bstat.p, s->pa.p, s->pb.p, p->margin, s->pov.p, s->psat.p,
                       s->pn.p, s->pdepth.p, s->mfc.p, s->mfr.p, s->mfoff.p, s->mfn.p);
    launch_emit_demand(st, kept, s->pa.p, s->pb.p, s->mfc.p, s->mfr.p, s->bfric.p,
                       p->self_friction, s->demand.p, s->rpc.p);
    CK(scan(s, s->demand.p, s->dscan.p, kept, s->dcnt.p + 1));
    launch_emit_succ(st, kept, s->mfc.p, s->rpc.p, s->dscan.p, s->row_cap - s->n_static_rows,
                     s->succ.p, s->reserve.p);
    CK(scan(s, s->succ.p, s->sscan.p, kept, s->dcnt.p + 2));
    CK(scan(s, s->reserve.p, s->rscan.p, kept, s->dcnt.p + 3));
    launch_emit_write(st, kept, s->pa.p, s->pb.p, s->mfr.p, s->mfn.p, s->mfoff.p, s->bfric.p,
                      p->self_friction, p->friction_static_mult, s->rpc.p, s->succ.p, s->dscan.p,
                      s->sscan.p, s->n_static_rows, s->g_static, s->geo.p, s->gst.p, s->hash_live,
                      s->hkeys.p, s->hstate.p, s->ml_mode ? s->hage.p : nullptr);
    CK(prof_end(s, 0.0));
    if (kept > 0) {
      // the recolor gate's conflict count, on this substep's pairs and the
      // previous colors, read back with the emission totals (one sync)
      launch_conflict_count(s);
      s->conflicts_ready = true;
      CK(readback(s, 2, 3));
      s->n_contacts = s->hcnt[2];
      s->n_active_rows = s->n_static_rows + s->hcnt[3];
      s->pool_count = s->n_contacts;
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  rep->broadphase_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
  return AVBD_OK;
}

int live_hash_slots(const avbd_solver* s, int entries) {
  // load factor at most 1/4: short linear probes for the inserts' CAS
  const long long want = std::max(4096LL, 4LL * entries + 1);
  return (int)std::min<long long>(s->hash_cap, want);
}

int ensure_hage(avbd_solver* s) {
  if (s->ml_mode == 0 || !s->flushed) return AVBD_OK;
  const int hc = std::max(s->hash_cap, 1);
  if (s->hage.p && s->hage.n >= (size_t)hc) return AVBD_OK;
  CK(s->hage.alloc(hc));
  launch_fill_int(s->st, s->hage.p, hc, 0);
  return AVBD_OK;
}

// Features of the first n groups (static, then emitted; bodies at the start
// of the substep) and, in mode 2, the network's initial dual state for them.
int ml_begin(avbd_solver* s, const avbd_params* p, int n) {
  CK(s->ml_feat.ensure((size_t)n * 24));
  CK(s->ml_targ.ensure((size_t)n * 4));
  CK(s->ml_mask.ensure(n));
  CK(prof_begin(s, kOther));
  launch_contact_features(s->st, n, 0, s->geo.p, s->gst.p, s->pose[s->cur].p, s->vel.p,
                          s->bstat.p, s->ml_feat.p);
  if (s->ml_mode == 2)
    launch_mlp_infer(s->st, n, s->ml_feat.p, s->ml_params.p, 0, s->gst.p, s->geo.p,
                     p->alpha * p->gamma, p->gamma, s->ml_scope == 1 ? 2 : 0, nullptr);
  CK(prof_end(s, (double)n * (96.0 + 48.0 + 96.0 + (s->ml_mode == 2 ? 96.0 + 48.0 : 0.0)),
              s->ml_mode == 2 ? 2 : 1));
  return AVBD_OK;
}

// tol >= 0: iterate until the dual's residual max |lambda_new - lambda| / k
// is <= tol, with p->iterations as the cap (one readback per iteration); the
// result equals a fixed-iteration substep with that many iterations.
// The stages of a direct substep, in the reference's launch order
// (solver_6dof.py:1269-1464): contacts, coloring, body->group lists,
// prelude, (primal sweep, dual) x iterations with the finalize after the
// last dual, post-stabilisation sweep, cache refresh. Each stage leaves what
// the next needs in s->sx.
enum { kStContacts = 0, kStColoring, kStCsr, kStPrelude, kStSweeps, kStEnd };
int primal_bpw(int variant);

int stage_contacts(avbd_solver* s, const avbd_params* p, avbd_report* rep) {
  std::memset(rep, 0, sizeof(*rep));
  rep->n_active_colors = s->n_active_colors;
  auto& x = s->sx;
  x = avbd_solver::StageCtx{};
  const int nb = s->nb;
  const bool fixed = p->fixed_capacity != 0;
  s->n_active_rows = s->n_static_rows;
  s->pool_count = 0;
  s->pair_count = 0;
  s->n_contacts = 0;
  s->adj_pairs = 0;
  s->conflicts_ready = false;
  x.n_active = s->n_static_rows;
  x.pool_reset = false;  // the pool grew this substep (see the cache refresh)
  rep->row_capacity = s->row_cap;
  if (p->self_collide && nb >= 2) {
    CK(emit_contacts(s, p, fixed, rep));
    if (fixed) {
      x.n_active = s->row_cap;
    } else {
      int raw = s->n_active_rows, cap = s->row_cap;
      if (raw > cap) {
        rep->row_overflow = 1;
        const int pool_before = s->pool_max;
        CK(grow_row_pool(s, std::max(2 * s->n_dyn_cap, (raw - s->n_static_rows) * 2),
                         s->g_static + s->n_contacts));
        x.pool_reset = s->pool_max > pool_before;
      }
      x.n_active = std::min(raw, cap);
    }
  }
  // learned warm start: features of the emitted contact groups, then either
  // the network's initial dual state (mode 2) or nothing more until the
  // targets at the end of the substep (mode 1)
  // groups: static (floor, pins) first, then this substep's box-box contacts
  x.n_ml = s->ml_mode == 0 ? 0
           : s->g_static + (p->self_collide ? std::min(s->n_contacts, s->pool_max) : 0);
  s->ml_n = 0;
  if (x.n_ml > 0) CK(ml_begin(s, p, x.n_ml));
  rep->n_active_rows = s->n_active_rows;
  rep->new_row_capacity = s->row_cap;
  rep->pool_count = s->pool_count;
  rep->pair_capacity = s->pair_cap;
  rep->n_contacts = std::min(s->n_contacts, s->pool_max);
  x.stage = kStColoring;
  return AVBD_OK;
}

int stage_coloring(avbd_solver* s, const avbd_params* p, int color_dirty, avbd_report* rep) {
  auto& x = s->sx;
  hipStream_t st = s->st;
  bool need = color_dirty != 0;
  if (!need) {
    if (s->conflicts_ready) {
      rep->color_conflicts = s->hcnt[4];  // counted with the emission
    } else if (s->adj_pairs == 0) {
      rep->color_conflicts = 0;
      if (s->n_scenes > 0) launch_fill_int(st, s->scene_dirty.p, s->n_scenes, 0);
    } else {
      CK(count_conflicts(s, &rep->color_conflicts));
    }
    need = rep->color_conflicts > 0;
  } else if (s->n_scenes > 0) {
    launch_fill_int(st, s->scene_dirty.p, s->n_scenes, 1);  // every scene recolors
  }
  if (need) {
    CK(recolor(s, p->coloring_mode));
    rep->recolored = 1;
  } else {
    s->colors_conflict_free = true;  // the conflict count just read is 0
  }
  rep->n_active_colors = s->n_active_colors;
  x.recolor = need;
  x.stage = kStCsr;
  (void)p;
  return AVBD_OK;
}

int stage_csr(avbd_solver* s) {
  auto& x = s->sx;
  // Without live rows the reference keeps the previous body->constraint
  // lists (solver_6dof.py:1366); if the color order changed since they were
  // built, map each color position to the list's old position.
  x.pos_map = nullptr;
  if (x.recolor) s->csr_order_stale = true;
  if (x.n_active > 0) {
    CK(build_csr(s));
  } else if (s->csr_built && s->csr_order_stale) {
    launch_tcsr_posmap(s->st, s->nb, s->cvals.p, s->pos_of.p, s->pos_map.p);
    x.pos_map = s->pos_map.p;
  }
  x.stage = kStPrelude;
  return AVBD_OK;
}

int stage_prelude(avbd_solver* s, const avbd_params* p, float tol) {
  auto& x = s->sx;
  const int nb = s->nb;
  x.n_groups = s->g_static + s->n_contacts;
  x.post = p->post_stabilize ? 1 : 0;
  CK(prof_begin(s, kPrelude));
  launch_substep_begin(s->st, nb, x.n_groups, s->pose[s->cur].p, s->pose[1 - s->cur].p,
                       s->poseI.p, s->inert.p, s->vel.p, s->pvel.p, s->bstat.p, s->Ilocal.p,
                       s->geo.p, s->gst.p, p->dt, p->gravity[0], p->gravity[1], p->gravity[2],
                       p->alpha, p->gamma, x.post, x.n_active > 0 ? 1 : 0,
                       reinterpret_cast<unsigned*>(s->dcnt.p + kEpochSlot));
  CK(prof_end(s, (double)nb * 188.0 + (double)s->n_contacts * 124.0 + (double)s->n_floor_groups * 88.0));
  x.max_lin = p->max_linear_speed;
  x.max_ang = p->max_angular_speed;
  if (!(x.max_lin > 0.0f && x.max_lin < kInf && x.max_ang > 0.0f && x.max_ang < kInf)) {
    x.max_lin = kInf;
    x.max_ang = kInf;
  }
  x.iterations = p->iterations;
  x.tol = tol;
  x.resid = 0.0f;
  const bool to_resid = tol >= 0.0f;
  x.nonempty = 0;
  for (int c = 0; c < s->n_active_colors; ++c) x.nonempty += (s->cstart[c + 1] > s->cstart[c]);
  x.pbytes = primal_sweep_bytes(s) / std::max(1, x.nonempty);
  // variants 6-10: every color of a sweep in one dataflow launch
  x.flow = s->primal_variant >= 6 && x.pos_map == nullptr && x.iterations + x.post < (int)4095;
  x.flow_bpw = primal_bpw(s->primal_variant);
  // the dual rides along in the sweep when no two neighbours share a color
  x.fuse = x.flow && s->colors_conflict_free && s->fuse_dual_mode == 1 && !to_resid;
  // or in trailing workgroups of the sweep's launch (any coloring)
  x.tail = x.flow && s->fuse_dual_mode == 2 && !to_resid;
  x.sweeps_done = 0;
  x.finalized = false;
  x.stage = kStSweeps;
  return AVBD_OK;
}

// Sweep `it` of the substep: the primal sweep, then (for it < iterations)
// the dual update unless it rode along in the sweep, then the finalize
// after the last dual.
int stage_sweep(avbd_solver* s, const avbd_params* p, int it) {
  auto& x = s->sx;
  hipStream_t st = s->st;
  const int I = x.iterations;
  const float dt = p->dt;
  float4* pw = s->pose[1 - s->cur].p;
  const int post_pass = (x.post && it == I) ? 1 : 0;
  // One timing interval per sweep (all colors back to back): per-launch
  // events would add their own gaps between the color launches.
  CK(prof_begin(s, kPrimal));
  int launched = 0;
  if (x.flow) {
    // A coloring with same-color neighbours (more than 32 colors needed):
    // one body per tile, each waiting for its same-color neighbours of lower
    // index, which is the serial order within a color.
    const bool ordered = !s->colors_conflict_free;
    launch_primal_sweep(st, s->n_active_colors, s->cstart, s->cvals.p, s->color.p, pw,
                        s->poseI.p, s->inert.p, s->bstat.p, s->cpos.p, s->tcsr.p, s->geo.p,
                        s->gst.p, dt, post_pass, s->flow_done.p,
                        reinterpret_cast<const unsigned*>(s->dcnt.p + kEpochSlot), it,
                        s->dcnt.p + kErrSlot, ordered ? 1 : x.flow_bpw,
                        x.fuse && it < I ? 1 : 0, p->beta, s->gst.p,
                        x.tail && it < I ? x.n_groups : 0,
                        s->flow_test_cap > 0 ? s->flow_never.p : nullptr, s->flow_test_cap,
                        nullptr, nullptr, nullptr, 0, nullptr, ordered);
    launched = x.nonempty;  // bodies of every color, in one launch
    CK(prof_end(s, x.pbytes * launched + (x.tail && it < I ? dual_sweep_bytes(s) : 0.0), 1));
  } else {
    for (int c = 0; c < s->n_active_colors; ++c) {
      int cnt = s->cstart[c + 1] - s->cstart[c];
      if (cnt <= 0) continue;
      launch_primal_color(st, cnt, s->cvals.p, s->cstart[c], pw, s->poseI.p, s->inert.p,
                          s->bstat.p, s->cpos.p, s->tcsr.p, x.pos_map, s->geo.p, s->gst.p, dt,
                          post_pass, s->primal_variant);
      ++launched;
    }
    CK(prof_end(s, x.pbytes * launched, launched));
  }
  if (it < I && !x.fuse && !x.tail) {
    const bool to_resid = x.tol >= 0.0f;
    float* rd = to_resid ? reinterpret_cast<float*>(s->dcnt.p + 10) : nullptr;
    if (rd) launch_fill_int(st, s->dcnt.p + 10, 1, 0);
    CK(prof_begin(s, kDual));
    launch_dual(st, x.n_groups, pw, s->geo.p, s->gst.p, p->beta, rd);
    CK(prof_end(s, dual_sweep_bytes(s)));
    if (rd) {
      CK(readback(s, 10, 1));
      std::memcpy(&x.resid, s->hcnt + 10, sizeof(float));
      if (x.resid <= x.tol) x.iterations = it + 1;  // this was the last dual
    }
  }
  if (it == x.iterations - 1) {
    CK(prof_begin(s, kOther));
    launch_finalize(st, s->nb, pw, s->pose[s->cur].p, s->vel.p, s->pvel.p, s->bstat.p, dt,
                    x.max_lin, x.max_ang);
    CK(prof_end(s, (double)s->nb * 132.0));
    x.finalized = true;
  }
  x.sweeps_done = it + 1;
  return AVBD_OK;
}

int stage_end(avbd_solver* s, const avbd_params* p) {
  auto& x = s->sx;
  hipStream_t st = s->st;
  if (x.n_ml > 0 && s->ml_mode == 1) {
    CK(prof_begin(s, kOther));
    launch_contact_targets(st, x.n_ml, 0, s->geo.p, s->gst.p, s->ml_feat.p,
                           p->alpha * p->gamma, p->gamma, s->ml_targ.p, s->ml_mask.p);
    CK(prof_end(s, (double)x.n_ml * (96.0 + 48.0 + 16.0 + 20.0)));
    s->ml_n = x.n_ml;
  }
  if (p->self_collide && s->hash_cap > 0) {
    CK(prof_begin(s, kOther));
    int entries = std::min(s->n_contacts, s->pool_max);
    s->hash_live = live_hash_slots(s, x.pool_reset ? 1 : entries);
    launch_fill_int(st, s->dcnt.p + kHashLiveSlot, 1, s->hash_live);
    launch_fill_int(st, s->hkeys.p, s->hash_live, kHashEmpty);
    launch_fill_int(st, s->howner.p, s->hash_live, -1);
    if (!x.pool_reset) {
      launch_hash_collect(st, entries, s->g_static, s->geo.p, s->gst.p, s->hash_live, s->hkeys.p,
                          s->howner.p, s->slot_of.p, s->hstate.p,
                          s->ml_mode ? s->hage.p : nullptr);
    } else if (entries > 0) {
      // the reference's refresh after a pool growth reads zeroed pool indices
      launch_hash_row0(st, s->geo.p, s->gst.p, s->hash_live, s->hkeys.p, s->hstate.p,
                       s->ml_mode ? s->hage.p : nullptr);
    }
    CK(prof_end(s, 0.0));
  }
  HIPCK(hipMemcpyAsync(s->hcnt + kErrSlot, s->dcnt.p + kErrSlot, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCK(hipGetLastError());
  s->cur = 1 - s->cur;
  s->last_post = x.post;
  CK(prof_collect(s));
  x.stage = kStContacts;
  return AVBD_OK;
}

int do_substep(avbd_solver* s, const avbd_params* p, int color_dirty, avbd_report* rep,
               float tol = -1.0f, int* iters_used = nullptr, float* resid_out = nullptr) {
  std::memset(rep, 0, sizeof(*rep));
  rep->n_active_colors = s->n_active_colors;
  if (s->nb == 0) return AVBD_OK;
  CK(stage_contacts(s, p, rep));
  CK(stage_coloring(s, p, color_dirty, rep));
  CK(stage_csr(s));
  CK(stage_prelude(s, p, tol));
  for (int it = 0; it < s->sx.iterations + s->sx.post; ++it) CK(stage_sweep(s, p, it));
  if (iters_used) *iters_used = s->sx.iterations;
  if (resid_out) *resid_out = s->sx.resid;
  return stage_end(s, p);
}

// ===========================================================================
// Substeps with every count and decision on the device
// ===========================================================================
// The substep of do_substep in the fixed-capacity mode, issued without any
// host read: launches are sized by capacity and each kernel takes its count
// from the solver's device counters; the recolor gate is a count that is 0
// when the coloring is kept; the sweep takes the color layout from the
// device. With the same inputs it performs the same operations in the same
// order, so it gives the same bits, and it can be captured as a HIP graph.

// Kept pairs, live groups and cache entries after the emission.
__global__ void k_after_emission(int* d, int pair_cap, int g_static, int pool_max) {
  if (threadIdx.x != 0) return;
  const int pairs = d[kPairSlot], groups = d[kGroupsSlot];
  d[kKeptSlot] = pairs < pair_cap ? pairs : pair_cap;
  d[kNGroupsSlot] = g_static + groups;
  d[kEntriesSlot] = groups < pool_max ? groups : pool_max;
}

// Kept pairs before the emission (the narrowphase and emission kernels'
// count).
__global__ void k_kept_pairs(int* d, int pair_cap) {
  if (threadIdx.x != 0) return;
  const int pairs = d[kPairSlot];
  d[kKeptSlot] = pairs < pair_cap ? pairs : pair_cap;
}

// The recolor gate (solver_6dof.py:1333-1356 with the coloring clean): a
// conflict row opens it.
__global__ void k_recolor_gate(int* d) {
  if (threadIdx.x != 0) return;
  const bool need = d[kConflictSlot] > 0;
  d[kGateAllSlot] = need ? 0x7fffffff : 0;
  d[kGateKeptSlot] = need ? d[kKeptSlot] : 0;
  if (need) d[kRecolorsSlot] += 1;
}

int primal_bpw(int variant) {
  return variant == 7 ? 1 : variant == 8 ? 4 : variant == 9 ? 8 : variant == 10 ? 3 : 2;
}

// What a graph-captured substep needs: the fixed-capacity semantics, the
// dataflow sweep with a separate dual or the dual in trailing workgroups of
// its launch, rows to solve, no learned warm start
// and no profiling. Buffers a direct substep would size on the fly are sized
// to capacity first (this allocates, so it runs before any capture).
int graph_eligible(avbd_solver* s, const avbd_params* p, bool* ok) {
  *ok = false;
  if (!s->graph_on || s->nb == 0 || p->fixed_capacity == 0 || s->ml_mode != 0 || s->prof ||
      s->primal_variant < 6 || s->primal_variant > 10 || s->fuse_dual_mode == 1 ||
      s->flow_test_cap != 0 || s->row_cap <= 0 || p->iterations + p->post_stabilize >= 4095)
    return AVBD_OK;
  const bool emit = p->self_collide && s->nb >= 2;
  if (emit) CK(ensure_pairs(s, std::max(256, 16 * s->nb)));
  const int pcap = std::max(s->pair_cap, 1);
  CK(s->nbr.ensure(2 * (size_t)pcap + 1));
  CK(s->asb_list.ensure((size_t)pcap));
  CK(s->tcsr.ensure((size_t)s->st_total + 2 * (size_t)(s->g_cap - s->g_static) + 3 * (size_t)s->nb + 1));
  CK(ensure_scan_ws(s, std::max<size_t>((size_t)pcap, (size_t)s->nb)));
  CK(s->pslots.ensure((size_t)s->nb * kPartnerSlots));
  *ok = true;
  return AVBD_OK;
}

void substep_device(avbd_solver* s, const avbd_params* p) {
  const int nb = s->nb;
  hipStream_t st = s->st;
  int* d = s->dcnt.p;
  const bool emit = p->self_collide && nb >= 2;
  const int pcap = s->pair_cap;
  const bool batched = s->n_scenes > 0;
  const int* kept = d + kKeptSlot;
  // ---- contacts (emit_contacts) ----
  if (emit) {
    float4* pc = s->pose[s->cur].p;
    launch_aabb(st, nb, pc, s->bstat.p, p->margin, s->lo.p, s->hi.p);
    const float4* tlo = s->lo.p;
    const float4* thi = s->hi.p;
    const int* scene = nullptr;
    if (batched) {
      launch_scene_shift(st, nb, s->lo.p, s->hi.p, s->scene_of.p, s->scene_grid, s->scene_pitch,
                         s->slo.p, s->shi.p);
      tlo = s->slo.p;
      thi = s->shi.p;
      scene = s->scene_of.p;
    }
    bvh_build(st, nb, tlo, thi, s->ws_bvh.p, s->ws_bvh.n);
    bvh_count_pairs(st, nb, s->lo.p, s->hi.p, tlo, thi, scene, s->bstat.p, s->ws_bvh.p,
                    s->bcount.p, s->pslots.p, s->broad_lanes);
    (void)scan(s, s->bcount.p, s->boff.p, nb, d + kPairSlot);
    hipLaunchKernelGGL(k_kept_pairs, dim3(1), dim3(64), 0, st, d, pcap);
    bvh_write_pairs(st, nb, s->lo.p, s->hi.p, tlo, thi, scene, s->bstat.p, s->ws_bvh.p,
                    s->bcount.p, s->pslots.p, s->boff.p, pcap, s->pa.p, s->pb.p);
    launch_narrowphase(st, pcap, pc, s->bstat.p, s->pa.p, s->pb.p, p->margin, s->pov.p,
                       s->psat.p, s->pn.p, s->pdepth.p, s->mfc.p, s->mfr.p, s->mfoff.p, s->mfn.p,
                       kept);
    launch_emit_demand(st, pcap, s->pa.p, s->pb.p, s->mfc.p, s->mfr.p, s->bfric.p,
                       p->self_friction, s->demand.p, s->rpc.p, kept);
    (void)scan(s, s->demand.p, s->dscan.p, pcap, d + 1, kept);
    launch_emit_succ(st, pcap, s->mfc.p, s->rpc.p, s->dscan.p, s->row_cap - s->n_static_rows,
                     s->succ.p, s->reserve.p, kept);
    (void)scan(s, s->succ.p, s->sscan.p, pcap, d + kGroupsSlot, kept);
    (void)scan(s, s->reserve.p, s->rscan.p, pcap, d + kRowsSlot, kept);
    launch_emit_write(st, pcap, s->pa.p, s->pb.p, s->mfr.p, s->mfn.p, s->mfoff.p, s->bfric.p,
                      p->self_friction, p->friction_static_mult, s->rpc.p, s->succ.p, s->dscan.p,
                      s->sscan.p, s->n_static_rows, s->g_static, s->geo.p, s->gst.p, s->hash_cap,
                      s->hkeys.p, s->hstate.p, nullptr, kept, d + kHashLiveSlot);
    // the recolor gate's conflict count
    launch_fill_int(st, d + kConflictSlot, 1, 0);
    if (batched) launch_fill_int(st, s->scene_dirty.p, s->n_scenes, 0);
    launch_count_conflicts(st, pcap, s->pa.p, s->pb.p, s->succ.p, s->rpc.p, s->color.p,
                           d + kConflictSlot, batched ? s->scene_of.p : nullptr,
                           batched ? s->scene_dirty.p : nullptr, kept);
  } else {
    launch_fill_int(st, d, kConflictSlot + 1, 0);
    if (batched) launch_fill_int(st, s->scene_dirty.p, s->n_scenes, 0);
  }
  hipLaunchKernelGGL(k_after_emission, dim3(1), dim3(64), 0, st, d, std::max(pcap, 0), s->g_static,
                     s->pool_max);
  // ---- recolor (recolor), gated ----
  hipLaunchKernelGGL(k_recolor_gate, dim3(1), dim3(64), 0, st, d);
  const int* gA = d + kGateAllSlot;
  const int* gK = d + kGateKeptSlot;
  const int* scene = batched ? s->scene_of.p : nullptr;
  const int* dirty = batched ? s->scene_dirty.p : nullptr;
  launch_fill_int(st, s->nbr_count.p, nb, 0, gA);
  launch_adj_count_pairs(st, pcap, s->pa.p, s->pb.p, s->succ.p, s->nbr_count.p, gK);
  (void)scan(s, s->nbr_count.p, s->nbr_start.p, nb, s->nbr_start.p + nb, gA);
  launch_fill_int(st, s->nbr_cursor.p, nb, 0, gA);
  launch_adj_fill_pairs(st, pcap, s->pa.p, s->pb.p, s->succ.p, s->nbr_start.p, s->nbr_cursor.p,
                        s->nbr.p, gK);
  launch_color_reset(st, nb, s->color.p, scene, dirty, gA);
  launch_fill_int(st, d + 6, 1, 0);
  launch_fill_int(st, d + 9, 1, 0);
  if (p->coloring_mode == AVBD_COLOR_JACOBI) {
    launch_color_greedy(st, nb, s->order_id.p, nullptr, s->nbr_start.p, s->nbr.p, s->color.p,
                        s->pub.p, d + 6, s->ticket.p, 0, d + 9, scene, dirty, gA);
    launch_fill_int(st, d + kUncoloredSlot, 1, 0);
    launch_color_resolve(st, nb, s->nbr_start.p, s->nbr.p, s->prio.p, s->color.p,
                         s->color_next.p, d + kUncoloredSlot, gA);
    launch_copy_int(st, s->color.p, s->color_next.p, nb, gA);
    launch_color_jacobi_tail(st, nb, s->nbr_start.p, s->nbr.p, s->prio.p, s->color.p,
                             s->color_next.p, d + kUncoloredSlot, d + 9, gA);
  } else {
    launch_color_greedy(st, nb, s->order_jp.p, s->pos_jp.p, s->nbr_start.p, s->nbr.p, s->color.p,
                        s->pub.p, d + 6, s->ticket.p, 1, d + 9, scene, dirty, gA);
  }
  launch_color_finish(st, nb, s->color.p, d + kCountsSlot, gA);
  color_bucket(st, nb, s->color.p, s->ckeys.p, s->cvals.p, s->ckeys_tmp.p, s->cvals_tmp.p,
               s->ws_color.p, gA);
  const int bpw = primal_bpw(s->primal_variant);
  launch_color_meta(st, d + kCountsSlot, bpw, d + kCstartSlot, d + kNColorsSlot, d + kTilesSlot);
  // ---- body -> group lists (build_csr) ----
  int* asb_cnt = s->tc_work.p;
  int* asb_groups = asb_cnt + nb;
  int* start_a = asb_groups + nb;
  int* end_a = start_a + nb;
  int* asb_cur = end_a + nb;
  launch_fill_int(st, s->tc_work.p, 5 * nb, 0);
  launch_tcsr_pairs(st, pcap, s->pa.p, s->pb.p, s->succ.p, asb_cnt, asb_groups, start_a, end_a,
                    kept);
  (void)scan(s, asb_cnt, s->asb_start.p, nb, s->asb_start.p + nb);
  launch_tcsr_asb_fill(st, pcap, s->pb.p, s->succ.p, s->asb_start.p, asb_cur, s->asb_list.p, kept);
  launch_tcsr_count(st, nb, s->cvals.p, s->st_count.p, asb_groups, start_a, end_a, s->sscan.p,
                    s->succ.p, s->padc.p, s->pos_of.p);
  (void)scan(s, s->padc.p, s->cpos.p, nb, s->cpos.p + nb);
  launch_tcsr_fill(st, nb, s->cvals.p, s->st_start.p, s->st_idx.p, s->asb_start.p, s->asb_list.p,
                   start_a, end_a, s->sscan.p, s->succ.p, s->pa.p, s->pb.p, s->g_static, s->cpos.p,
                   s->tcsr.p);
  // ---- predict + prelude, iterations, finalize ----
  float4* p0 = s->pose[s->cur].p;
  float4* pw = s->pose[1 - s->cur].p;
  const float dt = p->dt;
  const int post = p->post_stabilize ? 1 : 0;
  const int* ngroups = d + kNGroupsSlot;
  launch_substep_begin(st, nb, s->g_cap, p0, pw, s->poseI.p, s->inert.p, s->vel.p, s->pvel.p,
                       s->bstat.p, s->Ilocal.p, s->geo.p, s->gst.p, dt, p->gravity[0],
                       p->gravity[1], p->gravity[2], p->alpha, p->gamma, post, 1,
                       reinterpret_cast<unsigned*>(d + kEpochSlot), ngroups);
  float max_lin = p->max_linear_speed, max_ang = p->max_angular_speed;
  if (!(max_lin > 0.0f && max_lin < kInf && max_ang > 0.0f && max_ang < kInf)) {
    max_lin = kInf;
    max_ang = kInf;
  }
  const int I = p->iterations;
  const int max_tiles = (nb + bpw - 1) / bpw + kMaxColors;
  // the dual in trailing workgroups of the sweep's launch (sized by the group
  // capacity, the live count read on the device), or a launch of its own
  const bool tail = s->fuse_dual_mode == 2;
  for (int it = 0; it < I + post; ++it) {
    const int post_pass = (post && it == I) ? 1 : 0;
    launch_primal_sweep(st, 0, s->cstart, s->cvals.p, s->color.p, pw, s->poseI.p, s->inert.p,
                        s->bstat.p, s->cpos.p, s->tcsr.p, s->geo.p, s->gst.p, dt, post_pass,
                        s->flow_done.p, reinterpret_cast<const unsigned*>(d + kEpochSlot), it,
                        d + kErrSlot, bpw, 0, p->beta, s->gst.p, tail && it < I ? s->g_cap : 0,
                        nullptr, 0, d + kNColorsSlot, d + kCstartSlot, d + kTilesSlot, max_tiles,
                        ngroups);
    if (it < I && !tail)
      launch_dual(st, s->g_cap, pw, s->geo.p, s->gst.p, p->beta, nullptr, ngroups);
    if (it == I - 1)
      launch_finalize(st, nb, pw, p0, s->vel.p, s->pvel.p, s->bstat.p, dt, max_lin, max_ang);
  }
  // ---- warm-start cache refresh ----
  if (p->self_collide && s->hash_cap > 0) {
    launch_hash_size(st, d + kEntriesSlot, s->hash_cap, d + kHashLiveSlot);
    launch_fill_int(st, s->hkeys.p, s->hash_cap, kHashEmpty, d + kHashLiveSlot);
    launch_fill_int(st, s->howner.p, s->hash_cap, -1, d + kHashLiveSlot);
    launch_hash_collect(st, std::max(s->pool_max, 1), s->g_static, s->geo.p, s->gst.p,
                        s->hash_cap, s->hkeys.p, s->howner.p, s->slot_of.p, s->hstate.p, nullptr,
                        d + kEntriesSlot, d + kHashLiveSlot);
  }
}

// Everything a captured substep depends on besides device memory contents.
std::vector<char> graph_key(const avbd_solver* s, const avbd_params* p) {
  std::vector<char> k;
  auto put = [&k](const void* v, size_t n) {
    const char* c = static_cast<const char*>(v);
    k.insert(k.end(), c, c + n);
  };
  put(p, sizeof(*p));
  const long long ints[] = {s->nb, s->cur, s->n_static_rows, s->g_static, s->g_cap, s->row_cap,
                            s->pair_cap, s->pool_max, s->hash_cap, s->st_total,
                            s->primal_variant, s->n_scenes, s->scene_grid,
                            s->fuse_dual_mode, s->broad_lanes, (long long)g_alloc_gen};
  put(ints, sizeof(ints));
  put(&s->scene_pitch, sizeof(s->scene_pitch));
  const void* ptrs[] = {s->color.p, s->color_next.p, s->pose[0].p, s->pose[1].p};
  put(ptrs, sizeof(ptrs));
  return k;
}

void drop_graphs(avbd_solver* s) {
  for (int c = 0; c < 2; ++c) {
    if (s->gexec[c]) (void)hipGraphExecDestroy(s->gexec[c]);
    s->gexec[c] = nullptr;
    s->gkey[c].clear();
  }
}

// One substep through the graph of the current pose buffer, captured first
// when it is missing or stale.
int substep_graph(avbd_solver* s, const avbd_params* p) {
  const int c = s->cur;
  std::vector<char> key = graph_key(s, p);
  if (!s->gexec[c] || s->gkey[c] != key) {
    if (s->gexec[c]) (void)hipGraphExecDestroy(s->gexec[c]);
    s->gexec[c] = nullptr;
    hipGraph_t g = nullptr;
    // captured on the solver's own stream (a caller's stream may be the null
    // stream, which cannot capture) and launched on the current one
    const hipStream_t launch_st = s->st;
    s->st = s->own_st;
    const hipError_t e_begin = hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal);
    if (e_begin != hipSuccess) {
      s->st = launch_st;
      return set_err(AVBD_E_HIP, std::string("substep capture: ") + hipGetErrorString(e_begin));
    }
    substep_device(s, p);
    const hipError_t e_end = hipStreamEndCapture(s->st, &g);
    s->st = launch_st;
    if (e_end != hipSuccess || g == nullptr)
      return set_err(AVBD_E_HIP, std::string("substep capture failed: ") + hipGetErrorString(e_end));
    const hipError_t e_inst = hipGraphInstantiate(&s->gexec[c], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e_inst != hipSuccess) {
      s->gexec[c] = nullptr;
      return set_err(AVBD_E_HIP, std::string("substep graph: ") + hipGetErrorString(e_inst));
    }
    s->gkey[c] = key;
  }
  HIPCK(hipGraphLaunch(s->gexec[c], s->st));
  ++s->graph_launches;
  s->cur = 1 - c;
  s->last_post = p->post_stabilize ? 1 : 0;
  return AVBD_OK;
}

// The host's view after device-decided substeps: one copy of the counters.
int frame_end(avbd_solver* s, avbd_report* rep, int recolored_direct) {
  HIPCK(hipMemcpyAsync(s->hcnt, s->dcnt.p, kCounterSlots * sizeof(int), hipMemcpyDeviceToHost,
                       s->st));
  HIPCK(hipStreamSynchronize(s->st));
  ++s->syncs;
  CK(check_flow_error(s));
  const int* h = s->hcnt;
  if (h[6] != 0) return set_err(AVBD_E_STATE, "coloring dataflow did not converge");
  s->pair_count = h[kPairSlot];
  s->adj_pairs = std::min(h[kPairSlot], s->pair_cap);
  s->n_contacts = h[kGroupsSlot];
  s->pool_count = s->n_contacts;
  s->n_active_rows = s->n_static_rows + h[kRowsSlot];
  s->hash_live = h[kHashLiveSlot] > 0 ? h[kHashLiveSlot] : s->hash_live;
  s->n_active_colors = std::max(1, std::min(h[kNColorsSlot], kMaxColors));
  for (int c = 0; c <= kMaxColors; ++c) s->cstart[c] = h[kCstartSlot + c];
  s->colors_conflict_free = h[9] == 0;
  s->csr_built = true;
  s->csr_order_stale = false;
  const int recolors = h[kRecolorsSlot];
  std::memset(rep, 0, sizeof(*rep));
  rep->recolored = (recolors != s->recolors_seen || recolored_direct) ? 1 : 0;
  s->recolors_seen = recolors;
  rep->n_active_colors = s->n_active_colors;
  rep->n_active_rows = s->n_active_rows;
  rep->row_capacity = s->row_cap;
  rep->new_row_capacity = s->row_cap;
  rep->pair_count = s->pair_count;
  rep->pair_capacity = s->pair_cap;
  rep->pool_count = s->pool_count;
  rep->n_contacts = std::min(s->n_contacts, s->pool_max);
  rep->color_conflicts = h[kConflictSlot];
  return AVBD_OK;
}

bool isinf_pos(float v) { return v == kInf; }
bool isinf_neg(float v) { return v == -kInf; }

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* avbd_last_error(void) { return g_err.c_str(); }

int avbd_device_count(int* out) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *out = (e == hipSuccess) ? n : 0;
  return AVBD_OK;
}

int avbd_create(int device, avbd_solver** out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return set_err(AVBD_E_HIP, "no HIP device available");
  if (device < 0 || device >= n) return set_err(AVBD_E_ARG, "device ordinal out of range");
  HIPCK(hipSetDevice(device));
  avbd_solver* s = new avbd_solver();
  s->device = device;
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return set_err(AVBD_E_HIP, "hipStreamCreate failed");
  }
  s->own_st = s->st;
  if (hipHostMalloc((void**)&s->hcnt, kDcntSize * sizeof(int), hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(s->st);
    delete s;
    return set_err(AVBD_E_NOMEM, "hipHostMalloc failed");
  }
  std::memset(s->hcnt, 0, kDcntSize * sizeof(int));
  int r = s->dcnt.alloc(kDcntSize);
  if (r != AVBD_OK) {
    (void)hipHostFree(s->hcnt);
    (void)hipStreamDestroy(s->st);
    delete s;
    return r;
  }
  launch_fill_int(s->st, s->dcnt.p, kDcntSize, 0);
  *out = s;
  return AVBD_OK;
}

int avbd_destroy(avbd_solver* s) {
  if (!s) return AVBD_OK;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->st);
  drop_graphs(s);
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  if (s->hcnt) (void)hipHostFree(s->hcnt);
  (void)hipStreamDestroy(s->own_st);
  delete s;
  return AVBD_OK;
}

int avbd_flush(avbd_solver* s, const avbd_scene* sc) {
  if (!s || !sc) return set_err(AVBD_E_ARG, "null argument");
  HIPCK(hipSetDevice(s->device));
  const int nb = sc->n_bodies, ns = sc->n_static_rows;
  if (nb < 0 || ns < 0 || sc->n_dyn_capacity < 0 || sc->hash_cap < 0 || sc->pool_max < 0 ||
      sc->max_pairs < 0)
    return set_err(AVBD_E_ARG, "negative size in scene");
  HIPCK(hipStreamSynchronize(s->st));
  drop_graphs(s);
  s->nb = nb;
  s->cur = 0;
  const size_t nbb = std::max(nb, 1);
  // ---- bodies ----
  std::vector<float4> pose(2 * nbb), vel(2 * nbb), pvel(2 * nbb), bst(nbb);
  for (int i = 0; i < nb; ++i) {
    pose[2 * i] = make_float4(sc->x[3 * i], sc->x[3 * i + 1], sc->x[3 * i + 2], 0.f);
    pose[2 * i + 1] = make_float4(sc->q[4 * i], sc->q[4 * i + 1], sc->q[4 * i + 2], sc->q[4 * i + 3]);
    vel[2 * i] = make_float4(sc->v[3 * i], sc->v[3 * i + 1], sc->v[3 * i + 2], 0.f);
    vel[2 * i + 1] = make_float4(sc->w[3 * i], sc->w[3 * i + 1], sc->w[3 * i + 2], 0.f);
    pvel[2 * i] = make_float4(sc->prev_v[3 * i], sc->prev_v[3 * i + 1], sc->prev_v[3 * i + 2], 0.f);
    pvel[2 * i + 1] = make_float4(sc->prev_w[3 * i], sc->prev_w[3 * i + 1], sc->prev_w[3 * i + 2], 0.f);
    bst[i] = make_float4(sc->mass[i], sc->half_extents[3 * i], sc->half_extents[3 * i + 1],
                         sc->half_extents[3 * i + 2]);
  }
  CK(s->pose[0].alloc(2 * nbb)); CK(s->pose[1].alloc(2 * nbb));
  CK(s->poseI.alloc(2 * nbb)); CK(s->inert.alloc(3 * nbb)); CK(s->vel.alloc(2 * nbb));
  CK(s->pvel.alloc(2 * nbb)); CK(s->bstat.alloc(nbb)); CK(s->lo.alloc(nbb)); CK(s->hi.alloc(nbb));
  CK(s->Ilocal.alloc(9 * nbb)); CK(s->bfric.alloc(nbb)); CK(s->prio.alloc(nbb));
  CK(s->order_id.alloc(nbb)); CK(s->order_jp.alloc(nbb)); CK(s->pos_jp.alloc(nbb));
  if (nb > 0) {
    HIPCK(hipMemcpy(s->pose[0].p, pose.data(), 2 * nb * sizeof(float4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->pose[1].p, pose.data(), 2 * nb * sizeof(float4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->vel.p, vel.data(), 2 * nb * sizeof(float4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->pvel.p, pvel.data(), 2 * nb * sizeof(float4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->bstat.p, bst.data(), nb * sizeof(float4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->Ilocal.p, sc->inertia_local, 9 * (size_t)nb * sizeof(float), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->bfric.p, sc->friction, nb * sizeof(float), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->prio.p, sc->priority, nb * sizeof(float), hipMemcpyHostToDevice));
    std::vector<int> ord(nb);
    std::iota(ord.begin(), ord.end(), 0);
    HIPCK(hipMemcpy(s->order_id.p, ord.data(), nb * sizeof(int), hipMemcpyHostToDevice));
    // JP processing order: priority descending, index ascending on ties.
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return sc->priority[a] > sc->priority[b]; });
    HIPCK(hipMemcpy(s->order_jp.p, ord.data(), nb * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> pos(nb);
    for (int k = 0; k < nb; ++k) pos[ord[k]] = k;
    HIPCK(hipMemcpy(s->pos_jp.p, pos.data(), nb * sizeof(int), hipMemcpyHostToDevice));
  }
  // ---- static rows -> groups ----
  std::vector<GroupGeo> hg;
  std::vector<GroupSt> hs;
  int n_floor = 0;
  auto v3at = [](const float* a, int j) { return mk3(a[3 * j], a[3 * j + 1], a[3 * j + 2]); };
  for (int j = 0; j < ns;) {
    int t = sc->type[j];
    GroupGeo G;
    GroupSt S;
    std::memset(&G, 0, sizeof(G));
    std::memset(&S, 0, sizeof(S));
    if (t == kFloor) {
      if (!(isinf_neg(sc->fmin[j]) && sc->fmax[j] == 0.0f && isinf_pos(sc->fracture[j])))
        return set_err(AVBD_E_ARG, "floor row " + std::to_string(j) + " has non-default bounds");
      int nr = 1;
      if (j + 2 < ns && sc->type[j + 1] == kTangent &&
          sc->type[j + 2] == kTangent && sc->sibling[j + 1] == j && sc->sibling[j + 2] == j &&
          sc->partner[j + 1] == j + 2 && sc->partner[j + 2] == j + 1)
        nr = 3;
      v3 off = v3at(sc->off_a, j);
      if (nr == 3) {
        for (int r = 1; r < 3; ++r) {
          int q = j + r;
          v3 o2 = v3at(sc->off_a, q), an = v3at(sc->anchor, q);
          v3 want = r == 1 ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 0.f, 1.f);
          if (sc->body_a[q] != sc->body_a[j] || sc->body_b[q] >= 0 || o2.x != off.x ||
              o2.y != off.y || o2.z != off.z || an.x != want.x || an.y != want.y ||
              an.z != want.z || !(sc->stiffness[q] == sc->stiffness[j] ||
                                   (std::isnan(sc->stiffness[q]) && std::isnan(sc->stiffness[j]))) ||
              !isinf_pos(sc->fracture[q]) || sc->friction_d[q] != sc->friction_d[j + 1] ||
              sc->friction_s[q] != sc->friction_s[j + 1])
            return set_err(AVBD_E_ARG, "floor tangent row " + std::to_string(q) +
                                           " does not follow add_floor_contact_box layout");
        }
      }
      if (sc->body_b[j] >= 0)
        return set_err(AVBD_E_ARG, "floor row " + std::to_string(j) + " names a second body");
      G.hdr = make_int4(g_make(kGFloor, nr, 0), sc->body_a[j], -1, j);
      v3 an = v3at(sc->anchor, j);
      G.g0 = make_float4(an.x, an.y, an.z, sc->stiffness[j]);
      G.g1 = make_float4(off.x, off.y, off.z, nr == 3 ? sc->friction_d[j + 1] : 0.f);
      G.g2 = make_float4(0.f, 0.f, 0.f, nr == 3 ? sc->friction_s[j + 1] : 0.f);
      float lam[3] = {0, 0, 0}, pen[3] = {0, 0, 0};
      int bits = 0;
      for (int r = 0; r < nr; ++r) {
        lam[r] = sc->lambda[j + r];
        pen[r] = sc->penalty[j + r];
        if (sc->active[j + r]) bits |= (r == 0 ? kActN : (r == 1 ? kActT : kActB));
        if (sc->was_static[j + r]) bits |= (r == 0 ? kWasN : (r == 1 ? kWasT : kWasB));
      }
      S.s0 = make_float4(lam[0], lam[1], lam[2], 0.f);
      S.s1 = make_float4(pen[0], pen[1], pen[2], 0.f);
      S.s2 = make_float4(0.f, i2f(bits), 0.f, 0.f);
      ++n_floor;
      j += nr;
    } else if (t == kPin) {
      int axis = sc->body_b[j];
      if (axis < 0 || axis > 2) return set_err(AVBD_E_ARG, "pin row axis out of range");
      if (!(isinf_neg(sc->fmin[j]) && isinf_pos(sc->fmax[j])))
        return set_err(AVBD_E_ARG, "pin row " + std::to_string(j) + " has non-default bounds");
      v3 an = v3at(sc->anchor, j), off = v3at(sc->off_a, j);
      G.hdr = make_int4(g_make(kGPin, 1, 0), sc->body_a[j], axis, j);
      G.g0 = make_float4(an.x, an.y, an.z, sc->stiffness[j]);
      G.g1 = make_float4(off.x, off.y, off.z, 0.f);
      G.g2 = make_float4(0.f, 0.f, 0.f, sc->fracture[j]);
      int bits = (sc->active[j] ? kActN : 0) | (sc->was_static[j] ? kWasN : 0);
      S.s0 = make_float4(sc->lambda[j], 0.f, 0.f, 0.f);
      S.s1 = make_float4(sc->penalty[j], 0.f, 0.f, 0.f);
      S.s2 = make_float4(0.f, i2f(bits), 0.f, 0.f);
      j += 1;
    } else {
      return set_err(AVBD_E_ARG, "static row " + std::to_string(j) + " of type " +
                                     std::to_string(t) + " is not supported");
    }
    if (G.hdr.y < 0 || G.hdr.y >= nb) return set_err(AVBD_E_ARG, "row body index out of range");
    hg.push_back(G);
    hs.push_back(S);
  }
  s->n_static_rows = ns;
  s->g_static = (int)hg.size();
  s->n_floor_groups = n_floor;
  s->n_dyn_cap = sc->n_dyn_capacity;
  s->row_cap = ns + sc->n_dyn_capacity;
  s->g_cap = s->g_static + sc->n_dyn_capacity;
  CK(s->geo.alloc(std::max(s->g_cap, 1)));
  CK(s->gst.alloc(std::max(s->g_cap, 1)));
  if (s->g_static > 0) {
    HIPCK(hipMemcpy(s->geo.p, hg.data(), hg.size() * sizeof(GroupGeo), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(s->gst.p, hs.data(), hs.size() * sizeof(GroupSt), hipMemcpyHostToDevice));
  }
  // static CSR (each body's static groups in ascending order)
  s->h_st_count.assign(nbb, 0);
  for (auto& G : hg) s->h_st_count[G.hdr.y]++;
  s->h_st_start.assign(nbb + 1, 0);
  for (int i = 0; i < nb; ++i) s->h_st_start[i + 1] = s->h_st_start[i] + s->h_st_count[i];
  s->st_total = nb > 0 ? s->h_st_start[nb] : 0;
  std::vector<int> sidx(std::max(s->st_total, 1)), cur(s->h_st_start.begin(), s->h_st_start.end() - 1);
  for (int g = 0; g < (int)hg.size(); ++g) sidx[cur[hg[g].hdr.y]++] = g;
  CK(s->st_count.alloc(nbb)); CK(s->st_start.alloc(nbb + 1)); CK(s->st_idx.alloc(sidx.size()));
  HIPCK(hipMemcpy(s->st_count.p, s->h_st_count.data(), nbb * sizeof(int), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(s->st_start.p, s->h_st_start.data(), (nbb + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(s->st_idx.p, sidx.data(), sidx.size() * sizeof(int), hipMemcpyHostToDevice));
  // ---- pairs / hash / pools ----
  s->pair_cap = 0;
  s->pa.release();
  if (sc->max_pairs > 0) CK(ensure_pairs(s, sc->max_pairs));
  CK(s->bcount.alloc(nbb)); CK(s->boff.alloc(nbb));
  s->hash_cap = sc->hash_cap;
  s->pool_max = sc->pool_max;
  CK(s->hkeys.alloc(std::max(s->hash_cap, 1)));
  CK(s->howner.alloc(std::max(s->hash_cap, 1)));
  CK(s->hstate.alloc(2 * (size_t)std::max(s->hash_cap, 1)));
  CK(s->slot_of.alloc(std::max(s->pool_max, 1)));
  launch_fill_int(s->st, s->hkeys.p, std::max(s->hash_cap, 1), kHashEmpty);
  HIPCK(hipMemsetAsync(s->hstate.p, 0, 2 * (size_t)std::max(s->hash_cap, 1) * sizeof(float4), s->st));
  s->hash_live = std::min(s->hash_cap, 4096);
  s->hage.release();
  launch_fill_int(s->st, s->dcnt.p + kHashLiveSlot, 1, s->hash_live);
  launch_fill_int(s->st, s->dcnt.p + kRecolorsSlot, 1, 0);
  s->recolors_seen = 0;
  // ---- CSR + coloring ----
  CK(s->tc_work.alloc(5 * nbb)); CK(s->asb_start.alloc(nbb + 1)); CK(s->asb_list.alloc(1));
  CK(s->padc.alloc(nbb)); CK(s->cpos.alloc(nbb + 1)); CK(s->pos_of.alloc(nbb));
  CK(s->pos_map.alloc(nbb));
  CK(s->flow_done.alloc(nbb));
  launch_fill_int(s->st, s->flow_done.p, nbb, 0);
  // a flush starts the epoch over and clears a dataflow error of the
  // previous scene
  launch_fill_int(s->st, s->dcnt.p + kErrSlot, 1, 0);
  launch_fill_int(s->st, s->dcnt.p + kEpochSlot, 1, 0);
  s->hcnt[kErrSlot] = 0;
  s->csr_built = false;
  s->csr_order_stale = false;
  CK(s->tcsr.alloc(std::max(s->st_total, 1) + 3 * (size_t)nbb));
  HIPCK(hipMemsetAsync(s->cpos.p, 0, (nbb + 1) * sizeof(int), s->st));
  CK(s->color.alloc(nbb)); CK(s->color_next.alloc(nbb)); CK(s->nbr_count.alloc(nbb));
  CK(s->nbr_start.alloc(nbb + 1)); CK(s->nbr_cursor.alloc(nbb)); CK(s->nbr.alloc(1));
  CK(s->pub.alloc(nbb)); CK(s->ticket.alloc(1));
  CK(s->ckeys.alloc(nbb)); CK(s->ckeys_tmp.alloc(nbb)); CK(s->cvals.alloc(nbb));
  CK(s->cvals_tmp.alloc(nbb));
  CK(s->ws_bvh.alloc(bvh_workspace_bytes((int)nbb)));
  CK(s->ws_color.alloc(radix_workspace_bytes((int)nbb)));
  s->ws_scan_elems = 0;
  CK(ensure_scan_ws(s, std::max<size_t>(nbb, 16 * ((nbb + 1023) / 1024) + 16)));
  launch_fill_int(s->st, s->color.p, (int)nbb, -1);
  s->n_active_colors = kMaxColors;
  for (int c = 0; c <= kMaxColors; ++c) s->cstart[c] = 0;
  s->n_active_rows = ns;
  s->pair_count = s->pool_count = s->n_contacts = s->adj_pairs = 0;
  s->last_post = 0;
  s->n_scenes = 0;  // a flush replaces the scene; avbd_set_scenes splits it again
  s->flow_test_cap = 0;
  s->flushed = true;
  s->sx = avbd_solver::StageCtx{};
  CK(ensure_hage(s));
  HIPCK(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

int avbd_substep(avbd_solver* s, const avbd_params* p, int color_dirty, avbd_report* out) {
  if (!s || !p || !out) return set_err(AVBD_E_ARG, "null argument");
  if (!s->flushed) return set_err(AVBD_E_STATE, "substep before flush");
  if (s->sx.stage != kStContacts) return set_err(AVBD_E_STATE, "a staged substep is in progress");
  HIPCK(hipSetDevice(s->device));
  return do_substep(s, p, color_dirty, out);
}

int avbd_substep_to_residual(avbd_solver* s, const avbd_params* p, int color_dirty, float tol,
                             int* iterations_used, float* residual, avbd_report* out) {
  if (!s || !p || !out || !iterations_used || !residual)
    return set_err(AVBD_E_ARG, "null argument");
  if (!s->flushed) return set_err(AVBD_E_STATE, "substep before flush");
  if (!(tol >= 0.0f)) return set_err(AVBD_E_ARG, "tolerance must be >= 0");
  if (s->sx.stage != kStContacts) return set_err(AVBD_E_STATE, "a staged substep is in progress");
  HIPCK(hipSetDevice(s->device));
  return do_substep(s, p, color_dirty, out, tol, iterations_used, residual);
}

int avbd_step_frame(avbd_solver* s, const avbd_params* p, int n_substeps, int color_dirty,
                    avbd_report* out) {
  if (!s || !p || !out) return set_err(AVBD_E_ARG, "null argument");
  if (!s->flushed) return set_err(AVBD_E_STATE, "substep before flush");
  if (n_substeps < 1) return set_err(AVBD_E_ARG, "a frame needs at least one substep");
  if (s->sx.stage != kStContacts) return set_err(AVBD_E_STATE, "a staged substep is in progress");
  HIPCK(hipSetDevice(s->device));
  std::memset(out, 0, sizeof(*out));
  out->n_active_colors = s->n_active_colors;
  if (s->nb == 0) return AVBD_OK;
  bool ok = false;
  CK(graph_eligible(s, p, &ok));
  avbd_report rep{};
  int recolored_direct = 0;
  bool on_device = false;
  for (int k = 0; k < n_substeps; ++k) {
    const int dirty = (color_dirty && k == 0) ? 1 : 0;
    if (ok && !dirty) {
      // the device counters must describe the state the graph starts from:
      // after direct substeps they do (they are the same counters)
      CK(substep_graph(s, p));
      on_device = true;
    } else {
      CK(do_substep(s, p, dirty, &rep));
      recolored_direct |= rep.recolored;
    }
  }
  if (on_device) return frame_end(s, out, recolored_direct);
  *out = rep;
  out->recolored = recolored_direct;
  return AVBD_OK;
}

int avbd_sync_count(avbd_solver* s, long long* syncs, long long* graph_launches) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  if (syncs) *syncs = s->syncs;
  if (graph_launches) *graph_launches = s->graph_launches;
  return AVBD_OK;
}

// ---- one substep stage by stage -------------------------------------------
static int stage_guard(avbd_solver* s, const void* p, int want) {
  if (!s || !p) return set_err(AVBD_E_ARG, "null argument");
  if (!s->flushed) return set_err(AVBD_E_STATE, "substep before flush");
  if (s->sx.stage != want) return set_err(AVBD_E_STATE, "substep stage out of order");
  HIPCK(hipSetDevice(s->device));
  return AVBD_OK;
}

int avbd_stage_contacts(avbd_solver* s, const avbd_params* p, avbd_report* out) {
  CK(stage_guard(s, out ? (const void*)p : nullptr, kStContacts));
  if (s->nb == 0) {
    std::memset(out, 0, sizeof(*out));
    out->n_active_colors = s->n_active_colors;
    s->sx = avbd_solver::StageCtx{};
    s->sx.stage = kStColoring;
    return AVBD_OK;
  }
  const int r = stage_contacts(s, p, out);
  if (r != AVBD_OK) s->sx.stage = kStContacts;
  return r;
}

int avbd_stage_coloring(avbd_solver* s, const avbd_params* p, int color_dirty, avbd_report* out) {
  CK(stage_guard(s, out ? (const void*)p : nullptr, kStColoring));
  if (s->nb == 0) {
    s->sx.stage = kStCsr;
    return AVBD_OK;
  }
  const int r = stage_coloring(s, p, color_dirty, out);
  if (r != AVBD_OK) s->sx.stage = kStContacts;
  return r;
}

int avbd_stage_csr(avbd_solver* s) {
  CK(stage_guard(s, s, kStCsr));
  if (s->nb == 0) {
    s->sx.stage = kStPrelude;
    return AVBD_OK;
  }
  const int r = stage_csr(s);
  if (r != AVBD_OK) s->sx.stage = kStContacts;
  return r;
}

int avbd_stage_prelude(avbd_solver* s, const avbd_params* p) {
  CK(stage_guard(s, p, kStPrelude));
  if (s->nb == 0) {
    s->sx.iterations = p->iterations;
    s->sx.post = p->post_stabilize ? 1 : 0;
    s->sx.sweeps_done = 0;
    s->sx.stage = kStSweeps;
    return AVBD_OK;
  }
  const int r = stage_prelude(s, p, -1.0f);
  if (r != AVBD_OK) s->sx.stage = kStContacts;
  return r;
}

int avbd_stage_sweep(avbd_solver* s, const avbd_params* p, int sweep) {
  CK(stage_guard(s, p, kStSweeps));
  const int total = s->sx.iterations + s->sx.post;
  if (sweep != s->sx.sweeps_done || sweep >= total)
    return set_err(AVBD_E_ARG, "sweeps run in order, 0 .. iterations + post_stabilize - 1");
  if (s->nb == 0) {
    s->sx.sweeps_done = sweep + 1;
    return AVBD_OK;
  }
  const int r = stage_sweep(s, p, sweep);
  if (r != AVBD_OK) s->sx.stage = kStContacts;
  return r;
}

int avbd_stage_end(avbd_solver* s, const avbd_params* p) {
  CK(stage_guard(s, p, kStSweeps));
  if (s->sx.sweeps_done != s->sx.iterations + s->sx.post)
    return set_err(AVBD_E_STATE, "substep ended before its last sweep");
  if (s->nb == 0) {
    s->sx.stage = kStContacts;
    return AVBD_OK;
  }
  const int r = stage_end(s, p);
  s->sx.stage = kStContacts;
  return r;
}

// ---- streams and device views ---------------------------------------------
int avbd_get_stream(avbd_solver* s, void** stream) {
  if (!s || !stream) return set_err(AVBD_E_ARG, "null argument");
  *stream = reinterpret_cast<void*>(s->st);
  return AVBD_OK;
}

int avbd_set_stream(avbd_solver* s, void* stream) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  if (s->sx.stage != kStContacts) return set_err(AVBD_E_STATE, "a staged substep is in progress");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  s->st = stream ? reinterpret_cast<hipStream_t>(stream) : s->own_st;
  return AVBD_OK;
}

int avbd_device_views(avbd_solver* s, avbd_views* out) {
  if (!s || !out) return set_err(AVBD_E_ARG, "null argument");
  if (!s->flushed) return set_err(AVBD_E_STATE, "views before flush");
  std::memset(out, 0, sizeof(*out));
  out->n_bodies = s->nb;
  float* pose = reinterpret_cast<float*>(s->pose[s->cur].p);
  float* vel = reinterpret_cast<float*>(s->vel.p);
  out->x = pose;
  out->q = pose ? pose + 4 : nullptr;
  out->v = vel;
  out->w = vel ? vel + 4 : nullptr;
  out->body_stride = 8;
  out->n_groups = s->g_static + std::min(s->n_contacts, s->pool_max);
  float* gst = reinterpret_cast<float*>(s->gst.p);
  out->lambda = gst;
  out->penalty = gst ? gst + 4 : nullptr;
  out->state_bits = gst ? reinterpret_cast<int*>(gst + 9) : nullptr;
  out->group_stride = (int)(sizeof(GroupSt) / sizeof(float));
  out->header = reinterpret_cast<int*>(s->geo.p);
  out->header_stride = (int)(sizeof(GroupGeo) / sizeof(int));
  return AVBD_OK;
}

int avbd_read_bodies(avbd_solver* s, float* x, float* q, float* v, float* w, float* pv,
                     float* pw) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  HIPCK(hipSetDevice(s->device));
  int nb = s->nb;
  if (nb == 0) return AVBD_OK;
  std::vector<float4> a(2 * nb), b(2 * nb), c(2 * nb);
  HIPCK(hipMemcpyAsync(a.data(), s->pose[s->cur].p, 2 * nb * sizeof(float4), hipMemcpyDeviceToHost, s->st));
  HIPCK(hipMemcpyAsync(b.data(), s->vel.p, 2 * nb * sizeof(float4), hipMemcpyDeviceToHost, s->st));
  HIPCK(hipMemcpyAsync(c.data(), s->pvel.p, 2 * nb * sizeof(float4), hipMemcpyDeviceToHost, s->st));
  HIPCK(hipStreamSynchronize(s->st));
  CK(check_flow_error(s));
  for (int i = 0; i < nb; ++i) {
    if (x) { x[3 * i] = a[2 * i].x; x[3 * i + 1] = a[2 * i].y; x[3 * i + 2] = a[2 * i].z; }
    if (q) { q[4 * i] = a[2 * i + 1].x; q[4 * i + 1] = a[2 * i + 1].y; q[4 * i + 2] = a[2 * i + 1].z; q[4 * i + 3] = a[2 * i + 1].w; }
    if (v) { v[3 * i] = b[2 * i].x; v[3 * i + 1] = b[2 * i].y; v[3 * i + 2] = b[2 * i].z; }
    if (w) { w[3 * i] = b[2 * i + 1].x; w[3 * i + 1] = b[2 * i + 1].y; w[3 * i + 2] = b[2 * i + 1].z; }
    if (pv) { pv[3 * i] = c[2 * i].x; pv[3 * i + 1] = c[2 * i].y; pv[3 * i + 2] = c[2 * i].z; }
    if (pw) { pw[3 * i] = c[2 * i + 1].x; pw[3 * i + 1] = c[2 * i + 1].y; pw[3 * i + 2] = c[2 * i + 1].z; }
  }
  return AVBD_OK;
}

int avbd_write_bodies(avbd_solver* s, int field, int first, int n, const float* vals) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "write before flush");
  if (first < 0 || n < 0 || first + n > s->nb) return set_err(AVBD_E_ARG, "body range");
  if (n == 0) return AVBD_OK;
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  std::vector<float4> h(n);
  float4* base;
  int off;
  switch (field) {
    case 0: base = s->pose[s->cur].p; off = 0; break;
    case 1: base = s->pose[s->cur].p; off = 1; break;
    case 2: base = s->vel.p; off = 0; break;
    case 3: base = s->vel.p; off = 1; break;
    case 4: base = s->pvel.p; off = 0; break;
    case 5: base = s->pvel.p; off = 1; break;
    default: return set_err(AVBD_E_ARG, "unknown body field");
  }
  for (int k = 0; k < n; ++k) {
    if (field == 1) h[k] = make_float4(vals[4 * k], vals[4 * k + 1], vals[4 * k + 2], vals[4 * k + 3]);
    else h[k] = make_float4(vals[3 * k], vals[3 * k + 1], vals[3 * k + 2], 0.f);
  }
  HIPCK(hipMemcpy2DAsync(base + 2 * first + off, 2 * sizeof(float4), h.data(), sizeof(float4),
                         sizeof(float4), n, hipMemcpyHostToDevice, s->st));
  HIPCK(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

static int download_groups(avbd_solver* s, int ng, std::vector<GroupGeo>& G,
                           std::vector<GroupSt>& S) {
  G.resize(std::max(ng, 1));
  S.resize(std::max(ng, 1));
  if (ng > 0) {
    HIPCK(hipMemcpyAsync(G.data(), s->geo.p, ng * sizeof(GroupGeo), hipMemcpyDeviceToHost, s->st));
    HIPCK(hipMemcpyAsync(S.data(), s->gst.p, ng * sizeof(GroupSt), hipMemcpyDeviceToHost, s->st));
  }
  HIPCK(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

int avbd_read_rows(avbd_solver* s, int n, float* lam, float* pen, int* act, int* was, int* type,
                   int* ba, int* bb, float* ac0) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  for (int j = 0; j < n; ++j) {
    if (lam) lam[j] = 0.f;
    if (pen) pen[j] = 0.f;
    if (act) act[j] = 0;
    if (was) was[j] = 0;
    if (type) type[j] = 0;
    if (ba) ba[j] = 0;
    if (bb) bb[j] = 0;
    if (ac0) ac0[j] = 0.f;
  }
  int ng = s->g_static + s->n_contacts;
  std::vector<GroupGeo> G;
  std::vector<GroupSt> S;
  CK(download_groups(s, ng, G, S));
  CK(check_flow_error(s));
  for (int g = 0; g < ng; ++g) {
    int h = G[g].hdr.x, kind = g_kind(h), nr = g_nrows(h);
    int bits = f2i(S[g].s2.y);
    const float L[3] = {S[g].s0.x, S[g].s0.y, S[g].s0.z};
    const float K[3] = {S[g].s1.x, S[g].s1.y, S[g].s1.z};
    const float A[3] = {S[g].s0.w, S[g].s1.w, S[g].s2.x};
    for (int r = 0; r < nr; ++r) {
      int j = G[g].hdr.w + r;
      if (j < 0 || j >= n) continue;
      int rt = r == 0 ? (kind == kGFloor ? kFloor : (kind == kGBox ? kBoxBox : kPin)) : kTangent;
      if (lam) lam[j] = L[r];
      if (pen) pen[j] = K[r];
      if (act) act[j] = (bits >> r) & 1;
      if (was) was[j] = (bits >> (3 + r)) & 1;
      if (type) type[j] = rt;
      if (ba) ba[j] = G[g].hdr.y;
      if (bb) bb[j] = kind == kGFloor ? -1 : G[g].hdr.z;
      if (ac0) ac0[j] = (s->last_post && rt != kTangent) ? 0.0f : A[r];
    }
  }
  return AVBD_OK;
}

int avbd_write_static_rows(avbd_solver* s, int n, const float* lam, const float* pen,
                           const int* act, const int* was) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "write before flush");
  if (n > s->n_static_rows) return set_err(AVBD_E_ARG, "more rows than static rows");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  std::vector<GroupGeo> G;
  std::vector<GroupSt> S;
  CK(download_groups(s, s->g_static, G, S));
  for (int g = 0; g < s->g_static; ++g) {
    int nr = g_nrows(G[g].hdr.x);
    int bits = f2i(S[g].s2.y);
    float* L[3] = {&S[g].s0.x, &S[g].s0.y, &S[g].s0.z};
    float* K[3] = {&S[g].s1.x, &S[g].s1.y, &S[g].s1.z};
    for (int r = 0; r < nr; ++r) {
      int j = G[g].hdr.w + r;
      if (j >= n) continue;
      if (lam) *L[r] = lam[j];
      if (pen) *K[r] = pen[j];
      if (act) bits = act[j] ? (bits | (1 << r)) : (bits & ~(1 << r));
      if (was) bits = was[j] ? (bits | (8 << r)) : (bits & ~(8 << r));
    }
    S[g].s2.y = i2f(bits);
  }
  if (s->g_static > 0)
    HIPCK(hipMemcpy(s->gst.p, S.data(), s->g_static * sizeof(GroupSt), hipMemcpyHostToDevice));
  return AVBD_OK;
}

int avbd_read_colors(avbd_solver* s, int* c) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  if (s->nb == 0) return AVBD_OK;
  HIPCK(hipMemcpyAsync(c, s->color.p, s->nb * sizeof(int), hipMemcpyDeviceToHost, s->st));
  HIPCK(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

int avbd_write_colors(avbd_solver* s, const int* c, int n_active_colors) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "write before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  if (s->nb == 0) return AVBD_OK;
  for (int i = 0; i < s->nb; ++i)
    if (c[i] < 0 || c[i] >= kMaxColors) return set_err(AVBD_E_ARG, "color out of range");
  HIPCK(hipMemcpy(s->color.p, c, s->nb * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> cnt(kMaxColors, 0);
  for (int i = 0; i < s->nb; ++i) cnt[c[i]]++;
  s->cstart[0] = 0;
  for (int k = 0; k < kMaxColors; ++k) s->cstart[k + 1] = s->cstart[k] + cnt[k];
  s->csr_order_stale = true;
  color_bucket(s->st, s->nb, s->color.p, s->ckeys.p, s->cvals.p, s->ckeys_tmp.p, s->cvals_tmp.p,
               s->ws_color.p);
  HIPCK(hipStreamSynchronize(s->st));
  s->n_active_colors = n_active_colors;
  return AVBD_OK;
}

int avbd_read_counters(avbd_solver* s, int* nar, int* pc, int* poc) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  if (nar) *nar = s->n_active_rows;
  if (pc) *pc = s->pair_count;
  if (poc) *poc = s->pool_count;
  return AVBD_OK;
}

int avbd_count_color_conflicts(avbd_solver* s, int* out) {
  if (!s || !s->flushed) {
    *out = 0;
    return AVBD_OK;
  }
  HIPCK(hipSetDevice(s->device));
  return count_conflicts(s, out);
}

int avbd_pack_state(avbd_solver* s, float* bodies10, int n_rows, float* rows3) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  int nb = s->nb;
  std::vector<float> x(3 * (size_t)nb + 1), q(4 * (size_t)nb + 1), w(3 * (size_t)nb + 1);
  CK(avbd_read_bodies(s, x.data(), q.data(), nullptr, w.data(), nullptr, nullptr));
  for (int i = 0; i < nb; ++i) {
    float* o = bodies10 + 10 * (size_t)i;
    o[0] = x[3 * i]; o[1] = x[3 * i + 1]; o[2] = x[3 * i + 2];
    o[3] = q[4 * i]; o[4] = q[4 * i + 1]; o[5] = q[4 * i + 2]; o[6] = q[4 * i + 3];
    o[7] = w[3 * i]; o[8] = w[3 * i + 1]; o[9] = w[3 * i + 2];
  }
  if (n_rows > 0 && rows3) {
    std::vector<float> lam(n_rows);
    std::vector<int> act(n_rows), was(n_rows), ty(n_rows);
    CK(avbd_read_rows(s, n_rows, lam.data(), nullptr, act.data(), was.data(), ty.data(), nullptr,
                      nullptr, nullptr));
    for (int j = 0; j < n_rows; ++j) {
      rows3[3 * (size_t)j] = lam[j];
      rows3[3 * (size_t)j + 1] = (float)act[j];
      rows3[3 * (size_t)j + 2] = (float)(was[j] * 16 + ty[j]);
    }
  }
  return AVBD_OK;
}

int avbd_export_hash(avbd_solver* s, int cap, int* keys, float* st8, int* n) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  int hc = s->hash_cap > 0 ? s->hash_live : 0;  // the slots of the table as built
  *n = 0;
  if (hc <= 0) return AVBD_OK;
  std::vector<int> k(hc);
  std::vector<float4> v(2 * (size_t)hc);
  HIPCK(hipMemcpyAsync(k.data(), s->hkeys.p, hc * sizeof(int), hipMemcpyDeviceToHost, s->st));
  HIPCK(hipMemcpyAsync(v.data(), s->hstate.p, 2 * (size_t)hc * sizeof(float4), hipMemcpyDeviceToHost, s->st));
  HIPCK(hipStreamSynchronize(s->st));
  int m = 0;
  for (int sl = 0; sl < hc; ++sl) {
    if (k[sl] == kHashEmpty) continue;
    if (m < cap) {
      keys[m] = k[sl];
      float4 a = v[2 * (size_t)sl], b = v[2 * (size_t)sl + 1];
      float* o = st8 + 8 * (size_t)m;
      o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w; o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
    }
    ++m;
  }
  *n = m;
  return AVBD_OK;
}

int avbd_import_hash(avbd_solver* s, int n, const int* keys, const float* st8) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "write before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  if (s->hash_cap <= 0) return n == 0 ? AVBD_OK : set_err(AVBD_E_ARG, "no hash table");
  if (n > s->pool_max) return set_err(AVBD_E_ARG, "more cache entries than the pool holds");
  const int hc = live_hash_slots(s, n);
  s->hash_live = hc;
  std::vector<int> k(hc, kHashEmpty);
  std::vector<float4> v(2 * (size_t)hc, make_float4(0.f, 0.f, 0.f, 0.f));
  for (int e = 0; e < n; ++e) {
    int h = hash_pair_slot(keys[e], hc);
    for (int probe = 0; probe < hc; ++probe) {
      int sl = (h + probe) % hc;
      if (k[sl] == kHashEmpty || k[sl] == keys[e]) {
        k[sl] = keys[e];
        const float* o = st8 + 8 * (size_t)e;
        v[2 * (size_t)sl] = make_float4(o[0], o[1], o[2], o[3]);
        v[2 * (size_t)sl + 1] = make_float4(o[4], o[5], o[6], o[7]);
        break;
      }
    }
  }
  HIPCK(hipMemcpy(s->hkeys.p, k.data(), hc * sizeof(int), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(s->hstate.p, v.data(), 2 * (size_t)hc * sizeof(float4), hipMemcpyHostToDevice));
  if (s->hage.p) {
    launch_fill_int(s->st, s->hage.p, (int)s->hage.n, 0);
    HIPCK(hipStreamSynchronize(s->st));
  }
  return AVBD_OK;
}

int avbd_read_capacities(avbd_solver* s, int* mp, int* rc, int* pm, int* hc) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  if (mp) *mp = s->pair_cap;
  if (rc) *rc = s->row_cap;
  if (pm) *pm = s->pool_max;
  if (hc) *hc = s->hash_cap;
  return AVBD_OK;
}

int avbd_read_pairs(avbd_solver* s, int cap, int* a, int* b, int* ov, int* sat, float* n,
                    float* depth, int* count) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  *count = s->pair_count;
  int k = std::min(std::min(s->pair_count, s->pair_cap), cap);
  if (k <= 0) return AVBD_OK;
  std::vector<int> A(k), B(k), O(k), T(k);
  std::vector<float4> N(k);
  std::vector<float> D(k);
  HIPCK(hipMemcpy(A.data(), s->pa.p, k * sizeof(int), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(B.data(), s->pb.p, k * sizeof(int), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(O.data(), s->pov.p, k * sizeof(int), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(T.data(), s->psat.p, k * sizeof(int), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(N.data(), s->pn.p, k * sizeof(float4), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(D.data(), s->pdepth.p, k * sizeof(float), hipMemcpyDeviceToHost));
  for (int p = 0; p < k; ++p) {
    if (a) a[p] = A[p];
    if (b) b[p] = B[p];
    if (ov) ov[p] = O[p];
    if (sat) sat[p] = O[p] ? T[p] : -1;
    if (n) { n[3 * p] = O[p] ? N[p].x : 0.f; n[3 * p + 1] = O[p] ? N[p].y : 0.f; n[3 * p + 2] = O[p] ? N[p].z : 0.f; }
    if (depth) depth[p] = O[p] ? D[p] : 0.f;
  }
  return AVBD_OK;
}

int avbd_read_manifold(avbd_solver* s, int cap, int* count, int* refa, float* n, float* t,
                       float* b, float* offr, float* offi) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "read before flush");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  int k = std::min(std::min(s->pair_count, s->pair_cap), cap);
  if (k <= 0) return AVBD_OK;
  std::vector<int> C(k), R(k);
  std::vector<float4> N(k), O(8 * (size_t)k);
  HIPCK(hipMemcpy(C.data(), s->mfc.p, k * sizeof(int), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(R.data(), s->mfr.p, k * sizeof(int), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(N.data(), s->mfn.p, k * sizeof(float4), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(O.data(), s->mfoff.p, 8 * (size_t)k * sizeof(float4), hipMemcpyDeviceToHost));
  for (int p = 0; p < k; ++p) {
    int c = C[p];
    if (count) count[p] = c;
    if (refa) refa[p] = c ? R[p] : 0;
    v3 nn = c ? mk3(N[p].x, N[p].y, N[p].z) : mk3(0.f, 0.f, 0.f);
    v3 tt = c ? orthonormal_basis(nn) : mk3(0.f, 0.f, 0.f);
    v3 bb = c ? cross(nn, tt) : mk3(0.f, 0.f, 0.f);
    if (n) { n[3 * p] = nn.x; n[3 * p + 1] = nn.y; n[3 * p + 2] = nn.z; }
    if (t) { t[3 * p] = tt.x; t[3 * p + 1] = tt.y; t[3 * p + 2] = tt.z; }
    if (b) { b[3 * p] = bb.x; b[3 * p + 1] = bb.y; b[3 * p + 2] = bb.z; }
    for (int sl = 0; sl < 4; ++sl) {
      float4 r = sl < c ? O[8 * (size_t)p + sl] : make_float4(0.f, 0.f, 0.f, 0.f);
      float4 i = sl < c ? O[8 * (size_t)p + 4 + sl] : make_float4(0.f, 0.f, 0.f, 0.f);
      if (offr) { offr[3 * (4 * p + sl)] = r.x; offr[3 * (4 * p + sl) + 1] = r.y; offr[3 * (4 * p + sl) + 2] = r.z; }
      if (offi) { offi[3 * (4 * p + sl)] = i.x; offi[3 * (4 * p + sl) + 1] = i.y; offi[3 * (4 * p + sl) + 2] = i.z; }
    }
  }
  return AVBD_OK;
}

int avbd_set_option(avbd_solver* s, int option, int value) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  switch (option) {
    case AVBD_OPT_PRIMAL_VARIANT:
      if (value < 0 || value > 10) return set_err(AVBD_E_ARG, "primal variant must be 0 to 10");
      s->primal_variant = value;
      return AVBD_OK;
    case AVBD_OPT_FUSE_DUAL:
      if (value < 0 || value > 2) return set_err(AVBD_E_ARG, "dual placement must be 0 to 2");
      s->fuse_dual_mode = value;
      return AVBD_OK;
    case AVBD_OPT_LEARNED_SCOPE:
      if (value < 0 || value > 1) return set_err(AVBD_E_ARG, "learned scope must be 0 or 1");
      s->ml_scope = value;
      return AVBD_OK;
    case AVBD_OPT_GRAPH:
      if (value < 0 || value > 1) return set_err(AVBD_E_ARG, "graph option must be 0 or 1");
      s->graph_on = value != 0;
      if (!s->graph_on) {
        HIPCK(hipSetDevice(s->device));
        HIPCK(hipStreamSynchronize(s->st));
        drop_graphs(s);
      }
      return AVBD_OK;
    case AVBD_OPT_BROAD_LANES:
      if (value != 0 && value != 1 && value != 8 && value != 16 && value != 64)
        return set_err(AVBD_E_ARG, "broad-phase lanes must be 0, 1, 8, 16 or 64");
      s->broad_lanes = value;
      return AVBD_OK;
    case AVBD_OPT_FLOW_GIVE_UP_TEST:
      if (value < 0 || value > 4096) return set_err(AVBD_E_ARG, "give-up test cap must be 0 to 4096");
      if (value > 0) {
        if (!s->flushed) return set_err(AVBD_E_STATE, "give-up test before flush");
        HIPCK(hipSetDevice(s->device));
        CK(s->flow_never.ensure(std::max(s->nb, 1)));
        launch_fill_int(s->st, s->flow_never.p, std::max(s->nb, 1), 0);  // 0 is never a stamp
        HIPCK(hipStreamSynchronize(s->st));
      }
      s->flow_test_cap = value;
      return AVBD_OK;
    default:
      return set_err(AVBD_E_ARG, "unknown option");
  }
}

int avbd_profile_enable(avbd_solver* s, int on) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  s->prof = on != 0;
  for (int k = 0; k < kNumProf; ++k) {
    s->prof_ms[k] = 0;
    s->prof_launch[k] = 0;
    s->prof_bytes[k] = 0;
  }
  return AVBD_OK;
}

int avbd_profile_read(avbd_solver* s, int nc, double* ms, long long* launches, double* bytes) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  for (int k = 0; k < nc && k < kNumProf; ++k) {
    if (ms) ms[k] = s->prof_ms[k];
    if (launches) launches[k] = s->prof_launch[k];
    if (bytes) bytes[k] = s->prof_bytes[k];
  }
  return AVBD_OK;
}

// ---------------------------------------------------------------------------
// batched scenes
// ---------------------------------------------------------------------------
int avbd_set_scenes(avbd_solver* s, int n_scenes, const int* first_body, float pitch) {
  if (!s || !s->flushed) return set_err(AVBD_E_STATE, "set_scenes before flush");
  if (n_scenes < 1 || !first_body) return set_err(AVBD_E_ARG, "need at least one scene");
  if (first_body[0] != 0 || first_body[n_scenes] != s->nb)
    return set_err(AVBD_E_ARG, "scene ranges must cover bodies 0 .. n_bodies");
  for (int k = 0; k < n_scenes; ++k)
    if (first_body[k + 1] < first_body[k])
      return set_err(AVBD_E_ARG, "scene ranges must be non-decreasing");
  if (!(pitch > 0.0f) || !(pitch < 1.0e30f)) return set_err(AVBD_E_ARG, "pitch must be positive");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  if (n_scenes == 1) {
    s->n_scenes = 0;
    return AVBD_OK;
  }
  const int nbb = std::max(s->nb, 1);
  std::vector<int> sc(nbb, 0);
  for (int k = 0; k < n_scenes; ++k)
    for (int i = first_body[k]; i < first_body[k + 1]; ++i) sc[i] = k;
  CK(s->scene_of.alloc(nbb));
  CK(s->scene_dirty.alloc(n_scenes));
  CK(s->slo.alloc(nbb));
  CK(s->shi.alloc(nbb));
  HIPCK(hipMemcpy(s->scene_of.p, sc.data(), nbb * sizeof(int), hipMemcpyHostToDevice));
  launch_fill_int(s->st, s->scene_dirty.p, n_scenes, 1);
  int g = 1;
  while ((long long)g * g * g < n_scenes) ++g;
  s->scene_grid = g;
  s->scene_pitch = pitch;
  s->n_scenes = n_scenes;
  HIPCK(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

// ---------------------------------------------------------------------------
// learned warm start
// ---------------------------------------------------------------------------
int avbd_mlp_num_params(void) { return mlp_num_params(); }

int avbd_ml_set_mode(avbd_solver* s, int mode) {
  if (!s) return set_err(AVBD_E_ARG, "null solver");
  if (mode < 0 || mode > 2) return set_err(AVBD_E_ARG, "learned mode must be 0, 1 or 2");
  if (mode == 2 && !s->ml_have_params)
    return set_err(AVBD_E_STATE, "learned mode 2 needs parameters (avbd_ml_set_params)");
  HIPCK(hipSetDevice(s->device));
  s->ml_mode = mode;
  s->ml_n = 0;
  CK(ensure_hage(s));
  HIPCK(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

int avbd_ml_set_params(avbd_solver* s, const float* params, int n) {
  if (!s || !params) return set_err(AVBD_E_ARG, "null argument");
  if (n != mlp_num_params())
    return set_err(AVBD_E_ARG, "expected " + std::to_string(mlp_num_params()) + " parameters");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  CK(s->ml_params.ensure(n));
  HIPCK(hipMemcpy(s->ml_params.p, params, n * sizeof(float), hipMemcpyHostToDevice));
  s->ml_have_params = true;
  return AVBD_OK;
}

int avbd_ml_read_samples(avbd_solver* s, int cap, float* feat, float* targ, int* mask, int* n) {
  if (!s || !n) return set_err(AVBD_E_ARG, "null argument");
  if (s->ml_mode != 1) return set_err(AVBD_E_STATE, "samples are logged in learned mode 1 only");
  HIPCK(hipSetDevice(s->device));
  HIPCK(hipStreamSynchronize(s->st));
  *n = s->ml_n;
  const int m = std::min(cap, s->ml_n);
  if (m <= 0) return AVBD_OK;
  if (feat) HIPCK(hipMemcpy(feat, s->ml_feat.p, (size_t)m * 24 * sizeof(float), hipMemcpyDeviceToHost));
  if (targ) HIPCK(hipMemcpy(targ, s->ml_targ.p, (size_t)m * 4 * sizeof(float), hipMemcpyDeviceToHost));
  if (mask) HIPCK(hipMemcpy(mask, s->ml_mask.p, (size_t)m * sizeof(int), hipMemcpyDeviceToHost));
  return AVBD_OK;
}

namespace {

// device, stream and the sample arrays of one standalone network call
struct MlpJob {
  hipStream_t st = nullptr;
  DBuf<float> X, Y, P, G, L, Q;  // Q: per-workgroup gradient rows
  DBuf<int> M, I;
  ~MlpJob() {
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
  int open(int device, int n, const float* x, const float* y, const int* mask,
           const float* params) {
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
      return set_err(AVBD_E_HIP, "no HIP device available");
    if (device < 0 || device >= nd) return set_err(AVBD_E_ARG, "device ordinal out of range");
    if (n <= 0 || !x || !params) return set_err(AVBD_E_ARG, "empty sample set or null argument");
    HIPCK(hipSetDevice(device));
    HIPCK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int np = mlp_num_params();
    CK(X.alloc((size_t)n * 24));
    CK(P.alloc(np));
    HIPCK(hipMemcpy(X.p, x, (size_t)n * 24 * sizeof(float), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(P.p, params, np * sizeof(float), hipMemcpyHostToDevice));
    if (y) {
      CK(Y.alloc((size_t)n * 4));
      CK(M.alloc(n));
      HIPCK(hipMemcpy(Y.p, y, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
      if (mask) {
        HIPCK(hipMemcpy(M.p, mask, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
      } else {
        launch_fill_int(st, M.p, n, 15);
      }
    }
    return AVBD_OK;
  }
};

}  // namespace

// ---- avbd_mlp_grad without allocation ------------------------------------
// Staging for avbd_mlp_grad_noalloc in one static device array per GPU (in
// floats): samples x [kNoAllocMaxSamples][24], targets [..][4], masks [..],
// parameters, the gradient, the loss and the per-workgroup gradient rows,
// each region 256-byte aligned.
namespace {
constexpr int kNoAllocMaxSamples = 4096;
constexpr int kNoAllocMaxParams = 8192;
constexpr size_t kScrX = 0;
constexpr size_t kScrY = kScrX + (size_t)kNoAllocMaxSamples * 24;
constexpr size_t kScrM = kScrY + (size_t)kNoAllocMaxSamples * 4;
constexpr size_t kScrP = kScrM + (size_t)kNoAllocMaxSamples;
constexpr size_t kScrG = kScrP + kNoAllocMaxParams;
constexpr size_t kScrL = kScrG + kNoAllocMaxParams;
constexpr size_t kScrQ = kScrL + 64;        // per-workgroup gradient rows
constexpr size_t kScrFloats = kScrQ + (size_t)(kNoAllocMaxSamples / 64) * kNoAllocMaxParams;
// 659,520 floats = 2,638,080 bytes
static_assert(kScrY % 64 == 0 && kScrM % 64 == 0 && kScrP % 64 == 0 && kScrG % 64 == 0 &&
              kScrL % 64 == 0 && kScrQ % 64 == 0, "scratch regions stay 256-byte aligned");
}  // namespace

// The name `scratch` is a host-side stand-in in host code: its device
// address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) float scratch[kScrFloats];

namespace {
constexpr int kScrMaxDevices = 64;
std::mutex g_scr_mu;
float* g_scr_ptr[kScrMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current): looked
// up at the first call on that device and kept.
int scratch_ptr(int device, float** out) {
  std::lock_guard<std::mutex> lock(g_scr_mu);
  if (device < 0 || device >= kScrMaxDevices) return set_err(AVBD_E_ARG, "device ordinal out of range");
  if (g_scr_ptr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return set_err(AVBD_E_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    g_scr_ptr[device] = static_cast<float*>(p);
  }
  *out = g_scr_ptr[device];
  return AVBD_OK;
}
}  // namespace

int avbd_mlp_forward(int device, int n, const float* x, const float* params, float* y) {
  if (!y) return set_err(AVBD_E_ARG, "null output");
  MlpJob j;
  CK(j.open(device, n, x, nullptr, nullptr, params));
  CK(j.Y.alloc((size_t)n * 4));
  launch_mlp_infer(j.st, n, j.X.p, j.P.p, 0, nullptr, nullptr, 1.0f, 1.0f, 1, j.Y.p);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(y, j.Y.p, (size_t)n * 4 * sizeof(float), hipMemcpyDeviceToHost, j.st));
  HIPCK(hipStreamSynchronize(j.st));
  return AVBD_OK;
}

int avbd_mlp_grad(int device, int n, const float* x, const float* y, const int* mask,
                  const float* params, float huber, float* grad, float* loss) {
  if (!y || !grad || !loss) return set_err(AVBD_E_ARG, "null argument");
  MlpJob j;
  CK(j.open(device, n, x, y, mask, params));
  const int nw = mlp_num_weights();
  CK(j.G.alloc(nw));
  CK(j.L.alloc(1));
  CK(j.Q.alloc((size_t)((n + 63) / 64) * mlp_part_stride()));
  launch_mlp_train_tile(j.st, n, nullptr, j.X.p, j.Y.p, j.M.p, j.P.p, 1.0f / (float)n, huber, j.Q.p,
                        j.G.p, j.L.p);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(grad, j.G.p, nw * sizeof(float), hipMemcpyDeviceToHost, j.st));
  HIPCK(hipMemcpyAsync(loss, j.L.p, sizeof(float), hipMemcpyDeviceToHost, j.st));
  HIPCK(hipStreamSynchronize(j.st));
  return AVBD_OK;
}

int avbd_mlp_grad_noalloc(int device, int n, const float* x, const float* y, const int* mask,
                          const float* params, float huber, float* grad, float* loss) {
  if (!x || !y || !params || !grad || !loss) return set_err(AVBD_E_ARG, "null argument");
  if (n <= 0) return set_err(AVBD_E_ARG, "empty sample set");
  if (n > kNoAllocMaxSamples)
    return set_err(AVBD_E_ARG, "avbd_mlp_grad_noalloc: " + std::to_string(n) +
                                   " samples exceed its static scratch limit of " +
                                   std::to_string(kNoAllocMaxSamples) +
                                   " (avbd_mlp_grad takes any count)");
  const int np = mlp_num_params(), nw = mlp_num_weights();
  if (np > kNoAllocMaxParams || nw > kNoAllocMaxParams || mlp_part_stride() > kNoAllocMaxParams)
    return set_err(AVBD_E_STATE, "network larger than the static scratch layout");
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0)
    return set_err(AVBD_E_HIP, "no HIP device available");
  if (device < 0 || device >= nd) return set_err(AVBD_E_ARG, "device ordinal out of range");
  HIPCK(hipSetDevice(device));
  float* base = nullptr;
  CK(scratch_ptr(device, &base));
  float* X = base + kScrX;
  float* Y = base + kScrY;
  int* M = reinterpret_cast<int*>(base + kScrM);
  float* P = base + kScrP;
  float* G = base + kScrG;
  float* L = base + kScrL;
  float* Q = base + kScrQ;
  // the null stream: creating a stream would allocate; the copies below are
  // ordered with it
  hipStream_t st = nullptr;
  HIPCK(hipMemcpy(X, x, (size_t)n * 24 * sizeof(float), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(Y, y, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(P, params, (size_t)np * sizeof(float), hipMemcpyHostToDevice));
  if (mask) {
    HIPCK(hipMemcpy(M, mask, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  } else {
    launch_fill_int(st, M, n, 15);
  }
  // every value read below is written first in this call (the array keeps
  // the previous call's contents): the rows of the ceil(n / 64) workgroups,
  // then the gradient and the loss, which are overwritten with their sums
  launch_mlp_train_tile(st, n, nullptr, X, Y, M, P, 1.0f / (float)n, huber, Q, G, L);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpy(grad, G, (size_t)nw * sizeof(float), hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(loss, L, sizeof(float), hipMemcpyDeviceToHost));
  return AVBD_OK;
}

int avbd_mlp_train(int device, int n, const float* x, const float* y, const int* mask,
                   float* params, int steps, int batch, float lr, float huber, unsigned seed,
                   float* loss_hist) {
  if (!y || steps < 0 || batch <= 0) return set_err(AVBD_E_ARG, "bad training arguments");
  MlpJob j;
  CK(j.open(device, n, x, y, mask, params));
  const int nw = mlp_num_weights();
  const int b = std::min(batch, n);
  // minibatches: consecutive slices of seeded shuffles of 0..n-1
  std::vector<int> order(n), idx((size_t)std::max(steps, 1) * b);
  std::iota(order.begin(), order.end(), 0);
  unsigned long long rng = 0x9E3779B97F4A7C15ull ^ seed;
  auto next = [&rng]() {  // splitmix64
    unsigned long long z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  int used = n;
  for (size_t k = 0; k < idx.size(); ++k) {
    if (used == n) {
      for (int i = n - 1; i > 0; --i) std::swap(order[i], order[next() % (unsigned)(i + 1)]);
      used = 0;
    }
    idx[k] = order[used++];
  }
  CK(j.I.alloc(idx.size()));
  HIPCK(hipMemcpy(j.I.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  DBuf<float> m, v;
  CK(j.G.alloc(nw));
  CK(m.alloc(nw));
  CK(v.alloc(nw));
  CK(j.L.alloc(std::max(steps, 1)));
  CK(j.Q.alloc((size_t)((b + 63) / 64) * mlp_part_stride()));
  launch_fill_int(j.st, reinterpret_cast<int*>(m.p), nw, 0);
  launch_fill_int(j.st, reinterpret_cast<int*>(v.p), nw, 0);
  launch_fill_int(j.st, reinterpret_cast<int*>(j.L.p), std::max(steps, 1), 0);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double p1 = 1.0, p2 = 1.0;
  for (int t = 0; t < steps; ++t) {
    p1 *= b1;
    p2 *= b2;
    launch_mlp_train_tile(j.st, b, j.I.p + (size_t)t * b, j.X.p, j.Y.p, j.M.p, j.P.p,
                          1.0f / (float)b, huber, j.Q.p, j.G.p, j.L.p + t);
    launch_adam(j.st, nw, j.P.p, j.G.p, m.p, v.p, lr, b1, b2, (float)(1.0 - p1), (float)(1.0 - p2),
                eps);
  }
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(params, j.P.p, mlp_num_params() * sizeof(float), hipMemcpyDeviceToHost, j.st));
  if (loss_hist && steps > 0)
    HIPCK(hipMemcpyAsync(loss_hist, j.L.p, steps * sizeof(float), hipMemcpyDeviceToHost, j.st));
  HIPCK(hipStreamSynchronize(j.st));
  return AVBD_OK;
}

}  // extern "C"
