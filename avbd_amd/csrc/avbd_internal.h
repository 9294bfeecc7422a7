// avbd_internal.h — device data layout and kernel launchers of libavbd_hip.
//
// HBM layout (all arrays owned by the solver handle):
//   bodies  pose[2*i] = {x, y, z, 0}, pose[2*i+1] = {qx, qy, qz, qw}   (32 B)
//           pose0 / poseI: same layout for x_initial / x_inertial
//           vel[2*i] = {v, 0}, vel[2*i+1] = {w, 0}; pvel: previous v, w
//           inert[i] = 12 floats: I_world row-major in 3 float4 (w unused)
//           bstat[i] = {mass, he.x, he.y, he.z}; Ilocal[9*i]; bfric[i]
//   contact groups (one record per normal row plus its tangent rows):
//           GroupGeo (64 B, rewritten only at emit) + GroupSt (48 B, mutable)
//   A reference row j maps to (group g, slot r) with j = geo[g].hdr.w + r.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "avbd_fp.h"

namespace avbd {

// Group kinds.
constexpr int kGFloor = 0;   // FLOOR normal [+ world x / z tangents]
constexpr int kGBox = 1;     // BOX_BOX normal [+ t, b tangents]
constexpr int kGPin = 2;     // one PIN row

// hdr.x bit fields
__host__ __device__ inline int g_kind(int h) { return h & 3; }
__host__ __device__ inline int g_nrows(int h) { return ((h >> 2) & 3) + 1; }
__host__ __device__ inline int g_cidx(int h) { return (h >> 4) & 3; }
__host__ __device__ inline int g_make(int kind, int nrows, int cidx) {
  return kind | ((nrows - 1) << 2) | (cidx << 4);
}

// bit casts usable on host and device
__host__ __device__ inline float i2f(int v) { float f; __builtin_memcpy(&f, &v, 4); return f; }
__host__ __device__ inline int f2i(float f) { int v; __builtin_memcpy(&v, &f, 4); return v; }

// state bits (GroupSt.s2.y)
constexpr int kActN = 1, kActT = 2, kActB = 4, kWasN = 8, kWasT = 16, kWasB = 32;

struct __align__(16) GroupGeo {
  int4 hdr;     // {kind|nrows|cidx, body_a, body_b (axis for PIN, -1 floor), row_base}
  float4 g0;    // {anchor (floor: (0, y, 0); box: n; pin: world pt), stiffness}
  float4 g1;    // {off_a, mu_d}
  float4 g2;    // {off_b, mu_s (pin: fracture)}
};
struct __align__(16) GroupSt {
  float4 s0;    // {lam_n, lam_t, lam_b, ac0_n}
  float4 s1;    // {k_n, k_t, k_b, ac0_t}
  float4 s2;    // {ac0_b, bits (int), -, -}
};

// Element count of a launch sized by capacity: with a device count nd
// (substeps captured as a graph keep every count on the device) the kernel
// covers min(n, *nd) elements, else n.
__device__ __forceinline__ int dev_count(int n, const int* nd) {
  if (nd == nullptr) return n;
  const int m = *nd;
  return m < n ? m : n;
}

// ---- solve kernels (kernels_solve.hip) ----
void launch_substep_begin(hipStream_t st, int nb, int n_groups, float4* pose, float4* pose0,
                          float4* poseI, float4* inert, const float4* vel, const float4* pvel,
                          const float4* bstat, const float* Ilocal, GroupGeo* geo, GroupSt* gst,
                          float dt, float gx, float gy, float gz, float alpha, float gamma,
                          int post, int do_prelude, unsigned* epoch,
                          const int* nd_groups = nullptr);
void launch_primal_color(hipStream_t st, int count, const int* color_bodies, int base,
                         float4* pose, const float4* poseI, const float4* inert,
                         const float4* bstat, const int* cpos, const int2* csr,
                         const int* pos_map, const GroupGeo* geo, const GroupSt* gst, float dt,
                         int post_pass, int variant);
void launch_primal_sweep(hipStream_t st, int n_colors, const int* cstart, const int* color_bodies,
                         const int* color, float4* pose, const float4* poseI, const float4* inert,
                         const float4* bstat, const int* cpos, const int2* csr,
                         const GroupGeo* geo, const GroupSt* gst, float dt, int post_pass,
                         int* done, const unsigned* epoch, int sweep, int* err, int bpw,
                         int fuse_dual, float beta, GroupSt* gst_w, int tail_groups,
                         const int* done_rd = nullptr, int spin_cap = 0,
                         const int* meta_ncolors = nullptr, const int* meta_cstart = nullptr,
                         const int* meta_tiles = nullptr, int max_tiles = 0,
                         const int* tail_nd = nullptr, bool same_color_order = false);
void launch_dual(hipStream_t st, int n_groups, const float4* pose, const GroupGeo* geo,
                 GroupSt* gst, float beta, float* resid, const int* nd = nullptr);
void launch_finalize(hipStream_t st, int nb, const float4* pose, const float4* pose0,
                     float4* vel, float4* pvel, const float4* bstat, float dt, float max_lin,
                     float max_ang);

// ---- contact kernels (kernels_contacts.hip) ----
void launch_aabb(hipStream_t st, int nb, const float4* pose, const float4* bstat, float margin,
                 float4* lo, float4* hi);
int bvh_build(hipStream_t st, int nb, const float4* lo, const float4* hi, void* ws,
              size_t ws_bytes);
size_t bvh_workspace_bytes(int nb);
// tlo/thi: the boxes the tree was built over (lo/hi, or per-scene shifted
// copies); scene: per-body scene index or null (one scene)
// slots: nb x kPartnerSlots ints, each body's first partners sorted by the
// count pass and copied by the write pass
// lanes: traversal lanes per body (1 = one lane walking the escape links; 8,
// 16, 64 = a group sharing an LDS stack; 0 = bvh_count_lanes(nb))
constexpr int kPartnerSlots = 64;
void bvh_count_pairs(hipStream_t st, int nb, const float4* lo, const float4* hi,
                     const float4* tlo, const float4* thi, const int* scene, const float4* bstat,
                     const void* ws, int* counts, int* slots, int lanes);
int bvh_count_lanes(int nb);
void bvh_write_pairs(hipStream_t st, int nb, const float4* lo, const float4* hi,
                     const float4* tlo, const float4* thi, const int* scene, const float4* bstat,
                     const void* ws, const int* counts, const int* slots, const int* offsets,
                     int cap, int* pa, int* pb);
void launch_scene_shift(hipStream_t st, int nb, const float4* lo, const float4* hi,
                        const int* scene, int g, float pitch, float4* slo, float4* shi);
void radix_sort_pairs(hipStream_t st, unsigned* keys, int* vals, unsigned* keys_tmp,
                      int* vals_tmp, int n, void* ws);
size_t radix_workspace_bytes(int n);
void launch_narrowphase(hipStream_t st, int n_pairs, const float4* pose, const float4* bstat,
                        const int* pair_a, const int* pair_b, float margin, int* overlap,
                        int* sat_idx, float4* n_hat, float* depth, int* mf_count, int* mf_refa,
                        float4* mf_off /* 8 per pair: 4 ref, 4 inc */, float4* mf_n,
                        const int* nd = nullptr);
void launch_emit_demand(hipStream_t st, int n_pairs, const int* pair_a, const int* pair_b,
                        const int* mf_count, const int* mf_refa, const float* bfric,
                        float default_mu, int* demand, int* rpc_out, const int* nd = nullptr);
void launch_emit_succ(hipStream_t st, int n_pairs, const int* mf_count, const int* rpc,
                      const int* demand_scan, int cap_rows_after_static, int* succ,
                      int* reserve, const int* nd = nullptr);
void launch_emit_write(hipStream_t st, int n_pairs, const int* pair_a, const int* pair_b,
                       const int* mf_refa, const float4* mf_n, const float4* mf_off,
                       const float* bfric, float default_mu, float mult, const int* rpc,
                       const int* succ, const int* demand_scan, const int* succ_scan,
                       int n_static_rows, int g_static, GroupGeo* geo, GroupSt* gst,
                       int hash_cap, const int* hkeys, const float4* hstate, const int* hage,
                       const int* nd = nullptr, const int* hash_cap_dev = nullptr);
// p[0 .. min(n, *nd)) = v (all n without nd)
void launch_fill_int(hipStream_t st, int* p, int n, int v, const int* nd = nullptr);
void launch_copy_int(hipStream_t st, int* dst, const int* src, int n, const int* nd = nullptr);
void launch_hash_row0(hipStream_t st, const GroupGeo* geo, const GroupSt* gst, int hash_cap,
                      int* hkeys, float4* hstate, int* hage);
void launch_hash_collect(hipStream_t st, int n_entries, int g_first, const GroupGeo* geo,
                         const GroupSt* gst, int hash_cap, int* hkeys, int* howner,
                         int* slot_of, float4* hstate, int* hage, const int* nd = nullptr,
                         const int* hash_cap_dev = nullptr);
// Device-side sizes of a graph-captured substep (slots of the solver's
// device counters; see capi.hip): hash table size for `entries` live
// contacts, derived counts after emission, the recolor gate.
void launch_hash_size(hipStream_t st, const int* entries, int hash_cap, int* live);

// ---- graph kernels (kernels_graph.hip) ----
// Exclusive scan of n ints; returns the total through *total_dev (device int).
size_t scan_workspace_bytes(int n);
// Exclusive scan of n ints (in == out allowed). ws: scan_workspace_bytes(n),
// zero when first used and not shared with another stream; err (may be null)
// is set to 1 if a tile's look-back hit its spin cap. nd: device count (the
// first min(n, *nd) entries are scanned).
void exclusive_scan(hipStream_t st, const int* in, int* out, int n, int* total_dev, void* ws,
                    int* err, const int* nd = nullptr);
void launch_tcsr_pairs(hipStream_t st, int kept, const int* pa, const int* pb, const int* succ,
                       int* asb_cnt, int* asb_groups, int* start_a, int* end_a,
                       const int* nd = nullptr);
void launch_tcsr_asb_fill(hipStream_t st, int kept, const int* pb, const int* succ,
                          const int* asb_start, int* asb_cur, int* asb_list,
                          const int* nd = nullptr);
void launch_tcsr_count(hipStream_t st, int nb, const int* cvals, const int* st_count,
                       const int* asb_groups, const int* start_a, const int* end_a,
                       const int* sscan, const int* succ, int* padc, int* pos_of);
void launch_tcsr_posmap(hipStream_t st, int nb, const int* cvals, const int* pos_of, int* pos_map);
void launch_tcsr_fill(hipStream_t st, int nb, const int* cvals, const int* st_start,
                      const int* st_idx, const int* asb_start, int* asb_list, const int* start_a,
                      const int* end_a, const int* sscan, const int* succ, const int* pa,
                      const int* pb, int g_static, const int* cpos, int2* csr);
// Rows of the contact groups whose two bodies share a color, counted per
// pair: pair p emitted succ[p] groups of rpc[p] rows. scene / scene_dirty
// (batched scenes, else null): a conflict marks its scene.
void launch_count_conflicts(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                            const int* succ, const int* rpc, const int* color, int* out,
                            const int* scene, int* scene_dirty, const int* nd = nullptr);
void launch_color_reset(hipStream_t st, int nb, int* color, const int* scene,
                        const int* scene_dirty, const int* nd = nullptr);
void launch_adj_count(hipStream_t st, int n_dyn, int g_first, const GroupGeo* geo,
                      int* counts);
void launch_adj_fill(hipStream_t st, int n_dyn, int g_first, const GroupGeo* geo,
                     const int* starts, int* cursor, int* nbr);
void launch_adj_count_pairs(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                            const int* succ, int* counts, const int* nd = nullptr);
void launch_adj_fill_pairs(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                           const int* succ, const int* starts, int* cursor, int* nbr,
                           const int* nd = nullptr);
// scene_dirty (batched scenes, else null): bodies of clean scenes keep their color
void launch_color_greedy(hipStream_t st, int nb, const int* order, const int* pos,
                         const int* nbr_start, const int* nbr, int* color_in_out, int* pub,
                         int* err, int* ticket, int mode_jp, int* clamped, const int* scene,
                         const int* scene_dirty, const int* nd = nullptr);
void launch_color_resolve(hipStream_t st, int nb, const int* nbr_start, const int* nbr,
                          const float* prio, const int* color, int* color_next, int* uncolored,
                          const int* nd = nullptr);
// gate (may be null): a device count; 0 leaves counts and colors as they are
void launch_color_finish(hipStream_t st, int nb, int* color, int* counts,
                         const int* gate = nullptr);
// Jacobi rounds 2 .. MAX_COLORS of a recolor, on one workgroup, when round 1
// left bodies uncolored (*uncolored > 0); a no-op otherwise.
void launch_color_jacobi_tail(hipStream_t st, int nb, const int* nbr_start, const int* nbr,
                              const float* prio, int* color, int* color_next, int* uncolored,
                              int* clamped, const int* gate);
// Color starts, achieved color count and tiles of `bpw` bodies per color
// from the per-color counts (the layout of the solver's device counters).
void launch_color_meta(hipStream_t st, const int* counts, int bpw, int* cstart, int* n_active,
                       int* tile_start);
void launch_color_scatter(hipStream_t st, int nb, const int* color, const int* starts,
                          int* cursor, int* bodies);
void launch_sort_segments(hipStream_t st, int nb, const int* starts, int* idx);
int color_greedy_blocks(int device);

// learned warm start (kernels_learn.hip)
int mlp_num_params();
int mlp_num_weights();
void launch_contact_features(hipStream_t st, int n_dyn, int g_first, const GroupGeo* geo,
                             const GroupSt* gst, const float4* pose, const float4* vel,
                             const float4* bstat, float* feat);
void launch_contact_targets(hipStream_t st, int n_dyn, int g_first, const GroupGeo* geo,
                            const GroupSt* gst, const float* feat, float ag, float gamma,
                            float* targ, int* ymask);
void launch_mlp_infer(hipStream_t st, int n, const float* feat, const float* params, int g_first,
                      GroupSt* gst, const GroupGeo* geo, float ag, float gamma, int mode,
                      float* yout);
// Gradient and loss of n samples: `part` holds ceil(n / 64) rows of
// mlp_part_stride() floats (one per workgroup); grad and *loss are
// overwritten with the rows' sums.
int mlp_part_stride();
void launch_mlp_train_tile(hipStream_t st, int n, const int* idx, const float* X, const float* Y,
                           const int* ymask, const float* params, float inv_total, float huber,
                           float* part, float* grad, float* loss);
void launch_adam(hipStream_t st, int n, float* w, const float* g, float* m, float* v, float lr,
                 float b1, float b2, float c1, float c2, float eps);

}  // namespace avbd
