// kernels_particles.hip — the 3-DOF particle solver step on gfx950 and its
// C ABI (avbd_p_*, declared in include/avbd.h). Reference: Solver.step
// (src/avbd3d/solver.py:1124-1423) and its kernels (src/avbd3d/kernels.py).
//
// The host (avbd_amd/solver.py) owns the scene, the body coloring and the
// body -> row / body -> tet lists; the library holds the device state and
// issues one step: predict, warm start, alpha C0, then per iteration one
// primal launch per color (over that color's bodies only), the dual and the
// tet dual, and the velocity finalize.
//
// Layout: bodies are stored color by color (device index = position in the
// color-bucketed order), so a color's bodies are one contiguous index range
// and the primal indexes x, mass and its item lists without an indirection;
// rows, tets and items name bodies by device index, and the host-facing
// reads and writes map between the two orders. A row is one 64-byte record
// (PRow) holding its static fields and its state, so the primal's gather of
// a listed row touches one half cache line; a tet is an int4 of vertices
// plus a 16-byte state record. Rows keep the reference's indices. A contact's friction rows follow their normal
// row and are updated in the normal row's lane after it, which is the order
// a serial dual gives them (a friction row's bound reads its normal's new
// lambda).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/avbd.h"
#include "particles.h"

using namespace avbd;
using namespace avbd::particles;

namespace {

thread_local std::string p_err;
int perr(int code, const std::string& m) {
  p_err = m;
  return code;
}

#define PHIP(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return perr(AVBD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define PCK(expr)                 \
  do {                            \
    int r_ = (expr);              \
    if (r_ != AVBD_OK) return r_; \
  } while (0)

template <class T>
struct PBuf {
  T* p = nullptr;
  size_t n = 0;
  PBuf() = default;
  PBuf(const PBuf&) = delete;
  PBuf& operator=(const PBuf&) = delete;
  ~PBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int upload(const T* src, size_t count) {
    release();
    const size_t c = count ? count : 1;
    if (hipMalloc(&p, c * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return perr(AVBD_E_NOMEM, "hipMalloc failed");
    }
    n = c;
    if (count && src && hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return perr(AVBD_E_HIP, "hipMemcpy (upload) failed");
    return AVBD_OK;
  }
};

// One row: meta {type, body_a, body_b, sibling}, ar {anchor, rest},
// lim {stiffness, fmin, fmax, friction}, st {lambda, penalty, alpha C0,
// active (1 or 0)}.
struct alignas(64) PRow {
  int4 meta;
  float4 ar;
  float4 lim;
  float4 st;
};
static_assert(sizeof(PRow) == 64, "row record is 64 bytes");

__device__ __forceinline__ v3 ld3(const float4* a, int i) {
  const float4 t = a[i];
  return mk3(t.x, t.y, t.z);
}
__device__ __forceinline__ float4 st3(v3 v) { return make_float4(v.x, v.y, v.z, 0.0f); }
__device__ __forceinline__ bool is_pin(int t) {
  return t == kRowPinX || t == kRowPinY || t == kRowPinZ;
}
__device__ __forceinline__ bool has_offsets(int t) {
  return t == kRowSphere || t == kRowSphereBox || t == kRowBoxBox;
}
// body-local contact offsets of a body-body contact row (offs[2j], offs[2j+1])
__device__ __forceinline__ void row_offsets(const float4* __restrict__ offs, int j, int t, v3& oa,
                                            v3& ob) {
  if (has_offsets(t)) {
    oa = ld3(offs, 2 * j);
    ob = ld3(offs, 2 * j + 1);
  } else {
    oa = mk3(0.0f, 0.0f, 0.0f);
    ob = oa;
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// predict_inertial (kernels.py:161-207)
__global__ void __launch_bounds__(256)
k_p_predict(int nb, float4* __restrict__ x, const float4* __restrict__ v,
            const float4* __restrict__ prev_v, const float* __restrict__ mass, float dt, float gx,
            float gy, float gz, float4* __restrict__ initial, float4* __restrict__ inertial,
            int* __restrict__ epoch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) epoch[0] = epoch[0] + 1;  // step counter of the dataflow sweeps' stamps
  if (i >= nb) return;
  const v3 xi = ld3(x, i);
  initial[i] = st3(xi);
  if (mass[i] <= 0.0f) {
    inertial[i] = st3(xi);
    return;
  }
  const v3 g = mk3(gx, gy, gz);
  const v3 g_dt2 = mul(g, dt * dt);
  const v3 vi = ld3(v, i);
  inertial[i] = st3(add(add(xi, mul(vi, dt)), g_dt2));
  const v3 accel = divs(sub(vi, ld3(prev_v, i)), dt);
  const float g_norm = length(g);
  float w = 0.0f;
  if (g_norm > 0.0f) {
    const v3 g_hat = divs(g, g_norm);
    const float accel_ext = dot(accel, g_hat);
    w = clampf_(accel_ext / g_norm, 0.0f, 1.0f);
  }
  x[i] = st3(add(add(xi, mul(vi, dt)), mul(g_dt2, w)));
}

// warmstart_duals (kernels.py:210-236)
__global__ void __launch_bounds__(256)
k_p_warmstart(int nc, PRow* __restrict__ rows, float alpha, float gamma, int post) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nc) return;
  const int t = rows[j].meta.x;
  const float stiff = rows[j].lim.x;
  float4 st = rows[j].st;
  warmstart_row(t, stiff, alpha, gamma, post, st.x, st.y);
  rows[j].st = st;
}

// tet_warmstart_duals (kernels.py:815-836); tst = {rest, alpha C0, penalty, active}
__global__ void __launch_bounds__(256)
k_p_tet_warmstart(int nt, float4* __restrict__ tst, float* __restrict__ lam,
                  const float* __restrict__ stiff, float alpha, float gamma, int post) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  float p = clampf_(tst[t].z * gamma, kPenaltyMinTet, kPenaltyMax);
  if (post == 0) lam[t] = lam[t] * alpha * gamma;
  const float s = stiff[t];
  if (!(s != s) && s < kInf) p = fmin_(p, s);
  tst[t].z = p;
}

// cache_alpha_C0 (kernels.py:740-802)
__global__ void __launch_bounds__(256)
k_p_cache_c0(int nc, const float4* __restrict__ x, const float4* __restrict__ initial,
             PRow* __restrict__ rows, const float4* __restrict__ offs, float alpha) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nc) return;
  if (rows[j].st.w == 0.0f) {
    rows[j].st.z = 0.0f;
    return;
  }
  const int4 m = rows[j].meta;
  const float4 ar = rows[j].ar;
  const int t = m.x;
  const bool has_b = m.z >= 0 && !is_pin(t);
  const v3 an = mk3(ar.x, ar.y, ar.z);
  v3 oa, ob;
  row_offsets(offs, j, t, oa, ob);
  if (t == kRowTangent) {  // frame-start positions, alpha 1 (kernels.py:785-800)
    rows[j].st.z = row_C(t, ld3(initial, m.y), has_b ? ld3(initial, m.z) : mk3(0.f, 0.f, 0.f),
                         has_b, an, ar.w, oa, ob);
    return;
  }
  const float C0 = row_C(t, ld3(x, m.y), has_b ? ld3(x, m.z) : mk3(0.f, 0.f, 0.f), has_b, an, ar.w,
                         oa, ob);
  rows[j].st.z = alpha * C0;
}

// tet_cache_alpha_C0 (kernels.py:838-864)
__global__ void __launch_bounds__(256)
k_p_tet_cache_c0(int nt, const float4* __restrict__ x, const int4* __restrict__ tv,
                 float4* __restrict__ tst, float alpha) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const float4 ts = tst[t];
  if (ts.w == 0.0f) {
    tst[t].y = 0.0f;
    return;
  }
  const int4 q = tv[t];
  const float V = tet_volume(ld3(x, q.x), ld3(x, q.y), ld3(x, q.z), ld3(x, q.w));
  const float C0 = V / ts.x - 1.0f;
  tst[t].y = alpha * C0;
}

// One listed row's J, force and penalty seen from body i (kernels.py:395-530);
// false for an inactive row.
__device__ __forceinline__ bool row_item(const PRow* __restrict__ rows,
                                         const float4* __restrict__ offs,
                                         const float4* __restrict__ x, int j, int i, v3 xi,
                                         v3& J, float& f, float& kp) {
  const PRow r = rows[j];
  if (r.st.w == 0.0f) return false;
  const int t = r.meta.x;
  const bool has_b = r.meta.z >= 0 && !is_pin(t);
  const bool i_is_a = r.meta.y == i;
  const int other = has_b ? (i_is_a ? r.meta.z : r.meta.y) : i;
  v3 oa, ob;
  row_offsets(offs, j, t, oa, ob);
  float C;
  row_CJ(t, i_is_a, xi, ld3(x, other), has_b, mk3(r.ar.x, r.ar.y, r.ar.z), r.ar.w, oa, ob, C, J);
  const bool is_t = t == kRowTangent;
  kp = r.st.y;
  f = row_force(C, r.lim.x, r.st.z, r.st.x, kp, is_t, r.lim.w,
                is_t ? rows[r.meta.w].st.x : 0.0f, r.lim.y, r.lim.z);
  return true;
}

// One listed tet's J, force and penalty for its vertex `slot`
// (kernels.py:552-594); false for an inactive tet.
__device__ __forceinline__ bool tet_item(const int4* __restrict__ tv,
                                         const float4* __restrict__ tst,
                                         const float4* __restrict__ x, int enc, v3& J, float& f,
                                         float& kp) {
  const int tid = enc >> 2, slot = enc & 3;
  const float4 ts = tst[tid];
  if (ts.w == 0.0f) return false;
  const int4 q = tv[tid];
  kp = ts.z;
  tet_jac_force(ld3(x, q.x), ld3(x, q.y), ld3(x, q.z), ld3(x, q.w), slot, ts.x, ts.y, kp, J, f);
  return true;
}

// primal_update for the bodies of one color (kernels.py:335-608): one lane
// per body, rows in the body's list order, then its tets, then the 3x3 solve.
__global__ void __launch_bounds__(256)
k_p_primal(int count, int first, float4* __restrict__ x,
           const float4* __restrict__ inertial, const float* __restrict__ mass,
           const PRow* __restrict__ rows, const float4* __restrict__ offs,
           const int* __restrict__ con_start, const int* __restrict__ con_idx,
           const int4* __restrict__ tv,
           const float4* __restrict__ tst, const int* __restrict__ tet_start,
           const int* __restrict__ tet_idx, float dt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int i = first + k;
  const float m = mass[i];
  if (m <= 0.0f) return;
  const float mdt2 = m * (1.0f / (dt * dt));
  m33 lhs = mzero();
  lhs.m[0][0] = mdt2;
  lhs.m[1][1] = mdt2;
  lhs.m[2][2] = mdt2;
  const v3 xi = ld3(x, i);
  v3 rhs = mul(sub(xi, ld3(inertial, i)), mdt2);
  v3 J;
  float f, kp;
  for (int e = con_start[i]; e < con_start[i + 1]; ++e)
    if (row_item(rows, offs, x, con_idx[e], i, xi, J, f, kp))
      add_terms(lhs, rhs, make_terms(J, f, kp));
  for (int e = tet_start[i]; e < tet_start[i + 1]; ++e)
    if (tet_item(tv, tst, x, tet_idx[e], J, f, kp)) add_terms(lhs, rhs, make_terms(J, f, kp));
  x[i] = st3(sub(xi, mvmul(minverse(lhs), rhs)));
}

// The same primal update with G lanes per body: the lanes evaluate the
// body's rows and tets (the costly part: gathers and the row/tet math) in
// parallel into LDS, then lanes 0..8 of the group each add one entry of
// (rhs, lhs) over the items in list order, so every entry is summed in the
// serial order of kernels.py:400-599 and rounds exactly as it does there;
// lane 0 then solves. A skipped item (inactive row or tet) is stored as
// -0.0 terms: x + (-0.0) == x bit for bit for every x (signed zeros, inf
// and NaN included), so the sums take no branch. Items are taken in chunks
// of kChunk.
constexpr int kChunk = 64;

// A body's items in slot order (the color-bucketed body order), one int4
// each, rows first then tets, each in the body's list order:
//   row: {-(row + 1), other body (the body itself for one-body rows),
//         the normal row of a friction row (else -1), 0}
//   tet: {(tet << 2) | slot, the other three vertices in slot order}
// so the row record or tet state and the other bodies' positions are
// gathered side by side, one dependent load after the item.
// Dataflow sweep (all colors in one launch): a body's lower-color
// neighbours are exactly the device indices below its color's first index;
// it waits until each has published this sweep's stamp in `done` and reads
// its position through the agent-coherent path. Higher-color neighbours
// wait for this body before they move, so their positions are read as they
// stood at the start of the sweep.
struct PFlow {
  int* done;
  int stamp;
  int first;
  int* err;
};
constexpr int kPFlowSpinCap = 1 << 20;
constexpr int kPSweepStride = 4096;  // more sweeps a step than this: per-color launches

__device__ __forceinline__ float4 p_ld_agent(const float4* p) {
  const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
  const unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long b =
      __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float4(__uint_as_float((unsigned)a), __uint_as_float((unsigned)(a >> 32)),
                     __uint_as_float((unsigned)b), __uint_as_float((unsigned)(b >> 32)));
}
__device__ __forceinline__ void p_st_agent(float4* p, float4 v) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  __hip_atomic_store(q, (unsigned long long)__float_as_uint(v.x) |
                            ((unsigned long long)__float_as_uint(v.y) << 32),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, (unsigned long long)__float_as_uint(v.z) |
                                ((unsigned long long)__float_as_uint(v.w) << 32),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool FLOW>
__device__ __forceinline__ v3 other_pos(const float4* __restrict__ x, int d, const PFlow& fl) {
  if (!FLOW || d >= fl.first) return ld3(x, d);
  int spins = 0;
  while (__hip_atomic_load(fl.done + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != fl.stamp) {
    if (++spins > kPFlowSpinCap) {  // never resolved: flag it for the host and go on
      atomicExch(fl.err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  const float4 t = p_ld_agent(x + d);
  return mk3(t.x, t.y, t.z);
}

template <bool FLOW>
__device__ __forceinline__ bool slot_item(const PRow* __restrict__ rows,
                                          const float4* __restrict__ offs,
                                          const float4* __restrict__ tst,
                                          const float4* __restrict__ x, int4 it, int i, v3 xi,
                                          const PFlow& fl, v3& J, float& f, float& kp) {
  if (it.x < 0) {
    const int j = -it.x - 1;
    const PRow r = rows[j];
    const v3 xo = other_pos<FLOW>(x, it.y, fl);
    const float lam_sib = it.z >= 0 ? rows[it.z].st.x : 0.0f;
    if (r.st.w == 0.0f) return false;
    const int t = r.meta.x;
    const bool has_b = r.meta.z >= 0 && !is_pin(t);
    v3 oa, ob;
    row_offsets(offs, j, t, oa, ob);
    float C;
    row_CJ(t, r.meta.y == i, xi, xo, has_b, mk3(r.ar.x, r.ar.y, r.ar.z), r.ar.w, oa, ob, C, J);
    const bool is_t = t == kRowTangent;
    kp = r.st.y;
    f = row_force(C, r.lim.x, r.st.z, r.st.x, kp, is_t, r.lim.w, lam_sib, r.lim.y, r.lim.z);
    return true;
  }
  const int tid = it.x >> 2, slot = it.x & 3;
  const float4 ts = tst[tid];
  const v3 o1 = other_pos<FLOW>(x, it.y, fl), o2 = other_pos<FLOW>(x, it.z, fl),
           o3 = other_pos<FLOW>(x, it.w, fl);
  if (ts.w == 0.0f) return false;
  v3 p[4];
  int n = 0;
  for (int v = 0; v < 4; ++v) p[v] = v == slot ? xi : (n++ == 0 ? o1 : (n == 2 ? o2 : o3));
  kp = ts.z;
  tet_jac_force(p[0], p[1], p[2], p[3], slot, ts.x, ts.y, kp, J, f);
  return true;
}

// FLOW: one launch for all colors; block t is tile t of the color whose
// tile range holds it (tile_start / cstart, n_colors + 1 entries each). The
// sweep's stamp is step * kPSweepStride + sweep + 1 (mod 2^32), with the step
// counter read from `epoch` (advanced by the step's predict launch): two
// consecutive sweeps never share a stamp (fewer than kPSweepStride sweeps a
// step), graph replays included, and every body publishes every sweep, so a
// body's stamp from the previous sweep never passes for this one.
template <int G, bool FLOW>
__global__ void __launch_bounds__(256)
k_p_primal_grp(int count_arg, int first_arg, const int* __restrict__ slot_start,
               const int4* __restrict__ items, float4* __restrict__ x,
               const float4* __restrict__ inertial, const float* __restrict__ mass,
               const PRow* __restrict__ rows, const float4* __restrict__ offs,
               const float4* __restrict__ tst, float dt, int n_colors,
               const int* __restrict__ tile_start, const int* __restrict__ cstart,
               int* __restrict__ done, const int* __restrict__ epoch, int sweep, int* err) {
  constexpr int BPB = 256 / G;  // bodies per block
  __shared__ float terms[BPB][kChunk][9];
  __shared__ float acc[BPB][9];
  const int g = threadIdx.x % G, lb = threadIdx.x / G;
  int count = count_arg, first = first_arg, tile = blockIdx.x;
  PFlow fl{done, 0, 0, err};
  if (FLOW) {
    int c = 0;
    while (c + 1 < n_colors && tile >= tile_start[c + 1]) ++c;
    first = cstart[c];
    count = cstart[c + 1] - first;
    tile -= tile_start[c];
    fl.stamp = (int)((unsigned)*epoch * (unsigned)kPSweepStride + (unsigned)sweep + 1u);
    fl.first = first;
  }
  const int k = tile * BPB + lb;
  const bool live = k < count;
  const int i = first + (live ? k : 0);
  const int s0 = live ? slot_start[i] : 0, s1 = live ? slot_start[i + 1] : 0;
  const float m = live ? mass[i] : 0.0f;
  const bool dyn = live && m > 0.0f;
  const int n_items = dyn ? s1 - s0 : 0;
  const v3 xi = live ? ld3(x, i) : mk3(0.0f, 0.0f, 0.0f);
  // lane e < 9 owns entry e: rhs x, y, z, lhs 00 01 02 11 12 22
  float a = 0.0f;
  if (dyn && g < 9) {
    const float mdt2 = m * (1.0f / (dt * dt));
    const v3 rhs0 = mul(sub(xi, ld3(inertial, i)), mdt2);
    a = g == 0 ? rhs0.x : g == 1 ? rhs0.y : g == 2 ? rhs0.z
        : (g == 3 || g == 6 || g == 8) ? mdt2 : 0.0f;
  }
  for (int base = 0; __syncthreads_or(base < n_items); base += kChunk) {
    const int n_here = min(kChunk, n_items - base);
    for (int q = g; q < n_here; q += G) {
      v3 J;
      float f, kp;
      const bool ok = slot_item<FLOW>(rows, offs, tst, x, items[s0 + base + q], i, xi, fl, J,
                                      f, kp);
      float* dst = terms[lb][q];
      if (ok) {
        const Terms tm = make_terms(J, f, kp);
        dst[0] = tm.rhs.x;
        dst[1] = tm.rhs.y;
        dst[2] = tm.rhs.z;
        for (int e = 0; e < 6; ++e) dst[3 + e] = tm.l[e];
      } else {
        for (int e = 0; e < 9; ++e) dst[e] = -0.0f;
      }
    }
    __syncthreads();
    if (g < 9)
      for (int q = 0; q < n_here; ++q) a = a + terms[lb][q][g];
    __syncthreads();
  }
  if (g < 9) acc[lb][g] = a;
  __syncthreads();
  if (dyn && g == 0) {
    const float* c = acc[lb];
    m33 lhs;
    lhs.m[0][0] = c[3];
    lhs.m[0][1] = c[4];
    lhs.m[0][2] = c[5];
    lhs.m[1][0] = c[4];
    lhs.m[1][1] = c[6];
    lhs.m[1][2] = c[7];
    lhs.m[2][0] = c[5];
    lhs.m[2][1] = c[7];
    lhs.m[2][2] = c[8];
    const float4 xn = st3(sub(xi, mvmul(minverse(lhs), mk3(c[0], c[1], c[2]))));
    if (FLOW)
      p_st_agent(x + i, xn);
    else
      x[i] = xn;
  }
  if (FLOW && live && g == 0) {
    // publish after the position store has completed (a static body
    // publishes its unchanged position)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(done + i, fl.stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// dual_update of one row (kernels.py:611-701)
__device__ __forceinline__ void p_dual_one(int j, const float4* __restrict__ x,
                                           PRow* __restrict__ rows,
                                           const float4* __restrict__ offs,
                                           const float* __restrict__ fracture, float beta) {
  const PRow r = rows[j];
  if (r.st.w == 0.0f) return;
  const int t = r.meta.x;
  const bool has_b = r.meta.z >= 0 && !is_pin(t);
  v3 oa, ob;
  row_offsets(offs, j, t, oa, ob);
  const float C = row_C(t, ld3(x, r.meta.y), has_b ? ld3(x, r.meta.z) : mk3(0.f, 0.f, 0.f),
                        has_b, mk3(r.ar.x, r.ar.y, r.ar.z), r.ar.w, oa, ob);
  float l = r.st.x, p = r.st.y;
  const bool alive = dual_row(t, C, r.lim.x, r.st.z, l, p, r.lim.y, r.lim.z, fracture[j], r.lim.w,
                              t == kRowTangent ? rows[r.meta.w].st.x : 0.0f, beta);
  rows[j].st = make_float4(l, p, r.st.z, alive ? 1.0f : 0.0f);
}

// dual_update (kernels.py:611-701): a non-friction row's lane updates it and
// then the friction rows that follow it and name it as their sibling.
__device__ __forceinline__ void p_dual_lane(int j, int nc, const float4* __restrict__ x,
                                            PRow* __restrict__ rows,
                                            const float4* __restrict__ offs,
                                            const float* __restrict__ fracture, float beta) {
  if (j >= nc || rows[j].meta.x == kRowTangent) return;
  p_dual_one(j, x, rows, offs, fracture, beta);
  for (int k = j + 1; k < nc; ++k) {
    const int4 m = rows[k].meta;
    if (m.x != kRowTangent || m.w != j) break;
    p_dual_one(k, x, rows, offs, fracture, beta);
  }
}

// tet_dual_update (kernels.py:866-912)
__device__ __forceinline__ void p_tet_dual_lane(int t, int nt, const float4* __restrict__ x,
                                                const int4* __restrict__ tv,
                                                float4* __restrict__ tst, float* __restrict__ lam,
                                                const float* __restrict__ stiff,
                                                const float* __restrict__ fracture, float beta) {
  if (t >= nt) return;
  const float4 ts = tst[t];
  if (ts.w == 0.0f) return;
  const int4 q = tv[t];
  const float V = tet_volume(ld3(x, q.x), ld3(x, q.y), ld3(x, q.z), ld3(x, q.w));
  const float C = V / ts.x - 1.0f - ts.y;
  const float nl = ts.z * C;
  lam[t] = nl;
  if (fabs_(nl) >= fracture[t]) {
    lam[t] = 0.0f;
    tst[t] = make_float4(ts.x, ts.y, 0.0f, 0.0f);
    return;
  }
  const float upper = fmin_(kPenaltyMax, stiff[t]);
  tst[t].z = fmin_(ts.z + beta * fabs_(C), upper);
}

// Both duals of an iteration in one launch: the first row_blocks blocks take
// the rows, the rest the tets. They touch disjoint state, so the reference's
// two launches (solver.py:1318-1350) and this one give the same result.
__global__ void __launch_bounds__(256)
k_p_duals(int nc, int row_blocks, int nt, const float4* __restrict__ x, PRow* __restrict__ rows,
          const float4* __restrict__ offs, const float* __restrict__ fracture,
          const int4* __restrict__ tv, float4* __restrict__ tst, float* __restrict__ tet_lam,
          const float* __restrict__ tet_stiff, const float* __restrict__ tet_fracture,
          float beta) {
  if ((int)blockIdx.x < row_blocks)
    p_dual_lane(blockIdx.x * blockDim.x + threadIdx.x, nc, x, rows, offs, fracture, beta);
  else
    p_tet_dual_lane((blockIdx.x - row_blocks) * blockDim.x + threadIdx.x, nt, x, tv, tst, tet_lam,
                    tet_stiff, tet_fracture, beta);
}

// finalize_velocity + cap_velocity (kernels.py:707-737)
__global__ void __launch_bounds__(256)
k_p_finalize(int nb, const float4* __restrict__ x, const float4* __restrict__ initial,
             const float* __restrict__ mass, float dt, float4* __restrict__ v,
             float4* __restrict__ prev_v, int cap, float max_speed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  v3 vi = ld3(v, i);
  prev_v[i] = st3(vi);
  if (mass[i] > 0.0f) vi = divs(sub(ld3(x, i), ld3(initial, i)), dt);
  if (cap) {
    const float speed = length(vi);
    if (speed > max_speed) vi = mul(vi, max_speed / speed);
  }
  v[i] = st3(vi);
}

// Row state out to a contiguous staging array (for reads), and moved plane
// rows' {anchor, rest} in from one (for writes).
__global__ void __launch_bounds__(256)
k_p_rows_get_state(int nc, const PRow* __restrict__ rows, float4* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nc) out[j] = rows[j].st;
}

__global__ void __launch_bounds__(256)
k_p_rows_set_ar(int n, int first, const float4* __restrict__ in, PRow* __restrict__ rows) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) rows[first + k].ar = in[k];
}

constexpr int kB = 256;
inline dim3 grid(int n) { return dim3((n + kB - 1) / kB); }

}  // namespace

struct avbd_p_solver {
  int device = 0;
  hipStream_t st = nullptr;
  bool flushed = false;
  int nb = 0, nc = 0, nt = 0, n_colors = 0;
  std::vector<int> cstart;         // per-color device index ranges
  std::vector<int> order;          // device index -> body
  std::vector<int> pos;            // body -> device index
  std::vector<double> color_bytes;  // algorithmic bytes of one primal launch per color
  // one step captured as a graph, replayed while the buffers and the
  // parameters stay the same (a flush or new parameters capture again)
  bool use_graph = true;
  int primal_variant = 4;  // 0: lane per body, 1/2/3: 16/32/64 lanes per body,
                           // 4/5: 32/16 lanes per body, one dataflow launch per sweep
  bool graph_ok = false;
  avbd_p_params gparams{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // launch timing: one event pair per launch, per class (0 primal, 1 dual
  // and tet dual, 2 the rest)
  bool prof = false;
  std::vector<hipEvent_t> events;
  struct Pending {
    int cls;
    double bytes;
    hipEvent_t e0, e1;
  };
  std::vector<Pending> pending;
  size_t ev_next = 0;
  double prof_ms[3] = {0, 0, 0}, prof_bytes[3] = {0, 0, 0};
  long long prof_launch[3] = {0, 0, 0};
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_ok = false;
  }
  ~avbd_p_solver() {
    drop_graph();
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
  PBuf<float4> x, v, prev_v, initial, inertial;
  PBuf<float> mass, fracture;
  PBuf<PRow> rows;
  PBuf<float4> offs;         // 2 per row: body-local contact offsets (off_a, off_b)
  PBuf<float4> stage;        // n_rows entries: row state out, {anchor, rest} in
  std::vector<float4> h_ar;  // host copy of every row's {anchor, rest} (plane rows move)
  PBuf<int> con_start, con_idx;  // per device index: its rows (row indices)
  PBuf<int> slot_start;          // per device index: its range of `items`
  // dataflow sweep: tile starts for 32 / 16 lanes per body, color starts,
  // per-body stamps, sweep counter, error flag
  PBuf<int> d_tile_start32, d_tile_start16, d_cstart, done, epoch, flow_err;
  int n_tiles32 = 0, n_tiles16 = 0;
  PBuf<int4> items;      // per-body rows and tets in slot order (k_p_primal_grp)
  PBuf<int4> tv;
  PBuf<float4> tst;  // {rest volume, alpha C0, penalty, active}
  PBuf<float> tet_lam, tet_stiff, tet_fracture;
  PBuf<int> tet_start, tet_idx;
};

namespace {

std::vector<float4> pack3(const float* a, int n) {
  std::vector<float4> o((size_t)n);
  for (int i = 0; i < n; ++i) o[i] = make_float4(a[3 * i], a[3 * i + 1], a[3 * i + 2], 0.0f);
  return o;
}

int prof_collect(avbd_p_solver* s) {
  if (s->pending.empty()) return AVBD_OK;
  PHIP(hipStreamSynchronize(s->st));
  for (const auto& q : s->pending) {
    float ms = 0.0f;
    PHIP(hipEventElapsedTime(&ms, q.e0, q.e1));
    s->prof_ms[q.cls] += ms;
    s->prof_bytes[q.cls] += q.bytes;
    s->prof_launch[q.cls] += 1;
  }
  s->pending.clear();
  s->ev_next = 0;
  return AVBD_OK;
}

hipEvent_t next_event(avbd_p_solver* s) {
  if (s->ev_next == s->events.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    s->events.push_back(e);
  }
  return s->events[s->ev_next++];
}

// Issues one step's launches on the solver's stream (timed per launch when
// profiling; captured into a graph otherwise).
int enqueue_step(avbd_p_solver* s, const avbd_p_params* p) {
  const int nb = s->nb, nc = s->nc, nt = s->nt;
  hipStream_t st = s->st;
  const float dt = p->dt;
  auto timed = [&](int cls, double bytes, auto&& launch) -> int {
    if (!s->prof) {
      launch();
      return AVBD_OK;
    }
    if (s->pending.size() >= 8192) PCK(prof_collect(s));
    hipEvent_t e0 = next_event(s), e1 = next_event(s);
    if (!e0 || !e1) return perr(AVBD_E_HIP, "hipEventCreate failed");
    PHIP(hipEventRecord(e0, st));
    launch();
    PHIP(hipEventRecord(e1, st));
    s->pending.push_back({cls, bytes, e0, e1});
    return AVBD_OK;
  };
  hipLaunchKernelGGL(k_p_predict, grid(nb), dim3(kB), 0, st, nb, s->x.p, s->v.p, s->prev_v.p,
                     s->mass.p, dt, p->gravity[0], p->gravity[1], p->gravity[2], s->initial.p,
                     s->inertial.p, s->epoch.p);
  const int post = p->post_stabilize ? 1 : 0;
  auto cache = [&](float alpha) {
    if (nc > 0)
      hipLaunchKernelGGL(k_p_cache_c0, grid(nc), dim3(kB), 0, st, nc, s->x.p, s->initial.p,
                         s->rows.p, s->offs.p, alpha);
    if (nt > 0)
      hipLaunchKernelGGL(k_p_tet_cache_c0, grid(nt), dim3(kB), 0, st, nt, s->x.p, s->tv.p,
                         s->tst.p, alpha);
  };
  if (nc > 0) {
    hipLaunchKernelGGL(k_p_warmstart, grid(nc), dim3(kB), 0, st, nc, s->rows.p, p->alpha,
                       p->gamma, post);
    hipLaunchKernelGGL(k_p_cache_c0, grid(nc), dim3(kB), 0, st, nc, s->x.p, s->initial.p,
                       s->rows.p, s->offs.p, p->alpha);
  }
  if (nt > 0) {
    hipLaunchKernelGGL(k_p_tet_warmstart, grid(nt), dim3(kB), 0, st, nt, s->tst.p, s->tet_lam.p,
                       s->tet_stiff.p, p->alpha, p->gamma, post);
    hipLaunchKernelGGL(k_p_tet_cache_c0, grid(nt), dim3(kB), 0, st, nt, s->x.p, s->tv.p,
                       s->tst.p, p->alpha);
  }
  const int I = p->iterations;
  const bool capv = p->max_speed > 0.0f && p->max_speed < kInf;
  auto finalize = [&]() {
    hipLaunchKernelGGL(k_p_finalize, grid(nb), dim3(kB), 0, st, nb, s->x.p, s->initial.p,
                       s->mass.p, dt, s->v.p, s->prev_v.p, capv ? 1 : 0, p->max_speed);
  };
  for (int it = 0; it < I + post; ++it) {
    if (post && it == I) cache(0.0f);
    const bool flow = s->primal_variant >= 4 && I + post < kPSweepStride;
    if (flow) {
      double bytes = 0.0;
      for (double b : s->color_bytes) bytes += b;
      const bool g32 = s->primal_variant == 4;
      auto fkern = g32 ? k_p_primal_grp<32, true> : k_p_primal_grp<16, true>;
      PCK(timed(0, bytes, [&] {
        hipLaunchKernelGGL(fkern, dim3(g32 ? s->n_tiles32 : s->n_tiles16), dim3(kB), 0, st, 0, 0,
                           s->slot_start.p, s->items.p, s->x.p, s->inertial.p, s->mass.p,
                           s->rows.p, s->offs.p, s->tst.p, dt, s->n_colors,
                           g32 ? s->d_tile_start32.p : s->d_tile_start16.p, s->d_cstart.p,
                           s->done.p, s->epoch.p, it, s->flow_err.p);
      }));
    }
    for (int c = 0; c < s->n_colors && !flow; ++c) {
      const int cnt = s->cstart[c + 1] - s->cstart[c];
      if (cnt <= 0) continue;
      PCK(timed(0, s->color_bytes[c], [&] {
        if (s->primal_variant == 0) {
          hipLaunchKernelGGL(k_p_primal, grid(cnt), dim3(kB), 0, st, cnt, s->cstart[c],
                             s->x.p, s->inertial.p, s->mass.p,
                             s->rows.p, s->offs.p, s->con_start.p, s->con_idx.p, s->tv.p,
                             s->tst.p, s->tet_start.p, s->tet_idx.p, dt);
        } else {
          const int G = s->primal_variant == 1 ? 16 : s->primal_variant == 2 ? 32 : 64;
          const int bpb = kB / G;
          auto kern = G == 16 ? k_p_primal_grp<16, false>
                      : G == 32 ? k_p_primal_grp<32, false> : k_p_primal_grp<64, false>;
          hipLaunchKernelGGL(kern, dim3((cnt + bpb - 1) / bpb), dim3(kB), 0, st, cnt,
                             s->cstart[c], s->slot_start.p, s->items.p, s->x.p, s->inertial.p,
                             s->mass.p, s->rows.p, s->offs.p, s->tst.p, dt, 0, nullptr, nullptr,
                             nullptr, nullptr, 0, nullptr);
        }
      }));
    }
    if ((nc > 0 || nt > 0) && it < I) {
      const int rb = (nc + kB - 1) / kB, tb = (nt + kB - 1) / kB;
      PCK(timed(1, 0.0, [&] {
        hipLaunchKernelGGL(k_p_duals, dim3(rb + tb), dim3(kB), 0, st, nc, rb, nt, s->x.p,
                           s->rows.p, s->offs.p, s->fracture.p, s->tv.p, s->tst.p, s->tet_lam.p,
                           s->tet_stiff.p, s->tet_fracture.p, p->beta);
      }));
    }
    if (it == I - 1 && nt == 0) finalize();
  }
  if (nt > 0) finalize();
  PHIP(hipGetLastError());
  return AVBD_OK;
}

}  // namespace

extern "C" {

const char* avbd_p_last_error(void) { return p_err.c_str(); }

int avbd_p_create(int device, avbd_p_solver** out) {
  int n = 0;
  if (!out) return perr(AVBD_E_ARG, "null output");
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return perr(AVBD_E_HIP, "no HIP device available");
  if (device < 0 || device >= n) return perr(AVBD_E_ARG, "device ordinal out of range");
  PHIP(hipSetDevice(device));
  avbd_p_solver* s = new avbd_p_solver();
  s->device = device;
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return perr(AVBD_E_HIP, "hipStreamCreate failed");
  }
  *out = s;
  return AVBD_OK;
}

int avbd_p_destroy(avbd_p_solver* s) {
  if (!s) return AVBD_OK;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->st);
  (void)hipStreamDestroy(s->st);
  delete s;
  return AVBD_OK;
}

int avbd_p_flush(avbd_p_solver* s, const avbd_p_scene* sc) {
  if (!s || !sc) return perr(AVBD_E_ARG, "null argument");
  if (sc->n_bodies < 0 || sc->n_rows < 0 || sc->n_tets < 0 || sc->n_colors < 0)
    return perr(AVBD_E_ARG, "negative size");
  const int nb = sc->n_bodies, nc = sc->n_rows, nt = sc->n_tets;
  for (int j = 0; j < nc; ++j) {
    const int t = sc->type[j], a = sc->body_a[j], b = sc->body_b[j];
    const bool pin = t == kRowPinX || t == kRowPinY || t == kRowPinZ;
    const bool two = t == kRowDistance || t == kRowSphere || t == kRowSphereBox || t == kRowBoxBox;
    if (a < 0 || a >= nb || (!pin && b >= nb) || (two && b < 0))
      return perr(AVBD_E_ARG, "row " + std::to_string(j) + " names a body out of range");
    if (t == kRowTangent) {
      // a friction row follows its normal row, possibly after other friction
      // rows of the same normal (add_floor_contact's layout)
      const int sib = sc->sibling[j];
      bool ok = sib >= 0 && sib < j && sc->type[sib] != kRowTangent;
      for (int m = sib + 1; ok && m < j; ++m)
        ok = sc->type[m] == kRowTangent && sc->sibling[m] == sib;
      if (!ok)
        return perr(AVBD_E_ARG, "tangent row " + std::to_string(j) +
                                    " does not follow its normal row");
    }
    if (t != kRowPinX && t != kRowPinY && t != kRowPinZ && t != kRowDistance && t != kRowFloor &&
        t != kRowTangent && t != kRowPlane && t != kRowSphere && t != kRowSphereBox &&
        t != kRowBoxBox)
      return perr(AVBD_E_ARG, "row " + std::to_string(j) + " has an unsupported type");
  }
  for (int i = 0; i < nb; ++i)
    if (sc->color[i] < 0 || sc->color[i] >= sc->n_colors)
      return perr(AVBD_E_ARG, "body color out of range");
  // body -> row and body -> tet lists: every entry the primal reads must
  // name a row / tet / vertex that exists
  if (sc->con_start[0] != 0 || sc->tet_start[0] != 0)
    return perr(AVBD_E_ARG, "body lists must start at 0");
  for (int i = 0; i < nb; ++i)
    if (sc->con_start[i + 1] < sc->con_start[i] || sc->tet_start[i + 1] < sc->tet_start[i])
      return perr(AVBD_E_ARG, "body lists are not monotone");
  for (int k = 0; k < sc->con_start[nb]; ++k)
    if (sc->con_idx[k] < 0 || sc->con_idx[k] >= nc)
      return perr(AVBD_E_ARG, "body -> row list names a row out of range");
  for (int k = 0; k < sc->tet_start[nb]; ++k)
    if (sc->tet_idx[k] < 0 || (sc->tet_idx[k] >> 2) >= nt)
      return perr(AVBD_E_ARG, "body -> tet list names a tet out of range");
  for (int k = 0; k < 4 * nt; ++k)
    if (sc->tet_v[k] < 0 || sc->tet_v[k] >= nb)
      return perr(AVBD_E_ARG, "tet vertex out of range");
  PHIP(hipSetDevice(s->device));
  PHIP(hipStreamSynchronize(s->st));
  s->drop_graph();
  s->nb = nb;
  s->nc = nc;
  s->nt = nt;
  s->n_colors = sc->n_colors;
  // algorithmic bytes of a primal launch of k_p_primal_grp: per body its
  // slot entry, x, inertial, mass, item bounds and the x store; per row item
  // the item, the row record, the other body's x for two-body rows and the
  // normal's lambda for friction rows; per tet item the item, the tet state
  // and the other three vertices' x
  s->color_bytes.assign(sc->n_colors, 0.0);
  for (int i = 0; i < nb; ++i) {
    double b = 4 + 12 + 12 + 4 + 8 + 12;
    for (int k = sc->con_start[i]; k < sc->con_start[i + 1]; ++k) {
      const int j = sc->con_idx[k], t = sc->type[j];
      b += 16 + 64;
      if (t == kRowTangent) b += 4;
      if (sc->body_b[j] >= 0 && t != kRowPinX && t != kRowPinY && t != kRowPinZ) b += 12;
    }
    b += (double)(sc->tet_start[i + 1] - sc->tet_start[i]) * (16 + 16 + 36);
    s->color_bytes[sc->color[i]] += b;
  }
  // bodies bucketed by color (ascending body index inside a color)
  s->cstart.assign(sc->n_colors + 1, 0);
  for (int i = 0; i < nb; ++i) s->cstart[sc->color[i] + 1]++;
  for (int c = 0; c < sc->n_colors; ++c) s->cstart[c + 1] += s->cstart[c];
  std::vector<int>& order = s->order;
  std::vector<int>& pos = s->pos;
  order.assign((size_t)nb, 0);
  pos.assign((size_t)nb, 0);
  {
    std::vector<int> cur(s->cstart.begin(), s->cstart.end() - 1);
    for (int i = 0; i < nb; ++i) {
      pos[i] = cur[sc->color[i]]++;
      order[pos[i]] = i;
    }
  }
  auto dev3 = [&](const float* a) {  // a body array in device order
    std::vector<float4> o((size_t)nb);
    for (int k = 0; k < nb; ++k)
      o[k] = make_float4(a[3 * order[k]], a[3 * order[k] + 1], a[3 * order[k] + 2], 0.0f);
    return o;
  };
  const auto x = dev3(sc->x), v = dev3(sc->v), pv = dev3(sc->prev_v);
  PCK(s->x.upload(x.data(), nb));
  PCK(s->v.upload(v.data(), nb));
  PCK(s->prev_v.upload(pv.data(), nb));
  PCK(s->initial.upload(x.data(), nb));
  PCK(s->inertial.upload(x.data(), nb));
  {
    std::vector<float> m((size_t)nb);
    for (int k = 0; k < nb; ++k) m[k] = sc->mass[order[k]];
    PCK(s->mass.upload(m.data(), nb));
  }
  std::vector<PRow> rows((size_t)nc);
  s->h_ar.assign((size_t)nc, make_float4(0.f, 0.f, 0.f, 0.f));
  for (int j = 0; j < nc; ++j) {
    PRow& r = rows[j];
    r.meta = make_int4(sc->type[j], pos[sc->body_a[j]],
                       sc->body_b[j] >= 0 ? pos[sc->body_b[j]] : sc->body_b[j], sc->sibling[j]);
    r.ar = make_float4(sc->anchor[3 * j], sc->anchor[3 * j + 1], sc->anchor[3 * j + 2],
                       sc->rest[j]);
    r.lim = make_float4(sc->stiffness[j], sc->fmin[j], sc->fmax[j], sc->friction[j]);
    r.st = make_float4(sc->lambda[j], sc->penalty[j], 0.0f, sc->active[j] != 0 ? 1.0f : 0.0f);
    s->h_ar[j] = r.ar;
  }
  PCK(s->rows.upload(rows.data(), nc));
  {
    std::vector<float4> off((size_t)2 * nc, make_float4(0.f, 0.f, 0.f, 0.f));
    for (int j = 0; j < nc; ++j) {
      if (sc->off_a)
        off[2 * j] = make_float4(sc->off_a[3 * j], sc->off_a[3 * j + 1], sc->off_a[3 * j + 2], 0.f);
      if (sc->off_b)
        off[2 * j + 1] =
            make_float4(sc->off_b[3 * j], sc->off_b[3 * j + 1], sc->off_b[3 * j + 2], 0.f);
    }
    PCK(s->offs.upload(off.data(), off.size()));
  }
  PCK(s->stage.upload(nullptr, nc));
  {
    std::vector<int> t32((size_t)sc->n_colors + 1, 0), t16((size_t)sc->n_colors + 1, 0);
    for (int c = 0; c < sc->n_colors; ++c) {
      const int cnt = s->cstart[c + 1] - s->cstart[c];
      t32[c + 1] = t32[c] + (cnt + 7) / 8;
      t16[c + 1] = t16[c] + (cnt + 15) / 16;
    }
    s->n_tiles32 = t32[sc->n_colors];
    s->n_tiles16 = t16[sc->n_colors];
    PCK(s->d_tile_start32.upload(t32.data(), t32.size()));
    PCK(s->d_tile_start16.upload(t16.data(), t16.size()));
    PCK(s->d_cstart.upload(s->cstart.data(), s->cstart.size()));
    const std::vector<int> zero((size_t)std::max(nb, 1), 0);
    PCK(s->done.upload(zero.data(), nb));
    PCK(s->epoch.upload(zero.data(), 1));
    PCK(s->flow_err.upload(zero.data(), 1));
  }
  PCK(s->fracture.upload(sc->fracture, nc));
  // per-device-index row and tet lists (each body's lists in their order)
  auto regroup = [&](const int* start, const int* idx, PBuf<int>& d_start,
                     PBuf<int>& d_idx) -> int {
    std::vector<int> st((size_t)nb + 1, 0), ix;
    ix.reserve((size_t)start[nb]);
    for (int k = 0; k < nb; ++k) {
      ix.insert(ix.end(), idx + start[order[k]], idx + start[order[k] + 1]);
      st[k + 1] = (int)ix.size();
    }
    PCK(d_start.upload(st.data(), nb + 1));
    return d_idx.upload(ix.data(), ix.size());
  };
  PCK(regroup(sc->con_start, sc->con_idx, s->con_start, s->con_idx));
  std::vector<int4> tv((size_t)nt);
  std::vector<float4> tst((size_t)nt);
  for (int t = 0; t < nt; ++t) {
    tv[t] = make_int4(pos[sc->tet_v[4 * t]], pos[sc->tet_v[4 * t + 1]],
                      pos[sc->tet_v[4 * t + 2]], pos[sc->tet_v[4 * t + 3]]);
    tst[t] = make_float4(sc->tet_rest_vol[t], 0.0f, sc->tet_penalty[t],
                         sc->tet_active[t] != 0 ? 1.0f : 0.0f);
  }
  PCK(s->tv.upload(tv.data(), nt));
  PCK(s->tst.upload(tst.data(), nt));
  PCK(s->tet_lam.upload(sc->tet_lambda, nt));
  PCK(s->tet_stiff.upload(sc->tet_stiffness, nt));
  PCK(s->tet_fracture.upload(sc->tet_fracture, nt));
  PCK(regroup(sc->tet_start, sc->tet_idx, s->tet_start, s->tet_idx));
  {
    std::vector<int> ss((size_t)nb + 1, 0);
    std::vector<int4> it;
    it.reserve((size_t)sc->con_start[nb] + (size_t)sc->tet_start[nb]);
    for (int k = 0; k < nb; ++k) {
      const int i = order[k];
      for (int e = sc->con_start[i]; e < sc->con_start[i + 1]; ++e) {
        const int j = sc->con_idx[e], t = sc->type[j];
        const bool has_b = sc->body_b[j] >= 0 && t != kRowPinX && t != kRowPinY && t != kRowPinZ;
        const int other = has_b ? (sc->body_a[j] == i ? sc->body_b[j] : sc->body_a[j]) : i;
        it.push_back(make_int4(-j - 1, pos[other], t == kRowTangent ? sc->sibling[j] : -1, 0));
      }
      for (int e = sc->tet_start[i]; e < sc->tet_start[i + 1]; ++e) {
        const int enc = sc->tet_idx[e], tid = enc >> 2, slot = enc & 3;
        int o[3], n = 0;
        for (int v = 0; v < 4; ++v)
          if (v != slot) o[n++] = pos[sc->tet_v[4 * tid + v]];
        it.push_back(make_int4(enc, o[0], o[1], o[2]));
      }
      ss[k + 1] = (int)it.size();
    }
    PCK(s->slot_start.upload(ss.data(), nb + 1));
    PCK(s->items.upload(it.data(), it.size()));
  }
  s->flushed = true;
  return AVBD_OK;
}

// Solver.step without self-collision (solver.py:1124-1400).
int avbd_p_step(avbd_p_solver* s, const avbd_p_params* p) {
  if (!s || !p) return perr(AVBD_E_ARG, "null argument");
  if (!s->flushed) return perr(AVBD_E_STATE, "step before flush");
  if (p->iterations < 0) return perr(AVBD_E_ARG, "negative iteration count");
  if (s->nb == 0) return AVBD_OK;
  PHIP(hipSetDevice(s->device));
  if (!s->use_graph || s->prof) return enqueue_step(s, p);
  if (!s->graph_ok || std::memcmp(&s->gparams, p, sizeof(*p)) != 0) {
    s->drop_graph();
    PHIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(s, p);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s->st, &g);
    if (rc != AVBD_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess)
      return perr(AVBD_E_HIP, std::string("stream capture: ") + hipGetErrorString(ec));
    s->graph = g;
    PHIP(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
    s->gparams = *p;
    s->graph_ok = true;
  }
  PHIP(hipGraphLaunch(s->gexec, s->st));
  return AVBD_OK;
}

int avbd_p_set_option(avbd_p_solver* s, int option, int value) {
  if (!s) return perr(AVBD_E_ARG, "null solver");
  if (option == AVBD_P_OPT_GRAPH) {
    s->use_graph = value != 0;
  } else if (option == AVBD_P_OPT_PRIMAL_VARIANT) {
    if (value < 0 || value > 5) return perr(AVBD_E_ARG, "primal variant must be 0 to 5");
    s->primal_variant = value;
  } else {
    return perr(AVBD_E_ARG, "unknown option");
  }
  s->drop_graph();
  return AVBD_OK;
}

int avbd_p_profile_enable(avbd_p_solver* s, int on) {
  if (!s) return perr(AVBD_E_ARG, "null solver");
  PCK(prof_collect(s));
  s->prof = on != 0;
  for (int k = 0; k < 3; ++k) {
    s->prof_ms[k] = 0.0;
    s->prof_bytes[k] = 0.0;
    s->prof_launch[k] = 0;
  }
  return AVBD_OK;
}

int avbd_p_profile_read(avbd_p_solver* s, double* ms, long long* launches, double* bytes) {
  if (!s) return perr(AVBD_E_ARG, "null solver");
  PHIP(hipSetDevice(s->device));
  PCK(prof_collect(s));
  for (int k = 0; k < 3; ++k) {
    if (ms) ms[k] = s->prof_ms[k];
    if (launches) launches[k] = s->prof_launch[k];
    if (bytes) bytes[k] = s->prof_bytes[k];
  }
  return AVBD_OK;
}

int avbd_p_read(avbd_p_solver* s, float* x, float* v, float* prev_v, float* lambda,
                float* penalty, int* active, float* tet_lambda, float* tet_penalty,
                int* tet_active) {
  if (!s || !s->flushed) return perr(AVBD_E_STATE, "read before flush");
  PHIP(hipSetDevice(s->device));
  PHIP(hipStreamSynchronize(s->st));
  {
    int e = 0;
    PHIP(hipMemcpy(&e, s->flow_err.p, sizeof(int), hipMemcpyDeviceToHost));
    if (e != 0) {
      e = 0;
      PHIP(hipMemcpy(s->flow_err.p, &e, sizeof(int), hipMemcpyHostToDevice));
      return perr(AVBD_E_HIP, "a dataflow primal sweep waited past its spin limit");
    }
  }
  std::vector<float4> tmp((size_t)std::max(s->nb, 1));
  auto unpack = [&](const PBuf<float4>& b, float* out) -> int {
    if (!out || s->nb == 0) return AVBD_OK;
    PHIP(hipMemcpy(tmp.data(), b.p, s->nb * sizeof(float4), hipMemcpyDeviceToHost));
    for (int k = 0; k < s->nb; ++k) {
      const int i = s->order[k];
      out[3 * i] = tmp[k].x;
      out[3 * i + 1] = tmp[k].y;
      out[3 * i + 2] = tmp[k].z;
    }
    return AVBD_OK;
  };
  PCK(unpack(s->x, x));
  PCK(unpack(s->v, v));
  PCK(unpack(s->prev_v, prev_v));
  if (s->nc > 0 && (lambda || penalty || active)) {
    std::vector<float4> st((size_t)s->nc);
    hipLaunchKernelGGL(k_p_rows_get_state, grid(s->nc), dim3(kB), 0, s->st, s->nc, s->rows.p,
                       s->stage.p);
    PHIP(hipGetLastError());
    PHIP(hipMemcpyAsync(st.data(), s->stage.p, s->nc * sizeof(float4), hipMemcpyDeviceToHost,
                        s->st));
    PHIP(hipStreamSynchronize(s->st));
    for (int j = 0; j < s->nc; ++j) {
      if (lambda) lambda[j] = st[j].x;
      if (penalty) penalty[j] = st[j].y;
      if (active) active[j] = st[j].w != 0.0f ? 1 : 0;
    }
  }
  if (s->nt > 0) {
    if (tet_lambda)
      PHIP(hipMemcpy(tet_lambda, s->tet_lam.p, s->nt * sizeof(float), hipMemcpyDeviceToHost));
    if (tet_penalty || tet_active) {
      std::vector<float4> ts((size_t)s->nt);
      PHIP(hipMemcpy(ts.data(), s->tst.p, s->nt * sizeof(float4), hipMemcpyDeviceToHost));
      for (int t = 0; t < s->nt; ++t) {
        if (tet_penalty) tet_penalty[t] = ts[t].z;
        if (tet_active) tet_active[t] = ts[t].w != 0.0f ? 1 : 0;
      }
    }
  }
  return AVBD_OK;
}

int avbd_p_write_bodies(avbd_p_solver* s, int field, int first, int n, const float* vals) {
  if (!s || !s->flushed) return perr(AVBD_E_STATE, "write before flush");
  if (first < 0 || n < 0 || first + n > s->nb || !vals) return perr(AVBD_E_ARG, "bad range");
  PBuf<float4>* dst = field == 0 ? &s->x : field == 1 ? &s->v : nullptr;
  if (!dst) return perr(AVBD_E_ARG, "field must be 0 (x) or 1 (v)");
  PHIP(hipSetDevice(s->device));
  PHIP(hipStreamSynchronize(s->st));
  if (n == 0) return AVBD_OK;
  // bodies [first, first + n) sit at scattered device indices: rewrite the
  // whole array
  std::vector<float4> all((size_t)s->nb);
  PHIP(hipMemcpy(all.data(), dst->p, s->nb * sizeof(float4), hipMemcpyDeviceToHost));
  for (int m = 0; m < n; ++m)
    all[s->pos[first + m]] = make_float4(vals[3 * m], vals[3 * m + 1], vals[3 * m + 2], 0.0f);
  PHIP(hipMemcpy(dst->p, all.data(), s->nb * sizeof(float4), hipMemcpyHostToDevice));
  return AVBD_OK;
}

int avbd_p_write_rows(avbd_p_solver* s, int first, int n, const float* anchor, const float* rest) {
  if (!s || !s->flushed) return perr(AVBD_E_STATE, "write before flush");
  if (first < 0 || n < 0 || first + n > s->nc) return perr(AVBD_E_ARG, "bad range");
  if (n == 0 || (!anchor && !rest)) return AVBD_OK;
  PHIP(hipSetDevice(s->device));
  PHIP(hipStreamSynchronize(s->st));
  for (int k = 0; k < n; ++k) {
    float4& a = s->h_ar[first + k];
    if (anchor) {
      a.x = anchor[3 * k];
      a.y = anchor[3 * k + 1];
      a.z = anchor[3 * k + 2];
    }
    if (rest) a.w = rest[k];
  }
  // the {anchor, rest} field of rows [first, first + n): staged, then
  // scattered into the 64-byte records
  PHIP(hipMemcpyAsync(s->stage.p, s->h_ar.data() + first, n * sizeof(float4),
                      hipMemcpyHostToDevice, s->st));
  hipLaunchKernelGGL(k_p_rows_set_ar, grid(n), dim3(kB), 0, s->st, n, first, s->stage.p,
                     s->rows.p);
  PHIP(hipGetLastError());
  PHIP(hipStreamSynchronize(s->st));
  return AVBD_OK;
}

}  // extern "C"
