// kernels_contacts.hip — contact generation for gfx950.
//
//  k_aabb           compute_body_aabb_6dof (kernels_6dof.py:1922-1943)
//  LBVH             replaces wp.Bvh + bvh_broadphase_pairs (solver_6dof.py:1819-1848,
//                   kernels_6dof.py:1946-1978): 30-bit Morton codes, LSD radix
//                   sort, Karras (2012) hierarchy, bottom-up refit, escape-link
//                   (stackless) traversal. The pair list comes out in canonical
//                   (a asc, b asc) order: per-body counts, exclusive scan,
//                   per-body segment sort.
//  k_narrowphase    obb_sat_pairs + obb_contact_manifold_6dof fused
//                   (kernels_6dof.py:1991-2130, 2221-2537); the clip polygons
//                   live in registers instead of a global scratch buffer.
//  k_emit_*         gpu_pool_emit_rows (kernels_6dof.py:2695-2912) with the
//                   serial row-reservation order restated as scans.
//  k_hash_*         gpu_pool_hash_clear / collect (kernels_6dof.py:2915-2995);
//                   duplicate keys resolve to the highest pool index, the
//                   entry a serial upsert leaves behind.
#include "avbd_internal.h"

namespace avbd {

void radix_sort_pairs(hipStream_t st, unsigned* keys, int* vals, unsigned* keys_tmp,
                      int* vals_tmp, int n, void* ws);
size_t radix_workspace_bytes(int n);

__device__ __forceinline__ v3 f3(float4 a) { return mk3(a.x, a.y, a.z); }
__device__ __forceinline__ q4 f4q(float4 a) { return mk4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ float4 p4(v3 a) { return make_float4(a.x, a.y, a.z, 0.0f); }
// Column k of R, picked with selects (a runtime index into the matrix would
// put it in scratch memory).
__device__ __forceinline__ v3 colm(const m33& R, int k) {
  return k == 0   ? mk3(R.m[0][0], R.m[1][0], R.m[2][0])
         : k == 1 ? mk3(R.m[0][1], R.m[1][1], R.m[2][1])
                  : mk3(R.m[0][2], R.m[1][2], R.m[2][2]);
}

// ---------------------------------------------------------------------------
// AABB
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_aabb(int nb, const float4* __restrict__ pose, const float4* __restrict__ bstat, float margin,
       float4* __restrict__ lo, float4* __restrict__ hi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  v3 c = f3(pose[2 * i]);
  float4 bs = bstat[i];
  m33 R = qtomat(f4q(pose[2 * i + 1]));
  float ex = ((fabs_(R.m[0][0]) * bs.y + fabs_(R.m[0][1]) * bs.z) + fabs_(R.m[0][2]) * bs.w) + margin;
  float ey = ((fabs_(R.m[1][0]) * bs.y + fabs_(R.m[1][1]) * bs.z) + fabs_(R.m[1][2]) * bs.w) + margin;
  float ez = ((fabs_(R.m[2][0]) * bs.y + fabs_(R.m[2][1]) * bs.z) + fabs_(R.m[2][2]) * bs.w) + margin;
  lo[i] = p4(sub(c, mk3(ex, ey, ez)));
  hi[i] = p4(add(c, mk3(ex, ey, ez)));
}

void launch_aabb(hipStream_t st, int nb, const float4* pose, const float4* bstat, float margin,
                 float4* lo, float4* hi) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_aabb, dim3((nb + 255) / 256), dim3(256), 0, st, nb, pose, bstat, margin,
                     lo, hi);
}

// Batched scenes: each scene's boxes moved to its own cell of a g x g x g grid
// of pitch `pitch` for the tree. Rounding is monotonic (x <= y implies
// fl(x + d) <= fl(y + d)), so two boxes of one scene that overlap still
// overlap after the shift: the tree prunes conservatively, and the leaf test
// runs on the unshifted boxes.
__global__ void __launch_bounds__(256)
k_scene_shift(int nb, const float4* __restrict__ lo, const float4* __restrict__ hi,
              const int* __restrict__ scene, int g, float pitch, float4* __restrict__ slo,
              float4* __restrict__ shi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int sc = scene[i];
  const float ox = (float)(sc % g) * pitch, oy = (float)((sc / g) % g) * pitch,
              oz = (float)(sc / (g * g)) * pitch;
  const float4 a = lo[i], b = hi[i];
  slo[i] = make_float4(a.x + ox, a.y + oy, a.z + oz, 0.0f);
  shi[i] = make_float4(b.x + ox, b.y + oy, b.z + oz, 0.0f);
}

void launch_scene_shift(hipStream_t st, int nb, const float4* lo, const float4* hi,
                        const int* scene, int g, float pitch, float4* slo, float4* shi) {
  if (nb > 0)
    hipLaunchKernelGGL(k_scene_shift, dim3((nb + 255) / 256), dim3(256), 0, st, nb, lo, hi, scene,
                       g, pitch, slo, shi);
}

// ---------------------------------------------------------------------------
// LBVH
// ---------------------------------------------------------------------------
struct BvhWs {
  int n;
  int* bounds;          // 6 ordered ints: min xyz, max xyz of centroids
  unsigned* keys;       // n
  int* ids;             // n  (sorted leaf -> body)
  unsigned* keys_tmp;   // n
  int* ids_tmp;         // n
  int* left;            // n-1
  int* right;           // n-1
  int* parent;          // 2n-1
  int* escape;          // 2n-1
  int2* range;          // n-1: first and last sorted leaf under each internal node
  float4* nlo;          // 2n-1
  float4* nhi;          // 2n-1
  float4* b1;           // 2 * ceil(n/64): boxes of 64-leaf blocks (lo, hi)
  float4* b2;           // 2 * ceil(n/4096): boxes of 64-block superblocks
  // traversal records: internal node k at pk[2k] (lo, left child) and
  // pk[2k + 1] (hi, escape), 32 B in one cache line instead of four arrays;
  // leaf k at lf[k] (body, escape)
  float4* pk;           // 2 (n-1)
  int2* lf;             // n
  void* radix_ws;
};

constexpr int kRefitB = 64;  // leaves per block; blocks per superblock

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static BvhWs bvh_carve(void* base, int n) {
  BvhWs w;
  w.n = n;
  char* p = (char*)base;
  auto take = [&](size_t bytes) { char* r = p; p += align256(bytes); return (void*)r; };
  int nn = 2 * n - 1 > 0 ? 2 * n - 1 : 1;
  int ni = n - 1 > 0 ? n - 1 : 1;
  w.bounds = (int*)take(8 * sizeof(int));
  w.keys = (unsigned*)take(n * sizeof(unsigned));
  w.ids = (int*)take(n * sizeof(int));
  w.keys_tmp = (unsigned*)take(n * sizeof(unsigned));
  w.ids_tmp = (int*)take(n * sizeof(int));
  w.left = (int*)take(ni * sizeof(int));
  w.right = (int*)take(ni * sizeof(int));
  w.parent = (int*)take(nn * sizeof(int));
  w.escape = (int*)take(nn * sizeof(int));
  w.range = (int2*)take(ni * sizeof(int2));
  w.nlo = (float4*)take(nn * sizeof(float4));
  w.nhi = (float4*)take(nn * sizeof(float4));
  const int nb1 = (n + kRefitB - 1) / kRefitB + 1, nb2 = (nb1 + kRefitB - 1) / kRefitB + 1;
  w.b1 = (float4*)take(2 * (size_t)nb1 * sizeof(float4));
  w.b2 = (float4*)take(2 * (size_t)nb2 * sizeof(float4));
  w.pk = (float4*)take(2 * (size_t)ni * sizeof(float4));
  w.lf = (int2*)take((size_t)n * sizeof(int2));
  w.radix_ws = (void*)p;
  return w;
}

size_t bvh_workspace_bytes(int n) {
  size_t nn = 2 * (size_t)n, b = 0;
  const size_t nb1 = ((size_t)n + kRefitB - 1) / kRefitB + 1, nb2 = (nb1 + kRefitB - 1) / kRefitB + 1;
  b += align256(8 * sizeof(int));
  b += 4 * align256(n * sizeof(int));
  b += 2 * align256(n * sizeof(int));
  b += align256(n * sizeof(int2));
  b += 2 * align256(nn * sizeof(int));
  b += 2 * align256(nn * sizeof(float4));
  b += align256(2 * nb1 * sizeof(float4)) + align256(2 * nb2 * sizeof(float4));
  b += align256(2 * (size_t)n * sizeof(float4)) + align256((size_t)n * sizeof(int2));
  b += radix_workspace_bytes(n);
  return b + 4096;
}

__device__ __forceinline__ int ord_f(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float unord_f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void k_bounds_init(int* b) {
  if (threadIdx.x < 3) b[threadIdx.x] = 0x7fffffff;
  else if (threadIdx.x < 6) b[threadIdx.x] = (int)0x80000000;
}

// Grid-stride over the bodies with a bounded number of workgroups: the
// same-address atomics of the final step serialise, so there are at most
// kBoundsBlocks x 6 of them.
constexpr int kBoundsBlocks = 240;

__global__ void __launch_bounds__(256)
k_bounds(int n, const float4* __restrict__ lo, const float4* __restrict__ hi, int* b) {
  __shared__ int smin[3][256], smax[3][256];
  int mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int mx[3] = {(int)0x80000000, (int)0x80000000, (int)0x80000000};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 l = lo[i], h = hi[i];
    float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    for (int k = 0; k < 3; ++k) {
      const int o = ord_f(c[k]);
      mn[k] = min(mn[k], o);
      mx[k] = max(mx[k], o);
    }
  }
  for (int k = 0; k < 3; ++k) { smin[k][threadIdx.x] = mn[k]; smax[k][threadIdx.x] = mx[k]; }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      for (int k = 0; k < 3; ++k) {
        smin[k][threadIdx.x] = min(smin[k][threadIdx.x], smin[k][threadIdx.x + s]);
        smax[k][threadIdx.x] = max(smax[k][threadIdx.x], smax[k][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) {
      atomicMin(&b[k], smin[k][0]);
      atomicMax(&b[3 + k], smax[k][0]);
    }
  }
}

__device__ __forceinline__ unsigned expand10(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void __launch_bounds__(256)
k_morton(int n, const float4* __restrict__ lo, const float4* __restrict__ hi,
         const int* __restrict__ b, unsigned* __restrict__ keys, int* __restrict__ ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 l = lo[i], h = hi[i];
  float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
  unsigned q[3];
  for (int k = 0; k < 3; ++k) {
    float mn = unord_f(b[k]), mx = unord_f(b[3 + k]);
    float ext = mx - mn;
    float u = ext > 0.0f ? (c[k] - mn) / ext : 0.0f;
    u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    q[k] = (unsigned)u;
  }
  keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  ids[i] = i;
}

__device__ __forceinline__ int bvh_delta(const unsigned* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  unsigned a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((unsigned)(i ^ j));
  return __clz(a ^ b);
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees
// and k-d trees": internal node i covers a key range found by a binary search.
// Node ids: internal 0..n-2 (root 0), leaf k -> n-1+k.
__global__ void __launch_bounds__(256)
k_karras(int n, const unsigned* __restrict__ keys, int* __restrict__ left,
         int* __restrict__ right, int* __restrict__ parent, int2* __restrict__ range) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = bvh_delta(keys, n, i, i + 1) - bvh_delta(keys, n, i, i - 1) > 0 ? 1 : -1;
  int dmin = bvh_delta(keys, n, i, i - d);
  int lmax = 2;
  while (bvh_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (bvh_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = bvh_delta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (bvh_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int lo_i = i < j ? i : j, hi_i = i < j ? j : i;
  range[i] = make_int2(lo_i, hi_i);
  int L = (lo_i == gamma) ? (n - 1 + gamma) : gamma;
  int R = (hi_i == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
  left[i] = L;
  right[i] = R;
  parent[L] = i;
  parent[R] = i;
  if (i == 0) parent[0] = -1;
}

// Node boxes without a bottom-up walk: leaf boxes in sorted order, then the
// union of every 64 leaves (block) and of every 64 blocks (superblock); an
// internal node's box is the union over its sorted leaf range [first, last],
// read as whole superblocks and blocks in the middle and single leaves or
// blocks at the ends. min/max are exact, so the boxes equal a bottom-up
// union except possibly for the sign of a zero, which no overlap test sees.
__global__ void __launch_bounds__(256)
k_refit_leaves(int n, const float4* __restrict__ lo, const float4* __restrict__ hi,
               const int* __restrict__ ids, float4* __restrict__ nlo, float4* __restrict__ nhi) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int body = ids[k];
  nlo[n - 1 + k] = lo[body];
  nhi[n - 1 + k] = hi[body];
}

__device__ __forceinline__ void box_union(float4& a_lo, float4& a_hi, float4 b_lo, float4 b_hi) {
  a_lo = make_float4(fminf(a_lo.x, b_lo.x), fminf(a_lo.y, b_lo.y), fminf(a_lo.z, b_lo.z), 0.0f);
  a_hi = make_float4(fmaxf(a_hi.x, b_hi.x), fmaxf(a_hi.y, b_hi.y), fmaxf(a_hi.z, b_hi.z), 0.0f);
}

// One 64-lane wavefront per block: union of 64 consecutive boxes (src lo at
// src[k], hi at src_hi[k]) into dst[2 * blk] (lo), dst[2 * blk + 1] (hi).
__global__ void __launch_bounds__(64)
k_refit_block(int count, const float4* __restrict__ src_lo, const float4* __restrict__ src_hi,
              int stride, float4* __restrict__ dst) {
  const int blk = blockIdx.x, k = blk * kRefitB + threadIdx.x;
  float4 l = make_float4(kInf, kInf, kInf, 0.0f), h = make_float4(-kInf, -kInf, -kInf, 0.0f);
  if (k < count) {
    l = src_lo[(size_t)k * stride];
    h = src_hi[(size_t)k * stride];
  }
  for (int o = 32; o > 0; o >>= 1) {
    float4 ol = make_float4(__shfl_down(l.x, o, 64), __shfl_down(l.y, o, 64), __shfl_down(l.z, o, 64), 0.0f);
    float4 oh = make_float4(__shfl_down(h.x, o, 64), __shfl_down(h.y, o, 64), __shfl_down(h.z, o, 64), 0.0f);
    box_union(l, h, ol, oh);
  }
  if (threadIdx.x == 0) {
    dst[2 * blk] = l;
    dst[2 * blk + 1] = h;
  }
}

__global__ void __launch_bounds__(256)
k_refit_internal(int n, const int2* __restrict__ range, const float4* __restrict__ b1,
                 const float4* __restrict__ b2, float4* __restrict__ nlo, float4* __restrict__ nhi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int2 r = range[i];
  int a = r.x;
  const int e = r.y + 1;  // leaves [a, e)
  const float4* llo = nlo + (n - 1);
  const float4* lhi = nhi + (n - 1);
  float4 l = make_float4(kInf, kInf, kInf, 0.0f), h = make_float4(-kInf, -kInf, -kInf, 0.0f);
  constexpr int S = kRefitB * kRefitB;
  // leaves up to the next block boundary
  while (a < e && (a % kRefitB) != 0) { box_union(l, h, llo[a], lhi[a]); ++a; }
  // blocks up to the next superblock boundary
  while (a + kRefitB <= e && (a % S) != 0) {
    const int bk = a / kRefitB;
    box_union(l, h, b1[2 * bk], b1[2 * bk + 1]);
    a += kRefitB;
  }
  // whole superblocks
  while (a + S <= e) {
    const int sk = a / S;
    box_union(l, h, b2[2 * sk], b2[2 * sk + 1]);
    a += S;
  }
  // remaining whole blocks, then leaves
  while (a + kRefitB <= e) {
    const int bk = a / kRefitB;
    box_union(l, h, b1[2 * bk], b1[2 * bk + 1]);
    a += kRefitB;
  }
  while (a < e) { box_union(l, h, llo[a], lhi[a]); ++a; }
  nlo[i] = l;
  nhi[i] = h;
}

__global__ void __launch_bounds__(256)
k_escape(int n, const int* __restrict__ left, const int* __restrict__ right,
         const int* __restrict__ parent, int* __restrict__ escape,
         const float4* __restrict__ nlo, const float4* __restrict__ nhi,
         const int* __restrict__ ids, float4* __restrict__ pk, int2* __restrict__ lf) {
  int node = blockIdx.x * blockDim.x + threadIdx.x;
  int nn = 2 * n - 1;
  if (node >= nn) return;
  int e = node, out = -1;
  for (int guard = 0; e != 0 && guard < nn; ++guard) {
    int p = parent[e];
    if (p < 0 || p >= n - 1) break;
    if (left[p] == e) { out = right[p]; break; }
    e = p;
  }
  escape[node] = out;
  if (node < n - 1) {
    const float4 a = nlo[node], b = nhi[node];
    pk[2 * node] = make_float4(a.x, a.y, a.z, __int_as_float(left[node]));
    pk[2 * node + 1] = make_float4(b.x, b.y, b.z, __int_as_float(out));
  } else {
    lf[node - (n - 1)] = make_int2(ids[node - (n - 1)], out);
  }
}

int bvh_build(hipStream_t st, int n, const float4* lo, const float4* hi, void* ws, size_t) {
  BvhWs w = bvh_carve(ws, n);
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, w.bounds);
  hipLaunchKernelGGL(k_bounds, dim3(std::min((n + 255) / 256, kBoundsBlocks)), dim3(256), 0, st, n,
                     lo, hi, w.bounds);
  hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, st, n, lo, hi, w.bounds, w.keys,
                     w.ids);
  radix_sort_pairs(st, w.keys, w.ids, w.keys_tmp, w.ids_tmp, n, w.radix_ws);
  if (n > 1)
    hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, st, n, w.keys, w.left,
                       w.right, w.parent, w.range);
  hipLaunchKernelGGL(k_refit_leaves, dim3((n + 255) / 256), dim3(256), 0, st, n, lo, hi, w.ids,
                     w.nlo, w.nhi);
  if (n > 1) {
    const int nb1 = (n + kRefitB - 1) / kRefitB, nb2 = (nb1 + kRefitB - 1) / kRefitB;
    hipLaunchKernelGGL(k_refit_block, dim3(nb1), dim3(64), 0, st, n, w.nlo + (n - 1),
                       w.nhi + (n - 1), 1, w.b1);
    hipLaunchKernelGGL(k_refit_block, dim3(nb2), dim3(64), 0, st, nb1, w.b1, w.b1 + 1, 2, w.b2);
    hipLaunchKernelGGL(k_refit_internal, dim3((n - 1 + 255) / 256), dim3(256), 0, st, n, w.range,
                       w.b1, w.b2, w.nlo, w.nhi);
  }
  if (n > 1)
    hipLaunchKernelGGL(k_escape, dim3((2 * n - 1 + 255) / 256), dim3(256), 0, st, n, w.left,
                       w.right, w.parent, w.escape, w.nlo, w.nhi, w.ids, w.pk, w.lf);
  return 0;
}

__device__ __forceinline__ bool box_overlap(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  return alo.x <= bhi.x && blo.x <= ahi.x && alo.y <= bhi.y && blo.y <= ahi.y &&
         alo.z <= bhi.z && blo.z <= ahi.z;
}

// Boxes of the broad phase: lo/hi are the bodies' AABBs; the tree was built
// over tlo/thi (the same arrays, or per-scene shifted copies), and scene (or
// null) keeps pairs inside one scene.
struct BroadBoxes {
  const float4* lo;
  const float4* hi;
  const float4* tlo;
  const float4* thi;
  const int* scene;
};

__device__ __forceinline__ bool leaf_pair(const BroadBoxes& B, int i, int j, float4 qlo,
                                          float4 qhi, bool st_i, const float4* bstat) {
  return box_overlap(qlo, qhi, B.lo[j], B.hi[j]) && !(st_i && bstat[j].x <= 0.0f) &&
         (B.scene == nullptr || B.scene[j] == B.scene[i]);
}

// Traverse for body i; visit(j) for every body j > i with overlapping AABB that
// is not static-static. Returns the number visited.
template <bool kWrite>
__device__ int bvh_query(int n, int i, const BroadBoxes& B, const float4* __restrict__ bstat,
                         const BvhWs& w, int* out, int cap_left) {
  const float4 qlo = B.lo[i], qhi = B.hi[i], tqlo = B.tlo[i], tqhi = B.thi[i];
  bool st_i = bstat[i].x <= 0.0f;
  int cnt = 0;
  if (n == 1) return 0;
  int node = 0;
  int steps = 0, max_steps = 4 * n + 8;   // a DFS visits 2n-1 nodes once
  while (node != -1 && ++steps <= max_steps) {
    if (node >= n - 1) {
      const int2 L = w.lf[node - (n - 1)];
      const int j = L.x;
      if (j > i && leaf_pair(B, i, j, qlo, qhi, st_i, bstat)) {
        if (kWrite && cnt < cap_left) out[cnt] = j;
        ++cnt;
      }
      node = L.y;
    } else {
      const float4 a = w.pk[2 * node], b = w.pk[2 * node + 1];
      node = box_overlap(tqlo, tqhi, a, b) ? __float_as_int(a.w) : __float_as_int(b.w);
    }
  }
  return cnt;
}

// One traversal per body: counts every partner and keeps the first
// kPartnerSlots of them, sorted, in `slots` (nb x kPartnerSlots), so the
// write pass only copies them; a body with more partners is traversed again
// there.
__global__ void __launch_bounds__(256)
k_bvh_count(int n, BroadBoxes B, const float4* __restrict__ bstat, BvhWs w,
            int* __restrict__ counts, int* __restrict__ slots) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int* seg = slots + (size_t)i * kPartnerSlots;
  const int cnt = bvh_query<true>(n, i, B, bstat, w, seg, kPartnerSlots);
  if (cnt <= kPartnerSlots) {
    // canonical order: partners ascending (insertion sort)
    for (int a = 1; a < cnt; ++a) {
      int v = seg[a];
      int b = a - 1;
      while (b >= 0 && seg[b] > v) { seg[b + 1] = seg[b]; --b; }
      seg[b + 1] = v;
    }
  }
  counts[i] = cnt;
}

// The same traversal with G lanes of one wavefront per body. The group's lanes
// share a stack in LDS: every round each lane pops one node (the group takes
// up to G from the top), tests it, and an overlapping internal node pushes
// both children; the group's pushes are placed by a prefix count over its
// ballot bits. Every leaf whose ancestors all overlap the query is reached
// exactly once, as in the stackless walk, so the partner set is the same; a
// body's partners land in an LDS list (order of discovery) and are written to
// its slots by rank, ascending. When a push would not fit in the stack, the
// lane walks both children's subtrees itself with the escape links: from the
// left child until the parent's escape, the next node after the parent's
// subtree. Groups never leave their wavefront, so the shared arrays need no
// workgroup barrier: LDS instructions of one wavefront complete in order.
template <int G>
__global__ void __launch_bounds__(256)
k_bvh_count_grp(int n, BroadBoxes B, const float4* __restrict__ bstat, BvhWs w,
                int* __restrict__ counts, int* __restrict__ slots) {
  static_assert(G >= 2 && G <= 64 && (64 % G) == 0, "group within a wavefront");
  constexpr int kGroups = 256 / G;
  constexpr int kStack = G >= 64 ? 256 : (G >= 16 ? 192 : 128);
  __shared__ int stk[kGroups][kStack];
  __shared__ int found[kGroups][kPartnerSlots];
  __shared__ int nfound[kGroups];
  const int grp = threadIdx.x / G, r = threadIdx.x % G;
  const int i = blockIdx.x * kGroups + grp;
  if (i >= n) return;  // whole groups: G divides the wavefront
  int* seg = slots + (size_t)i * kPartnerSlots;
  if (n == 1) {
    if (r == 0) counts[i] = 0;
    return;
  }
  const int lane = threadIdx.x & 63, base = lane - r;
  const unsigned long long gmask =
      G == 64 ? ~0ull : (((1ull << G) - 1) << base);
  const unsigned long long below = (1ull << lane) - 1;
  const float4 qlo = B.lo[i], qhi = B.hi[i], tqlo = B.tlo[i], tqhi = B.thi[i];
  const bool st_i = bstat[i].x <= 0.0f;
  if (r == 0) {
    stk[grp][0] = 0;
    nfound[grp] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  int top = 1;  // the same in every lane of the group
  auto hit = [&](int j) {
    const int at = atomicAdd(&nfound[grp], 1);
    if (at < kPartnerSlots) found[grp][at] = j;
  };
  while (top > 0) {
    const int take = top < G ? top : G;
    const int node = r < take ? stk[grp][top - 1 - r] : -1;
    top -= take;
    int kid = -1, rkid = -1, esc = -1;
    if (node >= n - 1) {
      const int j = w.lf[node - (n - 1)].x;
      if (j > i && leaf_pair(B, i, j, qlo, qhi, st_i, bstat)) hit(j);
    } else if (node >= 0) {
      const float4 a = w.pk[2 * node], b = w.pk[2 * node + 1];
      const int rk = w.right[node];
      if (box_overlap(tqlo, tqhi, a, b)) {
        kid = __float_as_int(a.w);
        rkid = rk;
        esc = __float_as_int(b.w);
      }
    }
    const unsigned long long m = __ballot(kid >= 0) & gmask;
    const int before = __popcll(m & below), pushes = __popcll(m);
    const int room = (kStack - top) / 2;
    __builtin_amdgcn_wave_barrier();
    if (kid >= 0) {
      if (before < room) {
        stk[grp][top + 2 * before] = rkid;
        stk[grp][top + 2 * before + 1] = kid;
      } else {
        int v = kid;
        int steps = 0;
        const int max_steps = 4 * n + 8;
        while (v != esc && v != -1 && ++steps <= max_steps) {
          if (v >= n - 1) {
            const int2 L = w.lf[v - (n - 1)];
            if (L.x > i && leaf_pair(B, i, L.x, qlo, qhi, st_i, bstat)) hit(L.x);
            v = L.y;
          } else {
            const float4 a = w.pk[2 * v], b = w.pk[2 * v + 1];
            v = box_overlap(tqlo, tqhi, a, b) ? __float_as_int(a.w) : __float_as_int(b.w);
          }
        }
      }
    }
    top += 2 * (pushes < room ? pushes : room);
    __builtin_amdgcn_wave_barrier();
  }
  const int cnt = nfound[grp];
  if (r == 0) counts[i] = cnt;
  if (cnt <= kPartnerSlots) {
    // partners are distinct: a partner's rank is its position in ascending order
    for (int a = r; a < cnt; a += G) {
      const int v = found[grp][a];
      int rank = 0;
      for (int b = 0; b < cnt; ++b) rank += found[grp][b] < v ? 1 : 0;
      seg[rank] = v;
    }
  }
}

// Smallest overlapping partner of body i that is greater than `after`.
__device__ int bvh_next_partner(int n, int i, int after, const BroadBoxes& B,
                                const float4* __restrict__ bstat, const BvhWs& w) {
  const float4 qlo = B.lo[i], qhi = B.hi[i], tqlo = B.tlo[i], tqhi = B.thi[i];
  bool st_i = bstat[i].x <= 0.0f;
  int best = 0x7fffffff;
  int node = 0;
  int steps = 0, max_steps = 4 * n + 8;
  while (node != -1 && ++steps <= max_steps) {
    if (node >= n - 1) {
      const int2 L = w.lf[node - (n - 1)];
      const int j = L.x;
      if (j > i && j > after && j < best && leaf_pair(B, i, j, qlo, qhi, st_i, bstat)) best = j;
      node = L.y;
    } else {
      const float4 a = w.pk[2 * node], b = w.pk[2 * node + 1];
      node = box_overlap(tqlo, tqhi, a, b) ? __float_as_int(a.w) : __float_as_int(b.w);
    }
  }
  return best;
}

__global__ void __launch_bounds__(256)
k_bvh_write(int n, BroadBoxes B, const float4* __restrict__ bstat, BvhWs w,
            const int* __restrict__ counts, const int* __restrict__ slots,
            const int* __restrict__ offsets, int cap, int* __restrict__ pa,
            int* __restrict__ pb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int off = offsets[i];
  int room = cap - off;
  if (room <= 0 || n == 1) return;
  int cnt = counts[i];
  int* seg = pb + off;
  if (cnt <= room && cnt <= kPartnerSlots) {
    const int* src = slots + (size_t)i * kPartnerSlots;
    for (int a = 0; a < cnt; ++a) {
      seg[a] = src[a];
      pa[off + a] = i;
    }
  } else if (cnt <= room) {
    bvh_query<true>(n, i, B, bstat, w, seg, cnt);
    // Canonical order: this body's partners ascending (insertion sort).
    for (int a = 1; a < cnt; ++a) {
      int v = seg[a];
      int b = a - 1;
      while (b >= 0 && seg[b] > v) { seg[b + 1] = seg[b]; --b; }
      seg[b + 1] = v;
    }
    for (int a = 0; a < cnt; ++a) pa[off + a] = i;
  } else {
    // Pair pool ends inside this body's segment: keep its `room` smallest
    // partners, the prefix of the canonical list.
    int last = -1;
    for (int a = 0; a < room; ++a) {
      last = bvh_next_partner(n, i, last, B, bstat, w);
      seg[a] = last;
      pa[off + a] = i;
    }
  }
}

void bvh_count_pairs(hipStream_t st, int n, const float4* lo, const float4* hi,
                     const float4* tlo, const float4* thi, const int* scene, const float4* bstat,
                     const void* ws, int* counts, int* slots, int lanes) {
  BvhWs w = bvh_carve(const_cast<void*>(ws), n);
  const BroadBoxes B{lo, hi, tlo, thi, scene};
  if (lanes == 0) lanes = bvh_count_lanes(n);
  switch (lanes) {
    case 8:
      hipLaunchKernelGGL(k_bvh_count_grp<8>, dim3((n + 31) / 32), dim3(256), 0, st, n, B, bstat, w,
                         counts, slots);
      break;
    case 16:
      hipLaunchKernelGGL(k_bvh_count_grp<16>, dim3((n + 15) / 16), dim3(256), 0, st, n, B, bstat, w,
                         counts, slots);
      break;
    case 64:
      hipLaunchKernelGGL(k_bvh_count_grp<64>, dim3((n + 3) / 4), dim3(256), 0, st, n, B, bstat, w,
                         counts, slots);
      break;
    default:
      hipLaunchKernelGGL(k_bvh_count, dim3((n + 255) / 256), dim3(256), 0, st, n, B, bstat, w,
                         counts, slots);
  }
}

// Lanes per body by body count: few bodies fill few wavefronts, so their walks
// are split; with many, the chip is full and lanes that find nothing cost
// issue slots (measured: 10k-box pile best at 64, 98k pyramid at 8, 429k
// batched bodies at 1).
int bvh_count_lanes(int n) {
  return n <= 16384 ? 64 : (n <= 65536 ? 16 : (n <= 262144 ? 8 : 1));
}

void bvh_write_pairs(hipStream_t st, int n, const float4* lo, const float4* hi,
                     const float4* tlo, const float4* thi, const int* scene, const float4* bstat,
                     const void* ws, const int* counts, const int* slots, const int* offsets,
                     int cap, int* pa, int* pb) {
  BvhWs w = bvh_carve(const_cast<void*>(ws), n);
  const BroadBoxes B{lo, hi, tlo, thi, scene};
  hipLaunchKernelGGL(k_bvh_write, dim3((n + 255) / 256), dim3(256), 0, st, n, B, bstat, w, counts,
                     slots, offsets, cap, pa, pb);
}

// ---------------------------------------------------------------------------
// Narrow phase: SAT + manifold per pair
// ---------------------------------------------------------------------------
__device__ __forceinline__ float proj_radius(v3 L, v3 c0, v3 c1, v3 c2, v3 e) {
  return (e.x * fabs_(dot(L, c0)) + e.y * fabs_(dot(L, c1))) + e.z * fabs_(dot(L, c2));
}

__global__ void __launch_bounds__(128)
k_narrowphase(int n_pairs, const float4* __restrict__ pose, const float4* __restrict__ bstat,
              const int* __restrict__ pair_a, const int* __restrict__ pair_b, float margin,
              int* __restrict__ overlap, int* __restrict__ sat_out, float4* __restrict__ n_out,
              float* __restrict__ depth_out, int* __restrict__ mf_count,
              int* __restrict__ mf_refa, float4* __restrict__ mf_off, float4* __restrict__ mf_n,
              const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(n_pairs, nd)) return;
  int i = pair_a[p], j = pair_b[p];
  v3 cA = f3(pose[2 * i]), cB = f3(pose[2 * j]);
  float4 bsA = bstat[i], bsB = bstat[j];
  v3 eA = mk3(bsA.y, bsA.z, bsA.w), eB = mk3(bsB.y, bsB.z, bsB.w);
  m33 RA = qtomat(f4q(pose[2 * i + 1])), RB = qtomat(f4q(pose[2 * j + 1]));
  v3 t = sub(cB, cA);
  v3 A[3] = {colm(RA, 0), colm(RA, 1), colm(RA, 2)};
  v3 Bc[3] = {colm(RB, 0), colm(RB, 1), colm(RB, 2)};
  float best = 1.0e20f;
  int best_idx = -1;
  v3 best_axis = mk3(0.f, 0.f, 0.f);
  mf_count[p] = 0;
  bool sep = false;
  for (int k = 0; k < 3 && !sep; ++k) {
    v3 L = colm(RA, k);
    float ov = (comp(eA, k) + proj_radius(L, Bc[0], Bc[1], Bc[2], eB)) - fabs_(dot(t, L));
    float tdl = dot(t, L);
    if (ov < -margin) { sep = true; break; }
    if (ov < best) { best = ov; best_idx = k; best_axis = tdl > 0.0f ? neg(L) : L; }
  }
  for (int k = 0; k < 3 && !sep; ++k) {
    v3 L = colm(RB, k);
    float ov = (proj_radius(L, A[0], A[1], A[2], eA) + comp(eB, k)) - fabs_(dot(t, L));
    float tdl = dot(t, L);
    if (ov < -margin) { sep = true; break; }
    if (ov < best) { best = ov; best_idx = 3 + k; best_axis = tdl > 0.0f ? neg(L) : L; }
  }
  for (int ie = 0; ie < 3 && !sep; ++ie) {
    for (int je = 0; je < 3; ++je) {
      v3 L = cross(colm(RA, ie), colm(RB, je));
      float nl = length(L);
      if (nl < 1.0e-6f) continue;
      L = divs(L, nl);
      float ov = (proj_radius(L, A[0], A[1], A[2], eA) + proj_radius(L, Bc[0], Bc[1], Bc[2], eB)) -
                 fabs_(dot(t, L));
      float tdl = dot(t, L);
      if (ov < 0.0f) { sep = true; break; }
      if (ov < best) { best = ov; best_idx = 6 + 3 * ie + je; best_axis = tdl > 0.0f ? neg(L) : L; }
    }
  }
  if (sep) {
    overlap[p] = 0;
    return;
  }
  overlap[p] = 1;
  sat_out[p] = best_idx;
  n_out[p] = p4(best_axis);
  depth_out[p] = best;

  const int sat = best_idx;
  const v3 n_in = best_axis;
  float4* off = mf_off + 8 * (size_t)p;
  if (sat >= 6) {
    int e = sat - 6, kA = e / 3, kB = e % 3;
    v3 dA = colm(RA, kA), dB = colm(RB, kB);
    if (length(cross(dA, dB)) < 1.0e-4f) return;
    int kA1 = (kA + 1) % 3, kA2 = (kA + 2) % 3;
    v3 A1 = colm(RA, kA1), A2 = colm(RA, kA2);
    v3 toB = sub(cB, cA);
    float sbA = dot(toB, A1) >= 0.0f ? 1.0f : -1.0f;
    float scA = dot(toB, A2) >= 0.0f ? 1.0f : -1.0f;
    v3 midA = add(add(cA, mul(A1, sbA * comp(eA, kA1))), mul(A2, scA * comp(eA, kA2)));
    v3 P1 = sub(midA, mul(dA, comp(eA, kA)));
    v3 Q1 = add(midA, mul(dA, comp(eA, kA)));
    int kB1 = (kB + 1) % 3, kB2 = (kB + 2) % 3;
    v3 B1 = colm(RB, kB1), B2 = colm(RB, kB2);
    v3 toA = sub(cA, cB);
    float sbB = dot(toA, B1) >= 0.0f ? 1.0f : -1.0f;
    float scB = dot(toA, B2) >= 0.0f ? 1.0f : -1.0f;
    v3 midB = add(add(cB, mul(B1, sbB * comp(eB, kB1))), mul(B2, scB * comp(eB, kB2)));
    v3 P2 = sub(midB, mul(dB, comp(eB, kB)));
    v3 Q2 = add(midB, mul(dB, comp(eB, kB)));
    v3 d1 = sub(Q1, P1), d2 = sub(Q2, P2), r = sub(P1, P2);
    float a = dot(d1, d1), ev = dot(d2, d2), f = dot(d2, r);
    const float eps = 1.0e-12f;
    float sp = 0.0f, tp = 0.0f;
    if (a <= eps && ev <= eps) {
      sp = 0.0f; tp = 0.0f;
    } else if (a <= eps) {
      sp = 0.0f;
      tp = clampf_(f / fmax_(ev, eps), 0.0f, 1.0f);
    } else if (ev <= eps) {
      tp = 0.0f;
      float c = dot(d1, r);
      sp = clampf_(-c / a, 0.0f, 1.0f);
    } else {
      float c = dot(d1, r);
      float b = dot(d1, d2);
      float denom = a * ev - b * b;
      if (denom != 0.0f) sp = clampf_((b * f - c * ev) / denom, 0.0f, 1.0f);
      else sp = 0.0f;
      tp = (b * sp + f) / ev;
      if (tp < 0.0f) {
        tp = 0.0f;
        sp = clampf_(-c / a, 0.0f, 1.0f);
      } else if (tp > 1.0f) {
        tp = 1.0f;
        sp = clampf_((b - c) / a, 0.0f, 1.0f);
      }
    }
    v3 pA = add(P1, mul(d1, sp));
    v3 pB = add(P2, mul(d2, tp));
    float gap = dot(sub(pA, pB), n_in);
    if (gap > margin) return;
    off[0] = p4(mvmul(mT(RA), sub(pA, cA)));
    off[4] = p4(mvmul(mT(RB), sub(pB, cB)));
    mf_refa[p] = 1;
    mf_n[p] = p4(n_in);
    mf_count[p] = 1;
    return;
  }
  int ref_is_a = sat < 3 ? 1 : 0;
  int ref_axis;
  v3 n_used, ref_c, inc_c, e_ref, e_inc;
  m33 R_ref, R_inc;
  if (sat < 3) {
    ref_axis = sat; n_used = n_in; ref_c = cA; R_ref = RA; e_ref = eA; inc_c = cB; R_inc = RB; e_inc = eB;
  } else {
    ref_axis = sat - 3; n_used = neg(n_in); ref_c = cB; R_ref = RB; e_ref = eB; inc_c = cA; R_inc = RA; e_inc = eA;
  }
  v3 rcol = colm(R_ref, ref_axis);
  float rsign = dot(rcol, n_used) > 0.0f ? -1.0f : 1.0f;
  v3 rfn = mul(rcol, rsign);
  v3 rfc = add(ref_c, mul(rfn, comp(e_ref, ref_axis)));
  int ax1 = (ref_axis + 1) % 3, ax2 = (ref_axis + 2) % 3;
  v3 u_ref = colm(R_ref, ax1), v_ref = colm(R_ref, ax2);
  float eu = comp(e_ref, ax1), evv = comp(e_ref, ax2);
  int inc_axis = 0;
  float inc_sign = 1.0f, best_dot = -1.0e20f;
  for (int k = 0; k < 3; ++k) {
    float dpos = dot(colm(R_inc, k), n_used);
    if (dpos > best_dot) { best_dot = dpos; inc_axis = k; inc_sign = 1.0f; }
    float dneg = -dpos;
    if (dneg > best_dot) { best_dot = dneg; inc_axis = k; inc_sign = -1.0f; }
  }
  v3 ifn = mul(colm(R_inc, inc_axis), inc_sign);
  v3 ifc = add(inc_c, mul(ifn, comp(e_inc, inc_axis)));
  int iax1 = (inc_axis + 1) % 3, iax2 = (inc_axis + 2) % 3;
  v3 iu = colm(R_inc, iax1), iv = colm(R_inc, iax2);
  float ieu = comp(e_inc, iax1), iev = comp(e_inc, iax2);
  // the two clip polygons (up to 8 vertices each) live in LDS, one column
  // per lane: indexed by a running count, they would go to scratch memory
  __shared__ float clip[2][8][3][128];
  const int tl = threadIdx.x;
  auto bget = [&](int b, int k) {
    return mk3(clip[b][k][0][tl], clip[b][k][1][tl], clip[b][k][2][tl]);
  };
  auto bset = [&](int b, int k, v3 v) {
    clip[b][k][0][tl] = v.x;
    clip[b][k][1][tl] = v.y;
    clip[b][k][2][tl] = v.z;
  };
  bset(0, 0, add(add(ifc, mul(iu, ieu)), mul(iv, iev)));
  bset(0, 1, add(sub(ifc, mul(iu, ieu)), mul(iv, iev)));
  bset(0, 2, sub(sub(ifc, mul(iu, ieu)), mul(iv, iev)));
  bset(0, 3, sub(add(ifc, mul(iu, ieu)), mul(iv, iev)));
  int cur = 0, len = 4;
  for (int pl = 0; pl < 4; ++pl) {
    v3 ppt, pn;
    if (pl == 0) { ppt = add(rfc, mul(u_ref, eu)); pn = u_ref; }
    else if (pl == 1) { ppt = sub(rfc, mul(u_ref, eu)); pn = neg(u_ref); }
    else if (pl == 2) { ppt = add(rfc, mul(v_ref, evv)); pn = v_ref; }
    else { ppt = sub(rfc, mul(v_ref, evv)); pn = neg(v_ref); }
    if (len == 0) return;
    int out = 0;
    for (int iv2 = 0; iv2 < len; ++iv2) {
      v3 ap = bget(cur, iv2);
      v3 bp = bget(cur, (iv2 + 1) % len);
      float da = dot(sub(ap, ppt), pn);
      float db = dot(sub(bp, ppt), pn);
      if (da <= 0.0f) {
        if (out < 8) bset(1 - cur, out++, ap);
        if (db > 0.0f) {
          float tt = da / (da - db);
          if (out < 8) bset(1 - cur, out++, add(ap, mul(sub(bp, ap), tt)));
        }
      } else if (db <= 0.0f) {
        float tt = da / (da - db);
        if (out < 8) bset(1 - cur, out++, add(ap, mul(sub(bp, ap), tt)));
      }
    }
    cur = 1 - cur;
    len = out;
  }
  if (len == 0) return;
  // the four deepest clip points, in LDS for the same reason as the polygons
  __shared__ float bd_s[4][128];
  __shared__ int bi_s[4][128];
  for (int k = 0; k < 4; ++k) {
    bd_s[k][tl] = -1.0e30f;
    bi_s[k][tl] = -1;
  }
  int nbest = 0;
  for (int iv2 = 0; iv2 < len; ++iv2) {
    float ds = dot(sub(bget(cur, iv2), rfc), rfn);
    if (ds < margin) {
      float depth = -ds;
      if (nbest < 4) {
        bd_s[nbest][tl] = depth; bi_s[nbest][tl] = iv2; ++nbest;
      } else {
        int mp = 0;
        float mv = bd_s[0][tl];
        for (int k = 1; k < 4; ++k) if (bd_s[k][tl] < mv) { mv = bd_s[k][tl]; mp = k; }
        if (depth > mv) { bd_s[mp][tl] = depth; bi_s[mp][tl] = iv2; }
      }
    }
  }
  if (nbest == 0) return;
  for (int is = 0; is < nbest - 1; ++is) {
    int mp = is;
    for (int k = is + 1; k < nbest; ++k) if (bd_s[k][tl] > bd_s[mp][tl]) mp = k;
    if (mp != is) {
      float td = bd_s[is][tl]; bd_s[is][tl] = bd_s[mp][tl]; bd_s[mp][tl] = td;
      int ti = bi_s[is][tl]; bi_s[is][tl] = bi_s[mp][tl]; bi_s[mp][tl] = ti;
    }
  }
  m33 Rtr = mT(R_ref), Rti = mT(R_inc);
  for (int c = 0; c < nbest; ++c) {
    v3 pinc = bget(cur, bi_s[c][tl]);
    float ds = dot(sub(pinc, rfc), rfn);
    v3 pref = sub(pinc, mul(rfn, ds));
    off[c] = p4(mvmul(Rtr, sub(pref, ref_c)));
    off[4 + c] = p4(mvmul(Rti, sub(pinc, inc_c)));
  }
  mf_refa[p] = ref_is_a;
  mf_n[p] = p4(n_used);
  mf_count[p] = nbest;
}

void launch_narrowphase(hipStream_t st, int n_pairs, const float4* pose, const float4* bstat,
                        const int* pair_a, const int* pair_b, float margin, int* overlap,
                        int* sat_idx, float4* n_hat, float* depth, int* mf_count, int* mf_refa,
                        float4* mf_off, float4* mf_n, const int* nd) {
  if (n_pairs <= 0) return;
  hipLaunchKernelGGL(k_narrowphase, dim3((n_pairs + 127) / 128), dim3(128), 0, st, n_pairs, pose,
                     bstat, pair_a, pair_b, margin, overlap, sat_idx, n_hat, depth, mf_count,
                     mf_refa, mf_off, mf_n, nd);
}

// ---------------------------------------------------------------------------
// Row emission
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pair_mu(float mr, float mi, float default_mu) {
  float mu = default_mu;
  if (mr > 0.0f || mi > 0.0f) mu = sqrt_(mr * mi);
  return mu;
}

__global__ void __launch_bounds__(256)
k_emit_demand(int n_pairs, const int* __restrict__ pa, const int* __restrict__ pb,
              const int* __restrict__ mf_count, const int* __restrict__ mf_refa,
              const float* __restrict__ bfric, float default_mu, int* __restrict__ demand,
              int* __restrict__ rpc_out, const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(n_pairs, nd)) return;
  int nc = mf_count[p];
  int rpc = 3;
  if (nc > 0) {
    int ref = mf_refa[p] == 1 ? pa[p] : pb[p];
    int inc = mf_refa[p] == 1 ? pb[p] : pa[p];
    float mu = pair_mu(bfric[ref], bfric[inc], default_mu);
    rpc = mu <= 0.0f ? 1 : 3;
  }
  rpc_out[p] = rpc;
  demand[p] = nc * rpc;
}

void launch_emit_demand(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                        const int* mf_count, const int* mf_refa, const float* bfric,
                        float default_mu, int* demand, int* rpc_out, const int* nd) {
  if (n_pairs <= 0) return;
  hipLaunchKernelGGL(k_emit_demand, dim3((n_pairs + 255) / 256), dim3(256), 0, st, n_pairs, pa, pb,
                     mf_count, mf_refa, bfric, default_mu, demand, rpc_out, nd);
}

// Serial reservation order restated: a contact (p, c) fits iff every earlier
// contact fit and its rows end within capacity; after the first failure every
// later pair with contacts still reserves the rows of one failed attempt.
__global__ void __launch_bounds__(256)
k_emit_succ(int n_pairs, const int* __restrict__ mf_count, const int* __restrict__ rpc,
            const int* __restrict__ dscan, int avail, int* __restrict__ succ,
            int* __restrict__ reserve, const int* __restrict__ nd) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(n_pairs, nd)) return;
  int nc = mf_count[p], r = rpc[p];
  int s = 0;
  if (nc > 0) {
    long long room = (long long)avail - (long long)dscan[p];
    if (room > 0) {
      long long k = room / r;
      s = k < nc ? (int)k : nc;
    }
  }
  succ[p] = s;
  reserve[p] = s * r + (s < nc ? r : 0);
}

void launch_emit_succ(hipStream_t st, int n_pairs, const int* mf_count, const int* rpc,
                      const int* dscan, int avail, int* succ, int* reserve, const int* nd) {
  if (n_pairs <= 0) return;
  hipLaunchKernelGGL(k_emit_succ, dim3((n_pairs + 255) / 256), dim3(256), 0, st, n_pairs, mf_count,
                     rpc, dscan, avail, succ, reserve, nd);
}

__global__ void __launch_bounds__(256)
k_emit_write(int n_pairs, const int* __restrict__ pa, const int* __restrict__ pb,
             const int* __restrict__ mf_refa, const float4* __restrict__ mf_n,
             const float4* __restrict__ mf_off, const float* __restrict__ bfric, float default_mu,
             float mult, const int* __restrict__ rpc, const int* __restrict__ succ,
             const int* __restrict__ dscan, const int* __restrict__ sscan, int n_static_rows,
             int g_static, GroupGeo* __restrict__ geo, GroupSt* __restrict__ gst, int hash_cap,
             const int* __restrict__ hkeys, const float4* __restrict__ hstate,
             const int* __restrict__ hage, const int* __restrict__ nd,
             const int* __restrict__ hash_cap_dev) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dev_count(n_pairs, nd)) return;
  if (hash_cap_dev) hash_cap = *hash_cap_dev;  // the table as last built
  int ns = succ[p];
  if (ns == 0) return;
  int ref = mf_refa[p] == 1 ? pa[p] : pb[p];
  int inc = mf_refa[p] == 1 ? pb[p] : pa[p];
  float mu = pair_mu(bfric[ref], bfric[inc], default_mu);
  float mus = mu * mult;
  int r = rpc[p];
  int ka = ref < inc ? ref : inc, kb = ref < inc ? inc : ref;
  float4 n = mf_n[p];
  const float4* off = mf_off + 8 * (size_t)p;
  int gbase = g_static + sscan[p];
  int rbase = n_static_rows + dscan[p];
  for (int c = 0; c < ns; ++c) {
    GroupGeo G;
    G.hdr = make_int4(g_make(kGBox, r, c), ref, inc, rbase + c * r);
    G.g0 = make_float4(n.x, n.y, n.z, kInf);
    float4 oa = off[c], ob = off[4 + c];
    G.g1 = make_float4(oa.x, oa.y, oa.z, mu);
    G.g2 = make_float4(ob.x, ob.y, ob.z, mus);
    float ln = 0.f, lt = 0.f, lb = 0.f, kn = kPenaltyMin, kt = kPenaltyMinTangent,
          kbb = kPenaltyMinTangent;
    int wt = 0, wb = 0, age = 0;
    if (hash_cap > 0) {
      int key = encode_pair_key(ka, kb, c);
      int h = hash_pair_slot(key, hash_cap);
      for (int probe = 0; probe < hash_cap; ++probe) {
        int slot = (h + probe) % hash_cap;
        int sk = hkeys[slot];
        if (sk == key) {
          float4 s0 = hstate[2 * (size_t)slot], s1 = hstate[2 * (size_t)slot + 1];
          ln = s0.x; lt = s0.y; lb = s0.z; kn = s0.w; kt = s1.x; kbb = s1.y;
          wt = (int)s1.z; wb = (int)s1.w;
          if (hage) age = hage[slot] + 1;
          break;
        }
        if (sk == kHashEmpty) break;
      }
    }
    int bits = kActN;
    if (r == 3) {
      bits |= kActT | kActB;
      if (wt) bits |= kWasT;
      if (wb) bits |= kWasB;
    }
    GroupSt S;
    S.s0 = make_float4(ln, r == 3 ? lt : 0.0f, r == 3 ? lb : 0.0f, 0.0f);
    S.s1 = make_float4(kn, r == 3 ? kt : 0.0f, r == 3 ? kbb : 0.0f, 0.0f);
    // s2.z: substeps this contact has been cached (learned warm start only)
    S.s2 = make_float4(0.0f, __int_as_float(bits), __int_as_float(age), 0.0f);
    geo[gbase + c] = G;
    gst[gbase + c] = S;
  }
}

void launch_emit_write(hipStream_t st, int n_pairs, const int* pa, const int* pb,
                       const int* mf_refa, const float4* mf_n, const float4* mf_off,
                       const float* bfric, float default_mu, float mult, const int* rpc,
                       const int* succ, const int* dscan, const int* sscan, int n_static_rows,
                       int g_static, GroupGeo* geo, GroupSt* gst, int hash_cap, const int* hkeys,
                       const float4* hstate, const int* hage, const int* nd,
                       const int* hash_cap_dev) {
  if (n_pairs <= 0) return;
  hipLaunchKernelGGL(k_emit_write, dim3((n_pairs + 255) / 256), dim3(256), 0, st, n_pairs, pa, pb,
                     mf_refa, mf_n, mf_off, bfric, default_mu, mult, rpc, succ, dscan, sscan,
                     n_static_rows, g_static, geo, gst, hash_cap, hkeys, hstate, hage, nd,
                     hash_cap_dev);
}

// ---------------------------------------------------------------------------
// Warm-start cache refresh
// ---------------------------------------------------------------------------
// Fills and copies are grid-stride loops over at most kFillBlocks workgroups
// (a table of many millions of slots needs no more to reach full bandwidth).
constexpr int kFillBlocks = 4096;

__global__ void __launch_bounds__(256) k_fill_int(int* p, int n, int v, const int* nd) {
  n = dev_count(n, nd);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

void launch_fill_int(hipStream_t st, int* p, int n, int v, const int* nd) {
  if (n <= 0) return;
  const int blocks = std::min((n + 255) / 256, kFillBlocks);
  hipLaunchKernelGGL(k_fill_int, dim3(blocks), dim3(256), 0, st, p, n, v, nd);
}

__global__ void __launch_bounds__(256)
k_copy_int(int* __restrict__ dst, const int* __restrict__ src, int n, const int* nd) {
  n = dev_count(n, nd);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy_int(hipStream_t st, int* dst, const int* src, int n, const int* nd) {
  if (n <= 0) return;
  const int blocks = std::min((n + 255) / 256, kFillBlocks);
  hipLaunchKernelGGL(k_copy_int, dim3(blocks), dim3(256), 0, st, dst, src, n, nd);
}

// Slots of the warm-start table for `entries` live contacts: about four
// times the entries, at least 4096, at most the reference's capacity.
__global__ void k_hash_size(const int* entries, int hash_cap, int* live) {
  if (threadIdx.x != 0) return;
  const long long want = 4LL * (long long)*entries + 1;
  const long long w = want > 4096 ? want : 4096;
  *live = (int)(w < (long long)hash_cap ? w : (long long)hash_cap);
}

void launch_hash_size(hipStream_t st, const int* entries, int hash_cap, int* live) {
  hipLaunchKernelGGL(k_hash_size, dim3(1), dim3(64), 0, st, entries, hash_cap, live);
}

__global__ void __launch_bounds__(256)
k_hash_insert(int n, int g_first, const GroupGeo* __restrict__ geo,
              const GroupSt* __restrict__ gst, int hash_cap, int* hkeys, int* howner,
              int* slot_of, const int* __restrict__ nd, const int* __restrict__ hash_cap_dev) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= dev_count(n, nd)) return;
  if (hash_cap_dev) hash_cap = *hash_cap_dev;
  slot_of[e] = -1;
  GroupSt S = gst[g_first + e];
  int bits = __float_as_int(S.s2.y);
  if (!(bits & kActN)) return;
  int4 hd = geo[g_first + e].hdr;
  int a = hd.y, b = hd.z;
  int ka = a < b ? a : b, kb = a < b ? b : a;
  int key = encode_pair_key(ka, kb, g_cidx(hd.x));
  int h = hash_pair_slot(key, hash_cap);
  for (int probe = 0; probe < hash_cap; ++probe) {
    int slot = (h + probe) % hash_cap;
    int prev = atomicCAS(&hkeys[slot], kHashEmpty, key);
    if (prev == kHashEmpty || prev == key) {
      slot_of[e] = slot;
      atomicMax(&howner[slot], e);
      return;
    }
  }
}

__global__ void __launch_bounds__(256)
k_hash_store(int n, int g_first, const GroupGeo* __restrict__ geo,
             const GroupSt* __restrict__ gst, const int* __restrict__ howner,
             const int* __restrict__ slot_of, float4* __restrict__ hstate,
             int* __restrict__ hage, const int* __restrict__ nd) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= dev_count(n, nd)) return;
  int slot = slot_of[e];
  if (slot < 0 || howner[slot] != e) return;
  GroupSt S = gst[g_first + e];
  int nr = g_nrows(geo[g_first + e].hdr.x);
  int bits = __float_as_int(S.s2.y);
  float lt = 0.f, lb = 0.f, kt = kPenaltyMinTangent, kb = kPenaltyMinTangent;
  float wt = 0.f, wb = 0.f;
  if (nr == 3) {
    lt = S.s0.y; lb = S.s0.z; kt = S.s1.y; kb = S.s1.z;
    wt = (bits & kWasT) ? 1.0f : 0.0f;
    wb = (bits & kWasB) ? 1.0f : 0.0f;
  }
  hstate[2 * (size_t)slot] = make_float4(S.s0.x, lt, lb, S.s1.x);
  hstate[2 * (size_t)slot + 1] = make_float4(kt, kb, wt, wb);
  if (hage) hage[slot] = __float_as_int(S.s2.z);
}

// The cache refresh of a substep whose row-pool overflow grew the pool: the
// reference re-creates its pool index arrays zeroed (solver_6dof.py:1762-1768)
// before the refresh reads them (:1450-1464), so every pool entry names row
// 0 for its normal and both tangent rows and contact index 0. The table then
// holds one entry: row 0's bodies (key with contact 0) and row 0's state in
// all three row slots, if row 0 is active. Row 0 is group 0's first row.
__global__ void k_hash_row0(const GroupGeo* __restrict__ geo, const GroupSt* __restrict__ gst,
                            int hash_cap, int* hkeys, float4* hstate, int* hage) {
  if (threadIdx.x != 0) return;
  const GroupGeo G = geo[0];
  const GroupSt S = gst[0];
  const int bits = __float_as_int(S.s2.y);
  if (!(bits & kActN)) return;
  const int a = G.hdr.y, b = G.hdr.z;
  const int ka = a < b ? a : b, kb = a < b ? b : a;
  const int key = encode_pair_key(ka, kb, 0);
  const int slot = hash_pair_slot(key, hash_cap);
  const float lam = S.s0.x, k = S.s1.x, was = (bits & kWasN) ? 1.0f : 0.0f;
  hkeys[slot] = key;
  hstate[2 * (size_t)slot] = make_float4(lam, lam, lam, k);
  hstate[2 * (size_t)slot + 1] = make_float4(k, k, was, was);
  if (hage) hage[slot] = 0;
}

void launch_hash_row0(hipStream_t st, const GroupGeo* geo, const GroupSt* gst, int hash_cap,
                      int* hkeys, float4* hstate, int* hage) {
  hipLaunchKernelGGL(k_hash_row0, dim3(1), dim3(64), 0, st, geo, gst, hash_cap, hkeys, hstate,
                     hage);
}

void launch_hash_collect(hipStream_t st, int n, int g_first, const GroupGeo* geo,
                         const GroupSt* gst, int hash_cap, int* hkeys, int* howner, int* slot_of,
                         float4* hstate, int* hage, const int* nd, const int* hash_cap_dev) {
  if (n <= 0 || hash_cap <= 0) return;
  hipLaunchKernelGGL(k_hash_insert, dim3((n + 255) / 256), dim3(256), 0, st, n, g_first, geo, gst,
                     hash_cap, hkeys, howner, slot_of, nd, hash_cap_dev);
  hipLaunchKernelGGL(k_hash_store, dim3((n + 255) / 256), dim3(256), 0, st, n, g_first, geo, gst,
                     howner, slot_of, hstate, hage, nd);
}

}  // namespace avbd
