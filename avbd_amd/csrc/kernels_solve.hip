// kernels_solve.hip — per-substep solve kernels for gfx950.
//
//  k_substep_begin  predict_inertial_6dof + substep_prelude_6dof (+ the
//                   main-pass cache_alpha_C0) in one launch
//                   (kernels_6dof.py:122-187, 1565-1672)
//  k_primal_color   primal_update_6dof for one color (kernels_6dof.py:322-582)
//  k_dual           dual_update_6dof (kernels_6dof.py:1451-1548)
//  k_finalize       finalize_and_cap_6dof (kernels_6dof.py:1871-1901)
//
// Body poses are double-buffered: the substep reads x_initial from the
// buffer that holds the state at substep start and iterates on the other
// one, so predict (writer of the warm start) and the row prelude (reader of
// x_initial) share a launch without a race.
//
// Per body, constraint rows are visited in ascending row order (groups in
// CSR order, rows n -> t -> b inside a group) and accumulated serially, the
// order the reference's serial CPU launch uses; every float operation is the
// one in avbd_fp.h, so results round like the host restatement.
#include "avbd_internal.h"

namespace avbd {

__device__ __forceinline__ v3 f3(float4 a) { return mk3(a.x, a.y, a.z); }
__device__ __forceinline__ q4 f4q(float4 a) { return mk4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ float4 p4(v3 a) { return make_float4(a.x, a.y, a.z, 0.0f); }
__device__ __forceinline__ float4 p4q(q4 a) { return make_float4(a.x, a.y, a.z, a.w); }

// Component `k` of v (k = 0, 1, anything else -> z) without control flow.
// Selecting with a ternary chain let the compiler sink the pose load into a
// switch on k and leave its address undefined on one path; bit masks keep
// the full vector live and select with v_and/v_or.
__device__ __forceinline__ float pick3(v3 v, int k) {
  int m0 = -(int)(k == 0), m1 = -(int)(k == 1), m2 = ~(m0 | m1);
  int b = (__float_as_int(v.x) & m0) | (__float_as_int(v.y) & m1) | (__float_as_int(v.z) & m2);
  return __int_as_float(b);
}
// Unit axis e_k (k = 0, 1, anything else -> z).
__device__ __forceinline__ v3 axis3(int k) {
  return mk3(k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, (k != 0 && k != 1) ? 1.0f : 0.0f);
}

__device__ __forceinline__ m33 load_m33(const float4* p) {
  float4 r0 = p[0], r1 = p[1], r2 = p[2];
  m33 m;
  m.m[0][0] = r0.x; m.m[0][1] = r0.y; m.m[0][2] = r0.z;
  m.m[1][0] = r1.x; m.m[1][1] = r1.y; m.m[1][2] = r1.z;
  m.m[2][0] = r2.x; m.m[2][1] = r2.y; m.m[2][2] = r2.z;
  return m;
}

// ---------------------------------------------------------------------------
// predict + prelude
// ---------------------------------------------------------------------------
__device__ void predict_body(int i, const float4* __restrict__ pose_cur,
                             float4* __restrict__ pose_work, float4* __restrict__ poseI,
                             float4* __restrict__ inert, const float4* __restrict__ vel,
                             const float4* __restrict__ pvel, const float4* __restrict__ bstat,
                             const float* __restrict__ Ilocal, float dt, v3 g) {
  float m = bstat[i].x;
  v3 x = f3(pose_cur[2 * i]);
  q4 q = f4q(pose_cur[2 * i + 1]);
  m33 R = qtomat(q);
  m33 Il;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Il.m[r][c] = Ilocal[9 * i + 3 * r + c];
  m33 Iw = mmul(mmul(R, Il), mT(R));
  inert[3 * i + 0] = make_float4(Iw.m[0][0], Iw.m[0][1], Iw.m[0][2], 0.0f);
  inert[3 * i + 1] = make_float4(Iw.m[1][0], Iw.m[1][1], Iw.m[1][2], 0.0f);
  inert[3 * i + 2] = make_float4(Iw.m[2][0], Iw.m[2][1], Iw.m[2][2], 0.0f);
  if (m <= 0.0f) {
    poseI[2 * i] = p4(x);
    poseI[2 * i + 1] = p4q(q);
    pose_work[2 * i] = p4(x);
    pose_work[2 * i + 1] = p4q(q);
    return;
  }
  v3 g_dt2 = mul(g, dt * dt);
  v3 v = f3(vel[2 * i]);
  v3 w = f3(vel[2 * i + 1]);
  v3 pv = f3(pvel[2 * i]);
  v3 xin = add(add(x, mul(v, dt)), g_dt2);
  q4 dq = quat_from_rotvec(mul(w, dt));
  q4 qin = qmul(dq, q);
  v3 accel = divs(sub(v, pv), dt);
  float g_norm = length(g);
  float wgt = 0.0f;
  if (g_norm > 0.0f) {
    v3 g_hat = divs(g, g_norm);
    float a_ext = dot(accel, g_hat);
    wgt = clampf_(a_ext / g_norm, 0.0f, 1.0f);
  }
  v3 xw = add(add(x, mul(v, dt)), mul(g_dt2, wgt));
  poseI[2 * i] = p4(xin);
  poseI[2 * i + 1] = p4q(qin);
  pose_work[2 * i] = p4(xw);
  pose_work[2 * i + 1] = p4q(qin);
}

__device__ void prelude_group(int g, const float4* __restrict__ pose0, GroupGeo* __restrict__ geo,
                              GroupSt* __restrict__ gst, float alpha, float gamma, int post) {
  GroupGeo G = geo[g];
  GroupSt S = gst[g];
  const int h = G.hdr.x, kind = g_kind(h), nr = g_nrows(h);
  const int a = G.hdr.y, b = G.hdr.z;
  const float stiff = G.g0.w;
  int bits = __float_as_int(S.s2.y);
  float lam[3] = {S.s0.x, S.s0.y, S.s0.z};
  float kk[3] = {S.s1.x, S.s1.y, S.s1.z};
  float ac0[3] = {S.s0.w, S.s1.w, S.s2.x};
  const float main_alpha = post ? 1.0f : alpha;
  const bool finite_s = !(stiff != stiff) && stiff < kInf;
  // (1) warm-start decay of lambda and penalty (Eq. 19)
  for (int r = 0; r < nr; ++r) {
    float kf = r == 0 ? kPenaltyMin : kPenaltyMinTangent;
    float p = clampf_(kk[r] * gamma, kf, kPenaltyMax);
    if (finite_s) p = fmin_(p, stiff);
    kk[r] = p;
  }
  float lam_dec[3];
  for (int r = 0; r < nr; ++r) lam_dec[r] = post == 0 ? lam[r] * alpha * gamma : lam[r];
  // (2) static-friction state, evaluated on the lower tangent row with the
  // values serial row order gives it: sibling and itself already decayed,
  // partner not yet.
  if (nr == 3 && (bits & kActT) && (bits & kActN)) {
    float lam_n = fabs_(lam_dec[0]);
    int was;
    if (lam_n < 1.0e-9f) {
      was = 0;
    } else {
      float lt = lam_dec[1];
      float lb = lam[2];
      float ltb = sqrt_(lt * lt + lb * lb);
      was = ltb <= G.g2.w * lam_n ? 1 : 0;
    }
    bits = was ? (bits | kWasN) : (bits & ~kWasN);
  }
  // (3) alpha * C0 at x_initial (Eq. 18)
  v3 off_a = f3(G.g1);
  if (kind == kGFloor) {
    v3 x0 = f3(pose0[2 * a]);
    q4 q0 = f4q(pose0[2 * a + 1]);
    v3 rw = add(x0, qrotate(q0, off_a));
    float C0 = rw.y - G.g0.y;
    ac0[0] = (bits & kActN) ? (C0 > 0.0f ? 0.0f : main_alpha * C0) : 0.0f;
    if (nr == 3) {
      ac0[1] = (bits & kActT) ? dot(mk3(1.0f, 0.0f, 0.0f), rw) : 0.0f;
      ac0[2] = (bits & kActB) ? dot(mk3(0.0f, 0.0f, 1.0f), rw) : 0.0f;
    }
  } else if (kind == kGBox) {
    v3 xa = f3(pose0[2 * a]);
    q4 qa = f4q(pose0[2 * a + 1]);
    v3 xb = f3(pose0[2 * b]);
    q4 qb = f4q(pose0[2 * b + 1]);
    v3 ra = add(xa, qrotate(qa, off_a));
    v3 rb = add(xb, qrotate(qb, f3(G.g2)));
    v3 d = sub(ra, rb);
    v3 n = f3(G.g0);
    float C0 = dot(n, d);
    ac0[0] = (bits & kActN) ? (C0 > 0.0f ? 0.0f : main_alpha * C0) : 0.0f;
    if (nr == 3) {
      v3 t = orthonormal_basis(n);
      v3 bb = cross(n, t);
      ac0[1] = (bits & kActT) ? dot(t, d) : 0.0f;
      ac0[2] = (bits & kActB) ? dot(bb, d) : 0.0f;
    }
  } else {  // PIN
    v3 x0 = f3(pose0[2 * a]);
    q4 q0 = f4q(pose0[2 * a + 1]);
    v3 rw = add(x0, qrotate(q0, off_a));
    v3 anc = f3(G.g0);
    float C0 = pick3(rw, b) - pick3(anc, b);
    ac0[0] = (bits & kActN) ? main_alpha * C0 : 0.0f;
  }
  S.s0 = make_float4(lam_dec[0], nr > 1 ? lam_dec[1] : S.s0.y, nr > 2 ? lam_dec[2] : S.s0.z, ac0[0]);
  S.s1 = make_float4(kk[0], nr > 1 ? kk[1] : S.s1.y, nr > 2 ? kk[2] : S.s1.z, ac0[1]);
  S.s2 = make_float4(ac0[2], __int_as_float(bits), S.s2.z, S.s2.w);
  gst[g] = S;
}

__global__ void __launch_bounds__(256)
k_substep_begin(int nb, int n_groups, const float4* __restrict__ pose_cur,
                float4* __restrict__ pose_work, float4* __restrict__ poseI,
                float4* __restrict__ inert, const float4* __restrict__ vel,
                const float4* __restrict__ pvel, const float4* __restrict__ bstat,
                const float* __restrict__ Ilocal, GroupGeo* __restrict__ geo,
                GroupSt* __restrict__ gst, float dt, float gx, float gy, float gz, float alpha,
                float gamma, int post, int do_prelude, unsigned* __restrict__ epoch,
                const int* __restrict__ nd_groups) {
  int tid0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid0 == 0 && epoch) *epoch += 1u;  // stamps of this substep's sweeps
  n_groups = dev_count(n_groups, nd_groups);
  const int total = nb + (do_prelude ? n_groups : 0);
  // grid-stride: a launch sized by capacity runs a bounded grid
  for (int tid = tid0; tid < total; tid += gridDim.x * blockDim.x) {
    if (tid < nb)
      predict_body(tid, pose_cur, pose_work, poseI, inert, vel, pvel, bstat, Ilocal, dt,
                   mk3(gx, gy, gz));
    else
      prelude_group(tid - nb, pose_cur, geo, gst, alpha, gamma, post);
  }
}

// Workgroups of a launch sized by capacity with a device count (each thread
// strides over the elements).
constexpr int kCapBlocks = 8192;

void launch_substep_begin(hipStream_t st, int nb, int n_groups, float4* pose_cur,
                          float4* pose_work, float4* poseI, float4* inert, const float4* vel,
                          const float4* pvel, const float4* bstat, const float* Ilocal,
                          GroupGeo* geo, GroupSt* gst, float dt, float gx, float gy, float gz,
                          float alpha, float gamma, int post, int do_prelude, unsigned* epoch,
                          const int* nd_groups) {
  int total = nb + (do_prelude ? n_groups : 0);
  if (total <= 0) return;
  int blocks = (total + 255) / 256;
  if (nd_groups) blocks = std::min(blocks, kCapBlocks);
  hipLaunchKernelGGL(k_substep_begin, dim3(blocks), dim3(256), 0, st, nb, n_groups,
                     pose_cur, pose_work, poseI, inert, vel, pvel, bstat, Ilocal, geo, gst, dt,
                     gx, gy, gz, alpha, gamma, post, do_prelude, epoch, nd_groups);
}

// ---------------------------------------------------------------------------
// primal
// ---------------------------------------------------------------------------
// Values one constraint row contributes to the body's 6x6 system
// (primal_update_6dof loop body, kernels_6dof.py:507-564): Jacobian, the
// Eq. 14 LHS stiffness, the clamped force and the geometric-stiffness
// diagonal |f| G~ (used only when |f| > 0).
struct RowVals {
  v3 jl, ja;
  float k, f;
  v3 g;
};

__device__ __forceinline__ void row_values(RowVals& o, bool& gflag, v3 jl, v3 ja, float C, v3 nG,
                                           v3 r_self, bool hard, float kp, float lam, float ac0,
                                           bool is_t, float lam_sib, bool was_sib, float mu_d,
                                           float mu_s, float f_lo_n, float f_hi_n) {
  if (hard) C = C - ac0;
  float lam_eff = hard ? lam : 0.0f;
  float lam_plus = kp * C + lam_eff;
  float f_lo, f_hi;
  if (is_t) {
    float mu = was_sib ? mu_s : mu_d;
    float bound = mu * fabs_(lam_sib);
    f_lo = -bound;
    f_hi = bound;
  } else {
    f_lo = f_lo_n;
    f_hi = f_hi_n;
  }
  float f = clampf_(lam_plus, f_lo, f_hi);
  float k_lhs = kp;
  float absC = fabs_(C);
  // Eq. 14: the bound that clipped lambda+ over |C| (one division for
  // either bound)
  const bool below = lam_plus < f_lo;
  if (absC > 1.0e-12f && (below || lam_plus > f_hi)) {
    const float bound = below ? f_lo : f_hi;
    k_lhs = fabs_(bound - lam_plus) / absC;
  }
  o.jl = jl;
  o.ja = ja;
  o.k = k_lhs;
  o.f = f;
  float fmag = fabs_(f);
  gflag = fmag > 0.0f;
  o.g = gflag ? mul(geom_stiffness_diag(nG, r_self), fmag) : mk3(0.f, 0.f, 0.f);
}

// Rows of group G seen from body i, in row order. Returns a bit mask of the
// rows present and active (bit r) and sets gmask (bit r: |f| > 0).
// (ox, oq) is the pose of the other body of a box-box group (of the body
// itself for floor and pin groups, where it is not used).
__device__ __forceinline__ int group_rows(RowVals out[3], int& gmask, int i, v3 xi, q4 qi,
                                          const GroupGeo& G, const GroupSt& S, float4 ox,
                                          float4 oq, int post_pass) {
  const int h = G.hdr.x, kind = g_kind(h), nr = g_nrows(h);
  const int bits = __float_as_int(S.s2.y);
  const float stiff = G.g0.w;
  const bool hard = stiff >= kInf;
  const bool wasN = (bits & kWasN) != 0;
  const float lam_n = S.s0.x;
  v3 off_a = f3(G.g1);
  int mask = 0;
  gmask = 0;
  bool gf;
  if (kind == kGFloor || kind == kGBox) {
    // Floor corners and box-box contacts share one instruction stream: a
    // floor row is a contact whose "other side" is the world, with the fixed
    // directions y (normal), x and z (tangents). Each kind keeps its own
    // arithmetic: the floor normal C is (x.y + r.y) - anchor.y and floor
    // tangent C is t . (x + r), box rows use n, t, b . (w_a - w_b).
    const bool is_floor = (kind == kGFloor);
    const int a = G.hdr.y, b = G.hdr.z;
    const bool self_a = is_floor || (a == i);
    (void)b;
    const v3 off_b = f3(G.g2);
    const v3 off_self = self_a ? off_a : off_b;
    const v3 off_other = self_a ? off_b : off_a;
    const v3 r_self = qrotate(qi, off_self);
    const v3 r_o = qrotate(f4q(oq), off_other);
    const v3 ws = add(xi, r_self);
    const v3 wo = add(f3(ox), r_o);
    const v3 d = self_a ? sub(ws, wo) : sub(wo, ws);
    const v3 D = is_floor ? ws : d;
    const v3 gn = f3(G.g0);
    const v3 n = is_floor ? mk3(0.0f, 1.0f, 0.0f) : gn;
    if (bits & kActN) {
      const float C = is_floor ? ws.y - G.g0.y : dot(n, d);
      const v3 jl = self_a ? n : neg(n);
      const v3 ja = self_a ? cross(r_self, n) : neg(cross(r_self, n));
      row_values(out[0], gf, jl, ja, C, n, r_self, hard, S.s1.x, S.s0.x,
                 post_pass ? 0.0f : S.s0.w, false, 0.f, false, 0.f, 0.f, -kInf, 0.0f);
      mask |= 1;
      gmask |= gf ? 1 : 0;
    }
    if (nr == 3) {
      const v3 bt = orthonormal_basis(n);
      const v3 t = is_floor ? mk3(1.0f, 0.0f, 0.0f) : bt;
      const v3 bb = is_floor ? mk3(0.0f, 0.0f, 1.0f) : cross(n, bt);
      if (bits & kActT) {
        const v3 jl = self_a ? t : neg(t);
        const v3 ja = self_a ? cross(r_self, t) : neg(cross(r_self, t));
        row_values(out[1], gf, jl, ja, dot(t, D), t, r_self, hard, S.s1.y, S.s0.y, S.s1.w, true,
                   lam_n, wasN, G.g1.w, G.g2.w, 0.f, 0.f);
        mask |= 2;
        gmask |= gf ? 2 : 0;
      }
      if (bits & kActB) {
        const v3 jl = self_a ? bb : neg(bb);
        const v3 ja = self_a ? cross(r_self, bb) : neg(cross(r_self, bb));
        row_values(out[2], gf, jl, ja, dot(bb, D), bb, r_self, hard, S.s1.z, S.s0.z, S.s2.x,
                   true, lam_n, wasN, G.g1.w, G.g2.w, 0.f, 0.f);
        mask |= 4;
        gmask |= gf ? 4 : 0;
      }
    }
  } else {  // PIN
    if (bits & kActN) {
      const int axis = G.hdr.z;
      v3 r_self = qrotate(qi, off_a);
      v3 wp = add(xi, r_self);
      v3 anc = f3(G.g0);
      v3 n = axis3(axis);
      float C = pick3(wp, axis) - pick3(anc, axis);
      row_values(out[0], gf, n, cross(r_self, n), C, n, r_self, hard, S.s1.x, S.s0.x,
                 post_pass ? 0.0f : S.s0.w, false, 0.f, false, 0.f, 0.f, -kInf, kInf);
      mask |= 1;
      gmask |= gf ? 1 : 0;
    }
  }
  return mask;
}

struct Acc {
  m33 A, B, D;
  v3 rl, ra;
};

// Body-side terms of the system before any row (kernels_6dof.py:373-399).
__device__ __forceinline__ void acc_init(Acc& acc, float m, const m33& Iw, v3 xi, q4 qi,
                                         float4 xI, float4 qI, float inv_dt2) {
  acc.A = mzero();
  acc.A.m[0][0] = m * inv_dt2;
  acc.A.m[1][1] = m * inv_dt2;
  acc.A.m[2][2] = m * inv_dt2;
  acc.D = mscale(Iw, inv_dt2);
  acc.B = mzero();
  acc.rl = mul(sub(xi, f3(xI)), m * inv_dt2);
  q4 dq = qmul(qi, qconj(f4q(qI)));
  acc.ra = mvmul(Iw, mul(quat_to_rotvec(dq), inv_dt2));
}

__device__ __forceinline__ void acc_add(Acc& acc, const RowVals& v, bool gflag) {
  float l[3] = {v.jl.x, v.jl.y, v.jl.z};
  float an[3] = {v.ja.x, v.ja.y, v.ja.z};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc.A.m[r][c] = acc.A.m[r][c] + (l[r] * l[c]) * v.k;
      acc.B.m[r][c] = acc.B.m[r][c] + (an[r] * l[c]) * v.k;
      acc.D.m[r][c] = acc.D.m[r][c] + (an[r] * an[c]) * v.k;
    }
  }
  if (gflag) {
    acc.D.m[0][0] = acc.D.m[0][0] + v.g.x;
    acc.D.m[1][1] = acc.D.m[1][1] + v.g.y;
    acc.D.m[2][2] = acc.D.m[2][2] + v.g.z;
  }
  acc.rl = add(acc.rl, mul(v.jl, v.f));
  acc.ra = add(acc.ra, mul(v.ja, v.f));
}

// Schur-complement solve of [A B^T; B D] and the pose update
// (kernels_6dof.py:566-582).
__device__ __forceinline__ void solve_pose(const Acc& acc, v3 xi, q4 qi, float4& xo, float4& qo) {
  m33 A_inv = minverse(acc.A);
  m33 BAinv = mmul(acc.B, A_inv);
  v3 rhs = sub(acc.ra, mvmul(BAinv, acc.rl));
  m33 Sm = msub(acc.D, mmul(BAinv, mT(acc.B)));
  m33 S_inv = minverse(Sm);
  v3 d_th = mvmul(S_inv, rhs);
  v3 d_x = mvmul(A_inv, sub(acc.rl, mvmul(mT(acc.B), d_th)));
  xo = p4(sub(xi, d_x));
  q4 dq2 = quat_from_rotvec(neg(d_th));
  qo = p4q(qnormalize(qmul(dq2, qi)));
}

__device__ __forceinline__ void solve_and_apply(const Acc& acc, int i, v3 xi, q4 qi,
                                                float4* __restrict__ pose) {
  float4 xo, qo;
  solve_pose(acc, xi, qi, xo, qo);
  pose[2 * i] = xo;
  pose[2 * i + 1] = qo;
}

// Agent-scope relaxed accesses (sc1): coherent across the GPU's L2 slices,
// used for data exchanged between wavefronts of one launch.
__device__ __forceinline__ float4 ld_agent(const float4* p) {
  const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
  unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float4(__uint_as_float((unsigned)a), __uint_as_float((unsigned)(a >> 32)),
                     __uint_as_float((unsigned)b), __uint_as_float((unsigned)(b >> 32)));
}
__device__ __forceinline__ void st_agent(float4* p, float4 v) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  __hip_atomic_store(q, (unsigned long long)__float_as_uint(v.x) |
                            ((unsigned long long)__float_as_uint(v.y) << 32),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, (unsigned long long)__float_as_uint(v.z) |
                                ((unsigned long long)__float_as_uint(v.w) << 32),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane-per-body variant: each lane walks its body's groups serially.
__global__ void __launch_bounds__(64)
k_primal_color(int count, const int* __restrict__ color_bodies, int base,
               float4* __restrict__ pose, const float4* __restrict__ poseI,
               const float4* __restrict__ inert, const float4* __restrict__ bstat,
               const int* __restrict__ cpos, const int2* __restrict__ csr,
               const int* __restrict__ pos_map, const GroupGeo* __restrict__ geo,
               const GroupSt* __restrict__ gst, float dt, int post_pass) {
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= count) return;
  const int k = base + tid;
  const int kc = pos_map ? pos_map[k] : k;  // CSR position of this body
  int i = color_bodies[k];
  float m = bstat[i].x;
  if (m <= 0.0f) return;
  float inv_dt2 = 1.0f / (dt * dt);
  v3 xi = f3(pose[2 * i]);
  q4 qi = f4q(pose[2 * i + 1]);
  Acc acc;
  acc_init(acc, m, load_m33(inert + 3 * i), xi, qi, poseI[2 * i], poseI[2 * i + 1], inv_dt2);
  const int k0 = cpos[kc], k1 = cpos[kc + 1];
  for (int e = k0; e < k1; ++e) {
    const int2 ent = csr[e];
    if (ent.x < 0) break;  // padding ends the list
    GroupGeo G = geo[ent.x];
    GroupSt S = gst[ent.x];
    RowVals rv[3];
    int gmask;
    int mask = group_rows(rv, gmask, i, xi, qi, G, S, pose[2 * ent.y], pose[2 * ent.y + 1],
                          post_pass);
    for (int r = 0; r < 3; ++r)
      if (mask & (1 << r)) acc_add(acc, rv[r], (gmask >> r) & 1);
  }
  solve_and_apply(acc, i, xi, qi, pose);
}

// ---------------------------------------------------------------------------
// dual: one lane per group, rows n -> t -> b in order
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dual_row(float C, bool hard, float stiff, float& lam, float& kp,
                                         float ac0, float lo, float hi, float fracture,
                                         bool is_t, bool used_static, int act_bit, int& bits,
                                         float beta, float& rmax) {
  if (hard) C = C - ac0;
  float lam_eff = hard ? lam : 0.0f;
  float nl = clampf_(kp * C + lam_eff, lo, hi);
  // residual of this row: |lambda_new - lambda| / k with lambda, k before the
  // update; a non-finite state counts as an infinite residual
  float r = fabs_(nl - lam) / kp;
  r = r != r ? kInf : r;
  rmax = r > rmax ? r : rmax;
  lam = nl;
  if (fabs_(nl) >= fracture) {
    bits &= ~act_bit;
    lam = 0.0f;
    kp = 0.0f;
    return;
  }
  if (nl > lo && nl < hi) {
    float upper = fmin_(kPenaltyMax, stiff);
    kp = fmin_(kp + beta * fabs_(C), upper);
  } else if (is_t && used_static) {
    bits &= ~kWasN;
  }
}

// Dual update of one contact group (dual_update_6dof, kernels_6dof.py:1451-1548)
// from the final poses of its bodies a (xa, qa) and b (xb, qb; unused for
// floor and pin groups). Writes back only the state words that changed.
// resid (optional): atomically raised to the group's largest row residual.
__device__ __forceinline__ void dual_group(const GroupGeo& G, GroupSt S, GroupSt* __restrict__ out,
                                           float4 xa, float4 qa, float4 xb, float4 qb,
                                           float beta, float* resid = nullptr) {
  const int h = G.hdr.x, kind = g_kind(h), nr = g_nrows(h);
  const int b = G.hdr.z;
  int bits = __float_as_int(S.s2.y);
  if (!(bits & (kActN | kActT | kActB))) return;
  const float stiff = G.g0.w;
  const bool hard = stiff >= kInf;
  float lam[3] = {S.s0.x, S.s0.y, S.s0.z};
  float kk[3] = {S.s1.x, S.s1.y, S.s1.z};
  const float ac0[3] = {S.s0.w, S.s1.w, S.s2.x};
  float C[3] = {0.f, 0.f, 0.f};
  v3 off_a = f3(G.g1);
  if (kind == kGFloor) {
    v3 rw = add(f3(xa), qrotate(f4q(qa), off_a));
    C[0] = rw.y - G.g0.y;
    C[1] = dot(mk3(1.0f, 0.0f, 0.0f), rw);
    C[2] = dot(mk3(0.0f, 0.0f, 1.0f), rw);
  } else if (kind == kGBox) {
    v3 ra = add(f3(xa), qrotate(f4q(qa), off_a));
    v3 rb = add(f3(xb), qrotate(f4q(qb), f3(G.g2)));
    v3 d = sub(ra, rb);
    v3 n = f3(G.g0);
    C[0] = dot(n, d);
    if (nr == 3) {
      v3 t = orthonormal_basis(n);
      C[1] = dot(t, d);
      C[2] = dot(cross(n, t), d);
    }
  } else {
    v3 rw = add(f3(xa), qrotate(f4q(qa), off_a));
    v3 anc = f3(G.g0);
    C[0] = pick3(rw, b) - pick3(anc, b);
  }
  const int bits0 = bits;
  float rmax = 0.0f;
  // normal / pin row
  if (bits & kActN) {
    float hi0 = kind == kGPin ? kInf : 0.0f;
    float frac0 = kind == kGPin ? G.g2.w : kInf;
    dual_row(C[0], hard, stiff, lam[0], kk[0], ac0[0], -kInf, hi0, frac0, false, false, kActN,
             bits, beta, rmax);
  }
  // tangent rows: bounds from this launch's normal lambda and was_static
  for (int r = 1; r < nr; ++r) {
    int act = r == 1 ? kActT : kActB;
    if (!(bits & act)) continue;
    bool used_static = (bits & kWasN) != 0;
    float mu = used_static ? G.g2.w : G.g1.w;
    float bound = mu * fabs_(lam[0]);
    dual_row(C[r], hard, stiff, lam[r], kk[r], ac0[r], -bound, bound, kInf, true, used_static,
             act, bits, beta, rmax);
  }
  if (resid && rmax > 0.0f) atomicMax(reinterpret_cast<int*>(resid), __float_as_int(rmax));
  // write back what changed, compared bit for bit (a separated floor corner
  // keeps lambda = 0 and k, so most floor groups write nothing)
  const float4 n0 = make_float4(lam[0], lam[1], lam[2], S.s0.w);
  const float4 n1 = make_float4(kk[0], kk[1], kk[2], S.s1.w);
  if (__float_as_int(n0.x) != __float_as_int(S.s0.x) || __float_as_int(n0.y) != __float_as_int(S.s0.y) ||
      __float_as_int(n0.z) != __float_as_int(S.s0.z))
    out->s0 = n0;
  if (__float_as_int(n1.x) != __float_as_int(S.s1.x) || __float_as_int(n1.y) != __float_as_int(S.s1.y) ||
      __float_as_int(n1.z) != __float_as_int(S.s1.z))
    out->s1 = n1;
  if (bits != bits0) {
    float4 n2 = S.s2;
    n2.y = __int_as_float(bits);
    out->s2 = n2;
  }
}

// Tiled variant: one 64-lane wavefront per tile of BPW bodies of the color.
//  A. lanes evaluate one (body, group) incidence each -> row values in LDS
//  B. 30 lanes per body own one system entry each (A upper 6, B 9, D 9,
//     r_lin 3, r_ang 3) and add the row contributions in CSR order
//  C. one lane per body gathers its system and solves.
// The additions per entry happen in the same order as the lane-per-body
// variant, so both give identical bits.
//
// Phase B has no per-row conditions. A row that is absent, inactive or a
// padding slot is stored as jl = ja = +0, k = f = g = -0: each of its terms,
// (+0 * +0) * -0 or (+0 * -0) * 1, is -0, and x + (-0) == x bit for bit for
// every x (round to nearest), so adding it changes nothing. Likewise g = -0
// stands for "no geometric-stiffness term" (|f| == 0), and lanes that own an
// entry without such a term read a constant -0 field. Each body's slots are
// padded to a multiple of four so that phase B reads four slots per float4.
// Per row: jl 3, ja 3, k, f, g 3, then two constant fields (1.0 and -0.0),
// repeated in every row so that each operand's LDS address is the lane's
// field offset plus the row offset.
constexpr int kRowFields = 13;
constexpr int kFieldOne = 11;
constexpr int kFieldNegZero = 12;
constexpr int kSlotFields = 3 * kRowFields;
constexpr int kComps = 30;

// Fields that system entry ac (A upper 0-5, B 6-14, D 15-23, rl 24-26,
// ra 27-29) multiplies, per row:
//   A[r][c] += (jl_r * jl_c) * k      B[r][c] += (ja_r * jl_c) * k
//   D[r][c] += (ja_r * ja_c) * k, then += g_r on the diagonal
//   rl[r]   += (jl_r * f) * 1         ra[r]   += (ja_r * f) * 1
// (the factor 1.0 is exact). fz is k or the constant 1.0 field, fg is g_r
// or the constant -0 field.
__device__ __forceinline__ void entry_fields(int ac, int& fa, int& fb, int& fz, int& fg) {
  fz = 6;
  fg = kFieldNegZero;
  if (ac < 6) {
    const int rr = ac < 3 ? 0 : (ac < 5 ? 1 : 2);
    const int cc = ac < 3 ? ac : (ac < 5 ? ac - 2 : 2);
    fa = rr; fb = cc;
  } else if (ac < 15) {
    fa = 3 + (ac - 6) / 3; fb = (ac - 6) % 3;
  } else if (ac < 24) {
    const int cr = (ac - 15) / 3, cc = (ac - 15) % 3;
    fa = 3 + cr; fb = 3 + cc;
    if (cr == cc) fg = 8 + cr;
  } else if (ac < 27) {
    fa = ac - 24; fb = 7; fz = kFieldOne;
  } else {
    fa = 3 + (ac - 27); fb = 7; fz = kFieldOne;
  }
}

// Dataflow sweep: every color of one sweep in one launch. Workgroup t is
// tile t in color-major order (BPW consecutive positions of one color);
// before using a neighbour of a lower color a tile waits until that
// neighbour has published its pose of this sweep (done[j] == stamp), so each
// body reads exactly the poses that color-by-color launches would give it.
// A tile waits only on tiles with smaller indices. The dispatcher starts
// workgroups in index order, so those are already running or done and every
// wait ends; a spin cap with an error flag turns any violation of that into
// a reported error instead of a hang. (A ticket counter would not depend on
// dispatch order, but one device-wide atomic per tile serializes ~50k tiles
// per sweep: measured 2x slower than the color-by-color launches.)
struct SweepFlow {
  int n_colors;
  int n_tiles;
  int tile_start[kMaxColors + 1];   // first tile of each color
  int color_start[kMaxColors + 1];  // first position of each color
  const int* color;                 // body -> color
  int* done;                        // body -> stamp of the last sweep that moved it
  // The sweep's stamp is epoch * kSweepStride + sweep + 1 (mod 2^32): the
  // epoch is a device counter that k_substep_begin advances once per
  // substep, so replays of a captured substep get fresh stamps, and two
  // consecutive sweeps never share one. Every body publishes in every sweep,
  // so a stamp only has to differ from the previous sweep's.
  const unsigned* epoch;
  int sweep;
  int* err;
  // Waits poll done_rd (normally done) and give up after spin_cap polls,
  // setting *err. A host test points done_rd at an array nobody writes, with
  // a small cap, to run the give-up path without a hang.
  const int* done_rd;
  int spin_cap;
  // Color layout read from the device (graph-captured substeps): when
  // meta_ncolors is set, the color count, color starts and first tiles come
  // from these arrays instead of the fields above, and workgroups past the
  // last tile exit.
  const int* meta_ncolors;
  const int* meta_cstart;
  const int* meta_tiles;
  // Dual update fused into the sweep: each group is updated by the tile of
  // its higher-color body (the body itself for floor and pin groups), after
  // that tile's solve, when both bodies' poses of this sweep are final and
  // no other tile reads the group again in this sweep. Valid only for a
  // coloring without same-color neighbours.
  int fuse_dual;
  float beta;
  GroupSt* gst_w;
  // Dual update in trailing workgroups of the same launch: workgroups past
  // the last tile take 64 groups each, wait until both bodies of a group
  // have published this sweep's pose, and update it. Only the tiles of a
  // group's two bodies read its state, and both are done by then, so the
  // result is the separate dual launch's for any coloring. tail_nd
  // (graph-captured substeps): the live group count on the device, of which
  // tail_groups is the capacity.
  int tail_groups;
  const int* tail_nd;
};

constexpr int kFlowSpinCap = 1 << 20;
constexpr unsigned kSweepStride = 4096;  // sweeps per substep are fewer

__device__ __forceinline__ int sweep_stamp(const SweepFlow& fl) {
  return (int)(*fl.epoch * kSweepStride + (unsigned)fl.sweep + 1u);
}

// kSameColorOrder (one body per tile): a body also waits for its same-color
// neighbours of lower index, so a color that holds neighbours (a coloring
// past its 32 colors) is updated in body order, as a serial loop over the
// color does; tiles are in that order, so the waits still point backwards.
template <int BPW, bool kFlow, bool kSameColorOrder = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BPW >= 4 ? 3 : 1)))
k_primal_tiled(int count_arg, const int* __restrict__ color_bodies, int base_arg,
               float4* __restrict__ pose, const float4* __restrict__ poseI,
               const float4* __restrict__ inert, const float4* __restrict__ bstat,
               const int* __restrict__ cpos, const int2* __restrict__ csr,
               const GroupGeo* __restrict__ geo, const GroupSt* __restrict__ gst, float dt,
               int post_pass, SweepFlow fl) {
  // Row fields of the 64 slots of a pass, one row of 68 floats per field
  // (16-byte aligned so that four consecutive slots load as one float4).
  __shared__ __align__(16) float rows[kSlotFields][68];
  __shared__ float sys[BPW][kComps];
  __shared__ int b_i[BPW], b_id[BPW], b_off[BPW + 1];
  __shared__ float b_pose[BPW][7];
  const int lane = threadIdx.x;
  const int stamp = kFlow ? sweep_stamp(fl) : 0;
  if (kFlow && fl.tail_groups > 0 && (int)blockIdx.x >= fl.n_tiles) {
    const int g = ((int)blockIdx.x - fl.n_tiles) * 64 + lane;
    if (g < dev_count(fl.tail_groups, fl.tail_nd)) {
      const GroupGeo G = geo[g];
      const bool box = g_kind(G.hdr.x) == kGBox;
      const int a = G.hdr.y, b = box ? G.hdr.z : G.hdr.y;
      int spins = 0;
      while (__hip_atomic_load(fl.done_rd + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != stamp ||
             __hip_atomic_load(fl.done_rd + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != stamp) {
        if (++spins > fl.spin_cap) {
          atomicExch(fl.err, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      const float4 xa = ld_agent(pose + 2 * a), qa = ld_agent(pose + 2 * a + 1);
      float4 xb = xa, qb = qa;
      if (box) {
        xb = ld_agent(pose + 2 * b);
        qb = ld_agent(pose + 2 * b + 1);
      }
      dual_group(G, gst[g], fl.gst_w + g, xa, qa, xb, qb, fl.beta);
    }
    return;
  }
  int count = count_arg, base = base_arg, first = blockIdx.x * BPW, my_color = 0;
  if (kFlow && fl.meta_ncolors != nullptr) {
    const int t = blockIdx.x, nc = *fl.meta_ncolors;
    if (t >= fl.meta_tiles[nc]) return;
    int c = 0;
    while (c + 1 < nc && t >= fl.meta_tiles[c + 1]) ++c;
    my_color = c;
    base = fl.meta_cstart[c];
    count = fl.meta_cstart[c + 1] - base;
    first = (t - fl.meta_tiles[c]) * BPW;
  } else if (kFlow) {
    const int t = blockIdx.x;
    int c = 0;
    while (c + 1 < fl.n_colors && t >= fl.tile_start[c + 1]) ++c;
    my_color = c;
    base = fl.color_start[c];
    count = fl.color_start[c + 1] - base;
    first = (t - fl.tile_start[c]) * BPW;
  }
  const float inv_dt2 = 1.0f / (dt * dt);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    rows[kRowFields * r + kFieldOne][lane] = 1.0f;
    rows[kRowFields * r + kFieldNegZero][lane] = -0.0f;
  }
  // Tile slots: the CSR entries of the tile's positions, contiguous and
  // padded per body to a multiple of four. A pass gives each body a chunk of
  // kChunk consecutive slots of its own list (lanes [b * kChunk, (b + 1) *
  // kChunk) for body b), so that every body's slots are evaluated and
  // accumulated at the same pace and the accumulation of a pass costs
  // kChunk / 4 steps whatever the bodies' list lengths.
  constexpr int kChunk = (64 / BPW) & ~3;
  const int e0 = cpos[base + first];
  const int total = cpos[base + (first + BPW < count ? first + BPW : count)] - e0;
  const int pb = lane / kChunk, pq = lane % kChunk;  // this lane's body and slot in the chunk
  int q_lo = 0, q_hi = 0;                            // its body's CSR range
  if (pb < BPW) {
    const int s0 = first + pb < count ? first + pb : count;
    const int s1 = first + pb + 1 < count ? first + pb + 1 : count;
    q_lo = cpos[base + s0];
    q_hi = cpos[base + s1];
  }
  // Loads of the first pass go out before the body set-up below, so that
  // their latency overlaps it: list entry, group records and (when the
  // neighbour cannot move during this launch) the other body's pose.
  int slot = q_lo + pq;
  int2 ent = slot < q_hi ? csr[slot] : make_int2(-1, -1);
  GroupGeo Gq;
  GroupSt Sq;
  float4 oxq, oqq;
  int colq = 0, flagq = 0;  // dataflow: the other body's color and sweep stamp
  if (ent.x >= 0) {
    Gq = geo[ent.x];
    Sq = gst[ent.x];
    if (!kFlow) {
      oxq = pose[2 * ent.y];
      oqq = pose[2 * ent.y + 1];
    } else {
      colq = fl.color[ent.y];
      flagq = __hip_atomic_load(fl.done_rd + ent.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (lane <= BPW) {
    const int q = first + lane < count ? first + lane : count;
    b_off[lane] = cpos[base + q] - e0;
  }
  if (lane < BPW) {
    const int pos = first + lane;
    int i = -1, id = -1;
    if (pos < count) {
      id = i = color_bodies[base + pos];
      if (!(bstat[i].x > 0.0f)) i = -1;
    }
    b_i[lane] = i;
    b_id[lane] = id;
    if (kFlow && i < 0 && id >= 0) {
      // static body: its (unchanged) pose for the fused dual
      const float4 xp = pose[2 * id], qp = pose[2 * id + 1];
      b_pose[lane][0] = xp.x; b_pose[lane][1] = xp.y; b_pose[lane][2] = xp.z;
      b_pose[lane][3] = qp.x; b_pose[lane][4] = qp.y; b_pose[lane][5] = qp.z; b_pose[lane][6] = qp.w;
    }
    if (i >= 0) {
      float4 xp = pose[2 * i], qp = pose[2 * i + 1];
      b_pose[lane][0] = xp.x; b_pose[lane][1] = xp.y; b_pose[lane][2] = xp.z;
      b_pose[lane][3] = qp.x; b_pose[lane][4] = qp.y; b_pose[lane][5] = qp.z; b_pose[lane][6] = qp.w;
      Acc acc;
      acc_init(acc, bstat[i].x, load_m33(inert + 3 * i), mk3(xp.x, xp.y, xp.z),
               mk4(qp.x, qp.y, qp.z, qp.w), poseI[2 * i], poseI[2 * i + 1], inv_dt2);
      float* o = sys[lane];
      o[0] = acc.A.m[0][0]; o[1] = acc.A.m[0][1]; o[2] = acc.A.m[0][2];
      o[3] = acc.A.m[1][1]; o[4] = acc.A.m[1][2]; o[5] = acc.A.m[2][2];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          o[6 + 3 * r + c] = acc.B.m[r][c];
          o[15 + 3 * r + c] = acc.D.m[r][c];
        }
      o[24] = acc.rl.x; o[25] = acc.rl.y; o[26] = acc.rl.z;
      o[27] = acc.ra.x; o[28] = acc.ra.y; o[29] = acc.ra.z;
    }
  }
  __syncthreads();
  // passes: the longest list of the tile in chunks
  int longest = 0;
#pragma unroll
  for (int b = 0; b < BPW; ++b) {
    const int nb_slots = b_off[b + 1] - b_off[b];
    longest = nb_slots > longest ? nb_slots : longest;
  }
  const int n_pass = (longest + kChunk - 1) / kChunk;
  // Accumulator lanes: kChunk lanes per body, each owning kEpl entries
  // (entry ac + k * kChunk, k < kEpl, when below 30).
  constexpr int kLanes = kChunk;
  constexpr int kEpl = (kComps + kLanes - 1) / kLanes;
  const int ab = pb, al = pq;
  const bool body_ok = ab < BPW && b_i[ab < BPW ? ab : 0] >= 0;
  const int ab_slots = ab < BPW ? b_off[ab + 1] - b_off[ab] : 0;
  // With an even number of entries per lane, the first half of a lane's
  // entries are A and B entries (0-14), which take no geometric-stiffness
  // term (their g is the constant -0, and x + -0 == x), so their additions
  // of g are left out; the second half are D, r_lin and r_ang (15-29).
  constexpr bool kSplit = (kEpl % 2) == 0;
  constexpr int kHalf = kEpl / 2;
  int fa[kEpl], fb[kEpl], fz[kEpl], fg[kEpl], acs[kEpl];
  bool own[kEpl];
  float accv[kEpl];
  bool is_acc = false;
#pragma unroll
  for (int k = 0; k < kEpl; ++k) {
    const int ac = !kSplit ? al + k * kLanes
                           : (k < kHalf ? al + k * kLanes : 15 + al + (k - kHalf) * kLanes);
    const int lim = (kSplit && k < kHalf) ? 15 : kComps;
    own[k] = body_ok && ac < lim;
    acs[k] = ac < lim ? ac : 0;
    entry_fields(acs[k], fa[k], fb[k], fz[k], fg[k]);
    accv[k] = own[k] ? sys[ab][acs[k]] : 0.0f;
    is_acc = is_acc || own[k];
  }
  const bool self_ok = pb < BPW && b_i[pb < BPW ? pb : 0] >= 0;
  const int self_id = pb < BPW ? b_id[pb] : -1;
  for (int p = 0; p < n_pass; ++p) {
    // ---- A: row values of one incidence per lane ----
    const bool live = slot < q_hi;
    // list entry of this lane's slot in the next pass
    const int slot_next = slot + kChunk;
    const int2 ent_next = slot_next < q_hi ? csr[slot_next] : make_int2(-1, -1);
    if (kFlow) {
      // wait for this pass's lower-color neighbours (their color and stamp
      // were loaded with the group records), then read their poses through
      // the coherent path, after the stamp was seen
      const bool before_me =
          colq < my_color || (kSameColorOrder && colq == my_color && ent.y < self_id);
      bool need = ent.x >= 0 && ent.y >= 0 && ent.y != self_id && before_me && flagq != stamp;
      int spins = 0;
      while (need) {
        if (__hip_atomic_load(fl.done_rd + ent.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            stamp) {
          need = false;
        } else if (++spins > fl.spin_cap) {
          atomicExch(fl.err, 1);
          need = false;
        } else {
          __builtin_amdgcn_s_sleep(1);
        }
      }
      if (ent.x >= 0) {
        oxq = ld_agent(pose + 2 * ent.y);
        oqq = ld_agent(pose + 2 * ent.y + 1);
      }
    }
    if (live) {
      RowVals rv[3];
      int mask = 0, gmask = 0;
      if (ent.x >= 0 && self_ok) {
        v3 xi = mk3(b_pose[pb][0], b_pose[pb][1], b_pose[pb][2]);
        q4 qi = mk4(b_pose[pb][3], b_pose[pb][4], b_pose[pb][5], b_pose[pb][6]);
        mask = group_rows(rv, gmask, b_i[pb], xi, qi, Gq, Sq, oxq, oqq, post_pass);
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const bool on = (mask >> r) & 1;
        const bool gon = on && ((gmask >> r) & 1);
        float* o = &rows[kRowFields * r][0];
        o[0 * 68 + lane] = on ? rv[r].jl.x : 0.0f;
        o[1 * 68 + lane] = on ? rv[r].jl.y : 0.0f;
        o[2 * 68 + lane] = on ? rv[r].jl.z : 0.0f;
        o[3 * 68 + lane] = on ? rv[r].ja.x : 0.0f;
        o[4 * 68 + lane] = on ? rv[r].ja.y : 0.0f;
        o[5 * 68 + lane] = on ? rv[r].ja.z : 0.0f;
        o[6 * 68 + lane] = on ? rv[r].k : -0.0f;
        o[7 * 68 + lane] = on ? rv[r].f : -0.0f;
        o[8 * 68 + lane] = gon ? rv[r].g.x : -0.0f;
        o[9 * 68 + lane] = gon ? rv[r].g.y : -0.0f;
        o[10 * 68 + lane] = gon ? rv[r].g.z : -0.0f;
      }
    }
    // loads of the next pass, in flight during this pass's accumulation
    slot = slot_next;
    ent = ent_next;
    if (ent.x >= 0) {
      Gq = geo[ent.x];
      Sq = gst[ent.x];
      if (!kFlow) {
        oxq = pose[2 * ent.y];
        oqq = pose[2 * ent.y + 1];
      } else {
        colq = fl.color[ent.y];
        flagq = __hip_atomic_load(fl.done_rd + ent.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    __syncthreads();
    // ---- B: ordered accumulation of the body's chunk, four slots per step ----
    if (is_acc) {
      const int s0 = ab * kChunk;
      const int left = ab_slots - p * kChunk;
      const int s1 = s0 + (left < kChunk ? (left > 0 ? left : 0) : kChunk);
      for (int sb = s0; sb < s1; sb += 4) {
        __builtin_assume((sb & 3) == 0);  // chunks and list lengths are multiples of 4
#pragma unroll
        for (int k = 0; k < kEpl; ++k) {
          float4 X[3], Y[3], Z[3], Gd[3];
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const int rb = kRowFields * r;
            X[r] = *reinterpret_cast<const float4*>(&rows[rb + fa[k]][sb]);
            Y[r] = *reinterpret_cast<const float4*>(&rows[rb + fb[k]][sb]);
            Z[r] = *reinterpret_cast<const float4*>(&rows[rb + fz[k]][sb]);
            Gd[r] = *reinterpret_cast<const float4*>(&rows[rb + fg[k]][sb]);
          }
          float acc = accv[k];
          const bool no_g = kSplit && k < kHalf;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const float x = u == 0 ? X[r].x : u == 1 ? X[r].y : u == 2 ? X[r].z : X[r].w;
              const float y = u == 0 ? Y[r].x : u == 1 ? Y[r].y : u == 2 ? Y[r].z : Y[r].w;
              const float z = u == 0 ? Z[r].x : u == 1 ? Z[r].y : u == 2 ? Z[r].z : Z[r].w;
              const float g = u == 0 ? Gd[r].x : u == 1 ? Gd[r].y : u == 2 ? Gd[r].z : Gd[r].w;
              acc = no_g ? acc + (x * y) * z : (acc + (x * y) * z) + g;
            }
          }
          accv[k] = acc;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kEpl; ++k)
    if (own[k]) sys[ab][acs[k]] = accv[k];
  __syncthreads();
  // ---- C: solve ----
  if (lane < BPW && b_i[lane] >= 0) {
    const float* o = sys[lane];
    Acc acc;
    acc.A.m[0][0] = o[0]; acc.A.m[0][1] = o[1]; acc.A.m[0][2] = o[2];
    acc.A.m[1][0] = o[1]; acc.A.m[1][1] = o[3]; acc.A.m[1][2] = o[4];
    acc.A.m[2][0] = o[2]; acc.A.m[2][1] = o[4]; acc.A.m[2][2] = o[5];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        acc.B.m[r][c] = o[6 + 3 * r + c];
        acc.D.m[r][c] = o[15 + 3 * r + c];
      }
    acc.rl = mk3(o[24], o[25], o[26]);
    acc.ra = mk3(o[27], o[28], o[29]);
    const float* bp = b_pose[lane];
    float4 xo, qo;
    solve_pose(acc, mk3(bp[0], bp[1], bp[2]), mk4(bp[3], bp[4], bp[5], bp[6]), xo, qo);
    if (kFlow) {
      st_agent(pose + 2 * b_i[lane], xo);
      st_agent(pose + 2 * b_i[lane] + 1, qo);
      b_pose[lane][0] = xo.x; b_pose[lane][1] = xo.y; b_pose[lane][2] = xo.z;
      b_pose[lane][3] = qo.x; b_pose[lane][4] = qo.y; b_pose[lane][5] = qo.z; b_pose[lane][6] = qo.w;
    } else {
      pose[2 * b_i[lane]] = xo;
      pose[2 * b_i[lane] + 1] = qo;
    }
  }
  if (kFlow && lane < BPW && b_id[lane] >= 0) {
    // publish after the pose stores have completed (static bodies publish
    // an unchanged pose)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(fl.done + b_id[lane], stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (kFlow && fl.fuse_dual) {
    __syncthreads();  // final poses of the tile's bodies in b_pose
    for (int p0 = 0; p0 < total; p0 += 64) {
      const int j = p0 + lane;
      if (j >= total) continue;
      const int2 e = csr[e0 + j];
      if (e.x < 0) continue;
      int b = 0;
      while (b + 1 < BPW && j >= b_off[b + 1]) ++b;
      const int self = b_id[b];
      if (self < 0) continue;
      const bool owner = e.y == self || fl.color[e.y] < my_color;
      if (!owner) continue;
      const GroupGeo G = geo[e.x];
      const GroupSt S = fl.gst_w[e.x];
      const float4 xs = make_float4(b_pose[b][0], b_pose[b][1], b_pose[b][2], 0.0f);
      const float4 qs = make_float4(b_pose[b][3], b_pose[b][4], b_pose[b][5], b_pose[b][6]);
      if (e.y == self) {
        dual_group(G, S, fl.gst_w + e.x, xs, qs, xs, qs, fl.beta);
      } else {
        const float4 xo = ld_agent(pose + 2 * e.y), qo = ld_agent(pose + 2 * e.y + 1);
        if (G.hdr.y == self)
          dual_group(G, S, fl.gst_w + e.x, xs, qs, xo, qo, fl.beta);
        else
          dual_group(G, S, fl.gst_w + e.x, xo, qo, xs, qs, fl.beta);
      }
    }
  }
}

// One dataflow launch for a whole sweep (all colors); cstart holds the
// n_colors + 1 color starts; sweep is the sweep's index in its substep
// (below kSweepStride - 1).
void launch_primal_sweep(hipStream_t st, int n_colors, const int* cstart, const int* color_bodies,
                         const int* color, float4* pose, const float4* poseI, const float4* inert,
                         const float4* bstat, const int* cpos, const int2* csr,
                         const GroupGeo* geo, const GroupSt* gst, float dt, int post_pass,
                         int* done, const unsigned* epoch, int sweep, int* err, int bpw,
                         int fuse_dual, float beta, GroupSt* gst_w, int tail_groups,
                         const int* done_rd, int spin_cap, const int* meta_ncolors,
                         const int* meta_cstart, const int* meta_tiles, int max_tiles,
                         const int* tail_nd, bool same_color_order) {
  SweepFlow fl{};
  fl.tail_nd = tail_nd;
  fl.meta_ncolors = meta_ncolors;
  fl.meta_cstart = meta_cstart;
  fl.meta_tiles = meta_tiles;
  fl.done_rd = done_rd ? done_rd : done;
  fl.spin_cap = spin_cap > 0 ? spin_cap : kFlowSpinCap;
  fl.fuse_dual = fuse_dual;
  fl.tail_groups = tail_groups;
  fl.beta = beta;
  fl.gst_w = gst_w;
  fl.n_colors = n_colors;
  int t = 0;
  for (int c = 0; c < n_colors; ++c) {
    fl.tile_start[c] = t;
    fl.color_start[c] = cstart[c];
    t += (cstart[c + 1] - cstart[c] + bpw - 1) / bpw;
  }
  fl.tile_start[n_colors] = t;
  fl.color_start[n_colors] = cstart[n_colors];
  fl.n_tiles = t;
  fl.color = color;
  fl.done = done;
  fl.epoch = epoch;
  fl.sweep = sweep;
  fl.err = err;
  if (meta_ncolors != nullptr) {
    // the layout is on the device: launch the most tiles any coloring has
    t = max_tiles;
    fl.n_tiles = t;
  }
  if (t <= 0) return;
  const int grid = t + (tail_groups + 63) / 64;
  if (same_color_order && bpw == 1)
    hipLaunchKernelGGL((k_primal_tiled<1, true, true>), dim3(grid), dim3(64), 0, st, 0,
                       color_bodies, 0, pose, poseI, inert, bstat, cpos, csr, geo, gst, dt,
                       post_pass, fl);
  else if (bpw == 1)
    hipLaunchKernelGGL((k_primal_tiled<1, true>), dim3(grid), dim3(64), 0, st, 0, color_bodies, 0,
                       pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, fl);
  else if (bpw == 4)
    hipLaunchKernelGGL((k_primal_tiled<4, true>), dim3(grid), dim3(64), 0, st, 0, color_bodies, 0,
                       pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, fl);
  else if (bpw == 8)
    hipLaunchKernelGGL((k_primal_tiled<8, true>), dim3(grid), dim3(64), 0, st, 0, color_bodies, 0,
                       pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, fl);
  else if (bpw == 3)
    hipLaunchKernelGGL((k_primal_tiled<3, true>), dim3(grid), dim3(64), 0, st, 0, color_bodies, 0,
                       pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, fl);
  else
    hipLaunchKernelGGL((k_primal_tiled<2, true>), dim3(grid), dim3(64), 0, st, 0, color_bodies, 0,
                       pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, fl);
}

void launch_primal_color(hipStream_t st, int count, const int* color_bodies, int base,
                         float4* pose, const float4* poseI, const float4* inert,
                         const float4* bstat, const int* cpos, const int2* csr,
                         const int* pos_map, const GroupGeo* geo, const GroupSt* gst, float dt,
                         int post_pass, int variant) {
  if (count <= 0) return;
  // The tiled kernels read a tile's CSR entries as one contiguous range, so
  // a CSR laid out for an earlier color order (pos_map) goes lane per body.
  if (variant == 0 || pos_map) {
    hipLaunchKernelGGL(k_primal_color, dim3((count + 63) / 64), dim3(64), 0, st, count,
                       color_bodies, base, pose, poseI, inert, bstat, cpos, csr, pos_map, geo, gst,
                       dt, post_pass);
  } else {
    // Bodies per wavefront: 1, 2, 4, 8 for variants 1-4; variant 5 takes one
    // body per wavefront for colors of fewer than 4096 bodies (the launch
    // time of a small color is the critical path of one wavefront, which is
    // shorter with one body) and two otherwise.
    const int bpw = variant == 1 ? 1 : variant == 2 ? 2 : variant == 3 ? 4 : variant == 4 ? 8
                                     : (count < 4096 ? 1 : 2);
    const dim3 grid((count + bpw - 1) / bpw), blk(64);
    SweepFlow none{};
    if (bpw == 1)
      hipLaunchKernelGGL((k_primal_tiled<1, false>), grid, blk, 0, st, count, color_bodies, base,
                         pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, none);
    else if (bpw == 2)
      hipLaunchKernelGGL((k_primal_tiled<2, false>), grid, blk, 0, st, count, color_bodies, base,
                         pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, none);
    else if (bpw == 4)
      hipLaunchKernelGGL((k_primal_tiled<4, false>), grid, blk, 0, st, count, color_bodies, base,
                         pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, none);
    else
      hipLaunchKernelGGL((k_primal_tiled<8, false>), grid, blk, 0, st, count, color_bodies, base,
                         pose, poseI, inert, bstat, cpos, csr, geo, gst, dt, post_pass, none);
  }
}

__global__ void __launch_bounds__(256)
k_dual(int n_groups, const float4* __restrict__ pose, const GroupGeo* __restrict__ geo,
       GroupSt* __restrict__ gst, float beta, float* resid, const int* __restrict__ nd) {
  const int n = dev_count(n_groups, nd);
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
    const GroupGeo G = geo[g];
    const int kind = g_kind(G.hdr.x);
    const int a = G.hdr.y, b = G.hdr.z;
    const float4 xa = pose[2 * a], qa = pose[2 * a + 1];
    float4 xb = xa, qb = qa;
    if (kind == kGBox) {
      xb = pose[2 * b];
      qb = pose[2 * b + 1];
    }
    dual_group(G, gst[g], gst + g, xa, qa, xb, qb, beta, resid);
  }
}

void launch_dual(hipStream_t st, int n_groups, const float4* pose, const GroupGeo* geo,
                 GroupSt* gst, float beta, float* resid, const int* nd) {
  if (n_groups <= 0) return;
  int blocks = (n_groups + 255) / 256;
  if (nd) blocks = std::min(blocks, kCapBlocks);
  hipLaunchKernelGGL(k_dual, dim3(blocks), dim3(256), 0, st, n_groups, pose, geo, gst, beta, resid,
                     nd);
}

// ---------------------------------------------------------------------------
// finalize_and_cap
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_finalize(int nb, const float4* __restrict__ pose, const float4* __restrict__ pose0,
           float4* __restrict__ vel, float4* __restrict__ pvel, const float4* __restrict__ bstat,
           float dt, float max_lin, float max_ang) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  float4 v4 = vel[2 * i], w4 = vel[2 * i + 1];
  pvel[2 * i] = v4;
  pvel[2 * i + 1] = w4;
  if (bstat[i].x > 0.0f) {
    v3 nv = divs(sub(f3(pose[2 * i]), f3(pose0[2 * i])), dt);
    q4 dq = qmul(f4q(pose[2 * i + 1]), qconj(f4q(pose0[2 * i + 1])));
    v3 nw = divs(quat_to_rotvec(dq), dt);
    float sl = length(nv);
    if (sl > max_lin) nv = mul(nv, max_lin / sl);
    float sa = length(nw);
    if (sa > max_ang) nw = mul(nw, max_ang / sa);
    vel[2 * i] = p4(nv);
    vel[2 * i + 1] = p4(nw);
  }
}

void launch_finalize(hipStream_t st, int nb, const float4* pose, const float4* pose0,
                     float4* vel, float4* pvel, const float4* bstat, float dt, float max_lin,
                     float max_ang) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_finalize, dim3((nb + 255) / 256), dim3(256), 0, st, nb, pose, pose0, vel,
                     pvel, bstat, dt, max_lin, max_ang);
}

}  // namespace avbd
