// particles.h — per-row and per-body arithmetic of the 3-DOF particle solver
// (reference src/avbd3d/kernels.py), shared by the HIP kernels and the CPU
// checker so that both evaluate every row with the same operations in the
// same order (FP32, no contraction). Each function cites the reference lines
// it follows; loops over rows/bodies and the launch order live in the
// callers.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

#include "avbd_fp.h"

namespace avbd {
namespace particles {

// constraint types (kernels.py:28-70)
constexpr int kRowPinX = 0, kRowPinY = 1, kRowPinZ = 2, kRowDistance = 3, kRowFloor = 4,
              kRowSphere = 5, kRowTangent = 6, kRowSphereBox = 7, kRowBoxBox = 8, kRowPlane = 10;
constexpr float kPenaltyMinTet = 10.0f;  // kernels.py:79

AVBD_HD v3 axis_basis(int axis) {  // kernels.py:255-261
  return axis == 0 ? mk3(1.0f, 0.0f, 0.0f) : (axis == 1 ? mk3(0.0f, 1.0f, 0.0f)
                                                         : mk3(0.0f, 0.0f, 1.0f));
}

// ||pa - pb|| - rest (kernels.py:264-266)
AVBD_HD float eval_distance(v3 pa, v3 pb, float rest) { return length(sub(pa, pb)) - rest; }

// d/dpa ||pa - pb|| (kernels.py:320-331)
AVBD_HD v3 distance_jac(v3 pa, v3 pb) {
  const v3 d = sub(pa, pb);
  const float n = length(d);
  return n > 1.0e-12f ? divs(d, n) : mk3(0.0f, 0.0f, 0.0f);
}

// Gap and outward normal of a sphere (centre ps, radius r) against an
// axis-aligned box (centre pb, half-extents h): closest point on the box,
// or, with the centre inside, the face of least depth, ties to x then y
// (kernels.py:269-317). The normal points from the box to the sphere.
AVBD_HD void sphere_box(v3 ps, v3 pb, v3 h, float r, float& C, v3& n) {
  const v3 rel = sub(ps, pb);
  const v3 d = mk3(rel.x - clampf_(rel.x, -h.x, h.x), rel.y - clampf_(rel.y, -h.y, h.y),
                   rel.z - clampf_(rel.z, -h.z, h.z));
  const float dl = length(d);
  if (dl > 1.0e-9f) {
    n = divs(d, dl);
    C = dl - r;
    return;
  }
  const float dx = h.x - fabs_(rel.x), dy = h.y - fabs_(rel.y), dz = h.z - fabs_(rel.z);
  if (dx <= dy && dx <= dz) {
    n = mk3(rel.x < 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f);
    C = -(dx + r);
  } else if (dy <= dz) {
    n = mk3(0.0f, rel.y < 0.0f ? -1.0f : 1.0f, 0.0f);
    C = -(dy + r);
  } else {
    n = mk3(0.0f, 0.0f, rel.z < 0.0f ? -1.0f : 1.0f);
    C = -(dz + r);
  }
}

// One row's C (no stabilisation) as the dual and the alpha C0 cache see it
// (kernels.py:638-666, 760-802): xa, xb are the row's bodies' positions (xb
// unused without a second body); oa, ob the body-local contact offsets of
// the body-body contact rows.
AVBD_HD float row_C(int t, v3 xa, v3 xb, bool has_b, v3 anchor, float rest, v3 oa, v3 ob) {
  if (t == kRowPinX || t == kRowPinY || t == kRowPinZ) return comp(xa, t) - comp(anchor, t);
  if (t == kRowDistance) return eval_distance(xa, xb, rest);
  if (t == kRowFloor) return xa.y - anchor.y;
  if (t == kRowPlane) return dot(anchor, xa) - rest;
  if (t == kRowSphere) return eval_distance(add(xa, oa), add(xb, ob), rest);
  if (t == kRowSphereBox) {
    float C;
    v3 n;
    sphere_box(add(xa, oa), xb, anchor, rest, C, n);
    return C;
  }
  if (t == kRowBoxBox) return dot(anchor, sub(add(xa, oa), add(xb, ob)));
  if (t == kRowTangent) return has_b ? dot(anchor, sub(xa, xb)) : dot(anchor, xa);
  return 0.0f;
}

// C and dC/dx_i of a row seen from body i in the primal (kernels.py:395-505);
// xo is the other body's position (two-body rows); oa, ob as in row_C.
AVBD_HD void row_CJ(int t, bool i_is_a, v3 xi, v3 xo, bool has_b, v3 anchor, float rest, v3 oa,
                    v3 ob, float& C, v3& J) {
  if (t == kRowSphere) {
    const v3 pa = add(i_is_a ? xi : xo, oa), pb = add(i_is_a ? xo : xi, ob);
    const v3 pi = i_is_a ? pa : pb, po = i_is_a ? pb : pa;
    J = distance_jac(pi, po);
    C = eval_distance(pi, po, rest);
    return;
  }
  if (t == kRowSphereBox) {
    v3 n;
    sphere_box(add(i_is_a ? xi : xo, oa), i_is_a ? xo : xi, anchor, rest, C, n);
    J = i_is_a ? n : neg(n);
    return;
  }
  if (t == kRowBoxBox) {
    const v3 ra = add(i_is_a ? xi : xo, oa), rb = add(i_is_a ? xo : xi, ob);
    C = dot(anchor, sub(ra, rb));
    J = i_is_a ? anchor : neg(anchor);
    return;
  }
  if (t == kRowPinX || t == kRowPinY || t == kRowPinZ) {
    J = axis_basis(t);
    C = comp(xi, t) - comp(anchor, t);
  } else if (t == kRowDistance) {
    J = distance_jac(xi, xo);
    C = eval_distance(xi, xo, rest);
  } else if (t == kRowFloor) {
    J = mk3(0.0f, 1.0f, 0.0f);
    C = xi.y - anchor.y;
  } else if (t == kRowPlane) {
    J = anchor;
    C = dot(anchor, xi) - rest;
  } else if (t == kRowTangent) {
    J = anchor;
    if (has_b) {
      const v3 xa = i_is_a ? xi : xo, xb = i_is_a ? xo : xi;
      if (!i_is_a) J = neg(anchor);
      C = dot(anchor, sub(xa, xb));
    } else {
      C = dot(anchor, xi);
    }
  } else {
    J = mk3(0.0f, 0.0f, 0.0f);
    C = 0.0f;
  }
}

// Warm start of one row (kernels.py:210-236).
AVBD_HD void warmstart_row(int t, float stiff, float alpha, float gamma, int post, float& lam,
                           float& pen) {
  const float k_floor = t == kRowTangent ? kPenaltyMinTangent : kPenaltyMin;
  float p = clampf_(pen * gamma, k_floor, kPenaltyMax);
  if (post == 0) lam = lam * alpha * gamma;
  if (!(stiff != stiff) && stiff < kInf) p = fmin_(p, stiff);
  pen = p;
}

// Row force f = clamp(k C~ + lambda_eff, bounds) (kernels.py:507-530).
AVBD_HD float row_force(float C, float stiff, float ac0, float lam, float pen, bool is_t,
                        float mu, float lam_sib, float fmin, float fmax) {
  if (stiff >= kInf) C = C - ac0;
  const float lam_eff = stiff >= kInf ? lam : 0.0f;
  if (is_t) {
    const float bound = mu * fabs_(lam_sib);
    return clampf_(pen * C + lam_eff, -bound, bound);
  }
  return clampf_(pen * C + lam_eff, fmin, fmax);
}

// A row's or a tet's terms of the body system: J f for rhs and
// (J_r J_c) k for lhs (kernels.py:532-541, 595-599). lhs is symmetric (both
// products of an off-diagonal pair round alike), so six entries are kept:
// 00 01 02 11 12 22.
struct Terms {
  v3 rhs;
  float l[6];
};
AVBD_HD Terms make_terms(v3 J, float f, float pen) {
  Terms t;
  t.rhs = mul(J, f);
  t.l[0] = (J.x * J.x) * pen;
  t.l[1] = (J.x * J.y) * pen;
  t.l[2] = (J.x * J.z) * pen;
  t.l[3] = (J.y * J.y) * pen;
  t.l[4] = (J.y * J.z) * pen;
  t.l[5] = (J.z * J.z) * pen;
  return t;
}
AVBD_HD void add_terms(m33& lhs, v3& rhs, const Terms& t) {
  rhs = add(rhs, t.rhs);
  lhs.m[0][0] = lhs.m[0][0] + t.l[0];
  lhs.m[0][1] = lhs.m[0][1] + t.l[1];
  lhs.m[0][2] = lhs.m[0][2] + t.l[2];
  lhs.m[1][0] = lhs.m[1][0] + t.l[1];
  lhs.m[1][1] = lhs.m[1][1] + t.l[3];
  lhs.m[1][2] = lhs.m[1][2] + t.l[4];
  lhs.m[2][0] = lhs.m[2][0] + t.l[2];
  lhs.m[2][1] = lhs.m[2][1] + t.l[4];
  lhs.m[2][2] = lhs.m[2][2] + t.l[5];
}

// Row force and accumulation into (lhs, rhs) (kernels.py:507-541).
AVBD_HD void primal_add(m33& lhs, v3& rhs, v3 J, float C, float stiff, float ac0, float lam,
                        float pen, bool is_t, float mu, float lam_sib, float fmin, float fmax) {
  const float f = row_force(C, stiff, ac0, lam, pen, is_t, mu, lam_sib, fmin, fmax);
  add_terms(lhs, rhs, make_terms(J, f, pen));
}

// Signed volume ratio pieces of a tet (kernels.py:552-575).
AVBD_HD float tet_volume(v3 p0, v3 p1, v3 p2, v3 p3) {
  return dot(sub(p1, p0), cross(sub(p2, p0), sub(p3, p0))) * (1.0f / 6.0f);
}

// A tet's gradient in vertex `slot` and its force k C~ (kernels.py:552-594).
AVBD_HD void tet_jac_force(v3 p0, v3 p1, v3 p2, v3 p3, int slot, float V0, float ac0, float pen,
                           v3& J, float& f) {
  const v3 a_e = sub(p1, p0), b_e = sub(p2, p0), c_e = sub(p3, p0);
  const float inv_V0 = 1.0f / V0;
  const float V = dot(a_e, cross(b_e, c_e)) * (1.0f / 6.0f);
  const float C_t = V * inv_V0 - 1.0f - ac0;
  const float s_grad = (1.0f / 6.0f) * inv_V0;
  if (slot == 1) {
    J = mul(cross(b_e, c_e), s_grad);
  } else if (slot == 2) {
    J = mul(cross(c_e, a_e), s_grad);
  } else if (slot == 3) {
    J = mul(cross(a_e, b_e), s_grad);
  } else {
    const v3 g1 = mul(cross(b_e, c_e), s_grad), g2 = mul(cross(c_e, a_e), s_grad),
             g3 = mul(cross(a_e, b_e), s_grad);
    J = neg(add(add(g1, g2), g3));
  }
  f = pen * C_t;
}

// A tet's contribution to body `slot`'s system (kernels.py:552-599).
AVBD_HD void tet_primal_add(m33& lhs, v3& rhs, v3 p0, v3 p1, v3 p2, v3 p3, int slot, float V0,
                            float ac0, float pen) {
  v3 J;
  float f;
  tet_jac_force(p0, p1, p2, p3, slot, V0, ac0, pen, J, f);
  add_terms(lhs, rhs, make_terms(J, f, pen));
}

// Dual update of one row (kernels.py:668-701): returns false when the row
// fractured (it is then inactive with lambda = penalty = 0).
AVBD_HD bool dual_row(int t, float C, float stiff, float ac0, float& lam, float& pen, float fmin,
                      float fmax, float fracture, float mu, float lam_sib, float beta) {
  if (stiff >= kInf) C = C - ac0;
  const float lam_eff = stiff >= kInf ? lam : 0.0f;
  float lo = fmin, hi = fmax;
  if (t == kRowTangent) {
    const float bound = mu * fabs_(lam_sib);
    lo = -bound;
    hi = bound;
  }
  const float nl = clampf_(pen * C + lam_eff, lo, hi);
  lam = nl;
  if (fabs_(nl) >= fracture) {
    lam = 0.0f;
    pen = 0.0f;
    return false;
  }
  if (nl > lo && nl < hi) {
    const float upper = fmin_(kPenaltyMax, stiff);
    pen = fmin_(pen + beta * fabs_(C), upper);
  }
  return true;
}

}  // namespace particles
}  // namespace avbd
