// avbd_fp.h — FP32 scalar / vec3 / quat / mat33 primitives with a fixed
// evaluation order, compiled identically for the CDNA4 kernels (hipcc) and
// for the host-side checker (g++). Both builds use -ffp-contract=off and no
// fast-math, so every helper below rounds the same way on either side.
//
// The reference obtains these operations from warp-lang 1.13.0 intrinsics
// (wp.quat_rotate, wp.quat_to_matrix, wp.quat_inverse, wp.mul, wp.inverse,
// wp.normalize, wp.sin/cos/atan2; used at src/avbd3d/kernels_6dof.py:72-73,
// 96, 155, 174, 397, 431, 571-582). Warp's own rounding is not reproduced;
// the formulas are the textbook ones with the operation order spelled out.
//
// Transcendentals: sin/cos use Cody & Waite (1980) additive range reduction
// by pi/2 with a three-part constant, followed by the Taylor series of
// sin/cos on |r| <= pi/4 (truncation error < 1e-11, far below FP32 ulp).
// atan uses two argument halvings atan(t) = 2 atan(t / (1 + sqrt(1 + t^2)))
// followed by the Taylor series on |u| <= tan(pi/16).
#pragma once

#if defined(__HIPCC__)
#define AVBD_HD __host__ __device__ __forceinline__
#else
#define AVBD_HD inline
#endif

namespace avbd {

constexpr float kInf = __builtin_huge_valf();
constexpr float kPenaltyMin = 1.0e6f;         // kernels_6dof.py:50
constexpr float kPenaltyMinTangent = 1.0f;    // kernels_6dof.py:51
constexpr float kPenaltyMax = 1.0e9f;         // kernels_6dof.py:52

// Constraint type codes (kernels_6dof.py:44-47).
constexpr int kFloor = 0;
constexpr int kTangent = 1;
constexpr int kPin = 2;
constexpr int kBoxBox = 3;

constexpr int kMaxColors = 32;   // kernels_6dof.py:2582
constexpr int kJpRounds = 24;    // kernels_6dof.py:2585
constexpr int kHashEmpty = -1;   // kernels_6dof.py:2576

struct v3 { float x, y, z; };
struct q4 { float x, y, z, w; };          // xyzw, identity = (0,0,0,1)
struct m33 { float m[3][3]; };            // row-major

AVBD_HD v3 mk3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
AVBD_HD q4 mk4(float x, float y, float z, float w) { q4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

AVBD_HD float fabs_(float a) { return __builtin_fabsf(a); }
AVBD_HD float fmin_(float a, float b) { return a < b ? a : b; }
AVBD_HD float fmax_(float a, float b) { return a > b ? a : b; }
// clamp(x, lo, hi): lower bound first, then upper (NaN passes through).
AVBD_HD float clampf_(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Correctly rounded square root on both targets (hipcc's default
// -fhip-fp32-correctly-rounded-divide-sqrt; IEEE sqrtss on x86-64).
AVBD_HD float sqrt_(float a) { return __builtin_sqrtf(a); }

AVBD_HD v3 add(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
AVBD_HD v3 sub(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
AVBD_HD v3 neg(v3 a) { return mk3(-a.x, -a.y, -a.z); }
AVBD_HD v3 mul(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
AVBD_HD v3 divs(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
AVBD_HD float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
AVBD_HD v3 cross(v3 a, v3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
AVBD_HD float length(v3 a) { return sqrt_(dot(a, a)); }
AVBD_HD float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---- quaternions (Hamilton product, xyzw storage) ----
AVBD_HD q4 qmul(q4 a, q4 b) {
  // vector part: a.w*b.v + b.w*a.v + a.v x b.v ; scalar: a.w*b.w - a.v.b.v
  float x = ((a.w * b.x + b.w * a.x) + a.y * b.z) - a.z * b.y;
  float y = ((a.w * b.y + b.w * a.y) + a.z * b.x) - a.x * b.z;
  float z = ((a.w * b.z + b.w * a.z) + a.x * b.y) - a.y * b.x;
  float w = ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z;
  return mk4(x, y, z, w);
}
// Conjugate; equals the inverse for unit quaternions.
AVBD_HD q4 qconj(q4 a) { return mk4(-a.x, -a.y, -a.z, a.w); }
AVBD_HD float qlength(q4 a) {
  return sqrt_(((a.x * a.x + a.y * a.y) + a.z * a.z) + a.w * a.w);
}
AVBD_HD q4 qnormalize(q4 a) {
  float l = qlength(a);
  if (l > 1.0e-20f) {
    float inv = 1.0f / l;
    return mk4(a.x * inv, a.y * inv, a.z * inv, a.w * inv);
  }
  return mk4(0.0f, 0.0f, 0.0f, 1.0f);
}
// v' = v (2w^2 - 1) + 2 u (u.v) + 2 w (u x v)
AVBD_HD v3 qrotate(q4 q, v3 v) {
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * ((q.x * v.x + q.y * v.y) + q.z * v.z);
  float w2 = q.w * 2.0f;
  float rx = (v.x * c + q.x * d) + (q.y * v.z - q.z * v.y) * w2;
  float ry = (v.y * c + q.y * d) + (q.z * v.x - q.x * v.z) * w2;
  float rz = (v.z * c + q.z * d) + (q.x * v.y - q.y * v.x) * w2;
  return mk3(rx, ry, rz);
}
AVBD_HD m33 qtomat(q4 q) {
  float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  m33 r;
  r.m[0][0] = 1.0f - 2.0f * (yy + zz);
  r.m[0][1] = 2.0f * (xy - wz);
  r.m[0][2] = 2.0f * (xz + wy);
  r.m[1][0] = 2.0f * (xy + wz);
  r.m[1][1] = 1.0f - 2.0f * (xx + zz);
  r.m[1][2] = 2.0f * (yz - wx);
  r.m[2][0] = 2.0f * (xz - wy);
  r.m[2][1] = 2.0f * (yz + wx);
  r.m[2][2] = 1.0f - 2.0f * (xx + yy);
  return r;
}

// ---- mat33 ----
AVBD_HD m33 mzero() { m33 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = 0.0f; return r; }
AVBD_HD m33 mT(const m33& a) {
  m33 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
AVBD_HD m33 mmul(const m33& a, const m33& b) {
  m33 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
  return r;
}
AVBD_HD v3 mvmul(const m33& a, v3 v) {
  return mk3((a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
             (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
             (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z);
}
AVBD_HD m33 madd(const m33& a, const m33& b) {
  m33 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
  return r;
}
AVBD_HD m33 msub(const m33& a, const m33& b) {
  m33 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j];
  return r;
}
AVBD_HD m33 mscale(const m33& a, float s) {
  m33 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s;
  return r;
}
AVBD_HD float mdet(const m33& a) {
  return (a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1])
          - a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]))
         + a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
// Adjugate / determinant; zero matrix when the determinant vanishes.
AVBD_HD m33 minverse(const m33& a) {
  float det = mdet(a);
  m33 r = mzero();
  if (det == 0.0f) return r;
  float inv = 1.0f / det;
  r.m[0][0] = (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) * inv;
  r.m[0][1] = (a.m[0][2] * a.m[2][1] - a.m[0][1] * a.m[2][2]) * inv;
  r.m[0][2] = (a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1]) * inv;
  r.m[1][0] = (a.m[1][2] * a.m[2][0] - a.m[1][0] * a.m[2][2]) * inv;
  r.m[1][1] = (a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0]) * inv;
  r.m[1][2] = (a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2]) * inv;
  r.m[2][0] = (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]) * inv;
  r.m[2][1] = (a.m[0][1] * a.m[2][0] - a.m[0][0] * a.m[2][1]) * inv;
  r.m[2][2] = (a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]) * inv;
  return r;
}

// ---- transcendentals ----
// Cody-Waite split of pi/2: C1 has 8 significant bits and C2 11, so k*C1 and
// k*C2 are exact for |k| < 2^13; C3 carries the remainder.
constexpr float kPio2C1 = 0x1.92p+0f;
constexpr float kPio2C2 = 0x1.fb4p-12f;
constexpr float kPio2C3 = 0x1.4442d2p-24f;
constexpr float kTwoOverPi = 0x1.45f306p-1f;
constexpr float kPio2 = 0x1.921fb6p+0f;
constexpr float kPi = 0x1.921fb6p+1f;

// Taylor coefficients 1/n! rounded to FP32.
constexpr float kInvF3 = 0x1.555556p-3f, kInvF5 = 0x1.111112p-7f, kInvF7 = 0x1.a01a02p-13f,
                kInvF9 = 0x1.71de3ap-19f, kInvF11 = 0x1.ae6456p-26f;
constexpr float kInvF2 = 0x1.0p-1f, kInvF4 = 0x1.555556p-5f, kInvF6 = 0x1.6c16c2p-10f,
                kInvF8 = 0x1.a01a02p-16f, kInvF10 = 0x1.27e4fcp-22f, kInvF12 = 0x1.1eed8ep-29f;

AVBD_HD float sin_poly(float r) {
  float r2 = r * r;
  float p = kInvF9 - r2 * kInvF11;
  p = kInvF7 - r2 * p;
  p = kInvF5 - r2 * p;
  p = kInvF3 - r2 * p;
  return r - (r * r2) * p;
}
AVBD_HD float cos_poly(float r) {
  float r2 = r * r;
  float p = kInvF10 - r2 * kInvF12;
  p = kInvF8 - r2 * p;
  p = kInvF6 - r2 * p;
  p = kInvF4 - r2 * p;
  p = kInvF2 - r2 * p;
  return 1.0f - r2 * p;
}
// sin and cos of x; accurate to a few ulp for |x| < 2^13 * pi/2.
AVBD_HD void sincos_(float x, float* s, float* c) {
  float t = x * kTwoOverPi;
  t = t > 8192.0f ? 8192.0f : (t < -8192.0f ? -8192.0f : t);
  int k = (int)(t + (t >= 0.0f ? 0.5f : -0.5f));
  float kf = (float)k;
  float r = ((x - kf * kPio2C1) - kf * kPio2C2) - kf * kPio2C3;
  float sr = sin_poly(r), cr = cos_poly(r);
  switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}
AVBD_HD float sin_(float x) { float s, c; sincos_(x, &s, &c); return s; }
AVBD_HD float cos_(float x) { float s, c; sincos_(x, &s, &c); return c; }

// atan(t) for t in [0, 1].
AVBD_HD float atan01(float t) {
  float u = t / (1.0f + sqrt_(1.0f + t * t));
  u = u / (1.0f + sqrt_(1.0f + u * u));
  float u2 = u * u;
  // u - u^3/3 + u^5/5 - ... + u^13/13 (Horner)
  float p = 1.0f / 11.0f - u2 * (1.0f / 13.0f);
  p = 1.0f / 9.0f - u2 * p;
  p = 1.0f / 7.0f - u2 * p;
  p = 1.0f / 5.0f - u2 * p;
  p = 1.0f / 3.0f - u2 * p;
  float a = u - (u * u2) * p;
  return 4.0f * a;
}
AVBD_HD float atan2_(float y, float x) {
  float ax = fabs_(x), ay = fabs_(y);
  if (ax == 0.0f && ay == 0.0f) return 0.0f;
  float a;
  if (ay <= ax) a = atan01(ay / ax);
  else a = kPio2 - atan01(ax / ay);
  if (x < 0.0f) a = kPi - a;
  return y < 0.0f ? -a : a;
}

// ---- rotation-vector <-> quaternion (kernels_6dof.py:63-97) ----
// exp map: theta < 1e-9 -> (rv/2, 1); else (sin(theta/2)/theta * rv, cos(theta/2)).
// The result is not renormalised.
AVBD_HD q4 quat_from_rotvec(v3 rv) {
  float theta = length(rv);
  if (theta < 1.0e-9f) return mk4(0.5f * rv.x, 0.5f * rv.y, 0.5f * rv.z, 1.0f);
  float half = 0.5f * theta;
  float s, c;
  sincos_(half, &s, &c);
  s = s / theta;
  return mk4(s * rv.x, s * rv.y, s * rv.z, c);
}
// log map with the w >= 0 hemisphere; |v| < 1e-9 -> 0.
AVBD_HD v3 quat_to_rotvec(q4 q) {
  float qw = q.w, qx = q.x, qy = q.y, qz = q.z;
  if (qw < 0.0f) { qw = -qw; qx = -qx; qy = -qy; qz = -qz; }
  v3 qv = mk3(qx, qy, qz);
  float l = length(qv);
  if (l < 1.0e-9f) return mk3(0.0f, 0.0f, 0.0f);
  float angle = 2.0f * atan2_(l, qw);
  return mul(qv, angle / l);
}

// ---- constraint-row helpers shared by every consumer ----
// Column-norm diagonal of the rotational Hessian of C = n.(x + R off)
// (kernels_6dof.py:280-316).
AVBD_HD v3 geom_stiffness_diag(v3 n, v3 r) {
  float nr = dot(n, r);
  float h00 = n.x * r.x - nr;
  float h10 = 0.5f * (n.y * r.x + r.y * n.x);
  float h20 = 0.5f * (n.z * r.x + r.z * n.x);
  float g0 = sqrt_((h00 * h00 + h10 * h10) + h20 * h20);
  float h01 = 0.5f * (n.x * r.y + r.x * n.y);
  float h11 = n.y * r.y - nr;
  float h21 = 0.5f * (n.z * r.y + r.z * n.y);
  float g1 = sqrt_((h01 * h01 + h11 * h11) + h21 * h21);
  float h02 = 0.5f * (n.x * r.z + r.x * n.z);
  float h12 = 0.5f * (n.y * r.z + r.y * n.z);
  float h22 = n.z * r.z - nr;
  float g2 = sqrt_((h02 * h02 + h12 * h12) + h22 * h22);
  return mk3(g0, g1, g2);
}

// Duff et al. 2017 branch-free orthonormal basis (kernels_6dof.py:2208-2218).
AVBD_HD v3 orthonormal_basis(v3 n) {
  float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  return mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
}

// Pair-hash key and slot (kernels_6dof.py:2588-2603); int32 arithmetic wraps,
// so it is carried out in uint32.
AVBD_HD int encode_pair_key(int a, int b, int c) {
  unsigned int k = ((unsigned int)a * 16384u + (unsigned int)b) * 4u + (unsigned int)c;
  return (int)k;
}
AVBD_HD int hash_pair_slot(int key, int cap) {
  unsigned int h = (unsigned int)key * 73856093u;
  int hs = (int)h;
  hs = hs ^ (hs >> 16);          // arithmetic shift, as on int32
  hs = hs & 0x7fffffff;
  return hs % cap;
}

}  // namespace avbd
