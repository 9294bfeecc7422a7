// particles_oracle.cpp — CPU checker of the 3-DOF particle solver step.
// TEST INFRASTRUCTURE ONLY: loaded by tests/ (and the CPU baseline leg of
// bench.py), never by the avbd_amd package.
//
// A serial restatement of Solver.step without self-collision
// (src/avbd3d/solver.py:1124-1400) and its kernels (src/avbd3d/kernels.py),
// in the order the reference's default CPU device runs them: every kernel
// is a loop over its thread index, so the dual visits rows in index order and
// a friction row's cone bound reads its normal row's lambda of the same
// sweep. The primal is launched once per color over all bodies and gated on
// the body's color (kernels.py:384-386). Row arithmetic is shared with the
// HIP kernels through avbd_amd/csrc/particles.h; everything else (loops,
// gating, launch order, adjacency walk) is written here independently.
//
// Exports orc_p_* with the signatures of avbd_p_* (include/avbd.h).
#include <cstring>
#include <string>
#include <vector>

#include "../avbd_amd/csrc/particles.h"
#include "../include/avbd.h"

using namespace avbd;
using namespace avbd::particles;

namespace {

thread_local std::string g_err;
int fail(int code, const char* m) {
  g_err = m;
  return code;
}

struct POracle {
  int nb = 0, nc = 0, nt = 0, n_colors = 0;
  bool flushed = false;
  std::vector<v3> x, v, prev_v, initial, inertial, anchor, off_a, off_b;
  std::vector<float> mass, rest, stiff, fmin, fmax, fracture, friction, lam, pen, ac0;
  std::vector<int> color, type, ra, rb, active, sibling, con_start, con_idx;
  std::vector<int> tv;  // 4 per tet
  std::vector<float> tet_rest, tet_stiff, tet_fracture, tet_lam, tet_pen, tet_ac0;
  std::vector<int> tet_active, tet_start, tet_idx;
};

std::vector<v3> load3(const float* a, int n) {
  std::vector<v3> o((size_t)n);
  for (int i = 0; i < n; ++i) o[i] = mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
  return o;
}
template <class T>
std::vector<T> load(const T* a, int n) {
  return n > 0 ? std::vector<T>(a, a + n) : std::vector<T>();
}

bool is_pin(int t) { return t == kRowPinX || t == kRowPinY || t == kRowPinZ; }

// predict_inertial (kernels.py:161-207)
void predict(POracle& s, float dt, v3 g) {
  for (int i = 0; i < s.nb; ++i) {
    const float m = s.mass[i];
    s.initial[i] = s.x[i];
    if (m <= 0.0f) {
      s.inertial[i] = s.x[i];
      continue;
    }
    const v3 g_dt2 = mul(g, dt * dt);
    s.inertial[i] = add(add(s.x[i], mul(s.v[i], dt)), g_dt2);
    const v3 accel = divs(sub(s.v[i], s.prev_v[i]), dt);
    const float g_norm = length(g);
    float w = 0.0f;
    if (g_norm > 0.0f) {
      const v3 g_hat = divs(g, g_norm);
      w = clampf_(dot(accel, g_hat) / g_norm, 0.0f, 1.0f);
    }
    s.x[i] = add(add(s.x[i], mul(s.v[i], dt)), mul(g_dt2, w));
  }
}

// cache_alpha_C0 (kernels.py:740-802)
void cache_c0(POracle& s, float alpha) {
  for (int j = 0; j < s.nc; ++j) {
    if (s.active[j] == 0) {
      s.ac0[j] = 0.0f;
      continue;
    }
    const int t = s.type[j];
    const bool has_b = s.rb[j] >= 0 && !is_pin(t);
    if (t == kRowTangent) {
      const v3 xa = s.initial[s.ra[j]];
      s.ac0[j] = has_b ? dot(s.anchor[j], sub(xa, s.initial[s.rb[j]])) : dot(s.anchor[j], xa);
      continue;
    }
    const v3 xb = has_b ? s.x[s.rb[j]] : mk3(0.0f, 0.0f, 0.0f);
    s.ac0[j] = alpha * row_C(t, s.x[s.ra[j]], xb, has_b, s.anchor[j], s.rest[j], s.off_a[j],
                             s.off_b[j]);
  }
}

// tet_cache_alpha_C0 (kernels.py:838-864)
void tet_cache_c0(POracle& s, float alpha) {
  for (int t = 0; t < s.nt; ++t) {
    if (s.tet_active[t] == 0) {
      s.tet_ac0[t] = 0.0f;
      continue;
    }
    const int* q = &s.tv[4 * (size_t)t];
    const float V = tet_volume(s.x[q[0]], s.x[q[1]], s.x[q[2]], s.x[q[3]]);
    s.tet_ac0[t] = alpha * (V / s.tet_rest[t] - 1.0f);
  }
}

// primal_update for one body (kernels.py:335-608)
void primal_body(POracle& s, int i, float dt) {
  const float m = s.mass[i];
  if (m <= 0.0f) return;
  const float mdt2 = m * (1.0f / (dt * dt));
  m33 lhs = mzero();
  for (int d = 0; d < 3; ++d) lhs.m[d][d] = mdt2;
  v3 rhs = mul(sub(s.x[i], s.inertial[i]), mdt2);
  for (int k = s.con_start[i]; k < s.con_start[i + 1]; ++k) {
    const int j = s.con_idx[k];
    if (s.active[j] == 0) continue;
    const int t = s.type[j];
    const bool has_b = s.rb[j] >= 0 && !is_pin(t);
    const bool i_is_a = s.ra[j] == i;
    const v3 xo = has_b ? s.x[i_is_a ? s.rb[j] : s.ra[j]] : s.x[i];
    float C;
    v3 J;
    row_CJ(t, i_is_a, s.x[i], xo, has_b, s.anchor[j], s.rest[j], s.off_a[j], s.off_b[j], C, J);
    const bool is_t = t == kRowTangent;
    primal_add(lhs, rhs, J, C, s.stiff[j], s.ac0[j], s.lam[j], s.pen[j], is_t, s.friction[j],
               is_t ? s.lam[s.sibling[j]] : 0.0f, s.fmin[j], s.fmax[j]);
  }
  for (int k = s.tet_start[i]; k < s.tet_start[i + 1]; ++k) {
    const int enc = s.tet_idx[k];
    const int tid = enc >> 2, slot = enc & 3;
    if (s.tet_active[tid] == 0) continue;
    const int* q = &s.tv[4 * (size_t)tid];
    tet_primal_add(lhs, rhs, s.x[q[0]], s.x[q[1]], s.x[q[2]], s.x[q[3]], slot, s.tet_rest[tid],
                   s.tet_ac0[tid], s.tet_pen[tid]);
  }
  s.x[i] = sub(s.x[i], mvmul(minverse(lhs), rhs));
}

// dual_update, rows in index order (kernels.py:611-701)
void dual(POracle& s, float beta) {
  for (int j = 0; j < s.nc; ++j) {
    if (s.active[j] == 0) continue;
    const int t = s.type[j];
    const bool has_b = s.rb[j] >= 0 && !is_pin(t);
    const v3 xb = has_b ? s.x[s.rb[j]] : mk3(0.0f, 0.0f, 0.0f);
    const float C = row_C(t, s.x[s.ra[j]], xb, has_b, s.anchor[j], s.rest[j], s.off_a[j],
                          s.off_b[j]);
    const float lam_sib = t == kRowTangent ? s.lam[s.sibling[j]] : 0.0f;
    if (!dual_row(t, C, s.stiff[j], s.ac0[j], s.lam[j], s.pen[j], s.fmin[j], s.fmax[j],
                  s.fracture[j], s.friction[j], lam_sib, beta))
      s.active[j] = 0;
  }
}

// tet_dual_update (kernels.py:866-912)
void tet_dual(POracle& s, float beta) {
  for (int t = 0; t < s.nt; ++t) {
    if (s.tet_active[t] == 0) continue;
    const int* q = &s.tv[4 * (size_t)t];
    const float V = tet_volume(s.x[q[0]], s.x[q[1]], s.x[q[2]], s.x[q[3]]);
    const float C = V / s.tet_rest[t] - 1.0f - s.tet_ac0[t];
    const float nl = s.tet_pen[t] * C;
    s.tet_lam[t] = nl;
    if (fabs_(nl) >= s.tet_fracture[t]) {
      s.tet_active[t] = 0;
      s.tet_lam[t] = 0.0f;
      s.tet_pen[t] = 0.0f;
      continue;
    }
    s.tet_pen[t] = fmin_(s.tet_pen[t] + beta * fabs_(C), fmin_(kPenaltyMax, s.tet_stiff[t]));
  }
}

// finalize_velocity then cap_velocity (kernels.py:707-737)
void finalize(POracle& s, float dt, float max_speed) {
  for (int i = 0; i < s.nb; ++i) {
    s.prev_v[i] = s.v[i];
    if (s.mass[i] > 0.0f) s.v[i] = divs(sub(s.x[i], s.initial[i]), dt);
  }
  if (max_speed > 0.0f && max_speed < kInf) {
    for (int i = 0; i < s.nb; ++i) {
      const float speed = length(s.v[i]);
      if (speed > max_speed) s.v[i] = mul(s.v[i], max_speed / speed);
    }
  }
}

void copy3(const std::vector<v3>& a, float* out) {
  if (!out) return;
  for (size_t i = 0; i < a.size(); ++i) {
    out[3 * i] = a[i].x;
    out[3 * i + 1] = a[i].y;
    out[3 * i + 2] = a[i].z;
  }
}

}  // namespace

extern "C" {

const char* orc_p_last_error(void) { return g_err.c_str(); }

int orc_p_create(int /*device*/, void** out) {
  if (!out) return fail(AVBD_E_ARG, "null output");
  *out = new POracle();
  return AVBD_OK;
}

int orc_p_destroy(void* h) {
  delete static_cast<POracle*>(h);
  return AVBD_OK;
}

int orc_p_flush(void* h, const avbd_p_scene* sc) {
  if (!h || !sc) return fail(AVBD_E_ARG, "null argument");
  POracle& s = *static_cast<POracle*>(h);
  const int nb = sc->n_bodies, nc = sc->n_rows, nt = sc->n_tets;
  if (nb < 0 || nc < 0 || nt < 0) return fail(AVBD_E_ARG, "negative size");
  s.nb = nb;
  s.nc = nc;
  s.nt = nt;
  s.n_colors = sc->n_colors;
  s.x = load3(sc->x, nb);
  s.v = load3(sc->v, nb);
  s.prev_v = load3(sc->prev_v, nb);
  s.initial = s.x;
  s.inertial = s.x;
  s.mass = load(sc->mass, nb);
  s.color = load(sc->color, nb);
  s.type = load(sc->type, nc);
  s.ra = load(sc->body_a, nc);
  s.rb = load(sc->body_b, nc);
  s.anchor = load3(sc->anchor, nc);
  s.off_a = sc->off_a ? load3(sc->off_a, nc) : std::vector<v3>((size_t)nc, mk3(0.f, 0.f, 0.f));
  s.off_b = sc->off_b ? load3(sc->off_b, nc) : std::vector<v3>((size_t)nc, mk3(0.f, 0.f, 0.f));
  s.rest = load(sc->rest, nc);
  s.stiff = load(sc->stiffness, nc);
  s.fmin = load(sc->fmin, nc);
  s.fmax = load(sc->fmax, nc);
  s.fracture = load(sc->fracture, nc);
  s.sibling = load(sc->sibling, nc);
  s.friction = load(sc->friction, nc);
  s.lam = load(sc->lambda, nc);
  s.pen = load(sc->penalty, nc);
  s.active = load(sc->active, nc);
  s.ac0.assign((size_t)nc, 0.0f);
  s.con_start = load(sc->con_start, nb + 1);
  s.con_idx = load(sc->con_idx, nb > 0 ? s.con_start[nb] : 0);
  s.tv = load(sc->tet_v, 4 * nt);
  s.tet_rest = load(sc->tet_rest_vol, nt);
  s.tet_stiff = load(sc->tet_stiffness, nt);
  s.tet_fracture = load(sc->tet_fracture, nt);
  s.tet_lam = load(sc->tet_lambda, nt);
  s.tet_pen = load(sc->tet_penalty, nt);
  s.tet_active = load(sc->tet_active, nt);
  s.tet_ac0.assign((size_t)nt, 0.0f);
  s.tet_start = load(sc->tet_start, nb + 1);
  s.tet_idx = load(sc->tet_idx, nb > 0 ? s.tet_start[nb] : 0);
  s.flushed = true;
  return AVBD_OK;
}

// Solver.step (solver.py:1124-1400) without the self-collision prelude.
int orc_p_step(void* h, const avbd_p_params* p) {
  if (!h || !p) return fail(AVBD_E_ARG, "null argument");
  POracle& s = *static_cast<POracle*>(h);
  if (!s.flushed) return fail(AVBD_E_STATE, "step before flush");
  if (s.nb == 0) return AVBD_OK;
  const float dt = p->dt;
  const int post = p->post_stabilize ? 1 : 0;
  predict(s, dt, mk3(p->gravity[0], p->gravity[1], p->gravity[2]));
  if (s.nc > 0) {
    for (int j = 0; j < s.nc; ++j)  // warmstart_duals
      warmstart_row(s.type[j], s.stiff[j], p->alpha, p->gamma, post, s.lam[j], s.pen[j]);
    cache_c0(s, p->alpha);
  }
  if (s.nt > 0) {
    for (int t = 0; t < s.nt; ++t) {  // tet_warmstart_duals (kernels.py:815-836)
      float pk = clampf_(s.tet_pen[t] * p->gamma, kPenaltyMinTet, kPenaltyMax);
      if (post == 0) s.tet_lam[t] = s.tet_lam[t] * p->alpha * p->gamma;
      const float st = s.tet_stiff[t];
      if (!(st != st) && st < kInf) pk = fmin_(pk, st);
      s.tet_pen[t] = pk;
    }
    tet_cache_c0(s, p->alpha);
  }
  const int I = p->iterations;
  for (int it = 0; it < I + post; ++it) {
    if (post && it == I && s.nc > 0) cache_c0(s, 0.0f);
    if (post && it == I && s.nt > 0) tet_cache_c0(s, 0.0f);
    for (int c = 0; c < s.n_colors; ++c)
      for (int i = 0; i < s.nb; ++i)
        if (s.color[i] == c) primal_body(s, i, dt);
    if (s.nc > 0 && it < I) dual(s, p->beta);
    if (s.nt > 0 && it < I) tet_dual(s, p->beta);
    if (it == I - 1 && s.nt == 0) finalize(s, dt, p->max_speed);
  }
  if (s.nt > 0) finalize(s, dt, p->max_speed);
  return AVBD_OK;
}

int orc_p_read(void* h, float* x, float* v, float* prev_v, float* lambda, float* penalty,
               int* active, float* tet_lambda, float* tet_penalty, int* tet_active) {
  if (!h) return fail(AVBD_E_ARG, "null handle");
  POracle& s = *static_cast<POracle*>(h);
  if (!s.flushed) return fail(AVBD_E_STATE, "read before flush");
  copy3(s.x, x);
  copy3(s.v, v);
  copy3(s.prev_v, prev_v);
  if (lambda && s.nc) std::memcpy(lambda, s.lam.data(), s.nc * sizeof(float));
  if (penalty && s.nc) std::memcpy(penalty, s.pen.data(), s.nc * sizeof(float));
  if (active && s.nc) std::memcpy(active, s.active.data(), s.nc * sizeof(int));
  if (tet_lambda && s.nt) std::memcpy(tet_lambda, s.tet_lam.data(), s.nt * sizeof(float));
  if (tet_penalty && s.nt) std::memcpy(tet_penalty, s.tet_pen.data(), s.nt * sizeof(float));
  if (tet_active && s.nt) std::memcpy(tet_active, s.tet_active.data(), s.nt * sizeof(int));
  return AVBD_OK;
}

int orc_p_write_bodies(void* h, int field, int first, int n, const float* vals) {
  if (!h) return fail(AVBD_E_ARG, "null handle");
  POracle& s = *static_cast<POracle*>(h);
  if (!s.flushed) return fail(AVBD_E_STATE, "write before flush");
  if (first < 0 || n < 0 || first + n > s.nb || !vals) return fail(AVBD_E_ARG, "bad range");
  std::vector<v3>* dst = field == 0 ? &s.x : field == 1 ? &s.v : nullptr;
  if (!dst) return fail(AVBD_E_ARG, "field must be 0 (x) or 1 (v)");
  for (int k = 0; k < n; ++k) (*dst)[first + k] = mk3(vals[3 * k], vals[3 * k + 1], vals[3 * k + 2]);
  return AVBD_OK;
}

int orc_p_write_rows(void* h, int first, int n, const float* anchor, const float* rest) {
  if (!h) return fail(AVBD_E_ARG, "null handle");
  POracle& s = *static_cast<POracle*>(h);
  if (!s.flushed) return fail(AVBD_E_STATE, "write before flush");
  if (first < 0 || n < 0 || first + n > s.nc) return fail(AVBD_E_ARG, "bad range");
  for (int k = 0; k < n; ++k) {
    if (anchor) s.anchor[first + k] = mk3(anchor[3 * k], anchor[3 * k + 1], anchor[3 * k + 2]);
    if (rest) s.rest[first + k] = rest[k];
  }
  return AVBD_OK;
}

}  // extern "C"

// ---- probes of the shared row arithmetic (tests pin them against numerical
// gradients of the constraint functions) ----
extern "C" {

void orc_p_probe_row(int type, int i_is_a, const float* xi, const float* xo, int has_b,
                     const float* anchor, float rest, const float* oa, const float* ob, float* C,
                     float* J) {
  v3 j;
  float c;
  row_CJ(type, i_is_a != 0, mk3(xi[0], xi[1], xi[2]), mk3(xo[0], xo[1], xo[2]), has_b != 0,
         mk3(anchor[0], anchor[1], anchor[2]), rest, mk3(oa[0], oa[1], oa[2]),
         mk3(ob[0], ob[1], ob[2]), c, j);
  *C = c;
  J[0] = j.x;
  J[1] = j.y;
  J[2] = j.z;
}

void orc_p_probe_tet(const float* p, int slot, float V0, float ac0, float pen, float* lhs9,
                     float* rhs3) {
  m33 lhs = mzero();
  v3 rhs = mk3(0.0f, 0.0f, 0.0f);
  tet_primal_add(lhs, rhs, mk3(p[0], p[1], p[2]), mk3(p[3], p[4], p[5]), mk3(p[6], p[7], p[8]),
                 mk3(p[9], p[10], p[11]), slot, V0, ac0, pen);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) lhs9[3 * r + c] = lhs.m[r][c];
  rhs3[0] = rhs.x;
  rhs3[1] = rhs.y;
  rhs3[2] = rhs.z;
}

}  // extern "C"
