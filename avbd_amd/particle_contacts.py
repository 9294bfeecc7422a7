"""Body-body contact generation of the particle solver, run on the host once
per step before the flush, as the reference does (solver.py:400-783): an
all-pairs test over the colliding bodies, skipping pairs that already share
an active row, then by shape pair

    cube - cube          one box-box row per corner of the face overlap
                         rectangle (at most 4), normal on the axis of least
                         overlap (solver.py:484-601)
    sphere/pillar - cube one sphere-box row per sub-sphere of the non-box
                         body that is within reach (solver.py:400-482)
    otherwise            one sphere row per overlapping pair of
                         sub-spheres (solver.py:714-781)

each normal row followed by two friction rows along an orthonormal basis of
its normal when mu = sqrt(mu_a mu_b) > 0. The arithmetic is numpy float32 as
in the reference, so the emitted rows are the reference's rows.

Two properties of the reference are kept as they are:
  * only sphere rows (and their friction rows) are removed at the next
    step; box-box and sphere-box rows stay, and since the pair then shares
    an active row it is not tested again: its first contact persists;
  * the "already share an active row" test indexes the previous flush's
    active flags with the row's new index after the removal.
"""
from __future__ import annotations

import math

import numpy as np

SPHERE_CONTACT = 5
CONTACT_TANGENT = 6
SPHERE_BOX_CONTACT = 7
BOX_BOX_CONTACT = 8


def collision_spheres(shape, radius: float | None = None):
    """A body's collision geometry as (offset, radius) sub-spheres in its own
    frame (solver.py:50-98): a sphere is one; a cube its inscribed sphere
    plus 8 corner spheres of radius 0.15 r at +-0.85 r (r the least
    half-extent); a pillar (r, half-height h) a stack of n = max(2,
    ceil(h / r)) inscribed spheres from -(h - r) to h - r, or one when
    h <= r."""
    if radius is not None:
        return [((0.0, 0.0, 0.0), float(radius))]
    if shape is None:
        return [((0.0, 0.0, 0.0), 0.0)]
    if shape.kind == "sphere":
        return [((0.0, 0.0, 0.0), float(shape.size[0]))]
    if shape.kind == "cube":
        r = float(min(shape.size))
        c, rho = 0.85 * r, (1.0 - 0.85) * r
        return [((0.0, 0.0, 0.0), r)] + [((sx * c, sy * c, sz * c), rho)
                                         for sx in (-1.0, 1.0) for sy in (-1.0, 1.0)
                                         for sz in (-1.0, 1.0)]
    if shape.kind == "pillar":
        r, h = float(shape.size[0]), float(shape.size[1])
        if h <= r:
            return [((0.0, 0.0, 0.0), r)]
        n = max(2, int(math.ceil(h / r)))
        return [((0.0, -(h - r) + k * 2 * (h - r) / (n - 1), 0.0), r) for k in range(n)]
    return [((0.0, 0.0, 0.0), 0.0)]


def reach(spheres) -> float:
    """Largest |offset| + radius over the sub-spheres (solver.py:101-108)."""
    return max(math.sqrt(o[0] ** 2 + o[1] ** 2 + o[2] ** 2) + r for o, r in spheres) \
        if spheres else 0.0


def orthonormal_basis(n: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Tangent pair of a unit normal by the branchless construction of Duff
    et al., "Building an Orthonormal Basis, Revisited" (JCGT 2017), in
    float32 (solver.py:36-47)."""
    n = n / (np.linalg.norm(n) + 1e-20)
    sign = 1.0 if n[2] >= 0.0 else -1.0
    a = -1.0 / (sign + n[2])
    b = n[0] * n[1] * a
    t = np.array([1.0 + sign * n[0] * n[0] * a, sign * b, -sign * n[0]], dtype=np.float32)
    bt = np.array([b, sign + n[1] * n[1] * a, -n[1]], dtype=np.float32)
    return t, bt


def _v(a) -> tuple[float, float, float]:
    return (float(a[0]), float(a[1]), float(a[2]))


class ContactEmitter:
    """Appends contact rows to the solver's row list and records, per
    emitted contact, (key, normal row, friction row t, friction row b)."""

    def __init__(self, solver, row_cls):
        self.s = solver
        self.Row = row_cls
        self.pairs: list[tuple[tuple[int, int, int, int], int, int, int]] = []

    def _emit(self, key, a, b, normal_row, t_hat, b_hat, mu) -> None:
        rows = self.s._constraints
        n_idx = len(rows)
        rows.append(normal_row)
        t_idx = b_idx = -1
        if mu > 0.0:
            t_idx = len(rows)
            rows.append(self.Row(type=CONTACT_TANGENT, body_a=a, body_b=b,
                                 world_anchor=_v(t_hat), stiffness=math.inf, sibling=n_idx,
                                 friction=mu))
            b_idx = len(rows)
            rows.append(self.Row(type=CONTACT_TANGENT, body_a=a, body_b=b,
                                 world_anchor=_v(b_hat), stiffness=math.inf, sibling=n_idx,
                                 friction=mu))
        self.pairs.append((key, n_idx, t_idx, b_idx))

    def sphere_box(self, si_body: int, box: int, subs, hext, x: np.ndarray, margin: float,
                   mu: float) -> None:
        h = np.asarray(hext, dtype=np.float32)
        p_box = x[box]
        for k, (off, r) in enumerate(subs):
            o = np.asarray(off, dtype=np.float32)
            rel = (x[si_body] + o) - p_box
            d = rel - np.clip(rel, -h, h)
            dl = float(np.linalg.norm(d))
            if dl > r + margin:
                continue
            if dl > 1e-9:
                n = d / dl
            else:  # centre inside: face of least depth
                axis = int(np.argmin(h - np.abs(rel)))
                n = np.zeros(3, dtype=np.float32)
                n[axis] = float(np.sign(rel[axis]) or 1.0)
            t_hat, b_hat = orthonormal_basis(n)
            row = self.Row(type=SPHERE_BOX_CONTACT, body_a=si_body, body_b=box,
                           world_anchor=_v(h), rest=float(r), stiffness=math.inf,
                           fmin=-math.inf, fmax=0.0, off_a=_v(o), off_b=(0.0, 0.0, 0.0))
            self._emit((si_body, box, k, 0), si_body, box, row, t_hat, b_hat, mu)

    def box_box(self, i: int, j: int, hi, hj, x: np.ndarray, margin: float, mu: float) -> None:
        h_i = np.asarray(hi, dtype=np.float32)
        h_j = np.asarray(hj, dtype=np.float32)
        p_i, p_j = x[i], x[j]
        diff = p_i - p_j
        overlap = (h_i + h_j) - np.abs(diff)
        if np.any(overlap < -margin):
            return
        axis = int(np.argmin(overlap))
        sgn = 1.0 if diff[axis] >= 0.0 else -1.0
        n = np.zeros(3, dtype=np.float32)
        n[axis] = sgn
        u, w = [k for k in range(3) if k != axis]
        lo = np.array([max(p_i[u] - h_i[u], p_j[u] - h_j[u]),
                       max(p_i[w] - h_i[w], p_j[w] - h_j[w])], dtype=np.float32)
        hi_ = np.array([min(p_i[u] + h_i[u], p_j[u] + h_j[u]),
                        min(p_i[w] + h_i[w], p_j[w] + h_j[w])], dtype=np.float32)
        if hi_[0] <= lo[0] or hi_[1] <= lo[1]:
            return
        face_i = p_i[axis] - sgn * h_i[axis]
        face_j = p_j[axis] + sgn * h_j[axis]
        t_hat, b_hat = orthonormal_basis(n)
        for corner, (cu, cw) in enumerate(((lo[0], lo[1]), (hi_[0], lo[1]), (lo[0], hi_[1]),
                                           (hi_[0], hi_[1]))):
            wi = np.zeros(3, dtype=np.float32)
            wi[axis], wi[u], wi[w] = face_i, cu, cw
            wj = np.zeros(3, dtype=np.float32)
            wj[axis], wj[u], wj[w] = face_j, cu, cw
            row = self.Row(type=BOX_BOX_CONTACT, body_a=i, body_b=j, world_anchor=_v(n),
                           rest=0.0, stiffness=math.inf, fmin=-math.inf, fmax=0.0,
                           off_a=_v(wi - p_i), off_b=_v(wj - p_j))
            self._emit((i, j, axis * 4 + corner, 0), i, j, row, t_hat, b_hat, mu)

    def spheres(self, i: int, j: int, subs_i, subs_j, x: np.ndarray, margin: float,
                mu: float) -> None:
        for si, (off_i, r_i) in enumerate(subs_i):
            oi = np.asarray(off_i, dtype=np.float32)
            for sj, (off_j, r_j) in enumerate(subs_j):
                oj = np.asarray(off_j, dtype=np.float32)
                rest = float(r_i + r_j)
                d = (x[i] + oi) - (x[j] + oj)
                d2 = float(np.dot(d, d))
                if d2 > (rest + margin) ** 2:
                    continue
                dn = float(math.sqrt(d2))
                n = d / dn if dn > 1e-9 else np.array([0.0, 1.0, 0.0], dtype=np.float32)
                t_hat, b_hat = orthonormal_basis(n)
                row = self.Row(type=SPHERE_CONTACT, body_a=i, body_b=j, rest=rest,
                               stiffness=math.inf, fmin=-math.inf, fmax=0.0, off_a=_v(oi),
                               off_b=_v(oj))
                self._emit((i, j, si, sj), i, j, row, t_hat, b_hat, mu)


def rebuild_contact_pool(s, row_cls, x: np.ndarray, act: np.ndarray | None):
    """Strip last step's sphere rows, then emit this step's contacts from the
    positions `x` (solver.py:603-783). Returns the emitted (key, normal,
    t, b) records; the pool starts at the row count after the strip."""
    rows = s._constraints
    kept, new_index = [], {}
    for k, c in enumerate(rows):
        if c.type == SPHERE_CONTACT:
            continue
        if c.type == CONTACT_TANGENT and 0 <= c.sibling < len(rows) \
                and rows[c.sibling].type == SPHERE_CONTACT:
            continue
        new_index[k] = len(kept)
        kept.append(c)
    for c in kept:
        if c.type == CONTACT_TANGENT and c.sibling >= 0:
            c.sibling = new_index.get(c.sibling, -1)
    s._constraints = kept
    s._contact_pool_start = len(kept)

    linked = set()
    for k, c in enumerate(kept):
        if c.body_b < 0:
            continue
        if act is not None and k < len(act) and act[k] == 0:
            continue
        linked.add((min(c.body_a, c.body_b), max(c.body_a, c.body_b)))

    em = ContactEmitter(s, row_cls)
    radii = np.asarray(s._bodies_radius, dtype=np.float32)
    collide = np.asarray(s._bodies_collide, dtype=bool)
    margin = s._contact_margin
    ids = np.nonzero(collide)[0]
    for a_pos, i in enumerate(ids.tolist()):
        for j in ids[a_pos + 1:].tolist():
            if (i, j) in linked:
                continue
            reach_ij = float(radii[i] + radii[j])
            if float(np.sum((x[i] - x[j]) ** 2)) > (reach_ij + margin) ** 2:
                continue
            ki, kj = s._bodies_kind[i], s._bodies_kind[j]
            mu = math.sqrt(s._bodies_friction[i] * s._bodies_friction[j])
            if ki == "cube" and kj != "cube":
                em.sphere_box(j, i, s._bodies_collision_spheres[j], s._bodies_half_extents[i], x,
                              margin, mu)
            elif kj == "cube" and ki != "cube":
                em.sphere_box(i, j, s._bodies_collision_spheres[i], s._bodies_half_extents[j], x,
                              margin, mu)
            elif ki == "cube" and kj == "cube":
                em.box_box(i, j, s._bodies_half_extents[i], s._bodies_half_extents[j], x, margin,
                           mu)
            else:
                em.spheres(i, j, s._bodies_collision_spheres[i], s._bodies_collision_spheres[j],
                           x, margin, mu)
    return em.pairs
