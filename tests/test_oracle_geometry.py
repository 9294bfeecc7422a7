"""SAT and contact manifold of the checker against the reference's known
answers and a float64 brute-force 15-axis SAT."""
import math

import numpy as np
import pytest

from avbd_amd.solver_6dof import _obb_sat
from oracle.oracle import probe_pair

I4 = np.array([0, 0, 0, 1], np.float32)


def _R(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_sat_known_answer_axis_aligned_overlap():
    """test_solver_6dof.py:192-206: dy = 0.30, h = 0.2 -> A's y axis,
    overlap 0.1, n = (0, -1, 0)."""
    e = np.full(3, 0.2, np.float32)
    for res in (_obb_sat(np.zeros(3), np.eye(3), e, np.array([0, 0.3, 0]), np.eye(3), e),):
        assert res is not None
        idx, n, ov = res
        assert idx == 1 and abs(ov - 0.1) < 1e-4 and abs(n[1] + 1.0) < 1e-4
    r = probe_pair([0, 0, 0], I4, e, [0, 0.3, 0], I4, e)
    assert r["overlap"] == 1 and r["sat_idx"] == 1
    assert abs(r["depth"] - 0.1) < 1e-6 and np.allclose(r["n_hat"], [0, -1, 0])


def test_sat_known_answer_separated():
    """test_solver_6dof.py:209-217: gap 0.10 > margin -> no contact."""
    e = np.full(3, 0.2, np.float32)
    assert _obb_sat(np.zeros(3), np.eye(3), e, np.array([0, 0.5, 0]), np.eye(3), e) is None
    assert probe_pair([0, 0, 0], I4, e, [0, 0.5, 0], I4, e)["overlap"] == 0


def test_face_manifold_unit_cube_stack_four_corners():
    """A cube resting on another: 4 contacts at the top-face corners of the
    lower (reference) cube; the incident offsets sit on the upper cube's
    bottom face."""
    h = 0.5
    r = probe_pair([0, 0, 0], I4, [h, h, h], [0, 2 * h - 0.001, 0], I4, [h, h, h])
    assert r["overlap"] == 1 and r["sat_idx"] == 1 and r["count"] == 4 and r["ref_is_a"] == 1
    ref = np.round(r["off_ref"], 5)
    inc = np.round(r["off_inc"], 5)
    assert np.allclose(np.abs(r["off_ref"][:, [0, 2]]), h) and np.allclose(r["off_ref"][:, 1], h)
    assert np.allclose(np.abs(r["off_inc"][:, [0, 2]]), h) and np.allclose(r["off_inc"][:, 1], -h)
    assert len({tuple(v) for v in ref}) == 4 and len({tuple(v) for v in inc}) == 4


def test_face_manifold_reference_is_b_when_b_face_wins():
    h = 0.5
    small = [0.2, 0.2, 0.2]
    # B's face axis gives the smallest overlap when B is rotated 45 deg about y
    q = np.array([0, math.sin(math.pi / 8), 0, math.cos(math.pi / 8)], np.float32)
    r = probe_pair([0, 0, 0], q, [h, h, h], [0, h + 0.2 - 0.002, 0], I4, small)
    assert r["overlap"] == 1 and r["count"] >= 1
    assert r["sat_idx"] in (1, 4)


def test_edge_edge_contact_single_point():
    """Geometry of test_solver_6dof.py:388-422 brought into contact: an X-edge
    of A (rotated 45 deg about z) against an edge of B (45 deg about x)."""
    h = 0.25
    qa = np.array([0, 0, math.sin(math.pi / 8), math.cos(math.pi / 8)], np.float32)
    qb = np.array([math.sin(math.pi / 8), 0, 0, math.cos(math.pi / 8)], np.float32)
    d = 2 * h * math.sqrt(2) - 0.002
    r = probe_pair([0, 0, 0], qa, [h, h, h], [0, d, 0], qb, [h, h, h])
    assert r["overlap"] == 1
    assert r["sat_idx"] >= 6, r
    assert r["count"] == 1 and r["ref_is_a"] == 1
    pa = _R(qa) @ r["off_ref"][0]
    pb = np.array([0, d, 0]) + _R(qb) @ r["off_inc"][0]
    # closest points of the two edges: within the overlap along the normal
    assert np.linalg.norm(pa - pb) < 0.01
    assert abs(pa[1] - h * math.sqrt(2)) < 0.01 and abs(pb[1] - (d - h * math.sqrt(2))) < 0.01


def _sat64(cA, RA, eA, cB, RB, eB, margin):
    t = cB - cA
    cand = []
    for k in range(3):
        L = RA[:, k]
        ov = eA[k] + sum(eB[m] * abs(L @ RB[:, m]) for m in range(3)) - abs(t @ L)
        cand.append((k, ov, -margin, L, t @ L))
    for k in range(3):
        L = RB[:, k]
        ov = sum(eA[m] * abs(L @ RA[:, m]) for m in range(3)) + eB[k] - abs(t @ L)
        cand.append((3 + k, ov, -margin, L, t @ L))
    for i in range(3):
        for j in range(3):
            L = np.cross(RA[:, i], RB[:, j])
            nl = np.linalg.norm(L)
            if nl < 1e-6:
                continue
            L = L / nl
            ov = (sum(eA[m] * abs(L @ RA[:, m]) for m in range(3))
                  + sum(eB[m] * abs(L @ RB[:, m]) for m in range(3)) - abs(t @ L))
            cand.append((6 + 3 * i + j, ov, 0.0, L, t @ L))
    return cand


def test_sat_matches_float64_brute_force():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(3000):
        qa = rng.normal(size=4); qa /= np.linalg.norm(qa)
        qb = rng.normal(size=4); qb /= np.linalg.norm(qb)
        eA = rng.uniform(0.1, 0.4, 3); eB = rng.uniform(0.1, 0.4, 3)
        cA = np.zeros(3)
        cB = rng.normal(size=3); cB *= rng.uniform(0.1, 1.2) / np.linalg.norm(cB)
        qa32, qb32 = qa.astype(np.float32), qb.astype(np.float32)
        eA32, eB32, cB32 = eA.astype(np.float32), eB.astype(np.float32), cB.astype(np.float32)
        RA, RB = _R(qa32), _R(qb32)
        cand = _sat64(cA, RA, eA32.astype(np.float64), cB32.astype(np.float64), RB,
                      eB32.astype(np.float64), 0.005)
        r = probe_pair(cA, qa32, eA32, cB32, qb32, eB32)
        # skip decisions within FP32 reach of a threshold or a tie
        if any(abs(ov - thr) < 1e-4 for _, ov, thr, _, _ in cand):
            continue
        sep = any(ov < thr for _, ov, thr, _, _ in cand)
        if sep:
            assert r["overlap"] == 0
            checked += 1
            continue
        ovs = sorted(ov for _, ov, _, _, _ in cand)
        if len(ovs) > 1 and ovs[1] - ovs[0] < 1e-4:
            continue
        best = min(cand, key=lambda c: c[1])
        assert r["overlap"] == 1 and r["sat_idx"] == best[0]
        assert abs(r["depth"] - best[1]) < 1e-5
        n_ref = -best[3] if best[4] > 0 else best[3]
        assert np.allclose(r["n_hat"], n_ref, atol=1e-5)
        checked += 1
    assert checked > 2000
