"""Side-by-side stepping of the HIP library and the CPU checker."""
from __future__ import annotations

import numpy as np

from avbd_amd.solver_6dof import Solver6DOF
from oracle.oracle import oracle_solver_class


def pair(builder, **kw):
    """The same seeded scene built twice: HIP solver and checker."""
    return builder(solver_cls=Solver6DOF, **kw), builder(solver_cls=oracle_solver_class(), **kw)


def one_substep(s) -> None:
    """One _step_one with the substep dt that step() would use."""
    full = s.dt
    s.dt = full / s.substeps if s.substeps > 1 else full
    try:
        s._step_one()
    finally:
        s.dt = full


def live_rows(s) -> int:
    n, _, _ = s._h.read_counters()
    return min(n, s._gpu_pool_n_capacity)


def diff_report(g, o) -> dict:
    """Element-wise comparison of everything observable after a substep."""
    out = {}
    nb = len(g._x)
    bg, bo = g._h.read_bodies(nb), o._h.read_bodies(nb)
    for k in ("x", "q", "v", "w", "prev_v", "prev_w"):
        d = np.abs(bg[k].astype(np.float64) - bo[k].astype(np.float64))
        out["body_" + k] = float(d.max()) if d.size else 0.0
        out["body_" + k + "_exact"] = bool(np.array_equal(bg[k], bo[k]))
    cg, co = g._h.read_counters(), o._h.read_counters()
    out["counters_equal"] = cg == co
    out["counters"] = (cg, co)
    n = live_rows(g)
    rg, ro = g._h.read_rows(n), o._h.read_rows(n)
    for k in ("active", "was_static", "type", "body_a", "body_b"):
        out["rows_" + k + "_equal"] = bool(np.array_equal(rg[k], ro[k]))
    for k in ("lambda", "penalty"):
        a, b = rg[k].astype(np.float64), ro[k].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        rel[(a == b)] = 0.0
        out["rows_" + k + "_maxrel"] = float(rel.max()) if rel.size else 0.0
        out["rows_" + k + "_exact"] = bool(np.array_equal(rg[k], ro[k]))
    out["colors_equal"] = bool(np.array_equal(g._h.read_colors(nb), o._h.read_colors(nb)))
    out["n_active_colors"] = (g._n_active_colors, o._n_active_colors)
    if g._self_collide and nb >= 2:
        pg, po = g._h.read_pairs(), o._h.read_pairs()
        out["pairs_equal"] = (pg["count"] == po["count"] and np.array_equal(pg["a"], po["a"])
                              and np.array_equal(pg["b"], po["b"]))
        out["sat_equal"] = bool(np.array_equal(pg["overlap"], po["overlap"])
                                and np.array_equal(pg["sat_idx"], po["sat_idx"]))
        out["sat_n_exact"] = bool(np.array_equal(pg["n_hat"], po["n_hat"])
                                  and np.array_equal(pg["depth"], po["depth"]))
        mg, mo = g._h.read_manifold(), o._h.read_manifold()
        out["contact_counts_equal"] = bool(np.array_equal(mg["count"], mo["count"]))
        out["manifold_exact"] = all(np.array_equal(mg[k], mo[k]) for k in mg)
    return out


def assert_bit_exact(rep: dict, where: str = "") -> None:
    bad = [k for k, v in rep.items() if k.endswith(("_exact", "_equal")) and v is False]
    assert not bad, f"{where}: mismatch in {bad}: {rep}"


def snapshot(g) -> dict:
    """Everything that carries over from one substep to the next."""
    nb = len(g._x)
    keys, hstate = g._h.export_hash()
    return {
        "bodies": g._h.read_bodies(nb),
        "rows": g._h.read_rows(g._gpu_pool_n_static),
        "keys": keys,
        "hstate": hstate,
        "colors": g._h.read_colors(nb),
        "n_active_colors": g._n_active_colors,
        "color_dirty": g._color_dirty,
        "caps": g._h.read_capacities(),
        "bp_max_pairs": g._bp_max_pairs,
    }


def load_snapshot(o, snap: dict) -> None:
    """Put a flushed checker solver into the snapshot's state."""
    o._flush()
    for field, key in enumerate(("x", "q", "v", "w", "prev_v", "prev_w")):
        o._h.write_bodies(field, 0, snap["bodies"][key])
    r = snap["rows"]
    o._h.write_static_rows(r["lambda"], r["penalty"], r["active"], r["was_static"])
    max_pairs, row_cap, pool_max, _ = snap["caps"]
    o._h.set_capacities(max_pairs, row_cap, pool_max)
    o._gpu_pool_n_capacity = row_cap
    o._h.import_hash(snap["keys"], snap["hstate"])
    o._h.write_colors(snap["colors"], snap["n_active_colors"])
    o._color_dirty = snap["color_dirty"]
    o._n_active_colors = snap["n_active_colors"]
    o.num_active_colors = snap["n_active_colors"]
    o._bp_max_pairs = snap["bp_max_pairs"]


def within_tolerance(rep: dict, rel: float = 1e-5) -> list[str]:
    """Float observables beyond `rel` (relative to the checker) after a substep."""
    bad = []
    for k in ("rows_lambda_maxrel",):
        if rep.get(k, 0.0) > rel:
            bad.append(k)
    return bad
