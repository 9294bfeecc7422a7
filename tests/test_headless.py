"""The reference viewer's scene and headless driver (examples/viewer.py:91-194,
1691-1737), built on the checker so that these run without a GPU."""
import io

import numpy as np
import pytest

from avbd_amd.headless import build_scene, make_parser, run_headless
from oracle.oracle import oracle_solver_class


def _args(*argv):
    return make_parser().parse_args(list(argv))


def test_default_scene_sizes():
    s, boxes, pins = build_scene(_args(), solver_cls=oracle_solver_class())
    assert len(boxes) == 27 + 6 and pins == []
    assert len(s._rows) == 24 * 33           # floor rows on every body
    assert s.iterations == 4 and s.substeps == 8 and s.post_stabilize


def test_stress_scene_is_the_414_body_pile():
    # 8 x 8 towers of 6 cubes + 30 dominoes: the 414-body scene the
    # reference's README times (README.md:170)
    s, boxes, _ = build_scene(_args("--stress"), solver_cls=oracle_solver_class())
    assert len(boxes) == 8 * 8 * 6 + 30


def test_scene_draw_order_and_placement():
    """Cube positions follow the viewer's draws from default_rng(seed): one
    color draw per tower, then two jitter draws per cube."""
    seed = 7
    s, boxes, _ = build_scene(_args("--seed", str(seed)), solver_cls=oracle_solver_class())
    rng = np.random.default_rng(seed)
    origin = -(3 - 1) * 0.55 * 0.5
    k_body = 0
    for ix in range(3):
        for iz in range(3):
            rng.random()                                   # tower color
            for k in range(3):
                jx, jz = rng.uniform(-1e-3, 1e-3, size=2)
                want = (origin + ix * 0.55 + jx, 0.12 + 0.24 * k, origin + iz * 0.55 + jz)
                got = s._x[k_body]
                assert np.allclose(got, want, rtol=0, atol=1e-7), (ix, iz, k)
                k_body += 1
    # dominoes: 6 slabs, x spaced 0.10 around 0, z = 1.8, resting on the floor
    xs = [s._x[27 + i][0] for i in range(6)]
    assert np.allclose(xs, [-0.25, -0.15, -0.05, 0.05, 0.15, 0.25], atol=1e-7)
    assert all(abs(s._x[27 + i][1] - 0.12) < 1e-7 and abs(s._x[27 + i][2] - 1.8) < 1e-7
               for i in range(6))


def test_headless_printout_lines():
    lines = []
    res = run_headless(_args("--headless-warmup", "1", "--headless-frames", "2",
                             "--iterations", "2", "--substeps", "2"),
                       out=lines.append, solver_cls=oracle_solver_class())
    keys = [l.split(":")[0] for l in lines]
    assert keys == ["bodies", "static rows", "device", "substeps", "iterations", "post_stabilize",
                    "coloring mode", "colors (cap)", "colors (active)", "total rows (cap)",
                    "active rows", "avg step time", "p50 step time", "p95 step time",
                    "avg broadphase", "total 2 frames"]
    assert res["bodies"] == 33 and res["active_rows"] >= 792
