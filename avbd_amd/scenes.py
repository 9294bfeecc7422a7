"""Seeded scene generators for the benchmark workloads.

Each builder returns a configured ``Solver6DOF`` (not yet flushed). Random
draws come from ``numpy.random.default_rng(seed)`` in the order documented on
each function, so a seed fixes the scene bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from .solver_6dof import Solver6DOF


def random_orientation(rng: np.random.Generator) -> tuple[float, float, float, float]:
    """Uniform unit quaternion (xyzw) by Marsaglia's method, as the
    reference's viewer draws it (examples/viewer.py:50-65)."""
    while True:
        s1 = 2.0 * rng.random() - 1.0
        s2 = 2.0 * rng.random() - 1.0
        d1 = s1 * s1 + s2 * s2
        if d1 < 1.0:
            break
    while True:
        s3 = 2.0 * rng.random() - 1.0
        s4 = 2.0 * rng.random() - 1.0
        d2 = s3 * s3 + s4 * s4
        if d2 < 1.0:
            break
    s = math.sqrt((1.0 - d1) / d2)
    return (float(s1), float(s2), float(s3 * s), float(s4 * s))


def cube_stack(n: int = 10, seed: int = 0, iterations: int = 10, substeps: int = 4,
               solver_cls=Solver6DOF, **kw) -> Solver6DOF:
    """Vertical stack of `n` unit cubes (half-extent 0.5, 1000 kg) on the floor.
    Per cube k: x/z jitter U(-1 mm, 1 mm) (2 draws), then yaw U(-5, 5) deg."""
    rng = np.random.default_rng(seed)
    s = solver_cls(dt=1.0 / 60.0, iterations=iterations, substeps=substeps,
                   friction_static_mult=1.2, **kw)
    s.enable_self_collision(True, default_friction=0.5)
    for k in range(n):
        j = rng.uniform(-1e-3, 1e-3, size=2)
        yaw = math.radians(float(rng.uniform(-5.0, 5.0)))
        q = (0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2))
        b = s.add_box((float(j[0]), 0.5 + k, float(j[1])), (0.5, 0.5, 0.5), mass=1000.0,
                      orientation=q, friction=0.5)
        s.add_floor_contact_box(b, friction=0.5)
    return s


def random_pile(n: int = 10_000, seed: int = 0, iterations: int = 10, substeps: int = 4,
                lattice=(22, 22, 21), solver_cls=Solver6DOF, **kw) -> Solver6DOF:
    """Pile of `n` boxes dropped from a lattice: slots in (y, z, x) loop order,
    1 m spacing, centred in x/z, y = 1 + iy. Per box: jitter U(-5 cm, 5 cm)
    (3 draws), half-extents U(0.1, 0.4) (3 draws), Marsaglia orientation.
    Density 1000 kg/m^3, mu_d = 0.4, mu_s = 0.5."""
    nx, nz, ny = lattice
    if n > nx * ny * nz:
        raise ValueError("lattice too small for n boxes")
    rng = np.random.default_rng(seed)
    s = solver_cls(dt=1.0 / 60.0, iterations=iterations, substeps=substeps,
                   friction_static_mult=1.25, **kw)
    s.enable_self_collision(True, default_friction=0.4)
    k = 0
    for iy in range(ny):
        for iz in range(nz):
            for ix in range(nx):
                if k >= n:
                    return s
                j = rng.uniform(-0.05, 0.05, size=3)
                he = rng.uniform(0.1, 0.4, size=3)
                q = random_orientation(rng)
                pos = (ix - (nx - 1) / 2.0 + float(j[0]), 1.0 + iy + float(j[1]),
                       iz - (nz - 1) / 2.0 + float(j[2]))
                m = 1000.0 * 8.0 * float(he[0] * he[1] * he[2])
                b = s.add_box(pos, tuple(float(h) for h in he), mass=m, orientation=q,
                              friction=0.4)
                s.add_floor_contact_box(b, friction=0.4)
                k += 1
    return s


def pyramid(layers: int = 66, seed: int = 0, iterations: int = 20, substeps: int = 4,
            solver_cls=Solver6DOF, **kw) -> Solver6DOF:
    """Square pyramid of equal cubes (half-extent 0.25, 125 kg): layer L
    (0 = bottom) holds (layers-L)^2 cubes at x, z = (i - (n-1)/2) * 0.5,
    y = 0.25 + 0.5 L. Per cube: x/z jitter U(-1 mm, 1 mm) (2 draws).
    66 layers give 98,021 cubes. mu_d = 0.5, mu_s = 0.6."""
    rng = np.random.default_rng(seed)
    s = solver_cls(dt=1.0 / 60.0, iterations=iterations, substeps=substeps,
                   friction_static_mult=1.2, **kw)
    s.enable_self_collision(True, default_friction=0.5)
    h = 0.25
    for L in range(layers):
        n = layers - L
        for ix in range(n):
            for iz in range(n):
                j = rng.uniform(-1e-3, 1e-3, size=2)
                x = (ix - (n - 1) / 2.0) * 2 * h + float(j[0])
                z = (iz - (n - 1) / 2.0) * 2 * h + float(j[1])
                b = s.add_box((x, h + 2 * h * L, z), (h, h, h), mass=125.0, friction=0.5)
                s.add_floor_contact_box(b, friction=0.5)
    return s


def pyramid_body_count(layers: int) -> int:
    return layers * (layers + 1) * (2 * layers + 1) // 6
