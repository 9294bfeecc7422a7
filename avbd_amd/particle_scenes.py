"""Scene builders for the particle solver (avbd_amd.solver.Solver).

    chain(s, n)            pinned chain of n distance links (tests/test_solver.py:103-132)
    cloth(s, nx, nz)       grid of particles with structural and shear links, two
                           pinned corners (soft links: with hard links a flat
                           sheet pulled perpendicular to its links diverges,
                           the primal having no geometric-stiffness term,
                           kernels.py:522-526)
    soft_block(s, n)       tet lattice of a box, bound like bind_tet_mesh_to_solver
                           (src/avbd3d/deformable.py:552-659): one particle per
                           lattice vertex, one distance row per tet edge, floor
                           rows with friction on the surface vertices, one
                           tet-volume row per tet
    soft_blocks(s, k)      k such blocks side by side on one floor

The block is a regular grid of cells, each split into six tets around the
cell's main diagonal (the Freudenthal/Kuhn split), in place of the
reference's voxelised bunny (which needs a downloaded mesh).
"""
from __future__ import annotations

import math
import numpy as np

from .deformable import DeformableBody, TetMesh, bind_tet_mesh_to_solver, box_tet_mesh
from .scene import Body


def chain(s, n: int = 5, link: float = 0.5, top=(0.0, 5.0, 0.0), fracture: float = math.inf,
          stiffness: float = math.inf):
    bodies = [s.add_particle((top[0], top[1] - i * link, top[2]), 1.0) for i in range(n + 1)]
    s.add_pin(bodies[0], top, stiffness=math.inf)
    links = [s.add_distance(bodies[i], bodies[i + 1], rest=link, stiffness=stiffness,
                            fracture=fracture) for i in range(n)]
    return bodies, links


def cloth(s, nx: int, nz: int, spacing: float = 0.05, height: float = 2.0,
          mass: float = 0.05, stiffness: float = 2.0e3, shear: bool = True):
    """nx x nz particles in the y = height plane; rows: structural links along
    x and z, shear links on both diagonals; the two corners of the first row
    pinned. Returns the bodies as an (nx, nz) index array."""
    idx = np.zeros((nx, nz), np.int64)
    for i in range(nx):
        for k in range(nz):
            b = s.add_particle((i * spacing, height, k * spacing), mass, collide=False)
            idx[i, k] = b.index
    B = lambda i, k: Body(int(idx[i, k]))
    diag = spacing * math.sqrt(2.0)
    for i in range(nx):
        for k in range(nz):
            if i + 1 < nx:
                s.add_distance(B(i, k), B(i + 1, k), spacing, stiffness=stiffness)
            if k + 1 < nz:
                s.add_distance(B(i, k), B(i, k + 1), spacing, stiffness=stiffness)
            if shear and i + 1 < nx and k + 1 < nz:
                s.add_distance(B(i, k), B(i + 1, k + 1), diag, stiffness=stiffness)
                s.add_distance(B(i + 1, k), B(i, k + 1), diag, stiffness=stiffness)
    s.add_pin(B(0, 0), (0.0, height, 0.0))
    s.add_pin(B(0, nz - 1), (0.0, height, (nz - 1) * spacing))
    return idx


def box_lattice(n, size, center) -> TetMesh:
    """The tet mesh of a box (deformable.box_tet_mesh)."""
    return box_tet_mesh(n, size, center)


def soft_block(s, n=(3, 3, 3), size=(0.3, 0.3, 0.3), center=(0.0, 0.7, 0.0),
               density: float = 1000.0, edge_stiffness: float = 5.0e4,
               volume_stiffness: float = 1.0e4, friction: float = 0.5,
               floor_y: float | None = 0.0) -> DeformableBody:
    return bind_tet_mesh_to_solver(s, box_tet_mesh(n, size, center), density=density,
                                   edge_stiffness=edge_stiffness,
                                   volume_stiffness=volume_stiffness, floor_y=floor_y,
                                   friction=friction)


def soft_blocks(s, count: int, n=(6, 6, 6), size=(0.4, 0.4, 0.4), gap: float = 0.2,
                height: float = 0.7, **kw) -> list[DeformableBody]:
    """`count` soft blocks on a square grid in x/z, `gap` apart, each
    dropped from `height` (block centre) onto the floor y = 0."""
    side = int(math.ceil(math.sqrt(count)))
    pitch = max(size[0], size[2]) + gap
    out = []
    for k in range(count):
        cx, cz = (k % side) * pitch, (k // side) * pitch
        out.append(soft_block(s, n=n, size=size, center=(cx, height, cz), **kw))
    return out
