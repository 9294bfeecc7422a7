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
from dataclasses import dataclass

import numpy as np

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


@dataclass
class TetLattice:
    vertices: np.ndarray      # (V, 3) float32
    tets: np.ndarray          # (T, 4) int32
    volumes: np.ndarray       # (T,) signed rest volumes
    edges: np.ndarray         # (E, 2) unique, a < b
    surface_verts: np.ndarray


def box_lattice(n: tuple[int, int, int], size: tuple[float, float, float],
                center: tuple[float, float, float]) -> TetLattice:
    nx, ny, nz = n
    h = np.asarray(size, np.float64) / np.asarray(n, np.float64)
    lo = np.asarray(center, np.float64) - 0.5 * np.asarray(size, np.float64)
    vid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    verts = (lo + np.stack([gi, gj, gk], -1).reshape(-1, 3) * h).astype(np.float32)
    # the six monotone lattice paths from corner (0,0,0) to (1,1,1)
    paths = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for p in paths:
                    c = [0, 0, 0]
                    q = [vid(i, j, k)]
                    for ax in p:
                        c[ax] += 1
                        q.append(vid(i + c[0], j + c[1], k + c[2]))
                    tets.append(q)
    tets = np.asarray(tets, np.int32)
    P = verts.astype(np.float64)
    vols = np.einsum("ij,ij->i", P[tets[:, 1]] - P[tets[:, 0]],
                     np.cross(P[tets[:, 2]] - P[tets[:, 0]], P[tets[:, 3]] - P[tets[:, 0]])) / 6.0
    e = tets[:, [0, 0, 0, 1, 1, 2, 1, 2, 3, 2, 3, 3]].reshape(-1, 2, 6).transpose(0, 2, 1)
    e = np.sort(e.reshape(-1, 2), axis=1)
    edges = np.unique(e, axis=0).astype(np.int32)
    on = (gi == 0) | (gi == nx) | (gj == 0) | (gj == ny) | (gk == 0) | (gk == nz)
    surface = np.nonzero(on.reshape(-1))[0].astype(np.int32)
    return TetLattice(verts, tets, vols, edges, surface)


@dataclass
class SoftBlock:
    lattice: TetLattice
    indices: np.ndarray
    edge_constraints: list
    floor_constraints: list
    volume_constraints: list


def soft_block(s, n=(3, 3, 3), size=(0.3, 0.3, 0.3), center=(0.0, 0.7, 0.0),
               density: float = 1000.0, edge_stiffness: float = 5.0e4,
               volume_stiffness: float = 1.0e4, friction: float = 0.5,
               floor_y: float | None = 0.0) -> SoftBlock:
    lat = box_lattice(n, size, center)
    masses = np.zeros(len(lat.vertices), np.float32)
    for k in range(4):
        np.add.at(masses, lat.tets[:, k], np.abs(lat.volumes).astype(np.float32))
    masses = np.maximum(masses * np.float32(0.25 * density), np.float32(1e-6))
    bodies = [s.add_particle(tuple(float(c) for c in p), float(m), collide=False,
                             friction=friction) for p, m in zip(lat.vertices, masses)]
    edges = [s.add_distance(bodies[a], bodies[b],
                            rest=float(np.linalg.norm(lat.vertices[a] - lat.vertices[b])),
                            stiffness=edge_stiffness) for a, b in lat.edges.tolist()]
    floors = [] if floor_y is None else [
        s.add_floor_contact(bodies[v], floor_y=float(floor_y), friction=friction)
        for v in lat.surface_verts.tolist()]
    vols = [] if volume_stiffness <= 0.0 else [
        s.add_tet_volume(*(bodies[v] for v in t), stiffness=float(volume_stiffness))
        for t in lat.tets.tolist()]
    return SoftBlock(lat, np.array([b.index for b in bodies], np.int32), edges, floors, vols)


def soft_blocks(s, count: int, n=(6, 6, 6), size=(0.4, 0.4, 0.4), gap: float = 0.2,
                height: float = 0.7, **kw) -> list[SoftBlock]:
    """`count` soft blocks on a square grid in x/z, `gap` apart, each
    dropped from `height` (block centre) onto the floor y = 0."""
    side = int(math.ceil(math.sqrt(count)))
    pitch = max(size[0], size[2]) + gap
    out = []
    for k in range(count):
        cx, cz = (k % side) * pitch, (k // side) * pitch
        out.append(soft_block(s, n=n, size=size, center=(cx, height, cz), **kw))
    return out
