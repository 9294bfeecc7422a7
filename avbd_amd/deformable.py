"""Tet meshes bound to the particle solver (reference
src/avbd3d/deformable.py:196-216, 523-659): one particle per vertex, one
distance row per tet edge, floor rows on the surface vertices, one
tet-volume row per tet.

`bind_tet_mesh_to_solver` takes any `TetMesh`; `box_tet_mesh` builds one for
a box (a grid of cells, each split into six tets around its main diagonal).
The reference's bunny loader and render skin need a downloaded mesh and a
viewer and are not part of this package.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .scene import Body


@dataclass
class TetMesh:
    """vertices (N, 3); tets (M, 4) vertex indices; edges (E, 2) unique tet
    edges with i < j; surface_tris (K, 3) faces owned by one tet;
    surface_verts (S,) boundary vertices; volumes (M,) signed rest volumes."""
    vertices: np.ndarray
    tets: np.ndarray
    edges: np.ndarray
    surface_tris: np.ndarray
    surface_verts: np.ndarray
    volumes: np.ndarray


@dataclass
class DeformableBody:
    """Handle of a bound mesh: `indices` are the solver bodies of
    `tet.vertices` (query `solver.positions()[indices]`)."""
    tet: TetMesh
    bodies: list[Body]
    indices: np.ndarray
    edge_constraints: list
    floor_constraints: list
    volume_constraints: list = field(default_factory=list)


def box_tet_mesh(n=(3, 3, 3), size=(0.3, 0.3, 0.3), center=(0.0, 0.7, 0.0)) -> TetMesh:
    """A box of n cells per axis, each cell split into the six tets of the
    monotone lattice paths from its low corner to its high corner (the
    Freudenthal / Kuhn split)."""
    nx, ny, nz = n
    h = np.asarray(size, np.float64) / np.asarray(n, np.float64)
    lo = np.asarray(center, np.float64) - 0.5 * np.asarray(size, np.float64)
    gi, gj, gk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    verts = (lo + np.stack([gi, gj, gk], -1).reshape(-1, 3) * h).astype(np.float32)
    vid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
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
    tets = np.asarray(tets, np.int32).reshape(-1, 4)
    P = verts.astype(np.float64)
    vols = np.einsum("ij,ij->i", P[tets[:, 1]] - P[tets[:, 0]],
                     np.cross(P[tets[:, 2]] - P[tets[:, 0]], P[tets[:, 3]] - P[tets[:, 0]])) / 6.0
    e = tets[:, [0, 0, 0, 1, 1, 2, 1, 2, 3, 2, 3, 3]].reshape(-1, 2, 6).transpose(0, 2, 1)
    edges = np.unique(np.sort(e.reshape(-1, 2), axis=1), axis=0).astype(np.int32)
    faces = tets[:, [1, 2, 3, 0, 3, 2, 0, 1, 3, 0, 2, 1]].reshape(-1, 3)
    key, first, count = np.unique(np.sort(faces, axis=1), axis=0, return_index=True,
                                  return_counts=True)
    surface_tris = faces[np.sort(first[count == 1])].astype(np.int32)
    on = (gi == 0) | (gi == nx) | (gj == 0) | (gj == ny) | (gk == 0) | (gk == nz)
    surface_verts = np.nonzero(on.reshape(-1))[0].astype(np.int32)
    return TetMesh(verts, tets, edges, surface_tris, surface_verts, vols)


def bind_tet_mesh_to_solver(solver, tet: TetMesh, density: float = 1000.0,
                            edge_stiffness: float = 5.0e4, edge_fracture: float = math.inf,
                            volume_stiffness: float = 1.0e4, floor_y: float | None = 0.0,
                            friction: float = 0.5, surface_collide: bool = False,
                            surface_radius_scale: float = 0.35) -> DeformableBody:
    """Particles, edge rows, floor rows and tet rows of a mesh
    (deformable.py:552-659). Vertex mass = density x 1/4 of the |volume| of
    its tets (at least 1e-6); with `surface_collide` the surface vertices
    collide as spheres of radius surface_radius_scale x mean edge length."""
    masses = np.zeros(len(tet.vertices), dtype=np.float32)
    abs_vol = np.abs(tet.volumes)
    for k in range(4):
        np.add.at(masses, tet.tets[:, k], abs_vol)
    masses *= 0.25 * density
    masses = np.maximum(masses, 1e-6)
    mean_edge = 0.0
    if len(tet.edges):
        mean_edge = float(np.linalg.norm(tet.vertices[tet.edges[:, 0]]
                                         - tet.vertices[tet.edges[:, 1]], axis=1).mean())
    surface = set(int(v) for v in tet.surface_verts)
    r_surf = surface_radius_scale * mean_edge if mean_edge > 0 else 0.0
    bodies = []
    for i, p in enumerate(tet.vertices):
        on = i in surface
        bodies.append(solver.add_particle(
            position=tuple(float(v) for v in p), mass=float(masses[i]),
            collide=bool(surface_collide and on and r_surf > 0),
            radius=(r_surf if (surface_collide and on) else None), friction=friction))
    edges = [solver.add_distance(bodies[int(a)], bodies[int(b)],
                                 rest=float(np.linalg.norm(tet.vertices[int(a)]
                                                           - tet.vertices[int(b)])),
                                 stiffness=edge_stiffness, fracture=edge_fracture)
             for a, b in tet.edges]
    floors = [] if floor_y is None else [
        solver.add_floor_contact(bodies[int(v)], floor_y=float(floor_y), friction=friction)
        for v in tet.surface_verts]
    vols = [] if volume_stiffness <= 0.0 else [
        solver.add_tet_volume(*(bodies[int(v)] for v in t), stiffness=float(volume_stiffness))
        for t in tet.tets]
    return DeformableBody(tet=tet, bodies=bodies,
                          indices=np.array([b.index for b in bodies], np.int32),
                          edge_constraints=edges, floor_constraints=floors,
                          volume_constraints=vols)
