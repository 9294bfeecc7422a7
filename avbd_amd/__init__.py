"""avbd_amd — the AVBD solvers on AMD Instinct MI355X (gfx950).

Drop-in for ``avbd3d``: ``from avbd_amd import Solver6DOF, Solver`` (or
``import avbd_amd as avbd3d``). The 6-DOF substep and the 3-DOF particle
step run as hand-written HIP kernels in libavbd_hip.so, called through a
ctypes C ABI (include/avbd.h).
"""
from .solver_6dof import (  # noqa: F401
    BOX_BOX_CONTACT_6DOF,
    CONTACT_TANGENT_6DOF,
    FLOOR_CONTACT_6DOF,
    PIN_6DOF,
    RigidBody,
    Solver6DOF,
    box_inertia_local,
    box_inv_inertia_local,
)

from .deformable import DeformableBody, TetMesh, bind_tet_mesh_to_solver, box_tet_mesh  # noqa: F401
from .scene import Body, ConstraintHandle, Shape  # noqa: F401
from .solver import Solver  # noqa: F401

__all__ = [
    "Solver",
    "Body",
    "ConstraintHandle",
    "Shape",
    "TetMesh",
    "DeformableBody",
    "bind_tet_mesh_to_solver",
    "box_tet_mesh",
    "Solver6DOF",
    "RigidBody",
    "box_inertia_local",
    "box_inv_inertia_local",
    "FLOOR_CONTACT_6DOF",
    "CONTACT_TANGENT_6DOF",
    "PIN_6DOF",
    "BOX_BOX_CONTACT_6DOF",
]
