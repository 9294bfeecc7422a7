"""avbd_amd — the AVBD Solver6DOF substep on AMD Instinct MI355X (gfx950).

Drop-in for ``avbd3d``'s 6-DOF path: ``from avbd_amd import Solver6DOF``
(or ``import avbd_amd as avbd3d``). The substep runs as hand-written HIP
kernels in libavbd_hip.so, called through a ctypes C ABI (include/avbd.h).
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

__all__ = [
    "Solver6DOF",
    "RigidBody",
    "box_inertia_local",
    "box_inv_inertia_local",
    "FLOOR_CONTACT_6DOF",
    "CONTACT_TANGENT_6DOF",
    "PIN_6DOF",
    "BOX_BOX_CONTACT_6DOF",
]
