"""Scene-building handles of the particle solver (reference
src/avbd3d/scene.py:12-33): a body or constraint is an index into the
solver's arrays; ``Shape`` is the body's collision/visual primitive."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal


@dataclass
class Shape:
    """sphere: size = (radius,); cube: half-extents (hx, hy, hz);
    pillar: (radius, half-height)."""
    kind: Literal["sphere", "cube", "pillar"] = "sphere"
    size: tuple[float, ...] = (0.1,)
    color: tuple[float, float, float] = (0.4, 0.6, 0.9)


@dataclass
class Body:
    """A particle: its index in the body arrays."""
    index: int
    shape: Shape = field(default_factory=Shape)


@dataclass
class ConstraintHandle:
    """First row of a constraint and its row count (for a tet: its index in
    the tet pool)."""
    index: int
    rows: int
