"""The reference's particle Solver tests on the CPU checker: pins the
restatement to the reference's own physical known answers."""
import pytest

from oracle.oracle import particle_solver_class
from tests.particle_suite import *  # noqa: F401,F403


@pytest.fixture
def S():
    return particle_solver_class()
