"""The reference's Solver6DOF test suite on the CPU checker: pins the
restatement to the reference's own physical known answers."""
import pytest

from oracle.oracle import oracle_solver_class
from tests.reference_suite import *  # noqa: F401,F403


@pytest.fixture
def S():
    return oracle_solver_class()
