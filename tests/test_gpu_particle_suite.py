"""The reference's particle Solver tests on the HIP backend."""
import pytest

from avbd_amd.solver import Solver
from tests.particle_suite import *  # noqa: F401,F403

pytestmark = pytest.mark.gpu


@pytest.fixture
def S():
    return Solver
