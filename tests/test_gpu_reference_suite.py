"""The reference's Solver6DOF test suite on the HIP backend."""
import pytest

from avbd_amd import Solver6DOF
from tests.reference_suite import *  # noqa: F401,F403

pytestmark = pytest.mark.gpu


@pytest.fixture
def S():
    return Solver6DOF
