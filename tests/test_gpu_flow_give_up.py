"""The dataflow primal sweep's give-up path, forced from the host.

With AVBD_OPT_FLOW_GIVE_UP_TEST every wait of the sweep polls an array that
nobody writes and gives up after a few polls, as a wait that does not end
would at its spin cap. Every kernel still exits normally. The library must
report the error at the next sync, keep reporting it, and clear it at the
next flush.
"""
import numpy as np
import pytest

from avbd_amd import _native as N
from avbd_amd import scenes
from tests.parity import one_substep

pytestmark = pytest.mark.gpu


def test_give_up_is_reported_and_cleared_by_flush():
    s = scenes.pyramid(5, seed=3)   # several colors, so lower-color waits exist
    for _ in range(2):
        one_substep(s)
    assert np.isfinite(s.positions()).all()
    s._h.set_option(N.OPT_FLOW_GIVE_UP_TEST, 8)
    with pytest.raises(N.AvbdError, match="dataflow wait"):
        one_substep(s)
        s.positions()
    # sticky: every later read reports it too
    with pytest.raises(N.AvbdError, match="dataflow wait"):
        s._h.read_bodies(len(s._x))
    # a flush starts over: normal waits, no error
    sc, keep = s.make_scene_struct(s._flush_snapshot)
    s._h.flush(sc)
    del keep
    s._color_dirty = True
    for _ in range(2):
        one_substep(s)
    assert np.isfinite(s.positions()).all()


def test_give_up_option_is_validated():
    s = scenes.pyramid(3, seed=3)
    s._flush()
    with pytest.raises(N.AvbdError):
        s._h.set_option(N.OPT_FLOW_GIVE_UP_TEST, -1)
    with pytest.raises(N.AvbdError):
        s._h.set_option(N.OPT_FLOW_GIVE_UP_TEST, 1 << 20)
