"""The device-side boundary of the library.

* Stage entry points (avbd_stage_*, the reference's launch order at
  solver_6dof.py:1269-1464) give the bits of avbd_substep; out-of-order
  calls are refused and the substep can start again.
* Zero-copy torch views of x, q, v, omega, lambda and the row state read and
  write the solver's own device memory (solver_6dof.py:213-225 keeps them as
  device arrays).
* A caller's stream (a torch stream) orders the solver's work with torch's.
"""
import numpy as np
import pytest
import torch

from avbd_amd import _native as N
from avbd_amd import scenes
from tests.parity import assert_bit_exact, diff_report

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("build", [
    lambda **kw: scenes.pyramid(4, seed=21, iterations=6, **kw),
    lambda **kw: scenes.random_pile(90, seed=8, lattice=(5, 5, 4), iterations=5, **kw),
], ids=["pyramid", "pile"])
def test_stages_match_whole_substep(build):
    a, b = build(), build()
    for f in range(6):
        print(f"frame {f}", flush=True)
        a.step()
        b.step_staged()
        assert_bit_exact(diff_report(a, b), f"frame {f}")
    assert np.isfinite(b.positions()).all()


def test_stage_order_is_enforced():
    s = scenes.pyramid(3, seed=2, iterations=4)
    s.step()
    p = s._params(s.dt / s.substeps)
    with pytest.raises(N.AvbdError, match="out of order"):
        s._h.stage_sweep(p, 0)
    rep = s._h.stage_contacts(p)
    with pytest.raises(N.AvbdError, match="in progress"):
        s._h.substep(p, False)
    with pytest.raises(N.AvbdError, match="out of order"):
        s._h.stage_prelude(p)
    s._h.stage_coloring(p, False, rep)
    s._h.stage_csr()
    s._h.stage_prelude(p)
    with pytest.raises(N.AvbdError, match="in order"):
        s._h.stage_sweep(p, 1)
    for k in range(p.iterations + p.post_stabilize):
        s._h.stage_sweep(p, k)
    s._h.stage_end(p)
    s.step()   # whole substeps again
    assert np.isfinite(s.positions()).all()


def _flat_rows(views, n_rows):
    """The reference's flat row order from the group views, on the device."""
    h = views["header"]
    first = h[:, 3].long()
    nrows = ((h[:, 0] >> 2) & 3) + 1
    lam = torch.zeros(n_rows, dtype=torch.float32, device=h.device)
    act = torch.zeros(n_rows, dtype=torch.int32, device=h.device)
    for r in range(3):
        m = (nrows > r) & (first + r < n_rows)
        idx = first[m] + r
        lam[idx] = views["lam"][m, r]
        act[idx] = (views["state_bits"][m] >> r) & 1
    return lam, act


def test_torch_views_share_device_memory():
    s = scenes.random_pile(80, seed=4, lattice=(5, 4, 4), iterations=5)
    for _ in range(6):
        s.step()
    v = s.torch_views()
    assert v["x"].device.type == "cuda" and v["x"].shape == (80, 3)
    # reading: a torch op on the view equals the host readers
    assert torch.equal(v["x"].cpu(), torch.from_numpy(s.positions()))
    assert torch.equal(v["q"].cpu(), torch.from_numpy(s.orientations()))
    assert torch.equal(v["v"].cpu(), torch.from_numpy(s.v.numpy()))
    assert torch.equal(v["omega"].cpu(), torch.from_numpy(s.omega.numpy()))
    assert float(v["x"][:, 1].max()) == float(s.positions()[:, 1].max())
    n_live = s._h.read_counters()[0]
    lam, act = _flat_rows(v, n_live)
    rows = s._h.read_rows(n_live)
    assert torch.equal(lam.cpu(), torch.from_numpy(rows["lambda"]))
    assert torch.equal(act.cpu(), torch.from_numpy(rows["active"].astype(np.int32)))
    # writing through a view is writing the solver's state
    v["v"].zero_()
    v["omega"].zero_()
    torch.cuda.synchronize()
    assert not s.v.numpy().any() and not s.omega.numpy().any()
    v["x"][:, 1] += 1.0
    torch.cuda.synchronize()
    assert np.array_equal(s.positions(), v["x"].cpu().numpy())


def test_caller_stream_orders_solver_and_torch_work():
    s = scenes.pyramid(4, seed=5, iterations=6, device="cuda:0")   # resident frames
    ref = scenes.pyramid(4, seed=5, iterations=6, device="cuda:0")
    st = torch.cuda.Stream()
    s.set_stream(st)
    assert s.stream == st.cuda_stream
    with torch.cuda.stream(st):
        for f in range(5):
            s.step()
            ref.step()
            x2 = s.torch_views()["x"] * 2.0   # queued behind the frame on the same stream
            assert torch.equal(x2.cpu(), torch.from_numpy(ref.positions() * 2.0)), f
    s.set_stream(None)
    assert s.stream != st.cuda_stream
    s.step()
    ref.step()
    assert_bit_exact(diff_report(s, ref), "after switching back")
